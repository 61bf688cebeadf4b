// sweep_kernels.hpp -- the single-grid sweep entry points (instantiations
// of the kernel macros of sweep.hpp).  Included by exactly one translation
// unit (efs_engine.hip): the kernels have external linkage.
#pragma once
#include "hooks.hpp"  // SideWork, side_work
#include "sweep.hpp"

namespace efs {

EFS_SWEEP4_INTERIOR(sweep_e4, e4_march, false, false, false)
EFS_SWEEP4_INTERIOR(sweep_e4c, e4_march, false, false, true)
EFS_SWEEP4_INTERIOR(sweep_e4m, e4_march, false, true, false)
EFS_SWEEP4_INTERIOR(sweep_e4mc, e4_march, false, true, true)
EFS_SWEEP4_UPML(sweep_e4u, e4_march, false, false)
EFS_SWEEP4_UPML(sweep_e4uc, e4_march, false, true)
EFS_SWEEP4_FUSED(sweep_e4f, e4_march, false, false, false, float)
EFS_SWEEP4_FUSED(sweep_e4fc, e4_march, false, true, false, float)
EFS_SWEEP4_FUSED(sweep_e4fp, e4_march, false, false, true, float)
EFS_SWEEP4_FUSED(sweep_e4fcp, e4_march, false, true, true, float)
EFS_SWEEP4_PLAIN16(sweep_e4_h, e4_march, false, false)
EFS_SWEEP4_PLAIN16(sweep_e4c_h, e4_march, false, true)
EFS_SWEEP4_FUSED(sweep_e4f_h, e4_march, false, false, false, f16)
EFS_SWEEP4_FUSED(sweep_e4fc_h, e4_march, false, true, false, f16)
EFS_SWEEP4_FUSED(sweep_e4fp_h, e4_march, false, false, true, f16)
EFS_SWEEP4_FUSED(sweep_e4fcp_h, e4_march, false, true, true, f16)
EFS_SWEEP4_INTERIOR(sweep_e4x, e4_march, false, false, kCoefIdx16)
EFS_SWEEP4_INTERIOR(sweep_e4mx, e4_march, false, true, kCoefIdx16)
EFS_SWEEP4_UPML(sweep_e4ux, e4_march, false, kCoefIdx16)
EFS_SWEEP4_FUSED(sweep_e4fx, e4_march, false, kCoefIdx16, false, float)
EFS_SWEEP4_FUSED(sweep_e4fxp, e4_march, false, kCoefIdx16, true, float)
EFS_SWEEP4_PLAIN16(sweep_e4x_h, e4_march, false, kCoefIdx16)
EFS_SWEEP4_FUSED(sweep_e4fx_h, e4_march, false, kCoefIdx16, false, f16)
EFS_SWEEP4_FUSED(sweep_e4fxp_h, e4_march, false, kCoefIdx16, true, f16)
EFS_SWEEP4_INTERIOR(sweep_h4, h4_march, true, false, false)
EFS_SWEEP4_INTERIOR(sweep_h4c, h4_march, true, false, true)
EFS_SWEEP4_INTERIOR(sweep_h4m, h4_march, true, true, false)
EFS_SWEEP4_INTERIOR(sweep_h4mc, h4_march, true, true, true)
EFS_SWEEP4_UPML(sweep_h4u, h4_march, true, false)
EFS_SWEEP4_UPML(sweep_h4uc, h4_march, true, true)
EFS_SWEEP4_FUSED(sweep_h4f, h4_march, true, false, false, float)
EFS_SWEEP4_FUSED(sweep_h4fc, h4_march, true, true, false, float)
EFS_SWEEP4_FUSED(sweep_h4fp, h4_march, true, false, true, float)
EFS_SWEEP4_FUSED(sweep_h4fcp, h4_march, true, true, true, float)
EFS_SWEEP4_PLAIN16(sweep_h4_h, h4_march, true, false)
EFS_SWEEP4_PLAIN16(sweep_h4c_h, h4_march, true, true)
EFS_SWEEP4_FUSED(sweep_h4f_h, h4_march, true, false, false, f16)
EFS_SWEEP4_FUSED(sweep_h4fc_h, h4_march, true, true, false, f16)
EFS_SWEEP4_FUSED(sweep_h4fp_h, h4_march, true, false, true, f16)
EFS_SWEEP4_FUSED(sweep_h4fcp_h, h4_march, true, true, true, f16)
EFS_SWEEP4_INTERIOR(sweep_h4x, h4_march, true, false, kCoefIdx16)
EFS_SWEEP4_INTERIOR(sweep_h4mx, h4_march, true, true, kCoefIdx16)
EFS_SWEEP4_UPML(sweep_h4ux, h4_march, true, kCoefIdx16)
EFS_SWEEP4_FUSED(sweep_h4fx, h4_march, true, kCoefIdx16, false, float)
EFS_SWEEP4_FUSED(sweep_h4fxp, h4_march, true, kCoefIdx16, true, float)
EFS_SWEEP4_PLAIN16(sweep_h4x_h, h4_march, true, kCoefIdx16)
EFS_SWEEP4_FUSED(sweep_h4fx_h, h4_march, true, kCoefIdx16, false, f16)
EFS_SWEEP4_FUSED(sweep_h4fxp_h, h4_march, true, kCoefIdx16, true, f16)

// Each grid kernel and its side-work twin (NAME_s).
template <typename T>
using SweepFn = void (*)(EFS_SWEEP4_PARAMS(T));
template <typename T>
struct SweepPair {
    SweepFn<T> plain, side;
};
#define EFS_PAIR(N) {N, N##_s}
inline const SweepPair<float> kSweepPairs32[] = {
    EFS_PAIR(sweep_e4),   EFS_PAIR(sweep_e4c),  EFS_PAIR(sweep_e4m),  EFS_PAIR(sweep_e4mc), EFS_PAIR(sweep_e4f),
    EFS_PAIR(sweep_e4fc), EFS_PAIR(sweep_e4fp), EFS_PAIR(sweep_e4fcp), EFS_PAIR(sweep_e4x), EFS_PAIR(sweep_e4mx),
    EFS_PAIR(sweep_e4fx), EFS_PAIR(sweep_e4fxp), EFS_PAIR(sweep_h4),  EFS_PAIR(sweep_h4c),  EFS_PAIR(sweep_h4m),
    EFS_PAIR(sweep_h4mc), EFS_PAIR(sweep_h4f),  EFS_PAIR(sweep_h4fc), EFS_PAIR(sweep_h4fp), EFS_PAIR(sweep_h4fcp),
    EFS_PAIR(sweep_h4x),  EFS_PAIR(sweep_h4mx), EFS_PAIR(sweep_h4fx), EFS_PAIR(sweep_h4fxp)};
inline const SweepPair<f16> kSweepPairs16[] = {
    EFS_PAIR(sweep_e4_h),   EFS_PAIR(sweep_e4c_h),  EFS_PAIR(sweep_e4f_h),  EFS_PAIR(sweep_e4fc_h),
    EFS_PAIR(sweep_e4fp_h), EFS_PAIR(sweep_e4fcp_h), EFS_PAIR(sweep_e4x_h), EFS_PAIR(sweep_e4fx_h),
    EFS_PAIR(sweep_e4fxp_h), EFS_PAIR(sweep_h4_h),  EFS_PAIR(sweep_h4c_h),  EFS_PAIR(sweep_h4f_h),
    EFS_PAIR(sweep_h4fc_h), EFS_PAIR(sweep_h4fp_h), EFS_PAIR(sweep_h4fcp_h), EFS_PAIR(sweep_h4x_h),
    EFS_PAIR(sweep_h4fx_h), EFS_PAIR(sweep_h4fxp_h)};
#undef EFS_PAIR

// The twin of k (the kernel to launch when the grid carries side work), or
// null when k has none.
template <typename T, size_t N>
SweepFn<T> side_twin(SweepFn<T> k, const SweepPair<T> (&tab)[N]) {
    for (const SweepPair<T>& p : tab)
        if (p.plain == k) return p.side;
    return nullptr;
}
inline SweepFn<float> side_twin(SweepFn<float> k) { return side_twin(k, kSweepPairs32); }
inline SweepFn<f16> side_twin(SweepFn<f16> k) { return side_twin(k, kSweepPairs16); }

#undef EFS_SWEEP4_INTERIOR
#undef EFS_SWEEP4_UPML
#undef EFS_SWEEP4_FUSED
#undef EFS_SWEEP4_PLAIN16
#undef EFS_SWEEP4_PARAMS
#undef EFS_SWEEP4_SIDE
#undef EFS_SWEEP4_ARGS
#undef EFS_SWEEP4_KERNELS

}  // namespace efs
