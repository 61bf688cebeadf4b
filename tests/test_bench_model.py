"""The benchmark's algorithmic byte count per sweep launch (bench.py):
FP32 fields 60 B per cell; FP16 fields 9 x 2 B plus 6 coefficients, at 2 B
when they are indexed (byte or 16-bit indices) and 4 B as FP32 arrays;
60 B per UPML cell on top in every form."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_sweep_bytes_per_form():
    n, p = 1000, 100
    assert bench.sweep_bytes(n, p, "fp32", ("coef_idx",)) == 60 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp32") == 60 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("coef_idx",)) == 30 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("coef_idx16", "f16")) == 30 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("f16", "fused")) == 42 * n + 60 * p


def test_byte_model_names_the_form():
    assert bench.byte_model("fp32").startswith("60 B/cell")
    assert bench.byte_model("fp16", ("coef_idx",)).startswith("30 B/cell")
    assert bench.byte_model("fp16", ()).startswith("42 B/cell")


class _CountingEngine:
    def __init__(self):
        self.steps, self.energy_at = 0, []

    def iterate(self, k):
        self.steps += k

    def energy(self):
        self.energy_at.append(self.steps)
        return float(self.steps)


def test_timed_window_always_holds_an_energy_call():
    """The stress bench reads the energy every E steps and once more at the
    end of a window that is not a multiple of E, so a short driver window
    (e.g. 20 steps) still contains the reduction."""
    for K, E, want in ((20, 100, [20]), (2000, 100, list(range(100, 2001, 100))), (150, 100, [100, 150])):
        eng = _CountingEngine()
        got = bench.time_stress_window(eng, K, E)
        assert eng.steps == K
        assert eng.energy_at == want and got == [float(s) for s in want]


def test_timed_window_ends_on_an_nf2ff_fold():
    """The untimed lead-in (warmup + alignment) makes the window end on a
    fold of the K-step NF2FF buffer, so any window holds at least one."""
    kb = 64
    for W, K in ((5, 20), (100, 2000), (0, 1), (63, 64), (7, 128)):
        lead = W + (-(W + K)) % kb
        folds = (lead + K) // kb - lead // kb
        assert (lead + K) % kb == 0 and lead >= W
        assert folds >= 1 and folds in (K // kb, -(-K // kb))
