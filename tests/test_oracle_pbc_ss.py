"""Pins the oracle's steady-state detection and periodic boundaries.

Steady state (CPUFDTDEngine, src/webgpu-fdtd.mjs:1037-1095): the reference's
own known answers, tests/test_webgpu.mjs:2565-2674.

Periodic boundaries (WebGPUEngine only, src/webgpu-engine.mjs:475-584,
1976-2155; no reference test exists): with the field update switched off
(vv = ii = 1, vi = iv = 0) the source faces stay constant, so
- phase 0: the destination face is an exact copy of the source face
  (E: face N-2 -> face 0, all tangential cells; H: face 0 -> face N-2 over
  (N0-1) x (N1-1));
- Bloch phase p: the pair (destination, imaginary part) is the source
  rotated by p each step, z_n = exp(j p) (s + j imag_{n-1}), so with
  imag_0 = 0: imag_n = s sin p (1 - cos^n p) / (1 - cos p) and
  dest_n = s cos p - sin p imag_{n-1} -- a closed form the per-step code
  never evaluates.
"""
import numpy as np
import pytest

VARIANTS = [False, True]


def lin(dims, n, x, y, z):
    nx, ny, nz = dims
    return n * nx * ny * nz + x * ny * nz + y * nz + z


def free_space(orc, dims, fp64):
    N = int(np.prod(dims))
    co = {"vv": np.ones(3 * N), "vi": np.full(3 * N, 0.5), "ii": np.ones(3 * N), "iv": np.full(3 * N, 0.5)}
    return orc.Oracle(dims, co, fp64=fp64)


def static(orc, dims, fp64=False):
    N = int(np.prod(dims))
    co = {"vv": np.ones(3 * N), "vi": np.zeros(3 * N), "ii": np.ones(3 * N), "iv": np.zeros(3 * N)}
    return orc.Oracle(dims, co, fp64=fp64)


@pytest.mark.parametrize("fp64", VARIANTS)
def test_ss_period1(oracle_mod, fp64):
    dims = (4, 4, 4)
    e = free_space(oracle_mod, dims, fp64)
    v = np.zeros(3 * 64)
    i0 = lin(dims, 0, 1, 1, 1)
    v[i0] = 3.0
    e.uploadFields(v, np.zeros(3 * 64))
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 5, "threshold": 1e-6})
    e.ssSet(True, 0)
    e.accumulateEnergy()
    s = e.ssRead()
    assert abs(s["energy_period1"][0] - 9.0) < 1e-7
    assert abs(s["energy_period2"][0]) < 1e-10


@pytest.mark.parametrize("fp64", VARIANTS)
def test_ss_period2(oracle_mod, fp64):
    dims = (4, 4, 4)
    e = free_space(oracle_mod, dims, fp64)
    v = np.zeros(3 * 64)
    i0 = lin(dims, 0, 1, 1, 1)
    v[i0] = 2.0
    e.uploadFields(v, np.zeros(3 * 64))
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 3, "threshold": 1e-6})
    e.ssSet(True, 3)
    e.accumulateEnergy()
    s = e.ssRead()
    assert abs(s["energy_period2"][0] - 4.0) < 1e-7
    assert abs(s["energy_period1"][0]) < 1e-10
    assert s["currentSample"] == 4


@pytest.mark.parametrize("fp64", VARIANTS)
def test_ss_not_recording(oracle_mod, fp64):
    dims = (4, 4, 4)
    e = free_space(oracle_mod, dims, fp64)
    v = np.zeros(3 * 64)
    i0 = lin(dims, 0, 1, 1, 1)
    v[i0] = 100.0
    e.uploadFields(v, np.zeros(3 * 64))
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 5, "threshold": 1e-6})
    e.accumulateEnergy()
    s = e.ssRead()
    assert s["energy_period1"][0] == 0.0 and s["energy_period2"][0] == 0.0


def test_ss_convergence_rule(oracle_mod):
    """checkConvergence: both periods recorded and max |e2-e1|/|e1| < threshold
    (tests/test_webgpu.mjs:2620-2645, energies reached by accumulation)."""
    dims = (4, 4, 4)
    e = static(oracle_mod, dims)
    i0 = lin(dims, 0, 1, 1, 1)
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 2, "threshold": 1e-3})
    v = np.zeros(3 * 64)

    def run(a, b):
        e.ssSet(True, 0)
        e.configureSteadyState({"probe_idx": [i0], "periodSamples": 2, "threshold": 1e-3})
        e.ssSet(True, 0)
        for val in (a, a, b, b):
            v[i0] = val
            e.uploadFields(v, np.zeros(3 * 64))
            assert not e.checkConvergence() or e.ssRead()["currentSample"] >= 4
            e.accumulateEnergy()
        return e.checkConvergence()
    assert run(10.0, 10.0000002)        # ratio 4e-8 < 1e-3
    assert not run(10.0, 10.5)          # ratio ~0.1
    # fewer than 2 periods recorded: never converged
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 2, "threshold": 1e-3})
    e.ssSet(True, 0)
    e.accumulateEnergy()
    assert not e.checkConvergence()


def test_ss_records_in_step_before_the_voltage_update(oracle_mod):
    """Steady-state accumulation is the first pre-voltage hook: it sees the
    field of the previous step."""
    dims = (5, 5, 5)
    e = free_space(oracle_mod, dims, False)
    N = 125
    rng = np.random.default_rng(2)
    e.uploadFields(rng.normal(size=3 * N).astype(np.float32), rng.normal(size=3 * N).astype(np.float32))
    i0 = lin(dims, 1, 2, 2, 2)
    e.configureSteadyState({"probe_idx": [i0], "periodSamples": 3, "threshold": 1e-6})
    e.ssSet(True, 0)
    want = 0.0
    for _ in range(3):
        v = e.getFields()["volt"][i0]
        want = float(np.float32(want + float(v) * float(v)))
        e.step()
    assert e.ssRead()["energy_period1"][0] == want


def pbc_faces(dims, axis):
    t0, t1 = (axis + 1) % 3, (axis + 2) % 3
    return t0, t1, dims[t0], dims[t1]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pbc_phase0_is_an_exact_copy(oracle_mod, axis):
    dims = (6, 5, 7)
    N = int(np.prod(dims))
    rng = np.random.default_rng(10 + axis)
    v0 = rng.normal(size=3 * N).astype(np.float32)
    c0 = rng.normal(size=3 * N).astype(np.float32)
    e = static(oracle_mod, dims)
    e.uploadFields(v0, c0)
    e.configurePBC({"axes": [{"axis": axis, "phase": 0}]})
    e.step()
    f = e.getFields()
    t0, t1, f0, f1 = pbc_faces(dims, axis)
    vw, cw = v0.astype(np.float64).copy(), c0.astype(np.float64).copy()
    for i0 in range(f0):
        for i1 in range(f1):
            for comp in (t0, t1):
                ps, pd = [0, 0, 0], [0, 0, 0]
                ps[axis], pd[axis] = dims[axis] - 2, 0
                ps[t0] = pd[t0] = i0
                ps[t1] = pd[t1] = i1
                vw[lin(dims, comp, *pd)] = vw[lin(dims, comp, *ps)]
                if i0 < f0 - 1 and i1 < f1 - 1:
                    ps[axis], pd[axis] = 0, dims[axis] - 2
                    cw[lin(dims, comp, *pd)] = cw[lin(dims, comp, *ps)]
    np.testing.assert_array_equal(f["volt"], vw)
    np.testing.assert_array_equal(f["curr"], cw)


@pytest.mark.parametrize("axis", [0, 2])
def test_pbc_bloch_closed_form(oracle_mod, axis):
    dims = (6, 5, 7)
    N = int(np.prod(dims))
    rng = np.random.default_rng(20 + axis)
    v0 = rng.normal(size=3 * N).astype(np.float32)
    c0 = rng.normal(size=3 * N).astype(np.float32)
    phase = 0.7
    cs, sn = float(np.float32(np.cos(phase))), float(np.float32(np.sin(phase)))
    e = static(oracle_mod, dims, fp64=True)
    e.uploadFields(v0, c0)
    e.configurePBC({"axes": [{"axis": axis, "phase": phase}]})
    n = 9
    e.iterate(n)
    f = e.getFields()

    def imag(k, s):
        return s * sn * (1 - cs ** k) / (1 - cs)

    t0, t1, f0, f1 = pbc_faces(dims, axis)
    for i0 in range(f0):
        for i1 in range(f1):
            for comp in (t0, t1):
                ps, pd = [0, 0, 0], [0, 0, 0]
                ps[axis], pd[axis] = dims[axis] - 2, 0
                ps[t0] = pd[t0] = i0
                ps[t1] = pd[t1] = i1
                s = float(v0[lin(dims, comp, *ps)])
                want = s * cs - sn * imag(n - 1, s)
                assert abs(f["volt"][lin(dims, comp, *pd)] - want) <= 1e-12 * max(1.0, abs(s))
                if i0 < f0 - 1 and i1 < f1 - 1:
                    ps[axis], pd[axis] = 0, dims[axis] - 2
                    s = float(c0[lin(dims, comp, *ps)])
                    want = s * cs - sn * imag(n - 1, s)
                    assert abs(f["curr"][lin(dims, comp, *pd)] - want) <= 1e-12 * max(1.0, abs(s))
