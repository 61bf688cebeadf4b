"""Pins the CPU oracle against the reference's own known-answer tests.

Every case re-creates a section of tests/test_webgpu.mjs (reference) with the
same grid, coefficients, initial values and expected numbers; the line range
is cited on each test.  Both oracle variants (f32 = JavaScript semantics,
f64) must reproduce them.
"""
import numpy as np
import pytest

VARIANTS = [False, True]


def free_space(nx, ny, nz, vi=0.3, iv=0.3):
    """createFreeSpaceCoefficients (tests/test_webgpu.mjs:62-80)."""
    t = 3 * nx * ny * nz
    return {"vv": np.ones(t), "vi": np.full(t, vi), "ii": np.ones(t), "iv": np.full(t, iv)}


def IDX(n, x, y, z, Nx, Ny, Nz):
    return n * Nx * Ny * Nz + x * Ny * Nz + y * Nz + z


def make(orc, dims, fp64, coeffs=None):
    return orc.Oracle(dims, coeffs or free_space(*dims), fp64=fp64)


def poke(eng, volt=None, curr=None):
    f = eng.getFields()
    for i, v in (volt or {}).items():
        f["volt"][i] = v
    for i, v in (curr or {}).items():
        f["curr"][i] = v
    eng.uploadFields(f["volt"], f["curr"])


@pytest.mark.parametrize("fp64", VARIANTS)
def test_nijk_indexing(oracle_mod, fp64):
    """tests/test_webgpu.mjs:391-437"""
    e = make(oracle_mod, (5, 4, 3), fp64)
    assert e.idx(0, 0, 0, 0) == 0
    assert e.idx(1, 0, 0, 0) == 5 * 4 * 3
    assert e.idx(0, 1, 0, 0) == 4 * 3
    assert e.idx(0, 0, 1, 0) == 3
    assert e.idx(0, 0, 0, 1) == 1
    assert IDX(2, 4, 3, 2, 5, 4, 3) == 3 * 5 * 4 * 3 - 1


@pytest.mark.parametrize("fp64", VARIANTS)
def test_voltage_boundary_shift(oracle_mod, fp64):
    """tests/test_webgpu.mjs:443-489: backward neighbour clamped to self at 0."""
    N = (3, 3, 3)
    e = make(oracle_mod, N, fp64)
    curr = np.zeros(3 * 27)
    for n in range(3):
        for x in range(3):
            for y in range(3):
                for z in range(3):
                    curr[IDX(n, x, y, z, *N)] = (n + 1) * 0.1 * (x + y + z + 1)
    e.uploadFields(np.zeros(81), curr)
    e.updateVoltages()
    v = e.getFields()["volt"]
    c32 = curr.astype(np.float32).astype(np.float64) if not fp64 else curr
    assert abs(v[IDX(0, 0, 0, 0, *N)]) <= 1e-10
    curl = (c32[IDX(2, 1, 1, 1, *N)] - c32[IDX(2, 1, 0, 1, *N)] - c32[IDX(1, 1, 1, 1, *N)]
            + c32[IDX(1, 1, 1, 0, *N)])
    assert abs(v[IDX(0, 1, 1, 1, *N)] - 0.3 * curl) <= 1e-6


@pytest.mark.parametrize("fp64", VARIANTS)
def test_current_loop_bounds(oracle_mod, fp64):
    """tests/test_webgpu.mjs:495-561"""
    N = (4, 4, 4)
    e = make(oracle_mod, N, fp64)
    volt = np.zeros(3 * 64)
    for n in range(3):
        for x in range(4):
            for y in range(4):
                for z in range(4):
                    volt[IDX(n, x, y, z, *N)] = (n + 1) * 0.1
    e.uploadFields(volt, np.zeros(192))
    e.updateCurrents()
    c = e.getFields()["curr"]
    for x in range(4):
        for n in range(3):
            assert c[IDX(n, x, 3, 0, *N)] == 0
            assert c[IDX(n, x, 0, 3, *N)] == 0
    # gradient case: Hx(1,1,1) = 0.3 * curl
    e = make(oracle_mod, N, fp64)
    for n in range(3):
        for x in range(4):
            for y in range(4):
                for z in range(4):
                    volt[IDX(n, x, y, z, *N)] = n * 0.1 + x * 0.01 + y * 0.001 + z * 0.0001
    e.uploadFields(volt, np.zeros(192))
    e.updateCurrents()
    ez11 = 2 * 0.1 + 0.01 + 0.001 + 0.0001
    ez21 = 2 * 0.1 + 0.01 + 0.002 + 0.0001
    ey11 = 0.1 + 0.01 + 0.001 + 0.0001
    ey12 = 0.1 + 0.01 + 0.001 + 0.0002
    assert abs(e.getFields()["curr"][IDX(0, 1, 1, 1, *N)] - 0.3 * (ez11 - ez21 - ey11 + ey12)) <= 1e-6


@pytest.mark.parametrize("fp64", VARIANTS)
def test_full_timestep_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:567-631: Hy = 0.3, Hz = -0.3 after one step."""
    N = (8, 8, 8)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={IDX(0, 4, 4, 4, *N): 1.0})
    e.updateVoltages()
    e.updateCurrents()
    f = e.getFields()
    assert abs(f["volt"][IDX(0, 4, 4, 4, *N)] - 1.0) <= 1e-6
    assert abs(f["curr"][IDX(1, 4, 4, 4, *N)] - 0.3) <= 1e-6
    assert abs(f["curr"][IDX(2, 4, 4, 4, *N)] + 0.3) <= 1e-6
    for _ in range(3):
        e.updateVoltages()
        e.updateCurrents()
    ex = e.getFields()["volt"][: 512]
    assert np.count_nonzero(np.abs(ex) > 1e-10) > 1


@pytest.mark.parametrize("fp64", VARIANTS)
def test_excitation_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:637-710, incl. the beyond-length -> signal[0] rule."""
    N = (5, 5, 5)
    e = make(oracle_mod, N, fp64)
    e.configureExcitation({"signal": [1.0, 0.5, 0.25, 0.125], "delay": [0], "amp": [2.0], "dir": [0],
                           "pos": [2 * 25 + 2 * 5 + 2], "period": 0})
    g = IDX(0, 2, 2, 2, *N)
    for ts, want in ((0, 2.0), (2, 0.5), (10, 2.0)):
        e.uploadFields(np.zeros(375), np.zeros(375))
        e.numTS = ts
        e.applyExcitation()
        assert abs(e.getFields()["volt"][g] - want) <= 1e-6, ts
    # delay (tests/test_webgpu.mjs:716-756)
    e = make(oracle_mod, N, fp64)
    e.configureExcitation({"signal": [1, 2, 3, 4, 5], "delay": [3], "amp": [1.0], "dir": [1],
                           "pos": [25 + 5 + 1], "period": 0})
    g = IDX(1, 1, 1, 1, *N)
    for ts in (3, 1):
        e.uploadFields(np.zeros(375), np.zeros(375))
        e.numTS = ts
        e.applyExcitation()
        assert abs(e.getFields()["volt"][g] - 1.0) <= 1e-6


@pytest.mark.parametrize("fp64", VARIANTS)
def test_pec_invariant(oracle_mod, fp64):
    """tests/test_webgpu.mjs:762-804: vv=vi=0 on x=0 keeps E there at 0."""
    N = (6, 6, 6)
    c = free_space(*N)
    for n in range(3):
        for y in range(6):
            for z in range(6):
                i = IDX(n, 0, y, z, *N)
                c["vv"][i] = 0
                c["vi"][i] = 0
    e = make(oracle_mod, N, fp64, c)
    poke(e, volt={IDX(0, 3, 3, 3, *N): 1.0})
    e.iterate(5)
    v = e.getFields()["volt"]
    for n in range(3):
        assert np.max(np.abs(v[IDX(n, 0, 0, 0, *N): IDX(n, 0, 0, 0, *N) + 36])) <= 1e-10


@pytest.mark.parametrize("fp64", VARIANTS)
def test_pml_voltage_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:1165-1234: 0.5 / 1.5 / 2.25."""
    N = (8, 1, 1)
    e = make(oracle_mod, N, fp64)
    t = 9
    e.configurePML([{"startPos": [0, 0, 0], "numLines": [3, 1, 1], "vv": np.full(t, 0.8),
                     "vvfo": np.full(t, 0.2), "vvfn": np.full(t, 1.5), "ii": np.full(t, 0.9),
                     "iifo": np.full(t, 0.1), "iifn": np.full(t, 1.3)}])
    g = IDX(0, 1, 0, 0, *N)
    p = IDX(0, 1, 0, 0, 3, 1, 1)
    poke(e, volt={g: 2.0})
    vf, cf = e.pmlFlux(0)
    vf[p] = 0.5
    e.setPmlFlux(0, vf, cf)
    e.preVoltageUpdatePML()
    assert abs(e.getFields()["volt"][g] - 0.5) <= 1e-6
    assert abs(e.pmlFlux(0)[0][p] - 1.5) <= 1e-6
    e.updateVoltages()
    e.postVoltageUpdatePML()
    assert abs(e.pmlFlux(0)[0][p] - 0.5) <= 1e-6
    assert abs(e.getFields()["volt"][g] - 2.25) <= 1e-6


@pytest.mark.parametrize("fp64", VARIANTS)
def test_pml_current_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:1240-1300: 0.4 / 1.98 / 0.6 / 2.7."""
    N = (8, 2, 2)
    e = make(oracle_mod, N, fp64)
    t = 3 * 3 * 2 * 2
    e.configurePML([{"startPos": [0, 0, 0], "numLines": [3, 2, 2], "vv": np.ones(t), "vvfo": np.zeros(t),
                     "vvfn": np.ones(t), "ii": np.full(t, 0.7), "iifo": np.full(t, 0.3),
                     "iifn": np.full(t, 1.2)}])
    g = IDX(0, 1, 0, 0, *N)
    p = IDX(0, 1, 0, 0, 3, 2, 2)
    poke(e, curr={g: 3.0})
    vf, cf = e.pmlFlux(0)
    cf[p] = 0.4
    e.setPmlFlux(0, vf, cf)
    e.preCurrentUpdatePML()
    assert abs(e.getFields()["curr"][g] - 0.4) <= 1e-6
    assert abs(e.pmlFlux(0)[1][p] - 1.98) <= 1e-6
    poke(e, curr={g: 0.6})
    e.postCurrentUpdatePML()
    assert abs(e.pmlFlux(0)[1][p] - 0.6) <= 1e-6
    assert abs(e.getFields()["curr"][g] - 2.7) <= 1e-6


def _ade_cfg(orders):
    return {"orders": [{"numCells": 1, "hasLorentz": lor, "directions": [
        {"dir": 0, "pos_idx": [16 + 4 + 1], "v_int_ADE": [vi], "v_ext_ADE": [ve], "v_Lor_ADE": [vl]}]}
        for (lor, vi, ve, vl) in orders]}


@pytest.mark.parametrize("fp64", VARIANTS)
def test_ade_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:1921-2080 (Drude, Lorentz, multi-order, zero)."""
    N = (4, 4, 4)
    g = IDX(0, 1, 1, 1, *N)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 2.0})
    e.configureLorentz(_ade_cfg([(False, 0.8, 0.3, 0.0)]))
    e.updateVoltADE()
    assert abs(e.adeState(0)["volt_ADE"][0] - 0.6) <= 1e-7
    e.updateVoltADE()
    assert abs(e.adeState(0)["volt_ADE"][0] - (0.8 * 0.6 + 0.3 * 2.0)) <= 1e-6

    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 3.0})
    e.configureLorentz(_ade_cfg([(True, 0.9, 0.2, 0.1)]))
    e.updateVoltADE()
    s = e.adeState(0)
    assert abs(s["volt_Lor_ADE"][0]) <= 1e-7 and abs(s["volt_ADE"][0] - 0.6) <= 1e-7
    e.updateVoltADE()
    s = e.adeState(0)
    assert abs(s["volt_Lor_ADE"][0] - 0.06) <= 1e-7
    assert abs(s["volt_ADE"][0] - (0.9 * 0.6 + 0.2 * (3.0 - 0.06))) <= 1e-6

    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 1.0})
    e.configureLorentz(_ade_cfg([(False, 0.5, 0.4, 0.0), (False, 0.7, 0.2, 0.0)]))
    e.updateVoltADE()
    assert abs(e.adeState(0)["volt_ADE"][0] - 0.4) <= 1e-7
    assert abs(e.adeState(1)["volt_ADE"][0] - 0.2) <= 1e-7

    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 5.0})
    e.configureLorentz(_ade_cfg([(False, 0.0, 0.0, 0.0)]))
    e.updateVoltADE()
    assert abs(e.adeState(0)["volt_ADE"][0]) <= 1e-10


@pytest.mark.parametrize("fp64", VARIANTS)
def test_tfsf_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:2081-2198 (no delay, fractional delay, past end, period)."""
    N = (4, 4, 4)
    cases = [  # signal, field, delay_int, delay_frac, amp, period, numTS, expected
        ([1, 2, 3, 4, 5], IDX(0, 1, 1, 1, *N), 0, 0.0, 0.5, 0, 0, 0.5),
        ([0, 1, 2, 3, 4], IDX(1, 2, 2, 2, *N), 1, 0.25, 1.0, 0, 3, 2.25),
        ([1, 2, 3], IDX(0, 1, 1, 1, *N), 0, 0.0, 1.0, 0, 10, 0.0),
        ([10, 20, 30, 40, 50], IDX(0, 1, 1, 1, *N), 0, 0.0, 1.0, 3, 5, 30.0),
    ]
    for sig, fi, di, df, amp, per, ts, want in cases:
        e = make(oracle_mod, N, fp64)
        e.configureTFSF({"signal": sig, "period": per, "voltagePoints": [
            {"field_idx": fi, "delay_int": di, "delay_frac": df, "amp": amp}]})
        e.numTS = ts
        e.applyTFSFVoltage()
        assert abs(e.getFields()["volt"][fi] - want) <= 1e-7, (sig, ts)
    # current points use the same rule on the current field (createTFSFTestConfig, :376-383)
    e = make(oracle_mod, N, fp64)
    e.configureTFSF({"signal": [1, 0.5, 0.25, 0], "period": 0,
                     "voltagePoints": [{"field_idx": 3, "delay_int": 0, "delay_frac": 0, "amp": 1}],
                     "currentPoints": [{"field_idx": 7, "delay_int": 1, "delay_frac": 0.25, "amp": 0.5}]})
    e.numTS = 2
    e.applyTFSFCurrent()
    assert abs(e.getFields()["curr"][7] - 0.5 * (0.75 * 0.5 + 0.25 * 0.25)) <= 1e-7


def _rlc(field_idx, type_flag, **kw):
    d = {"field_idx": field_idx, "direction": 0, "type_flag": type_flag}
    for k in ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2"):
        d[k] = kw.get(k, 0.0)
    return d


@pytest.mark.parametrize("fp64", VARIANTS)
def test_rlc_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:2199-2330 (parallel, series, history, mixed)."""
    N = (4, 4, 4)
    lin = 16 + 4 + 1
    g = IDX(0, 1, 1, 1, *N)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 5.0})
    e.configureRLC({"elements": [_rlc(lin, 0, i2v=0.1, ilv=0.2)]})
    e.updateRLC()
    assert abs(e.rlcState()["VIl"][0]) <= 1e-10
    e.updateRLC()
    assert abs(e.rlcState()["VIl"][0] - 0.1 * 0.2 * 5.0) <= 1e-7

    e = make(oracle_mod, N, fp64)
    poke(e, volt={g: 2.0})
    e.configureRLC({"elements": [_rlc(lin, 1, vvd=0.5, vv2=0.1, ib0=0.3)]})
    e.updateRLC()
    assert abs(e.getFields()["volt"][g] - 1.0) <= 1e-7
    assert abs(e.rlcState()["Jn"][0] - 0.3) <= 1e-7

    e = make(oracle_mod, N, fp64)
    e.configureRLC({"elements": [_rlc(0, 0)]})
    hist = []
    for v in (1.0, 2.0, 3.0):
        poke(e, volt={0: v})
        e.updateRLC()
        hist.append(e.rlcState()["Vdn"][:3].copy())
    assert list(hist[0][:1]) == [1.0]
    assert list(hist[1][:2]) == [2.0, 1.0]
    assert list(hist[2]) == [3.0, 2.0, 1.0]

    e = make(oracle_mod, N, fp64)
    poke(e, volt={0: 1.0, 1: 2.0})
    e.configureRLC({"elements": [_rlc(0, 0, i2v=0.5, ilv=0.3), _rlc(1, 1, vvd=0.8, ib0=0.4)]})
    e.updateRLC()
    assert abs(e.rlcState()["VIl"][0]) <= 1e-10
    assert abs(e.getFields()["volt"][1] - 1.6) <= 1e-7


@pytest.mark.parametrize("fp64", VARIANTS)
def test_mur_known_values(oracle_mod, fp64):
    """tests/test_webgpu.mjs:2330-2520 (3-phase, multi-point, dual-component)."""
    N = (6, 4, 4)
    nrm, sh = IDX(0, 0, 1, 1, *N), IDX(0, 1, 1, 1, *N)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={nrm: 1.0, sh: 2.0})
    e.configureMur({"coeff": 0.3, "normal_idx": [nrm], "shifted_idx": [sh]})
    e.murPreVoltage()
    poke(e, volt={sh: 3.5})
    e.murPostVoltage()
    e.murApply()
    assert abs(e.getFields()["volt"][nrm] - (1.7 + 0.3 * 3.5)) <= 1e-6

    N = (4, 4, 4)
    n0, n1 = IDX(0, 0, 0, 0, *N), IDX(0, 0, 1, 0, *N)
    s0, s1 = IDX(0, 1, 0, 0, *N), IDX(0, 1, 1, 0, *N)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={n0: 10.0, n1: 20.0, s0: 15.0, s1: 25.0})
    e.configureMur({"coeff": 0.4, "normal_idx": [n0, n1], "shifted_idx": [s0, s1]})
    e.murPreVoltage()
    poke(e, volt={s0: 16.0, s1: 26.0})
    e.murPostVoltage()
    e.murApply()
    v = e.getFields()["volt"]
    assert abs(v[n0] - (15.0 - 4.0 + 0.4 * 16.0)) <= 1e-5
    assert abs(v[n1] - (25.0 - 8.0 + 0.4 * 26.0)) <= 1e-5

    N = (6, 6, 6)
    nP, sP = IDX(1, 0, 2, 2, *N), IDX(1, 1, 2, 2, *N)
    nPP, sPP = IDX(2, 0, 2, 2, *N), IDX(2, 1, 2, 2, *N)
    e = make(oracle_mod, N, fp64)
    poke(e, volt={nP: 1.0, sP: 2.0, nPP: 3.0, sPP: 4.0})
    e.configureMur({"coeff_nyP": [0.3], "coeff_nyPP": [0.5], "normal_idx_nyP": [nP], "shifted_idx_nyP": [sP],
                    "normal_idx_nyPP": [nPP], "shifted_idx_nyPP": [sPP]})
    e.murPreVoltage()
    poke(e, volt={sP: 5.0, sPP: 6.0})
    e.murPostVoltage()
    e.murApply()
    v = e.getFields()["volt"]
    assert abs(v[nP] - (1.7 + 0.3 * 5.0)) <= 1e-6
    assert abs(v[nPP] - (2.5 + 0.5 * 6.0)) <= 1e-6


def test_mur_coefficient_closed_form():
    """tests/test_webgpu.mjs:2360-2380: (c dT - d)/(c dT + d) = 0.5."""
    c, dT, d = 3e8, 1e-12, 1e-4
    assert abs((c * dT - d) / (c * dT + d) - 0.5) <= 1e-10
