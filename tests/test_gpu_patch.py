"""The patch-antenna workload (86 x 87 x 71 graded mesh, Mur on six faces,
lumped port, NF2FF at 2.4 GHz) on the HIP engine against the CPU oracle.

Tolerances: fields and port traces relative L2 <= 1e-5 (FP32 engines);
NF2FF accumulators 2e-5 (FP32 phase arithmetic over the sampled steps).
The port chain then gives the same S11 from both traces.
"""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu

STEPS = 300


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def patch():
    from efs_amd import workloads
    return workloads.patch_antenna()


def run(eng, w, steps):
    from efs_amd import ports, workloads
    slices = workloads.configure_patch(eng, w, max_samples=steps)
    eng.iterate(steps)
    r = eng.readProbeGatherBuffered()
    if isinstance(r, dict):
        data, ns, stride = r["data"], r["numSamples"], r["stride"]
    else:
        data, ns, stride = r.ravel(), r.shape[0], r.shape[1]
    tr = ports.probe_traces(data, ns, stride, w["probes"], slices, 1, w["dT"])
    return tr


def test_patch_matches_oracle(patch, oracle_mod):
    from efs_amd import engine, ports
    engine.load_library()
    say("patch", patch["gridSize"], "steps", STEPS)
    g = engine.HIPEngine()
    assert g.initGPU()
    tg = run(g, patch, STEPS)
    gf = g.getFields()
    ga = g.readNF2FFAccumulation()
    o = oracle_mod.Oracle(patch["gridSize"], patch["coefficients"])
    to = run(o, patch, STEPS)
    of = o.getFields()
    oe, oh = o.readNF2FF()
    for key in ("volt", "curr"):
        err = rel_l2(np.asarray(gf[key], np.float64), np.asarray(of[key], np.float64))
        say(key, err)
        assert err <= 1e-5
    for k in range(2):
        np.testing.assert_allclose(tg[k][0], to[k][0], rtol=0, atol=0)
        err = rel_l2(tg[k][1], to[k][1])
        say("probe", k, err)
        assert err <= 1e-5
    errE = rel_l2(np.asarray(ga["accumE"][0], np.float64), oe[0])
    errH = rel_l2(np.asarray(ga["accumH"][0], np.float64), oh[0])
    say("nf2ff", errE, errH)
    assert errE <= 2e-5 and errH <= 2e-5
    freqs = np.linspace(1e9, 3e9, 21)
    s_g = ports.port_s11(tg[0], tg[1], freqs)
    s_o = ports.port_s11(to[0], to[1], freqs)
    assert np.max(np.abs(s_g - s_o)) <= 1e-4 * np.max(np.abs(s_o))
    flags = g.pathFlags()
    assert "vec4" in flags  # Nz = 71 runs the 4-wide sweeps on rows padded to 72
    g.destroy()


def test_patch_reference_checks(patch):
    """The reference's own checks of this antenna (examples/
    patch_antenna_test.mjs:56-90) on the full 30,000-step HIP run: S11
    below -10 dB somewhere in 1-3 GHz, the resonance (deepest S11) between
    1.8 and 2.5 GHz, and Re(Zin) > 0 there.  Then the directivity at
    2.4 GHz: the surface Poynting power and the far-field intensity agree,
    and Dmax lies within 1 dB of the native 6.6 dBi the reference reports."""
    from efs_amd import engine, farfield, ports, workloads
    engine.load_library()
    steps = patch["nTS"]
    say("patch full run", patch["gridSize"], steps)
    g = engine.HIPEngine()
    assert g.initGPU()
    slices = workloads.configure_patch(g, patch, nf2ff=True, max_samples=steps)
    g.iterate(steps)
    # directivity at the NF2FF frequency (2.4 GHz) on the reference's grid
    theta, phi = farfield.sphere_grid()
    ff = g.computeNF2FFfarField({"pointMeta": patch["pointMeta"], "theta": theta, "phi": phi,
                                 "center": [0, 0, 0], "frequency": float(patch["freqs"][0]), "radius": 1.0})
    acc = g.readNF2FFAccumulation()
    P = patch["nf2ff"]["numPoints"]
    p_surf = farfield.rad_power(acc["accumE"][0], acc["accumH"][0], patch["pointMeta"], P)
    p_sphere = farfield.intensity_power(ff, theta, phi)
    d = farfield.directivity(ff, len(theta), len(phi), p_surf)
    say("Dmax", d["DmaxdBi"], "dBi; P_surface", p_surf, "P_sphere", p_sphere)
    # the surface flux and the far-field intensity carry the same power
    assert abs(p_surf - p_sphere) <= 0.05 * p_surf
    # the reference's native openEMS result for this antenna is 6.6 dBi
    # (STATUS.md:100); the mesh timestep and operator details differ here
    assert 5.6 <= d["DmaxdBi"] <= 7.6
    r = g.readProbeGatherBuffered()
    tr = ports.probe_traces(r["data"], r["numSamples"], r["stride"], patch["probes"], slices, 1, patch["dT"])
    f = np.linspace(1e9, 3e9, 401)
    s = ports.s11_at(tr[0], tr[1], f)
    k = int(np.argmin(s["s11_dB"]))
    say("min S11", s["s11_dB"][k], "dB at", f[k] / 1e9, "GHz, Zin", s["Z"][k])
    assert s["s11_dB"][k] < -10
    assert 1.8e9 <= f[k] <= 2.5e9
    assert s["Z"][k].real > 0
    g.destroy()


def test_patch_from_openems_xml(patch):
    """The same antenna read from the openEMS XML the reference's script API
    writes (efs_amd/openems_xml.py) passes the reference's own checks on the
    full HIP run, with the resonance within 1 % of the builder's run.  Zin
    differs (about 34 ohm against 46 ohm): the script API widens the current
    probe box two mesh lines to each side of the port (app/ems-api.mjs:
    630-662), so its loop also encloses the displacement current of the
    fringing field around the feed."""
    from efs_amd import engine, openems_xml, ports, workloads

    import openems_docs
    engine.load_library()
    w = openems_xml.build(openems_docs.patch_xml(patch))
    steps = w["nTS"]
    f = np.linspace(1e9, 3e9, 401)
    res = []
    for cfg in (patch, w):
        say("patch", "xml" if cfg is w else "builder", steps)
        g = engine.HIPEngine()
        assert g.initGPU()
        slices = workloads.configure_patch(g, cfg, nf2ff=False, max_samples=steps)
        g.iterate(steps)
        r = g.readProbeGatherBuffered()
        tr = ports.probe_traces(r["data"], r["numSamples"], r["stride"], cfg["probes"], slices, 1, cfg["dT"])
        g.destroy()
        s = ports.s11_at(tr[0], tr[1], f)
        k = int(np.argmin(s["s11_dB"]))
        say("min S11", s["s11_dB"][k], "dB at", f[k] / 1e9, "GHz, Zin", s["Z"][k])
        res.append((f[k], s["s11_dB"][k], s["Z"][k]))
    (fb, _, _), (fx, sx, zx) = res
    assert sx < -10 and 1.8e9 <= fx <= 2.5e9 and zx.real > 0
    assert abs(fx - fb) <= 0.01 * fb
