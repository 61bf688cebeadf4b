"""Pins the oracle's NF2FF accumulation, far field, energy and buffered
probe gather against what mathematics fixes, independently of the oracle's
own code path:

- NF2FF accumulation (src/webgpu-engine.mjs:637-800, 2729-2838): with the
  field update switched off (vv = ii = 1, vi = iv = 0) the surface fields are
  constant in time, so each accumulator is a field value times a finite
  geometric series, sum_{t=1..K} exp(-j w t dT) (E) or exp(-j w (t+1/2) dT)
  (H); the Hann window expands into three such series.  A field linear in
  x, y, z makes the cell interpolation equal the field at the face / edge
  midpoint, which pins every neighbour index.
- Far field (src/webgpu-engine.mjs:802-986): single surface points against
  the textbook radiation integrals (Balanis, Antenna Theory, 12-10/12-12):
  J = n x H, M = -n x E, N = J A exp(+j k r.r'), E_theta =
  -j k exp(-j k r)/(4 pi r) (L_phi + eta N_theta), E_phi = +j k exp(-j k r)/
  (4 pi r) (L_theta - eta N_phi), with numpy cross products and complex
  arithmetic; and the Hertzian-dipole directivity 1.5.
- Energy (src/webgpu-engine.mjs:3039-3117): brute-force sum over the
  (Nx-1)(Ny-1)(Nz-1) interior.
- Buffered gather (src/webgpu-engine.mjs:2615-2687): each stored row equals
  the fields read after the step it belongs to.
"""
import numpy as np
import pytest

C0 = 299792458.0
ETA = 376.73031346177066


def static_oracle(orc, dims, fp64=True):
    N = int(np.prod(dims))
    co = {"vv": np.ones(3 * N), "vi": np.zeros(3 * N), "ii": np.ones(3 * N), "iv": np.zeros(3 * N)}
    return orc.Oracle(dims, co, fp64=fp64)


def linear_fields(dims, a, b):
    """volt[n] = a[n].(x, y, z) + a[n][3], curr likewise from b."""
    nx, ny, nz = dims
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    v = np.concatenate([(a[n][0] * x + a[n][1] * y + a[n][2] * z + a[n][3]).ravel() for n in range(3)])
    c = np.concatenate([(b[n][0] * x + b[n][1] * y + b[n][2] * z + b[n][3]).ravel() for n in range(3)])
    return v, c


def geom(w, dT, K, shift=0.0):
    """sum_{t=1..K} exp(-j w (t + shift) dT), closed form."""
    q = np.exp(-1j * w * dT)
    if abs(1 - q) < 1e-15:
        return K * np.exp(-1j * w * shift * dT)
    return np.exp(-1j * w * shift * dT) * q * (1 - q ** K) / (1 - q)


@pytest.mark.parametrize("window", [0, 1])
def test_nf2ff_accumulation_closed_form(oracle_mod, window):
    dims = (7, 6, 5)
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (3, 4))
    b = rng.uniform(-1, 1, (3, 4))
    o = static_oracle(oracle_mod, dims)
    v, c = linear_fields(dims, a, b)
    o.uploadFields(v, c)
    pel = [np.full(d, 2.0) for d in dims]    # primal edges
    dul = [np.full(d, 0.5) for d in dims]    # dual edges (equal: interpolation at the midpoint)
    pts = np.array([[1, 2, 3], [4, 1, 0], [0, 0, 0], [6, 2, 2], [3, 5, 1], [2, 3, 4]], np.uint32)
    omegas = np.array([2 * np.pi * 1.3e9, 2 * np.pi * 4.1e9, 0.0])
    dT, K, M = 1.7e-11, 23, 40
    o.configureNF2FF(pts.ravel(), omegas, dT, pel, dul, maxTS=M, windowType=window)
    o.iterate(K)
    accE, accH = o.readNF2FF()
    # expected interpolated fields
    nx, ny, nz = dims
    for p, (ix, iy, iz) in enumerate(pts):
        Ee = np.zeros(3)
        if ix < nx - 1 and iy < ny - 1 and iz < nz - 1:
            centre = [(ix, iy + .5, iz + .5), (ix + .5, iy, iz + .5), (ix + .5, iy + .5, iz)]
            Ee = np.array([a[n][:3] @ np.array(centre[n]) + a[n][3] for n in range(3)]) / 2.0
        He = np.zeros(3)
        pos = (ix, iy, iz)
        for n in range(3):
            r = np.array(pos, float)
            if pos[n] < dims[n] - 1:
                r[n] += 0.5                   # midpoint between the two dual edges
            He[n] = (b[n][:3] @ r + b[n][3]) / 0.5
        for f, w in enumerate(omegas):
            if window == 0:
                sE, sH = geom(w, dT, K), geom(w, dT, K, 0.5)
                norm = 1.0
            else:
                # Hann: 0.5 - 0.25 e^{+j W t} - 0.25 e^{-j W t}, W = 2 pi / (M-1) / dT in w units
                W = 2 * np.pi / (M - 1) / dT
                sE = 0.5 * geom(w, dT, K) - 0.25 * geom(w - W, dT, K) - 0.25 * geom(w + W, dT, K)
                # H phase uses t + 1/2 while the window uses t: shift only the carrier
                sH = (0.5 * geom(w, dT, K) - 0.25 * geom(w - W, dT, K) - 0.25 * geom(w + W, dT, K)) \
                    * np.exp(-1j * w * 0.5 * dT)
                norm = 2.0 / (M - 1)        # 1 / sum_{n=0}^{M-1} hann(n)
            for n in range(3):
                gotE = accE[f, 6 * p + 2 * n] + 1j * accE[f, 6 * p + 2 * n + 1]
                gotH = accH[f, 6 * p + 2 * n] + 1j * accH[f, 6 * p + 2 * n + 1]
                wantE = Ee[n] * sE * dT * norm
                wantH = He[n] * sH * dT * norm
                scale = max(abs(wantE), 1e-3 * dT)
                assert abs(gotE - wantE) <= 1e-9 * scale + 1e-22, (p, f, n, gotE, wantE)
                scale = max(abs(wantH), 1e-3 * dT)
                assert abs(gotH - wantH) <= 1e-9 * scale + 1e-22, (p, f, n, gotH, wantH)


def far_closed_form(accE, accH, meta, theta, phi, center, freq, radius):
    k = 2 * np.pi * freq / C0
    out = []
    for th in theta:
        for ph in phi:
            rhat = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
            that = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
            phat = np.array([-np.sin(ph), np.cos(ph), 0.0])
            Nv = np.zeros(3, complex)
            Lv = np.zeros(3, complex)
            for p in range(len(meta) // 8):
                m = meta[8 * p: 8 * p + 8]
                nvec = np.zeros(3)
                nvec[int(m[3])] = m[4]
                E = accE[6 * p: 6 * p + 6: 2] + 1j * accE[6 * p + 1: 6 * p + 6: 2]
                H = accH[6 * p: 6 * p + 6: 2] + 1j * accH[6 * p + 1: 6 * p + 6: 2]
                J = np.cross(nvec, H)
                Mc = -np.cross(nvec, E)
                ph_ = np.exp(1j * k * rhat @ (m[:3] - center)) * m[5]
                Nv += J * ph_
                Lv += Mc * ph_
            pre = 1j * k * np.exp(-1j * k * radius) / (4 * np.pi * radius)
            Et = -pre * (Lv @ phat + ETA * (Nv @ that))
            Ep = pre * (Lv @ that - ETA * (Nv @ phat))
            out.append([(abs(Et) ** 2 + abs(Ep) ** 2) / (2 * ETA), Et.real, Et.imag, Ep.real, Ep.imag])
    return np.array(out).ravel()


def test_farfield_textbook_radiation_integrals(oracle_mod):
    rng = np.random.default_rng(8)
    P = 9
    meta = np.zeros(8 * P)
    for p in range(P):
        meta[8 * p: 8 * p + 3] = rng.uniform(-0.03, 0.03, 3)
        meta[8 * p + 3] = p % 3
        meta[8 * p + 4] = 1.0 if p % 2 else -1.0
        meta[8 * p + 5] = rng.uniform(1e-6, 4e-6)
    accE = rng.normal(size=6 * P)
    accH = rng.normal(size=6 * P) / ETA
    theta = np.linspace(0.1, 3.0, 7)
    phi = np.linspace(0.0, 6.0, 5)
    center = np.array([0.004, -0.002, 0.001])
    freq, radius = 2.4e9, 1.0
    got = oracle_mod.farfield(accE, accH, meta, theta, phi, center, freq, radius)
    want = far_closed_form(accE, accH, meta, theta, phi, center, freq, radius)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_farfield_hertzian_dipole_directivity(oracle_mod):
    """A z-directed current element (one point, normal y, H_x chosen so
    J = n x H points along z) radiates sin^2(theta): directivity 1.5."""
    meta = np.array([0, 0, 0, 1, 1.0, 1e-6, 0, 0], float)   # normal +y
    accE = np.zeros(6)
    accH = np.zeros(6)
    accH[0] = -1.0      # y x (-x) = +z
    theta = (np.arange(181) + 0.5) * np.pi / 181
    phi = (np.arange(36) + 0.5) * 2 * np.pi / 36
    out = oracle_mod.farfield(accE, accH, meta, theta, phi, np.zeros(3), 1e9, 1.0).reshape(181, 36, 5)
    U = out[:, :, 0]
    dth, dph = np.pi / 181, 2 * np.pi / 36
    Prad = (U * np.sin(theta)[:, None]).sum() * dth * dph
    D = 4 * np.pi * U.max() / Prad
    assert abs(D - 1.5) < 1e-3
    # and the pattern is sin^2(theta), independent of phi
    np.testing.assert_allclose(U / U.max(), np.broadcast_to((np.sin(theta) ** 2 / np.sin(theta).max() ** 2)[:, None],
                                                            U.shape), rtol=1e-9)


def test_energy_brute_force(oracle_mod):
    dims = (6, 5, 7)
    rng = np.random.default_rng(1)
    o = static_oracle(oracle_mod, dims)
    N = int(np.prod(dims))
    v, c = rng.normal(size=3 * N), rng.normal(size=3 * N)
    o.uploadFields(v, c)
    V = v.reshape(3, *dims)[:, :-1, :-1, :-1]
    Cc = c.reshape(3, *dims)[:, :-1, :-1, :-1]
    want = float((V ** 2).sum() + (Cc ** 2).sum())
    assert abs(o.energy() - want) <= 1e-12 * want
    # the last planes do not count
    v2 = v.reshape(3, *dims).copy()
    v2[:, -1, :, :] += 100.0
    v2[:, :, -1, :] += 100.0
    v2[:, :, :, -1] += 100.0
    o.uploadFields(v2.ravel(), c)
    assert abs(o.energy() - want) <= 1e-12 * want


def test_gather_rows_are_fields_after_their_step(oracle_mod):
    from cases import configure, random_case
    case = random_case(5, (8, 7, 6), pml_width=2, mur=True)
    idx = np.array([0, 17, 3 * 336 - 1, 400, 77], np.uint32)
    interval, max_samples, steps = 3, 4, 17

    def make():
        o = oracle_mod.Oracle(case["dims"], case["coeffs"])
        configure(o, case)
        return o
    a = make()
    a.configureProbeGatherBuffered(idx, max_samples, interval)
    a.iterate(steps)
    rows = a.readProbeGatherBuffered()
    b = make()
    want = []
    for s in range(steps):
        b.iterate(1)
        if (s + 1) % interval == 0 and len(want) < max_samples:
            f = b.getFields()
            want.append(np.stack([f["volt"][idx], f["curr"][idx]], 1).ravel())
    assert rows.shape == (max_samples, 2 * len(idx))
    np.testing.assert_array_equal(rows, np.array(want))
