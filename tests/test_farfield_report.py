"""Far-field reporting (efs_amd/farfield.py) and the NF2FF surface builder.

Pins:
- rad_power against the textbook Poynting flux 0.5 Re(E x H*) . n A,
  computed with numpy cross products, for every normal direction and sign;
- directivity of a Hertzian dipole: the oracle's far field of a single
  z-directed current element gives Dmax = 1.5 (1.76 dBi) with the power
  from integrating U over the sphere (intensity_power);
- the cuts read the dBi map at phi = 0 / 90 degrees (and phi + 180 for
  negative theta), exactly on grid angles;
- the surface builder follows the reference's point order and face area
  rule: the areas of each face sum to the face's area, and the reference's
  2-degree sphere grid is 91 x 181 (the patch case's far-field grid).
"""
import numpy as np
import pytest

from efs_amd import farfield, operator as op


def test_rad_power_is_the_poynting_flux():
    rng = np.random.default_rng(4)
    P = 60
    meta = np.zeros((P, 8))
    meta[:, 3] = rng.integers(0, 3, P)
    meta[:, 4] = rng.choice([-1.0, 1.0], P)
    meta[:, 5] = rng.uniform(1e-6, 3e-6, P)
    E = rng.normal(size=(P, 3)) + 1j * rng.normal(size=(P, 3))
    H = rng.normal(size=(P, 3)) + 1j * rng.normal(size=(P, 3))
    accE = np.stack([E.real, E.imag], axis=2).ravel()
    accH = np.stack([H.real, H.imag], axis=2).ravel()
    got = farfield.rad_power(accE, accH, meta.ravel(), P)
    nhat = np.zeros((P, 3))
    nhat[np.arange(P), meta[:, 3].astype(int)] = meta[:, 4]
    want = np.sum(0.5 * np.real(np.sum(np.cross(E, np.conj(H)) * nhat, axis=1)) * meta[:, 5])
    assert abs(got - want) <= 1e-12 * np.sum(np.abs(want))


def test_hertzian_dipole_directivity(oracle_mod):
    meta = np.array([0, 0, 0, 1, 1.0, 1e-6, 0, 0], float)  # normal +y, J = n x H along z
    accE = np.zeros(6)
    accH = np.zeros(6)
    accH[0] = -1.0
    theta, phi = farfield.sphere_grid(1, 2)
    ff = oracle_mod.farfield(accE, accH, meta, theta, phi, np.zeros(3), 1e9, 1.0)
    p = farfield.intensity_power(ff, theta, phi)
    d = farfield.directivity(ff, len(theta), len(phi), p)
    assert abs(d["Dmax"] - 1.5) < 2e-3
    assert abs(d["DmaxdBi"] - 10 * np.log10(1.5)) < 6e-3


def test_cuts_read_the_map():
    theta, phi = farfield.sphere_grid()
    nT, nP = len(theta), len(phi)
    rng = np.random.default_rng(8)
    e = rng.uniform(0.1, 1.0, nT * nP)
    d = 5.0
    tdeg, xz, yz = farfield.cuts(e, d, nT, nP)
    emax = e.max()
    dmap = 20 * np.log10(e / emax) + d
    for k, t in enumerate(tdeg):
        i = abs(t) // 2
        if t >= 0:
            assert xz[k] == dmap[i * nP + 0] and yz[k] == dmap[i * nP + 45]
        else:
            assert xz[k] == dmap[i * nP + 90] and yz[k] == dmap[i * nP + 135]


def test_surface_builder_order_and_areas():
    mesh = op.Mesh(np.cumsum(np.r_[0, np.linspace(1, 2, 11)]) * 1e-3, np.arange(9) * 1.5e-3,
                   np.cumsum(np.r_[0, np.linspace(2, 1, 10)]) * 1e-3)
    lo, hi = [2, 1, 3], [9, 6, 8]
    pts, meta, _, _ = op.nf2ff_box(mesh, lo, hi)
    pts = pts.reshape(-1, 3)
    meta = meta.reshape(-1, 8)
    k = 0
    for ax in range(3):
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for side in (lo[ax], hi[ax]):
            n = (hi[a1] - lo[a1] + 1) * (hi[a2] - lo[a2] + 1)
            face = pts[k:k + n]
            assert np.all(face[:, ax] == side)
            # x, y, z loop order: the linear index increases along the face
            lin = (face[:, 0] * 100 + face[:, 1]) * 100 + face[:, 2]
            assert np.all(np.diff(lin) > 0)
            span1 = mesh.lines[a1][hi[a1]] - mesh.lines[a1][lo[a1]]
            span2 = mesh.lines[a2][hi[a2]] - mesh.lines[a2][lo[a2]]
            assert abs(meta[k:k + n, 5].astype(np.float64).sum() - span1 * span2) <= 1e-6 * span1 * span2
            k += n
    assert k == len(pts)
    theta, phi = farfield.sphere_grid()
    assert (len(theta), len(phi)) == (91, 181)


def test_circular_components(oracle_mod):
    """Two crossed current elements (x and y, equal amplitude, +-90 degrees
    apart) radiate a circularly polarised wave along z: one of the two
    components vanishes there and the other carries all of |E|; the phase
    sign picks which (the labels are the reference's).  Everywhere
    |E_cprh|^2 + |E_cplh|^2 = |E_theta|^2 + |E_phi|^2."""
    theta, phi = farfield.sphere_grid(10, 10)
    out = {}
    for q in (1.0, -1.0):
        # normal +z: J = n x H, so J_x = -H_y and J_y = H_x; J_y = j q J_x
        meta = np.array([0, 0, 0, 2, 1.0, 1e-6, 0, 0], float)
        accE = np.zeros(6)
        accH = np.zeros(6)
        accH[2] = -1.0          # H_y re -> J_x = 1
        accH[1] = q             # H_x im -> J_y = j q
        ff = oracle_mod.farfield(accE, accH, meta, theta, phi, np.zeros(3), 1e9, 1.0)
        rh, lh = farfield.circular(ff)
        f = np.asarray(ff).reshape(-1, 5)
        np.testing.assert_allclose(rh ** 2 + lh ** 2, np.sum(f[:, 1:] ** 2, axis=1), rtol=1e-12, atol=1e-40)
        out[q] = (rh[0], lh[0])  # theta = 0
    (rh_p, lh_p), (rh_m, lh_m) = out[1.0], out[-1.0]
    assert min(rh_p, lh_p) < 1e-9 * max(rh_p, lh_p)
    assert min(rh_m, lh_m) < 1e-9 * max(rh_m, lh_m)
    assert (rh_p > lh_p) != (rh_m > lh_m)


def _dipole_fields(P, p_hat, src, k):
    """Exact fields (near and far zone, exp(+j w t)) of a unit current element
    I dl = 1 along p_hat at src."""
    z0 = 376.73031346177066
    r = P - src
    R = np.linalg.norm(r, axis=1)[:, None]
    rh = r / R
    pr = (rh @ p_hat)[:, None]
    e = np.exp(-1j * k * R)
    E = (z0 / (4 * np.pi)) * e * (-(1j * k / R + 1 / R ** 2 + 1 / (1j * k * R ** 3)) * (p_hat[None, :] - pr * rh)
                                  + 2 * (1 / R ** 2 + 1 / (1j * k * R ** 3)) * pr * rh)
    H = (1 / (4 * np.pi)) * e * (1j * k / R + 1 / R ** 2) * np.cross(p_hat[None, :], rh)
    return E, H


def _box(lo, hi, n):
    pts, meta = [], []
    for ax in range(3):
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for side, sgn in ((lo[ax], -1.0), (hi[ax], 1.0)):
            u, v = np.linspace(lo[a1], hi[a1], n), np.linspace(lo[a2], hi[a2], n)
            du, dv = np.full(n, u[1] - u[0]), np.full(n, v[1] - v[0])
            du[[0, -1]] /= 2
            dv[[0, -1]] /= 2
            U, V = np.meshgrid(u, v, indexing="ij")
            p = np.zeros((n * n, 3))
            p[:, ax], p[:, a1], p[:, a2] = side, U.ravel(), V.ravel()
            m = np.zeros((n * n, 8))
            m[:, :3], m[:, 3], m[:, 4], m[:, 5] = p, ax, sgn, np.outer(du, dv).ravel()
            pts.append(p)
            meta.append(m)
    return np.concatenate(pts), np.concatenate(meta)


def _acc(F):
    return np.stack([F.real, F.imag], axis=2).ravel()


@pytest.mark.parametrize("kind", ["PEC", "PMC"])
@pytest.mark.parametrize("pdir", [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0)])
def test_mirror_is_image_theory(oracle_mod, kind, pdir):
    """A current element at height h above the plane z = 0, its fields on a
    closed box above the plane, far field with a PEC / PMC mirror: equal to
    the far field of the element and its image (PEC: horizontal currents
    reversed; PMC: vertical reversed) from a box enclosing both, to 1e-3 of
    the peak, with imageNormals set.  The default, the reference's rule
    (mirrored faces keep their normals, src/nf2ff.mjs:377-398), misses this
    by up to half the peak."""
    f = 1e9
    k = 2 * np.pi * f / 299792458.0
    h, a = 0.06, 0.03
    p_hat = np.array(pdir)
    theta, phi = np.linspace(0.01, np.pi - 0.01, 25), np.array([0.0, np.pi / 4, np.pi / 2])
    pts, meta = _box([-a, -a, h - a], [a, a, h + a], 31)
    E, H = _dipole_fields(pts, p_hat, np.array([0, 0, h]), k)
    img = np.array([-p_hat[0], -p_hat[1], p_hat[2]]) if kind == "PEC" else np.array([p_hat[0], p_hat[1], -p_hat[2]])
    pts2, meta2 = _box([-a, -a, -h - a], [a, a, h + a], 61)
    E1, H1 = _dipole_fields(pts2, p_hat, np.array([0, 0, h]), k)
    E2, H2 = _dipole_fields(pts2, img, np.array([0, 0, -h]), k)
    ref = oracle_mod.farfield(_acc(E1 + E2), _acc(H1 + H2), meta2.ravel(), theta, phi, np.zeros(3), f, 1.0)
    mir = {"type": kind, "direction": 2, "position": 0.0, "imageNormals": True}
    got = oracle_mod.farfield(_acc(E), _acc(H), meta.ravel(), theta, phi, np.zeros(3), f, 1.0, mirror=mir)
    U, Ur = got.reshape(-1, 5)[:, 0], ref.reshape(-1, 5)[:, 0]
    assert np.max(np.abs(U - Ur)) < 1e-3 * Ur.max()
    old = oracle_mod.farfield(_acc(E), _acc(H), meta.ravel(), theta, phi, np.zeros(3), f, 1.0,
                              mirror=dict(mir, imageNormals=False)).reshape(-1, 5)[:, 0]
    assert np.max(np.abs(old - Ur)) > 0.02 * Ur.max()
