"""The threaded slab builder gives the whole-grid builder's numbers, and the
large mixed-precision workload keeps its port array clear of the boxes."""
import numpy as np

from efs_amd import operator as op
from efs_amd import workloads


def test_threaded_builder_same_numbers():
    rng = np.random.default_rng(1)
    mesh = op.Mesh(np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 40)]) * 1e-3, np.arange(23) * 1e-3,
                   np.arange(19) * 0.7e-3)
    mat = op.Materials(mesh)
    mat.eps[...] = rng.uniform(1, 10, mat.eps.shape)
    mat.kappa[...] = rng.uniform(0, 0.05, mat.eps.shape)
    mat.mu[...] = rng.uniform(1, 3, mat.eps.shape)
    mat.pec |= rng.random(mat.pec.shape) < 0.05
    mat.port_g[...] = rng.random(mat.port_g.shape) * (rng.random(mat.port_g.shape) < 0.02)
    dT = mesh.timestep()
    a = op.build_coefficients(mesh, mat, dT)
    for threads, chunk in ((4, 3), (3, 2), (8, None)):
        b = op.build_coefficients_threaded(mesh, mat, dT, threads=threads, chunk=chunk)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])


def test_large_array_structure():
    w = workloads.large_array(dims=(96, 96, 72), port_pitch=8, n_boxes=20, nfreq=4, n_ts=100, threads=4)
    assert len(w["probes"]) == 32 and len(w["excitation"]["pos"]) == 16  # 16 one-edge ports
    co = w["coefficients"]
    N = 96 * 96 * 72
    vi = co["vi"].reshape(3, 96, 96, 72)
    # every port edge is a live, lumped-loaded edge (vv < 1 from the 50-ohm
    # conductance) between two PEC dipole arms
    for p in w["excitation"]["pos"]:
        x, y, z = p // (96 * 72), (p // 72) % 96, p % 72
        assert vi[2, x, y, z] > 0 and co["vv"][2 * N + p] < 1.0
        assert vi[2, x, y, z + 1] == 0 and vi[2, x, y, z - 1] == 0
    assert len(w["nf2ff"]["omega"]) == 4 and w["n_pml_cells"] > 0
