"""Mixed-precision dense stage (hbem_sigma_sweep_mixed): FP32 LU of T on the
matrix cores, every solve of the randomized SVD refined against the FP64 T.

Bar: sigma within 1e-8 relative of the FP64 sweep with the same R and q (the
path's stated mixed-precision bar); k whose refinement does not converge are
recomputed in FP64 and must then equal the FP64 sweep bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-8


@pytest.mark.parametrize("proj", [True, False])
def test_mixed_equals_fp64_small(H, proj):
    mesh = H.star_mesh(40)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 8)
    R = H.gaussian_test_matrix(80, 2, 3)
    ks = np.linspace(1.0, 9.0, 70)
    ref = H.sigma_sweep(bd, ks, 1.0, 16.0, R, 2, proj=proj)
    sig, used = H.sigma_sweep_mixed(bd, ks, 1.0, 16.0, R, 2, refine=3, proj=proj)
    err = np.abs(sig - ref) / ref
    print("proj", proj, "max rel err", err.max(), "fp64 fallbacks", int(used.sum()), flush=True)
    assert err.max() <= MIXED_TOL
    assert not used.all()
    assert np.array_equal(sig[used], ref[used])


def test_mixed_without_refinement_falls_back_to_fp64(H):
    """refine = 0 certifies nothing: every k is recomputed with the FP64
    factorisation and equals the FP64 sweep bit for bit."""
    mesh = H.random_polygon_mesh(2, 9, 60)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 8)
    R = H.gaussian_test_matrix(120, 2, 1)
    ks = np.linspace(2.0, 12.0, 40)
    ref = H.sigma_sweep(bd, ks, 1.0, 9.0, R, 2)
    sig, used = H.sigma_sweep_mixed(bd, ks, 1.0, 9.0, R, 2, refine=0)
    assert used.all() and np.array_equal(sig, ref)


def test_mixed_polygon2000_grid(H):
    """The headline workload: 2000-panel polygon (seed 3), order 11, c_i = 20,
    128 k in [5, 40]; mixed sigma per k within 1e-8 of FP64."""
    mesh = H.random_polygon_mesh(3, 12, 2000)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 11)
    R = H.gaussian_test_matrix(4000, 2, 3)
    ks = np.linspace(5.0, 40.0, 128)
    ref = H.sigma_sweep(bd, ks, 1.0, 20.0, R, 2)
    sig, used = H.sigma_sweep_mixed(bd, ks, 1.0, 20.0, R, 2, refine=3)
    err = np.abs(sig - ref) / ref
    print("polygon2000 max rel err %.3e, fp64 fallbacks %d" % (err.max(), int(used.sum())), flush=True)
    assert err.max() <= MIXED_TOL
    assert used.sum() < ks.size // 4


@pytest.mark.parametrize("nc", [3, 5])
def test_mixed_panelled_solves_more_right_hand_sides(H, nc):
    """A 700-panel polygon (T of order 1400: the solves run in panels), order
    11, c_i = 20, 32 k in [5, 20], with nc = 3 (each panel's coupling beside
    the next panel's solve) and nc = 5 (each coupling as a launch of its own):
    mixed sigma per k within 1e-8 of FP64 with the same test matrix."""
    mesh = H.random_polygon_mesh(7, 10, 700)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 11)
    R = H.gaussian_test_matrix(1400, nc, 9)
    ks = np.linspace(5.0, 20.0, 32)
    ref = H.sigma_sweep(bd, ks, 1.0, 20.0, R, 2)
    sig, used = H.sigma_sweep_mixed(bd, ks, 1.0, 20.0, R, 2, refine=3)
    err = np.abs(sig - ref) / ref
    print("nc %d max rel err %.3e, fp64 fallbacks %d" % (nc, err.max(), int(used.sum())), flush=True)
    assert err.max() <= MIXED_TOL
