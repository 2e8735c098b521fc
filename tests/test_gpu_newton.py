"""Batched d2sigma/dk2 for the Newton refinement (hbem_sigma_sweep_der2).

* Parity: for every k the batched (sigma, sigma', sigma'') equals the per-k
  API path -- gen_sol_op_2nd_der's projected T, T', T'' and
  randomized_svd::sv_der2 (randsvd.cpp:171-216) -- at converged subspace
  iteration (the sweep takes the Wielandt vector from the subspace's
  singular pair, the per-k API the reference's pivoted-QR column; they
  coincide once sigma has converged), and the oracle's sv_der2.  Bars: sigma
  1e-10, sigma' 1e-8, sigma'' 1e-7 relative to max(1, |value|) (the bordered
  2n x 2n solve amplifies the rounding of T_p, which the two paths form
  differently).
* The resonance-search workload at full panel count: the unit square in 1000
  line panels (40 grid points on [5, 7]) and a random 10-vertex polygon (seed
  4) in 1000 panels (200 grid points on [5, 5.5]), Gauss order 11, c_i = 20,
  c_o = 1, as two independent instances.  Newton with the batched sigma''
  must find the roots Brent on dsigma/dk finds, to 1e-8 relative.  The
  polygon's dsigma/dk changes sign every 0.009 to 0.05 in k near k = 5, so
  its grid step (0.0025) keeps one root per bracket: with several, Brent and
  Newton may each converge to a different one, both correct.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(b))


@pytest.mark.parametrize("proj", [True, False])
def test_der2_sweep_matches_per_k_api(H, proj):
    mesh = H.star_mesh(40)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 8)
    R = H.gaussian_test_matrix(80, 2, 7)
    ks = np.array([1.5, 2.75, 4.0, 6.5])
    q = 200     # converged at every k (at k = 6.5, q = 30 still leaves 5e-6 between the two vectors)
    s, d1, d2 = H.sigma_sweep_der2(bd, ks, 1.0, 16.0, R, q, proj=proj)
    s0, d10 = H.sigma_sweep(bd, ks, 1.0, 16.0, R, q, der=2, proj=proj)
    assert np.array_equal(s, s0) and np.array_equal(d1, d10)
    so = H.SolutionsOperator(bd, proj)
    for i, k in enumerate(ks):
        T, Td, Tdd = so.gen_sol_op_2nd_der(k, 1.0, 16.0)
        ref = H.randomized_svd.sv_der2(T, Td, Tdd, R, q)
        assert close(s[i], ref[0], 1e-10) and close(d1[i], ref[1], 1e-8) and close(d2[i], ref[2], 1e-7), \
            (k, (s[i], d1[i], d2[i]), ref)


def test_der2_sweep_vs_oracle_converged(H, O):
    mesh = H.square_mesh(0.5, 12)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 11)
    R = H.gaussian_test_matrix(96, 2, 0)
    ks = np.array([2.0, 3.0])
    s, d1, d2 = H.sigma_sweep_der2(bd, ks, 1.0, 20.0, R, 30)
    S = O.SolOp(mesh, 11)
    for i, k in enumerate(ks):
        T, Td, Tdd = S.gen_sol_op(k, 1.0, 20.0, d=2)
        ref = O.sv_der2(T, Td, Tdd, R, 30)
        assert close(s[i], ref[0], 1e-10) and close(d1[i], ref[1], 1e-8) and close(d2[i], ref[2], 1e-6), \
            (k, (s[i], d1[i], d2[i]), ref)


def resonance_instances(H):
    sq = H.square_mesh(1.0, 250, center=(0.5, 0.5))
    poly = H.random_polygon_mesh(4, 10, 1000)
    return {"square1000": (sq, dict(k_min=5.0, k_max=7.0, n_points=40)),
            "polygon1000": (poly, dict(k_min=5.0, k_max=5.5, n_points=200))}


@pytest.mark.parametrize("name", ["square1000", "polygon1000"])
def test_newton_roots_equal_brent_roots_full_size(H, name):
    mesh, grid = resonance_instances(H)[name]
    assert mesh.getNumPanels() == 1000
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 11)
    R = H.gaussian_test_matrix(2000, 2, 4)
    from helmholtzbem_amd import resonances as RZ
    kw = dict(c_o=1.0, c_i=20.0, R=R, q=2, q_refine=20, tol=1e-11, **grid)
    brent = RZ.find_minima(bd, method="brent", **kw)
    newton = RZ.find_minima(bd, method="newton", **kw)
    print(name, "brent", [r for r, _, _ in brent], "newton", [r for r, _, _ in newton], flush=True)
    assert len(brent) >= 1 and len(newton) == len(brent)
    for (kb, sb, _), (kn, sn, _) in zip(brent, newton):
        assert abs(kb - kn) <= 1e-8 * kb, (kb, kn)
        assert abs(sb - sn) <= 1e-8 * sb
