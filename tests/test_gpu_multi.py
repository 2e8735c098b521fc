"""Many scatterers in one batched computation (hbem_sigma_sweep_multi) on the
many-instance workload: 64 random star-shaped polygons from one seeded
stream (6-16 vertices, radii in [0.5, 1.5], N in [100, 1000] panels split
over the edges by length), c_i uniform in [4, 25], c_o = 1, 32 wavenumbers
uniform in [2, 20] each, Gauss order 12.

* Every instance's sigma from the batched call equals its own
  hbem_sigma_sweep bit for bit (its T padded with the identity to the
  bucket's order factorises to the same bits), and so do the results of
  another bucket granularity.
* The oracle's sigma (same R, q = 2) at two k of three instances with
  N <= 300: <= 1e-10 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORDER, Q = 12, 2


@pytest.fixture(scope="module")
def workload(H):
    sp = H.ContinuousSpace1()
    items = []
    for i, (mesh, c_i, ks) in enumerate(H.random_polygon_instances(2)):
        bd = H.BuilderData(mesh, sp, sp, ORDER)
        items.append((bd, ks, 1.0, c_i, H.gaussian_test_matrix(2 * mesh.getNumPanels(), 2, 100 + i)))
    print("instances", len(items), "panels", min(it[0].getNumPanels() for it in items), "..",
          max(it[0].getNumPanels() for it in items), flush=True)
    return items


def test_batched_equals_each_instance_alone(workload, H):
    batched = H.sigma_sweep_batched(workload, Q)
    for i, ((bd, ks, c_o, c_i, R), s) in enumerate(zip(workload, batched)):
        alone = H.sigma_sweep(bd, ks, c_o, c_i, R, Q)
        bd.release_work()
        assert np.array_equal(s, alone), (i, bd.getNumPanels(), np.abs(s - alone).max())
        assert np.all(np.isfinite(s)) and np.all(s > 0)


@pytest.mark.parametrize("nc", [3, 5])
def test_batched_equals_alone_with_more_right_hand_sides(workload, H, nc):
    """As above with nc = 3 (the panelled solves' couplings beside the panel
    solves) and nc = 5 (couplings as launches of their own), on six instances:
    the three largest, whose solves run in panels either way, and three of at
    most 300 panels, which run the one-kernel solve alone and the panelled
    solve when padded to the bucket's order."""
    by_size = sorted(workload, key=lambda it: it[0].getNumPanels())
    pick = [it for it in by_size if it[0].getNumPanels() <= 300][:3] + by_size[-3:]
    sub = [(bd, ks, c_o, c_i, H.gaussian_test_matrix(2 * bd.getNumPanels(), nc, 200 + j))
           for j, (bd, ks, c_o, c_i, _) in enumerate(pick)]
    batched = H.sigma_sweep_batched(sub, Q)
    for (bd, ks, c_o, c_i, R), s in zip(sub, batched):
        alone = H.sigma_sweep(bd, ks, c_o, c_i, R, Q)
        bd.release_work()
        assert np.array_equal(s, alone), (nc, bd.getNumPanels(), np.abs(s - alone).max())


def test_bucket_granularity_does_not_change_bits(workload, H):
    sub = workload[:16]
    a = H.sigma_sweep_batched(sub, Q, bucket=64)
    b = H.sigma_sweep_batched(sub, Q, bucket=512)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_batched_vs_oracle_small_instances(workload, H, O):
    small = [it for it in workload if it[0].getNumPanels() <= 300][:3]
    assert len(small) == 3
    batched = H.sigma_sweep_batched(small, Q)
    for (bd, ks, c_o, c_i, R), s in zip(small, batched):
        ref = O.SolOp(bd.mesh, ORDER).sigma_sweep(ks[[0, 31]], c_o, c_i, R, Q)
        assert np.all(np.abs(s[[0, 31]] - ref) <= 1e-10 * ref), (s[[0, 31]], ref)


def test_graph_replays_equal_eager(workload, H, monkeypatch):
    """The device phase of a batched call is captured into a hipGraph on first
    use and replayed by later calls; replays, with eager work and host
    synchronisations between them, give the eager call's bits."""
    sub = workload[:12]
    monkeypatch.setenv("HBEM_NO_GRAPH", "1")
    eager = H.sigma_sweep_batched(sub, Q)
    monkeypatch.delenv("HBEM_NO_GRAPH")
    for rep in range(3):
        got = H.sigma_sweep_batched(sub, Q)
        bd, ks, c_o, c_i, R = sub[rep]
        H.sigma_sweep(bd, ks, c_o, c_i, R, Q)
        bd.release_work()
        for x, y in zip(got, eager):
            assert np.array_equal(x, y), rep


def test_autotuned_bucket(workload, H, tmp_path):
    """The autotuner times the candidate granularities on the instances and
    caches its choice; any choice gives the same bits."""
    from helmholtzbem_amd import autotune
    sub = workload[:8]
    cache = str(tmp_path / "tune.json")
    g = autotune.tune_bucket(sub, Q, cache=cache, repeats=1)
    assert g in autotune.CANDIDATES
    a = H.sigma_sweep_batched(sub, Q, bucket="auto", tune_cache=cache)
    b = H.sigma_sweep_batched(sub, Q, bucket=64)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_batched_k_count_not_a_multiple_of_32(H):
    """40 wavenumbers per instance: the slot blocks of 32 are padded (the
    padded slots are computed and never written); three small polygons of
    uneven size, batched == each alone bit for bit."""
    sp = H.ContinuousSpace1()
    rng = np.random.default_rng(11)
    items = []
    for i, N in enumerate((37, 52, 90)):
        mesh = H.random_polygon_mesh(20 + i, 7, N)
        bd = H.BuilderData(mesh, sp, sp, ORDER)
        ks = np.sort(rng.uniform(2.0, 12.0, 40))
        items.append((bd, ks, 1.0, 6.0 + i, H.gaussian_test_matrix(2 * N, 2, 40 + i)))
    batched = H.sigma_sweep_batched(items, Q, bucket=64)
    for (bd, ks, c_o, c_i, R), s in zip(items, batched):
        alone = H.sigma_sweep(bd, ks, c_o, c_i, R, Q)
        assert s.shape == (40,) and np.all(np.isfinite(s)) and np.all(s > 0)
        assert np.array_equal(s, alone), np.abs(s - alone).max()
