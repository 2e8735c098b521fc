"""Many scatterers of uneven size at once (the instance batches of the
resonance drivers), on a small stand-in: 6 random polygons with N in
[24, 90], Gauss order 8, 6 wavenumbers each.  sigma_sweep_multi issues every
instance's batched sweep from its own host thread onto its own device
stream.  The results must be the same bits as one instance at a time, and
the oracle's sigma_min to the path's bar (<= 1e-10 relative with the same
test matrix R)."""
import numpy as np
import pytest


def instances(H, count=6, seed=2):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = int(rng.integers(24, 90))
        mesh = H.random_polygon_mesh(100 + i, int(rng.integers(6, 12)), N)
        sp = H.ContinuousSpace1()
        bd = H.BuilderData(mesh, sp, sp, 8)
        ks = np.sort(rng.uniform(2.0, 8.0, 6))
        c_i = float(rng.uniform(4.0, 25.0))
        R = H.gaussian_test_matrix(2 * N, 2, i)
        out.append((bd, ks, 1.0, c_i, R))
    return out


@pytest.mark.gpu
def test_concurrent_sweeps_equal_serial_and_oracle(H, O):
    items = instances(H)
    multi = H.sigma_sweep_multi(items, q=2)
    for (bd, ks, c_o, c_i, R), sig in zip(items, multi):
        serial = H.sigma_sweep(bd, ks, c_o, c_i, R, 2)
        assert np.array_equal(sig, serial)
        ref = O.SolOp(bd.mesh, 8).sigma_sweep(ks, c_o, c_i, R, 2)
        assert np.abs(sig - ref).max() <= 1e-10 * np.abs(ref).max()
    der = H.sigma_sweep_multi(items[:3], q=2, der=1)
    for (bd, ks, c_o, c_i, R), (s, ds) in zip(items[:3], der):
        s1, ds1 = H.sigma_sweep(bd, ks, c_o, c_i, R, 2, der=1)
        assert np.array_equal(s, s1) and np.array_equal(ds, ds1)
