"""Benchmark of the sigma_min(k) path on one MI355X.

Workload (the configuration the metric is quoted on): smooth star curve
r(t) = 1 + 0.25 cos 5t split into 400 equal-parameter panels (800 x 800 complex
transmission matrix), c_i = 16, c_o = 1, Gauss order 16, 256 real k =
linspace(1, 10); per k the operators are assembled with their first and second
k-derivatives (FP64), T, T', T'' are formed, sigma_min is estimated by the
randomized SVD (2 Gaussian columns, 2 subspace iterations) and dsigma/dk is
evaluated from the singular pair.

One step = one pass over the 256-point k grid.  value = sigma_min(k)
evaluations per second over the timed steps (inputs -- geometry, quadrature,
projection factor, test matrix -- are resident on the device before timing).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
# HBM bytes per launch of the general-pair assembly kernel at this workload:
# rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (KiB per dispatch, scaled by
# the counters' reading of a plain 1 GiB copy), from
# profiles/k_assemble_x2_general_star400_hbm_traffic.txt
GENERAL_KERNEL_TRAFFIC_BYTES = 3.2686e10
N_PANELS, ORDER, C_I, C_O, NK, K_MIN, K_MAX, NC, Q = 400, 16, 16.0, 1.0, 256, 1.0, 10.0, 2, 2


def cpu_baseline(mesh, R, threads):
    """The oracle (CPU FP64 restatement) on a bounded sample: one k of the
    workload -- gen_sol_op with d = 2 (both refraction indices, T, T', T''
    projected as the reference does) plus sv -- on `threads` host threads."""
    from oracle import oracle as O
    S = O.SolOp(mesh, ORDER)
    k = 0.5 * (K_MIN + K_MAX)
    t0 = time.perf_counter()
    Ts = S.gen_sol_op(k, C_O, C_I, d=2, nthreads=threads)
    O.sv(Ts[0], R, Q)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "sigma_min(k) evals/s", "cores": threads, "kind": "port",
            "sample": "1 k of the workload (k=5.5): gen_sol_op d=2 (c_i and c_o assembly, T/T'/T'' projected) "
                      "+ sv; dsigma/dk not included; %.1f s" % dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 times), the reference "
                         "dispatch for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "dispatches": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the sigma_min(k) sweep would shard over wavenumbers)", file=sys.stderr)
        sys.exit(2)

    import helmholtzbem_amd as H

    mesh = H.star_mesh(N_PANELS)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, ORDER)
    n = 2 * N_PANELS
    ks = np.linspace(K_MIN, K_MAX, NK)
    R = H.gaussian_test_matrix(n, NC, 1)

    for _ in range(args.warmup):
        H.sigma_sweep(bd, ks, C_O, C_I, R, Q, der=2)
    H.profile_enable(bd, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sig, dsig = H.sigma_sweep(bd, ks, C_O, C_I, R, Q, der=2)   # synchronous: returns after its stream drained
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_gen, n_gen, fl_gen = H.profile_query(bd, 0)
    ms_adj, n_adj, fl_adj = H.profile_query(bd, 1)
    ms_coi, n_coi, fl_coi = H.profile_query(bd, 2)
    ms_lu, n_lu, fl_lu = H.profile_query(bd, 3)
    H.profile_enable(bd, False)
    if not (np.all(np.isfinite(sig)) and np.all(np.isfinite(dsig)) and np.all(sig > 0)):
        raise RuntimeError("non-finite or non-positive sigma in the benchmark sweep")

    evals = NK * args.steps
    value = evals / elapsed
    asm_s = (ms_gen + ms_adj + ms_coi) * 1e-3
    pairs_per_k = 2 * N_PANELS * (N_PANELS + 1) // 2
    avg_launch_s = ms_gen * 1e-3 / max(n_gen, 1)
    achieved = (fl_gen / max(n_gen, 1)) / avg_launch_s / 1e12 if n_gen else 0.0
    out = {
        "metric": "sigma_min(k) evals/s",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "star r(t)=1+0.25cos5t, 400 panels (800x800 complex T), c_i=16, c_o=1, "
                               "Gauss order 16, 256 k in linspace(1,10), d=2 operators, sigma_min + dsigma/dk per k, "
                               "rSVD nc=2 q=2, FP64",
                   "panels": N_PANELS, "order": ORDER, "nk": NK, "global_batch": NK, "parallelism": "single-gpu"},
        "panel_pair_evals_per_s": pairs_per_k * evals / asm_s if asm_s > 0 else None,
        "phase_seconds_per_step": {"assembly_general": ms_gen * 1e-3 / args.steps,
                                   "assembly_adjacent": ms_adj * 1e-3 / args.steps,
                                   "assembly_coinciding": ms_coi * 1e-3 / args.steps,
                                   "lu": ms_lu * 1e-3 / args.steps},
        "roofline": {"kernel": "k_assemble_x2<general, d=2> (general panel pairs, 256 k x 2 c per launch)",
                     "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": GENERAL_KERNEL_TRAFFIC_BYTES,
                     "avg_launch_ms": avg_launch_s * 1e3, "launches": n_gen,
                     "flops_per_launch": fl_gen / max(n_gen, 1)},
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(mesh, R, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
