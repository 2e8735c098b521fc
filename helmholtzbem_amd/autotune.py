"""Autotuning of the many-scatterer sweep (sigma_sweep_batched).

The tunable is the bucket granularity g: instances whose orders 2N round up
to the same multiple of g share one assembly launch per pair class and one
batched LU / randomized SVD, their T padded with the identity to that
multiple.  Small g wastes less work on padding; large g gives fewer, larger
batches (more matrices per launch, fewer launches).  The padding does not
change any result bit, so the choice is purely one of speed.

tune_bucket() times each candidate on the caller's own instances (after one
untimed call that also captures the graphs) and keeps the fastest; the choice
is cached in a JSON file under a key made of what determines the work:
quadrature order, k-grid length, test columns, q, projection and the sorted
panel counts.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from typing import Dict, Optional, Sequence

CANDIDATES = (64, 128, 256, 512)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "helmholtzbem_amd", "autotune.json")


def signature(items, q: int, proj: bool) -> str:
    """Key of the tuning cache: the sizes that decide the work of a call."""
    bd0, ks0, _, _, R0 = items[0]
    panels = sorted(int(it[0].getNumPanels()) for it in items)
    desc = {"order": int(bd0.getGaussQROrder()), "nk": len(ks0), "nc": int(R0.shape[1]), "q": int(q),
            "proj": bool(proj), "panels": panels}
    return hashlib.sha1(json.dumps(desc, sort_keys=True).encode()).hexdigest()


def _load(path: str) -> Dict[str, dict]:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, data: Dict[str, dict]) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_bucket(items, q: int = 2, proj: bool = True, candidates: Sequence[int] = CANDIDATES,
                cache: Optional[str] = DEFAULT_CACHE, repeats: int = 2, force: bool = False) -> int:
    """The fastest bucket granularity for sigma_sweep_batched on these
    instances (cached per signature in `cache`; None disables the cache)."""
    from . import sigma_sweep_batched
    items = list(items)
    key = signature(items, q, proj)
    data = _load(cache) if cache else {}
    if not force and key in data:
        return int(data[key]["bucket"])
    times = {}
    for g in candidates:
        sigma_sweep_batched(items, q, proj, bucket=g)          # first call: capture, allocation
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            sigma_sweep_batched(items, q, proj, bucket=g)
            best = min(best, time.perf_counter() - t0)
        times[g] = best
    choice = min(times, key=times.get)
    if cache:
        data = _load(cache)
        data[key] = {"bucket": int(choice), "seconds": {str(g): t for g, t in times.items()},
                     "instances": len(items)}
        _store(cache, data)
    return int(choice)
