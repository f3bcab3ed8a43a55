"""Launch-shape autotuner.

The integrator kernels are compiled for several (lanes per chain, elements
per lane) tilings; which one is fastest depends on D, on the number of chains
(how many waves fill the 256 CUs) and on the precision.  ``tune`` times a few
short persistent launches per candidate on the caller's device and keeps the
fastest in a process-wide table keyed by (kernel, D, chains, dtype); the
table is read-only after an entry is written (guarded by a lock).
"""
import threading
import time

import torch

_table = {}
_lock = threading.Lock()

LANES_PER_CHAIN = (4, 8, 16, 32, 64)


def lookup(kind, dim, chains, dtype):
    with _lock:
        return _table.get((kind, int(dim), int(chains), str(dtype)))


def tune(kind, dim, chains, dtype, run, candidates=LANES_PER_CHAIN, repeats=3):
    """Pick the fastest ``lanes_per_chain`` for ``run(lanes_per_chain)``.

    ``run`` launches a short representative workload and returns nothing;
    candidates the library rejects (size outside the compiled tiles) are
    skipped.  Returns (best_lpc, {lpc: seconds}).
    """
    key = (kind, int(dim), int(chains), str(dtype))
    with _lock:
        if key in _table:
            return _table[key], {}
    timings = {}
    for lpc in candidates:
        try:
            run(lpc)  # compile/warm
        except Exception:
            continue
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run(lpc)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        timings[lpc] = best
    if not timings:
        raise RuntimeError(f"no launch shape of {kind} supports dim={dim}")
    best_lpc = min(timings, key=timings.get)
    with _lock:
        _table.setdefault(key, best_lpc)
    return best_lpc, timings
