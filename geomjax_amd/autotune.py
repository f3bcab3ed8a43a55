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


# ---------------------------------------------------------------------------
# Tuned logistic kernels: workgroup size per chain (1 or 2 waves) for the
# posteriors of up to 32 and of 33..64 features, chosen per problem.
# ---------------------------------------------------------------------------

enabled = True  # False: every call takes the library default without timing
_logistic = {}


def _problem_key(model, chains):
    from .models import LogisticBatch

    if isinstance(model, LogisticBatch):
        shape = tuple((m.N, m.D, n) for m, n in zip(model.models, model.counts))
    else:
        shape = ((model.N, model.D, int(chains)),)
    return shape, int(chains)


def logistic_waves_field(model, chains, step_size=0.05, num_integration_steps=4, max_iters=6):
    """The C ABI waves_per_chain field for the tuned path on ``model`` with
    ``chains`` chains: the cached choice, tuned on first use (timed with the
    first call's step size, number of steps and fixed-point cap) unless tuning
    is disabled or the current stream is being captured (then the library
    default, 0)."""
    key = _problem_key(model, chains)
    with _lock:
        if key in _logistic:
            return _logistic[key]
    if not enabled or torch.cuda.is_current_stream_capturing():
        return 0
    return tune_logistic(model, chains, num_integration_steps=max(1, int(num_integration_steps)),
                         max_iters=int(max_iters), step_size=float(step_size))[0]


LOGISTIC_WAVES = (0x12, 0x11, 0x22, 0x21)  # bits 0-3: D <= 32, bits 4-7: 33..64 features
LOGISTIC_DEFAULT = 0x12  # the library's default shape (kDefaultWavesPerChain)
LOGISTIC_MARGIN = 0.98   # another shape replaces the default only when at least 2 % faster


def tune_logistic(model, chains, repeats=3, num_integration_steps=4, max_iters=6, step_size=0.05, transitions=2):
    """Time a persistent launch of ``transitions`` transitions of the tuned
    RMHMC kernel per waves_per_chain candidate on the caller's device
    (positions drawn near the origin, the device PRNG; nothing of the caller's
    state is touched), the best of ``repeats`` after one untimed launch, and
    cache the fastest for (problem, chains); the library default stays unless
    another shape is faster by more than the timing noise (LOGISTIC_MARGIN),
    so the choice does not flip between runs on near-ties.  Candidates that
    differ only in a width the problem does not have are not timed twice.
    Returns (field, {field: seconds})."""
    import functools

    from .random import PRNGKey
    from .rmhmc import implicit_midpoint, init, sample

    key = _problem_key(model, chains)
    with _lock:
        if key in _logistic:
            return _logistic[key], {}
    widths = {1 if d <= 32 else (2 if d <= 64 else 4) for (_, d, _) in key[0]}
    cands, seen = [], set()
    for f in LOGISTIC_WAVES:
        sig = tuple((f >> (4 * (nb - 1))) & 0xF for nb in (1, 2) if nb in widths)
        if sig and sig not in seen:
            seen.add(sig)
            cands.append(f)
    if not cands:  # only wide posteriors: the four-wave kernel has no choice
        with _lock:
            _logistic.setdefault(key, 0)
        return 0, {}
    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device=dev)
    g.manual_seed(12345)
    q = 0.1 * torch.randn(int(chains), model.dim, generator=g, device=dev, dtype=torch.float32)
    if hasattr(model, "counts"):  # zero the coordinates past each posterior's D
        mask = torch.zeros_like(q)
        c0 = 0
        for m, n in zip(model.models, model.counts):
            mask[c0:c0 + n, :m.D] = 1.0
            c0 += n
        q = q * mask
    state = init(q, model)
    integ = functools.partial(implicit_midpoint, max_iters=max_iters)

    def run(field):
        sample(PRNGKey(7, 0), state, model, step_size, _metric(), num_integration_steps, int(transitions),
               integrator=integ, store_samples=False, waves_per_chain=(field & 0xF, field >> 4))

    timings = {}
    for f in cands:
        run(f)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run(f)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        timings[f] = best
    best_f = min(timings, key=timings.get)
    if LOGISTIC_DEFAULT in timings and timings[best_f] > LOGISTIC_MARGIN * timings[LOGISTIC_DEFAULT]:
        best_f = LOGISTIC_DEFAULT
    with _lock:
        _logistic.setdefault(key, best_f)
    return best_f, timings


def _metric():
    from .metrics import FisherMetric

    return FisherMetric()
