"""Autotuner of batched plans.

A batched plan's schedule depends on planning knobs that the library reads
when the plan is created (lfx_set_tuning_ex in the C ABI):

    levels_per_pass  levels fused per tiled pass (register tiles take 1..4)
    line_limit       longest band that register-tiled plans hand to the
                     whole-line (LDS) kernels; shorter bands finish there
    side_streams     side streams that run a stage's LDS launches next to
                     its register launch
    tile             tile length of LDS-tiled passes (float64 and other plans
                     without register tiles)
    across           ND calls: strided float32 axes of at most 512 samples
                     read in place by the across-line kernel (0 off, 1 on,
                     -1 default: on)
    strip            register window (strip) of float32 plans in samples:
                     1024 or 2048 (0: default, 2048)

tune_batch() builds the plans of a signal list for every candidate setting,
times them on the device with the caller's data, and returns the fastest.
The reference has no counterpart (its only knob is the thread count).
"""
import contextlib
import ctypes
import itertools
import time

from . import _lib

FIELDS = ("tile", "levels_per_pass", "side_streams", "line_limit", "across", "strip")


class LfxTuning(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int64), ("levels_per_pass", ctypes.c_int32), ("side_streams", ctypes.c_int32),
                ("line_limit", ctypes.c_int64), ("across", ctypes.c_int32), ("strip", ctypes.c_int32)]


def _bind():
    L = _lib.lib()
    if not hasattr(L, "_tune_bound"):
        L.lfx_set_tuning_ex.argtypes = [ctypes.POINTER(LfxTuning)]
        L.lfx_set_tuning_ex.restype = ctypes.c_int
        L.lfx_get_tuning.argtypes = [ctypes.POINTER(LfxTuning)]
        L.lfx_get_tuning.restype = ctypes.c_int
        L._tune_bound = True
    return L


def get_tuning():
    """The calling thread's planning knobs (0 / -1 = library default)."""
    t = LfxTuning()
    _lib.check(_bind().lfx_get_tuning(ctypes.byref(t)))
    return {f: getattr(t, f) for f in FIELDS}


def set_tuning(tile=0, levels_per_pass=0, side_streams=-1, line_limit=0, across=-1, strip=0):
    """Planning knobs for plans created afterwards on this thread."""
    t = LfxTuning(int(tile), int(levels_per_pass), int(side_streams), int(line_limit), int(across), int(strip))
    _lib.check(_bind().lfx_set_tuning_ex(ctypes.byref(t)))


@contextlib.contextmanager
def tuning(**knobs):
    """Temporarily set planning knobs (unnamed ones keep their defaults)."""
    old = get_tuning()
    cfg = {"tile": 0, "levels_per_pass": 0, "side_streams": -1, "line_limit": 0, "across": -1, "strip": 0}
    cfg.update(knobs)
    set_tuning(**cfg)
    try:
        yield
    finally:
        set_tuning(**old)


def default_space(dtype, chain=False):
    """Candidate knob values searched by tune_batch for a storage type.
    chain=True: the plans run as a BatchChain, and "chunks" (signal chunks of
    the pipeline; 1 = no pipelining) is searched too."""
    import torch
    if dtype == torch.float64:
        space = {"tile": [0, 2048, 8192], "levels_per_pass": [0, 4], "side_streams": [-1, 0]}
        if chain:
            space = {"chunks": [2, 1], "tile": [0, 2048, 8192], "levels_per_pass": [0, 4]}
        return space
    if chain:
        return {"chunks": [2, 1, 3], "levels_per_pass": [0, 3], "line_limit": [0, 32], "strip": [0, 1024]}
    return {"levels_per_pass": [0, 3], "line_limit": [0, 32, 1024], "side_streams": [-1, 0], "strip": [0, 1024]}


def _time_plans(plans, x, bufs, reps, warmup):
    import torch
    stream = torch.cuda.current_stream(x.device)

    def run():
        if len(plans) == 1 and hasattr(plans[0], "transforms"):  # a BatchChain
            plans[0].execute(x, bufs)
            return
        src = x
        for p, b in zip(plans, bufs):
            p.execute(src, b)
            src = b

    for _ in range(warmup):
        run()
    torch.cuda.synchronize(x.device)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        run()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def build(signals, transforms, knobs, mode=0, dtype=None):
    """The executables a candidate stands for: a list of BatchPlans (one per
    transform) or, when the knobs name "chunks", one BatchChain."""
    from .batch import BatchChain, BatchPlan
    knobs = dict(knobs)
    chunks = knobs.pop("chunks", None)
    with tuning(**knobs):
        if chunks is not None:
            return [BatchChain(signals, transforms, mode=mode, dtype=dtype, chunks=chunks)]
        return [BatchPlan(signals, tf, mode=mode, dtype=dtype) for tf in transforms]


def tune_batch(signals, transforms, x, mode=0, space=None, reps=5, warmup=2, verbose=False, min_gain=0.02,
               rounds=2, chain=False):
    """Search planning knobs for batched plans of `signals`.

    transforms: transform types run in sequence (x -> out_0 -> out_1 ...),
    e.g. (Forward, ForwardAdjoint). x: the packed input (device tensor, the
    plans' storage type). Returns (best knobs dict, [(knobs, ms), ...]) with
    the candidates in the order timed; the library defaults are always the
    first candidate, so the result is never slower than them on this data.
    The candidates are timed in `rounds` round-robin rounds (plans rebuilt
    each round) and each keeps its fastest round, so that the first timings
    of a run (clocks, caches still settling) do not count against the
    candidate timed first. Another candidate is chosen only when it beats
    the defaults by at least min_gain (a fraction): closer timings are
    within run-to-run noise, and keeping the defaults then keeps the chosen
    plans the same from run to run."""
    import torch
    space = dict(space if space is not None else default_space(x.dtype, chain))
    keys = list(space)
    cands = [dict(zip(keys, values)) for values in itertools.product(*(space[k] for k in keys))]
    best_ms = [float("inf")] * len(cands)
    bufs = [torch.empty_like(x) for _ in transforms]
    for rnd in range(max(1, int(rounds))):
        for i, knobs in enumerate(cands):
            t0 = time.perf_counter()
            plans = build(signals, transforms, knobs, mode=mode, dtype=x.dtype)
            try:
                ms = _time_plans(plans, x, bufs, reps, warmup)
            finally:
                for p in plans:
                    p.close()
            best_ms[i] = min(best_ms[i], ms)
            if verbose:
                print(f"[tune] round {rnd}: {knobs}: {ms:.3f} ms (plans {time.perf_counter() - t0:.2f} s)", flush=True)
    results = list(zip(cands, best_ms))
    fastest = min(results, key=lambda r: r[1])
    best = fastest[0] if fastest[1] < results[0][1] * (1.0 - min_gain) else results[0][0]
    return best, results
