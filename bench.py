"""Benchmark: batched forward + forward-adjoint multi-level DWT on one MI355X.

Workload (the headline configuration): 4096 1-D float32 signals, lengths
round(2^U(10, 20)), each with a wavelet drawn uniformly from the 33 non-Lazy
predefined wavelets and a level count drawn uniformly from [3, max_level];
values iid N(0, 1); Zero boundary; total capped at 2^30 samples. One step =
forward transform of every signal, then the forward-adjoint of the result:
two batched plans run as one chain pipelined over two signal chunks (each
plan's big stage, the finest levels, runs chunk after chunk while the other
chunk's coarse-level stages run beside it). Inputs are resident in HBM before
the timed region.

Reports Gsamples/s (2 x samples per step / time); the HBM roofline of the
dominant kernel (its big-stage launches), timed with HIP events recorded around
each launch on the stream it runs on, over eager steps of the same chain; its
HBM traffic from two
rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters
read on a device copy of known size; the plain per-level kernels (one launch
per level) for comparison; and the CPU restatement of the reference timed on
this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-plain] [--no-traffic]

Secondary workloads (one JSON line each; not the headline):
    --workload c1   one float64 signal of 65536 samples, Daubechies2, 8 levels, forward
                    then inverse (latency-bound), with the CPU restatement beside it
    --workload c2   4096^2 float32 sum of Gaussians, CDF 9/7, 5 levels, F / I / FA,
                    separable and level-by-level, fused and plain
    --workload c3   512^3 float32, CDF 9/7 levels (4, 4, 4): forward, soft-threshold
                    loss and backward (the forward-adjoint) through autograd
    --workload c5   16 x 8192^2 1/f^2 fields stored bf16 / fp16 / fp32, CDF 9/7, 6
                    levels, forward + inverse: errors against an fp64 run and GB/s
"""
import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "fwd+adjoint DWT Gsamples/s and % HBM roofline, fused/batched vs plain, MI355X"
HBM_PEAK = 8.0e12
SEED = 20261201 + 3  # fourth workload of the benchmark set
NSIG = 4096
CAP = 1 << 30
KIND_NAMES = {  # (plan transform, launch kind) -> kernel symbol prefix ({pl}: samples per lane)
    (0, "register"): "void lfx::k_reg<0, {pl}>(",
    (2, "register"): "void lfx::k_reg<2, {pl}>(",
    (0, "lds"): "void lfx::k_batch<float, float, true>(",
    (2, "lds"): "void lfx::k_batch<float, float, false>(",
}


def make_workload(seed=SEED, nsig=NSIG):
    import lifted_amd as la
    g = np.random.default_rng(seed)
    wavelets = [int(w) for w in la.Wavelet if w != la.Wavelet.Lazy]
    sigs, off = [], 0
    for _ in range(nsig):
        n = int(round(2.0 ** g.uniform(10.0, 20.0)))
        w = wavelets[int(g.integers(0, len(wavelets)))]
        ml = la.max_level(w, n)
        L = int(g.integers(3, max(3, ml) + 1))
        if off + n > CAP:
            break
        sigs.append((off, n, w, L))
        off += n
    return sigs, off


def time_plans(plans, x, bufs, steps, warmup, per_launch=False):
    """Run the plans in sequence (x -> bufs[0] -> bufs[1] ...). Returns
    (ms per step, per-stage average ms [(plan, stage, ms)], per-launch records).
    Stage times come from events on the caller's stream; with per_launch, each
    kernel launch is also bracketed by events on its own stream."""
    import torch
    stream = torch.cuda.current_stream()
    infos = [[p.launch_info(s) for s in range(p.num_stages)] for p in plans]

    def new_ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)  # creates the underlying event
        return e

    def one_step(stage_ev=None, launch_ev=None):
        src = x
        for pi, p in enumerate(plans):
            for s in range(p.num_stages):
                if launch_ev is not None:
                    p.execute_stage_timed(s, src, bufs[pi], launch_ev[pi][s])
                else:
                    p.execute_stage(s, src, bufs[pi])
                if stage_ev is not None:
                    stage_ev[pi][s].record(stream)
            src = bufs[pi]

    for _ in range(warmup):
        one_step()
    torch.cuda.synchronize()
    sev = [[[torch.cuda.Event(enable_timing=True) for _ in range(p.num_stages)] for p in plans] for _ in range(steps)]
    lev = None
    if per_launch:
        lev = [[[[(new_ev(), new_ev()) for _ in infos[pi][s]] for s in range(p.num_stages)]
                for pi, p in enumerate(plans)] for _ in range(steps)]
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for k in range(steps):
        starts[k].record(stream)
        one_step(sev[k], lev[k] if lev else None)
    t1.record(stream)
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    stage_ms = []
    for pi, p in enumerate(plans):
        for s in range(p.num_stages):
            acc = 0.0
            for k in range(steps):
                prev = starts[k] if (pi == 0 and s == 0) else (
                    sev[k][pi][s - 1] if s > 0 else sev[k][pi - 1][plans[pi - 1].num_stages - 1])
                acc += prev.elapsed_time(sev[k][pi][s])
            stage_ms.append((pi, s, acc / steps))
    launches = []
    if per_launch:
        for pi, p in enumerate(plans):
            for s in range(p.num_stages):
                for li, info in enumerate(infos[pi][s]):
                    ms = sum(lev[k][pi][s][li][0].elapsed_time(lev[k][pi][s][li][1]) for k in range(steps)) / steps
                    launches.append(dict(info, plan=pi, stage=s, ms=ms))
    return total_ms / steps, stage_ms, launches


def time_chain(chain, x, bufs, steps, warmup):
    """Eager chain executions with HIP events around every launch (on the
    stream each launch goes to). Returns (ms per step, per-launch records
    averaged over the steps)."""
    import torch
    stream = torch.cuda.current_stream()
    infos = chain.launch_info()

    def new_ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)  # creates the underlying event
        return e

    for _ in range(warmup):
        chain.execute(x, bufs)
    torch.cuda.synchronize()
    lev = [[(new_ev(), new_ev()) for _ in infos] for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for k in range(steps):
        chain.execute_timed(x, bufs, lev[k])
    t1.record(stream)
    torch.cuda.synchronize()
    launches = []
    for i, info in enumerate(infos):
        ms = sum(lev[k][i][0].elapsed_time(lev[k][i][1]) for k in range(steps)) / steps
        launches.append(dict(info, ms=ms))
    return t0.elapsed_time(t1) / steps, launches


def chain_dominant(chain, launches, strip=0):
    """The kernel (transform, launch kind) with the most time per step, over
    its big-stage launches (the finest levels of every signal: most of the
    kernel's bytes; its coarse-level launches run beside the other chunk's big
    stage and are listed separately in `launches`). strip: the plans'
    register window (0: the default, 2048 samples)."""
    by = {}
    for r in launches:
        key = (chain.transforms[r["transform_index"]], r["kind"])
        d = by.setdefault(key, {"ms": 0.0})
        d["ms"] += r["ms"]
    key = max(by, key=lambda k: by[k]["ms"])
    big = [r for r in launches if (chain.transforms[r["transform_index"]], r["kind"]) == key and r["big"]]
    if not big:
        big = [r for r in launches if (chain.transforms[r["transform_index"]], r["kind"]) == key]
    n = len(big)
    pl = 16 if strip == 1024 else 32
    return {"symbol": KIND_NAMES.get(key, str(key)).format(pl=pl), "launches_per_step": n,
            "alg_bytes_per_launch": sum(r["alg_bytes"] for r in big) / n,
            "avg_launch_ms": sum(r["ms"] for r in big) / n, "items": [r["items"] for r in big],
            "ms_per_step_all_launches": by[key]["ms"]}


def time_graph(run, steps, warmup):
    """ms per step of `run` (enqueues one step on the current stream)
    captured once in a hipGraph and replayed."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            run()
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def measure_copies(x, reps=5):
    """GB/s (read + write) of copies of x: torch copy_, and lfx_stream_copy
    patterns 0 (float4 per thread) and 1 (2048-float window per wavefront)."""
    import ctypes
    import torch
    from lifted_amd import _lib
    L = _lib.lib()
    L.lfx_stream_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
    L.lfx_stream_copy.restype = ctypes.c_int
    a = torch.empty_like(x)
    st = torch.cuda.current_stream()
    nbytes = x.numel() * x.element_size()
    runs = {"torch_copy": (lambda: a.copy_(x), nbytes),
            "float4_per_thread": (lambda: _lib.check(L.lfx_stream_copy(x.data_ptr(), a.data_ptr(), nbytes, 0,
                                                                       st.cuda_stream)), nbytes // 16 * 16),
            "window_per_wave": (lambda: _lib.check(L.lfx_stream_copy(x.data_ptr(), a.data_ptr(), nbytes, 1,
                                                                     st.cuda_stream)), nbytes // 8192 * 8192)}
    out = {}
    for name, (fn, moved) in runs.items():
        fn()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(st)
        for _ in range(reps):
            fn()
        c1.record(st)
        torch.cuda.synchronize()
        out[name] = round(2 * moved * reps / (c0.elapsed_time(c1) * 1e-3) / 1e9, 1)
    del a
    return out


def _pmc_pass(counter, child_args, tmp, grids=None):
    """One rocprofv3 counter pass over a child run of this script; returns
    {kernel symbol: (dispatches, summed counter value)}, counting only
    dispatches whose grid size (work-items) is in `grids` when given."""
    out = os.path.join(tmp, counter + "_" + child_args[-1].lstrip("-"))
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, "-u", os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=150)
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 {counter} exited {r.returncode}: {r.stderr[-400:]}")
    path = None
    for root, _, files in os.walk(out):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if path is None:
        raise RuntimeError(f"rocprofv3 {counter}: no counter file")
    agg = {}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            if grids is not None and int(float(row.get("Grid_Size", 0))) not in grids:
                continue
            k = row["Kernel_Name"]
            n, v = agg.get(k, (0, 0.0))
            agg[k] = (n + 1, v + float(row["Counter_Value"]))
    return agg


def measure_traffic(total, symbol, tuning, grids=None):
    """HBM bytes per launch of `symbol` for one step of the workload (the
    chain built with the planning knobs `tuning`, as in the timed steps; only
    the dispatches whose grid size is in `grids` -- the big-stage launches --
    when given), from FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled by the ratio of known bytes to the counter on a device copy of the
    input size."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    tmp = tempfile.mkdtemp(dir="/tmp", prefix="lfx_pmc_")
    try:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            copy = _pmc_pass(counter, ["--copy-calib"], tmp)
            ck = [k for k in copy if "copyBuffer" in k or "copy" in k.lower()]
            n, v = copy[ck[0]]
            factor = (total * 4 / 1024.0) / (v / n)
            work = _pmc_pass(counter, ["--pmc-child", "--tuning-json", json.dumps(tuning)], tmp, grids)
            wk = [k for k in work if k.startswith(symbol)]
            if not wk:
                raise RuntimeError(f"{symbol} not in the {counter} pass")
            n, v = work[wk[0]]
            res[counter] = (v / n * 1024.0 * factor, factor)
        traffic = res["FETCH_SIZE"][0] + res["WRITE_SIZE"][0]
        return traffic, {"fetch_bytes": round(res["FETCH_SIZE"][0]), "write_bytes": round(res["WRITE_SIZE"][0]),
                         "fetch_factor": round(res["FETCH_SIZE"][1], 4), "write_factor": round(res["WRITE_SIZE"][1], 4)}
    except Exception as e:  # the counters are a report, not the measurement
        return None, f"traffic not measured: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_model():
    """CPU model name of this host (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_cores():
    """Cores this process may run on: its CPU affinity, capped at the 16-core
    host share of one GPU (the GPU hosts give each GPU's jobs 16 cores,
    whatever nproc reports)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16)), os.cpu_count() or 1


def _median_timed(fn, reps=5, warmup=1):
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return float(np.median(times))


def cpu_baseline(sigs, total, x_host, reps=5):
    """The CPU restatement of the reference timed on this host: 1 warm-up and
    the median of `reps` runs, on all the cores this process may use (a
    dynamic parallel-for over signals: the reference runs one line on one
    thread) and on one core (a subset of the signals)."""
    import oracle as O
    threads, nproc = cpu_cores()
    y = np.empty_like(x_host)
    z = np.empty_like(x_host)

    def run(sub, th):
        O.batch(x_host, sub, O.FORWARD, O.ZERO, out=y, n_threads=th)
        O.batch(y, sub, O.FORWARD_ADJOINT, O.ZERO, out=z, n_threads=th)

    dt = _median_timed(lambda: run(sigs, threads), reps)
    # one core on the first sixteenth of the signals (about 1 s per run)
    sub = sigs[: len(sigs) // 16]
    sub_n = sum(s[1] for s in sub)
    dt1 = _median_timed(lambda: run(sub, 1), reps)
    model = cpu_model()
    return {"value": round(2 * total / dt / 1e9, 4), "unit": "Gsamples/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": nproc,
            "sample": f"the full workload ({len(sigs)} signals, {total} samples), forward then forward-adjoint, "
                      f"1 warm-up then the median of {reps} runs ({dt:.2f} s each) on {threads} threads (all cores "
                      f"of this process's share; nproc {nproc}), {model}; CPU restatement of the reference "
                      f"algorithm (its Highway build is unavailable offline), dynamic parallel-for over signals",
            "single_core": {"value": round(2 * sub_n / dt1 / 1e9, 4), "cores": 1,
                            "sample": f"first {len(sub)} signals ({sub_n} samples), 1 warm-up then the median of "
                                      f"{reps} runs ({dt1:.2f} s each)"}}


def pmc_child(tuning):
    """One step of the workload, untimed (run under rocprofv3 counter passes),
    with the chain built from the planning knobs the timed steps used."""
    import torch
    import lifted_amd as la
    from lifted_amd.batch import BatchChain
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    sigs, total = make_workload()
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    x = torch.randn(total, generator=gen, device=dev, dtype=torch.float32)
    bufs = [torch.empty_like(x), torch.empty_like(x)]
    knobs = {k: v for k, v in tuning.items() if k != "chunks"}
    chain = BatchChain(sigs, (la.Transform.Forward, la.Transform.ForwardAdjoint), dtype=torch.float32,
                       chunks=tuning.get("chunks", 2), tuning=knobs)
    chain.execute(x, bufs)
    torch.cuda.synchronize()
    print("pmc child done", flush=True)


def _time_calls(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def workload_c1(steps, warmup, cpu):
    """One float64 signal of 65536 N(0,1) samples, Daubechies2, 8 levels,
    Zero boundary: forward then inverse (the reference's CPU-runnable case);
    latency-bound on a GPU."""
    import torch
    import lifted_amd as la
    x_np = np.random.default_rng(20261201).standard_normal(65536)
    x = torch.tensor(x_np, device="cuda")
    y = torch.empty_like(x)
    z = torch.empty_like(x)

    def fi():
        la.lwt_forward(x, la.Wavelet.Daubechies2, 0, level=[8], out=y)
        la.lwt_inverse(y, la.Wavelet.Daubechies2, 0, level=[8], out=z)

    ms = _time_calls(fi, steps * 10, warmup)
    res = {"workload": "c1: 1-D float64 N(0,1), n = 65536, Daubechies2, 8 levels, Zero BC, forward then inverse",
           "ms_F_plus_I": round(ms, 4), "Gsamples_per_s": round(2 * 65536 / (ms * 1e-3) / 1e9, 3),
           "reconstruction_max_abs_err": float((z - x).abs().max())}
    if cpu:
        import oracle as O
        reps = 200
        t = time.perf_counter()
        for _ in range(reps):
            yc = O.lwt(x_np, 2, O.FORWARD, O.ZERO, level=[8])
            O.lwt(yc, 2, O.INVERSE, O.ZERO, level=[8])
        dt = (time.perf_counter() - t) / reps
        res["cpu_baseline"] = {"ms": round(dt * 1e3, 4), "cores": 1, "kind": "port",
                               "sample": f"forward + inverse, mean of {reps} runs (a single line runs on one "
                                         "thread in the reference)"}
    return res


def workload_c2(steps, warmup, cpu):
    """4096^2 float32, sum of 32 Gaussians + 0.01 N(0,1) noise; CDF 9/7, 5 levels."""
    import torch
    import lifted_amd as la
    rng = np.random.default_rng(20261201 + 1)
    n = 4096
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float32)
    img = np.zeros((n, n), np.float32)
    for _ in range(32):
        cy, cx = rng.uniform(0, n, 2)
        sig = np.exp(rng.uniform(np.log(64), np.log(1024)))
        amp = rng.standard_normal()
        img += (amp * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))).astype(np.float32)
    img += 0.01 * rng.standard_normal((n, n)).astype(np.float32)
    x = torch.tensor(img, device="cuda")
    out = torch.empty_like(x)
    N = x.numel()
    rows = []
    for mode_name, level in (("separable", [5, 5]), ("level-by-level", 5)):
        for tf in (0, 1, 2):
            for plain in (False, True):
                ms = _time_calls(lambda: la._lwt(x, la.Wavelet.CDF9_7, 0, tf, (0, 1), level, out=out, plain=plain),
                                 steps, warmup)
                rows.append({"mode": mode_name, "transform": ["F", "I", "FA"][tf], "plain": plain,
                             "ms": round(ms, 4), "GBps_compulsory": round(2 * N * 4 / (ms * 1e-3) / 1e9, 1)})
    res = {"workload": "c2: 4096^2 float32 sum of 32 Gaussians + noise, CDF9_7, 5 levels, Zero BC",
           "rows": rows,
           "note": "GBps_compulsory = 2 N 4 B / t (each sample read and written once); a one-pass-per-axis "
                   "design moves 2 N 4 B per axis pass, so its ceiling is half of that"}
    for r in rows:
        if not r["plain"]:
            p = [q for q in rows if q["mode"] == r["mode"] and q["transform"] == r["transform"] and q["plain"]][0]
            r["vs_plain"] = round(p["ms"] / r["ms"], 3)
    if cpu:
        import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        t = time.perf_counter()
        O.lwt(img, la.Wavelet.CDF9_7, O.FORWARD, O.ZERO, level=[5, 5], n_threads=threads)
        res["cpu_baseline"] = {"ms": round((time.perf_counter() - t) * 1e3, 1), "cores": threads, "kind": "port",
                               "sample": "separable forward, once"}
    return res


def workload_c3(steps, warmup):
    """512^3 float32 N(0,1), CDF 9/7 levels (4, 4, 4): forward, loss, backward."""
    import torch
    import lifted_amd.autograd as AG
    g = torch.Generator(device="cuda").manual_seed(20261201 + 2)
    x = torch.randn(512, 512, 512, device="cuda", generator=g, requires_grad=True)
    N = x.numel()

    def step():
        c = AG.forward(x, 100907, 0, axes=(0, 1, 2), level=(4, 4, 4))
        loss = (torch.sign(c) * torch.clamp(c.abs() - 0.1, min=0.0)).abs().sum()
        x.grad = None
        loss.backward()

    ms = _time_calls(step, steps, warmup)
    c = AG.forward(x.detach(), 100907, 0, axes=(0, 1, 2), level=(4, 4, 4))
    ms_f = _time_calls(lambda: AG.forward(x.detach(), 100907, 0, axes=(0, 1, 2), level=(4, 4, 4)), steps, warmup)
    import lifted_amd as la
    ms_a = _time_calls(lambda: la.lwt_forward_adjoint(c, 100907, 0, axes=(0, 1, 2), level=[4, 4, 4]), steps, warmup)
    return {"workload": "c3: 512^3 float32 N(0,1), CDF9_7 levels (4,4,4), Zero BC, autograd F + loss + FA",
            "ms_per_iteration": round(ms, 3), "forward_ms": round(ms_f, 3), "adjoint_ms": round(ms_a, 3),
            "GBps_compulsory_forward": round(2 * N * 4 / (ms_f * 1e-3) / 1e9, 1),
            "GBps_compulsory_adjoint": round(2 * N * 4 / (ms_a * 1e-3) / 1e9, 1),
            "GBps_per_axis_pass_forward": round(3 * 2 * N * 4 / (ms_f * 1e-3) / 1e9, 1)}


def workload_c5(steps, warmup):
    """16 x 8192^2 1/f^2 fields, CDF 9/7, 6 levels (level-by-level), F + I."""
    import torch
    import lifted_amd as la
    B, n = 16, 8192
    g = torch.Generator(device="cuda").manual_seed(20261201 + 4)
    x64 = torch.empty(B, n, n, dtype=torch.float64, device="cuda")
    ky = torch.fft.fftfreq(n, device="cuda").view(n, 1)
    kx = torch.fft.rfftfreq(n, device="cuda").view(1, n // 2 + 1)
    k2 = ky * ky + kx * kx
    filt = torch.where(k2 > 0, 1.0 / k2, torch.zeros_like(k2))
    for b in range(B):
        w = torch.randn(n, n, device="cuda", generator=g)
        f = torch.fft.irfft2(torch.fft.rfft2(w) * filt, s=(n, n))
        x64[b] = ((f - f.mean()) / f.std()).double()
        del w, f
    axes = (1, 2)
    ref = la.lwt_forward(x64, 100907, 0, axes=axes, level=6)
    rows, cases = [], []
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16), ("fp16", torch.float16)):
        xs = x64.to(dt)
        y = la.lwt_forward(xs, 100907, 0, axes=axes, level=6)
        z = la.lwt_inverse(y, 100907, 0, axes=axes, level=6)
        e_coef = float((y.double() - ref).norm() / ref.norm())
        e_coef_max = float((y.double() - ref).abs().max())
        e_coef_max_rel = e_coef_max / float(ref.abs().max())
        e_rt = float((z.double() - xs.double()).norm() / xs.double().norm())
        e_rt64 = float((z.double() - x64).norm() / x64.norm())
        del y, z
        rows.append({"storage": name, "coef_rel_err_vs_fp64": e_coef, "coef_max_abs_err_vs_fp64": e_coef_max,
                     "coef_max_err_rel_to_max_coef": e_coef_max_rel, "roundtrip_rel_err": e_rt,
                     "roundtrip_rel_err_vs_fp64_input": e_rt64})
        cases.append((xs, torch.empty_like(xs), torch.empty_like(xs)))
    torch.cuda.empty_cache()

    def fi(xs, yo, zo):
        la.lwt_forward(xs, 100907, 0, axes=axes, level=6, out=yo)
        la.lwt_inverse(yo, 100907, 0, axes=axes, level=6, out=zo)

    # the storage types are timed in turns (one F + I each per round, median of
    # the rounds), so that run order (clocks, caches) favours none of them
    for c in cases:
        for _ in range(max(warmup, 1)):
            fi(*c)
    torch.cuda.synchronize()
    times = [[] for _ in cases]
    for _ in range(max(steps, 1)):
        for i, c in enumerate(cases):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fi(*c)
            b.record()
            b.synchronize()
            times[i].append(a.elapsed_time(b))
    for row, (xs, _, _), ts in zip(rows, cases, times):
        ms = sorted(ts)[len(ts) // 2]
        es = xs.element_size()
        row.update({"ms_F_plus_I": round(ms, 3),
                    "GBps_compulsory": round(4 * xs.numel() * es / (ms * 1e-3) / 1e9, 1),
                    "Gsamples_per_s": round(2 * xs.numel() / (ms * 1e-3) / 1e9, 2)})
    del cases
    torch.cuda.empty_cache()
    return {"workload": "c5: 16 x 8192^2 1/f^2 fields (unit variance), CDF9_7, level=6 (level by level), "
                        "Zero BC, forward + inverse; errors are relative L2 against the same kernels in fp64",
            "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a device copy of the workload's input size (HBM counter calibration)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tuning-json", default="{}", help=argparse.SUPPRESS)
    ap.add_argument("--workload", choices=["c4", "c1", "c2", "c3", "c5"], default="c4",
                    help="c4 (default): the batched headline; c1 / c2 / c3 / c5: secondary workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path: "
              "every signal is independent, so the workload would shard by signal with no exchange.",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child(json.loads(args.tuning_json))
        return
    if args.workload != "c4":
        import torch
        torch.cuda.set_device(0)
        steps = min(args.steps, 10)
        fn = {"c1": lambda: workload_c1(steps, 2, not args.no_cpu),
              "c2": lambda: workload_c2(steps, 2, not args.no_cpu), "c3": lambda: workload_c3(steps, 2),
              "c5": lambda: workload_c5(steps, 1)}[args.workload]
        print(json.dumps(fn()))
        return
    import torch
    import lifted_amd as la
    from lifted_amd.batch import BatchPlan
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    sigs, total = make_workload()
    if args.copy_calib:
        a = torch.empty(total, device=dev, dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": total * 4, "copy_bytes_written": total * 4, "launches": 3}))
        return
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED)
    x = torch.randn(total, generator=gen, device=dev, dtype=torch.float32)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    # planning knobs chosen by the autotuner on this data (defaults first):
    # the forward and forward-adjoint plans as one pipelined chain, with the
    # chunk count searched too
    from lifted_amd.batch import BatchChain
    from lifted_amd.tune import tune_batch
    TF = (la.Transform.Forward, la.Transform.ForwardAdjoint)
    t_tune = time.perf_counter()
    best, tried = tune_batch(sigs, TF, x, reps=5, warmup=2, chain=True)
    t_tune = time.perf_counter() - t_tune
    print(f"[bench] autotune {t_tune:.1f}s: " + ", ".join(f"{k}={ms:.3f}ms" for k, ms in tried), file=sys.stderr,
          flush=True)
    knobs = {k: v for k, v in best.items() if k != "chunks"}
    t_plan = time.perf_counter()
    chain = BatchChain(sigs, TF, dtype=torch.float32, chunks=best.get("chunks", 2), tuning=knobs)
    t_plan = time.perf_counter() - t_plan
    bufs = [y, z]
    print(f"[bench] {len(sigs)} signals, {total} samples, chain of {len(TF)} plans over {chain.num_chunks} "
          f"chunks built in {t_plan:.1f}s", file=sys.stderr, flush=True)
    # eager chain steps with HIP events around every launch (per-kernel roofline)
    eager_ms, launches = time_chain(chain, x, bufs, args.steps, args.warmup)
    dom = chain_dominant(chain, launches, best.get("strip", 0))
    achieved = dom["alg_bytes_per_launch"] / (dom["avg_launch_ms"] * 1e-3)
    print(f"[bench] eager: {eager_ms:.3f} ms/step, {2 * total / (eager_ms * 1e-3) / 1e9:.2f} Gsamples/s; dominant "
          f"{dom['symbol']} big stages {dom['avg_launch_ms']:.3f} ms/launch, {achieved / 1e9:.0f} GB/s", file=sys.stderr,
          flush=True)
    eager_out = z.clone()
    # the timed steps: the chain captured once as a hipGraph and replayed K
    # times (the way a serving loop runs it); its output must equal the eager
    # steps' bit for bit, else the eager steps are reported
    z.zero_()
    step_mode = (f"forward + forward-adjoint plans as one chain pipelined over {chain.num_chunks} signal chunks, "
                 "captured once in a hipGraph, replayed K times (output checked equal to the eager steps)")
    try:
        ms = time_graph(lambda: chain.execute(x, bufs), args.steps, args.warmup)
        if not torch.equal(z, eager_out):
            raise RuntimeError("graph replay output differs from the eager steps")
        print(f"[bench] graph replay: {ms:.3f} ms/step, {2 * total / (ms * 1e-3) / 1e9:.2f} Gsamples/s "
              "(output equal to eager)", file=sys.stderr, flush=True)
    except RuntimeError as e:
        ms = eager_ms
        step_mode = f"eager chain steps (hipGraph replay not used: {e})"
        print(f"[bench] {step_mode}", file=sys.stderr, flush=True)
    del eager_out
    value = 2 * total / (ms * 1e-3) / 1e9
    # the same two plans run one after the other (no pipelining): stage times
    fwd = BatchPlan(sigs, la.Transform.Forward, dtype=torch.float32, tuning=knobs)
    adj = BatchPlan(sigs, la.Transform.ForwardAdjoint, dtype=torch.float32, tuning=knobs)
    plans = [fwd, adj]
    seq_ms, stage_ms, _ = time_plans(plans, x, bufs, args.steps, args.warmup)
    try:
        seq_graph_ms = time_graph(lambda: (fwd.execute(x, y), adj.execute(y, z)), args.steps, args.warmup)
    except RuntimeError:
        seq_graph_ms = None
    print(f"[bench] plans in turn (no pipelining): {seq_ms:.3f} ms/step eager, {seq_graph_ms} graph",
          file=sys.stderr, flush=True)
    # copies of the input size: the achievable streaming rate on this device
    # (torch's copy, and the library's reference copies: one float4 per thread,
    # one 2048-float window per wavefront like the register tiles)
    copy_gbs = measure_copies(x)
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gsamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched 1-D fwd+adjoint: 4096 float32 N(0,1) signals, lengths 2^U(10,20), "
                               "random non-Lazy wavelet, levels U[3,max], Zero BC",
                   "signals": len(sigs), "samples": total, "seed": SEED},
        "hbm_frac_whole_step": round(4 * total * 4 / (ms * 1e-3) / HBM_PEAK, 4),
        "tuning": {"chosen": best, "candidates": [{"knobs": k, "ms_per_step": round(t, 4)} for k, t in tried],
                   "seconds": round(t_tune, 1)},
        "copy_GBps": copy_gbs,
        "step_mode": step_mode,
        "eager": {"ms_per_step": round(eager_ms, 4), "Gsamples_per_s": round(2 * total / (eager_ms * 1e-3) / 1e9, 3),
                  "note": "the same chain launched eagerly with HIP events around every launch (roofline source)"},
        "unpipelined": {"eager_ms_per_step": round(seq_ms, 4),
                        "graph_ms_per_step": None if seq_graph_ms is None else round(seq_graph_ms, 4),
                        "note": "the forward plan, then the forward-adjoint plan, each over all signals"},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "kernel": dom["symbol"].rstrip("(") + " (big-stage launches)",
                     "launches_per_step": dom["launches_per_step"],
                     "avg_launch_ms": round(dom["avg_launch_ms"], 4),
                     "alg_bytes_per_launch": round(dom["alg_bytes_per_launch"]),
                     "items_per_launch": dom["items"]},
        "stages_unpipelined": [{"plan": "fwd" if p == 0 else "adj", "stage": s, "ms": round(t, 4),
                                **plans[p].stage_info(s)} for p, s, t in stage_ms],
        "launches": [{"transform": int(chain.transforms[r["transform_index"]]), "chunk": r["chunk"],
                      "stage": r["stage"], "big": r["big"], "kind": r["kind"], "threads": r["threads"],
                      "items": r["items"], "alg_bytes": round(r["alg_bytes"]), "ms": round(r["ms"], 4)}
                     for r in launches],
    }
    if not args.no_plain:
        # plain = one launch per level: the same register-tile kernels run one
        # level per pass (the band round-trips through HBM between levels),
        # short bands one level per LDS launch; as plans in turn and as the same
        # pipelined chain, both graph-replayed like the fused step (the faster
        # one is the comparison). The LDS-only plain plans (the generic step
        # interpreter, one level per launch) are a secondary figure.
        pf = BatchPlan(sigs, la.Transform.Forward, dtype=torch.float32, plain=True)
        pa = BatchPlan(sigs, la.Transform.ForwardAdjoint, dtype=torch.float32, plain=True)
        pms_eager, _, _ = time_plans([pf, pa], x, [y, z], args.plain_steps, 1)
        try:
            pms_plans = time_graph(lambda: (pf.execute(x, y), pa.execute(y, z)), args.plain_steps, 1)
        except RuntimeError:
            pms_plans = pms_eager
        res["plain_stages"] = pf.num_stages + pa.num_stages
        pf.close()
        pa.close()
        pc = BatchChain(sigs, TF, dtype=torch.float32, chunks=chain.num_chunks, plain=True)
        try:
            pms_chain = time_graph(lambda: pc.execute(x, bufs), args.plain_steps, 1)
        except RuntimeError:
            pms_chain = float("inf")
        pc.close()
        pms = min(pms_plans, pms_chain)
        res["plain_ms_per_step"] = round(pms, 4)
        res["plain_detail"] = {"plans_graph_ms": round(pms_plans, 4), "plans_eager_ms": round(pms_eager, 4),
                               "chain_graph_ms": round(pms_chain, 4)}
        res["vs_plain"] = round(pms / ms, 3)
        res["plain_note"] = ("one launch per level with the same register-tile kernels (levels_per_pass = 1), "
                             "short bands one level per LDS launch; the faster of plans in turn and the pipelined "
                             "chain, both hipGraph-replayed like the fused step")
        print(f"[bench] plain (register tiles, one level per launch): plans {pms_plans:.3f} ms, chain "
              f"{pms_chain:.3f} ms (graph)", file=sys.stderr, flush=True)
        qf = BatchPlan(sigs, la.Transform.Forward, dtype=torch.float32, plain=True, lds_only=True)
        qa = BatchPlan(sigs, la.Transform.ForwardAdjoint, dtype=torch.float32, plain=True, lds_only=True)
        qms, _, _ = time_plans([qf, qa], x, [y, z], args.plain_steps, 1)
        res["plain_lds_ms_per_step"] = round(qms, 4)
        res["vs_plain_lds"] = round(qms / ms, 3)
        print(f"[bench] plain LDS-only: {qms:.3f} ms/step eager", file=sys.stderr, flush=True)
        qf.close()
        qa.close()
    chain.close()
    fwd.close()
    adj.close()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(sigs, total, x.cpu().numpy())
        print(f"[bench] cpu: {res['cpu_baseline']['value']} Gsamples/s", file=sys.stderr, flush=True)
    if not args.no_traffic:
        del x, y, z
        torch.cuda.empty_cache()
        traffic, note = measure_traffic(total, dom["symbol"], best, grids={64 * i for i in dom["items"]})
        res["roofline"]["traffic"] = None if traffic is None else round(traffic)
        res["traffic_detail"] = note
        print(f"[bench] traffic: {traffic} ({note})", file=sys.stderr, flush=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
