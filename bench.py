"""Benchmark: batched forward + forward-adjoint multi-level DWT on one MI355X.

Workload (the headline configuration): 4096 1-D float32 signals, lengths
round(2^U(10, 20)), each with a wavelet drawn uniformly from the 33 non-Lazy
predefined wavelets and a level count drawn uniformly from [3, max_level];
values iid N(0, 1); Zero boundary. One step = forward transform of every
signal, then the forward-adjoint of the result (two batched plans, four stage
launches). Inputs are resident in HBM before the timed region.

Reports Gsamples/s (2 x samples per step / time), the HBM roofline of the
dominant kernel (measured with HIP events on the launch stream), the plain
per-level kernels (one launch per level) for comparison, and the CPU
restatement of the reference timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "fwd+adjoint DWT Gsamples/s and % HBM roofline, fused/batched vs plain, MI355X"
HBM_PEAK = 8.0e12
SEED = 20261201 + 3  # fourth workload of the benchmark set
NSIG = 4096
CAP = 1 << 30


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


def time_plans(plans, x, bufs, steps, warmup):
    """Run the plans in sequence (x -> bufs[0] -> bufs[1] ...), timing each
    stage with HIP events on the launch stream. Returns (ms per step, per-stage
    average ms list of (plan index, stage, ms))."""
    import torch
    stream = torch.cuda.current_stream()

    def one_step(events=None):
        src = x
        for pi, p in enumerate(plans):
            for s in range(p.num_stages):
                p.execute_stage(s, src, bufs[pi])
                if events is not None:
                    events[pi][s].record(stream)
            src = bufs[pi]

    for _ in range(warmup):
        one_step()
    torch.cuda.synchronize()
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(p.num_stages)] for p in plans] for _ in range(steps)]
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for k in range(steps):
        starts[k].record(stream)
        one_step(ev[k])
    t1.record(stream)
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    stage_ms = []
    for pi, p in enumerate(plans):
        for s in range(p.num_stages):
            acc = 0.0
            for k in range(steps):
                prev = starts[k] if (pi == 0 and s == 0) else (
                    ev[k][pi][s - 1] if s > 0 else ev[k][pi - 1][plans[pi - 1].num_stages - 1])
                acc += prev.elapsed_time(ev[k][pi][s])
            stage_ms.append((pi, s, acc / steps))
    return total_ms / steps, stage_ms


def cpu_baseline(sigs, total, x_host):
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    y = np.empty_like(x_host)
    z = np.empty_like(x_host)
    t = time.perf_counter()
    O.batch(x_host, sigs, O.FORWARD, O.ZERO, out=y, n_threads=threads)
    O.batch(y, sigs, O.FORWARD_ADJOINT, O.ZERO, out=z, n_threads=threads)
    dt = time.perf_counter() - t
    return {"value": round(2 * total / dt / 1e9, 4), "unit": "Gsamples/s", "cores": threads, "kind": "port",
            "sample": f"the full workload ({len(sigs)} signals, {total} samples), forward then forward-adjoint, "
                      f"once, {dt:.2f} s, CPU restatement of the reference (its Highway build needs network "
                      f"access); dynamic parallel-for over signals"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a device copy of the workload's input size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no multi-GPU path: "
              "every signal is independent, so the workload would shard by signal with no exchange.",
              file=sys.stderr)
        sys.exit(2)
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
    t_plan = time.perf_counter()
    fwd = BatchPlan(sigs, la.Transform.Forward, dtype=torch.float32)
    adj = BatchPlan(sigs, la.Transform.ForwardAdjoint, dtype=torch.float32)
    t_plan = time.perf_counter() - t_plan
    print(f"[bench] {len(sigs)} signals, {total} samples, plans in {t_plan:.1f}s, stages "
          f"{fwd.num_stages}+{adj.num_stages}", file=sys.stderr, flush=True)
    ms, stage_ms = time_plans([fwd, adj], x, [y, z], args.steps, args.warmup)
    plans = [fwd, adj]
    dom = max(stage_ms, key=lambda t: t[2])
    info = plans[dom[0]].stage_info(dom[1])
    achieved = info["alg_bytes"] / (dom[2] * 1e-3)
    value = 2 * total / (ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gsamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched 1-D fwd+adjoint: 4096 float32 N(0,1) signals, lengths 2^U(10,20), "
                               "random non-Lazy wavelet, levels U[3,max], Zero BC",
                   "signals": len(sigs), "samples": total, "seed": SEED},
        "hbm_frac_whole_step": round(4 * total * 4 / (ms * 1e-3) / HBM_PEAK, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "kernel": f"k_batch stage {dom[1]} of the {'forward' if dom[0] == 0 else 'adjoint'} plan",
                     "kernel_ms": round(dom[2], 4), "alg_bytes": info["alg_bytes"]},
        "stages": [{"plan": "fwd" if p == 0 else "adj", "stage": s, "ms": round(t, 4),
                    **plans[p].stage_info(s)} for p, s, t in stage_ms],
    }
    print(f"[bench] fused: {ms:.3f} ms/step, {value:.2f} Gsamples/s", file=sys.stderr, flush=True)
    if not args.no_plain:
        pf = BatchPlan(sigs, la.Transform.Forward, dtype=torch.float32, plain=True)
        pa = BatchPlan(sigs, la.Transform.ForwardAdjoint, dtype=torch.float32, plain=True)
        pms, _ = time_plans([pf, pa], x, [y, z], args.plain_steps, 1)
        res["plain_ms_per_step"] = round(pms, 4)
        res["vs_plain"] = round(pms / ms, 3)
        res["plain_stages"] = pf.num_stages + pa.num_stages
        print(f"[bench] plain: {pms:.3f} ms/step ({pf.num_stages}+{pa.num_stages} launches)", file=sys.stderr,
              flush=True)
        pf.close()
        pa.close()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(sigs, total, x.cpu().numpy())
        print(f"[bench] cpu: {res['cpu_baseline']['value']} Gsamples/s", file=sys.stderr, flush=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
