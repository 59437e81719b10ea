"""Batched plans (many packed 1-D signals of uneven length / wavelet / level
count) against the oracle, per signal, fused and plain (one level per stage)."""
import itertools

import numpy as np
import pytest

from helpers import WAVELETS, assert_parity


def line_bands(n, L):
    from lifted_amd.layout import line_bands as lb
    return lb(n, L)

pytestmark = pytest.mark.gpu


def make_signals(seed, nsig, lo, hi, oracle):
    g = np.random.default_rng(seed)
    sigs, off = [], 0
    for _ in range(nsig):
        n = int(round(2 ** g.uniform(np.log2(lo), np.log2(hi))))
        w = int(g.choice([w for w in WAVELETS if w != 0]))
        L = int(g.integers(1, max(2, oracle.max_level(w, n)) + 1))
        sigs.append((off, n, w, L))
        off += n
    return sigs, off


@pytest.mark.parametrize("variant", ["fused", "fused_strip1024", "plain", "plain_lds"])
@pytest.mark.parametrize("tf", [0, 1, 2, 3])
def test_batch_matches_oracle(gpu, oracle, tf, variant):
    """Fused plans (register windows of 2048 samples, and of 1024 with the
    strip knob), plain plans (one launch per level, register tiles) and plain
    LDS-only plans against the oracle per signal, bit for bit."""
    import torch
    from lifted_amd.batch import BatchPlan
    from lifted_amd.tune import tuning
    plain = variant.startswith("plain")
    sigs, total = make_signals(10 + tf, 40, 16, 300000, oracle)
    for bc in (0, 3):
        x = np.random.default_rng(1).standard_normal(total).astype(np.float32)
        with tuning(strip=1024 if variant == "fused_strip1024" else 0):
            plan = BatchPlan(sigs, tf, mode=bc, dtype=torch.float32, plain=plain, lds_only=variant == "plain_lds")
        if plain:
            assert plan.num_stages >= max(L for _, _, _, L in sigs)
        y = plan.execute(torch.from_numpy(x).to(gpu)).cpu().numpy()
        for off, n, w, L in sigs:
            ref = oracle.lwt(x[off:off + n], w, tf, bc, level=[L])
            got = y[off:off + n]
            assert_parity(got, ref, f"off={off} n={n} w={w} L={L} tf={tf} bc={bc} {variant}",
                          bands=list(line_bands(n, L).values()))


@pytest.mark.parametrize("line", [16, 64])
def test_register_tiles_down_to_short_bands(gpu, oracle, line):
    """Zero boundary, float32: with a whole-line threshold of 16 or 64 samples
    the register tiles keep running passes of masked windows on short bands
    (an extra stage) before the whole-line items; every transform type still
    equals the oracle per signal, at each signal's maximum level."""
    import torch
    from lifted_amd.batch import BatchPlan
    from lifted_amd.tune import tuning
    g = np.random.default_rng(40 + line)
    sigs, off = [], 0
    for _ in range(36):
        n = int(round(2 ** g.uniform(6, 18)))
        w = int(g.choice([w for w in WAVELETS if w != 0]))
        sigs.append((off, n, w, max(1, oracle.max_level(w, n))))
        off += n
    x = np.random.default_rng(2).standard_normal(off).astype(np.float32)
    xd = torch.from_numpy(x).to(gpu)
    for tf in (0, 1, 2, 3):
        with tuning(line_limit=line):
            plan = BatchPlan(sigs, tf, mode=0, dtype=torch.float32)
        y = plan.execute(xd).cpu().numpy()
        for o, n, w, L in sigs:
            ref = oracle.lwt(x[o:o + n], w, tf, 0, level=[L])
            assert_parity(y[o:o + n], ref, f"line={line} tf={tf} n={n} w={w} L={L}",
                          bands=list(line_bands(n, L).values()))


def test_batch_fp64_and_graph_replay(gpu, oracle):
    """fp64 batch; a captured graph of forward + forward-adjoint replays
    identically to eager execution (several replays, eager work in between)."""
    import torch
    from lifted_amd.batch import BatchPlan
    sigs, total = make_signals(99, 24, 500, 100000, oracle)
    x = torch.from_numpy(np.random.default_rng(2).standard_normal(total)).to(gpu)
    fp = BatchPlan(sigs, 0, dtype=torch.float64)
    fa = BatchPlan(sigs, 2, dtype=torch.float64)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    fp.execute(x, y)
    fa.execute(y, z)
    ye, ze = y.clone(), z.clone()
    xn = x.cpu().numpy()
    for off, n, w, L in sigs:
        r = oracle.lwt(xn[off:off + n], w, 0, 0, level=[L])
        assert_parity(ye.cpu().numpy()[off:off + n], r, f"fp64 n={n} w={w} L={L}")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    y.zero_()
    z.zero_()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fp.execute(x, y)
            fa.execute(y, z)
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(3):
        y.zero_()
        z.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ye) and torch.equal(z, ze)


def test_batch_fp32_graph_replay(gpu, oracle):
    """float32 plans (register launches on the caller's stream, LDS launches
    on the plan's side streams) captured in one graph: three replays, with
    eager work in between, equal the eager result bit for bit."""
    import torch
    from lifted_amd.batch import BatchPlan
    sigs, total = make_signals(7, 64, 100, 200000, oracle)
    x = torch.from_numpy(np.random.default_rng(4).standard_normal(total).astype(np.float32)).to(gpu)
    fp = BatchPlan(sigs, 0, dtype=torch.float32)
    fa = BatchPlan(sigs, 2, dtype=torch.float32)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    fp.execute(x, y)
    fa.execute(y, z)
    ye, ze = y.clone(), z.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fp.execute(x, y)
            fa.execute(y, z)
    torch.cuda.current_stream().wait_stream(s)
    for r in range(3):
        y.fill_(float(r + 1))
        z.fill_(-1.0)
        w = (x * 2.0).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ye) and torch.equal(z, ze), r
        assert torch.isfinite(w)


def test_autotuned_plans_match_oracle(gpu, oracle):
    """tune_batch times candidate planning knobs (defaults first) and the plans
    built with the chosen knobs -- and with every candidate's knobs -- still
    equal the oracle per signal."""
    import torch
    from lifted_amd.batch import BatchPlan
    from lifted_amd.tune import tune_batch
    sigs, total = make_signals(21, 48, 200, 150000, oracle)
    xn = np.random.default_rng(5).standard_normal(total).astype(np.float32)
    x = torch.from_numpy(xn).to(gpu)
    best, tried = tune_batch(sigs, (0, 2), x, reps=1, warmup=1)
    assert tried[0][0] == {"levels_per_pass": 0, "line_limit": 0, "side_streams": -1, "strip": 0}
    assert {k["strip"] for k, _ in tried} == {0, 1024}
    assert best in [k for k, _ in tried]
    for knobs, _ in tried:
        for tf in (0, 2):
            y = BatchPlan(sigs, tf, dtype=torch.float32, tuning=knobs).execute(x).cpu().numpy()
            for off, n, w, L in sigs:
                ref = oracle.lwt(xn[off:off + n], w, tf, 0, level=[L])
                assert_parity(y[off:off + n], ref, f"{knobs} tf={tf} n={n} w={w} L={L}")


def test_stream_copy_reference_is_exact(gpu):
    """lfx_stream_copy (the benchmark's streaming-rate reference): both access
    patterns copy the rounded-down byte count exactly and leave the rest."""
    import ctypes
    import torch
    from lifted_amd import _lib
    L = _lib.lib()
    L.lfx_stream_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
    n = 3 * 2048 * 5 + 700
    x = torch.randn(n, device=gpu)
    for pattern, unit in ((0, 4), (1, 2048)):
        y = torch.full_like(x, -7.0)
        _lib.check(L.lfx_stream_copy(x.data_ptr(), y.data_ptr(), n * 4, pattern, None))
        torch.cuda.synchronize()
        m = n // unit * unit
        assert torch.equal(y[:m], x[:m]) and bool((y[m:] == -7.0).all()), pattern
    with pytest.raises(_lib.LfxError):
        _lib.check(L.lfx_stream_copy(x.data_ptr(), y.data_ptr(), 16, 2, None))


@pytest.mark.parametrize("storage", ["bfloat16", "float16"])
def test_batch_16bit_storage(gpu, oracle, storage):
    """16-bit storage, fp32 arithmetic: intermediate s bands stay fp32 in the
    plan's scratch, so every output equals the fp32 oracle on the 16-bit input,
    rounded once to the storage type (bit for bit)."""
    import torch
    from lifted_amd.batch import BatchPlan
    tdt = getattr(torch, storage)
    sigs, total = make_signals(31, 24, 100, 120000, oracle)
    x = torch.from_numpy(np.random.default_rng(8).standard_normal(total).astype(np.float32)).to(tdt)
    x32 = x.float().numpy()
    for tf in (0, 1, 2, 3):
        y = BatchPlan(sigs, tf, dtype=tdt).execute(x.to(gpu)).float().cpu()
        for off, n, w, L in sigs:
            ref = torch.from_numpy(oracle.lwt(x32[off:off + n], w, tf, 0, level=[L])).to(tdt).float()
            assert torch.equal(y[off:off + n], ref), (storage, tf, off, n, w, L)


def test_batch_custom_programs_periodic_and_max_level(gpu, oracle):
    """Caller-defined step programs in a batch, the Periodic boundary on lines
    that fit one workgroup and on longer ones, and levels = 0 (max level),
    against the oracle."""
    import torch
    import lifted_amd as la
    from helpers import MOCK_WAVELETS, oracle_steps
    from lifted_amd.batch import BatchPlan
    cw = la.CustomWavelet(MOCK_WAVELETS["MultiStep"], width=8)
    st = oracle_steps(MOCK_WAVELETS["MultiStep"])
    g = np.random.default_rng(9)
    lens = [37, 300, 4000, 1000, 129]
    sigs, off = [], 0
    for i, n in enumerate(lens):
        sigs.append((off, n, cw.id if i % 2 else 100907, 0 if i == 2 else 2))
        off += n
    x = g.standard_normal(off).astype(np.float32)
    for bc, tf in itertools.product((2, 0, 4), (0, 2)):
        y = BatchPlan(sigs, tf, mode=bc, dtype=torch.float32).execute(torch.from_numpy(x).to(gpu)).cpu().numpy()
        for (o, n, w, L) in sigs:
            if w == cw.id:
                ref = oracle.lwt(x[o:o + n], 0, tf, bc, level=[L], steps=st, width=8)
            else:
                ref = oracle.lwt(x[o:o + n], w, tf, bc, level=[L if L else oracle.max_level(w, n)])
            assert_parity(y[o:o + n], ref, f"bc={bc} tf={tf} n={n} w={w} L={L}")

    # Periodic signals longer than the LDS limit (bands in the plan's global
    # scratch), power-of-two and other lengths, with short ones beside them,
    # float32 and float64, fused and plain
    g2 = np.random.default_rng(19)
    lens = [300000, 65536, 1000, (1 << 18) + 3, 40000]
    sigs2, off = [], 0
    for i, n in enumerate(lens):
        sigs2.append((off, n, [100907, 2, 10204, 3, 10608][i], [3, 8, 4, 12, 0][i]))
        off += n
    for dt, tf, plain in itertools.product((np.float32, np.float64), (0, 1, 2, 3), (False, True)):
        x2 = g2.standard_normal(off).astype(dt)
        tdt = torch.float32 if dt == np.float32 else torch.float64
        y2 = BatchPlan(sigs2, tf, mode=2, dtype=tdt, plain=plain).execute(torch.from_numpy(x2).to(gpu)).cpu().numpy()
        for (o, n, w, L) in sigs2:
            ref = oracle.lwt(x2[o:o + n], w, tf, 2, level=[L if L else oracle.max_level(w, n)])
            assert_parity(y2[o:o + n], ref, f"Periodic {dt.__name__} tf={tf} plain={plain} n={n} w={w} L={L}")


@pytest.mark.parametrize("chunks", [1, 2, 3])
def test_chain_pipelined_equals_plans(gpu, oracle, chunks):
    """A chain of Forward then ForwardAdjoint (and Inverse then
    InverseAdjoint) pipelined over 1, 2 or 3 signal chunks equals the plans
    run one after the other bit for bit, and the oracle per signal; Zero and
    Symmetric boundaries, float32 and float64; a captured graph of the chain
    replays identically three times with eager work in between."""
    import torch
    from lifted_amd.batch import BatchChain, BatchPlan
    sigs, total = make_signals(60 + chunks, 48, 100, 200000, oracle)
    for dt, bc, tfs in itertools.product((np.float32, np.float64), (0, 3), ((0, 2), (1, 3))):
        tdt = torch.float32 if dt == np.float32 else torch.float64
        xn = np.random.default_rng(chunks).standard_normal(total).astype(dt)
        x = torch.from_numpy(xn).to(gpu)
        ch = BatchChain(sigs, tfs, mode=bc, dtype=tdt, chunks=chunks)
        assert ch.num_chunks == chunks
        outs = ch.execute(x)
        torch.cuda.synchronize()
        y = BatchPlan(sigs, tfs[0], mode=bc, dtype=tdt).execute(x)
        z = BatchPlan(sigs, tfs[1], mode=bc, dtype=tdt).execute(y)
        assert torch.equal(outs[0], y) and torch.equal(outs[1], z), (dt, bc, tfs)
        yh, zh = outs[0].cpu().numpy(), outs[1].cpu().numpy()
        for off, n, w, L in sigs[::5]:
            ry = oracle.lwt(xn[off:off + n], w, tfs[0], bc, level=[L])
            assert_parity(yh[off:off + n], ry, f"chain step 0 {dt.__name__} bc={bc} n={n} w={w}")
            assert_parity(zh[off:off + n], oracle.lwt(ry, w, tfs[1], bc, level=[L]),
                          f"chain step 1 {dt.__name__} bc={bc} n={n} w={w}")
    # graph capture of the chain (float32, Zero boundary)
    x = torch.from_numpy(np.random.default_rng(8).standard_normal(total).astype(np.float32)).to(gpu)
    ch = BatchChain(sigs, (0, 2), mode=0, dtype=torch.float32, chunks=chunks)
    outs = [torch.empty_like(x), torch.empty_like(x)]
    ch.execute(x, outs)
    ref = [o.clone() for o in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            ch.execute(x, outs)
    torch.cuda.current_stream().wait_stream(s)
    for r in range(3):
        outs[0].fill_(float(r))
        outs[1].fill_(-1.0)
        w = (x * 3.0).sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(outs[0], ref[0]) and torch.equal(outs[1], ref[1]), r
        assert torch.isfinite(w)
