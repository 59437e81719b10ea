"""The GPU kernel bodies compiled for the host (one lane, AddressSanitizer on
exact-size LDS buffers) against the oracle.

This checks the tiling logic of long lines -- windows, halos, valid-range
tracking, boundary folds at tile edges, multi-pass schedules -- without a GPU:
the line-length limit is lowered so that short lines take the tiled path.
Results must be bit-identical to the oracle (same operations, same order).
"""
import concurrent.futures
import itertools
import os
import subprocess

import numpy as np
import pytest

from helpers import MOCK_WAVELETS, WAVELETS, WIDTHS, oracle_steps

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "cpu_sim", "build", "lfx_sim")

# Every table wavelet and mock program by default (about 40 s on 8 cores);
# LFX_QUICK_TESTS=1 keeps a subset chosen to reach every step shape:
# Daubechies1 (unit steps, sqrt2 scale), Daubechies2 (the 1-D workload),
# Symlet5 (negative scale), Bior2_2 (repeat steps), Bior3_1 (two s-updates),
# Bior6_8 (widest stencil, width 21) and CDF9_7; and six mock programs with
# negative, zero and positive offsets of every step kind.
QUICK = os.environ.get("LFX_QUICK_TESTS") == "1"
SIM_WAVELETS = [1, 2, 105, 10202, 10301, 10608, 100907] if QUICK else [w for w in WAVELETS if w != 0]
SIM_MOCKS = [m for m in ("DUpdateNegOff", "SUpdatePosOff", "SUnitUpdateNegOff", "DRepeatUpdatePosOff", "MultiScale",
                         "MultiStep") if m in MOCK_WAVELETS] if QUICK else list(MOCK_WAVELETS)
JOBS = min(8, os.cpu_count() or 1)


def run_cases(cases):
    """Run case callables on a thread pool (each spends its time in a
    subprocess or in the oracle's C code); re-raise the first failure."""
    with concurrent.futures.ThreadPoolExecutor(JOBS) as ex:
        for f in [ex.submit(c) for c in cases]:
            f.result()


@pytest.fixture(scope="module")
def sim():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpu_sim")], check=True)
    return SIM


def run_sim(sim, w, tf, bc, x, L, limit, tile, k, steps=None, env=None):
    dt = 1 if x.dtype == np.float64 else 0
    args = [sim, "custom" if steps else str(w), str(tf), str(bc), str(dt), str(len(x)), str(L), str(limit),
            str(tile), str(k)]
    if steps:
        args.append(steps)
    r = subprocess.run(args, input=x.tobytes(), capture_output=True,
                       env=None if env is None else dict(os.environ, **env))
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    return np.frombuffer(r.stdout, dtype=x.dtype), r.stderr.decode()


def ok_bands(bc, n, L):
    m = n
    for _ in range(L):
        nd, ns = m // 2, m - m // 2
        if bc != 0 and (nd == 0 or (bc == 4 and (nd < 2 or ns < 2))):
            return False
        m = ns
    return True


def _case(sim, oracle, w, tf, bc, x, L, limit, tile, k, msg, ref_level=None, steps=None, spec=None, check=None,
          env=None):
    """One simulator run against the oracle (bit-exact)."""
    def run():
        ref = oracle.lwt(x, w if steps is None else 0, tf, bc, level=[L] if ref_level is None else ref_level,
                         steps=steps)
        got, info = run_sim(sim, w, tf, bc, x, L, limit=limit, tile=tile, k=k, steps=spec, env=env)
        np.testing.assert_array_equal(got, ref, err_msg=f"{msg}\n{info}")
        if check is not None:
            check(info)
    return run


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_tiled_all_wavelets(sim, oracle, w):
    g = np.random.default_rng(w)
    cases = []
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (3001, 2048)):
        L = min(7, max(1, oracle.max_level(w, n)))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n)
        cases.append(_case(sim, oracle, w, tf, bc, x, L, 200, 512, 0, f"w={w} bc={bc} tf={tf} n={n}",
                           check=lambda info: None if "pass" in info else pytest.fail(info)))
    run_cases(cases)


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_tiled_one_sample_end_tiles(sim, oracle, w):
    """Last tiles a single sample wide: a window touching the band end holds
    only the halo, so mirrored reads (Symmetric / Reflect) and adjoint folds
    must be covered by the halo rule, with plain LDS tiles and with register
    passes whose band-end pieces are cut from the tail."""
    g = np.random.default_rng(w + 3)
    cases = []
    for bc, tf, n in itertools.product((1, 3, 4), (0, 1, 2, 3), (2561, 4097)):
        L = min(4, oracle.max_level(w, n))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n).astype(np.float32)
        for tile, k in ((512, 0), (64, 2)):
            cases.append(_case(sim, oracle, w, tf, bc, x, L, 200, tile, k, f"w={w} bc={bc} tf={tf} n={n} tile={tile}"))
    run_cases(cases)


@pytest.mark.parametrize("tile,k,limit", [(128, 1, 100), (256, 2, 60), (1024, 4, 300), (64, 0, 50)])
def test_tiled_shapes_cdf97_float32(sim, oracle, tile, k, limit):
    g = np.random.default_rng(tile + k)
    cases = []
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (1777, 4096, 999)):
        x = g.standard_normal(n).astype(np.float32)
        cases.append(_case(sim, oracle, 100907, tf, bc, x, 6, limit, tile, k, f"bc={bc} tf={tf} n={n}"))
    run_cases(cases)


def _spec(steps):
    parts = []
    for s in oracle_steps(steps):
        tgt = 1 if s["target"] in ("d", 1) else 0
        taps = s.get("taps", [])
        parts.append(",".join(str(v) for v in [s["kind"], tgt, s["offset"], s.get("sign", 1), s["ntaps"]] +
                              [repr(float(t)) for t in taps]))
    return ";".join(parts)


@pytest.mark.parametrize("name", SIM_MOCKS)
def test_tiled_mock_wavelets(sim, oracle, name):
    """The reference's mock step programs (large positive / negative offsets
    exercise folds that land next to tile edges)."""
    st = oracle_steps(MOCK_WAVELETS[name])
    spec = _spec(MOCK_WAVELETS[name])
    g = np.random.default_rng(5)
    cases = []
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (1531, 1024)):
        x = g.standard_normal(n)
        cases.append(_case(sim, oracle, 0, tf, bc, x, 3, 100, 256, 0, f"{name} bc={bc} tf={tf} n={n}", steps=st,
                           spec=spec))
    run_cases(cases)


@pytest.mark.parametrize("name", SIM_MOCKS)
def test_whole_line_all_bcs_mock(sim, oracle, name):
    """Whole-line path (no tiling) with all five BCs, including Periodic folds
    at non-power-of-two lengths whose targets land mid-band."""
    st = oracle_steps(MOCK_WAVELETS[name])
    spec = _spec(MOCK_WAVELETS[name])
    g = np.random.default_rng(6)
    cases = []
    for bc, tf, n in itertools.product(range(5), (0, 1, 2, 3), (153, 510, 1000, 37)):
        x = g.standard_normal(n)
        cases.append(_case(sim, oracle, 0, tf, bc, x, 2, 1 << 20, 0, 0, f"{name} bc={bc} tf={tf} n={n}", steps=st,
                           spec=spec))
    run_cases(cases)


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_whole_line_all_bcs(sim, oracle, w):
    g = np.random.default_rng(w + 1)
    cases = []
    for bc, tf, n in itertools.product(range(5), (0, 1, 2, 3), (1000, 333)):
        L = min(4, max(1, oracle.max_level(w, n)))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n).astype(np.float32)
        cases.append(_case(sim, oracle, w, tf, bc, x, L, 1 << 20, 0, 0, f"w={w} bc={bc} tf={tf} n={n}"))
    run_cases(cases)


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_register_tiles_all_wavelets(sim, oracle, w):
    """Passes of four levels whose interior runs as register tiles (64-lane
    model of a wavefront) and whose band ends run as LDS tiles."""
    g = np.random.default_rng(w + 7)
    cases = []
    for bc, tf, n in itertools.product((0, 3), (0, 1, 2, 3), (30011, 24576)):
        L = min(6, oracle.max_level(w, n))
        x = g.standard_normal(n).astype(np.float32)
        cases.append(_case(sim, oracle, w, tf, bc, x, L, 2000, 8192, 4, f"w={w} bc={bc} tf={tf} n={n}"))
    run_cases(cases)


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_register_tiles_zero_boundary_whole_band(sim, oracle, w):
    """Zero boundary: register windows overlap the band ends and run masked,
    so register tiles cover whole bands, down to bands of a few hundred
    samples (one window holds the band) -- no LDS pieces at the ends."""
    g = np.random.default_rng(w + 13)
    cases = []
    for tf, n in itertools.product((0, 1, 2, 3), (5003, 700, 1024, 333)):
        L = min(9, oracle.max_level(w, n))
        x = g.standard_normal(n).astype(np.float32)

        def whole(info, n=n):
            if n >= 1024:
                assert "reg [0," in info, info
        cases.append(_case(sim, oracle, w, tf, 0, x, L, 64, 8192, 4, f"w={w} tf={tf} n={n}", check=whole))
    run_cases(cases)


@pytest.mark.parametrize("w", [1, 2, 10, 10301, 10608, 100503, 100907])
def test_register_tiles_1024_sample_windows(sim, oracle, w):
    """Register windows of 1024 samples (16 per lane, the strip knob's other
    setting): interior tiles with LDS band ends (Symmetric) and masked windows
    over the band ends (Zero), passes of one to four levels, all transforms."""
    g = np.random.default_rng(w + 17)
    cases = []
    for bc, tf, n, L in itertools.product((0, 3), (0, 1, 2, 3), (20011, 4096), (3, 6)):
        L = min(L, oracle.max_level(w, n))
        x = g.standard_normal(n).astype(np.float32)

        def has_reg(info):
            assert "reg [" in info, info
        cases.append(_case(sim, oracle, w, tf, bc, x, L, 600 if bc else 64, 8192, 4,
                           f"w={w} bc={bc} tf={tf} n={n} L={L} strip=1024", check=has_reg,
                           env={"LFX_SIM_STRIP": "1024"}))
    run_cases(cases)


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_register_tiles_fewer_levels(sim, oracle, w):
    """Register passes of one to three levels (level counts that are not a
    multiple of four; the last pass of a 5-level transform is one level)."""
    g = np.random.default_rng(w + 11)
    n = 20011
    cases = []
    for L, tf in itertools.product((1, 2, 3, 5), (0, 1, 2, 3)):
        if L > oracle.max_level(w, n):
            continue
        x = g.standard_normal(n).astype(np.float32)

        def has_reg(info):
            assert "reg [" in info, info
        cases.append(_case(sim, oracle, w, tf, 3, x, L, 600, 8192, 4, f"w={w} L={L} tf={tf}", check=has_reg))
    run_cases(cases)


def run_sim_rows(sim, w, tf, x2d, L, limit, tile, k):
    """R equal lines through the simulator's row mode (float32)."""
    R, n = x2d.shape
    args = [sim, str(w), str(tf), "0", "0", str(n), str(L), str(limit), str(tile), str(k)]
    env = dict(os.environ, LFX_SIM_ROWS=str(R))
    r = subprocess.run(args, input=np.ascontiguousarray(x2d).tobytes(), capture_output=True, env=env)
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    return np.frombuffer(r.stdout, dtype=np.float32).reshape(R, n), r.stderr.decode()


@pytest.mark.parametrize("w", SIM_WAVELETS)
def test_row_mode_equal_short_lines(sim, oracle, w):
    """Row mode (four equal lines per wavefront, one per 16-lane row, reads
    past a row's ends are zeros): whole bands of at most 512 samples, Zero
    boundary, groups of four lines and a partial group, bit-exact per line."""
    g = np.random.default_rng(w + 17)
    cases = []
    for tf, n, R in itertools.product((0, 1, 2, 3), (512, 333, 100), (6,)):
        L = min(6, oracle.max_level(w, n))
        if L < 1:
            continue
        x = g.standard_normal((R, n)).astype(np.float32)

        def case(tf=tf, n=n, x=x, L=L):
            got, info = run_sim_rows(sim, w, tf, x, L, limit=32, tile=8192, k=4)
            for r in range(x.shape[0]):
                ref = oracle.lwt(x[r], w, tf, 0, level=[L])
                np.testing.assert_array_equal(got[r], ref, err_msg=f"w={w} tf={tf} n={n} L={L} line {r}\n{info}")
            if n >= 100 and w in (2, 100907):
                assert "rows [0," in info, info
        cases.append(case)
    run_cases(cases)


def test_xcd_block_renumbering(sim):
    """The register and window-copy launches renumber workgroups per XCD
    (lfx_xcd.hpp): every grid size up to 5000 must map onto itself one to one,
    with each XCD's workgroups numbered consecutively."""
    r = subprocess.run([os.path.join(os.path.dirname(sim), "xcd_check")], capture_output=True)
    assert r.returncode == 0 and r.stdout.decode().strip() == "ok", r.stdout.decode() + r.stderr.decode()


def test_boundary_maps_closed_form(sim):
    """The kernels' Symmetric / Reflect maps are closed forms (periods 2n and
    2n - 2, no loop). They must equal the reference's iterated mirrors
    (src/lifted-common.hpp:750-777) at every index |i| <= 4n + 3 for n >= 1
    (Symmetric) and n >= 2 (Reflect), and return at once on degenerate bands."""
    r = subprocess.run([os.path.join(os.path.dirname(sim), "bcmap_check")], capture_output=True, timeout=60)
    out = r.stdout.decode()
    assert r.returncode == 0 and out.startswith("ok"), out + r.stderr.decode()
    print(out.strip())


def _lvl2d_sim(w, tf, x, L, storage=0):
    sim2 = os.path.join(HERE, "cpu_sim", "build", "sim_lvl2d")
    B, R, C = x.shape
    data = x.astype(np.float32).tobytes() if storage == 0 else x.tobytes()
    r = subprocess.run([sim2, str(w), str(tf), str(storage), str(B), str(R), str(C), str(L)], input=data,
                       capture_output=True)
    if r.returncode == 3:
        return None, r.stderr.decode()
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    dt = np.float32 if storage == 0 else np.uint16
    return np.frombuffer(r.stdout, dtype=dt).reshape(x.shape), r.stderr.decode()


def test_two_axis_level_tiles_match_oracle(sim, oracle):
    """The two-axis level tiles (lfx_lvl2d.hpp) on the 64-lane CPU model,
    driven by the library's plan (every tile of every level, with the
    workspace ping-pong): level-by-level transforms over the last two axes of
    batched 2-D arrays, odd and even sizes, band-end tiles, all transforms,
    bit for bit against the oracle; shapes whose bands are too short are
    refused (the library then takes the chained path)."""
    g = np.random.default_rng(2024)
    cases = [(100907, (1, 300, 200), 1), (100907, (2, 300, 200), 3), (2, (1, 129, 257), 4), (10608, (1, 500, 260), 2),
             (1, (2, 64, 64), 3), (3, (3, 97, 131), 3), (105, (1, 250, 333), 3), (10202, (2, 65, 66), 4),
             (10301, (1, 120, 90), 2), (100907, (1, 1000, 70), 5)]
    planned = 0
    for w, shape, L in cases:
        x = g.standard_normal(shape).astype(np.float32)
        for tf in (0, 1, 2, 3):
            got, info = _lvl2d_sim(w, tf, x, L)
            if got is None:
                continue
            planned += 1
            ref = oracle.lwt(x, w, tf, 0, axes=[1, 2], level=L)
            np.testing.assert_array_equal(got, ref, err_msg=f"w={w} shape={shape} L={L} tf={tf}\n{info}")
    assert planned >= 30
    # too short for the tiles: refused
    got, info = _lvl2d_sim(100503, 0, g.standard_normal((1, 77, 1000)).astype(np.float32), 5)
    assert got is None and "no plan" in info


def test_two_axis_level_tiles_bf16(sim, oracle):
    """bf16 storage through the two-axis tiles: every level rounds its result
    once to bf16 (its ss quadrant goes through a bf16 workspace), which is the
    oracle applied one level at a time with that rounding."""
    import torch
    from helpers import oracle_level_by_level_rounded, round_to
    x = round_to(np.random.default_rng(5).standard_normal((2, 200, 150)), "bfloat16")
    xb = torch.from_numpy(x).to(torch.bfloat16).view(torch.uint16).numpy()
    for tf in (0, 1):
        got, info = _lvl2d_sim(100907, tf, xb, 3, storage=2)
        assert got is not None, info
        gotf = torch.from_numpy(got.copy()).view(torch.bfloat16).float().numpy()
        ref = oracle_level_by_level_rounded(oracle, x, 100907, tf, 0, [1, 2], 3, "bfloat16")
        np.testing.assert_array_equal(gotf, ref, err_msg=f"tf={tf}")


def _across_sim(w, tf, x, L, storage=0):
    sim3 = os.path.join(HERE, "cpu_sim", "build", "sim_across")
    B, N, C = x.shape
    data = x.astype(np.float32).tobytes() if storage == 0 else x.tobytes()
    r = subprocess.run([sim3, str(w), str(tf), str(storage), str(B), str(N), str(C), str(L)], input=data,
                       capture_output=True)
    if r.returncode == 3:
        return None, r.stderr.decode()
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    dt = np.float32 if storage == 0 else np.uint16
    return np.frombuffer(r.stdout, dtype=dt).reshape(x.shape), r.stderr.decode()


def test_across_lines_match_oracle(sim, oracle):
    """The across-line kernel (lfx_across.hpp) on the 64-lane CPU model with a
    model of its LDS slab (same padded layout), driven by the library's
    planning: lines along a strided axis, blocks of 16 neighbouring lines
    (partial last blocks), every pass of up to four levels in place in the
    slab, all transforms, bit for bit against the oracle; lines too short for
    the row tiles' passes are refused."""
    g = np.random.default_rng(77)
    cases = [(100907, (2, 512, 40), 4), (100907, (1, 300, 33), 6), (2, (3, 257, 70), 5), (10608, (1, 500, 31), 2),
             (1, (2, 64, 64), 3), (105, (1, 129, 96), 3), (10301, (1, 90, 5), 1), (10, (2, 512, 17), 8)]
    planned = 0
    for w, shape, L in cases:
        x = g.standard_normal(shape).astype(np.float32)
        for tf in (0, 1, 2, 3):
            got, info = _across_sim(w, tf, x, L)
            if got is None:
                continue
            planned += 1
            ref = oracle.lwt(x, w, tf, 0, axes=[1], level=[L])
            np.testing.assert_array_equal(got, ref, err_msg=f"w={w} shape={shape} L={L} tf={tf}\n{info}")
    assert planned >= 24
    # a band too short for a pass's stencil: refused
    got, info = _across_sim(10, 0, g.standard_normal((1, 40, 8)).astype(np.float32), 3)
    assert got is None and "no plan" in info
