"""The GPU kernel bodies compiled for the host (one lane, AddressSanitizer on
exact-size LDS buffers) against the oracle.

This checks the tiling logic of long lines -- windows, halos, valid-range
tracking, boundary folds at tile edges, multi-pass schedules -- without a GPU:
the line-length limit is lowered so that short lines take the tiled path.
Results must be bit-identical to the oracle (same operations, same order).
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

from helpers import MOCK_WAVELETS, WAVELETS, WIDTHS, oracle_steps

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "cpu_sim", "build", "lfx_sim")


@pytest.fixture(scope="module")
def sim():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpu_sim")], check=True)
    return SIM


def run_sim(sim, w, tf, bc, x, L, limit, tile, k, steps=None):
    dt = 1 if x.dtype == np.float64 else 0
    args = [sim, "custom" if steps else str(w), str(tf), str(bc), str(dt), str(len(x)), str(L), str(limit),
            str(tile), str(k)]
    if steps:
        args.append(steps)
    r = subprocess.run(args, input=x.tobytes(), capture_output=True)
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


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_tiled_all_wavelets(sim, oracle, w):
    g = np.random.default_rng(w)
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (3001, 2048)):
        L = min(7, max(1, oracle.max_level(w, n)))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        got, info = run_sim(sim, w, tf, bc, x, L, limit=200, tile=512, k=0)
        assert "pass" in info, info
        np.testing.assert_array_equal(got, ref, err_msg=f"w={w} bc={bc} tf={tf} n={n}\n{info}")


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_tiled_one_sample_end_tiles(sim, oracle, w):
    """Last tiles a single sample wide: a window touching the band end holds
    only the halo, so mirrored reads (Symmetric / Reflect) and adjoint folds
    must be covered by the halo rule, with plain LDS tiles and with register
    passes whose band-end pieces are cut from the tail."""
    g = np.random.default_rng(w + 3)
    for bc, tf, n in itertools.product((1, 3, 4), (0, 1, 2, 3), (2561, 4097)):
        L = min(4, oracle.max_level(w, n))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        for tile, k in ((512, 0), (64, 2)):
            got, info = run_sim(sim, w, tf, bc, x, L, limit=200, tile=tile, k=k)
            np.testing.assert_array_equal(got, ref, err_msg=f"w={w} bc={bc} tf={tf} n={n} tile={tile}\n{info}")


@pytest.mark.parametrize("tile,k,limit", [(128, 1, 100), (256, 2, 60), (1024, 4, 300), (64, 0, 50)])
def test_tiled_shapes_cdf97_float32(sim, oracle, tile, k, limit):
    g = np.random.default_rng(tile + k)
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (1777, 4096, 999)):
        L = 6
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, 100907, tf, bc, level=[L])
        got, info = run_sim(sim, 100907, tf, bc, x, L, limit=limit, tile=tile, k=k)
        np.testing.assert_array_equal(got, ref, err_msg=f"bc={bc} tf={tf} n={n}\n{info}")


def _spec(steps):
    parts = []
    for s in oracle_steps(steps):
        tgt = 1 if s["target"] in ("d", 1) else 0
        taps = s.get("taps", [])
        parts.append(",".join(str(v) for v in [s["kind"], tgt, s["offset"], s.get("sign", 1), s["ntaps"]] +
                              [repr(float(t)) for t in taps]))
    return ";".join(parts)


@pytest.mark.parametrize("name", list(MOCK_WAVELETS))
def test_tiled_mock_wavelets(sim, oracle, name):
    """The reference's mock step programs (large positive / negative offsets
    exercise folds that land next to tile edges)."""
    st = oracle_steps(MOCK_WAVELETS[name])
    spec = _spec(MOCK_WAVELETS[name])
    g = np.random.default_rng(5)
    for bc, tf, n in itertools.product((0, 1, 3, 4), (0, 1, 2, 3), (1531, 1024)):
        x = g.standard_normal(n)
        ref = oracle.lwt(x, 0, tf, bc, level=[3], steps=st)
        got, info = run_sim(sim, 0, tf, bc, x, 3, limit=100, tile=256, k=0, steps=spec)
        np.testing.assert_array_equal(got, ref, err_msg=f"{name} bc={bc} tf={tf} n={n}\n{info}")


@pytest.mark.parametrize("name", list(MOCK_WAVELETS))
def test_whole_line_all_bcs_mock(sim, oracle, name):
    """Whole-line path (no tiling) with all five BCs, including Periodic folds
    at non-power-of-two lengths whose targets land mid-band."""
    st = oracle_steps(MOCK_WAVELETS[name])
    spec = _spec(MOCK_WAVELETS[name])
    g = np.random.default_rng(6)
    for bc, tf, n in itertools.product(range(5), (0, 1, 2, 3), (153, 510, 1000, 37)):
        x = g.standard_normal(n)
        ref = oracle.lwt(x, 0, tf, bc, level=[2], steps=st)
        got, info = run_sim(sim, 0, tf, bc, x, 2, limit=1 << 20, tile=0, k=0, steps=spec)
        np.testing.assert_array_equal(got, ref, err_msg=f"{name} bc={bc} tf={tf} n={n}\n{info}")


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_whole_line_all_bcs(sim, oracle, w):
    g = np.random.default_rng(w + 1)
    for bc, tf, n in itertools.product(range(5), (0, 1, 2, 3), (1000, 333)):
        L = min(4, max(1, oracle.max_level(w, n)))
        if not ok_bands(bc, n, L):
            continue
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        got, info = run_sim(sim, w, tf, bc, x, L, limit=1 << 20, tile=0, k=0)
        np.testing.assert_array_equal(got, ref, err_msg=f"w={w} bc={bc} tf={tf} n={n}")


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_register_tiles_all_wavelets(sim, oracle, w):
    """Passes of four levels whose interior runs as register tiles (64-lane
    model of a wavefront) and whose band ends run as LDS tiles."""
    g = np.random.default_rng(w + 7)
    for bc, tf, n in itertools.product((0, 3), (0, 1, 2, 3), (30011, 24576)):
        L = min(6, oracle.max_level(w, n))
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, w, tf, bc, level=[L])
        got, info = run_sim(sim, w, tf, bc, x, L, limit=2000, tile=8192, k=4)
        np.testing.assert_array_equal(got, ref, err_msg=f"w={w} bc={bc} tf={tf} n={n}\n{info}")


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_register_tiles_zero_boundary_whole_band(sim, oracle, w):
    """Zero boundary: register windows overlap the band ends and run masked,
    so register tiles cover whole bands, down to bands of a few hundred
    samples (one window holds the band) -- no LDS pieces at the ends."""
    g = np.random.default_rng(w + 13)
    for tf, n in itertools.product((0, 1, 2, 3), (5003, 700, 1024, 333)):
        L = min(9, oracle.max_level(w, n))
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, w, tf, 0, level=[L])
        got, info = run_sim(sim, w, tf, 0, x, L, limit=64, tile=8192, k=4)
        np.testing.assert_array_equal(got, ref, err_msg=f"w={w} tf={tf} n={n}\n{info}")
        if n >= 1024:
            assert "reg [0," in info, info


@pytest.mark.parametrize("w", [w for w in WAVELETS if w != 0])
def test_register_tiles_fewer_levels(sim, oracle, w):
    """Register passes of one to three levels (level counts that are not a
    multiple of four; the last pass of a 5-level transform is one level)."""
    g = np.random.default_rng(w + 11)
    n = 20011
    for L, tf in itertools.product((1, 2, 3, 5), (0, 1, 2, 3)):
        if L > oracle.max_level(w, n):
            continue
        x = g.standard_normal(n).astype(np.float32)
        ref = oracle.lwt(x, w, tf, 3, level=[L])
        got, info = run_sim(sim, w, tf, 3, x, L, limit=600, tile=8192, k=4)
        assert "reg [" in info, info
        np.testing.assert_array_equal(got, ref, err_msg=f"w={w} L={L} tf={tf}\n{info}")
