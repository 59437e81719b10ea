"""Coefficient-layout helpers (lifted_amd.layout) against the oracle: every
element belongs to exactly one band, and each band holds what composing
single-level transforms puts there (src/lifted-drivers-inl.hpp:117-141 for the
separable layout, :532-540 for the level-by-level sub-boxes)."""
import itertools

import numpy as np
import pytest

from helpers import fill_sin
from lifted_amd import layout

F = 0


@pytest.mark.parametrize("shape,axes,levels", [((37,), [0], [4]), ((20, 33), [1, 0], [2, 3]),
                                                ((9, 12, 17), [2, 0], [3, 1]), ((16, 16), [0, 1], [4, 4])])
def test_separable_bands_partition_and_content(oracle, shape, axes, levels):
    x = fill_sin(int(np.prod(shape)), -4.0, 7.0).reshape(shape)
    out = oracle.lwt(x, 100907, F, oracle.SYMMETRIC, axes=axes, level=levels)
    bands = layout.separable_bands(shape, axes, levels)
    cover = np.zeros(shape, int)
    for idx in bands.values():
        cover[idx] += 1
    assert np.all(cover == 1)
    assert len(bands) == np.prod([L + 1 for L in levels])
    # along each axis, d_l of the multi-level line == d of one level applied to s_(l-1)
    if len(axes) == 1:
        a, L = axes[0], levels[0]
        s = x
        for l in range(1, L + 1):
            one = oracle.lwt(s, 100907, F, oracle.SYMMETRIC, axes=[a], level=[1])
            b1 = layout.line_bands(s.shape[a], 1)
            idx_d = [slice(None)] * len(shape)
            idx_d[a] = b1[("d", 1)]
            np.testing.assert_allclose(out[bands[(("d", l),)]], one[tuple(idx_d)], rtol=0, atol=1e-12)
            idx_s = [slice(None)] * len(shape)
            idx_s[a] = b1["s"]
            s = np.ascontiguousarray(one[tuple(idx_s)])
        np.testing.assert_allclose(out[bands[("s",)]], s, rtol=0, atol=1e-12)


@pytest.mark.parametrize("shape,axes,level", [((40, 36), [0, 1], 2), ((25, 30, 7), [1, 2], 3), ((64,), [0], 5)])
def test_level_by_level_bands_partition_and_content(oracle, shape, axes, level):
    x = fill_sin(int(np.prod(shape)), -3.0, 9.0).reshape(shape)
    out = oracle.lwt(x, 3, F, oracle.ZERO, axes=axes, level=level)
    approx, details = layout.level_by_level_bands(shape, axes, level)
    cover = np.zeros(shape, int)
    cover[approx] += 1
    for lvl in details:
        assert len(lvl) == 2 ** len(axes) - 1
        for idx in lvl.values():
            cover[idx] += 1
    assert np.all(cover == 1)
    # level l's detail boxes == one level on the level-(l-1) approximation box
    box = x
    for l, lvl in enumerate(details, start=1):
        one = oracle.lwt(np.ascontiguousarray(box), 3, F, oracle.ZERO, axes=axes, level=1)
        a1, d1 = layout.level_by_level_bands(box.shape, axes, 1)
        for key, idx in lvl.items():
            np.testing.assert_allclose(out[idx], one[d1[0][key]], rtol=0, atol=1e-12)
        box = one[a1]
    np.testing.assert_allclose(out[approx], box, rtol=0, atol=1e-12)


def test_line_bands_lengths():
    assert layout.band_lengths(31, 3) == [(16, 15), (8, 8), (4, 4)]
    b = layout.line_bands(31, 3)
    assert b["s"] == slice(0, 4) and b[("d", 3)] == slice(4, 8) and b[("d", 2)] == slice(8, 16)
    assert b[("d", 1)] == slice(16, 31)
    with pytest.raises(ValueError):
        layout.separable_bands((4, 4), [0, 1], [1])
    with pytest.raises(ValueError):
        layout.separable_bands((4, 4), [2], [1])
