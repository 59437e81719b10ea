"""Where each band of a transformed array lives (the coefficient layouts).

Separable path (a sequence of levels, one per axis): along each transformed
axis, a line of n samples transformed with L levels holds

    [ s_L | d_L | d_(L-1) | ... | d_1 ]

where level l splits the band s_(l-1) of length m (s_0 = the line) into
ns_l = ceil(m / 2) approximation and nd_l = floor(m / 2) detail samples, and
d_l occupies [ns_l, ns_(l-1)) (src/lifted-drivers-inl.hpp:117-141 of the
reference). A band of the ND array is one band choice per transformed axis.

Level-by-level path (one level count for all axes): level l transforms the
sub-box [0, n_(l-1)) of every transformed axis (n_0 = the shape), so each
level leaves 2^k - 1 detail boxes (k transformed axes) and the final
approximation is the box [0, n_L) (src/lifted-drivers-inl.hpp:532-540).

All functions return index tuples (slices) that select views of a numpy
array or torch tensor of the transformed shape; nothing is copied.
"""


def band_lengths(n, levels):
    """[(ns_l, nd_l) for l = 1..levels] of a line of n samples."""
    out = []
    m = int(n)
    for _ in range(int(levels)):
        ns, nd = m - m // 2, m // 2
        out.append((ns, nd))
        m = ns
    return out


def line_bands(n, levels):
    """Slices of one transformed line: {"s": s_L, ("d", l): d_l for l = 1..L}."""
    lens = band_lengths(n, levels)
    bands = {}
    hi = int(n)
    for l, (ns, _) in enumerate(lens, start=1):
        bands[("d", l)] = slice(ns, hi)
        hi = ns
    bands["s"] = slice(0, hi)
    return bands


def _norm_axes(axes, ndim):
    if axes is None:
        return list(range(ndim))
    if isinstance(axes, int):
        axes = [axes]
    out = []
    for a in axes:
        a = int(a)
        if a < -ndim or a >= ndim:
            raise ValueError(f"axis {a} out of range for {ndim} dimensions")
        out.append(a % ndim)
    return out


def separable_bands(shape, axes, levels):
    """Bands of a separable transform: {key: index tuple}, key = one band
    label per transformed axis ("s" or ("d", l)), in `axes` order."""
    import itertools
    axes = _norm_axes(axes, len(shape))
    if len(levels) != len(axes):
        raise ValueError("levels must be the same length as axes")
    per_axis = [list(line_bands(shape[a], L).items()) for a, L in zip(axes, levels)]
    out = {}
    for combo in itertools.product(*per_axis):
        idx = [slice(None)] * len(shape)
        for a, (_, sl) in zip(axes, combo):
            idx[a] = sl
        out[tuple(lbl for lbl, _ in combo)] = tuple(idx)
    return out


def level_by_level_bands(shape, axes, level):
    """Bands of a level-by-level transform: (approx, details) where approx is
    the index tuple of the final approximation box and details[l - 1] maps a
    key (one of "s" / "d" per transformed axis, not all "s") to the index
    tuple of that level-l detail box."""
    import itertools
    axes = _norm_axes(axes, len(shape))
    box = [int(s) for s in shape]
    details = []
    for _ in range(int(level)):
        ns = {a: box[a] - box[a] // 2 for a in axes}
        lvl = {}
        for combo in itertools.product("sd", repeat=len(axes)):
            if all(c == "s" for c in combo):
                continue
            idx = [slice(0, box[k]) for k in range(len(shape))]
            for a, c in zip(axes, combo):
                idx[a] = slice(0, ns[a]) if c == "s" else slice(ns[a], box[a])
            lvl[combo] = tuple(idx)
        details.append(lvl)
        for a in axes:
            box[a] = ns[a]
    approx = tuple(slice(0, box[k]) for k in range(len(shape)))
    return approx, details


def bands(shape, axes=None, level=1):
    """Dispatch like lwt_forward: a sequence `level` selects the separable
    layout (separable_bands), an int the level-by-level one
    (level_by_level_bands)."""
    if isinstance(level, int):
        return level_by_level_bands(shape, axes, level)
    return separable_bands(shape, axes, list(level))
