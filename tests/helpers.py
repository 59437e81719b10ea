"""Shared helpers of the test suite (inputs, tolerances, reference-test mocks)."""
import numpy as np

WAVELETS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 104, 105, 106, 202, 203, 10103, 10105, 10202, 10204, 10206,
            10208, 10301, 10303, 10305, 10307, 10309, 10402, 10404, 10406, 10505, 10608, 100503, 100907]
WIDTHS = {0: 1, 1: 2, 2: 4, 3: 6, 4: 8, 5: 10, 6: 12, 7: 14, 8: 18, 9: 18, 10: 20, 104: 8, 105: 10, 106: 12,
          202: 12, 203: 18, 10103: 6, 10105: 10, 10202: 5, 10204: 9, 10206: 13, 10208: 17, 10301: 4,
          10303: 8, 10305: 12, 10307: 16, 10309: 20, 10402: 7, 10404: 11, 10406: 15, 10505: 14,
          10608: 21, 100503: 5, 100907: 9}
# orthogonal families: Daubechies, Symlets, Coiflets
ORTHOGONAL = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 104, 105, 106, 202, 203]
BCS = [0, 1, 2, 3, 4]
TRANSFORMS = [0, 1, 2, 3]


def fill_sin(n, x0=-10.0, xf=10.0, scale=1.0, dtype=np.float64):
    """The reference's test input (tests/test_helpers.h:22-28)."""
    dx = (xf - x0) / (n - 1)
    return (np.sin(x0 + dx * np.arange(n)) * scale).astype(dtype)


def rng(seed):
    return np.random.default_rng(seed)


def max_ulp_diff(a, b):
    """Max distance in units of the last place, scaled by max |b| (fp32/fp64)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    if scale == 0.0:
        return 0.0 if np.array_equal(a, b) else np.inf
    ulp = np.spacing(np.asarray(scale, dtype=b.dtype))
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))) / float(ulp))


# Every parity comparison of the session: (test id, what, bit-exact, worst
# per-band error in units of the band's own ulp). tests/conftest.py prints a
# summary of it at the end of the run (also under -q).
PARITY = []


def band_slices(shape, axes, level):
    """Index tuples of every coefficient band of a transform's output: the
    separable layout for a sequence `level`, the level-by-level sub-boxes for
    an int (lifted_amd.layout)."""
    from lifted_amd import layout
    ndim = len(shape)
    axes = list(range(ndim)) if axes is None else ([axes] if isinstance(axes, int) else list(axes))
    if isinstance(level, (int, np.integer)):
        approx, details = layout.level_by_level_bands(shape, axes, int(level))
        return [approx] + [idx for lvl in details for idx in lvl.values()]
    return list(layout.separable_bands(shape, axes, [int(v) for v in level]).values())


def band_ulp_error(got, ref, bands=None):
    """Worst error over the bands, each in units of the last place of that
    band's own max |ref| (so a small detail band is not hidden behind a large
    approximation band); bands=None: the whole array is one band."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    worst = 0.0
    for idx in (bands if bands is not None else [Ellipsis]):
        g = got[idx].astype(np.float64)
        r = ref[idx]
        if r.size == 0:
            continue
        scale = float(np.max(np.abs(r)))
        diff = float(np.max(np.abs(g - r.astype(np.float64))))
        if diff == 0.0:
            continue
        if scale == 0.0:
            return np.inf
        worst = max(worst, diff / float(np.spacing(np.asarray(scale, dtype=r.dtype))))
    return worst


def assert_parity(got, ref, what, bands=None, exact=True, tol_ulp=4.0):
    """GPU result against the oracle. exact=True (the GPU and the oracle run
    the same operations in the same order): bit for bit. exact=False: at most
    tol_ulp units of the last place of each band's own max |ref|. The measured
    worst per-band error is recorded for the session summary either way."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, f"{what}: shape {got.shape} != {ref.shape}"
    same = bool(np.array_equal(got, ref))
    err = 0.0 if same else band_ulp_error(got, ref, bands)
    import os
    PARITY.append((os.environ.get("PYTEST_CURRENT_TEST", "?").split(" ")[0], what, same, err))
    if exact:
        assert same, f"{what}: not bit-exact, worst {err:.3g} ulp of a band's max |ref|"
    else:
        assert err <= tol_ulp, f"{what}: {err:.3g} ulp of a band's max |ref| (bar {tol_ulp})"
    return err


# Mock wavelets of the reference's step tests (tests/test_helpers.h:102-360),
# as step programs: kind, target, offset, taps / sign / n / value.
def _g(t, off, *taps):
    return {"kind": "generic", "target": t, "offset": off, "taps": list(taps)}


def _u(t, off, sign):
    return {"kind": "unit", "target": t, "offset": off, "sign": sign}


def _r(t, n, off, v):
    return {"kind": "repeat", "target": t, "offset": off, "n": n, "value": v}


MOCK_WAVELETS = {
    "SUpdateZeroOff": [_g("s", 0, 1.2)],
    "SUpdateNegOff": [_g("s", -5, 1.2, 1.5, 1.6)],
    "SUpdatePosOff": [_g("s", 3, 1.2, 1.5, 1.6)],
    "SUpdate": [_g("s", -2, 1.2, 1.3, 1.5, -1.2, -3.2)],
    "DUpdateZeroOff": [_g("d", 0, 0.5)],
    "DUpdateNegOff": [_g("d", -5, 1.2, 1.5, 1.6)],
    "DUpdatePosOff": [_g("d", 3, 1.2, 1.5, 1.6)],
    "DUpdate": [_g("d", -2, 0.5, 1.203, 4.2, 33.213)],
    "SUnitUpdateZeroOff": [_u("s", 0, +1)],
    "SUnitUpdatePosOff": [_u("s", 3, -1)],
    "SUnitUpdateNegOff": [_u("s", -3, +1)],
    "DUnitUpdateZeroOff": [_u("d", 0, -1)],
    "DUnitUpdatePosOff": [_u("d", 3, +1)],
    "DUnitUpdateNegOff": [_u("d", -3, -1)],
    "SRepeatUpdateZeroOff": [_r("s", 1, 0, 0.5)],
    "SRepeatUpdateNegOff": [_r("s", 3, -4, 0.5)],
    "SRepeatUpdatePosOff": [_r("s", 3, 2, 0.5)],
    "SRepeatUpdate": [_r("s", 3, -1, 0.5)],
    "DRepeatUpdateZeroOff": [_r("d", 1, 0, 0.5)],
    "DRepeatUpdateNegOff": [_r("d", 3, -4, 0.5)],
    "DRepeatUpdatePosOff": [_r("d", 3, 2, 0.5)],
    "DRepeatUpdate": [_r("d", 3, -1, 0.5)],
    "Scale": [{"kind": "scale", "values": [0.124]}],
    "MultiScale": [{"kind": "scale", "values": [5.23, 0.124]}],
    "MultiStep": [_g("s", -5, 1.2, 1.3, 1.5, -1.2, -3.2), _g("d", -5, 0.5, 1.203, 4.2, 33.213),
                  {"kind": "scale", "values": [0.124]}],
}


def oracle_steps(steps):
    """Mock step dicts -> the oracle's step records."""
    out = []
    for s in steps:
        if s["kind"] == "generic":
            out.append({"kind": 0, "target": s["target"], "offset": s["offset"], "ntaps": len(s["taps"]),
                        "taps": s["taps"]})
        elif s["kind"] == "unit":
            out.append({"kind": 1, "target": s["target"], "offset": s["offset"], "ntaps": 1, "sign": s["sign"]})
        elif s["kind"] == "repeat":
            out.append({"kind": 2, "target": s["target"], "offset": s["offset"], "ntaps": s["n"],
                        "taps": [s["value"]]})
        else:
            v = list(s["values"])
            if len(v) == 1:
                v = v * 2
            out.append({"kind": 3, "target": "s", "offset": 0, "ntaps": 0, "taps": v})
    return out


def c_strides(shape):
    st = [1] * len(shape)
    for i in range(len(shape) - 2, -1, -1):
        st[i] = st[i + 1] * shape[i + 1]
    return st


def round_to(x, storage):
    """float32 values rounded to a 16-bit storage type ("bfloat16" /
    "float16", round to nearest even) and widened back to float32."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(getattr(torch, storage)).float().numpy()


def oracle_level_by_level_rounded(oracle, x, wavelet, transform, bc, axes, level, storage):
    """What the level-by-level path computes for 16-bit storage with several
    axes: every level reads its sub-box [0, n_(l-1)) of the transformed axes
    from the storage type, runs one level along each axis in order in float32,
    and rounds the sub-box once to the storage type (Forward / InverseAdjoint
    walk the levels fine -> coarse, Inverse / ForwardAdjoint coarse -> fine,
    src/lifted-drivers-inl.hpp:488-610). x holds values of the storage type
    (as float32); returns float32 values of the storage type."""
    x = np.array(x, dtype=np.float32, copy=True)
    boxes, box = [], list(x.shape)
    for _ in range(int(level)):
        boxes.append(tuple(box))
        for a in axes:
            box[a] = box[a] - box[a] // 2
    order = range(level) if transform in (0, 3) else reversed(range(level))
    for l in order:
        idx = tuple(slice(0, b) for b in boxes[l])
        sub = np.ascontiguousarray(x[idx])
        x[idx] = round_to(oracle.lwt(sub, wavelet, transform, bc, axes=list(axes), level=1), storage)
    return x


def field_1f2(shape, seed):
    """Seeded 1/f^2 random fields over the last two dims (white noise, Fourier
    amplitudes times 1/|k|^2 with k = 0 zeroed), each scaled to zero mean and
    unit variance; float32."""
    g = np.random.default_rng(seed)
    ny, nx = shape[-2], shape[-1]
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k2 = ky * ky + kx * kx
    filt = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    out = np.empty(shape, np.float32)
    for idx in np.ndindex(*shape[:-2]):
        f = np.fft.irfft2(np.fft.rfft2(g.standard_normal((ny, nx))) * filt, s=(ny, nx))
        out[idx] = ((f - f.mean()) / f.std()).astype(np.float32)
    return out
