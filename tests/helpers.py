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
