"""C ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, resolves coefficients exactly like the oracle, and rejects
invalid descriptors before touching the device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import WAVELETS, WIDTHS

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols(header=None):
    """Functions declared in include/*.h (or in one header)."""
    syms = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if not fn.endswith(".h") or (header is not None and fn != header):
            continue
        text = open(os.path.join(REPO, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lfx_\w+)\s*\(", text):
            syms.add(m.group(1))
    return syms


def _check_exports(path, syms, L):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = sorted(s for s in syms if s not in exported)
    assert not missing, missing
    for s in syms:
        getattr(L, s)


def test_library_exports_every_declared_symbol():
    """liblifted_hip.so exports every function of include/lifted_hip.h, and
    liblifted_cpu.so every function of include/lifted_cpu.h."""
    from lifted_amd import _lib
    syms = _declared_symbols("lifted_hip.h")
    assert len(syms) >= 10
    _check_exports(_lib.LIB_PATH, syms, _lib.lib())
    cpu = _declared_symbols("lifted_cpu.h")
    assert len(cpu) >= 3
    _check_exports(_lib.CPU_LIB_PATH, cpu, _lib.cpu_lib())
    assert _declared_symbols() == syms | cpu


def test_max_level_matches_oracle(oracle):
    import lifted_amd as la
    for w in WAVELETS:
        for n in (1, 2, 3, 7, 16, 31, 64, 100, 511, 512, 4096, 65536, 1 << 20):
            if WIDTHS[w] == 1:
                with pytest.raises(la.LfxError):
                    la.max_level(w, n)
                continue
            assert la.max_level(w, n) == oracle.max_level(w, n), (w, n)
    assert la.max_level(2, 65536) == 14
    assert la.max_level(100907, 4096) == 9 and la.max_level(100907, 512) == 6 and la.max_level(100907, 64) == 3


@pytest.mark.parametrize("tf", [0, 1, 2, 3])
def test_resolved_programs_match_oracle_coefficients(oracle, tf):
    """The GPU's per-transform program (order, band, offsets, signed taps)
    agrees with the oracle's step data for every wavelet, in both precisions."""
    import lifted_amd as la
    fwd_stepping = tf in (0, 3)
    normal = tf in (0, 1)
    fwd_tf = tf in (0, 2)
    for w in WAVELETS:
        for dt, st in ((np.float32, 0), (np.float64, 1)):
            rows = oracle.wavelet_dump(w, dt)
            prog = la.program_dump(w, tf, st)
            assert len(prog) == len(rows)
            order = list(range(len(rows))) if fwd_stepping else list(range(len(rows)))[::-1]
            for p, k in zip(prog, order):
                r = rows[k]
                kind = int(r[0])
                assert int(p[0]) == kind
                if kind == 3:
                    sm = r[5] if fwd_tf else r[6]
                    dm = r[8] if fwd_tf else r[7]
                    assert p[9] == sm and p[10] == dm
                    continue
                N, off = int(r[3]), int(r[2])
                assert int(p[1]) == (1 if (int(r[1]) == 1) == normal else 0)
                assert int(p[4]) == (off if normal else -(N - 1 + off))
                assert int(p[8]) == off
                if kind == 0:
                    taps = r[5:5 + N] if normal else r[5:5 + N][::-1]
                    sign = 1.0 if fwd_tf else -1.0
                    assert list(p[9:9 + N]) == list(sign * taps)
                elif kind == 2:
                    assert p[9] == (r[5] if fwd_tf else -r[5])
                else:
                    assert int(p[3]) == (0 if (r[4] > 0) == fwd_tf else 1)


def test_invalid_descriptors_are_rejected_without_device():
    import lifted_amd as la
    from lifted_amd import _lib
    L = _lib.lib()
    base = dict(wavelet=2, transform=0, bc=0, shape=[8], stride_in=[1], stride_out=[1], axes=[0], levels=[1],
                separable=True, storage=0)

    def call(**kw):
        a = dict(base, **kw)
        d = _lib.make_desc(a["wavelet"], a["transform"], a["bc"], a["shape"], a["stride_in"], a["stride_out"],
                           a["axes"], a["levels"], a["separable"], a["storage"])
        return L.lfx_lwt(d, None, None, None)

    assert call(transform=4) == -1
    assert call(bc=5) == -1
    assert call(storage=7) == -1
    assert call(axes=[1]) == -1
    assert call(levels=[-1]) == -1
    assert call(wavelet=12345) == -1
    assert b"unknown wavelet" in L.lfx_last_error()


def test_python_api_argument_errors():
    import lifted_amd as la
    x = np.zeros((4, 4), np.float32)
    with pytest.raises(KeyError):
        la.lwt_forward(x, "db99")
    with pytest.raises(ValueError):
        la.lwt_forward(x, "db2", axes=2)
    with pytest.raises(TypeError):
        la.lwt_forward(x, "db2", axes=[0, 1], level=[1])
    with pytest.raises(TypeError):
        la.lwt_forward(x.astype(np.int32), "db2")
    with pytest.raises(ValueError):
        la.lwt_forward(x, "db2", out=np.zeros((4, 4), np.float64))
    with pytest.raises(ValueError):
        la.lwt_forward(x, "db2", out=np.zeros((4, 5), np.float32))
    with pytest.raises(KeyError):
        la.lwt_forward(x, "db2", mode="Bogus")


def test_custom_program_validation():
    import lifted_amd as la
    with pytest.raises(la.LfxError):
        la.CustomWavelet([{"kind": "generic", "target": "d", "offset": 0, "taps": [0.0] * 17}])
    w = la.CustomWavelet([{"kind": "generic", "target": "d", "offset": -1, "taps": [0.5, 0.5]}], width=3)
    p = la.program_dump(w, 1)
    assert p.shape[0] == 1 and p[0, 9] == -0.5 and p[0, 10] == -0.5


def test_cpp_dropin_header_compiles_and_throws():
    """include/lifted.hpp: a reference-style caller compiles against it and
    links the library; invalid arguments throw std::invalid_argument before
    any device work."""
    d = os.path.join(REPO, "tests", "cpp_shim")
    subprocess.run(["make", "-s", "-C", d], check=True)
    r = subprocess.run([os.path.join(d, "build", "shim_main")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors_thrown 3/3 max_level_db2_65536 14" in r.stdout


def test_tuning_knobs_round_trip_and_validation():
    """lfx_set_tuning_ex / lfx_get_tuning (no device needed): values round
    trip, out-of-range values are rejected, the context manager restores."""
    from lifted_amd import tune
    from lifted_amd._lib import LfxError
    base = tune.get_tuning()
    with tune.tuning(levels_per_pass=3, line_limit=1024, side_streams=0):
        assert tune.get_tuning() == {"tile": 0, "levels_per_pass": 3, "side_streams": 0, "line_limit": 1024,
                                     "across": -1, "strip": 0}
    for v in (0, 1):
        with tune.tuning(across=v):
            assert tune.get_tuning()["across"] == v
    for v in (1024, 2048):
        with tune.tuning(strip=v):
            assert tune.get_tuning()["strip"] == v
    assert tune.get_tuning() == base
    for bad in ({"tile": 8}, {"levels_per_pass": 13}, {"side_streams": 9}, {"line_limit": 5}, {"across": 2},
                {"across": -2}, {"strip": 512}, {"strip": 4096}):
        with pytest.raises(LfxError):
            tune.set_tuning(**bad)
    assert tune.get_tuning() == base
