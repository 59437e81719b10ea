"""Build lifted_amd/liblifted_hip.so for gfx950 with hipcc (in-tree).

    python -m lifted_amd.build        (or lifted_amd.build.build())
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liblifted_hip.so")
OBJDIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("LFX_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter",
]
# device code: explicit fma builtins only (contraction off), no packed fp32 ops
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("lfx_kernels.hip", True),
    ("lfx_reg_tf0.hip", True),
    ("lfx_reg_tf1.hip", True),
    ("lfx_reg_tf2.hip", True),
    ("lfx_reg_tf3.hip", True),
    ("lfx_api.cpp", False),
    ("lfx_table.cpp", False),
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    hipcc = _hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    objs, cmds = [], []
    for src, device in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, "-c"] + COMMON
        cmd += DEVICE if device else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
        cmd += [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
        objs.append(obj)
    # compile the translation units in parallel
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
