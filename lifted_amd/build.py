"""Build lifted_amd/liblifted_hip.so for gfx950 with hipcc (in-tree).

    python -m lifted_amd.build            full rebuild (what build() does by default)
    python -m lifted_amd.build --incr     recompile only objects whose sources or
                                          headers changed (compiler dep files)
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

# (source, device code, object name, defines): the kernel families compile once
# per transform type (LFX_TF; the register windows also per samples per lane,
# LFX_PL, the two-axis level kernels per storage type, LFX_ST), so that their
# wavelet instantiations build in parallel
SOURCES = [
    ("lfx_kernels.hip", True, "lfx_kernels", []),
    ("lfx_noalloc.hip", True, "lfx_noalloc", []),
]
for _tf in range(4):
    SOURCES += [
        ("lfx_reg_kernels.hip", True, f"lfx_reg_tf{_tf}", [f"-DLFX_TF={_tf}", "-DLFX_ROWS=0", "-DLFX_PL=32"]),
        ("lfx_reg_kernels.hip", True, f"lfx_reg_tf{_tf}_pl16", [f"-DLFX_TF={_tf}", "-DLFX_ROWS=0", "-DLFX_PL=16"]),
        ("lfx_reg_kernels.hip", True, f"lfx_rows_tf{_tf}", [f"-DLFX_TF={_tf}", "-DLFX_ROWS=1"]),
        ("lfx_across_kernels.hip", True, f"lfx_across_tf{_tf}", [f"-DLFX_TF={_tf}"]),
    ]
    SOURCES += [
        ("lfx_lvl2d_kernels.hip", True, f"lfx_lvl2d_tf{_tf}_st{_st}", [f"-DLFX_TF={_tf}", f"-DLFX_ST={_st}"])
        for _st in (0, 2, 3)  # float32, bfloat16, float16 storage
    ]
SOURCES += [
    ("lfx_api.cpp", False, "lfx_api", []),
    ("lfx_table.cpp", False, "lfx_table", []),
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _up_to_date(obj, dep):
    """True when obj exists and is newer than every file its dep file lists."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    t = os.path.getmtime(obj)
    with open(dep) as fh:
        text = fh.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    if not files:
        return False
    for f in files:
        if not os.path.exists(f) or os.path.getmtime(f) > t:
            return False
    return True


def build(verbose=False, incremental=False):
    hipcc = _hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    objs, cmds = [], []
    for src, device, name, defines in SOURCES:
        base = os.path.join(OBJDIR, name)
        obj, dep = base + ".o", base + ".d"
        objs.append(obj)
        if incremental and _up_to_date(obj, dep):
            continue
        cmd = [hipcc, "-c"] + COMMON + defines
        cmd += DEVICE if device else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
        # dependency list for incremental builds
        cmd += ["-MMD", "-MF", dep]
        cmd += [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
    # compile the translation units in parallel
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    if cmds:
        with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
            for r in ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
                if r.returncode != 0:
                    sys.stderr.write(r.stdout + r.stderr)
                    raise subprocess.CalledProcessError(r.returncode, r.args)
    if incremental and not cmds and os.path.exists(OUT):
        return OUT
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


CPU_OUT = os.path.join(HERE, "liblifted_cpu.so")


def build_cpu(verbose=False):
    """Build lifted_amd/liblifted_cpu.so: the CPU product path (the kernels'
    line bodies compiled for the host with g++, explicit FMAs, no contraction)."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O3", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared",
           "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas", "-Wno-ignored-attributes",
           os.path.join(CSRC, "lfx_cpu.cpp"), os.path.join(CSRC, "lfx_table.cpp"), "-o", CPU_OUT + ".tmp",
           "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise subprocess.CalledProcessError(r.returncode, cmd)
    os.replace(CPU_OUT + ".tmp", CPU_OUT)
    return CPU_OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, incremental="--incr" in sys.argv)
    build_cpu(verbose="-v" in sys.argv)
    print(OUT)
    print(CPU_OUT)
