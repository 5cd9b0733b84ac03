"""Build libdgamd.so (HIP, gfx950) in-tree with hipcc.

The library is compiled every time build() is called; sources are compiled
in parallel into dgcuda_amd/_build/ and linked into dgcuda_amd/libdgamd.so;
the command-line driver dgcuda_amd/dgrun links against it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgamd.so")
CLI = os.path.join(HERE, "dgrun")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["dg_kernels.hip", "dg_adjoint.hip", "dg_ensemble.hip", "dg_api.cpp"]

# Packed FP32 instructions stay disabled (-packed-fp32-ops) for every HIP file.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, extra):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + HIPFLAGS + extra + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def gen_tables():
    """regenerate dg_tabvals.h (compile-time reference-element operators) with
    the host generator; the output is deterministic"""
    exe = os.path.join(OUT, "gen_tabvals")
    cxx = os.environ.get("CXX", "g++")
    r = subprocess.run([cxx, "-O2", "-o", exe, os.path.join(CSRC, "gen_tabvals.cpp")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("table generator failed to compile")
    tmp = os.path.join(OUT, "dg_tabvals.h")
    subprocess.run([exe, tmp], check=True)
    dst = os.path.join(CSRC, "dg_tabvals.h")
    new = open(tmp).read()
    if not os.path.exists(dst) or open(dst).read() != new:
        with open(dst, "w") as f:
            f.write(new)


def build(verbose=False, extra=None):
    extra = list(extra or [])
    os.makedirs(OUT, exist_ok=True)
    gen_tables()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    with ThreadPoolExecutor(max_workers=min(len(srcs), 4)) as ex:
        for src, obj, r in ex.map(lambda s: _compile(s, extra), srcs):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError("hipcc failed on %s" % src)
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    build_cli()
    return LIB


def build_cli():
    """dgrun, the reference's command line over the C ABI (host code only)"""
    cxx = os.environ.get("CXX", "g++")
    tmp = CLI + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", tmp, os.path.join(CSRC, "dg_cli.cpp"),
           "-L" + HERE, "-ldgamd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("dgrun failed to build")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
