"""Build recipe for libbv.so (gfx950 kernels + host engine + host
pre-processing) and the test_dvecchia_batch driver, with hipcc and g++
directly.  Usage: python -m blockvecchia_amd.build  (or build_all());
python -m blockvecchia_amd.build --trace builds lib/libbv_trace.so only."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "blockvecchia_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", "-ffp-contract=off",
             "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(trace=False):
    """libbv.so, or with trace=True libbv_trace.so: the same library whose
    factorization kernels record every wavefront's start and end
    (bv_trace_read in include/bv.h)."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    tag = "_trace" if trace else ""
    objs = []
    for src in ["bv_kernels.hip", "bv_api.hip"]:
        obj = os.path.join(OBJDIR, src + tag + ".o")
        extra = ["-DBV_TRACE"] if trace and src == "bv_kernels.hip" else []
        _run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ["bv_prep.cpp"]:
        obj = os.path.join(OBJDIR, src + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, "libbv%s.so" % tag)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, out)
    return out


def build_driver():
    os.makedirs(BINDIR, exist_ok=True)
    src = os.path.join(CSRC, "test_dvecchia_batch.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(BINDIR, "test_dvecchia_batch")
    _run(["g++"] + CXX_FLAGS + [src, "-o", out, "-L" + LIBDIR, "-lbv",
                                "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    return out


def build_all():
    lib = build_lib()
    drv = build_driver()
    return lib, drv


if __name__ == "__main__":
    if "--trace" in sys.argv[1:]:
        build_lib(trace=True)
    else:
        build_all()
    sys.exit(0)
