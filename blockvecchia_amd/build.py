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
# covariance kinds of the register-resident kernels (bv_matern.h, bv_device.h):
# nu = 1/2, 3/2, 5/2, general nu, A read from the workspace
RES_KINDS = (0, 1, 2, 3, 7)
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


def _run_parallel(cmds, workers=None):
    """Run independent compile commands at once (at most `workers`, default
    the usable CPUs up to 8); raise if any fails."""
    workers = workers or max(1, min(8, len(os.sched_getaffinity(0))))
    pending, running, failed = list(cmds), [], []
    while pending or running:
        while pending and len(running) < workers:
            cmd = pending.pop(0)
            print("+ " + " ".join(cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
        cmd, proc = running.pop(0)
        if proc.wait() != 0:
            failed.append(cmd)
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


def build_lib(trace=False):
    """libbv.so, or with trace=True libbv_trace.so: the same library whose
    factorization kernels record every wavefront's start and end
    (bv_trace_read in include/bv.h)."""
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    tag = "_trace" if trace else ""
    jobs = []
    for src in ["bv_kernels.hip", "bv_api.hip"]:
        obj = os.path.join(OBJDIR, src + tag + ".o")
        extra = ["-DBV_TRACE"] if trace and src == "bv_kernels.hip" else []
        jobs.append(([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj], obj))
    # the register-resident kernels: one object per covariance kind
    for kind in RES_KINDS:
        obj = os.path.join(OBJDIR, "bv_resident_k%d.o" % kind)
        jobs.append(([HIPCC] + HIP_FLAGS + ["-DBV_RES_KIND=%d" % kind, "-c",
                                            os.path.join(CSRC, "bv_resident.hip"), "-o", obj], obj))
    for src in ["bv_prep.cpp"]:
        obj = os.path.join(OBJDIR, src + ".o")
        jobs.append((["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], obj))
    _run_parallel([j[0] for j in jobs])
    objs = [j[1] for j in jobs]
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
