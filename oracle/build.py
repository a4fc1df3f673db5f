"""Build recipe for the CPU oracle (liboracle.so) and the LAPACK chain
(liblapackchain.so, the CPU baseline).  Test infrastructure only: loaded by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def _cc(src, out, extra=()):
    tmp = out + ".tmp"
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c11",
           os.path.join(HERE, src), "-o", tmp, "-lm"] + list(extra)
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_oracle():
    out = _cc("oracle.c", os.path.join(HERE, "liboracle.so"))
    # the LAPACK (MKL, opened at run time) chain timed as the CPU baseline
    _cc("lapack_chain.c", os.path.join(HERE, "liblapackchain.so"), ["-ldl"])
    return out


if __name__ == "__main__":
    build_oracle()
