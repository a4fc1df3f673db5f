"""Build recipe for the CPU oracle (liboracle.so).  Test infrastructure only:
loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_oracle():
    out = os.path.join(HERE, "liboracle.so")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c11",
           os.path.join(HERE, "oracle.c"), "-o", tmp, "-lm"]
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build_oracle()
