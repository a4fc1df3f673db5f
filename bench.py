"""Benchmark: Block Vecchia FP64 log-likelihood evaluations per second on one
MI355X, at the reference's performance shape (README.md:40-46):

  n = 20,000 locations iid U[0,1]^2 (seed 0); Matern sigma^2 = 1.5,
  range = 0.1, nu = 0.5, no nugget; observations y = L z with L the dense
  FP64 Cholesky of the full covariance; k-means into 1,500 blocks; cs = 300
  nearest conditioning points per block.

A step is one full likelihood evaluation (covariance generation, per-block
factorization, log-determinant and quadratic form, reduction, the 8-byte
result back on the host) with the dataset resident in HBM.  Prints one JSON
line with the driver's contract fields plus `roofline` (the factorization
kernel's FP64 rate, HIP-event timed) and `cpu_baseline` (the CPU port of the
same chain, timed on this host's cores).

    python bench.py [--gpus 1] [--steps 200] [--warmup 50] [--no-cpu-baseline]

The defaults average over 200 evaluations after 50 warm-up: 20-evaluation
windows of separate processes on one box spread from 0.76 to 0.81 ms per
evaluation, 200-evaluation ones from 0.74 to 0.75 ms.

HBM traffic of the factorization launches (the `traffic` field, per
evaluation) comes from two rocprofv3 counter passes over this script with the
factorization configuration fixed by environment (the autotuner's choice),
each with a plain device copy of known size appended as the calibration
(--calib-copy-mib), summarised by --traffic-from into
profiles/factorization_n20000_cs300_traffic.json; the commands are in the
profile summary next to it.  It is reported only when the configuration in
use matches the one it was measured with.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Vecchia llh evals/s + p50/p99 latency at 1e-10 rel err; FP64 TFLOP/s vs peak"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(n=20000, bc=1500, cs=300, theta=(1.5, 0.1, 0.5), seed=0):
    import blockvecchia_amd as bv
    from blockvecchia_amd import datagen
    x, y = datagen.uniform_locations(n, seed)
    t0 = time.time()
    try:
        import torch
        dev = "cuda" if torch.cuda.is_available() else None
    except Exception:
        dev = None
    obs = datagen.sample_dense(x, y, theta[0], theta[1], theta[2], seed=seed, device=dev)
    log("observations (dense Cholesky sampling, %s): %.1f s" % (dev or "cpu", time.time() - t0))
    t0 = time.time()
    blocks = bv.partition(x, y, obs, bc, cs, seed=seed)
    log("partition (k-means + kNN): %.1f s, %d blocks" % (time.time() - t0, blocks.batch_count))
    return blocks


def traffic_from_profiles(config):
    """HBM bytes per evaluation of the factorization launches from the
    committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy-kernel calibration), or None when it was measured for another
    factorization configuration than the one in use."""
    p = os.path.join(ROOT, "profiles", "factorization_n20000_cs300_traffic.json")
    if not os.path.exists(p) or config is None:
        return None
    with open(p) as f:
        d = json.load(f)
    # one entry per measured configuration; the kernels (and so their
    # traffic) depend on the first four entries of a configuration only, the
    # fifth says whether they are replayed from a graph
    for e in d.get("entries", [d]):
        if list(e.get("config", []))[:4] == list(config)[:4]:
            return e.get("hbm_bytes_per_launch")
    return None


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "factorization_n20000_cs300_traffic.json")


def calibration_copy(mib, repeats=3):
    """A plain streaming copy of `mib` MiB on the device (an elementwise
    kernel, so the counters see it; a same-dtype copy_ may go to a DMA
    engine): the known-size reference for FETCH_SIZE / WRITE_SIZE.  These are
    the last three elementwise dispatches of the run."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(repeats):
        torch.mul(src, 1.0, out=dst)  # a streaming kernel: reads n doubles, writes n
    torch.cuda.synchronize()


FACTOR_PREFIXES = ("void bv_block_kernel<", "void bv_step_kernel<", "void bv_step2_kernel<",
                   "void bv_pair_kernel<", "void bv_res_kernel<",
                   "void bv_block_kernel_1w<", "bv_step_parts_kernel",
                   "void bv_gen_kernel<", "bv_cheb_build_kernel")


def summarize_traffic(fetch_csv, write_csv, mib, config):
    """HBM bytes per evaluation of the factorization launches (bv_block_kernel,
    or the bv_step_kernel panel steps with their partial sums, plus the
    covariance generation launch when one runs) from rocprofv3 FETCH_SIZE /
    WRITE_SIZE CSVs, per evaluation = summed over those dispatches / the
    number of bv_reduce_kernel dispatches; scaled by copy bytes / the copy's
    counter."""
    import csv

    def per_eval(path, counter):
        fac, copy, nred = {}, {}, set()
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name, did, v = r["Kernel_Name"], int(r["Dispatch_Id"]), float(r["Counter_Value"])
            if name.startswith(FACTOR_PREFIXES):
                fac[did] = fac.get(did, 0.0) + v
            elif name.startswith("bv_reduce_kernel"):
                nred.add(did)
            elif "copy" in name.lower() or "elementwise" in name.lower():
                copy[did] = copy.get(did, 0.0) + v
        ids = sorted(copy)[-3:]  # the calibration copies run last
        return sum(fac.values()) / len(nred), sum(copy[i] for i in ids) / len(ids), len(nred)

    f, fc, ne = per_eval(fetch_csv, "FETCH_SIZE")
    w, wc, _ = per_eval(write_csv, "WRITE_SIZE")
    copy_bytes = mib * (1 << 20)
    f_scale = copy_bytes / (fc * 1024.0)
    w_scale = copy_bytes / (wc * 1024.0)
    fetch = f * 1024.0 * f_scale
    write = w * 1024.0 * w_scale
    d = {
        "workload": "n=20000 bc=1500 cs=300 (1478 blocks), nu=1/2",
        "config": list(config),
        "evaluations": ne,
        "fetch_kib_per_eval": f, "write_kib_per_eval": w,
        "copy_bytes": copy_bytes,
        "copy_fetch_kib": fc, "copy_write_kib": wc,
        "fetch_scale": f_scale, "write_scale": w_scale,
        "hbm_fetch_bytes_per_launch": fetch, "hbm_write_bytes_per_launch": write,
        "hbm_bytes_per_launch": fetch + write,
    }
    entries = []
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as fh:
            old = json.load(fh)
        entries = [e for e in old.get("entries", [old])
                   if list(e.get("config", []))[:4] != list(config)[:4]]
    entries.append(d)
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump({"entries": entries}, fh, indent=1)
    print(json.dumps(d, indent=1))


def eng_pmax(blocks, cs):
    """Panel steps of the largest block: Mp / 32 with Mp = roundup(m + 1, 32),
    m = cs + bs (block 0: bs alone)."""
    bn = blocks.batch_num
    m = max(int(bn[0]), int(bn[1:].max()) + cs if len(bn) > 1 else 0)
    return ((m + 1 + 31) // 32 * 32) // 32


def _with_obs(blocks, obs):
    import copy
    b = copy.copy(blocks)
    b.obs = np.ascontiguousarray(np.asarray(obs), dtype=np.float64)
    return b


def cpu_info():
    """The host's CPU model, logical cores and the cores this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable}


def measure_uneven(steps=10, warmup=2):
    """Uneven batching: n = 200,000 points, 70% uniform and 30% from 20
    Gaussian clusters (sd 0.02), k-means into 10,000 blocks (1 to ~100
    points each), cs = 400; Matern sigma^2 = 1, range 0.05, nu in {3/2, 5/2};
    observations from the Vecchia sequential simulation with cs = 400 over
    the same partition (GPU, bv_simulate).  Per nu: evaluations/s, latency,
    the factorization's FP64 rate, and how many blocks are not numerically
    positive definite (no nugget, as the reference)."""
    import blockvecchia_amd as bv
    from blockvecchia_amd import datagen, flops

    n, bc, cs = 200000, 10000, 400
    x, y = datagen.clustered_locations(n, 0)
    t0 = time.time()
    blocks = bv.partition(x, y, np.zeros(n), bc, cs, seed=0, device=0)
    bn = blocks.batch_num
    log("uneven: partition (GPU k-means, host kNN) %.1f s, %d blocks, sizes %d..%d (block 0: %d)"
        % (time.time() - t0, blocks.batch_count, bn[1:].min(), bn[1:].max(), bn[0]))
    F = flops.eval_flops(bn, cs)
    out = []
    for nu in (1.5, 2.5):
        theta = [1.0, 0.05, nu]
        t0 = time.time()
        obs = datagen.sample_partition(blocks, *theta, seed=0)
        log("uneven nu=%g: observations (GPU Vecchia simulation, cs=400) %.1f s, %d levels, "
            "nugget %g" % (nu, time.time() - t0, obs.levels, obs.nugget))
        eng = bv.Engine(_with_obs(blocks, obs))
        for _ in range(warmup):
            eng.llh(theta)
        lat, kms = [], []
        t_start = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            v = eng.llh(theta)
            lat.append(time.perf_counter() - t1)
            kms.append(eng.last_kernel_ms())
        el = time.perf_counter() - t_start
        info = eng.block_terms()[2]
        kern = float(np.mean(kms))
        achieved = F / (kern * 1e-3) / 1e12
        out.append({
            "nu": nu, "evals_per_s": steps / el, "ms_per_eval": 1e3 * el / steps,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3)},
            "llh": v, "blocks_not_positive_definite": int(np.count_nonzero(info)),
            "simulation": {"levels": int(obs.levels), "nugget": float(obs.nugget)},
            "factor_config": list(eng.factor_config()),
            "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "kernel_ms": kern, "flops_per_eval": F},
        })
        log("uneven nu=%g: %.3f ms/eval (kernels %.3f ms), %.2f TFLOP/s (%.1f%% of peak), %d "
            "non-PD blocks" % (nu, out[-1]["ms_per_eval"], kern, achieved,
                               100 * achieved / FP64_PEAK_TFLOPS,
                               out[-1]["blocks_not_positive_definite"]))
        eng.close()
    return {"workload": "n=200000 (70% uniform, 30% in 20 clusters sd 0.02), k-means bc=10000, "
                        "cs=400, sigma2=1 range=0.05, Vecchia-simulated observations (cs=400)",
            "n": n, "bc": bc, "cs": cs, "batch_count": int(blocks.batch_count),
            "steps": steps, "warmup": warmup, "per_nu": out}


def measure_mle_eval(steps=20, warmup=3):
    """Single evaluations at the maximum-likelihood workload's shape: the
    MLE's n = 100,000 uniform points and Block Vecchia observations (truth
    (1.0, 0.08, 0.8), simulated with cs = 400), partitioned into bc = 5,000
    blocks with cs = 150, evaluated at the MLE's estimate (general nu) with
    the autotuned form: evaluations/s, latency, the factorization's FP64
    rate."""
    import blockvecchia_amd as bv
    from blockvecchia_amd import datagen, flops

    n, bc, cs = 100000, 5000, 150
    truth = [1.0, 0.08, 0.8]
    theta = [0.790758, 0.0691475, 0.800678]
    x, y = datagen.uniform_locations(n, 0)
    obs = np.asarray(datagen.sample_vecchia(x, y, *truth, cs=400, bc=bc, seed=0))
    blocks = bv.partition(x, y, obs, bc, cs, seed=0, device=0)
    F = flops.eval_flops(blocks.batch_num, cs)
    eng = bv.Engine(blocks)
    for _ in range(warmup):
        eng.llh(theta)
    lat, kms = [], []
    t_start = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        v = eng.llh(theta)
        lat.append(time.perf_counter() - t1)
        kms.append(eng.last_kernel_ms())
    el = time.perf_counter() - t_start
    kern = float(np.mean(kms))
    achieved = F / (kern * 1e-3) / 1e12
    r = {"workload": "MLE shape: n=100000 uniform, Vecchia-simulated observations (truth "
                     "(1.0, 0.08, 0.8), cs=400), bc=5000, cs=150, evaluated at the MLE "
                     "estimate (0.790758, 0.0691475, 0.800678)",
         "n": n, "bc": bc, "cs": cs, "batch_count": int(blocks.batch_count),
         "steps": steps, "warmup": warmup, "evals_per_s": steps / el,
         "ms_per_eval": 1e3 * el / steps,
         "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                        "p99": float(np.percentile(lat, 99) * 1e3)},
         "llh": v, "factor_config": list(eng.factor_config()),
         "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                      "kernel_ms": kern, "flops_per_eval": F}}
    eng.close()
    log("mle-eval: %.3f ms/eval (kernels %.3f ms), %.2f TFLOP/s (%.1f%% of peak), config %s" % (
        r["ms_per_eval"], kern, achieved, 100 * achieved / FP64_PEAK_TFLOPS, r["factor_config"]))
    return r


def run_mle_eval(args):
    """--workload mle-eval: single evaluations at the MLE's shape, one JSON
    line (value: evaluations/s)."""
    r = measure_mle_eval(args.steps, args.warmup)
    print(json.dumps({
        "metric": METRIC, "value": r["evals_per_s"], "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_eval"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": r["workload"], "n": r["n"], "bc": r["bc"], "cs": r["cs"],
                   "batch_count": r["batch_count"], "parallelism": "single-gpu"},
        "latency_ms": r["latency_ms"], "llh": r["llh"], "factor_config": r["factor_config"],
        "roofline": r["roofline"]}))


def run_uneven(args):
    """--workload uneven: the uneven-batching configuration alone, one JSON
    line (value: evaluations/s at nu = 3/2)."""
    r = measure_uneven(args.steps, args.warmup)
    out = r["per_nu"]
    print(json.dumps({
        "metric": METRIC, "value": out[0]["evals_per_s"], "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": out[0]["ms_per_eval"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uneven batching: " + r["workload"] + "; value at nu=3/2",
                   "n": r["n"], "bc": r["bc"], "cs": r["cs"], "batch_count": r["batch_count"],
                   "parallelism": "single-gpu"},
        "per_nu": out}))


def run_serving_c5(args):
    """Serving: 32 datasets (n from {5k, 20k, 50k, 200k}, cs from {30, 60,
    150, 300}, n/13 blocks, nu from {1/2, 3/2, 5/2}), 4,096 requests near each
    dataset's truth arriving by Poisson at 0.6 times the server's capacity;
    served against plain on the same arrival stream.  One JSON line."""
    from blockvecchia_amd import serving as srv
    r = srv.benchmark_c5(log=log)
    log("serving c5 (open loop): %.0f evals/s vs plain %.0f (x%.2f), p50 %.2f ms p99 %.2f ms, "
        "plain p50 %.2f ms p99 %.2f ms; closed loop %.0f vs %.0f evals/s (x%.2f); max rel "
        "diff %.2e" % (
            r["evals_per_s"], r["plain_evals_per_s"], r["speedup_vs_plain"],
            r["latency_ms"]["p50"], r["latency_ms"]["p99"], r["plain_latency_ms"]["p50"],
            r["plain_latency_ms"]["p99"], r["closed_loop"]["evals_per_s"],
            r["closed_loop"]["plain_evals_per_s"], r["closed_loop"]["speedup_vs_plain"],
            r["max_rel_diff_vs_alone"]))
    print(json.dumps({
        "metric": METRIC, "value": r["evals_per_s"], "unit": "evals/s", "n_gpus": 1,
        "steps": r["requests"], "warmup": 512, "ms_per_step": 1e3 / r["evals_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (model draws: dense for n <= 20k, Block Vecchia above)",
        "config": {"workload": "serving: 32 datasets, n in {5k,20k,50k,200k}, cs in "
                               "{30,60,150,300}, bc=n/13, nu in {1/2,3/2,5/2}; 4096 Poisson "
                               "requests at 0.6x server capacity",
                   "parallelism": "single-gpu"},
        "latency_ms": r["latency_ms"], "serving": r}), flush=True)


def measure_mle():
    """Full maximum-likelihood run with general smoothness: n = 100,000
    uniform, truth (1.0, 0.08, 0.8), observations from the Vecchia
    sequential simulation with cs = 400 (GPU, its own partition into 5,000
    blocks); the driver (blockvecchia_amd/bin/test_dvecchia_batch) partitions
    into bc = 5,000 blocks with cs = 150 and maximises from 0.1:0.1:0.1 with
    xtol 1e-4.  Wall-clock to convergence, evaluations, estimates and the
    final log-likelihood."""
    import subprocess
    import tempfile

    from blockvecchia_amd import datagen

    n, bc, cs = 100000, 5000, 150
    theta = [1.0, 0.08, 0.8]
    x, y = datagen.uniform_locations(n, 0)
    t0 = time.time()
    obs = np.asarray(datagen.sample_vecchia(x, y, *theta, cs=400, bc=bc, seed=0))
    log("mle: observations (GPU Vecchia simulation, cs=400): %.1f s" % (time.time() - t0))
    driver = os.path.join(ROOT, "blockvecchia_amd", "bin", "test_dvecchia_batch")
    with tempfile.TemporaryDirectory() as tmp:
        datagen.write_xy_csv(os.path.join(tmp, "xy.csv"), x, y)
        datagen.write_obs_csv(os.path.join(tmp, "obs.csv"), obs)
        cmd = [driver, "--xy_path", os.path.join(tmp, "xy.csv"), "--obs_path",
               os.path.join(tmp, "obs.csv"), "--num_loc", str(n), "--vecchia_bc", str(bc),
               "--vecchia_cs", str(cs), "--kernel_init", "0.1:0.1:0.1", "--ikernel", "?:?:?",
               "--tol", "4", "--omp_threads", "16", "--seed", "0"]
        log("mle driver: " + " ".join(cmd))
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit("driver failed (%d): %s" % (r.returncode, r.stderr[-2000:]))
        logs = os.listdir(os.path.join(tmp, "log"))
        vals = [v.strip() for v in
                open(os.path.join(tmp, "log", logs[0])).read().strip().splitlines()[1].split(",")]
    iters, est, llh, opt_s = int(vals[0]), [float(v) for v in vals[1:4]], float(vals[4]), \
        float(vals[5])
    log("mle: %d evaluations, %.2f s to convergence, estimate %s, llh %.10g" % (
        iters, opt_s, est, llh))
    return {"workload": "MLE: n=100000 uniform, truth (1.0, 0.08, 0.8), Vecchia-simulated "
                        "observations (cs=400), bc=5000, cs=150, init 0.1:0.1:0.1, xtol 1e-4",
            "n": n, "bc": bc, "cs": cs, "optimization_s": opt_s, "evaluations": iters,
            "ms_per_eval": 1e3 * opt_s / iters, "estimate": est, "truth": theta, "llh": llh,
            "driver_wall_s": wall}


def run_mle(args):
    r = measure_mle()
    print(json.dumps({
        "metric": "MLE wall-clock to convergence (driver, general nu)", "value": r["optimization_s"],
        "unit": "s", "n_gpus": 1, "steps": r["evaluations"], "warmup": 0,
        "ms_per_step": r["ms_per_eval"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": r["workload"], "n": r["n"], "bc": r["bc"], "cs": r["cs"],
                   "parallelism": "single-gpu"},
        "evaluations": r["evaluations"], "estimate": r["estimate"], "truth": r["truth"],
        "llh": r["llh"], "driver_wall_s": r["driver_wall_s"]}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving scenarios (many sessions on resident datasets)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other configurations reported under extra keys of the "
                         "default run (uneven batching, the general-nu MLE, many-dataset "
                         "serving)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="append a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two rocprofv3 counter CSVs into the traffic JSON and exit")
    ap.add_argument("--workload", choices=["perf", "uneven", "mle", "mle-eval", "serving"],
                    default="perf",
                    help="perf: the n=20000 cs=300 evaluation (default); uneven: n=200000 "
                         "clustered, cs=400, nu 3/2 and 5/2; mle: the driver's full "
                         "maximum-likelihood run at n=100000, general nu; mle-eval: single "
                         "evaluations at that run's shape and estimate")
    ap.add_argument("--nu", type=float, default=0.5,
                    help="smoothness evaluated on the perf workload (its data stay the "
                         "nu=1/2 draw): 0.5, 1.5, 2.5 closed forms, else the device K_nu")
    ap.add_argument("--config", default=None,
                    help="with --traffic-from: the factorization configuration the counters "
                         "were taken with, as variant,stepped,chunks,streams")
    args = ap.parse_args()
    if args.traffic_from:
        cfg = [int(v) for v in (args.config or "0,0,1,1").split(",")]
        summarize_traffic(args.traffic_from[0], args.traffic_from[1], args.calib_copy_mib or 512,
                          cfg)
        return
    if args.gpus != 1:
        print("bench.py: one GPU only (the likelihood would shard by blocks across GPUs, "
              "cost-balanced contiguous ranges, with a single 8-byte sum at the end); "
              "--gpus %d is not run" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven":
        return run_uneven(args)
    if args.workload == "serving":
        return run_serving_c5(args)
    if args.workload == "mle":
        return run_mle(args)
    if args.workload == "mle-eval":
        return run_mle_eval(args)

    import blockvecchia_amd as bv
    from blockvecchia_amd import flops

    n, bc, cs = 20000, 1500, 300
    blocks = make_workload(n, bc, cs, [1.5, 0.1, 0.5])
    # the data are the nu = 1/2 draw; --nu changes only the smoothness evaluated
    theta = [1.5, 0.1, args.nu]
    closed = args.nu in (0.5, 1.5, 2.5)
    nu_desc = ("nu=%g closed form" % args.nu) if closed else \
        ("nu=%g: x^nu K_nu(x) from a per-evaluation Chebyshev table (bv_cheb_build_kernel), "
         "entries generated inside the factorization" % args.nu)
    F = flops.eval_flops(blocks.batch_num, cs)
    eng = bv.Engine(blocks)
    for _ in range(args.warmup):
        eng.llh(theta)
    st, bad = eng.status()
    if st != bv.BV_OK:
        raise SystemExit("evaluation failed: status %d block %d" % (st, bad))
    cfg = eng.factor_config()
    if cfg is not None and cfg[1] == 0 and cfg[0] == 3:
        kname = ("bv_res_kernel (register-resident factor, one workgroup of four wavefronts "
                 "per block, %s%s)" % (nu_desc, ", HIP graph replay" if cfg[4] else ""))
    elif cfg is None or cfg[1] == 0:
        nw = {0: "four wavefronts", 1: "one wavefront", 2: "two wavefronts"}[cfg[0] if cfg else 0]
        kname = "bv_block_kernel (%s, %s per block)" % (nu_desc, nw)
    elif cfg[1] == 3:
        kname = "bv_pair_kernel (paired panels, lookahead) x %d launches (%s, %d chunk(s)%s)" % (
            (eng_pmax(blocks, cs) + 1) // 2, nu_desc, cfg[2], ", HIP graph replay" if cfg[4] else "")
    else:
        kname = "%s x %d panel-step launches (%s, %d chunk(s)%s)" % (
            "bv_step2_kernel (lookahead)" if cfg[1] == 2 else "bv_step_kernel",
            eng_pmax(blocks, cs), nu_desc, cfg[2], ", HIP graph replay" if cfg[4] else "")
    log("factorization configuration (autotuned): %s" % (cfg,))
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        torch = None
    lat = []
    kms = []
    llh = None
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        llh = eng.llh(theta)
        lat.append(time.perf_counter() - t0)
        kms.append(eng.last_kernel_ms())
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    if args.calib_copy_mib:
        calibration_copy(args.calib_copy_mib)
    evals_per_s = args.steps / elapsed
    lat_ms = np.array(lat) * 1e3
    kern_ms = float(np.mean(kms))
    achieved = F / (kern_ms * 1e-3) / 1e12
    log("llh %.17g  %.3f ms/eval  kernel %.3f ms  %.2f TFLOP/s" %
        (llh, 1e3 * elapsed / args.steps, kern_ms, achieved))

    cpu = None
    rel_err = None
    if not closed and not args.no_cpu_baseline:
        # the host chain covers the closed-form kinds (the oracle's K_nu is a
        # long-double quadrature: minutes per evaluation)
        log("cpu baseline skipped for general nu")
    if closed and not args.no_cpu_baseline:
        import oracle
        info = cpu_info()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["usable_cores"]
        threads = min(threads, info["usable_cores"])
        # FP64 agreement with the oracle's op-by-op restatement (one evaluation)
        ov = oracle.llh(blocks, theta, threads=threads)
        rel_err = abs(llh - ov) / abs(ov)
        if oracle.lapack_available():
            oracle.lapack_llh(blocks, theta, threads=threads)  # warm-up
            ts = []
            lv = None
            for _ in range(args.cpu_repeats):
                t0 = time.perf_counter()
                lv = oracle.lapack_llh(blocks, theta, threads=threads)
                ts.append(time.perf_counter() - t0)
            med = float(np.median(ts))
            cpu = {"value": 1.0 / med, "unit": "evals/s", "cores": threads, "kind": "port",
                   "library": oracle.lapack_library(), "cpu_model": info["cpu_model"],
                   "nproc": info["nproc"], "usable_cores": info["usable_cores"],
                   "rel_err_vs_gpu": abs(lv - llh) / abs(llh),
                   "sample": "the full n=20000 evaluation (all %d blocks), median of %d after "
                             "one warm-up: the reference's per-block chain (dpotrf, dtrsm, dgemm, "
                             "dnrm2; oracle/lapack_chain.c) through MKL, OpenMP over blocks "
                             "with one MKL thread each" % (blocks.batch_count, args.cpu_repeats)}
            log("cpu baseline (MKL) %.3f s/eval on %d of %d cores (%s); gpu vs oracle rel err "
                "%.3e" % (med, threads, info["nproc"], info["cpu_model"], rel_err))
        else:
            log("cpu baseline: MKL (libmkl_rt.so) not found; not timed")

    serving = None
    if not args.no_serving:
        from blockvecchia_amd import serving as srv
        t0 = time.perf_counter()
        serving = srv.scenarios()
        for name, r in serving.items():
            log("serving %s: %.0f evals/s vs plain %.0f (x%.2f), p50 %.2f ms p99 %.2f ms" %
                (name, r["evals_per_s"], r["plain_evals_per_s"], r["speedup_vs_plain"],
                 r["latency_ms"]["p50"], r["latency_ms"]["p99"]))
        log("serving scenarios: %.1f s" % (time.perf_counter() - t0))

    extra = {}
    if not args.no_extra:
        for key, fn in (("uneven", lambda: measure_uneven(10, 2)), ("mle", measure_mle),
                        ("mle_eval", lambda: measure_mle_eval(20, 3)),
                        ("serving_c5", lambda: __import__("blockvecchia_amd.serving",
                                                          fromlist=["x"]).benchmark_c5(log=log))):
            t0 = time.perf_counter()
            try:
                extra[key] = fn()
            except Exception as e:  # report and keep the main line
                extra[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("%s: %.1f s" % (key, time.perf_counter() - t0))

    traffic = traffic_from_profiles(cfg) if args.nu == 0.5 else None
    nlaunch = {0: 1, 1: eng_pmax(blocks, cs), 2: eng_pmax(blocks, cs),
               3: (eng_pmax(blocks, cs) + 1) // 2}[cfg[1] if cfg else 0]
    if cfg and cfg[1] == 0 and cfg[0] == 3:
        nlaunch = 2  # the register-resident kernel's two builds, on two streams
    if cfg and cfg[1] and cfg[2] > 1 and cfg[3] != 0:
        nlaunch *= cfg[2]
    limiter = None
    pipe = None
    try:
        pipe = bv.fp64_pipe_probe(0)
        log("fp64 pipe probe (cycles per repetition): %s" %
            ", ".join("%s %.1f" % kv for kv in pipe.items()))
    except Exception as e:  # report and keep the main line
        log("fp64 pipe probe failed: %s" % e)
    if traffic:
        hbm_rate = traffic / (kern_ms * 1e-3) / 1e9
        if cfg and cfg[1] == 0 and cfg[0] == 3:
            reading = ("the factor stays on chip (register-resident form): %.2f GB per evaluation "
                       "below L2 (%.0f%% of HBM peak), the matrix cores at %.0f%%; what bounds it is "
                       "its one wavefront per SIMD (the registers hold the factor): a wavefront's "
                       "FP64 vector arithmetic (covariance generation, diagonal sweeps) does not "
                       "overlap its own matrix-core products, only another wavefront's "
                       "(fp64_pipe_probe), and the 16-step diagonal sweep per tile column is a "
                       "serial chain (profiles/bv_res_kernel_n20000_cs300.md)"
                       % (traffic / 1e9, 100 * hbm_rate / HBM_PEAK_GBS,
                          100 * achieved / FP64_PEAK_TFLOPS))
        else:
            reading = ("algorithmically MFMA-bound (%.1f GFLOP on 2.3 MB of inputs), measured "
                       "memory-bound: the factor tiles make %.2f GB of HBM round trips per "
                       "evaluation at %.0f%% of HBM peak while the matrix cores run at %.0f%%; "
                       "waves wait on memory results for 42-49%% of their cycles "
                       "(profiles/factorization_counters_n20000_cs300.md)"
                       % (F / 1e9, traffic / 1e9, 100 * hbm_rate / HBM_PEAK_GBS,
                          100 * achieved / FP64_PEAK_TFLOPS))
        limiter = {"hbm_bytes_per_eval": traffic, "hbm_GBps": hbm_rate,
                   "hbm_frac": hbm_rate / HBM_PEAK_GBS, "reading": reading}
    out = {
        "metric": METRIC,
        "value": evals_per_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "n=20000 iid U[0,1]^2 seed 0, Matern sigma2=1.5 range=0.1 nu=0.5, "
                        "y from dense FP64 Cholesky, k-means bc=1500, cs=300"
                        + ("" if args.nu == 0.5 else "; evaluated at nu=%g" % args.nu),
            "n": n, "bc": bc, "cs": cs, "batch_count": int(blocks.batch_count),
            "parallelism": "single-gpu",
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99))},
        "llh": llh,
        "rel_err_vs_cpu": rel_err,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "measured_limiter": limiter,
            "fp64_pipe_probe": pipe,
            "kernel": kname,
            "config": cfg,
            "flops_per_eval": F,
            "launches_per_eval": nlaunch,
            "kernel_ms": kern_ms,
        },
        "cpu_baseline": cpu,
        "serving": serving,
    }
    out.update(extra)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
