"""Serving benchmark: many optimizer sessions on several resident datasets.

Each session is bound to one dataset and submits its next parameter vector
as soon as its previous result returns (closed loop, like an optimizer
waiting for its objective).  The server merges whatever is queued into shared
launches on several lanes.  The plain form evaluates the same request
sequence one evaluation at a time with bv_llh.  Reported: evaluations/s of
both, the served p50/p99 latency from submit to result, the plain per-call
p50/p99, and the largest relative difference between a served result and the
same evaluation alone.

The open-loop scenario (benchmark_poisson) instead sends requests at Poisson
arrival times at a fixed offered rate, to the server and, in arrival order,
to the plain form; latency counts from arrival, so it includes queueing.
"""
import time

import numpy as np

import blockvecchia_amd as bv
from blockvecchia_amd import datagen

# (n, cs, smoothness nu): block count n/13, as the reference's perf shapes.
# Many small datasets: single evaluations leave most of the GPU idle.
SMALL_DATASETS = [
    (1000, 30, 0.5), (1000, 60, 1.5), (2000, 30, 0.5), (2000, 60, 2.5),
    (3000, 60, 0.5), (3000, 30, 1.5), (5000, 60, 0.5), (5000, 30, 0.5),
]
# Mixed sizes up to the n = 20k / cs = 300 shape: the large evaluations
# already fill the GPU on their own.
DEFAULT_DATASETS = [
    (2000, 30, 0.5), (2000, 60, 1.5), (5000, 60, 0.5), (5000, 150, 2.5),
    (10000, 150, 0.5), (10000, 60, 1.5), (20000, 300, 0.5), (20000, 150, 0.5),
]


def build_datasets(specs=DEFAULT_DATASETS, seed=0):
    """Observations are drawn from the model at the true theta (dense FP64
    Cholesky, on the GPU when torch sees one): white noise under a smooth
    Matern makes the quadratic form ill-conditioned, and then results that
    differ only in the order of FP64 operations differ far beyond rounding."""
    try:
        import torch
        device = "cuda" if torch.cuda.is_available() else None
    except ImportError:
        device = None
    engines, truths = [], []
    for i, (n, cs, nu) in enumerate(specs):
        x, y = datagen.uniform_locations(n, seed + i)
        obs = datagen.sample_dense(x, y, 1.5, 0.1, nu, seed=seed + i, device=device)
        blocks = bv.partition(x, y, obs, max(1, n // 13), cs, seed=0)
        engines.append(bv.Engine(blocks))
        truths.append(np.array([1.5, 0.1, nu]))
    return engines, truths


def make_requests(truths, sessions, per_session, seed=1):
    """Per session: dataset index and its sequence of parameter vectors
    (the true theta with a few percent jitter on variance and range)."""
    rng = np.random.default_rng(seed)
    plan = []
    for s in range(sessions):
        d = s % len(truths)
        ths = []
        for _ in range(per_session):
            th = truths[d].copy()
            th[0] *= 1.0 + 0.05 * rng.standard_normal()
            th[1] *= 1.0 + 0.05 * rng.standard_normal()
            ths.append(th)
        plan.append((d, ths))
    return plan


def run_plain(engines, plan):
    """One evaluation at a time, sessions interleaved round-robin."""
    order = []
    per = len(plan[0][1])
    for k in range(per):
        for s, (d, ths) in enumerate(plan):
            order.append((s, k, d, ths[k]))
    lat = []
    out = {}
    t0 = time.perf_counter()
    for s, k, d, th in order:
        a = time.perf_counter()
        out[(s, k)] = engines[d].llh(th)
        lat.append(time.perf_counter() - a)
    wall = time.perf_counter() - t0
    return out, wall, np.array(lat)


def run_served(engines, plan, lanes=4, ws_bytes=2 << 30, max_evals=64):
    srv = bv.Server(lanes=lanes, ws_bytes_per_lane=ws_bytes, max_evals_per_batch=max_evals)
    ids = [srv.attach(e) for e in engines]
    nxt = [0] * len(plan)
    t_sub = {}
    out = {}
    lat = []

    def submit(s):
        d, ths = plan[s]
        k = nxt[s]
        tag = s * 1_000_000 + k
        t_sub[tag] = time.perf_counter()
        srv.submit(ids[d], ths[k], tag)
        nxt[s] += 1

    t0 = time.perf_counter()
    for s in range(len(plan)):
        submit(s)
    total = sum(len(ths) for _, ths in plan)
    done = 0
    while done < total:
        res = srv.poll(wait=True)
        now = time.perf_counter()
        for tag, v, st in res:
            s, k = divmod(tag, 1_000_000)
            out[(s, k)] = v
            lat.append(now - t_sub.pop(tag))
            done += 1
            if nxt[s] < len(plan[s][1]):
                submit(s)
    wall = time.perf_counter() - t0
    srv.close()
    return out, wall, np.array(lat)


def poisson_requests(truths, count, rate, seed=2):
    """count requests (dataset index, theta near its truth) with Poisson
    arrival times at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    reqs = []
    for i in range(count):
        d = int(rng.integers(len(truths)))
        th = truths[d].copy()
        th[0] *= 1.0 + 0.05 * rng.standard_normal()
        th[1] *= 1.0 + 0.05 * rng.standard_normal()
        reqs.append((float(t[i]), d, th))
    return reqs


def run_open_plain(engines, reqs):
    """Requests evaluated one at a time in arrival order, each as soon as it
    has arrived and the previous one is done; latency from arrival."""
    lat = np.empty(len(reqs))
    out = {}
    t0 = time.perf_counter()
    for i, (ta, d, th) in enumerate(reqs):
        while time.perf_counter() - t0 < ta:
            pass
        out[i] = engines[d].llh(th)
        lat[i] = time.perf_counter() - t0 - ta
    return out, time.perf_counter() - t0, lat


def run_open_served(engines, reqs, lanes=4, ws_bytes=2 << 30, max_evals=64):
    """Requests submitted to the server at their arrival times, results
    collected as lanes finish; latency from arrival."""
    srv = bv.Server(lanes=lanes, ws_bytes_per_lane=ws_bytes, max_evals_per_batch=max_evals)
    ids = [srv.attach(e) for e in engines]
    lat = np.empty(len(reqs))
    out = {}
    nxt = 0
    done = 0
    t0 = time.perf_counter()
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            srv.submit(ids[reqs[nxt][1]], reqs[nxt][2], nxt)
            nxt += 1
        res = srv.poll(wait=False)
        if not res and nxt == len(reqs):
            res = srv.poll(wait=True)
        now = time.perf_counter() - t0
        for tag, v, st in res:
            out[tag] = v
            lat[tag] = now - reqs[tag][0]
            done += 1
    wall = time.perf_counter() - t0
    srv.close()
    return out, wall, lat


def benchmark_poisson(count=4096, load=0.6, lanes=4, specs=SMALL_DATASETS, seed=2):
    """Open-loop arrivals: `count` requests over the datasets at a Poisson
    rate of `load` times the server's closed-loop capacity, served and
    plain on the same arrival stream.  Past its own capacity the plain form
    queues, so its latency grows with the run; the served form keeps up."""
    engines, truths = build_datasets(specs)
    plan = make_requests(truths, 64, 16)
    run_served(engines, plan, lanes)
    _, t_cap, _ = run_served(engines, plan, lanes)
    capacity = sum(len(ths) for _, ths in plan) / t_cap
    rate = load * capacity
    reqs = poisson_requests(truths, count, rate, seed)
    run_open_served(engines, reqs, lanes)                # warm-up: lane buffers at full size
    served, t_s, lat_s = run_open_served(engines, reqs, lanes)
    plain, t_p, lat_p = run_open_plain(engines, reqs)
    worst = max(abs(served[k] - plain[k]) / abs(plain[k]) for k in plain)
    for e in engines:
        e.close()
    return {
        "requests": count, "offered_rate_per_s": rate, "server_capacity_per_s": capacity,
        "evals_per_s": count / t_s, "plain_evals_per_s": count / t_p,
        "speedup_vs_plain": t_p / t_s,
        "latency_ms": {"p50": float(np.percentile(lat_s, 50) * 1e3),
                       "p99": float(np.percentile(lat_s, 99) * 1e3)},
        "plain_latency_ms": {"p50": float(np.percentile(lat_p, 50) * 1e3),
                             "p99": float(np.percentile(lat_p, 99) * 1e3)},
        "max_rel_diff_vs_alone": worst,
    }


def benchmark(sessions=32, per_session=16, lanes=4, ws_bytes=2 << 30, max_evals=64,
              specs=DEFAULT_DATASETS):
    engines, truths = build_datasets(specs)
    plan = make_requests(truths, sessions, per_session)
    # warm-up: both paths over the whole plan once (autotuning, lane buffers)
    run_plain(engines, plan)
    run_served(engines, plan, lanes, ws_bytes, max_evals)
    plain, t_plain, lat_plain = run_plain(engines, plan)
    served, t_served, lat_served = run_served(engines, plan, lanes, ws_bytes, max_evals)
    worst = max(abs(served[k] - plain[k]) / abs(plain[k]) for k in plain)
    nreq = len(plain)
    for e in engines:
        e.close()
    return {
        "requests": nreq,
        "sessions": sessions,
        "datasets": [{"n": n, "cs": cs, "nu": nu, "bc": max(1, n // 13)} for n, cs, nu in specs],
        "lanes": lanes,
        "evals_per_s": nreq / t_served,
        "plain_evals_per_s": nreq / t_plain,
        "speedup_vs_plain": t_plain / t_served,
        "latency_ms": {"p50": float(np.percentile(lat_served, 50) * 1e3),
                       "p99": float(np.percentile(lat_served, 99) * 1e3)},
        "plain_latency_ms": {"p50": float(np.percentile(lat_plain, 50) * 1e3),
                             "p99": float(np.percentile(lat_plain, 99) * 1e3)},
        "max_rel_diff_vs_alone": worst,
    }


def build_c5_datasets(count=32, seed=5, log=None):
    """The many-dataset serving configuration: `count` datasets with n drawn
    from {5k, 20k, 50k, 200k} uniform on [0,1]^2, cs from {30, 60, 150,
    300}, n/13 k-means blocks, Matern nu from {1/2, 3/2, 5/2}, sigma^2 = 1.5,
    range 0.1.  Observations are draws from the Block Vecchia model of the
    dataset's own partition (GPU sequential simulation, bv_simulate).  White
    noise would make the quadratic forms so ill-conditioned that two orders of
    the same FP64 operations disagree far beyond rounding."""
    rng = np.random.default_rng(seed)
    specs, engines, truths = [], [], []
    for i in range(count):
        n = int(rng.choice([5000, 20000, 50000, 200000]))
        cs = int(rng.choice([30, 60, 150, 300]))
        nu = float(rng.choice([0.5, 1.5, 2.5]))
        x, y = datagen.uniform_locations(n, seed + 100 + i)
        blocks = bv.partition(x, y, np.zeros(n), max(1, n // 13), cs, seed=0, device=0)
        blocks.obs = np.ascontiguousarray(
            datagen.sample_partition(blocks, 1.5, 0.1, nu, seed=seed + i), dtype=np.float64)
        engines.append(bv.Engine(blocks))
        truths.append(np.array([1.5, 0.1, nu]))
        specs.append((n, cs, nu))
        if log:
            log("c5 dataset %d: n=%d cs=%d nu=%g" % (i, n, cs, nu))
    return specs, engines, truths


def benchmark_c5(count=4096, load=0.6, lanes=4, seed=5, log=None):
    """Open-loop serving of the many-dataset configuration: 32 datasets
    (build_c5_datasets), `count` requests with a theta near each dataset's
    truth arriving by Poisson at `load` times the server's closed-loop
    capacity, served and, on the same arrival stream, plain (one evaluation
    at a time); evaluations/s and p50/p99 latency from arrival."""
    t0 = time.perf_counter()
    specs, engines, truths = build_c5_datasets(seed=seed, log=log)
    if log:
        log("c5 datasets: %.1f s" % (time.perf_counter() - t0))
    for e, th in zip(engines, truths):
        e.llh(th)                                         # autotune each dataset
    plan = make_requests(truths, 32, 8)
    run_served(engines, plan, lanes)
    _, t_cap, lat_cl = run_served(engines, plan, lanes)
    _, t_pcl, lat_pcl = run_plain(engines, plan)
    nplan = sum(len(ths) for _, ths in plan)
    capacity = nplan / t_cap
    rate = load * capacity
    reqs = poisson_requests(truths, count, rate, seed)
    run_open_served(engines, reqs[:512], lanes)          # warm-up
    served, t_s, lat_s = run_open_served(engines, reqs, lanes)
    if log:
        log("c5 served: %.1f s" % t_s)
    plain, t_p, lat_p = run_open_plain(engines, reqs)
    if log:
        log("c5 plain: %.1f s" % t_p)
    worst = max(abs(served[k] - plain[k]) / abs(plain[k]) for k in plain)
    identical = sum(1 for k in plain if served[k] == plain[k])
    # per dataset: served and plain results of one request run the dataset's
    # tuned factorization form, so they agree bit for bit
    per_ds = {}
    for k in plain:
        d = reqs[k][1]
        per_ds[d] = max(per_ds.get(d, 0.0), abs(served[k] - plain[k]) / abs(plain[k]))
    for e in engines:
        e.close()
    return {
        "datasets": [{"n": n, "cs": cs, "nu": nu, "bc": max(1, n // 13)} for n, cs, nu in specs],
        "requests": count, "lanes": lanes,
        "offered_rate_per_s": rate, "server_capacity_per_s": capacity,
        "evals_per_s": count / t_s, "plain_evals_per_s": count / t_p,
        "speedup_vs_plain": t_p / t_s,
        "latency_ms": {"p50": float(np.percentile(lat_s, 50) * 1e3),
                       "p99": float(np.percentile(lat_s, 99) * 1e3)},
        "plain_latency_ms": {"p50": float(np.percentile(lat_p, 50) * 1e3),
                             "p99": float(np.percentile(lat_p, 99) * 1e3)},
        "max_rel_diff_vs_alone": worst,
        "results_identical_to_alone": identical,
        "max_rel_diff_by_dataset": [per_ds.get(d, 0.0) for d in range(len(specs))],
        "closed_loop": {"sessions": 32, "requests": nplan,
                        "evals_per_s": capacity, "plain_evals_per_s": nplan / t_pcl,
                        "speedup_vs_plain": t_pcl / t_cap,
                        "latency_ms": {"p50": float(np.percentile(lat_cl, 50) * 1e3),
                                       "p99": float(np.percentile(lat_cl, 99) * 1e3)}},
    }


def scenarios():
    """The two serving scenarios reported by bench.py."""
    return {
        "small_datasets": benchmark(sessions=64, per_session=16, specs=SMALL_DATASETS),
        "mixed_sizes": benchmark(sessions=32, per_session=16, specs=DEFAULT_DATASETS),
        "poisson_small_datasets": benchmark_poisson(),
    }


if __name__ == "__main__":
    import json
    print(json.dumps(scenarios()))
