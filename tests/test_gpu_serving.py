"""Request server: evaluations from several datasets and covariance kinds,
merged into shared launches on several lanes, give each evaluation's bv_llh
result bit for bit (every evaluation runs its dataset's tuned factorization
form, in fixed reduction orders), and agree with the oracle at the stated
1e-10 relative tolerance; requests submitted while batches are on the GPU
are admitted and complete.  Runs on the GPU through the C ABI."""
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen
import oracle

from conftest import LLH_RTOL

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[serving] " + msg, flush=True)
    sys.stdout.flush()


def _datasets():
    specs = [(600, 40, 20, 1), (900, 60, 30, 2), (1200, 80, 60, 3), (1500, 100, 25, 4)]
    out = []
    for n, bc, cs, seed in specs:
        x, y = datagen.uniform_locations(n, seed)
        # observations drawn from the model (dense FP64 Cholesky on the CPU)
        obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=seed)
        out.append(bv.partition(x, y, obs, bc, cs, seed=0))
    return out


def _requests(nds):
    thetas = [[1.5, 0.1, 0.5], [1.0, 0.08, 1.5], [1.2, 0.05, 2.5], [1.0, 0.07, 0.8]]
    reqs = []
    tag = 0
    for rep in range(3):
        for d in range(nds):
            th = list(thetas[(d + rep) % len(thetas)])
            th[0] *= 1.0 + 0.05 * rep
            reqs.append((tag, d, th))
            tag += 1
    return reqs


def test_server_matches_alone_and_oracle():
    blocks = _datasets()
    engines = [bv.Engine(b) for b in blocks]
    reqs = _requests(len(blocks))
    _say("alone: %d requests" % len(reqs))
    alone = {tag: engines[d].llh(th) for tag, d, th in reqs}
    # a small workspace budget forces several batches per lane
    srv = bv.Server(lanes=2, ws_bytes_per_lane=3e6, max_evals_per_batch=5)
    ids = [srv.attach(e) for e in engines]
    got = {}
    _say("mix: submit half, poll, submit the rest while batches run")
    half = len(reqs) // 2
    for tag, d, th in reqs[:half]:
        srv.submit(ids[d], th, tag)
    for t, v, st in srv.poll(wait=False):
        got[t] = (v, st)
    for tag, d, th in reqs[half:]:
        srv.submit(ids[d], th, tag)
        for t, v, st in srv.poll(wait=False):
            got[t] = (v, st)
    while srv.pending() > 0:
        for t, v, st in srv.poll(wait=True):
            got[t] = (v, st)
    assert sorted(got) == sorted(alone)
    for tag, d, th in reqs:
        v, st = got[tag]
        assert st == bv.BV_OK
        assert v == alone[tag], (tag, v, alone[tag])
    _say("oracle check (the well-conditioned requests: nu <= 1.5)")
    for tag, d, th in reqs:
        if th[2] > 1.5:
            continue  # smooth and misspecified: checked alone vs mix only
        ref = oracle.llh(blocks[d], th)
        assert abs(got[tag][0] - ref) <= LLH_RTOL * abs(ref), (tag, got[tag][0], ref)
    srv.close()
    for e in engines:
        e.close()


def test_server_large_request_runs_alone_and_rejects_bad_theta():
    blocks = _datasets()[:2]
    engines = [bv.Engine(b) for b in blocks]
    srv = bv.Server(lanes=1, ws_bytes_per_lane=1.0, max_evals_per_batch=4)
    ids = [srv.attach(e) for e in engines]
    with pytest.raises(RuntimeError):
        srv.submit(ids[0], [1.0, -0.1, 0.5], 99)
    srv.submit(ids[0], [1.5, 0.1, 0.5], 1)
    srv.submit(ids[1], [1.5, 0.1, 0.5], 2)
    got = {}
    while srv.pending() > 0:
        for t, v, st in srv.poll(wait=True):
            got[t] = v
    for tag, e in ((1, engines[0]), (2, engines[1])):
        alone = e.llh([1.5, 0.1, 0.5])
        assert got[tag] == alone, (tag, got[tag], alone)
    srv.close()
    for e in engines:
        e.close()


def test_server_own_path_requests(monkeypatch):
    """Evaluations whose dataset's tuned path is a stepped form and whose
    workspace passes the size threshold run alone on that path (the
    dataset's stream, items and configuration) instead of in a merged fused
    launch; at most one at a time per dataset, the rest wait in the queue.
    Results equal bv_llh's bit for bit (the same kernels), mixed with merged
    requests of a small dataset (whose fused-form result is bv_llh's bit for
    bit too)."""
    monkeypatch.setenv("BV_PATH", "stepped2")
    monkeypatch.setenv("BV_CHUNKS", "2")
    monkeypatch.setenv("BV_CHUNK_STREAMS", "2")
    big = [bv.Engine(b) for b in _datasets()[2:]]
    monkeypatch.setenv("BV_PATH", "fused")
    small = bv.Engine(_datasets()[0])
    monkeypatch.setenv("BV_SERVER_OWN_BYTES", "1")
    engines = big + [small]
    reqs = []
    tag = 0
    for rep in range(4):
        for d in range(len(engines)):
            reqs.append((tag, d, [1.5 * (1 + 0.02 * rep), 0.1 * (1 - 0.01 * rep), 0.5]))
            tag += 1
    alone = {t: engines[d].llh(th) for t, d, th in reqs}
    srv = bv.Server(lanes=3, ws_bytes_per_lane=3e7, max_evals_per_batch=4)
    ids = [srv.attach(e) for e in engines]
    _say("own path: %d requests over %d datasets" % (len(reqs), len(engines)))
    got = {}
    for t, d, th in reqs:
        srv.submit(ids[d], th, t)
        for tt, v, st in srv.poll(wait=False):
            got[tt] = (v, st)
    while srv.pending() > 0:
        for tt, v, st in srv.poll(wait=True):
            got[tt] = (v, st)
    srv.close()
    assert sorted(got) == sorted(alone)
    for t, d, th in reqs:
        v, st = got[t]
        assert st == bv.BV_OK
        assert v == alone[t], (t, v, alone[t])
    for e in engines:
        e.close()


def _serve(engines, reqs, lanes, ws_bytes, max_evals):
    """Submit (tag, dataset, theta) requests, polling while submitting, and
    collect every result."""
    srv = bv.Server(lanes=lanes, ws_bytes_per_lane=ws_bytes, max_evals_per_batch=max_evals)
    ids = [srv.attach(e) for e in engines]
    got = {}
    for t, d, th in reqs:
        srv.submit(ids[d], th, t)
        for tt, v, st in srv.poll(wait=False):
            got[tt] = (v, st)
    while srv.pending() > 0:
        for tt, v, st in srv.poll(wait=True):
            got[tt] = (v, st)
    srv.close()
    return got


@pytest.mark.parametrize("forms", [("stepped", "stepped2", "paired", "fused"),
                                   ("paired", "paired", "stepped2", "stepped"),
                                   ("res", "fused", "res", "paired")])
def test_server_merges_stepped_forms(monkeypatch, forms):
    """Datasets fixed to each factorization form, below the own-path size:
    their requests merge into shared batches (one group of step launches per
    form and covariance kind, over blocks of several datasets and
    evaluations) and every result equals bv_llh on its dataset bit for bit,
    for closed-form and general smoothness."""
    blocks = _datasets()
    engines = []
    for b, form in zip(blocks, forms):
        if form == "res":  # the register-resident kernel, a one-launch form
            monkeypatch.setenv("BV_PATH", "fused")
            monkeypatch.setenv("BV_FACTOR_KERNEL", "res")
        else:
            monkeypatch.setenv("BV_PATH", form)
            monkeypatch.delenv("BV_FACTOR_KERNEL", raising=False)
        monkeypatch.setenv("BV_CHUNKS", "2" if form not in ("fused", "res") else "1")
        monkeypatch.setenv("BV_CHUNK_STREAMS", "0")
        engines.append(bv.Engine(b))
    thetas = [[1.5, 0.1, 0.5], [1.0, 0.08, 1.5], [1.2, 0.1, 0.8], [1.4, 0.09, 0.5]]
    reqs = []
    tag = 0
    for rep in range(4):
        for d in range(len(engines)):
            th = list(thetas[(d + rep) % len(thetas)])
            th[0] *= 1.0 + 0.03 * rep
            reqs.append((tag, d, th))
            tag += 1
    alone = {t: engines[d].llh(th) for t, d, th in reqs}
    _say("merged forms %s: %d requests" % (forms, len(reqs)))
    got = _serve(engines, reqs, lanes=2, ws_bytes=2e8, max_evals=8)
    assert sorted(got) == sorted(alone)
    for t, d, th in reqs:
        v, st = got[t]
        assert st == bv.BV_OK and v == alone[t], (t, forms[d], th, v, alone[t])
    for t, d, th in reqs[:len(engines)]:
        ref = oracle.llh(blocks[d], th)
        assert abs(got[t][0] - ref) <= LLH_RTOL * abs(ref), (t, got[t][0], ref)
    for e in engines:
        e.close()


def test_server_many_dataset_shapes():
    """Datasets of the many-dataset serving configuration's shapes, autotuned:
    n = 50,000 with cs = 60 at nu = 5/2 (smooth and dense: ill-conditioned,
    so any difference in the order of operations would show far above
    rounding), n = 20,000 with cs = 300 and with cs = 150, n = 5,000 with
    cs = 30; n/13 blocks each, observations from the Block Vecchia model of
    the partition.  Requests near each truth, several lanes: every served
    result equals bv_llh on its dataset bit for bit."""
    specs = [(50000, 60, 2.5, 11), (20000, 300, 0.5, 12), (20000, 150, 1.5, 13),
             (5000, 30, 0.5, 14)]
    engines, truths = [], []
    for n, cs, nu, seed in specs:
        _say("dataset n=%d cs=%d nu=%g" % (n, cs, nu))
        x, y = datagen.uniform_locations(n, seed)
        b0 = bv.partition(x, y, np.zeros(n), max(1, n // 13), cs, seed=0, device=0)
        obs = datagen.sample_block_vecchia(b0, 1.0, 0.1, nu, seed=seed)
        b = bv.Blocks(b0.x, b0.y, np.asarray(obs, dtype=np.float64), b0.batch_num, b0.nbr, cs)
        engines.append(bv.Engine(b))
        truths.append([1.0, 0.1, nu])
    rng = np.random.default_rng(5)
    reqs = []
    tag = 0
    for rep in range(6):
        for d in range(len(engines)):
            th = list(truths[d])
            th[0] *= 1.0 + 0.05 * rng.standard_normal()
            th[1] *= 1.0 + 0.05 * rng.standard_normal()
            reqs.append((tag, d, th))
            tag += 1
    alone = {t: engines[d].llh(th) for t, d, th in reqs}
    cfgs = [e.factor_config() for e in engines]
    _say("tuned configurations %s; serving %d requests" % (cfgs, len(reqs)))
    got = _serve(engines, reqs, lanes=3, ws_bytes=1.2e9, max_evals=16)
    assert sorted(got) == sorted(alone)
    worst = 0.0
    for t, d, th in reqs:
        v, st = got[t]
        if np.isfinite(alone[t]):
            worst = max(worst, abs(v - alone[t]) / abs(alone[t]))
        assert v == alone[t], (t, specs[d], th, v, alone[t])
    _say("max rel diff served vs alone %.3e" % worst)
    for e in engines:
        e.close()
