"""Request server: evaluations from several datasets and covariance kinds,
merged into shared launches on two lanes, give the same log-likelihood as each
evaluation alone (bv_llh) and as the oracle, at the stated 1e-10 relative
tolerance; requests submitted while batches are on the GPU are admitted and
complete.  Runs on the GPU through the C ABI."""
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
        assert abs(v - alone[tag]) <= LLH_RTOL * abs(alone[tag]), (tag, v, alone[tag])
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
        assert abs(got[tag] - alone) <= LLH_RTOL * abs(alone), (tag, got[tag], alone)
    srv.close()
    for e in engines:
        e.close()


def test_server_own_path_requests(monkeypatch):
    """Evaluations whose dataset's tuned path is a stepped form and whose
    workspace passes the size threshold run alone on that path (the
    dataset's stream, items and configuration) instead of in a merged fused
    launch; at most one at a time per dataset, the rest wait in the queue.
    Results equal bv_llh's bit for bit (the same kernels), mixed with merged
    requests of a small dataset."""
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
        if d < len(big):
            assert v == alone[t], (t, v, alone[t])
        else:
            assert abs(v - alone[t]) <= LLH_RTOL * abs(alone[t]), (t, v, alone[t])
    for e in engines:
        e.close()
