"""Single evaluations replayed from a captured HIP graph give the same
log-likelihood, to the last bit, as the same kernels launched one by one,
for each factorization configuration, across changes of theta, with other
evaluations of the program (eager launches, a second context) and
synchronisations between the replays."""
import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen

pytestmark = pytest.mark.gpu

CONFIGS = {
    "fused": {"BV_PATH": "fused"},
    "fused_1w": {"BV_FACTOR_KERNEL": "1w"},
    "res": {"BV_FACTOR_KERNEL": "res"},
    "stepped": {"BV_PATH": "stepped"},
    "stepped_2chunks": {"BV_PATH": "stepped", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "2"},
    "stepped2_2chunks": {"BV_PATH": "stepped2", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "2"},
    "paired_2chunks": {"BV_PATH": "paired", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "2"},
    "stepped2_2chunks_skewed": {"BV_PATH": "stepped2", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "0"},
}


def _engine(n, cs, seed):
    x, y = datagen.uniform_locations(n, seed)
    obs = np.random.default_rng(seed).standard_normal(n)
    return bv.Engine(bv.partition(x, y, obs, max(1, n // 13), cs, seed=0))


@pytest.mark.parametrize("name", sorted(CONFIGS))
@pytest.mark.parametrize("nu", [0.5, 1.5, 0.8])
def test_graph_replays_match_eager(monkeypatch, name, nu):
    for k, v in CONFIGS[name].items():
        monkeypatch.setenv(k, v)
    print(f"graph replay: {name} nu={nu} n=3000 cs=60", flush=True)
    eng = _engine(3000, 60, 1)
    other = _engine(1000, 30, 2)
    thetas = [np.array([1.5 * (1 + 0.03 * i), 0.1 * (1 - 0.02 * i), nu]) for i in range(4)]
    eng.set_graph(False)
    eager = [eng.llh(t) for t in thetas]
    assert all(np.isfinite(eager))
    eng.set_graph(True)
    for rep in range(3):
        for i, t in enumerate(thetas):
            got = eng.llh(t)
            assert got == eager[i], (rep, i, got, eager[i])
            # the program's own eager work between replays
            other.llh(np.array([1.0, 0.1, 0.5]))
            pair, _ = eng.llh_many(np.stack([t, thetas[(i + 1) % 4]]))
            assert np.isfinite(pair).all()
    eng.close()
    other.close()


def test_graph_follows_buffer_changes():
    """A call with more evaluations reallocates the buffers; the next
    single evaluation captures again and still matches."""
    print("graph replay: buffer change n=2000 cs=30", flush=True)
    eng = _engine(2000, 30, 3)
    t = np.array([1.2, 0.08, 0.5])
    eng.set_graph(False)
    ref = eng.llh(t)
    eng.set_graph(True)
    assert eng.llh(t) == ref
    eng.llh_many(np.stack([t] * 5))
    assert eng.llh(t) == ref
    eng.close()
