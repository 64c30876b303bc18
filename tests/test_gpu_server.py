"""The serving loop: every request's NAMs equal the oracle's, and are the
same when the request runs alone and when it shares steps with others
(continuous batching, several steps in flight, admission by HBM budget)."""
import numpy as np
import pytest

import gpu_index_test_amd as G
from oracle import oracle_lib as O
from tests.helpers import length_reads, n_reads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    g, lens = G.synth_genome(400_000, n_contigs=3, gc=0.41, repeat_fraction=0.08, n_families=20,
                             family_len=(300, 3000), copies=(2, 40), divergence=(0.0, 0.02),
                             tandem_fraction=0.03, tandem_unit=(1, 12), seed=91)
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens, f=0.0002, n_threads=8)
    oidx = O.Index.build_arrays(g, lens, O.params_from_read_length(150), n_threads=8)
    rng = np.random.default_rng(92)
    s, o, l, _ = G.synth_reads(g, lens, 1500, lengths=(50, 100, 150, 250, 300),
                               weights=(0.1, 0.2, 0.4, 0.2, 0.1), paired=True, sub_rate=0.005,
                               indel_rate=0.0005, seed=93)
    reads = [s[int(o[i]): int(o[i]) + int(l[i])].tobytes() for i in range(len(l))]
    reads += length_reads(rng) + n_reads(rng)
    seqs, offs, ls = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, seqs, offs, ls, 2 * oidx.filter_cutoff, 1000, 8)
    # requests: consecutive read ranges of log-uniform sizes 1..300
    sizes, first, at = [], [], 0
    while at < len(ls):
        k = int(min(len(ls) - at, max(1, round(np.exp(rng.uniform(0, np.log(300)))))))
        first.append(at)
        sizes.append(k)
        at += k
    return idx, (seqs, offs, ls), (eo, en, ef, efl), np.array(first), np.array(sizes)


def _request_expected(exp, f, k):
    eo, en, ef, efl = exp
    a, b = int(eo[f]), int(eo[f + k])
    return (eo[f: f + k + 1] - eo[f]), en[a:b], ef[f: f + k], efl[f: f + k]


def _check(res, exp):
    offs, nams, frac, flags = exp
    assert np.array_equal(res.nam_offsets, offs)
    assert res.nams.tobytes() == nams.tobytes()
    assert res.nonrepetitive_fraction.tobytes() == frac.tobytes()
    assert np.array_equal(res.flags & 1, flags & 1)


@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_alone_and_mixed_match_oracle(served, mode):
    idx, (seqs, offs, ls), exp, first, sizes = served
    srv = G.Server(idx, slots=3, max_step_reads=512, mode=mode)
    try:
        sub = lambda f, k: srv.submit(seqs, offs[f: f + k], ls[f: f + k])
        alone = []
        for f, k in zip(first, sizes):
            res, lat = srv.wait(sub(f, k))
            assert lat >= 0
            _check(res, _request_expected(exp, f, k))
            alone.append(res)
        # all at once: requests share steps and steps overlap on the device
        reqs = [sub(f, k) for f, k in zip(first, sizes)]
        for (f, k), rq, a in zip(zip(first, sizes), reqs, alone):
            res, _ = srv.wait(rq)
            assert res.nams.tobytes() == a.nams.tobytes()
            assert np.array_equal(res.nam_offsets, a.nam_offsets)
            _check(res, _request_expected(exp, f, k))
        st = srv.stats()
        assert st["requests"] == 2 * len(first)
        assert st["steps"] < 2 * len(first), "queued requests must share steps"
    finally:
        srv.close()


def test_replay_trace(served):
    idx, (seqs, offs, ls), exp, first, sizes = served
    srv = G.Server(idx, slots=4, max_step_reads=2048)
    try:
        rng = np.random.default_rng(94)
        arrival = np.cumsum(rng.exponential(0.002, size=len(first)))
        lat, wall = srv.replay(seqs, offs, ls, arrival, first, sizes)
        assert np.all(lat >= 0) and wall >= arrival[-1]
        assert srv.stats()["reads"] == int(sizes.sum())
    finally:
        srv.close()


def test_oversized_request_is_split(served):
    """A request of more reads than a step takes runs as several steps and
    comes back whole, equal to the oracle; so does one larger than the HBM
    budget. Only a read whose own step exceeds the budget is refused."""
    idx, (seqs, offs, ls), exp, first, sizes = served
    n = len(ls)
    srv = G.Server(idx, slots=3, max_step_reads=64)
    try:
        res, _ = srv.wait(srv.submit(seqs, offs, ls))
        _check(res, _request_expected(exp, 0, n))
        assert srv.stats()["steps"] >= n // 64
    finally:
        srv.close()
    srv = G.Server(idx, slots=2, hbm_budget=1 << 20)
    try:
        res, _ = srv.wait(srv.submit(seqs, offs, ls))
        _check(res, _request_expected(exp, 0, n))
    finally:
        srv.close()
    srv = G.Server(idx, slots=2, hbm_budget=8 << 10)
    try:
        with pytest.raises(G.GisError) as e:
            srv.submit(seqs, offs[:3], ls[:3])
        assert e.value.code == 8
    finally:
        srv.close()


def test_small_budget_steps_wait(served):
    """A budget below slots x step size: steps wait for budget (and idle
    slots give their arrays back) instead of failing; every request equals
    the oracle and the charged bytes never exceed the budget."""
    idx, (seqs, offs, ls), exp, first, sizes = served
    budget = 24 << 20
    srv = G.Server(idx, slots=4, max_step_reads=1024, hbm_budget=budget)
    try:
        reqs = [(f, k, srv.submit(seqs, offs[f: f + k], ls[f: f + k])) for f, k in zip(first, sizes)]
        for f, k, h in reqs:
            res, _ = srv.wait(h)
            _check(res, _request_expected(exp, f, k))
        assert srv.stats()["peak_bytes"] <= budget
    finally:
        srv.close()


def test_one_waiter_per_request(served):
    """A second caller waiting for a request that already has a waiter is
    refused (or finds it gone); the first gets the results."""
    import threading
    idx, (seqs, offs, ls), exp, first, sizes = served
    srv = G.Server(idx, slots=2, max_step_reads=4096)
    try:
        h = srv.submit(seqs, offs, ls)
        got = {}
        t = threading.Thread(target=lambda: got.setdefault("r", srv.wait(h)))
        t.start()
        import time
        time.sleep(0.05)
        with pytest.raises(G.GisError) as e:
            srv.wait(h, timeout_ms=0)
        assert e.value.code in (1, 6)
        t.join()
        _check(got["r"][0], _request_expected(exp, 0, len(ls)))
    finally:
        srv.close()
