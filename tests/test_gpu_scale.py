"""Full-size parity of the benchmark configurations: the HIP path against the
oracle on the same seeded inputs at the sizes the benchmarks run.

* 100 Mbp genome (8 contigs, GC 41%, 5% repeats): the GPU-built index equals
  the oracle's; 1M single-end 150 bp reads through the host-to-host path in
  65,536-read batches (both forms); 2M pairs of 50-300 bp reads in
  131,072-read batches, the batched form under the autotuned configuration
  and the per-stage form.
* 3.1 Gbp genome (24 contigs, 45% repeats, 2% tandem repeats), index built
  on the GPU: a full 262,144-read batch of the 10M pairs and 65,536 reads
  drawn across all of them, both forms, against the oracle over the
  downloaded table; the serving loop on the same index with requests of
  1-8,192 reads of the 50-300 bp mixture, single-end and paired, alone and
  mixed.

Every comparison is bit-exact: NAM offsets and records, nonrepetitive
fractions and rescue decisions."""
import os

import numpy as np
import pytest

import gpu_index_test_amd as G
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
MIX = dict(lengths=(50, 100, 150, 250, 300), weights=(0.1, 0.2, 0.4, 0.2, 0.1))


def _compare(res, exp):
    eo, en, ef, efl = exp
    assert np.array_equal(res.nam_offsets, eo), "nam offsets differ"
    if res.nams.tobytes() != en.tobytes():
        bad = next(r for r in range(len(eo) - 1)
                   if res.nams[int(eo[r]): int(eo[r + 1])].tobytes() != en[int(eo[r]): int(eo[r + 1])].tobytes())
        raise AssertionError(f"NAM records differ, first at read {bad}")
    assert res.nonrepetitive_fraction.tobytes() == ef.tobytes(), "nonrepetitive fractions differ"
    assert np.array_equal(res.flags & 1, efl & 1), "rescue decisions differ"


def _oracle(oidx, s, o, l):
    return O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, THREADS)


def _pair_blocks(offs, lens, batch):
    """Adjacent mates -> per batch: the batch's mate 1 reads, then its mate 2
    reads (the pair step's layout)."""
    n = len(lens)
    order = np.empty(n, dtype=np.int64)
    for lo in range(0, n, batch):
        hi = min(n, lo + batch)
        k = (hi - lo) // 2
        order[lo: lo + k] = np.arange(lo, hi, 2)
        order[lo + k: hi] = np.arange(lo + 1, hi, 2)
    return offs[order], lens[order]


# ------------------------------------------------------------ 100 Mbp genome

@pytest.fixture(scope="module")
def genome100(ctx):
    g, lens = G.synth_genome(100_000_000, n_contigs=8, gc=0.41, repeat_fraction=0.05, n_families=500,
                             family_len=(1000, 10000), copies=(2, 50), divergence=(0.01, 0.05), seed=1)
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens, f=0.0002, n_threads=THREADS)
    oidx = O.Index.build_arrays(g, lens, O.params_from_read_length(150), n_threads=THREADS)
    yield g, lens, idx, oidx
    oidx.free()
    idx.ctx.close()


@pytest.mark.timeout(600)
def test_100mbp_index_matches_oracle(genome100):
    g, lens, idx, oidx = genome100
    info = idx.info()
    assert (info.bits, info.filter_cutoff, info.n_randstrobes) == (oidx.bits, oidx.filter_cutoff, oidx.n)
    assert info.bits == 23
    assert idx.randstrobes.tobytes() == oidx.randstrobes.tobytes()
    assert np.array_equal(idx.randstrobe_start_indices, oidx.starts)


@pytest.mark.timeout(600)
def test_1m_single_end_150bp_all_reads(genome100):
    """1M single-end 150 bp reads (0.5% substitutions, 0.05% indels), every
    read, host memory to host memory in 65,536-read batches."""
    g, lens, idx, oidx = genome100
    s, o, l, _ = G.synth_reads(g, lens, 1_000_000, lengths=(150,), weights=(1.0,), sub_rate=0.005,
                               indel_rate=0.0005, seed=2)
    exp = _oracle(oidx, s, o, l)
    assert (exp[3] & 1).sum() > 0, "the reads must exercise the rescue path"
    pin = (G.pinned_copy(s), G.pinned_copy(o), G.pinned_copy(l))
    for mode in (G.MODE_PLAIN, G.MODE_FUSED):
        res, ms = G.find_nams_stream(idx, *pin, batch_reads=65536, lanes=2, mode=mode)
        assert len(ms) == 16
        _compare(res, exp)


@pytest.mark.timeout(600)
def test_2m_pairs_mixed_lengths_all_reads(genome100):
    """2M pairs of 50-300 bp reads (insert N(400, 50)), every read, in
    131,072-read batches laid out for the pair step: the batched form under
    the autotuned configuration and the per-stage form."""
    g, lens, idx, oidx = genome100
    s, o, l, _ = G.synth_reads(g, lens, 2_000_000, paired=True, insert=(400.0, 50.0), sub_rate=0.005,
                               indel_rate=0.0005, seed=2, **MIX)
    o, l = _pair_blocks(o, l, 131072)
    exp = _oracle(oidx, s, o, l)
    m = G.Mapper(idx, 131072, int((o[:131072] + l[:131072]).max()))
    m.upload(s, o[:131072], l[:131072])
    best, _, n_tried = m.autotune()
    m.close()
    assert n_tried > 1
    pin = (G.pinned_copy(s), G.pinned_copy(o), G.pinned_copy(l))
    for mode in (G.MODE_FUSED, G.MODE_PLAIN):
        res, ms = G.find_nams_stream(idx, *pin, batch_reads=131072, lanes=2, mode=mode)
        assert len(ms) == 31
        _compare(res, exp)


# ------------------------------------------------------------ 3.1 Gbp genome

@pytest.fixture(scope="module")
def genome3g(ctx):
    g, lens = G.synth_genome(3_100_000_000, n_contigs=24, gc=0.41, repeat_fraction=0.45, n_families=5000,
                             family_len=(300, 6000), copies=(2, 50), divergence=(0.02, 0.20),
                             tandem_fraction=0.02, tandem_unit=(2, 50), contig_len=(50_000_000, 250_000_000),
                             seed=1)
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens, f=0.0002, n_threads=THREADS)
    info = idx.info()
    oidx = O.Index.wrap(idx.randstrobes, idx.randstrobe_start_indices, info.bits, info.filter_cutoff,
                        O.params_from_read_length(150))
    idx._host = None  # the oracle keeps its own reference to the table
    yield g, lens, idx, oidx
    oidx.free()
    idx.ctx.close()


@pytest.mark.timeout(900)
def test_3g_index_batch_and_sample(genome3g):
    """10M pairs of 150 bp: the first 262,144-read batch and 65,536 reads
    drawn uniformly across all 20M, both forms, against the oracle over the
    table built on the GPU."""
    g, lens, idx, oidx = genome3g
    info = idx.info()
    assert info.bits == 28 and info.n_randstrobes > 500_000_000
    s, o, l, _ = G.synth_reads(g, lens, 10_000_000, lengths=(150,), weights=(1.0,), paired=True,
                               insert=(400.0, 50.0), sub_rate=0.005, indel_rate=0.0005, seed=2)
    pick = np.random.default_rng(7).choice(len(l) - 262144, size=65536, replace=False) + 262144
    for sel in (np.arange(262144), np.sort(pick)):
        so, sl = o[sel], l[sel]
        exp = _oracle(oidx, s, so, sl)
        # the batch from pinned memory; the sample, spread over all the
        # reads, from pageable memory
        end = int((so + sl).max())
        pin = (G.pinned_copy(s[:end]), G.pinned_copy(so), G.pinned_copy(sl)) if len(sel) == 262144 else (s, so, sl)
        for mode in (G.MODE_FUSED, G.MODE_PLAIN):
            res, _ = G.find_nams_stream(idx, *pin, batch_reads=262144, lanes=1, mode=mode)
            _compare(res, exp)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_3g_serving_alone_and_mixed(genome3g, mode):
    """The serving loop on the 3.1 Gbp index: requests of 1-8,192 reads
    (log-uniform) of the 50-300 bp mixture, single-end and paired half and
    half; each request's NAMs equal the oracle's alone and when all requests
    are in flight together."""
    g, lens, idx, oidx = genome3g
    se = G.synth_reads(g, lens, 40_000, sub_rate=0.005, indel_rate=0.0005, seed=3, **MIX)
    pe = G.synth_reads(g, lens, 20_000, paired=True, insert=(400.0, 50.0), sub_rate=0.005,
                       indel_rate=0.0005, seed=2, **MIX)
    s = np.concatenate([se[0], pe[0]])
    o = np.concatenate([se[1], pe[1] + np.uint64(len(se[0]))])
    l = np.concatenate([se[2], pe[2]])
    n_se = len(se[2])
    exp = _oracle(oidx, s, o, l)
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(24):
        k = int(np.clip(np.exp(rng.uniform(0, np.log(8192))), 1, 8192))
        if i % 2:
            k += k & 1
            f = n_se + 2 * int(rng.integers(0, (len(l) - n_se - k) // 2))
        else:
            f = int(rng.integers(0, n_se - k))
        reqs.append((f, k))
    reqs.append((0, 8192))

    def expected(f, k):
        eo, en, ef, efl = exp
        return (eo[f: f + k + 1] - eo[f], en[int(eo[f]): int(eo[f + k])], ef[f: f + k], efl[f: f + k])

    srv = G.Server(idx, slots=4, max_step_reads=16384, mode=mode)
    try:
        for f, k in reqs:
            res, _ = srv.wait(srv.submit(s, o[f: f + k], l[f: f + k]))
            _compare(res, expected(f, k))
        handles = [srv.submit(s, o[f: f + k], l[f: f + k]) for f, k in reqs]
        for (f, k), h in zip(reqs, handles):
            res, _ = srv.wait(h)
            _compare(res, expected(f, k))
        assert srv.stats()["steps"] < 2 * len(reqs)
    finally:
        srv.close()
