"""Parity of the pair step (gis_batch_pairs, gis_pairs_from_nams) with the
oracle's: NAM lists in the reference order, pair types and per-type lists,
top_dropoff and MAPQ per list, joint NAM pairs. Everything compares bit for
bit (the dropoff is one float division; MAPQ is an integer)."""
import numpy as np
import pytest

import gpu_index_test_amd as G
from oracle import oracle_lib as O
from tests.helpers import mutate, rand_seq, revcomp
from tests.test_pairs_oracle import nam, nams, rand_list

pytestmark = pytest.mark.gpu


def _compare(res, exp, offs):
    assert np.array_equal(res.nam_offsets, np.asarray(offs, dtype=np.uint64))
    assert res.nams.tobytes() == exp["ord"].tobytes()
    assert np.array_equal(res.types, exp["types"])
    got_lists = res.type_lists()
    for t in range(5):
        assert np.array_equal(got_lists[t], exp["type_lists"][t]), t
    assert res.top_dropoff.tobytes() == exp["top_dropoff"].tobytes()
    assert np.array_equal(res.mapq, exp["mapq"])
    assert np.array_equal(res.joint_offsets, exp["joint_offsets"])
    assert res.joint.tobytes() == exp["joint"].tobytes()


def _flatten(lists):
    offs = np.zeros(len(lists) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(x) for x in lists])
    flat = np.concatenate(lists) if lists else np.zeros(0, dtype=O.NAM)
    return offs, np.ascontiguousarray(flat, dtype=O.NAM)


def _proper_rich_pair(rng, n):
    """Mate lists where most NAMs pair up (one contig, opposite strands,
    diagonals within the insert limit), so the joint scan meets its
    2 max_tries + 1 limit."""
    a, b = [], []
    for i in range(n):
        # strictly increasing reference starts keep the NAMs of a list distinct
        d = 1000 + 37 * i + int(rng.integers(0, 30))
        a.append(nam(i, 60 + i, d, d + 60, int(rng.integers(3, 8)), 4, 0,
                     score=float(rng.integers(10, 90)), nam_id=i))
        b.append(nam(i, 60 + i, d + 300, d + 360, int(rng.integers(3, 8)), 4, 1,
                     score=float(rng.integers(10, 90)), nam_id=i))
    return nams(*a), nams(*b)


def test_pairs_from_nams_matches_oracle(ctx):
    rng = np.random.default_rng(101)
    n_pairs = 700
    m1, m2 = [], []
    for p in range(n_pairs):
        kind = p % 7
        if kind == 0:
            a, b = rand_list(rng, 0), rand_list(rng, int(rng.integers(0, 4)))
        elif kind == 1:
            a, b = _proper_rich_pair(rng, int(rng.integers(5, 40)))
        elif kind == 2:
            # long lists: several 64-NAM tiles per list
            a, b = rand_list(rng, int(rng.integers(60, 300)), 3), rand_list(rng, int(rng.integers(1, 200)), 3)
        elif kind == 3:
            # equal entries (NAMs differing only in ref_id)
            x = nam(0, 30, 100, 130, 5, ref_id=0, score=50)
            y = nam(0, 30, 100, 130, 5, ref_id=1, score=50)
            z = nam(0, 30, 100, 130, 3, ref_id=2, is_rc=1, score=30)
            a, b = nams(x, y), nams(z)
        else:
            a, b = rand_list(rng, int(rng.integers(0, 9)), 2), rand_list(rng, int(rng.integers(0, 9)), 2)
        # any order within a list
        m1.append(a[rng.permutation(len(a))])
        m2.append(b[rng.permutation(len(b))])
    for lst in m1 + m2:
        keys = {(int(x["is_rc"]), int(x["ref_id"]), int(x["query_start"]), int(x["ref_start"])) for x in lst}
        assert len(keys) == len(lst)
    offs, flat = _flatten(m1 + m2)
    for mt in (1, 3, 20, 256):
        exp = O.pairs_batch(offs, flat, n_pairs, max_tries=mt)
        res = G.pairs_from_nams(ctx, offs, flat, n_pairs, max_tries=mt)
        _compare(res, exp, offs)
    # without the joint step, and with other thresholds
    exp = O.pairs_batch(offs, flat, n_pairs, dropoff=0.8, mu=120.0, sigma=35.5, joint=False)
    res = G.pairs_from_nams(ctx, offs, flat, n_pairs, dropoff_threshold=0.8, mu=120.0, sigma=35.5,
                            joint=False)
    _compare(res, exp, offs)


def test_pairs_from_nams_edge_cases(ctx):
    # no pairs at all; only empty lists; one pair
    for lists, n in (([], 0), ([rand_list(np.random.default_rng(0), 0)] * 4, 2)):
        offs, flat = _flatten(lists)
        res = G.pairs_from_nams(ctx, offs, flat, n)
        exp = O.pairs_batch(offs, flat, n)
        _compare(res, exp, offs)
        assert list(res.type_counts) == [n, 0, 0, 0, 0]
    # repeated (is_rc, ref_id, query_start, ref_start) within a list is refused
    a = nams(nam(0, 30, 100, 130, 5), nam(0, 30, 100, 130, 5))
    offs, flat = _flatten([a, nams(nam(0, 30, 300, 330, 5, is_rc=1))])
    with pytest.raises(G.GisError):
        G.pairs_from_nams(ctx, offs, flat, 1)
    with pytest.raises(G.GisError):
        G.pairs_from_nams(ctx, offs, flat, 1, max_tries=0)


def _paired_genome():
    g, lens = G.synth_genome(400_000, n_contigs=2, gc=0.42, repeat_fraction=0.1, n_families=12,
                             family_len=(300, 2500), copies=(2, 30), divergence=(0.0, 0.02),
                             tandem_fraction=0.02, tandem_unit=(1, 10), seed=23)
    return g, lens


@pytest.fixture(scope="module")
def paired_setup(ctx):
    g, lens = _paired_genome()
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens, f=0.0002, n_threads=8)
    oidx = O.Index.build_arrays(g, lens, O.params_from_read_length(150), n_threads=8)
    return g, lens, idx, oidx


def _mates_split(s, o, l):
    """Adjacent mates -> the reference's layout (all mate 1, then all mate 2)."""
    reads = [s[int(o[i]): int(o[i]) + int(l[i])].tobytes() for i in range(len(l))]
    return reads[0::2] + reads[1::2]


def test_batch_pairs_on_mapped_pairs(paired_setup):
    g, lens, idx, oidx = paired_setup
    n_pairs = 3000
    s, o, l, _ = G.synth_reads(g, lens, n_pairs, lengths=(150,), paired=True, sub_rate=0.01,
                               indel_rate=0.001, insert=(350.0, 60.0), seed=31)
    reads = _mates_split(s, o, l)
    # unrelated mates and a few unmappable reads: types 0, 1, 2 and 4 occur
    rng = np.random.default_rng(4)
    for p in range(0, n_pairs, 37):
        reads[p] = rand_seq(rng, 150)
    for p in range(5, n_pairs, 41):
        reads[n_pairs + p] = rand_seq(rng, 150)
    for p in range(11, n_pairs, 53):
        reads[n_pairs + p] = reads[n_pairs + (p * 7) % n_pairs]
    S, O_, L = G.pack_reads(reads)
    eo, en, _, _ = O.map_batch(oidx, S, O_, L, 2 * oidx.filter_cutoff, 1000, 8)
    m = G.Mapper(idx, len(reads), int(L.sum()))
    m.upload(S, O_, L)
    for mode in (G.MODE_PLAIN, G.MODE_FUSED):
        m.run(mode=mode)
        for mt in (20, 5):
            res = m.pairs(max_tries=mt)
            exp = O.pairs_batch(eo, en, n_pairs, max_tries=mt)
            _compare(res, exp, eo)
    counts = res.type_counts
    assert sum(counts) == n_pairs and all(c > 0 for c in counts), counts
    # a later download returns the same results
    again = m.pairs_download()
    assert again.joint.tobytes() == res.joint.tobytes()
    m.close()


def test_batch_pairs_repeat_pairs(ctx):
    """Mates inside high-copy elements: lists of hundreds of NAMs with
    (strand, ref) groups beyond one 64-NAM tile."""
    rng = np.random.default_rng(57)
    elem = rand_seq(rng, 1200)
    parts = [rand_seq(rng, 5000)]
    for c in range(80):
        parts.append(mutate(rng, elem, 0.003))
        parts.append(rand_seq(rng, 300))
    seqs = [b"".join(parts), rand_seq(rng, 30000)]
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(G.References(seqs), p)
    idx.populate(f=0.0002, n_threads=4)
    oidx = O.Index.build(seqs, O.params_from_read_length(150), n_threads=4)
    n_pairs = 200
    m1, m2 = [], []
    for _ in range(n_pairs):
        c = 0 if rng.random() < 0.8 else 1
        st = int(rng.integers(0, len(seqs[c]) - 500))
        frag = seqs[c][st: st + int(rng.integers(250, 480))]
        m1.append(mutate(rng, frag[:150], 0.004))
        m2.append(revcomp(mutate(rng, frag[-150:], 0.004)))
    reads = m1 + m2
    S, O_, L = G.pack_reads(reads)
    eo, en, _, _ = O.map_batch(oidx, S, O_, L, 2 * oidx.filter_cutoff, 1000, 4)
    assert int(np.max(np.diff(eo))) > 128   # lists beyond two tiles
    m = G.Mapper(idx, len(reads), int(L.sum()))
    m.upload(S, O_, L)
    m.run()
    res = m.pairs()
    _compare(res, O.pairs_batch(eo, en, n_pairs), eo)
    m.close()


def test_batch_pairs_state_errors(paired_setup):
    g, lens, idx, oidx = paired_setup
    s, o, l, _ = G.synth_reads(g, lens, 21, lengths=(150,), seed=3)
    m = G.Mapper(idx, 21, int(l.sum()))
    m.upload(s, o, l)
    with pytest.raises(G.GisError):
        m.pairs()                      # not run
    m.run()
    with pytest.raises(G.GisError):
        m.pairs()                      # odd read count
    m2 = G.Mapper(idx, 20, int(l.sum()), share=m)
    m2.upload(s, o[:20], l[:20])
    m2.run()
    with pytest.raises(G.GisError):
        m.pairs()                      # another batch ran on the workspace
    res = m2.pairs()
    assert len(res.types) == 10
    m2.close()
    m.close()
