"""Parity of the HIP path with the CPU oracle on the same seeded inputs.

Every comparison is bit-exact (integer/byte work; the one float field, the
NAM score, is an int->float conversion and the nonrepetitive fraction a
float division, both compared bit for bit)."""
import numpy as np
import pytest

import gpu_index_test_amd as G
from oracle import oracle_lib as O
from tests.helpers import (HAND_NAMS, hand_case, length_reads, mutate, n_reads, rand_seq, revcomp,
                           tie_heavy_reads, tiny_index_arrays)

pytestmark = pytest.mark.gpu


def _genome():
    # 600 kbp, 3 contigs, repeat families (filtered seeds, rescue) and tandem
    # repeats (tied syncmer windows)
    g, lens = G.synth_genome(600_000, n_contigs=3, gc=0.41, repeat_fraction=0.08, n_families=20,
                             family_len=(300, 3000), copies=(2, 40), divergence=(0.0, 0.02),
                             tandem_fraction=0.03, tandem_unit=(1, 12), seed=11)
    return g, lens


@pytest.fixture(scope="module")
def setup(ctx):
    g, lens = _genome()
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)  # a context of its own
    idx.populate_arrays(g, lens, f=0.0002, n_threads=8)
    op = O.params_from_read_length(150)
    oidx = O.Index.build_arrays(g, lens, op, n_threads=8)
    return g, lens, p, idx, op, oidx


def _read_set(g, lens, n=400, seed=5):
    rng = np.random.default_rng(seed)
    s, o, l, _ = G.synth_reads(g, lens, n, lengths=(50, 100, 150, 250, 300),
                               weights=(0.1, 0.2, 0.4, 0.2, 0.1), sub_rate=0.005,
                               indel_rate=0.0005, seed=seed)
    reads = [s[int(o[i]): int(o[i]) + int(l[i])].tobytes() for i in range(n)]
    reads += tie_heavy_reads(rng) + n_reads(rng) + length_reads(rng)
    # exact copies of tandem-repeat and repeat-family regions of the genome
    gb = g.tobytes()
    for st in rng.integers(0, len(gb) - 400, size=60):
        reads.append(gb[st: st + 150])
        reads.append(revcomp(gb[st + 7: st + 157]))
    return reads


def test_index_build_matches_oracle(setup):
    g, lens, p, idx, op, oidx = setup
    info = idx.info()
    assert info.bits == oidx.bits
    assert info.filter_cutoff == oidx.filter_cutoff
    assert info.n_randstrobes == oidx.n
    assert idx.randstrobes.tobytes() == oidx.randstrobes.tobytes()
    assert np.array_equal(idx.randstrobe_start_indices, oidx.starts)
    assert info.distinct_strobemers == oidx.distinct_strobemers


def _adversarial_reference():
    """Contigs with N runs, a 3 kbp homopolymer, lower case, short tandem
    repeats (tied syncmer windows) and contigs shorter than w_max."""
    from tests.helpers import rand_seq
    rng = np.random.default_rng(31)
    g, lens = G.synth_genome(200_000, n_contigs=2, gc=0.41, repeat_fraction=0.05, n_families=5,
                             family_len=(300, 2000), copies=(2, 10), tandem_fraction=0.08,
                             tandem_unit=(1, 6), seed=12)
    gb = g.tobytes()
    seqs = [gb[: int(lens[0])], gb[int(lens[0]):]]
    s0 = bytearray(seqs[0])
    s0[5000:5040] = b"N" * 40
    s0[20000:23000] = b"A" * 3000
    s0[40000:40100] = bytes(s0[40000:40100]).lower()
    s0[60000:60001] = b"N"
    seqs[0] = bytes(s0)
    seqs += [rand_seq(rng, 30), rand_seq(rng, 5), rand_seq(rng, 2500), b"ACGTN" * 3, b"AC" * 600]
    return seqs


def test_device_seeding_matches_oracle(ctx):
    """The reference seeded on the device (gis_reference_randstrobes, the
    seeding gis_index_build runs) gives the oracle's records, in (contig,
    position) order, also inside tandem repeats and homopolymers (tied
    windows replayed from their anchors), around N runs and for contigs
    shorter than w_max; an index built from the same contigs equals the
    oracle's index."""
    import ctypes as C
    from gpu_index_test_amd import _native as N
    seqs = _adversarial_reference()
    p = G.IndexParameters.from_read_length(150)
    bufs = [C.create_string_buffer(s, len(s)) for s in seqs]
    arr = (C.c_void_p * len(bufs))(*[C.cast(b, C.c_void_p) for b in bufs])
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    n = C.c_uint64()
    L = N.lib()
    assert L.gis_reference_randstrobes(ctx.h, arr, N.ptr(lens), len(seqs), C.byref(p.c), None, 0,
                                       C.byref(n)) == 5
    got = np.empty(max(1, n.value), dtype=G.REF_RANDSTROBE)
    ctx.check(L.gis_reference_randstrobes(ctx.h, arr, N.ptr(lens), len(seqs), C.byref(p.c), N.ptr(got),
                                          n.value, C.byref(n)))
    got = got[: n.value]
    op = O.params_from_read_length(150)
    oidx = O.Index.build(seqs, op, n_threads=2)
    order = np.lexsort((got["packed"], got["position"], got["hash"]))
    assert got[order].tobytes() == oidx.randstrobes.tobytes()
    key = (got["packed"].astype(np.uint64) >> np.uint64(8)) * np.uint64(2 ** 32) + got["position"]
    nz = key[got["hash"] != 0]
    assert np.all(np.diff(nz.astype(np.int64)) > 0)
    idx = G.StrobemerIndex(G.References(seqs), p)
    idx.populate()
    assert idx.randstrobes.tobytes() == oidx.randstrobes.tobytes()
    assert np.array_equal(idx.randstrobe_start_indices, oidx.starts)
    assert idx.filter_cutoff == oidx.filter_cutoff
    oidx.free()
    idx.ctx.close()


def test_syncmers_parity(setup):
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 200)
    s, o, l = G.pack_reads(reads)
    offs, sy = G.syncmers_batch(idx, s, o, l)
    for r, seq in enumerate(reads):
        exp = O.syncmers(seq, op)
        got = sy[int(offs[r]): int(offs[r + 1])]
        assert got.tobytes() == exp.tobytes(), (r, len(seq), seq[:60])


def test_randstrobes_parity(setup):
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 300)
    s, o, l = G.pack_reads(reads)
    offs, rs = G.randstrobes_batch(idx, s, o, l)
    for r, seq in enumerate(reads):
        exp = O.randstrobes_query(seq, op)
        got = rs[int(offs[r]): int(offs[r + 1])]
        assert len(got) == len(exp), (r, len(seq))
        for f in ("hash", "start", "end", "is_reverse"):
            assert np.array_equal(got[f], exp[f]), (r, f, len(seq))
    # randstrobes_query(seq, IndexParameters): no index involved
    for seq in reads[::7]:
        exp = O.randstrobes_query(seq, op)
        got = G.randstrobes_query(seq, p)
        assert len(got) == len(exp)
        for f in ("hash", "start", "end", "is_reverse"):
            assert np.array_equal(got[f], exp[f]), (f, len(seq))


@pytest.mark.parametrize("read_length", [50, 100, 125, 250, 300, 400])
def test_randstrobes_every_profile(setup, read_length):
    """Seeding under the other read-length profiles: k 20-23, s 16-18 (s-mer
    codes of 16 bases and of more), t 3-4, their w_min/w_max and max_dist;
    reads with and without N, tie-heavy reads, lengths around the windows."""
    g, lens, _, _, _, _ = setup
    p = G.IndexParameters.from_read_length(read_length)
    op = O.params_from_read_length(read_length)
    reads = _read_set(g, lens, 150, seed=read_length)
    s, o, l = G.pack_reads(reads)
    ctx = G.default_context()
    cap = int(2 * np.maximum(l.astype(np.int64) - p.syncmer.k + 1, 0).sum()) or 1
    out = np.empty(cap, dtype=G.QUERY_RANDSTROBE)
    oo = np.empty(len(l) + 1, dtype=np.uint64)
    nout = G._native.C.c_uint64()
    N = G._native
    ctx.check(ctx.L.gis_randstrobes_params(ctx.h, N.C.byref(p.c), N.ptr(s), N.ptr(o), N.ptr(l), len(l),
                                           N.ptr(oo), N.ptr(out), cap, N.C.byref(nout)))
    for r, seq in enumerate(reads):
        exp = O.randstrobes_query(seq, op)
        got = out[int(oo[r]): int(oo[r + 1])]
        assert len(got) == len(exp), (read_length, r, len(seq))
        for f in ("hash", "start", "end", "is_reverse"):
            assert np.array_equal(got[f], exp[f]), (read_length, r, f, len(seq))


def test_lookup_parity(setup):
    g, lens, p, idx, op, oidx = setup
    rng = np.random.default_rng(3)
    tab = oidx.randstrobes["hash"]
    keys = np.concatenate([
        tab[rng.integers(0, len(tab), 5000)],                 # present
        tab[:3], tab[-3:],                                    # first hash (never findable) and last
        rng.integers(0, 2 ** 63, 5000, dtype=np.uint64) * 2 + 1,  # mostly absent
        tab[rng.integers(0, len(tab), 500)] + 1,              # near misses
        np.array([0, 2 ** 64 - 1], dtype=np.uint64),
    ]).astype(np.uint64)
    got = idx.find_many(keys)
    exp = np.array([oidx.find(int(k)) for k in keys], dtype=np.uint64)
    assert np.array_equal(got, exp)


def _expected_hits(oidx, qrs, rescue, rescue_cutoff, hard=1000):
    """Hits of one read from the oracle's find/is_filtered/get_count and its
    table: the add_to_hits_per_ref walk (step2/nam.cu:45-62)."""
    rs = oidx.randstrobes
    k = oidx.params.k

    def walk(pos, qs, qe):
        out = []
        key = int(rs["hash"][pos])
        md = 2 ** 31 - 1
        x = pos
        while x < len(rs) and int(rs["hash"][x]) == key:
            rspan = (int(rs["packed"][x]) & 0xFF) + k
            d = abs((qe - qs) - rspan)
            if d <= md:
                out.append((int(rs["packed"][x]) >> 8, qs, qe, int(rs["position"][x]),
                            int(rs["position"][x]) + rspan))
                md = d
            x += 1
        return out

    hits = {0: [], 1: []}
    if not rescue:
        for q in qrs:
            pos = oidx.find(int(q["hash"]))
            if pos == O.NOT_FOUND or oidx.is_filtered(pos):
                continue
            hits[int(q["is_reverse"])] += walk(pos, int(q["start"]), int(q["end"]))
    else:
        for strand in (0, 1):
            items = []
            for q in qrs[qrs["is_reverse"] == strand]:
                pos = oidx.find(int(q["hash"]))
                if pos != O.NOT_FOUND:
                    items.append((oidx.get_count(pos), int(q["start"]), int(q["end"]), pos))
            items.sort()
            taken = []
            for c, qs, qe, pos in items:
                if (c > rescue_cutoff and len(taken) >= 5) or c > hard:
                    break
                taken.append((qs, qe, pos))
            for qs, qe, pos in sorted(taken):
                hits[strand] += walk(pos, qs, qe)
    out = []
    for strand in (0, 1):
        # grouped by ref, (query_start, ref_start) order within a ref
        for ref, qs, qe, rs0, re0 in sorted(hits[strand], key=lambda h: (h[0], h[1], h[3])):
            out.append((qs, qe, rs0, re0, ref, strand))
    return out


def test_hits_parity(setup):
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 150, seed=9)
    s, o, l = G.pack_reads(reads)
    offs, hits = G.hits_batch(idx, s, o, l)
    rcut = 2 * oidx.filter_cutoff
    for r, seq in enumerate(reads):
        _, fr, fl = O.map_read(oidx, seq, rcut)
        qrs = O.randstrobes_query(seq, op)
        exp = _expected_hits(oidx, qrs, bool(fl & 1), rcut)
        got = hits[int(offs[r]): int(offs[r + 1])]
        got_t = [tuple(int(x) for x in h) for h in got]
        assert got_t == exp, (r, len(got_t), len(exp))


def _compare_nams(got, exp, frac_g, frac_e, flags_g, flags_e):
    assert np.array_equal(got[0], exp[0]), "nam offsets differ"
    assert got[1].tobytes() == exp[1].tobytes(), "NAM records differ"
    assert frac_g.tobytes() == frac_e.tobytes(), "nonrepetitive fractions differ"
    assert np.array_equal(flags_g & 1, flags_e & 1), "rescue decisions differ"


@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_find_nams_parity(setup, mode):
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 1500, seed=21)
    s, o, l = G.pack_reads(reads)
    res = G.find_nams_batch(idx, s, o, l, mode=mode)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    if not np.array_equal(res.nam_offsets, eo) or res.nams.tobytes() != en.tobytes():
        for r in range(len(reads)):
            a = res.read_nams(r)
            b = en[int(eo[r]): int(eo[r + 1])]
            assert a.tobytes() == b.tobytes(), (r, len(reads[r]), a, b)
    _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                  res.flags, efl)
    assert (efl & 1).sum() > 0, "the read set must exercise the rescue path"


def test_find_nams_qrs_parity(setup):
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 120, seed=33)
    rcut = 2 * oidx.filter_cutoff
    for seq in reads:
        qrs = O.randstrobes_query(seq, op)
        fr, nams = G.find_nams(qrs, idx)
        efr, enams, _ = O.find_nams(oidx, qrs, canonical=True)
        assert np.float32(fr).tobytes() == np.float32(efr).tobytes()
        assert nams.tobytes() == enams.tobytes()
        rn = G.find_nams_rescue(qrs, idx, rcut)
        ern = O.find_nams_rescue(oidx, qrs, rcut, 1000, canonical=True)
        assert rn.tobytes() == ern.tobytes()
    # many reads' query randstrobes in one call, both decisions
    qs = [O.randstrobes_query(seq, op) for seq in reads]
    for rescue in (0, 1):
        res = G._qrs_call(idx, qs, rescue, rcut, 1000)
        for r, q in enumerate(qs):
            if rescue:
                exp = O.find_nams_rescue(oidx, q, rcut, 1000, canonical=True)
            else:
                _, exp, _ = O.find_nams(oidx, q, canonical=True)
            assert res.read_nams(r).tobytes() == exp.tobytes(), (rescue, r)
    # a query span the packed device form cannot hold is refused
    for field, value in (("start", None), ("end", 70000)):
        bad = O.randstrobes_query(reads[0], op).copy()
        bad[field][0] = bad["end"][0] + 1 if value is None else value
        with pytest.raises(G.GisError):
            G.find_nams(bad, idx)


def test_find_nams_composition_worked_example(ctx):
    """The find_nams case worked by hand from the reference's text (the
    running min_diff filter of a run, a filtered and an absent seed, both
    strands, the merge and its scores; tests/helpers.py) on the GPU path: the
    table uploaded as is, the query randstrobes through gis_find_nams_qrs."""
    ents, q = hand_case()
    arr, starts = tiny_index_arrays(ents, 8)
    idx = G.StrobemerIndex(None, G.IndexParameters.from_read_length(150))
    idx.upload(arr, starts, 8, 5)
    fr, nams = G.find_nams(q, idx)
    assert np.float32(fr) == np.float32(3) / np.float32(4)
    got = [(int(n["is_rc"]), int(n["query_start"]), int(n["query_end"]), int(n["ref_start"]), int(n["ref_end"]),
            int(n["n_hits"]), float(n["score"])) for n in nams]
    assert got == HAND_NAMS
    assert [int(n["nam_id"]) for n in nams] == list(range(len(HAND_NAMS)))
    idx.ctx.close()


def test_bucket_line_boundaries(ctx):
    """The lookup reads a bucket of up to seven records from its line and
    searches the table for larger ones: buckets of 7 and 8 equal records, a
    run one past the filter cutoff, several hashes in one bucket, empty
    buckets and misses, through gis_lookup_batch and find_nams, against the
    oracle over the same table."""
    def h(top, low):
        return (top << 56) | low

    ents = [(h(0x10, 5), 100 + i, 10 + i) for i in range(7)]             # 7 records: the line
    ents += [(h(0x20, 5), 200 + i, 12) for i in range(8)]                # 8: the table search
    ents += [(h(0x30, 5), 300 + i, 14) for i in range(9)]                # 9 > cutoff 8: filtered
    ents += [(h(0x40, v), 400 + v, 16) for v in (3, 7, 11)]              # three hashes, one bucket
    ents += [(h(0x50, 9), 500 + i, 20 + (i % 3)) for i in range(6)] + [(h(0x50, 10), 560, 18)]
    ents += [(h(0x60, v), 600 + v, 22) for v in range(12)]               # 12 hashes, one bucket
    arr, starts = tiny_index_arrays(ents, 8)
    op = O.params_from_read_length(150)
    oidx = O.Index.wrap(arr, starts, 8, 8, op)
    idx = G.StrobemerIndex(None, G.IndexParameters.from_read_length(150))
    idx.upload(arr, starts, 8, 8)
    present = sorted({e[0] for e in ents})
    absent = [h(0x10, 4), h(0x10, 6), h(0x20, 4), h(0x40, 5), h(0x40, 12), h(0x60, 12), h(0x70, 1), 1, 0,
              2 ** 64 - 1]
    keys = np.array(present + absent, dtype=np.uint64)
    got = idx.find_many(keys)
    exp = np.array([oidx.find(int(k)) for k in keys], dtype=np.uint64)
    assert np.array_equal(got, exp)
    # every key as a forward and a reverse query randstrobe of one read,
    # with spans that make the running minimum keep some records and not others
    q = np.zeros(2 * len(keys), dtype=O.QUERY_RANDSTROBE)
    q["hash"] = np.concatenate([keys, keys])
    q["start"] = np.tile(np.arange(len(keys), dtype=np.uint32) * 2, 2)
    q["end"] = q["start"] + 40
    q["is_reverse"] = np.repeat([0, 1], len(keys))
    fr, nams = G.find_nams(q, idx)
    efr, enams, _ = O.find_nams(oidx, q, canonical=True)
    assert np.float32(fr).tobytes() == np.float32(efr).tobytes()
    assert nams.tobytes() == enams.tobytes()
    idx.ctx.close()


def test_sti_roundtrip(setup, tmp_path):
    g, lens, p, idx, op, oidx = setup
    path = tmp_path / "ref.r150.sti"
    idx.write(path)
    raw = path.read_bytes()
    assert raw[:4] == b"STI\x01"
    assert np.frombuffer(raw[4:8], "<u4")[0] == 2
    rsv = int(np.frombuffer(raw[8:16], "<u8")[0])
    o = 16 + rsv
    cutoff, bits, r, k, s_, l_, u_, q, md = np.frombuffer(raw[o: o + 36], "<i4")
    assert (cutoff, bits) == (oidx.filter_cutoff, oidx.bits)
    assert (r, k, s_, l_, u_, q, md) == (150, 20, 16, 1, 7, 255, 80)
    o += 36
    n = int(np.frombuffer(raw[o: o + 8], "<u8")[0])
    o += 8
    assert raw[o: o + 16 * n] == oidx.randstrobes.tobytes()
    o += 16 * n
    m = int(np.frombuffer(raw[o: o + 8], "<u8")[0])
    assert m == (1 << bits) + 1
    assert raw[o + 8:] == oidx.starts.tobytes()
    idx2 = G.StrobemerIndex(None, p, ctx=idx.ctx)
    idx2.read(path)
    assert idx2.randstrobes.tobytes() == oidx.randstrobes.tobytes()
    # and back to the original index for the other tests
    idx.read(path)


def test_c1_scale_parity(ctx):
    """5 Mbp uniform genome, 20 000 single-end 150 bp reads, 50% reverse
    complemented, 0.1% substitutions, 0.01% indels."""
    g, lens = G.synth_genome(5_000_000, n_contigs=1, gc=0.5, seed=1)
    p = G.IndexParameters.from_read_length(150)
    idx = G.StrobemerIndex(None, p)
    idx.populate_arrays(g, lens)
    oidx = O.Index.build_arrays(g, lens, O.params_from_read_length(150), n_threads=8)
    assert idx.randstrobes.tobytes() == oidx.randstrobes.tobytes()
    s, o, l, _ = G.synth_reads(g, lens, 20_000, sub_rate=0.001, indel_rate=0.0001, seed=2)
    for mode in (G.MODE_PLAIN, G.MODE_FUSED):
        res = G.find_nams_batch(idx, s, o, l, mode=mode)
        eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
        _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                      res.flags, efl)
    idx.ctx.close()


def test_shared_context_detects_replaced_index(setup):
    g, lens, p, idx, op, oidx = setup
    other = G.StrobemerIndex(None, p, ctx=idx.ctx)
    other.upload(oidx.randstrobes, oidx.starts, oidx.bits, oidx.filter_cutoff)
    with pytest.raises(G.GisError):
        idx.find(1)
    idx.upload(oidx.randstrobes, oidx.starts, oidx.bits, oidx.filter_cutoff)
    assert idx.randstrobes.tobytes() == oidx.randstrobes.tobytes()


@pytest.mark.parametrize("mode,lanes,per", [(G.MODE_PLAIN, 1, 4096), (G.MODE_PLAIN, 3, 97),
                                            (G.MODE_FUSED, 2, 333)])
def test_stream_host_to_host_matches_oracle(setup, mode, lanes, per):
    """gis_find_nams_stream: reads in pinned host memory cut into batches
    (the last one ragged), several in flight, NAM arrays copied back into one
    output in read order; equal to the oracle, and the same on a second call
    that reuses the lanes, their graphs and the output arrays."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 1200, seed=81)
    s, o, l = G.pack_reads(reads)
    # reads out of storage order within the pool (the stream derives each
    # batch's base from its smallest offset)
    perm = np.random.default_rng(82).permutation(len(l))
    o, l = o[perm], l[perm]
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    ps, po, pl = G.pinned_copy(s), G.pinned_copy(o), G.pinned_copy(l)
    res, ms = G.find_nams_stream(idx, ps, po, pl, batch_reads=per, lanes=lanes, mode=mode)
    _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef, res.flags, efl)
    assert len(ms) == -(-len(l) // per) and np.all(ms > 0)
    st0 = G.stream_copy_stats(idx)
    res2, _ = G.find_nams_stream(idx, ps, po, pl, batch_reads=per, lanes=lanes, mode=mode, out=res)
    _compare_nams((res2.nam_offsets, res2.nams), (eo, en), res2.nonrepetitive_fraction, ef, res2.flags, efl)
    # pinned memory: every copy ran on the two engines, one per direction
    st1 = G.stream_copy_stats(idx)
    assert st1["h2d_engine"] and st1["d2h_engine"] and not st1["h2d_engine"] & st1["d2h_engine"], st1
    assert st1["refused"] == 0 and st1["hip_copies"] == st0["hip_copies"], st1
    assert st1["engine_copies"] > st0["engine_copies"]
    # pageable host memory, for the reads and for the NAMs, takes the same
    # path, its copies through HIP
    res3, _ = G.find_nams_stream(idx, s, o, l, batch_reads=per, lanes=lanes, mode=mode, pinned=False)
    _compare_nams((res3.nam_offsets, res3.nams), (eo, en), res3.nonrepetitive_fraction, ef, res3.flags, efl)
    assert G.stream_copy_stats(idx)["hip_copies"] > st1["hip_copies"]


def test_lookup_bytes_match_oracle(setup):
    """The lookup kernel's algorithmic byte count (the roofline's numerator)
    equals the oracle's count over the same reads' query randstrobes; the
    lookup count equals the randstrobe count."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 800, seed=71)
    s, o, l = G.pack_reads(reads)
    m = G.Mapper(idx, len(l), len(s))
    m.upload(s, o, l)
    m.run(mode=G.MODE_PLAIN)
    got, n_lookups, streams = m.lookup_bytes()
    qrs = [O.randstrobes_query(seq, op) for seq in reads]
    exp = sum(O.lookup_bytes(oidx, q) for q in qrs)
    assert n_lookups == sum(len(q) for q in qrs)
    assert got == exp
    assert streams >= 12 * n_lookups  # each query read (hash + packed span), at least
    m.close()


@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_replaced_index_drops_captured_graphs(ctx, mode):
    """Reads mapped with graph replays on index A, then a differently sized
    index B (another genome, other bits) loaded on the same context: the
    context's batch must not replay the launch sequence captured with A's
    (freed) table; the same reads give the oracle's NAMs on B."""
    ctx2 = G.Context(0)
    p = G.IndexParameters.from_read_length(150)
    ga, la = G.synth_genome(300_000, n_contigs=2, gc=0.45, seed=31)
    gb, lb = G.synth_genome(1_500_000, n_contigs=3, gc=0.41, repeat_fraction=0.05, n_families=10,
                            family_len=(300, 2000), copies=(2, 20), seed=32)
    s, o, l, _ = G.synth_reads(gb, lb, 3000, lengths=(100, 150, 250), weights=(0.3, 0.5, 0.2),
                               sub_rate=0.005, indel_rate=0.0005, seed=33)
    ia = G.StrobemerIndex(None, p, ctx=ctx2)
    ia.populate_arrays(ga, la)
    bits_a = ia.info().bits
    for _ in range(2):  # capture, then replay
        G.find_nams_batch(ia, s, o, l, mode=mode, graph=True)
    ib = G.StrobemerIndex(None, p, ctx=ctx2)
    ib.populate_arrays(gb, lb)
    assert ib.info().bits != bits_a
    oidx = O.Index.build_arrays(gb, lb, O.params_from_read_length(150), n_threads=8)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    for _ in range(2):
        res = G.find_nams_batch(ib, s, o, l, mode=mode, graph=True)
        _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                      res.flags, efl)
    oidx.free()
    ctx2.close()


@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_pool_overflow_reruns(setup, mode, monkeypatch):
    """Tiny initial hit/NAM pools: every pool overflows, grows to what the
    device allocators asked for, and the batch runs again."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 300, seed=44)
    s, o, l = G.pack_reads(reads)
    monkeypatch.setenv("GIS_TEST_SMALL_POOLS", "1")
    ctx2 = G.Context(0)
    idx2 = G.StrobemerIndex(None, p, ctx=ctx2)
    idx2.populate_arrays(g, lens)
    res = G.find_nams_batch(idx2, s, o, l, mode=mode)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                  res.flags, efl)
    ctx2.close()


def test_fused_spills_and_mapper(setup):
    """The fused form hands long and hit-heavy reads to the per-stage kernels;
    a device-resident batch (upload / run / download) gives the same NAMs in
    both forms."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 400, seed=55)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    m = G.Mapper(idx, len(l), len(s))
    m.upload(s, o, l)
    for mode in (G.MODE_FUSED, G.MODE_PLAIN, G.MODE_FUSED):
        m.run(mode=mode)
        res = m.download()
        _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                      res.flags, efl)
        if mode == G.MODE_FUSED:
            spilled = (res.flags & 0x80) != 0
            assert spilled.any(), "long reads must take the per-stage path"
            assert (~spilled).sum() > 0.8 * len(l)
            assert spilled[np.asarray(l) > 320].all()
    m.close()


@pytest.mark.parametrize("mode", [G.MODE_PLAIN, G.MODE_FUSED])
def test_graph_replays_match_eager(setup, mode):
    """A batch's captured launch sequence replayed several times, with eager
    (non-graph) work of the same context and synchronisations in between;
    every replay equals the eager result and the oracle."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 500, seed=66)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    m = G.Mapper(idx, len(l), len(s))
    m.upload(s, o, l)
    other = _read_set(g, lens, 200, seed=67)
    s2, o2, l2 = G.pack_reads(other)
    for rep in range(4):
        m.run(mode=mode, graph=True)
        res = m.download()
        _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                      res.flags, efl)
        eager = G.find_nams_batch(idx, s2, o2, l2, mode=mode, graph=False)
        assert len(eager.nam_offsets) == len(l2) + 1
    eager_same = G.find_nams_batch(idx, s, o, l, mode=mode, graph=False)
    assert eager_same.nams.tobytes() == en.tobytes()
    m.close()


def test_cxx_facade_matches_oracle(setup, tmp_path):
    """A C++ caller of include/gis.hpp (per-read randstrobes_query ->
    find_nams -> rescue decision -> find_nams_rescue, one Batch, the pair
    step) gets the oracle's NAMs and pair results."""
    import os
    import subprocess
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 150, seed=21)
    reads = [r for r in reads if len(r) > 0 and b"\n" not in r]
    reads = reads[: len(reads) // 2 * 2]  # whole pairs for the pair step
    fa = tmp_path / "refs.fa"
    with open(fa, "wb") as fh:
        pos = 0
        for i, L in enumerate(lens):
            fh.write(b">contig%d description\n" % i)
            seq = g[pos: pos + int(L)].tobytes()
            for j in range(0, len(seq), 80):
                fh.write(seq[j: j + 80] + b"\n")
            pos += int(L)
    rd = tmp_path / "reads.txt"
    rd.write_bytes(b"\n".join(reads) + b"\n")
    exe = os.path.join(os.path.dirname(G.__file__), "_lib", "facade_check")
    out = str(tmp_path / "res")
    subprocess.run([exe, str(fa), str(rd), out], check=True, timeout=120)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 4)
    for suffix in ("single", "threads", "batch"):
        raw = open(out + "." + suffix, "rb").read()
        n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
        offs = np.frombuffer(raw[8: 8 + 8 * (n + 1)], dtype=np.uint64)
        nams = np.frombuffer(raw[8 + 8 * (n + 1):], dtype=G.NAM)
        assert n == len(reads)
        assert np.array_equal(offs, eo), suffix
        assert nams.tobytes() == en.tobytes(), suffix
    ep = O.pairs_batch(eo, en, len(reads) // 2)
    raw = open(out + ".pairs", "rb").read()
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    assert np.array_equal(np.frombuffer(raw[8: 8 + 8 * (n + 1)], dtype=np.uint64), eo)
    assert raw[8 + 8 * (n + 1):] == ep["ord"].tobytes()
    raw = open(out + ".types", "rb").read()
    npairs = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    assert raw[8: 8 + npairs] == ep["types"].tobytes()
    nj = int(np.frombuffer(raw[8 + npairs: 16 + npairs], dtype=np.uint64)[0])
    assert raw[16 + npairs:] == ep["joint"].tobytes() and nj == len(ep["joint"])


FUSED_CONFIGS = [dict(waves_per_block=w, length_split=sp, queue=q, chunk=ch, blocks_per_cu=bpc)
                 for w in (2, 4, 8) for sp in (0, 1)
                 for (q, ch, bpc) in ((0, 0, 0), (0, 0, 2), (1, 3, 2))]


def test_fused_configs_match_oracle(setup):
    """Every launch configuration of the fused form (waves per block, length
    split into a 160-base and a 320-base launch, static or dynamic work
    queue, persistent grid) gives the oracle's NAMs."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 600, seed=71)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    saved = idx.ctx.fused_config
    try:
        for cfg in FUSED_CONFIGS:
            idx.ctx.fused_config = cfg
            for graph in (False, True):
                res = G.find_nams_batch(idx, s, o, l, mode=G.MODE_FUSED, graph=graph)
                _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                              res.flags, efl)
    finally:
        idx.ctx.fused_config = saved


def test_fused_config_validation(setup):
    g, lens, p, idx, op, oidx = setup
    saved = idx.ctx.fused_config
    bad = [dict(waves_per_block=3), dict(length_split=2), dict(queue=1, chunk=0, blocks_per_cu=2),
           dict(queue=1, chunk=8, blocks_per_cu=0), dict(blocks_per_cu=65), dict(queue=2)]
    for b in bad:
        with pytest.raises(G.GisError):
            idx.ctx.fused_config = dict(saved, **b)
        assert idx.ctx.fused_config == saved


def test_autotune_keeps_results(setup):
    """The autotuner tries every candidate on the uploaded reads and keeps the
    fastest; NAMs stay the oracle's under the chosen configuration."""
    g, lens, p, idx, op, oidx = setup
    reads = _read_set(g, lens, 800, seed=72)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8)
    saved = idx.ctx.fused_config
    m = G.Mapper(idx, len(l), len(s))
    try:
        m.upload(s, o, l)
        best, ms, n = m.autotune()
        assert n >= 18 and ms > 0
        assert idx.ctx.fused_config == best
        for graph in (False, True):
            m.run(mode=G.MODE_FUSED, graph=graph)
            res = m.download()
            _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                          res.flags, efl)
    finally:
        m.close()
        idx.ctx.fused_config = saved


def _repeat_genome(rng):
    """Random sequence with a long tandem repeat (11 bp unit, 1% mutations)
    and 60 near-identical copies of a 1.5 kbp element on both contigs: seeds
    there have hundreds of hits, so rescued reads have many hits per
    (strand, ref) group and many NAMs open at once."""
    unit = rand_seq(rng, 11)
    tandem = bytearray((unit * 3000)[:30000])
    for i in range(len(tandem)):
        if rng.random() < 0.01:
            tandem[i] = b"ACGT"[rng.integers(4)]
    elem = rand_seq(rng, 1500)
    parts0, parts1 = [rand_seq(rng, 20000), bytes(tandem)], [rand_seq(rng, 20000)]
    for c in range(60):
        (parts0 if c % 2 else parts1).append(mutate(rng, elem, 0.002))
        (parts0 if c % 2 else parts1).append(rand_seq(rng, 700))
    return [b"".join(parts0), b"".join(parts1)]


def test_heavy_repeat_reads(ctx):
    """Reads inside a tandem repeat and a high-copy element: hit lists far
    beyond the merge's LDS budget (radix grouping) and groups with more than
    64 NAMs open at once (the wide list), in both forms."""
    rng = np.random.default_rng(81)
    seqs = _repeat_genome(rng)
    p = G.IndexParameters.from_read_length(150)
    refs = G.References(seqs)
    idx = G.StrobemerIndex(refs, p)
    idx.populate(f=0.0002, n_threads=4)
    oidx = O.Index.build(seqs, O.params_from_read_length(150), n_threads=4)
    reads = []
    for _ in range(300):
        c = int(rng.integers(2))
        st = int(rng.integers(0, len(seqs[c]) - 150))
        r = mutate(rng, seqs[c][st: st + 150], 0.005)
        reads.append(revcomp(r) if rng.integers(2) else r)
    s, o, l = G.pack_reads(reads)
    eo, en, ef, efl = O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 4)
    for mode in (G.MODE_PLAIN, G.MODE_FUSED):
        res = G.find_nams_batch(idx, s, o, l, mode=mode)
        _compare_nams((res.nam_offsets, res.nams), (eo, en), res.nonrepetitive_fraction, ef,
                      res.flags, efl)
    # the read set must reach the wide form: some group has more than 64
    # NAMs spanning one query position
    most = 0
    for r in range(len(reads)):
        nams = en[int(eo[r]): int(eo[r + 1])]
        for key in set(zip(nams["is_rc"].tolist(), nams["ref_id"].tolist())):
            g = nams[(nams["is_rc"] == key[0]) & (nams["ref_id"] == key[1])]
            for x in np.unique(g["query_start"]):
                most = max(most, int(((g["query_start"] <= x) & (g["query_end"] >= x)).sum()))
    assert most > 64, most
    assert (efl & 1).sum() > 10


def test_shared_workspace_batches(setup):
    """Batches that share one device workspace give their own NAMs, one run
    and download at a time; results overwritten by another batch's run are
    refused."""
    g, lens, p, idx, op, oidx = setup
    sets = [_read_set(g, lens, 200, seed=s) for s in (101, 102)]
    packs = [G.pack_reads(r) for r in sets]
    exp = [O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8) for s, o, l in packs]
    a = G.Mapper(idx, len(packs[0][2]), len(packs[0][0]))
    b = G.Mapper(idx, len(packs[1][2]), len(packs[1][0]), share=a)
    try:
        a.upload(*packs[0])
        b.upload(*packs[1])
        for mode in (G.MODE_PLAIN, G.MODE_FUSED):
            for m, e in ((a, exp[0]), (b, exp[1]), (a, exp[0])):
                m.run(mode=mode)
                res = m.download()
                _compare_nams((res.nam_offsets, res.nams), (e[0], e[1]), res.nonrepetitive_fraction, e[2],
                              res.flags, e[3])
        a.run()
        b.run()
        with pytest.raises(G.GisError):
            a.download()
    finally:
        b.close()
        a.close()


def test_overlapped_batches(setup):
    """Batches on two lanes (a stream and workspace each, a second batch
    sharing each lane) launched so that runs of both lanes are in flight at
    once give the oracle's NAMs, graph-replayed and eager, in both forms; a
    lane's workspace refuses a second launch, a run, an upload and a wait of
    another batch while its run is in flight."""
    g, lens, p, idx, op, oidx = setup
    sets = [_read_set(g, lens, 150, seed=s) for s in (111, 112, 113, 114)]
    packs = [G.pack_reads(r) for r in sets]
    exp = [O.map_batch(oidx, s, o, l, 2 * oidx.filter_cutoff, 1000, 8) for s, o, l in packs]
    size = (max(len(pk[2]) for pk in packs), max(len(pk[0]) for pk in packs))
    a0 = G.Mapper(idx, *size, own_stream=True)
    b0 = G.Mapper(idx, *size, own_stream=True)
    a1 = G.Mapper(idx, *size, share=a0)
    b1 = G.Mapper(idx, *size, share=b0)
    ms = [a0, b0, a1, b1]

    def check(i):
        res = ms[i].download()
        e = exp[i]
        _compare_nams((res.nam_offsets, res.nams), (e[0], e[1]), res.nonrepetitive_fraction, e[2],
                      res.flags, e[3])

    try:
        for m, pk in zip(ms, packs):
            m.upload(*pk)
        for mode in (G.MODE_PLAIN, G.MODE_FUSED):
            for graph in (True, False, True):
                a0.launch(mode=mode, graph=graph)
                b0.launch(mode=mode, graph=graph)
                a0.wait()
                check(0)
                a1.launch(mode=mode, graph=graph)
                b0.wait()
                check(1)
                b1.launch(mode=mode, graph=graph)
                a1.wait()
                check(2)
                b1.wait()
                check(3)
        a0.launch()
        for bad in (a1.launch, a0.launch, a0.run, a1.wait, lambda: a1.upload(*packs[2])):
            with pytest.raises(G.GisError):
                bad()
        a0.wait()
        check(0)
        with pytest.raises(G.GisError):
            a0.wait()
    finally:
        for m in reversed(ms):
            m.close()


def test_sti_rejects_damaged_files(setup, tmp_path):
    """A .sti whose bucket starts leave the table, or that is cut short, is
    refused on the host before anything reaches the GPU."""
    g, lens, p, idx, op, oidx = setup
    path = tmp_path / "ok.sti"
    idx.write(path)
    raw = bytearray(path.read_bytes())
    nb = (1 << oidx.bits) + 1
    bad = bytearray(raw)
    pos = len(bad) - 8 * nb + 8 * (nb // 2)
    bad[pos: pos + 8] = np.uint64(oidx.n + 5).tobytes()
    # a reserved chunk size that would wrap the read position around
    wrap = bytearray(raw)
    wrap[8:16] = np.uint64(2 ** 64 - 8).tobytes()
    # the table itself: two records out of order; a record whose top hash
    # bits leave its bucket (the lookups' assumptions, checked on the device)
    rec0 = 16 + int(np.frombuffer(raw[8:16], "<u8")[0]) + 36 + 8
    mid = oidx.n // 2
    swapped = bytearray(raw)
    a, b = rec0 + 16 * mid, rec0 + 16 * (mid + 1)
    swapped[a: a + 16], swapped[b: b + 16] = raw[b: b + 16], raw[a: a + 16]
    moved = bytearray(raw)
    last = rec0 + 16 * (oidx.n - 1)
    moved[last: last + 8] = np.uint64(int(oidx.randstrobes["hash"][0])).tobytes()
    other = G.StrobemerIndex(None, p)  # a context of its own
    assert oidx.randstrobes["hash"][mid] != oidx.randstrobes["hash"][mid + 1]
    for name, data in (("starts.sti", bad), ("short.sti", raw[: len(raw) // 2]), ("magic.sti", b"XX" + raw[2:]),
                       ("wrap.sti", wrap), ("unsorted.sti", swapped), ("bucket.sti", moved)):
        f = tmp_path / name
        f.write_bytes(bytes(data))
        with pytest.raises(G.GisError) as e:
            other.read(f)
        assert e.value.code == 2, name
    other.read(path)
    assert other.randstrobes.tobytes() == oidx.randstrobes.tobytes()
