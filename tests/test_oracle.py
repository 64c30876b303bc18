"""Pins the CPU oracle (oracle/oracle.c) against what fixes each function:
the xxhash library, hand-derived tie cases, textbook definitions, brute force
and closed forms. The reference ships no golden vectors for this path and its
CPU units cannot be compiled here without stand-in headers, so these checks
are the oracle's anchor."""
import struct

import numpy as np
import pytest
import xxhash

from oracle import oracle_lib as O
from tests import spec
from tests.helpers import HAND_NAMS, hand_case, n_reads, rand_seq, revcomp, tie_heavy_reads, tiny_index_arrays

P150 = None


def p150():
    global P150
    if P150 is None:
        P150 = O.params_from_read_length(150)
    return P150


# ---------------------------------------------------------------- hashing

def test_xxh64_matches_library():
    rng = np.random.default_rng(0)
    vals = list(rng.integers(0, 2 ** 63, 2000, dtype=np.uint64) * 2 + 1) + [0, 1, 2 ** 64 - 1, 2 ** 32]
    for v in vals:
        v = int(v)
        assert O.xxh64(v) == xxhash.xxh64_intdigest(struct.pack("<Q", v), seed=0), v


def _xxh64_inverse(y):
    """Inverse of the single-word xxh64 (each step is a bijection of u64)."""
    M = 2 ** 64
    P1, P2, P3, P4, P5 = (0x9E3779B185EBCA87, 0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9,
                          0x85EBCA77C2B2AE63, 0x27D4EB2F165667C5)

    def unxorshift(x, s):
        r = x
        for _ in range(64 // s + 1):
            r = x ^ (r >> s)
        return r

    def rotr(x, r):
        return ((x >> r) | (x << (64 - r))) % M

    x = unxorshift(y, 32)
    x = x * pow(P3, -1, M) % M
    x = unxorshift(x, 29)
    x = x * pow(P2, -1, M) % M
    x = unxorshift(x, 33)
    x = (x - P4) * pow(P1, -1, M) % M
    x = rotr(x, 27)
    x ^= (P5 + 8) % M
    x = x * pow(P1, -1, M) % M
    x = rotr(x, 31)
    return x * pow(P2, -1, M) % M


def test_xxh64_inverse_and_all_ones_preimage():
    rng = np.random.default_rng(1)
    for v in rng.integers(0, 2 ** 63, 200, dtype=np.uint64):
        assert _xxh64_inverse(O.xxh64(int(v))) == int(v)
    # the only u64 hashing to ~0 is not a 2s-bit code for any s <= 31, so an
    # s-mer hash never equals ~0 (the NONE window state is unreachable)
    pre = _xxh64_inverse(2 ** 64 - 1)
    assert O.xxh64(pre) == 2 ** 64 - 1
    assert pre >= 4 ** 31


def test_nt4_table():
    for c in range(256):
        exp = {65: 0, 97: 0, 67: 1, 99: 1, 71: 2, 103: 2, 84: 3, 116: 3, 85: 3, 117: 3}.get(c, 4)
        assert O.nt4(c) == exp


# ---------------------------------------------------------------- parameters

# (read length) -> (canonical r, k, s, t, l, u, q, max_dist, w_min, w_max), worked
# out from the read-length profiles (step2/indexparameters.cpp:37-45)
PROFILE_CASES = {
    1: (50, 20, 16, 3, -3, 2, 255, 20, 1, 6), 50: (50, 20, 16, 3, -3, 2, 255, 20, 1, 6),
    90: (50, 20, 16, 3, -3, 2, 255, 20, 1, 6), 91: (100, 20, 16, 3, -2, 2, 255, 30, 2, 6),
    110: (100, 20, 16, 3, -2, 2, 255, 30, 2, 6), 111: (125, 20, 16, 3, -1, 4, 255, 55, 3, 8),
    135: (125, 20, 16, 3, -1, 4, 255, 55, 3, 8), 136: (150, 20, 16, 3, 1, 7, 255, 80, 5, 11),
    150: (150, 20, 16, 3, 1, 7, 255, 80, 5, 11), 175: (150, 20, 16, 3, 1, 7, 255, 80, 5, 11),
    176: (250, 20, 16, 3, 4, 13, 255, 180, 8, 17), 275: (250, 20, 16, 3, 4, 13, 255, 180, 8, 17),
    276: (300, 22, 18, 3, 2, 12, 255, 230, 6, 16), 375: (300, 22, 18, 3, 2, 12, 255, 230, 6, 16),
    376: (400, 23, 17, 4, 2, 12, 255, 255, 5, 15), 10000: (400, 23, 17, 4, 2, 12, 255, 255, 5, 15),
}


@pytest.mark.parametrize("r", sorted(PROFILE_CASES))
def test_params_profiles(r):
    assert O.params_from_read_length(r).as_tuple() == PROFILE_CASES[r]


def test_params_overrides_and_errors():
    p = O.params_from_read_length(150, k=24, s=18, c=4, max_seed_len=100)
    # k - s + 1 = 7, 24 // 7 = 3: w_min = 3 + 1, w_max = 3 + 7; max_dist = 100 - 24; q = 2^4 - 1
    assert p.as_tuple() == (150, 24, 18, 4, 1, 7, 15, 76, 4, 10)
    with pytest.raises(ValueError):
        O.params_from_read_length(150, k=7)
    with pytest.raises(ValueError):
        O.params_from_read_length(150, k=33)
    with pytest.raises(ValueError):
        O.params_from_read_length(150, k=20, s=17)   # k - s odd
    with pytest.raises(ValueError):
        O.params_from_read_length(150, s=21)         # s > k
    with pytest.raises(ValueError):
        O.params_from_read_length(150, max_seed_len=300)  # max_dist 280 > 255
    with pytest.raises(ValueError):
        O.params_from_read_length(150, l=5, u=1)     # w_min > w_max


# ---------------------------------------------------------------- syncmers

def test_syncmers_textbook_on_unique_minima():
    rng = np.random.default_rng(2)
    p = p150()
    for _ in range(40):
        seq = rand_seq(rng, int(rng.integers(20, 400)), gc=float(rng.uniform(0.3, 0.7)))
        got = {int(x["position"]) for x in O.syncmers(seq, p)}
        ref = spec.unique_min_syncmers(seq, p.k, p.s)
        assert ref, "random reads have tie-free windows"
        for pos, is_sync in ref.items():
            assert (pos in got) == is_sync, pos


def test_syncmer_hashes_are_canonical_kmer_hashes():
    rng = np.random.default_rng(3)
    p = p150()
    for _ in range(20):
        seq = rand_seq(rng, 200)
        for x in O.syncmers(seq, p):
            pos = int(x["position"])
            kmer = seq[pos: pos + p.k]
            codes = [spec.code(c) for c in kmer]
            assert int(x["hash"]) == spec.xxh(spec.canonical(codes, 0, p.k))
        # canonical: the reverse complement has the same syncmer k-mers
        fw = sorted(int(x["hash"]) for x in O.syncmers(seq, p))
        rc = sorted(int(x["hash"]) for x in O.syncmers(revcomp(seq), p))
        common = set(fw) & set(rc)
        assert len(common) >= 0.9 * len(fw)


def test_homopolymer_tie_rule():
    """All s-mers equal: first window leftmost (offset 0), then the minimum is
    popped every 5 windows and re-found rightmost (offset 4), so the t-1 = 2
    offset recurs at windows 3, 8, 13, ..."""
    p = p150()
    for L in (20, 23, 24, 60, 150):
        got = [int(x["position"]) for x in O.syncmers(b"A" * L, p)]
        assert got == list(range(3, L - p.k + 1, 5)), L
        assert all(int(x["hash"]) == spec.xxh(0) for x in O.syncmers(b"A" * L, p))


def test_dinucleotide_tie_rule():
    """ACAC...: s-mer hashes alternate hA (even starts) / hC (odd starts).
    Worked through the rules: if hA < hC the syncmers are 2, 6, 10, ...; else
    3, 7, 11, ..."""
    p = p150()
    seq = (b"AC" * 100)[:150]
    _, h = spec.smer_hashes(seq, p.s)
    hA, hC = h[0], h[1]
    exp = list(range(2, 150 - p.k + 1, 4)) if hA < hC else list(range(3, 150 - p.k + 1, 4))
    assert [int(x["position"]) for x in O.syncmers(seq, p)] == exp


def test_syncmers_tie_heavy_and_n_against_window_model():
    rng = np.random.default_rng(4)
    p = p150()
    reads = tie_heavy_reads(rng) + n_reads(rng)
    for seq in reads:
        got = [(int(x["hash"]), int(x["position"])) for x in O.syncmers(seq, p)]
        assert got == spec.syncmers(seq, p.k, p.s), seq[:40]
    # other parameter sets (w = 3, 7, 1)
    for r, kw in ((300, {}), (400, {}), (150, {"k": 20, "s": 20})):
        pp = O.params_from_read_length(r, **kw)
        for seq in reads[:15] + [rand_seq(rng, 300)]:
            got = [(int(x["hash"]), int(x["position"])) for x in O.syncmers(seq, pp)]
            assert got == spec.syncmers(seq, pp.k, pp.s)


def test_n_resets():
    p = p150()
    rng = np.random.default_rng(5)
    a, b = rand_seq(rng, 100), rand_seq(rng, 100)
    joined = O.syncmers(a + b"N" + b, p)
    left = O.syncmers(a, p)
    right = O.syncmers(b, p)
    exp = [(int(x["hash"]), int(x["position"])) for x in left] + \
          [(int(x["hash"]), int(x["position"]) + 101) for x in right]
    assert [(int(x["hash"]), int(x["position"])) for x in joined] == exp


# ---------------------------------------------------------------- randstrobes

def _check_partner_choice(syn, rs_list, p, L, reverse):
    """Each randstrobe pairs strobe1 with the candidate of minimal
    popcount((h1 ^ h2) & q); the earliest such candidate wins; strobe1 itself
    when no candidate lies within max_dist."""
    if reverse:
        syn = [(h, L - pos - p.k) for h, pos in reversed(syn)]
    n = len(syn)
    assert len(rs_list) == max(0, n - p.w_min)
    for i, r in enumerate(rs_list):
        h1, p1 = syn[i]
        cands = [j for j in range(i + p.w_min, min(i + p.w_max, n - 1) + 1)]
        ok = []
        for j in cands:
            if syn[j][1] > p1 + p.max_dist:
                break
            ok.append(j)
        if not ok:
            j = i
        else:
            best = min(bin((h1 ^ syn[j][0]) & p.q).count("1") for j in ok)
            j = next(j for j in ok if bin((h1 ^ syn[j][0]) & p.q).count("1") == best)
        assert int(r["hash"]) == (h1 + syn[j][0]) % 2 ** 64
        assert int(r["start"]) == p1
        assert int(r["end"]) == syn[j][1] + p.k
        assert int(r["is_reverse"]) == int(reverse)


def test_randstrobes_query_partner_rule():
    rng = np.random.default_rng(6)
    for r in (150, 50, 300):
        p = O.params_from_read_length(r)
        reads = [rand_seq(rng, int(rng.integers(0, 350))) for _ in range(25)] + tie_heavy_reads(rng)
        for seq in reads:
            rs = O.randstrobes_query(seq, p)
            if len(seq) < p.w_max:
                assert len(rs) == 0
                continue
            syn = spec.syncmers(seq, p.k, p.s)
            if not syn:
                assert len(rs) == 0
                continue
            nf = max(0, len(syn) - p.w_min)
            _check_partner_choice(syn, rs[:nf], p, len(seq), False)
            _check_partner_choice(syn, rs[nf:], p, len(seq), True)


# ---------------------------------------------------------------- index

def _spec_index(contigs, p, bits, f):
    ents = []
    for ref, seq in enumerate(contigs):
        syn = spec.syncmers(seq, p.k, p.s)
        if len(seq) < p.w_max:
            continue
        n = len(syn)
        for i in range(max(0, n - p.w_min)):
            cands, ok = range(i + p.w_min, min(i + p.w_max, n - 1) + 1), []
            for j in cands:
                if syn[j][1] > syn[i][1] + p.max_dist:
                    break
                ok.append(j)
            j = i
            if ok:
                best = min(bin((syn[i][0] ^ syn[j][0]) & p.q).count("1") for j in ok)
                j = next(j for j in ok if bin((syn[i][0] ^ syn[j][0]) & p.q).count("1") == best)
            ents.append(((syn[i][0] + syn[j][0]) % 2 ** 64, syn[i][1], (ref << 8) + syn[j][1] - syn[i][1]))
    ents.sort()
    arr = np.array(ents, dtype=[("hash", "<u8"), ("position", "<u4"), ("packed", "<u4")])
    h = arr["hash"]
    # bucket b starts at the first entry whose top bits are >= b, except that
    # the very first hash's run is skipped (starts are only recorded when the
    # hash changes)
    first_run = int(np.searchsorted(h, h[0], side="right")) if len(h) else 0
    top = (h >> np.uint64(64 - bits)).astype(np.int64)
    starts = np.searchsorted(top[first_run:], np.arange((1 << bits) + 1), side="left") + first_run
    starts[-1] = len(h)
    _, counts = np.unique(h, return_counts=True)
    unique = len(counts)
    rep = np.sort(counts[counts > 1])[::-1]
    cut_i = int(np.float32(unique) * np.float32(f))
    cutoff = 30 if len(rep) == 0 else int(min(100, max(30, rep[cut_i] if cut_i < len(rep) else rep[-1])))
    return arr, starts.astype(np.uint64), cutoff


def test_index_build_against_spec():
    rng = np.random.default_rng(7)
    unit = rand_seq(rng, 400)
    contigs = [rand_seq(rng, 3000) + unit + rand_seq(rng, 500) + unit,
               rand_seq(rng, 20) + b"N" + rand_seq(rng, 2000),
               unit + rand_seq(rng, 10) + (b"AC" * 200) + rand_seq(rng, 700), b"ACGT" * 3]
    p = p150()
    for bits, f in ((8, 0.0002), (10, 0.1), (9, 0.5)):
        idx = O.Index.build(contigs, p, bits=bits, f=f, n_threads=2)
        arr, starts, cutoff = _spec_index(contigs, p, bits, f)
        assert idx.randstrobes.tobytes() == arr.tobytes()
        assert np.array_equal(idx.starts, starts)
        assert idx.filter_cutoff == cutoff


def test_pick_bits_and_determinism():
    rng = np.random.default_rng(8)
    contigs = [rand_seq(rng, 50_000) for _ in range(3)]
    p = p150()
    a = O.Index.build(contigs, p, n_threads=1)
    b = O.Index.build(contigs, p, n_threads=4)
    assert a.randstrobes.tobytes() == b.randstrobes.tobytes()
    # bits = int(log2(150000 / 5)) - 1 = 13
    assert a.bits == 13
    assert O.lib().orc_pick_bits(10, O.C.byref(p)) == 8
    assert O.lib().orc_pick_bits(5 * 2 ** 40, O.C.byref(p)) == 31


def test_find_count_filter_against_numpy():
    rng = np.random.default_rng(9)
    unit = rand_seq(rng, 300)
    contigs = [rand_seq(rng, 20000) + unit * 6 + rand_seq(rng, 5000)]
    p = p150()
    idx = O.Index.build(contigs, p, bits=10)
    h = idx.randstrobes["hash"]
    first_run_end = int(np.searchsorted(h, h[0], side="right"))
    keys = np.concatenate([h, h + np.uint64(1), rng.integers(0, 2 ** 63, 500, dtype=np.uint64)])
    for key in keys:
        key = int(key)
        lo = int(np.searchsorted(h, key, side="left"))
        exp = lo if lo < len(h) and int(h[lo]) == key and lo >= first_run_end else O.NOT_FOUND
        got = idx.find(key)
        assert got == exp, key
        if got != O.NOT_FOUND:
            hi = int(np.searchsorted(h, key, side="right"))
            assert idx.get_count(got) == hi - lo
            nxt = int(h[got + idx.filter_cutoff]) if got + idx.filter_cutoff < len(h) else O.NOT_FOUND
            assert idx.is_filtered(got) == (nxt == key)
    assert idx.find(int(h[0])) == O.NOT_FOUND  # the first hash is never findable


def _lookup_bytes_model(h, starts, bits, cutoff, keys):
    """The byte model of the lookup stage from the sorted hash column alone:
    16 for the bucket bounds, 16 per record a linear scan of the bucket reads
    up to the first hash >= key (all records when none is), 8 for a found
    seed's filter probe, and for an unfiltered seed 16 per run record past
    the first plus 8 for the hash that ends the run."""
    total = 0
    for key in keys:
        key = int(key)
        b = key >> (64 - bits)
        s, e = int(starts[b]), int(starts[b + 1])
        lt = int(np.searchsorted(h[s:e], np.uint64(key), side="left"))
        total += 16 + 16 * (lt + 1 if lt < e - s else e - s)
        if lt == e - s or int(h[s + lt]) != key:
            continue
        pos = s + lt
        total += 8
        probe = int(h[pos + cutoff]) if pos + cutoff < len(h) else 2 ** 64 - 1
        if probe == key:
            continue
        run = int(np.searchsorted(h, np.uint64(key), side="right")) - pos
        total += 16 * (run - 1) + 8
    return total


def test_lookup_bytes_against_model():
    """orc_lookup_bytes (the algorithmic bytes of the lookup roofline) against
    the byte model computed from numpy searches, on buckets of one to dozens
    of records, filtered and unfiltered runs, absent keys and the first hash
    (whose bucket never holds it)."""
    rng = np.random.default_rng(12)
    unit = rand_seq(rng, 300)
    contigs = [rand_seq(rng, 30000) + unit * 8 + rand_seq(rng, 4000) + unit * 3]
    p = p150()
    for bits in (8, 12):
        idx = O.Index.build(contigs, p, bits=bits)
        h = idx.randstrobes["hash"]
        keys = np.concatenate([h[::3], h[:4], h[::7] + np.uint64(1),
                               rng.integers(0, 2 ** 63, 400, dtype=np.uint64) * 2 + 1])
        qrs = np.zeros(len(keys), dtype=O.QUERY_RANDSTROBE)
        qrs["hash"] = keys
        exp = _lookup_bytes_model(h, idx.starts, idx.bits, idx.filter_cutoff, keys)
        assert O.lookup_bytes(idx, qrs) == exp, bits
        idx.free()


# ---------------------------------------------------------------- merging

def _random_hits(rng, n, n_refs=3):
    hits, refs = [], []
    qs_vals = sorted(rng.choice(np.arange(0, 300), size=min(n, 120), replace=False))
    for qs in qs_vals:
        span = int(rng.integers(20, 60))
        for _ in range(int(rng.integers(1, 4))):
            ref = int(rng.integers(0, n_refs))
            diag = int(rng.choice([1000, 1003, 5000, int(rng.integers(0, 9000))]))
            rs = diag + int(qs)
            hits.append((int(qs), int(qs) + span, rs, rs + span + int(rng.integers(-3, 4))))
            refs.append(ref)
    arr = np.array(hits, dtype=O.HIT)
    # one entry per (ref, query_start, ref_start)
    _, keep = np.unique(np.stack([refs, arr["query_start"], arr["ref_start"]]), axis=1, return_index=True)
    keep = np.sort(keep)
    return np.array(refs, dtype=np.int32)[keep], arr[keep]


def test_hit_outer_and_nam_outer_merges_agree():
    rng = np.random.default_rng(10)
    for trial in range(300):
        refs, hits = _random_hits(rng, int(rng.integers(1, 80)))
        perm = rng.permutation(len(hits))
        for is_rc in (0, 1):
            a = O.canonicalize(O.merge_hits(refs[perm], hits[perm], 20, is_rc, 0))
            b = O.canonicalize(O.merge_hits(refs, hits, 20, is_rc, 1))
            assert a.tobytes() == b.tobytes(), trial


def test_merge_against_spec_model():
    """The oracle's hit-outer merge against the model written from the
    reference's merge rules (tests/spec.py), over random hit sets with
    crowded diagonals (extend, contain and reopen all occur) and flushes."""
    rng = np.random.default_rng(13)
    multi = 0
    for trial in range(400):
        refs, hits = _random_hits(rng, int(rng.integers(1, 90)), n_refs=int(rng.integers(1, 4)))
        perm = rng.permutation(len(hits))
        for is_rc in (0, 1):
            got = O.canonicalize(O.merge_hits(refs[perm], hits[perm], 20, is_rc, 0))
            exp = spec.canonical_nams(spec.merge_hits_into_nams(refs, hits, 20, is_rc))
            assert len(got) == len(exp), trial
            for g, e in zip(got, exp):
                gt = (int(g["nam_id"]), int(g["query_start"]), int(g["query_end"]),
                      int(g["query_prev_hit_startpos"]), int(g["ref_start"]), int(g["ref_end"]),
                      int(g["ref_prev_hit_startpos"]), int(g["n_hits"]), int(g["ref_id"]),
                      np.float32(g["score"]), int(g["is_rc"]))
                assert gt == e, (trial, gt, e)
                multi += int(g["n_hits"]) > 1
    assert multi > 100, "the hit sets must make multi-hit NAMs"


def test_merge_worked_example():
    # three chained hits on one diagonal, one contained hit, one unrelated
    hits = np.array([(0, 30, 100, 130), (10, 40, 110, 140), (20, 50, 120, 150),
                     (22, 45, 122, 145), (5, 35, 905, 935)], dtype=O.HIT)
    refs = np.zeros(5, dtype=np.int32)
    nams = O.canonicalize(O.merge_hits(refs, hits, 20, 0, 0))
    assert len(nams) == 2
    a, b = nams
    # chain: query 0..50, ref 100..150, 4 hits (the contained one counts)
    assert (a["query_start"], a["query_end"], a["ref_start"], a["ref_end"], a["n_hits"]) == (0, 50, 100, 150, 4)
    assert (a["query_prev_hit_startpos"], a["ref_prev_hit_startpos"]) == (22, 122)
    assert a["score"] == np.float32(4 * (2 * 50 - 50))
    assert (b["query_start"], b["ref_start"], b["n_hits"], b["score"]) == (5, 905, 1, np.float32(30))
    assert (a["nam_id"], b["nam_id"]) == (0, 1)


def test_score_degenerate_spans():
    # query span 20, ref span 60: 2*20 - 60 < 0 -> score 1
    hits = np.array([(0, 20, 0, 60)], dtype=O.HIT)
    nams = O.merge_hits(np.zeros(1, np.int32), hits, 20, 1, 0)
    assert nams[0]["score"] == np.float32(1.0) and nams[0]["is_rc"] == 1


# ---------------------------------------------------------------- whole read

def test_error_free_reads_find_their_origin():
    rng = np.random.default_rng(11)
    genome = rand_seq(rng, 200_000)
    p = p150()
    idx = O.Index.build([genome], p, n_threads=2)
    for _ in range(100):
        pos = int(rng.integers(0, len(genome) - 150))
        read = genome[pos: pos + 150]
        rc = bool(rng.integers(2))
        if rc:
            read = revcomp(read)
        nams, frac, flags = O.map_read(idx, read, 2 * idx.filter_cutoff)
        best = nams[np.argmax(nams["score"])]
        assert int(best["is_rc"]) == int(rc)
        assert int(best["ref_start"]) - int(best["query_start"]) == pos
        assert frac == 1.0 and flags == 0
        qrs = O.randstrobes_query(read, p)
        strand_n = int((qrs["is_reverse"] == int(rc)).sum())
        # every randstrobe whose partner window is not cut by the read's end
        # equals the genome's randstrobe there and joins the NAM
        assert strand_n - (p.w_max - p.w_min + 1) <= int(best["n_hits"]) <= strand_n


def _tiny_index(entries, filter_cutoff, bits=8):
    arr, starts = tiny_index_arrays(entries, bits)
    return O.Index.wrap(arr, starts, bits, filter_cutoff, p150())


def _qrs(hashes, strand=0):
    q = np.zeros(len(hashes), dtype=O.QUERY_RANDSTROBE)
    q["hash"] = hashes
    q["start"] = np.arange(len(hashes)) * 3
    q["end"] = q["start"] + 30
    q["is_reverse"] = strand
    return q


def test_find_nams_composition_worked_example():
    ents, q = hand_case()
    idx = _tiny_index(ents, filter_cutoff=5, bits=8)
    frac, nams, nh = O.find_nams(idx, q, canonical=True)
    assert np.float32(frac) == np.float32(3) / np.float32(4)
    got = [(int(n["is_rc"]), int(n["query_start"]), int(n["query_end"]), int(n["ref_start"]), int(n["ref_end"]),
            int(n["n_hits"]), float(n["score"])) for n in nams]
    assert got == HAND_NAMS
    assert [int(n["nam_id"]) for n in nams] == list(range(5))
    assert all(int(n["ref_id"]) == 0 for n in nams)


@pytest.mark.parametrize("n_good,n_rep,rescued", [(7, 3, True), (8, 2, False), (6, 4, True),
                                                  (10, 0, False), (0, 0, True)])
def test_rescue_decision_at_point_seven(n_good, n_rep, rescued):
    """A read whose randstrobes hit n_good unique and n_rep repetitive index
    entries: frac = good/total in float, compared with 0.7 in double, so 7/10
    (0.699999988) is rescued; no hits at all is rescued too."""
    rng = np.random.default_rng(12)
    p = p150()
    while True:
        read = rand_seq(rng, 150)
        qrs = O.randstrobes_query(read, p)
        hs, cnt = np.unique(qrs["hash"], return_counts=True)
        single = [int(h) for h, c in zip(hs, cnt) if c == 1]
        if len(single) >= 10:
            break
    chosen = [single[i] for i in rng.permutation(len(single))[: n_good + n_rep]]
    ents = []
    for i, hh in enumerate(chosen[:n_good]):
        ents.append((hh, 1000 * i, 10))
    for i, hh in enumerate(chosen[n_good:]):
        for c in range(4):          # 4 copies > filter_cutoff 2: filtered
            ents.append((hh, 50000 + 1000 * i + c, 10))
    idx = _tiny_index(ents, filter_cutoff=2)
    frac, nams, nh = O.find_nams(idx, qrs)
    total = n_good + n_rep
    assert np.float32(frac) == (np.float32(n_good) / np.float32(total) if total else np.float32(1.0))
    _, frac2, flags = O.map_read(idx, read, 2 * idx.filter_cutoff)
    assert np.float32(frac2) == np.float32(frac)
    assert bool(flags & O.FLAG_RESCUED) == rescued


def test_rescue_takes_the_documented_prefix():
    """find_nams_rescue keeps seeds in (count, query_start) order while
    count <= hard_cutoff and (count <= rescue_cutoff or fewer than 5 taken)."""
    counts = [1, 2, 40, 40, 70, 90, 120, 300, 900, 1500]
    rng = np.random.default_rng(13)
    hashes = [int(x) for x in rng.integers(2 ** 40, 2 ** 62, len(counts), dtype=np.uint64)]
    ents = []
    for i, (hh, c) in enumerate(zip(hashes, counts)):
        for j in range(c):
            ents.append((hh, 10_000_000 * i + 100 * j, 10))
    idx = _tiny_index(ents, filter_cutoff=30, bits=10)
    order = rng.permutation(len(counts))
    q = _qrs([hashes[i] for i in order])
    for rescue_cutoff, hard in ((60, 1000), (10, 1000), (500, 1000), (60, 100)):
        nams = O.find_nams_rescue(idx, q, rescue_cutoff, hard)
        taken = set()
        srt = sorted(range(len(counts)), key=lambda i: (counts[order[i]], int(q["start"][i])))
        for rank, i in enumerate(srt):
            c = counts[order[i]]
            if (c > rescue_cutoff and rank >= 5) or c > hard:
                break
            taken.add(int(q["start"][i]))
        # every hit is its own NAM here (different ref positions 100 apart,
        # query spans 30 < 100): one NAM per kept run entry along the min_diff rule
        got_qs = set(int(x) for x in nams["query_start"])
        assert got_qs == taken, (rescue_cutoff, hard)
