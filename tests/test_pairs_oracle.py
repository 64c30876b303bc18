"""Pins the oracle's NAM consumer for paired reads (oracle/oracle.c:
orc_sort_nams_reference, orc_top_dropoff, orc_get_mapq,
orc_is_proper_nam_pair, orc_pair_type, orc_best_scoring_nam_pairs) against
what fixes each function: Python's own sort on the comparators' key tuples,
closed forms worked by hand, brute-force enumeration of proper pairs, and a
hand-traced quicksort for the order of equal entries."""
import math

import numpy as np
import pytest

from oracle import oracle_lib as O


def nam(qs, qe, rs, re, n_hits, ref_id=0, is_rc=0, score=None, nam_id=0):
    a = np.zeros(1, dtype=O.NAM)
    a["nam_id"], a["query_start"], a["query_end"] = nam_id, qs, qe
    a["ref_start"], a["ref_end"], a["n_hits"] = rs, re, n_hits
    a["ref_id"], a["is_rc"] = ref_id, is_rc
    a["score"] = float(n_hits * 10 if score is None else score)
    return a[0]


def nams(*xs):
    return np.array(list(xs), dtype=O.NAM)


def rand_list(rng, n, n_refs=3):
    """n NAMs with distinct (is_rc, ref_id, query_start, ref_start), as
    find_nams produces them, with many equal secondary fields."""
    seen, out = set(), []
    while len(out) < n:
        is_rc, ref = int(rng.integers(0, 2)), int(rng.integers(0, n_refs))
        qs, rs = int(rng.integers(0, 40)), int(rng.integers(0, 2000))
        if (is_rc, ref, qs, rs) in seen:
            continue
        seen.add((is_rc, ref, qs, rs))
        qe = qs + int(rng.integers(20, 60))
        re = rs + int(rng.integers(20, 60))
        nh = int(rng.integers(1, 6))
        out.append(nam(qs, qe, rs, re, nh, ref, is_rc, score=nh * 7.5, nam_id=len(out)))
    return nams(*out) if out else np.zeros(0, dtype=O.NAM)


# ------------------------------------------------------------ reference order

def _key1(n):
    return (int(n["is_rc"]), int(n["ref_id"]), -int(n["n_hits"]), int(n["query_end"]),
            int(n["query_start"]), int(n["ref_end"]), int(n["ref_start"]))


def _key2(n):
    d = max(0, int(n["ref_start"]) - int(n["query_start"]))
    return (int(n["is_rc"]), int(n["ref_id"]), d, -int(n["n_hits"]), int(n["query_end"]),
            int(n["query_start"]), int(n["ref_end"]), int(n["ref_start"]))


@pytest.mark.parametrize("mate,key", [(1, _key1), (2, _key2)])
def test_reference_order_is_python_sort_on_the_comparator_keys(mate, key):
    rng = np.random.default_rng(11 + mate)
    for n in list(range(0, 12)) + [40, 200]:
        for _ in range(3):
            a = rand_list(rng, n)
            got = O.sort_nams_reference(a, mate)
            want = sorted(a, key=key)
            assert [int(x["nam_id"]) for x in got] == [int(x["nam_id"]) for x in want]


def test_mate2_order_uses_the_diagonal_and_mate1_does_not():
    # same strand and reference: mate 1 sorts by n_hits first, mate 2 by
    # max(0, ref_start - query_start) first
    a = nam(0, 30, 500, 530, 2, nam_id=0)      # diagonal 500, 2 hits
    b = nam(10, 40, 100, 130, 5, nam_id=1)     # diagonal 90, 5 hits
    c = nam(50, 80, 20, 50, 3, nam_id=2)       # diagonal max(0, -30) = 0
    assert [x["nam_id"] for x in O.sort_nams_reference(nams(a, b, c), 1)] == [1, 2, 0]
    assert [x["nam_id"] for x in O.sort_nams_reference(nams(a, b, c), 2)] == [2, 1, 0]


# ------------------------------------------------------------ dropoff and MAPQ

def test_top_dropoff_closed_forms():
    assert O.top_dropoff(nams(nam(0, 1, 0, 1, 2), nam(0, 1, 5, 6, 9))) == 1.0     # n_max.n_hits <= 2
    assert O.top_dropoff(nams(nam(0, 1, 0, 1, 10))) == 0.0                       # one NAM
    assert O.top_dropoff(nams(nam(0, 1, 0, 1, 10), nam(0, 1, 5, 6, 7))) == np.float32(0.7)
    # the list's first NAM is the denominator even when it is not the largest
    assert O.top_dropoff(nams(nam(0, 1, 0, 1, 4), nam(0, 1, 5, 6, 10))) == np.float32(2.5)


def test_get_mapq_closed_forms():
    one = nams(nam(0, 30, 0, 30, 4, score=20.0))
    assert O.get_mapq(one) == 60                       # a single NAM
    assert O.get_mapq(nams()) == 60
    # 40 (1 - 15/20) min(4/10, 1) ln 20 = 11.98 -> 11
    l = nams(nam(0, 30, 0, 30, 4, score=20.0), nam(0, 30, 90, 120, 2, score=15.0))
    assert O.get_mapq(l) == 11
    assert int(40 * 0.25 * 0.4 * math.log(20)) == 11
    # capped at 60: 40 (1 - 50/100) * 1 * ln 100 = 92.1
    l = nams(nam(0, 30, 0, 30, 20, score=100.0), nam(0, 30, 90, 120, 2, score=50.0))
    assert O.get_mapq(l) == 60
    # equal scores -> 0
    l = nams(nam(0, 30, 0, 30, 20, score=100.0), nam(0, 30, 90, 120, 2, score=100.0))
    assert O.get_mapq(l) == 0
    # a second score above the first gives -92, which the uint8 return wraps to 164
    l = nams(nam(0, 30, 0, 30, 10, score=10.0), nam(0, 30, 90, 120, 2, score=20.0))
    assert int(40 * (1 - 2.0) * 1.0 * math.log(10)) == -92
    assert O.get_mapq(l) == 256 - 92


# ------------------------------------------------------------ proper pairs, types

def test_is_proper_nam_pair_boundaries():
    # r1 ---> <--- r2: nam2 reverse, a <= b, b - a < mu + 10 sigma = 1300
    n1 = nam(10, 40, 1010, 1040, 3, ref_id=2, is_rc=0)          # a = 1000
    ok = nam(5, 35, 2304, 2334, 3, ref_id=2, is_rc=1)           # b = 2299
    edge = nam(5, 35, 2305, 2335, 3, ref_id=2, is_rc=1)         # b - a = 1300
    assert O.is_proper_nam_pair(n1, ok)
    assert not O.is_proper_nam_pair(n1, edge)
    assert O.is_proper_nam_pair(n1, edge, mu=301.0)             # 1301 > 1300
    assert not O.is_proper_nam_pair(n1, nam(5, 35, 2304, 2334, 3, ref_id=1, is_rc=1))  # other contig
    assert not O.is_proper_nam_pair(n1, nam(5, 35, 2304, 2334, 3, ref_id=2, is_rc=0))  # same strand
    assert not O.is_proper_nam_pair(n1, nam(5, 35, 900, 930, 3, ref_id=2, is_rc=1))    # b < a
    # r2 ---> <--- r1: nam1 reverse, b <= a
    assert O.is_proper_nam_pair(nam(0, 30, 1500, 1530, 3, 2, 1), nam(0, 30, 1000, 1030, 3, 2, 0))
    # diagonal clamped at 0: a = max(0, 5 - 40) = 0
    assert O.is_proper_nam_pair(nam(40, 70, 5, 35, 3, 0, 0), nam(0, 30, 100, 130, 3, 0, 1))


def test_pair_types():
    good1 = nam(10, 40, 1010, 1040, 9, ref_id=2, is_rc=0)
    good2 = nam(5, 35, 1300, 1330, 9, ref_id=2, is_rc=1)
    weak = nam(0, 30, 50000, 50030, 3, ref_id=2, is_rc=0)
    assert O.pair_type(nams(), nams()) == 0
    assert O.pair_type(nams(good1), nams()) == 1
    assert O.pair_type(nams(), nams(good2)) == 2
    assert O.pair_type(nams(good1), nams(good2)) == 3
    assert O.pair_type(nams(good1, weak), nams(good2)) == 3       # 3/9 < 0.5
    strong = nam(0, 30, 70000, 70030, 5, ref_id=2, is_rc=0)
    assert O.pair_type(nams(good1, strong), nams(good2)) == 4     # 5/9 >= 0.5
    far = nam(5, 35, 9000, 9030, 9, ref_id=2, is_rc=1)
    assert O.pair_type(nams(good1), nams(far)) == 4               # not a proper pair
    two = nam(10, 40, 1010, 1040, 2, ref_id=2, is_rc=0)
    assert O.pair_type(nams(two), nams(good2)) == 4               # n_hits <= 2: dropoff 1.0


# ------------------------------------------------------------ joint NAM pairs

def _proper(a, b):
    return O.is_proper_nam_pair(a, b)


def test_joint_pairs_hand_worked():
    # R1: A (8 hits, proper with X), B (3 hits, proper with nothing), C (7 hits, proper with Y)
    A = nam(0, 30, 1000, 1030, 8, 1, 0, score=80)
    B = nam(0, 30, 90000, 90030, 3, 1, 0, score=30)
    C = nam(2, 32, 5000, 5030, 7, 1, 0, score=70)
    X = nam(0, 30, 1200, 1230, 6, 1, 1, score=60)
    Y = nam(0, 30, 5300, 5330, 6, 1, 1, score=60)
    got = O.best_scoring_nam_pairs(nams(A, B, C), nams(X, Y))
    # row A: (A,X)=14 pushed, (A,Y) not proper; row B: 3+6=9 >= 7, no proper;
    # row C: (C,X) 13 not proper, (C,Y)=13 pushed. best = 14.
    # unpaired R1: A added, B: 3 < 14 // 2 = 7 -> stop. R2: all added.
    assert [tuple(x) for x in got] == [(14, 0, 0), (13, 2, 1)]


def test_joint_pairs_break_on_half_best():
    A = nam(0, 30, 1000, 1030, 10, 0, 0, score=100)
    X = nam(0, 30, 1200, 1230, 10, 0, 1, score=100)
    low = nam(0, 30, 1100, 1130, 1, 0, 1, score=10)
    B = nam(5, 35, 1005, 1035, 3, 0, 0, score=30)
    # row A: (A,X)=20 pushed; (A,low)=11 >= 10 and proper -> pushed
    # row B: (B,X)=13 >= 10, proper -> pushed; (B,low)=4 < 10 -> break
    got = O.best_scoring_nam_pairs(nams(A, B), nams(X, low))
    assert sorted((int(s), int(i), int(j)) for s, i, j in got) == [(11, 0, 1), (13, 1, 0), (20, 0, 0)]


def test_joint_pairs_unpaired_threshold_uses_integer_division():
    # no proper pairs: best1 = nams1[0].n_hits = 7, threshold 7 // 2 = 3
    R1 = nams(nam(0, 30, 10, 40, 7, 0, 0, score=70), nam(1, 31, 500, 530, 3, 0, 0, score=30),
              nam(2, 32, 900, 930, 2, 0, 0, score=20), nam(3, 33, 990, 1020, 9, 0, 0, score=90))
    R2 = nams(nam(0, 30, 10, 40, 4, 3, 1, score=40))
    got = O.best_scoring_nam_pairs(R1, R2)
    # R1 scan: 7 ok, 3 ok (3 >= 3), 2 < 3 -> stop (the 9 after it is never seen)
    assert sorted((int(s), int(i), int(j)) for s, i, j in got) == [(3, 1, -1), (4, -1, 0), (7, 0, -1)]


def test_joint_pairs_enumerate_every_proper_pair_when_nothing_breaks():
    rng = np.random.default_rng(5)
    for trial in range(60):
        n1, n2 = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        R1 = [nam(int(rng.integers(0, 20)), 60, int(rng.integers(0, 3000)), 0, 5, int(rng.integers(0, 2)),
                  0, score=float(rng.integers(1, 100)), nam_id=i) for i in range(n1)]
        R2 = [nam(int(rng.integers(0, 20)), 60, int(rng.integers(0, 3000)), 0, 5, int(rng.integers(0, 2)),
                  1, score=float(rng.integers(1, 100)), nam_id=j) for j in range(n2)]
        got = O.best_scoring_nam_pairs(nams(*R1), nams(*R2), max_tries=20)
        want = {(i, j) for i in range(n1) for j in range(n2) if _proper(R1[i], R2[j])}
        paired = {(int(i), int(j)) for _, i, j in got if i >= 0 and j >= 0}
        assert paired == want
        # every NAM in no proper pair appears once on its own (all have 5 hits)
        assert {int(i) for _, i, j in got if j < 0} == set(range(n1)) - {i for i, _ in want}
        assert {int(j) for _, i, j in got if i < 0} == set(range(n2)) - {j for _, j in want}


def test_joint_pairs_stop_after_2_max_tries_plus_one():
    # every (i, j) is proper with the same joint hits
    R1 = nams(*[nam(0, 30, 1000 + i, 1030 + i, 5, 0, 0, nam_id=i) for i in range(10)])
    R2 = nams(*[nam(0, 30, 1100 + j, 1130 + j, 5, 0, 1, nam_id=j) for j in range(10)])
    for mt in (1, 2, 3, 7):
        got = O.best_scoring_nam_pairs(R1, R2, max_tries=mt)
        paired = [(int(i), int(j)) for _, i, j in got if i >= 0 and j >= 0]
        assert len(paired) == 2 * mt + 1
        # row-major order of discovery: the first 2 mt + 1 cells
        assert sorted(paired) == [(k // 10, k % 10) for k in range(2 * mt + 1)]


def _entry_key(e, R1, R2):
    """Sort key of a joint entry: score descending, then each NAM's score
    descending, is_rc, query_end, query_start, ref_end, ref_start (an absent
    NAM is all zero with ref_start -1)."""
    s, i, j = (int(x) for x in e)

    def f(n):
        if n is None:
            return (-0.0, 0, 0, 0, 0, -1)
        return (-float(n["score"]), int(n["is_rc"]), int(n["query_end"]), int(n["query_start"]),
                int(n["ref_end"]), int(n["ref_start"]))
    return (-s,) + f(R1[i] if i >= 0 else None) + f(R2[j] if j >= 0 else None)


def test_joint_pairs_sorted_by_the_comparator():
    rng = np.random.default_rng(9)
    for trial in range(100):
        R1 = rand_list(rng, int(rng.integers(1, 15)), n_refs=2)
        R2 = rand_list(rng, int(rng.integers(1, 15)), n_refs=2)
        R1 = O.sort_nams_reference(R1, 1)
        R2 = O.sort_nams_reference(R2, 2)
        got = O.best_scoring_nam_pairs(R1, R2, max_tries=int(rng.integers(1, 8)))
        keys = [_entry_key(e, R1, R2) for e in got]
        assert keys == sorted(keys)


def test_joint_pairs_equal_entries_keep_the_quicksort_order():
    # X and Y differ only in ref_id, which the comparator does not look at:
    # entries (5, 0, -1) and (5, 1, -1) compare equal. Traced by hand: the
    # first partition (pivot (3, -1, 0)) leaves [e0, e1, pivot]; the second
    # (range [0, 1], pivot e1) finds e0 not less than e1 and swaps them.
    X = nam(0, 30, 100, 130, 5, ref_id=0, is_rc=0, score=50)
    Y = nam(0, 30, 100, 130, 5, ref_id=1, is_rc=0, score=50)
    Z = nam(0, 30, 100, 130, 3, ref_id=2, is_rc=1, score=30)
    got = O.best_scoring_nam_pairs(nams(X, Y), nams(Z))
    assert [tuple(int(v) for v in e) for e in got] == [(5, 1, -1), (5, 0, -1), (3, -1, 0)]


def test_pairs_batch_agrees_with_the_single_pair_functions():
    rng = np.random.default_rng(21)
    n_pairs = 40
    lists = [rand_list(rng, int(rng.integers(0, 6)), n_refs=2) for _ in range(2 * n_pairs)]
    offs = np.zeros(2 * n_pairs + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(x) for x in lists])
    flat = np.concatenate(lists)
    r = O.pairs_batch(offs, flat, n_pairs)
    for p in range(n_pairs):
        a = O.sort_nams_reference(lists[p], 1)
        b = O.sort_nams_reference(lists[n_pairs + p], 2)
        assert r["types"][p] == O.pair_type(a, b)
        js = r["joint"][int(r["joint_offsets"][p]): int(r["joint_offsets"][p + 1])]
        if len(a) and len(b):
            assert js.tobytes() == O.best_scoring_nam_pairs(a, b).tobytes()
        else:
            assert len(js) == 0
