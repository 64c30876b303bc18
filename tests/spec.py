"""Specification-level models used to pin the oracle.

These are written from the definitions (canonical open syncmers with the
reference's tie rules stated window by window, randstrobe partner choice by
minimal popcount, the bucket index), in plain Python over recomputed windows,
with the independent `xxhash` library for hashing -- not from the oracle's
code."""
import struct

import numpy as np
import xxhash

MAX64 = 2 ** 64 - 1
NONE = "none"


def xxh(v):
    return xxhash.xxh64_intdigest(struct.pack("<Q", v), seed=0)


def code(c):
    return {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3, ord("U"): 3,
            ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3, ord("u"): 3}.get(c, 4)


def canonical(codes, j, n):
    f = 0
    r = 0
    for m in range(n):
        c = codes[j + m]
        f = (f << 2) | c
        r |= (3 - c) << (2 * m)
    return min(f, r)


def smer_hashes(seq, s):
    codes = [code(c) for c in seq]
    out = []
    for j in range(len(seq) - s + 1):
        if any(c > 3 for c in codes[j: j + s]):
            out.append(None)
        else:
            out.append(xxh(canonical(codes, j, s)))
    return codes, out


def syncmers(seq, k, s):
    """[(hash, position)] under the reference's window rules."""
    w = k - s + 1
    t = (k - s) // 2 + 1
    codes, h = smer_hashes(seq, s)
    out = []
    state = None
    for p in range(len(seq) - k + 1):
        if any(c > 3 for c in codes[p: p + k]):
            state = None
            continue
        vals = h[p: p + w]
        if p == 0 or codes[p - 1] > 3:
            m = min(vals)
            state = NONE if m == MAX64 else vals.index(m)          # leftmost minimum
        elif state == 0:
            m = min(vals)
            state = max(i for i, v in enumerate(vals) if v == m)   # rightmost minimum
        elif state == NONE:
            state = w - 1 if vals[-1] < MAX64 else NONE
        else:
            old = h[p - 1 + state]
            state = w - 1 if vals[-1] < old else state - 1
        if state == t - 1:
            out.append((xxh(canonical(codes, p, k)), p))
    return out


def unique_min_syncmers(seq, k, s):
    """Textbook canonical open syncmers on windows whose minimum is unique:
    {position: is_syncmer}; tied windows are left out."""
    w = k - s + 1
    t = (k - s) // 2 + 1
    codes, h = smer_hashes(seq, s)
    res = {}
    for p in range(len(seq) - k + 1):
        vals = h[p: p + w]
        if any(v is None for v in vals):
            continue
        m = min(vals)
        if vals.count(m) == 1:
            res[p] = vals.index(m) == t - 1
    return res


def merge_hits_into_nams(ref_ids, hits, k, is_rc):
    """NAMs of one strand's hits, from the rules of merge_hits_into_nams
    (step2/nam.cu:347-513): per reference (ascending ref id here; the
    reference's map order only permutes the output), hits in (query_start,
    ref_start) order; a hit joins the first open NAM, in the order NAMs were
    opened, whose previous hit starts before it on both sequences and whose
    end it does not pass on either, when it either advances both ends or
    lies inside both; otherwise it opens a NAM. Once the query start has
    moved more than k past the last flush point, the open NAMs ending before
    the hit are closed in opening order. Closed NAMs are numbered in closing
    order and scored n_hits * (2 min_span - max_span), or 1 when that is not
    positive. Returns (nam_id, qs, qe, qprev, rs, re, rprev, n_hits, ref,
    score, is_rc) tuples in closing order; score is a float32."""
    by_ref = {}
    for r, h in zip(ref_ids, hits):
        by_ref.setdefault(int(r), []).append(tuple(int(x) for x in h))
    out = []

    def close(n):
        qspan, rspan = n["qe"] - n["qs"], n["re"] - n["rs"]
        d = 2 * min(qspan, rspan) - max(qspan, rspan)
        score = np.float32(n["hits"] * d) if d > 0 else np.float32(1)
        out.append((len(out), n["qs"], n["qe"], n["qp"], n["rs"], n["re"], n["rp"], n["hits"], n["ref"],
                    score, is_rc))

    for ref in sorted(by_ref):
        open_nams = []
        last_flush = 0
        for qs, qe, rs, re in sorted(by_ref[ref], key=lambda h: (h[0], h[2])):
            for n in open_nams:
                inside = n["qp"] < qs <= n["qe"] and n["rp"] < rs <= n["re"]
                advances = qe > n["qe"] and re > n["re"]
                contained = qe <= n["qe"] and re <= n["re"]
                if inside and (advances or contained):
                    if advances:
                        n["qe"], n["re"] = qe, re
                    n["qp"], n["rp"] = qs, rs
                    n["hits"] += 1
                    break
            else:
                open_nams.append({"qs": qs, "qe": qe, "rs": rs, "re": re, "qp": qs, "rp": rs, "hits": 1,
                                  "ref": ref})
            if qs > last_flush + k:
                for n in [n for n in open_nams if n["qe"] < qs]:
                    close(n)
                open_nams = [n for n in open_nams if n["qe"] >= qs]
                last_flush = qs
        for n in open_nams:
            close(n)
    return out


def canonical_nams(nams):
    """(is_rc, ref, query_start, ref_start) order, nam_id = rank."""
    s = sorted(nams, key=lambda n: (n[10], n[8], n[1], n[4]))
    return [(i,) + tuple(n[1:]) for i, n in enumerate(s)]
