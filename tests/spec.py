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
