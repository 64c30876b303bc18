"""Shared inputs for the tests: deterministic sequences that exercise the
edge cases of the seeding path (ties in the syncmer window, N bases, short
and long reads, lowercase)."""
import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def rand_seq(rng, n, gc=0.5):
    p = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
    return ACGT[rng.choice(4, size=n, p=p)].tobytes()


def revcomp(s):
    t = bytes.maketrans(b"ACGTacgtNn", b"TGCAtgcaNn")
    return s.translate(t)[::-1]


def tie_heavy_reads(rng, L=150):
    """Reads whose k-mer windows contain repeated canonical s-mers."""
    out = []
    out.append(b"A" * L)
    out.append(b"C" * L)
    out.append((b"AC" * L)[:L])
    out.append((b"AT" * L)[:L])       # palindromic dinucleotide
    out.append((b"ACG" * L)[:L])
    out.append((b"AGCT" * L)[:L])
    out.append((b"GATC" * L)[:L])
    for unit in (2, 3, 4, 5, 7, 11):
        u = rand_seq(rng, unit)
        out.append((u * (L // unit + 1))[:L])
    # repeats embedded in random flanks
    for unit in (1, 2, 3, 4):
        u = rand_seq(rng, unit)
        mid = (u * 60)[:40 + unit]
        a = rand_seq(rng, 50)
        out.append((a + mid + rand_seq(rng, L))[:L])
    # homopolymer runs of different lengths
    for run in (16, 17, 18, 19, 20, 21, 25, 40):
        out.append((rand_seq(rng, 40) + b"T" * run + rand_seq(rng, L))[:L])
    return out


def n_reads(rng, L=150):
    out = []
    base = rand_seq(rng, L)
    for pos in (0, 1, 5, 19, 20, 21, 50, 75, 130, 149):
        s = bytearray(base)
        s[pos] = ord("N")
        out.append(bytes(s))
    s = bytearray(base)
    for pos in range(10, L, 23):
        s[pos] = ord("N")
    out.append(bytes(s))
    out.append(b"N" * L)
    s = bytearray(base)
    for pos in range(0, L, 7):
        s[pos] = ord("R")  # IUPAC ambiguity code: not a base
    out.append(bytes(s))
    out.append(base.lower())
    out.append(base[:60].lower() + base[60:])
    out.append(base.replace(b"T", b"U"))
    return out


def length_reads(rng):
    lens = [0, 1, 10, 15, 16, 19, 20, 21, 24, 25, 30, 31, 40, 50, 63, 64, 65, 100, 127, 128,
            129, 150, 199, 250, 300, 319, 320, 321, 400, 500, 1000, 2500]
    return [rand_seq(rng, L) for L in lens]


def mutate(rng, s, sub=0.01):
    b = bytearray(s)
    for i in range(len(b)):
        if rng.random() < sub:
            b[i] = ACGT[rng.integers(4)]
    return bytes(b)
