"""Shared inputs for the tests: deterministic sequences that exercise the
edge cases of the seeding path (ties in the syncmer window, N bases, short
and long reads, lowercase)."""
import numpy as np

from oracle.oracle_lib import QUERY_RANDSTROBE, REF_RANDSTROBE

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


def tiny_index_arrays(entries, bits=8):
    """RefRandstrobe table and bucket starts from (hash, position, packed)
    entries, with a leading dummy entry (the first hash of a table is never
    findable: bucket starts are pushed only at hash changes,
    step2/index.cu:206-210)."""
    ents = sorted([(1, 0, 0)] + list(entries))
    arr = np.array(ents, dtype=REF_RANDSTROBE)
    h = arr["hash"]
    first_run = int(np.searchsorted(h, h[0], side="right"))
    top = (h >> np.uint64(64 - bits)).astype(np.int64)
    starts = np.searchsorted(top[first_run:], np.arange((1 << bits) + 1)) + first_run
    starts[-1] = len(h)
    return arr, starts.astype(np.uint64)


# The find_nams composition worked by hand from the reference's text
# (step2/nam.cu:45-62 add_to_hits_per_ref, 347-513 merge_hits_into_nams,
# 553-747 find_nams; k = 20, filter_cutoff 5). Index, all on reference 0:
#   H1: positions 1000, 2000, 3000, 4000, 5000 with strobe2 offsets 25, 17,
#       24, 23, 20 (reference spans 45, 37, 44, 43, 40)
#   H2: position 1010, offset 30 (span 50)
#   H3: six entries at 7000.. (run of 6 > filter_cutoff: filtered)
# Query randstrobes: H1 at 10..50 and H2 at 20..60 forward, H3 at 30..70
# forward, H4 (absent) at 40..80 forward, H2 at 25..65 reverse.
# H1's run against query span 40: differences 5, 3, 4, 3, 0 against the
# running minimum (<=): 1000 kept (min 5), 2000 kept (3), 3000 dropped
# (4 > 3), 4000 kept (3 <= 3), 5000 kept (0). Forward hits sorted by
# (query_start, ref_start): (10,50,1000,1045) (10,50,2000,2037)
# (10,50,4000,4043) (10,50,5000,5040) (20,60,1010,1060); the first four open
# NAMs (query_prev 10 is not < 10), the fifth extends the 1000 NAM (both ends
# advance). Scores n_hits * (2 min_span - max_span): 2 * (100 - 60) = 80,
# 74 - 40 = 34, 80 - 43 = 37, 80 - 40 = 40; the reverse NAM 80 - 50 = 30.
# found 4 (H1, H2, H3, H2), good 3: fraction 3/4.
H1, H2, H3, H4 = (0x0123456789ABCDE1, 0x3456789ABCDEF012, 0x56789ABCDEF01234, 0x789ABCDEF0123456)
HAND_NAMS = [  # is_rc, query_start, query_end, ref_start, ref_end, n_hits, score
    (0, 10, 60, 1000, 1060, 2, 80.0), (0, 10, 50, 2000, 2037, 1, 34.0), (0, 10, 50, 4000, 4043, 1, 37.0),
    (0, 10, 50, 5000, 5040, 1, 40.0), (1, 25, 65, 1010, 1060, 1, 30.0)]


def hand_case():
    ents = [(H1, 1000, 25), (H1, 2000, 17), (H1, 3000, 24), (H1, 4000, 23), (H1, 5000, 20), (H2, 1010, 30)]
    ents += [(H3, 7000 + c, 10) for c in range(6)]
    q = np.zeros(5, dtype=QUERY_RANDSTROBE)
    q["hash"] = [H1, H2, H3, H4, H2]
    q["start"] = [10, 20, 30, 40, 25]
    q["end"] = [50, 60, 70, 80, 65]
    q["is_reverse"] = [0, 0, 0, 0, 1]
    return ents, q
