"""ctypes binding of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, by the smoke check and by bench.py's CPU-baseline leg; the
product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

ORC_DEFAULT = -(2 ** 31)
NOT_FOUND = 2 ** 64 - 1
FLAG_RESCUED = 1

QUERY_RANDSTROBE = np.dtype([("hash", "<u8"), ("start", "<u4"), ("end", "<u4"),
                             ("is_reverse", "u1"), ("pad", "u1", (7,))])
REF_RANDSTROBE = np.dtype([("hash", "<u8"), ("position", "<u4"), ("packed", "<u4")])
SYNCMER = np.dtype([("hash", "<u8"), ("position", "<u8")])
HIT = np.dtype([("query_start", "<i4"), ("query_end", "<i4"), ("ref_start", "<i4"),
                ("ref_end", "<i4")])
NAM_PAIR = np.dtype([("score", "<i4"), ("i1", "<i4"), ("i2", "<i4")])
NAM = np.dtype([("nam_id", "<i4"), ("query_start", "<i4"), ("query_end", "<i4"),
                ("query_prev_hit_startpos", "<i4"), ("ref_start", "<i4"), ("ref_end", "<i4"),
                ("ref_prev_hit_startpos", "<i4"), ("n_hits", "<i4"), ("ref_id", "<i4"),
                ("score", "<f4"), ("is_rc", "u1"), ("pad", "u1", (3,))])


class Params(C.Structure):
    _fields_ = [("canonical_read_length", C.c_int32), ("k", C.c_int32), ("s", C.c_int32),
                ("t_syncmer", C.c_int32), ("l", C.c_int32), ("u", C.c_int32),
                ("q", C.c_uint64), ("max_dist", C.c_int32), ("w_min", C.c_uint32),
                ("w_max", C.c_uint32)]

    def as_tuple(self):
        return (self.canonical_read_length, self.k, self.s, self.t_syncmer, self.l, self.u,
                self.q, self.max_dist, self.w_min, self.w_max)


class _Index(C.Structure):
    _fields_ = [("randstrobes", C.c_void_p), ("n", C.c_uint64), ("starts", C.c_void_p),
                ("bits", C.c_int32), ("filter_cutoff", C.c_uint32), ("params", Params),
                ("distinct_strobemers", C.c_uint64), ("tot_occur_once", C.c_uint64),
                ("tot_high_ab", C.c_uint64), ("tot_mid_ab", C.c_uint64),
                ("index_cutoff", C.c_uint64), ("owns", C.c_int)]


_lib = None
_P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_native():
    """Build the oracle for the CPU this process runs on (-march=native) and
    load that build instead of the portable one. Must be called before the
    first use of lib(). Returns the compile flags in effect."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, timeout=300,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    except (subprocess.SubprocessError, OSError):
        return "-O3 -march=x86-64-v3"
    LIB_PATH = os.path.join(_HERE, "_build", "liboracle_native.so")
    return "-O3 -march=native"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_xxh64": (C.c_uint64, [C.c_uint64]),
            "orc_nt4": (C.c_int, [C.c_ubyte]),
            "orc_params_from_read_length": (C.c_int, [C.c_int] * 7 + [C.POINTER(Params)]),
            "orc_params_make": (C.c_int, [C.c_int] * 7 + [C.POINTER(Params)]),
            "orc_syncmers": (C.c_size_t, [_P, C.c_size_t, C.POINTER(Params), _P, C.c_size_t]),
            "orc_randstrobes_query": (C.c_size_t, [_P, C.c_size_t, C.POINTER(Params), _P, C.c_size_t]),
            "orc_pick_bits": (C.c_int, [C.c_uint64, C.POINTER(Params)]),
            "orc_index_build": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(Params), C.c_int, C.c_float,
                                          C.c_int, C.POINTER(C.POINTER(_Index))]),
            "orc_index_wrap": (C.c_int, [_P, C.c_uint64, _P, C.c_int, C.c_uint32, C.POINTER(Params),
                                         C.POINTER(C.POINTER(_Index))]),
            "orc_index_free": (None, [C.POINTER(_Index)]),
            "orc_find": (C.c_uint64, [C.POINTER(_Index), C.c_uint64]),
            "orc_is_filtered": (C.c_int, [C.POINTER(_Index), C.c_uint64]),
            "orc_get_count": (C.c_uint32, [C.POINTER(_Index), C.c_uint64]),
            "orc_find_nams": (C.c_long, [C.POINTER(_Index), _P, C.c_size_t, _P, C.c_size_t,
                                         C.POINTER(C.c_float), C.POINTER(C.c_uint64)]),
            "orc_find_nams_rescue": (C.c_long, [C.POINTER(_Index), _P, C.c_size_t, C.c_uint,
                                                C.c_uint, _P, C.c_size_t]),
            "orc_merge_hits": (C.c_long, [_P, _P, C.c_size_t, C.c_int, C.c_int, C.c_int, _P,
                                          C.c_size_t]),
            "orc_canonicalize": (None, [_P, C.c_size_t]),
            "orc_map_read": (C.c_long, [C.POINTER(_Index), _P, C.c_size_t, C.c_uint, C.c_uint, _P,
                                        C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_uint8)]),
            "orc_map_batch": (C.c_int, [C.POINTER(_Index), _P, _P, _P, C.c_size_t, C.c_uint,
                                        C.c_uint, C.c_int, _P, _P, _P, _P, _P]),
            "orc_lookup_bytes": (C.c_int, [C.POINTER(_Index), _P, C.c_size_t, C.POINTER(C.c_uint64)]),
            "orc_sort_nams_reference": (None, [_P, C.c_size_t, C.c_int]),
            "orc_top_dropoff": (C.c_float, [_P, C.c_size_t]),
            "orc_get_mapq": (C.c_uint8, [_P, C.c_size_t, _P]),
            "orc_is_proper_nam_pair": (C.c_int, [_P, _P, C.c_float, C.c_float]),
            "orc_pair_type": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, C.c_float, C.c_float,
                                        C.c_float]),
            "orc_best_scoring_nam_pairs": (C.c_long, [_P, C.c_size_t, _P, C.c_size_t, C.c_float,
                                                      C.c_float, C.c_int, _P, C.c_size_t]),
            "orc_pairs_batch": (C.c_int, [_P, _P, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                          C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def xxh64(x):
    return lib().orc_xxh64(C.c_uint64(x))


def nt4(c):
    return lib().orc_nt4(c)


def params_from_read_length(r, k=ORC_DEFAULT, s=ORC_DEFAULT, l=ORC_DEFAULT, u=ORC_DEFAULT,
                            c=ORC_DEFAULT, max_seed_len=ORC_DEFAULT):
    p = Params()
    rc = lib().orc_params_from_read_length(r, k, s, l, u, c, max_seed_len, C.byref(p))
    if rc:
        raise ValueError(f"bad parameter ({rc})")
    return p


def params_make(r, k, s, l, u, q, max_dist):
    p = Params()
    rc = lib().orc_params_make(r, k, s, l, u, q, max_dist, C.byref(p))
    if rc:
        raise ValueError(f"bad parameter ({rc})")
    return p


def _bytes(seq):
    return seq.encode() if isinstance(seq, str) else bytes(seq)


def syncmers(seq, p):
    b = _bytes(seq)
    out = np.empty(len(b) + 1, dtype=SYNCMER)
    n = lib().orc_syncmers(b, len(b), C.byref(p), _p(out), len(out))
    return out[:n]


def randstrobes_query(seq, p):
    b = _bytes(seq)
    cap = 2 * (len(b) + 1)
    out = np.empty(cap, dtype=QUERY_RANDSTROBE)
    n = lib().orc_randstrobes_query(b, len(b), C.byref(p), _p(out), cap)
    return out[:n]


class Index:
    def __init__(self, handle, keep=None):
        self.h = handle
        self._keep = keep
        ix = handle.contents
        self.n, self.bits, self.filter_cutoff = ix.n, ix.bits, ix.filter_cutoff
        self.params = ix.params
        self.distinct_strobemers = ix.distinct_strobemers
        self.tot_occur_once, self.tot_high_ab, self.tot_mid_ab = ix.tot_occur_once, ix.tot_high_ab, ix.tot_mid_ab
        self.index_cutoff = ix.index_cutoff
        nb = (1 << ix.bits) + 1
        self.randstrobes = np.ctypeslib.as_array(
            (C.c_uint8 * (16 * max(ix.n, 1))).from_address(ix.randstrobes)).view(REF_RANDSTROBE)[: ix.n] \
            if ix.n else np.empty(0, dtype=REF_RANDSTROBE)
        self.starts = np.ctypeslib.as_array((C.c_uint64 * nb).from_address(ix.starts))

    @classmethod
    def build(cls, seqs, p, bits=-1, f=0.0002, n_threads=4):
        bs = [_bytes(s) for s in seqs]
        bufs = [C.create_string_buffer(b, len(b)) for b in bs]
        arr = (C.c_void_p * max(1, len(bs)))(*[C.cast(b, C.c_void_p) for b in bufs])
        lens = np.array([len(b) for b in bs], dtype=np.uint64)
        out = C.POINTER(_Index)()
        rc = lib().orc_index_build(arr, _p(lens), len(bs), C.byref(p), bits, C.c_float(f),
                                   n_threads, C.byref(out))
        if rc:
            raise ValueError(f"index build failed ({rc})")
        return cls(out)

    @classmethod
    def build_arrays(cls, genome, contig_lens, p, bits=-1, f=0.0002, n_threads=4):
        contig_lens = np.ascontiguousarray(contig_lens, dtype=np.uint64)
        starts = np.concatenate([[0], np.cumsum(contig_lens)[:-1]]).astype(np.uint64)
        arr = (C.c_void_p * len(contig_lens))(*[genome.ctypes.data + int(s) for s in starts])
        out = C.POINTER(_Index)()
        rc = lib().orc_index_build(arr, _p(contig_lens), len(contig_lens), C.byref(p), bits,
                                   C.c_float(f), n_threads, C.byref(out))
        if rc:
            raise ValueError(f"index build failed ({rc})")
        return cls(out)

    @classmethod
    def wrap(cls, randstrobes, starts, bits, filter_cutoff, p):
        rs = np.ascontiguousarray(randstrobes, dtype=REF_RANDSTROBE)
        st = np.ascontiguousarray(starts, dtype=np.uint64)
        out = C.POINTER(_Index)()
        rc = lib().orc_index_wrap(_p(rs), len(rs), _p(st), bits, filter_cutoff, C.byref(p),
                                  C.byref(out))
        if rc:
            raise ValueError(f"wrap failed ({rc})")
        return cls(out, keep=(rs, st))

    def free(self):
        if self.h:
            lib().orc_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def find(self, key):
        return lib().orc_find(self.h, C.c_uint64(key))

    def is_filtered(self, pos):
        return bool(lib().orc_is_filtered(self.h, pos))

    def get_count(self, pos):
        return lib().orc_get_count(self.h, pos)


def find_nams(idx, qrs, canonical=False):
    qrs = np.ascontiguousarray(qrs, dtype=QUERY_RANDSTROBE)
    cap = 64 + 16 * len(qrs) * 8
    while True:
        out = np.empty(cap, dtype=NAM)
        fr = C.c_float()
        nh = C.c_uint64()
        n = lib().orc_find_nams(idx.h, _p(qrs), len(qrs), _p(out), cap, C.byref(fr), C.byref(nh))
        if n >= 0:
            break
        cap *= 8
    out = out[:n].copy()
    if canonical:
        canonicalize(out)
    return fr.value, out, nh.value


def find_nams_rescue(idx, qrs, rescue_cutoff, hard_cutoff=1000, canonical=False):
    qrs = np.ascontiguousarray(qrs, dtype=QUERY_RANDSTROBE)
    cap = 64 + 16 * len(qrs) * 8
    while True:
        out = np.empty(cap, dtype=NAM)
        n = lib().orc_find_nams_rescue(idx.h, _p(qrs), len(qrs), rescue_cutoff, hard_cutoff,
                                       _p(out), cap)
        if n >= 0:
            break
        cap *= 8
    out = out[:n].copy()
    if canonical:
        canonicalize(out)
    return out


def merge_hits(ref_ids, hits, k, is_rc, nam_outer):
    ref_ids = np.ascontiguousarray(ref_ids, dtype=np.int32)
    hits = np.ascontiguousarray(hits, dtype=HIT)
    cap = len(hits) + 1
    out = np.empty(cap, dtype=NAM)
    n = lib().orc_merge_hits(_p(ref_ids), _p(hits), len(hits), k, int(is_rc), int(nam_outer),
                             _p(out), cap)
    return out[:n].copy()


def canonicalize(nams):
    nams = np.ascontiguousarray(nams)
    lib().orc_canonicalize(_p(nams), len(nams))
    return nams


def map_read(idx, seq, rescue_cutoff, hard_cutoff=1000):
    b = _bytes(seq)
    cap = 64 + 32 * len(b)
    while True:
        out = np.empty(cap, dtype=NAM)
        fr = C.c_float()
        fl = C.c_uint8()
        n = lib().orc_map_read(idx.h, b, len(b), rescue_cutoff, hard_cutoff, _p(out), cap,
                               C.byref(fr), C.byref(fl))
        if n >= 0:
            break
        cap *= 8
    return out[:n].copy(), fr.value, fl.value


def map_batch(idx, seqs, offsets, lens, rescue_cutoff, hard_cutoff=1000, n_threads=4):
    """Whole path for many reads -> (nam_offsets, nams, frac, flags), NAMs in
    canonical order."""
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = len(lens)
    counts = np.empty(max(n, 1), dtype=np.uint32)
    frac = np.empty(max(n, 1), dtype=np.float32)
    flags = np.empty(max(n, 1), dtype=np.uint8)
    L = lib()
    L.orc_map_batch(idx.h, _p(seqs), _p(offsets), _p(lens), n, rescue_cutoff, hard_cutoff,
                    n_threads, None, None, _p(counts), _p(frac), _p(flags))
    noff = np.zeros(n + 1, dtype=np.uint64)
    noff[1:] = np.cumsum(counts[:n], dtype=np.uint64)
    nams = np.empty(max(int(noff[-1]), 1), dtype=NAM)
    L.orc_map_batch(idx.h, _p(seqs), _p(offsets), _p(lens), n, rescue_cutoff, hard_cutoff,
                    n_threads, _p(noff), _p(nams), _p(counts), _p(frac), _p(flags))
    return noff, nams[: int(noff[-1])], frac[:n], flags[:n]


def time_map_batch(idx, seqs, offsets, lens, rescue_cutoff, hard_cutoff=1000, n_threads=4):
    """Timed count-only pass (the CPU baseline)."""
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = len(lens)
    counts = np.empty(max(n, 1), dtype=np.uint32)
    import time
    t0 = time.perf_counter()
    lib().orc_map_batch(idx.h, _p(seqs), _p(offsets), _p(lens), n, rescue_cutoff, hard_cutoff,
                        n_threads, None, None, _p(counts), None, None)
    return time.perf_counter() - t0


def lookup_bytes(idx, qrs):
    qrs = np.ascontiguousarray(qrs, dtype=QUERY_RANDSTROBE)
    b = C.c_uint64()
    lib().orc_lookup_bytes(idx.h, _p(qrs), len(qrs), C.byref(b))
    return b.value


# ---- NAM consumer of paired reads ----

def _nam_arr(nams):
    return np.ascontiguousarray(nams, dtype=NAM)


def sort_nams_reference(nams, mate):
    a = _nam_arr(nams).copy()
    lib().orc_sort_nams_reference(_p(a), len(a), mate)
    return a


def top_dropoff(nams):
    a = _nam_arr(nams)
    return lib().orc_top_dropoff(_p(a), len(a))


def get_mapq(nams, n_max=None):
    a = _nam_arr(nams)
    m = _nam_arr([n_max] if n_max is not None else a[:1])
    return lib().orc_get_mapq(_p(a), len(a), _p(m))


def is_proper_nam_pair(nam1, nam2, mu=300.0, sigma=100.0):
    a, b = _nam_arr([nam1]), _nam_arr([nam2])
    return bool(lib().orc_is_proper_nam_pair(_p(a), _p(b), mu, sigma))


def pair_type(nams1, nams2, dropoff=0.5, mu=300.0, sigma=100.0):
    a, b = _nam_arr(nams1), _nam_arr(nams2)
    return lib().orc_pair_type(_p(a), len(a), _p(b), len(b), dropoff, mu, sigma)


def best_scoring_nam_pairs(nams1, nams2, mu=300.0, sigma=100.0, max_tries=20):
    a, b = _nam_arr(nams1), _nam_arr(nams2)
    cap = 4 * max_tries + 1
    out = np.empty(cap, dtype=NAM_PAIR)
    n = lib().orc_best_scoring_nam_pairs(_p(a), len(a), _p(b), len(b), mu, sigma, max_tries,
                                         _p(out), cap)
    if n < 0:
        raise ValueError("empty NAM list or bad max_tries")
    return out[:n].copy()


def pairs_batch(nam_offsets, nams, n_pairs, dropoff=0.5, mu=300.0, sigma=100.0, max_tries=20,
                joint=True):
    """The pair step on 2 n_pairs lists -> dict with ord (NAMs in the
    reference order), types, type_lists, top_dropoff, mapq, joint_offsets,
    joint."""
    noff = np.ascontiguousarray(nam_offsets, dtype=np.uint64)
    a = _nam_arr(nams)
    ord_ = np.empty(max(len(a), 1), dtype=NAM)
    types = np.empty(max(n_pairs, 1), dtype=np.uint8)
    drop = np.empty(max(2 * n_pairs, 1), dtype=np.float32)
    mapq = np.empty(max(2 * n_pairs, 1), dtype=np.uint8)
    stride = 4 * max_tries + 1
    jc = np.zeros(max(n_pairs, 1), dtype=np.uint32)
    jout = np.empty(max(n_pairs * stride if joint else 1, 1), dtype=NAM_PAIR)
    rc = lib().orc_pairs_batch(_p(noff), _p(a), n_pairs, dropoff, mu, sigma, max_tries, int(joint),
                               _p(ord_), _p(types), _p(drop), _p(mapq), _p(jc), _p(jout))
    if rc:
        raise RuntimeError(f"orc_pairs_batch failed ({rc})")
    types = types[:n_pairs]
    joff = np.zeros(n_pairs + 1, dtype=np.uint64)
    if joint and n_pairs:
        joff[1:] = np.cumsum(jc[:n_pairs], dtype=np.uint64)
        sel = np.concatenate([np.arange(p * stride, p * stride + jc[p]) for p in range(n_pairs)])
        jflat = jout[sel.astype(np.int64)] if len(sel) else np.empty(0, dtype=NAM_PAIR)
    else:
        jflat = np.empty(0, dtype=NAM_PAIR)
    return {
        "ord": ord_[: len(a)], "types": types,
        "type_lists": [np.nonzero(types == t)[0].astype(np.uint32) for t in range(5)],
        "top_dropoff": drop[: 2 * n_pairs], "mapq": mapq[: 2 * n_pairs],
        "joint_offsets": joff, "joint": jflat,
    }
