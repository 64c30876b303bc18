"""MI355X-native strobealign seeding path (syncmers -> randstrobes -> index
lookup -> NAMs) on hand-written gfx950 kernels.

This module is the host-side mirror of the reference's C++ interface for the
path -- ``IndexParameters``, ``StrobemerIndex``, ``randstrobes_query``,
``find_nams``, ``find_nams_rescue`` (step2/indexparameters.hpp:72-95,
step2/index.hpp:37-183, step2/randstrobes.hpp:107, step2/nam.hpp:11-20) --
over the C ABI of ``libgis.so`` (include/gis.h). Whole batches go through
``Mapper``. Every call runs on the GPU; without the built library (or a GPU)
calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import _native as N
from ._native import (FLAG_NO_RANDSTROBES, FLAG_RESCUED, GIS_DEFAULT, GIS_NOT_FOUND, HIT,
                      MODE_FUSED, MODE_PLAIN, NAM, NAM_PAIR, QUERY_RANDSTROBE, REF_RANDSTROBE,
                      SYNCMER, GisError)

__all__ = [
    "IndexParameters", "References", "Context", "StrobemerIndex", "Mapper", "MapResult", "Server",
    "randstrobes_query", "find_nams", "find_nams_rescue", "synth_genome", "synth_reads",
    "pack_reads", "GisError", "NAM", "QUERY_RANDSTROBE", "REF_RANDSTROBE", "SYNCMER", "HIT",
    "GIS_NOT_FOUND", "FLAG_RESCUED", "FLAG_NO_RANDSTROBES", "MODE_PLAIN", "MODE_FUSED",
    "PairResult", "pairs_from_nams", "NAM_PAIR", "FastqReader", "FastqChunk",
    "find_nams_stream", "stream_copy_stats", "pinned_empty", "pinned_copy",
]


def _check(code, ctx=None):
    if code != 0:
        msg = N.lib().gis_last_error(ctx).decode() if ctx else ""
        raise GisError(code, msg)


class _Syncmer:
    def __init__(self, k, s, t):
        self.k, self.s, self.t_syncmer = k, s, t


class _Randstrobe:
    def __init__(self, l, u, q, max_dist, w_min, w_max):
        self.l, self.u, self.q, self.max_dist, self.w_min, self.w_max = l, u, q, max_dist, w_min, w_max


class IndexParameters:
    """IndexParameters (step2/indexparameters.hpp:72-95)."""
    DEFAULT = GIS_DEFAULT

    def __init__(self, canonical_read_length, k, s, l, u, q, max_dist):
        p = N.Params()
        _check(N.lib().gis_params_make(canonical_read_length, k, s, l, u, q, max_dist, C.byref(p)))
        self._set(p)

    def _set(self, p):
        self.c = p
        self.canonical_read_length = p.canonical_read_length
        self.syncmer = _Syncmer(p.k, p.s, p.t_syncmer)
        self.randstrobe = _Randstrobe(p.l, p.u, p.q, p.max_dist, p.w_min, p.w_max)

    @classmethod
    def from_read_length(cls, read_length, k=GIS_DEFAULT, s=GIS_DEFAULT, l=GIS_DEFAULT,
                         u=GIS_DEFAULT, c=GIS_DEFAULT, max_seed_len=GIS_DEFAULT):
        p = N.Params()
        _check(N.lib().gis_params_from_read_length(read_length, k, s, l, u, c, max_seed_len,
                                                   C.byref(p)))
        obj = cls.__new__(cls)
        obj._set(p)
        return obj

    @classmethod
    def _from_c(cls, p):
        obj = cls.__new__(cls)
        q = N.Params()
        C.memmove(C.byref(q), C.byref(p), C.sizeof(N.Params))
        obj._set(q)
        return obj

    def as_tuple(self):
        p = self.c
        return (p.canonical_read_length, p.k, p.s, p.t_syncmer, p.l, p.u, p.q, p.max_dist,
                p.w_min, p.w_max)

    def __eq__(self, other):
        return isinstance(other, IndexParameters) and self.as_tuple() == other.as_tuple()

    def filename_extension(self):
        """".r<len>.sti" when nothing was overridden (step2/indexparameters.cpp:116-125)"""
        if self == IndexParameters.from_read_length(self.canonical_read_length):
            return f".r{self.canonical_read_length}.sti"
        return ".sti"

    def __repr__(self):
        p = self.c
        return (f"IndexParameters(r={p.canonical_read_length}, k={p.k}, s={p.s}, "
                f"t_syncmer={p.t_syncmer}, l={p.l}, u={p.u}, q={p.q}, max_dist={p.max_dist}, "
                f"w_min={p.w_min}, w_max={p.w_max})")


class References:
    """Reference sequences (step2/refs.hpp:53-91): names, sequences, lengths."""

    def __init__(self, sequences, names=None):
        self.sequences = [s.encode() if isinstance(s, str) else bytes(s) for s in sequences]
        self.names = list(names) if names is not None else [f"contig{i}" for i in range(len(sequences))]
        if len(self.names) != len(self.sequences):
            raise ValueError("lengths do not match")
        self.lengths = [len(s) for s in self.sequences]

    def total_length(self):
        return sum(self.lengths)

    def size(self):
        return len(self.sequences)

    @classmethod
    def from_fasta(cls, path):
        """FASTA reader: names cut at the first space, sequences upper-cased
        with c & ~32 (step2/refs.cpp:32-79)."""
        names, seqs, cur, name = [], [], [], None
        with open(path, "rb") as fh:
            first = fh.read(1)
            if first != b">":
                raise GisError(10, "FASTA file must begin with '>' character")
            fh.seek(0)
            for line in fh:
                line = line.rstrip(b"\r\n")
                if line.startswith(b">"):
                    if cur:
                        seqs.append(bytes(x & ~32 & 0xFF for x in b"".join(cur)))
                        names.append(name)
                    name = line[1:].split(b" ")[0].decode()
                    cur = []
                else:
                    cur.append(line)
            if cur:
                seqs.append(bytes(x & ~32 & 0xFF for x in b"".join(cur)))
                names.append(name)
        return cls(seqs, names)


class Context:
    """A device context (one HIP stream, serialised calls)."""

    def __init__(self, device=0):
        self.L = N.lib()
        h = C.c_void_p()
        code = self.L.gis_ctx_create(device, C.byref(h))
        if code != 0:
            raise GisError(code, f"cannot create a context on device {device}")
        self.h = h

    def check(self, code):
        _check(code, self.h)

    @property
    def fused_config(self):
        """Launch configuration of the fused form (waves per block, length
        split, queue discipline, chunk, persistent blocks per CU)."""
        c = N.FusedConfig()
        self.check(self.L.gis_fused_config_get(self.h, C.byref(c)))
        return c.as_dict()

    @fused_config.setter
    def fused_config(self, cfg):
        c = N.FusedConfig()
        cur = self.fused_config
        cur.update(cfg)
        for k, v in cur.items():
            setattr(c, k, int(v))
        self.check(self.L.gis_fused_config_set(self.h, C.byref(c)))

    def close(self):
        if getattr(self, "h", None):
            self.L.gis_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class StrobemerIndex:
    """StrobemerIndex (step2/index.hpp:37-183), resident in HBM.

    A context holds one resident index; each StrobemerIndex gets a context of
    its own unless one is passed, and loading another index on a shared
    context invalidates this one (its calls then raise)."""

    def __init__(self, references, parameters, bits=-1, ctx=None):
        self.ctx = ctx or Context(0)
        self.references = references
        self.parameters = parameters
        self._bits_arg = bits
        self._host = None

    def _claim(self):
        self.ctx._owner = id(self)
        self._host = None

    def _check_owner(self):
        if getattr(self.ctx, "_owner", None) != id(self):
            raise GisError(6, "another index was loaded on this context after this one")

    # --- construction / persistence ---
    def populate(self, f=0.0002, n_threads=None):
        refs = self.references
        seqs = [s if isinstance(s, bytes) else bytes(s) for s in refs.sequences]
        bufs = [C.create_string_buffer(s, len(s)) for s in seqs]
        arr = (C.c_void_p * max(1, len(bufs)))(*[C.cast(b, C.c_void_p) for b in bufs])
        lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        n_threads = n_threads or min(16, os.cpu_count() or 1)
        self.ctx.check(self.ctx.L.gis_index_build(self.ctx.h, arr, N.ptr(lens), len(seqs),
                                                  C.byref(self.parameters.c), self._bits_arg,
                                                  C.c_float(f), n_threads))
        self._claim()

    def populate_arrays(self, genome, contig_lens, f=0.0002, n_threads=None):
        """populate() from one concatenated uint8 genome array (no copies)."""
        contig_lens = np.ascontiguousarray(contig_lens, dtype=np.uint64)
        starts = np.concatenate([[0], np.cumsum(contig_lens)[:-1]]).astype(np.uint64)
        base = genome.ctypes.data
        arr = (C.c_void_p * len(contig_lens))(*[base + int(s) for s in starts])
        n_threads = n_threads or min(16, os.cpu_count() or 1)
        self.ctx.check(self.ctx.L.gis_index_build(self.ctx.h, arr, N.ptr(contig_lens),
                                                  len(contig_lens), C.byref(self.parameters.c),
                                                  self._bits_arg, C.c_float(f), n_threads))
        self._claim()

    def upload(self, randstrobes, starts, bits, filter_cutoff):
        randstrobes = np.ascontiguousarray(randstrobes, dtype=REF_RANDSTROBE)
        starts = np.ascontiguousarray(starts, dtype=np.uint64)
        self.ctx.check(self.ctx.L.gis_index_upload(self.ctx.h, N.ptr(randstrobes), len(randstrobes),
                                                   N.ptr(starts), bits, filter_cutoff,
                                                   C.byref(self.parameters.c)))
        self._claim()

    def read(self, filename):
        self.ctx.check(self.ctx.L.gis_index_read_sti(self.ctx.h, os.fsencode(filename),
                                                     C.byref(self.parameters.c)))
        self._claim()

    def write(self, filename):
        self._check_owner()
        self.ctx.check(self.ctx.L.gis_index_write_sti(self.ctx.h, os.fsencode(filename)))

    # --- properties ---
    def info(self):
        self._check_owner()
        i = N.IndexInfo()
        self.ctx.check(self.ctx.L.gis_index_info_get(self.ctx.h, C.byref(i)))
        return i

    @property
    def bits(self):
        return self.info().bits

    @property
    def filter_cutoff(self):
        return self.info().filter_cutoff

    def size(self):
        return self.info().n_randstrobes

    def k(self):
        return self.parameters.syncmer.k

    def _download(self):
        if self._host is None:
            i = self.info()
            rs = np.empty(i.n_randstrobes, dtype=REF_RANDSTROBE)
            st = np.empty((1 << i.bits) + 1, dtype=np.uint64)
            self.ctx.check(self.ctx.L.gis_index_download(self.ctx.h, N.ptr(rs), N.ptr(st)))
            self._host = (rs, st)
        return self._host

    @property
    def randstrobes(self):
        return self._download()[0]

    @property
    def randstrobe_start_indices(self):
        return self._download()[1]

    # --- lookups (on the GPU) ---
    def find_many(self, keys):
        self._check_owner()
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(len(keys), dtype=np.uint64)
        self.ctx.check(self.ctx.L.gis_lookup_batch(self.ctx.h, N.ptr(keys), len(keys), N.ptr(out)))
        return out

    def find(self, key):
        return int(self.find_many(np.array([key], dtype=np.uint64))[0])

    def end(self):
        return GIS_NOT_FOUND

    # host-side accessors over the downloaded table
    def get_hash(self, position):
        rs = self.randstrobes
        return int(rs["hash"][position]) if position < len(rs) else GIS_NOT_FOUND

    def is_filtered(self, position):
        return self.get_hash(position) == self.get_hash(position + self.filter_cutoff)

    def get_strobe1_position(self, position):
        return int(self.randstrobes["position"][position])

    def strobe2_offset(self, position):
        return int(self.randstrobes["packed"][position]) & 0xFF

    def reference_index(self, position):
        return int(self.randstrobes["packed"][position]) >> 8


def pack_reads(reads):
    """Concatenate reads (str/bytes) -> (uint8 seqs, u64 offsets, u32 lens)."""
    bs = [r.encode() if isinstance(r, str) else bytes(r) for r in reads]
    lens = np.array([len(b) for b in bs], dtype=np.uint32)
    offsets = np.zeros(len(bs), dtype=np.uint64)
    if len(bs) > 1:
        offsets[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    seqs = np.frombuffer(b"".join(bs) or b"\0", dtype=np.uint8).copy()
    return seqs, offsets, lens


def _seq_args(seqs, offsets, lens):
    seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    return seqs, offsets, lens


def syncmers_batch(index, seqs, offsets, lens):
    """Canonical syncmers of every read -> (offsets[n+1], SYNCMER array)."""
    index._check_owner()
    ctx = index.ctx
    seqs, offsets, lens = _seq_args(seqs, offsets, lens)
    n = len(lens)
    cap = int(np.maximum(lens.astype(np.int64) - index.parameters.syncmer.k + 1, 0).sum())
    out = np.empty(max(cap, 1), dtype=SYNCMER)
    oo = np.empty(n + 1, dtype=np.uint64)
    nout = C.c_uint64()
    ctx.check(ctx.L.gis_syncmers_batch(ctx.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), n,
                                       N.ptr(oo), N.ptr(out), cap, C.byref(nout)))
    return oo, out[: nout.value]


def randstrobes_batch(index, seqs, offsets, lens):
    """randstrobes_query of every read -> (offsets[n+1], QUERY_RANDSTROBE array)."""
    index._check_owner()
    ctx = index.ctx
    seqs, offsets, lens = _seq_args(seqs, offsets, lens)
    n = len(lens)
    cap = int(2 * np.maximum(lens.astype(np.int64) - index.parameters.syncmer.k + 1, 0).sum())
    out = np.empty(max(cap, 1), dtype=QUERY_RANDSTROBE)
    oo = np.empty(n + 1, dtype=np.uint64)
    nout = C.c_uint64()
    ctx.check(ctx.L.gis_randstrobes_batch(ctx.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), n,
                                          N.ptr(oo), N.ptr(out), cap, C.byref(nout)))
    return oo, out[: nout.value]


def randstrobes_query(seq, parameters):
    """randstrobes_query(seq, parameters) (step2/randstrobes.hpp:107) for one
    read, on the GPU. ``parameters`` is an IndexParameters (seeded on the
    default context, no index needed) or a StrobemerIndex (its context and
    parameters)."""
    s, o, l = pack_reads([seq])
    if isinstance(parameters, StrobemerIndex):
        return randstrobes_batch(parameters, s, o, l)[1]
    ctx = default_context()
    n = len(l)
    cap = int(2 * max(len(seq) - parameters.syncmer.k + 1, 0)) or 1
    out = np.empty(cap, dtype=QUERY_RANDSTROBE)
    oo = np.empty(n + 1, dtype=np.uint64)
    nout = C.c_uint64()
    ctx.check(ctx.L.gis_randstrobes_params(ctx.h, C.byref(parameters.c), N.ptr(s), N.ptr(o), N.ptr(l), n,
                                           N.ptr(oo), N.ptr(out), cap, C.byref(nout)))
    return out[: nout.value]


def _map_options(rescue_level=2, hard_cutoff=1000, mode=N.MODE_PLAIN, graph=True):
    o = N.MapOptions()
    o.rescue_level, o.hard_cutoff, o.mode = rescue_level, hard_cutoff, mode
    o.flags = 0 if graph else N.RUN_NO_GRAPH
    return o


class MapResult:
    def __init__(self, nam_offsets, nams, frac, flags):
        self.nam_offsets, self.nams, self.nonrepetitive_fraction, self.flags = nam_offsets, nams, frac, flags

    def read_nams(self, r):
        return self.nams[int(self.nam_offsets[r]): int(self.nam_offsets[r + 1])]


def _nam_out(n_reads, cap):
    offs = np.empty(n_reads + 1, dtype=np.uint64)
    nams = np.empty(max(cap, 1), dtype=NAM)
    frac = np.empty(max(n_reads, 1), dtype=np.float32)
    flags = np.empty(max(n_reads, 1), dtype=np.uint8)
    o = N.NamOut()
    o.nam_offsets, o.nams, o.nam_capacity = offs.ctypes.data, nams.ctypes.data, cap
    o.nonrepetitive_fraction, o.flags = frac.ctypes.data, flags.ctypes.data
    return o, offs, nams, frac, flags


def _call_with_nam_out(ctx, n_reads, fn, cap=None):
    cap = cap if cap is not None else max(16, 8 * n_reads)
    while True:
        o, offs, nams, frac, flags = _nam_out(n_reads, cap)
        code = fn(o)
        if code == 5 and o.n_nams > cap:   # GIS_ERR_CAPACITY: retry with the reported size
            cap = int(o.n_nams)
            continue
        ctx.check(code)
        return MapResult(offs, nams[: o.n_nams], frac[:n_reads], flags[:n_reads])


def find_nams_batch(index, seqs, offsets, lens, rescue_level=2, hard_cutoff=1000,
                    mode=N.MODE_PLAIN, graph=True):
    """Whole path for a batch of reads: per-stage launches (MODE_PLAIN) or
    the fused kernel with the per-stage kernels taking its spilled reads
    (MODE_FUSED)."""
    index._check_owner()
    ctx = index.ctx
    seqs, offsets, lens = _seq_args(seqs, offsets, lens)
    n = len(lens)
    opt = _map_options(rescue_level, hard_cutoff, mode, graph)
    return _call_with_nam_out(ctx, n, lambda o: ctx.L.gis_find_nams_batch(
        ctx.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), n, C.byref(opt), C.byref(o)))


class _PinnedBlock:
    """A block of pinned host memory (gis_host_alloc), freed with its last
    array."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        _check(N.lib().gis_host_alloc(nbytes, C.byref(p)))
        self.p, self.nbytes = p.value, nbytes

    def __del__(self):
        try:
            N.lib().gis_host_free(self.p)
        except Exception:
            pass


def pinned_empty(n, dtype):
    """An uninitialised numpy array of n elements in pinned host memory:
    copies between it and HBM run asynchronously (gis_find_nams_stream)."""
    dtype = np.dtype(dtype)
    nbytes = max(1, int(n) * dtype.itemsize)
    blk = _PinnedBlock(nbytes)
    raw = (C.c_uint8 * nbytes).from_address(blk.p)
    raw._owner = blk
    return np.frombuffer(raw, dtype=np.uint8, count=int(n) * dtype.itemsize).view(dtype)


def pinned_copy(a):
    """A pinned copy of a contiguous numpy array."""
    a = np.ascontiguousarray(a)
    out = pinned_empty(a.size, a.dtype).reshape(a.shape)
    out[...] = a
    return out


def find_nams_stream(index, seqs, offsets, lens, batch_reads=65536, lanes=2, rescue_level=2, hard_cutoff=1000,
                     mode=N.MODE_PLAIN, graph=True, out=None, pinned=True):
    """Host-to-host mapping of many reads (gis_find_nams_stream): batches of
    batch_reads, up to `lanes` in flight, each batch's reads copied to HBM,
    mapped and its NAM arrays copied back. out: the MapResult of an earlier
    call over as many reads, whose arrays are reused; else fresh arrays are
    allocated, pinned unless pinned=False. -> (MapResult, per-batch latency
    in ms from the first copy's issue to the last copy's completion)."""
    alloc = pinned_empty if pinned else (lambda n, dt: np.empty(max(int(n), 1), dtype=dt))
    index._check_owner()
    ctx = index.ctx
    seqs, offsets, lens = _seq_args(seqs, offsets, lens)
    n = len(lens)
    opt = _map_options(rescue_level, hard_cutoff, mode, graph)
    nb = max(1, -(-n // batch_reads))
    ms = np.zeros(nb, dtype=np.float64)
    full = out._full if out is not None and len(out._full[0]) == n + 1 else None
    cap = len(full[1]) if full is not None else max(16, 12 * n)
    while True:
        if full is not None:
            offs, nams, frac, flags = full
        else:
            offs, nams = alloc(n + 1, np.uint64), alloc(cap, NAM)
            frac, flags = alloc(max(n, 1), np.float32), alloc(max(n, 1), np.uint8)
        o = N.NamOut()
        o.nam_offsets, o.nams, o.nam_capacity = offs.ctypes.data, nams.ctypes.data, cap
        o.nonrepetitive_fraction, o.flags = frac.ctypes.data, flags.ctypes.data
        code = ctx.L.gis_find_nams_stream(ctx.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), n, batch_reads,
                                          lanes, C.byref(opt), C.byref(o), N.ptr(ms))
        if code == 5 and o.n_nams > cap:
            cap, full = int(o.n_nams), None
            continue
        ctx.check(code)
        res = MapResult(offs, nams[: o.n_nams], frac[:n], flags[:n])
        res._full = (offs, nams, frac, flags)
        return res, ms


def stream_copy_stats(index):
    """How find_nams_stream's copies ran on the index's context
    (gis_stream_copy_stats): the SDMA engine bits per direction (0 = through
    HIP) and the counts of copies run on them, through HIP, and refused."""
    st = N.CopyStats()
    index.ctx.check(index.ctx.L.gis_stream_copy_stats(index.ctx.h, C.byref(st)))
    return st.as_dict()


def hits_batch(index, seqs, offsets, lens, rescue_level=2, hard_cutoff=1000):
    index._check_owner()
    ctx = index.ctx
    seqs, offsets, lens = _seq_args(seqs, offsets, lens)
    n = len(lens)
    opt = _map_options(rescue_level, hard_cutoff)
    cap = max(64, 64 * n)
    while True:
        oo = np.empty(n + 1, dtype=np.uint64)
        out = np.empty(cap, dtype=HIT)
        nout = C.c_uint64()
        code = ctx.L.gis_hits_batch(ctx.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), n,
                                    C.byref(opt), N.ptr(oo), N.ptr(out), cap, C.byref(nout))
        if code == 5 and nout.value > cap:
            cap = int(nout.value)
            continue
        ctx.check(code)
        return oo, out[: nout.value]


def _qrs_call(index, query_randstrobes_list, rescue, rescue_cutoff, hard_cutoff):
    index._check_owner()
    ctx = index.ctx
    arrs = [np.ascontiguousarray(q, dtype=QUERY_RANDSTROBE) for q in query_randstrobes_list]
    qoff = np.zeros(len(arrs) + 1, dtype=np.uint64)
    qoff[1:] = np.cumsum([len(a) for a in arrs])
    allq = np.concatenate(arrs) if arrs else np.empty(0, dtype=QUERY_RANDSTROBE)
    allq = np.ascontiguousarray(allq)
    n = len(arrs)
    return _call_with_nam_out(ctx, n, lambda o: ctx.L.gis_find_nams_qrs(
        ctx.h, N.ptr(allq), N.ptr(qoff), n, rescue, rescue_cutoff, hard_cutoff, C.byref(o)))


def find_nams(query_randstrobes, index):
    """find_nams (step2/nam.hpp:11-14): (nonrepetitive fraction, NAMs in
    canonical order)."""
    r = _qrs_call(index, [query_randstrobes], 0, 0, 1000)
    return float(r.nonrepetitive_fraction[0]), r.read_nams(0)


def find_nams_rescue(query_randstrobes, index, rescue_cutoff, hard_cutoff=1000):
    """find_nams_rescue (step2/nam.hpp:16-20): NAMs in canonical order."""
    return _qrs_call(index, [query_randstrobes], 1, rescue_cutoff, hard_cutoff).read_nams(0)


def _pair_params(dropoff_threshold=0.5, max_tries=20, mu=300.0, sigma=100.0, joint=True):
    pp = N.PairParams()
    pp.dropoff_threshold, pp.max_tries, pp.mu, pp.sigma = dropoff_threshold, max_tries, mu, sigma
    pp.flags = N.PAIRS_JOINT if joint else 0
    return pp


class PairResult:
    """The pair step's output for n pairs (mate 1 of pair p is list p, mate 2
    list n + p): NAM lists in the reference order, types 0..4, pair ids by
    type, top_dropoff and MAPQ per list, joint NAM pairs per pair."""

    def __init__(self, n_pairs, nam_offsets, nams, types, type_ids, type_counts, top_dropoff, mapq,
                 joint_offsets, joint):
        self.n_pairs = n_pairs
        self.nam_offsets, self.nams = nam_offsets, nams
        self.types, self.type_ids, self.type_counts = types, type_ids, type_counts
        self.top_dropoff, self.mapq = top_dropoff, mapq
        self.joint_offsets, self.joint = joint_offsets, joint

    def list_nams(self, r):
        return self.nams[int(self.nam_offsets[r]): int(self.nam_offsets[r + 1])]

    def pair_joint(self, p):
        return self.joint[int(self.joint_offsets[p]): int(self.joint_offsets[p + 1])]

    def type_lists(self):
        """The five per-type pair id lists (step2/main.cu:4641-4646)."""
        out, at = [], 0
        for c in self.type_counts:
            out.append(self.type_ids[at: at + c])
            at += c
        return out


def _call_with_pair_out(ctx, n_pairs, n_nams, fn, joint_cap):
    while True:
        offs = np.empty(2 * n_pairs + 1, dtype=np.uint64)
        nams = np.empty(max(n_nams, 1), dtype=NAM)
        types = np.empty(max(n_pairs, 1), dtype=np.uint8)
        tids = np.empty(max(n_pairs, 1), dtype=np.uint32)
        drop = np.empty(max(2 * n_pairs, 1), dtype=np.float32)
        mapq = np.empty(max(2 * n_pairs, 1), dtype=np.uint8)
        joff = np.empty(n_pairs + 1, dtype=np.uint64)
        joint = np.empty(max(joint_cap, 1), dtype=N.NAM_PAIR)
        o = N.PairOut()
        o.nam_offsets, o.nams, o.nam_capacity = offs.ctypes.data, nams.ctypes.data, n_nams
        o.types, o.type_ids = types.ctypes.data, tids.ctypes.data
        o.top_dropoff, o.mapq = drop.ctypes.data, mapq.ctypes.data
        o.joint_offsets, o.joint, o.joint_capacity = joff.ctypes.data, joint.ctypes.data, joint_cap
        code = fn(o)
        if code == 5 and (o.n_joint > joint_cap or o.n_nams > n_nams):
            joint_cap, n_nams = max(joint_cap, int(o.n_joint)), max(n_nams, int(o.n_nams))
            continue
        ctx.check(code)
        return PairResult(n_pairs, offs, nams[: o.n_nams], types[:n_pairs], tids[:n_pairs],
                          list(o.type_counts), drop[: 2 * n_pairs], mapq[: 2 * n_pairs], joff,
                          joint[: o.n_joint])


def pairs_from_nams(ctx, nam_offsets, nams, n_pairs, dropoff_threshold=0.5, max_tries=20, mu=300.0,
                    sigma=100.0, joint=True):
    """The pair step on NAM lists in host memory (2 n_pairs lists)."""
    nam_offsets = np.ascontiguousarray(nam_offsets, dtype=np.uint64)
    nams = np.ascontiguousarray(nams, dtype=NAM)
    pp = _pair_params(dropoff_threshold, max_tries, mu, sigma, joint)
    return _call_with_pair_out(ctx, n_pairs, len(nams), lambda o: ctx.L.gis_pairs_from_nams(
        ctx.h, N.ptr(nam_offsets), N.ptr(nams), n_pairs, C.byref(pp), C.byref(o)),
        4 * n_pairs + 16)


class Mapper:
    """Device-resident batches: upload reads, run the path, download NAMs."""

    def __init__(self, index, max_reads, max_bases, share=None, own_stream=False):
        """share: another Mapper whose device workspace (and stream) this one
        runs on (batches run one at a time; download before running another).
        own_stream: without share, a workspace on a stream of its own, so that
        launch() of this mapper overlaps launches of mappers on other streams."""
        index._check_owner()
        self.index, self.ctx = index, index.ctx
        h = C.c_void_p()
        if share is None and own_stream:
            self.ctx.check(self.ctx.L.gis_batch_create_async(self.ctx.h, max_reads, max_bases, C.byref(h)))
        elif share is None:
            self.ctx.check(self.ctx.L.gis_batch_create(self.ctx.h, max_reads, max_bases, C.byref(h)))
        else:
            self._share = share
            self.ctx.check(self.ctx.L.gis_batch_create_shared(self.ctx.h, max_reads, max_bases, share.h,
                                                              C.byref(h)))
        self.h = h

    def upload(self, seqs, offsets, lens):
        seqs, offsets, lens = _seq_args(seqs, offsets, lens)
        self._keep = (seqs, offsets, lens)
        self.n = len(lens)
        self.ctx.check(self.ctx.L.gis_batch_upload(self.h, N.ptr(seqs), N.ptr(offsets),
                                                   N.ptr(lens), self.n))

    def run(self, rescue_level=2, hard_cutoff=1000, mode=N.MODE_PLAIN, graph=True):
        self.index._check_owner()
        opt = _map_options(rescue_level, hard_cutoff, mode, graph)
        self.ctx.check(self.ctx.L.gis_batch_run(self.h, C.byref(opt)))

    def launch(self, rescue_level=2, hard_cutoff=1000, mode=N.MODE_PLAIN, graph=True):
        """Enqueue the run() and return; wait() completes it."""
        self.index._check_owner()
        opt = _map_options(rescue_level, hard_cutoff, mode, graph)
        self.ctx.check(self.ctx.L.gis_batch_launch(self.h, C.byref(opt)))

    def wait(self):
        self.ctx.check(self.ctx.L.gis_batch_wait(self.h))

    def nam_count(self):
        v = C.c_uint64()
        _check(self.ctx.L.gis_batch_nam_count(self.h, C.byref(v)), None)
        return v.value

    def download(self):
        cap = self.nam_count()
        return _call_with_nam_out(self.ctx, self.n,
                                  lambda o: self.ctx.L.gis_batch_download(self.h, C.byref(o)), cap)

    def pairs(self, dropoff_threshold=0.5, max_tries=20, mu=300.0, sigma=100.0, joint=True):
        """Run the pair step on the last run's NAMs (reads [0, n) are mate 1,
        [n, 2n) mate 2) and download it -> PairResult."""
        pp = _pair_params(dropoff_threshold, max_tries, mu, sigma, joint)
        self.ctx.check(self.ctx.L.gis_batch_pairs(self.h, C.byref(pp)))
        return self.pairs_download()

    def pairs_download(self):
        n_pairs = self.n // 2
        return _call_with_pair_out(self.ctx, n_pairs, self.nam_count(), lambda o: self.ctx.L.gis_batch_pairs_download(
            self.h, C.byref(o)), 4 * n_pairs + 16)

    def autotune(self, rescue_level=2, hard_cutoff=1000):
        """Time every fused-form configuration on the uploaded reads and keep
        the fastest in the context -> (config dict, best ms, candidates tried)."""
        opt = _map_options(rescue_level, hard_cutoff, N.MODE_FUSED, False)
        best, ms, n = N.FusedConfig(), C.c_double(), C.c_uint32()
        self.ctx.check(self.ctx.L.gis_batch_autotune(self.h, C.byref(opt), C.byref(best), C.byref(ms),
                                                     C.byref(n)))
        return best.as_dict(), ms.value, n.value

    def lookup_bytes(self):
        """Of the last MODE_PLAIN run: (algorithmic table bytes of the lookup
        stage in the reference search's byte model, lookups, bytes the lookup
        kernel streamed beside the table)."""
        b, q, s = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.gis_batch_lookup_bytes(self.h, C.byref(b), C.byref(q)))
        self.ctx.check(self.ctx.L.gis_batch_lookup_streams(self.h, C.byref(s)))
        return b.value, q.value, s.value

    def stage_times(self):
        ms = (C.c_double * 16)()
        n = C.c_int()
        names = C.create_string_buffer(512)
        self.ctx.check(self.ctx.L.gis_batch_stage_times(self.h, ms, 16, C.byref(n), names, 512))
        keys = names.value.decode().split(",") if n.value else []
        return dict(zip(keys, list(ms)[: n.value]))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.gis_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (continuous batching) over a resident index: submit
    requests of any size, wait for their NAMs, or replay a timed trace.
    Requests are admitted by an HBM budget, coalesced into steps replayed
    from captured HIP graphs, with up to `slots` steps in flight."""

    def __init__(self, index, slots=6, max_step_reads=131072, hbm_budget=0, mode=N.MODE_PLAIN,
                 graph=True, rescue_level=2, hard_cutoff=1000):
        index._check_owner()
        self.index, self.ctx = index, index.ctx
        cfg = N.ServerConfig()
        cfg.slots, cfg.max_step_reads, cfg.hbm_budget = slots, max_step_reads, hbm_budget
        cfg.options = _map_options(rescue_level, hard_cutoff, mode, graph)
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.gis_server_create(self.ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, seqs, offsets, lens):
        seqs, offsets, lens = _seq_args(seqs, offsets, lens)
        rid = C.c_uint64()
        _check(self.ctx.L.gis_server_submit(self.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), len(lens),
                                            C.byref(rid)), None)
        return rid.value, len(lens)

    def wait(self, request, timeout_ms=-1):
        """-> (MapResult, latency ms) of a request returned by submit()."""
        rid, n = request
        lat = C.c_double()
        cap = max(16, 16 * n)
        while True:
            o, offs, nams, frac, flags = _nam_out(n, cap)
            code = self.ctx.L.gis_server_wait(self.h, rid, timeout_ms, C.byref(o), C.byref(lat))
            if code == 5 and o.n_nams > cap:
                cap = int(o.n_nams)
                continue
            _check(code, None)
            return MapResult(offs, nams[: o.n_nams], frac[:n], flags[:n]), lat.value

    def replay(self, seqs, offsets, lens, arrival_s, first, count):
        """Timed trace -> (latency ms per request, wall seconds)."""
        seqs, offsets, lens = _seq_args(seqs, offsets, lens)
        arrival_s = np.ascontiguousarray(arrival_s, dtype=np.float64)
        first = np.ascontiguousarray(first, dtype=np.uint32)
        count = np.ascontiguousarray(count, dtype=np.uint32)
        lat = np.empty(max(1, len(arrival_s)), dtype=np.float64)
        wall = C.c_double()
        _check(self.ctx.L.gis_server_replay(self.h, N.ptr(seqs), N.ptr(offsets), N.ptr(lens), len(arrival_s),
                                            N.ptr(arrival_s), N.ptr(first), N.ptr(count), N.ptr(lat),
                                            C.byref(wall)), None)
        return lat[: len(arrival_s)], wall.value

    def submit_fastq(self, reader, keep=True):
        """Take the next chunk of a FastqReader and submit it as one request
        -> (request, FastqChunk), or (None, None) at the end of the input.
        keep=False skips copying the chunk's reads and names to Python and
        returns the record count in place of the FastqChunk."""
        rid = C.c_uint64()
        c = N.FastqChunk()
        code = self.ctx.L.gis_server_submit_fastq(self.h, reader.h, C.byref(rid), C.byref(c))
        if code:
            raise GisError(code, reader._last_error() if code in (7, 11) else "")
        if c.n_records == 0:
            return None, None
        return (rid.value, c.n_reads), (FastqChunk._from_c(c) if keep else c.n_records)

    def stats(self):
        st = N.ServerStats()
        _check(self.ctx.L.gis_server_stats_get(self.h, C.byref(st)), None)
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.gis_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- FASTQ ingestion ----------------

class FastqChunk:
    """One chunk of records: concatenated reads (with two files, mate 1 of
    record i is read i and mate 2 read n_records + i) and their names."""

    def __init__(self, index, first_record, n_records, seqs, offsets, lens, names, name_offsets):
        self.index, self.first_record, self.n_records = index, first_record, n_records
        self.seqs, self.offsets, self.lens = seqs, offsets, lens
        self.names, self.name_offsets = names, name_offsets

    @classmethod
    def _from_c(cls, c):
        n = c.n_reads

        def arr(p, count, dt):
            if not count:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(count,)).copy()
        noff = arr(c.name_offsets, n + 1, np.uint64)
        return cls(c.index, c.first_record, c.n_records, arr(c.seqs, c.bases, np.uint8),
                   arr(c.offsets, n, np.uint64), arr(c.lens, n, np.uint32),
                   arr(c.names, int(noff[-1]) if n else 0, np.uint8), noff)

    @property
    def n_reads(self):
        return len(self.lens)

    def read(self, i):
        o = int(self.offsets[i])
        return self.seqs[o: o + int(self.lens[i])].tobytes()

    def name(self, i):
        return self.names[int(self.name_offsets[i]): int(self.name_offsets[i + 1])].tobytes()


class FastqReader:
    """Chunked FASTQ reader (one file, or two mates files): a reader thread
    and parser threads work ahead; iterate to get FastqChunk objects in file
    order."""

    def __init__(self, path1, path2=None, chunk_bytes=0, threads=0):
        h = C.c_void_p()
        code = N.lib().gis_fastq_open(os.fsencode(path1), os.fsencode(path2) if path2 else None, chunk_bytes,
                                      threads, C.byref(h))
        if code:
            raise GisError(code, f"cannot open {path1}" + (f" / {path2}" if path2 else ""))
        self.h = h

    def _last_error(self):
        return N.lib().gis_fastq_last_error(self.h).decode()

    def next_chunk(self):
        c = N.FastqChunk()
        code = N.lib().gis_fastq_next(self.h, C.byref(c))
        if code:
            raise GisError(code, self._last_error())
        return FastqChunk._from_c(c) if c.n_records else None

    def __iter__(self):
        while True:
            c = self.next_chunk()
            if c is None:
                return
            yield c

    def close(self):
        if getattr(self, "h", None):
            N.lib().gis_fastq_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- synthetic inputs ----------------

def synth_genome(total_length, n_contigs=1, gc=0.5, repeat_fraction=0.0, n_families=0,
                 family_len=(1000, 10000), copies=(2, 50), divergence=(0.01, 0.05),
                 tandem_fraction=0.0, tandem_unit=(2, 50), contig_len=(0, 0), seed=1):
    """Deterministic genome -> (uint8 array of ACGT, u64 contig lengths)."""
    s = N.GenomeSpec()
    s.total_length, s.n_contigs, s.gc = total_length, n_contigs, gc
    s.repeat_fraction, s.n_families = repeat_fraction, n_families
    s.family_min_len, s.family_max_len = family_len
    s.copies_min, s.copies_max = copies
    s.div_min, s.div_max = divergence
    s.tandem_fraction = tandem_fraction
    s.tandem_unit_min, s.tandem_unit_max = tandem_unit
    s.contig_min_len, s.contig_max_len = contig_len
    s.seed = seed
    seq = np.empty(total_length, dtype=np.uint8)
    lens = np.empty(n_contigs, dtype=np.uint64)
    _check(N.lib().gis_synth_genome(C.byref(s), N.ptr(seq), N.ptr(lens)))
    return seq, lens


def synth_reads(genome, contig_lens, n_reads, lengths=(150,), weights=(1.0,), paired=False,
                sub_rate=0.001, indel_rate=0.0001, rc_fraction=0.5, insert=(400.0, 50.0), seed=2):
    """Deterministic reads -> (seqs, offsets, lens, truth dict)."""
    s = N.ReadSpec()
    s.n_reads, s.paired = n_reads, 1 if paired else 0
    for i, (L, w) in enumerate(zip(lengths, weights)):
        s.lengths[i], s.weights[i] = L, w
    s.n_lengths = len(lengths)
    s.sub_rate, s.indel_rate, s.rc_fraction = sub_rate, indel_rate, rc_fraction
    s.insert_mean, s.insert_sd = insert
    s.seed = seed
    n_out = n_reads * (2 if paired else 1)
    cap = n_out * max(lengths)
    seqs = np.empty(max(cap, 1), dtype=np.uint8)
    offsets = np.empty(n_out, dtype=np.uint64)
    lens = np.empty(n_out, dtype=np.uint32)
    tc = np.empty(n_out, dtype=np.uint32)
    tp = np.empty(n_out, dtype=np.uint32)
    tr = np.empty(n_out, dtype=np.uint8)
    contig_lens = np.ascontiguousarray(contig_lens, dtype=np.uint64)
    genome = np.ascontiguousarray(genome, dtype=np.uint8)
    _check(N.lib().gis_synth_reads(C.byref(s), N.ptr(genome), N.ptr(contig_lens), len(contig_lens),
                                   N.ptr(seqs), N.ptr(offsets), N.ptr(lens), N.ptr(tc), N.ptr(tp),
                                   N.ptr(tr)))
    used = int(offsets[-1] + lens[-1]) if n_out else 0
    return seqs[:max(used, 1)], offsets, lens, {"contig": tc, "pos": tp, "rc": tr}
