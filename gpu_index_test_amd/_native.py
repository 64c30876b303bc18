"""ctypes binding of libgis.so (the C ABI declared in include/gis.h).

The library is built in-tree (``make -C gpu_index_test_amd``); loading fails
loudly when it is missing -- there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgis.so")

GIS_DEFAULT = -(2 ** 31)
GIS_NOT_FOUND = 2 ** 64 - 1
GIS_MAX_READ_LEN = 65535
FLAG_RESCUED = 1
FLAG_NO_RANDSTROBES = 2
MODE_PLAIN = 0
MODE_FUSED = 1
RUN_NO_GRAPH = 1

STATUS = {
    0: "GIS_OK", 1: "GIS_ERR_BAD_PARAMETER", 2: "GIS_ERR_INVALID_INDEX", 3: "GIS_ERR_HIP",
    4: "GIS_ERR_NOMEM", 5: "GIS_ERR_CAPACITY", 6: "GIS_ERR_STATE", 7: "GIS_ERR_IO",
    8: "GIS_ERR_BUDGET", 9: "GIS_ERR_CANCELLED", 10: "GIS_ERR_INVALID_FASTA",
    11: "GIS_ERR_INVALID_INPUT",
}

# numpy views of the C structs (byte layouts of the reference types)
QUERY_RANDSTROBE = np.dtype([("hash", "<u8"), ("start", "<u4"), ("end", "<u4"),
                             ("is_reverse", "u1"), ("pad", "u1", (7,))])
REF_RANDSTROBE = np.dtype([("hash", "<u8"), ("position", "<u4"), ("packed", "<u4")])
SYNCMER = np.dtype([("hash", "<u8"), ("position", "<u8")])
NAM = np.dtype([("nam_id", "<i4"), ("query_start", "<i4"), ("query_end", "<i4"),
                ("query_prev_hit_startpos", "<i4"), ("ref_start", "<i4"), ("ref_end", "<i4"),
                ("ref_prev_hit_startpos", "<i4"), ("n_hits", "<i4"), ("ref_id", "<i4"),
                ("score", "<f4"), ("is_rc", "u1"), ("pad", "u1", (3,))])
HIT = np.dtype([("query_start", "<i4"), ("query_end", "<i4"), ("ref_start", "<i4"),
                ("ref_end", "<i4"), ("ref_id", "<i4"), ("is_rc", "<i4")])
NAM_PAIR = np.dtype([("score", "<i4"), ("i1", "<i4"), ("i2", "<i4")])
PAIRS_JOINT = 1
assert QUERY_RANDSTROBE.itemsize == 24 and REF_RANDSTROBE.itemsize == 16
assert NAM.itemsize == 44 and SYNCMER.itemsize == 16 and HIT.itemsize == 24


class Params(C.Structure):
    _fields_ = [("canonical_read_length", C.c_int32), ("k", C.c_int32), ("s", C.c_int32),
                ("t_syncmer", C.c_int32), ("l", C.c_int32), ("u", C.c_int32),
                ("q", C.c_uint64), ("max_dist", C.c_int32), ("w_min", C.c_uint32),
                ("w_max", C.c_uint32), ("reserved_", C.c_int32)]


class IndexInfo(C.Structure):
    _fields_ = [("n_randstrobes", C.c_uint64), ("bits", C.c_int32), ("filter_cutoff", C.c_uint32),
                ("params", Params), ("distinct_strobemers", C.c_uint64),
                ("tot_occur_once", C.c_uint64), ("tot_high_ab", C.c_uint64),
                ("tot_mid_ab", C.c_uint64), ("index_cutoff", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class NamOut(C.Structure):
    _fields_ = [("nam_offsets", C.c_void_p), ("nams", C.c_void_p), ("nam_capacity", C.c_uint64),
                ("nonrepetitive_fraction", C.c_void_p), ("flags", C.c_void_p),
                ("n_nams", C.c_uint64)]


class CopyStats(C.Structure):
    _fields_ = [("h2d_engine", C.c_uint32), ("d2h_engine", C.c_uint32), ("engine_copies", C.c_uint64),
                ("hip_copies", C.c_uint64), ("refused", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MapOptions(C.Structure):
    _fields_ = [("rescue_level", C.c_int32), ("hard_cutoff", C.c_uint32), ("mode", C.c_int32),
                ("flags", C.c_int32)]


class FusedConfig(C.Structure):
    _fields_ = [("waves_per_block", C.c_int32), ("length_split", C.c_int32), ("queue", C.c_int32),
                ("chunk", C.c_int32), ("blocks_per_cu", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_step_reads", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("options", MapOptions)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("reads", C.c_uint64), ("steps", C.c_uint64),
                ("busy_s", C.c_double), ("peak_bytes", C.c_uint64)]


class GenomeSpec(C.Structure):
    _fields_ = [("total_length", C.c_uint64), ("n_contigs", C.c_uint32), ("gc", C.c_double),
                ("repeat_fraction", C.c_double), ("n_families", C.c_uint32),
                ("family_min_len", C.c_uint32), ("family_max_len", C.c_uint32),
                ("copies_min", C.c_uint32), ("copies_max", C.c_uint32),
                ("div_min", C.c_double), ("div_max", C.c_double),
                ("tandem_fraction", C.c_double), ("tandem_unit_min", C.c_uint32),
                ("tandem_unit_max", C.c_uint32), ("contig_min_len", C.c_uint32),
                ("contig_max_len", C.c_uint32), ("seed", C.c_uint64)]


class ReadSpec(C.Structure):
    _fields_ = [("n_reads", C.c_uint64), ("paired", C.c_int32), ("lengths", C.c_uint32 * 8),
                ("weights", C.c_double * 8), ("n_lengths", C.c_uint32), ("sub_rate", C.c_double),
                ("indel_rate", C.c_double), ("rc_fraction", C.c_double),
                ("insert_mean", C.c_double), ("insert_sd", C.c_double), ("seed", C.c_uint64)]


class PairParams(C.Structure):
    _fields_ = [("dropoff_threshold", C.c_float), ("max_tries", C.c_int32), ("mu", C.c_float),
                ("sigma", C.c_float), ("flags", C.c_uint32)]


class PairOut(C.Structure):
    _fields_ = [("nam_offsets", C.c_void_p), ("nams", C.c_void_p), ("nam_capacity", C.c_uint64),
                ("types", C.c_void_p), ("type_ids", C.c_void_p), ("type_counts", C.c_uint32 * 5),
                ("top_dropoff", C.c_void_p), ("mapq", C.c_void_p), ("joint_offsets", C.c_void_p),
                ("joint", C.c_void_p), ("joint_capacity", C.c_uint64), ("n_nams", C.c_uint64),
                ("n_joint", C.c_uint64)]


class FastqChunk(C.Structure):
    _fields_ = [("n_records", C.c_uint32), ("n_reads", C.c_uint32), ("seqs", C.c_void_p),
                ("offsets", C.c_void_p), ("lens", C.c_void_p), ("names", C.c_void_p),
                ("name_offsets", C.c_void_p), ("index", C.c_uint64), ("first_record", C.c_uint64),
                ("bases", C.c_uint64)]


# name -> (restype, argtypes); every name must be exported by libgis.so
_P = C.c_void_p
_U32, _U64, _I32 = C.c_uint32, C.c_uint64, C.c_int32
PROTOTYPES = {
    "gis_params_from_read_length": (C.c_int, [C.c_int] * 7 + [C.POINTER(Params)]),
    "gis_params_make": (C.c_int, [C.c_int] * 7 + [C.POINTER(Params)]),
    "gis_map_options_default": (None, [C.POINTER(MapOptions)]),
    "gis_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "gis_ctx_destroy": (None, [_P]),
    "gis_last_error": (C.c_char_p, [_P]),
    "gis_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gis_index_build": (C.c_int, [_P, _P, _P, _U32, C.POINTER(Params), C.c_int, C.c_float, C.c_int]),
    "gis_index_upload": (C.c_int, [_P, _P, _U64, _P, C.c_int, _U32, C.POINTER(Params)]),
    "gis_reference_randstrobes": (C.c_int, [_P, _P, _P, _U32, C.POINTER(Params), _P, _U64, C.POINTER(_U64)]),
    "gis_index_read_sti": (C.c_int, [_P, C.c_char_p, C.POINTER(Params)]),
    "gis_index_write_sti": (C.c_int, [_P, C.c_char_p]),
    "gis_index_info_get": (C.c_int, [_P, C.POINTER(IndexInfo)]),
    "gis_index_download": (C.c_int, [_P, _P, _P]),
    "gis_syncmers_batch": (C.c_int, [_P, _P, _P, _P, _U32, _P, _P, _U64, C.POINTER(_U64)]),
    "gis_randstrobes_batch": (C.c_int, [_P, _P, _P, _P, _U32, _P, _P, _U64, C.POINTER(_U64)]),
    "gis_lookup_batch": (C.c_int, [_P, _P, _U64, _P]),
    "gis_hits_batch": (C.c_int, [_P, _P, _P, _P, _U32, C.POINTER(MapOptions), _P, _P, _U64,
                                 C.POINTER(_U64)]),
    "gis_find_nams_batch": (C.c_int, [_P, _P, _P, _P, _U32, C.POINTER(MapOptions),
                                      C.POINTER(NamOut)]),
    "gis_find_nams_qrs": (C.c_int, [_P, _P, _P, _U32, C.c_int, _U32, _U32, C.POINTER(NamOut)]),
    "gis_batch_create": (C.c_int, [_P, _U32, _U64, C.POINTER(_P)]),
    "gis_batch_create_shared": (C.c_int, [_P, _U32, _U64, _P, C.POINTER(_P)]),
    "gis_batch_destroy": (None, [_P]),
    "gis_batch_upload": (C.c_int, [_P, _P, _P, _P, _U32]),
    "gis_batch_run": (C.c_int, [_P, C.POINTER(MapOptions)]),
    "gis_batch_create_async": (C.c_int, [_P, _U32, _U64, C.POINTER(_P)]),
    "gis_batch_launch": (C.c_int, [_P, C.POINTER(MapOptions)]),
    "gis_batch_wait": (C.c_int, [_P]),
    "gis_batch_download": (C.c_int, [_P, C.POINTER(NamOut)]),
    "gis_batch_stage_times": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int]),
    "gis_batch_lookup_bytes": (C.c_int, [_P, C.POINTER(_U64), C.POINTER(_U64)]),
    "gis_batch_lookup_streams": (C.c_int, [_P, C.POINTER(_U64)]),
    "gis_find_nams_stream": (C.c_int, [_P, _P, _P, _P, _U64, _U32, _U32, C.POINTER(MapOptions),
                                       C.POINTER(NamOut), _P]),
    "gis_stream_copy_stats": (C.c_int, [_P, C.POINTER(CopyStats)]),
    "gis_host_alloc": (C.c_int, [_U64, C.POINTER(_P)]),
    "gis_host_free": (None, [_P]),
    "gis_lookup_batch_noalloc": (C.c_int, [_P, _P, _U64, _P]),
    "gis_randstrobes_params": (C.c_int, [_P, C.POINTER(Params), _P, _P, _P, _U32, _P, _P, _U64, C.POINTER(_U64)]),
    "gis_batch_nam_count": (C.c_int, [_P, C.POINTER(_U64)]),
    "gis_fused_config_get": (C.c_int, [_P, C.POINTER(FusedConfig)]),
    "gis_fused_config_set": (C.c_int, [_P, C.POINTER(FusedConfig)]),
    "gis_batch_autotune": (C.c_int, [_P, C.POINTER(MapOptions), C.POINTER(FusedConfig),
                                     C.POINTER(C.c_double), C.POINTER(_U32)]),
    "gis_server_create": (C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "gis_server_destroy": (None, [_P]),
    "gis_server_submit": (C.c_int, [_P, _P, _P, _P, _U32, C.POINTER(_U64)]),
    "gis_server_wait": (C.c_int, [_P, _U64, C.c_int, C.POINTER(NamOut), C.POINTER(C.c_double)]),
    "gis_server_replay": (C.c_int, [_P, _P, _P, _P, _U32, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "gis_server_stats_get": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "gis_pair_params_default": (None, [C.POINTER(PairParams)]),
    "gis_batch_pairs": (C.c_int, [_P, C.POINTER(PairParams)]),
    "gis_batch_pairs_download": (C.c_int, [_P, C.POINTER(PairOut)]),
    "gis_pairs_from_nams": (C.c_int, [_P, _P, _P, _U32, C.POINTER(PairParams), C.POINTER(PairOut)]),
    "gis_fastq_open": (C.c_int, [C.c_char_p, C.c_char_p, _U64, _U32, C.POINTER(_P)]),
    "gis_fastq_next": (C.c_int, [_P, C.POINTER(FastqChunk)]),
    "gis_fastq_last_error": (C.c_char_p, [_P]),
    "gis_fastq_close": (None, [_P]),
    "gis_server_submit_fastq": (C.c_int, [_P, _P, C.POINTER(_U64), C.POINTER(FastqChunk)]),
    "gis_synth_genome": (C.c_int, [C.POINTER(GenomeSpec), _P, _P]),
    "gis_synth_reads": (C.c_int, [C.POINTER(ReadSpec), _P, _P, _U32, _P, _P, _P, _P, _P, _P]),
}

_lib = None


class GisError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{STATUS.get(code, code)}: {msg}")


def lib():
    """Load libgis.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Address of a contiguous numpy array (None for empty/None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data if a.size else None
