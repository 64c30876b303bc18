"""The allocation-free lookup (gis_lookup_batch_noalloc) through the C ABI:
its positions equal those of gis_lookup_batch on the same hashes, on two
calls in a row (the second with fewer hashes, over the array the first one
left), and more hashes than its array holds are refused."""
import ctypes as C

import numpy as np
import pytest

import gpu_index_test_amd as G
from gpu_index_test_amd import _native as N

pytestmark = pytest.mark.gpu


def _lookup(L, ctx, fn, keys):
    out = np.full(len(keys), 7, dtype=np.uint64)
    ctx.check(getattr(L, fn)(ctx.h, N.ptr(keys), len(keys), N.ptr(out)))
    return out


def test_noalloc_lookup_matches_lookup_batch(ctx):
    g, lens = G.synth_genome(800_000, n_contigs=2, gc=0.41, repeat_fraction=0.05, n_families=10,
                             family_len=(300, 2000), copies=(2, 30), seed=41)
    idx = G.StrobemerIndex(None, G.IndexParameters.from_read_length(150))
    idx.populate_arrays(g, lens)
    c, L = idx.ctx, N.lib()
    rng = np.random.default_rng(42)
    tab = idx.randstrobes["hash"]
    big = np.concatenate([tab[rng.integers(0, len(tab), 150_000)],
                          rng.integers(0, 2 ** 63, 100_000, dtype=np.uint64) * 2 + 1,
                          tab[:3], np.array([0, 2 ** 64 - 1], dtype=np.uint64)]).astype(np.uint64)
    small = big[::37].copy()
    for keys in (big, small):
        exp = _lookup(L, c, "gis_lookup_batch", keys)
        got = _lookup(L, c, "gis_lookup_batch_noalloc", keys)
        assert np.array_equal(got, exp)
        assert (exp != G.GIS_NOT_FOUND).sum() > len(keys) // 3
    assert L.gis_lookup_batch_noalloc(c.h, None, 0, None) == 0
    # one past the array's capacity (GIS_NOALLOC_LOOKUP_MAX = 2^20): refused
    # on the host before any copy
    too_many = np.zeros((1 << 20) + 1, dtype=np.uint64)
    out = np.empty_like(too_many)
    assert L.gis_lookup_batch_noalloc(c.h, N.ptr(too_many), len(too_many), N.ptr(out)) == 5
    assert b"maximum" in L.gis_last_error(c.h)
    c.close()
