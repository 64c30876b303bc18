// Core arithmetic of the seeding path shared by the gfx950 kernels and the
// host-side index builder (compiled by hipcc for both).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gis.h"

#define GIS_HD __host__ __device__ __forceinline__

namespace gis {

// xxh64 of a single 64-bit word with seed 0 (step2/hash.hpp:106-119): the
// XXH64 finalisation for an 8-byte input. On gfx950 each 64-bit multiply is a
// short v_mul_lo/v_mul_hi/v_mad sequence.
GIS_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

GIS_HD uint64_t xxh64_u64(uint64_t v) {
    constexpr uint64_t K1 = 0x9E3779B185EBCA87ULL, K2 = 0xC2B2AE3D27D4EB4FULL,
                       K3 = 0x165667B19E3779F9ULL, K4 = 0x85EBCA77C2B2AE63ULL,
                       K5 = 0x27D4EB2F165667C5ULL;
    uint64_t acc = (K5 + 8) ^ (rotl64(v * K2, 31) * K1);
    acc = rotl64(acc, 27) * K1 + K4;
    acc = (acc ^ (acc >> 33)) * K2;
    acc = (acc ^ (acc >> 29)) * K3;
    return acc ^ (acc >> 32);
}

// ASCII -> 2-bit base, 4 for anything but ACGTU (either case),
// step2/randstrobes.hpp:22-43.
GIS_HD int nt4(unsigned char c) {
    unsigned u = (unsigned)(c | 0x20);  // fold case for letters
    int r = 4;
    r = u == 'a' ? 0 : r;
    r = u == 'c' ? 1 : r;
    r = u == 'g' ? 2 : r;
    r = (u == 't' || u == 'u') ? 3 : r;
    // non-letters that fold onto a letter code (e.g. 0x01|0x20 = '!') are not bases
    return ((c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z')) ? r : 4;
}

// Window state of the canonical-syncmer test. A window of w = k-s+1 s-mer
// hashes has a tracked minimum position (offset 0..w-1) or NONE; the update
// rules are those of SyncmerIterator::next (step2/randstrobes.cu:38-63):
//   first window of a segment: leftmost minimum (NONE if every hash is ~0);
//   slide, tracked minimum left the window: rightmost minimum of the new window;
//   slide, tracked NONE: newest if its hash < ~0, else NONE;
//   slide otherwise: newest if strictly below the tracked value, else keep.
// A window whose minimum is unique and below ~0 has that minimum as its state
// whatever came before, so the recurrence only matters across windows with
// tied minima.
constexpr uint8_t ST_NONE = 0xFD, ST_UNKNOWN = 0xFE, ST_INVALID = 0xFF;

// state of window x (x >= 1) after sliding from window x-1 in state r;
// h points at the s-mer hashes, window x holds h[x .. x+w-1].
template <typename H>
GIS_HD uint8_t slide(uint8_t r, const H &h, int x, int w) {
    const uint64_t v = h[x + w - 1];
    if (r == 0) {
        int best = w - 1;
        uint64_t bv = v;
        for (int o = w - 2; o >= 0; o--) {
            uint64_t c = h[x + o];
            if (c < bv) { bv = c; best = o; }
        }
        return (uint8_t)best;
    }
    if (r == ST_NONE) return v < UINT64_MAX ? (uint8_t)(w - 1) : ST_NONE;
    const uint64_t old = h[x - 1 + r];
    return v < old ? (uint8_t)(w - 1) : (uint8_t)(r - 1);
}

// State of a window computable without history: segment-first windows and
// windows with a unique minimum below ~0. Returns ST_UNKNOWN otherwise.
template <typename H>
GIS_HD uint8_t anchor_state(const H &h, int p, int w, bool seg_first) {
    uint64_t minv = UINT64_MAX;
    int first = 0, eq = 0;
    for (int o = 0; o < w; o++) {
        uint64_t v = h[p + o];
        if (v < minv) { minv = v; first = o; eq = 1; }
        else if (v == minv) eq++;
    }
    if (seg_first) return minv == UINT64_MAX ? ST_NONE : (uint8_t)first;
    if (eq == 1 && minv != UINT64_MAX) return (uint8_t)first;
    return ST_UNKNOWN;
}

// Device copy of the seeding parameters
struct SeedParams {
    int k, s, t, w;          // w = k - s + 1 s-mers per k-mer
    uint32_t w_min, w_max;
    uint64_t max_dist_u;     // max_dist as added to a position (unsigned)
    uint64_t q;
    uint64_t kmask, smask;
};

inline SeedParams make_seed_params(const gis_params &p) {
    SeedParams sp;
    sp.k = p.k;
    sp.s = p.s;
    sp.t = p.t_syncmer;
    sp.w = p.k - p.s + 1;
    sp.w_min = p.w_min;
    sp.w_max = p.w_max;
    sp.max_dist_u = (uint64_t)(uint32_t)p.max_dist;
    sp.q = p.q;
    sp.kmask = p.k >= 32 ? ~0ULL : (1ULL << (2 * p.k)) - 1;
    sp.smask = p.s >= 32 ? ~0ULL : (1ULL << (2 * p.s)) - 1;
    return sp;
}

// Score of a finished NAM, step2/nam.cu:459-464
GIS_HD float nam_score(int qs, int qe, int rs, int re, int n_hits) {
    int qspan = qe - qs, rspan = re - rs;
    int mx = qspan > rspan ? qspan : rspan, mn = qspan < rspan ? qspan : rspan;
    return (2 * mn - mx) > 0 ? (float)(n_hits * (2 * mn - mx)) : 1.0f;
}

}  // namespace gis
