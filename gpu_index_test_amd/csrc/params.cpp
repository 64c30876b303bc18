// IndexParameters of the seeding path (step2/indexparameters.{hpp,cpp}).
#include <cmath>
#include <cstdint>

#include "../../include/gis.h"

namespace {

// Read-length profiles, step2/indexparameters.cpp:37-45:
// {canonical read length, longest read length it covers, k, s - k, l, u}
struct LengthProfile {
    int canonical, upto, k, s_minus_k, l, u;
};
constexpr LengthProfile kProfiles[] = {
    {50, 90, 20, -4, -3, 2},    {100, 110, 20, -4, -2, 2},  {125, 135, 20, -4, -1, 4},
    {150, 175, 20, -4, 1, 7},   {250, 275, 20, -4, 4, 13},  {300, 375, 22, -4, 2, 12},
    {400, 2147483647, 23, -6, 2, 12},
};

}  // namespace

// IndexParameters(r, k, s, l, u, q, max_dist) with the checks of
// SyncmerParameters::verify and RandstrobeParameters::verify
// (step2/indexparameters.hpp:24-34, 61-68, 80-85).
extern "C" int gis_params_make(int canonical_read_length, int k, int s, int l, int u, int q,
                               int max_dist, gis_params *out) {
    if (!out) return GIS_ERR_BAD_PARAMETER;
    if (k <= 7 || k > 32) return GIS_ERR_BAD_PARAMETER;  // "k not in [8,32]"
    if (s > k || s < 1) return GIS_ERR_BAD_PARAMETER;    // s >= 1: an empty s-mer has no meaning
    if ((k - s) % 2 != 0) return GIS_ERR_BAD_PARAMETER;
    const int per_window = k / (k - s + 1);
    const int wmin = per_window + l;
    const unsigned w_min = wmin > 0 ? (unsigned)wmin : 0u;
    const unsigned w_max = (unsigned)(per_window + u);
    if (max_dist > 255) return GIS_ERR_BAD_PARAMETER;
    if (w_min > w_max) return GIS_ERR_BAD_PARAMETER;
    gis_params p{};
    p.canonical_read_length = canonical_read_length;
    p.k = k;
    p.s = s;
    p.t_syncmer = (k - s) / 2 + 1;
    p.l = l;
    p.u = u;
    p.q = (uint64_t)(int64_t)q;
    p.max_dist = max_dist;
    p.w_min = w_min;
    p.w_max = w_max;
    *out = p;
    return GIS_OK;
}

// IndexParameters::from_read_length, step2/indexparameters.cpp:51-83
extern "C" int gis_params_from_read_length(int read_length, int k, int s, int l, int u, int c,
                                           int max_seed_len, gis_params *out) {
    int canonical = 50;
    for (const LengthProfile &p : kProfiles) {
        if (read_length > p.upto) continue;
        if (k == GIS_DEFAULT) k = p.k;
        if (s == GIS_DEFAULT) s = k + p.s_minus_k;
        if (l == GIS_DEFAULT) l = p.l;
        if (u == GIS_DEFAULT) u = p.u;
        canonical = p.canonical;
        break;
    }
    int max_dist;
    if (max_seed_len == GIS_DEFAULT) {
        max_dist = canonical - 70 > k ? canonical - 70 : k;
        max_dist = max_dist < 255 ? max_dist : 255;
    } else {
        max_dist = max_seed_len - k;
    }
    const int bits_c = c == GIS_DEFAULT ? 8 : c;
    const int q = (int)(std::pow(2.0, (double)bits_c) - 1.0);
    return gis_params_make(canonical, k, s, l, u, q, max_dist, out);
}

extern "C" void gis_map_options_default(gis_map_options *opt) {
    if (!opt) return;
    opt->rescue_level = 2;   // CommandLineOptions::rescue_level, step2/cmdline.hpp:56
    opt->hard_cutoff = 1000; // find_nams_rescue, step2/nam.cu:964
    opt->mode = GIS_MODE_PLAIN;
    opt->flags = 0;
}
