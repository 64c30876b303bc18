// Deterministic synthetic genomes and reads for the seeding benchmarks.
//
// Genomes: contigs of i.i.d. bases at a given GC fraction, overlaid with
// interspersed repeat families (a random segment copied several times with
// substitutions at a per-family divergence) and tandem repeats (a short unit
// repeated over a region). Reads: uniform positions, optional reverse
// complement, substitutions and 1-bp indels; paired-end mates are drawn from a
// fragment of normal length, FR orientation.
//
// Everything is a function of the spec's seed: the same spec gives the same
// bytes on every machine.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gis.h"

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9E3779B97F4A7C15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }  // bias is irrelevant here
    uint64_t range(uint64_t lo, uint64_t hi) { return lo + below(hi - lo + 1); }
    double normal(double mu, double sd) {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return mu + sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline char base_gc(Rng &r, double gc) {
    double x = r.uniform();
    if (x < gc * 0.5) return 'G';
    if (x < gc) return 'C';
    if (x < gc + (1.0 - gc) * 0.5) return 'A';
    return 'T';
}

inline char other_base(Rng &r, char b) {
    static const char ACGT[4] = {'A', 'C', 'G', 'T'};
    for (;;) {
        char c = ACGT[r.below(4)];
        if (c != b) return c;
    }
}

inline char comp(char c) {
    switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return 'N';
    }
}

uint64_t mix(uint64_t a, uint64_t b) {
    uint64_t z = a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL);
    z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ULL;
    return z ^ (z >> 29);
}

}  // namespace

extern "C" int gis_synth_genome(const gis_genome_spec *spec, char *seq, uint64_t *contig_lens) {
    if (!spec || !seq || !contig_lens || spec->n_contigs == 0 || spec->total_length == 0)
        return GIS_ERR_BAD_PARAMETER;
    const uint32_t nc = spec->n_contigs;
    const uint64_t total = spec->total_length;
    Rng r(spec->seed);

    // contig lengths: equal split, or uniform draws in [min,max] rescaled to the total
    std::vector<uint64_t> lens(nc);
    if (spec->contig_min_len == 0 || spec->contig_max_len == 0) {
        for (uint32_t i = 0; i < nc; i++) lens[i] = total / nc + (i < total % nc ? 1 : 0);
    } else {
        std::vector<double> w(nc);
        double sw = 0;
        for (uint32_t i = 0; i < nc; i++) {
            w[i] = (double)r.range(spec->contig_min_len, spec->contig_max_len);
            sw += w[i];
        }
        uint64_t acc = 0;
        for (uint32_t i = 0; i < nc; i++) {
            lens[i] = (uint64_t)(w[i] / sw * (double)total);
            acc += lens[i];
        }
        lens[nc - 1] += total - acc;
    }
    std::vector<uint64_t> start(nc + 1, 0);
    for (uint32_t i = 0; i < nc; i++) start[i + 1] = start[i] + lens[i];
    for (uint32_t i = 0; i < nc; i++) contig_lens[i] = lens[i];

    // i.i.d. background, 1 Mbp blocks generated in parallel from per-block streams
    const uint64_t block = 1 << 20;
    const uint64_t nblocks = (total + block - 1) / block;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([&, t]() {
            for (uint64_t b = t; b < nblocks; b += nt) {
                Rng rb(mix(spec->seed, b + 1));
                uint64_t lo = b * block, hi = std::min(total, lo + block);
                for (uint64_t i = lo; i < hi; i++) seq[i] = base_gc(rb, spec->gc);
            }
        });
    }
    for (auto &x : th) x.join();

    // interspersed repeat families, placed until the repeat budget is spent
    uint64_t budget = (uint64_t)(spec->repeat_fraction * (double)total);
    if (budget > 0 && spec->n_families > 0) {
        struct Fam { std::string s; uint32_t copies; double div; };
        std::vector<Fam> fams(spec->n_families);
        double want = 0;
        for (auto &f : fams) {
            uint64_t len = r.range(spec->family_min_len, spec->family_max_len);
            f.s.resize(len);
            for (auto &c : f.s) c = base_gc(r, spec->gc);
            f.copies = (uint32_t)r.range(spec->copies_min, spec->copies_max);
            f.div = spec->div_min + (spec->div_max - spec->div_min) * r.uniform();
            want += (double)len * f.copies;
        }
        double scale = want > 0 ? (double)budget / want : 0;
        for (auto &f : fams) {
            uint32_t c = (uint32_t)std::llround(f.copies * scale);
            f.copies = std::max(spec->copies_min, std::min(f.copies, c));
        }
        uint64_t placed = 0;
        for (uint32_t round = 0; placed < budget; round++) {
            bool any = false;
            for (auto &f : fams) {
                if (round >= f.copies || placed >= budget) continue;
                any = true;
                uint32_t ci = (uint32_t)r.below(nc);
                uint64_t L = f.s.size();
                if (lens[ci] <= L) continue;
                uint64_t pos = start[ci] + r.below(lens[ci] - L);
                for (uint64_t j = 0; j < L; j++) {
                    char c = f.s[j];
                    seq[pos + j] = r.uniform() < f.div ? other_base(r, c) : c;
                }
                placed += L;
            }
            if (!any) break;
        }
    }

    // tandem repeats: regions of 100..2000 bp made of a 2..50 bp unit with 1% noise
    uint64_t tbudget = (uint64_t)(spec->tandem_fraction * (double)total);
    uint64_t tplaced = 0;
    while (tplaced < tbudget) {
        uint32_t ci = (uint32_t)r.below(nc);
        uint64_t L = r.range(100, 2000);
        if (lens[ci] <= L) break;
        uint32_t ulen = (uint32_t)r.range(spec->tandem_unit_min ? spec->tandem_unit_min : 2,
                                          spec->tandem_unit_max ? spec->tandem_unit_max : 50);
        char unit[64];
        for (uint32_t j = 0; j < ulen && j < 64; j++) unit[j] = base_gc(r, spec->gc);
        uint64_t pos = start[ci] + r.below(lens[ci] - L);
        for (uint64_t j = 0; j < L; j++) {
            char c = unit[j % ulen];
            seq[pos + j] = r.uniform() < 0.01 ? other_base(r, c) : c;
        }
        tplaced += L;
    }
    return GIS_OK;
}

namespace {

// Copy a mutated version of src (read direction) into dst until len bases are
// produced; src must hold enough bases.
void mutate_into(Rng &r, const char *src, uint64_t src_avail, char *dst, uint32_t len,
                 double sub, double indel) {
    uint32_t o = 0;
    uint64_t i = 0;
    while (o < len) {
        char c = i < src_avail ? src[i] : 'A';
        double x = r.uniform();
        if (x < indel * 0.5) {         // insertion
            dst[o++] = "ACGT"[r.below(4)];
            continue;
        } else if (x < indel) {        // deletion
            i++;
            continue;
        }
        if (r.uniform() < sub) c = other_base(r, c);
        dst[o++] = c;
        i++;
    }
}

}  // namespace

extern "C" int gis_synth_reads(const gis_read_spec *spec, const char *genome,
                               const uint64_t *contig_lens, uint32_t n_contigs, char *seqs,
                               uint64_t *offsets, uint32_t *lens, uint32_t *truth_contig,
                               uint32_t *truth_pos, uint8_t *truth_rc) {
    if (!spec || !genome || !contig_lens || n_contigs == 0 || spec->n_lengths == 0 ||
        spec->n_lengths > 8)
        return GIS_ERR_BAD_PARAMETER;
    Rng r(spec->seed);
    std::vector<uint64_t> start(n_contigs + 1, 0);
    for (uint32_t i = 0; i < n_contigs; i++) start[i + 1] = start[i] + contig_lens[i];
    const uint64_t total = start[n_contigs];
    double wsum = 0;
    for (uint32_t i = 0; i < spec->n_lengths; i++) wsum += spec->weights[i];
    auto draw_len = [&]() -> uint32_t {
        double x = r.uniform() * wsum;
        for (uint32_t i = 0; i < spec->n_lengths; i++) {
            if (x < spec->weights[i]) return spec->lengths[i];
            x -= spec->weights[i];
        }
        return spec->lengths[spec->n_lengths - 1];
    };
    auto pick_contig = [&](uint64_t need, uint32_t *ci) -> bool {
        for (int tries = 0; tries < 1000; tries++) {
            uint64_t g = r.below(total);
            uint32_t c = (uint32_t)(std::upper_bound(start.begin(), start.end(), g) - start.begin() - 1);
            if (contig_lens[c] > need + 64) { *ci = c; return true; }
        }
        return false;
    };
    std::vector<char> tmp, fwd;
    uint64_t off = 0, out = 0;
    for (uint64_t n = 0; n < spec->n_reads; n++) {
        uint32_t L = draw_len();
        if (!spec->paired) {
            uint32_t ci;
            if (!pick_contig((uint64_t)L * 2, &ci)) return GIS_ERR_BAD_PARAMETER;
            uint64_t span = (uint64_t)L + L / 8 + 16;
            uint64_t pos = r.below(contig_lens[ci] - span);
            bool rc = r.uniform() < spec->rc_fraction;
            const char *src = genome + start[ci] + pos;
            tmp.resize(L);
            mutate_into(r, src, span, tmp.data(), L, spec->sub_rate, spec->indel_rate);
            char *dst = seqs + off;
            if (rc) for (uint32_t j = 0; j < L; j++) dst[j] = comp(tmp[L - 1 - j]);
            else memcpy(dst, tmp.data(), L);
            offsets[out] = off; lens[out] = L;
            if (truth_contig) truth_contig[out] = ci;
            if (truth_pos) truth_pos[out] = (uint32_t)pos;
            if (truth_rc) truth_rc[out] = rc;
            off += L; out++;
        } else {
            uint64_t frag = (uint64_t)std::max(
                (double)L + 1.0, std::round(r.normal(spec->insert_mean, spec->insert_sd)));
            uint32_t ci;
            if (!pick_contig(frag + L, &ci)) return GIS_ERR_BAD_PARAMETER;
            uint64_t span = frag + L / 8 + 16;
            uint64_t pos = r.below(contig_lens[ci] - span);
            const char *src = genome + start[ci] + pos;
            bool flip = r.uniform() < 0.5;  // fragment from the reverse strand
            fwd.resize(frag);
            // fragment in its own orientation
            if (!flip) memcpy(fwd.data(), src, frag);
            else for (uint64_t j = 0; j < frag; j++) fwd[j] = comp(src[frag - 1 - j]);
            // mate 1: first L bases of the fragment; mate 2: rc of the last L bases
            tmp.resize(L);
            mutate_into(r, fwd.data(), frag, tmp.data(), L, spec->sub_rate, spec->indel_rate);
            memcpy(seqs + off, tmp.data(), L);
            offsets[out] = off; lens[out] = L;
            if (truth_contig) truth_contig[out] = ci;
            if (truth_pos) truth_pos[out] = (uint32_t)(flip ? pos + frag - L : pos);
            if (truth_rc) truth_rc[out] = flip;
            off += L; out++;
            std::vector<char> rcf(frag);
            for (uint64_t j = 0; j < frag; j++) rcf[j] = comp(fwd[frag - 1 - j]);
            mutate_into(r, rcf.data(), frag, tmp.data(), L, spec->sub_rate, spec->indel_rate);
            memcpy(seqs + off, tmp.data(), L);
            offsets[out] = off; lens[out] = L;
            if (truth_contig) truth_contig[out] = ci;
            if (truth_pos) truth_pos[out] = (uint32_t)(flip ? pos : pos + frag - L);
            if (truth_rc) truth_rc[out] = !flip;
            off += L; out++;
        }
    }
    return GIS_OK;
}
