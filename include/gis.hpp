// C++ façade over the gis C ABI with the reference's names and semantics:
// IndexParameters (step2/indexparameters.hpp:72-95), References
// (step2/refs.hpp:53-91), StrobemerIndex (step2/index.hpp:37-183),
// randstrobes_query (step2/randstrobes.hpp:107), find_nams and
// find_nams_rescue (step2/nam.hpp:11-20), the pair step after them and the
// FASTQ reader in front of them. Header-only; link against
// libgis.so. Errors surface as the reference's exception types
// (step2/exceptions.hpp:7-25), all derived from gis::Error.
//
// Differences a caller sees: the index is resident on a GPU, so it is built
// or loaded through a device context; NAM lists come back in canonical order
// (is_rc, ref_id, query_start, ref_start) with nam_id = rank, which is the
// reference's NAM set with its implementation-defined order fixed.
#ifndef GIS_HPP
#define GIS_HPP

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "gis.h"

namespace gis {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};
class BadParameter : public Error {
  public:
    explicit BadParameter(const std::string &m) : Error(GIS_ERR_BAD_PARAMETER, m) {}
};
class InvalidIndexFile : public Error {
  public:
    explicit InvalidIndexFile(const std::string &m) : Error(GIS_ERR_INVALID_INDEX, m) {}
};
class InvalidFasta : public Error {
  public:
    explicit InvalidFasta(const std::string &m) : Error(GIS_ERR_INVALID_FASTA, m) {}
};

inline void check(int code, const gis_ctx *ctx = nullptr) {
    if (code == GIS_OK) return;
    const std::string msg = ctx ? gis_last_error(ctx) : std::string("gis call failed");
    switch (code) {
    case GIS_ERR_BAD_PARAMETER: throw BadParameter(msg);
    case GIS_ERR_INVALID_INDEX: throw InvalidIndexFile(msg);
    case GIS_ERR_INVALID_FASTA: throw InvalidFasta(msg);
    default: throw Error(code, msg);
    }
}

using Nam = gis_nam;
using QueryRandstrobe = gis_query_randstrobe;
using RefRandstrobe = gis_ref_randstrobe;
using QueryRandstrobeVector = std::vector<QueryRandstrobe>;

// IndexParameters: syncmer (k, s, t_syncmer) and randstrobe (l, u, q,
// max_dist, w_min, w_max) parameters.
struct IndexParameters {
    static constexpr int DEFAULT = GIS_DEFAULT;
    gis_params p{};

    IndexParameters() = default;
    IndexParameters(int canonical_read_length, int k, int s, int l, int u, int q, int max_dist) {
        check(gis_params_make(canonical_read_length, k, s, l, u, q, max_dist, &p));
    }
    static IndexParameters from_read_length(int read_length, int k = DEFAULT, int s = DEFAULT,
                                            int l = DEFAULT, int u = DEFAULT, int c = DEFAULT,
                                            int max_seed_len = DEFAULT) {
        IndexParameters r;
        check(gis_params_from_read_length(read_length, k, s, l, u, c, max_seed_len, &r.p));
        return r;
    }
    int canonical_read_length() const { return p.canonical_read_length; }
    int k() const { return p.k; }
    int s() const { return p.s; }
    int t_syncmer() const { return p.t_syncmer; }
    uint64_t q() const { return p.q; }
    int max_dist() const { return p.max_dist; }
    unsigned w_min() const { return p.w_min; }
    unsigned w_max() const { return p.w_max; }
    bool operator==(const IndexParameters &o) const {
        return p.canonical_read_length == o.p.canonical_read_length && p.k == o.p.k &&
               p.s == o.p.s && p.t_syncmer == o.p.t_syncmer && p.l == o.p.l && p.u == o.p.u &&
               p.q == o.p.q && p.max_dist == o.p.max_dist && p.w_min == o.p.w_min &&
               p.w_max == o.p.w_max;
    }
    bool operator!=(const IndexParameters &o) const { return !(*this == o); }
};

// References: names, sequences (upper-cased), lengths.
struct References {
    std::vector<std::string> sequences;
    std::vector<std::string> names;
    std::vector<size_t> lengths;

    References() = default;
    References(std::vector<std::string> seqs, std::vector<std::string> nms)
        : sequences(std::move(seqs)), names(std::move(nms)) {
        if (names.size() != sequences.size()) throw BadParameter("lengths do not match");
        for (auto &s : sequences) lengths.push_back(s.size());
    }
    size_t total_length() const {
        size_t t = 0;
        for (size_t l : lengths) t += l;
        return t;
    }
    size_t size() const { return sequences.size(); }

    // FASTA reader: names cut at the first space, bases upper-cased with
    // c & ~32 (step2/refs.cpp:32-79).
    static References from_fasta(const std::string &path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw Error(GIS_ERR_IO, "cannot open " + path);
        References r;
        std::string line, cur, name;
        bool have = false;
        if (in.peek() != '>') throw InvalidFasta("FASTA file must begin with '>' character");
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (!line.empty() && line[0] == '>') {
                if (have && !cur.empty()) {
                    r.sequences.push_back(cur);
                    r.names.push_back(name);
                }
                name = line.substr(1, line.find(' ') == std::string::npos ? std::string::npos
                                                                           : line.find(' ') - 1);
                cur.clear();
                have = true;
            } else {
                for (char c : line) cur.push_back((char)(c & ~32));
            }
        }
        if (have && !cur.empty()) {
            r.sequences.push_back(cur);
            r.names.push_back(name);
        }
        for (auto &s : r.sequences) r.lengths.push_back(s.size());
        return r;
    }
};

// A device context: one HIP stream and one resident index.
class Context {
  public:
    explicit Context(int device = 0) { check(gis_ctx_create(device, &h_)); }
    ~Context() {
        if (h_) gis_ctx_destroy(h_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    Context(Context &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    gis_ctx *get() const { return h_; }

  private:
    gis_ctx *h_ = nullptr;
};

// StrobemerIndex, resident in HBM of the context's device. As in the
// reference it keeps a reference to its References (which must outlive it);
// the parameters are copied (read() may replace them with the file's).
class StrobemerIndex {
  public:
    StrobemerIndex(const References &references, const IndexParameters &parameters,
                   int bits = -1, int device = 0)
        : references_(references), parameters(parameters), bits_arg_(bits), ctx_(device) {}

    void populate(float f, size_t n_threads) {
        std::vector<const char *> seqs;
        std::vector<uint64_t> lens;
        for (auto &s : references_.sequences) {
            seqs.push_back(s.data());
            lens.push_back(s.size());
        }
        check(gis_index_build(ctx_.get(), seqs.data(), lens.data(), (uint32_t)seqs.size(),
                              &parameters.p, bits_arg_, f, (int)n_threads),
              ctx_.get());
        refresh();
    }
    void read(const std::string &filename) {
        check(gis_index_read_sti(ctx_.get(), filename.c_str(), &parameters.p), ctx_.get());
        refresh();
    }
    void write(const std::string &filename) const {
        check(gis_index_write_sti(ctx_.get(), filename.c_str()), ctx_.get());
    }

    // find over many keys at once (one GPU launch); GIS_NOT_FOUND when absent
    std::vector<uint64_t> find(const std::vector<uint64_t> &keys) const {
        std::vector<uint64_t> out(keys.size());
        check(gis_lookup_batch(ctx_.get(), keys.data(), keys.size(), out.data()), ctx_.get());
        return out;
    }
    size_t find(uint64_t key) const { return (size_t)find(std::vector<uint64_t>{key})[0]; }
    size_t end() const { return (size_t)GIS_NOT_FOUND; }

    // accessors over a host copy of the table (fetched on first use)
    uint64_t get_hash(size_t position) const {
        host();
        return position < randstrobes_.size() ? randstrobes_[position].hash : UINT64_MAX;
    }
    bool is_filtered(size_t position) const {
        return get_hash(position) == get_hash(position + info_.filter_cutoff);
    }
    unsigned get_count(size_t position) const {
        host();
        const uint64_t h = randstrobes_[position].hash;
        const uint64_t e = starts_[(h >> (64 - info_.bits)) + 1];
        unsigned c = 0;
        for (size_t x = position; x < e && randstrobes_[x].hash == h; x++) c++;
        return c;
    }
    unsigned get_strobe1_position(size_t position) const {
        host();
        return randstrobes_[position].position;
    }
    int strobe2_offset(size_t position) const {
        host();
        return (int)(randstrobes_[position].packed & 0xff);
    }
    unsigned reference_index(size_t position) const {
        host();
        return randstrobes_[position].packed >> 8;
    }
    int k() const { return parameters.p.k; }
    int bits() const { return info_.bits; }
    unsigned filter_cutoff() const { return info_.filter_cutoff; }
    size_t size() const { return (size_t)info_.n_randstrobes; }
    const std::vector<RefRandstrobe> &randstrobes() const {
        host();
        return randstrobes_;
    }
    const std::vector<uint64_t> &randstrobe_start_indices() const {
        host();
        return starts_;
    }
    gis_ctx *context() const { return ctx_.get(); }

    const References &references() const { return references_; }

  private:
    const References &references_;

  public:
    IndexParameters parameters;

  private:
    void refresh() {
        check(gis_index_info_get(ctx_.get(), &info_), ctx_.get());
        randstrobes_.clear();
        starts_.clear();
    }
    void host() const {
        if (!randstrobes_.empty() || info_.n_randstrobes == 0) return;
        randstrobes_.resize(info_.n_randstrobes);
        starts_.resize(((size_t)1 << info_.bits) + 1);
        check(gis_index_download(ctx_.get(), randstrobes_.data(), starts_.data()), ctx_.get());
    }
    int bits_arg_;
    Context ctx_;
    gis_index_info info_{};
    mutable std::vector<RefRandstrobe> randstrobes_;
    mutable std::vector<uint64_t> starts_;
};

// randstrobes_query (step2/randstrobes.hpp:107): forward randstrobes, then
// reverse-complement ones. The GPU needs a device, so the index's context
// (and its parameters) are used.
inline QueryRandstrobeVector randstrobes_query(std::string_view seq, const StrobemerIndex &index) {
    const uint64_t off = 0;
    const uint32_t len = (uint32_t)seq.size();
    const int64_t n_k = (int64_t)seq.size() - index.k() + 1;
    const uint64_t cap = n_k > 0 ? 2 * (uint64_t)n_k : 1;
    QueryRandstrobeVector out(cap);
    uint64_t offs[2];
    uint64_t n_out = 0;
    check(gis_randstrobes_batch(index.context(), seq.data(), &off, &len, 1, offs, out.data(), cap,
                                &n_out),
          index.context());
    out.resize(n_out);
    return out;
}

// randstrobes_query(seq, parameters) (step2/randstrobes.hpp:107) as the
// reference declares it: no index needed. Each calling thread seeds on a
// context of its own (device 0), so threads do not wait for each other.
inline QueryRandstrobeVector randstrobes_query(std::string_view seq, const IndexParameters &parameters) {
    thread_local Context ctx(0);
    const uint64_t off = 0;
    const uint32_t len = (uint32_t)seq.size();
    const int64_t n_k = (int64_t)seq.size() - parameters.k() + 1;
    const uint64_t cap = n_k > 0 ? 2 * (uint64_t)n_k : 1;
    QueryRandstrobeVector out(cap);
    uint64_t offs[2];
    uint64_t n_out = 0;
    check(gis_randstrobes_params(ctx.get(), &parameters.p, seq.data(), &off, &len, 1, offs, out.data(), cap, &n_out),
          ctx.get());
    out.resize(n_out);
    return out;
}

namespace detail {
inline std::vector<Nam> qrs_call(const QueryRandstrobeVector &q, const StrobemerIndex &index,
                                 int rescue, unsigned rescue_cutoff, unsigned hard_cutoff,
                                 float *frac) {
    const uint64_t qoff[2] = {0, q.size()};
    uint64_t noff[2];
    float fr = 0;
    uint8_t fl = 0;
    std::vector<Nam> nams(16 + q.size());
    for (;;) {
        gis_nam_out o{noff, nams.data(), nams.size(), &fr, &fl, 0};
        const int code = gis_find_nams_qrs(index.context(), q.data(), qoff, 1, rescue,
                                           rescue_cutoff, hard_cutoff, &o);
        if (code == GIS_ERR_CAPACITY && o.n_nams > nams.size()) {
            nams.resize(o.n_nams);
            continue;
        }
        check(code, index.context());
        nams.resize(o.n_nams);
        break;
    }
    if (frac) *frac = fr;
    return nams;
}
}  // namespace detail

// find_nams (step2/nam.hpp:11-14): (nonrepetitive fraction, NAMs)
inline std::pair<float, std::vector<Nam>> find_nams(const QueryRandstrobeVector &query_randstrobes,
                                                    const StrobemerIndex &index) {
    float frac = 0;
    auto nams = detail::qrs_call(query_randstrobes, index, 0, 0, 1000, &frac);
    return {frac, std::move(nams)};
}

// find_nams_rescue (step2/nam.hpp:16-20), with the count limit of
// step2/nam.cu:964 as the default hard cutoff
inline std::vector<Nam> find_nams_rescue(const QueryRandstrobeVector &query_randstrobes,
                                         const StrobemerIndex &index, unsigned rescue_cutoff,
                                         unsigned hard_cutoff = 1000) {
    return detail::qrs_call(query_randstrobes, index, 1, rescue_cutoff, hard_cutoff, nullptr);
}

// Per-read NAM lists of a whole batch, as the GPU driver consumes them.
struct BatchNams {
    std::vector<uint64_t> offsets;  // n_reads + 1
    std::vector<Nam> nams;
    std::vector<float> nonrepetitive_fraction;
    std::vector<uint8_t> flags;     // GIS_FLAG_*
};

// A device-resident batch: upload reads once, run the path (seeding through
// NAMs, with the caller's rescue decision), download NAMs. With own_stream the
// batch runs on a stream of its own, and launch() / wait() overlap its run
// with runs of batches on other streams.
class Batch {
  public:
    Batch(const StrobemerIndex &index, uint32_t max_reads, uint64_t max_bases, bool own_stream = false)
        : index_(index) {
        check(own_stream ? gis_batch_create_async(index.context(), max_reads, max_bases, &b_)
                         : gis_batch_create(index.context(), max_reads, max_bases, &b_),
              index.context());
    }
    ~Batch() {
        if (b_) gis_batch_destroy(b_);
    }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    void upload(const std::vector<std::string_view> &reads) {
        seqs_.clear();
        offs_.clear();
        lens_.clear();
        for (auto r : reads) {
            offs_.push_back(seqs_.size());
            lens_.push_back((uint32_t)r.size());
            seqs_.append(r.data(), r.size());
        }
        check(gis_batch_upload(b_, seqs_.data(), offs_.data(), lens_.data(), (uint32_t)reads.size()),
              index_.context());
    }
    void run(const gis_map_options *opt = nullptr) {
        gis_map_options o;
        gis_map_options_default(&o);
        check(gis_batch_run(b_, opt ? opt : &o), index_.context());
    }
    void launch(const gis_map_options *opt = nullptr) {
        gis_map_options o;
        gis_map_options_default(&o);
        check(gis_batch_launch(b_, opt ? opt : &o), index_.context());
    }
    void wait() { check(gis_batch_wait(b_), index_.context()); }
    BatchNams download() {
        uint64_t total = 0;
        check(gis_batch_nam_count(b_, &total), index_.context());
        BatchNams r;
        const size_t n = lens_.size();
        r.offsets.resize(n + 1);
        r.nams.resize(total ? total : 1);
        r.nonrepetitive_fraction.resize(n ? n : 1);
        r.flags.resize(n ? n : 1);
        gis_nam_out o{r.offsets.data(), r.nams.data(), r.nams.size(),
                      r.nonrepetitive_fraction.data(), r.flags.data(), 0};
        check(gis_batch_download(b_, &o), index_.context());
        r.nams.resize(o.n_nams);
        r.nonrepetitive_fraction.resize(n);
        r.flags.resize(n);
        return r;
    }

    // The pair step on the last run's NAMs (reads p and n + p are the mates of
    // pair p): gpu_sort_nams, gpu_pre_cal_type with its per-type lists,
    // top_dropoff, get_mapq and get_best_scoring_nam_pairs
    // (step2/main.cu:870-910, 1719-1830, 3449-3472, 3552-3581, 4636-4646).
    struct Pairs {
        BatchNams lists;                      // NAM lists in the reference order
        std::vector<uint8_t> types;           // 0 .. 4 per pair
        std::vector<std::vector<uint32_t>> by_type;  // the five types[] lists
        std::vector<float> top_dropoff;       // per list
        std::vector<uint8_t> mapq;            // per list
        std::vector<uint64_t> joint_offsets;  // n_pairs + 1
        std::vector<gis_nam_pair> joint;      // NamPair{score, i1, i2}
    };
    Pairs pairs(const gis_pair_params *pp = nullptr) {
        gis_pair_params d;
        gis_pair_params_default(&d);
        d.flags = GIS_PAIRS_JOINT;
        check(gis_batch_pairs(b_, pp ? pp : &d), index_.context());
        const size_t n = lens_.size() / 2;
        uint64_t total = 0;
        check(gis_batch_nam_count(b_, &total), index_.context());
        Pairs r;
        r.lists.offsets.resize(2 * n + 1);
        r.lists.nams.resize(total ? total : 1);
        r.types.resize(n ? n : 1);
        std::vector<uint32_t> ids(n ? n : 1);
        r.top_dropoff.resize(n ? 2 * n : 1);
        r.mapq.resize(n ? 2 * n : 1);
        r.joint_offsets.resize(n + 1);
        gis_pair_out o{};
        o.nam_offsets = r.lists.offsets.data();
        o.nams = r.lists.nams.data();
        o.nam_capacity = r.lists.nams.size();
        o.types = r.types.data();
        o.type_ids = ids.data();
        o.top_dropoff = r.top_dropoff.data();
        o.mapq = r.mapq.data();
        o.joint_offsets = r.joint_offsets.data();
        // first call sizes the joint entries
        int rc = gis_batch_pairs_download(b_, &o);
        if (rc == GIS_ERR_CAPACITY || o.n_joint) {
            r.joint.resize(o.n_joint ? o.n_joint : 1);
            o.joint = r.joint.data();
            o.joint_capacity = r.joint.size();
            rc = gis_batch_pairs_download(b_, &o);
        }
        check(rc, index_.context());
        r.lists.nams.resize(o.n_nams);
        r.joint.resize(o.n_joint);
        r.types.resize(n);
        r.top_dropoff.resize(2 * n);
        r.mapq.resize(2 * n);
        size_t at = 0;
        for (int t = 0; t < 5; t++) {
            r.by_type.emplace_back(ids.begin() + at, ids.begin() + at + o.type_counts[t]);
            at += o.type_counts[t];
        }
        return r;
    }

  private:
    const StrobemerIndex &index_;
    gis_batch *b_ = nullptr;
    std::string seqs_;
    std::vector<uint64_t> offs_;
    std::vector<uint32_t> lens_;
};

// FASTQ records in chunks (one file, or two mates files: mate 1 of record i
// is read i and mate 2 read n_records + i of a chunk), read and formatted
// ahead by a reader thread and parser threads (producer_pe_fastq_task and
// chunkFormat, step2/main.cu:65-84, 4212-4240).
class FastqReader {
  public:
    explicit FastqReader(const std::string &path1, const std::string &path2 = "", uint64_t chunk_bytes = 0) {
        const int rc = gis_fastq_open(path1.c_str(), path2.empty() ? nullptr : path2.c_str(), chunk_bytes, 0, &f_);
        if (rc) throw Error(rc, "cannot open " + path1 + (path2.empty() ? "" : " / " + path2));
    }
    ~FastqReader() {
        if (f_) gis_fastq_close(f_);
    }
    FastqReader(const FastqReader &) = delete;
    FastqReader &operator=(const FastqReader &) = delete;
    // the next chunk (valid until the next call); false at the end of the input
    bool next(gis_fastq_chunk &chunk) {
        const int rc = gis_fastq_next(f_, &chunk);
        if (rc) throw Error(rc, gis_fastq_last_error(f_));
        return chunk.n_records != 0;
    }

  private:
    gis_fastq *f_ = nullptr;
};

}  // namespace gis

#endif
