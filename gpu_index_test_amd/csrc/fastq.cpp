// FASTQ ingestion: a chunked producer/consumer in front of the path (the
// reference's producer_pe_fastq_task and chunk formatting, step2/main.cu:65-84,
// 4212-4240; chunk reading step2/FastxStream.h:305-364,
// step2/FastxStream.cpp:499-640; record formatting step2/Formater.cpp:264-290,
// 404-425; names step2/main.cu:4121-4138).
//
// One reader thread pulls raw bytes (plain or gzip, through zlib) in chunks
// of about chunk_bytes and cuts them after the last complete record: records
// are four lines, so the cut is found by counting newlines from the start of
// the file (AVX2 compare + popcount over 64-byte blocks), never by guessing
// at '@'. With two files the chunks of both mates are cut to the same record
// count. Parser threads format chunks in parallel into concatenated reads and
// names; the consumer receives them in file order. At most `ahead` chunks are
// read but not yet consumed.
//
// Records follow the reference's formatting: lines end at '\n' only (a '\r'
// of CRLF files stays in the line, as in neoGetLine), the read is the second
// line, the name is the first line without its leading '@' up to the first
// space. Unlike the reference, malformed input is reported instead of read
// past: a record must start with '@', its third line with '+', a file must
// end on a complete record (trailing empty lines are ignored), mates files
// must hold the same number of records, and reads are at most 65,535 bases.
#include <immintrin.h>
#include <zlib.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gis.h"

namespace {

struct Parsed {
    uint64_t index = 0, first_record = 0;
    uint32_t n_records = 0, n_reads = 0;
    std::vector<char> seqs, names;
    std::vector<uint64_t> offsets, name_offsets;
    std::vector<uint32_t> lens;
    int status = GIS_OK;
    std::string err;
};

struct Raw {
    uint64_t index = 0, first_record = 0;
    uint32_t records = 0;
    std::string part[2];
};

// newline mask of 64 bytes
inline uint64_t nl_mask(const char *p) {
    const __m256i nl = _mm256_set1_epi8('\n');
    const __m256i a = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(p));
    const __m256i b = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(p + 32));
    const uint32_t lo = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(a, nl));
    const uint32_t hi = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(b, nl));
    return (uint64_t)hi << 32 | lo;
}

uint64_t count_newlines(const char *p, size_t n) {
    uint64_t c = 0;
    size_t i = 0;
    for (; i + 64 <= n; i += 64) c += (uint64_t)__builtin_popcountll(nl_mask(p + i));
    for (; i < n; i++) c += p[i] == '\n';
    return c;
}

// offset just past the k-th newline (k >= 1); n if there are fewer
size_t after_newline(const char *p, size_t n, uint64_t k) {
    size_t i = 0;
    for (; i + 64 <= n; i += 64) {
        uint64_t m = nl_mask(p + i);
        const uint64_t c = (uint64_t)__builtin_popcountll(m);
        if (c >= k) {
            while (--k) m &= m - 1;
            return i + (size_t)__builtin_ctzll(m) + 1;
        }
        k -= c;
    }
    for (; i < n; i++)
        if (p[i] == '\n' && --k == 0) return i + 1;
    return n;
}

// The rest of a file (everything after the last cut) without the empty
// lines ("\n" or "\r\n") that follow its last record. A line is added
// back when the records would otherwise miss their last line: an empty
// quality line of an empty read.
size_t end_of_records(const std::string &s) {
    size_t n = s.size();
    int stripped = 0;
    for (;;) {
        if (n >= 1 && s[n - 1] == '\n' && (n == 1 || s[n - 2] == '\n')) {
            n -= 1;
        } else if (n >= 2 && s[n - 1] == '\n' && s[n - 2] == '\r' && (n == 2 || s[n - 3] == '\n')) {
            n -= 2;
        } else {
            break;
        }
        stripped++;
    }
    if (stripped && count_newlines(s.data(), n) % 4 == 3) n += s[n] == '\r' ? 2 : 1;
    return n;
}

}  // namespace

struct gis_fastq {
    gzFile f[2] = {nullptr, nullptr};
    std::string path[2];
    int n_files = 1;
    size_t chunk_bytes = 4u << 20;
    size_t ahead = 8;
    std::thread reader;
    std::vector<std::thread> parsers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Raw> raw;
    std::map<uint64_t, Parsed> done;
    uint64_t n_cut = 0;      // chunks cut by the reader
    uint64_t next_out = 0;   // next chunk the consumer takes
    bool reader_done = false, stop = false;
    // an error of the reader, reported after the chunks cut before it
    int read_status = GIS_OK;
    std::string read_err;
    Parsed current;
    std::string err;  // message of the last failed call
};

namespace {

// append up to `want` bytes of file k to s; false at end of file
bool read_more(gis_fastq *q, int k, std::string &s, size_t want, int &status, std::string &err) {
    const size_t old = s.size();
    s.resize(old + want);
    size_t got = 0;
    while (got < want) {
        const unsigned ask = (unsigned)std::min<size_t>(want - got, 1u << 30);
        const int r = gzread(q->f[k], &s[old + got], ask);
        if (r < 0) {
            int zerr = 0;
            const char *m = gzerror(q->f[k], &zerr);
            status = GIS_ERR_IO;
            err = q->path[k] + ": read error: " + (m ? m : "");
            s.resize(old + got);
            return false;
        }
        if (r == 0) break;
        got += (size_t)r;
    }
    s.resize(old + got);
    return got == want;
}

void reader_main(gis_fastq *q) {
    std::string carry[2];
    bool eof[2] = {false, false};
    uint64_t first_record = 0;
    int status = GIS_OK;
    std::string err;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(q->mu);
            q->cv.wait(lk, [&] { return q->stop || q->n_cut - q->next_out < q->ahead; });
            if (q->stop) break;
        }
        Raw rc;
        uint64_t recs[2] = {0, 0};
        // read until both parts hold a complete record (a record longer than
        // a chunk makes the read longer)
        size_t want = q->chunk_bytes;
        for (;;) {
            for (int k = 0; k < q->n_files; k++) {
                std::string &s = rc.part[k];
                if (s.empty()) s.swap(carry[k]);
                if (!eof[k] && s.size() < want) {
                    if (!read_more(q, k, s, want - s.size(), status, err)) eof[k] = true;
                    if (status) break;
                }
                if (eof[k]) {
                    // a last line without '\n' ends the file
                    if (!s.empty() && s.back() != '\n') s.push_back('\n');
                    s.resize(end_of_records(s));
                }
                recs[k] = count_newlines(s.data(), s.size()) / 4;
            }
            if (status) break;
            const uint64_t r = q->n_files == 2 ? std::min(recs[0], recs[1]) : recs[0];
            const bool all_eof = eof[0] && (q->n_files == 1 || eof[1]);
            if (r > 0 || all_eof) break;
            want *= 2;
        }
        if (status) break;
        const uint64_t r = q->n_files == 2 ? std::min(recs[0], recs[1]) : recs[0];
        if (r == 0) {
            // end of input: whatever is left is an incomplete record or a
            // mates file with more records than the other
            const bool left0 = !rc.part[0].empty(), left1 = q->n_files == 2 && !rc.part[1].empty();
            if (left0 || left1) {
                status = GIS_ERR_INVALID_INPUT;
                if (q->n_files == 2 && (recs[0] || recs[1]))
                    err = "mates files hold different numbers of records (" + q->path[recs[0] ? 0 : 1] +
                          " has more after record " + std::to_string(first_record) + ")";
                else
                    err = q->path[left0 ? 0 : 1] + ": incomplete record after record " +
                          std::to_string(first_record);
            }
            break;
        }
        for (int k = 0; k < q->n_files; k++) {
            const size_t cut = after_newline(rc.part[k].data(), rc.part[k].size(), 4 * r);
            carry[k].assign(rc.part[k], cut, std::string::npos);
            rc.part[k].resize(cut);
        }
        rc.records = (uint32_t)r;
        rc.first_record = first_record;
        first_record += r;
        {
            std::lock_guard<std::mutex> lk(q->mu);
            rc.index = q->n_cut++;
            q->raw.push_back(std::move(rc));
        }
        q->cv.notify_all();
    }
    std::lock_guard<std::mutex> lk(q->mu);
    q->read_status = status;
    q->read_err = err;
    q->reader_done = true;
    q->cv.notify_all();
}

// one file's records of a chunk appended to out; false on malformed input
bool format_part(const std::string &s, uint32_t records, uint64_t first_record, const std::string &path,
                 Parsed &out) {
    const char *p = s.data();
    const char *end = p + s.size();
    for (uint32_t r = 0; r < records; r++) {
        const char *line[4];
        size_t len[4];
        for (int k = 0; k < 4; k++) {
            const char *nl = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
            line[k] = p;
            len[k] = (size_t)(nl - p);
            p = nl + 1;
        }
        const std::string where = path + ": record " + std::to_string(first_record + r + 1);
        if (len[0] == 0 || line[0][0] != '@') {
            out.status = GIS_ERR_INVALID_INPUT;
            out.err = where + " does not start with '@'";
            return false;
        }
        if (len[2] == 0 || line[2][0] != '+') {
            out.status = GIS_ERR_INVALID_INPUT;
            out.err = where + ": third line does not start with '+'";
            return false;
        }
        if (len[1] > GIS_MAX_READ_LEN) {
            out.status = GIS_ERR_INVALID_INPUT;
            out.err = where + ": read longer than " + std::to_string(GIS_MAX_READ_LEN) + " bases";
            return false;
        }
        out.offsets.push_back(out.seqs.size());
        out.lens.push_back((uint32_t)len[1]);
        out.seqs.insert(out.seqs.end(), line[1], line[1] + len[1]);
        // ConvertNeo2KSeq: drop a leading '@', stop at the first space
        const char *nm = line[0] + 1;
        size_t nl = len[0] - 1;
        const void *sp = std::memchr(line[0], ' ', len[0]);
        if (sp) nl = (size_t)(static_cast<const char *>(sp) - nm);
        out.names.insert(out.names.end(), nm, nm + nl);
        out.name_offsets.push_back(out.names.size());
    }
    return true;
}

void parser_main(gis_fastq *q) {
    for (;;) {
        Raw rc;
        {
            std::unique_lock<std::mutex> lk(q->mu);
            q->cv.wait(lk, [&] { return q->stop || !q->raw.empty() || q->reader_done; });
            if (q->stop) return;
            if (q->raw.empty()) {
                if (q->reader_done) return;
                continue;
            }
            rc = std::move(q->raw.front());
            q->raw.pop_front();
        }
        Parsed out;
        out.index = rc.index;
        out.first_record = rc.first_record;
        out.n_records = rc.records;
        out.n_reads = rc.records * (uint32_t)q->n_files;
        size_t bytes = rc.part[0].size() + rc.part[1].size();
        out.seqs.reserve(bytes / 2);
        out.offsets.reserve(out.n_reads);
        out.lens.reserve(out.n_reads);
        out.name_offsets.reserve(out.n_reads + 1);
        out.name_offsets.push_back(0);
        for (int k = 0; k < q->n_files; k++)
            if (!format_part(rc.part[k], rc.records, rc.first_record, q->path[k], out)) break;
        {
            std::lock_guard<std::mutex> lk(q->mu);
            q->done.emplace(out.index, std::move(out));
        }
        q->cv.notify_all();
    }
}

}  // namespace

extern "C" int gis_fastq_open(const char *path1, const char *path2, uint64_t chunk_bytes, uint32_t threads,
                              gis_fastq **out) {
    if (!path1 || !out) return GIS_ERR_BAD_PARAMETER;
    *out = nullptr;
    gis_fastq *q = nullptr;
    try {
        q = new gis_fastq();
    } catch (const std::bad_alloc &) {
        return GIS_ERR_NOMEM;
    }
    q->n_files = path2 ? 2 : 1;
    q->path[0] = path1;
    if (path2) q->path[1] = path2;
    if (chunk_bytes) q->chunk_bytes = (size_t)std::max<uint64_t>(chunk_bytes, 4096);
    for (int k = 0; k < q->n_files; k++) {
        q->f[k] = gzopen(q->path[k].c_str(), "rb");
        if (!q->f[k]) {
            for (int j = 0; j < k; j++) gzclose(q->f[j]);
            delete q;
            return GIS_ERR_IO;
        }
        gzbuffer(q->f[k], 1u << 20);
    }
    unsigned hw = std::thread::hardware_concurrency();
    unsigned n_par = threads ? threads : std::max(1u, std::min(8u, hw ? hw : 1u));
    q->ahead = std::max<size_t>(4, 2 * n_par);
    try {
        q->reader = std::thread(reader_main, q);
        for (unsigned t = 0; t < n_par; t++) q->parsers.emplace_back(parser_main, q);
    } catch (...) {
        gis_fastq_close(q);
        return GIS_ERR_NOMEM;
    }
    *out = q;
    return GIS_OK;
}

extern "C" int gis_fastq_next(gis_fastq *q, gis_fastq_chunk *out) {
    if (!q || !out) return GIS_ERR_BAD_PARAMETER;
    std::memset(out, 0, sizeof *out);
    Parsed next;
    {
        std::unique_lock<std::mutex> lk(q->mu);
        q->cv.wait(lk, [&] {
            return q->done.count(q->next_out) || (q->reader_done && q->next_out >= q->n_cut);
        });
        auto it = q->done.find(q->next_out);
        if (it == q->done.end()) {
            // every chunk is out: the end, or the reader's error
            out->index = q->next_out;
            if (q->read_status) {
                q->err = q->read_err;
                return q->read_status;
            }
            return GIS_OK;
        }
        next = std::move(it->second);
        q->done.erase(it);
        q->next_out++;
    }
    q->cv.notify_all();
    if (next.status) {
        q->err = next.err;
        return next.status;
    }
    q->current = std::move(next);
    Parsed &c = q->current;
    out->n_records = c.n_records;
    out->n_reads = c.n_reads;
    out->seqs = c.seqs.data();
    out->offsets = c.offsets.data();
    out->lens = c.lens.data();
    out->names = c.names.data();
    out->name_offsets = c.name_offsets.data();
    out->index = c.index;
    out->first_record = c.first_record;
    out->bases = c.seqs.size();
    return GIS_OK;
}

extern "C" const char *gis_fastq_last_error(const gis_fastq *q) { return q ? q->err.c_str() : "null reader"; }

extern "C" void gis_fastq_close(gis_fastq *q) {
    if (!q) return;
    {
        std::lock_guard<std::mutex> lk(q->mu);
        q->stop = true;
    }
    q->cv.notify_all();
    if (q->reader.joinable()) q->reader.join();
    for (auto &t : q->parsers)
        if (t.joinable()) t.join();
    for (int k = 0; k < q->n_files; k++)
        if (q->f[k]) gzclose(q->f[k]);
    delete q;
}

extern "C" int gis_server_submit_fastq(gis_server *s, gis_fastq *q, uint64_t *id, gis_fastq_chunk *chunk) {
    if (!s || !q || !id || !chunk) return GIS_ERR_BAD_PARAMETER;
    *id = 0;
    const int rc = gis_fastq_next(q, chunk);
    if (rc || chunk->n_records == 0) return rc;
    return gis_server_submit(s, chunk->seqs, chunk->offsets, chunk->lens, chunk->n_reads, id);
}
