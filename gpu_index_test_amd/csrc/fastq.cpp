// FASTQ ingestion: a chunked producer/consumer in front of the path (the
// reference's producer_pe_fastq_task and chunk formatting, step2/main.cu:65-84,
// 4212-4240; chunk reading step2/FastxStream.h:305-364,
// step2/FastxStream.cpp:499-640; record formatting step2/Formater.cpp:264-290,
// 404-425; names step2/main.cu:4121-4138).
//
// One reader thread pulls raw bytes (plain or gzip, through zlib) in chunks
// of about chunk_bytes and cuts them after the last complete record: records
// are four lines, so the cut is found by counting newlines from the start of
// the file (AVX2 compare over 64-byte blocks), never by guessing at '@'. Two
// mates files are read and scanned side by side and cut to the same record
// count. Parser threads format chunks in parallel into concatenated reads and
// names; the consumer receives them in file order. At most `ahead` chunks are
// read but not yet consumed, and their buffers are recycled.
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
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gis.h"

namespace {

// growable byte buffer that never zero-fills (recycled between chunks, so
// steady state touches no fresh pages)
struct Bytes {
    std::unique_ptr<char[]> p;
    size_t cap = 0, len = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        const size_t c = std::max(n, cap + cap / 2);
        std::unique_ptr<char[]> q(new char[c]);
        if (len) std::memcpy(q.get(), p.get(), len);
        p.swap(q);
        cap = c;
    }
    void append(const char *src, size_t n) {
        reserve(len + n);
        std::memcpy(p.get() + len, src, n);
        len += n;
    }
};

struct Parsed {
    uint64_t index = 0, first_record = 0;
    uint32_t n_records = 0, n_reads = 0;
    Bytes seqs, names;
    std::vector<uint64_t> offsets, name_offsets;
    std::vector<uint32_t> lens;
    int status = GIS_OK;
    std::string err;
    void reset() {
        seqs.len = names.len = 0;
        offsets.clear();
        name_offsets.clear();
        lens.clear();
        status = GIS_OK;
        err.clear();
    }
};

struct Raw {
    uint64_t index = 0, first_record = 0;
    uint32_t records = 0;
    Bytes part[2];
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

// complete records of p[0, n) (every fourth newline ends one): their count,
// with the offset just past each record's last newline in ends
uint64_t scan_records(const char *p, size_t n, std::vector<size_t> &ends) {
    ends.clear();
    uint64_t lines = 0;
    size_t i = 0;
    for (; i + 64 <= n; i += 64) {
        uint64_t m = nl_mask(p + i);
        while (m) {
            if ((++lines & 3) == 0) ends.push_back(i + (size_t)__builtin_ctzll(m) + 1);
            m &= m - 1;
        }
    }
    for (; i < n; i++)
        if (p[i] == '\n' && (++lines & 3) == 0) ends.push_back(i + 1);
    return ends.size();
}

// The rest of a file (everything after the last cut) without the empty
// lines ("\n" or "\r\n") that follow its last record. A line is added back
// when the records would otherwise miss their last line: an empty quality
// line of an empty read.
size_t end_of_records(const char *s, size_t len) {
    size_t n = len;
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
    if (stripped && count_newlines(s, n) % 4 == 3) n += s[n] == '\r' ? 2 : 1;
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
    std::deque<std::unique_ptr<Raw>> raw;              // cut, waiting for a parser
    std::vector<std::unique_ptr<Raw>> raw_free;        // recycled
    std::map<uint64_t, std::unique_ptr<Parsed>> done;  // parsed, waiting for the consumer
    std::vector<std::unique_ptr<Parsed>> parsed_free;  // recycled
    uint64_t n_cut = 0;      // chunks cut by the reader
    uint64_t next_out = 0;   // next chunk the consumer takes
    bool reader_done = false, stop = false;
    // an error of the reader, reported after the chunks cut before it
    int read_status = GIS_OK;
    std::string read_err;
    std::unique_ptr<Parsed> current;  // the chunk handed out last
    std::string err;                  // message of the last failed call
};

namespace {

// append up to `want` bytes of file k to b; false at end of file
bool read_more(gis_fastq *q, int k, Bytes &b, size_t want, int &status, std::string &err) {
    b.reserve(b.len + want + 1);
    size_t got = 0;
    while (got < want) {
        const unsigned ask = (unsigned)std::min<size_t>(want - got, 1u << 30);
        const int r = gzread(q->f[k], b.p.get() + b.len + got, ask);
        if (r < 0) {
            int zerr = 0;
            const char *m = gzerror(q->f[k], &zerr);
            status = GIS_ERR_IO;
            err = q->path[k] + ": read error: " + (m ? m : "");
            b.len += got;
            return false;
        }
        if (r == 0) break;
        got += (size_t)r;
    }
    b.len += got;
    return got == want;
}

void reader_main(gis_fastq *q) {
    Bytes carry[2];
    std::vector<size_t> ends[2];
    bool eof[2] = {false, false};
    uint64_t first_record = 0;
    int status = GIS_OK;
    std::string err;
    for (;;) {
        std::unique_ptr<Raw> rc;
        {
            std::unique_lock<std::mutex> lk(q->mu);
            q->cv.wait(lk, [&] { return q->stop || q->n_cut - q->next_out < q->ahead; });
            if (q->stop) break;
            if (!q->raw_free.empty()) {
                rc = std::move(q->raw_free.back());
                q->raw_free.pop_back();
            }
        }
        if (!rc) rc.reset(new Raw());
        uint64_t recs[2] = {0, 0};
        for (int k = 0; k < q->n_files; k++) {
            rc->part[k].len = 0;
            rc->part[k].append(carry[k].p.get(), carry[k].len);
            carry[k].len = 0;
        }
        // read until both parts hold a complete record (a record longer than
        // a chunk makes the read longer); two mates files are read and
        // scanned by two threads, as the reference reads its left and right
        // chunks
        size_t want = q->chunk_bytes;
        int st[2] = {GIS_OK, GIS_OK};
        std::string er[2];
        auto fill = [&](int k) {
            Bytes &b = rc->part[k];
            if (!eof[k] && b.len < want) {
                if (!read_more(q, k, b, want - b.len, st[k], er[k])) eof[k] = true;
                if (st[k]) return;
            }
            if (eof[k]) {
                // a last line without '\n' ends the file
                if (b.len && b.p[b.len - 1] != '\n') b.append("\n", 1);
                b.len = end_of_records(b.p.get(), b.len);
            }
            recs[k] = scan_records(b.p.get(), b.len, ends[k]);
        };
        for (;;) {
            if (q->n_files == 2) {
                std::thread t(fill, 1);
                fill(0);
                t.join();
            } else {
                fill(0);
            }
            if (st[0] || st[1]) {
                status = st[0] ? st[0] : st[1];
                err = st[0] ? er[0] : er[1];
                break;
            }
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
            const bool left0 = rc->part[0].len != 0, left1 = q->n_files == 2 && rc->part[1].len != 0;
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
            Bytes &b = rc->part[k];
            const size_t cut = ends[k][r - 1];
            carry[k].append(b.p.get() + cut, b.len - cut);
            b.len = cut;
        }
        rc->records = (uint32_t)r;
        rc->first_record = first_record;
        first_record += r;
        {
            std::lock_guard<std::mutex> lk(q->mu);
            rc->index = q->n_cut++;
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

void malformed(Parsed &out, const std::string &path, uint64_t record, const char *what) {
    out.status = GIS_ERR_INVALID_INPUT;
    out.err = path + ": record " + std::to_string(record) + what;
}

// one file's records of a chunk appended to out; false on malformed input
bool format_part(const Bytes &b, uint32_t records, uint64_t first_record, const std::string &path,
                 Parsed &out) {
    const char *p = b.p.get();
    const char *end = p + b.len;
    for (uint32_t r = 0; r < records; r++) {
        const char *line[4];
        size_t len[4];
        for (int k = 0; k < 4; k++) {
            const char *nl = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
            line[k] = p;
            len[k] = (size_t)(nl - p);
            p = nl + 1;
        }
        if (len[0] == 0 || line[0][0] != '@') {
            malformed(out, path, first_record + r + 1, " does not start with '@'");
            return false;
        }
        if (len[2] == 0 || line[2][0] != '+') {
            malformed(out, path, first_record + r + 1, ": third line does not start with '+'");
            return false;
        }
        if (len[1] > GIS_MAX_READ_LEN) {
            malformed(out, path, first_record + r + 1, ": read longer than 65535 bases");
            return false;
        }
        out.offsets.push_back(out.seqs.len);
        out.lens.push_back((uint32_t)len[1]);
        out.seqs.append(line[1], len[1]);
        // ConvertNeo2KSeq: drop a leading '@', stop at the first space
        const char *nm = line[0] + 1;
        size_t nl = len[0] - 1;
        const void *sp = std::memchr(line[0], ' ', len[0]);
        if (sp) nl = (size_t)(static_cast<const char *>(sp) - nm);
        out.names.append(nm, nl);
        out.name_offsets.push_back(out.names.len);
    }
    return true;
}

void parser_main(gis_fastq *q) {
    for (;;) {
        std::unique_ptr<Raw> rc;
        std::unique_ptr<Parsed> out;
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
            if (!q->parsed_free.empty()) {
                out = std::move(q->parsed_free.back());
                q->parsed_free.pop_back();
            }
        }
        if (!out) out.reset(new Parsed());
        out->reset();
        out->index = rc->index;
        out->first_record = rc->first_record;
        out->n_records = rc->records;
        out->n_reads = rc->records * (uint32_t)q->n_files;
        out->seqs.reserve(rc->part[0].len + rc->part[1].len);
        out->offsets.reserve(out->n_reads);
        out->lens.reserve(out->n_reads);
        out->name_offsets.reserve(out->n_reads + 1);
        out->name_offsets.push_back(0);
        for (int k = 0; k < q->n_files; k++)
            if (!format_part(rc->part[k], rc->records, rc->first_record, q->path[k], *out)) break;
        {
            std::lock_guard<std::mutex> lk(q->mu);
            q->done.emplace(out->index, std::move(out));
            q->raw_free.push_back(std::move(rc));
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
    std::unique_ptr<Parsed> next;
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
        if (q->current) q->parsed_free.push_back(std::move(q->current));
    }
    q->cv.notify_all();
    if (next->status) {
        q->err = next->err;
        std::lock_guard<std::mutex> lk(q->mu);
        q->parsed_free.push_back(std::move(next));
        return GIS_ERR_INVALID_INPUT;
    }
    q->current = std::move(next);
    Parsed &c = *q->current;
    out->n_records = c.n_records;
    out->n_reads = c.n_reads;
    out->seqs = c.seqs.p.get();
    out->offsets = c.offsets.data();
    out->lens = c.lens.data();
    out->names = c.names.p.get();
    out->name_offsets = c.name_offsets.data();
    out->index = c.index;
    out->first_record = c.first_record;
    out->bases = c.seqs.len;
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
