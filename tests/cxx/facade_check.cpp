// Drives the C++ façade (include/gis.hpp) the way a reference caller would:
// StrobemerIndex::populate, then per read randstrobes_query -> find_nams ->
// the caller's rescue decision -> find_nams_rescue (step2/main.cu:3209-3216),
// and the same reads again through one device-resident Batch, then (for an
// even read count, read p and n + p as the mates of pair p) the pair step.
//
// The per-read path runs again from 8 threads at once (the reference's
// re-entrant find_nams on N workers), and randstrobes_query with bare
// IndexParameters must equal randstrobes_query with the index.
//
// usage: facade_check <refs.fa> <reads.txt> <out_prefix>
// writes <out_prefix>.single, <out_prefix>.threads and <out_prefix>.batch, each:
//   u64 n_reads, u64 offsets[n_reads + 1], gis_nam nams[offsets[n_reads]]
// and, for an even read count, <out_prefix>.pairs (the same layout, lists in
// the reference order) and <out_prefix>.types:
//   u64 n_pairs, u8 types[n_pairs], u64 n_joint, gis_nam_pair joint[n_joint]
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gis.hpp"

static void write_out(const std::string &path, const std::vector<uint64_t> &offs,
                      const std::vector<gis::Nam> &nams) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const uint64_t n = offs.size() - 1;
    std::fwrite(&n, 8, 1, f);
    std::fwrite(offs.data(), 8, offs.size(), f);
    if (!nams.empty()) std::fwrite(nams.data(), sizeof(gis::Nam), nams.size(), f);
    std::fclose(f);
}

int main(int argc, char **argv) {
    if (argc != 4) {
        std::cerr << "usage: facade_check <refs.fa> <reads.txt> <out_prefix>\n";
        return 2;
    }
    const char *stage = "references";
    try {
        const gis::References refs = gis::References::from_fasta(argv[1]);
        const auto params = gis::IndexParameters::from_read_length(150);
        stage = "populate";
        gis::StrobemerIndex index(refs, params);
        index.populate(0.0002f, 4);
        stage = "per-read find_nams";
        std::vector<std::string> reads;
        {
            std::ifstream in(argv[2]);
            std::string line;
            while (std::getline(in, line))
                if (!line.empty()) reads.push_back(line);
        }
        const unsigned rescue_cutoff = 2 * index.filter_cutoff();
        std::vector<uint64_t> offs{0};
        std::vector<gis::Nam> all;
        for (const auto &r : reads) {
            const auto q = gis::randstrobes_query(r, index);
            auto [frac, nams] = gis::find_nams(q, index);
            if ((double)frac < 0.7 || nams.empty()) nams = gis::find_nams_rescue(q, index, rescue_cutoff);
            all.insert(all.end(), nams.begin(), nams.end());
            offs.push_back(all.size());
        }
        write_out(std::string(argv[3]) + ".single", offs, all);

        stage = "randstrobes_query with IndexParameters";
        for (const auto &r : reads) {
            const auto a = gis::randstrobes_query(r, index);
            const auto b = gis::randstrobes_query(r, params);
            if (a.size() != b.size() ||
                (!a.empty() && std::memcmp(a.data(), b.data(), a.size() * sizeof(gis::QueryRandstrobe)) != 0)) {
                std::cerr << "randstrobes_query(seq, IndexParameters) differs from randstrobes_query(seq, index)\n";
                return 3;
            }
        }

        stage = "per-read find_nams on 8 threads";
        std::vector<std::vector<gis::Nam>> per(reads.size());
        std::vector<std::thread> workers;
        std::vector<std::string> errors(8);
        for (int t = 0; t < 8; t++) {
            workers.emplace_back([&, t] {
                try {
                    for (size_t i = t; i < reads.size(); i += 8) {
                        const auto q = gis::randstrobes_query(reads[i], params);
                        auto [frac, nams] = gis::find_nams(q, index);
                        if ((double)frac < 0.7 || nams.empty()) nams = gis::find_nams_rescue(q, index, rescue_cutoff);
                        per[i] = std::move(nams);
                    }
                } catch (const gis::Error &e) {
                    errors[t] = e.what();
                }
            });
        }
        for (auto &w : workers) w.join();
        for (const auto &e : errors)
            if (!e.empty()) throw std::runtime_error("worker: " + e);
        std::vector<uint64_t> toffs{0};
        std::vector<gis::Nam> tall;
        for (const auto &v : per) {
            tall.insert(tall.end(), v.begin(), v.end());
            toffs.push_back(tall.size());
        }
        write_out(std::string(argv[3]) + ".threads", toffs, tall);

        uint64_t bases = 0;
        std::vector<std::string_view> views;
        for (const auto &r : reads) {
            bases += r.size();
            views.emplace_back(r);
        }
        stage = "batch";
        gis::Batch batch(index, (uint32_t)reads.size(), bases);
        batch.upload(views);
        batch.run();
        const gis::BatchNams res = batch.download();
        write_out(std::string(argv[3]) + ".batch", res.offsets, res.nams);
        if (reads.size() % 2 == 0) {
            // reads p and n + p as the mates of pair p: the pair step
            stage = "pairs";
            const auto pr = batch.pairs();
            write_out(std::string(argv[3]) + ".pairs", pr.lists.offsets, pr.lists.nams);
            FILE *f = std::fopen((std::string(argv[3]) + ".types").c_str(), "wb");
            if (!f) throw std::runtime_error("cannot write the pair types");
            const uint64_t n = pr.types.size(), nj = pr.joint.size();
            std::fwrite(&n, 8, 1, f);
            std::fwrite(pr.types.data(), 1, n, f);
            std::fwrite(&nj, 8, 1, f);
            if (nj) std::fwrite(pr.joint.data(), sizeof(gis_nam_pair), nj, f);
            std::fclose(f);
        }
        std::cout << "facade_check: " << reads.size() << " reads, " << all.size() << " NAMs\n";
    } catch (const gis::Error &e) {
        std::cerr << "gis error " << e.code() << " in " << stage << ": " << e.what() << "\n";
        return 1;
    } catch (const std::exception &e) {
        std::cerr << "error in " << stage << ": " << e.what() << "\n";
        return 1;
    }
    return 0;
}
