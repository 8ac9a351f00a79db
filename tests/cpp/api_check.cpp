// api_check -- drives the drop-in C++ API (HVCache, cosine, encode,
// cosine_search) on inputs from a text file and writes the results as raw
// little-endian arrays, for tests/test_gpu_cpp_api.py to compare with the
// golden vectors of the reference's own build; `api_check prims` drives the
// host primitives alone (tests/test_cpp_prims.py).
//
// input file: sections "SEQS n", "REFS n", "QUERIES n", one sequence per line
// output file: counts of get(seq) [n][HV_DIM] int16, cosine matrix [n][n]
// float, precompute windows per ref (count int32 + rows int16), similarity
// [nq][nr] (float score, int64 pos), encode() window counts per seq (int32
// count + rows), cosine_search top-3 of get(seq[i]) against all [n][3]
// (float score, int32 index).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../bacterial-classification_amd/host/GPU.h"
#include "../../bacterial-classification_amd/host/HV.h"
#include "../../bacterial-classification_amd/host/encoder.h"

static std::vector<std::string> section(std::istream& in) {
    std::string tag;
    size_t n = 0;
    in >> tag >> n;
    std::vector<std::string> v(n);
    std::string line;
    std::getline(in, line);
    for (auto& s : v) std::getline(in, s);
    return v;
}

template <typename T>
static void put(std::FILE* f, const T* p, size_t n) {
    std::fwrite(p, sizeof(T), n, f);
}

template <typename T>
static void get(std::FILE* f, T* p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
}

// prims mode: the host primitives (no device).  Input (raw little-endian):
// int32 n, then n x (int64 shift, HV l, HV r) for mult; int32 n, then n x
// (int32 b, HV_16 acc, b HVs) for superpose_batch; int32 n, then n x (int32
// b, HV_16 acc, b old HVs, b new HVs) for replace_batch.  Output: empty_hv,
// get_base_hv of "ACGTNx", each mult result (then the same with out
// aliasing l), each superpose / replace result, then kmers.size() (int64)
// and the HV of every k-mer over ACGT in lexicographic order (zeros when
// absent).
static int prims_main(const char* in_path, const char* out_path) {
    std::FILE* in = std::fopen(in_path, "rb");
    std::FILE* f = std::fopen(out_path, "wb");
    if (!in || !f) return 2;
    auto read_hv = [&]() { HV h(HV_WORDS); get(in, h.data(), HV_WORDS); return h; };
    const HV e = empty_hv();
    put(f, e.data(), HV_WORDS);
    for (char c : std::string("ACGTNx")) put(f, get_base_hv(c).data(), HV_WORDS);
    int32_t n = 0;
    get(in, &n, 1);
    std::vector<HV> alias;
    for (int i = 0; i < n; ++i) {
        int64_t s = 0;
        get(in, &s, 1);
        HV l = read_hv(), r = read_hv(), o;
        mult(l, r, o, (size_t)s);
        put(f, o.data(), HV_WORDS);
        mult(l, r, l, (size_t)s);  // out aliasing l
        alias.push_back(l);
    }
    for (const auto& h : alias) put(f, h.data(), HV_WORDS);
    get(in, &n, 1);
    for (int i = 0; i < n; ++i) {
        int32_t b = 0;
        get(in, &b, 1);
        HV_16 acc(HV_DIM);
        get(in, acc.data(), HV_DIM);
        std::vector<HV> hv;
        for (int j = 0; j < b; ++j) hv.push_back(read_hv());
        superpose_batch(acc, hv.data(), (size_t)b);
        put(f, acc.data(), HV_DIM);
    }
    get(in, &n, 1);
    for (int i = 0; i < n; ++i) {
        int32_t b = 0;
        get(in, &b, 1);
        HV_16 acc(HV_DIM);
        get(in, acc.data(), HV_DIM);
        std::vector<HV> olds, news;
        for (int j = 0; j < b; ++j) olds.push_back(read_hv());
        for (int j = 0; j < b; ++j) news.push_back(read_hv());
        std::vector<const HV*> op, np;
        for (int j = 0; j < b; ++j) { op.push_back(&olds[j]); np.push_back(&news[j]); }
        replace_batch(acc, op.data(), np.data(), (size_t)b);
        put(f, acc.data(), HV_DIM);
    }
    HVCache cache(16);
    const int64_t size = (int64_t)cache.kmers.size();
    put(f, &size, 1);
    const char acgt[4] = {'A', 'C', 'G', 'T'};
    std::string s(KMER_SIZE, 'A');
    const HV zero(HV_WORDS, 0);
    for (size_t i = 0; i < (size_t(1) << (2 * KMER_SIZE)); ++i) {
        for (size_t p = 0; p < KMER_SIZE; ++p) s[p] = acgt[(i >> (2 * (KMER_SIZE - 1 - p))) & 3];
        auto it = cache.kmers.find(s);
        put(f, (it == cache.kmers.end() ? zero : it->second).data(), HV_WORDS);
    }
    std::fclose(in);
    std::fclose(f);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::string(argv[1]) == "prims") return prims_main(argv[2], argv[3]);
    if (argc != 3) {
        std::cerr << "usage: api_check input.txt output.bin | api_check prims in.bin out.bin\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    const auto seqs = section(in), refs = section(in), queries = section(in);
    std::FILE* f = std::fopen(argv[2], "wb");
    HVCache cache(1u << 20);
    std::vector<HV_16> hv;
    for (const auto& s : seqs) hv.push_back(cache.get(s));
    for (const auto& h : hv) put(f, h.data(), HV_DIM);
    for (size_t i = 0; i < hv.size(); ++i)
        for (size_t j = 0; j < hv.size(); ++j) {
            const float c = cosine(hv[i], hv[j]);
            put(f, &c, 1);
        }
    for (const auto& r : refs) {
        HVCache c(1u << 20);
        c.precompute(r);
        std::vector<HV_16> wins;
        if (r.size() >= SEQUENCE_LENGTH)
            for (size_t x = 0; x == 0 || x + SEQUENCE_LENGTH < r.size(); x += SIM_STEP)
                wins.push_back(c.get(r.substr(x, SEQUENCE_LENGTH)));
        const int32_t n = (int32_t)wins.size();
        put(f, &n, 1);
        for (const auto& w : wins) put(f, w.data(), HV_DIM);
    }
    for (const auto& q : queries)
        for (const auto& r : refs) {
            HVCache c(1u << 20);
            c.precompute(r);
            const SIM s = c.similarity(q, r);
            const int64_t pos = (int64_t)s.pos;
            put(f, &s.score, 1);
            put(f, &pos, 1);
        }
    const auto enc = encode({}, seqs);
    for (const auto& wins : enc) {
        const int32_t n = (int32_t)wins.size();
        put(f, &n, 1);
        for (const auto& w : wins) put(f, w.data(), HV_DIM);
    }
    std::vector<HV_16*> ptrs;
    for (auto& h : hv) ptrs.push_back(&h);
    const auto res = cosine_search(ptrs, ptrs, 3);
    for (auto* p : ptrs) {
        const auto& m = res.at(p);
        for (const auto& [score, q] : m) {
            const int32_t idx = (int32_t)(q - &hv[0]);
            put(f, &score, 1);
            put(f, &idx, 1);
        }
    }
    std::fclose(f);
    return 0;
}
