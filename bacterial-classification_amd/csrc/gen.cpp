// gen.cpp -- deterministic synthetic reference sets and reads.
//
// Ancestors are uniform ACGT; each descendant copies its ancestor with
// per-base indels (half insertion of a uniform base, half deletion) and
// substitutions to one of the three other bases at a per-descendant rate;
// reads copy a uniform window of a uniform descendant with per-read
// substitution errors.  Every ancestor / descendant / read draws from its own
// std::mt19937_64 whose seed mixes (spec seed, purpose, index) with
// splitmix64, so output does not depend on thread count or call order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hdc_amd.h"

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

uint64_t stream_seed(uint64_t seed, uint64_t purpose, uint64_t index) {
    return splitmix64(seed ^ splitmix64((purpose << 48) ^ index));
}

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double u01() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
    int64_t uint_in(int64_t a, int64_t b) {  // inclusive
        if (b <= a) return a;
        int64_t v = a + (int64_t)(u01() * (double)(b - a + 1));
        return v > b ? b : v;
    }
    char base() { return "ACGT"[g() >> 62]; }
    char other(char c) {
        static const char* const others[4] = {"CGT", "AGT", "ACT", "ACG"};
        int i = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3;
        return others[i][uint_in(0, 2)];
    }
};

struct Dataset {
    std::vector<std::string> refs;
    std::vector<std::string> reads;
    std::vector<int32_t> truth;
};

template <typename F>
void parallel_for(int64_t n, F&& f) {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 32) t = 32;
    if (n < 64) t = 1;
    std::vector<std::thread> th;
    for (unsigned w = 0; w < t; ++w)
        th.emplace_back([&, w] {
            for (int64_t i = w; i < n; i += t) f(i);
        });
    for (auto& x : th) x.join();
}

bool valid(const hvk_gen_spec& s) {
    return s.n_ancestors > 0 && s.desc_per_ancestor > 0 && s.anc_len_min > 0 &&
           s.anc_len_max >= s.anc_len_min && s.div_min >= 0 && s.div_max >= s.div_min &&
           s.indel_rate >= 0 && s.indel_rate < 1 && s.n_reads >= 0 && s.read_len_min > 0 &&
           s.read_len_max >= s.read_len_min && s.err_min >= 0 && s.err_max >= s.err_min &&
           (int64_t)s.n_ancestors * s.desc_per_ancestor < (int64_t)1 << 31;
}

void generate(const hvk_gen_spec& s, Dataset& out) {
    std::vector<std::string> anc(s.n_ancestors);
    parallel_for(s.n_ancestors, [&](int64_t a) {
        Rng r(stream_seed(s.seed, 1, (uint64_t)a));
        const int64_t len = r.uint_in(s.anc_len_min, s.anc_len_max);
        std::string x((size_t)len, 'A');
        for (auto& c : x) c = r.base();
        anc[a] = std::move(x);
    });
    const int64_t n_refs = (int64_t)s.n_ancestors * s.desc_per_ancestor;
    out.refs.assign((size_t)n_refs, std::string());
    parallel_for(n_refs, [&](int64_t i) {
        Rng r(stream_seed(s.seed, 2, (uint64_t)i));
        const std::string& a = anc[(size_t)(i / s.desc_per_ancestor)];
        const double div = s.div_min + (s.div_max - s.div_min) * r.u01();
        std::string x;
        x.reserve(a.size() + a.size() / 16 + 8);
        for (char c : a) {
            const double u = r.u01();
            if (u < 0.5 * s.indel_rate) continue;              // deletion
            if (u < s.indel_rate) x.push_back(r.base());       // insertion before c
            x.push_back(r.u01() < div ? r.other(c) : c);
        }
        if (x.empty()) x.push_back(r.base());
        out.refs[(size_t)i] = std::move(x);
    });
    out.reads.assign((size_t)s.n_reads, std::string());
    out.truth.assign((size_t)s.n_reads, 0);
    parallel_for(s.n_reads, [&](int64_t i) {
        Rng r(stream_seed(s.seed, 3, (uint64_t)i));
        const int64_t ref = r.uint_in(0, n_refs - 1);
        const std::string& src = out.refs[(size_t)ref];
        int64_t len;
        if (s.read_len_log_uniform) {
            const double lo = std::log((double)s.read_len_min), hi = std::log((double)s.read_len_max);
            len = (int64_t)std::llround(std::exp(lo + (hi - lo) * r.u01()));
            len = std::clamp<int64_t>(len, s.read_len_min, s.read_len_max);
        } else {
            len = r.uint_in(s.read_len_min, s.read_len_max);
        }
        len = std::min<int64_t>(len, (int64_t)src.size());
        const int64_t start = r.uint_in(0, (int64_t)src.size() - len);
        const double err = s.err_min + (s.err_max - s.err_min) * r.u01();
        std::string x = src.substr((size_t)start, (size_t)len);
        for (auto& c : x)
            if (r.u01() < err) c = r.other(c);
        out.reads[(size_t)i] = std::move(x);
        out.truth[(size_t)i] = (int32_t)ref;
    });
}

std::mutex g_mu;
hvk_gen_spec g_spec;
bool g_have = false;
Dataset g_data;

const Dataset* cached(const hvk_gen_spec& s) {
    if (!g_have || std::memcmp(&g_spec, &s, sizeof(s)) != 0) {
        g_data = Dataset();
        generate(s, g_data);
        g_spec = s;
        g_have = true;
    }
    return &g_data;
}

}  // namespace

extern "C" int hvk_gen_size(const hvk_gen_spec* spec, int64_t* ref_bases, int64_t* read_bases) {
    if (!spec || !valid(*spec)) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    hvk_gen_spec s;
    std::memset(&s, 0, sizeof(s));
    s = *spec;
    const Dataset* d = cached(s);
    int64_t a = 0, b = 0;
    for (const auto& x : d->refs) a += (int64_t)x.size();
    for (const auto& x : d->reads) b += (int64_t)x.size();
    if (ref_bases) *ref_bases = a;
    if (read_bases) *read_bases = b;
    return HVK_OK;
}

extern "C" int hvk_gen_fill(const hvk_gen_spec* spec, char* refs, int64_t* ref_offsets, char* reads,
                            int64_t* read_offsets, int32_t* read_truth) {
    if (!spec || !valid(*spec)) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    hvk_gen_spec s;
    std::memset(&s, 0, sizeof(s));
    s = *spec;
    const Dataset* d = cached(s);
    if (refs && ref_offsets) {
        int64_t o = 0;
        for (size_t i = 0; i < d->refs.size(); ++i) {
            ref_offsets[i] = o;
            std::memcpy(refs + o, d->refs[i].data(), d->refs[i].size());
            o += (int64_t)d->refs[i].size();
        }
        ref_offsets[d->refs.size()] = o;
    }
    if (reads && read_offsets) {
        int64_t o = 0;
        for (size_t i = 0; i < d->reads.size(); ++i) {
            read_offsets[i] = o;
            std::memcpy(reads + o, d->reads[i].data(), d->reads[i].size());
            o += (int64_t)d->reads[i].size();
        }
        read_offsets[d->reads.size()] = o;
    }
    if (read_truth) std::memcpy(read_truth, d->truth.data(), sizeof(int32_t) * d->truth.size());
    // release the cache once filled
    g_data = Dataset();
    g_have = false;
    return HVK_OK;
}
