// sanitize_check -- runs the host-side code under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`, the
// reference's Debug flags: src/CMakeLists.txt:27,39-42): the drop-in's host
// primitives (host/HV.cpp) and taxonomy store (host/hierarchy.cpp), and the
// C oracle (oracle/hdc_oracle.c, oracle/descend_oracle.cpp), on small inputs
// that include the edge cases (empty and short sequences, N / IUPAC / lower
// case, ring and word boundaries of mult, zero norms, k > n, ties).  It also
// checks a few results against each other, so a run that is clean but wrong
// fails too.  Usage: sanitize_check <scratch directory>.  Prints "ok" and
// exits 0; a sanitizer report or a mismatch exits non-zero.
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../bacterial-classification_amd/host/HV.h"
#include "../../bacterial-classification_amd/host/hierarchy.h"
#include "../../oracle/hdc_oracle.h"

extern "C" void orc_descend(int D, const int32_t* proto, const int32_t* child_off,
                            const int32_t* child_idx, const int32_t* roots, int n_roots,
                            const int16_t* q, int n, int beam, float thr, int levels,
                            int32_t* path, float* pscore, int32_t* len, int32_t* exit_level,
                            int threads);

static int g_fail = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                  \
        }                                                                \
    } while (0)

// uniform ACGT; with noisy, 1 base in 25 is one of N R Y a c g t
static std::string random_seq(std::mt19937_64& g, size_t n, bool noisy) {
    const char* acgt = "ACGT";
    const char* other = "NRYacgt";
    std::string s(n, 'A');
    for (auto& c : s) c = (noisy && g() % 25 == 0) ? other[g() % 7] : acgt[g() % 4];
    return s;
}

static void host_primitives() {
    // host/HV.cpp at the compile-time configuration (HV_DIM, KMER_SIZE)
    const HV e = empty_hv();
    CHECK(e.size() == HV_WORDS);
    const char* syms = "ACGTNx";
    std::vector<uint64_t> ob(HV_WORDS);
    for (const char* p = syms; *p; ++p) {
        orc_base_hv(HV_DIM, *p, ob.data());
        CHECK(std::memcmp(ob.data(), get_base_hv(*p).data(), 8 * HV_WORDS) == 0);
    }
    std::mt19937_64 g(1);
    HV l(HV_WORDS), r(HV_WORDS), o;
    for (auto& w : l) w = g();
    for (auto& w : r) w = g();
    l.back() &= e.back();
    r.back() &= e.back();
    for (size_t s : {size_t(0), size_t(1), size_t(63), size_t(64), size_t(65), size_t(HV_DIM - 1),
                     size_t(HV_DIM), size_t(64 * HV_WORDS - 1), size_t(64 * HV_WORDS + 3)}) {
        mult(l, r, o, s);
        orc_mult(HV_DIM, l.data(), r.data(), ob.data(), s);
        CHECK(std::memcmp(ob.data(), o.data(), 8 * HV_WORDS) == 0);
        HV a = l;
        mult(a, r, a, s);  // aliasing
        CHECK(a == o);
    }
    std::vector<HV> batch(16, HV(HV_WORDS));
    for (auto& h : batch) {
        for (auto& w : h) w = g();
        h.back() &= e.back();
    }
    HV_16 acc = make_accumulator(), ref(HV_DIM, 0);
    superpose_batch(acc, batch.data(), 16);
    std::vector<const uint64_t*> bp;
    for (auto& h : batch) bp.push_back(h.data());
    orc_superpose_batch(HV_DIM, ref.data(), bp.data(), 16);
    CHECK(acc == ref);
    std::vector<const HV*> olds, news;
    for (int i = 0; i < 8; ++i) {
        olds.push_back(&batch[i]);
        news.push_back(&batch[8 + i]);
    }
    replace_batch(acc, olds.data(), news.data(), 8);
    std::vector<const uint64_t*> op(bp.begin(), bp.begin() + 8), np(bp.begin() + 8, bp.end());
    orc_replace_batch(HV_DIM, ref.data(), op.data(), np.data(), 8);
    CHECK(acc == ref);
    HV_16 zero(HV_DIM, 0);
    CHECK(cosine(zero, acc) == -1.0f);
    CHECK(cosine(acc, acc) == orc_cosine(HV_DIM, acc.data(), acc.data(), 1));
}

static void host_hierarchy(const std::filesystem::path& dir) {
    // host/hierarchy.cpp on a small taxonomy (header line skipped, a space in
    // a name, a category without a FASTA file), then appends
    std::filesystem::create_directories(dir / "seqs");
    {
        std::ofstream t(dir / "tax.tsv");
        t << "seq_id\tleaf\ttaxonomy\n"
          << "a\ts__x\tk__B; p__P one; c__C1; o__O; f__F; g__G; s__x\n"
          << "b\ts__y\tk__B; p__P one; c__C2; o__O2; f__F2; g__G2; s__y\n\n";
        std::ofstream f(dir / "seqs" / "c__C1.fasta");
        f << ">c1_0\nACGTACGT\nACGT\n>c1_1\nTTTT\n";
    }
    Hierarchy h((dir / "seqs").string(), (dir / "tax.tsv").string());
    const auto& s = h.get_sequences("c__C1");
    CHECK(s.size() == 2 && s[0] == "ACGTACGTACGT" && s[1] == "TTTT");
    CHECK(h.get_sequences("c__C2").empty());
    CHECK(h.get_order("p__").size() == 1);
    CHECK(h.get_children("p__P_one") != nullptr || h.get_children("p__P one") != nullptr);
    h.append_sequence("c__C2", "GGGG");
    h.append_sequence("c__C1", "CCCC");
    CHECK(h.get_sequences("c__C1").size() == 3);
}

static void oracle_functions() {
    std::mt19937_64 g(7);
    for (int D : {100, 192, 8192}) {
        const int k = D == 100 ? 5 : 10;
        std::vector<std::string> seqs = {"", "A", std::string(k - 1, 'C'), std::string(k, 'G'),
                                         random_seq(g, 300, true), random_seq(g, 260, false)};
        std::vector<int16_t> c((size_t)seqs.size() * D);
        for (size_t i = 0; i < seqs.size(); ++i)
            orc_compute(D, k, seqs[i].data(), seqs[i].size(), c.data() + i * D);
        std::vector<int32_t> c32(D);
        orc_compute32(D, k, seqs[4].data(), seqs[4].size(), c32.data());
        for (int d = 0; d < D; ++d) CHECK(c32[d] == c[4 * (size_t)D + d]);
        const int nq = 3, ndb = (int)seqs.size(), kk = 8;  // k > n: padded
        std::vector<int32_t> idx((size_t)nq * kk);
        std::vector<double> cs((size_t)nq * kk);
        orc_topk16(D, c.data() + 3 * (size_t)D, nq, c.data(), ndb, kk, idx.data(), cs.data(), 2);
        CHECK(idx[kk - 1] == -1);
        std::vector<int16_t> wins((size_t)64 * D);
        const std::string ref = random_seq(g, 400, true);
        const int nw = orc_precompute_windows(D, k, 125, 10, ref.data(), ref.size(), wins.data(), 64);
        CHECK(nw > 20);
        size_t pos = 0;
        const std::string q = ref.substr(40, 125);
        const float sc = orc_precompute_similarity(D, k, 125, 10, q.data(), q.size(), ref.data(),
                                                   ref.size(), &pos);
        CHECK(sc > 0.0f);
        std::string reads = seqs[4] + seqs[5];
        const int64_t off[3] = {0, (int64_t)seqs[4].size(), (int64_t)reads.size()};
        std::vector<int64_t> dbn(ndb);
        for (int i = 0; i < ndb; ++i)
            dbn[i] = orc_dot16(D, c.data() + (size_t)i * D, c.data() + (size_t)i * D);
        std::vector<int32_t> ci(2 * 3);
        std::vector<double> cc(2 * 3);
        orc_classify_reads(D, k, reads.data(), off, 2, c.data(), dbn.data(), ndb, 3, ci.data(),
                           cc.data(), 2);
        CHECK(ci[0] == 4 && ci[3] == 5);  // each read finds itself
    }
    // a three-level tree: root 0 -> {1, 2}; 1 -> {3, 4}; 2 -> {5}; leaves list themselves
    const int D = 256, n_cat = 6;
    std::vector<int32_t> proto((size_t)n_cat * D);
    for (auto& v : proto) v = (int32_t)(g() % 41) - 20;
    const int32_t off[n_cat + 1] = {0, 2, 4, 5, 6, 7, 8};
    const int32_t idx[8] = {1, 2, 3, 4, 5, 3, 4, 5};
    const int32_t roots[1] = {0};
    std::vector<int16_t> q((size_t)4 * D);
    for (auto& v : q) v = (int16_t)((int)(g() % 21) - 10);
    std::vector<int32_t> path(4 * 3), len(4), ex(4);
    std::vector<float> ps(4 * 3);
    orc_descend(D, proto.data(), off, idx, roots, 1, q.data(), 4, 2, -1.0f, 3, path.data(),
                ps.data(), len.data(), ex.data(), 2);
    for (int r = 0; r < 4; ++r) CHECK(path[r * 3] == 0 && len[r] == 3);
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: sanitize_check <scratch directory>\n");
        return 2;
    }
    host_primitives();
    host_hierarchy(argv[1]);
    oracle_functions();
    if (g_fail) return 1;
    std::printf("ok\n");
    return 0;
}
