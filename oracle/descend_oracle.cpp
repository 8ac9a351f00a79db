// descend_oracle.cpp -- CPU restatement of identify's rank-by-rank beam
// (targets/identify.cpp:89-202) with one integer prototype per category.
//
// TEST INFRASTRUCTURE ONLY: the checker of hvk_descend.  Only tests/ and
// bench.py's CPU baseline load it.
//
// Written in C++ so that the candidate sort is the host's own std::sort with
// identify's comparator (targets/identify.cpp:183-184): for equal scores the
// order it leaves is what decides which candidates the beam keeps.
//
// Scoring: the reference scores a candidate category by the max over its
// reference sequences of HVCache::similarity (targets/identify.cpp:150-157);
// a category that holds one sequence no longer than the similarity window
// scores cosine(read, that sequence) (src/encoder.cpp:198-203).  Here every
// category holds one prototype bundle (a leaf's sequence bundle, a parent's
// sum of its children's) and scores cosine(read, prototype) with the
// reference's formula (src/HV.cpp:219-242) on exact int64 sums.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

struct Cand {
    std::vector<int32_t> path;
    std::vector<float> path_scores;
    float score;
};

// src/HV.cpp:219-242 with int32 prototype entries
float ref_cosine(int D, const int16_t* q, const int32_t* p) {
    int64_t dot = 0, nl = 0, nr = 0;
    for (int d = 0; d < D; ++d) {
        const int64_t a = q[d], b = p[d];
        dot += a * b;
        nl += a * a;
        nr += b * b;
    }
    if (nl == 0 || nr == 0) return -1.0f;
    return static_cast<float>(2.0 * ((dot / (std::sqrt((double)nl) * std::sqrt((double)nr))) - 0.5f));
}

void descend_one(int D, const int32_t* proto, const int32_t* child_off, const int32_t* child_idx,
                 const int32_t* roots, int n_roots, const int16_t* q, int beam, float thr,
                 int levels, int32_t* path, float* pscore, int32_t* len, int32_t* exit_level) {
    std::vector<Cand> cur{Cand{{}, {}, -1.0f}};
    for (int x = 0; x < levels; ++x) {
        std::vector<Cand> next;
        for (const Cand& s : cur) {
            // Sequence::get_children (targets/identify.cpp:28-35): the roots
            // for an empty path, else the candidate list of the last category
            const int32_t* list = roots;
            int n = n_roots;
            if (!s.path.empty()) {
                const int c = s.path.back();
                list = child_idx + child_off[c];
                n = child_off[c + 1] - child_off[c];
            }
            for (int i = 0; i < n; ++i) {
                Cand c{s.path, s.path_scores, -1.0f};
                c.path.push_back(list[i]);
                next.push_back(std::move(c));
            }
        }
        for (Cand& c : next)  // targets/identify.cpp:150-165
            c.score = std::max(c.score, ref_cosine(D, q, proto + (int64_t)c.path.back() * D));
        for (Cand& c : next) c.path_scores.push_back(c.score);
        std::sort(next.begin(), next.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
        if (next.empty() || next[0].score < thr) {  // targets/identify.cpp:186-197
            *exit_level = x;
            *len = next.empty() ? 0 : (int32_t)next[0].path.size();
            for (int i = 0; i < *len; ++i) {
                path[i] = next[0].path[i];
                pscore[i] = next[0].path_scores[i];
            }
            return;
        }
        size_t keep = 0;
        while (keep < next.size() && keep < (size_t)beam && next[keep].score >= thr) ++keep;
        next.resize(keep);
        cur = std::move(next);
    }
    *exit_level = -1;
    *len = (int32_t)cur[0].path.size();
    for (int i = 0; i < *len; ++i) {
        path[i] = cur[0].path[i];
        pscore[i] = cur[0].path_scores[i];
    }
}

}  // namespace

extern "C" {

/* One read per call of descend_one, reads split over `threads` threads.
 * proto: [n_cat][D] int32; q: [n][D] int16; outputs as hvk_descend:
 * path / pscore [n][levels], len [n], exit_level [n] (-1: ran all levels). */
void orc_descend(int D, const int32_t* proto, const int32_t* child_off, const int32_t* child_idx,
                 const int32_t* roots, int n_roots, const int16_t* q, int n, int beam, float thr,
                 int levels, int32_t* path, float* pscore, int32_t* len, int32_t* exit_level,
                 int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([=] {
            for (int r = t; r < n; r += threads)
                descend_one(D, proto, child_off, child_idx, roots, n_roots, q + (int64_t)r * D,
                            beam, thr, levels, path + (int64_t)r * levels,
                            pscore + (int64_t)r * levels, len + r, exit_level + r);
        });
    for (auto& th : pool) th.join();
}

/* Candidate lists in the order the reference's Hierarchy hands them out:
 * the tree is read from taxonomy_index.tsv into the same container types,
 * filled in the same order (src/hierarchy.cpp:18-54), so iterating them
 * gives the same order as get_order("p__") (:110-116) and the children
 * vectors (:118-123).  names: n category names separated by '\n' (ids are
 * their positions); a category without children lists itself
 * (targets/identify.cpp:30-35).  Returns the number of list entries, or -1
 * when `cap` is too small. */
int orc_tax_order(const char* tsv_path, const char* names_blob, int n, int32_t* child_off,
                  int32_t* child_idx, int cap, int32_t* roots, int* n_roots) {
    std::unordered_map<std::string, int> id;
    {
        std::stringstream ns(names_blob);
        std::string nm;
        int i = 0;
        while (i < n && std::getline(ns, nm, '\n')) id.emplace(nm, i++);
    }
    std::unordered_map<std::string, std::unordered_set<std::string>> tmp;
    std::ifstream in(tsv_path);
    std::string line;
    std::getline(in, line);  // header
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::string seq_id, leaf, tax;
        {
            std::stringstream ss(line);
            std::getline(ss, seq_id, '\t');
            std::getline(ss, leaf, '\t');
            std::getline(ss, tax);
        }
        std::vector<std::string> ranks;
        std::string tok;
        std::stringstream ts(tax);
        while (std::getline(ts, tok, ';')) {
            tok.erase(0, tok.find_first_not_of(" "));
            if (tok.find("__") != std::string::npos &&
                !(tok.size() >= 2 && tok.compare(tok.size() - 2, 2, "__") == 0)) {
                for (char& ch : tok)
                    if (ch == ' ') ch = '_';
                ranks.push_back(tok);
            }
        }
        for (size_t i = 1; i < ranks.size(); ++i) tmp[ranks[i - 1]].insert(ranks[i]);
    }
    std::unordered_map<std::string, std::vector<std::string>> children;
    for (auto& [p, kids] : tmp) children[p] = std::vector<std::string>(kids.begin(), kids.end());
    int nr = 0;
    for (const auto& [p, kids] : children)
        if (p.compare(0, 3, "p__") == 0 && id.count(p)) roots[nr++] = id[p];
    *n_roots = nr;
    std::vector<std::string> names(n);
    for (auto& [nm, i] : id) names[i] = nm;
    int m = 0;
    child_off[0] = 0;
    for (int c = 0; c < n; ++c) {
        auto it = children.find(names[c]);
        if (it == children.end()) {
            if (m >= cap) return -1;
            child_idx[m++] = c;
        } else {
            for (const std::string& k : it->second) {
                if (m >= cap) return -1;
                auto f = id.find(k);
                child_idx[m++] = f == id.end() ? -1 : f->second;
            }
        }
        child_off[c + 1] = m;
    }
    return m;
}

/* Scores of one read against a list of categories, the reference's formula
 * (a probe for tests). */
void orc_descend_scores(int D, const int32_t* proto, const int16_t* q, const int32_t* cats, int n,
                        float* out) {
    for (int i = 0; i < n; ++i) out[i] = ref_cosine(D, q, proto + (int64_t)cats[i] * D);
}

}  // extern "C"
