// identify -- drop-in for the reference's targets/identify.cpp: beam search
// (width 5, threshold 0 on the 2c - 1 score) down the taxonomy, scoring a
// read against a category by its best SEQUENCE_LENGTH window over the
// category's reference sequences.  Same usage, logs and output format.
//
// The reference scores each (read, window) pair with a CPU cosine; here the
// cache is touched in the reference's order (so the same entries are evicted
// and every window holds the reference's value), and each category's cosines
// are one GPU top-1 search of its candidate reads against its windows (the
// max over refs of the max over windows is the top-1 window).  The
// containers are the reference's where their iteration order reaches the
// output (reads, categories, beams).
#include <algorithm>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../host/HV.h"
#include "../host/encoder.h"
#include "../host/hierarchy.h"
#include "../host/session.h"

namespace {

constexpr float kThreshold = 0.0f;
constexpr int kBeamWidth = 5;
constexpr int kLevels = 6;

struct Beam {
    std::string sequence;
    std::vector<std::string> path;
    std::vector<float> path_scores;
    float score;

    std::vector<std::string> children(Hierarchy& tree) const {
        if (path.empty()) return tree.get_order("p__");
        const auto* kids = tree.get_children(path.back());
        if (!kids) return {path.back()};
        return *kids;
    }
};

// The scores of one category's candidates: the reference's loop
// (targets/identify.cpp:147-158) -- for each reference sequence,
// precompute(ref), then similarity(read, ref) for each candidate in order --
// replayed on the cache with the same touches (get(window) then get(read) per
// window; get(ref) then get(read) for a ref of at most SEQUENCE_LENGTH), and
// the cosines evaluated on the GPU.  The windows the first candidate sees
// are the rows W; a candidate that saw the same value (get_tagged's vid) at
// every window and one value of its read is scored by a top-1 search of its
// read against W.  A candidate that did not (an eviction in between) keeps
// its own (read value, window value) pairs, scored group by group.
void score_category(HVCache& cache, hvk_host::Session& gpu, const std::vector<std::string>& refs,
                    const std::vector<Beam*>& cands) {
    const size_t nc = cands.size();
    std::vector<HV_16> W, X, Q;  // window rows, extra window rows, read rows
    std::vector<uint64_t> wvid;
    std::map<std::pair<std::string, uint64_t>, int> q_row;
    std::vector<int> qrow(nc, -1);
    std::vector<uint64_t> qvid(nc, 0);
    std::vector<char> irregular(nc, 0);
    std::vector<std::vector<std::pair<int, int>>> pairs(nc);  // (Q row, W row or ~X row)
    std::vector<size_t> wbase;
    auto read_row = [&](const std::string& seq, uint64_t vid, const HV_16& v) {
        auto it = q_row.find({seq, vid});
        if (it != q_row.end()) return it->second;
        Q.push_back(v);
        q_row.emplace(std::make_pair(seq, vid), (int)Q.size() - 1);
        return (int)Q.size() - 1;
    };
    // candidate c becomes irregular before window p of ref j: its pairs so far
    auto materialize = [&](size_t c, size_t j, size_t p) {
        irregular[c] = 1;
        for (size_t jj = 0; jj <= j; ++jj) {
            const size_t end = jj < j ? (jj + 1 < wbase.size() ? wbase[jj + 1] : W.size()) : wbase[j] + p;
            for (size_t w = wbase[jj]; w < end; ++w) pairs[c].push_back({qrow[c], (int)w});
        }
    };
    for (size_t j = 0; j < refs.size(); ++j) {
        const std::string& ref = refs[j];
        cache.precompute(ref);
        wbase.push_back(W.size());
        const bool whole = SEQUENCE_LENGTH >= ref.size();
        for (size_t c = 0; c < nc; ++c) {
            const std::string& seq = cands[c]->sequence;
            size_t p = 0;
            for (size_t x = 0; whole ? x == 0 : x + SEQUENCE_LENGTH < ref.size();
                 x += SIM_STEP, ++p) {
                uint64_t wv = 0;
                const HV_16& w = cache.get_tagged(whole ? ref : ref.substr(x, SEQUENCE_LENGTH), wv);
                int wrow;
                if (c == 0) {
                    W.push_back(w);
                    wvid.push_back(wv);
                    wrow = (int)W.size() - 1;
                } else if (!irregular[c] && wv == wvid[wbase[j] + p]) {
                    wrow = (int)(wbase[j] + p);
                } else {
                    if (!irregular[c]) materialize(c, j, p);
                    X.push_back(w);
                    wrow = ~(int)(X.size() - 1);
                }
                uint64_t qv = 0;
                const HV_16& q = cache.get_tagged(seq, qv);
                if (qrow[c] < 0) {
                    qrow[c] = read_row(seq, qv, q);
                    qvid[c] = qv;
                } else if (!irregular[c] && qv != qvid[c]) {
                    materialize(c, j, p);
                }
                if (irregular[c]) pairs[c].push_back({read_row(seq, qv, q), wrow});
            }
        }
    }
    if (W.empty()) return;
    auto score_of_c = [](double c) { return (float)(2.0 * (c - 0.5)); };
    // regular candidates: one top-1 search of their reads against W
    std::vector<int> regular_rows;
    std::map<int, int> slot;  // Q row -> query slot
    for (size_t c = 0; c < nc; ++c)
        if (!irregular[c] && qrow[c] >= 0 && slot.emplace(qrow[c], (int)regular_rows.size()).second)
            regular_rows.push_back(qrow[c]);
    if (!regular_rows.empty()) {
        std::vector<const HV_16*> qp, wp;
        for (int r : regular_rows) qp.push_back(&Q[r]);
        for (auto& w : W) wp.push_back(&w);
        const auto top = gpu.search(gpu.upload_counts(qp), gpu.upload_counts(wp), 1);
        for (size_t c = 0; c < nc; ++c)
            if (!irregular[c] && qrow[c] >= 0)
                cands[c]->score = std::max(cands[c]->score, score_of_c(top.cos[slot[qrow[c]]]));
    }
    // irregular candidates: per read value, a top-1 search over its windows
    for (size_t c = 0; c < nc; ++c) {
        if (!irregular[c]) continue;
        std::map<int, std::vector<const HV_16*>> by_q;
        for (auto [qr, wr] : pairs[c]) by_q[qr].push_back(wr >= 0 ? &W[wr] : &X[~wr]);
        for (auto& [qr, wp] : by_q) {
            const auto top = gpu.search(gpu.upload_counts({&Q[qr]}), gpu.upload_counts(wp), 1);
            cands[c]->score = std::max(cands[c]->score, score_of_c(top.cos[0]));
        }
    }
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc != 5)
        throw std::runtime_error("Usage: identify sequence-directory taxonomy-file input-fasta output-file");
    Hierarchy tree(argv[1], argv[2]);
    HVCache cache;
    auto& gpu = hvk_host::Session::instance();
    std::unordered_map<std::string, std::vector<std::string>> outputs;

    // reads
    std::vector<std::string> reads;
    std::unordered_map<std::string, std::string> name_of;
    if (std::filesystem::exists(argv[3])) {
        std::ifstream in(argv[3]);
        std::string line, cur, name;
        while (std::getline(in, line)) {
            if (line.rfind(">", 0) == 0) {
                if (!cur.empty()) {
                    reads.push_back(cur);
                    name_of.emplace(cur, name);
                    cur.clear();
                }
                name = line.substr(1);
            } else {
                cur += line;
            }
        }
        if (!cur.empty()) {
            reads.push_back(cur);
            name_of.emplace(cur, name);
        }
    }
    std::cout << "Loaded " << reads.size() << " sequences\n";
    for (auto& s : reads) std::cout << "  len=" << s.size() << " name=" << name_of[s] << "\n";
    cache.prefetch(reads);

    std::unordered_map<std::string, std::vector<Beam>> beams;
    for (auto& s : reads) beams.emplace(s, std::vector<Beam>{Beam{s, {}, {}, -1.0f}});

    for (int level = 0; level < kLevels; ++level) {
        std::cout << "\n=== TAXONOMY LEVEL " << level << " ===\n";
        std::unordered_map<std::string, std::vector<Beam*>> by_category;
        for (auto& [target, cur] : beams) {
            std::cout << "Target len=" << target.size() << " beam=" << cur.size() << "\n";
            std::vector<Beam> next;
            size_t total = 0;
            for (auto& b : cur) total += b.children(tree).size();
            next.reserve(total);
            for (auto& b : cur) {
                const auto kids = b.children(tree);
                std::cout << "  Expanding [";
                for (auto& p : b.path) std::cout << p << " ";
                std::cout << "] (" << std::to_string(b.score) << ") -> " << kids.size() << " children\n";
                for (auto& child : kids) {
                    by_category[child];
                    std::vector<std::string> path = b.path;
                    path.push_back(child);
                    next.push_back(Beam{b.sequence, std::move(path), b.path_scores, -1.0f});
                    by_category[child].push_back(&next.back());
                }
            }
            cur = std::move(next);
        }

        // scoring: the reference's cache touches, cosines on the GPU
        for (auto& [child, cands] : by_category) {
            const auto& refs = tree.get_sequences(child);
            std::cout << " Scoring category " << child << " (" << cands.size() << " candidates)\n";
            score_category(cache, gpu, refs, cands);
        }

        for (auto& [target, bs] : beams)
            for (auto& b : bs) b.path_scores.push_back(b.score);

        // prune
        for (auto it = beams.begin(); it != beams.end();) {
            auto& target = it->first;
            auto& bs = it->second;
            std::cout << "  Pre-prune: " << bs.size() << "\n";
            std::sort(bs.begin(), bs.end(), [](auto& a, auto& b) { return a.score > b.score; });
            if (bs.empty() || bs[0].score < kThreshold) {
                std::cout << "  EARLY EXIT (best=" << (bs.empty() ? -1.0f : bs[0].score) << ")\n";
                outputs.emplace(target, bs.empty() ? std::vector<std::string>{} : bs[0].path);
                it = beams.erase(it);
                continue;
            }
            size_t keep = 0;
            while (keep < bs.size() && keep < (size_t)kBeamWidth && bs[keep].score >= kThreshold)
                ++keep;
            std::cout << "  Keeping " << keep << " candidates\n";
            bs.resize(keep);
            ++it;
        }
    }

    for (auto& [target, bs] : beams) {
        std::cout << "FINAL target len=" << target.size() << " score=" << bs[0].score << "\n";
        outputs[target] = bs[0].path;
    }

    std::ostream* os = &std::cout;
    std::ofstream file;
    if (std::filesystem::exists(argv[4])) {
        file.open(argv[4]);
        if (file) os = &file;
    }
    std::ostream& out = *os;
    out << "\nRESULTS\n";
    for (auto& [key, output] : outputs) {
        out << "\n" << name_of[key] << ":";
        const auto& bs = beams.count(key) ? beams[key] : std::vector<Beam>{};
        if (!bs.empty()) {
            const auto& best = bs[0];
            for (size_t i = 0; i < best.path.size(); ++i)
                out << " " << best.path[i] << " (" << best.path_scores[i] << ")";
        }
    }
    out << "\n";
}
