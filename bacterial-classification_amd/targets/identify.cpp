// identify -- drop-in for the reference's targets/identify.cpp: beam search
// (width 5, threshold 0 on the 2c - 1 score) down the taxonomy, scoring a
// read against a category by its best SEQUENCE_LENGTH window over the
// category's reference sequences.  Same usage, logs and output format.
//
// The reference scores each (read, window) pair with a CPU cosine; here each
// category is one GPU top-1 search of its candidate reads against all of its
// windows (the max over refs of the max over windows is the top-1 window).
// Window bundles come from HVCache::precompute, called per reference in the
// reference's order, so cache hits and sliding-window values match.  The
// containers are the reference's where their iteration order reaches the
// output (reads, categories, beams).
#include <algorithm>
#include <filesystem>
#include <fstream>
#include <iostream>
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

// float(2c - 1) from an FP64 cosine, as cosine() forms it
float score_of(double c) { return (float)(2.0 * (c - 0.5)); }

// All window bundles similarity(read, ref) scans for every ref of a category,
// gathered through the cache in the reference's order.
void category_windows(HVCache& cache, const std::vector<std::string>& refs,
                      std::vector<HV_16>& out) {
    for (const auto& ref : refs) {
        cache.precompute(ref);
        if (SEQUENCE_LENGTH >= ref.size()) {
            out.push_back(cache.get(ref));
            continue;
        }
        for (size_t x = 0; x + SEQUENCE_LENGTH < ref.size(); x += SIM_STEP)
            out.push_back(cache.get(ref.substr(x, SEQUENCE_LENGTH)));
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

        // scoring: one GPU top-1 search per category
        for (auto& [child, cands] : by_category) {
            const auto& refs = tree.get_sequences(child);
            std::cout << " Scoring category " << child << " (" << cands.size() << " candidates)\n";
            std::vector<HV_16> windows;
            category_windows(cache, refs, windows);
            if (windows.empty() || cands.empty()) continue;
            std::vector<const HV_16*> qrows;
            std::unordered_map<std::string, int> row_of;
            for (Beam* b : cands) {
                if (row_of.emplace(b->sequence, (int)qrows.size()).second)
                    qrows.push_back(&cache.get(b->sequence));
            }
            std::vector<HV_16> qcopy;
            qcopy.reserve(qrows.size());
            for (const HV_16* p : qrows) qcopy.push_back(*p);
            std::vector<const HV_16*> qp, wp;
            for (auto& q : qcopy) qp.push_back(&q);
            for (auto& w : windows) wp.push_back(&w);
            const auto qd = gpu.upload_counts(qp);
            const auto wd = gpu.upload_counts(wp);
            const auto top = gpu.search(qd, wd, 1);
            for (Beam* b : cands) {
                const float s = score_of(top.cos[row_of[b->sequence]]);
                b->score = std::max(b->score, s);
            }
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
