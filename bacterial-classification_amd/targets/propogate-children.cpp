// propogate-children -- drop-in for the reference's
// targets/propogate-children.cpp: for each rank p..s and each parent, every
// descendant sequence whose best 2c - 1 score against the parent's sequences
// (and those already added to it) is below 0.75 is appended to the parent's
// FASTA.  Same usage, progress output and histogram.
//
// The descendants' scores against the parent's sequences at the start of the
// parent are one GPU top-1 search; the sequentially growing set of sequences
// added to the parent is scored with one more top-1 search per descendant.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/HV.h"
#include "../host/encoder.h"
#include "../host/hierarchy.h"
#include "../host/session.h"

namespace {

constexpr double kThreshold = 0.75;
const std::vector<std::string> kOrder = {"k__", "p__", "c__", "o__", "f__", "g__", "s__"};

void collect(Hierarchy& tree, const std::string& node, std::vector<std::string>& out) {
    const auto* kids = tree.get_children(node);
    if (!kids) return;
    for (const auto& child : *kids) {
        const auto& seqs = tree.get_sequences(child);
        out.insert(out.end(), seqs.begin(), seqs.end());
        collect(tree, child, out);
    }
    std::cout << "\n\nFound " << std::to_string(out.size()) << " sequences in " << node;
}

float score_of(double c) { return (float)(2.0 * (c - 0.5)); }

// best 2c - 1 of q against rows (top-1 on the GPU), -inf for no rows
float best_against(hvk_host::Session& gpu, const std::vector<const HV_16*>& q,
                   const std::vector<const HV_16*>& rows, std::vector<float>& out) {
    out.assign(q.size(), -INFINITY);
    if (rows.empty() || q.empty()) return -INFINITY;
    const auto t = gpu.search(gpu.upload_counts(q), gpu.upload_counts(rows), 1);
    for (size_t i = 0; i < q.size(); ++i) out[i] = score_of(t.cos[i]);
    return out[0];
}

}  // namespace

int main(int argc, char* argv[]) {
    HVCache cache;
    std::cout << "\nCache size is " << std::to_string(CACHE_SIZE) << " entries ("
              << std::to_string(CACHE_BYTES) << "b)\n";
    std::vector<int> counts(101, 0);
    int total = 0;
    if (argc != 3) throw std::runtime_error("Usage: propogate-children sequence-directory taxonomy-file");
    Hierarchy tree(argv[1], argv[2]);
    auto& gpu = hvk_host::Session::instance();

    for (size_t x = 1; x < kOrder.size(); ++x) {
        for (const auto& parent : tree.get_order(kOrder[x])) {
            const auto& parent_seqs = tree.get_sequences(parent);
            std::vector<std::string> desc;
            collect(tree, parent, desc);
            cache.prefetch(desc);
            // descendants' bundles and their best score against the parent set
            std::vector<HV_16> dh;
            dh.reserve(desc.size());
            for (const auto& d : desc) dh.push_back(cache.get(d));
            std::vector<HV_16> ph;
            for (const auto& p : parent_seqs) ph.push_back(cache.get(p));
            std::vector<const HV_16*> dp, pp;
            for (auto& h : dh) dp.push_back(&h);
            for (auto& h : ph) pp.push_back(&h);
            std::vector<float> vs_parent;
            best_against(gpu, dp, pp, vs_parent);

            std::vector<HV_16> added;
            for (size_t i = 0; i < desc.size(); ++i) {
                std::cout << "\rAnalyzing " << parent << " " << std::to_string(i) << "/"
                          << std::to_string(desc.size()) << "        " << std::flush;
                float best = std::max(0.0f, vs_parent[i]);
                if (!added.empty()) {
                    std::vector<const HV_16*> ap;
                    for (auto& h : added) ap.push_back(&h);
                    std::vector<float> s;
                    best = std::max(best, best_against(gpu, {&dh[i]}, ap, s));
                }
                counts[(int)std::floor(best)]++;
                if (best < kThreshold) {
                    tree.append_sequence(parent, desc[i]);
                    added.push_back(dh[i]);
                    total += 1;
                }
            }
        }
    }
    std::cout << "\n\nbucket,count";
    for (int b = 0; b < 100; ++b) std::cout << "\n" << std::to_string(b / 100) << "," << std::to_string(counts[b]);
    std::cout << "\n\nPropogated: " << std::to_string(total) << "\n";
}
