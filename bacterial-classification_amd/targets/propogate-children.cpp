// propogate-children -- drop-in for the reference's
// targets/propogate-children.cpp: for each rank p..s and each parent, every
// descendant sequence whose best 2c - 1 score against the parent's sequences
// (and those already added to it) is below 0.75 is appended to the parent's
// FASTA.  Same usage, progress output and histogram.
//
// The descendants' bundles are encoded once and stay on the device; so does
// the set of sequences added to the parent, which grows by device row copies.
// Descendants are scored in blocks of kBlock: one top-1 search of the block
// against the parent's sequences and one against the sequences added before
// the block, plus the block against itself (every pair, exact), so the
// sequential acceptance inside the block runs on the host with no further
// device work.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <string_view>
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

constexpr int kBlock = 256;

// best 2c - 1 of each row of q against r (top-1 on the GPU), -inf for no rows
std::vector<float> best_against(hvk_host::Session& gpu, const hvk_bundles& q,
                                const hvk_bundles& r) {
    std::vector<float> out((size_t)q.n, -INFINITY);
    if (q.n == 0 || r.n == 0) return out;
    const auto t = gpu.search(q, r, 1);
    for (int i = 0; i < q.n; ++i) out[i] = score_of(t.cos[i]);
    return out;
}

}  // namespace

int main(int argc, char* argv[]) {
    std::cout << "\nCache size is " << std::to_string(CACHE_SIZE) << " entries ("
              << std::to_string(CACHE_BYTES) << "b)\n";
    std::vector<int> counts(101, 0);
    int total = 0;
    if (argc != 3) throw std::runtime_error("Usage: propogate-children sequence-directory taxonomy-file");
    Hierarchy tree(argv[1], argv[2]);
    auto& gpu = hvk_host::Session::instance();

    for (size_t x = 1; x < kOrder.size(); ++x) {
        for (const auto& parent : tree.get_order(kOrder[x])) {
            const std::vector<std::string> parent_seqs = tree.get_sequences(parent);
            std::vector<std::string> desc;
            collect(tree, parent, desc);
            const int n = (int)desc.size();
            // bundles on the device: descendants, the parent's sequences, and
            // room for every descendant in the added set
            std::vector<std::string_view> dv(desc.begin(), desc.end()),
                pv(parent_seqs.begin(), parent_seqs.end());
            hvk_host::DeviceBundles dh = gpu.encode(gpu.pack(dv));
            hvk_host::DeviceBundles ph = gpu.encode(gpu.pack(pv));
            hvk_host::DeviceBundles added = gpu.alloc(std::max(n, 1));
            // hi limbs wherever any bundle needs them (exact either way)
            const bool two = dh.needs_two_limbs() || ph.needs_two_limbs();
            const std::vector<float> vs_parent =
                best_against(gpu, hvk_host::Session::rows(dh, 0, n, two),
                             hvk_host::Session::rows(ph, 0, ph.n, two));
            std::vector<int> block_added;  // block offsets of this block's acceptances
            for (int i0 = 0; i0 < n; i0 += kBlock) {
                const int i1 = std::min(n, i0 + kBlock), m = i1 - i0;
                const hvk_bundles blk = hvk_host::Session::rows(dh, i0, i1, two);
                const std::vector<float> vs_added =
                    best_against(gpu, blk, hvk_host::Session::rows(added, 0, added.n, two));
                // every pair of the block: s[a][b] = score of descendant i0+a vs i0+b
                const auto t = gpu.search(blk, blk, m);
                std::vector<float> pair((size_t)m * m, -INFINITY);
                for (int a2 = 0; a2 < m; ++a2)
                    for (int j = 0; j < m; ++j) {
                        const int b2 = t.idx[(size_t)a2 * m + j];
                        if (b2 >= 0) pair[(size_t)a2 * m + b2] = score_of(t.cos[(size_t)a2 * m + j]);
                    }
                block_added.clear();
                for (int a2 = 0; a2 < m; ++a2) {
                    const int i = i0 + a2;
                    std::cout << "\rAnalyzing " << parent << " " << std::to_string(i) << "/"
                              << std::to_string(n) << "        " << std::flush;
                    float best = std::max(0.0f, vs_parent[i]);
                    best = std::max(best, vs_added[a2]);
                    for (int b2 : block_added) best = std::max(best, pair[(size_t)a2 * m + b2]);
                    counts[(int)std::floor(best)]++;
                    if (best < kThreshold) {
                        tree.append_sequence(parent, desc[i]);
                        block_added.push_back(a2);
                        total += 1;
                    }
                }
                for (int b2 : block_added) gpu.append_row(added, dh, i0 + b2);
            }
        }
    }
    std::cout << "\n\nbucket,count";
    for (int b = 0; b < 100; ++b) std::cout << "\n" << std::to_string(b / 100) << "," << std::to_string(counts[b]);
    std::cout << "\n\nPropogated: " << std::to_string(total) << "\n";
}
