// encoder.h -- drop-in for the reference's src/encoder.h (:1-42): HVCache,
// an LRU cache of sequence bundles, with the same members and semantics.
// Bundles are computed on the GPU (libhdc_amd); the LRU bookkeeping is
// host-side and follows the reference's order of operations, so cache hits,
// evictions and the sliding-window values of precompute() match it.
#pragma once
#include <list>
#include <string>
#include <unordered_map>
#include <vector>

#include "HV.h"

constexpr size_t CACHE_BYTES = 1'000'000'000;
constexpr size_t HV_MEM = HV_DIM * 2;
constexpr size_t CACHE_SIZE = CACHE_BYTES / HV_MEM;

struct SIM {
    size_t pos;
    float score;
};

class HVCache {
public:
    // kmers holds every k-mer over ACGT and its bound HV, as the reference's
    // table (src/encoder.cpp:47-74), for channel <= 10 (4^10 entries, ~1.3 GB
    // at D = 10000).  Above, it stays empty: the reference's table would take
    // ~5.4 GB at k = 11 and ~23 GB at k = 12, and nothing here reads it (the
    // GPU binds k-mers from the base HVs).
    explicit HVCache(size_t max_entries = CACHE_SIZE, size_t channel = KMER_SIZE);

    const HV_16& get(const std::string& seq);

    // Windows of SEQUENCE_LENGTH at SIM_STEP, cached with the reference's
    // incremental update (src/encoder.cpp:134-196).
    void precompute(const std::string& seq);

    // Best window of reference for query (src/encoder.cpp:198-215); the
    // window cosines are evaluated on the GPU in one batch.
    SIM similarity(const std::string& query, const std::string& reference);

    // Extension: computes the bundles of seqs in one GPU batch; later get()
    // misses on them use these values instead of a launch each.  The LRU is
    // not touched, so its state stays the reference's.
    void prefetch(const std::vector<std::string>& seqs);

    // Extension: get() that also names the value it returns: vid = 0 when
    // the entry holds compute(seq) (inserted by a get() miss), else a number
    // unique to the one precompute() insertion whose sliding-window value it
    // holds.  Equal (seq, vid) means equal values.
    const HV_16& get_tagged(const std::string& seq, uint64_t& vid);

    std::unordered_map<std::string, HV> kmers;

private:
    struct Entry {
        HV_16 hv;
        std::list<std::string>::iterator lru_it;
        uint64_t vid;
    };

    size_t capacity;
    std::list<std::string> lru;
    std::unordered_map<std::string, Entry> map;
    size_t channel;
    std::unordered_map<std::string, HV_16> prefetched;
    uint64_t next_vid = 0;

    HV_16 compute(const std::string& seq);
    void insert(const std::string& key, HV_16 hv, uint64_t vid);
};
