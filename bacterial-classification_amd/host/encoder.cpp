// encoder.cpp -- HVCache over the GPU encoder (semantics of
// src/encoder.cpp:47-215).
#include "encoder.h"

#include <cmath>
#include <stdexcept>
#include <string_view>

#include "session.h"

namespace {

const char kAlphabet[4] = {'A', 'C', 'G', 'T'};

char comp(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return 'N';
    }
}

bool acgt(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

// float(2c - 1) from an FP64 cosine c, as cosine() forms it (src/HV.cpp:239)
float two_c_minus_one(double c) { return (float)(2.0 * (c - 0.5)); }

}  // namespace

HVCache::HVCache(size_t max_entries, size_t channel_) : capacity(max_entries), channel(channel_) {
    if (channel > 10) return;
    // every k-mer s maps to the HV of min(s, revcomp(s)) bound by mult
    const uint64_t total = 1ull << (2 * channel);
    std::string s(channel, 'A'), rc(channel, 'A');
    for (uint64_t i = 0; i < total; ++i) {
        uint64_t x = i;
        for (size_t p = channel; p-- > 0;) {
            s[p] = kAlphabet[x & 3];
            x >>= 2;
        }
        for (size_t p = 0; p < channel; ++p) rc[channel - 1 - p] = comp(s[p]);
        const std::string& canon = rc < s ? rc : s;
        auto it = kmers.find(canon);
        if (it == kmers.end()) {
            HV h = empty_hv();
            for (size_t p = 0; p < channel; ++p) mult(h, get_base_hv(canon[p]), h, p);
            it = kmers.emplace(canon, std::move(h)).first;
        }
        kmers.emplace(s, it->second);
    }
}

HV_16 HVCache::compute(const std::string& seq) {
    auto pf = prefetched.find(seq);
    if (pf != prefetched.end()) {
        HV_16 h = std::move(pf->second);
        prefetched.erase(pf);
        return h;
    }
    auto& s = hvk_host::Session::instance();
    const hvk_host::DevicePacked p = s.pack({std::string_view(seq)});
    return std::move(s.download_counts(s.encode(p))[0]);
}

void HVCache::prefetch(const std::vector<std::string>& seqs) {
    std::vector<std::string_view> todo;
    for (const auto& q : seqs)
        if (!map.count(q) && !prefetched.count(q)) todo.emplace_back(q);
    if (todo.empty()) return;
    auto& s = hvk_host::Session::instance();
    const hvk_host::DevicePacked p = s.pack(todo);
    std::vector<HV_16> rows = s.download_counts(s.encode(p));
    for (size_t i = 0; i < todo.size(); ++i) prefetched.emplace(std::string(todo[i]), std::move(rows[i]));
}

void HVCache::insert(const std::string& key, HV_16 hv, uint64_t vid) {
    if (map.size() >= capacity) {
        map.erase(lru.back());
        lru.pop_back();
    }
    lru.push_front(key);
    map.emplace(key, Entry{std::move(hv), lru.begin(), vid});
}

const HV_16& HVCache::get_tagged(const std::string& seq, uint64_t& vid) {
    auto it = map.find(seq);
    if (it != map.end()) {
        lru.splice(lru.begin(), lru, it->second.lru_it);
        vid = it->second.vid;
        return it->second.hv;
    }
    insert(seq, compute(seq), 0);
    vid = 0;
    return map.find(seq)->second.hv;
}

const HV_16& HVCache::get(const std::string& seq) {
    uint64_t vid;
    return get_tagged(seq, vid);
}

// Windows w = 0, 1, ... start at SIM_STEP * w.  Window 0 is get() of its
// substring.  Window w >= 1, unless its string is already cached, is window
// w-1 plus the bundle of the SIM_STEP k-mers at positions
// SIM_STEP*w + SEQUENCE_LENGTH - SIM_STEP - k + x minus those at
// SIM_STEP*(w-1) + x (x < SIM_STEP), a k-mer outside ACGT counting as the
// all-ones HV (+1 in every dimension) -- the reference's sliding update
// (src/encoder.cpp:156-195), whose new positions sit one base before the
// window's last k-mers.  Both step bundles come from one GPU launch of
// windows of SIM_STEP + k - 1 bases (which skip non-ACGT k-mers), corrected
// by the count of non-ACGT k-mers in each range.
void HVCache::precompute(const std::string& seq) {
    const size_t len = seq.size();
    if (len < SEQUENCE_LENGTH) return;
    const size_t k = channel;
    std::vector<int64_t> starts;
    std::vector<int> z;  // (#invalid new) - (#invalid old) per step
    std::vector<char> valid(len >= k ? len - k + 1 : 0, 0);
    for (size_t i = 0; i + k <= len; ++i) {
        bool ok = true;
        for (size_t j = 0; j < k; ++j) ok = ok && acgt(seq[i + j]);
        valid[i] = ok;
    }
    for (size_t index = SIM_STEP; index + SEQUENCE_LENGTH < len; index += SIM_STEP) {
        const size_t nw = index + SEQUENCE_LENGTH - SIM_STEP - k, od = index - SIM_STEP;
        starts.push_back((int64_t)nw);
        starts.push_back((int64_t)od);
        int zi = 0;
        for (size_t x = 0; x < SIM_STEP; ++x) zi += (!valid[nw + x]) - (!valid[od + x]);
        z.push_back(zi);
    }
    HV_16 window = get(seq.substr(0, SEQUENCE_LENGTH));
    if (starts.empty()) return;
    auto& s = hvk_host::Session::instance();
    const hvk_host::DevicePacked p = s.pack({std::string_view(seq)});
    const std::vector<HV_16> d =
        s.download_counts(s.encode_windows(p, starts, (int)(SIM_STEP + k - 1)));
    size_t step = 0;
    for (size_t index = SIM_STEP; index + SEQUENCE_LENGTH < len; index += SIM_STEP, ++step) {
        const std::string key = seq.substr(index, SEQUENCE_LENGTH);
        auto it = map.find(key);
        if (it != map.end()) {
            lru.splice(lru.begin(), lru, it->second.lru_it);
            window = it->second.hv;
            continue;
        }
        const HV_16& add = d[2 * step];
        const HV_16& sub = d[2 * step + 1];
        for (size_t j = 0; j < HV_DIM; ++j)
            window[j] = (int16_t)(window[j] + add[j] - sub[j] + z[step]);
        insert(key, window, ++next_vid);
    }
}

SIM HVCache::similarity(const std::string& l, const std::string& r) {
    SIM sim{0, -1.0f};
    if (SEQUENCE_LENGTH >= r.size()) {
        // cosine(get(l), get(r)) in the reference: GCC evaluates the
        // arguments right to left, so get(r) touches the cache first
        const HV_16 ref = get(r);
        const HV_16& q = get(l);
        sim.score = cosine(q, ref);
        return sim;
    }
    // the windows through get(), each followed by get(l), the reference's
    // order of cache touches (src/encoder.cpp:205-207).  get(l) returns the
    // same value at every window unless an eviction in between replaced a
    // sliding-window value by the exact one; consecutive windows that saw the
    // same value of l form one segment.
    std::vector<HV_16> wins, qs;
    std::vector<size_t> pos, seg_end;  // segment i: windows [seg_end[i-1], seg_end[i])
    uint64_t last_vid = ~uint64_t(0);
    for (size_t x = 0; x + SEQUENCE_LENGTH < r.size(); x += SIM_STEP) {
        wins.push_back(get(r.substr(x, SEQUENCE_LENGTH)));
        pos.push_back(x);
        uint64_t vid = 0;
        const HV_16& q = get_tagged(l, vid);
        if (qs.empty() || vid != last_vid) {
            if (!qs.empty()) seg_end.push_back(wins.size() - 1);
            qs.push_back(q);
            last_vid = vid;
        }
    }
    seg_end.push_back(wins.size());
    auto& s = hvk_host::Session::instance();
    size_t w0 = 0;
    for (size_t g = 0; g < qs.size(); ++g) {
        const size_t w1 = seg_end[g], n = w1 - w0;
        std::vector<const HV_16*> wp;
        for (size_t w = w0; w < w1; ++w) wp.push_back(&wins[w]);
        const hvk_host::DeviceBundles qd = s.upload_counts({&qs[g]});
        const hvk_host::DeviceBundles wd = s.upload_counts(wp);
        // The reference keeps the first window whose float score is the
        // maximum.  Candidates come ordered by (c desc, index asc); distinct c
        // may round to the same float, so take the earliest window among
        // those whose float equals the best one.
        int kk = (int)std::min<size_t>(n, 29);
        hvk_host::Session::TopK t = s.search(qd, wd, kk);
        const float best = two_c_minus_one(t.cos[0]);
        auto earliest = [&](int m, int& j) {
            size_t first = (size_t)t.idx[0];
            for (j = 1; j < m && two_c_minus_one(t.cos[j]) == best; ++j)
                first = std::min(first, (size_t)t.idx[j]);
            return first;
        };
        int j = 0;
        size_t first = earliest(kk, j);
        if (j == kk && kk < (int)n) {
            // the float tie runs past the candidate list: rank every window
            kk = (int)n;
            t = s.search(qd, wd, kk);
            first = earliest(kk, j);
        }
        if (best > sim.score) {  // segments in window order: the earliest maximum
            sim.score = best;
            sim.pos = pos[w0 + first];
        }
        w0 = w1;
    }
    return sim;
}
