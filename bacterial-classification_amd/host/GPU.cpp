// GPU.cpp -- the src/GPU.cuh entry points over the device session.
#include "GPU.h"

#include <stdexcept>
#include <string_view>

#include "session.h"

std::vector<std::vector<HV_16>> encode(const std::vector<int16_t>& kmer_map,
                                       const std::vector<std::string>& sequences) {
    (void)kmer_map;
    auto& s = hvk_host::Session::instance();
    std::vector<std::string_view> views(sequences.begin(), sequences.end());
    const hvk_host::DevicePacked p = s.pack(views);
    std::vector<int64_t> starts;
    std::vector<int> per_seq(sequences.size(), 0);
    for (size_t i = 0; i < sequences.size(); ++i) {
        const int64_t len = (int64_t)sequences[i].size();
        if (len < SEQUENCE_LENGTH) continue;
        const int64_t nw = (len - SEQUENCE_LENGTH) / SIM_STEP + 1;
        for (int64_t w = 0; w < nw; ++w) starts.push_back(p.host_offsets[i] + w * SIM_STEP);
        per_seq[i] = (int)nw;
    }
    std::vector<std::vector<HV_16>> out(sequences.size());
    if (starts.empty()) return out;
    const hvk_host::DeviceBundles b = s.encode_windows(p, starts, SEQUENCE_LENGTH);
    std::vector<HV_16> rows = s.download_counts(b);
    size_t o = 0;
    for (size_t i = 0; i < sequences.size(); ++i)
        for (int w = 0; w < per_seq[i]; ++w) out[i].push_back(std::move(rows[o++]));
    return out;
}

std::unordered_map<HV_16*, std::vector<std::pair<float, HV_16*>>> cosine_search(
    const std::vector<HV_16*>& inputs, const std::vector<HV_16*>& db, int k) {
    std::unordered_map<HV_16*, std::vector<std::pair<float, HV_16*>>> result;
    if (k > (int)db.size()) k = (int)db.size();
    if (inputs.empty() || k <= 0) {
        for (HV_16* p : inputs) result[p];
        return result;
    }
    auto& s = hvk_host::Session::instance();
    const hvk_host::DeviceBundles q = s.upload_counts({inputs.begin(), inputs.end()});
    const hvk_host::DeviceBundles r = s.upload_counts({db.begin(), db.end()});
    const hvk_host::Session::TopK t = s.search(q, r, k);
    result.reserve(inputs.size());
    for (size_t i = 0; i < inputs.size(); ++i) {
        std::vector<std::pair<float, HV_16*>> m((size_t)k);
        for (int j = 0; j < k; ++j) {
            const int idx = t.idx[i * k + j];
            m[j] = {(float)t.cos[i * k + j], db[idx]};
        }
        result[inputs[i]] = std::move(m);
    }
    return result;
}
