// Checks hvk::lsort::sort_desc against the host's std::sort with the
// comparator identify uses (a.score > b.score, targets/identify.cpp:183-184):
// both must leave every candidate at the same position, ties included.
// Inputs: random sizes 0..300 with few distinct values (many ties), sorted,
// reversed, constant and organ-pipe arrays, and McIlroy's adversary ("A
// Killer Adversary for Quicksort", 1999), which drives introsort to its
// depth limit so the heapsort branch is compared too.
// Prints "ok <cases> heap=<cases that reached heapsort>" and exits 0, or
// prints the first mismatch and exits 1.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

static long g_heap_calls = 0;
#define HVK_LSORT_PROBE_HEAP() (++g_heap_calls)
#include "../../bacterial-classification_amd/csrc/lsort.h"

struct Cand {
    float score;
    int id;
};

static bool check(const std::vector<float>& s, long& cases) {
    const int n = (int)s.size();
    std::vector<Cand> ref(n);
    for (int i = 0; i < n; ++i) ref[i] = {s[i], i};
    std::sort(ref.begin(), ref.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
    for (int stride : {1, 3}) {  // contiguous, and interleaved as in LDS columns
        std::vector<float> ms((size_t)n * stride + 1, -9.0f);
        std::vector<int> mc((size_t)n * stride + 1, -9);
        for (int i = 0; i < n; ++i) {
            ms[(size_t)i * stride] = s[i];
            mc[(size_t)i * stride] = i;
        }
        hvk::lsort::sort_desc(ms.data(), mc.data(), n, stride);
        ++cases;
        for (int i = 0; i < n; ++i)
            if (ref[i].id != mc[(size_t)i * stride] || ref[i].score != ms[(size_t)i * stride]) {
                std::printf("mismatch n=%d stride=%d at %d: std %d (%g) mine %d (%g)\n", n, stride,
                            i, ref[i].id, ref[i].score, mc[(size_t)i * stride],
                            ms[(size_t)i * stride]);
                return false;
            }
        for (int i = 0; i < n; ++i)  // the gaps are untouched
            for (int g = 1; g < stride; ++g)
                if (ms[(size_t)i * stride + g] != -9.0f || mc[(size_t)i * stride + g] != -9)
                    return false;
    }
    return true;
}

// McIlroy's adversary: values are fixed lazily while std::sort runs, so the
// candidate pivot always ends up extreme.  Sorting ascending on -score is
// the same comparison sequence as sorting descending on score.
static std::vector<float> adversary(int n) {
    std::vector<int> val(n, n);  // n = "gas"
    int nsolid = 0, candidate = 0;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    auto less = [&](int x, int y) {
        if (val[x] == n && val[y] == n) {
            if (x == candidate) val[x] = nsolid++;
            else val[y] = nsolid++;
        }
        if (val[x] == n) candidate = x;
        else if (val[y] == n) candidate = y;
        return val[x] < val[y];
    };
    std::sort(idx.begin(), idx.end(), less);
    std::vector<float> s(n);
    for (int i = 0; i < n; ++i) s[i] = -(float)val[i];
    return s;
}

int main() {
    long cases = 0, heap_cases = 0;
    std::mt19937_64 rng(12345);
    auto run = [&](const std::vector<float>& s) {
        const long before = g_heap_calls;
        if (!check(s, cases)) return false;
        if (g_heap_calls > before) ++heap_cases;
        return true;
    };
    for (int it = 0; it < 200000; ++it) {
        const int n = (int)(rng() % 301);
        const int distinct = 1 + (int)(rng() % 8);
        std::vector<float> s(n);
        for (auto& v : s) v = (float)(rng() % distinct) * 0.25f - 0.5f;
        if (!run(s)) return 1;
    }
    for (int n = 0; n <= 300; ++n) {
        std::vector<float> up(n), down(n), same(n, 0.5f), pipe(n);
        for (int i = 0; i < n; ++i) {
            up[i] = (float)i;
            down[i] = (float)(n - i);
            pipe[i] = (float)std::min(i, n - 1 - i);
        }
        if (!run(up) || !run(down) || !run(same) || !run(pipe)) return 1;
        std::vector<float> adv = adversary(n);
        if (!run(adv)) return 1;
        std::vector<float> adv_ties(adv);  // the same order with pairs of equal values
        for (auto& v : adv_ties) v = (float)(int)(v / 2);
        if (!run(adv_ties)) return 1;
    }
    std::printf("ok %ld heap=%ld\n", cases, heap_cases);
    return heap_cases > 0 ? 0 : 1;
}
