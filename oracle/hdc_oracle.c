/*
 * hdc_oracle.c -- CPU restatement of the reference classifier's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see hdc_oracle.h).  Plain C, built
 * by oracle/Makefile into oracle/liboracle.so.  Citations are paths inside the
 * reference repository.
 */
#include "hdc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* std::mt19937_64, restated from the engine's published definition          */
/* (C++ [rand.eng.mers] / [rand.predef]: w=64 n=312 m=156 r=31 ...).         */
/* The reference draws its base HVs from it: src/HV.cpp:20-35.               */
/* ------------------------------------------------------------------------ */
#define MT_N 312
#define MT_M 156
typedef struct { uint64_t x[MT_N]; int i; } mt64;

static void mt64_seed(mt64* g, uint64_t seed) {
    g->x[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        g->x[i] = 6364136223846793005ULL * (g->x[i - 1] ^ (g->x[i - 1] >> 62)) + (uint64_t)i;
    g->i = MT_N;
}

static uint64_t mt64_next(mt64* g) {
    const uint64_t UPPER = 0xFFFFFFFF80000000ULL, LOWER = 0x7FFFFFFFULL;
    const uint64_t A = 0xB5026F5AA96619E9ULL;
    if (g->i >= MT_N) {
        for (int j = 0; j < MT_N; ++j) {
            uint64_t y = (g->x[j] & UPPER) | (g->x[(j + 1) % MT_N] & LOWER);
            uint64_t v = g->x[(j + MT_M) % MT_N] ^ (y >> 1);
            if (y & 1) v ^= A;
            g->x[j] = v;
        }
        g->i = 0;
    }
    uint64_t z = g->x[g->i++];
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
    z ^= (z << 37) & 0xFFF7EEE000000000ULL;
    z ^= z >> 43;
    return z;
}

void orc_mt19937_64(uint64_t seed, uint64_t* out, size_t n) {
    mt64 g;
    mt64_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) out[i] = mt64_next(&g);
}

/* ------------------------------------------------------------------------ */
/* HV primitives                                                            */
/* ------------------------------------------------------------------------ */
int orc_words(int D) { return (D + 63) / 64; }

static uint64_t tail_mask(int D) {
    int excess = orc_words(D) * 64 - D;
    return excess ? (~0ULL) >> excess : ~0ULL;
}

/* src/HV.cpp:9-18 */
void orc_empty_hv(int D, uint64_t* out) {
    int W = orc_words(D);
    for (int i = 0; i < W; ++i) out[i] = ~0ULL;
    out[W - 1] &= tail_mask(D);
}

/* src/HV.cpp:20-35 (make_hv) and :37-51 (seeds A0 C1 T2 G3, default = N) */
void orc_base_hv(int D, char symbol, uint64_t* out) {
    int seed;
    switch (symbol) {
        case 'A': seed = 0; break;
        case 'C': seed = 1; break;
        case 'T': seed = 2; break;
        case 'G': seed = 3; break;
        default: orc_empty_hv(D, out); return;
    }
    int W = orc_words(D);
    orc_mt19937_64((uint64_t)seed, out, (size_t)W);
    out[W - 1] &= tail_mask(D);
}

/* src/HV.cpp:98-147: output word i takes ring bits [64i+s, 64i+s+64) of r
 * (ring = 64*W bits, word index wraps), XORs l, then masks the tail. */
void orc_mult(int D, const uint64_t* l, const uint64_t* r, uint64_t* out, size_t k_bits) {
    int W = orc_words(D);
    size_t ws = (k_bits >> 6) % (size_t)W;
    unsigned s = (unsigned)(k_bits & 63);
    for (int i = 0; i < W; ++i) {
        size_t a = (ws + (size_t)i) % (size_t)W;
        size_t b = (a + 1) % (size_t)W;
        uint64_t rw = s ? ((r[a] >> s) | (r[b] << (64 - s))) : r[a];
        out[i] = l[i] ^ rw;
    }
    out[W - 1] &= tail_mask(D);
}

/* src/encoder.cpp:10-25 */
static char rc_base(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
        default: return 'N';
    }
}

static int is_acgt(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

/* base HV cache per D (oracle is single-D per call; recompute is cheap) */
typedef struct { int D; uint64_t* hv[5]; } base_set;

static void base_set_init(base_set* b, int D) {
    int W = orc_words(D);
    const char sym[5] = {'A', 'C', 'G', 'T', 'N'};
    b->D = D;
    for (int i = 0; i < 5; ++i) {
        b->hv[i] = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W);
        orc_base_hv(D, sym[i], b->hv[i]);
    }
}
static void base_set_free(base_set* b) {
    for (int i = 0; i < 5; ++i) free(b->hv[i]);
}
static const uint64_t* base_of(const base_set* b, char c) {
    switch (c) {
        case 'A': return b->hv[0];
        case 'C': return b->hv[1];
        case 'G': return b->hv[2];
        case 'T': return b->hv[3];
        default: return b->hv[4];
    }
}

/* src/encoder.cpp:47-74: canonical = (rc < s) ? rc : s (string order);
 * binder = empty_hv(); binder = mult(binder, base[canon[i]], i) for i<k. */
static int kmer_hv_with(const base_set* b, int D, int k, const char* kmer, uint64_t* out) {
    char canon[64], rc[64] = {0};
    if (k <= 0 || k > 64) return 0;
    for (int i = 0; i < k; ++i)
        if (!is_acgt(kmer[i])) return 0;
    for (int i = 0; i < k; ++i) rc[k - 1 - i] = rc_base(kmer[i]);
    const char* c = (memcmp(rc, kmer, (size_t)k) < 0) ? rc : kmer;
    memcpy(canon, c, (size_t)k);
    orc_empty_hv(D, out);
    for (int i = 0; i < k; ++i) orc_mult(D, out, base_of(b, canon[i]), out, (size_t)i);
    return 1;
}

int orc_kmer_hv(int D, int k, const char* kmer, uint64_t* out) {
    if (k <= 0 || k > 32) return 0;
    base_set b;
    base_set_init(&b, D);
    int ok = kmer_hv_with(&b, D, k, kmer, out);
    base_set_free(&b);
    return ok;
}

/* src/HV.cpp:156-216: acc[d] += -n + 2 * #{b : bit d of hv_b set} (int16) */
void orc_superpose_batch(int D, int16_t* acc, const uint64_t* const* hvs, size_t n) {
    for (int d = 0; d < D; ++d) {
        int cnt = 0;
        for (size_t b = 0; b < n; ++b) cnt += (int)((hvs[b][d >> 6] >> (d & 63)) & 1);
        acc[d] = (int16_t)(acc[d] + (int16_t)(-(int)n));
        acc[d] = (int16_t)(acc[d] + (int16_t)(cnt * 2));
    }
}

/* src/HV.h:21-106: acc[d] += 2*new_bit - 2*old_bit for each pair */
void orc_replace_batch(int D, int16_t* acc, const uint64_t* const* olds,
                       const uint64_t* const* news, size_t n) {
    for (size_t b = 0; b < n; ++b)
        for (int d = 0; d < D; ++d) {
            int nb = (int)((news[b][d >> 6] >> (d & 63)) & 1);
            int ob = (int)((olds[b][d >> 6] >> (d & 63)) & 1);
            acc[d] = (int16_t)(acc[d] + (int16_t)(2 * nb - 2 * ob));
        }
}

/* src/encoder.cpp:106-132: batches of 16 valid k-mers into superpose_batch;
 * k-mers with a char outside ACGT are not in the table and are skipped. */
static int compute_with(const base_set* b, int D, int k, const char* seq, size_t len,
                        int16_t* counts16, int32_t* counts32) {
    enum { BATCH = 16 };
    int W = orc_words(D);
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W * BATCH);
    const uint64_t* ptrs[BATCH];
    int nb = 0, total = 0;
    if (counts16) memset(counts16, 0, sizeof(int16_t) * (size_t)D);
    if (counts32) memset(counts32, 0, sizeof(int32_t) * (size_t)D);
    for (size_t j = 0; j + (size_t)k <= len; ++j) {
        if (!kmer_hv_with(b, D, k, seq + j, buf + (size_t)nb * W)) continue;
        ptrs[nb] = buf + (size_t)nb * W;
        ++nb;
        ++total;
        if (nb == BATCH) {
            if (counts16) orc_superpose_batch(D, counts16, ptrs, (size_t)nb);
            if (counts32)
                for (int d = 0; d < D; ++d)
                    for (int q = 0; q < nb; ++q)
                        counts32[d] += ((ptrs[q][d >> 6] >> (d & 63)) & 1) ? 1 : -1;
            nb = 0;
        }
    }
    if (nb) {
        if (counts16) orc_superpose_batch(D, counts16, ptrs, (size_t)nb);
        if (counts32)
            for (int d = 0; d < D; ++d)
                for (int q = 0; q < nb; ++q)
                    counts32[d] += ((ptrs[q][d >> 6] >> (d & 63)) & 1) ? 1 : -1;
    }
    free(buf);
    return total;
}

int orc_compute(int D, int k, const char* seq, size_t len, int16_t* counts) {
    base_set b;
    base_set_init(&b, D);
    int n = compute_with(&b, D, k, seq, len, counts, NULL);
    base_set_free(&b);
    return n;
}

int orc_compute32(int D, int k, const char* seq, size_t len, int32_t* counts) {
    base_set b;
    base_set_init(&b, D);
    int n = compute_with(&b, D, k, seq, len, NULL, counts);
    base_set_free(&b);
    return n;
}

/* ------------------------------------------------------------------------ */
/* cosine                                                                   */
/* ------------------------------------------------------------------------ */
int64_t orc_dot16(int D, const int16_t* l, const int16_t* r) {
    int64_t s = 0;
    for (int d = 0; d < D; ++d) s += (int64_t)l[d] * (int64_t)r[d];
    return s;
}

/* src/HV.cpp:219-242 with NORM_COSINE (install.sh:3), else :243-252 */
float orc_cosine(int D, const int16_t* l, const int16_t* r, int norm_mode) {
    int64_t dot = 0, nl = 0, nr = 0;
    for (int d = 0; d < D; ++d) {
        int64_t a = l[d], b = r[d];
        dot += a * b;
        nl += a * a;
        nr += b * b;
    }
    if (!norm_mode) return (float)dot;
    if (nl == 0 || nr == 0) return -1.0f;
    return (float)(2.0 * ((double)dot / (sqrt((double)nl) * sqrt((double)nr)) - 0.5));
}

/* c in FP64 as the reference forms it before its 2c-1 map (src/HV.cpp:239);
 * a zero norm gives 0.0, i.e. 2c-1 = -1 as the reference returns. */
double orc_cos64_from(int64_t dot, int64_t nl, int64_t nr) {
    if (nl == 0 || nr == 0) return 0.0;
    return (double)dot / (sqrt((double)nl) * sqrt((double)nr));
}

/* ------------------------------------------------------------------------ */
/* brute-force top-k: order (c desc, index asc)                              */
/* ------------------------------------------------------------------------ */
static void topk_insert(int k, int32_t* idx, double* val, int* cnt, int32_t i, double v) {
    int n = *cnt;
    if (n == k) {
        if (v < val[n - 1] || (v == val[n - 1] && i > idx[n - 1])) return;
        n = k - 1;
    }
    int p = n;
    while (p > 0 && (val[p - 1] < v || (val[p - 1] == v && idx[p - 1] > i))) {
        val[p] = val[p - 1];
        idx[p] = idx[p - 1];
        --p;
    }
    val[p] = v;
    idx[p] = i;
    *cnt = n + 1 > k ? k : n + 1;
}

typedef struct {
    int D, k, nq, ndb, t, nt;
    const int16_t *q, *db;
    int64_t* dbn;
    int32_t* out_idx;
    double* out_cos;
} topk_job;

static void* topk_worker(void* arg) {
    topk_job* j = (topk_job*)arg;
    for (int qi = j->t; qi < j->nq; qi += j->nt) {
        const int16_t* q = j->q + (size_t)qi * j->D;
        int64_t qn = orc_dot16(j->D, q, q);
        int cnt = 0;
        int32_t* oi = j->out_idx + (size_t)qi * j->k;
        double* ov = j->out_cos + (size_t)qi * j->k;
        for (int r = 0; r < j->ndb; ++r) {
            const int16_t* row = j->db + (size_t)r * j->D;
            double c = orc_cos64_from(orc_dot16(j->D, q, row), qn, j->dbn[r]);
            topk_insert(j->k, oi, ov, &cnt, r, c);
        }
        for (int p = cnt; p < j->k; ++p) { oi[p] = -1; ov[p] = -INFINITY; }
    }
    return NULL;
}

void orc_topk16(int D, const int16_t* q, int nq, const int16_t* db, int ndb,
                int k, int32_t* out_idx, double* out_cos, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    int64_t* dbn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ndb > 0 ? ndb : 1));
    for (int r = 0; r < ndb; ++r) dbn[r] = orc_dot16(D, db + (size_t)r * D, db + (size_t)r * D);
    pthread_t th[256];
    topk_job jobs[256];
    if (n_threads > 256) n_threads = 256;
    for (int t = 0; t < n_threads; ++t) {
        jobs[t] = (topk_job){D, k, nq, ndb, t, n_threads, q, db, dbn, out_idx, out_cos};
        pthread_create(&th[t], NULL, topk_worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    free(dbn);
}

/* ------------------------------------------------------------------------ */
/* HVCache::precompute / ::similarity (src/encoder.cpp:76-104,134-215)       */
/* One fresh cache shared by a precompute(ref) and a similarity(q, ref).    */
/* ------------------------------------------------------------------------ */
typedef struct { char* key; size_t klen; int16_t* hv; } centry;
typedef struct { centry* e; int n, cap, D, k; base_set b; } ocache;

static void oc_init(ocache* c, int D, int k) {
    c->e = NULL; c->n = c->cap = 0; c->D = D; c->k = k;
    base_set_init(&c->b, D);
}
static void oc_free(ocache* c) {
    for (int i = 0; i < c->n; ++i) { free(c->e[i].key); free(c->e[i].hv); }
    free(c->e);
    base_set_free(&c->b);
}
static int16_t* oc_find(ocache* c, const char* s, size_t len) {
    for (int i = 0; i < c->n; ++i)
        if (c->e[i].klen == len && memcmp(c->e[i].key, s, len) == 0) return c->e[i].hv;
    return NULL;
}
static int16_t* oc_put(ocache* c, const char* s, size_t len, const int16_t* hv) {
    if (c->n == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 64;
        c->e = (centry*)realloc(c->e, sizeof(centry) * (size_t)c->cap);
    }
    centry* e = &c->e[c->n++];
    e->key = (char*)malloc(len ? len : 1);
    memcpy(e->key, s, len);
    e->klen = len;
    e->hv = (int16_t*)malloc(sizeof(int16_t) * (size_t)c->D);
    memcpy(e->hv, hv, sizeof(int16_t) * (size_t)c->D);
    return e->hv;
}
/* HVCache::get (src/encoder.cpp:76-104) without eviction */
static const int16_t* oc_get(ocache* c, const char* s, size_t len) {
    int16_t* h = oc_find(c, s, len);
    if (h) return h;
    int16_t* tmp = (int16_t*)malloc(sizeof(int16_t) * (size_t)c->D);
    compute_with(&c->b, c->D, c->k, s, len, tmp, NULL);
    h = oc_put(c, s, len, tmp);
    free(tmp);
    return h;
}

/* src/encoder.cpp:134-196 */
static void oc_precompute(ocache* c, int window, int step, const char* seq, size_t len) {
    int D = c->D, k = c->k, W = orc_words(D);
    if (len < (size_t)window) return;
    size_t nk = len - (size_t)k + 1;
    uint64_t* hvs = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W * nk);
    uint64_t* ones = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W);
    const uint64_t** ptr = (const uint64_t**)malloc(sizeof(uint64_t*) * nk);
    orc_empty_hv(D, ones);
    for (size_t i = 0; i < nk; ++i)
        ptr[i] = kmer_hv_with(&c->b, D, k, seq + i, hvs + i * (size_t)W) ? hvs + i * (size_t)W : ones;
    int16_t* win = (int16_t*)malloc(sizeof(int16_t) * (size_t)D);
    memcpy(win, oc_get(c, seq, len < (size_t)window ? len : (size_t)window), sizeof(int16_t) * (size_t)D);
    const uint64_t* oldb[64];
    const uint64_t* newb[64];
    for (size_t index = (size_t)step; index + (size_t)window < len; index += (size_t)step) {
        int16_t* hit = oc_find(c, seq + index, (size_t)window);
        if (hit) { memcpy(win, hit, sizeof(int16_t) * (size_t)D); continue; }
        for (int x = 0; x < step; ++x) {
            oldb[x] = ptr[index + (size_t)x - (size_t)step];
            newb[x] = ptr[index + (size_t)window + (size_t)x - (size_t)(step + k)];
        }
        orc_replace_batch(D, win, oldb, newb, (size_t)step);
        oc_put(c, seq + index, (size_t)window, win);
    }
    free(win); free(ptr); free(ones); free(hvs);
}

/* src/encoder.cpp:198-215 */
static float oc_similarity(ocache* c, int window, int step, const char* q, size_t ql,
                           const char* r, size_t rl, size_t* pos) {
    int D = c->D;
    size_t best_pos = 0;
    float best = -1.0f;
    if ((size_t)window >= rl) {
        best = orc_cosine(D, oc_get(c, q, ql), oc_get(c, r, rl), 1);
    } else {
        for (size_t x = 0; x + (size_t)window < rl; x += (size_t)step) {
            const int16_t* ref = oc_get(c, r + x, (size_t)window);
            float s = orc_cosine(D, oc_get(c, q, ql), ref, 1);
            if (s > best) { best = s; best_pos = x; }
        }
    }
    if (pos) *pos = best_pos;
    return best;
}

float orc_precompute_similarity(int D, int k, int window, int step,
                                const char* query, size_t qlen,
                                const char* ref, size_t rlen, size_t* pos_out) {
    ocache c;
    oc_init(&c, D, k);
    oc_precompute(&c, window, step, ref, rlen);
    float s = oc_similarity(&c, window, step, query, qlen, ref, rlen, pos_out);
    oc_free(&c);
    return s;
}

int orc_precompute_windows(int D, int k, int window, int step,
                           const char* ref, size_t rlen, int16_t* out, int max_windows) {
    ocache c;
    oc_init(&c, D, k);
    oc_precompute(&c, window, step, ref, rlen);
    int n = 0;
    if (rlen >= (size_t)window) {
        for (size_t x = 0; (x == 0 || x + (size_t)window < rlen) && n < max_windows; x += (size_t)step) {
            const int16_t* h = oc_find(&c, ref + x, (size_t)window);
            if (!h) break;
            memcpy(out + (size_t)n * D, h, sizeof(int16_t) * (size_t)D);
            ++n;
        }
    }
    oc_free(&c);
    return n;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: the reference's bit-packed path, one read at a time:       */
/* compute() (k-mer binding + ctz bundling, src/encoder.cpp:106-132,        */
/* src/HV.cpp:156-216) then cosine() against every DB row (src/HV.cpp:      */
/* 219-242, three int64 accumulators) keeping the top-k.                    */
/* ------------------------------------------------------------------------ */
static void superpose_ctz(int D, int16_t* acc, const uint64_t* const* hvs, size_t n) {
    /* the reference's ctz form of superpose_batch (src/HV.cpp:156-216) */
    int W = orc_words(D), full = D / 64;
    for (int w = 0; w < full; ++w) {
        int16_t* blk = acc + w * 64;
        uint8_t cnt[64] = {0};
        for (int i = 0; i < 64; ++i) blk[i] = (int16_t)(blk[i] - (int16_t)n);
        for (size_t b = 0; b < n; ++b) {
            uint64_t x = hvs[b][w];
            while (x) { cnt[__builtin_ctzll(x)]++; x &= x - 1; }
        }
        for (int i = 0; i < 64; ++i) blk[i] = (int16_t)(blk[i] + cnt[i] * 2);
    }
    int tail = D - full * 64;
    if (tail) {
        int16_t* blk = acc + full * 64;
        uint8_t cnt[64] = {0};
        for (int i = 0; i < tail; ++i) blk[i] = (int16_t)(blk[i] - (int16_t)n);
        for (size_t b = 0; b < n; ++b) {
            uint64_t x = hvs[b][W - 1];
            while (x) { unsigned bit = (unsigned)__builtin_ctzll(x); if (bit >= (unsigned)tail) break; cnt[bit]++; x &= x - 1; }
        }
        for (int i = 0; i < tail; ++i) blk[i] = (int16_t)(blk[i] + cnt[i] * 2);
    }
}

typedef struct {
    int D, k, n_reads, ndb, topk, t, nt;
    const char* seqs; const int64_t* off;
    const int16_t* db; const int64_t* dbn;
    int32_t* out_idx; double* out_cos;
    const base_set* b;
} cls_job;

static void* cls_worker(void* arg) {
    cls_job* j = (cls_job*)arg;
    int D = j->D, W = orc_words(D);
    int16_t* acc = (int16_t*)malloc(sizeof(int16_t) * (size_t)D);
    uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)W * 16);
    const uint64_t* ptrs[16];
    for (int r = j->t; r < j->n_reads; r += j->nt) {
        const char* s = j->seqs + j->off[r];
        size_t len = (size_t)(j->off[r + 1] - j->off[r]);
        memset(acc, 0, sizeof(int16_t) * (size_t)D);
        int nb = 0;
        for (size_t p = 0; p + (size_t)j->k <= len; ++p) {
            if (!kmer_hv_with(j->b, D, j->k, s + p, buf + (size_t)nb * W)) continue;
            ptrs[nb] = buf + (size_t)nb * W;
            if (++nb == 16) { superpose_ctz(D, acc, ptrs, 16); nb = 0; }
        }
        if (nb) superpose_ctz(D, acc, ptrs, (size_t)nb);
        int cnt = 0;
        int32_t* oi = j->out_idx + (size_t)r * j->topk;
        double* ov = j->out_cos + (size_t)r * j->topk;
        for (int q = 0; q < j->ndb; ++q) {
            const int16_t* row = j->db + (size_t)q * D;
            int64_t dot = 0, nl = 0, nr = 0;
            for (int d = 0; d < D; ++d) {
                int64_t a = acc[d], bb = row[d];
                dot += a * bb; nl += a * a; nr += bb * bb;
            }
            double c = orc_cos64_from(dot, nl, nr);
            topk_insert(j->topk, oi, ov, &cnt, q, c);
        }
        for (int p = cnt; p < j->topk; ++p) { oi[p] = -1; ov[p] = -INFINITY; }
    }
    free(buf);
    free(acc);
    return NULL;
}

double orc_classify_reads(int D, int k, const char* seqs, const int64_t* offsets,
                          int n_reads, const int16_t* db, const int64_t* db_sqnorm,
                          int ndb, int topk, int32_t* out_idx, double* out_cos,
                          int n_threads) {
    (void)db_sqnorm;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    base_set b;
    base_set_init(&b, D);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pthread_t th[256];
    cls_job jobs[256];
    for (int t = 0; t < n_threads; ++t) {
        jobs[t] = (cls_job){D, k, n_reads, ndb, topk, t, n_threads, seqs, offsets, db,
                            db_sqnorm, out_idx, out_cos, &b};
        pthread_create(&th[t], NULL, cls_worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    base_set_free(&b);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
