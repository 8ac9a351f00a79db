/*
 * hdc_amd.h -- C ABI of the MI355X (gfx950) read-classification hot path:
 * k-mer hyperdimensional encoding of DNA sequences and exact cosine top-k
 * search against a reference prototype database.
 *
 * Every entry point takes plain pointers and sizes.  Pointers named d_* are
 * device memory owned by the caller; outputs are caller-allocated, as with
 * the reference's launchers.  A stream argument is a hipStream_t (NULL = the
 * default stream); work is enqueued on it and the call returns without
 * synchronising unless stated.  No C++ exception crosses this boundary: every
 * call returns an hvk status code (HVK_OK = 0).
 *
 * Reference interfaces replaced (paths in the reference repository):
 *   src/config.h:5-17              compile-time D, k        -> hvk_ctx_create
 *   src/HV.cpp:20-51               base HVs, get_base_hv    -> hvk_base_hv
 *   src/encoder.cpp:47-74          4^k k-mer table (ctor)   -> device base table
 *   src/encode.cu:169-185          pack_sequence            -> hvk_pack
 *   src/encode.cu:136 (launch_hdc_encode), src/encoder.cpp:106-132
 *                                  (HVCache::compute)        -> hvk_encode
 *   src/cosine.cu:441 (launch_cosine_topk), :172 (launch_cosine_argmax)
 *                                  -> hvk_search_plain / hvk_search_mfma
 *   src/cosine.cu:35-50            compute_inv_norms_kernel -> norms from hvk_encode
 */
#ifndef HDC_AMD_H
#define HDC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define HVK_OK               0
#define HVK_ERR_ARG          1  /* bad argument (null, size, unsupported D/k) */
#define HVK_ERR_HIP          2  /* a HIP runtime call failed */
#define HVK_ERR_RANGE        3  /* values outside the exact-arithmetic range */
#define HVK_ERR_NOMEM        4  /* device allocation failed */
#define HVK_ERR_NODEVICE     5  /* no usable gfx950 device */

typedef struct hvk_ctx hvk_ctx;
typedef void* hvk_stream; /* hipStream_t */

/* Human-readable text for a status code (static storage). */
const char* hvk_status_string(int status);
/* Text of the last HIP error seen by this thread's calls (static storage). */
const char* hvk_last_error(void);

/* ---- context ----------------------------------------------------------
 * Binds dimension D (bits per hypervector) and k-mer size k on the current
 * device: draws the base HVs exactly as the reference (std::mt19937_64 with
 * seed A:0 C:1 T:2 G:3, W = ceil(D/64) words, tail masked) and uploads the
 * table of their ring rotations by 0..k-1 bits used by the encoder.
 * Supported: 64 <= D <= 32768, 1 <= k <= 16.
 * Replaces src/config.h:5-17 and HVCache::HVCache (src/encoder.cpp:47-74). */
int hvk_ctx_create(int dim, int kmer, hvk_ctx** out);
int hvk_ctx_destroy(hvk_ctx* ctx);
/* D, k, words per binary HV, and the row stride (elements) hvk_encode uses
 * for limb/count outputs when the caller passes ld = 0. */
int hvk_ctx_info(const hvk_ctx* ctx, int* dim, int* kmer, int* words64, int* row_ld);
/* get_base_hv (src/HV.cpp:43-51): W words for 'A','C','G','T'; any other
 * character gives the all-ones HV (src/HV.cpp:9-18). */
int hvk_base_hv(const hvk_ctx* ctx, char symbol, uint64_t* host_out);

/* Kernel timing: when enabled, hvk_encode and the search calls record HIP
 * events on their stream around their main kernel; hvk_ctx_last_timing
 * waits for those events and returns the durations in milliseconds (a value
 * is -1 when that kernel has not run since timing was enabled). */
int hvk_ctx_set_timing(hvk_ctx* ctx, int enable);
int hvk_ctx_last_timing(hvk_ctx* ctx, float* encode_ms, float* search_ms);

/* Calibration (no counterpart in the reference; for measurement only): runs
 * int8 MFMAs (v_mfma_i32_32x32x32_i8) from registers on every CU of the
 * current device, waves_per_simd (1 or 2) waves per SIMD, iters rounds of 64
 * MFMAs per wave, on operands drawn ~ round(N(0, q_sd)) and round(N(0, r_sd))
 * clamped to +-127 (sd 0: all-zero operands), with no memory traffic.
 * *ops (nullable) receives the integer operations launched (2 per MAC).
 * d_sink: one device int32 that keeps the products live.  Time the launch
 * with events on `stream` to get the card's int8 MFMA rate for that operand
 * distribution under its power and clock limits. */
int hvk_calib_mfma_i8(int waves_per_simd, int iters, int q_sd, int r_sd, int32_t* d_sink,
                      double* ops, hvk_stream stream);

/* Tuning of the encode and search launches (no counterpart in the
 * reference; used by the tile autotuner and the tests).  Results do not
 * depend on it.
 * HVK_TUNE_SPLITS: number of reference slices per query tile of the MFMA
 * search, 0 = automatic (the default), else 1..64.
 * HVK_TUNE_SELECT_KIB: workspace budget (KiB) of the selection path for
 * k > 29, which bounds how many queries one chunk holds; 0 = 1 GiB.
 * HVK_TUNE_SHAPE: tile of hvk_search_mfma_bounded for k <= 5, as queries x
 * references per workgroup: 0 = 256 x 256 (default), 1 = 128 x 256,
 * 2 = 256 x 128, 3 = 128 x 512, 4 = 512 x 128, 5 = 256 x 256 on a 4 x 2
 * wave grid.
 * HVK_TUNE_ENCODER: 0 = automatic (the default: one wave per sequence where
 * it applies), 1 = the workgroup-per-sequence encoders only.
 * HVK_TUNE_DESCENT: 0 = hvk_descend and hvk_batch_descend score a read whose
 * every count fits int8 from an int8 copy of its row, against candidates
 * whose every |p| <= 32639 split into two int8 digits (the default); other
 * pairs use the int16 rows.  1 = int16 rows only.  The scores are the same
 * exact dots either way.
 * HVK_TUNE_GROUP_SLICES: reference slices that share one XCD's group of
 * workgroups in the MFMA search (the group holds CUs-per-XCD / slices query
 * tiles, so that the XCD's L2 serves each query tile to that many slices and
 * each reference tile to that many query tiles); 0 = automatic (8, or all
 * slices when fewer), else 1..64. */
#define HVK_TUNE_SPLITS 1
#define HVK_TUNE_SELECT_KIB 2
#define HVK_TUNE_SHAPE 3
#define HVK_TUNE_ENCODER 4
#define HVK_TUNE_DESCENT 5
#define HVK_TUNE_GROUP_SLICES 6
int hvk_ctx_set_tuning(hvk_ctx* ctx, int param, int value);

/* ---- packing (host) -----------------------------------------------------
 * Sequences are concatenated: sequence i is seqs[offsets[i] .. offsets[i+1]).
 * packed: 2-bit codes A0 C1 G2 T3, base p at bits 2*(p%16) of word p/16.
 * invalid: bit p%32 of word p/32 set when base p is not one of "ACGT"
 * (lower case and IUPAC codes included), so k-mers over it are skipped as
 * HVCache::compute skips them (src/encoder.cpp:115-117).
 * hvk_pack_words gives the uint32 count for each array (includes padding).
 * Replaces pack_sequence (src/encode.cu:169-185). */
int64_t hvk_pack_words(int64_t total_bases);
int hvk_pack(const char* seqs, const int64_t* offsets, int n_seqs,
             uint32_t* packed, uint32_t* invalid);

/* ---- encode -----------------------------------------------------------
 * One bundle per sequence with HVCache::compute semantics
 * (src/encoder.cpp:106-132): count[d] = sum over the sequence's all-ACGT
 * k-mers of (2*bit_d(HV(kmer)) - 1), HV(kmer) bound from the canonical
 * k-mer (src/encoder.cpp:47-74).  Bit-exact with the reference.
 *
 * d_packed/d_invalid: hvk_pack output on the device.  d_offsets: n_seqs+1
 * int64 base offsets on the device.  max_len: longest sequence (bases).
 * Outputs (each nullable; row stride ld elements, ld = 0 -> ctx row_ld,
 * padding columns [D, ld) are written as zero):
 *   d_counts  int16 [n][ld]   counts (int16 wrap as the reference's HV_16)
 *   d_lo,d_hi int8  [n][ld]   limbs: count = 256*hi + lo, lo in [-128,127]
 *   d_sqnorm  int64 [n]       exact sum of count^2
 *   d_maxabs  int32 [n]       max |count|
 *   d_nkmers  int32 [n]       number of k-mers bundled
 * Replaces launch_hdc_encode (src/encode.cu:136) / encode (src/GPU.cuh:6). */
int hvk_encode(hvk_ctx* ctx, const uint32_t* d_packed, const uint32_t* d_invalid,
               const int64_t* d_offsets, int n_seqs, int max_len,
               int16_t* d_counts, int8_t* d_lo, int8_t* d_hi, int64_t ld,
               int64_t* d_sqnorm, int32_t* d_maxabs, int32_t* d_nkmers,
               hvk_stream stream);

/* Outputs of the encoder (every pointer nullable; rows of ld elements,
 * ld = 0 -> ctx row_ld; padding columns [D, ld) are written as zero):
 *   counts  int16 [n][ld]  counts (int16 wrap as the reference's HV_16)
 *   lo, hi  int8  [n][ld]  limbs: count = 256*hi + lo, lo in [-128,127]
 *   clamp8  int8  [n][ld]  count clamped to [-127, 127]
 *   sqnorm  int64 [n]      exact sum of count^2
 *   maxabs  int32 [n]      max |count|
 *   nkmers  int32 [n]      number of k-mers bundled
 *   resid1  int32 [n]      sum of |count - clamp8| (L1 norm of the residual) */
typedef struct {
    int16_t* counts;
    int8_t* lo;
    int8_t* hi;
    int8_t* clamp8;
    int64_t ld;
    int64_t* sqnorm;
    int32_t* maxabs;
    int32_t* nkmers;
    int32_t* resid1;
} hvk_encode_out;

/* hvk_encode with every output of hvk_encode_out. */
int hvk_encode_ex(hvk_ctx* ctx, const uint32_t* d_packed, const uint32_t* d_invalid,
                  const int64_t* d_offsets, int n_seqs, int max_len,
                  const hvk_encode_out* out, hvk_stream stream);

/* Windowed encoding: one bundle per window of window_len bases starting at
 * base offset d_starts[i] of the packed sequence buffer, each with
 * HVCache::compute semantics of the window's substring.  Replaces the
 * reference's windowed GPU encoder hdc_encode / launch_hdc_encode
 * (src/encode.cu:69-165) and the exact windows of HVCache::precompute
 * (src/encoder.cpp:134-196). */
int hvk_encode_windows(hvk_ctx* ctx, const uint32_t* d_packed, const uint32_t* d_invalid,
                       const int64_t* d_starts, int n_windows, int window_len,
                       const hvk_encode_out* out, hvk_stream stream);

/* Search operands from int16 count rows the caller already holds (HV_16
 * vectors of the C++ API): fills the lo/hi/clamp8 rows and the per-row
 * sqnorm/maxabs/resid1 of out (out->counts must be NULL; nullable outputs are
 * skipped).  d_counts rows have stride out->ld (0 -> ctx row_ld); columns
 * [D, ld) are treated as zero.  Replaces compute_inv_norms_kernel
 * (src/cosine.cu:35-50). */
int hvk_prepare_counts(hvk_ctx* ctx, const int16_t* d_counts, int n, const hvk_encode_out* out,
                       hvk_stream stream);

/* ---- search -----------------------------------------------------------
 * Top-k cosine search of n_q queries against n_r references.  The order is
 * (c desc, reference index asc) with c = dot / (sqrt(|q|^2) sqrt(|r|^2))
 * evaluated in FP64 from the exact integer dot, and c = 0 when a norm is
 * zero (the reference's cosine returns 2c-1 = -1 there, src/HV.cpp:236-237).
 * Candidates are selected in FP32; every candidate whose FP32 margin is not
 * conclusive is re-scored exactly, so the indices equal the FP64 brute-force
 * oracle's.  Outputs per query (row-major [n_q][k]):
 *   d_idx int32 reference index (-1 past n_r), d_score float c (the scale of
 *   the reference's cosine_search, src/cosine.cu:270), d_cos64 (nullable)
 *   double c.
 * d_n_rescored (nullable, int32 on the device): count of queries that took
 * the exact re-scan.  Any k >= 1, as the reference's launch_cosine_topk
 * (src/cosine.cu:450, which clamps k to num_db): for k <= 29 the search
 * kernels keep per-query candidate lists; above, every pair's exact cosine
 * goes to a key matrix (chunks of queries, at most 1 GiB of workspace) and
 * a radix select and sort per query give the top k (nothing re-scored).
 * Rows past n_r hold -1 / -inf.
 *
 * Plain form (VALU int16 dot): rows of int16 counts with stride ld (0 ->
 * ctx row_ld).  Requires max|q| * max|r| * 64 < 2^31; the caller passes the
 * max |count| of each side (from hvk_encode's d_maxabs) and gets
 * HVK_ERR_RANGE otherwise.
 * Replaces launch_cosine_topk (src/cosine.cu:441). */
int hvk_search_plain(hvk_ctx* ctx,
                     const int16_t* d_q, const int64_t* d_q_sqnorm, int n_q, int q_maxabs,
                     const int16_t* d_r, const int64_t* d_r_sqnorm, int n_r, int r_maxabs,
                     int64_t ld, int k,
                     int32_t* d_idx, float* d_score, double* d_cos64,
                     int32_t* d_n_rescored, hvk_stream stream);

/* MFMA form: int8 limb GEMM on the matrix cores with int32 accumulation,
 * limbs recombined exactly in int64 (value = 256*hi + lo).  d_q_hi / d_r_hi
 * may be NULL when every |value| on that side is <= 127 (one limb); bipolar
 * (+1/-1) HVs pass as one limb.  Limb rows have stride ld (0 -> ctx row_ld).
 * Replaces launch_cosine_topk (src/cosine.cu:441). */
int hvk_search_mfma(hvk_ctx* ctx,
                    const int8_t* d_q_lo, const int8_t* d_q_hi, const int64_t* d_q_sqnorm, int n_q,
                    const int8_t* d_r_lo, const int8_t* d_r_hi, const int64_t* d_r_sqnorm, int n_r,
                    int64_t ld, int k,
                    int32_t* d_idx, float* d_score, double* d_cos64,
                    int32_t* d_n_rescored, hvk_stream stream);

/* One side of a bounded search: rows as hvk_encode_ex writes them. */
typedef struct {
    const int8_t* clamp8;   /* [n][ld] counts clamped to [-127, 127] */
    const int8_t* lo;       /* [n][ld] exact limbs, value = 256*hi + lo */
    const int8_t* hi;       /* NULL when every |count| <= 127 */
    const int64_t* sqnorm;  /* [n] */
    const int32_t* maxabs;  /* [n] */
    const int32_t* resid1;  /* [n] sum |count - clamp8| */
    int n;
} hvk_bundles;

/* MFMA form with one limb per side and an exact bound (same results as
 * hvk_search_mfma).  The matrix cores multiply clamped counts (one int8 limb
 * each side, half the work of the two-limb form); for each pair the FP32
 * score is raised by an upper bound on what the clamps removed,
 *   |dot - dot_clamped| <= min(max|q|,127) resid1(r) + resid1(q) max|r|,
 * so the candidates kept are those whose upper bound ranks high.  Their exact
 * dots are recomputed from the limbs and re-scored in FP64, with the same
 * margin check as hvk_search_mfma; reads that fail it take a second bounded
 * pass with 32-entry lists (d_n_rescored counts them), and what still fails
 * takes the exact FP64 re-scan. */
int hvk_search_mfma_bounded(hvk_ctx* ctx, const hvk_bundles* q, const hvk_bundles* r, int64_t ld,
                            int k, int32_t* d_idx, float* d_score, double* d_cos64,
                            int32_t* d_n_rescored, hvk_stream stream);

/* Allocation-free variant of hvk_search_mfma_bounded (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * hipGraph capture, serving loops and harnesses that forbid hipMalloc.  Same
 * arguments and results, k <= 29; the search buffers live in a fixed device
 * array of HVK_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
 * ends, so no call allocates or frees.  Sizes whose buffers do not fit return
 * HVK_ERR_RANGE (about 4096 queries x 262,144 references at k <= 29 fit).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams or threads) race and must be serialised. */
#define HVK_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int hvk_search_mfma_bounded_noalloc(hvk_ctx* ctx, const hvk_bundles* q, const hvk_bundles* r,
                                    int64_t ld, int k, int32_t* d_idx, float* d_score,
                                    double* d_cos64, int32_t* d_n_rescored, hvk_stream stream);

/* ---- hierarchical descent ----------------------------------------------
 * identify's rank-by-rank beam (targets/identify.cpp:89-202) over a taxonomy
 * whose categories each carry one integer prototype bundle (a leaf's is its
 * sequence's bundle, a parent's the sum of its children's), scored with the
 * reference's cosine (src/HV.cpp:219-242) from exact int64 dots:
 *   score = float(2 * (dot / (sqrt(|q|^2) * sqrt(|p|^2)) - 0.5)), -1 on a zero norm.
 * Each level expands every beam entry into its candidate list, scores all
 * candidates, sorts them by score with std::sort's order for equal scores
 * (comparator a.score > b.score, targets/identify.cpp:183-184), and then
 * either exits (no candidate, or the best below `threshold`: the read's
 * result is the best candidate's path, reported with its exit level) or
 * keeps the first min(beam, #candidates >= threshold) entries.
 *
 * Categories are 0..n_cat-1.  Prototype p = 65536*hi + lo with lo in
 * [-32768, 32767], so |p| < 2^31.  The candidate lists are host arrays
 * (validated and uploaded by the call): the candidates after category c are
 * child_idx[child_off[c] .. child_off[c+1]) in order, and a category without
 * children must list itself (Sequence::get_children returns {path.back()},
 * targets/identify.cpp:30-35); the level-0 candidates are roots[0..n_roots)
 * in the order Hierarchy::get_order("p__") returns them
 * (src/hierarchy.cpp:110-116). */
typedef struct {
    int n_cat;
    int64_t ld;                  /* row stride (elements) of d_lo / d_hi */
    const int16_t* d_lo;         /* [n_cat][ld] */
    const int16_t* d_hi;         /* [n_cat][ld] */
    const int64_t* d_sqnorm;     /* [n_cat] exact sum of p^2 */
    int32_t max_abs_limb;        /* max |lo|, |hi| over all rows (<= 32768) */
    const int32_t* child_off;    /* host, [n_cat + 1] */
    const int32_t* child_idx;    /* host, [child_off[n_cat]] */
    const int32_t* roots;        /* host, [n_roots] */
    int n_roots;
    int64_t max_abs_proto;       /* max |p| over the roots' rows, 0 = unknown; when below
                                    2^30 and n_roots <= 64 the level-0 scores run as an
                                    int8-limb GEMM on the matrix cores */
    const int32_t* d_row_maxabs; /* [n_cat] max |p| of each row, nullable: a row below
                                    2^15 has a zero hi limb, which is then skipped (without
                                    it, rows whose d_sqnorm is below 2^30 are) */
} hvk_tree;

/* Outputs per read (device, caller-allocated): path[levels] category ids and
 * path_score[levels] scores of the result path, its length, and the level at
 * which the read exited early (-1 when it ran all levels).  q_maxabs bounds
 * |q| of every read (the encoder's maxabs).  beam: 1..16 (identify uses 5);
 * levels: 1..32 (identify walks 6).  HVK_ERR_RANGE when the int32 partial
 * sums could overflow (|q| * max_abs_limb * 2 must stay below 2^31). */
int hvk_descend(hvk_ctx* ctx, const hvk_tree* tree, const int16_t* d_q_counts,
                const int64_t* d_q_sqnorm, int n_q, int64_t q_ld, int q_maxabs, int beam,
                float threshold, int levels, int32_t* d_path, float* d_path_score,
                int32_t* d_path_len, int32_t* d_exit_level, hvk_stream stream);

/* Prototype construction.  hvk_tree_place writes the int32 rows
 * proto[cat[i]] = counts[i] (host cat[n]); hvk_tree_sum sets, for each
 * parent p listed in h_parents (host, in order), proto[p] = sum of the rows
 * of its candidate list child_idx[child_off[p] .. child_off[p+1]) (call it
 * rank by rank from the leaves up); a sum outside int32 sets *d_overflow
 * (nullable) to 1.  hvk_tree_limbs splits int32 rows into the lo/hi limbs
 * and exact squared norms hvk_descend reads, each row's max |limb|, and
 * (d_maxabs_proto, nullable) each row's max |p|, the tree's d_row_maxabs. */
int hvk_tree_place(hvk_ctx* ctx, const int16_t* d_counts, int64_t counts_ld, int n,
                   const int32_t* h_cat, int32_t* d_proto, int64_t ld, int n_cat,
                   hvk_stream stream);
int hvk_tree_sum(hvk_ctx* ctx, int32_t* d_proto, int64_t ld, int n_cat, const int32_t* h_child_off,
                 const int32_t* h_child_idx, const int32_t* h_parents, int n_parents,
                 int32_t* d_overflow, hvk_stream stream);
int hvk_tree_limbs(hvk_ctx* ctx, const int32_t* d_proto, int n, int64_t ld, int16_t* d_lo,
                   int16_t* d_hi, int64_t* d_sqnorm, int32_t* d_maxabs_limb,
                   int32_t* d_maxabs_proto, hvk_stream stream);

/* ---- prototype database file ---------------------------------------------
 * A versioned binary file of encoded reference prototypes, so references are
 * encoded once (no counterpart in the reference, which re-encodes its FASTA
 * files on every run through HVCache, src/hierarchy.cpp:56-81 and
 * src/encoder.cpp:76-104).  Layout, little-endian:
 *   128-byte header: magic "HVKPDB01", version (1), header size, D, k, row
 *   stride ld, row count n, array flags, names size, FNV-1a 64 of the first
 *   120 header bytes;
 *   then, each 4096-byte aligned, the arrays the flags name, in this order:
 *   counts int16 [n][ld], lo int8 [n][ld], hi int8 [n][ld], clamp8 int8
 *   [n][ld], sqnorm int64 [n], maxabs int32 [n], resid1 int32 [n],
 *   nkmers int32 [n], then the names ('\n'-separated, optional).
 * hvk_db_open maps the file read-only and validates magic, version, header
 * checksum and size; the arrays are then readable in place (host pointers
 * into the mapping) or copied to device buffers with hvk_db_upload. */
#define HVK_DB_COUNTS 1u
#define HVK_DB_LO     2u
#define HVK_DB_HI     4u
#define HVK_DB_CLAMP8 8u
#define HVK_DB_SUMS   16u  /* sqnorm, maxabs, resid1, nkmers */
typedef struct hvk_db hvk_db;
/* Writes a database from host arrays (`arrays` in host memory, ld = its row
 * stride; arrays not named by `flags` may be NULL).  names: optional blob of
 * names_len bytes. */
int hvk_db_write(const char* path, int dim, int kmer, int64_t ld, int n, uint32_t flags,
                 const hvk_encode_out* arrays, const char* names, int64_t names_len);
int hvk_db_open(const char* path, hvk_db** out);
int hvk_db_close(hvk_db* db);
int hvk_db_info(const hvk_db* db, int* dim, int* kmer, int64_t* ld, int* n, uint32_t* flags,
                int64_t* names_len);
/* Host view of one array (HVK_DB_COUNTS ... HVK_DB_CLAMP8, or 32 + i for the
 * row sums sqnorm, maxabs, resid1, nkmers; 64 for the names); NULL when absent. */
const void* hvk_db_array(const hvk_db* db, int which, int64_t* bytes);
/* Copies the arrays present in both the file and `dst` (device buffers, row
 * stride dst->ld or the file's ld when 0) to the device; HVK_ERR_ARG when
 * ctx's D, k differ from the file's.  Synchronous. */
int hvk_db_upload(const hvk_db* db, const hvk_ctx* ctx, const hvk_encode_out* dst);

/* ---- batching form -------------------------------------------------------
 * No counterpart in the reference, whose reads go one at a time through
 * HVCache::get and cosine (targets/identify.cpp:150-157).  Classifies a large
 * batch of reads of uneven length against a database, top-k, with results
 * identical to hvk_encode_ex followed by hvk_search_mfma_bounded:
 *  - reads are ordered by length, longest first, and cut into chunks of at
 *    most `chunk` reads, so every encoder launch holds reads of similar
 *    length (its workgroups finish together);
 *  - each call copies the packed reads into batch-owned staging buffers; a
 *    chunk runs encode -> MFMA search -> finalize on batch-owned buffers,
 *    then a scatter (outside the graph) writes its rows' results to the
 *    reads' own output rows.  The first call that meets a chunk (same
 *    position, size, k, settings, counter width of its longest read and
 *    database) runs it directly and captures the same launches into a
 *    hipGraph; later calls replay it, whatever their read and output
 *    buffers.  At most 64 graphs are kept (least recently used dropped);
 *  - hvk_batch_autotune times candidate settings (search form: bounded one
 *    limb or two limbs; the bounded form's tile shape, reads x references
 *    per workgroup from 128 to 512 each, see HVK_TUNE_SHAPE; reference
 *    slices per read tile; chunk size) on a sample of the batch and keeps the
 *    fastest.  Results never depend on the settings.
 * A batch owns a private context for D, k, whose workspaces back the
 * captured graphs; `db` may come from any context of the same D, k and is
 * read, not copied: it must stay valid and unchanged while graphs that use
 * it are replayed.  One batch must not be used from two threads at once. */
typedef struct hvk_batch hvk_batch;
#define HVK_BATCH_FORM_BOUNDED  0
#define HVK_BATCH_FORM_TWO_LIMB 1
int hvk_batch_create(int dim, int kmer, int chunk, hvk_batch** out);
int hvk_batch_destroy(hvk_batch* b);
/* form: HVK_BATCH_FORM_*; shape: the bounded form's tile for top-5 (the
 * values of HVK_TUNE_SHAPE, 0..5; other k use shape 0); splits: 0 =
 * automatic, else 1..64; chunk >= 256.  Changing a setting drops the
 * captured graphs. */
int hvk_batch_config(hvk_batch* b, int form, int shape, int splits, int chunk);
int hvk_batch_get_config(const hvk_batch* b, int* form, int* shape, int* splits, int* chunk,
                         int* n_graphs);
/* d_packed / d_invalid: hvk_pack layout of all reads (device); h_offsets:
 * n+1 base offsets (host).  Outputs [n][k] (device); d_cos64 and
 * d_n_rescored (the number of reads re-scored by the exact passes, summed
 * over chunks) are nullable.  Any k >= 1 (see hvk_search_plain).  Enqueued
 * on `stream`. */
int hvk_batch_classify(hvk_batch* b, const uint32_t* d_packed, const uint32_t* d_invalid,
                       const int64_t* h_offsets, int n, const hvk_bundles* db, int k,
                       int32_t* d_idx, float* d_score, double* d_cos64, int32_t* d_n_rescored,
                       hvk_stream stream);
/* Times candidate settings on the first min(sample, n) reads of the batch
 * (one direct run, then two timed graph replays, the faster counted, into
 * the caller's output buffers) and applies the fastest; *ms_per_mread
 * (nullable) gets its time per million reads.  Coordinate search: every
 * tile shape (top-5) at automatic slices and half the sample per chunk, then
 * 4, 8 and 16 slices on the best shape, then chunks of 1/4 and all of the
 * sample (at least 256 reads), and the two-limb form.  table (nullable,
 * *table_rows rows of 5 floats on entry): one row per candidate timed --
 * form, shape, splits, chunk, ms per million reads; *table_rows is set to
 * the rows written.  Synchronises `stream`. */
int hvk_batch_autotune(hvk_batch* b, const uint32_t* d_packed, const uint32_t* d_invalid,
                       const int64_t* h_offsets, int n, const hvk_bundles* db, int k,
                       int32_t* d_idx, float* d_score, int sample, float* ms_per_mread,
                       float* table, int* table_rows, hvk_stream stream);
/* hvk_descend in the batching form: the reads (hvk_pack layout, device;
 * h_offsets host) are ordered by length and cut into chunks; each chunk is
 * encoded (int16 counts) and descended by launches captured into a hipGraph
 * once and replayed, and its results go to the reads' own rows (outputs as
 * hvk_descend's: d_path / d_path_score [n][levels], d_path_len,
 * d_exit_level [n]).  The partial-sum bound uses the longest read's k-mer
 * count.  Same results as hvk_encode_ex + hvk_descend.  The tree's device
 * arrays are read at every replay (the candidate lists are uploaded by each
 * call); they must stay valid while the call runs.  Enqueued on `stream`. */
int hvk_batch_descend(hvk_batch* b, const hvk_tree* tree, const uint32_t* d_packed,
                      const uint32_t* d_invalid, const int64_t* h_offsets, int n, int beam,
                      float threshold, int levels, int32_t* d_path, float* d_path_score,
                      int32_t* d_path_len, int32_t* d_exit_level, hvk_stream stream);

/* ---- synthetic workloads (host, deterministic) -------------------------
 * Ancestors of uniform ACGT, descendants by substitution + indels, and reads
 * sampled from descendants with substitution errors; every draw comes from
 * std::mt19937_64 streams derived from `seed`, so a spec always yields the
 * same data.  Two calls: *_size then *_fill. */
typedef struct {
    uint64_t seed;
    int32_t n_ancestors;       /* e.g. 2000 */
    int32_t desc_per_ancestor; /* e.g. 50 -> n_refs = n_ancestors * this */
    int32_t anc_len_min, anc_len_max;
    double div_min, div_max;   /* per-reference substitution rate range */
    double indel_rate;         /* per-base indel probability (half ins, half del) */
    int32_t n_reads;
    int32_t read_len_min, read_len_max;
    int32_t read_len_log_uniform; /* 0: uniform lengths, 1: log-uniform */
    double err_min, err_max;   /* per-read substitution rate range */
} hvk_gen_spec;

/* total bases of the references / reads the spec produces */
int hvk_gen_size(const hvk_gen_spec* spec, int64_t* ref_bases, int64_t* read_bases);
/* ref_offsets: n_refs+1; read_offsets: n_reads+1; read_truth: source
 * reference of each read (nullable). */
int hvk_gen_fill(const hvk_gen_spec* spec, char* refs, int64_t* ref_offsets,
                 char* reads, int64_t* read_offsets, int32_t* read_truth);

#ifdef __cplusplus
}
#endif
#endif /* HDC_AMD_H */
