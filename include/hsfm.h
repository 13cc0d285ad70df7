/*
 * hsfm.h -- C ABI of the MI355X (gfx950) batched query path of haystackfm.
 *
 * This is the drop-in boundary.  The reference crate (sriram98v/haystackfm)
 * exposes its GPU query path as Rust methods; each entry point below replaces
 * one of them and is what that crate's FFI layer binds (see INTEGRATION.md for
 * the extern "C" block and the Rust wrappers).  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns 0 on success, a negative HSFM_E* code on error;
 *     hsfm_last_error() then returns a thread-local message.  This maps to the
 *     reference's FmIndexError::GpuError(String) (src/error.rs:26-29) and, for
 *     malformed index bytes, FmIndexError::DeserializeError (src/error.rs:18-20).
 *   - Query bytes are the reference's encoded IUPAC codes ($=0, A..T=1..4, N=5,
 *     R..V=6..15; src/alphabet.rs:18-52); any other byte matches nothing.
 *   - Queries are CSR: qflat (concatenated codes) + qoff[nq+1] (u64 offsets).
 *   - Host-buffer entry points copy inputs to the device and results back.
 *     *_device entry points take device pointers and leave results in
 *     library-owned device memory that stays valid until the next call on the
 *     same index.  Calls on one index are serialized internally.
 */
#ifndef HSFM_H
#define HSFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSFM_OK 0
#define HSFM_EINVAL (-1)    /* bad argument or inconsistent index view      */
#define HSFM_EDESER (-2)    /* malformed FmIndex byte stream                */
#define HSFM_EDEVICE (-3)   /* HIP runtime / device error                   */
#define HSFM_ENOMEM (-4)    /* device or host allocation failed             */
#define HSFM_EINDEX (-5)    /* index invariant violated at query time       */

typedef struct hsfm_device hsfm_device; /* GpuContext (src/gpu/mod.rs:19-66)       */
typedef struct hsfm_index hsfm_index;   /* device-resident FmIndex / BidirFmIndex */
typedef struct hsfm_lsi hsfm_lsi;       /* learned seed index of an hsfm_index     */

/*
 * Borrowed view of one FmIndex, field for field as serialized by
 * FmIndex::to_bytes (src/fm_index/serialize.rs:55-67; OccTable
 * src/occ/mod.rs:44-78; SampledSuffixArray src/suffix_array/mod.rs:54-63;
 * LookupTable src/fm_index/lookup.rs:19-27).  compat_mask[b] is the bitmask of
 * reference symbols that alphabet_fns.compatible_fn(b) returns
 * (src/alphabet.rs:203-212), so custom Alphabet impls keep their semantics.
 */
typedef struct {
    uint32_t c_array[16];
    uint8_t occ_num_lanes, occ_num_planes;
    uint8_t occ_symbol_to_lane[16], occ_lane_to_symbol[16];
    const uint32_t *occ_superblock_checkpoints; uint64_t occ_superblock_len;
    const uint8_t *occ_block_data; uint64_t occ_block_data_len;
    uint64_t occ_block_stride;
    uint32_t occ_encoding; /* 0 = Bitplane, 1 = OneHot (occ/mod.rs:93-97) */
    uint32_t occ_text_len;
    const uint8_t *sa_word_data; uint64_t sa_word_data_len;
    const uint32_t *sa_vals; uint64_t sa_vals_len;
    uint32_t sa_sample_rate, sa_text_len;
    uint32_t text_len, num_sequences;
    const uint32_t *seq_boundaries; uint64_t seq_boundaries_len;
    uint32_t has_lookup, lookup_depth;
    const uint8_t *lookup_core_symbols; uint64_t lookup_core_len;
    const uint32_t *lookup_intervals; uint64_t lookup_intervals_len; /* (lo,hi) pairs */
    uint8_t alphabet_tag;
    uint16_t compat_mask[256];
} hsfm_index_view;

/* Device context.  Replaces GpuContext::new + context_cache::get_or_init
 * (src/gpu/mod.rs:26-66, src/gpu/context_cache.rs:11-19). */
int hsfm_device_open(int ordinal, hsfm_device **out);
int hsfm_device_close(hsfm_device *dev);

/* Zero-copy parse of an FmIndex::to_bytes stream into a view whose pointers
 * point into `bytes`.  Replaces the bincode half of FmIndex::from_bytes
 * (src/fm_index/serialize.rs:32-52); unknown alphabet tags are rejected as
 * there.  compat_mask is filled for the built-in tags 0 (IupacDna) and
 * 1 (ExactDna) (src/alphabet.rs:240-254). */
int hsfm_index_view_from_bytes(const uint8_t *bytes, uint64_t len, hsfm_index_view *out);

/* Upload one FmIndex (rev == NULL) or a BidirFmIndex (fwd + rev) and convert
 * it to the device layout.  Replaces the per-call uploads of
 * locate_batch_gpu (src/gpu/locate.rs:69-111), FindIndexBuffers::new
 * (src/gpu/mem_find.rs:56-102) and ResolveIndexBuffers::new
 * (src/gpu/mem_resolve.rs:44-84): here the index is uploaded once and stays
 * resident.  The views are only read during the call. */
int hsfm_index_upload(hsfm_device *dev, const hsfm_index_view *fwd, const hsfm_index_view *rev,
                      hsfm_index **out);
int hsfm_index_free(hsfm_index *idx);

/* FmIndex::build_cpu + FmIndex::to_bytes (src/fm_index/mod.rs:117-195,
 * serialize.rs:13-26), or with bidir != 0 BidirFmIndex::build_cpu + to_bytes
 * ([u32 LE fwd_len][fwd][rev], bidir_index.rs:53-80, 183-192), computed on the
 * device: the same bytes, compacted Occ lanes included (the reference's GPU
 * build leaves them out, src/occ/gpu.rs:94-105).  text: the concatenated codes
 * with a 0 sentinel ending each sequence (concatenate_sequences,
 * src/alphabet.rs:375-393), at most 2^31 - 2 codes; bounds[nseq]: exclusive
 * ends; headers: '\0'-separated names or NULL ("seq_{i}" for empty ones);
 * encoding 0 = Bitplane, 1 = OneHot; tag 0 = IupacDna, 1 = ExactDna.  *out is
 * freed with hsfm_free_result. */
int hsfm_build_index_device(hsfm_device *dev, const uint8_t *text, uint64_t n, const uint32_t *bounds,
                            uint32_t nseq, const char *headers, uint32_t sample_rate, uint32_t lookup_depth,
                            uint32_t encoding, uint8_t tag, int bidir, uint8_t **out, uint64_t *out_len);

/* ---- batched locate ----------------------------------------------------- */
typedef struct {
    uint32_t num_queries;
    uint64_t *offsets;  /* num_queries + 1 */
    uint32_t *hits;     /* 2 * offsets[num_queries]: (seq_idx, pos_in_seq) */
} hsfm_locate_result;

/* locate_batch_gpu (src/gpu/locate.rs:52-56) / FmIndex::locate_gpu
 * (src/fm_index/query.rs:367-384): per query, the (seq_idx, pos) of every
 * occurrence, with FmIndex::locate's semantics (query.rs:22-39, 219-277)
 * including lookup-table seeding; hits of a query are in ascending SA-row
 * order.  Free the result with hsfm_free_result. */
int hsfm_locate_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                      hsfm_locate_result **out);
/* FmIndex::locate_positions / locate_positions_many (src/fm_index/query.rs:46-208)
 * over a batch: per query, the absolute text offsets (into the concatenated
 * text, sentinels included) of every occurrence, in ascending SA-row order
 * (locate_positions' order; locate_positions_many leaves it open).  Free the
 * result with hsfm_free_result. */
typedef struct {
    uint32_t num_queries;
    uint64_t *offsets;   /* num_queries + 1 */
    uint32_t *positions; /* offsets[num_queries] absolute text positions */
} hsfm_positions_result;
int hsfm_locate_positions_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                hsfm_positions_result **out);
/* FmIndex::count over a batch (src/fm_index/query.rs:9-14); counts[nq]. */
int hsfm_count_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                     uint32_t *counts);

typedef struct {
    uint32_t num_queries;
    uint64_t total_hits;
    const uint64_t *d_offsets; /* device, num_queries + 1 */
    const uint32_t *d_hits;    /* device, 2 * total_hits */
} hsfm_locate_device_result;
/* Device-resident variant of hsfm_locate_batch: d_qflat / d_qoff are device
 * pointers; results stay in device memory owned by idx. */
int hsfm_locate_device(hsfm_index *idx, const uint8_t *d_qflat, const uint64_t *d_qoff,
                       uint32_t nq, hsfm_locate_device_result *out);

/* ---- SMEM / MEM ---------------------------------------------------------- */
#define HSFM_MODE_SMEM 0 /* find_smems_gpu (src/fm_index/bidir_index.rs:225-245) */
#define HSFM_MODE_MEM 1  /* find_mems_gpu  (src/fm_index/bidir_index.rs:256-276) */

/* One MemHit (src/gpu/mem_hit.rs:6-19): one record per (MEM, SA interval),
 * match_count = interval size, positions = the first min(size, max_hits) rows
 * of the interval in SA order, truncated = size > max_hits
 * (bidir_index.rs:452-468, 307-359).  Records of a query are sorted by
 * (query_start, query_end, interval start). */
typedef struct {
    uint32_t query_start, query_end, match_count, truncated;
    uint64_t pos_offset; /* into positions (pairs) */
    uint32_t pos_len, _reserved;
} hsfm_mem_hit;

typedef struct {
    uint32_t num_queries;
    uint64_t *hit_offsets;  /* num_queries + 1, into hits */
    hsfm_mem_hit *hits;
    uint32_t *positions;    /* 2 * total: (ref_id, offset_within_ref) */
    uint64_t num_hits, num_positions;
} hsfm_mem_result;

/* BidirFmIndex::find_smems_gpu / find_mems_gpu (bidir_index.rs:225-276) with
 * the CPU definitions find_smems / find_mems (src/fm_index/smem.rs:71-124).
 * ref_boundaries[i] = exclusive end of reference i in the concatenated text;
 * n_refs == 0 skips position resolution (positions empty), as the reference
 * does.  idx must have been uploaded with a reverse index. */
int hsfm_find_mems_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                         uint32_t min_len, int mode, const uint32_t *ref_boundaries,
                         uint32_t n_refs, uint32_t max_hits, hsfm_mem_result **out);

typedef struct {
    uint32_t num_queries;
    uint64_t num_hits, num_positions;
    const uint64_t *d_hit_offsets; /* device, num_queries + 1 */
    const hsfm_mem_hit *d_hits;    /* device */
    const uint32_t *d_positions;   /* device, 2 * num_positions */
} hsfm_mem_device_result;
/* Device-resident variant: d_qflat, d_qoff and d_ref_boundaries are device
 * pointers; results stay in device memory owned by idx. */
int hsfm_find_mems_device(hsfm_index *idx, const uint8_t *d_qflat, const uint64_t *d_qoff,
                          uint32_t nq, uint32_t min_len, int mode,
                          const uint32_t *d_ref_boundaries, uint32_t n_refs, uint32_t max_hits,
                          hsfm_mem_device_result *out);

/* ---- batched MEM pipeline (uneven reads) ----------------------------------
 * The same results as hsfm_find_mems_device, produced batch by batch: fn
 * receives the device-resident result of reads [r0, r0 + batch_reads), in
 * order (hit offsets and position offsets relative to the batch, valid only
 * during the call).  A window of window_batches consecutive batches runs
 * through one captured HIP graph (find, second-level passes, positions,
 * resolve, reference map): its persistent kernel takes the work items of
 * every batch of the window from one device queue, so that a batch's last
 * items do not leave the GPU idle.  Each buffer set keeps its instantiated
 * graph and reuses it on later calls while the captured buffers, index views
 * and configuration are unchanged.  Reference analogue: the query chunking of
 * resolve_mem_hits_gpu (bidir_index.rs:418-439) under output_budget_u32
 * (gpu/mod.rs:68-82).  Needs min_len > 12 (seed-table restarts). */
typedef struct {
    uint32_t batch_reads;  /* reads per batch (65536)                          */
    int autotune;          /* time the candidate configurations on window 0    */
    int main_kernel;       /* -1 = tuned/cached/default; 0..2 thread per read
                              with 8/16/32-interval LDS sets, 3 wave per read  */
    int resolve_fanout;    /* 0 = tuned/cached/default; else 1, 2 or 4 walks
                              per thread                                       */
    int use_graph;         /* capture each batch into one HIP graph            */
    int batch_sets;        /* windows in flight, each with its own buffers,
                              graph and stream (0 = 3, at most 8)            */
    uint32_t window_batches; /* batches per graph (0: enough work items for
                              about sixteen per lane of the persistent grid;
                              at most 64)                                     */
} hsfm_batch_config;
typedef struct {
    uint32_t batches, fallback_batches; /* batches, those re-run host-driven   */
    int main_kernel, resolve_fanout;    /* configuration used                  */
    uint64_t records, positions;
    double total_ms;                    /* HIP events around the batch loop    */
    double tune_total_ms;
    double tune_ms[6];  /* window-0 time of main kernels 0..3 (fan-out 1), then
                           the best main kernel with fan-out 2 and 4           */
    uint32_t window_batches, windows; /* batches per graph, graphs launched   */
    uint32_t graphs_captured;         /* graphs captured by this call (0: every
                                         set reused its graph from an earlier call) */
} hsfm_batch_stats;
/* The index stays locked while fn runs: an entry point called from fn on the
 * same index (hsfm_index_free included) returns HSFM_EINVAL without doing
 * anything.  Other indexes and the device helpers below may be used. */
typedef void (*hsfm_batch_fn)(void *user, uint32_t first_read, uint32_t num_reads,
                              const hsfm_mem_device_result *result);
int hsfm_find_mems_batched(hsfm_index *idx, const uint8_t *d_qflat, const uint64_t *d_qoff,
                           uint32_t nq, uint32_t min_len, int mode,
                           const uint32_t *d_ref_boundaries, uint32_t n_refs, uint32_t max_hits,
                           const hsfm_batch_config *cfg, hsfm_batch_fn fn, void *user,
                           hsfm_batch_stats *stats);
/* Device memory helpers for callers without a device allocator of their own
 * (inputs of the *_device entry points, results inside a batch callback). */
int hsfm_device_alloc(hsfm_device *dev, uint64_t bytes, void **out);
int hsfm_device_free(hsfm_device *dev, void *d_ptr);
int hsfm_host_to_device(hsfm_device *dev, void *d_dst, const void *src, uint64_t bytes);
int hsfm_device_to_host(hsfm_device *dev, void *dst, const void *d_src, uint64_t bytes);

/* ---- both strands ------------------------------------------------------- */
/* Reverse complement of every query, for searching the other strand: query i
 * of the output is query i read backwards with each code complemented (A<->T,
 * C<->G, R<->Y, K<->M, B<->V, D<->H; N, S, W, $ and bytes >= 16 unchanged); the
 * offsets are the input's.  No counterpart in the reference, which searches
 * the given strand only (src/alphabet.rs:149-168 fixes the base sets the
 * complement maps); both-strand callers run the forward and the output batch
 * through the same entry points.  d_* are device pointers, out_flat holds
 * qoff[nq] bytes. */
int hsfm_revcomp_device(hsfm_device *dev, const uint8_t *d_qflat, const uint64_t *d_qoff, uint32_t nq,
                        uint8_t *d_out);
/* Host-array variant of hsfm_revcomp_device. */
int hsfm_revcomp_batch(hsfm_device *dev, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq, uint8_t *out);

/* ---- misc ---------------------------------------------------------------- */
void hsfm_free_result(void *result);
const char *hsfm_last_error(void);

/* Statistics of the last call on idx (fallback path usage, walk-guard hits). */
typedef struct {
    uint64_t fallback_reads;   /* reads that left the register fast path   */
    uint64_t capacity_retries; /* relaunches with larger per-wave set space */
    uint64_t guard_hits;       /* LF walks stopped by the sample-rate guard */
    uint64_t records;          /* MEM interval records emitted               */
    double main_kernel_ms;     /* search / MEM-find kernel, first pass (HIP events) */
    double resolve_kernel_ms;  /* LF-walk resolve kernel (HIP events)             */
    uint64_t occ_probes;       /* 64 B records the main kernel's algorithm fetched */
    uint64_t diag[4];          /* MEM thread path: deferrals by seed union, by
                                  extension, by record slots; plain restarts.
                                  locate: [0] = reads of the workgroup pass       */
    double fallback_kernel_ms; /* kernels that re-ran the reads the main kernel
                                  deferred (larger sets)                          */
    uint64_t sa_full_rows;     /* rows of the device-built full suffix array that
                                  resolve reads (0: LF walks to the samples)      */
} hsfm_stats;
int hsfm_last_stats(hsfm_index *idx, hsfm_stats *out);
/* Seeding of the restarts of the one-shot MEM calls on idx (hsfm_find_mems_batch
 * / _device); every mode gives the same results.  HSFM_SEED_TABLES: packed
 * 12-/14-/15-mer tables built from idx (default).  HSFM_SEED_PLAIN: plain
 * backward search from the full interval, the reference's find_smems restart
 * (src/fm_index/smem.rs:154-222).  HSFM_SEED_LEARNED: the learned seed index
 * lsi (built on idx, k <= 25) answers the restart at j when q[j-k..j) holds at
 * most three ambiguity codes: the set of q[j-k..j) (every interval of its
 * IUPAC-compatible text windows), or, when that set is empty, the longest
 * occurring suffix of q[j-k..j), which is L(j); plain search otherwise (j < k,
 * more ambiguity codes, or an empty set with L(j) <= j - min_len);
 * HSFM_SEED_LEARNED_TABLES: the same with the tables as fallback.  The
 * learned modes keep lsi attached until it is freed (which restores
 * HSFM_SEED_TABLES).  The batched pipeline always seeds from the tables.
 * No counterpart in the reference. */
#define HSFM_SEED_TABLES 0
#define HSFM_SEED_PLAIN 1
#define HSFM_SEED_LEARNED 2
#define HSFM_SEED_LEARNED_TABLES 3
int hsfm_mem_seed_mode(hsfm_index *idx, int mode, hsfm_lsi *lsi);
/* Reads whose work items the main kernel of the last one-shot MEM call on idx
 * (hsfm_find_mems_batch / _device) deferred to the workgroup passes, ascending
 * and distinct: min(cap, *count) ids are written.  No counterpart in the
 * reference; for parity checks and profiling of the deferred path.
 * HSFM_EINVAL if the last call on idx was not a one-shot MEM call. */
int hsfm_last_deferred(hsfm_index *idx, uint32_t *ids, uint64_t cap, uint64_t *count);

/* ---- learned seed index ---------------------------------------------------
 * New component (no reference entry point): replaces the first k exact
 * backward-search steps of an exact-ACGT k-mer.  Its answer is that of plain
 * search, extend_left over the k-mer (bidir.rs:112-130), as LookupTable::build
 * computes it for depth-k prefixes (lookup.rs:37-91).  Keys of these calls
 * pack the k-mer two bits per base, first base most significant (A=0, C=1,
 * G=2, T=3).  For k <= 25 the same index also answers MEM restarts
 * (hsfm_mem_seed_mode): it holds, besides the k-mers, the base strings of every
 * text window with ambiguity codes and the windows cut short by contig starts,
 * sorted by their reversed codes, so a lookup gives a window's whole IUPAC set
 * or its longest occurring suffix. */
typedef struct {
    uint32_t k;           /* k-mer length, 1..31                            */
    uint32_t leaves;      /* stage-2 leaves, a power of two (4096): leaf =
                             the key's leading log2(leaves) bits            */
    uint32_t hidden;      /* stage-1 MLP hidden units (<= 64)               */
    uint32_t train_steps; /* Adam steps of stage 1; 0 keeps f(x) = x        */
    uint32_t sample;      /* training sample size (keys, evenly strided)    */
    float lr;             /* Adam learning rate                             */
    uint32_t lookup;      /* leaf placement of the lookup kernel: 0 = chosen
                             by the build (LDS when the leaves fit and at
                             most two fence levels), 1 = global, 2 = LDS
                             (needs leaves <= 4096)                          */
} hsfm_lsi_config;
typedef struct {
    uint64_t num_keys;    /* distinct present exact-ACGT k-mers             */
    uint64_t entries;     /* sorted entries: the k-mers, plus (k <= 25) the
                             base strings of text windows holding ambiguity
                             codes and the windows cut short by contig starts */
    uint64_t variant_entries; /* entries of ambiguity-code windows           */
    uint64_t short_entries;   /* entries of windows shorter than k           */
    uint32_t max_err;     /* largest per-leaf error bound (slots)           */
    uint32_t fence_levels; /* fence levels of the bounded search (0..7)     */
    double mean_window;   /* mean search window 2*err+2 over the entries    */
    double mse_initial, mse_final; /* stage-1 training loss before / after  */
    double enum_ms, train_ms, fit_ms;
    double extra_ms;      /* ambiguity and short entries, sort, run bounds  */
    uint32_t mem_seeding; /* 1: usable for MEM restarts (hsfm_mem_seed_mode) */
    uint32_t pad2;
} hsfm_lsi_info;
typedef struct {
    double kernel_ms;     /* HIP events around the kernel                  */
    uint64_t probes;      /* learned: 64 B units the bounded search loaded  *
                           * (fence and key rounds); plain search: 0        */
} hsfm_seed_stats;
/* Build from idx itself: every present k-mer, its interval, the restart
 * entries (k <= 25, when the device text is available), both stages. */
int hsfm_lsi_build(hsfm_index *idx, const hsfm_lsi_config *cfg, hsfm_lsi **out);
int hsfm_lsi_get_info(const hsfm_lsi *lsi, hsfm_lsi_info *out);
/* Per key (device arrays): d_out[4i..4i+4) = {fwd_lo, fwd_hi, rev_lo, found}. */
int hsfm_lsi_seed_device(hsfm_lsi *lsi, const uint64_t *d_keys, uint64_t n, uint32_t *d_out,
                         hsfm_seed_stats *stats);
/* hsfm_lsi_seed_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  No
 * counterpart in the reference.  Its only scratch, the load counter, lives in
 * a 256-byte static device array each GPU holds until the process ends; calls
 * that overlap on one GPU race on it.  n <= 2^32 - 1 keys per call; *probes
 * (optional, host) receives the call's loads in 64-byte units, which
 * synchronizes the index stream.  Same results as hsfm_lsi_seed_device. */
int hsfm_lsi_seed_device_noalloc(hsfm_lsi *lsi, const uint64_t *d_keys, uint64_t n, uint32_t *d_out,
                                 uint64_t *probes);
/* The same answer by plain exact search (k left extensions per key). */
int hsfm_kmer_search_device(hsfm_index *idx, uint32_t k, const uint64_t *d_keys, uint64_t n,
                            uint32_t *d_out, hsfm_seed_stats *stats);
/* Host-array variants of the two calls above. */
int hsfm_lsi_seed(hsfm_lsi *lsi, const uint64_t *keys, uint64_t n, uint32_t *out, hsfm_seed_stats *stats);
int hsfm_kmer_search(hsfm_index *idx, uint32_t k, const uint64_t *keys, uint64_t n, uint32_t *out);
/* Stage-1 loss gradient of the training kernel on host data (for checks):
 * params = w1[H], b1[H], w2[H], b2; grad_out has the same layout. */
int hsfm_lsi_grad(hsfm_device *dev, const float *params, uint32_t hidden, const float *x,
                  const float *y, uint32_t n, float *grad_out, float *loss_out);
int hsfm_lsi_free(hsfm_lsi *lsi);

#ifdef __cplusplus
}
#endif
#endif
