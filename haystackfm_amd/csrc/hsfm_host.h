// hsfm_host.h -- host-side objects shared by the translation units of libhsfm.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "hsfm_internal.h"

struct hsfm_index;

struct hsfm_device {
    int ordinal = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
};

namespace hsfm {

// Thread-local error message (hsfm_last_error).
int set_error(int code, const char *fmt, ...);
#define HSFM_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::hsfm::set_error(HSFM_EDEVICE, "%s failed: %s (%s:%d)", #call,              \
                                     hipGetErrorString(e_), __FILE__, __LINE__);                \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct DirBuffers {
    DevBuf occ, sb, samp, sa_vals, lookup;
    DevBuf sa_full;           // every row's SA value (see ensure_sa_full)
    DevBuf text;              // the text as lanes, from sa_full
    DevBuf isa;               // row of every text position, from sa_full
    DevBuf text4;             // text as base-set nibbles (DevDir::text4), from text
    bool sa_full_tried = false;
    DevDir view{};
};

// Scratch of the query pipelines, owned by one index.
struct Workspace {
    DevBuf qcopy, qoffcopy, refb;            // host-API staging of inputs
    DevBuf cnt, niv, ivs, status;             // locate fast path
    DevBuf ovf_list, counters;                // overflow queue + atomic counters
    DevBuf fb_sets, fb_ivs, fb_meta;          // fallback sets and results
    DevBuf hit_off, rows, hits;               // resolve
    DevBuf scan_tmp;
    DevBuf mem_recs, mem_cnt, mem_off, mem_sorted, mem_hits, pos_off, positions, mem_flags, read_list;
    DevBuf wave_sets, mem_slots, mem_status;
    DevBuf items, item_cnt, item_first, read_off; // MEM work items (read segments)
    DevBuf item_tot;                          // {work items, flat query length} of a MEM call
    DevBuf qnib;                              // queries as base-set nibbles
    DevBuf defer_ids;                         // work items the MEM main kernel deferred
    DevBuf part_cnt;                          // record counts of the parts of split deferred items
    uint64_t *h_pinned = nullptr;             // small readbacks
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}; // main, resolve, fallback kernels
};

// Index whose batch callback runs on this thread: hsfm_find_mems_batched
// holds the index lock across the callback, so entry points called from it on
// the same index return HSFM_EINVAL (HSFM_LOCK_INDEX) instead of deadlocking.
extern thread_local const hsfm_index *tl_callback_index;

// Pinned host words of an index (Workspace::h_pinned): small read-backs at
// [0, 64), the batched pipeline's per-set read-back slots at 64 (k + 1) and
// its per-set batch arguments at kPinnedArgs + 8 k (k < kMaxBatchSets).
constexpr size_t kPinnedBytes = 16384;
constexpr size_t kPinnedArgs = 1024;
struct BatchBufs; // buffers of the batched MEM pipeline (hsfm_mem.hip)
struct LsiView;   // learned-seed lookups inside the MEM kernel (hsfm_lsi_dev.h)
// The lookup view of a learned seed index built on idx, for MEM restarts;
// HSFM_EINVAL if it cannot seed them exactly.
int lsi_mem_view(const hsfm_lsi *x, hsfm_index *idx, LsiView *out);
uint32_t lsi_k(const hsfm_lsi *x);
void free_batch_bufs(BatchBufs *b);

} // namespace hsfm

struct hsfm_index {
    hsfm_device *dev = nullptr;
    hsfm::DirBuffers fwd, rev;
    bool has_rev = false;
    hsfm::DevConst h_const{};
    hsfm::DevBuf d_const;
    hsfm::DevBuf seq_bounds;
    uint32_t nseq = 0;
    std::mutex mu;
    hsfm::Workspace ws;
    hsfm_stats stats{};
    // work items the MEM main kernel deferred in the last call on the index
    // (ws.defer_ids; item table in ws.items when reads were split), valid
    // only when that call was a one-shot MEM call (hsfm_last_deferred)
    uint64_t defer_n = 0;
    bool defer_valid = false, defer_items = false;
    // restart seeding of the one-shot MEM calls (hsfm_mem_seed_mode)
    int mem_seed_mode = HSFM_SEED_TABLES;
    hsfm_lsi *mem_lsi = nullptr;
    // bidirectional seed table of every ACGT F-mer (built on first MEM call)
    hsfm::DevBuf seed_off, seed_iv, seed_lsl;
    uint32_t seed_F = 0;
    // packed one-load records of the 12-, 14- and 15-mer tables (the 14- and
    // 15-mers with their own interval lists)
    hsfm::DevBuf seed_rec12, seed_rec14, seed_iv14, seed_rec15, seed_iv15;
    // batched MEM pipeline: configuration picked by the autotuner
    int tuned_main = -1, tuned_rpt = -1;
    uint32_t tuned_min_len = 0, tuned_batch = 0;
    // its buffers, kept across calls (grown when a call needs more): set 0 on
    // the device stream, further sets with streams of their own, so that
    // consecutive batches overlap
    static constexpr int kMaxBatchSets = 8;
    hsfm::BatchBufs *batch[kMaxBatchSets] = {};
    hipStream_t bstream[kMaxBatchSets] = {};
};

namespace hsfm {
// Exclusive scan of u32 counts into u64 offsets (n+1 entries, last = total).
int exclusive_scan_u32_u64(hsfm_index *idx, const uint32_t *d_in, uint64_t *d_out, uint32_t n);
// Read back `count` u64 values from device memory (synchronizes the stream).
int readback_u64(hsfm_index *idx, const uint64_t *d_src, uint64_t *h_dst, int count);
int fill_u32(hsfm_index *idx, uint32_t *d, uint64_t n, uint32_t v);
// HIP events bracketing the main (slot 0) and resolve (slot 1) kernels.
int event_record(hsfm_index *idx, int slot, bool end);
double event_ms(hsfm_index *idx, int slot);
// Table of the bidirectional intervals of every ACGT F-mer's compatible strings
// (kept on the index until it is freed).
int build_seed_table(hsfm_index *idx, uint32_t F);
// Packed records of the F-mer table (F = 12 or 14), one 16-byte load per F-mer:
// {fwd_lo, fwd_hi, rev_lo, meta} when the F-mer has one compatible interval,
// else {offset lo, offset hi, 0, meta} into the interval list; meta = lsl |
// count << 5.  Builds the 12-mer table first.
int build_seed_packed(hsfm_index *idx, uint32_t F);
// Seed length of the persistent MEM kernel for min_len (12 or 14; 0 = none).
uint32_t mem_seed_length(hsfm_index *idx, uint32_t min_len);
// The forward direction's full suffix array (4 bytes per row), derived once on
// the device by LF walks between samples; resolve then reads it instead of
// walking.  Left unset (walks stay in use) if the walks do not cover every row
// within the sample rate or the memory is not available.
int ensure_sa_full(hsfm_index *idx);
// Resolve rows[0, min(cap, *d_nrows)) with rpt (1, 2, 4) interleaved walks per
// thread on a grid of `grid` blocks (no host-side row count needed).
int launch_resolve_fan(hsfm_index *idx, const DevDir &d, const uint32_t *d_rows, uint64_t cap,
                       const uint64_t *d_nrows, const uint32_t *d_bounds, uint32_t nbounds, uint32_t *d_out_pairs,
                       unsigned long long *d_guard, int rpt, uint32_t grid);

// Workgroups of a persistent kernel: as many as stay resident on every CU.
template <typename K>
inline uint32_t persist_grid(hsfm_index *idx, K kernel, uint32_t block) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)idx->dev->num_cus;
}


// Row expansion + LF-walk resolve + boundary map, shared by locate and MEM.
//   rows of segment s: [lo[s], lo[s] + len[s]) written from dst offsets off[s]
int launch_resolve(hsfm_index *idx, const DevDir &d, const uint32_t *d_rows, uint64_t nrows,
                   const uint32_t *d_bounds, uint32_t nbounds, uint32_t *d_out_pairs,
                   unsigned long long *d_guard);
} // namespace hsfm

// Serialize the calls on one index (see tl_callback_index).
#define HSFM_LOCK_INDEX(idx, name)                                                                    \
    if (::hsfm::tl_callback_index == (idx))                                                           \
        return ::hsfm::set_error(HSFM_EINVAL, "%s: called from a batch callback on the same index",   \
                                 name);                                                               \
    std::lock_guard<std::mutex> hsfm_index_lock_((idx)->mu)
