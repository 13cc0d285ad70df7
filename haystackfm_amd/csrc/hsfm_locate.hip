// hsfm_locate.hip -- batched count / locate on the gfx950 device layout.
//
// Pipeline (replaces locate_batch_gpu, src/gpu/locate.rs:52-277):
//   1. k_locate_search: one thread per read runs FmIndex::backward_search
//      (src/fm_index/query.rs:219-277) with lookup-table seeding, holding up to
//      kLocateRegCap SA intervals in registers.  Reads whose interval set grows
//      past that are queued for step 2 (the reference GPU instead drops
//      intervals beyond 16, locate_search.wgsl:177).
//   2. k_locate_wave: one wave per queued read keeps the whole interval set in
//      global memory and extends it cooperatively; sets larger than the per-wave
//      capacity are re-run with 4x the capacity until they fit, so no interval
//      is ever dropped.
//   3. per-read hit counts -> exclusive scan -> k_expand writes the SA rows of
//      every read in ascending row order.
//   4. k_resolve (hsfm_runtime.hip) LF-walks each row and maps it to
//      (seq_idx, pos_in_seq).
// Intervals are unions of rows; merging adjacent intervals (query.rs:390-407)
// does not change the row set, so it is not performed.
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>

#include <vector>

#include "hsfm_host.h"

namespace hsfm {

constexpr uint32_t KR = kLocateRegCap;
constexpr uint8_t kNivOverflow = 0xff;

struct LocateCounters {            // device counters (u64 each)
    unsigned long long n_ovf;      // reads the register kernel gives to the group pass
    unsigned long long work;       // wave path work index
    unsigned long long fb_used;    // fallback interval slots reserved
    unsigned long long n_retry;    // reads that need a larger wave capacity
    unsigned long long guard;      // walk guard hits
    unsigned long long probes;     // records fetched by the main search kernel
    unsigned long long n_wave;     // reads the group pass gives to the wave pass
    unsigned long long pad[1];
};

__device__ __forceinline__ void load_const_lds(const DevConst *cst, uint16_t *s_compat, int8_t *s_digit,
                                               uint32_t *s_clane) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        s_compat[i] = cst->compat_lanes[i];
        s_digit[i] = cst->lookup_digit[i];
    }
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
}

// Lookup seeding (query.rs:225-242, lookup.rs:101-111): the leftmost of the
// last `depth` codes is the least significant base-radix digit.
// Returns 0 = not seeded, 1 = seeded (lo,hi), 2 = seeded empty.
__device__ __forceinline__ int lookup_seed(const DevDir &d, const int8_t *s_digit, uint32_t depth, uint32_t radix,
                                           const uint8_t *p, uint32_t m, uint32_t &lo, uint32_t &hi) {
    if (depth == 0 || m < depth || !d.lookup) return 0;
    uint64_t idx = 0;
    for (int t = (int)depth - 1; t >= 0; t--) {
        const int dg = s_digit[p[m - depth + t]];
        if (dg < 0) return 0;
        idx = idx * radix + (uint64_t)dg;
    }
    if (idx >= d.lookup_len) idx = d.lookup_len - 1;
    const uint2 iv = d.lookup[idx];
    if (iv.x >= iv.y) return 2;
    lo = iv.x < d.n ? iv.x : d.n;
    hi = iv.y < d.n ? iv.y : d.n;
    return 1;
}

__device__ __forceinline__ uint32_t locate_read(const DevDir &d, const DevConst *cst, const uint8_t *q,
                                                const uint64_t *qoff, uint32_t r, uint32_t *cnt, uint8_t *niv,
                                                uint2 *ivs, uint32_t *ovf_list, LocateCounters *ctr,
                                                const uint16_t *s_compat, const int8_t *s_digit,
                                                const uint32_t *s_clane, const uint4 *seed_rec, uint32_t F,
                                                const uint8_t *s_expand);

__global__ __launch_bounds__(256) void k_locate_search(DevDir d, const DevConst *__restrict__ cst,
                                                       const uint8_t *__restrict__ q,
                                                       const uint64_t *__restrict__ qoff, uint32_t nq,
                                                       uint32_t *__restrict__ cnt, uint8_t *__restrict__ niv,
                                                       uint2 *__restrict__ ivs, uint32_t *__restrict__ ovf_list,
                                                       LocateCounters *__restrict__ ctr,
                                                       const uint4 *__restrict__ seed_rec, uint32_t F) {
    __shared__ uint16_t s_compat[256];
    __shared__ int8_t s_digit[256];
    __shared__ uint8_t s_expand[256];
    __shared__ uint32_t s_clane[16];
    load_const_lds(cst, s_compat, s_digit, s_clane);
    for (uint32_t x = threadIdx.x; x < 256; x += blockDim.x) s_expand[x] = cst->seed_expand[x];
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t probes = 0;
    if (r < nq)
        probes = locate_read(d, cst, q, qoff, r, cnt, niv, ivs, ovf_list, ctr, s_compat, s_digit, s_clane, seed_rec,
                             F, s_expand);
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) probes += __shfl_xor(probes, dd, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&ctr->probes, (unsigned long long)probes);
}


// A single row lo with q[i..m) matched at text position sa_full[lo]: the
// backward search from here follows that one occurrence (its only child is
// the row of the previous text position, when the query code there is
// compatible with the text), so the read matches iff the text left of the
// occurrence is compatible with q[0..i), and then its one row is the row of
// text position sa_full[lo] - i.  Compares 8 codes per load pair.
__device__ __forceinline__ void locate_finish_text(const DevDir &d, const uint16_t *s_compat,
                                                   const uint8_t *s_expand, const uint8_t *p, uint32_t i,
                                                   uint32_t &lo, uint32_t &hi, uint32_t &ns, uint32_t &probes) {
    uint32_t tp = d.sa_full[lo]; // text position of q[i]
    probes++;
    uint32_t tq = i;
    bool ok = true;
    if (d.text4 && d.qnib16) {
        // up to 16 codes per load pair, leftwards: the aligned 16-code group of
        // text nibbles holding text position tp - 1 and the aligned 16 query
        // bytes holding q[tq - 1], compared as base sets in registers (a code
        // matches while its set and the text lane's share a base; the
        // sentinel's set is empty, so the comparison stops there as the index
        // extension does)
        const unsigned long long tab = d.qnib16;
        while (tq > 0 && ok) {
            if (tp == 0) { // nothing left of the text start
                ok = false;
                break;
            }
            const uint32_t tl = tp - 1;
            const uintptr_t qa = reinterpret_cast<uintptr_t>(p + tq - 1);
            const uint2 tw = *reinterpret_cast<const uint2 *>(d.text4 + ((tl >> 3) & ~1u));
            const uint4 qw = *reinterpret_cast<const uint4 *>(qa & ~(uintptr_t)15);
            probes++;
            const uint32_t to = tl & 15u, qo = (uint32_t)(qa & 15);
            const uint32_t lim = min(min(to + 1, qo + 1), tq);
            // query bytes -> nibbles (byte j of the group -> nibble j); bytes
            // of 16 and more have no base
            const uint32_t qb[4] = {qw.x, qw.y, qw.z, qw.w};
            unsigned long long qn = 0;
#pragma unroll
            for (uint32_t j = 0; j < 16; j++) {
                const uint32_t c = (qb[j >> 2] >> (8 * (j & 3))) & 0xffu;
                qn |= (c < 16 ? (tab >> (4 * c)) & 15ull : 0ull) << (4 * j);
            }
            // position tl at the top nibble of t, q[tq - 1] at the top of q:
            // nibble 15 - k holds the k-th code leftwards of both
            const unsigned long long t = ((unsigned long long)tw.x | ((unsigned long long)tw.y << 32)) << (4 * (15 - to));
            const unsigned long long x = t & (qn << (4 * (15 - qo)));
            const unsigned long long z = ~(x | (x >> 1) | (x >> 2) | (x >> 3)) & 0x1111111111111111ull;
            // leading codes that match: up to the highest zero nibble
            const uint32_t run = z ? (uint32_t)__clzll((long long)z) >> 2 : 16u;
            const uint32_t adv = run < lim ? run : lim;
            if (run < lim) ok = false;
            tq -= adv;
            tp -= adv;
        }
    } else {
        while (tq > 0 && ok) {
            if (tp == 0) { // nothing left of the text start
                ok = false;
                break;
            }
            const uint32_t tl = tp - 1;
            const uintptr_t qa = reinterpret_cast<uintptr_t>(p + tq - 1);
            const uint2 tw = *reinterpret_cast<const uint2 *>(d.text + (tl & ~7u));
            const uint2 qw = *reinterpret_cast<const uint2 *>(qa & ~(uintptr_t)7);
            probes++;
            const uint32_t to = tl & 7u, qo = (uint32_t)(qa & 7);
            const uint32_t lim = min(min(to + 1, qo + 1), tq);
            const uint64_t t64 = (uint64_t)tw.x | ((uint64_t)tw.y << 32), q64 = (uint64_t)qw.x | ((uint64_t)qw.y << 32);
            uint32_t adv = 0;
#pragma unroll
            for (uint32_t k = 0; k < 8; k++) {
                if (k < lim && ok) {
                    const uint32_t lane = (uint32_t)(t64 >> (8 * (to - k))) & 0xffu;
                    const uint32_t code = (uint32_t)(q64 >> (8 * (qo - k))) & 0xffu;
                    if ((s_compat[code] >> lane) & 1u) adv++;
                    else ok = false;
                }
            }
            tq -= adv;
            tp -= adv;
        }
    }
    if (ok) {
        lo = d.isa[tp];
        hi = lo + 1;
        ns = 1;
    } else {
        ns = 0;
    }
}

// Backward search of one read (the body of k_locate_search); returns the
// number of 64 B records it fetched.
__device__ __forceinline__ uint32_t locate_read(const DevDir &d, const DevConst *cst, const uint8_t *q,
                                                const uint64_t *qoff, uint32_t r, uint32_t *cnt, uint8_t *niv,
                                                uint2 *ivs, uint32_t *ovf_list, LocateCounters *ctr,
                                                const uint16_t *s_compat, const int8_t *s_digit,
                                                const uint32_t *s_clane, const uint4 *seed_rec, uint32_t F,
                                                const uint8_t *s_expand) {
    uint32_t probes = 0;
    const uint64_t b0 = qoff[r];
    const uint32_t m = (uint32_t)(qoff[r + 1] - b0);
    const uint8_t *p = q + b0;
    uint32_t lo[KR], hi[KR], ns = 1;
#pragma unroll
    for (uint32_t j = 0; j < KR; j++) { lo[j] = 0xffffffffu; hi[j] = 0xffffffffu; }
    lo[0] = 0;
    hi[0] = d.n;
    int start = (int)m;
    const int sd = lookup_seed(d, s_digit, cst->lookup_depth, cst->lookup_radix, p, m, lo[0], hi[0]);
    if (sd) probes++;
    if (sd == 2) ns = 0;
    else if (sd == 1) start = (int)(m - cst->lookup_depth);
    else if (seed_rec && m >= F) {
        // the lookup cannot seed (an ambiguity code among the last codes):
        // the packed F-mer records of the expansions of q[m-F..m) (one
        // ambiguity code: at most four, loaded together) give the set of
        // every compatible string when each holds a single interval
        SeedWin sw;
        if (seed_window(p + m - F, F, s_expand, sw, 1)) {
            uint4 R[4];
            bool val[4];
#pragma unroll
            for (uint32_t c = 0; c < 4; c++) {
                uint32_t ix = 0;
                val[c] = c < seed_ncombo(sw) && seed_combo(sw, c, ix);
                R[c] = val[c] ? seed_rec[ix] : make_uint4(0, 0, 0, 0);
            }
            bool ok = true;
            uint32_t k2 = 0;
#pragma unroll
            for (uint32_t c = 0; c < 4; c++) {
                if (!val[c]) continue;
                probes++;
                const uint32_t l = seed_rec_lsl(R[c].w), n2 = seed_rec_cnt(R[c].w);
                if (l != F || n2 == 0) continue;
                if (n2 > 1) ok = false;
                bool dup = false;
#pragma unroll
                for (uint32_t t = 0; t < KR; t++) dup |= t < k2 && lo[t] == R[c].x;
                if (!dup && k2 < KR) {
#pragma unroll
                    for (uint32_t t = 0; t < KR; t++)
                        if (t == k2) { lo[t] = R[c].x; hi[t] = R[c].y; }
                    k2++;
                }
            }
            if (ok) {
                ns = k2;
                start = (int)(m - F);
            } else {
                lo[0] = 0;
                hi[0] = d.n;
#pragma unroll
                for (uint32_t t = 1; t < KR; t++) { lo[t] = 0xffffffffu; hi[t] = 0xffffffffu; }
            }
        }
    }
    const bool textmode = d.isa != nullptr && d.text != nullptr && d.sa_full != nullptr;
    if (textmode && ns == 1 && hi[0] - lo[0] == 1 && start > 0) {
        locate_finish_text(d, s_compat, s_expand, p, (uint32_t)start, lo[0], hi[0], ns, probes);
        start = 0;
    }
    for (int i = start - 1; i >= 0 && ns; i--) {
        const uint32_t cm = s_compat[p[i]];
        uint32_t nlo[KR], nhi[KR], nn = 0;
        bool ovf = false;
#pragma unroll
        for (uint32_t j = 0; j < KR; j++) { nlo[j] = 0xffffffffu; nhi[j] = 0xffffffffu; }
#pragma unroll
        for (uint32_t j = 0; j < KR; j++) {
            if (j < ns) {
                probes += 1u + ((lo[j] >> kRecShift) != (hi[j] >> kRecShift) ? 1u : 0u);
                uint32_t cl[16], ch[16];
                pair_counts(d, lo[j], hi[j], cl, ch); // one superblock row for both ends
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    if (((cm >> k) & 1u) && ch[k] > cl[k]) {
                        const uint32_t a = s_clane[k] + cl[k], b = s_clane[k] + ch[k];
                        if (nn >= KR) ovf = true;
#pragma unroll
                        for (uint32_t t = 0; t < KR; t++)
                            if (t == nn) { nlo[t] = a; nhi[t] = b; }
                        nn++;
                    }
                }
            }
        }
        if (ovf) {
            niv[r] = kNivOverflow;
            cnt[r] = 0;
            const unsigned long long at = atomicAdd(&ctr->n_ovf, 1ull);
            ovf_list[at] = r;
            return probes;
        }
#pragma unroll
        for (uint32_t j = 0; j < KR; j++) { lo[j] = nlo[j]; hi[j] = nhi[j]; }
        ns = nn;
        if (textmode && ns == 1 && hi[0] - lo[0] == 1 && i > 0) {
            locate_finish_text(d, s_compat, s_expand, p, (uint32_t)i, lo[0], hi[0], ns, probes);
            break;
        }
    }
    // sort the (<= KR) disjoint intervals by start; empty slots hold 0xffffffff
#pragma unroll
    for (uint32_t a = 0; a < KR; a++)
#pragma unroll
        for (uint32_t b = 0; b + 1 < KR - a; b++) {
            const bool sw = lo[b] > lo[b + 1];
            const uint32_t tl = sw ? lo[b + 1] : lo[b], th = sw ? hi[b + 1] : hi[b];
            const uint32_t ul = sw ? lo[b] : lo[b + 1], uh = sw ? hi[b] : hi[b + 1];
            lo[b] = tl; hi[b] = th; lo[b + 1] = ul; hi[b + 1] = uh;
        }
    uint32_t total = 0;
#pragma unroll
    for (uint32_t j = 0; j < KR; j++)
        if (j < ns) {
            total += hi[j] - lo[j];
            ivs[(size_t)r * KR + j] = make_uint2(lo[j], hi[j]);
        }
    niv[r] = (uint8_t)ns;
    cnt[r] = total;
    return probes;
}

// ---------------------------------------------------------------------------
// Second pass over the reads the register kernel gave up on, one group of G
// lanes per read: the same backward search with room for G intervals, one per
// lane, so a step costs one dependent round trip for the whole set.  Final
// sets of up to KR intervals go to ivs; larger ones are published in row order
// to fb_ivs (fb_meta = offset, count; niv = kNivOverflow); reads whose sets
// outgrow G (or the published list) are queued for the workgroup kernel.
// Reads the lookup table cannot seed start from the packed F-mer records of
// the expansions of their last F codes (one ambiguity code) when those hold
// at most G intervals together.  A step's children are counted per lane, placed by a
// prefix sum over the group (sub-wave shuffles) and gathered through the
// group's LDS list; seeding and the text finish run on lane 0.  Every
// decision (set size, give-up, loop index) is uniform over the group.
// ---------------------------------------------------------------------------
// One read of the group pass (G lanes of one wave).
template <uint32_t G>
__device__ __forceinline__ void locate_group_read(const DevDir &d, const DevConst *__restrict__ cst,
                                                  const uint8_t *__restrict__ q, const uint64_t *__restrict__ qoff,
                                                  uint32_t r, uint32_t *__restrict__ cnt, uint8_t *__restrict__ niv,
                                                  uint2 *__restrict__ ivs, uint2 *__restrict__ fb_meta,
                                                  uint2 *__restrict__ fb_ivs, uint64_t fb_cap,
                                                  uint32_t *__restrict__ ovf_list, LocateCounters *__restrict__ ctr,
                                                  const uint4 *__restrict__ seed_rec, const uint4 *__restrict__ seed_iv,
                                                  uint32_t F, const uint16_t *s_compat, const int8_t *s_digit,
                                                  const uint8_t *s_expand, const uint32_t *s_clane, uint32_t *glo,
                                                  uint32_t *ghi, uint32_t gl) {
    const uint64_t b0 = qoff[r];
    const uint32_t m = (uint32_t)(qoff[r + 1] - b0);
    const uint8_t *p = q + b0;

    // seeding on lane 0, into the group's list
    uint32_t ns = 0, give = 0;
    int start = (int)m;
    if (gl == 0) {
        uint32_t lo0 = 0, hi0 = d.n;
        ns = 1;
        const int sd = lookup_seed(d, s_digit, cst->lookup_depth, cst->lookup_radix, p, m, lo0, hi0);
        bool listed = false;
        if (sd == 2) ns = 0;
        else if (sd == 1) start = (int)(m - cst->lookup_depth);
        else if (seed_rec && m >= F) {
            SeedWin sw;
            if (seed_window(p + m - F, F, s_expand, sw, 1)) {
                uint32_t k2 = 0;
                bool ok = true;
                for (uint32_t c = 0; c < seed_ncombo(sw) && ok; c++) {
                    uint32_t ix = 0;
                    if (!seed_combo(sw, c, ix)) continue;
                    const uint4 R = seed_rec[ix];
                    const uint32_t l = seed_rec_lsl(R.w), n2 = seed_rec_cnt(R.w);
                    if (l != F || n2 == 0) continue;
                    const uint64_t o = (uint64_t)R.x | ((uint64_t)R.y << 32);
                    for (uint32_t e = 0; e < n2 && ok; e++) {
                        const uint2 iv =
                            n2 == 1 ? make_uint2(R.x, R.y) : make_uint2(seed_iv[o + e].x, seed_iv[o + e].y);
                        bool dup = false;
                        for (uint32_t u = 0; u < k2 && !dup; u++) dup = glo[u] == iv.x;
                        if (dup) continue;
                        if (k2 >= G) { ok = false; break; }
                        glo[k2] = iv.x;
                        ghi[k2] = iv.y;
                        k2++;
                    }
                }
                if (ok) {
                    ns = k2;
                    start = (int)(m - F);
                    listed = true;
                } else {
                    give = 1;
                }
            }
        }
        if (!listed && ns == 1) {
            glo[0] = lo0;
            ghi[0] = hi0;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    ns = __shfl(ns, 0, G);
    give = __shfl(give, 0, G);
    start = __shfl(start, 0, G);
    uint32_t mlo = 0xffffffffu, mhi = 0xffffffffu;
    if (gl < ns) {
        mlo = glo[gl];
        mhi = ghi[gl];
    }
    const bool textmode = d.isa != nullptr && d.text != nullptr && d.sa_full != nullptr;
    uint32_t tprobes = 0;
    // Rows held by the group's set, saturated at G + 1 (every lane gets the sum).
    auto group_rows = [&]() {
        uint32_t v = gl < ns ? min(mhi - mlo, G + 1u) : 0u;
#pragma unroll
        for (uint32_t dd = G / 2; dd >= 1; dd >>= 1) v += __shfl_xor(v, dd, G);
        return v;
    };
    // At most G rows with q[at..m) matched: every row is one text occurrence,
    // and backward search from a row follows that occurrence alone (its only
    // child is the row of the previous text position, when the query code is
    // compatible with the text there).  So each lane takes one row and
    // compares q[0..at) with the text left of it; the survivors' rows (the
    // inverse suffix array at their text starts) are the final set, one row
    // per interval, in the group's list.
    auto finish_rows = [&](uint32_t at, uint32_t nrows) {
        uint32_t sz = gl < ns ? mhi - mlo : 0u, off = sz;
#pragma unroll
        for (uint32_t dd = 1; dd < G; dd <<= 1) {
            const uint32_t y = __shfl_up(off, dd, G);
            if (gl >= dd) off += y;
        }
        off -= sz;
        for (uint32_t u = 0; u < sz; u++) glo[off + u] = mlo + u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint32_t rlo = gl < nrows ? glo[gl] : 0u, rhi = rlo + 1u, rns = gl < nrows ? 1u : 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (rns) locate_finish_text(d, s_compat, s_expand, p, at, rlo, rhi, rns, tprobes);
        uint32_t pos = rns;
#pragma unroll
        for (uint32_t dd = 1; dd < G; dd <<= 1) {
            const uint32_t y = __shfl_up(pos, dd, G);
            if (gl >= dd) pos += y;
        }
        ns = __shfl(pos, G - 1, G);
        if (rns) glo[pos - 1] = rlo;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        mlo = gl < ns ? glo[gl] : 0xffffffffu;
        mhi = gl < ns ? mlo + 1u : 0xffffffffu;
    };
    if (!give && textmode && ns && start > 0) {
        const uint32_t nrows = group_rows();
        if (nrows <= G) {
            finish_rows((uint32_t)start, nrows);
            start = 0;
        }
    }
    for (int i = start - 1; i >= 0 && ns && !give; i--) {
        const uint32_t cm = s_compat[p[i]];
        uint32_t cl[16], ch[16], c = 0;
        if (gl < ns) {
            pair_counts(d, mlo, mhi, cl, ch);
#pragma unroll
            for (int k = 0; k < 16; k++) c += (((cm >> k) & 1u) && ch[k] > cl[k]) ? 1u : 0u;
        }
        uint32_t inc = c;
#pragma unroll
        for (uint32_t dd = 1; dd < G; dd <<= 1) {
            const uint32_t y = __shfl_up(inc, dd, G);
            if (gl >= dd) inc += y;
        }
        const uint32_t tot = __shfl(inc, G - 1, G);
        if (tot > G) {
            give = 1;
            break;
        }
        if (gl < ns) {
            uint32_t w = inc - c;
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (((cm >> k) & 1u) && ch[k] > cl[k]) {
                    glo[w] = s_clane[k] + cl[k];
                    ghi[w] = s_clane[k] + ch[k];
                    w++;
                }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        ns = tot;
        mlo = gl < ns ? glo[gl] : 0xffffffffu;
        mhi = gl < ns ? ghi[gl] : 0xffffffffu;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (textmode && ns && i > 0) {
            const uint32_t nrows = group_rows();
            if (nrows <= G) {
                finish_rows((uint32_t)i, nrows);
                break;
            }
        }
    }
    if (!give) {
        // row order: rank among the group's intervals (disjoint, distinct lo)
        uint32_t rank = 0, size = gl < ns ? mhi - mlo : 0u;
#pragma unroll
        for (uint32_t u = 0; u < G; u++) {
            const uint32_t lu = __shfl(mlo, u, G);
            rank += (u < ns && lu < mlo) ? 1u : 0u;
        }
#pragma unroll
        for (uint32_t dd = G / 2; dd >= 1; dd >>= 1) size += __shfl_xor(size, dd, G);
        if (ns <= KR) {
            if (gl < ns) ivs[(size_t)r * KR + rank] = make_uint2(mlo, mhi);
            if (gl == 0) {
                niv[r] = (uint8_t)ns;
                cnt[r] = size;
            }
            return;
        }
        unsigned long long slot = 0;
        if (gl == 0) slot = atomicAdd(&ctr->fb_used, (unsigned long long)ns);
        slot = __shfl(slot, 0, G);
        if (slot + ns <= fb_cap) {
            if (gl < ns) fb_ivs[slot + rank] = make_uint2(mlo, mhi);
            if (gl == 0) {
                fb_meta[r] = make_uint2((uint32_t)slot, ns);
                niv[r] = kNivOverflow;
                cnt[r] = size;
            }
            return;
        }
    }
    if (gl == 0) {
        niv[r] = kNivOverflow;
        cnt[r] = 0;
        ovf_list[atomicAdd(&ctr->n_wave, 1ull)] = r;
    }
}

template <uint32_t G>
__global__ __launch_bounds__(256) void k_locate_search_group(DevDir d, const DevConst *__restrict__ cst,
                                                             const uint8_t *__restrict__ q,
                                                             const uint64_t *__restrict__ qoff,
                                                             const uint32_t *__restrict__ list, uint32_t nlist,
                                                             uint32_t *__restrict__ cnt, uint8_t *__restrict__ niv,
                                                             uint2 *__restrict__ ivs, uint2 *__restrict__ fb_meta,
                                                             uint2 *__restrict__ fb_ivs, uint64_t fb_cap,
                                                             uint32_t *__restrict__ ovf_list,
                                                             LocateCounters *__restrict__ ctr,
                                                             const uint4 *__restrict__ seed_rec,
                                                             const uint4 *__restrict__ seed_iv, uint32_t F) {
    static_assert(G >= 16 && G <= 64 && (G & (G - 1)) == 0, "group size: a power of two in 16..64");
    constexpr uint32_t NG = 256 / G;
    __shared__ uint16_t s_compat[256];
    __shared__ int8_t s_digit[256];
    __shared__ uint8_t s_expand[256];
    __shared__ uint32_t s_clane[16];
    __shared__ uint32_t s_lo[NG][G], s_hi[NG][G];
    load_const_lds(cst, s_compat, s_digit, s_clane);
    for (uint32_t x = threadIdx.x; x < 256; x += blockDim.x) s_expand[x] = cst->seed_expand[x];
    __syncthreads();
    const uint32_t gl = threadIdx.x % G, gi = threadIdx.x / G;
    // a persistent grid: the register kernel's count of reads comes from the
    // device, so no host round trip sits between the two passes
    const uint32_t nl = (uint32_t)ctr->n_ovf;
    for (uint32_t t = blockIdx.x * NG + gi; t < nl; t += gridDim.x * NG)
        locate_group_read<G>(d, cst, q, qoff, list[t], cnt, niv, ivs, fb_meta, fb_ivs, fb_cap, ovf_list, ctr,
                             seed_rec, seed_iv, F, s_compat, s_digit, s_expand, s_clane, s_lo[gi], s_hi[gi], gl);
}

// One wave per queued read.  Set buffers: per wave 2 * cap (lo,hi) pairs.
__global__ __launch_bounds__(64) void k_locate_wave(DevDir d, const DevConst *__restrict__ cst,
                                                    const uint8_t *__restrict__ q,
                                                    const uint64_t *__restrict__ qoff,
                                                    const uint32_t *__restrict__ list, uint32_t nlist,
                                                    uint32_t cap, uint2 *__restrict__ sets,
                                                    uint32_t *__restrict__ cnt, uint2 *__restrict__ fb_meta,
                                                    uint2 *__restrict__ fb_ivs, uint64_t fb_cap,
                                                    uint32_t *__restrict__ retry_list,
                                                    LocateCounters *__restrict__ ctr,
                                                    const unsigned long long *__restrict__ nlist_dev,
                                                    const uint4 *__restrict__ seed_rec,
                                                    const uint4 *__restrict__ seed_iv, uint32_t F) {
    __shared__ uint16_t s_compat[256];
    __shared__ int8_t s_digit[256];
    __shared__ uint8_t s_expand[256];
    __shared__ uint32_t s_clane[16];
    load_const_lds(cst, s_compat, s_digit, s_clane);
    for (uint32_t x = threadIdx.x; x < 256; x += blockDim.x) s_expand[x] = cst->seed_expand[x];
    __syncthreads();
    const uint32_t lane = threadIdx.x;
    const bool textmode = d.isa != nullptr && d.text != nullptr && d.sa_full != nullptr && cap >= 64;
    uint2 *bufA = sets + (size_t)blockIdx.x * 2 * cap;
    uint2 *bufB = bufA + cap;
    if (nlist_dev) nlist = (uint32_t)*nlist_dev; // the group pass's count, without a host round trip
    for (;;) {
        uint32_t wi = 0;
        if (lane == 0) wi = (uint32_t)atomicAdd(&ctr->work, 1ull);
        wi = __shfl(wi, 0, 64);
        if (wi >= nlist) break;
        const uint32_t r = list[wi];
        const uint64_t b0 = qoff[r];
        const uint32_t m = (uint32_t)(qoff[r + 1] - b0);
        const uint8_t *p = q + b0;
        uint2 *A = bufA, *B = bufB;
        uint32_t ns = 1, lo0 = 0, hi0 = d.n;
        int start = (int)m;
        const int sd = lookup_seed(d, s_digit, cst->lookup_depth, cst->lookup_radix, p, m, lo0, hi0);
        if (sd == 2) ns = 0;
        else if (sd == 1) start = (int)(m - cst->lookup_depth);
        bool listed = false;
        if (sd == 0 && seed_rec && m >= F) {
            // the lookup cannot seed (ambiguity codes among the last codes): the
            // set of q[m-F..m) is the union of the F-mer table entries of its
            // expansions (up to five ambiguity codes, spread over the lanes),
            // an interval shared by several expansions kept once
            WideWin sw;
            if (wide_window(p + m - F, F, s_expand, sw)) {
                const uint32_t nc = 1u << (2 * sw.namb);
                uint32_t total = 0;
                bool over = false;
                for (uint32_t x0 = 0; x0 < nc && !over; x0 += 64) {
                    const uint32_t xi = x0 + lane;
                    uint32_t ix = 0, c = 0;
                    uint64_t o0 = 0;
                    uint2 one = make_uint2(0u, 0u);
                    if (xi < nc && wide_combo(sw, xi, ix)) {
                        const uint4 R = seed_rec[ix];
                        if (seed_rec_lsl(R.w) == F) {
                            c = seed_rec_cnt(R.w);
                            o0 = (uint64_t)R.x | ((uint64_t)R.y << 32);
                            one = make_uint2(R.x, R.y);
                        }
                    }
                    const uint32_t incl = wave_incl_scan(c), tot = __shfl(incl, 63, 64);
                    if (total + tot > cap) {
                        over = true;
                    } else {
                        const uint32_t at = total + incl - c;
                        if (c == 1) A[at] = one;
                        else
                            for (uint32_t t = 0; t < c; t++) A[at + t] = make_uint2(seed_iv[o0 + t].x, seed_iv[o0 + t].y);
                        total += tot;
                    }
                }
                if (!over) {
                    __syncthreads();
                    uint32_t kept = total;
                    if (sw.namb) {
                        kept = 0;
                        for (uint32_t base = 0; base < total; base += 64) {
                            const uint32_t t = base + lane;
                            bool keep = t < total;
                            uint2 iv = make_uint2(0u, 0u);
                            if (keep) {
                                iv = A[t];
                                for (uint32_t u = 0; u < t && keep; u++) keep = A[u].x != iv.x;
                            }
                            const uint32_t incl = wave_incl_scan(keep ? 1u : 0u);
                            if (keep) B[kept + incl - 1] = iv;
                            kept += __shfl(incl, 63, 64);
                        }
                        __syncthreads();
                        uint2 *tmp = A; A = B; B = tmp;
                    }
                    ns = kept;
                    start = (int)(m - F);
                    listed = true;
                }
            }
        }
        if (!listed && lane == 0) A[0] = make_uint2(lo0, hi0);
        __syncthreads();
        bool too_big = false;
        // At most 64 rows with q[at..m) matched: every row is one text
        // occurrence that backward search follows alone, so each lane takes a
        // row and compares q[0..at) with the text left of it; the survivors
        // (one row each) become the final set in A.  Returns true when done.
        auto finish_rows = [&](uint32_t at) -> bool {
            uint32_t part = 0;
            for (uint32_t j = lane; j < ns; j += 64) part += min(A[j].y - A[j].x, 65u);
            const uint32_t rows = wave_sum(part);
            if (rows > 64) return false;
            const uint2 iv = lane < ns ? A[lane] : make_uint2(0u, 0u); // rows <= 64 implies ns <= 64
            const uint32_t sz = iv.y - iv.x, off = wave_incl_scan(sz) - sz;
            for (uint32_t u = 0; u < sz; u++) B[off + u] = make_uint2(iv.x + u, 0u);
            __syncthreads();
            uint32_t rlo = lane < rows ? B[lane].x : 0u, rhi = rlo + 1u, rns = lane < rows ? 1u : 0u, pr = 0;
            if (rns) locate_finish_text(d, s_compat, s_expand, p, at, rlo, rhi, rns, pr);
            const uint32_t pos = wave_incl_scan(rns);
            ns = __shfl(pos, 63, 64);
            if (rns) A[pos - 1] = make_uint2(rlo, rlo + 1u);
            __syncthreads();
            return true;
        };
        if (textmode && ns && start > 0 && finish_rows((uint32_t)start)) start = 0;
        for (int i = start - 1; i >= 0 && ns && !too_big; i--) {
            const uint32_t cm = s_compat[p[i]];
            uint32_t total = 0;
            for (uint32_t base = 0; base < ns; base += 64) {
                const uint32_t j = base + lane;
                uint32_t cl[16], ch[16], cc = 0;
                if (j < ns) {
                    const uint2 iv = A[j];
                    pair_counts(d, iv.x, iv.y, cl, ch);
#pragma unroll
                    for (int k = 0; k < 16; k++) cc += (((cm >> k) & 1u) && ch[k] > cl[k]) ? 1u : 0u;
                }
                const uint32_t incl = wave_incl_scan(cc);
                const uint32_t tot = __shfl(incl, 63, 64);
                if (total + tot > cap) { too_big = true; break; }
                if (j < ns && cc) {
                    uint32_t w = total + incl - cc;
#pragma unroll
                    for (int k = 0; k < 16; k++)
                        if (((cm >> k) & 1u) && ch[k] > cl[k]) {
                            B[w++] = make_uint2(s_clane[k] + cl[k], s_clane[k] + ch[k]);
                        }
                }
                total += tot;
            }
            __syncthreads();
            uint2 *t = A; A = B; B = t;
            ns = too_big ? ns : total;
            if (!too_big && textmode && ns && i > 0 && finish_rows((uint32_t)i)) break;
        }
        if (too_big) {
            if (lane == 0) {
                const unsigned long long at = atomicAdd(&ctr->n_retry, 1ull);
                retry_list[at] = r;
            }
            continue;
        }
        // publish the set; it is put in row order by a segmented sort afterwards
        uint64_t slot = 0;
        if (lane == 0) slot = atomicAdd(&ctr->fb_used, (unsigned long long)ns);
        slot = __shfl(slot, 0, 64);
        if (slot + ns > fb_cap) {
            if (lane == 0) {
                const unsigned long long at = atomicAdd(&ctr->n_retry, 1ull);
                retry_list[at] = r;
            }
            continue;
        }
        uint32_t part = 0;
        for (uint32_t j = lane; j < ns; j += 64) {
            const uint2 iv = A[j];
            part += iv.y - iv.x;
            fb_ivs[slot + j] = iv; // ordered afterwards by k_fb_sort_* + segmented sort
        }
        const uint32_t count = wave_sum(part);
        if (lane == 0) {
            cnt[r] = count;
            fb_meta[r] = make_uint2((uint32_t)slot, ns);
            // slot may exceed 32 bits only beyond 4G intervals; fb_cap bounds it
        }
        __syncthreads();
    }
}

// Segments of the wave-path reads and 64-bit sort keys (start << 32 | end) of
// their intervals; intervals are disjoint, so key order is row order.
__global__ void k_fb_segments(const uint32_t *__restrict__ reads, uint32_t nreads, const uint2 *__restrict__ fb_meta,
                              uint32_t *__restrict__ seg_begin, uint32_t *__restrict__ seg_end) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nreads) return;
    const uint2 mt = fb_meta[reads[i]];
    seg_begin[i] = mt.x;
    seg_end[i] = mt.x + mt.y;
}
__global__ void k_fb_keys(const uint2 *__restrict__ iv, uint64_t n, unsigned long long *__restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = ((unsigned long long)iv[i].x << 32) | iv[i].y;
}
__global__ void k_fb_unkeys(const unsigned long long *__restrict__ keys, uint64_t n, uint2 *__restrict__ iv) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        iv[i] = make_uint2((uint32_t)(keys[i] >> 32), (uint32_t)keys[i]);
}

// Write the SA rows of every read at its hit offset, in ascending row order.
__global__ __launch_bounds__(256) void k_expand(uint32_t nq, const uint8_t *__restrict__ niv,
                                                const uint2 *__restrict__ ivs, const uint2 *__restrict__ fb_meta,
                                                const uint2 *__restrict__ fb_ivs,
                                                const uint64_t *__restrict__ hit_off, uint32_t *__restrict__ rows) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nq) return;
    uint64_t w = hit_off[r];
    const uint64_t e = hit_off[r + 1];
    if (w == e) return;
    const uint8_t k = niv[r];
    const uint2 *src;
    uint32_t ns;
    if (k == kNivOverflow) {
        const uint2 mt = fb_meta[r];
        src = fb_ivs + mt.x;
        ns = mt.y;
    } else {
        src = ivs + (size_t)r * KR;
        ns = k;
    }
    for (uint32_t j = 0; j < ns && w < e; j++) {
        const uint2 iv = src[j];
        for (uint32_t row = iv.x; row < iv.y && w < e; row++) rows[w++] = row;
    }
}

// Runs steps 1-2 and leaves per-read counts in ws.cnt (plus intervals).
// total (optional): the scan of the hit counts into ws.hit_off runs behind
// the passes, so its total comes back with their counters in one read-back
// (redone after capacity retries, which change counts).
static int locate_search_phase(hsfm_index *idx, const uint8_t *d_q, const uint64_t *d_qoff, uint32_t nq,
                               uint64_t *total = nullptr) {
    Workspace &ws = idx->ws;
    hipStream_t s = idx->dev->stream;
    const DevDir &d = idx->fwd.view;
    int rc;
    if ((rc = ws.cnt.ensure((size_t)nq * 4 + 4))) return rc;
    if ((rc = ws.niv.ensure((size_t)nq + 4))) return rc;
    if ((rc = ws.ivs.ensure((size_t)nq * KR * 8 + 8))) return rc;
    if ((rc = ws.ovf_list.ensure((size_t)nq * 4 + 4))) return rc;
    if ((rc = ws.read_list.ensure((size_t)nq * 4 + 4))) return rc;
    if ((rc = ws.counters.ensure(sizeof(LocateCounters)))) return rc;
    if ((rc = ws.fb_meta.ensure((size_t)nq * 8 + 8))) return rc;
    LocateCounters *ctr = ws.counters.as<LocateCounters>();
    if ((rc = fill_u32(idx, ws.counters.as<uint32_t>(), sizeof(LocateCounters) / 4, 0))) return rc;
    // final sets of more than KR intervals are published to fb_ivs
    if ((rc = ws.fb_ivs.ensure(((size_t)nq * 2 + 65536) * 8))) return rc;
    // 1) register kernel over every read (at most KR intervals per read)
    // 2) the reads it gives up on: 16 lanes per read, room for 16 intervals
    // 3) what still outgrows those: one wave per read
    // (a pipelined persistent ring kernel, for every read or as the second
    // level, and a thread-per-read 16-interval second level were measured
    // slower and are not kept)
    if (nq && idx->fwd.view.n > (1u << 20) && (rc = build_seed_packed(idx, 12))) return rc;
    uint64_t h[8];
    // register kernel -> group pass -> wave pass without host round trips
    // between them (each takes its read count from the device counters), then
    // one read-back of the counters
    const uint32_t wgrid = (uint32_t)idx->dev->num_cus * 8;
    uint32_t cap = 4096;
    if ((rc = ws.wave_sets.ensure((size_t)wgrid * 2 * cap * 8))) return rc;
    uint64_t fb_cap = ws.fb_ivs.cap / 8;
    if (nq) {
        if ((rc = event_record(idx, 0, false))) return rc;
        hipLaunchKernelGGL(k_locate_search, dim3((nq + 255) / 256), dim3(256), 0, s, d,
                           idx->d_const.as<DevConst>(), d_q, d_qoff, nq, ws.cnt.as<uint32_t>(),
                           ws.niv.as<uint8_t>(), ws.ivs.as<uint2>(), ws.ovf_list.as<uint32_t>(), ctr,
                           idx->seed_rec12.as<uint4>(), idx->seed_rec12.p ? 12u : 0u);
        HSFM_HIP(hipGetLastError());
        if ((rc = event_record(idx, 0, true))) return rc;
        if ((rc = event_record(idx, 2, false))) return rc;
        hipLaunchKernelGGL(k_locate_search_group<16>, dim3(persist_grid(idx, k_locate_search_group<16>, 256)),
                           dim3(256), 0, s, d, idx->d_const.as<DevConst>(), d_q, d_qoff,
                           ws.ovf_list.as<uint32_t>(), nq, ws.cnt.as<uint32_t>(), ws.niv.as<uint8_t>(),
                           ws.ivs.as<uint2>(), ws.fb_meta.as<uint2>(), ws.fb_ivs.as<uint2>(), fb_cap,
                           ws.read_list.as<uint32_t>(), ctr, idx->seed_rec12.as<uint4>(), idx->seed_iv.as<uint4>(),
                           idx->seed_rec12.p ? 12u : 0u);
        HSFM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_locate_wave, dim3(wgrid), dim3(64), 0, s, d, idx->d_const.as<DevConst>(), d_q, d_qoff,
                           ws.read_list.as<uint32_t>(), 0u, cap, ws.wave_sets.as<uint2>(), ws.cnt.as<uint32_t>(),
                           ws.fb_meta.as<uint2>(), ws.fb_ivs.as<uint2>(), fb_cap, ws.ovf_list.as<uint32_t>(), ctr,
                           &ctr->n_wave, idx->seed_rec12.as<uint4>(), idx->seed_iv.as<uint4>(),
                           idx->seed_rec12.p ? 12u : 0u);
        HSFM_HIP(hipGetLastError());
    }
    if (total) {
        if ((rc = ws.hit_off.ensure(((size_t)nq + 1) * 8))) return rc;
        if ((rc = exclusive_scan_u32_u64(idx, ws.cnt.as<uint32_t>(), ws.hit_off.as<uint64_t>(), nq))) return rc;
        if (!ws.h_pinned) HSFM_HIP(hipHostMalloc((void **)&ws.h_pinned, kPinnedBytes, hipHostMallocDefault));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned, ctr, 64, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned + 8, ws.hit_off.as<uint64_t>() + nq, 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
        memcpy(h, ws.h_pinned, 64);
        *total = ws.h_pinned[8];
    } else if ((rc = readback_u64(idx, reinterpret_cast<const uint64_t *>(ctr), h, 8))) {
        return rc;
    }
    idx->stats.occ_probes = h[5]; // the main kernel's own probes
    idx->stats.fallback_reads = h[0];
    const uint64_t n_wave = h[6];
    idx->stats.diag[0] = n_wave; // reads the workgroup-per-read pass takes
    uint64_t nlist = h[3], used = h[2];
    if (nlist) {
        // reads that outgrew their wave's set space or found the output list
        // full (rare): keep the wave pass's read list for the row-order sort,
        // grow both and re-run only those reads, until none is left
        if ((rc = ws.status.ensure(n_wave * 4 + 4))) return rc;
        HSFM_HIP(hipMemcpyAsync(ws.status.p, ws.read_list.p, n_wave * 4, hipMemcpyDeviceToDevice, s));
        const uint32_t *list = ws.ovf_list.as<uint32_t>();
        uint32_t *retry = ws.read_list.as<uint32_t>();
        while (nlist) {
            idx->stats.capacity_retries++;
            if (cap >= (1u << 28))
                return set_error(HSFM_ENOMEM, "interval set of a single read exceeds 2^28 intervals");
            cap *= 4;
            const uint64_t want = used + nlist * 1024;
            const uint64_t grown_cap = fb_cap * 2 > want ? fb_cap * 2 : want;
            // per-wave set space is 2 * cap intervals; keep the total under 4 GiB
            uint64_t waves = nlist < (uint64_t)wgrid ? nlist : (uint64_t)wgrid;
            const uint64_t wave_limit = (4ull << 30) / (2ull * cap * 8);
            if (waves > wave_limit) waves = wave_limit;
            if (waves == 0) waves = 1;
            if ((rc = ws.wave_sets.ensure((size_t)waves * 2 * cap * 8))) return rc;
            if (ws.fb_ivs.cap < grown_cap * 8) {
                // keep the intervals already published by earlier passes
                DevBuf grown;
                if ((rc = grown.ensure(grown_cap * 8))) return rc;
                if (ws.fb_ivs.p) HSFM_HIP(hipMemcpyAsync(grown.p, ws.fb_ivs.p, ws.fb_ivs.cap, hipMemcpyDeviceToDevice, s));
                HSFM_HIP(hipStreamSynchronize(s));
                ws.fb_ivs.release();
                ws.fb_ivs = grown;
            }
            fb_cap = ws.fb_ivs.cap / 8;
            if ((rc = fill_u32(idx, reinterpret_cast<uint32_t *>(&ctr->work), 2, 0))) return rc;
            if ((rc = fill_u32(idx, reinterpret_cast<uint32_t *>(&ctr->n_retry), 2, 0))) return rc;
            hipLaunchKernelGGL(k_locate_wave, dim3((uint32_t)waves), dim3(64), 0, s, d, idx->d_const.as<DevConst>(),
                               d_q, d_qoff, list, (uint32_t)nlist, cap, ws.wave_sets.as<uint2>(),
                               ws.cnt.as<uint32_t>(), ws.fb_meta.as<uint2>(), ws.fb_ivs.as<uint2>(), fb_cap, retry,
                               ctr, nullptr, idx->seed_rec12.as<uint4>(), idx->seed_iv.as<uint4>(),
                               idx->seed_rec12.p ? 12u : 0u);
            HSFM_HIP(hipGetLastError());
            if ((rc = readback_u64(idx, reinterpret_cast<const uint64_t *>(ctr), h, 8))) return rc;
            nlist = h[3];
            used = h[2];
            const uint32_t *next_list = retry;
            retry = const_cast<uint32_t *>(list);
            list = next_list;
        }
        if (total) { // retried reads changed their counts
            if ((rc = exclusive_scan_u32_u64(idx, ws.cnt.as<uint32_t>(), ws.hit_off.as<uint64_t>(), nq))) return rc;
            if ((rc = readback_u64(idx, ws.hit_off.as<uint64_t>() + nq, total, 1))) return rc;
        }
    }
    if (n_wave) {
        // put every wave-path read's intervals in row order (FmIndex::locate order)
        const uint32_t nfb = (uint32_t)n_wave;
        const uint32_t *wave_reads = idx->stats.capacity_retries ? ws.status.as<uint32_t>() : ws.read_list.as<uint32_t>();
        if ((rc = ws.fb_sets.ensure(used * 16 + (uint64_t)nfb * 8 + 64))) return rc;
        unsigned long long *k_in = ws.fb_sets.as<unsigned long long>();
        unsigned long long *k_out = k_in + used;
        uint32_t *seg_b = reinterpret_cast<uint32_t *>(k_out + used);
        uint32_t *seg_e = seg_b + nfb;
        hipLaunchKernelGGL(k_fb_segments, dim3((nfb + 255) / 256), dim3(256), 0, s, wave_reads, nfb,
                           ws.fb_meta.as<uint2>(), seg_b, seg_e);
        HSFM_HIP(hipGetLastError());
        const uint32_t kb = (uint32_t)((used + 255) / 256 < 65536 ? (used + 255) / 256 : 65536);
        hipLaunchKernelGGL(k_fb_keys, dim3(kb ? kb : 1), dim3(256), 0, s, ws.fb_ivs.as<uint2>(), used, k_in);
        HSFM_HIP(hipGetLastError());
        size_t tmp = 0;
        HSFM_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, k_in, k_out, (int)used, (int)nfb, seg_b,
                                                             seg_e, 0, 64, s));
        if ((rc = ws.scan_tmp.ensure(tmp))) return rc;
        // intervals outside the sorted segments (already in row order) pass through
        HSFM_HIP(hipMemcpyAsync(k_out, k_in, used * 8, hipMemcpyDeviceToDevice, s));
        HSFM_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(ws.scan_tmp.p, tmp, k_in, k_out, (int)used, (int)nfb,
                                                             seg_b, seg_e, 0, 64, s));
        hipLaunchKernelGGL(k_fb_unkeys, dim3(kb ? kb : 1), dim3(256), 0, s, k_out, used, ws.fb_ivs.as<uint2>());
        HSFM_HIP(hipGetLastError());
    }
    if (idx->stats.fallback_reads && (rc = event_record(idx, 2, true))) return rc;
    return HSFM_OK;
}

} // namespace hsfm

using namespace hsfm;

static int locate_device_impl(hsfm_index *idx, const uint8_t *d_q, const uint64_t *d_qoff, uint32_t nq,
                              hsfm_locate_device_result *out) {
    Workspace &ws = idx->ws;
    hipStream_t s = idx->dev->stream;
    idx->stats = hsfm_stats{};
    idx->defer_valid = false;
    int rc = ensure_sa_full(idx);
    if (rc) return rc;
    uint64_t total = 0;
    if ((rc = locate_search_phase(idx, d_q, d_qoff, nq, &total))) return rc;
    if ((rc = ws.rows.ensure(total * 4 + 4))) return rc;
    if ((rc = ws.hits.ensure(total * 8 + 8))) return rc;
    LocateCounters *ctr = ws.counters.as<LocateCounters>();
    if (total) {
        hipLaunchKernelGGL(k_expand, dim3((nq + 255) / 256), dim3(256), 0, s, nq, ws.niv.as<uint8_t>(),
                           ws.ivs.as<uint2>(), ws.fb_meta.as<uint2>(), ws.fb_ivs.as<uint2>(),
                           ws.hit_off.as<uint64_t>(), ws.rows.as<uint32_t>());
        HSFM_HIP(hipGetLastError());
        if ((rc = launch_resolve(idx, idx->fwd.view, ws.rows.as<uint32_t>(), total, idx->seq_bounds.as<uint32_t>(),
                                 idx->nseq, ws.hits.as<uint32_t>(), &ctr->guard)))
            return rc;
    }
    // resolve from the full suffix array makes no LF walk: no guard to read back
    uint64_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (total && !idx->fwd.view.sa_full) {
        if ((rc = readback_u64(idx, reinterpret_cast<const uint64_t *>(ctr), h, 8))) return rc;
    } else {
        HSFM_HIP(hipStreamSynchronize(s)); // results complete on return
    }
    idx->stats.guard_hits = h[4];
    idx->stats.main_kernel_ms = event_ms(idx, 0);
    idx->stats.resolve_kernel_ms = total ? event_ms(idx, 1) : 0.0;
    idx->stats.fallback_kernel_ms = idx->stats.fallback_reads ? event_ms(idx, 2) : 0.0;
    if (h[4]) return set_error(HSFM_EINDEX, "%llu LF walks exceeded the sample rate: index is inconsistent", (unsigned long long)h[4]);
    out->num_queries = nq;
    out->total_hits = total;
    out->d_offsets = ws.hit_off.as<uint64_t>();
    out->d_hits = ws.hits.as<uint32_t>();
    return HSFM_OK;
}

static int stage_queries(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq) {
    Workspace &ws = idx->ws;
    hipStream_t s = idx->dev->stream;
    const uint64_t tot = qoff[nq];
    for (uint32_t i = 0; i < nq; i++)
        if (qoff[i + 1] < qoff[i]) return set_error(HSFM_EINVAL, "query offsets not monotone at %u", i);
    int rc;
    if ((rc = ws.qcopy.ensure(tot + 16))) return rc;
    if ((rc = ws.qoffcopy.ensure(((size_t)nq + 1) * 8))) return rc;
    if (tot) HSFM_HIP(hipMemcpyAsync(ws.qcopy.p, qflat, tot, hipMemcpyHostToDevice, s));
    HSFM_HIP(hipMemcpyAsync(ws.qoffcopy.p, qoff, ((size_t)nq + 1) * 8, hipMemcpyHostToDevice, s));
    return HSFM_OK;
}

extern "C" int hsfm_locate_device(hsfm_index *idx, const uint8_t *d_qflat, const uint64_t *d_qoff, uint32_t nq,
                                  hsfm_locate_device_result *out) {
    if (!idx || !out || (nq && (!d_qflat || !d_qoff))) return set_error(HSFM_EINVAL, "hsfm_locate_device: null argument");
    HSFM_LOCK_INDEX(idx, "hsfm_locate_device");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    return locate_device_impl(idx, d_qflat, d_qoff, nq, out);
}

extern "C" int hsfm_locate_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                 hsfm_locate_result **out) {
    if (!idx || !out || !qoff || (nq && !qflat)) return set_error(HSFM_EINVAL, "hsfm_locate_batch: null argument");
    HSFM_LOCK_INDEX(idx, "hsfm_locate_batch");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    int rc = stage_queries(idx, qflat, qoff, nq);
    if (rc) return rc;
    hsfm_locate_device_result dr;
    if ((rc = locate_device_impl(idx, idx->ws.qcopy.as<uint8_t>(), idx->ws.qoffcopy.as<uint64_t>(), nq, &dr))) return rc;
    const size_t hdr = (sizeof(hsfm_locate_result) + 15) & ~(size_t)15;
    const size_t bytes = hdr + ((size_t)nq + 1) * 8 + dr.total_hits * 8;
    char *blk = (char *)malloc(bytes);
    if (!blk) return set_error(HSFM_ENOMEM, "host allocation of %zu bytes failed", bytes);
    hsfm_locate_result *res = (hsfm_locate_result *)blk;
    res->num_queries = nq;
    res->offsets = (uint64_t *)(blk + hdr);
    res->hits = (uint32_t *)(blk + hdr + ((size_t)nq + 1) * 8);
    hipStream_t s = idx->dev->stream;
    hipError_t e = hipMemcpyAsync(res->offsets, dr.d_offsets, ((size_t)nq + 1) * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && dr.total_hits)
        e = hipMemcpyAsync(res->hits, dr.d_hits, dr.total_hits * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free(blk);
        return set_error(HSFM_EDEVICE, "copy of locate results failed: %s", hipGetErrorString(e));
    }
    *out = res;
    return HSFM_OK;
}

extern "C" int hsfm_locate_positions_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                           hsfm_positions_result **out) {
    if (!idx || !out || !qoff || (nq && !qflat))
        return set_error(HSFM_EINVAL, "hsfm_locate_positions_batch: null argument");
    HSFM_LOCK_INDEX(idx, "hsfm_locate_positions_batch");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    int rc = stage_queries(idx, qflat, qoff, nq);
    if (rc) return rc;
    hsfm_locate_device_result dr;
    if ((rc = locate_device_impl(idx, idx->ws.qcopy.as<uint8_t>(), idx->ws.qoffcopy.as<uint64_t>(), nq, &dr))) return rc;
    const size_t hdr = (sizeof(hsfm_positions_result) + 15) & ~(size_t)15;
    const size_t bytes = hdr + ((size_t)nq + 1) * 8 + dr.total_hits * 8;
    char *blk = (char *)malloc(bytes);
    if (!blk) return set_error(HSFM_ENOMEM, "host allocation of %zu bytes failed", bytes);
    hsfm_positions_result *res = (hsfm_positions_result *)blk;
    res->num_queries = nq;
    res->offsets = (uint64_t *)(blk + hdr);
    // (seq_idx, offset) pairs land behind the offsets, then fold in place into
    // absolute positions: the start of sequence s is the end of sequence s - 1
    // (seq_boundaries are cumulative exclusive ends, sentinels included)
    uint32_t *pairs = (uint32_t *)(blk + hdr + ((size_t)nq + 1) * 8);
    res->positions = pairs;
    std::vector<uint32_t> ends(idx->nseq);
    hipStream_t s = idx->dev->stream;
    hipError_t e = hipMemcpyAsync(res->offsets, dr.d_offsets, ((size_t)nq + 1) * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && dr.total_hits)
        e = hipMemcpyAsync(pairs, dr.d_hits, dr.total_hits * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && idx->nseq)
        e = hipMemcpyAsync(ends.data(), idx->seq_bounds.p, (size_t)idx->nseq * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free(blk);
        return set_error(HSFM_EDEVICE, "copy of locate results failed: %s", hipGetErrorString(e));
    }
    for (uint64_t t = 0; t < dr.total_hits; t++) {
        const uint32_t sq = pairs[2 * t], off = pairs[2 * t + 1];
        const uint32_t start = sq == 0 || sq > idx->nseq ? 0u : ends[sq - 1];
        pairs[t] = start + off; // t <= 2t: each write lands on a pair already read
    }
    *out = res;
    return HSFM_OK;
}

extern "C" int hsfm_count_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                uint32_t *counts) {
    if (!idx || !counts || !qoff || (nq && !qflat)) return set_error(HSFM_EINVAL, "hsfm_count_batch: null argument");
    HSFM_LOCK_INDEX(idx, "hsfm_count_batch");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    idx->stats = hsfm_stats{};
    idx->defer_valid = false;
    int rc = stage_queries(idx, qflat, qoff, nq);
    if (rc) return rc;
    if ((rc = locate_search_phase(idx, idx->ws.qcopy.as<uint8_t>(), idx->ws.qoffcopy.as<uint64_t>(), nq))) return rc;
    hipStream_t s = idx->dev->stream;
    if (nq) HSFM_HIP(hipMemcpyAsync(counts, idx->ws.cnt.p, (size_t)nq * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    return HSFM_OK;
}
