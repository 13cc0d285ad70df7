// hsfm_lsi_dev.h -- device layout and lookup steps of the learned seed index,
// shared by its own kernels (hsfm_lsi.hip) and the MEM kernel's learned-seed
// restarts (hsfm_mem.hip).
//
// Entries are sorted 64-bit keys.  A k-mer's key holds its codes two bits per
// base (A=0 .. T=3) from the top of the word down, LAST base first, so the
// sorted order is the order of the reversed k-mers and k-mers that share a
// suffix are neighbours; the low 64 - 2k bits are entry metadata:
//   full entry  : a text window of k codes.  Its payload is the bidirectional
//                 interval of that window; bit 0 set when the window holds an
//                 ambiguity code (the key is then one of the window's
//                 expansions into bases), clear for an exact-ACGT window.
//   short entry : a window of r < k codes cut on its left by a contig start or
//                 a symbol no base is compatible with; its key holds r codes
//                 (zeros below), kLsiShort is set, and G / H below give the
//                 neighbour bounds of the longest-suffix answer (no payload).
// Among equal codes, exact entries sort first, then ambiguity entries, then
// short ones.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsfm {

constexpr uint32_t kLsiMaxHidden = 64;
// Fence level v holds every 2^(5 + 4v)-th key.  A range of at most 2^(9 + 4v)
// + 1 keys spans at most 17 entries of level v + 1 ... of level v, and level
// 0 leaves at most 33 keys.  Slots are u32, so a window of 2 max_err + 2 <=
// 2^33 keys spans at most 17 entries of level 6 (stride 2^29): seven levels.
constexpr int kLsiMaxFences = 7;
__host__ __device__ constexpr int lsi_fence_shift(int v) { return 5 + 4 * v; }
// Fence levels needed for a largest per-leaf error emax: none when the widest
// range [lo, hi] (hi - lo <= 2 emax + 1) already fits the last round (hi - lo
// <= 32), else as many as make the coarsest level v span it in at most 17
// entries, (2 emax + 1) >> shift(v) <= 16.
inline int lsi_fence_levels(uint64_t emax) {
    if (2ull * emax + 1ull <= 32ull) return 0;
    int nl = 1;
    while (nl < kLsiMaxFences && ((2ull * emax + 1ull) >> lsi_fence_shift(nl - 1)) > 16) nl++;
    return nl;
}
struct LsiFences {
    const unsigned long long *f[kLsiMaxFences];
};

// metadata bits (k <= kLsiMemMaxK leaves at least 14 of them)
constexpr uint32_t kLsiMemMaxK = 25;
constexpr unsigned long long kLsiVariant = 1ull;       // full entry of an ambiguity-code window
constexpr unsigned long long kLsiShort = 1ull << 13;   // short entry
constexpr int kLsiGShift = 8, kLsiHShift = 3;          // 5-bit G and H of a short entry
// payload .w: bit 0 = exact entry; bits 8.. = on the first full entry of a
// run of equal codes, the number of full entries in the run
constexpr uint32_t kLsiPayExact = 1u;

struct LsiLeaf {   // 16 bytes
    uint32_t s0;   // first slot of the leaf (of the next leaf for an empty one)
    uint32_t err;  // max |pred - slot| over the leaf's entries
    float a, b;    // slot - s0 ~ a * f(u) + b, a >= 0
};

struct LsiModel {
    const float *p; // stage 1: w1[H], b1[H], w2[H], b2 (w1, w2 >= 0: f is non-decreasing)
                    // of H clipped ramps
    uint32_t hidden, leaves, k;
    uint64_t nkeys;           // entries
    uint32_t nfence;          // fence levels in use (0..kLsiMaxFences)
    uint32_t lbits;           // log2 leaves: leaf = the key's top lbits bits
    unsigned long long smask; // the code bits of a key (top 2k bits)
};

// 2-bit digit reversal: a k-mer packed first base most significant in the low
// 2k bits -> its key (last base in the top two bits)
__host__ __device__ __forceinline__ unsigned long long lsi_key_of_fwd(unsigned long long v) {
    const unsigned long long y = __builtin_bitreverse64(v);
    return ((y >> 1) & 0x5555555555555555ull) | ((y & 0x5555555555555555ull) << 1);
}

// f(x) = b2 + sum_i w2_i clamp(w1_i x + b1_i, 0, 1): with w1, w2 >= 0 every
// step (fma, clamp, fma) is non-decreasing in x, rounding included
__device__ __forceinline__ float lsi_stage1(const float *p, uint32_t H, float x) {
    float f = p[3 * H];
    for (uint32_t i = 0; i < H; i++) f = fmaf(p[2 * H + i], fminf(fmaxf(fmaf(p[i], x, p[H + i]), 0.f), 1.f), f);
    return f;
}
__device__ __forceinline__ uint32_t lsi_leaf(const LsiModel &md, unsigned long long key) {
    return md.lbits ? (uint32_t)((key & md.smask) >> (64 - md.lbits)) : 0u;
}
// position of key inside its leaf, in [0, 1) with 24 bits
__device__ __forceinline__ float lsi_u(const LsiModel &md, unsigned long long key) {
    return (float)(uint32_t)(((key & md.smask) << md.lbits) >> 40) * 5.9604644775390625e-8f;
}
// predicted slot of key (identical in the error pass and lookups); non-
// decreasing in the key inside a leaf
__device__ __forceinline__ long long lsi_pred(const LsiLeaf &lf, float fu) {
    return (long long)lf.s0 + (long long)floorf(fmaf(lf.a, fu, lf.b));
}
// Search range of a lookup: the first slot whose key is >= w (the lower bound
// LB) lies in [lo, hi + 1].  For a key x of the leaf at slot s, |pred(x) - s|
// <= err; pred is non-decreasing, so for keys x_i < w <= x_{i+1} of the leaf
// LB = i + 1 is in [pred(w) - err, pred(w) + err + 1], and a w below (above)
// every key of its leaf has LB = s0 (one past the leaf) inside the same range
// once lo is raised to s0.
__device__ __forceinline__ void lsi_range(const LsiModel &md, const LsiLeaf &f, float fu, long long &lo,
                                          long long &hi) {
    const long long pred = lsi_pred(f, fu), N = (long long)md.nkeys;
    lo = pred - (long long)f.err;
    lo = lo < (long long)f.s0 ? (long long)f.s0 : lo;
    lo = lo > N ? N : lo;
    hi = pred + (long long)f.err + 1;
    hi = hi > N - 1 ? N - 1 : hi;
    hi = hi < lo - 1 ? lo - 1 : hi;
}

// Last e in [a, z] with f[e] < key, or j if there is none (f ascending), for
// a range of at most 17 entries: one round of 16-byte pairs from an even index
// (f holds one entry past z).  loads += 64-byte units.
__device__ __forceinline__ long long lsi_last_below(const unsigned long long *__restrict__ f, long long a,
                                                    long long z, long long j, unsigned long long key,
                                                    uint32_t &loads) {
    constexpr int kPairs = 9;
    if (a <= z) {
        const long long b = a & ~1ll;
        const ulonglong2 *fp = reinterpret_cast<const ulonglong2 *>(f + b);
        uint32_t c = 0;
#pragma unroll
        for (int u = 0; u < kPairs; u++) {
            const long long e = b + 2 * u;
            ulonglong2 v = make_ulonglong2(~0ull, ~0ull);
            if (e <= z) v = fp[u];
            c += (e >= a && e <= z && v.x < key ? 1u : 0u) + (e + 1 >= a && e + 1 <= z && v.y < key ? 1u : 0u);
        }
        loads += (uint32_t)((((z - b) / 2 + 1) * 16 + 63) / 64);
        if (c) j = a + c - 1;
    }
    return j;
}

// One fence level: narrows [lo, hi] keeping LB in [lo, hi + 1].
__device__ __forceinline__ void lsi_fence_step(const unsigned long long *__restrict__ f, int sh, long long &lo,
                                               long long &hi, unsigned long long w, uint32_t &loads) {
    const long long j = lsi_last_below(f, (lo >> sh) + 1, hi >> sh, lo >> sh, w, loads);
    lo = lo > (j << sh) ? lo : (j << sh);
    hi = hi < ((j + 1) << sh) ? hi : ((j + 1) << sh);
}

// Keys of the last round: slots [max(lo - 1, 0), min(hi + 1, N - 1)] from the
// even slot b, nine 16-byte pairs from b + 18 part.  Accumulates the number
// of keys below w, the last of them (pv) and the first key >= w (sv, hs).
__device__ __forceinline__ void lsi_keys_part(const unsigned long long *__restrict__ keys, long long first,
                                              long long last, long long b, unsigned long long w, uint32_t &below,
                                              unsigned long long &pv, unsigned long long &sv, bool &hs,
                                              uint32_t &loads) {
    const ulonglong2 *kp = reinterpret_cast<const ulonglong2 *>(keys + b);
#pragma unroll
    for (int u = 0; u < 9; u++) {
        const long long a0 = b + 2 * u;
        ulonglong2 v = make_ulonglong2(~0ull, ~0ull);
        if (a0 <= last) v = kp[u];
        const bool in0 = a0 >= first && a0 <= last, in1 = a0 + 1 >= first && a0 + 1 <= last;
        if (in0 && v.x < w) { below++; pv = v.x; }
        if (in0 && v.x >= w && !hs) { sv = v.x; hs = true; }
        if (in1 && v.y < w) { below++; pv = v.y; }
        if (in1 && v.y >= w && !hs) { sv = v.y; hs = true; }
    }
    const long long end = b + 17 < last ? b + 17 : last;
    if (end >= b) loads += (uint32_t)((((end - b) / 2 + 1) * 16 + 63) / 64);
}

// Code digits two keys share from the top (k when the codes are equal).
__device__ __forceinline__ uint32_t lsi_lcs(const LsiModel &md, unsigned long long a, unsigned long long b) {
    const unsigned long long x = (a ^ b) & md.smask;
    const uint32_t c = x ? (uint32_t)__clzll((long long)x) >> 1 : 32u;
    return c < md.k ? c : md.k;
}
// G / H bound of an entry (k for a full entry)
__device__ __forceinline__ uint32_t lsi_gbound(const LsiModel &md, unsigned long long e) {
    return (e & kLsiShort) ? (uint32_t)(e >> kLsiGShift) & 31u : md.k;
}
__device__ __forceinline__ uint32_t lsi_hbound(const LsiModel &md, unsigned long long e) {
    return (e & kLsiShort) ? (uint32_t)(e >> kLsiHShift) & 31u : md.k;
}
// The entry at LB is a full entry with w's codes.
__device__ __forceinline__ bool lsi_present(const LsiModel &md, unsigned long long w, unsigned long long sv, bool hs) {
    return hs && (sv & md.smask) == w && !(sv & kLsiShort);
}
// Longest suffix of an absent window w that occurs compatibly in the text:
// the largest min(r, common suffix) over every entry, which the neighbours of
// w give through their G / H bounds.
__device__ __forceinline__ uint32_t lsi_suffix_len(const LsiModel &md, unsigned long long w, unsigned long long pv,
                                                   bool hp, unsigned long long sv, bool hs) {
    uint32_t a = 0, b = 0;
    if (hp) {
        const uint32_t l = lsi_lcs(md, w, pv), g = lsi_gbound(md, pv);
        a = l < g ? l : g;
    }
    if (hs) {
        const uint32_t l = lsi_lcs(md, w, sv), h = lsi_hbound(md, sv);
        b = l < h ? l : h;
    }
    return a > b ? a : b;
}

// The whole search of one key (fence rounds, then the last round): LB and
// its neighbours.
__device__ __forceinline__ long long lsi_search(const LsiModel &md, const unsigned long long *__restrict__ keys,
                                                const LsiFences &fc, const LsiLeaf &f, float fu, unsigned long long w,
                                                uint32_t &probes, unsigned long long &pv, unsigned long long &sv,
                                                bool &hs) {
    long long lo, hi;
    lsi_range(md, f, fu, lo, hi);
    if (lo <= hi) {
#pragma unroll
        for (int lv = kLsiMaxFences - 1; lv >= 0; lv--)
            if (lv < (int)md.nfence) lsi_fence_step(fc.f[lv], lsi_fence_shift(lv), lo, hi, w, probes);
    }
    const long long N = (long long)md.nkeys;
    const long long first = lo > 0 ? lo - 1 : 0, last = hi + 1 < N - 1 ? hi + 1 : N - 1;
    uint32_t below = 0;
    hs = false;
    pv = 0;
    sv = ~0ull;
    const long long b = first & ~1ll;
    lsi_keys_part(keys, first, last, b, w, below, pv, sv, hs, probes);
    if (!hs && b + 18 <= last) lsi_keys_part(keys, first, last, b + 18, w, below, pv, sv, hs, probes);
    return first + below;
}

// Everything a MEM kernel needs to run learned-seed lookups (passed by value).
struct LsiView {
    LsiModel md;
    LsiFences fc;
    const unsigned long long *keys; // sorted entries, padded by two keys
    const uint4 *pay;               // {fwd_lo, fwd_hi, rev_lo, flags | run << 8} per entry
    const LsiLeaf *leaves;
};

} // namespace hsfm
