// hsfm_lsi_dev.h -- device layout and lookup steps of the learned seed index,
// shared by its own kernels (hsfm_lsi.hip) and the MEM kernel's learned-seed
// restarts (hsfm_mem.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsfm {

constexpr uint32_t kLsiMaxHidden = 64;
// Fence level v holds every 2^(5 + 4v)-th key.  A range of at most 2^(9 + 4v)
// + 1 keys spans at most 17 entries of level v + 1 ... of level v, and level
// 0 leaves at most 33 keys.  Slots are u32, so a window of 2 max_err + 1 <=
// 2^33 keys spans at most 17 entries of level 6 (stride 2^29): seven levels.
constexpr int kLsiMaxFences = 7;
__host__ __device__ constexpr int lsi_fence_shift(int v) { return 5 + 4 * v; }
// Fence levels needed for a largest per-leaf error emax: the coarsest level v
// must span the widest window in at most 17 entries, (2 emax) >> shift <= 16;
// levels 0 and 1 are always built.
inline int lsi_fence_levels(uint64_t emax) {
    int nl = 2;
    while (nl < kLsiMaxFences && ((2ull * emax) >> lsi_fence_shift(nl - 1)) > 16) nl++;
    return nl;
}
struct LsiFences {
    const unsigned long long *f[kLsiMaxFences];
};

struct LsiLeaf {   // 16 bytes
    uint32_t s0;   // first slot of the leaf
    uint32_t err;  // max |pred - slot| over the leaf's keys
    float a, b;    // slot - s0 ~ a * f(u) + b
};

struct LsiModel {
    const float *p; // stage 1: w1[H], b1[H], w2[H], b2
    uint32_t hidden, leaves, k;
    uint64_t nkeys;
    uint32_t nfence;  // fence levels in use (2..kLsiMaxFences)
    uint32_t ubits;   // key bits below the leaf index (2k - log2 L, or 0)
    uint32_t upshift; // log2 L - 2k when the leaves outnumber the keys' bits
};

// key / 4^k in [0, 1) in FP32 (the stage-1 training feature)
__device__ __forceinline__ float lsi_feature(unsigned long long key, uint32_t k) {
    const int sh = (int)(2 * k) - 24;
    return sh > 0 ? (float)(uint32_t)(key >> sh) * 5.9604644775390625e-8f   // 2^-24
                  : (float)(uint32_t)key * ldexpf(1.f, -(int)(2 * k));
}
__device__ __forceinline__ float lsi_stage1(const float *p, uint32_t H, float x) {
    float f = p[3 * H];
    for (uint32_t i = 0; i < H; i++) f = fmaf(p[2 * H + i], fmaxf(fmaf(p[i], x, p[H + i]), 0.f), f);
    return f;
}
__device__ __forceinline__ uint32_t lsi_leaf(const LsiModel &md, unsigned long long key) {
    return (uint32_t)((key >> md.ubits) << md.upshift);
}
// position of key inside its leaf, in [0, 1) with 24 bits
__device__ __forceinline__ float lsi_u(const LsiModel &md, unsigned long long key) {
    if (md.ubits == 0) return 0.f;
    const unsigned long long v = key & ((1ull << md.ubits) - 1ull);
    return md.ubits > 24 ? (float)(uint32_t)(v >> (md.ubits - 24)) * 5.9604644775390625e-8f
                         : (float)(uint32_t)v * ldexpf(1.f, -(int)md.ubits);
}
// predicted slot of key (identical in the error pass and lookups)
__device__ __forceinline__ long long lsi_pred(const LsiLeaf &lf, float fu) {
    return (long long)lf.s0 + (long long)floorf(fmaf(lf.a, fu, lf.b));
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

// Slot of key among the sorted keys, or -1 if it is absent: the bounded
// search of a lookup (fence rounds, then the last 33 keys).
__device__ __forceinline__ long long lsi_find_slot(const LsiModel &md, const unsigned long long *__restrict__ keys,
                                                   const LsiFences &fc, const LsiLeaf &f, float fu,
                                                   unsigned long long key, uint32_t &probes) {
    const long long pred = lsi_pred(f, fu);
    long long lo = pred - (long long)f.err, hi = pred + (long long)f.err; // inclusive window
    lo = lo < 0 ? 0 : lo;
    hi = hi >= (long long)md.nkeys ? (long long)md.nkeys - 1 : hi;
    // lower bound of key in [lo, hi].  At each fence level (stride 2^sh
    // keys) the last entry j in (lo >> sh, hi >> sh] below key, else
    // lo >> sh, leaves the lower bound in [max(lo, j << sh), min(hi, (j + 1)
    // << sh)] or just past hi.  The window spans at most 17 entries of the
    // coarsest level in use (the build picks the level count so); each level
    // then leaves at most 17 entries of the next and level 0 at most 33 keys,
    // read as 16-byte pairs in one round.
    constexpr int kPairs = 17; // 33 keys from an even index: 17 pairs
    if (lo <= hi) {
#pragma unroll
        for (int lv = kLsiMaxFences - 1; lv >= 0; lv--) {
            if (lv >= (int)md.nfence) continue;
            const int sh = lsi_fence_shift(lv);
            const long long j = lsi_last_below(fc.f[lv], (lo >> sh) + 1, hi >> sh, lo >> sh, key, probes);
            lo = lo > (j << sh) ? lo : (j << sh);
            hi = hi < ((j + 1) << sh) ? hi : ((j + 1) << sh);
        }
    }
    if (lo > hi) return -1;
    const long long b = lo & ~1ll; // pairs from an even index (the key buffer is padded by one key)
    probes += (uint32_t)((((hi - b) / 2 + 1) * 16 + 63) / 64); // the round's bytes in 64-byte units
    const ulonglong2 *kp = reinterpret_cast<const ulonglong2 *>(keys + b);
    uint32_t below = 0;
    bool hit = false;
#pragma unroll
    for (int u = 0; u < kPairs; u++) {
        const long long a0 = b + 2 * u;
        ulonglong2 v = make_ulonglong2(~0ull, ~0ull);
        if (a0 <= hi) v = kp[u];
        const bool in0 = a0 >= lo && a0 <= hi, in1 = a0 + 1 >= lo && a0 + 1 <= hi;
        below += (in0 && v.x < key ? 1u : 0u) + (in1 && v.y < key ? 1u : 0u);
        hit |= (in0 && v.x == key) || (in1 && v.y == key);
    }
    return hit ? lo + below : -1;
}

// Everything a MEM kernel needs to run learned-seed lookups (passed by value).
struct LsiView {
    LsiModel md;
    LsiFences fc;
    const unsigned long long *keys; // sorted present k-mers, padded by one key
    const uint4 *pay;               // {fwd_lo, fwd_hi, rev_lo, flags}; flags bit 0: the k-mer is
                                    // compatible with a text window holding a non-ACGT symbol
    const LsiLeaf *leaves;
};
constexpr uint32_t kLsiFlagAmbiguous = 1u;

} // namespace hsfm
