// hsfm_lsi.hip -- learned seed index on gfx950.
//
// For an exact-ACGT k-mer W (k <= 31; 21 by default) the index answers the
// bidirectional SA interval {fwd_lo, fwd_hi, rev_lo} that plain backward
// search gives (extend_left over W from the full interval, bidir.rs:112-130,
// the exact-code extension of LookupTable::build, lookup.rs:37-91), or "absent".
//
// Keys pack W two bits per base, first base most significant, so sorting the
// keys sorts the k-mers in suffix-array order.  Build, all on the device:
//   1. enumerate every present k-mer from the index itself, level by level:
//      the node of cS is one exact left extension of the node of S; children
//      are placed digit-major, which keeps each level sorted by key
//   2. stage 1 (FP32 MLP, x -> f(x) ~ slot / N, the key CDF) trained by
//      Adam on a sample of (key, slot) pairs with the forward/backward kernels
//      below
//   3. stage 2: leaf = the key's leading log2(L) bits (L a power of two), so
//      leaves are aligned with k-mer prefixes; per leaf a least-squares map
//      slot ~ s0 + a * f(u) + b of the stage-1 output at u, the key's
//      position inside its leaf in [0, 1) (float64 sums, FP32 parameters),
//      and the maximum error of its keys.  The key distribution is close to
//      a product of per-base frequencies, so the CDF inside a prefix has the
//      shape of the global CDF: the trained f bends each leaf's map along it
//      (an untrained f(x) = x leaves one straight line per leaf)
// Lookup: route, predict, then a binary search of the sorted keys restricted
// to [pred - err, pred + err] through fence levels of every 2^5, 2^9, 2^13 ...
// -th key (as many as the widest window needs; each level narrows the range
// to 17 entries of the next) and a last round of at most 33 keys, each level
// one round of independent loads; exact because err bounds every present key
// and an absent key is absent from every window.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>

#include "hsfm_host.h"
#include "hsfm_lsi_dev.h"

namespace hsfm {

// ---- 1. enumeration ------------------------------------------------------
// Exact lanes of A, C, G, T (lane_sym inverse); 0xff if the index lacks one.
__device__ __forceinline__ void base_lanes(const DevConst *cst, uint32_t *s_bl) {
    if (threadIdx.x < 4) {
        uint32_t l = 0xffu;
        for (uint32_t j = 0; j < cst->num_lanes; j++)
            if (cst->lane_sym[j] == threadIdx.x + 1) l = j;
        s_bl[threadIdx.x] = l;
    }
}

// Children of every parent node (key of S, {fwd_lo, fwd_hi, rev_lo}): cS for
// c in ACGT.  COUNT: flags[d * P + p] = 1 if child d exists.  Else: write the
// child at off[d][p] + base[d] (digit-major: sorted by key).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_lsi_expand(DevDir fwd, const DevConst *__restrict__ cst,
                                                    const unsigned long long *__restrict__ pkey,
                                                    const uint4 *__restrict__ piv, uint32_t P, uint32_t level,
                                                    uint8_t *__restrict__ flags, const uint32_t *__restrict__ off,
                                                    uint4 base, unsigned long long *__restrict__ ckey,
                                                    uint4 *__restrict__ civ) {
    __shared__ uint32_t s_bl[4], s_clane[16];
    base_lanes(cst, s_bl);
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const uint4 iv = piv[p];
    uint32_t cl[16], ch[16];
    pair_counts(fwd, iv.x, iv.y, cl, ch);
    uint32_t below[17];
    below[0] = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) below[j + 1] = below[j] + (ch[j] - cl[j]);
    const unsigned long long pk = pkey[p];
#pragma unroll
    for (uint32_t d = 0; d < 4; d++) {
        const uint32_t l = s_bl[d];
        uint32_t a = 0, b = 0, acc = 0;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if ((uint32_t)j == l) { a = cl[j]; b = ch[j]; acc = below[j]; }
        const bool ok = l < 16 && b > a;
        if (!WRITE) {
            flags[(size_t)d * P + p] = ok ? 1 : 0;
        } else if (ok) {
            uint32_t cs = 0;
#pragma unroll
            for (int j = 0; j < 16; j++)
                if ((uint32_t)j == l) cs = s_clane[j];
            const uint32_t bd = d == 0 ? base.x : d == 1 ? base.y : d == 2 ? base.z : base.w;
            const size_t at = (size_t)bd + off[(size_t)d * P + p];
            ckey[at] = ((unsigned long long)d << (2 * level)) | pk;
            civ[at] = make_uint4(cs + a, cs + b, iv.z + acc, 0u);
        }
    }
}

struct U8ToU32 {
    __device__ __forceinline__ uint32_t operator()(uint8_t v) const { return v; }
};

// ---- 2. stage-1 training -----------------------------------------------
// Gradient of mean((f - y)^2) over n samples, accumulated into g[3H+1]
// (zeroed by the caller); loss sum into *loss.
__global__ __launch_bounds__(256) void k_lsi_grad(const float *__restrict__ p, uint32_t H,
                                                  const float *__restrict__ x, const float *__restrict__ y,
                                                  uint32_t n, float *__restrict__ g, float *__restrict__ loss) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    __shared__ float s_g[3 * kLsiMaxHidden + 2];
    const uint32_t np = 3 * H + 1;
    for (uint32_t i = threadIdx.x; i < np + 1; i += blockDim.x) {
        if (i < np) s_p[i] = p[i];
        s_g[i] = 0.f;
    }
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = t < n;
    const float xi = act ? x[t] : 0.f, yi = act ? y[t] : 0.f;
    float f = s_p[3 * H];
    for (uint32_t i = 0; i < H; i++) f = fmaf(s_p[2 * H + i], fmaxf(fmaf(s_p[i], xi, s_p[H + i]), 0.f), f);
    const float e = act ? f - yi : 0.f;
    const float df = 2.f * e / (float)n;
    // per-parameter contributions, reduced over the wave then over the block
    auto reduce_add = [&](float v, uint32_t slot) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&s_g[slot], v);
    };
    for (uint32_t i = 0; i < H; i++) {
        const float pre = fmaf(s_p[i], xi, s_p[H + i]);
        const float h = fmaxf(pre, 0.f);
        const float dh = pre > 0.f ? df * s_p[2 * H + i] : 0.f;
        reduce_add(dh * xi, i);          // w1
        reduce_add(dh, H + i);           // b1
        reduce_add(df * h, 2 * H + i);   // w2
    }
    reduce_add(df, 3 * H);               // b2
    reduce_add(e * e, np);               // loss
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) atomicAdd(&g[i], s_g[i]);
    if (threadIdx.x == 0) atomicAdd(loss, s_g[np]);
}

__global__ void k_lsi_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, uint32_t np, float lr, float b1t, float b2t) {
    const uint32_t i = threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + 1e-8f);
}

// training sample: every stride-th slot
__global__ void k_lsi_sample(const unsigned long long *__restrict__ keys, uint64_t nkeys, uint64_t stride,
                             uint32_t ns, uint32_t k, float *__restrict__ x, float *__restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const uint64_t s = (uint64_t)i * stride;
    x[i] = lsi_feature(keys[s], k);
    y[i] = (float)((double)s / (double)nkeys);
}

// ---- 3. stage 2 ----------------------------------------------------------
constexpr uint32_t kLsiChunk = 64; // keys per thread in the leaf passes

__global__ void k_lsi_leaf_init(LsiLeaf *lf, uint32_t *cnt, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    LsiLeaf z;
    z.s0 = 0xffffffffu; z.err = 0; z.a = 0.f; z.b = 0.f;
    lf[i] = z;
    cnt[i] = 0;
}

// per leaf first slot and key count (chunk-aggregated atomics; keys sorted)
__global__ __launch_bounds__(256) void k_lsi_route(LsiModel md, const unsigned long long *__restrict__ keys,
                                                   LsiLeaf *__restrict__ lf, uint32_t *__restrict__ cnt) {
    const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLsiChunk;
    uint32_t cur = 0xffffffffu, c = 0, s0 = 0;
    for (uint64_t s = c0; s < c0 + kLsiChunk && s < md.nkeys; s++) {
        const uint32_t l = lsi_leaf(md, keys[s]);
        if (l != cur) {
            if (c) {
                atomicMin(&lf[cur].s0, s0);
                atomicAdd(&cnt[cur], c);
            }
            cur = l; c = 0; s0 = (uint32_t)s;
        }
        c++;
    }
    if (c) {
        atomicMin(&lf[cur].s0, s0);
        atomicAdd(&cnt[cur], c);
    }
}

// least-squares sums per leaf of (f(u), slot - s0): sx, sy, sxx, sxy (float64)
__global__ __launch_bounds__(256) void k_lsi_moments(LsiModel md, const unsigned long long *__restrict__ keys,
                                                     const LsiLeaf *__restrict__ lf, double *__restrict__ mom) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    __syncthreads();
    const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLsiChunk;
    uint32_t cur = 0xffffffffu;
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (uint64_t s = c0; s < c0 + kLsiChunk && s < md.nkeys; s++) {
        const unsigned long long key = keys[s];
        const uint32_t l = lsi_leaf(md, key);
        if (l != cur) {
            if (cur != 0xffffffffu) {
                atomicAdd(&mom[4 * cur + 0], sx); atomicAdd(&mom[4 * cur + 1], sy);
                atomicAdd(&mom[4 * cur + 2], sxx); atomicAdd(&mom[4 * cur + 3], sxy);
            }
            cur = l; sx = sy = sxx = sxy = 0;
        }
        const double dx = (double)lsi_stage1(s_p, md.hidden, lsi_u(md, key)), dy = (double)(s - lf[l].s0);
        sx += dx; sy += dy; sxx += dx * dx; sxy += dx * dy;
    }
    if (cur != 0xffffffffu) {
        atomicAdd(&mom[4 * cur + 0], sx); atomicAdd(&mom[4 * cur + 1], sy);
        atomicAdd(&mom[4 * cur + 2], sxx); atomicAdd(&mom[4 * cur + 3], sxy);
    }
}

__global__ void k_lsi_solve(LsiLeaf *__restrict__ lf, const uint32_t *__restrict__ cnt,
                            const double *__restrict__ mom, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    LsiLeaf f = lf[i];
    const uint32_t c = cnt[i];
    if (c == 0) {
        f.s0 = 0; f.a = 0.f; f.b = 0.f;
    } else {
        const double n = c, sx = mom[4 * i], sy = mom[4 * i + 1], sxx = mom[4 * i + 2], sxy = mom[4 * i + 3];
        const double den = n * sxx - sx * sx;
        double a = 0, b = sy / n;
        if (c >= 2 && den > 0) {
            a = (n * sxy - sx * sy) / den;
            b = (sy - a * sx) / n;
        }
        if (!isfinite(a) || !isfinite(b)) {
            a = 0;
            b = sy / n;
        }
        f.a = (float)a;
        f.b = (float)b;
    }
    lf[i] = f;
}

__global__ __launch_bounds__(256) void k_lsi_err(LsiModel md, const unsigned long long *__restrict__ keys,
                                                 LsiLeaf *__restrict__ lf) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    __syncthreads();
    const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLsiChunk;
    uint32_t cur = 0xffffffffu, e = 0;
    for (uint64_t s = c0; s < c0 + kLsiChunk && s < md.nkeys; s++) {
        const unsigned long long key = keys[s];
        const uint32_t l = lsi_leaf(md, key);
        if (l != cur) {
            if (cur != 0xffffffffu) atomicMax(&lf[cur].err, e);
            cur = l; e = 0;
        }
        const long long d = lsi_pred(lf[l], lsi_stage1(s_p, md.hidden, lsi_u(md, key))) - (long long)s;
        const uint32_t ad = (uint32_t)(d < 0 ? -d : d);
        e = ad > e ? ad : e;
    }
    if (cur != 0xffffffffu) atomicMax(&lf[cur].err, e);
}

// ---- lookup ----------------------------------------------------------------
// fence[j] = keys[j << sh]; entries past the last key hold the largest key value
__global__ void k_lsi_fence(const unsigned long long *__restrict__ keys, uint64_t nkeys, int sh,
                            unsigned long long *__restrict__ fence, uint64_t nf) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nf) fence[j] = (j << sh) < nkeys ? keys[j << sh] : ~0ull;
}

// out[4i..4i+4) = {fwd_lo, fwd_hi, rev_lo, found}; probes += the bounded
// search's loads in 64-byte units (the bytes of each round of paired loads).
// One key per lane (a quad of lanes per key, loading whole lines together, ran
// 3.5x slower: fewer keys in flight); a final round of 33 keys (16-key fences
// with rounds of at most nine loads, for eight waves per SIMD, ran 8% slower).
__device__ __forceinline__ uint4 lsi_lookup_key(const LsiModel &md, const unsigned long long *__restrict__ keys,
                                                const LsiFences &fc, const uint4 *__restrict__ pay,
                                                const LsiLeaf &f, float fu, unsigned long long key,
                                                uint32_t &probes) {
    const long long slot = lsi_find_slot(md, keys, fc, f, fu, key, probes);
    if (slot < 0) return make_uint4(0, 0, 0, 0);
    const uint4 v = pay[slot];
    return make_uint4(v.x, v.y, v.z, 1u);
}

// Keys compatible with a text window that holds a symbol other than A, C, G,
// T get kLsiFlagAmbiguous in their payload: the IUPAC set of such a k-mer is
// more than its exact interval.  ls.qb[lane] holds the query bases (bit b:
// base b+1) compatible with the lane's symbol; ls.exact[lane] is set for the
// lanes of A, C, G and T.  One thread per text position p holding another
// symbol takes the windows [s, s + k) with s in (p - k, p] whose first such
// symbol is p, expands each (every combination of compatible bases) and
// flags the expansions that are present keys.  ctr[0] counts flagged
// expansions, ctr[1] windows with more than `cap` expansions (left
// unflagged: the flags cannot be trusted then).
constexpr uint32_t kLsiFlagMaxExpansions = 4096; // per window (six N)
constexpr uint32_t kLsiMemMaxK = 25;               // k-mer windows the MEM kernel decodes
struct LsiLaneSets {
    uint8_t qb[16];
    uint8_t exact[16];
};
__global__ __launch_bounds__(256) void k_lsi_flag(LsiModel md, const unsigned long long *__restrict__ keys,
                                                  LsiFences fc, uint4 *__restrict__ pay,
                                                  const LsiLeaf *__restrict__ lf, const uint8_t *__restrict__ text,
                                                  uint32_t n, LsiLaneSets ls, uint32_t cap,
                                                  unsigned long long *__restrict__ ctr) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    __syncthreads();
    const uint32_t K = md.k;
    unsigned long long flagged = 0, over = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t tp = text[p] & 15u;
        if (ls.exact[tp] || ls.qb[tp] == 0) continue; // an exact base, or nothing matches here
        const uint64_t s_lo = p + 1 >= K ? p + 1 - K : 0;
        for (uint64_t s0 = s_lo; s0 <= p && s0 + K <= n; s0++) {
            // the window's first non-exact symbol must be p; no position without
            // compatible bases; expansion count
            bool ok = true;
            uint32_t cnt = 1;
            for (uint32_t i = 0; i < K && ok; i++) {
                const uint32_t t = text[s0 + i] & 15u;
                if (s0 + i < p && !ls.exact[t]) ok = false;
                else if (ls.qb[t] == 0) ok = false;
                else if (cnt <= cap) cnt *= (uint32_t)__popc(ls.qb[t]);
            }
            if (!ok) continue;
            if (cnt > cap) {
                over++;
                continue;
            }
            for (uint32_t e = 0; e < cnt; e++) {
                unsigned long long key = 0;
                uint32_t r = e;
                for (uint32_t i = 0; i < K; i++) { // mixed-radix digit of each position
                    const uint32_t m = ls.qb[text[s0 + i] & 15u];
                    const uint32_t c = (uint32_t)__popc(m);
                    uint32_t pick = r % c, mm = m;
                    r /= c;
                    for (uint32_t z = 0; z < pick; z++) mm &= mm - 1;
                    key = (key << 2) | (unsigned long long)(__ffs(mm) - 1);
                }
                uint32_t pr = 0;
                const long long slot = lsi_find_slot(md, keys, fc, lf[lsi_leaf(md, key)],
                                                     lsi_stage1(s_p, md.hidden, lsi_u(md, key)), key, pr);
                if (slot >= 0) {
                    atomicOr(&pay[slot].w, kLsiFlagAmbiguous);
                    flagged++;
                }
            }
        }
    }
    if (flagged) atomicAdd(&ctr[0], flagged);
    if (over) atomicAdd(&ctr[1], over);
}

__global__ __launch_bounds__(256) void k_lsi_lookup(LsiModel md, const unsigned long long *__restrict__ keys,
                                                    LsiFences fc, const uint4 *__restrict__ pay,
                                                    const LsiLeaf *__restrict__ lf,
                                                    const unsigned long long *__restrict__ q, uint64_t nq,
                                                    uint4 *__restrict__ out, unsigned long long *__restrict__ steps) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t probes = 0;
    if (t < nq && !md.nkeys) out[t] = make_uint4(0, 0, 0, 0);
    if (t < nq && md.nkeys) {
        const unsigned long long key = q[t];
        const LsiLeaf f = lf[lsi_leaf(md, key)];
        out[t] = lsi_lookup_key(md, keys, fc, pay, f, lsi_stage1(s_p, md.hidden, lsi_u(md, key)), key, probes);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) probes += __shfl_xor(probes, d, 64);
    if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps, (unsigned long long)probes);
}

// The same lookup with the leaves in LDS (at most kLsiLdsLeaves): one
// 1024-thread workgroup per CU copies them once and strides over the keys, so
// the leaf costs an LDS read instead of a global round in each key's chain.
constexpr uint32_t kLsiLdsLeaves = 4096;
// keys per call of hsfm_lsi_seed_device_noalloc: one thread each, 256 per block
constexpr uint64_t kLsiNoAllocMaxKeys = (1ull << 32) - 1;
__global__ __launch_bounds__(1024) void k_lsi_lookup_lds(LsiModel md, const unsigned long long *__restrict__ keys,
                                                         LsiFences fc, const uint4 *__restrict__ pay,
                                                         const LsiLeaf *__restrict__ lf,
                                                         const unsigned long long *__restrict__ q, uint64_t nq,
                                                         uint4 *__restrict__ out, unsigned long long *__restrict__ steps) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    __shared__ LsiLeaf s_lf[kLsiLdsLeaves];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    for (uint32_t i = threadIdx.x; i < md.leaves; i += blockDim.x) s_lf[i] = lf[i];
    __syncthreads();
    uint32_t probes = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nq; t += stride) {
        if (!md.nkeys) {
            out[t] = make_uint4(0, 0, 0, 0);
            continue;
        }
        const unsigned long long key = q[t];
        const LsiLeaf f = s_lf[lsi_leaf(md, key)];
        out[t] = lsi_lookup_key(md, keys, fc, pay, f, lsi_stage1(s_p, md.hidden, lsi_u(md, key)), key, probes);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) probes += __shfl_xor(probes, d, 64);
    if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps, (unsigned long long)probes);
}

// Plain exact search of k-mer keys: k exact left extensions from the full
// interval (the reference algorithm the learned index replaces).
__global__ __launch_bounds__(256) void k_kmer_search(DevDir fwd, const DevConst *__restrict__ cst, uint32_t k,
                                                     const unsigned long long *__restrict__ q, uint64_t nq,
                                                     uint4 *__restrict__ out) {
    __shared__ uint32_t s_bl[4], s_clane[16];
    base_lanes(cst, s_bl);
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const unsigned long long key = q[t];
    uint4 iv = make_uint4(0, fwd.n, 0, 0);
    bool found = true;
    for (uint32_t step = 0; step < k && found; step++) {
        const uint32_t d = (uint32_t)(key >> (2 * step)) & 3u; // last base first
        const uint32_t l = s_bl[d];
        uint32_t cl[16], ch[16];
        pair_counts(fwd, iv.x, iv.y, cl, ch);
        uint32_t a = 0, b = 0, acc = 0, cs = 0, run = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if ((uint32_t)j == l) { a = cl[j]; b = ch[j]; acc = run; cs = s_clane[j]; }
            run += ch[j] - cl[j];
        }
        found = l < 16 && b > a;
        iv = make_uint4(cs + a, cs + b, iv.z + acc, 0u);
    }
    out[t] = found ? make_uint4(iv.x, iv.y, iv.z, 1u) : make_uint4(0, 0, 0, 0);
}

} // namespace hsfm

using namespace hsfm;

struct hsfm_lsi {
    hsfm_index *idx = nullptr;
    hsfm_lsi_config cfg{};
    hsfm_lsi_info info{};
    DevBuf keys, pay, leaves, params;
    DevBuf fence[kLsiMaxFences]; // level v: keys[j << lsi_fence_shift(v)] for every j
    uint32_t nfence = 0;         // levels built
    LsiModel model() const {
        uint32_t lbits = 0;
        while ((1u << lbits) < cfg.leaves) lbits++;
        const uint32_t kbits = 2 * cfg.k;
        return LsiModel{params.as<float>(), cfg.hidden, cfg.leaves, cfg.k, info.num_keys, nfence,
                        lbits <= kbits ? kbits - lbits : 0u, lbits > kbits ? lbits - kbits : 0u};
    }
    LsiFences fences() const {
        LsiFences fc{};
        for (int v = 0; v < kLsiMaxFences; v++) fc.f[v] = fence[v].as<unsigned long long>();
        return fc;
    }
};

namespace {

int lsi_scan_u8(hsfm_index *idx, const uint8_t *in, uint32_t *out, uint32_t n) {
    hipStream_t s = idx->dev->stream;
    if (n == 0) return HSFM_OK;
    hipcub::TransformInputIterator<uint32_t, U8ToU32, const uint8_t *> it(in, U8ToU32());
    size_t tmp = 0;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, (int)n, s));
    int rc = idx->ws.scan_tmp.ensure(tmp);
    if (rc) return rc;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(idx->ws.scan_tmp.p, tmp, it, out, (int)n, s));
    return HSFM_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// Enumerate every present exact-ACGT k-mer: sorted keys + {fwd_lo, fwd_hi, rev_lo}.
int lsi_enumerate(hsfm_lsi *x) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    int rc;
    DevBuf pk, pi;
    if ((rc = pk.ensure(8)) || (rc = pi.ensure(16))) return rc;
    const unsigned long long k0 = 0;
    const uint4 root = make_uint4(0, idx->fwd.view.n, 0, 0);
    HSFM_HIP(hipMemcpyAsync(pk.p, &k0, 8, hipMemcpyHostToDevice, s));
    HSFM_HIP(hipMemcpyAsync(pi.p, &root, 16, hipMemcpyHostToDevice, s));
    uint64_t P = 1;
    for (uint32_t level = 0; level < x->cfg.k; level++) {
        if (P > 0x7fffffffull) {
            pk.release(); pi.release();
            return set_error(HSFM_EINVAL, "learned seed index: %llu nodes at one level", (unsigned long long)P);
        }
        DevBuf flags, off, ck, ci;
        if ((rc = flags.ensure(4 * P)) || (rc = off.ensure(4 * P * 4 + 16))) return rc;
        const dim3 grid((uint32_t)((P + 255) / 256));
        hipLaunchKernelGGL(k_lsi_expand<false>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                           pk.as<unsigned long long>(), pi.as<uint4>(), (uint32_t)P, level, flags.as<uint8_t>(),
                           nullptr, make_uint4(0, 0, 0, 0), nullptr, nullptr);
        HSFM_HIP(hipGetLastError());
        uint32_t tot[4];
        uint4 base = make_uint4(0, 0, 0, 0);
        for (int d = 0; d < 4; d++)
            if ((rc = lsi_scan_u8(idx, flags.as<uint8_t>() + (size_t)d * P, off.as<uint32_t>() + (size_t)d * P,
                                  (uint32_t)P)))
                return rc;
        // totals: last offset + last flag of each digit
        for (int d = 0; d < 4; d++) {
            uint32_t lo = 0;
            uint8_t lf = 0;
            HSFM_HIP(hipMemcpyAsync(&lo, off.as<uint32_t>() + (size_t)d * P + P - 1, 4, hipMemcpyDeviceToHost, s));
            HSFM_HIP(hipMemcpyAsync(&lf, flags.as<uint8_t>() + (size_t)d * P + P - 1, 1, hipMemcpyDeviceToHost, s));
            HSFM_HIP(hipStreamSynchronize(s));
            tot[d] = lo + lf;
        }
        base.x = 0; base.y = tot[0]; base.z = tot[0] + tot[1]; base.w = tot[0] + tot[1] + tot[2];
        const uint64_t C = (uint64_t)base.w + tot[3];
        if ((rc = ck.ensure(C * 8 + 8)) || (rc = ci.ensure(C * 16 + 16))) return rc;
        hipLaunchKernelGGL(k_lsi_expand<true>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                           pk.as<unsigned long long>(), pi.as<uint4>(), (uint32_t)P, level, nullptr,
                           off.as<uint32_t>(), base, ck.as<unsigned long long>(), ci.as<uint4>());
        HSFM_HIP(hipGetLastError());
        HSFM_HIP(hipStreamSynchronize(s));
        flags.release(); off.release(); pk.release(); pi.release();
        pk = ck; pi = ci;
        P = C;
        if (P == 0) break;
    }
    x->keys = pk;
    x->pay = pi;
    x->info.num_keys = P;
    return HSFM_OK;
}

int lsi_train_and_fit(hsfm_lsi *x) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    const uint32_t H = x->cfg.hidden, L = x->cfg.leaves, np = 3 * H + 1;
    const uint64_t N = x->info.num_keys;
    int rc;
    // stage-1 init: hidden unit i = relu(x - i/H); output = unit 0 (f(x) = x)
    float hp[3 * kLsiMaxHidden + 1];
    for (uint32_t i = 0; i < H; i++) {
        hp[i] = 1.f;
        hp[H + i] = -(float)i / (float)H;
        hp[2 * H + i] = i == 0 ? 1.f : 0.f;
    }
    hp[3 * H] = 0.f;
    if ((rc = x->params.ensure(np * 4))) return rc;
    HSFM_HIP(hipMemcpyAsync(x->params.p, hp, np * 4, hipMemcpyHostToDevice, s));
    DevBuf xs, ys, gm, lossb;
    const uint32_t ns = (uint32_t)(N < x->cfg.sample ? N : x->cfg.sample);
    if ((rc = xs.ensure((size_t)ns * 4 + 4)) || (rc = ys.ensure((size_t)ns * 4 + 4)) ||
        (rc = gm.ensure(np * 4 * 3)) || (rc = lossb.ensure(8)))
        return rc;
    float *g = gm.as<float>(), *m = g + np, *v = m + np;
    HSFM_HIP(hipMemsetAsync(gm.p, 0, np * 4 * 3, s));
    hipEvent_t e0, e1;
    HSFM_HIP(hipEventCreate(&e0));
    HSFM_HIP(hipEventCreate(&e1));
    HSFM_HIP(hipEventRecord(e0, s));
    if (ns) {
        const uint64_t stride = N / ns;
        hipLaunchKernelGGL(k_lsi_sample, dim3((ns + 255) / 256), dim3(256), 0, s, x->keys.as<unsigned long long>(), N,
                           stride ? stride : 1, ns, x->cfg.k, xs.as<float>(), ys.as<float>());
        HSFM_HIP(hipGetLastError());
        float b1t = 1.f, b2t = 1.f;
        for (uint32_t step = 0; step <= x->cfg.train_steps; step++) {
            HSFM_HIP(hipMemsetAsync(g, 0, np * 4, s));
            HSFM_HIP(hipMemsetAsync(lossb.p, 0, 4, s));
            hipLaunchKernelGGL(k_lsi_grad, dim3((ns + 255) / 256), dim3(256), 0, s, x->params.as<float>(), H,
                               xs.as<float>(), ys.as<float>(), ns, g, lossb.as<float>());
            HSFM_HIP(hipGetLastError());
            if (step == 0 || step == x->cfg.train_steps) {
                float l = 0.f;
                HSFM_HIP(hipMemcpyAsync(&l, lossb.p, 4, hipMemcpyDeviceToHost, s));
                HSFM_HIP(hipStreamSynchronize(s));
                if (step == 0) x->info.mse_initial = l / ns;
                else x->info.mse_final = l / ns;
            }
            if (step == x->cfg.train_steps) break;
            b1t *= 0.9f;
            b2t *= 0.999f;
            hipLaunchKernelGGL(k_lsi_adam, dim3(1), dim3(256), 0, s, x->params.as<float>(), m, v, g, np, x->cfg.lr,
                               b1t, b2t);
            HSFM_HIP(hipGetLastError());
        }
        if (x->cfg.train_steps == 0) x->info.mse_final = x->info.mse_initial;
    }
    HSFM_HIP(hipEventRecord(e1, s));
    x->info.train_ms = elapsed_ms(e0, e1);
    xs.release(); ys.release(); gm.release(); lossb.release();

    // stage 2: leaves are the keys' leading log2(L) bits
    HSFM_HIP(hipEventRecord(e0, s));
    DevBuf cnt, mom;
    if ((rc = x->leaves.ensure((size_t)L * sizeof(LsiLeaf))) || (rc = cnt.ensure((size_t)L * 4)) ||
        (rc = mom.ensure((size_t)L * 32)))
        return rc;
    HSFM_HIP(hipMemsetAsync(mom.p, 0, (size_t)L * 32, s));
    hipLaunchKernelGGL(k_lsi_leaf_init, dim3((L + 255) / 256), dim3(256), 0, s, x->leaves.as<LsiLeaf>(),
                       cnt.as<uint32_t>(), L);
    const LsiModel md = x->model();
    const uint64_t chunks = (N + kLsiChunk - 1) / kLsiChunk;
    const dim3 cg((uint32_t)((chunks + 255) / 256));
    if (N) {
        hipLaunchKernelGGL(k_lsi_route, cg, dim3(256), 0, s, md, x->keys.as<unsigned long long>(),
                           x->leaves.as<LsiLeaf>(), cnt.as<uint32_t>());
        hipLaunchKernelGGL(k_lsi_moments, cg, dim3(256), 0, s, md, x->keys.as<unsigned long long>(),
                           x->leaves.as<LsiLeaf>(), mom.as<double>());
    }
    hipLaunchKernelGGL(k_lsi_solve, dim3((L + 255) / 256), dim3(256), 0, s, x->leaves.as<LsiLeaf>(),
                       cnt.as<uint32_t>(), mom.as<double>(), L);
    if (N)
        hipLaunchKernelGGL(k_lsi_err, cg, dim3(256), 0, s, md, x->keys.as<unsigned long long>(),
                           x->leaves.as<LsiLeaf>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipEventRecord(e1, s));
    x->info.fit_ms = elapsed_ms(e0, e1);
    std::vector<LsiLeaf> h(L);
    std::vector<uint32_t> hc(L);
    HSFM_HIP(hipMemcpyAsync(h.data(), x->leaves.p, (size_t)L * sizeof(LsiLeaf), hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)L * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    double wsum = 0;
    uint32_t emax = 0;
    for (uint32_t i = 0; i < L; i++) {
        emax = h[i].err > emax ? h[i].err : emax;
        wsum += (double)hc[i] * (2.0 * h[i].err + 1.0);
    }
    x->info.max_err = emax;
    x->info.mean_window = N ? wsum / (double)N : 0.0;
    cnt.release(); mom.release();
    // fence levels: as many as the widest window (2 max_err + 1 keys) needs
    // to span at most 17 entries of the coarsest; each padded past the last
    // key (pair loads read one entry ahead)
    x->nfence = (uint32_t)lsi_fence_levels(emax);
    x->info.fence_levels = x->nfence;
    for (int lv = 0; lv < (int)x->nfence; lv++) {
        const int sh = lsi_fence_shift(lv);
        DevBuf &f = x->fence[lv];
        const uint64_t nf = (N >> sh) + 2;
        if ((rc = f.ensure(nf * 8))) return rc;
        hipLaunchKernelGGL(k_lsi_fence, dim3((uint32_t)((nf + 255) / 256)), dim3(256), 0, s,
                           x->keys.as<unsigned long long>(), N, sh, f.as<unsigned long long>(), nf);
        HSFM_HIP(hipGetLastError());
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return HSFM_OK;
}

// Flags of the keys whose IUPAC set is more than their exact interval
// (k_lsi_flag); sets info.mem_seeding when the flags are complete.
int lsi_flag(hsfm_lsi *x) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    const DevConst &hc = idx->h_const;
    LsiLaneSets ls{};
    bool any = false;
    for (uint32_t l = 0; l < 16 && l < hc.num_lanes; l++) {
        uint32_t m = 0;
        for (uint32_t b = 0; b < 4; b++) m |= ((hc.compat_lanes[b + 1] >> l) & 1u) << b;
        ls.qb[l] = (uint8_t)m;
        ls.exact[l] = hc.lane_sym[l] >= 1 && hc.lane_sym[l] <= 4;
        any |= !ls.exact[l] && m != 0;
    }
    x->info.flagged = 0;
    x->info.mem_seeding = 0;
    if (x->cfg.k > kLsiMemMaxK) return HSFM_OK; // the MEM kernel reads windows of at most kLsiMemMaxK codes
    if (!any || x->info.num_keys == 0) { // no text window can hold another compatible symbol
        x->info.mem_seeding = 1;
        return HSFM_OK;
    }
    int rc;
    if ((rc = ensure_sa_full(idx))) return rc;
    if (!idx->fwd.view.text) return HSFM_OK; // no device text: flags unknown, no MEM seeding
    DevBuf ctr;
    if ((rc = ctr.ensure(16))) return rc;
    HSFM_HIP(hipMemsetAsync(ctr.p, 0, 16, s));
    hipEvent_t e0, e1;
    HSFM_HIP(hipEventCreate(&e0));
    HSFM_HIP(hipEventCreate(&e1));
    HSFM_HIP(hipEventRecord(e0, s));
    const uint32_t n = idx->fwd.view.n;
    const uint32_t grid = (uint32_t)idx->dev->num_cus * 16;
    hipLaunchKernelGGL(k_lsi_flag, dim3(grid), dim3(256), 0, s, x->model(), x->keys.as<unsigned long long>(),
                       x->fences(), x->pay.as<uint4>(), x->leaves.as<LsiLeaf>(), idx->fwd.view.text, n, ls,
                       kLsiFlagMaxExpansions, ctr.as<unsigned long long>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipEventRecord(e1, s));
    uint64_t h[2] = {0, 0};
    HSFM_HIP(hipMemcpyAsync(h, ctr.p, 16, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    x->info.flag_ms = elapsed_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    x->info.flagged = h[0];
    x->info.mem_seeding = h[1] == 0 ? 1u : 0u;
    return HSFM_OK;
}

} // namespace

namespace hsfm {
int lsi_mem_view(const hsfm_lsi *x, hsfm_index *idx, LsiView *out) {
    if (!x) return set_error(HSFM_EINVAL, "learned-seed restarts need a learned seed index");
    if (x->idx != idx) return set_error(HSFM_EINVAL, "the learned seed index was built on another index");
    if (!x->info.mem_seeding)
        return set_error(HSFM_EINVAL, "this learned seed index cannot seed MEM restarts exactly (k > %u, or the "
                                      "text's ambiguity windows could not all be expanded)", kLsiMemMaxK);
    out->md = x->model();
    out->fc = x->fences();
    out->keys = x->keys.as<unsigned long long>();
    out->pay = x->pay.as<uint4>();
    out->leaves = x->leaves.as<LsiLeaf>();
    return HSFM_OK;
}
uint32_t lsi_k(const hsfm_lsi *x) { return x->cfg.k; }
} // namespace hsfm

extern "C" int hsfm_lsi_build(hsfm_index *idx, const hsfm_lsi_config *cfg, hsfm_lsi **out) {
    if (!idx || !cfg || !out) return set_error(HSFM_EINVAL, "hsfm_lsi_build: null argument");
    if (cfg->k == 0 || cfg->k > 31) return set_error(HSFM_EINVAL, "hsfm_lsi_build: k must be in 1..31");
    if (cfg->hidden == 0 || cfg->hidden > kLsiMaxHidden)
        return set_error(HSFM_EINVAL, "hsfm_lsi_build: hidden must be in 1..%u", kLsiMaxHidden);
    if (cfg->leaves == 0 || cfg->leaves > (1u << 24) || (cfg->leaves & (cfg->leaves - 1)))
        return set_error(HSFM_EINVAL, "hsfm_lsi_build: leaves must be a power of two in 1..2^24");
    if (cfg->lookup > 2 || (cfg->lookup == 2 && cfg->leaves > kLsiLdsLeaves))
        return set_error(HSFM_EINVAL, "hsfm_lsi_build: lookup must be 0, 1 or 2 (2 needs at most %u leaves)",
                         kLsiLdsLeaves);
    HSFM_LOCK_INDEX(idx, "hsfm_lsi_build");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    hsfm_lsi *x = new hsfm_lsi();
    x->idx = idx;
    x->cfg = *cfg;
    hipEvent_t e0, e1;
    HSFM_HIP(hipEventCreate(&e0));
    HSFM_HIP(hipEventCreate(&e1));
    HSFM_HIP(hipEventRecord(e0, idx->dev->stream));
    int rc = lsi_enumerate(x);
    HSFM_HIP(hipEventRecord(e1, idx->dev->stream));
    x->info.enum_ms = elapsed_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = lsi_train_and_fit(x);
    if (!rc) rc = lsi_flag(x);
    if (rc) {
        x->keys.release(); x->pay.release(); x->leaves.release(); x->params.release(); for (DevBuf &f : x->fence) f.release();
        delete x;
        return rc;
    }
    *out = x;
    return HSFM_OK;
}

extern "C" int hsfm_lsi_get_info(const hsfm_lsi *x, hsfm_lsi_info *out) {
    if (!x || !out) return set_error(HSFM_EINVAL, "hsfm_lsi_get_info: null argument");
    *out = x->info;
    return HSFM_OK;
}

extern "C" int hsfm_lsi_free(hsfm_lsi *x) {
    if (!x) return HSFM_OK;
    {
        HSFM_LOCK_INDEX(x->idx, "hsfm_lsi_free");
        if (x->idx->mem_lsi == x) { // back to table seeding
            x->idx->mem_lsi = nullptr;
            x->idx->mem_seed_mode = HSFM_SEED_TABLES;
        }
    }
    x->keys.release(); x->pay.release(); x->leaves.release(); x->params.release(); for (DevBuf &f : x->fence) f.release();
    delete x;
    return HSFM_OK;
}

namespace {
// Lookup of n device keys; the caller holds idx->mu.
int lsi_seed_impl(hsfm_lsi *x, const uint64_t *d_keys, uint64_t n, uint32_t *d_out, hsfm_seed_stats *st) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    int rc;
    if ((rc = idx->ws.counters.ensure(64))) return rc;
    HSFM_HIP(hipMemsetAsync(idx->ws.counters.p, 0, 8, s));
    const LsiModel md = x->model();
    if ((rc = event_record(idx, 0, false))) return rc;
    // leaves in LDS when they fit and two fence levels suffice (measured:
    // 100 Mbp, 4096 leaves, 0.66 ms against 0.88 ms; 1 Gbp, whose windows need
    // a third level, 1.16 ms against 1.12 ms)
    const bool lds = x->cfg.lookup == 2 || (x->cfg.lookup == 0 && x->cfg.leaves <= kLsiLdsLeaves && x->nfence == 2);
    const unsigned long long *dk = reinterpret_cast<const unsigned long long *>(d_keys);
    if (n && lds) {
        const uint64_t need = (n + 1023) / 1024, cus = (uint64_t)idx->dev->num_cus;
        hipLaunchKernelGGL(k_lsi_lookup_lds, dim3((uint32_t)(need < cus ? need : cus)), dim3(1024), 0, s, md,
                           x->keys.as<unsigned long long>(), x->fences(), x->pay.as<uint4>(),
                           x->leaves.as<LsiLeaf>(), dk, n, reinterpret_cast<uint4 *>(d_out),
                           idx->ws.counters.as<unsigned long long>());
    } else if (n) {
        hipLaunchKernelGGL(k_lsi_lookup, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, md,
                           x->keys.as<unsigned long long>(), x->fences(), x->pay.as<uint4>(),
                           x->leaves.as<LsiLeaf>(), dk, n, reinterpret_cast<uint4 *>(d_out),
                           idx->ws.counters.as<unsigned long long>());
    }
    HSFM_HIP(hipGetLastError());
    if ((rc = event_record(idx, 0, true))) return rc;
    if (st) {
        uint64_t probes = 0;
        if ((rc = readback_u64(idx, idx->ws.counters.as<uint64_t>(), &probes, 1))) return rc;
        st->kernel_ms = event_ms(idx, 0);
        st->probes = probes;
    }
    return HSFM_OK;
}

// Plain exact search of n device keys; the caller holds idx->mu.
int kmer_search_impl(hsfm_index *idx, uint32_t k, const uint64_t *d_keys, uint64_t n, uint32_t *d_out,
                     hsfm_seed_stats *st) {
    hipStream_t s = idx->dev->stream;
    int rc;
    if ((rc = event_record(idx, 0, false))) return rc;
    if (n)
        hipLaunchKernelGGL(k_kmer_search, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, idx->fwd.view,
                           idx->d_const.as<DevConst>(), k, reinterpret_cast<const unsigned long long *>(d_keys), n,
                           reinterpret_cast<uint4 *>(d_out));
    HSFM_HIP(hipGetLastError());
    if ((rc = event_record(idx, 0, true))) return rc;
    if (st) {
        st->kernel_ms = event_ms(idx, 0);
        st->probes = 0;
    }
    return HSFM_OK;
}
} // namespace

// ---- allocation-free lookup -----------------------------------------------
// The scratch of hsfm_lsi_seed_device_noalloc: word 0 accumulates the call's
// bounded-search loads.  Each GPU holds its own copy of the array until the
// process ends; host code uses the address hipGetSymbolAddress gives on that
// device, never the array's name.  Calls that overlap on one GPU share it.
constexpr uint32_t kLsiScratchWords = 32; // 256 bytes
static __device__ __align__(256) unsigned long long scratch[kLsiScratchWords];
constexpr int kLsiScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch_ptr[kLsiScratchDevices];

__global__ void k_lsi_zero(unsigned long long *p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

namespace {
// Device address of `scratch` on `ordinal` (that device current), looked up
// once per device.
int lsi_scratch(int ordinal, unsigned long long **out) {
    if (ordinal < 0 || ordinal >= kLsiScratchDevices)
        return set_error(HSFM_EINVAL, "allocation-free lookup: device ordinal %d out of range", ordinal);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ordinal]) {
        void *p = nullptr;
        HSFM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[ordinal] = static_cast<unsigned long long *>(p);
    }
    *out = g_scratch_ptr[ordinal];
    return HSFM_OK;
}
} // namespace

extern "C" int hsfm_lsi_seed_device_noalloc(hsfm_lsi *x, const uint64_t *d_keys, uint64_t n, uint32_t *d_out,
                                            uint64_t *probes) {
    if (!x || (n && (!d_keys || !d_out)))
        return set_error(HSFM_EINVAL, "hsfm_lsi_seed_device_noalloc: null argument");
    if (n > kLsiNoAllocMaxKeys)
        return set_error(HSFM_EINVAL, "hsfm_lsi_seed_device_noalloc: %llu keys, at most %llu per call",
                         (unsigned long long)n, (unsigned long long)kLsiNoAllocMaxKeys);
    hsfm_index *idx = x->idx;
    HSFM_LOCK_INDEX(idx, "hsfm_lsi_seed_device_noalloc");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    unsigned long long *sc = nullptr;
    int rc;
    if ((rc = lsi_scratch(idx->dev->ordinal, &sc))) return rc;
    hipStream_t s = idx->dev->stream;
    hipLaunchKernelGGL(k_lsi_zero, dim3(1), dim3(64), 0, s, sc, kLsiScratchWords);
    const LsiModel md = x->model();
    const bool lds = x->cfg.lookup == 2 || (x->cfg.lookup == 0 && x->cfg.leaves <= kLsiLdsLeaves && x->nfence == 2);
    const unsigned long long *dk = reinterpret_cast<const unsigned long long *>(d_keys);
    if (n && lds) {
        const uint64_t need = (n + 1023) / 1024, cus = (uint64_t)idx->dev->num_cus;
        hipLaunchKernelGGL(k_lsi_lookup_lds, dim3((uint32_t)(need < cus ? need : cus)), dim3(1024), 0, s, md,
                           x->keys.as<unsigned long long>(), x->fences(), x->pay.as<uint4>(),
                           x->leaves.as<LsiLeaf>(), dk, n, reinterpret_cast<uint4 *>(d_out), sc);
    } else if (n) {
        hipLaunchKernelGGL(k_lsi_lookup, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, md,
                           x->keys.as<unsigned long long>(), x->fences(), x->pay.as<uint4>(),
                           x->leaves.as<LsiLeaf>(), dk, n, reinterpret_cast<uint4 *>(d_out), sc);
    }
    HSFM_HIP(hipGetLastError());
    if (probes) {
        HSFM_HIP(hipMemcpyAsync(probes, sc, 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
    }
    return HSFM_OK;
}

extern "C" int hsfm_lsi_seed_device(hsfm_lsi *x, const uint64_t *d_keys, uint64_t n, uint32_t *d_out,
                                    hsfm_seed_stats *st) {
    if (!x || (n && (!d_keys || !d_out))) return set_error(HSFM_EINVAL, "hsfm_lsi_seed_device: null argument");
    HSFM_LOCK_INDEX(x->idx, "hsfm_lsi_seed_device");
    HSFM_HIP(hipSetDevice(x->idx->dev->ordinal));
    return lsi_seed_impl(x, d_keys, n, d_out, st);
}

extern "C" int hsfm_kmer_search_device(hsfm_index *idx, uint32_t k, const uint64_t *d_keys, uint64_t n,
                                       uint32_t *d_out, hsfm_seed_stats *st) {
    if (!idx || (n && (!d_keys || !d_out))) return set_error(HSFM_EINVAL, "hsfm_kmer_search_device: null argument");
    if (k == 0 || k > 31) return set_error(HSFM_EINVAL, "hsfm_kmer_search_device: k must be in 1..31");
    HSFM_LOCK_INDEX(idx, "hsfm_kmer_search_device");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    return kmer_search_impl(idx, k, d_keys, n, d_out, st);
}

extern "C" int hsfm_lsi_grad(hsfm_device *dev, const float *params, uint32_t hidden, const float *x, const float *y,
                             uint32_t n, float *grad_out, float *loss_out) {
    if (!dev || !params || !grad_out || !loss_out || (n && (!x || !y)))
        return set_error(HSFM_EINVAL, "hsfm_lsi_grad: null argument");
    if (hidden == 0 || hidden > kLsiMaxHidden) return set_error(HSFM_EINVAL, "hsfm_lsi_grad: bad hidden size");
    HSFM_HIP(hipSetDevice(dev->ordinal));
    const uint32_t np = 3 * hidden + 1;
    float *dp = nullptr, *dx = nullptr, *dy = nullptr, *dg = nullptr;
    HSFM_HIP(hipMalloc((void **)&dp, np * 4));
    HSFM_HIP(hipMalloc((void **)&dg, (np + 1) * 4));
    HSFM_HIP(hipMalloc((void **)&dx, (size_t)n * 4 + 4));
    HSFM_HIP(hipMalloc((void **)&dy, (size_t)n * 4 + 4));
    hipStream_t s = dev->stream;
    HSFM_HIP(hipMemcpyAsync(dp, params, np * 4, hipMemcpyHostToDevice, s));
    if (n) {
        HSFM_HIP(hipMemcpyAsync(dx, x, (size_t)n * 4, hipMemcpyHostToDevice, s));
        HSFM_HIP(hipMemcpyAsync(dy, y, (size_t)n * 4, hipMemcpyHostToDevice, s));
    }
    HSFM_HIP(hipMemsetAsync(dg, 0, (np + 1) * 4, s));
    if (n)
        hipLaunchKernelGGL(k_lsi_grad, dim3((n + 255) / 256), dim3(256), 0, s, dp, hidden, dx, dy, n, dg, dg + np);
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipMemcpyAsync(grad_out, dg, np * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipMemcpyAsync(loss_out, dg + np, 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if (n) *loss_out /= (float)n;
    (void)hipFree(dp); (void)hipFree(dg); (void)hipFree(dx); (void)hipFree(dy);
    return HSFM_OK;
}

// Host-array variants: stage keys and results through the index workspace,
// under one hold of the index lock (the staging buffers are shared with the
// other entry points of the index).
namespace {
int stage_seed(hsfm_index *idx, const uint64_t *keys, uint64_t n, uint64_t **d_keys, uint32_t **d_out) {
    int rc;
    if ((rc = idx->ws.qcopy.ensure(n * 8 + 8)) || (rc = idx->ws.hits.ensure(n * 16 + 16))) return rc;
    if (n) HSFM_HIP(hipMemcpyAsync(idx->ws.qcopy.p, keys, n * 8, hipMemcpyHostToDevice, idx->dev->stream));
    *d_keys = idx->ws.qcopy.as<uint64_t>();
    *d_out = idx->ws.hits.as<uint32_t>();
    return HSFM_OK;
}
int unstage_seed(hsfm_index *idx, uint32_t *out, uint64_t n) {
    if (n) HSFM_HIP(hipMemcpyAsync(out, idx->ws.hits.p, n * 16, hipMemcpyDeviceToHost, idx->dev->stream));
    HSFM_HIP(hipStreamSynchronize(idx->dev->stream));
    return HSFM_OK;
}
} // namespace

extern "C" int hsfm_lsi_seed(hsfm_lsi *x, const uint64_t *keys, uint64_t n, uint32_t *out, hsfm_seed_stats *st) {
    if (!x || (n && (!keys || !out))) return set_error(HSFM_EINVAL, "hsfm_lsi_seed: null argument");
    uint64_t *dk;
    uint32_t *dout;
    int rc;
    HSFM_LOCK_INDEX(x->idx, "hsfm_lsi_seed");
    HSFM_HIP(hipSetDevice(x->idx->dev->ordinal));
    if ((rc = stage_seed(x->idx, keys, n, &dk, &dout)) || (rc = lsi_seed_impl(x, dk, n, dout, st))) return rc;
    return unstage_seed(x->idx, out, n);
}

extern "C" int hsfm_kmer_search(hsfm_index *idx, uint32_t k, const uint64_t *keys, uint64_t n, uint32_t *out) {
    if (!idx || (n && (!keys || !out))) return set_error(HSFM_EINVAL, "hsfm_kmer_search: null argument");
    if (k == 0 || k > 31) return set_error(HSFM_EINVAL, "hsfm_kmer_search: k must be in 1..31");
    uint64_t *dk;
    uint32_t *dout;
    int rc;
    HSFM_LOCK_INDEX(idx, "hsfm_kmer_search");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    if ((rc = stage_seed(idx, keys, n, &dk, &dout)) || (rc = kmer_search_impl(idx, k, dk, n, dout, nullptr)))
        return rc;
    return unstage_seed(idx, out, n);
}
