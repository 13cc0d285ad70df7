// hsfm_lsi.hip -- learned seed index on gfx950.
//
// For a window W of k codes (k <= 31; 21 by default) the index answers what
// the MEM restarts and exact k-mer seeding need without k backward-search
// steps:
//   * exact lookup: for an exact-ACGT k-mer, the bidirectional SA interval
//     {fwd_lo, fwd_hi, rev_lo} of plain backward search (extend_left over W
//     from the full interval, bidir.rs:112-130, the exact-code extension of
//     LookupTable::build, lookup.rs:37-91), or "absent";
//   * restart lookup (k <= 25): for an all-base window q[j-k..j), every
//     interval of its IUPAC-compatible set (the set find_smems extends,
//     smem.rs:154-222), or, when the set is empty, the length of the longest
//     suffix of W with a compatible occurrence (so L(j) = j - that length).
// Entries (hsfm_lsi_dev.h) are sorted keys that hold the codes last base
// first, so k-mers sharing a suffix are neighbours:
//   1. exact entries: every present exact-ACGT k-mer, enumerated from the
//      index itself level by level (the node of cS is one exact left extension
//      of the node of S; children placed parent by parent, which keeps each
//      level sorted);
//   2. ambiguity entries: every distinct text window of k codes that holds an
//      ambiguity code, expanded into the base strings compatible with it, each
//      with the window's interval;
//   3. short entries: windows of r < k codes after a contig start or a
//      symbol no base matches, with the neighbour bounds of their r;
//   4. stage 1, an FP32 MLP x -> f(x) ~ slot / N (the key CDF) of H clipped
//      ReLU ramps, trained by Adam on a sample of (key, slot) pairs with the
//      forward/backward kernels below and kept non-decreasing (w1, w2
//      projected to >= 0; a sum of non-negative ramps is any monotone
//      piecewise-linear curve);
//   5. stage 2: leaf = the key's top log2(L) bits (the last bases of the
//      k-mer); per leaf a least-squares map slot ~ s0 + a * f(u) + b (a >= 0)
//      of the stage-1 output at u, the key's position inside its leaf, and the
//      maximum error of its entries.  The key distribution is close to a
//      product of per-base frequencies, so the CDF inside a leaf has the shape
//      of the global CDF: the trained f bends each leaf's map along it (an
//      untrained f(x) = x leaves one straight line per leaf).
// Lookup: route, predict, then the lower bound of the key inside the error
// window through fence levels of every 2^5, 2^9, 2^13 ... -th key (as many as
// the widest window needs; each level narrows the range to 17 entries of the
// next) and a last round of at most 35 keys, each level one round of
// independent loads.  The prediction is non-decreasing inside a leaf, so the
// window holds the lower bound of any key, present or not, and its two
// neighbours give the longest-suffix answer.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>

#include "hsfm_host.h"
#include "hsfm_lsi_dev.h"

namespace hsfm {

// ---- 1. exact entries ------------------------------------------------------
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
// c in ACGT, c in the key's digit below S's.  COUNT: nch[p] = number of
// children.  WRITE: the children of p at off[p], in c order (sorted).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_lsi_expand(DevDir fwd, const DevConst *__restrict__ cst,
                                                    const unsigned long long *__restrict__ pkey,
                                                    const uint4 *__restrict__ piv, uint32_t P, uint32_t level,
                                                    uint8_t *__restrict__ nch, const uint32_t *__restrict__ off,
                                                    unsigned long long *__restrict__ ckey, uint4 *__restrict__ civ) {
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
    const unsigned long long pk = WRITE ? pkey[p] : 0ull;
    uint32_t at = WRITE ? off[p] : 0u, cnt = 0;
#pragma unroll
    for (uint32_t d = 0; d < 4; d++) {
        const uint32_t l = s_bl[d];
        uint32_t a = 0, b = 0, acc = 0, cs = 0;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if ((uint32_t)j == l) { a = cl[j]; b = ch[j]; acc = below[j]; cs = s_clane[j]; }
        if (l < 16 && b > a) {
            cnt++;
            if (WRITE) {
                ckey[at] = pk | ((unsigned long long)d << (62 - 2 * level));
                civ[at] = make_uint4(cs + a, cs + b, iv.z + acc, kLsiPayExact);
                at++;
            }
        }
    }
    if (!WRITE) nch[p] = (uint8_t)cnt;
}

struct U8ToU32 {
    __device__ __forceinline__ uint32_t operator()(uint8_t v) const { return v; }
};

// ---- 2./3. ambiguity and short entries -------------------------------------
// ls.qb[lane]: the query bases (bit b: base b+1) compatible with the lane's
// symbol (0: no base, the lane cuts windows); ls.exact[lane]: the lane of A,
// C, G or T.
struct LsiLaneSets {
    uint8_t qb[16];
    uint8_t exact[16];
};
constexpr uint32_t kLsiMaxExpansions = 4096; // per window (six N)

// Windows of K codes that hold an ambiguity code and no cutting symbol, each
// taken once: by the thread of its first ambiguity position p.  PASS 0 counts
// them (ctr[0]) and those with more than cap expansions (ctr[1]).  PASS 1
// computes each one's interval (K exact-lane extend_left steps of its own
// codes) and keeps it when its start is the interval's first row (one window
// per distinct string): rec[] = {start, fwd_lo, fwd_hi, rev_lo}, rexp[] =
// its expansion count, at ctr[2]++ (capacity rcap).
template <int PASS>
__global__ __launch_bounds__(256) void k_lsi_windows(DevDir fwd, const DevConst *__restrict__ cst, uint32_t K,
                                                     LsiLaneSets ls, uint32_t cap, unsigned long long *__restrict__ ctr,
                                                     uint4 *__restrict__ rec, uint32_t *__restrict__ rexp,
                                                     uint64_t rcap) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint8_t *__restrict__ text = fwd.text;
    const uint32_t n = fwd.n;
    unsigned long long nwin = 0, over = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t tp = text[p] & 15u;
        if (ls.exact[tp] || ls.qb[tp] == 0) continue;
        for (uint64_t s = p;; s--) {
            if (s < p && !ls.exact[text[s] & 15u]) break; // p is not the first ambiguity code
            if (s + K <= n) {
                bool ok = true;
                uint32_t ne = 1;
                for (uint32_t i = 0; i < K && ok; i++) {
                    const uint32_t m = ls.qb[text[s + i] & 15u];
                    if (m == 0) ok = false;
                    else if (ne <= cap) ne *= (uint32_t)__popc(m);
                }
                if (ok) {
                    if (PASS == 0) {
                        nwin++;
                        if (ne > cap) over++;
                    } else {
                        uint32_t lo = 0, hi = n, rl = 0;
                        for (int i = (int)K - 1; i >= 0; i--) {
                            const uint32_t L = text[s + (uint32_t)i] & 15u;
                            uint32_t cl[16], ch[16];
                            pair_counts(fwd, lo, hi, cl, ch);
                            uint32_t a = 0, b = 0, acc = 0, run = 0, cs = 0;
#pragma unroll
                            for (int j = 0; j < 16; j++) {
                                if ((uint32_t)j == L) { a = cl[j]; b = ch[j]; acc = run; cs = s_clane[j]; }
                                run += ch[j] - cl[j];
                            }
                            lo = cs + a;
                            hi = cs + b;
                            rl += acc;
                        }
                        if (fwd.isa[s] == lo) {
                            const unsigned long long at = atomicAdd(&ctr[2], 1ull);
                            if (at < rcap) {
                                rec[at] = make_uint4((uint32_t)s, lo, hi, rl);
                                rexp[at] = ne;
                            }
                        }
                    }
                }
            }
            if (s == 0 || s + K == p + 1) break;
        }
    }
    if (PASS == 0) {
        if (nwin) atomicAdd(&ctr[0], nwin);
        if (over) atomicAdd(&ctr[1], over);
    }
}

// The base strings compatible with text[s, s + r) as keys: expansion e picks,
// at each position, the (e's mixed-radix digit)-th compatible base.
__device__ __forceinline__ unsigned long long lsi_expansion_key(const uint8_t *__restrict__ text, uint64_t s,
                                                                uint32_t r, const LsiLaneSets &ls, uint32_t e) {
    unsigned long long key = 0;
    for (uint32_t i = 0; i < r; i++) {
        const uint32_t m = ls.qb[text[s + i] & 15u];
        const uint32_t c = (uint32_t)__popc(m);
        const uint32_t pick = e % c;
        e /= c;
        uint32_t mm = m;
        for (uint32_t z = 0; z < pick; z++) mm &= mm - 1;
        key |= (unsigned long long)(__ffs(mm) - 1) << (64 - 2 * r + 2 * i);
    }
    return key;
}

// Ambiguity entries of the kept windows: window i writes its rexp[i]
// expansions from off[i].
__global__ __launch_bounds__(256) void k_lsi_window_entries(const uint8_t *__restrict__ text, uint32_t K,
                                                            LsiLaneSets ls, const uint4 *__restrict__ rec,
                                                            const uint32_t *__restrict__ rexp,
                                                            const uint64_t *__restrict__ off, uint64_t nrec,
                                                            unsigned long long *__restrict__ keys,
                                                            uint4 *__restrict__ pay) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrec;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 r = rec[i];
        const uint64_t o = off[i];
        for (uint32_t e = 0; e < rexp[i]; e++) {
            keys[o + e] = lsi_expansion_key(text, r.x, K, ls, e) | kLsiVariant;
            pay[o + e] = make_uint4(r.y, r.z, r.w, 0u);
        }
    }
}

// Short entries: for every contig start s (position 0, or one past a symbol
// no base matches) the windows [s, s + r), r < K, that hold no cutting
// symbol, expanded.  PASS 0 counts the entries (ctr[0]) and windows with more
// than cap expansions (ctr[1]); PASS 1 writes them from out[ctr[2]++].
template <int PASS>
__global__ __launch_bounds__(256) void k_lsi_short(const uint8_t *__restrict__ text, uint32_t n, uint32_t K,
                                                   LsiLaneSets ls, uint32_t cap, unsigned long long *__restrict__ ctr,
                                                   unsigned long long *__restrict__ keys, uint4 *__restrict__ pay,
                                                   uint64_t kcap) {
    unsigned long long cnt = 0, over = 0;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (uint64_t)gridDim.x * blockDim.x) {
        if (s > 0 && ls.qb[text[s - 1] & 15u] != 0) continue; // not a contig start
        uint32_t ne = 1;
        for (uint32_t r = 1; r < K && s + r <= n; r++) {
            const uint32_t m = ls.qb[text[s + r - 1] & 15u];
            if (m == 0) break;
            if (ne <= cap) ne *= (uint32_t)__popc(m);
            if (ne > cap) {
                over++;
                break;
            }
            if (PASS == 0) {
                cnt += ne;
            } else {
                const unsigned long long at = atomicAdd(&ctr[2], (unsigned long long)ne);
                const unsigned long long meta =
                    kLsiShort | ((unsigned long long)r << kLsiGShift) | ((unsigned long long)r << kLsiHShift);
                for (uint32_t e = 0; e < ne && at + e < kcap; e++) {
                    keys[at + e] = lsi_expansion_key(text, s, r, ls, e) | meta;
                    pay[at + e] = make_uint4(0u, 0u, 0u, 0u);
                }
            }
        }
    }
    if (PASS == 0) {
        if (cnt) atomicAdd(&ctr[0], cnt);
        if (over) atomicAdd(&ctr[1], over);
    }
}

__global__ void k_lsi_iota(uint32_t *v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}
__global__ void k_lsi_gather(const uint4 *__restrict__ in, const uint32_t *__restrict__ perm, uint64_t n,
                             uint4 *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}
__global__ void k_lsi_pad(unsigned long long *keys, uint64_t n) {
    if (threadIdx.x < 2) keys[n + threadIdx.x] = ~0ull;
}

// Full entries: the first of every run of equal codes records the run's
// length in its payload.
__global__ __launch_bounds__(256) void k_lsi_runs(const unsigned long long *__restrict__ keys, uint64_t n,
                                                  unsigned long long smask, uint4 *__restrict__ pay) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long key = keys[i];
        if (key & kLsiShort) continue;
        const unsigned long long w = key & smask;
        if (i > 0 && (keys[i - 1] & smask) == w) continue;
        uint64_t j = i + 1;
        while (j < n && (keys[j] & smask) == w && !(keys[j] & kLsiShort)) j++;
        pay[i].w |= (uint32_t)(j - i) << 8;
    }
}

// Neighbour bounds of short entries.  With entries x_0 < x_1 < ... of reach
// r_m (k for full entries), G_i = max over m <= i of min(r_m, lcs(x_m, x_i))
// satisfies G_i = max(r_i, min(G_{i-1}, lcs(x_{i-1}, x_i))) (the common
// prefix of sorted strings is the minimum along the run between them), and
// H likewise from the right; a full entry's own term makes its bound k.  One
// thread per run of consecutive short entries, G pass then H pass.
template <bool G>
__global__ __launch_bounds__(256) void k_lsi_bounds(unsigned long long *__restrict__ keys, uint64_t n, LsiModel md) {
    const int sh = G ? kLsiGShift : kLsiHShift;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!(keys[i] & kLsiShort)) continue;
        const uint64_t nb = G ? i - 1 : i + 1; // neighbour towards the run's outside
        const bool edge = G ? i == 0 : i + 1 == n;
        if (!edge && (keys[nb] & kLsiShort)) continue; // not the first entry of its run (in this direction)
        uint32_t bound = edge ? 0u : md.k;
        unsigned long long prev = edge ? 0ull : keys[nb];
        for (uint64_t j = i;; j = G ? j + 1 : j - 1) {
            const unsigned long long x = keys[j];
            if (!(x & kLsiShort)) break;
            const uint32_t r = (uint32_t)(x >> sh) & 31u;
            const uint32_t l = lsi_lcs(md, prev, x);
            uint32_t b = bound < l ? bound : l;
            if (j == i && edge) b = 0;
            b = b > r ? b : r;
            keys[j] = (x & ~(31ull << sh)) | ((unsigned long long)b << sh);
            bound = b;
            prev = x;
            if (G ? j + 1 == n : j == 0) break;
        }
    }
}

// ---- 4. stage-1 training -----------------------------------------------
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
    const float f = lsi_stage1(s_p, H, xi);
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
        const float h = fminf(fmaxf(pre, 0.f), 1.f);
        const float dh = pre > 0.f && pre < 1.f ? df * s_p[2 * H + i] : 0.f;
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

// One Adam step, then w1 and w2 projected onto >= 0 (f stays non-decreasing).
__global__ void k_lsi_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, uint32_t H, float lr, float b1t, float b2t) {
    const uint32_t i = threadIdx.x, np = 3 * H + 1;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float pi = p[i] - lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + 1e-8f);
    if (i < H || (i >= 2 * H && i < 3 * H)) pi = fmaxf(pi, 0.f);
    p[i] = pi;
}

// training sample, every stride-th slot: the key's position inside its leaf
// (u) and its rank there as a fraction of the leaf's entries -- the in-leaf
// CDF that every leaf's map then scales
__global__ void k_lsi_sample(const unsigned long long *__restrict__ keys, uint64_t stride, uint32_t ns, LsiModel md,
                             const LsiLeaf *__restrict__ lf, const uint32_t *__restrict__ cnt, float *__restrict__ x,
                             float *__restrict__ y) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const uint64_t s = (uint64_t)i * stride;
    const unsigned long long key = keys[s];
    const uint32_t l = lsi_leaf(md, key);
    x[i] = lsi_u(md, key);
    y[i] = (float)((double)(s - lf[l].s0) / (double)cnt[l]);
}

// ---- 5. stage 2 ----------------------------------------------------------
constexpr uint32_t kLsiChunk = 64; // keys per thread in the leaf passes

// entries per leaf (chunk-aggregated atomics; keys sorted)
__global__ __launch_bounds__(256) void k_lsi_route(LsiModel md, const unsigned long long *__restrict__ keys,
                                                   uint32_t *__restrict__ cnt) {
    const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLsiChunk;
    uint32_t cur = 0xffffffffu, c = 0;
    for (uint64_t s = c0; s < c0 + kLsiChunk && s < md.nkeys; s++) {
        const uint32_t l = lsi_leaf(md, keys[s]);
        if (l != cur) {
            if (c) atomicAdd(&cnt[cur], c);
            cur = l;
            c = 0;
        }
        c++;
    }
    if (c) atomicAdd(&cnt[cur], c);
}

// s0 of every leaf from the scanned counts (an empty leaf gets the next
// leaf's first slot)
__global__ void k_lsi_leaf_init(LsiLeaf *lf, const uint32_t *first, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    LsiLeaf z;
    z.s0 = first[i]; z.err = 0; z.a = 0.f; z.b = 0.f;
    lf[i] = z;
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

// per leaf: the least-squares line with a >= 0 (a negative slope becomes the
// mean), so the prediction stays non-decreasing in the key
__global__ void k_lsi_solve(LsiLeaf *__restrict__ lf, const uint32_t *__restrict__ cnt,
                            const double *__restrict__ mom, uint32_t L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    LsiLeaf f = lf[i];
    const uint32_t c = cnt[i];
    if (c == 0) {
        f.a = 0.f; f.b = 0.f;
    } else {
        const double n = c, sx = mom[4 * i], sy = mom[4 * i + 1], sxx = mom[4 * i + 2], sxy = mom[4 * i + 3];
        const double den = n * sxx - sx * sx;
        double a = 0, b = sy / n;
        if (c >= 2 && den > 0) {
            a = (n * sxy - sx * sy) / den;
            b = (sy - a * sx) / n;
        }
        if (!isfinite(a) || !isfinite(b) || a < 0) {
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

// A key above every entry of its leaf (or any key of an empty leaf) has its
// lower bound one past the leaf, at the next leaf's s0, which [pred - err,
// pred + err + 1] holds only if err also covers how far the leaf predicts
// past that slot.  The prediction is non-decreasing inside a leaf, so the
// leaf's largest key (its leaf bits, every code bit below them set) bounds
// it: err >= pred(top) - end.  One thread per leaf, after k_lsi_err.
__global__ __launch_bounds__(256) void k_lsi_err_top(LsiModel md, LsiLeaf *__restrict__ lf) {
    __shared__ float s_p[3 * kLsiMaxHidden + 1];
    for (uint32_t i = threadIdx.x; i < 3 * md.hidden + 1; i += blockDim.x) s_p[i] = md.p[i];
    __syncthreads();
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= md.leaves) return;
    const long long end = l + 1 < md.leaves ? (long long)lf[l + 1].s0 : (long long)md.nkeys;
    const unsigned long long top = (md.lbits ? (unsigned long long)l << (64 - md.lbits) : 0ull) |
                                   (md.smask & (~0ull >> md.lbits));
    const long long d = lsi_pred(lf[l], lsi_stage1(s_p, md.hidden, lsi_u(md, top))) - end;
    if (d > 0 && (unsigned long long)d > lf[l].err) lf[l].err = (uint32_t)d;
}

// ---- lookup ----------------------------------------------------------------
// fence[j] = keys[j << sh]; entries past the last key hold the largest key value
__global__ void k_lsi_fence(const unsigned long long *__restrict__ keys, uint64_t nkeys, int sh,
                            unsigned long long *__restrict__ fence, uint64_t nf) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nf) fence[j] = (j << sh) < nkeys ? keys[j << sh] : ~0ull;
}

// Exact lookup of one k-mer (packed first base most significant):
// {fwd_lo, fwd_hi, rev_lo, 1} of its exact entry, or zeros.  probes += the
// bounded search's loads in 64-byte units (the bytes of each round of paired
// loads).  One key per lane (a quad of lanes per key, loading whole lines
// together, ran 3.5x slower: fewer keys in flight).
__device__ __forceinline__ uint4 lsi_lookup_key(const LsiModel &md, const unsigned long long *__restrict__ keys,
                                                const LsiFences &fc, const uint4 *__restrict__ pay,
                                                const LsiLeaf &f, unsigned long long w, const float *s_p,
                                                uint32_t &probes) {
    if (!md.nkeys) return make_uint4(0, 0, 0, 0);
    unsigned long long pv, sv;
    bool hs;
    const long long lb = lsi_search(md, keys, fc, f, lsi_stage1(s_p, md.hidden, lsi_u(md, w)), w, probes, pv, sv, hs);
    if (!lsi_present(md, w, sv, hs)) return make_uint4(0, 0, 0, 0);
    const uint4 v = pay[lb];
    return (v.w & kLsiPayExact) ? make_uint4(v.x, v.y, v.z, 1u) : make_uint4(0, 0, 0, 0);
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
    if (t < nq) {
        const unsigned long long w = lsi_key_of_fwd(q[t]) & md.smask;
        out[t] = lsi_lookup_key(md, keys, fc, pay, lf[lsi_leaf(md, w)], w, s_p, probes);
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
        const unsigned long long w = lsi_key_of_fwd(q[t]) & md.smask;
        out[t] = lsi_lookup_key(md, keys, fc, pay, s_lf[lsi_leaf(md, w)], w, s_p, probes);
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
        return LsiModel{params.as<float>(), cfg.hidden, cfg.leaves, cfg.k, info.entries, nfence, lbits,
                        ~0ull << (64 - 2 * cfg.k)};
    }
    LsiFences fences() const {
        LsiFences fc{};
        for (int v = 0; v < kLsiMaxFences; v++) fc.f[v] = fence[v].as<unsigned long long>();
        return fc;
    }
    void release() {
        keys.release(); pay.release(); leaves.release(); params.release();
        for (DevBuf &f : fence) f.release();
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

int lsi_scan_u32(hsfm_index *idx, const uint32_t *in, uint32_t *out, uint32_t n) {
    hipStream_t s = idx->dev->stream;
    if (n == 0) return HSFM_OK;
    size_t tmp = 0;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
    int rc = idx->ws.scan_tmp.ensure(tmp);
    if (rc) return rc;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(idx->ws.scan_tmp.p, tmp, in, out, (int)n, s));
    return HSFM_OK;
}

// Releases local device buffers on every return (DevBuf frees only on release()).
template <int N> struct Releaser {
    DevBuf *b[N];
    ~Releaser() {
        for (DevBuf *x : b) x->release();
    }
};

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

uint32_t grid_of(uint64_t n, uint32_t cap = 65536) {
    const uint64_t b = (n + 255) / 256;
    return (uint32_t)(b < 1 ? 1 : b > cap ? cap : b);
}

// Every present exact-ACGT k-mer: sorted keys + {fwd_lo, fwd_hi, rev_lo, exact}.
int lsi_enumerate(hsfm_lsi *x, DevBuf &out_keys, DevBuf &out_pay, uint64_t &out_n) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    int rc;
    DevBuf pk, pi, nch, off, ck, ci;
    Releaser<6> rel{{&pk, &pi, &nch, &off, &ck, &ci}};
    if ((rc = pk.ensure(8)) || (rc = pi.ensure(16))) return rc;
    const unsigned long long k0 = 0;
    const uint4 root = make_uint4(0, idx->fwd.view.n, 0, 0);
    HSFM_HIP(hipMemcpyAsync(pk.p, &k0, 8, hipMemcpyHostToDevice, s));
    HSFM_HIP(hipMemcpyAsync(pi.p, &root, 16, hipMemcpyHostToDevice, s));
    uint64_t P = 1;
    for (uint32_t level = 0; level < x->cfg.k; level++) {
        if (P > 0x7fffffffull)
            return set_error(HSFM_EINVAL, "learned seed index: %llu nodes at one level", (unsigned long long)P);
        if ((rc = nch.ensure(P)) || (rc = off.ensure(4 * P + 16))) return rc;
        const dim3 grid((uint32_t)((P + 255) / 256));
        hipLaunchKernelGGL(k_lsi_expand<false>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                           pk.as<unsigned long long>(), pi.as<uint4>(), (uint32_t)P, level, nch.as<uint8_t>(),
                           nullptr, nullptr, nullptr);
        HSFM_HIP(hipGetLastError());
        if ((rc = lsi_scan_u8(idx, nch.as<uint8_t>(), off.as<uint32_t>(), (uint32_t)P))) return rc;
        uint32_t lo = 0;
        uint8_t lf = 0;
        HSFM_HIP(hipMemcpyAsync(&lo, off.as<uint32_t>() + P - 1, 4, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(&lf, nch.as<uint8_t>() + P - 1, 1, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
        const uint64_t C = (uint64_t)lo + lf;
        if ((rc = ck.ensure(C * 8 + 16)) || (rc = ci.ensure(C * 16 + 16))) return rc;
        hipLaunchKernelGGL(k_lsi_expand<true>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                           pk.as<unsigned long long>(), pi.as<uint4>(), (uint32_t)P, level, nullptr,
                           off.as<uint32_t>(), ck.as<unsigned long long>(), ci.as<uint4>());
        HSFM_HIP(hipGetLastError());
        HSFM_HIP(hipStreamSynchronize(s));
        nch.release(); off.release(); pk.release(); pi.release();
        pk = ck; pi = ci;
        ck = DevBuf{}; ci = DevBuf{};
        P = C;
        if (P == 0) break;
    }
    out_keys = pk;
    out_pay = pi;
    pk = DevBuf{};
    pi = DevBuf{};
    out_n = P;
    return HSFM_OK;
}

// Lane sets of the index's alphabet for the restart entries; false if its
// exact lanes are not compatible with exactly their own base (then the
// exact entries are not the sets of all-base windows).
bool lsi_lane_sets(const DevConst &hc, LsiLaneSets &ls) {
    memset(&ls, 0, sizeof(ls));
    bool ok = true;
    for (uint32_t l = 0; l < 16 && l < hc.num_lanes; l++) {
        uint32_t m = 0;
        for (uint32_t b = 0; b < 4; b++) m |= ((hc.compat_lanes[b + 1] >> l) & 1u) << b;
        ls.qb[l] = (uint8_t)m;
        ls.exact[l] = hc.lane_sym[l] >= 1 && hc.lane_sym[l] <= 4;
        if (ls.exact[l] && m != (1u << (hc.lane_sym[l] - 1))) ok = false;
    }
    return ok;
}

// Entries of the restart lookups beyond the exact k-mers, merged with them in
// sorted order; sets info.mem_seeding.  Leaves the exact entries alone (and
// MEM seeding off) when the index cannot have them.
int lsi_restart_entries(hsfm_lsi *x, DevBuf &ek, DevBuf &ep, uint64_t E) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    const uint32_t K = x->cfg.k;
    LsiLaneSets ls;
    x->keys = ek;
    x->pay = ep;
    ek = DevBuf{};
    ep = DevBuf{};
    x->info.entries = E;
    x->info.mem_seeding = 0;
    if (K > kLsiMemMaxK || !lsi_lane_sets(idx->h_const, ls)) return HSFM_OK;
    int rc;
    if ((rc = ensure_sa_full(idx))) return rc;
    const DevDir &fwd = idx->fwd.view;
    if (!fwd.text || !fwd.isa) return HSFM_OK; // no device text: exact lookups only
    const uint32_t n = fwd.n;
    const uint32_t grid = (uint32_t)idx->dev->num_cus * 16;
    DevBuf ctr, rec, rexp, roff, k0, k1, v0, v1, p0, p1, tb;
    Releaser<11> rel{{&ctr, &rec, &rexp, &roff, &k0, &k1, &v0, &v1, &p0, &p1, &tb}};
    if ((rc = ctr.ensure(64))) return rc;
    HSFM_HIP(hipMemsetAsync(ctr.p, 0, 64, s));
    unsigned long long *c = ctr.as<unsigned long long>();
    hipLaunchKernelGGL(k_lsi_windows<0>, dim3(grid), dim3(256), 0, s, fwd, idx->d_const.as<DevConst>(), K, ls,
                       kLsiMaxExpansions, c, nullptr, nullptr, 0ull);
    hipLaunchKernelGGL(k_lsi_short<0>, dim3(grid), dim3(256), 0, s, fwd.text, n, K, ls, kLsiMaxExpansions, c + 3,
                       nullptr, nullptr, 0ull);
    HSFM_HIP(hipGetLastError());
    uint64_t h[8];
    HSFM_HIP(hipMemcpyAsync(h, c, 64, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if (h[1] || h[4]) return HSFM_OK; // a window with too many expansions: the sets cannot all be stored
    const uint64_t nwin = h[0], S = h[3];
    // kept windows and their expansions
    if ((rc = rec.ensure(nwin * 16 + 16)) || (rc = rexp.ensure(nwin * 4 + 16)) || (rc = roff.ensure(nwin * 8 + 16)))
        return rc;
    HSFM_HIP(hipMemsetAsync(ctr.p, 0, 64, s));
    if (nwin)
        hipLaunchKernelGGL(k_lsi_windows<1>, dim3(grid), dim3(256), 0, s, fwd, idx->d_const.as<DevConst>(), K, ls,
                           kLsiMaxExpansions, c, rec.as<uint4>(), rexp.as<uint32_t>(), (uint64_t)nwin);
    HSFM_HIP(hipGetLastError());
    uint64_t nrec = 0;
    HSFM_HIP(hipMemcpyAsync(&nrec, c + 2, 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if (nrec > nwin) return set_error(HSFM_EDEVICE, "learned seed index: window count changed between passes");
    if (nrec > 0xffffffffull) return HSFM_OK;
    if (nrec && (rc = exclusive_scan_u32_u64(idx, rexp.as<uint32_t>(), roff.as<uint64_t>(), (uint32_t)nrec)))
        return rc;
    uint64_t V = 0;
    if (nrec) HSFM_HIP(hipMemcpyAsync(&V, roff.as<uint64_t>() + nrec, 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    const uint64_t T = E + V + S;
    if (T > 0xfffffffeull || T > 0x7fffffffull) return HSFM_OK; // slots are 32-bit; sort sizes are int
    x->info.variant_entries = V;
    x->info.short_entries = S;
    if (V + S == 0) {
        x->info.mem_seeding = 1;
        return HSFM_OK;
    }
    // all entries: exact ones, then the new ones, sorted together
    if ((rc = k0.ensure(T * 8 + 16)) || (rc = k1.ensure(T * 8 + 16)) || (rc = v0.ensure(T * 4 + 16)) ||
        (rc = v1.ensure(T * 4 + 16)) || (rc = p0.ensure(T * 16 + 16)))
        return rc;
    if (E) {
        HSFM_HIP(hipMemcpyAsync(k0.p, x->keys.p, E * 8, hipMemcpyDeviceToDevice, s));
        HSFM_HIP(hipMemcpyAsync(p0.p, x->pay.p, E * 16, hipMemcpyDeviceToDevice, s));
    }
    HSFM_HIP(hipStreamSynchronize(s));
    x->keys.release();
    x->pay.release();
    if (nrec)
        hipLaunchKernelGGL(k_lsi_window_entries, dim3(grid_of(nrec)), dim3(256), 0, s, fwd.text, K, ls,
                           rec.as<uint4>(), rexp.as<uint32_t>(), roff.as<uint64_t>(), nrec,
                           k0.as<unsigned long long>() + E, p0.as<uint4>() + E);
    HSFM_HIP(hipMemsetAsync(ctr.p, 0, 64, s));
    if (S)
        hipLaunchKernelGGL(k_lsi_short<1>, dim3(grid), dim3(256), 0, s, fwd.text, n, K, ls, kLsiMaxExpansions, c,
                           k0.as<unsigned long long>() + E + V, p0.as<uint4>() + E + V, (uint64_t)S);
    hipLaunchKernelGGL(k_lsi_iota, dim3(grid_of(T)), dim3(256), 0, s, v0.as<uint32_t>(), T);
    HSFM_HIP(hipGetLastError());
    rec.release(); rexp.release(); roff.release();
    hipcub::DoubleBuffer<unsigned long long> kb(k0.as<unsigned long long>(), k1.as<unsigned long long>());
    hipcub::DoubleBuffer<uint32_t> vb(v0.as<uint32_t>(), v1.as<uint32_t>());
    size_t tmp = 0;
    HSFM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kb, vb, (int)T, 0, 64, s));
    if ((rc = tb.ensure(tmp))) return rc;
    HSFM_HIP(hipcub::DeviceRadixSort::SortPairs(tb.p, tmp, kb, vb, (int)T, 0, 64, s));
    HSFM_HIP(hipStreamSynchronize(s));
    tb.release();
    DevBuf &ks = kb.Current() == k0.as<unsigned long long>() ? k0 : k1;
    DevBuf &kother = kb.Current() == k0.as<unsigned long long>() ? k1 : k0;
    DevBuf &vs = vb.Current() == v0.as<uint32_t>() ? v0 : v1;
    DevBuf &vother = vb.Current() == v0.as<uint32_t>() ? v1 : v0;
    kother.release();
    vother.release();
    if ((rc = p1.ensure(T * 16 + 16))) return rc;
    hipLaunchKernelGGL(k_lsi_gather, dim3(grid_of(T)), dim3(256), 0, s, p0.as<uint4>(), vs.as<uint32_t>(), T,
                       p1.as<uint4>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipStreamSynchronize(s));
    p0.release();
    vs.release();
    x->keys = ks;
    ks = DevBuf{};
    x->pay = p1;
    p1 = DevBuf{};
    x->info.entries = T;
    x->info.mem_seeding = 1;
    return HSFM_OK;
}

// Run lengths of full entries, neighbour bounds of short entries; pads the
// key buffer.
int lsi_finish_entries(hsfm_lsi *x) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    const uint64_t N = x->info.entries;
    int rc;
    if (x->keys.cap < N * 8 + 16) { // room for the two pad keys
        DevBuf nk;
        if ((rc = nk.ensure(N * 8 + 16))) return rc;
        if (N) HSFM_HIP(hipMemcpyAsync(nk.p, x->keys.p, N * 8, hipMemcpyDeviceToDevice, s));
        HSFM_HIP(hipStreamSynchronize(s));
        x->keys.release();
        x->keys = nk;
    }
    if (!x->pay.p && (rc = x->pay.ensure(16))) return rc;
    hipLaunchKernelGGL(k_lsi_pad, dim3(1), dim3(64), 0, s, x->keys.as<unsigned long long>(), N);
    const LsiModel md = x->model();
    if (N) {
        hipLaunchKernelGGL(k_lsi_runs, dim3(grid_of(N)), dim3(256), 0, s, x->keys.as<unsigned long long>(), N,
                           md.smask, x->pay.as<uint4>());
        if (x->info.short_entries) {
            hipLaunchKernelGGL(k_lsi_bounds<true>, dim3(grid_of(N)), dim3(256), 0, s,
                               x->keys.as<unsigned long long>(), N, md);
            hipLaunchKernelGGL(k_lsi_bounds<false>, dim3(grid_of(N)), dim3(256), 0, s,
                               x->keys.as<unsigned long long>(), N, md);
        }
    }
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipStreamSynchronize(s));
    return HSFM_OK;
}

int lsi_train_and_fit(hsfm_lsi *x) {
    hsfm_index *idx = x->idx;
    hipStream_t s = idx->dev->stream;
    const uint32_t H = x->cfg.hidden, L = x->cfg.leaves, np = 3 * H + 1;
    const uint64_t N = x->info.entries;
    int rc;
    // stage-1 init: ramp i rises over [i/H, (i+1)/H], each weighted 1/H, so
    // f(x) = x; training moves the slope of every segment (w2 >= 0)
    float hp[3 * kLsiMaxHidden + 1];
    for (uint32_t i = 0; i < H; i++) {
        hp[i] = (float)H;
        hp[H + i] = -(float)i;
        hp[2 * H + i] = 1.f / (float)H;
    }
    hp[3 * H] = 0.f;
    if ((rc = x->params.ensure(np * 4))) return rc;
    HSFM_HIP(hipMemcpyAsync(x->params.p, hp, np * 4, hipMemcpyHostToDevice, s));
    DevBuf xs, ys, gm, lossb, cnt, first, mom;
    Releaser<7> rel{{&xs, &ys, &gm, &lossb, &cnt, &first, &mom}};
    // leaves: entries per leaf and each leaf's first slot (the model has no
    // part in routing, so they come first and give the training targets)
    if ((rc = x->leaves.ensure((size_t)L * sizeof(LsiLeaf))) || (rc = cnt.ensure((size_t)L * 4 + 4)) ||
        (rc = first.ensure((size_t)L * 4 + 4)) || (rc = mom.ensure((size_t)L * 32)))
        return rc;
    HSFM_HIP(hipMemsetAsync(cnt.p, 0, (size_t)L * 4, s));
    const LsiModel md0 = x->model();
    const uint64_t chunks = (N + kLsiChunk - 1) / kLsiChunk;
    const dim3 cg((uint32_t)((chunks + 255) / 256));
    if (N)
        hipLaunchKernelGGL(k_lsi_route, cg, dim3(256), 0, s, md0, x->keys.as<unsigned long long>(), cnt.as<uint32_t>());
    if ((rc = lsi_scan_u32(idx, cnt.as<uint32_t>(), first.as<uint32_t>(), L))) return rc;
    hipLaunchKernelGGL(k_lsi_leaf_init, dim3((L + 255) / 256), dim3(256), 0, s, x->leaves.as<LsiLeaf>(),
                       first.as<uint32_t>(), L);
    HSFM_HIP(hipGetLastError());
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
        hipLaunchKernelGGL(k_lsi_sample, dim3((ns + 255) / 256), dim3(256), 0, s, x->keys.as<unsigned long long>(),
                           stride ? stride : 1, ns, md0, x->leaves.as<LsiLeaf>(), cnt.as<uint32_t>(), xs.as<float>(),
                           ys.as<float>());
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
            hipLaunchKernelGGL(k_lsi_adam, dim3(1), dim3(256), 0, s, x->params.as<float>(), m, v, g, H, x->cfg.lr,
                               b1t, b2t);
            HSFM_HIP(hipGetLastError());
        }
        if (x->cfg.train_steps == 0) x->info.mse_final = x->info.mse_initial;
    }
    HSFM_HIP(hipEventRecord(e1, s));
    x->info.train_ms = elapsed_ms(e0, e1);
    xs.release(); ys.release(); gm.release(); lossb.release();

    // stage 2: per leaf the least-squares map of the trained f, its error
    HSFM_HIP(hipEventRecord(e0, s));
    HSFM_HIP(hipMemsetAsync(mom.p, 0, (size_t)L * 32, s));
    const LsiModel md = x->model();
    if (N)
        hipLaunchKernelGGL(k_lsi_moments, cg, dim3(256), 0, s, md, x->keys.as<unsigned long long>(),
                           x->leaves.as<LsiLeaf>(), mom.as<double>());
    hipLaunchKernelGGL(k_lsi_solve, dim3((L + 255) / 256), dim3(256), 0, s, x->leaves.as<LsiLeaf>(),
                       cnt.as<uint32_t>(), mom.as<double>(), L);
    if (N)
        hipLaunchKernelGGL(k_lsi_err, cg, dim3(256), 0, s, md, x->keys.as<unsigned long long>(),
                           x->leaves.as<LsiLeaf>());
    hipLaunchKernelGGL(k_lsi_err_top, dim3((L + 255) / 256), dim3(256), 0, s, md, x->leaves.as<LsiLeaf>());
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
        wsum += (double)hc[i] * (2.0 * h[i].err + 2.0);
    }
    x->info.max_err = emax;
    x->info.mean_window = N ? wsum / (double)N : 0.0;
    cnt.release(); first.release(); mom.release();
    // fence levels: as many as the widest window (2 max_err + 2 keys) needs
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

} // namespace

namespace hsfm {
int lsi_mem_view(const hsfm_lsi *x, hsfm_index *idx, LsiView *out) {
    if (!x) return set_error(HSFM_EINVAL, "learned-seed restarts need a learned seed index");
    if (x->idx != idx) return set_error(HSFM_EINVAL, "the learned seed index was built on another index");
    if (!x->info.mem_seeding)
        return set_error(HSFM_EINVAL, "this learned seed index cannot seed MEM restarts exactly (k > %u, an "
                                      "alphabet whose bases match other lanes, no device text, or text windows "
                                      "with more than %u base strings)", kLsiMemMaxK, kLsiMaxExpansions);
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
    hipStream_t s = idx->dev->stream;
    hipEvent_t e0, e1, e2;
    HSFM_HIP(hipEventCreate(&e0));
    HSFM_HIP(hipEventCreate(&e1));
    HSFM_HIP(hipEventCreate(&e2));
    HSFM_HIP(hipEventRecord(e0, s));
    DevBuf ek, ep;
    uint64_t E = 0;
    int rc = lsi_enumerate(x, ek, ep, E);
    HSFM_HIP(hipEventRecord(e1, s));
    x->info.num_keys = E;
    x->info.enum_ms = elapsed_ms(e0, e1);
    if (!rc && E > 0xfffffffeull) rc = set_error(HSFM_EINVAL, "learned seed index: more than 2^32 - 2 k-mers");
    if (!rc) rc = lsi_restart_entries(x, ek, ep, E);
    ek.release();
    ep.release();
    if (!rc) rc = lsi_finish_entries(x);
    HSFM_HIP(hipEventRecord(e2, s));
    x->info.extra_ms = elapsed_ms(e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (!rc) rc = lsi_train_and_fit(x);
    if (rc) {
        x->release();
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
    x->release();
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
    // leaves in LDS when they fit and at most two fence levels are needed (measured:
    // 100 Mbp, 4096 leaves, 0.66 ms against 0.88 ms; 1 Gbp, whose windows need
    // a third level, 1.16 ms against 1.12 ms)
    const bool lds = x->cfg.lookup == 2 || (x->cfg.lookup == 0 && x->cfg.leaves <= kLsiLdsLeaves && x->nfence <= 2);
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
    const bool lds = x->cfg.lookup == 2 || (x->cfg.lookup == 0 && x->cfg.leaves <= kLsiLdsLeaves && x->nfence <= 2);
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
