// hsfm_mem.hip -- SMEM / MEM finding, resolve and reference mapping on gfx950.
//
// Output contract (BidirFmIndex::find_smems / find_mems, src/fm_index/smem.rs:71-124,
// in the MemHit shape of resolve_mem_hits_gpu, src/fm_index/bidir_index.rs:391-514):
// the MEMs of a query are the pairs (i, R(i)) where R(i) is the end of the longest
// IUPAC-compatible match starting at i, R(i) - i >= min_len and i is left-maximal
// (i == 0 or R(i-1) < R(i)).  Because R is non-decreasing, no such MEM contains
// another, so the containment filter of find_smems (smem.rs:87-103) keeps them
// all: find_smems and find_mems return the same spans, counts and positions
// (tests pin this on the oracle, including the randomized smem.rs:805-870 case).
//
// Algorithm per read (one wave, interval sets cooperatively extended):
//   s0 = 0, E = 0
//   loop: j = max(s0 + min_len, E + 1); stop if j > |q|
//     backward: extend the set of q[k..j) left from k = j while it stays non-empty
//               -> k = L(j), the start of the longest match ending at j
//     if k > j - min_len: no MEM of length >= min_len starts in [s0, k); s0 = k
//     else: extend right from j -> e = R(k); emit (k, e) with its intervals;
//           s0 = k + 1, E = e
// Every interval is one concrete (IUPAC-compatible) string, so the bidirectional
// pairing of bidir.rs:80-130 holds; count_smaller_than (bidir.rs:136-143) runs
// over every lane below the extension symbol.
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "hsfm_host.h"
#include "hsfm_lsi_dev.h"

namespace hsfm {

struct MemCounters {
    unsigned long long work;
    unsigned long long rec_used;
    unsigned long long n_retry;
    unsigned long long guard;
    unsigned long long probes;   // records fetched by the main kernel
    unsigned long long diag[4];  // deferrals: seed union, extension, slots; plain restarts
    unsigned long long pwork;    // work index of k_mem_persist
    unsigned long long pos_ovf;  // records whose rows did not fit the position buffers
    unsigned long long pad[5];
};

struct MemRec { // 32 bytes
    uint32_t read, local, qs, qe, lo, hi, pass, pad;
};

// Inclusive prefix sum over the TB threads of a block (TB a multiple of 64);
// *total receives the block sum.  s_part holds TB/64 + 1 words.
template <int TB>
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *s_part, uint32_t &total) {
    const uint32_t incl = wave_incl_scan(v);
    if (TB == 64) {
        total = __shfl(incl, 63, 64);
        return incl;
    }
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) s_part[wv] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t u = 0; u < (uint32_t)TB / 64; u++) {
        const uint32_t x = s_part[u];
        before += u < wv ? x : 0u;
        all += x;
    }
    __syncthreads();
    total = all;
    return before + incl;
}
// Value of thread 0, for every thread of the block.
template <int TB, typename T>
__device__ __forceinline__ T block_bcast(T v, T *s_b) {
    if (TB == 64) return __shfl(v, 0, 64);
    if (threadIdx.x == 0) *s_b = v;
    __syncthreads();
    const T x = *s_b;
    __syncthreads();
    return x;
}

// Extend every interval of A (ns entries) by query code c, TB intervals at a
// time.  LEFT: P -> cP on the forward index; else P -> Pc on the reverse
// index.  Writes the non-empty children to B and returns their number; sets
// too_big when cap is exceeded.
template <bool LEFT, int TB>
__device__ uint32_t group_extend(const DevDir &D, const uint32_t *s_clane, uint32_t cm, const uint4 *A, uint32_t ns,
                                 uint4 *B, uint32_t cap, bool &too_big, uint32_t *s_part) {
    uint32_t total = 0;
    for (uint32_t base = 0; base < ns; base += TB) {
        const uint32_t j = base + threadIdx.x;
        uint32_t cl[16], ch[16], cc = 0;
        uint4 iv = make_uint4(0, 0, 0, 0);
        if (j < ns) {
            iv = A[j];
            const uint32_t size = iv.y - iv.x;
            const uint32_t xlo = LEFT ? iv.x : iv.z;
            const uint32_t xhi = LEFT ? iv.y : iv.z + size;
            pair_counts(D, xlo, xhi, cl, ch);
#pragma unroll
            for (int k = 0; k < 16; k++) cc += (((cm >> k) & 1u) && ch[k] > cl[k]) ? 1u : 0u;
        }
        uint32_t tot;
        const uint32_t incl = block_incl_scan<TB>(cc, s_part, tot);
        if (total + tot > cap) {
            too_big = true;
            return total;
        }
        if (j < ns && cc) {
            uint32_t w = total + incl - cc;
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const uint32_t d = ch[k] - cl[k];
                if (((cm >> k) & 1u) && ch[k] > cl[k]) {
                    if (LEFT) B[w++] = make_uint4(s_clane[k] + cl[k], s_clane[k] + ch[k], iv.z + acc, 0u);
                    else B[w++] = make_uint4(iv.x + acc, iv.x + acc + d, s_clane[k] + cl[k], 0u);
                }
                acc += d;
            }
        }
        total += tot;
    }
    return total;
}

// One workgroup of TB threads per read.  LDS: both sets in LDS (kWaveLdsCap
// intervals each); otherwise in global memory, cap intervals each.  Restarts
// take q[j-F..j) from the seed table when min_len > F (up to kSeedAmb
// ambiguity codes expanded).
constexpr uint32_t kWaveLdsCap = 1024;
// Threads per read for reads the persistent kernel defers: their sets are
// wider than its ring, so a step extends up to 256 intervals at once instead
// of 64 per dependent round.
constexpr int kDeferTB = 256;
// Parts per deferred item in the first workgroup pass (at most; see k_mem_wave)
// and the pass_of mark of an item one of whose parts failed.
constexpr uint32_t kDeferParts = 64;
constexpr uint32_t kPartFailed = 0xfffffffeu;
constexpr uint32_t kPartCntMax = 65536; // part counts a batch set holds (workgroups of the deferred pass)

template <bool LDS, int TB>
__global__ __launch_bounds__(TB) void k_mem_wave(DevDir fwd, DevDir rev, const DevConst *__restrict__ cst,
                                                 SeedView sv, const uint8_t *__restrict__ q,
                                                 const uint64_t *__restrict__ qoff,
                                                 const uint32_t *__restrict__ list, uint32_t nlist,
                                                 const uint4 *__restrict__ items, uint32_t min_len,
                                                 uint32_t cap, uint4 *__restrict__ sets, MemRec *__restrict__ recs,
                                                 uint64_t rec_cap, uint32_t *__restrict__ mem_cnt,
                                                 uint32_t *__restrict__ pass_of, uint32_t pass,
                                                 uint32_t *__restrict__ retry_list, MemCounters *__restrict__ ctr,
                                                 const unsigned long long *__restrict__ d_nlist,
                                                 uint32_t *__restrict__ part_cnt, uint32_t parts_max) {
    __shared__ uint16_t s_compat[256];
    __shared__ uint8_t s_expand[256];
    __shared__ uint32_t s_clane[16];
    __shared__ uint4 s_sets[LDS ? 2 * kWaveLdsCap : 1];
    __shared__ uint32_t s_part[TB / 64 + 1];
    __shared__ unsigned long long s_b64;
    __shared__ uint32_t s_b32, s_lsl;
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        s_compat[i] = cst->compat_lanes[i];
        s_expand[i] = cst->seed_expand[i];
    }
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint32_t lane = threadIdx.x;
    const uint32_t n = fwd.n;
    const uint32_t F = sv.F;
    if (LDS) cap = kWaveLdsCap;
    if (d_nlist) nlist = *d_nlist < nlist ? (uint32_t)*d_nlist : nlist; // count left by the previous kernel
    // A short list (the few items the persistent kernel defers) would leave
    // most workgroups idle while each item runs one long dependent chain: its
    // MEM starts are split into P parts (a power of two, at most parts_max,
    // with every part on a workgroup of this grid), each swept like an item of
    // its own.  Part counts go to part_cnt[work index]; the scatter places a
    // part's records behind those of the item's earlier parts.
    uint32_t P = 1;
    if (part_cnt)
        while (P * 2 <= parts_max && (uint64_t)nlist * P * 2 <= gridDim.x) P *= 2;
    const uint32_t nwork = nlist * P; // P > 1 only when nlist * P <= gridDim.x
    uint4 *bufA = LDS ? s_sets : sets + (size_t)blockIdx.x * 2 * cap;
    uint4 *bufB = bufA + cap;
    for (;;) {
        uint32_t wi = 0;
        if (lane == 0) wi = (uint32_t)atomicAdd(&ctr->work, 1ull);
        wi = block_bcast<TB>(wi, &s_b32);
        if (wi >= nwork) break;
        const uint32_t li = wi / P, part = wi % P;
        const uint32_t w = list ? list[li] : li;
        const uint4 it = items ? items[w] : make_uint4(w, 0u, 0xffffffffu, 0u);
        const uint32_t r = it.x;
        const uint64_t b0 = qoff[r];
        const uint32_t m = (uint32_t)(qoff[r + 1] - b0);
        const uint8_t *p = q + b0;
        uint32_t ia = it.y, ib = it.z < m ? it.z : m; // starts owned by this item
        if (P > 1) { // starts owned by this part
            const uint64_t span = ib > ia ? ib - ia : 0u;
            const uint32_t a = ia + (uint32_t)(span * part / P);
            ib = ia + (uint32_t)(span * (part + 1) / P);
            ia = a;
        }
        uint32_t s0 = ia, E = 0, nrec = 0;
        bool too_big = false;
        while (min_len && !too_big) {
            const uint32_t j = (s0 + min_len > E + 1) ? s0 + min_len : E + 1;
            if (j > m || s0 >= ib) break;
            uint4 *A = bufA, *B = bufB;
            uint32_t ns = 1, k = j;
            bool seeded = false;
            if (F && min_len > F && j >= F) {
                // q[j-F..j) from the table: the union of its expansions' entries,
                // the expansions spread over the workgroup's threads
                WideWin sw;
                if (wide_window(p + j - F, F, s_expand, sw)) {
                    const uint32_t namb = sw.namb, nc = 1u << (2 * namb);
                    if (lane == 0) s_lsl = 0;
                    __syncthreads();
                    uint32_t total = 0;
                    for (uint32_t x0 = 0; x0 < nc; x0 += TB) {
                        const uint32_t xi = x0 + lane;
                        uint32_t idx = 0, cnt = 0;
                        uint64_t o0 = 0;
                        uint4 one = make_uint4(0, 0, 0, 0);
                        if (xi < nc && wide_combo(sw, xi, idx)) {
                            if (sv.rec) { // packed record: one load per expansion
                                const uint4 R = sv.rec[idx];
                                const uint32_t l = seed_rec_lsl(R.w);
                                atomicMax(&s_lsl, l);
                                if (l == F) {
                                    cnt = seed_rec_cnt(R.w);
                                    o0 = (uint64_t)R.x | ((uint64_t)R.y << 32);
                                    one = make_uint4(R.x, R.y, R.z, 0u);
                                }
                            } else {
                                const uint32_t l = sv.lsl[idx];
                                atomicMax(&s_lsl, l);
                                if (l == F) {
                                    o0 = sv.off[idx];
                                    cnt = (uint32_t)(sv.off[idx + 1] - o0);
                                }
                            }
                        }
                        uint32_t tot;
                        const uint32_t at = total + block_incl_scan<TB>(cnt, s_part, tot) - cnt;
                        if (sv.rec && cnt == 1) {
                            if (at < cap) A[at] = one;
                        } else {
                            for (uint32_t t = 0; t < cnt && at + t < cap; t++) A[at + t] = sv.iv[o0 + t];
                        }
                        total += tot;
                    }
                    __syncthreads();
                    const uint32_t lslmax = s_lsl;
                    __syncthreads();
                    if (lslmax < F) { // q[j-lslmax-1..j) does not occur: L(j) = j - lslmax
                        s0 = j - lslmax;
                        continue;
                    }
                    if (total > cap) {
                        too_big = true;
                        break;
                    }
                    __syncthreads();
                    ns = total;
                    if (namb) { // drop entries repeated across expansions (same fwd_lo)
                        uint32_t kept = 0;
                        for (uint32_t base = 0; base < total; base += TB) {
                            const uint32_t t = base + lane;
                            bool keep = t < total;
                            uint4 iv = make_uint4(0, 0, 0, 0);
                            if (keep) {
                                iv = A[t];
                                for (uint32_t u = 0; u < t && keep; u++) keep = A[u].x != iv.x;
                            }
                            uint32_t tot;
                            const uint32_t incl = block_incl_scan<TB>(keep ? 1u : 0u, s_part, tot);
                            if (keep) B[kept + incl - 1] = iv;
                            kept += tot;
                        }
                        __syncthreads();
                        uint4 *tmp = A; A = B; B = tmp;
                        ns = kept;
                    }
                    k = j - F;
                    seeded = true;
                }
            }
            if (!seeded && lane == 0) A[0] = make_uint4(0, n, 0, 0);
            __syncthreads();
            while (k > 0) {
                if (fwd.text && fwd.sa_full && fwd.isa && ns == 1 && A[0].y - A[0].x == 1) {
                    // a single row: its only occurrence extends left while the
                    // text is compatible (64 codes per step, first wave; the
                    // sentinels stop it); the row of the longer string is the
                    // inverse suffix array at its text start
                    if (threadIdx.x < 64) {
                        const uint32_t tk = fwd.sa_full[A[0].x]; // text position of query code k
                        uint32_t kk = k;
                        for (;;) {
                            const uint32_t back = (k - kk) + 1u + lane; // code kk-1-lane sits at tk - back
                            bool ok = lane < kk && back <= tk;
                            if (ok) ok = (s_compat[p[kk - 1u - lane]] >> fwd.text[tk - back]) & 1u;
                            const uint64_t bad = __ballot(!ok);
                            const uint32_t adv = bad ? (uint32_t)__ffsll((unsigned long long)bad) - 1u : 64u;
                            kk -= adv;
                            if (adv < 64) break;
                        }
                        if (threadIdx.x == 0) {
                            const uint32_t row = fwd.isa[tk - (k - kk)];
                            A[0] = make_uint4(row, row + 1u, 0u, 0u);
                            s_b32 = kk;
                        }
                    }
                    __syncthreads();
                    k = s_b32;
                    __syncthreads();
                    break;
                }
                const uint32_t c2 = group_extend<true, TB>(fwd, s_clane, s_compat[p[k - 1]], A, ns, B, cap, too_big,
                                                           s_part);
                __syncthreads();
                if (too_big || c2 == 0) break;
                uint4 *t = A; A = B; B = t;
                ns = c2;
                k--;
            }
            if (too_big) break;
            if (k + min_len > j) { // k > j - min_len
                s0 = k;
                continue;
            }
            uint32_t e = j;
            while (e < m) {
                if (fwd.text && fwd.sa_full && ns == 1 && A[0].y - A[0].x == 1) {
                    // a single row: its only occurrence continues while the text
                    // is compatible (64 codes per step, first wave); the row and
                    // so the record stay the same
                    if (threadIdx.x < 64) {
                        const uint32_t toff = fwd.sa_full[A[0].x] - k; // text position of code e: toff + e
                        for (;;) {
                            const uint32_t x = e + lane;
                            bool ok = x < m && toff + x < fwd.n; // the sentinel at n - 1 stops it first
                            if (ok) ok = (s_compat[p[x]] >> fwd.text[toff + x]) & 1u;
                            const uint64_t bad = __ballot(!ok);
                            const uint32_t adv = bad ? (uint32_t)__ffsll((unsigned long long)bad) - 1u : 64u;
                            e += adv;
                            if (adv < 64) break;
                        }
                        if (threadIdx.x == 0) s_b32 = e;
                    }
                    __syncthreads();
                    e = s_b32;
                    __syncthreads();
                    break;
                }
                const uint32_t c2 = group_extend<false, TB>(rev, s_clane, s_compat[p[e]], A, ns, B, cap, too_big,
                                                            s_part);
                __syncthreads();
                if (too_big || c2 == 0) break;
                uint4 *t = A; A = B; B = t;
                ns = c2;
                e++;
            }
            if (too_big) break;
            if (k >= ib) break; // starts in a later segment
            if (k < ia) {       // starts in an earlier segment: not ours, scan on
                s0 = ia;
                E = e;
                continue;
            }
            // emit (k, e): one record per interval, sorted by fwd_lo
            unsigned long long slot = 0;
            if (lane == 0) slot = atomicAdd(&ctr->rec_used, (unsigned long long)ns);
            slot = block_bcast<TB>(slot, &s_b64);
            if (slot + ns > rec_cap) {
                // mark the part of the reservation that fits as dead records
                for (uint64_t t = slot + lane; t < slot + ns && t < rec_cap; t += TB) {
                    recs[t].read = 0xffffffffu;
                    recs[t].pass = 0xffffffffu;
                }
                too_big = true;
                break;
            }
            for (uint32_t t = lane; t < ns; t += TB) {
                const uint4 iv = A[t];
                uint32_t rank = 0;
                for (uint32_t u = 0; u < ns; u++) rank += A[u].x < iv.x ? 1u : 0u;
                MemRec rc;
                rc.read = w;
                rc.local = nrec + rank;
                rc.qs = k;
                rc.qe = e;
                rc.lo = iv.x;
                rc.hi = iv.y;
                rc.pass = pass;
                rc.pad = P > 1 ? (part << 24) | (wi + 1u) : 0u; // wi + 1 <= gridDim.x < 2^24
                recs[slot + rank] = rc;
            }
            nrec += ns;
            s0 = k + 1;
            E = e;
            __syncthreads();
        }
        if (lane == 0) {
            if (P > 1) {
                // the item succeeds in this pass when every part does: a failed
                // part marks it (the first one lists it for the retry pass,
                // which runs the whole item), a successful part adds its count
                // and sets the pass unless a failure came first
                part_cnt[wi] = too_big ? 0u : nrec;
                if (too_big) {
                    if (atomicExch(&pass_of[w], kPartFailed) != kPartFailed) {
                        const unsigned long long at = atomicAdd(&ctr->n_retry, 1ull);
                        retry_list[at] = w;
                    }
                } else {
                    atomicAdd(&mem_cnt[w], nrec);
                    atomicCAS(&pass_of[w], 0xffffffffu, pass);
                }
            } else if (too_big) {
                const unsigned long long at = atomicAdd(&ctr->n_retry, 1ull);
                retry_list[at] = w;
            } else {
                mem_cnt[w] = nrec;
                pass_of[w] = pass;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Seed table: for every ACGT F-mer W (index = sum digit(W[t]) * 4^t, leftmost
// code least significant as in lookup.rs:101-111) the bidirectional intervals
// of all compatible concrete strings (empty when W does not occur) and lsl(W),
// the length of W's longest occurring suffix.  Built level by level on the
// device: the node for a string of length l is its first code prepended to the
// node of its suffix of length l-1 (one left extension, bidir.rs:112-130).
// ---------------------------------------------------------------------------
template <bool WRITE>
__global__ __launch_bounds__(256) void k_seed_level(DevDir fwd, const DevConst *__restrict__ cst,
                                                    const uint64_t *__restrict__ poff, const uint4 *__restrict__ piv,
                                                    uint32_t nchild, uint32_t *__restrict__ cnt,
                                                    const uint64_t *__restrict__ coff, uint4 *__restrict__ civ) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchild) return;
    const uint32_t L = cst->num_lanes;
    const uint32_t cm = cst->compat_lanes[1 + (c & 3)];
    const uint32_t p = c >> 2;
    uint32_t total = 0;
    uint64_t w = WRITE ? coff[c] : 0;
    for (uint64_t i = poff[p]; i < poff[p + 1]; i++) {
        const uint4 iv = piv[i];
        uint32_t cl[16], ch[16];
        interval_counts<true>(fwd, L, iv, cl, ch);
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t d = ch[k] - cl[k];
            if (((cm >> k) & 1u) && ch[k] > cl[k]) {
                if (WRITE) civ[w++] = bidir_child<true>(iv, s_clane[k], cl[k], ch[k], acc);
                total++;
            }
            acc += d;
        }
    }
    if (!WRITE) cnt[c] = total;
}

__global__ void k_seed_lsl(uint32_t nchild, uint32_t level, const uint32_t *__restrict__ cnt,
                           const uint8_t *__restrict__ plsl, uint8_t *__restrict__ clsl) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchild) return;
    clsl[c] = cnt[c] ? (uint8_t)level : plsl[c >> 2];
}

__global__ void k_seed_root(uint64_t *off, uint4 *iv, uint8_t *lsl, uint32_t n) {
    off[0] = 0;
    off[1] = 1;
    iv[0] = make_uint4(0, n, 0, 0);
    lsl[0] = 0;
}


// ---------------------------------------------------------------------------
// Persistent MEM finder: the min_len-skip algorithm of k_mem_wave run as a
// per-lane state machine, one work item (read segment) per lane.  Restarts seed q[j-F..j) from the packed F-mer table
// (F = 12 or 14): one record load gives lsl, the interval count and, for most
// F-mers, the single interval.  Every pass of the main loop performs at most one
// Occ probe per lane (one interval of the current set, or one seed-table entry
// streamed into its first left extension), so lanes of a wave stay on the same
// code path whatever phase their read is in; a lane that finishes a read takes
// the next one from a device work queue.  The interval sets of a lane live in
// a ring of PR LDS slots: the current set at [head, head+ns), its children
// appended behind it, so a set may hold up to PR intervals less the parents
// still unprocessed; reads that outgrow the ring are deferred.
// ---------------------------------------------------------------------------
enum : uint32_t {
    SM_FETCH = 0, // take the next read from the queue
    SM_CAND,      // next restart point j
    SM_RESTART,   // seed the restart at j from the tables, or start it unseeded
    SM_SEEDX,     // next seed-table expansion of q[j-F..j)
    SM_ENDLEFT,   // left extension ended at k
    SM_EMIT,      // right extension ended at e: write the records
    SM_LSI_NEXT,  // a learned lookup's base string is done: the next one, or the answer
    SM_SEEDE,     // probe: one seed entry, extended left by q[j-F-1]
    SM_SEEDR,     // load: the packed seed record of one expansion
    SM_TSA,       // load: text position of a single-row set (right extension)
    SM_TEXT,      // load: 8 text and 8 query codes, compared directly
    SM_LEFT,      // probe: one interval of the set of q[k..j), by q[k-1]
    SM_RIGHT,     // probe: one interval of the set of q[k..e), by q[e]
    SM_LSI_LEAF,  // load: the learned index's leaf of the k-mer q[j-K..j)
    SM_LSI_FENCE, // loads: one fence level of its bounded search
    SM_LSI_KEYS,  // loads: up to 18 keys of its last round (two passes at most)
    SM_LSI_PAY,   // load: the payload of one entry of the run with the k-mer's codes
    SM_DONE
};

// The K (<= 25) codes of q[g..g+K) as a learned-index key template (two bits
// per base from the top of the word, last code first; 0 at ambiguity codes)
// and the ambiguity codes, at most three: positions in bits 5a.., base sets
// in bits 15 + 4a.., their number in bits 27..  False if a code has no base
// or there are more ambiguity codes.
__device__ __forceinline__ bool lsi_window_amb(const uint32_t *__restrict__ nib, uint64_t g, uint32_t K,
                                               unsigned long long &tmpl, uint32_t &amb) {
    const uint32_t *b = nib + (g >> 3);
    const uint64_t lo = (uint64_t)b[0] | ((uint64_t)b[1] << 32), hi = (uint64_t)b[2] | ((uint64_t)b[3] << 32);
    const uint32_t sh = 4 * (uint32_t)(g & 7);
    uint64_t x0 = sh ? (lo >> sh) | (hi << (64 - sh)) : lo; // codes 0..15 of the window
    uint64_t x1 = hi >> sh;                                 // codes 16..
    bool ok = true;
    unsigned long long k = 0;
    uint32_t a = 0, na = 0;
#pragma unroll
    for (uint32_t i = 0; i < 25; i++) {
        if (i < K) {
            const uint32_t nbl = (uint32_t)((i < 16 ? x0 >> (4 * i) : x1 >> (4 * (i - 16))) & 15u);
            if ((nbl & (nbl - 1)) == 0) {
                ok &= nbl != 0;
                k |= (unsigned long long)((nbl >> 1) - (nbl >> 3)) << (64 - 2 * K + 2 * i); // 1, 2, 4, 8 -> 0..3
            } else {
                ok &= na < 3;
                if (na < 3) a |= (i << (5 * na)) | (nbl << (15 + 4 * na));
                na++;
            }
        }
    }
    tmpl = k;
    amb = a | ((na < 3 ? na : 3u) << 27);
    return ok;
}
// base strings of a window: the product of its ambiguity codes' base counts
__device__ __forceinline__ uint32_t lsi_num_expansions(uint32_t amb) {
    uint32_t c = 1;
    for (uint32_t a = 0; a < (amb >> 27); a++) c *= (uint32_t)__popc((amb >> (15 + 4 * a)) & 15u);
    return c;
}
// key of expansion x: the x-th combination of the ambiguity codes' bases
__device__ __forceinline__ unsigned long long lsi_expansion(unsigned long long tmpl, uint32_t amb, uint32_t K,
                                                            uint32_t x) {
    for (uint32_t a = 0; a < (amb >> 27); a++) {
        uint32_t m = (amb >> (15 + 4 * a)) & 15u;
        const uint32_t c = (uint32_t)__popc(m), pick = x % c;
        x /= c;
        for (uint32_t z = 0; z < pick; z++) m &= m - 1;
        const uint32_t pos = (amb >> (5 * a)) & 31u;
        tmpl |= (unsigned long long)(__ffs(m) - 1) << (64 - 2 * K + 2 * pos);
    }
    return tmpl;
}

// Base-set nibbles of the flat query array (eight codes per word; words past
// the end hold zero nibbles).
__global__ __launch_bounds__(256) void k_pack_bases(const uint8_t *__restrict__ q, uint64_t len,
                                                    const DevConst *__restrict__ cst, uint32_t *__restrict__ nib,
                                                    uint64_t nwords) {
    __shared__ uint8_t s_expand[256];
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_expand[i] = cst->seed_expand[i];
    __syncthreads();
    // two words (16 codes) per thread; whole 16-byte groups with one load when
    // the array is 16-byte aligned
    const bool vec = (reinterpret_cast<uintptr_t>(q) & 15) == 0;
#pragma unroll 4
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * t < nwords;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t g0 = t * 16;
        uint32_t b[4] = {0, 0, 0, 0};
        if (vec && g0 + 16 <= len) {
            const uint4 v = *reinterpret_cast<const uint4 *>(q + g0);
            b[0] = v.x; b[1] = v.y; b[2] = v.z; b[3] = v.w;
        } else {
            for (uint32_t u = 0; u < 16; u++)
                if (g0 + u < len) b[u >> 2] |= (uint32_t)q[g0 + u] << (8 * (u & 3));
        }
        uint32_t w0 = 0, w1 = 0;
#pragma unroll
        for (uint32_t u = 0; u < 16; u++) {
            const uint32_t nb = g0 + u < len ? base_nibble((b[u >> 2] >> (8 * (u & 3))) & 0xffu, s_expand) : 0u;
            if (u < 8) w0 |= nb << (4 * u);
            else w1 |= nb << (4 * (u - 8));
        }
        if (2 * t + 1 < nwords) *reinterpret_cast<uint2 *>(nib + 2 * t) = make_uint2(w0, w1);
        else nib[2 * t] = w0;
    }
}

int pack_bases(hsfm_index *idx, const uint8_t *d_q, uint64_t len, DevBuf &out) {
    const uint64_t nwords = (len + 7) / 8 + 4; // +4: the window reads three words from any start
    int rc = out.ensure(nwords * 4);
    if (rc) return rc;
    // one resident round of blocks, each thread striding over several groups
    uint64_t blocks = (nwords / 2 + 256) / 256;
    if (blocks > (uint64_t)idx->dev->num_cus * 8) blocks = (uint64_t)idx->dev->num_cus * 8;
    hipLaunchKernelGGL(k_pack_bases, dim3((uint32_t)blocks), dim3(256), 0, idx->dev->stream, d_q, len,
                       idx->d_const.as<DevConst>(), out.as<uint32_t>(), nwords);
    HSFM_HIP(hipGetLastError());
    return HSFM_OK;
}

constexpr uint32_t kPersistChunk = 64; // work indices a wave claims per atomic
// 256-thread blocks of the ring kernel fit three per CU by LDS; the second
// launch bound keeps the registers within the same three waves per SIMD.

template <int PR, int PB, bool LSI = false>
__global__ __launch_bounds__(PB, PB >= 256 ? 3 : 1) void k_mem_persist(DevDir fwd, DevDir rev, const DevConst *__restrict__ cst,
                                                    SeedView sv, const uint8_t *__restrict__ q,
                                                    const uint32_t *__restrict__ qnib,
                                                    const uint64_t *__restrict__ qoff,
                                                    const uint32_t *__restrict__ list, uint32_t nlist,
                                                    const uint4 *__restrict__ items,
                                                    uint32_t min_len, uint4 *__restrict__ slots,
                                                    uint32_t *__restrict__ mem_cnt, uint8_t *__restrict__ status,
                                                    uint32_t *__restrict__ defer_list,
                                                    MemCounters *__restrict__ ctr, LsiView lv, uint32_t lsik) {
    static_assert((PR & (PR - 1)) == 0, "ring size must be a power of two");
    __shared__ uint32_t s_lo[PR][PB], s_hi[PR][PB], s_rl[PR][PB];
    constexpr uint32_t RM = PR - 1;
    __shared__ uint16_t s_compat[256];
    __shared__ uint8_t s_expand[256];
    __shared__ uint32_t s_clane[16];
    // per wave: the unclaimed rest [s_qb, s_qb + s_ql) of its chunk of work indices
    __shared__ uint32_t s_qb[PB / 64], s_ql[PB / 64];
    __shared__ float s_lp[LSI ? 3 * kLsiMaxHidden + 1 : 1]; // stage 1 of the learned index
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        s_compat[i] = cst->compat_lanes[i];
        s_expand[i] = cst->seed_expand[i];
    }
    if (LSI && lsik)
        for (uint32_t i = threadIdx.x; i < 3 * lv.md.hidden + 1; i += blockDim.x) s_lp[i] = lv.md.p[i];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    if (threadIdx.x < PB / 64) {
        s_qb[threadIdx.x] = 0;
        s_ql[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = tid >> 6;
    const uint32_t n = fwd.n;
    const uint32_t F = sv.F;
    const bool seeding = F != 0 && min_len > F;
    // right extensions of a single row compare the text directly: the row's
    // only occurrence continues while the text lane is compatible with the
    // query code (no code is compatible with the sentinel, so it stops there
    // as the index extension does), and the row stays the same
    const bool textmode = fwd.text != nullptr && fwd.sa_full != nullptr;
    // compare 16 codes per load pair as base-set nibbles (text4 against the
    // query nibble stream) where the alphabet allows it, else 8 as bytes
    const bool nibtext = fwd.text4 != nullptr;

    uint32_t st = SM_FETCH, w = 0, r = 0, m = 0, ia = 0, ib = 0, s0 = 0, E = 0, j = 0, k = 0, e = 0;
    uint32_t ns = 0, nn = 0, t = 0, head = 0, nrec = 0, slot_cap = 0;
    uint32_t xi = 0, lslmax = 0, probes = 0, restarts = 0, sidx = 0;
    SeedWin sw{0, 0, 0x111u, 0};
    uint64_t slot_base = 0, ient = 0, iend = 0;
    const uint8_t *p = q;
    uint4 piv = make_uint4(0, 0, 0, 0);
    // learned-seed lookup of q[j-K..j): key of the current base string, search
    // range [llo, lhi) (slots < 2^32; the lower bound lies in [llo, lhi]),
    // fence level, key-round part or run entry, keys below the key so far and
    // the last of them, the first key not below it (~0: none yet); the
    // window's key template, ambiguity codes and longest occurring suffix so
    // far live in piv, the base string's number in t, the set built so far in
    // the ring (ns intervals).  They alias state that is dead from SM_CAND
    // until the lookup ends (seed entries and windows, the right end, the
    // interval being extended), which keeps the kernel within its register
    // budget.
    uint64_t &lkey = ient, &lpv = iend;
    uint32_t &llo = xi, &lhi = lslmax, &llv = sidx, &lbelow = e;
    unsigned long long lsv = ~0ull;

    for (;;) {
        // ---- control: move this lane to a memory state (or SM_DONE) ----
        while (st < SM_SEEDE) {
            if (st == SM_FETCH) {
                // work indices come from the wave's chunk; one device atomic
                // claims the next kPersistChunk when the chunk runs out, so the
                // shared counter sees one atomic per chunk, not one per read
                const uint64_t act = __ballot(1);
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                const uint32_t need = (uint32_t)__popcll(act);
                uint32_t cb = s_qb[wv], cl = s_ql[wv], wi;
                if (need <= cl) {
                    wi = cb + rank;
                    cb += need;
                    cl -= need;
                } else {
                    uint32_t nb = 0;
                    if (__lane_id() == leader) nb = (uint32_t)atomicAdd(&ctr->pwork, (unsigned long long)kPersistChunk);
                    nb = __shfl(nb, (int)leader, 64);
                    wi = rank < cl ? cb + rank : nb + (rank - cl);
                    cb = nb + (need - cl);
                    cl = kPersistChunk - (need - cl);
                }
                if (__lane_id() == leader) {
                    s_qb[wv] = cb;
                    s_ql[wv] = cl;
                }
                if (wi >= nlist) { st = SM_DONE; break; }
                w = list ? list[wi] : wi;
                uint4 it = make_uint4(w, 0u, 0xffffffffu, 0u);
                if (items) it = items[w];
                r = it.x;
                const uint64_t b0 = qoff[r];
                m = (uint32_t)(qoff[r + 1] - b0);
                p = q + b0;
                ia = it.y;
                ib = it.z < m ? it.z : m;
                slot_base = ((b0 + ia) >> 3) + 8ull * w;
                slot_cap = (uint32_t)(((b0 + ib) >> 3) - ((b0 + ia) >> 3)) + 8u;
                s0 = ia; E = 0; nrec = 0;
                st = SM_CAND;
            } else if (st == SM_CAND || st == SM_RESTART) {
                // SM_RESTART: a learned-seed lookup that did not give the set
                if (st == SM_CAND) {
                    j = (s0 + min_len > E + 1) ? s0 + min_len : E + 1;
                    if (min_len == 0 || j > m || s0 >= ib) { // no MEM starts in [s0, ib)
                        status[w] = 0;
                        mem_cnt[w] = nrec;
                        st = SM_FETCH;
                        continue;
                    }
                    // learned seed of q[j-K..j) (at most three ambiguity codes,
                    // one lookup per base string): the union of the base
                    // strings' sets, or, when every one is empty, the longest
                    // occurring suffix of any of them, which gives L(j)
                    unsigned long long tmpl = 0;
                    uint32_t amb = 0;
                    if (LSI && lsik && j >= lsik &&
                        lsi_window_amb(qnib, (uint64_t)(p - q) + j - lsik, lsik, tmpl, amb)) {
                        piv = make_uint4((uint32_t)tmpl, (uint32_t)(tmpl >> 32), amb, 0u);
                        t = 0;
                        ns = 0;
                        lkey = lsi_expansion(tmpl, amb, lsik, 0);
                        st = SM_LSI_LEAF;
                        continue;
                    }
                }
                // up to three ambiguity codes (64 expansions, one record load
                // each); windows with more restart unseeded
                bool usable = false;
                if (seeding && j > F) {
                    const uint64_t g = (uint64_t)(p - q) + j - F; // flat position of the window
                    uint32_t w0, w1, w2;
                    seed_window_nib_load(qnib, g, w0, w1, w2);
                    usable = seed_window_nib(g, F, w0, w1, w2, sw, kSeedAmb);
                }
                nn = 0;
                t = 0;
                if (usable) {
                    ns = 0; // seed entries come from the table, not the ring
                    xi = 0;
                    lslmax = 0;
                    st = SM_SEEDX;
                } else {
                    restarts++;
                    s_lo[head][tid] = 0;
                    s_hi[head][tid] = n;
                    s_rl[head][tid] = 0;
                    ns = 1; k = j;
                    piv = make_uint4(0, n, 0, 0);
                    st = SM_LEFT;
                }
            } else if (st == SM_SEEDX) {
                uint32_t idx = 0;
                while (xi < seed_ncombo(sw) && !seed_combo(sw, xi, idx)) xi++;
                if (xi >= seed_ncombo(sw)) {
                    if (lslmax < F) { // q[j-lslmax-1..j) does not occur: L(j) = j - lslmax
                        s0 = j - lslmax;
                        st = SM_CAND;
                    } else if (nn == 0) { // q[j-F..j) occurs, q[j-F-1..j) does not: L(j) = j - F
                        s0 = j - F;
                        st = SM_CAND;
                    } else {
                        ns = nn; nn = 0; t = 0;
                        k = j - F - 1;
                        if (k == 0) st = SM_ENDLEFT;
                        else {
                            piv = make_uint4(s_lo[head][tid], s_hi[head][tid], s_rl[head][tid], 0u);
                            st = SM_LEFT;
                        }
                    }
                    continue;
                }
                xi++;
                sidx = idx; // its record is loaded with the other lanes' probes
                st = SM_SEEDR;
            } else if (st == SM_ENDLEFT) {
                if (k + min_len > j) {
                    s0 = k;
                    st = SM_CAND;
                    continue;
                }
                e = j; t = 0; nn = 0;
                if (e < m) {
                    piv = make_uint4(s_lo[head][tid], s_hi[head][tid], s_rl[head][tid], 0u);
                    st = (textmode && ns == 1 && piv.y - piv.x == 1) ? SM_TSA : SM_RIGHT;
                } else {
                    st = SM_EMIT;
                }
            } else if (LSI && st == SM_LSI_NEXT) {
                // this base string is done: the next one, or the window's answer
                const unsigned long long tmpl = (unsigned long long)piv.x | ((unsigned long long)piv.y << 32);
                if (++t < lsi_num_expansions(piv.z)) {
                    lkey = lsi_expansion(tmpl, piv.z, lsik, t);
                    st = SM_LSI_LEAF;
                } else if (ns > 0) { // the set of q[j-K..j)
                    nn = 0;
                    t = 0;
                    k = j - lsik;
                    if (k == 0) {
                        st = SM_ENDLEFT;
                    } else {
                        piv = make_uint4(s_lo[head][tid], s_hi[head][tid], s_rl[head][tid], 0u);
                        st = SM_LEFT;
                    }
                } else {
                    // no compatible occurrence: L(j) = j - (longest occurring
                    // suffix); below min_len no MEM starts before it, else the
                    // restart searches for its set
                    const uint32_t lam = piv.w;
                    if (lam < min_len) {
                        s0 = j - lam;
                        st = SM_CAND;
                    } else {
                        st = SM_RESTART;
                    }
                }
            } else { // SM_EMIT
                if (k >= ib) { // starts in a later segment of the read
                    status[w] = 0;
                    mem_cnt[w] = nrec;
                    st = SM_FETCH;
                    continue;
                }
                if (k < ia) { // starts in an earlier segment: not ours, scan on
                    s0 = ia;
                    E = e;
                    st = SM_CAND;
                    continue;
                }
                if (nrec + ns > slot_cap) {
                    status[w] = 1;
                    mem_cnt[w] = 0;
                    defer_list[atomicAdd(&ctr->n_retry, 1ull)] = w;
                    atomicAdd(&ctr->diag[2], 1ull);
                    st = SM_FETCH;
                    continue;
                }
                for (uint32_t u = 0; u < ns; u++) {
                    const uint32_t lo = s_lo[(head + u) & RM][tid];
                    uint32_t rank = 0;
                    for (uint32_t w = 0; w < ns; w++) rank += s_lo[(head + w) & RM][tid] < lo ? 1u : 0u;
                    slots[slot_base + nrec + rank] = make_uint4(k, e, lo, s_hi[(head + u) & RM][tid]);
                }
                nrec += ns;
                s0 = k + 1;
                E = e;
                st = SM_CAND;
            }
        }
        if (__ballot(st != SM_DONE) == 0) break;
        if (st == SM_DONE) continue;

        // ---- memory pass: a seed record, or one interval extended by one code.
        // The record load is issued first and consumed after the probes, so
        // its latency overlaps theirs instead of stalling the wave on its own.
        const bool srec = st == SM_SEEDR, stext = st == SM_TSA || st == SM_TEXT;
        const bool lsis = LSI && st >= SM_LSI_LEAF && st <= SM_LSI_PAY;
        uint4 R = make_uint4(0u, 0u, 0u, 0u);
        bool lhit = false;
        if (srec) {
            R = sv.rec[sidx];
        } else if (LSI && st == SM_LSI_LEAF) {
            R = *reinterpret_cast<const uint4 *>(lv.leaves + lsi_leaf(lv.md, lkey));
            probes++;
        } else if (LSI && st == SM_LSI_FENCE) {
            const unsigned long long *f = lv.fc.f[0];
#pragma unroll
            for (uint32_t v = 1; v < kLsiMaxFences; v++) f = llv == v ? lv.fc.f[v] : f;
            long long lo = llo, hi = (long long)lhi - 1;
            lsi_fence_step(f, lsi_fence_shift((int)llv), lo, hi, lkey, probes);
            llo = (uint32_t)lo;
            lhi = (uint32_t)(hi + 1);
        } else if (LSI && st == SM_LSI_KEYS) {
            // part llv of the last round: slots [llo - 1, lhi] (at most 35
            // keys), 18 from the even slot below the first per part
            const long long first = llo > 0 ? (long long)llo - 1 : 0;
            const long long last = (long long)lhi < (long long)lv.md.nkeys - 1 ? (long long)lhi
                                                                              : (long long)lv.md.nkeys - 1;
            bool hs = lsv != ~0ull;
            unsigned long long pv = lpv, sv = lsv;
            lsi_keys_part(lv.keys, first, last, (first & ~1ll) + 18 * (long long)llv, lkey, lbelow, pv, sv, hs,
                          probes);
            lpv = pv;
            lsv = sv;
            lhit = !hs && llv == 0 && (first & ~1ll) + 18 <= last; // read the second part
        } else if (LSI && st == SM_LSI_PAY) {
            R = lv.pay[llo + llv];
            probes++;
        } else if (st == SM_TSA) {
            R.x = fwd.sa_full[piv.x];
        } else if (st == SM_TEXT) {
            if (nibtext) { // the aligned 16-code groups holding text position sidx + e and query code e
                const uint2 tw = *reinterpret_cast<const uint2 *>(fwd.text4 + (((sidx + e) >> 3) & ~1u));
                const uint2 qw = *reinterpret_cast<const uint2 *>(qnib + ((((uint64_t)(p - q) + e) >> 3) & ~1ull));
                R = make_uint4(tw.x, tw.y, qw.x, qw.y);
            } else {
                const uint2 tw = *reinterpret_cast<const uint2 *>(fwd.text + ((sidx + e) & ~7u));
                const uint2 qw =
                    *reinterpret_cast<const uint2 *>(reinterpret_cast<uintptr_t>(p + e) & ~(uintptr_t)7);
                R = make_uint4(tw.x, tw.y, qw.x, qw.y);
            }
        }
        bool ovf = false;
        if (!srec && !stext && !lsis) {
            const bool left = st != SM_RIGHT;
            const uint32_t size = piv.y - piv.x;
            const uint32_t xlo = left ? piv.x : piv.z;
            const uint32_t xhi = left ? piv.y : piv.z + size;
            const uint32_t qpos = st == SM_SEEDE ? j - F - 1 : (left ? k - 1 : e);
            const uint32_t cm = s_compat[p[qpos]];
            const DevDir &D = left ? fwd : rev;
            probes += 1u + ((xlo >> kRecShift) != (xhi >> kRecShift) ? 1u : 0u);
            uint32_t cl[16], ch[16];
            pair_counts(D, xlo, xhi, cl, ch);
            const uint32_t cbase = head + ns; // children start behind the current set
            // slots not held by parents still to be processed (this one is in piv)
            const uint32_t room = st == SM_SEEDE ? PR : PR - (ns - t - 1);
            const bool dedup = st == SM_SEEDE;
            uint32_t acc = 0;
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const uint32_t d = ch[kk] - cl[kk];
                if (((cm >> kk) & 1u) && d != 0u) {
                    const uint32_t ca = s_clane[kk] + cl[kk];
                    const uint4 c = left ? make_uint4(ca, ca + d, piv.z + acc, 0u)
                                         : make_uint4(piv.x + acc, piv.x + acc + d, ca, 0u);
                    bool dup = false;
                    if (dedup)
                        for (uint32_t u = 0; u < nn; u++) dup |= s_lo[(cbase + u) & RM][tid] == c.x;
                    if (!dup) {
                        if (nn + 1 > room) {
                            ovf = true;
                        } else {
                            const uint32_t w = (cbase + nn) & RM;
                            s_lo[w][tid] = c.x;
                            s_hi[w][tid] = c.y;
                            s_rl[w][tid] = c.z;
                            nn++;
                        }
                    }
                }
                acc += d;
            }
        }

        // ---- advance ----
        if (srec) {
            // one packed record: lsl, interval count and the interval itself
            // when it is the only one
            probes++;
            const uint32_t l = seed_rec_lsl(R.w), cnt = seed_rec_cnt(R.w);
            lslmax = l > lslmax ? l : lslmax;
            // the query code left of the window: when its bases have no child
            // under the F-mer (extension mask), its first left extension is
            // empty and the expansion adds nothing to the set
            const uint32_t xe = j > F ? (uint32_t)s_expand[p[j - F - 1]] : 0u;
            const uint32_t xb = (xe & 0x10u) ? 1u << (xe & 3u) : (xe & 15u);
            if (l != F || cnt == 0 || (xb && (seed_rec_mask(R.w) & xb) == 0)) {
                st = SM_SEEDX;
            } else if (cnt == 1) {
                ient = 0;
                iend = 1;
                piv = make_uint4(R.x, R.y, R.z, 0u);
                st = SM_SEEDE;
            } else {
                ient = (uint64_t)R.x | ((uint64_t)R.y << 32);
                iend = ient + cnt;
                piv = sv.iv[ient];
                probes++;
                st = SM_SEEDE;
            }
        } else if (st == SM_TSA) {
            probes++;
            sidx = R.x - k; // text position of query code e is sidx + e
            st = SM_TEXT;
        } else if (st == SM_TEXT && nibtext) {
            probes++;
            // codes e, e+1, ... match while text and query nibbles share a base
            // (text nibbles are 0 at sentinels and past the end): the first
            // nibble of (t & q) that is zero ends the match
            const uint32_t to = (sidx + e) & 15u;
            const uint32_t qo = (uint32_t)(((uint64_t)(p - q) + e) & 15u);
            const uint64_t tw = ((uint64_t)R.x | ((uint64_t)R.y << 32)) >> (4 * to);
            const uint64_t qw = ((uint64_t)R.z | ((uint64_t)R.w << 32)) >> (4 * qo);
            const uint32_t lim = min(16u - max(to, qo), m - e); // codes both words hold (>= 1)
            const uint64_t x = tw & qw;
            const uint64_t ones = 0x1111111111111111ull;
            uint64_t miss = ~(x | (x >> 1) | (x >> 2) | (x >> 3)) & ones;
            if (lim < 16) miss |= 1ull << (4 * lim);
            const uint32_t adv = miss ? (uint32_t)__ffsll((unsigned long long)miss) >> 2 : 16u; // (bit + 1) / 4
            e += adv < lim ? adv : lim;
            if (adv < lim || e >= m) st = SM_EMIT;
        } else if (st == SM_TEXT) {
            probes++;
            const uint32_t to = (sidx + e) & 7u;
            const uint32_t qo = (uint32_t)(reinterpret_cast<uintptr_t>(p + e) & 7);
            const uint64_t tw = (uint64_t)R.x | ((uint64_t)R.y << 32), qw = (uint64_t)R.z | ((uint64_t)R.w << 32);
            const uint32_t lim = min(min(8u - to, 8u - qo), m - e);
            uint32_t adv = 0;
            bool go = true;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                if (i < lim && go) {
                    const uint32_t tl = (uint32_t)(tw >> (8 * (to + i))) & 0xffu;
                    const uint32_t qc = (uint32_t)(qw >> (8 * (qo + i))) & 0xffu;
                    if ((s_compat[qc] >> tl) & 1u) adv++;
                    else go = false;
                }
            }
            e += adv;
            if (!go || e >= m) st = SM_EMIT;
        } else if (LSI && st == SM_LSI_LEAF) {
            LsiLeaf lf;
            lf.s0 = R.x;
            lf.err = R.y;
            lf.a = __uint_as_float(R.z);
            lf.b = __uint_as_float(R.w);
            long long lo, hi;
            lsi_range(lv.md, lf, lsi_stage1(s_lp, lv.md.hidden, lsi_u(lv.md, lkey)), lo, hi);
            llo = (uint32_t)lo;
            lhi = (uint32_t)(hi + 1);
            lbelow = 0;
            lsv = ~0ull;
            if (lo <= hi && lv.md.nfence) {
                llv = lv.md.nfence - 1;
                st = SM_LSI_FENCE;
            } else {
                llv = 0;
                st = SM_LSI_KEYS;
            }
        } else if (LSI && st == SM_LSI_FENCE) {
            if (llv == 0) st = SM_LSI_KEYS;
            else llv--;
        } else if (LSI && st == SM_LSI_KEYS) {
            if (lhit) {
                llv = 1;
            } else {
                const uint32_t lb = (llo > 0 ? llo - 1 : 0u) + lbelow; // the first slot whose key is not below
                bool next = true;
                if (lsi_present(lv.md, lkey, lsv, lsv != ~0ull)) {
                    llo = lb;
                    llv = 0;
                    st = SM_LSI_PAY;
                    next = false;
                } else {
                    // no compatible occurrence of this base string: its longest
                    // occurring suffix
                    const uint32_t lam = lsi_suffix_len(lv.md, lkey, lpv, lbelow > 0, lsv, lsv != ~0ull);
                    piv.w = lam > piv.w ? lam : piv.w;
                }
                if (next) st = SM_LSI_NEXT;
            }
        } else if (LSI && st == SM_LSI_PAY) {
            // one interval of the set: the run of full entries with the base
            // string's codes (its length on the first entry); base strings of
            // one window can share text windows, whose intervals are kept once
            if (llv == 0) lhi = R.w >> 8;
            bool dup = false;
            if (piv.z >> 27)
                for (uint32_t u = 0; u < ns; u++) dup |= s_lo[(head + u) & RM][tid] == R.x;
            if (!dup && ns >= PR) {
                status[w] = 1;
                mem_cnt[w] = 0;
                defer_list[atomicAdd(&ctr->n_retry, 1ull)] = w;
                atomicAdd(&ctr->diag[1], 1ull);
                st = SM_FETCH;
            } else {
                if (!dup) {
                    const uint32_t wi = (head + ns) & RM;
                    s_lo[wi][tid] = R.x;
                    s_hi[wi][tid] = R.y;
                    s_rl[wi][tid] = R.z;
                    ns++;
                }
                if (++llv >= lhi) st = SM_LSI_NEXT;
            }
        } else if (ovf) {
            status[w] = 1;
            mem_cnt[w] = 0;
            defer_list[atomicAdd(&ctr->n_retry, 1ull)] = w;
            atomicAdd(&ctr->diag[st == SM_SEEDE ? 0 : 1], 1ull);
            st = SM_FETCH;
        } else if (st == SM_SEEDE) {
            ient++;
            if (ient < iend) {
                piv = sv.iv[ient];
                probes++;
            } else {
                st = SM_SEEDX;
            }
        } else if (++t < ns) {
            const uint32_t w = (head + t) & RM;
            piv = make_uint4(s_lo[w][tid], s_hi[w][tid], s_rl[w][tid], 0u);
        } else if (nn == 0) {
            st = st == SM_LEFT ? SM_ENDLEFT : SM_EMIT;
        } else {
            head = (head + ns) & RM; ns = nn; nn = 0; t = 0;
            bool more;
            if (st == SM_LEFT) { k--; more = k > 0; if (!more) st = SM_ENDLEFT; }
            else { e++; more = e < m; if (!more) st = SM_EMIT; }
            if (more) {
                piv = make_uint4(s_lo[head][tid], s_hi[head][tid], s_rl[head][tid], 0u);
                if (st == SM_RIGHT && textmode && ns == 1 && piv.y - piv.x == 1) st = SM_TSA;
            }
        }
    }
    if (restarts) atomicAdd(&ctr->diag[3], (unsigned long long)restarts);
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) probes += __shfl_xor(probes, dd, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&ctr->probes, (unsigned long long)probes);
}

// Copy the records of thread-path reads from their slots to the final order.
__global__ __launch_bounds__(256) void k_mem_gather(uint32_t nw, const uint64_t *__restrict__ qoff,
                                                    const uint4 *__restrict__ items,
                                                    const uint8_t *__restrict__ status, const uint32_t *__restrict__ mem_cnt,
                                                    const uint64_t *__restrict__ mem_off, const uint4 *__restrict__ slots,
                                                    uint4 *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw || status[w] != 0) return;
    const uint4 it = items ? items[w] : make_uint4(w, 0u, 0u, 0u);
    const uint64_t base = ((qoff[it.x] + it.y) >> 3) + 8ull * w, o = mem_off[w];
    for (uint32_t i = 0; i < mem_cnt[w]; i++) out[o + i] = slots[base + i];
}

// Work items: a read of m codes is split into ceil(m / seg) items, item s
// owning the MEMs that start in [s * seg, (s + 1) * seg); every read has at
// least one item.  Count pass, then fill at the scanned offsets.
__global__ void k_items_count(uint32_t nq, const uint64_t *__restrict__ qoff, uint32_t seg,
                              uint32_t *__restrict__ cnt) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nq) return;
    const uint64_t m = qoff[r + 1] - qoff[r];
    cnt[r] = m <= seg ? 1u : (uint32_t)((m + seg - 1) / seg);
}
// Work items of all reads together into total[0] (zeroed by the caller) and
// the flat length of the queries into total[1], for one read-back.
__global__ __launch_bounds__(256) void k_items_total(uint32_t nq, const uint64_t *__restrict__ qoff, uint32_t seg,
                                                     unsigned long long *__restrict__ total) {
    unsigned long long c = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nq; r += gridDim.x * blockDim.x) {
        const uint64_t m = qoff[r + 1] - qoff[r];
        c += m <= seg ? 1u : (m + seg - 1) / seg;
    }
    // one atomic per block (atomics on one address serialize in L2)
    __shared__ unsigned long long s_w[4];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long b = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (b) atomicAdd(total, b);
        if (blockIdx.x == 0) total[1] = qoff[nq];
    }
}
__global__ void k_items_fill(uint32_t nq, const uint64_t *__restrict__ qoff, uint32_t seg,
                             const uint64_t *__restrict__ first, uint4 *__restrict__ items) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nq) return;
    const uint32_t m = (uint32_t)(qoff[r + 1] - qoff[r]);
    const uint64_t f = first[r];
    const uint32_t ns = (uint32_t)(first[r + 1] - f);
    for (uint32_t s = 0; s < ns; s++) {
        const uint32_t a = s * seg, b = s + 1 == ns ? m : a + seg;
        items[f + s] = make_uint4(r, a, b, 0u);
    }
}
// hit offsets per read from the offsets per item (first[r] = first item of r)
__global__ void k_read_off(uint32_t nq, const uint64_t *__restrict__ first, const uint64_t *__restrict__ item_off,
                           uint64_t *__restrict__ read_off) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nq) return;
    read_off[r] = item_off[first[r]];
}

// Place the records of successful attempts at their final position.
// A reservation that did not fit the record buffer leaves dead slots (pass
// ~0, or never written): skip them before using their read id.
// Records of an item split into parts (pad = part << 24 | work index + 1) go
// behind the records of the item's earlier parts (part_cnt, by work index).
__global__ __launch_bounds__(256) void k_mem_scatter(const MemRec *__restrict__ recs, uint64_t nrec, uint32_t nq,
                                                     const uint32_t *__restrict__ pass_of,
                                                     const uint64_t *__restrict__ mem_off, uint4 *__restrict__ out,
                                                     const unsigned long long *__restrict__ d_nrec,
                                                     const uint32_t *__restrict__ part_cnt) {
    if (d_nrec) nrec = *d_nrec < nrec ? *d_nrec : nrec;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrec; t += (uint64_t)gridDim.x * blockDim.x) {
        const MemRec rc = recs[t];
        if (rc.pass == 0xffffffffu || rc.read >= nq || pass_of[rc.read] != rc.pass) continue;
        uint64_t base = 0;
        if (rc.pad && part_cnt) {
            const uint32_t wi = (rc.pad & 0xffffffu) - 1u, part = rc.pad >> 24;
            for (uint32_t i = 1; i <= part && i <= wi; i++) base += part_cnt[wi - i];
        }
        out[mem_off[rc.read] + base + rc.local] = make_uint4(rc.qs, rc.qe, rc.lo, rc.hi);
    }
}

// Per record: positions to resolve = min(size, max_hits) (bidir_index.rs:315-337).
__global__ __launch_bounds__(256) void k_mem_npos(const uint4 *__restrict__ recs, uint64_t nrec, uint32_t max_hits,
                                                  uint32_t resolve, uint32_t *__restrict__ npos,
                                                  const uint64_t *__restrict__ d_nrec) {
    if (d_nrec) nrec = *d_nrec < nrec ? *d_nrec : nrec; // capacity-sized buffers: the count is on the device
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrec; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 r = recs[t];
        const uint32_t size = r.w - r.z;
        npos[t] = resolve ? (size < max_hits ? size : max_hits) : 0u;
    }
}

__global__ __launch_bounds__(256) void k_mem_hits(const uint4 *__restrict__ recs, uint64_t nrec, uint32_t max_hits,
                                                  const uint64_t *__restrict__ pos_off, hsfm_mem_hit *__restrict__ hits,
                                                  uint32_t *__restrict__ rows, const uint64_t *__restrict__ d_nrec,
                                                  uint64_t pos_cap, unsigned long long *__restrict__ ovf) {
    if (d_nrec) nrec = *d_nrec < nrec ? *d_nrec : nrec;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrec; t += (uint64_t)gridDim.x * blockDim.x) {
        if (pos_off[t + 1] > pos_cap) { // positions would overrun a capacity-sized buffer
            atomicAdd(ovf, 1ull);
            continue;
        }
        const uint4 r = recs[t];
        const uint32_t size = r.w - r.z;
        const uint64_t po = pos_off[t], pl = pos_off[t + 1] - po;
        hsfm_mem_hit h;
        h.query_start = r.x;
        h.query_end = r.y;
        h.match_count = size;
        h.truncated = size > max_hits ? 1u : 0u;
        h.pos_offset = po;
        h.pos_len = (uint32_t)pl;
        h._reserved = 0;
        hits[t] = h;
        for (uint32_t i = 0; i < (uint32_t)pl; i++) rows[po + i] = r.z + i;
    }
}

} // namespace hsfm

using namespace hsfm;

namespace hsfm {
// Build the seed table of all ACGT F-mers (k_seed_level, level by level).
// One level of the seed table: the children of level l-1 (poff, piv, plsl)
// into (coff, civ, clsl); cnt receives the children's interval counts.
static int seed_level(hsfm_index *idx, uint32_t l, DevBuf &poff, DevBuf &piv, DevBuf &plsl, DevBuf &cnt,
                      DevBuf &coff, DevBuf &civ, DevBuf &clsl) {
    hipStream_t s = idx->dev->stream;
    const uint32_t nchild = 1u << (2 * l);
    int rc;
    if ((rc = cnt.ensure((size_t)nchild * 4)) || (rc = coff.ensure(((size_t)nchild + 1) * 8)) ||
        (rc = clsl.ensure(nchild)))
        return rc;
    const dim3 grid((nchild + 255) / 256);
    hipLaunchKernelGGL(k_seed_level<false>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                       poff.as<uint64_t>(), piv.as<uint4>(), nchild, cnt.as<uint32_t>(), nullptr, nullptr);
    HSFM_HIP(hipGetLastError());
    if ((rc = exclusive_scan_u32_u64(idx, cnt.as<uint32_t>(), coff.as<uint64_t>(), nchild))) return rc;
    uint64_t total = 0;
    if ((rc = readback_u64(idx, coff.as<uint64_t>() + nchild, &total, 1))) return rc;
    if ((rc = civ.ensure(total * 16 + 16))) return rc;
    hipLaunchKernelGGL(k_seed_level<true>, grid, dim3(256), 0, s, idx->fwd.view, idx->d_const.as<DevConst>(),
                       poff.as<uint64_t>(), piv.as<uint4>(), nchild, nullptr, coff.as<uint64_t>(), civ.as<uint4>());
    HSFM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_seed_lsl, grid, dim3(256), 0, s, nchild, l, cnt.as<uint32_t>(), plsl.as<uint8_t>(),
                       clsl.as<uint8_t>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipStreamSynchronize(s));
    return HSFM_OK;
}

int build_seed_table(hsfm_index *idx, uint32_t F) {
    if (idx->seed_F == F) return HSFM_OK;
    hipStream_t s = idx->dev->stream;
    DevBuf poff, piv, plsl;
    int rc;
    if ((rc = poff.ensure(16)) || (rc = piv.ensure(16)) || (rc = plsl.ensure(4))) return rc;
    hipLaunchKernelGGL(k_seed_root, dim3(1), dim3(1), 0, s, poff.as<uint64_t>(), piv.as<uint4>(), plsl.as<uint8_t>(),
                       idx->fwd.view.n);
    HSFM_HIP(hipGetLastError());
    DevBuf cnt;
    for (uint32_t l = 1; l <= F; l++) {
        DevBuf coff, civ, clsl;
        if ((rc = seed_level(idx, l, poff, piv, plsl, cnt, coff, civ, clsl))) {
            coff.release(); civ.release(); clsl.release();
            poff.release(); piv.release(); plsl.release(); cnt.release();
            return rc;
        }
        poff.release(); piv.release(); plsl.release();
        poff = coff; piv = civ; plsl = clsl;
        coff = DevBuf(); civ = DevBuf(); clsl = DevBuf();
    }
    cnt.release();
    idx->seed_off.release(); idx->seed_iv.release(); idx->seed_lsl.release(); idx->seed_rec12.release();
    idx->seed_off = poff; idx->seed_iv = piv; idx->seed_lsl = plsl;
    idx->seed_F = F;
    return HSFM_OK;
}

// Packed record of every F-mer c from its level arrays (count = off[c+1] -
// off[c]); a count beyond the meta field sets *bad.
__global__ __launch_bounds__(256) void k_seed_pack(DevDir fwd, const DevConst *__restrict__ cst, uint32_t nchild,
                                                   const uint64_t *__restrict__ off, const uint4 *__restrict__ iv,
                                                   const uint8_t *__restrict__ lsl, uint4 *__restrict__ rec,
                                                   uint32_t *__restrict__ bad) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchild) return;
    const uint64_t a = off[c], cnt = off[c + 1] - a;
    if (cnt >= (1ull << (32 - kSeedCntShift))) {
        *bad = 1u;
        rec[c] = make_uint4(0, 0, 0, 0);
        return;
    }
    uint32_t meta = (uint32_t)lsl[c] | ((uint32_t)cnt << kSeedCntShift);
    if (cnt == 1) {
        // which query bases extend the one interval left (the child rule of
        // k_seed_level / the MEM kernels: a compatible lane with rows)
        const uint4 v = iv[a];
        uint32_t cl[16], ch[16];
        interval_counts<true>(fwd, cst->num_lanes, v, cl, ch);
        uint32_t mask = 0;
        for (uint32_t b = 0; b < 4; b++) {
            const uint32_t cm = cst->compat_lanes[1 + b];
            for (int k = 0; k < 16; k++)
                if (((cm >> k) & 1u) && ch[k] > cl[k]) mask |= 1u << b;
        }
        rec[c] = make_uint4(v.x, v.y, v.z, meta | (mask << kSeedLslBits));
    } else {
        meta |= 15u << kSeedLslBits;
        rec[c] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), 0u, meta);
    }
}

static int seed_pack(hsfm_index *idx, uint32_t F, const DevBuf &off, const DevBuf &iv, const DevBuf &lsl,
                     DevBuf &rec) {
    hipStream_t s = idx->dev->stream;
    const uint32_t nchild = 1u << (2 * F);
    DevBuf flag;
    int rc;
    if ((rc = flag.ensure(4))) return rc;
    if ((rc = rec.ensure((size_t)nchild * 16))) {
        flag.release();
        return rc;
    }
    // on any failure the records are released, so that a later call builds them again
    uint32_t bad = 0;
    rc = fill_u32(idx, flag.as<uint32_t>(), 1, 0);
    if (!rc) {
        hipLaunchKernelGGL(k_seed_pack, dim3((nchild + 255) / 256), dim3(256), 0, s, idx->fwd.view,
                           idx->d_const.as<DevConst>(), nchild, off.as<uint64_t>(), iv.as<uint4>(),
                           lsl.as<uint8_t>(), rec.as<uint4>(), flag.as<uint32_t>());
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = set_error(HSFM_EDEVICE, "seed records of %u-mers: %s", F, hipGetErrorString(e));
    }
    flag.release();
    if (!rc && bad) rc = set_error(HSFM_EINVAL, "a %u-mer has more compatible intervals than the seed record holds", F);
    if (rc) rec.release();
    return rc;
}

int build_seed_packed(hsfm_index *idx, uint32_t F) {
    int rc;
    if ((rc = build_seed_table(idx, 12))) return rc;
    if (F == 12) {
        if (idx->seed_rec12.p) return HSFM_OK;
        return seed_pack(idx, 12, idx->seed_off, idx->seed_iv, idx->seed_lsl, idx->seed_rec12);
    }
    if (F != 14 && F != 15) return set_error(HSFM_EINVAL, "seed length %u not supported", F);
    DevBuf &rec = F == 14 ? idx->seed_rec14 : idx->seed_rec15;
    DevBuf &ivs = F == 14 ? idx->seed_iv14 : idx->seed_iv15;
    if (rec.p) return HSFM_OK;
    // levels 13..F from the 12-mer table (which stays as it is)
    DevBuf cnt, poff, piv, plsl, coff, civ, clsl;
    for (uint32_t l = 13; l <= F && !rc; l++) {
        rc = seed_level(idx, l, l == 13 ? idx->seed_off : poff, l == 13 ? idx->seed_iv : piv,
                        l == 13 ? idx->seed_lsl : plsl, cnt, coff, civ, clsl);
        if (l > 13) {
            poff.release(); piv.release(); plsl.release();
        }
        poff = coff; piv = civ; plsl = clsl;
        coff = DevBuf(); civ = DevBuf(); clsl = DevBuf();
    }
    cnt.release();
    if (!rc) rc = seed_pack(idx, F, poff, piv, plsl, rec);
    poff.release(); plsl.release();
    if (rc) {
        piv.release();
        return rc;
    }
    ivs.release();
    ivs = piv;
    return HSFM_OK;
}

// Packed-record view of the FP-mer table (built by build_seed_packed); sv is
// the 12-mer table's view.  FP = 0: no seeding.
static SeedView packed_view(hsfm_index *idx, uint32_t FP, const SeedView &sv) {
    if (FP == 15) return SeedView{nullptr, idx->seed_iv15.as<uint4>(), nullptr, 15u, idx->seed_rec15.as<uint4>()};
    if (FP == 14) return SeedView{nullptr, idx->seed_iv14.as<uint4>(), nullptr, 14u, idx->seed_rec14.as<uint4>()};
    SeedView v = sv;
    if (FP == 12) v.rec = idx->seed_rec12.as<uint4>();
    else v.F = 0;
    return v;
}

// The packed table of FP-mers, or of 14-mers when the 15-mer table (17 GB)
// does not fit in device memory: the seed length used (0: none).
static int packed_seeds(hsfm_index *idx, uint32_t &FP) {
    if (!FP) return HSFM_OK;
    int rc = build_seed_packed(idx, FP);
    if (rc == HSFM_ENOMEM && FP == 15) {
        FP = 14;
        rc = build_seed_packed(idx, FP);
    }
    return rc;
}

// 14-mer seeds pay off (fewer, smaller restart sets) once 12-mers repeat by
// chance: texts of 2^22 positions and more.  From 2^24 positions the 15-mer
// table (2^30 records, 17 GB) also takes the first left extension of every
// restart whose 14-mer occurs.  HSFM_SEED_F=12 / 14 / 15 overrides.  Seeds
// are shorter than min_len.
uint32_t mem_seed_length(hsfm_index *idx, uint32_t min_len) {
    const uint32_t n = idx->fwd.view.n;
    uint32_t F = n >= (1u << 24) ? 15u : n >= (1u << 22) ? 14u : 12u;
    const char *e = getenv("HSFM_SEED_F");
    if (e && *e) {
        const long v = strtol(e, nullptr, 10);
        F = v == 15 ? 15u : v == 14 ? 14u : 12u;
    }
    if (F == 15 && min_len <= 15) F = 14;
    if (F == 14 && min_len <= 14) F = 12;
    return min_len > 12 ? F : 0u;
}
} // namespace hsfm

// Codes per MEM work item (HSFM_SEGMENT overrides; 0 keeps whole reads).
static uint32_t mem_segment_length() {
    const char *e = getenv("HSFM_SEGMENT");
    if (!e || !*e) return 256;
    const long v = strtol(e, nullptr, 10);
    return v <= 0 ? 0u : (v < 32 ? 32u : (uint32_t)v);
}

static int mem_device_impl(hsfm_index *idx, const uint8_t *d_q, const uint64_t *d_qoff, uint32_t nq, uint32_t min_len,
                           const uint32_t *d_refb, uint32_t n_refs, uint32_t max_hits, hsfm_mem_device_result *out) {
    Workspace &ws = idx->ws;
    hipStream_t s = idx->dev->stream;
    idx->stats = hsfm_stats{};
    idx->defer_valid = false;
    int rc;
    uint64_t h[16];
    if ((rc = ensure_sa_full(idx))) return rc;

    // path choice: the persistent kernel (one lane per work item) for min_len
    // >= 12, with its deferred items on the workgroup kernel; shorter min_len
    // (and HSFM_MEM_PATH=wave, which tests hold equal) run every read on the
    // workgroup kernel.  (A thread-per-read kernel with 8-interval sets and a
    // fully pipelined variant of the persistent kernel measured slower and are
    // not kept.)
    const uint32_t F = 12;
    const char *path = getenv("HSFM_MEM_PATH");
    const bool use_thread = nq > 0 && min_len >= F && !(path && strcmp(path, "wave") == 0);
    uint32_t FP = mem_seed_length(idx, min_len); // seed length of the persistent kernel

    // work items: reads longer than seg codes are split so that one long read
    // does not leave a single lane working after all others are done
    const uint32_t seg = mem_segment_length();
    uint64_t W = nq;
    const uint4 *items = nullptr;
    uint64_t total_len = 0; // flat length of the queries (read back with the item count)
    if (nq && !(seg && nq) && (rc = readback_u64(idx, d_qoff + nq, &total_len, 1))) return rc;
    if (seg && nq) {
        // item count and flat length in one round trip; the per-read counts,
        // their scan and the item table only when some read is split
        if ((rc = ws.item_tot.ensure(16)) || (rc = fill_u32(idx, ws.item_tot.as<uint32_t>(), 4, 0))) return rc;
        const uint32_t tg = std::min<uint32_t>((nq + 255) / 256, idx->dev->num_cus);
        hipLaunchKernelGGL(k_items_total, dim3(tg), dim3(256), 0, s, nq, d_qoff, seg,
                           ws.item_tot.as<unsigned long long>());
        HSFM_HIP(hipGetLastError());
        if ((rc = readback_u64(idx, ws.item_tot.as<uint64_t>(), h, 2))) return rc;
        W = h[0];
        total_len = h[1];
        if (W > 0xffffffffull - 16) return set_error(HSFM_EINVAL, "too many read segments (%llu)", (unsigned long long)W);
        if (W > nq) {
            if ((rc = ws.item_cnt.ensure((size_t)nq * 4 + 4)) || (rc = ws.item_first.ensure(((size_t)nq + 1) * 8)))
                return rc;
            hipLaunchKernelGGL(k_items_count, dim3((nq + 255) / 256), dim3(256), 0, s, nq, d_qoff, seg,
                               ws.item_cnt.as<uint32_t>());
            HSFM_HIP(hipGetLastError());
            if ((rc = exclusive_scan_u32_u64(idx, ws.item_cnt.as<uint32_t>(), ws.item_first.as<uint64_t>(), nq)))
                return rc;
            if ((rc = ws.items.ensure(W * 16))) return rc;
            hipLaunchKernelGGL(k_items_fill, dim3((nq + 255) / 256), dim3(256), 0, s, nq, d_qoff, seg,
                               ws.item_first.as<uint64_t>(), ws.items.as<uint4>());
            HSFM_HIP(hipGetLastError());
            items = ws.items.as<uint4>();
        }
    }
    if ((rc = ws.mem_cnt.ensure(W * 4 + 4))) return rc;
    if ((rc = ws.mem_flags.ensure(W * 4 + 4))) return rc;
    if ((rc = ws.ovf_list.ensure(W * 4 + 4))) return rc;
    if ((rc = ws.read_list.ensure(W * 4 + 4))) return rc;
    // [0]: the persistent kernel (and the resolve guard), [1]: the workgroup passes
    if ((rc = ws.counters.ensure(2 * sizeof(MemCounters)))) return rc;
    if ((rc = ws.mem_off.ensure((W + 1) * 8))) return rc;
    MemCounters *ctr = ws.counters.as<MemCounters>();
    if ((rc = fill_u32(idx, ws.counters.as<uint32_t>(), 2 * sizeof(MemCounters) / 4, 0))) return rc;
    MemCounters *wc = ctr + 1;
    // the persistent kernel writes every item's count (zero when it defers
    // the item), so the counts start at zero only for the workgroup path
    if (!use_thread && (rc = fill_u32(idx, ws.mem_cnt.as<uint32_t>(), W, 0))) return rc;
    if ((rc = fill_u32(idx, ws.mem_flags.as<uint32_t>(), W, 0xffffffffu))) return rc;
    const uint32_t *list = nullptr; // identity list
    uint64_t nlist = W;
    const unsigned long long *first_count = nullptr; // device count of the first workgroup pass's list
    SeedView sv{nullptr, nullptr, nullptr, 0};
    SeedView svw = sv; // seeds of the workgroup passes
    uint32_t *retry = ws.ovf_list.as<uint32_t>();
    if (use_thread) {
        if ((rc = build_seed_table(idx, F))) return rc;
        const uint64_t nslots = (total_len >> 3) + 8ull * W + 8;
        if ((rc = ws.mem_slots.ensure(nslots * 16 + 16))) return rc;
        if ((rc = pack_bases(idx, d_q, total_len, ws.qnib))) return rc;
        if ((rc = ws.mem_status.ensure(W + 16))) return rc;
        sv = SeedView{idx->seed_off.as<uint64_t>(), idx->seed_iv.as<uint4>(), idx->seed_lsl.as<uint8_t>(), F};
        if ((rc = packed_seeds(idx, FP))) return rc;
        SeedView svp = packed_view(idx, FP, sv); // FP = 0 (min_len 12): no seeding in the persistent kernel
        svw = svp;
        // restart seeding chosen by hsfm_mem_seed_mode
        LsiView lv{};
        uint32_t lsik = 0;
        const int smode = idx->mem_seed_mode;
        if (smode == HSFM_SEED_LEARNED || smode == HSFM_SEED_LEARNED_TABLES) {
            if ((rc = lsi_mem_view(idx->mem_lsi, idx, &lv))) return rc;
            lsik = lsi_k(idx->mem_lsi);
        }
        if (smode == HSFM_SEED_PLAIN || smode == HSFM_SEED_LEARNED) {
            svp.F = 0; // restarts not seeded by the tables
            svw = svp;
        }
        if ((rc = event_record(idx, 0, false))) return rc;
        {
            // the learned-seed states are compiled into their own instance, so
            // the other modes run the kernel without them
            auto kern = lsik ? k_mem_persist<16, 256, true> : k_mem_persist<16, 256, false>;
            const uint32_t g1 = persist_grid(idx, kern, 256);
            hipLaunchKernelGGL(kern, dim3(g1), dim3(256), 0, s, idx->fwd.view, idx->rev.view,
                               idx->d_const.as<DevConst>(), svp, d_q, ws.qnib.as<uint32_t>(), d_qoff,
                               (const uint32_t *)nullptr, (uint32_t)W,
                               items, min_len, ws.mem_slots.as<uint4>(), ws.mem_cnt.as<uint32_t>(),
                               ws.mem_status.as<uint8_t>(), ws.ovf_list.as<uint32_t>(), ctr, lv, lsik);
        }
        HSFM_HIP(hipGetLastError());
        if ((rc = event_record(idx, 0, true))) return rc;
        if ((rc = event_record(idx, 2, false))) return rc;
        // the deferred items go to the workgroup pass with no host round trip:
        // it reads their number from the persistent kernel's counter, and the
        // counters of both come back together after it
        first_count = &ctr->n_retry;
        nlist = W; // a bound; the count is on the device
        list = ws.ovf_list.as<uint32_t>();
        retry = ws.read_list.as<uint32_t>();
    }

    // records of the workgroup passes: eight per listed item (the count of a
    // device-counted list is not known here, so its bound, every item)
    uint64_t rec_cap = nlist * 8 + 4096;
    if (ws.mem_recs.cap / sizeof(MemRec) > rec_cap) rec_cap = ws.mem_recs.cap / sizeof(MemRec);
    // First pass: sets in LDS (kWaveLdsCap intervals); reads that outgrow them
    // re-run with global-memory sets of 4x the capacity per pass.
    uint32_t cap = kWaveLdsCap;
    uint32_t pass = 0;
    uint64_t rec_used = 0;
    uint64_t nhits = 0;
    bool scanned = false; // mem_off holds the scan of the final counts
    // reads deferred by the persistent kernel: kDeferTB threads each; every
    // read on the workgroup path (no persistent kernel): one wave each
    const uint32_t lds_waves = use_thread ? persist_grid(idx, k_mem_wave<true, kDeferTB>, kDeferTB)
                                          : persist_grid(idx, k_mem_wave<true, 64>, 64);
    // part counts of the deferred items split by the first workgroup pass
    // (one per workgroup of its grid)
    if ((rc = ws.part_cnt.ensure((size_t)lds_waves * 4 + 4))) return rc;
    while (nlist) {
        const bool lds = cap == kWaveLdsCap;
        uint64_t waves = lds ? lds_waves : (uint64_t)idx->dev->num_cus * 2;
        if (waves > nlist) waves = nlist;
        const uint64_t wave_limit = (4ull << 30) / (2ull * cap * 16);
        if (waves > wave_limit) waves = wave_limit;
        if (waves == 0) waves = 1;
        if (!lds && (rc = ws.wave_sets.ensure((size_t)waves * 2 * cap * 16))) return rc;
        if (ws.mem_recs.cap < rec_cap * sizeof(MemRec)) {
            DevBuf grown;
            if ((rc = grown.ensure(rec_cap * sizeof(MemRec)))) return rc;
            const size_t kept = ws.mem_recs.p ? ws.mem_recs.cap : 0;
            if (kept) HSFM_HIP(hipMemcpyAsync(grown.p, ws.mem_recs.p, kept, hipMemcpyDeviceToDevice, s));
            // slots reserved past the old end were never written: the new part
            // starts as dead records (read = pass = ~0), which the scatter skips
            if ((rc = fill_u32(idx, reinterpret_cast<uint32_t *>(static_cast<char *>(grown.p) + kept),
                               (grown.cap - kept) / 4, 0xffffffffu)))
                return rc;
            HSFM_HIP(hipStreamSynchronize(s));
            ws.mem_recs.release();
            ws.mem_recs = grown;
        }
        rec_cap = ws.mem_recs.cap / sizeof(MemRec);
        if ((rc = fill_u32(idx, reinterpret_cast<uint32_t *>(&wc->work), 2, 0))) return rc;
        if ((rc = fill_u32(idx, reinterpret_cast<uint32_t *>(&wc->n_retry), 2, 0))) return rc;
        const unsigned long long *dcount = pass == 0 ? first_count : nullptr;
        const bool timed = !use_thread && pass == 0;
        if (timed && (rc = event_record(idx, 0, false))) return rc;
        if (lds && use_thread)
            hipLaunchKernelGGL((k_mem_wave<true, kDeferTB>), dim3((uint32_t)waves), dim3(kDeferTB), 0, s,
                               idx->fwd.view, idx->rev.view, idx->d_const.as<DevConst>(), svw, d_q, d_qoff, list,
                               (uint32_t)nlist, items, min_len, cap, (uint4 *)nullptr, ws.mem_recs.as<MemRec>(),
                               rec_cap, ws.mem_cnt.as<uint32_t>(), ws.mem_flags.as<uint32_t>(), pass, retry, wc,
                               dcount, dcount ? ws.part_cnt.as<uint32_t>() : nullptr, kDeferParts);
        else if (lds)
            hipLaunchKernelGGL((k_mem_wave<true, 64>), dim3((uint32_t)waves), dim3(64), 0, s, idx->fwd.view,
                               idx->rev.view,
                               idx->d_const.as<DevConst>(), svw, d_q, d_qoff, list, (uint32_t)nlist, items, min_len, cap,
                               (uint4 *)nullptr, ws.mem_recs.as<MemRec>(), rec_cap, ws.mem_cnt.as<uint32_t>(),
                               ws.mem_flags.as<uint32_t>(), pass, retry, wc, dcount, (uint32_t *)nullptr, 1u);
        else
            hipLaunchKernelGGL((k_mem_wave<false, 1024>), dim3((uint32_t)waves), dim3(1024), 0, s, idx->fwd.view,
                               idx->rev.view, idx->d_const.as<DevConst>(), svw, d_q, d_qoff, list, (uint32_t)nlist,
                               items, min_len, cap, ws.wave_sets.as<uint4>(), ws.mem_recs.as<MemRec>(), rec_cap,
                               ws.mem_cnt.as<uint32_t>(), ws.mem_flags.as<uint32_t>(), pass, retry, wc, dcount,
                               (uint32_t *)nullptr, 1u);
        HSFM_HIP(hipGetLastError());
        if (timed && (rc = event_record(idx, 0, true))) return rc;
        if (use_thread && (rc = event_record(idx, 2, true))) return rc;
        // the scan of the hit counts runs behind the pass, so its total comes
        // back with the pass counters in one round trip (redone after retries)
        if ((rc = exclusive_scan_u32_u64(idx, ws.mem_cnt.as<uint32_t>(), ws.mem_off.as<uint64_t>(), (uint32_t)W)))
            return rc;
        if (!ws.h_pinned) HSFM_HIP(hipHostMalloc((void **)&ws.h_pinned, kPinnedBytes, hipHostMallocDefault));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned, wc, 64, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned + 8, ws.mem_off.as<uint64_t>() + W, 8, hipMemcpyDeviceToHost, s));
        if (dcount) HSFM_HIP(hipMemcpyAsync(ws.h_pinned + 16, ctr, 128, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
        memcpy(h, ws.h_pinned, 64);
        const uint64_t nretry = h[2];
        rec_used = h[1];
        if (dcount) { // the persistent kernel's counters: deferred items, accesses, deferral causes
            const uint64_t *hd = ws.h_pinned + 16;
            const uint64_t ndef = hd[2];
            idx->stats.occ_probes = hd[4];
            for (int i = 0; i < 4; i++) idx->stats.diag[i] = hd[5 + i];
            idx->stats.fallback_reads = ndef;
            // keep the deferred items for hsfm_last_deferred (a retry pass rewrites ovf_list)
            if ((rc = ws.defer_ids.ensure(ndef * 4 + 4))) return rc;
            if (ndef)
                HSFM_HIP(hipMemcpyAsync(ws.defer_ids.p, ws.ovf_list.p, ndef * 4, hipMemcpyDeviceToDevice, s));
            idx->defer_n = ndef;
            idx->defer_items = items != nullptr;
            idx->defer_valid = true;
        }
        if (nretry == 0) {
            nhits = ws.h_pinned[8];
            scanned = true;
            break;
        }
        idx->stats.capacity_retries++;
        if (cap >= (1u << 26)) return set_error(HSFM_ENOMEM, "interval set of a single read exceeds 2^26 intervals");
        // sets that outgrow the LDS are usually far larger: go straight to
        // 65536 intervals, then 4x per further pass
        cap = lds ? 65536 : cap * 4;
        // grow the record buffer only when the retried reads may not fit it
        // (a retry for set capacity alone keeps it)
        const uint64_t want = rec_used + nretry * 4096;
        if (want > rec_cap) rec_cap = rec_cap * 2 > want ? rec_cap * 2 : want;
        // next pass runs only the retried reads
        uint32_t *next_list = retry;
        retry = (next_list == ws.ovf_list.as<uint32_t>()) ? ws.read_list.as<uint32_t>() : ws.ovf_list.as<uint32_t>();
        list = next_list;
        nlist = nretry;
        pass++;
    }
    if (!use_thread) idx->stats.fallback_reads = 0;
    if (!scanned) {
        if ((rc = exclusive_scan_u32_u64(idx, ws.mem_cnt.as<uint32_t>(), ws.mem_off.as<uint64_t>(), (uint32_t)W)))
            return rc;
        if ((rc = readback_u64(idx, ws.mem_off.as<uint64_t>() + W, &nhits, 1))) return rc;
    }
    idx->stats.records = nhits;
    if ((rc = ws.mem_sorted.ensure(nhits * 16 + 16))) return rc;
    if (use_thread) {
        hipLaunchKernelGGL(k_mem_gather, dim3((uint32_t)((W + 255) / 256)), dim3(256), 0, s, (uint32_t)W, d_qoff, items,
                           ws.mem_status.as<uint8_t>(), ws.mem_cnt.as<uint32_t>(), ws.mem_off.as<uint64_t>(),
                           ws.mem_slots.as<uint4>(), ws.mem_sorted.as<uint4>());
        HSFM_HIP(hipGetLastError());
    }
    if (rec_used > rec_cap) rec_used = rec_cap; // reservations past the buffer were never written
    if (rec_used) {
        hipLaunchKernelGGL(k_mem_scatter, dim3((uint32_t)((rec_used + 255) / 256)), dim3(256), 0, s,
                           ws.mem_recs.as<MemRec>(), rec_used, (uint32_t)W, ws.mem_flags.as<uint32_t>(),
                           ws.mem_off.as<uint64_t>(), ws.mem_sorted.as<uint4>(), nullptr,
                           ws.part_cnt.as<uint32_t>());
        HSFM_HIP(hipGetLastError());
    }
    if ((rc = ws.cnt.ensure(nhits * 4 + 4))) return rc;
    if ((rc = ws.pos_off.ensure((nhits + 1) * 8))) return rc;
    if ((rc = ws.mem_hits.ensure(nhits * sizeof(hsfm_mem_hit) + 64))) return rc;
    if (nhits) {
        hipLaunchKernelGGL(k_mem_npos, dim3((uint32_t)((nhits + 255) / 256)), dim3(256), 0, s,
                           ws.mem_sorted.as<uint4>(), nhits, max_hits, n_refs ? 1u : 0u, ws.cnt.as<uint32_t>(),
                           nullptr);
        HSFM_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan_u32_u64(idx, ws.cnt.as<uint32_t>(), ws.pos_off.as<uint64_t>(), (uint32_t)nhits))) return rc;
    // The position count stays on the device: the row and position buffers
    // keep the capacity of earlier calls (at least two rows per record), the
    // hits kernel counts the records whose rows would not fit, resolve runs
    // over the device count, and one read-back after it gives the count, the
    // overflow and the walk guard.  An overflow grows the buffers to the count
    // and runs the two kernels again.
    uint64_t npos = 0;
    uint64_t pos_cap = std::min<uint64_t>(ws.rows.cap / 4, ws.positions.cap / 8);
    if (pos_cap < 2 * nhits + 64) {
        pos_cap = 2 * nhits + 64;
        if ((rc = ws.rows.ensure(pos_cap * 4)) || (rc = ws.positions.ensure(pos_cap * 8))) return rc;
    }
    for (;;) {
        if (nhits) {
            hipLaunchKernelGGL(k_mem_hits, dim3((uint32_t)((nhits + 255) / 256)), dim3(256), 0, s,
                               ws.mem_sorted.as<uint4>(), nhits, max_hits, ws.pos_off.as<uint64_t>(),
                               ws.mem_hits.as<hsfm_mem_hit>(), ws.rows.as<uint32_t>(), nullptr, pos_cap,
                               &ctr->pos_ovf);
            HSFM_HIP(hipGetLastError());
        }
        idx->stats.sa_full_rows = idx->fwd.view.sa_full ? idx->fwd.view.n : 0;
        if (nhits && n_refs) {
            if ((rc = event_record(idx, 1, false))) return rc;
            if ((rc = launch_resolve_fan(idx, idx->fwd.view, ws.rows.as<uint32_t>(), pos_cap,
                                         ws.pos_off.as<uint64_t>() + nhits, d_refb, n_refs,
                                         ws.positions.as<uint32_t>(), &ctr->guard, 4, idx->dev->num_cus * 8)))
                return rc;
            if ((rc = event_record(idx, 1, true))) return rc;
        }
        if (!ws.h_pinned) HSFM_HIP(hipHostMalloc((void **)&ws.h_pinned, kPinnedBytes, hipHostMallocDefault));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned, ws.pos_off.as<uint64_t>() + nhits, 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned + 1, &ctr->guard, 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(ws.h_pinned + 2, &ctr->pos_ovf, 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
        npos = ws.h_pinned[0];
        const uint64_t guard = ws.h_pinned[1], ovf = ws.h_pinned[2];
        idx->stats.guard_hits = guard;
        if (guard) return set_error(HSFM_EINDEX, "%llu LF walks exceeded the sample rate: index is inconsistent", (unsigned long long)guard);
        if (ovf == 0) break;
        pos_cap = npos;
        if ((rc = ws.rows.ensure(pos_cap * 4 + 4)) || (rc = ws.positions.ensure(pos_cap * 8 + 8)) ||
            (rc = fill_u32(idx, reinterpret_cast<uint32_t *>(&ctr->pos_ovf), 2, 0)))
            return rc;
    }
    idx->stats.main_kernel_ms = event_ms(idx, 0);
    idx->stats.resolve_kernel_ms = npos ? event_ms(idx, 1) : 0.0;
    idx->stats.fallback_kernel_ms = idx->stats.fallback_reads ? event_ms(idx, 2) : 0.0;
    out->num_queries = nq;
    out->num_hits = nhits;
    out->num_positions = npos;
    out->d_hit_offsets = ws.mem_off.as<uint64_t>();
    if (items) { // offsets per read from the offsets per item
        if ((rc = ws.read_off.ensure(((size_t)nq + 1) * 8))) return rc;
        hipLaunchKernelGGL(k_read_off, dim3((nq + 256) / 256), dim3(256), 0, s, nq, ws.item_first.as<uint64_t>(),
                           ws.mem_off.as<uint64_t>(), ws.read_off.as<uint64_t>());
        HSFM_HIP(hipGetLastError());
        out->d_hit_offsets = ws.read_off.as<uint64_t>();
    }
    out->d_hits = ws.mem_hits.as<hsfm_mem_hit>();
    out->d_positions = ws.positions.as<uint32_t>();
    return HSFM_OK;
}

extern "C" int hsfm_find_mems_device(hsfm_index *idx, const uint8_t *d_qflat, const uint64_t *d_qoff, uint32_t nq,
                                     uint32_t min_len, int mode, const uint32_t *d_ref_boundaries, uint32_t n_refs,
                                     uint32_t max_hits, hsfm_mem_device_result *out) {
    if (!idx || !out || (nq && (!d_qflat || !d_qoff)) || (n_refs && !d_ref_boundaries))
        return set_error(HSFM_EINVAL, "hsfm_find_mems_device: null argument");
    if (!idx->has_rev) return set_error(HSFM_EINVAL, "MEM search needs an index uploaded with its reverse index");
    if (mode != HSFM_MODE_SMEM && mode != HSFM_MODE_MEM) return set_error(HSFM_EINVAL, "unknown mode %d", mode);
    HSFM_LOCK_INDEX(idx, "hsfm_find_mems_device");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    return mem_device_impl(idx, d_qflat, d_qoff, nq, min_len, d_ref_boundaries, n_refs, max_hits, out);
}

extern "C" int hsfm_find_mems_batch(hsfm_index *idx, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                    uint32_t min_len, int mode, const uint32_t *ref_boundaries, uint32_t n_refs,
                                    uint32_t max_hits, hsfm_mem_result **out) {
    if (!idx || !out || !qoff || (nq && !qflat) || (n_refs && !ref_boundaries))
        return set_error(HSFM_EINVAL, "hsfm_find_mems_batch: null argument");
    if (!idx->has_rev) return set_error(HSFM_EINVAL, "MEM search needs an index uploaded with its reverse index");
    if (mode != HSFM_MODE_SMEM && mode != HSFM_MODE_MEM) return set_error(HSFM_EINVAL, "unknown mode %d", mode);
    HSFM_LOCK_INDEX(idx, "hsfm_find_mems_batch");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    Workspace &ws = idx->ws;
    hipStream_t s = idx->dev->stream;
    for (uint32_t i = 0; i < nq; i++)
        if (qoff[i + 1] < qoff[i]) return set_error(HSFM_EINVAL, "query offsets not monotone at %u", i);
    const uint64_t tot = qoff[nq];
    int rc;
    if ((rc = ws.qcopy.ensure(tot + 16))) return rc;
    if ((rc = ws.qoffcopy.ensure(((size_t)nq + 1) * 8))) return rc;
    if ((rc = ws.refb.ensure((size_t)n_refs * 4 + 4))) return rc;
    if (tot) HSFM_HIP(hipMemcpyAsync(ws.qcopy.p, qflat, tot, hipMemcpyHostToDevice, s));
    HSFM_HIP(hipMemcpyAsync(ws.qoffcopy.p, qoff, ((size_t)nq + 1) * 8, hipMemcpyHostToDevice, s));
    if (n_refs) HSFM_HIP(hipMemcpyAsync(ws.refb.p, ref_boundaries, (size_t)n_refs * 4, hipMemcpyHostToDevice, s));
    hsfm_mem_device_result dr;
    if ((rc = mem_device_impl(idx, ws.qcopy.as<uint8_t>(), ws.qoffcopy.as<uint64_t>(), nq, min_len,
                              ws.refb.as<uint32_t>(), n_refs, max_hits, &dr)))
        return rc;
    const size_t hdr = (sizeof(hsfm_mem_result) + 15) & ~(size_t)15;
    const size_t b_off = ((size_t)nq + 1) * 8;
    const size_t b_hits = dr.num_hits * sizeof(hsfm_mem_hit);
    const size_t bytes = hdr + b_off + b_hits + dr.num_positions * 8;
    char *blk = (char *)malloc(bytes);
    if (!blk) return set_error(HSFM_ENOMEM, "host allocation of %zu bytes failed", bytes);
    hsfm_mem_result *res = (hsfm_mem_result *)blk;
    res->num_queries = nq;
    res->hit_offsets = (uint64_t *)(blk + hdr);
    res->hits = (hsfm_mem_hit *)(blk + hdr + b_off);
    res->positions = (uint32_t *)(blk + hdr + b_off + b_hits);
    res->num_hits = dr.num_hits;
    res->num_positions = dr.num_positions;
    hipError_t e = hipMemcpyAsync(res->hit_offsets, dr.d_hit_offsets, b_off, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && b_hits) e = hipMemcpyAsync(res->hits, dr.d_hits, b_hits, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && dr.num_positions)
        e = hipMemcpyAsync(res->positions, dr.d_positions, dr.num_positions * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free(blk);
        return set_error(HSFM_EDEVICE, "copy of MEM results failed: %s", hipGetErrorString(e));
    }
    *out = res;
    return HSFM_OK;
}

// ===========================================================================
// Batched MEM pipeline for uneven reads: fixed-size batches of B reads, each
// run by one captured HIP graph (find -> second-level passes -> gather/scatter
// -> positions -> resolve -> reference map) whose kernels take every size they
// need from device counters, so a batch costs one graph launch and one host
// wait (to hand its results to the caller).  The graph reads the batch's
// offsets from a fixed staging buffer that a copy outside the graph refills.
// Reads that outgrow the in-graph capacities send their batch through the
// host-driven path instead.
// ===========================================================================
namespace hsfm {

constexpr uint32_t kBatchGlobalCap = 65536; // intervals per set in the in-graph global pass
constexpr uint32_t kBatchGlobalWaves = 64;  // workgroups of that pass
enum : int { BM_PERSIST8 = 0, BM_PERSIST16, BM_PERSIST32, BM_WAVE, BM_COUNT };

// The batch's read offsets from the call's: reads [r0, r0 + nb) of
// args = {r0, nb}, then zero-length padding reads up to B (nb <= B and
// r0 + nb <= nq, checked by the host).
__global__ void k_stage_qoff(const uint64_t *__restrict__ qoff, const uint64_t *__restrict__ args,
                             uint64_t *__restrict__ bq, uint32_t B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > B) return;
    const uint64_t r0 = args[0], nb = args[1];
    bq[i] = qoff[r0 + (i <= nb ? i : nb)];
}
__global__ void k_npos_total(const uint64_t *mem_off, uint32_t B, const uint64_t *pos_off, uint64_t pos_cap,
                             uint64_t *tot) {
    const uint64_t t = pos_off[mem_off[B]];
    *tot = t < pos_cap ? t : pos_cap;
}
// dead pass numbers in the whole record buffer (stale records never match)
__global__ void k_fill_u64(uint64_t *d, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

struct BatchGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    void release() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        x = nullptr;
        g = nullptr;
    }
};

struct BatchBufs {
    uint32_t B = 0, W = 0, seg = 0; // reads, work items (capacity), codes per item
    uint32_t Bc = 0, Wb = 1;        // reads per callback batch, batches per window (B = Bc Wb)
    uint64_t slot_cap = 0, rec_cap = 0, hit_cap = 0, pos_cap = 0;
    DevBuf bq_off, slots, mem_cnt, status, pass_of, l1, l2, l3, ctr, recs, wsets;
    DevBuf mem_off, sorted, npos, pos_off, hits, rows, positions, misc, tmp;
    DevBuf items, item_cnt, item_first, read_off;
    DevBuf qnib; // all queries as base-set nibbles (read by every batch's graph)
    DevBuf bargs; // {first read, reads} of the batch, copied in before each launch
    DevBuf sub_off, slices; // per-batch read offsets and {first hit}, {first position} of each batch of a window
    DevBuf bcount;          // set-up: work items per callback batch and read offsets at batch starts
    DevBuf cscan;           // block sums of the device-count scan
    DevBuf part_cnt;        // record counts of the parts of split deferred items (kPartCntMax)
    const void *recs_dead = nullptr; // recs as last filled with dead records throughout
    // the set's instantiated graph, kept across calls while everything it
    // captured (pointers, sizes, configuration) is unchanged
    BatchGraph graph;
    std::vector<uint64_t> graph_key;
    const uint64_t *qoff_all = nullptr; // the call's read offsets (nq + 1)
    size_t tmp_bytes = 0;
    void release() {
        DevBuf *all[] = {&bq_off, &slots, &mem_cnt, &status, &pass_of, &l1, &l2, &l3, &ctr, &recs, &wsets,
                         &mem_off, &sorted, &npos, &pos_off, &hits, &rows, &positions, &misc, &tmp,
                         &items, &item_cnt, &item_first, &read_off, &qnib, &bargs, &sub_off, &slices, &bcount,
                         &cscan, &part_cnt};
        graph.release();
        recs_dead = nullptr;
        graph_key.clear();
        for (DevBuf *b : all) b->release();
    }
    // every buffer address of the set (a graph captures them)
    void addresses(std::vector<uint64_t> &k) const {
        const DevBuf *all[] = {&bq_off, &slots, &mem_cnt, &status, &pass_of, &l1, &l2, &l3, &ctr, &recs, &wsets,
                               &mem_off, &sorted, &npos, &pos_off, &hits, &rows, &positions, &misc, &tmp,
                               &items, &item_cnt, &item_first, &read_off, &qnib, &bargs, &sub_off, &slices,
                               &cscan, &part_cnt};
        for (const DevBuf *b : all) k.push_back(reinterpret_cast<uint64_t>(b->p));
    }
};
void free_batch_bufs(BatchBufs *b) {
    if (!b) return;
    b->release();
    delete b;
}
// items [item_first[B], W) of a batch: empty (read 0, no starts)
__global__ void k_pad_items(uint4 *items, const uint64_t *first_end, uint32_t W) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < W && i >= *first_end) items[i] = make_uint4(0, 0, 0, 0);
}
// misc: [0] positions-overflow count, [1] npos total, [2] walk guard
struct BatchMisc {
    unsigned long long pos_ovf, npos, guard, pad;
};

struct ToU64 {
    __device__ __forceinline__ uint64_t operator()(uint32_t v) const { return v; }
};
__global__ void k_zero_u64(uint64_t *p) { *p = 0; }

// Exclusive prefix sums of in[0, n) into out[0, n] (out[n] = the total) for
// n = min(*d_n, cap) read on the device: a graph cannot size a library scan by
// a device count, and a scan over the whole capacity reads and writes the
// capacity.  kCScanBlocks blocks each take a chunk of [0, n): their sums, one
// block's scan of the sums, then each chunk in tiles of 256 from its base.
constexpr uint32_t kCScanBlocks = 512;
__global__ __launch_bounds__(256) void k_cscan_sums(const uint32_t *__restrict__ in, uint64_t cap,
                                                    const uint64_t *__restrict__ d_n, uint64_t *__restrict__ part) {
    __shared__ uint64_t s_w[4];
    const uint64_t n = *d_n < cap ? *d_n : cap;
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    uint64_t sum = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) sum += in[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}
__global__ __launch_bounds__(kCScanBlocks) void k_cscan_part(uint64_t *__restrict__ part) {
    __shared__ uint64_t s_w[kCScanBlocks / 64];
    const uint64_t v = part[threadIdx.x];
    uint64_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t x = __shfl_up(incl, d, 64);
        if ((threadIdx.x & 63) >= (uint32_t)d) incl += x;
    }
    if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (uint32_t u = 0; u < kCScanBlocks / 64; u++) {
        before += u < (threadIdx.x >> 6) ? s_w[u] : 0ull;
        all += s_w[u];
    }
    __syncthreads();
    part[threadIdx.x] = before + incl - v;
    if (threadIdx.x == 0) part[kCScanBlocks] = all;
}
__global__ __launch_bounds__(256) void k_cscan_apply(const uint32_t *__restrict__ in, uint64_t cap,
                                                     const uint64_t *__restrict__ d_n,
                                                     const uint64_t *__restrict__ part, uint64_t *__restrict__ out) {
    __shared__ uint32_t s_part[256 / 64 + 1];
    const uint64_t n = *d_n < cap ? *d_n : cap;
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    uint64_t base = part[blockIdx.x];
    for (uint64_t t0 = lo; t0 < hi; t0 += 256) {
        const uint64_t i = t0 + threadIdx.x;
        const uint32_t v = i < hi ? in[i] : 0u;
        uint32_t tile = 0;
        const uint32_t incl = block_incl_scan<256>(v, s_part, tile);
        if (i < hi) out[i] = base + incl - v;
        base += tile;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = part[gridDim.x];
}
// The records of the previous window of this buffer set back to dead (its
// counters still hold how many it reserved; the rest of the buffer has been
// dead since the set's buffers were allocated).
__global__ void k_recs_reset(uint64_t *r0, uint64_t *r1, uint64_t half, const unsigned long long *used0,
                             const unsigned long long *used1) {
    const uint64_t n0 = *used0 < half ? *used0 : half, n1 = *used1 < half ? *used1 : half;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n0; i += stride) r0[i] = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n1; i += stride) r1[i] = ~0ull;
}

static int count_scan(hsfm_index *idx, BatchBufs &bb, const uint32_t *in, uint64_t *out, uint64_t cap,
                      const uint64_t *d_n) {
    hipStream_t s = idx->dev->stream;
    uint64_t *part = bb.cscan.as<uint64_t>();
    hipLaunchKernelGGL(k_cscan_sums, dim3(kCScanBlocks), dim3(256), 0, s, in, cap, d_n, part);
    hipLaunchKernelGGL(k_cscan_part, dim3(1), dim3(kCScanBlocks), 0, s, part);
    hipLaunchKernelGGL(k_cscan_apply, dim3(kCScanBlocks), dim3(256), 0, s, in, cap, d_n, part, out);
    HSFM_HIP(hipGetLastError());
    return HSFM_OK;
}

static int batch_scan(hsfm_index *idx, BatchBufs &bb, const uint32_t *in, uint64_t *out, uint64_t n) {
    hipStream_t s = idx->dev->stream;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, out);
    if (!n) return HSFM_OK;
    hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> it(in, ToU64());
    size_t tmp = bb.tmp_bytes;
    HSFM_HIP(hipcub::DeviceScan::InclusiveSum(bb.tmp.p, tmp, it, out + 1, (int)n, s));
    return HSFM_OK;
}

// Per callback batch of Bc reads: items[b] = its work items (reads of more
// than seg codes split), bound[b] = the offset of its first read, bound[ncb]
// = the total length; items zeroed by the caller.  Each thread sums a chunk
// of consecutive reads (at most two batches when Bc >= the chunk).
__global__ __launch_bounds__(256) void k_batch_counts(const uint64_t *__restrict__ qoff, uint32_t nq, uint32_t Bc,
                                                      uint32_t seg, unsigned long long *__restrict__ items,
                                                      uint64_t *__restrict__ bound, uint32_t ncb) {
    constexpr uint64_t kChunk = 256;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * kChunk;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kChunk; c < nq; c += stride) {
        uint32_t cur = 0xffffffffu;
        unsigned long long acc = 0;
        for (uint64_t r = c; r < c + kChunk && r < nq; r++) {
            const uint32_t b = (uint32_t)(r / Bc);
            if (b != cur) {
                if (acc) atomicAdd(&items[cur], acc);
                cur = b;
                acc = 0;
            }
            const uint64_t m = qoff[r + 1] - qoff[r];
            acc += m <= seg ? 1ull : (m + seg - 1) / seg;
            if (r % Bc == 0) bound[b] = qoff[r];
        }
        if (acc) atomicAdd(&items[cur], acc);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) bound[ncb] = qoff[nq];
}

// A window of Wb batches (Bc reads each; the window's read count nb in
// bargs[1]): out[j] = batch j's first hit, out[Wb + 1 + j] = its first
// position (j = 0..Wb; batches past nb start at the window's end).
constexpr uint32_t kMaxWindowBatches = 64;
__global__ void k_batch_slices(const uint64_t *__restrict__ read_off, const uint64_t *__restrict__ pos_off,
                               uint64_t hit_cap, const uint64_t *__restrict__ bargs, uint32_t Bc, uint32_t Wb,
                               uint64_t *__restrict__ out) {
    const uint32_t j = threadIdx.x;
    if (j > Wb) return;
    const uint64_t nb = bargs[1];
    uint64_t r = (uint64_t)j * Bc;
    r = r < nb ? r : nb;
    uint64_t hs = read_off[r];
    hs = hs < hit_cap ? hs : hit_cap;
    out[j] = hs;
    out[Wb + 1 + j] = pos_off[hs];
}
// Each batch's result relative to itself: read offsets from its first hit
// (Bc + 1 per batch in sub_off) and position offsets of its hits from its
// first position.
__global__ __launch_bounds__(256) void k_batch_rebase(const uint64_t *__restrict__ read_off,
                                                      const uint64_t *__restrict__ bargs, uint32_t Bc, uint32_t Wb,
                                                      const uint64_t *__restrict__ slices,
                                                      uint64_t *__restrict__ sub_off, hsfm_mem_hit *__restrict__ hits,
                                                      const uint64_t *__restrict__ nhits_p, uint64_t hit_cap) {
    __shared__ uint64_t s_hs[kMaxWindowBatches + 1], s_ps[kMaxWindowBatches + 1];
    if (threadIdx.x <= Wb) {
        s_hs[threadIdx.x] = slices[threadIdx.x];
        s_ps[threadIdx.x] = slices[Wb + 1 + threadIdx.x];
    }
    __syncthreads();
    const uint64_t nb = bargs[1];
    uint64_t nh = *nhits_p;
    nh = nh < hit_cap ? nh : hit_cap;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= nb; r += stride) {
        if (r < nb) { // read r: entry r - j Bc of batch j
            const uint64_t j = r / Bc;
            sub_off[j * (Bc + 1) + (r - j * Bc)] = read_off[r] - s_hs[j];
        }
        if (r > 0 && (r % Bc == 0 || r == nb)) { // the closing entry of the batch that ends at r
            const uint64_t j = (r - 1) / Bc;
            sub_off[j * (Bc + 1) + (r - j * Bc)] = read_off[r] - s_hs[j];
        }
    }
    for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += stride) {
        uint32_t j = 0; // the last batch starting at or before hit h
        for (uint32_t t = 1; t < Wb; t++) j = s_hs[t] <= h ? t : j;
        hits[h].pos_offset -= s_ps[j];
    }
}

// The launches of one batch (captured into a graph, or run eagerly).
static int batch_record(hsfm_index *idx, BatchBufs &bb, const SeedView &sv, const SeedView &svp,
                        const uint8_t *d_q, const uint32_t *qnib, uint32_t min_len,
                        const uint32_t *d_refb, uint32_t n_refs, uint32_t max_hits, int main_kernel, int rpt) {
    hipStream_t s = idx->dev->stream;
    const uint32_t B = bb.B, W = bb.W;
    const uint4 *items = bb.items.as<uint4>();
    MemCounters *c = bb.ctr.as<MemCounters>();
    const DevConst *cst = idx->d_const.as<DevConst>();
    const uint64_t *bq = bb.bq_off.as<uint64_t>();
    int rc;
    // stage the batch: its read offsets (from bargs), then its work items and
    // empty items up to the capacity
    hipLaunchKernelGGL(k_stage_qoff, dim3((B + 256) / 256), dim3(256), 0, s, bb.qoff_all, bb.bargs.as<uint64_t>(),
                       bb.bq_off.as<uint64_t>(), B);
    hipLaunchKernelGGL(k_items_count, dim3((B + 255) / 256), dim3(256), 0, s, B, bb.bq_off.as<uint64_t>(), bb.seg,
                       bb.item_cnt.as<uint32_t>());
    if ((rc = batch_scan(idx, bb, bb.item_cnt.as<uint32_t>(), bb.item_first.as<uint64_t>(), B))) return rc;
    hipLaunchKernelGGL(k_items_fill, dim3((B + 255) / 256), dim3(256), 0, s, B, bb.bq_off.as<uint64_t>(), bb.seg,
                       bb.item_first.as<uint64_t>(), bb.items.as<uint4>());
    hipLaunchKernelGGL(k_pad_items, dim3((bb.W + 255) / 256), dim3(256), 0, s, bb.items.as<uint4>(),
                       bb.item_first.as<uint64_t>() + B, bb.W);
    HSFM_HIP(hipGetLastError());
    const uint64_t half = bb.rec_cap / 2;
    MemRec *r0 = bb.recs.as<MemRec>(), *r1 = r0 + half;
    // the previous window's records back to dead, before its counters are cleared
    hipLaunchKernelGGL(k_recs_reset, dim3(idx->dev->num_cus * 4), dim3(256), 0, s, reinterpret_cast<uint64_t *>(r0),
                       reinterpret_cast<uint64_t *>(r1), half, &c[1].rec_used, &c[2].rec_used);
    if ((rc = fill_u32(idx, bb.ctr.as<uint32_t>(), 3 * sizeof(MemCounters) / 4, 0)) ||
        (rc = fill_u32(idx, bb.misc.as<uint32_t>(), sizeof(BatchMisc) / 4, 0)) ||
        (rc = fill_u32(idx, bb.mem_cnt.as<uint32_t>(), W, 0)) ||
        (rc = fill_u32(idx, bb.pass_of.as<uint32_t>(), W, 0xffffffffu)) ||
        (rc = fill_u32(idx, bb.status.as<uint32_t>(), (W + 3) / 4, main_kernel == BM_WAVE ? 0x01010101u : 0u)))
        return rc;
    const uint32_t lds_waves = persist_grid(idx, k_mem_wave<true, 64>, 64);
    const unsigned long long *second_count = &c[0].n_retry;
    if (main_kernel == BM_WAVE) {
        // one wave per read for every read
        hipLaunchKernelGGL((k_mem_wave<true, 64>), dim3(lds_waves), dim3(64), 0, s, idx->fwd.view, idx->rev.view, cst,
                           svp, d_q, bq, (const uint32_t *)nullptr, W, items, min_len, kWaveLdsCap, (uint4 *)nullptr, r0,
                           half,
                           bb.mem_cnt.as<uint32_t>(), bb.pass_of.as<uint32_t>(), 0u, bb.l2.as<uint32_t>(), &c[1],
                           (const unsigned long long *)nullptr, (uint32_t *)nullptr, 1u);
    } else {
        uint4 *slots = bb.slots.as<uint4>();
        uint32_t *cnt = bb.mem_cnt.as<uint32_t>();
        uint8_t *st = bb.status.as<uint8_t>();
        if (main_kernel == BM_PERSIST8)
            hipLaunchKernelGGL((k_mem_persist<8, 256>), dim3(persist_grid(idx, k_mem_persist<8, 256>, 256)), dim3(256),
                               0, s, idx->fwd.view, idx->rev.view, cst, svp, d_q, qnib, bq, (const uint32_t *)nullptr, W,
                               items, min_len, slots, cnt, st, bb.l1.as<uint32_t>(), &c[0], LsiView{}, 0u);
        else if (main_kernel == BM_PERSIST32)
            hipLaunchKernelGGL((k_mem_persist<32, 128>), dim3(persist_grid(idx, k_mem_persist<32, 128>, 128)),
                               dim3(128), 0, s, idx->fwd.view, idx->rev.view, cst, svp, d_q, qnib, bq,
                               (const uint32_t *)nullptr, W, items, min_len, slots, cnt, st, bb.l1.as<uint32_t>(),
                               &c[0], LsiView{}, 0u);
        else
            hipLaunchKernelGGL((k_mem_persist<16, 256>), dim3(persist_grid(idx, k_mem_persist<16, 256>, 256)),
                               dim3(256), 0, s, idx->fwd.view, idx->rev.view, cst, svp, d_q, qnib, bq,
                               (const uint32_t *)nullptr, W, items, min_len, slots, cnt, st, bb.l1.as<uint32_t>(),
                               &c[0], LsiView{}, 0u);
        HSFM_HIP(hipGetLastError());
        // deferred reads: kDeferTB threads each, sets in LDS, split into parts
        // when they are few
        const uint32_t dg = persist_grid(idx, k_mem_wave<true, kDeferTB>, kDeferTB);
        hipLaunchKernelGGL((k_mem_wave<true, kDeferTB>), dim3(dg), dim3(kDeferTB), 0, s,
                           idx->fwd.view, idx->rev.view, cst, svp, d_q, bq, bb.l1.as<uint32_t>(), W, items, min_len,
                           kWaveLdsCap, (uint4 *)nullptr, r0, half, cnt, bb.pass_of.as<uint32_t>(), 0u,
                           bb.l2.as<uint32_t>(), &c[1], second_count,
                           dg <= kPartCntMax ? bb.part_cnt.as<uint32_t>() : nullptr, kDeferParts);
    }
    HSFM_HIP(hipGetLastError());
    // sets beyond the LDS: 1024 threads per read, global sets
    hipLaunchKernelGGL((k_mem_wave<false, 1024>), dim3(kBatchGlobalWaves), dim3(1024), 0, s, idx->fwd.view,
                       idx->rev.view, cst, svp, d_q, bq, bb.l2.as<uint32_t>(), W, items, min_len, kBatchGlobalCap,
                       bb.wsets.as<uint4>(), r1, half, bb.mem_cnt.as<uint32_t>(), bb.pass_of.as<uint32_t>(), 1u,
                       bb.l3.as<uint32_t>(), &c[2], &c[1].n_retry, (uint32_t *)nullptr, 1u);
    HSFM_HIP(hipGetLastError());
    uint64_t *mem_off = bb.mem_off.as<uint64_t>();
    if ((rc = batch_scan(idx, bb, bb.mem_cnt.as<uint32_t>(), mem_off, W))) return rc;
    uint4 *sorted = bb.sorted.as<uint4>();
    if (main_kernel != BM_WAVE)
        hipLaunchKernelGGL(k_mem_gather, dim3((W + 255) / 256), dim3(256), 0, s, W, bq, items,
                           bb.status.as<uint8_t>(), bb.mem_cnt.as<uint32_t>(), mem_off, bb.slots.as<uint4>(), sorted);
    hipLaunchKernelGGL(k_read_off, dim3((B + 256) / 256), dim3(256), 0, s, B, bb.item_first.as<uint64_t>(), mem_off,
                       bb.read_off.as<uint64_t>());
    const dim3 hg(idx->dev->num_cus * 8); // grid-stride over the device counts
    hipLaunchKernelGGL(k_mem_scatter, hg, dim3(256), 0, s, r0, half, W, bb.pass_of.as<uint32_t>(), mem_off, sorted,
                       &c[1].rec_used, (const uint32_t *)bb.part_cnt.as<uint32_t>());
    hipLaunchKernelGGL(k_mem_scatter, hg, dim3(256), 0, s, r1, half, W, bb.pass_of.as<uint32_t>(), mem_off, sorted,
                       &c[2].rec_used, (const uint32_t *)nullptr);
    HSFM_HIP(hipGetLastError());
    const dim3 cg(idx->dev->num_cus * 8);
    hipLaunchKernelGGL(k_mem_npos, cg, dim3(256), 0, s, sorted, bb.hit_cap, max_hits, n_refs ? 1u : 0u,
                       bb.npos.as<uint32_t>(), mem_off + W);
    if ((rc = count_scan(idx, bb, bb.npos.as<uint32_t>(), bb.pos_off.as<uint64_t>(), bb.hit_cap, mem_off + W)))
        return rc;
    BatchMisc *misc = bb.misc.as<BatchMisc>();
    hipLaunchKernelGGL(k_mem_hits, cg, dim3(256), 0, s, sorted, bb.hit_cap, max_hits, bb.pos_off.as<uint64_t>(),
                       bb.hits.as<hsfm_mem_hit>(), bb.rows.as<uint32_t>(), mem_off + W, bb.pos_cap, &misc->pos_ovf);
    hipLaunchKernelGGL(k_npos_total, dim3(1), dim3(1), 0, s, mem_off, W, bb.pos_off.as<uint64_t>(), bb.pos_cap,
                       reinterpret_cast<uint64_t *>(&misc->npos));
    HSFM_HIP(hipGetLastError());
    if (n_refs)
        if ((rc = launch_resolve_fan(idx, idx->fwd.view, bb.rows.as<uint32_t>(), bb.pos_cap,
                                     reinterpret_cast<const uint64_t *>(&misc->npos), d_refb, n_refs,
                                     bb.positions.as<uint32_t>(), &misc->guard, rpt, idx->dev->num_cus * 8)))
            return rc;
    if (bb.Wb > 1) { // the window's batches, each relative to itself
        hipLaunchKernelGGL(k_batch_slices, dim3(1), dim3(64), 0, s, bb.read_off.as<uint64_t>(),
                           bb.pos_off.as<uint64_t>(), bb.hit_cap, bb.bargs.as<uint64_t>(), bb.Bc, bb.Wb,
                           bb.slices.as<uint64_t>());
        hipLaunchKernelGGL(k_batch_rebase, dim3(1024), dim3(256), 0, s, bb.read_off.as<uint64_t>(),
                           bb.bargs.as<uint64_t>(), bb.Bc, bb.Wb, bb.slices.as<uint64_t>(), bb.sub_off.as<uint64_t>(),
                           bb.hits.as<hsfm_mem_hit>(), mem_off + W, bb.hit_cap);
        HSFM_HIP(hipGetLastError());
    }
    return HSFM_OK;
}

static int batch_capture(hsfm_index *idx, BatchBufs &bb, const SeedView &sv, const SeedView &svp,
                         const uint8_t *d_q, const uint32_t *qnib, uint32_t min_len,
                         const uint32_t *d_refb, uint32_t n_refs, uint32_t max_hits, int main_kernel, int rpt,
                         BatchGraph &out) {
    hipStream_t s = idx->dev->stream;
    HSFM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = batch_record(idx, bb, sv, svp, d_q, qnib, min_len, d_refb, n_refs, max_hits, main_kernel, rpt);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(HSFM_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    out.g = g;
    HSFM_HIP(hipGraphInstantiate(&out.x, g, nullptr, nullptr, 0));
    return HSFM_OK;
}

// The set's graph for this call: the cached one when nothing it captured has
// changed, else a new capture (replacing it).
static int batch_graph(hsfm_index *idx, BatchBufs &bb, const SeedView &sv, const SeedView &svp,
                       const uint8_t *d_q, const uint32_t *qnib, uint32_t min_len, const uint32_t *d_refb,
                       uint32_t n_refs, uint32_t max_hits, int main_kernel, int rpt, hipGraphExec_t &out,
                       uint32_t &captured) {
    std::vector<uint64_t> key = {
        reinterpret_cast<uint64_t>(d_q), reinterpret_cast<uint64_t>(qnib), reinterpret_cast<uint64_t>(d_refb),
        reinterpret_cast<uint64_t>(bb.qoff_all), n_refs, min_len, max_hits, (uint64_t)main_kernel, (uint64_t)rpt,
        bb.B, bb.W, bb.Bc, bb.Wb, bb.seg, bb.slot_cap, bb.rec_cap, bb.hit_cap, bb.pos_cap, bb.tmp_bytes,
        reinterpret_cast<uint64_t>(sv.off), reinterpret_cast<uint64_t>(sv.iv), reinterpret_cast<uint64_t>(sv.lsl), sv.F,
        reinterpret_cast<uint64_t>(svp.rec), reinterpret_cast<uint64_t>(svp.off), reinterpret_cast<uint64_t>(svp.iv),
        svp.F, reinterpret_cast<uint64_t>(idx->d_const.p)};
    for (const DevDir *d : {&idx->fwd.view, &idx->rev.view})
        key.insert(key.end(), {reinterpret_cast<uint64_t>(d->occ), reinterpret_cast<uint64_t>(d->sb),
                               reinterpret_cast<uint64_t>(d->samp), reinterpret_cast<uint64_t>(d->sa_vals),
                               reinterpret_cast<uint64_t>(d->lookup), d->n, reinterpret_cast<uint64_t>(d->sa_full),
                               reinterpret_cast<uint64_t>(d->text), reinterpret_cast<uint64_t>(d->isa),
                               reinterpret_cast<uint64_t>(d->text4), d->qnib16, d->num_samples, d->sample_rate,
                               d->num_lanes, d->lookup_len});
    bb.addresses(key);
    if (!bb.graph.x || key != bb.graph_key) {
        bb.graph.release();
        bb.graph_key.clear();
        int rc = batch_capture(idx, bb, sv, svp, d_q, qnib, min_len, d_refb, n_refs, max_hits, main_kernel, rpt,
                               bb.graph);
        if (rc) return rc;
        bb.graph_key = key;
        captured++;
    }
    out = bb.graph.x;
    return HSFM_OK;
}

// Device buffers of one batch set for the capacities in bb (kept and grown).
static int batch_alloc(BatchBufs &bb) {
    const size_t Wn = bb.W;
    const uint32_t B = bb.B;
    int rc;
    if ((rc = bb.bq_off.ensure(((size_t)B + 1) * 8)) || (rc = bb.bargs.ensure(64)) || (rc = bb.slots.ensure(bb.slot_cap * 16)) ||
        (rc = bb.mem_cnt.ensure(Wn * 4 + 4)) || (rc = bb.status.ensure(Wn + 8)) ||
        (rc = bb.pass_of.ensure(Wn * 4 + 4)) || (rc = bb.l1.ensure(Wn * 4 + 4)) ||
        (rc = bb.l2.ensure(Wn * 4 + 4)) || (rc = bb.l3.ensure(Wn * 4 + 4)) ||
        (rc = bb.items.ensure(Wn * 16 + 16)) || (rc = bb.item_cnt.ensure((size_t)B * 4 + 4)) ||
        (rc = bb.item_first.ensure(((size_t)B + 1) * 8)) || (rc = bb.read_off.ensure(((size_t)B + 1) * 8)) ||
        (rc = bb.ctr.ensure(3 * sizeof(MemCounters))) || (rc = bb.recs.ensure(bb.rec_cap * sizeof(MemRec))) ||
        (rc = bb.wsets.ensure((size_t)kBatchGlobalWaves * 2 * kBatchGlobalCap * 16)) ||
        (rc = bb.mem_off.ensure((Wn + 1) * 8)) || (rc = bb.sorted.ensure(bb.hit_cap * 16)) ||
        (rc = bb.npos.ensure(bb.hit_cap * 4 + 4)) || (rc = bb.pos_off.ensure((bb.hit_cap + 1) * 8)) ||
        (rc = bb.hits.ensure(bb.hit_cap * sizeof(hsfm_mem_hit))) || (rc = bb.rows.ensure(bb.pos_cap * 4)) ||
        (rc = bb.positions.ensure(bb.pos_cap * 8)) || (rc = bb.misc.ensure(sizeof(BatchMisc))) ||
        (rc = bb.tmp.ensure(bb.tmp_bytes + 256)) || (rc = bb.sub_off.ensure(((size_t)B + bb.Wb + 1) * 8)) ||
        (rc = bb.slices.ensure((2ull * bb.Wb + 2) * 8 + 16)) || (rc = bb.cscan.ensure((kCScanBlocks + 1) * 8)) ||
        (rc = bb.part_cnt.ensure((size_t)kPartCntMax * 4)))
        return rc;
    return HSFM_OK;
}
// A record buffer new to the set starts dead throughout (later windows reset
// only what the previous one reserved) with zero counters.
static int batch_recs_dead(hsfm_index *idx, BatchBufs &bb) {
    if (bb.recs_dead == bb.recs.p) return HSFM_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, idx->dev->stream, bb.recs.as<uint64_t>(), bb.rec_cap * 4,
                       ~0ull);
    HSFM_HIP(hipGetLastError());
    int rc = fill_u32(idx, bb.ctr.as<uint32_t>(), 3 * sizeof(MemCounters) / 4, 0);
    if (rc) return rc;
    bb.recs_dead = bb.recs.p;
    return HSFM_OK;
}

// The batch's first read and size for the next launch (the graph stages the
// offsets and work items itself).  The pinned slot is rewritten only after
// the previous batch's wait, so the copy has run by then.
static int batch_set_args(hsfm_index *idx, BatchBufs &bb, uint32_t r0, uint32_t nb, hipStream_t s, int k) {
    if (!idx->ws.h_pinned) HSFM_HIP(hipHostMalloc((void **)&idx->ws.h_pinned, kPinnedBytes, hipHostMallocDefault));
    uint64_t *a = idx->ws.h_pinned + kPinnedArgs + 8 * k; // this set's batch arguments
    a[0] = r0;
    a[1] = nb;
    HSFM_HIP(hipMemcpyAsync(bb.bargs.p, a, 16, hipMemcpyHostToDevice, s));
    return HSFM_OK;
}

} // namespace hsfm

extern "C" int hsfm_find_mems_batched(hsfm_index *idx, const uint8_t *d_q, const uint64_t *d_qoff, uint32_t nq,
                                      uint32_t min_len, int mode, const uint32_t *d_refb, uint32_t n_refs,
                                      uint32_t max_hits, const hsfm_batch_config *cfg, hsfm_batch_fn fn, void *user,
                                      hsfm_batch_stats *st) {
    if (!idx || !cfg || (nq && (!d_q || !d_qoff)) || (n_refs && !d_refb))
        return set_error(HSFM_EINVAL, "hsfm_find_mems_batched: null argument");
    if (mode != HSFM_MODE_SMEM && mode != HSFM_MODE_MEM) return set_error(HSFM_EINVAL, "bad mode %d", mode);
    if (!idx->has_rev) return set_error(HSFM_EINVAL, "MEM search needs an index uploaded with its reverse index");
    if (cfg->batch_reads == 0) return set_error(HSFM_EINVAL, "batch_reads must be positive");
    const uint32_t F = 12;
    if (min_len <= F) return set_error(HSFM_EINVAL, "the batched pipeline needs min_len > %u", F);
    if (tl_callback_index == idx)
        return set_error(HSFM_EINVAL, "hsfm_find_mems_batched: called from a batch callback on the same index");
    std::unique_lock<std::mutex> lk(idx->mu);
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    hipStream_t s = idx->dev->stream;
    int rc;
    hsfm_batch_stats stats{};
    if ((rc = ensure_sa_full(idx))) return rc;
    if ((rc = build_seed_table(idx, F))) return rc;
    SeedView sv{idx->seed_off.as<uint64_t>(), idx->seed_iv.as<uint4>(), idx->seed_lsl.as<uint8_t>(), F};
    uint32_t FP = mem_seed_length(idx, min_len);
    if ((rc = packed_seeds(idx, FP))) return rc;
    const SeedView svp = packed_view(idx, FP, sv);
    // capacities from per-batch counts computed on the device (work items of
    // every callback batch, read offsets at batch starts)
    const uint32_t Bc = cfg->batch_reads < nq ? cfg->batch_reads : (nq ? nq : 1);
    const uint32_t ncb = nq ? (nq + Bc - 1) / Bc : 0;
    const uint32_t seg0 = mem_segment_length() ? mem_segment_length() : 0xffffffffu;
    if (!idx->batch[0]) idx->batch[0] = new BatchBufs();
    BatchBufs &bb = *idx->batch[0]; // kept across calls: ensure() below only grows it
    std::vector<uint64_t> bitems(ncb), bound((size_t)ncb + 1, 0);
    if (ncb) {
        if ((rc = bb.bcount.ensure(((size_t)2 * ncb + 1) * 8))) return rc;
        unsigned long long *d_items = bb.bcount.as<unsigned long long>();
        uint64_t *d_bound = bb.bcount.as<uint64_t>() + ncb;
        HSFM_HIP(hipMemsetAsync(d_items, 0, (size_t)ncb * 8, s));
        const uint64_t chunks = ((uint64_t)nq + 255) / 256;
        const uint32_t grid = (uint32_t)((chunks + 255) / 256 < 4096 ? (chunks + 255) / 256 : 4096);
        hipLaunchKernelGGL(k_batch_counts, dim3(grid ? grid : 1), dim3(256), 0, s, d_qoff, nq, Bc, seg0, d_items,
                           d_bound, ncb);
        HSFM_HIP(hipGetLastError());
        HSFM_HIP(hipMemcpyAsync(bitems.data(), d_items, (size_t)ncb * 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipMemcpyAsync(bound.data(), d_bound, ((size_t)ncb + 1) * 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
    }
    // callback batches of Bc reads; graphs over windows of Wb batches, so that
    // the persistent kernel has about eight work items per lane (a 64k-read
    // batch alone leaves most lanes idle while its last items finish; 2M
    // uneven reads: 38 ms with one batch per graph, 18 ms with 11, 17 ms with
    // 16, against 10 ms for one call over all reads)
    uint32_t Wb = cfg->window_batches;
    if (!Wb) {
        uint64_t items = 0;
        for (uint32_t b = 0; b < ncb; b++) items += bitems[b];
        const double per_read = nq ? (double)items / nq : 1.0;
        const double lanes = (double)persist_grid(idx, k_mem_persist<16, 256>, 256) * 256.0;
        // about sixteen work items per lane: the persistent kernel's end (its
        // last items) is then a small part of the window
        Wb = (uint32_t)ceil(16.0 * lanes / (per_read * (double)Bc));
    }
    Wb = Wb < 1 ? 1 : (Wb > kMaxWindowBatches ? kMaxWindowBatches : Wb);
    if (Wb > ncb) Wb = ncb ? ncb : 1;
    if ((uint64_t)Bc * Wb > 0x7fffffffull) Wb = (uint32_t)(0x7fffffffull / Bc) ? (uint32_t)(0x7fffffffull / Bc) : 1;
    const uint32_t B = (uint64_t)Bc * Wb < nq ? Bc * Wb : (nq ? nq : 1);
    stats.window_batches = Wb;
    bb.B = B;
    bb.Bc = Bc;
    bb.Wb = Wb;
    bb.qoff_all = d_qoff;
    if ((rc = pack_bases(idx, d_q, bound[ncb], bb.qnib))) {
        bb.release();
        return rc;
    }
    bb.seg = seg0;
    // item capacity and slot ranges over all windows (slot index of item w:
    // ((qoff + start) >> 3) + 8 w)
    uint64_t max_w = 0, max_slots = 0, max_batch_slots = 0;
    for (uint32_t b0 = 0; b0 < ncb; b0 += Wb) {
        const uint32_t b1 = b0 + Wb < ncb ? b0 + Wb : ncb;
        const uint32_t r0 = b0 * Bc, r1 = b1 * Bc < nq ? b1 * Bc : nq;
        uint64_t w = B - (r1 - r0); // padding reads: one item each
        for (uint32_t b = b0; b < b1; b++) w += bitems[b];
        max_w = w > max_w ? w : max_w;
    }
    if (max_w > 0x7fffffffull) return set_error(HSFM_EINVAL, "too many read segments in a batch");
    bb.W = (uint32_t)(max_w ? max_w : 1);
    for (uint32_t b0 = 0; b0 < ncb; b0 += Wb) {
        const uint32_t b1 = b0 + Wb < ncb ? b0 + Wb : ncb;
        const uint64_t sl = (bound[b1] >> 3) + 8ull * bb.W + 16;
        max_slots = sl > max_slots ? sl : max_slots;
        const uint64_t v = (bound[b1] >> 3) - (bound[b0] >> 3) + 9ull * bb.W + 8;
        max_batch_slots = v > max_batch_slots ? v : max_batch_slots;
    }
    bb.slot_cap = max_slots;
    bb.rec_cap = 2ull * (16ull * B + 65536);
    bb.hit_cap = max_batch_slots + bb.rec_cap;
    bb.pos_cap = 4 * bb.hit_cap + 65536;
    {
        size_t t1 = 0, t2 = 0;
        hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> it(nullptr, ToU64());
        HSFM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, t1, it, (uint64_t *)nullptr, (int)bb.W, s));
        bb.tmp_bytes = t1 > t2 ? t1 : t2;
    }
    if ((rc = batch_alloc(bb)) || (rc = batch_recs_dead(idx, bb))) {
        bb.release();
        return rc;
    }
    hipEvent_t e0, e1;
    HSFM_HIP(hipEventCreate(&e0));
    HSFM_HIP(hipEventCreate(&e1));
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return code;
    };
    // configuration: explicit, cached, or tuned on the first batch
    int main_kernel = cfg->main_kernel, rpt = cfg->resolve_fanout;
    const bool cached = idx->tuned_main >= 0 && idx->tuned_min_len == min_len && idx->tuned_batch == B;
    if (main_kernel < 0 || rpt <= 0) {
        if (cached && !cfg->autotune) {
            main_kernel = main_kernel < 0 ? idx->tuned_main : main_kernel;
            rpt = rpt <= 0 ? idx->tuned_rpt : rpt;
        } else if (cfg->autotune && nq) {
            HSFM_HIP(hipEventRecord(e0, s));
            const uint32_t nb0 = B < nq ? B : nq;
            auto time_cfg = [&](int mk, int rp, double &ms) -> int {
                BatchGraph gr;
                int r2 = batch_capture(idx, bb, sv, svp, d_q, bb.qnib.as<uint32_t>(), min_len, d_refb, n_refs, max_hits, mk, rp, gr);
                if (r2) return r2;
                ms = 1e30;
                for (int rep = 0; rep < 2; rep++) {
                    if ((r2 = batch_set_args(idx, bb, 0, nb0, s, 0))) break;
                    hipEvent_t a, b;
                    HSFM_HIP(hipEventCreate(&a));
                    HSFM_HIP(hipEventCreate(&b));
                    HSFM_HIP(hipEventRecord(a, s));
                    HSFM_HIP(hipGraphLaunch(gr.x, s));
                    HSFM_HIP(hipEventRecord(b, s));
                    float t = 0.f;
                    HSFM_HIP(hipEventSynchronize(b));
                    HSFM_HIP(hipEventElapsedTime(&t, a, b));
                    (void)hipEventDestroy(a);
                    (void)hipEventDestroy(b);
                    if (rep) ms = t;
                }
                gr.release();
                return r2;
            };
            double best = 1e30;
            int bm = BM_PERSIST16;
            for (int mk = 0; mk < BM_COUNT; mk++) {
                double ms;
                if ((rc = time_cfg(mk, 1, ms))) return cleanup(rc);
                stats.tune_ms[mk] = ms;
            }
            // the 16-interval ring unless another main kernel is clearly faster
            // (the ring of 8 defers more items, whose passes then contend with
            // the next window's main kernel)
            best = stats.tune_ms[BM_PERSIST16];
            for (int mk = 0; mk < BM_COUNT; mk++)
                if (stats.tune_ms[mk] < 0.95 * best) { best = stats.tune_ms[mk]; bm = mk; }
            int br = 1;
            const int rps[2] = {2, 4};
            for (int i = 0; i < 2; i++) {
                double ms;
                if ((rc = time_cfg(bm, rps[i], ms))) return cleanup(rc);
                stats.tune_ms[BM_COUNT + i] = ms;
                if (ms < best) { best = ms; br = rps[i]; }
            }
            main_kernel = cfg->main_kernel >= 0 ? cfg->main_kernel : bm;
            rpt = cfg->resolve_fanout > 0 ? cfg->resolve_fanout : br;
            idx->tuned_main = bm;
            idx->tuned_rpt = br;
            idx->tuned_min_len = min_len;
            idx->tuned_batch = B;
            HSFM_HIP(hipEventRecord(e1, s));
            HSFM_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            HSFM_HIP(hipEventElapsedTime(&t, e0, e1));
            stats.tune_total_ms = t;
        } else {
            main_kernel = main_kernel < 0 ? BM_PERSIST16 : main_kernel;
            rpt = rpt <= 0 ? 1 : rpt;
        }
    }
    if (main_kernel < 0 || main_kernel >= BM_COUNT || (rpt != 1 && rpt != 2 && rpt != 4))
        return cleanup(set_error(HSFM_EINVAL, "bad batch configuration (main %d, fan-out %d)", main_kernel, rpt));
    stats.main_kernel = main_kernel;
    stats.resolve_fanout = rpt;
    // Several buffer sets, each with its own graph and stream, when there is
    // more than one batch: batch b + 1 runs while batch b's tail drains (a
    // 64k-read batch has fewer work items than the persistent kernel has
    // lanes, so its last items leave most of the GPU idle) and while the host
    // hands batch b to the callback.  Batches complete and reach the callback
    // in order.
    constexpr int KS = hsfm_index::kMaxBatchSets;
    hipGraphExec_t gx[KS] = {};
    BatchBufs *set[KS] = {};
    hipStream_t sst[KS] = {};
    set[0] = &bb;
    sst[0] = s;
    const uint32_t nbatches = nq ? (nq + B - 1) / B : 0;
    int nsets = 1;
    if (cfg->use_graph && nbatches > 1) {
        const int v = cfg->batch_sets > 0 ? cfg->batch_sets : 3; // 3 measured best of 1-4
        nsets = v > KS ? KS : v;
        if ((uint32_t)nsets > nbatches) nsets = (int)nbatches;
    }
    // (each batch keeps the whole persistent grid: a grid split between the
    // batches in flight measured slower, 2M uneven reads: 54.5 ms with three
    // thirds against 47 ms with three whole grids)
    if (cfg->use_graph && (rc = batch_graph(idx, bb, sv, svp, d_q, bb.qnib.as<uint32_t>(), min_len, d_refb, n_refs,
                                            max_hits, main_kernel, rpt, gx[0], stats.graphs_captured)))
        return cleanup(rc);
    for (int k = 1; k < nsets; k++) {
        if (!idx->batch[k]) idx->batch[k] = new BatchBufs();
        BatchBufs &bk = *idx->batch[k];
        bk.B = bb.B; bk.W = bb.W; bk.seg = bb.seg; bk.Bc = bb.Bc; bk.Wb = bb.Wb;
        bk.slot_cap = bb.slot_cap; bk.rec_cap = bb.rec_cap; bk.hit_cap = bb.hit_cap; bk.pos_cap = bb.pos_cap;
        bk.tmp_bytes = bb.tmp_bytes;
        bk.qoff_all = d_qoff;
        if ((rc = batch_alloc(bk)) || (rc = batch_recs_dead(idx, bk))) return cleanup(rc);
        if (!idx->bstream[k] && hipStreamCreateWithFlags(&idx->bstream[k], hipStreamNonBlocking) != hipSuccess)
            return cleanup(set_error(HSFM_EDEVICE, "hipStreamCreate failed"));
        set[k] = &bk;
        sst[k] = idx->bstream[k];
        if ((rc = batch_graph(idx, bk, sv, svp, d_q, bb.qnib.as<uint32_t>(), min_len, d_refb, n_refs, max_hits,
                              main_kernel, rpt, gx[k], stats.graphs_captured)))
            return cleanup(rc);
    }
    HSFM_HIP(hipEventRecord(e0, s));
    for (int k = 1; k < nsets; k++) HSFM_HIP(hipStreamWaitEvent(sst[k], e0, 0)); // after the set-up on s
    struct Pending {
        uint32_t r0, nb;
        int k;
    };
    Pending pend[KS];
    int npend = 0;
    // wait for batch pd, read its counters back (one round trip) and hand it on
    auto finish = [&](const Pending &pd) -> int {
        BatchBufs &b = *set[pd.k];
        const hipStream_t st = sst[pd.k];
        MemCounters *c = b.ctr.as<MemCounters>();
        const uint64_t *m = b.misc.as<uint64_t>();
        uint64_t h[8 + 2 * (kMaxWindowBatches + 1)];
        uint64_t *hp = idx->ws.h_pinned + 64 * (pd.k + 1); // this set's read-back slots
        const size_t nsl = b.Wb > 1 ? 2 * ((size_t)b.Wb + 1) : 0;
        if (hipMemcpyAsync(hp, &c[2].n_retry, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hp + 1, m, 24, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hp + 4, b.mem_off.as<uint64_t>() + b.W, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            (nsl && hipMemcpyAsync(hp + 8, b.slices.p, nsl * 8, hipMemcpyDeviceToHost, st) != hipSuccess) ||
            hipStreamSynchronize(st) != hipSuccess)
            return set_error(HSFM_EDEVICE, "batch result read-back failed");
        memcpy(h, hp, (8 + nsl) * 8);
        const uint64_t late = h[0], pos_ovf = h[1], npos = h[2], guard = h[3], nhits = h[4];
        if (guard)
            return set_error(HSFM_EINDEX, "%llu LF walks exceeded the sample rate: index is inconsistent",
                             (unsigned long long)guard);
        stats.windows++;
        const bool redo = late || pos_ovf || nhits > b.hit_cap;
        for (uint32_t j = 0; j * b.Bc < pd.nb; j++) {
            const uint32_t r0 = pd.r0 + j * b.Bc, nb = pd.nb - j * b.Bc < b.Bc ? pd.nb - j * b.Bc : b.Bc;
            stats.batches++;
            hsfm_mem_device_result res{};
            if (redo) {
                // outgrew the in-graph capacities: this batch through the host-driven path
                stats.fallback_batches++;
                // (with the index lock still held: its result lives in the index's
                // workspace until the callback below has read it)
                const int r2 = mem_device_impl(idx, d_q, d_qoff + r0, nb, min_len, d_refb, n_refs, max_hits, &res);
                if (r2) return r2;
            } else if (b.Wb > 1) {
                const uint64_t *hs = h + 8, *ps = h + 8 + b.Wb + 1;
                res.num_queries = nb;
                res.num_hits = hs[j + 1] - hs[j];
                res.num_positions = n_refs ? ps[j + 1] - ps[j] : 0;
                res.d_hit_offsets = b.sub_off.as<uint64_t>() + (size_t)j * (b.Bc + 1);
                res.d_hits = b.hits.as<hsfm_mem_hit>() + hs[j];
                res.d_positions = b.positions.as<uint32_t>() + 2 * ps[j];
            } else {
                res.num_queries = nb;
                res.num_hits = nhits;
                res.num_positions = n_refs ? npos : 0;
                res.d_hit_offsets = b.read_off.as<uint64_t>();
                res.d_hits = b.hits.as<hsfm_mem_hit>();
                res.d_positions = b.positions.as<uint32_t>();
            }
            stats.records += res.num_hits;
            stats.positions += res.num_positions;
            if (fn) {
                // the lock stays held: calls on this index from fn return HSFM_EINVAL
                const hsfm_index *outer = tl_callback_index;
                tl_callback_index = idx;
                fn(user, r0, nb, &res);
                tl_callback_index = outer;
            }
        }
        return HSFM_OK;
    };
    uint32_t kb = 0;
    for (uint32_t r0 = 0; r0 < nq && !rc; r0 += B, kb++) {
        const uint32_t nb = r0 + B < nq ? B : nq - r0;
        const int k = (int)(kb % (uint32_t)nsets);
        if (npend == nsets) { // the oldest pending batch holds set k: finish it first
            rc = finish(pend[0]);
            for (int x = 1; x < npend; x++) pend[x - 1] = pend[x];
            npend--;
            if (rc) break;
        }
        if ((rc = batch_set_args(idx, *set[k], r0, nb, sst[k], k))) break;
        if (cfg->use_graph) {
            if (hipGraphLaunch(gx[k], sst[k]) != hipSuccess) {
                rc = set_error(HSFM_EDEVICE, "hipGraphLaunch failed");
                break;
            }
        } else if ((rc = batch_record(idx, bb, sv, svp, d_q, bb.qnib.as<uint32_t>(), min_len, d_refb, n_refs,
                                      max_hits, main_kernel, rpt))) {
            break;
        }
        pend[npend++] = Pending{r0, nb, k};
    }
    for (int x = 0; x < npend && !rc; x++) rc = finish(pend[x]);
    for (int k = 1; k < nsets; k++) (void)hipStreamSynchronize(sst[k]);
    HSFM_HIP(hipEventRecord(e1, s));
    HSFM_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    HSFM_HIP(hipEventElapsedTime(&t, e0, e1));
    stats.total_ms = t;
    if (st) *st = stats;
    return cleanup(rc);
}

extern "C" int hsfm_mem_seed_mode(hsfm_index *idx, int mode, hsfm_lsi *lsi) {
    if (!idx) return set_error(HSFM_EINVAL, "hsfm_mem_seed_mode: null index");
    if (mode < HSFM_SEED_TABLES || mode > HSFM_SEED_LEARNED_TABLES)
        return set_error(HSFM_EINVAL, "hsfm_mem_seed_mode: unknown mode %d", mode);
    HSFM_LOCK_INDEX(idx, "hsfm_mem_seed_mode");
    if (mode == HSFM_SEED_LEARNED || mode == HSFM_SEED_LEARNED_TABLES) {
        LsiView lv;
        int rc = lsi_mem_view(lsi, idx, &lv);
        if (rc) return rc;
    }
    idx->mem_seed_mode = mode;
    idx->mem_lsi = (mode == HSFM_SEED_LEARNED || mode == HSFM_SEED_LEARNED_TABLES) ? lsi : nullptr;
    return HSFM_OK;
}

extern "C" int hsfm_last_deferred(hsfm_index *idx, uint32_t *ids, uint64_t cap, uint64_t *count) {
    if (!idx || !count || (cap && !ids)) return set_error(HSFM_EINVAL, "hsfm_last_deferred: null argument");
    HSFM_LOCK_INDEX(idx, "hsfm_last_deferred");
    if (!idx->defer_valid)
        return set_error(HSFM_EINVAL, "hsfm_last_deferred: the last call on this index was not a one-shot MEM call");
    HSFM_HIP(hipSetDevice(idx->dev->ordinal));
    hipStream_t s = idx->dev->stream;
    std::vector<uint32_t> w(idx->defer_n);
    if (idx->defer_n)
        HSFM_HIP(hipMemcpyAsync(w.data(), idx->ws.defer_ids.p, idx->defer_n * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if (idx->defer_items) { // item -> its read
        for (uint32_t &v : w) {
            uint4 it;
            HSFM_HIP(hipMemcpy(&it, idx->ws.items.as<uint4>() + v, sizeof(uint4), hipMemcpyDeviceToHost));
            v = it.x;
        }
    }
    std::sort(w.begin(), w.end());
    w.erase(std::unique(w.begin(), w.end()), w.end());
    *count = w.size();
    for (uint64_t i = 0; i < w.size() && i < cap; i++) ids[i] = w[i];
    return HSFM_OK;
}

extern "C" int hsfm_device_alloc(hsfm_device *dev, uint64_t bytes, void **out) {
    if (!dev || !out) return set_error(HSFM_EINVAL, "hsfm_device_alloc: null argument");
    HSFM_HIP(hipSetDevice(dev->ordinal));
    *out = nullptr;
    if (bytes && hipMalloc(out, bytes) != hipSuccess)
        return set_error(HSFM_ENOMEM, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
    return HSFM_OK;
}
extern "C" int hsfm_device_free(hsfm_device *dev, void *p) {
    if (!dev) return set_error(HSFM_EINVAL, "hsfm_device_free: null device");
    HSFM_HIP(hipSetDevice(dev->ordinal));
    if (p) HSFM_HIP(hipFree(p));
    return HSFM_OK;
}
extern "C" int hsfm_host_to_device(hsfm_device *dev, void *d_dst, const void *src, uint64_t bytes) {
    if (!dev || (bytes && (!d_dst || !src))) return set_error(HSFM_EINVAL, "hsfm_host_to_device: null argument");
    if (!bytes) return HSFM_OK;
    HSFM_HIP(hipSetDevice(dev->ordinal));
    HSFM_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, dev->stream));
    HSFM_HIP(hipStreamSynchronize(dev->stream));
    return HSFM_OK;
}
extern "C" int hsfm_device_to_host(hsfm_device *dev, void *dst, const void *d_src, uint64_t bytes) {
    if (!dev || (bytes && (!dst || !d_src))) return set_error(HSFM_EINVAL, "hsfm_device_to_host: null argument");
    if (!bytes) return HSFM_OK;
    HSFM_HIP(hipSetDevice(dev->ordinal));
    HSFM_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, dev->stream));
    HSFM_HIP(hipStreamSynchronize(dev->stream));
    return HSFM_OK;
}
