// hsfm_construct.hip -- FmIndex construction on the device.
//
// The byte stream of FmIndex::build_cpu + to_bytes (src/fm_index/mod.rs:117-195,
// serialize.rs:13-26) and BidirFmIndex::build_cpu + to_bytes
// (bidir_index.rs:53-80, 183-192), computed on the GPU:
//   suffix array   prefix doubling: suffixes ranked by their first 12 codes
//                  (5 bits each, 0 past the end so a suffix that runs out sorts
//                  first, as a shorter slice does in build_suffix_array_naive,
//                  suffix_array/cpu.rs:20-25), then by (rank[i], rank[i + h])
//                  pairs for h = 12, 24, 48 ... until every rank is distinct;
//                  each round one radix sort of 64-bit keys
//   BWT            BWT[i] = T[SA[i] - 1], T[n - 1] for SA[i] = 0 (bwt/cpu.rs:14-25)
//   Occ            compact lanes of the symbols present, one wave per 64-position
//                  block (ballots give the lane bit vectors), lane counts scanned
//                  over the blocks, records [sb u32 x L][delta u16 x L]
//                  [planes or one-hot u64] (occ/cpu.rs:23-114, occ/mod.rs:108-171)
//   sampled SA     rows with SA % rate == 0 or a sequence start, one wave per
//                  word, 14-byte records [word][sb][delta] and the values in row
//                  order (suffix_array/mod.rs:71-118)
//   lookup table   the BFS of LookupTable::build level by level over dense
//                  arrays, ranks on the record layout (lookup.rs:37-91,
//                  occ/mod.rs:330-363)
// The reference builds its GPU index with compacted lanes missing
// (occ/gpu.rs:94-105); this one produces build_cpu's bytes exactly.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "hsfm_host.h"

namespace hsfm {
namespace {

constexpr uint32_t kPackCodes = 12; // codes in the first sort key (5 bits each)

__global__ void k_pack_first(const uint8_t *__restrict__ t, uint32_t n, unsigned long long *__restrict__ key,
                             uint32_t *__restrict__ val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long k = 0;
#pragma unroll
        for (uint32_t d = 0; d < kPackCodes; d++) {
            const uint64_t p = i + d;
            const unsigned long long v = p < n ? (unsigned long long)(t[p] & 15u) + 1ull : 0ull;
            k = (k << 5) | v;
        }
        key[i] = k;
        val[i] = (uint32_t)i;
    }
}

// flag[i] = 1 where sorted key i starts a new group
__global__ void k_group_flags(const unsigned long long *__restrict__ key, uint32_t n, uint32_t *__restrict__ flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// rank[sa[i]] = group number of sorted position i (1-based)
__global__ void k_scatter_rank(const uint32_t *__restrict__ sa, const uint32_t *__restrict__ grp, uint32_t n,
                               uint32_t *__restrict__ rank) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        rank[sa[i]] = grp[i];
}

__global__ void k_pair_keys(const uint32_t *__restrict__ rank, uint32_t n, uint32_t h,
                            unsigned long long *__restrict__ key, uint32_t *__restrict__ val) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = i + h;
        key[i] = ((unsigned long long)rank[i] << 32) | (j < n ? rank[j] : 0u);
        val[i] = (uint32_t)i;
    }
}

__global__ void k_bwt(const uint8_t *__restrict__ t, const uint32_t *__restrict__ sa, uint32_t n,
                      uint8_t *__restrict__ bwt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = sa[i];
        bwt[i] = t[s == 0 ? n - 1 : s - 1] & 15u;
    }
}

__global__ void k_histogram(const uint8_t *__restrict__ t, uint32_t n, unsigned long long *__restrict__ cnt) {
    __shared__ uint32_t s[16];
    if (threadIdx.x < 16) s[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&s[t[i] & 15u], 1u);
    __syncthreads();
    if (threadIdx.x < 16 && s[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], (unsigned long long)s[threadIdx.x]);
}

struct LaneMap {
    uint8_t s2l[16]; // symbol -> lane (0xff: absent)
    uint32_t L, W, enc, stride;
};

// Per 64-position block and lane: the count of the lane's symbol
// (cnt[lane * nb + block]).  One wave per block.
__global__ __launch_bounds__(256) void k_occ_counts(const uint8_t *__restrict__ bwt, uint32_t n, uint64_t nb,
                                                    LaneMap lm, uint32_t *__restrict__ cnt) {
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; b < nb; b += waves) {
        const uint64_t p = 64 * b + __lane_id();
        const uint32_t lane = p < n ? lm.s2l[bwt[p]] : 0xffu;
        for (uint32_t k = 0; k < lm.L; k++) {
            const uint64_t m = __ballot(lane == k);
            if (__lane_id() == 0) cnt[(uint64_t)k * nb + b] = (uint32_t)__popcll(m);
        }
    }
}

__device__ void occ_record(const uint8_t *__restrict__ bwt, uint32_t n, uint64_t nb, const LaneMap &lm,
                           const uint32_t *__restrict__ cum, uint8_t *__restrict__ rec, uint64_t b);

// Block records from the exclusive scans of the counts (cum[lane * nb + b]).
__global__ __launch_bounds__(256) void k_occ_records(const uint8_t *__restrict__ bwt, uint32_t n, uint64_t nb,
                                                     LaneMap lm, const uint32_t *__restrict__ cum,
                                                     uint8_t *__restrict__ rec) {
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; b < nb; b += waves)
        occ_record(bwt, n, nb, lm, cum, rec, b);
}

// The record of block b (all 64 lanes of the wave).
__device__ void occ_record(const uint8_t *__restrict__ bwt, uint32_t n, uint64_t nb, const LaneMap &lm,
                           const uint32_t *__restrict__ cum, uint8_t *__restrict__ rec, uint64_t b) {
    const uint64_t p = 64 * b + __lane_id();
    const uint32_t lane = p < n ? lm.s2l[bwt[p]] : 0xffu;
    uint64_t words[16];
    uint32_t sb[16], dl[16];
    const uint64_t sb0 = b & ~7ull; // first block of the superblock
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
        words[k] = 0;
        sb[k] = 0;
        dl[k] = 0;
    }
    for (uint32_t k = 0; k < lm.L; k++) {
        const uint64_t m = __ballot(lane == k);
        sb[k] = cum[(uint64_t)k * nb + sb0];
        dl[k] = cum[(uint64_t)k * nb + b] - sb[k];
        if (lm.enc == 0) {
            for (uint32_t q = 0; q < lm.W; q++)
                if ((k >> q) & 1u) words[q] |= m;
        } else {
            words[k] = m;
        }
    }
    uint8_t *r = rec + b * lm.stride;
    for (uint32_t t = __lane_id(); t < lm.stride; t += 64) {
        uint32_t v;
        if (t < 4 * lm.L) {
            uint32_t x = 0;
            for (uint32_t k = 0; k < 16; k++)
                if (k == t / 4) x = sb[k];
            v = (x >> (8 * (t % 4))) & 0xffu;
        } else if (t < 6 * lm.L) {
            const uint32_t u = t - 4 * lm.L;
            uint32_t x = 0;
            for (uint32_t k = 0; k < 16; k++)
                if (k == u / 2) x = dl[k];
            v = (x >> (8 * (u % 2))) & 0xffu;
        } else {
            const uint32_t u = t - 6 * lm.L;
            uint64_t x = 0;
            for (uint32_t k = 0; k < 16; k++)
                if (k == u / 8) x = words[k];
            v = (uint32_t)(x >> (8 * (u % 8))) & 0xffu;
        }
        r[t] = (uint8_t)v;
    }
}

__global__ void k_occ_checkpoints(const uint32_t *__restrict__ cum, uint64_t nb, uint32_t L, uint64_t nsb,
                                  uint32_t *__restrict__ sb) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsb * L;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = i / L, k = i % L;
        sb[i] = cum[k * nb + 8 * s];
    }
}

__device__ __forceinline__ bool is_start(uint32_t v, const uint32_t *__restrict__ starts, uint32_t ns) {
    uint32_t lo = 0, hi = ns;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < ns && starts[lo] == v;
}

// Sample words: bit j of word w set when row 64w + j is sampled; wcnt[w] = popcount.
__global__ __launch_bounds__(256) void k_sa_words(const uint32_t *__restrict__ sa, uint32_t n, uint64_t nw,
                                                  uint32_t rate, const uint32_t *__restrict__ starts, uint32_t ns,
                                                  unsigned long long *__restrict__ word, uint32_t *__restrict__ wcnt) {
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nw; w += waves) {
        const uint64_t i = 64 * w + __lane_id();
        bool mk = false;
        if (i < n) {
            const uint32_t v = sa[i];
            mk = v % rate == 0 || is_start(v, starts, ns);
        }
        const uint64_t m = __ballot(mk);
        if (__lane_id() == 0) {
            word[w] = m;
            wcnt[w] = (uint32_t)__popcll(m);
        }
    }
}

__device__ void sa_record(const uint32_t *__restrict__ sa, const unsigned long long *__restrict__ word,
                          const uint32_t *__restrict__ wcum, uint8_t *__restrict__ rec, uint32_t *__restrict__ vals,
                          uint64_t w);

// 14-byte word records and the sampled values in row order.
__global__ __launch_bounds__(256) void k_sa_records(const uint32_t *__restrict__ sa, uint32_t n, uint64_t nw,
                                                    const unsigned long long *__restrict__ word,
                                                    const uint32_t *__restrict__ wcum, uint8_t *__restrict__ rec,
                                                    uint32_t *__restrict__ vals) {
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nw; w += waves)
        sa_record(sa, word, wcum, rec, vals, w);
}

// The 14-byte record and sampled values of word w (all 64 lanes of the wave).
__device__ void sa_record(const uint32_t *__restrict__ sa, const unsigned long long *__restrict__ word,
                          const uint32_t *__restrict__ wcum, uint8_t *__restrict__ rec, uint32_t *__restrict__ vals,
                          uint64_t w) {
    const uint32_t lid = __lane_id();
    const uint64_t m = word[w];
    const uint32_t base = wcum[w & ~7ull], cur = wcum[w];
    if ((m >> lid) & 1ull) {
        const uint32_t below = (uint32_t)__popcll(m & ((1ull << lid) - 1ull));
        vals[cur + below] = sa[64 * w + lid];
    }
    if (lid < 14) {
        const uint32_t d = cur - base;
        uint32_t v;
        if (lid < 8) v = (uint32_t)(m >> (8 * lid)) & 0xffu;
        else if (lid < 12) v = (base >> (8 * (lid - 8))) & 0xffu;
        else v = (d >> (8 * (lid - 12))) & 0xffu;
        rec[w * 14 + lid] = (uint8_t)v;
    }
}

struct RefOcc {
    const uint8_t *rec;
    uint8_t s2l[16];
    uint32_t L, P, enc, stride;
};

__device__ __forceinline__ uint64_t rd_u64(const uint8_t *p) {
    uint64_t v = 0;
#pragma unroll
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}
__device__ __forceinline__ uint32_t rd_u32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// rank(c, i) on the reference record layout (occ/mod.rs:330-363)
__device__ uint32_t ref_rank(const RefOcc &o, uint32_t c, uint32_t i) {
    if (i == 0 || c >= 16 || o.s2l[c] == 0xffu) return 0;
    const uint32_t lane = o.s2l[c], pos = i - 1, off = pos & 63u;
    const uint8_t *r = o.rec + (uint64_t)(pos >> 6) * o.stride;
    uint64_t m = ~0ull;
    if (o.enc == 1) {
        m = rd_u64(r + 6 * o.L + 8 * lane);
    } else {
        for (uint32_t q = 0; q < o.P; q++) {
            const uint64_t pv = rd_u64(r + 6 * o.L + 8 * q);
            m &= ((lane >> q) & 1u) ? pv : ~pv;
        }
    }
    const uint64_t mask = off == 63 ? ~0ull : ((1ull << (off + 1)) - 1ull);
    const uint32_t delta = (uint32_t)r[4 * o.L + 2 * lane] | ((uint32_t)r[4 * o.L + 2 * lane + 1] << 8);
    return rd_u32(r + 4 * lane) + delta + (uint32_t)__popcll(m & mask);
}

// One BFS level of LookupTable::build over dense arrays: child idx = parent * 4 + digit
// (the newly consumed code is the least significant digit); a child exists when its
// parent's interval is non-empty and its symbol occurs in the text.
__global__ void k_lookup_level(RefOcc o, const uint32_t *__restrict__ c_arr, uint32_t n, const uint2 *__restrict__ prev,
                               uint64_t count, uint2 *__restrict__ out) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const uint2 pv = prev ? prev[idx >> 2] : make_uint2(0u, n);
    const uint32_t sym = (uint32_t)(idx & 3u) + 1u;
    const uint32_t up = sym + 1 < 16 ? c_arr[sym + 1] : n;
    uint2 r = make_uint2(0u, 0u);
    if (pv.x < pv.y && up > c_arr[sym]) r = make_uint2(c_arr[sym] + ref_rank(o, sym, pv.x), c_arr[sym] + ref_rank(o, sym, pv.y));
    out[idx] = r;
}

// ---- host -------------------------------------------------------------------
struct Writer {
    std::vector<uint8_t> b;
    void put(const void *p, size_t k) {
        const uint8_t *q = static_cast<const uint8_t *>(p);
        b.insert(b.end(), q, q + k);
    }
    void u8(uint8_t v) { put(&v, 1); }
    void u32(uint32_t v) { put(&v, 4); }
    void u64(uint64_t v) { put(&v, 8); }
};

uint32_t grid_for(uint64_t work, uint32_t block = 256) {
    const uint64_t g = (work + block - 1) / block;
    return (uint32_t)(g < 65536 ? (g ? g : 1) : 65536);
}

template <class T>
int scan_exclusive(hipStream_t s, const T *in, T *out, uint64_t n, DevBuf &tmp) {
    size_t bytes = 0;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
    int rc = tmp.ensure(bytes + 16);
    if (rc) return rc;
    HSFM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s));
    return HSFM_OK;
}

// Suffix array of d_text[0, n) by prefix doubling into sa (device).
int suffix_sort(hipStream_t s, const uint8_t *d_text, uint32_t n, DevBuf &sa) {
    DevBuf k0, k1, v1, rank, grp, tmp;
    int rc;
    if ((rc = k0.ensure((size_t)n * 8)) || (rc = k1.ensure((size_t)n * 8)) || (rc = sa.ensure((size_t)n * 4 + 16)) ||
        (rc = v1.ensure((size_t)n * 4)) || (rc = rank.ensure((size_t)n * 4)) || (rc = grp.ensure((size_t)n * 4 + 8)))
        return rc;
    hipcub::DoubleBuffer<unsigned long long> keys(k0.as<unsigned long long>(), k1.as<unsigned long long>());
    hipcub::DoubleBuffer<uint32_t> vals(sa.as<uint32_t>(), v1.as<uint32_t>());
    hipLaunchKernelGGL(k_pack_first, dim3(grid_for(n)), dim3(256), 0, s, d_text, n, keys.Current(), vals.Current());
    HSFM_HIP(hipGetLastError());
    int end_bit = 5 * (int)kPackCodes;
    for (uint64_t h = kPackCodes;; h *= 2) {
        size_t bytes = 0;
        HSFM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, vals, (int)n, 0, end_bit, s));
        if ((rc = tmp.ensure(bytes + 16))) return rc;
        HSFM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys, vals, (int)n, 0, end_bit, s));
        // group starts (into rank, free until the scatter) -> group numbers
        hipLaunchKernelGGL(k_group_flags, dim3(grid_for(n)), dim3(256), 0, s, keys.Current(), n, rank.as<uint32_t>());
        HSFM_HIP(hipGetLastError());
        size_t sb = 0;
        HSFM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, sb, rank.as<uint32_t>(), grp.as<uint32_t>(), (int)n, s));
        if ((rc = tmp.ensure(sb + 16))) return rc;
        HSFM_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, sb, rank.as<uint32_t>(), grp.as<uint32_t>(), (int)n, s));
        uint32_t groups = 0;
        HSFM_HIP(hipMemcpyAsync(&groups, grp.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
        if (groups == n || h >= n) break; // every suffix ranked apart: vals holds the suffix array
        hipLaunchKernelGGL(k_scatter_rank, dim3(grid_for(n)), dim3(256), 0, s, vals.Current(), grp.as<uint32_t>(), n,
                           rank.as<uint32_t>());
        hipLaunchKernelGGL(k_pair_keys, dim3(grid_for(n)), dim3(256), 0, s, rank.as<uint32_t>(), n, (uint32_t)h,
                           keys.Current(), vals.Current());
        HSFM_HIP(hipGetLastError());
        int bits = 0;
        while (bits < 32 && (1ull << bits) <= groups) bits++;
        end_bit = 32 + bits;
    }
    if (vals.Current() != sa.as<uint32_t>())
        HSFM_HIP(hipMemcpyAsync(sa.p, vals.Current(), (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    HSFM_HIP(hipStreamSynchronize(s));
    return HSFM_OK;
}

// One FmIndex as FmIndex::build_cpu builds it, serialized (to_bytes).
int build_one(hsfm_device *dev, const uint8_t *text, uint32_t n, const std::vector<uint32_t> &bounds,
              const std::vector<std::string> &headers, uint32_t rate, uint32_t depth, uint32_t enc, uint8_t tag,
              Writer &w) {
    hipStream_t s = dev->stream;
    int rc;
    DevBuf d_text, sa, bwt, cnt, cum, rec, sbk, words, wcnt, wcum, srec, svals, starts, hist, tmp;
    if ((rc = d_text.ensure((size_t)n + 16))) return rc;
    HSFM_HIP(hipMemcpyAsync(d_text.p, text, n, hipMemcpyHostToDevice, s));
    // C array (c_array.rs:14-31)
    if ((rc = hist.ensure(16 * 8))) return rc;
    HSFM_HIP(hipMemsetAsync(hist.p, 0, 16 * 8, s));
    hipLaunchKernelGGL(k_histogram, dim3(grid_for(n, 256) < 4096 ? grid_for(n, 256) : 4096), dim3(256), 0, s,
                       d_text.as<uint8_t>(), n, hist.as<unsigned long long>());
    HSFM_HIP(hipGetLastError());
    uint64_t h16[16];
    HSFM_HIP(hipMemcpyAsync(h16, hist.p, 16 * 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    uint32_t carr[16];
    uint64_t acc = 0;
    for (int c = 0; c < 16; c++) {
        carr[c] = (uint32_t)acc;
        acc += h16[c];
    }
    // suffix array and BWT
    if ((rc = suffix_sort(s, d_text.as<uint8_t>(), n, sa))) return rc;
    if ((rc = bwt.ensure((size_t)n + 64))) return rc;
    hipLaunchKernelGGL(k_bwt, dim3(grid_for(n)), dim3(256), 0, s, d_text.as<uint8_t>(), sa.as<uint32_t>(), n,
                       bwt.as<uint8_t>());
    HSFM_HIP(hipGetLastError());
    // sampled SA
    const uint64_t nw = ((uint64_t)n + 63) / 64;
    std::vector<uint32_t> st{0};
    for (size_t i = 0; i + 1 < bounds.size(); i++) st.push_back(bounds[i]);
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    if ((rc = starts.ensure(st.size() * 4)) || (rc = words.ensure(nw * 8)) || (rc = wcnt.ensure(nw * 4 + 4)) ||
        (rc = wcum.ensure(nw * 4 + 4)) || (rc = srec.ensure(nw * 14 + 16)))
        return rc;
    HSFM_HIP(hipMemcpyAsync(starts.p, st.data(), st.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sa_words, dim3(grid_for(nw * 64)), dim3(256), 0, s, sa.as<uint32_t>(), n, nw, rate,
                       starts.as<uint32_t>(), (uint32_t)st.size(), words.as<unsigned long long>(),
                       wcnt.as<uint32_t>());
    HSFM_HIP(hipGetLastError());
    if ((rc = scan_exclusive<uint32_t>(s, wcnt.as<uint32_t>(), wcum.as<uint32_t>(), nw + 1, tmp))) return rc;
    uint32_t nvals = 0;
    HSFM_HIP(hipMemcpyAsync(&nvals, wcum.as<uint32_t>() + nw, 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if ((rc = svals.ensure((size_t)nvals * 4 + 4))) return rc;
    hipLaunchKernelGGL(k_sa_records, dim3(grid_for(nw * 64)), dim3(256), 0, s, sa.as<uint32_t>(), n, nw,
                       words.as<unsigned long long>(), wcum.as<uint32_t>(), srec.as<uint8_t>(), svals.as<uint32_t>());
    HSFM_HIP(hipGetLastError());
    sa.release();
    words.release();
    // Occ: lanes of the symbols present (the BWT holds the text's symbols)
    LaneMap lm{};
    memset(lm.s2l, 0xff, 16);
    uint8_t l2s[16] = {0};
    uint32_t L = 0;
    for (int c = 0; c < 16; c++)
        if (h16[c]) {
            lm.s2l[c] = (uint8_t)L;
            l2s[L] = (uint8_t)c;
            L++;
        }
    uint32_t P = 0;
    if (L > 1)
        while ((1u << P) < L) P++;
    lm.L = L;
    lm.enc = enc;
    lm.W = enc == 0 ? P : L;
    lm.stride = 6 * L + 8 * lm.W;
    const uint64_t nb = ((uint64_t)n + 63) / 64, nsb = (nb + 7) / 8;
    if ((rc = cnt.ensure((size_t)L * nb * 4 + 4)) || (rc = cum.ensure((size_t)L * nb * 4 + 4)) ||
        (rc = rec.ensure(nb * lm.stride + 16)) || (rc = sbk.ensure(nsb * L * 4 + 4)))
        return rc;
    hipLaunchKernelGGL(k_occ_counts, dim3(grid_for(nb * 64)), dim3(256), 0, s, bwt.as<uint8_t>(), n, nb, lm,
                       cnt.as<uint32_t>());
    HSFM_HIP(hipGetLastError());
    for (uint32_t k = 0; k < L; k++)
        if ((rc = scan_exclusive<uint32_t>(s, cnt.as<uint32_t>() + (size_t)k * nb, cum.as<uint32_t>() + (size_t)k * nb,
                                           nb, tmp)))
            return rc;
    hipLaunchKernelGGL(k_occ_records, dim3(grid_for(nb * 64)), dim3(256), 0, s, bwt.as<uint8_t>(), n, nb, lm,
                       cum.as<uint32_t>(), rec.as<uint8_t>());
    hipLaunchKernelGGL(k_occ_checkpoints, dim3(grid_for(nsb * L)), dim3(256), 0, s, cum.as<uint32_t>(), nb, L, nsb,
                       sbk.as<uint32_t>());
    HSFM_HIP(hipGetLastError());
    cnt.release();
    cum.release();
    bwt.release();
    // lookup table
    std::vector<uint32_t> lk;
    if (depth) {
        DevBuf lv[2], d_c;
        if ((rc = d_c.ensure(64))) return rc;
        HSFM_HIP(hipMemcpyAsync(d_c.p, carr, 64, hipMemcpyHostToDevice, s));
        RefOcc o{};
        o.rec = rec.as<uint8_t>();
        memcpy(o.s2l, lm.s2l, 16);
        o.L = L;
        o.P = P;
        o.enc = enc;
        o.stride = lm.stride;
        uint64_t count = 1;
        const uint2 *prev = nullptr;
        for (uint32_t l = 1; l <= depth; l++) {
            count *= 4;
            DevBuf &cur = lv[l & 1];
            if ((rc = cur.ensure(count * 8))) return rc;
            hipLaunchKernelGGL(k_lookup_level, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s, o,
                               d_c.as<uint32_t>(), n, prev, count, cur.as<uint2>());
            HSFM_HIP(hipGetLastError());
            prev = cur.as<uint2>();
        }
        lk.resize(count * 2);
        HSFM_HIP(hipMemcpyAsync(lk.data(), prev, count * 8, hipMemcpyDeviceToHost, s));
        HSFM_HIP(hipStreamSynchronize(s));
    }
    // serialize in SerializableFmIndex field order (serialize.rs:55-67)
    std::vector<uint32_t> h_sbk(nsb * L);
    if (!h_sbk.empty()) HSFM_HIP(hipMemcpyAsync(h_sbk.data(), sbk.p, h_sbk.size() * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    w.put(carr, 64);
    w.u8((uint8_t)L);
    w.u8((uint8_t)P);
    w.put(lm.s2l, 16);
    w.put(l2s, 16);
    w.u64(h_sbk.size());
    w.put(h_sbk.data(), h_sbk.size() * 4);
    const uint64_t nrec = nb * lm.stride;
    w.u64(nrec);
    size_t at = w.b.size();
    w.b.resize(at + nrec);
    if (nrec) HSFM_HIP(hipMemcpyAsync(w.b.data() + at, rec.p, nrec, hipMemcpyDeviceToHost, s));
    w.u64(lm.stride);
    w.u32(enc);
    w.u32(n);
    w.u64(nw * 14);
    at = w.b.size();
    w.b.resize(at + nw * 14);
    HSFM_HIP(hipMemcpyAsync(w.b.data() + at, srec.p, nw * 14, hipMemcpyDeviceToHost, s));
    w.u64(nvals);
    at = w.b.size();
    w.b.resize(at + (size_t)nvals * 4);
    if (nvals) HSFM_HIP(hipMemcpyAsync(w.b.data() + at, svals.p, (size_t)nvals * 4, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    w.u32(rate);
    w.u32(n);
    w.u32(n);
    w.u32((uint32_t)bounds.size());
    w.u64(bounds.size());
    w.put(bounds.data(), bounds.size() * 4);
    w.u64(headers.size());
    for (const std::string &h : headers) {
        w.u64(h.size());
        w.put(h.data(), h.size());
    }
    if (depth) {
        w.u8(1);
        w.u32(depth);
        const uint8_t core[4] = {1, 2, 3, 4};
        w.u64(4);
        w.put(core, 4);
        w.u64(lk.size() / 2);
        w.put(lk.data(), lk.size() * 4);
    } else {
        w.u8(0);
    }
    w.u8(tag);
    return HSFM_OK;
}

} // namespace
} // namespace hsfm

using namespace hsfm;

extern "C" int hsfm_build_index_device(hsfm_device *dev, const uint8_t *text, uint64_t n, const uint32_t *bounds,
                                       uint32_t nseq, const char *headers, uint32_t sample_rate, uint32_t lookup_depth,
                                       uint32_t encoding, uint8_t tag, int bidir, uint8_t **out, uint64_t *out_len) {
    if (!dev || !text || !bounds || !out || !out_len) return set_error(HSFM_EINVAL, "hsfm_build_index_device: null argument");
    if (n == 0 || n >= 0x7fffffffull) return set_error(HSFM_EINVAL, "hsfm_build_index_device: text length must be in 1..2^31-2");
    if (nseq == 0 || bounds[nseq - 1] != n)
        return set_error(HSFM_EINVAL, "hsfm_build_index_device: the last boundary must equal the text length");
    for (uint32_t i = 0; i < nseq; i++)
        if (bounds[i] == 0 || text[bounds[i] - 1] != 0 || (i && bounds[i] <= bounds[i - 1]))
            return set_error(HSFM_EINVAL, "hsfm_build_index_device: sequence %u does not end in a sentinel", i);
    if (sample_rate == 0 || lookup_depth > 15 || encoding > 1 || tag > 1)
        return set_error(HSFM_EINVAL, "hsfm_build_index_device: bad sample rate, lookup depth, encoding or tag");
    for (uint64_t i = 0; i < n; i++)
        if (text[i] > 15) return set_error(HSFM_EINVAL, "hsfm_build_index_device: code %u at %llu", text[i], (unsigned long long)i);
    HSFM_HIP(hipSetDevice(dev->ordinal));
    std::vector<uint32_t> b(bounds, bounds + nseq);
    std::vector<std::string> hs;
    const char *hp = headers;
    for (uint32_t i = 0; i < nseq; i++) { // empty header -> "seq_{i}" (fm_index/mod.rs:126-137)
        if (hp && *hp) {
            hs.emplace_back(hp);
            hp += strlen(hp) + 1;
        } else {
            hs.push_back("seq_" + std::to_string(i));
            if (hp) hp++;
        }
    }
    Writer w;
    int rc;
    if (!bidir) {
        if ((rc = build_one(dev, text, (uint32_t)n, b, hs, sample_rate, lookup_depth, encoding, tag, w))) return rc;
    } else {
        Writer fw, rw;
        if ((rc = build_one(dev, text, (uint32_t)n, b, hs, sample_rate, lookup_depth, encoding, tag, fw))) return rc;
        // reverse index: the text without its last sentinel, reversed, as one
        // unnamed sequence; no lookup table (bidir_index.rs:66-77, 523-531)
        std::vector<uint8_t> rt(n);
        for (uint64_t i = 0; i + 1 < n; i++) rt[i] = text[n - 2 - i];
        rt[n - 1] = 0;
        if ((rc = build_one(dev, rt.data(), (uint32_t)n, {(uint32_t)n}, {"seq_0"}, sample_rate, 0, encoding, tag, rw)))
            return rc;
        w.u32((uint32_t)fw.b.size());
        w.put(fw.b.data(), fw.b.size());
        w.put(rw.b.data(), rw.b.size());
    }
    uint8_t *p = static_cast<uint8_t *>(malloc(w.b.size() ? w.b.size() : 1));
    if (!p) return set_error(HSFM_ENOMEM, "hsfm_build_index_device: host allocation of %zu bytes", w.b.size());
    memcpy(p, w.b.data(), w.b.size());
    *out = p;
    *out_len = w.b.size();
    return HSFM_OK;
}
