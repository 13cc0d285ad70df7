// hsfm_runtime.hip -- device context, errors, buffers, scans and the shared
// LF-walk resolve kernel of libhsfm.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>

#include "hsfm_host.h"

namespace hsfm {

static thread_local std::string g_err;

thread_local const hsfm_index *tl_callback_index = nullptr;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return HSFM_OK;
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    want = (want + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError(); // a later launch check must not report this failure again
        return set_error(HSFM_ENOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    cap = want;
    return HSFM_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

__global__ void k_fill_u32(uint32_t *d, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

int fill_u32(hsfm_index *idx, uint32_t *d, uint64_t n, uint32_t v) {
    if (n == 0) return HSFM_OK;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, idx->dev->stream, d, n, v);
    HSFM_HIP(hipGetLastError());
    return HSFM_OK;
}

int event_record(hsfm_index *idx, int slot, bool end) {
    hipEvent_t &e = idx->ws.ev[2 * slot + (end ? 1 : 0)];
    if (!e) HSFM_HIP(hipEventCreate(&e));
    HSFM_HIP(hipEventRecord(e, idx->dev->stream));
    return HSFM_OK;
}

double event_ms(hsfm_index *idx, int slot) {
    hipEvent_t a = idx->ws.ev[2 * slot], b = idx->ws.ev[2 * slot + 1];
    if (!a || !b) return 0.0;
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

struct CastU64 {
    __host__ __device__ uint64_t operator()(uint32_t v) const { return (uint64_t)v; }
};

__global__ void k_zero_first(uint64_t *d) { d[0] = 0; }

int exclusive_scan_u32_u64(hsfm_index *idx, const uint32_t *d_in, uint64_t *d_out, uint32_t n) {
    hipStream_t s = idx->dev->stream;
    hipLaunchKernelGGL(k_zero_first, dim3(1), dim3(1), 0, s, d_out);
    HSFM_HIP(hipGetLastError());
    if (n == 0) return HSFM_OK;
    hipcub::TransformInputIterator<uint64_t, CastU64, const uint32_t *> it(d_in, CastU64());
    size_t tmp = 0;
    HSFM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, it, d_out + 1, (int)n, s));
    int rc = idx->ws.scan_tmp.ensure(tmp);
    if (rc) return rc;
    HSFM_HIP(hipcub::DeviceScan::InclusiveSum(idx->ws.scan_tmp.p, tmp, it, d_out + 1, (int)n, s));
    return HSFM_OK;
}

int readback_u64(hsfm_index *idx, const uint64_t *d_src, uint64_t *h_dst, int count) {
    hipStream_t s = idx->dev->stream;
    if (!idx->ws.h_pinned) HSFM_HIP(hipHostMalloc((void **)&idx->ws.h_pinned, kPinnedBytes, hipHostMallocDefault));
    HSFM_HIP(hipMemcpyAsync(idx->ws.h_pinned, d_src, (size_t)count * 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    memcpy(h_dst, idx->ws.h_pinned, (size_t)count * 8);
    return HSFM_OK;
}

// ---------------------------------------------------------------------------
// Resolve: one thread per SA row.  LF-walk to the nearest sampled row
// (src/fm_index/query.rs:282-291, same walk as locate_resolve.wgsl:138-175 and
// mem_resolve.wgsl:123-135), then the boundary partition point of
// FmIndex::map_position (query.rs:342-357) / ref_map.wgsl:31-59.  The walk is
// bounded by sample_rate steps (the reference GPU guard); a valid index never
// reaches it, a counter records it if it does.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resolve(DevDir d, const DevConst *__restrict__ cst,
                                                 const uint32_t *__restrict__ rows, uint64_t nrows,
                                                 const uint32_t *__restrict__ bounds, uint32_t nb,
                                                 uint32_t *__restrict__ out,
                                                 unsigned long long *__restrict__ guard) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const uint32_t nm1 = d.n - 1;
    uint32_t row = rows[t];
    row = row < nm1 ? row : nm1;
    uint32_t steps = 0, pos = 0xffffffffu;
    if (d.sa_full) pos = d.sa_full[row];
    else for (;;) {
        const uint4 w = d.samp[row >> 6];
        RecView r;
        load_rec(d, row, r);
        const uint64_t bits = (uint64_t)w.x | ((uint64_t)w.y << 32);
        const uint32_t off = row & 63;
        if ((bits >> off) & 1ull) {
            uint32_t rk = w.z + popc64(bits & ((1ull << off) - 1));
            rk = rk < d.num_samples ? rk : d.num_samples - 1;
            pos = d.sa_vals[rk] + steps;
            break;
        }
        if (steps >= d.sample_rate) {
            atomicAdd(guard, 1ull);
            break;
        }
        uint32_t lane, rank;
        lf_lane_rank(r, row, lane, rank);
        const uint32_t nx = s_clane[lane] + rank;
        row = nx < nm1 ? nx : nm1;
        steps++;
    }
    uint32_t id = 0xffffffffu, o = 0xffffffffu;
    if (pos != 0xffffffffu) {
        uint32_t lo = 0, hi = nb;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (bounds[mid] <= pos) lo = mid + 1;
            else hi = mid;
        }
        id = lo;
        o = pos - (lo ? bounds[lo - 1] : 0u);
    }
    out[2 * t] = id;
    out[2 * t + 1] = o;
}

// Resolve with a tunable fan-out: each thread walks RPT rows in lockstep
// (independent chains, more loads in flight per thread), grid-stride over
// rows[0, min(cap, *d_nrows)), so a fixed launch serves any row count.
template <int RPT>
__global__ __launch_bounds__(256) void k_resolve_fan(DevDir d, const DevConst *__restrict__ cst,
                                                     const uint32_t *__restrict__ rows, uint64_t cap,
                                                     const uint64_t *__restrict__ d_nrows,
                                                     const uint32_t *__restrict__ bounds, uint32_t nb,
                                                     uint32_t *__restrict__ out, unsigned long long *__restrict__ guard) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint64_t nrows = d_nrows ? (*d_nrows < cap ? *d_nrows : cap) : cap;
    const uint32_t nm1 = d.n - 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * RPT;
    for (uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * RPT; base < nrows; base += stride) {
        uint32_t row[RPT], steps[RPT], pos[RPT];
        bool live[RPT];
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            live[j] = base + j < nrows;
            const uint32_t r0 = live[j] ? rows[base + j] : 0u;
            row[j] = r0 < nm1 ? r0 : nm1;
            steps[j] = 0;
            pos[j] = 0xffffffffu;
        }
        bool any = true;
        if (d.sa_full) {
#pragma unroll
            for (int j = 0; j < RPT; j++) pos[j] = live[j] ? d.sa_full[row[j]] : 0xffffffffu;
            any = false;
        }
        while (any) {
            any = false;
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                if (!live[j]) continue;
                const uint4 w = d.samp[row[j] >> 6];
                RecView r;
                load_rec(d, row[j], r);
                const uint64_t bits = (uint64_t)w.x | ((uint64_t)w.y << 32);
                const uint32_t off = row[j] & 63;
                if ((bits >> off) & 1ull) {
                    uint32_t rk = w.z + popc64(bits & ((1ull << off) - 1));
                    rk = rk < d.num_samples ? rk : d.num_samples - 1;
                    pos[j] = d.sa_vals[rk] + steps[j];
                    live[j] = false;
                    continue;
                }
                if (steps[j] >= d.sample_rate) {
                    atomicAdd(guard, 1ull);
                    live[j] = false;
                    continue;
                }
                uint32_t lane, rank;
                lf_lane_rank(r, row[j], lane, rank);
                const uint32_t nx = s_clane[lane] + rank;
                row[j] = nx < nm1 ? nx : nm1;
                steps[j]++;
                any = true;
            }
        }
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            if (base + j >= nrows) continue;
            uint32_t id = 0xffffffffu, o = 0xffffffffu;
            if (pos[j] != 0xffffffffu) {
                uint32_t lo = 0, hi = nb;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (bounds[mid] <= pos[j]) lo = mid + 1;
                    else hi = mid;
                }
                id = lo;
                o = pos[j] - (lo ? bounds[lo - 1] : 0u);
            }
            out[2 * (base + j)] = id;
            out[2 * (base + j) + 1] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// Full suffix array of the forward direction.  Every row's resolve value
// (src/fm_index/query.rs:282-291) is the sample reached by its LF walk plus the
// steps taken, so walking LF backward from each sampled row writes the values
// of all rows between it and the previous sample of the same sequence (the
// walk stops before stepping over a sentinel: sequence starts are sampled).
// The rows after the last sample of a sequence are those on the walk from
// the sequence's sentinel row, which k_sa_full_sent resolves the same way and
// then fills.  A walk longer than the sample rate marks the table unusable.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sa_full(DevDir d, const DevConst *__restrict__ cst, uint32_t sent_lane,
                                                 uint32_t nblk, uint32_t *__restrict__ sa,
                                                 unsigned long long *__restrict__ bad) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint4 w = d.samp[b];
    uint64_t bits = (uint64_t)w.x | ((uint64_t)w.y << 32);
    uint32_t rk = w.z;
    const uint32_t n = d.n;
    while (bits) {
        const uint32_t bit = (uint32_t)__ffsll((unsigned long long)bits) - 1u;
        bits &= bits - 1;
        const uint32_t rs = b * 64 + bit;
        const uint32_t rkc = rk < d.num_samples ? rk : d.num_samples - 1;
        rk++;
        if (rs >= n) break;
        uint32_t pos = d.sa_vals[rkc], cur = rs;
        sa[rs] = pos;
        for (uint32_t steps = 0;; steps++) {
            RecView r;
            load_rec(d, cur, r);
            uint32_t lane, rank;
            lf_lane_rank(r, cur, lane, rank);
            if (lane == sent_lane) break; // the previous position is a sentinel
            if (steps >= d.sample_rate || pos == 0) {
                atomicAdd(bad, 1ull);
                break;
            }
            cur = s_clane[lane] + rank;
            pos--;
            uint32_t v;
            if (cur >= n || sa_sampled(d, cur, v)) break;
            sa[cur] = pos;
        }
    }
}

__global__ __launch_bounds__(256) void k_sa_full_sent(DevDir d, const DevConst *__restrict__ cst, uint32_t r0,
                                                      uint32_t nrows, uint32_t *__restrict__ sa,
                                                      unsigned long long *__restrict__ bad) {
    __shared__ uint32_t s_clane[16];
    if (threadIdx.x < 16) s_clane[threadIdx.x] = cst->c_lane[threadIdx.x];
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const uint32_t x = r0 + t, n = d.n;
    uint32_t v;
    if (x >= n || sa_sampled(d, x, v)) return;
    uint32_t cur = x, steps = 0;
    for (;;) { // resolve x
        if (sa_sampled(d, cur, v)) break;
        if (steps >= d.sample_rate) {
            atomicAdd(bad, 1ull);
            return;
        }
        RecView r;
        load_rec(d, cur, r);
        uint32_t lane, rank;
        lf_lane_rank(r, cur, lane, rank);
        cur = s_clane[lane] + rank;
        if (cur >= n) {
            atomicAdd(bad, 1ull);
            return;
        }
        steps++;
    }
    const uint32_t px = v + steps;
    cur = x;
    for (uint32_t i = 0; i < steps; i++) { // the rows on the way
        sa[cur] = px - i;
        RecView r;
        load_rec(d, cur, r);
        uint32_t lane, rank;
        lf_lane_rank(r, cur, lane, rank);
        cur = s_clane[lane] + rank;
    }
}

// Text lanes from the full suffix array: the first symbol of row r is the lane
// whose C range holds r (lane counts from the record at n), and row r starts
// at text position sa[r].
__global__ __launch_bounds__(256) void k_text_from_sa(DevDir d, const DevConst *__restrict__ cst,
                                                      const uint32_t *__restrict__ sa, uint8_t *__restrict__ text,
                                                      uint32_t *__restrict__ isa) {
    __shared__ uint32_t s_lo[16], s_hi[16];
    if (threadIdx.x == 0) {
        RecView rv;
        load_rec(d, d.n, rv);
        uint32_t cnt[16];
        lane_counts(rv, d.n, 16, cnt);
        for (int l = 0; l < 16; l++) {
            s_lo[l] = cst->c_lane[l];
            s_hi[l] = (uint32_t)l < cst->num_lanes ? cst->c_lane[l] + cnt[l] : 0u;
        }
    }
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.n) return;
    uint32_t lane = 0;
#pragma unroll
    for (int l = 0; l < 16; l++)
        if (r >= s_lo[l] && r < s_hi[l]) lane = (uint32_t)l;
    const uint32_t p = sa[r];
    if (p < d.n) {
        text[p] = (uint8_t)lane;
        if (isa) isa[p] = r;
    }
}

// Base-set nibbles of the text (DevDir::text4): word i from the lanes of
// positions [8i, 8i + 8); positions past n give 0.
__global__ __launch_bounds__(256) void k_text_nib(const uint8_t *__restrict__ text, uint32_t n,
                                                  const DevConst *__restrict__ cst, uint32_t *__restrict__ text4,
                                                  uint32_t nwords) {
    __shared__ uint8_t s_nib[16];
    if (threadIdx.x < 16) s_nib[threadIdx.x] = cst->text_nib[threadIdx.x];
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x) {
        uint32_t w = 0;
        for (uint32_t u = 0; u < 8; u++) {
            const uint64_t pos = 8ull * i + u;
            if (pos < n) w |= (uint32_t)s_nib[text[pos] & 15u] << (4 * u);
        }
        text4[i] = w;
    }
}

__global__ void k_count_value(const uint32_t *__restrict__ a, uint64_t n, uint32_t v,
                              unsigned long long *__restrict__ out) {
    uint32_t c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        c += a[i] == v ? 1u : 0u;
    if (c) atomicAdd(out, (unsigned long long)c);
}

// DevDir::qnib16 from the alphabet's query-byte base sets (with text4 only)
static void set_qnib16(const DevConst &hc, DirBuffers &db) {
    db.view.qnib16 = 0;
    if (!db.view.text4) return;
    unsigned long long t16 = 0;
    for (int b = 0; b < 256; b++) {
        const uint32_t qn = (b >= 1 && b <= 4) ? 1u << (b - 1) : (uint32_t)(hc.seed_expand[b] & 15u);
        if (b < 16) t16 |= (unsigned long long)qn << (4 * b);
        else if (qn) return;
    }
    db.view.qnib16 = t16;
}

int ensure_sa_full(hsfm_index *idx) {
    DirBuffers &db = idx->fwd;
    const char *env = getenv("HSFM_SA_FULL"); // "0": keep the LF walks (for comparisons)
    if (env && env[0] == '0') {
        db.view.sa_full = nullptr;
        db.view.text = nullptr;
        db.view.isa = nullptr;
        db.view.text4 = nullptr;
        db.view.qnib16 = 0;
        return HSFM_OK;
    }
    if (db.sa_full_tried) {
        db.view.sa_full = db.sa_full.as<uint32_t>();
        db.view.text = db.text.as<uint8_t>();
        db.view.isa = db.isa.as<uint32_t>();
        db.view.text4 = db.text4.as<uint32_t>();
        set_qnib16(idx->h_const, db);
        return HSFM_OK;
    }
    db.sa_full_tried = true;
    const uint32_t n = db.view.n;
    if (n == 0 || db.view.sample_rate == 0 || db.view.num_samples == 0) return HSFM_OK;
    // the sentinel's lane and its rows [C[$], C[$] + count)
    const DevConst &hc = idx->h_const;
    uint32_t sl = 16;
    for (uint32_t l = 0; l < hc.num_lanes && l < 16; l++)
        if (hc.lane_sym[l] == 0) sl = l;
    uint32_t s0 = 0, s1 = 0;
    if (sl < 16) {
        s0 = hc.c_lane[sl];
        s1 = n;
        for (uint32_t l = 0; l < hc.num_lanes && l < 16; l++)
            if (hc.c_lane[l] > s0 && hc.c_lane[l] < s1) s1 = hc.c_lane[l];
    }
    hipStream_t s = idx->dev->stream;
    DevBuf flag;
    if (db.sa_full.ensure((size_t)n * 4 + 16) || flag.ensure(16)) { // no memory: keep the walks
        db.sa_full.release();
        flag.release();
        return HSFM_OK;
    }
    int rc;
    if ((rc = fill_u32(idx, db.sa_full.as<uint32_t>(), n, 0xffffffffu)) ||
        (rc = fill_u32(idx, flag.as<uint32_t>(), 4, 0)))
        return rc;
    const uint32_t nblk = (n + 63) / 64;
    unsigned long long *d_bad = flag.as<unsigned long long>();
    hipLaunchKernelGGL(k_sa_full, dim3((nblk + 255) / 256), dim3(256), 0, s, db.view, idx->d_const.as<DevConst>(), sl,
                       nblk, db.sa_full.as<uint32_t>(), d_bad);
    HSFM_HIP(hipGetLastError());
    if (s1 > s0) {
        hipLaunchKernelGGL(k_sa_full_sent, dim3((s1 - s0 + 255) / 256), dim3(256), 0, s, db.view,
                           idx->d_const.as<DevConst>(), s0, s1 - s0, db.sa_full.as<uint32_t>(), d_bad);
        HSFM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_count_value, dim3(1024), dim3(256), 0, s, db.sa_full.as<uint32_t>(), (uint64_t)n,
                       0xffffffffu, d_bad + 1);
    HSFM_HIP(hipGetLastError());
    uint64_t h[2] = {0, 0};
    if ((rc = readback_u64(idx, reinterpret_cast<const uint64_t *>(d_bad), h, 2))) return rc;
    flag.release();
    if (h[0] || h[1]) { // a walk ran past the sample rate or a row was not reached
        db.sa_full.release();
        return HSFM_OK;
    }
    db.view.sa_full = db.sa_full.as<uint32_t>();
    // the text as lanes (right extensions of single rows compare it directly)
    // and the inverse (the row of a text position: locate finishes single
    // rows on the text and needs the row of the occurrence it ends on)
    if (!db.text.ensure((size_t)n + 16)) {
        if (db.isa.ensure((size_t)n * 4 + 16)) db.isa.release(); // without memory for it: text only
        if ((rc = fill_u32(idx, db.text.as<uint32_t>(), ((size_t)n + 16) / 4, 0))) return rc;
        hipLaunchKernelGGL(k_text_from_sa, dim3((n + 255) / 256), dim3(256), 0, s, db.view,
                           idx->d_const.as<DevConst>(), db.sa_full.as<uint32_t>(), db.text.as<uint8_t>(),
                           db.isa.as<uint32_t>());
        HSFM_HIP(hipGetLastError());
        // and as base-set nibbles (16 codes per 8-byte load) where the alphabet allows
        const uint32_t nwords = (uint32_t)(((uint64_t)n + 7) / 8 + 4);
        if (hc.nib_text && !db.text4.ensure((size_t)nwords * 4)) {
            const uint32_t blocks = (nwords + 255) / 256 < 65536 ? (nwords + 255) / 256 : 65536;
            hipLaunchKernelGGL(k_text_nib, dim3(blocks), dim3(256), 0, s, db.text.as<uint8_t>(), n,
                               idx->d_const.as<DevConst>(), db.text4.as<uint32_t>(), nwords);
            HSFM_HIP(hipGetLastError());
        }
        HSFM_HIP(hipStreamSynchronize(s));
        db.view.text = db.text.as<uint8_t>();
        db.view.isa = db.isa.as<uint32_t>();
        db.view.text4 = db.text4.as<uint32_t>();
        set_qnib16(hc, db);
    } else {
        db.text.release();
    }
    return HSFM_OK;
}

int launch_resolve_fan(hsfm_index *idx, const DevDir &d, const uint32_t *d_rows, uint64_t cap,
                       const uint64_t *d_nrows, const uint32_t *d_bounds, uint32_t nbounds, uint32_t *d_out_pairs,
                       unsigned long long *d_guard, int rpt, uint32_t grid) {
    hipStream_t s = idx->dev->stream;
    const DevConst *c = idx->d_const.as<DevConst>();
    if (rpt == 4)
        hipLaunchKernelGGL(k_resolve_fan<4>, dim3(grid), dim3(256), 0, s, d, c, d_rows, cap, d_nrows, d_bounds,
                           nbounds, d_out_pairs, d_guard);
    else if (rpt == 2)
        hipLaunchKernelGGL(k_resolve_fan<2>, dim3(grid), dim3(256), 0, s, d, c, d_rows, cap, d_nrows, d_bounds,
                           nbounds, d_out_pairs, d_guard);
    else
        hipLaunchKernelGGL(k_resolve_fan<1>, dim3(grid), dim3(256), 0, s, d, c, d_rows, cap, d_nrows, d_bounds,
                           nbounds, d_out_pairs, d_guard);
    HSFM_HIP(hipGetLastError());
    return HSFM_OK;
}

int launch_resolve(hsfm_index *idx, const DevDir &d, const uint32_t *d_rows, uint64_t nrows,
                   const uint32_t *d_bounds, uint32_t nbounds, uint32_t *d_out_pairs,
                   unsigned long long *d_guard) {
    idx->stats.sa_full_rows = d.sa_full ? d.n : 0;
    if (nrows == 0) return HSFM_OK;
    const uint64_t blocks = (nrows + 255) / 256;
    int rc = event_record(idx, 1, false);
    if (rc) return rc;
    hipLaunchKernelGGL(k_resolve, dim3((uint32_t)blocks), dim3(256), 0, idx->dev->stream, d,
                       idx->d_const.as<DevConst>(), d_rows, nrows, d_bounds, nbounds, d_out_pairs,
                       d_guard);
    HSFM_HIP(hipGetLastError());
    return event_record(idx, 1, true);
}

} // namespace hsfm

using namespace hsfm;

extern "C" const char *hsfm_last_error(void) { return g_err.c_str(); }

extern "C" int hsfm_device_open(int ordinal, hsfm_device **out) {
    if (!out) return set_error(HSFM_EINVAL, "hsfm_device_open: null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(HSFM_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (ordinal < 0 || ordinal >= n)
        return set_error(HSFM_EINVAL, "device ordinal %d out of range (%d devices)", ordinal, n);
    HSFM_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    HSFM_HIP(hipGetDeviceProperties(&prop, ordinal));
    hsfm_device *d = new hsfm_device();
    d->ordinal = ordinal;
    d->num_cus = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return set_error(HSFM_EDEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = d;
    return HSFM_OK;
}

extern "C" int hsfm_device_close(hsfm_device *dev) {
    if (!dev) return HSFM_OK;
    (void)hipSetDevice(dev->ordinal);
    if (dev->stream) (void)hipStreamDestroy(dev->stream);
    delete dev;
    return HSFM_OK;
}

extern "C" void hsfm_free_result(void *r) {
    // Results are single malloc blocks (header + payload).
    free(r);
}

extern "C" int hsfm_last_stats(hsfm_index *idx, hsfm_stats *out) {
    if (!idx || !out) return set_error(HSFM_EINVAL, "hsfm_last_stats: null argument");
    *out = idx->stats;
    return HSFM_OK;
}

// ---------------------------------------------------------------------------
// Reverse complement of a query batch (both-strand searching).  One thread per
// output code: its query by binary search over the offsets, then the mirrored
// input code through the IUPAC complement (the code whose base set is the
// complemented base set: A<->T, C<->G, R<->Y, K<->M, B<->V, D<->H; N, S, W and
// $ map to themselves; bytes >= 16 pass unchanged).
// ---------------------------------------------------------------------------
namespace {
__constant__ uint8_t c_comp[16] = {0, 4, 3, 2, 1, 5, 7, 6, 8, 9, 11, 10, 15, 14, 13, 12};

__global__ __launch_bounds__(256) void k_revcomp(const uint8_t *__restrict__ q, const uint64_t *__restrict__ qoff,
                                                 uint32_t nq, uint64_t total, uint8_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t lo = 0, hi = nq; // the query r with qoff[r] <= i < qoff[r + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (qoff[mid] <= i) lo = mid;
        else hi = mid;
    }
    const uint64_t b = qoff[lo], e = qoff[lo + 1];
    const uint8_t c = q[b + (e - 1 - i)];
    out[i] = c < 16 ? c_comp[c] : c;
}
} // namespace

extern "C" int hsfm_revcomp_device(hsfm_device *dev, const uint8_t *d_qflat, const uint64_t *d_qoff, uint32_t nq,
                                   uint8_t *d_out) {
    if (!dev || (nq && (!d_qflat || !d_qoff || !d_out)))
        return set_error(HSFM_EINVAL, "hsfm_revcomp_device: null argument");
    if (nq == 0) return HSFM_OK;
    HSFM_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = dev->stream;
    uint64_t ends[2] = {0, 0};
    HSFM_HIP(hipMemcpyAsync(ends, d_qoff, 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipMemcpyAsync(ends + 1, d_qoff + nq, 8, hipMemcpyDeviceToHost, s));
    HSFM_HIP(hipStreamSynchronize(s));
    if (ends[0] != 0 || ends[1] < ends[0]) return set_error(HSFM_EINVAL, "hsfm_revcomp_device: bad offsets");
    const uint64_t total = ends[1];
    if (total) {
        hipLaunchKernelGGL(k_revcomp, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, d_qflat, d_qoff, nq,
                           total, d_out);
        HSFM_HIP(hipGetLastError());
    }
    HSFM_HIP(hipStreamSynchronize(s));
    return HSFM_OK;
}

extern "C" int hsfm_revcomp_batch(hsfm_device *dev, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                                  uint8_t *out) {
    if (!dev || !qoff || (nq && (!qflat || !out))) return set_error(HSFM_EINVAL, "hsfm_revcomp_batch: null argument");
    if (nq == 0) return HSFM_OK;
    if (qoff[0] != 0) return set_error(HSFM_EINVAL, "hsfm_revcomp_batch: offsets must start at 0");
    for (uint32_t r = 0; r < nq; r++)
        if (qoff[r + 1] < qoff[r]) return set_error(HSFM_EINVAL, "hsfm_revcomp_batch: offsets must not decrease");
    const uint64_t total = qoff[nq];
    HSFM_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = dev->stream;
    uint8_t *dq = nullptr, *dout = nullptr;
    uint64_t *doff = nullptr;
    int rc = HSFM_OK;
    if (hipMalloc((void **)&dq, total + 1) != hipSuccess || hipMalloc((void **)&dout, total + 1) != hipSuccess ||
        hipMalloc((void **)&doff, ((size_t)nq + 1) * 8) != hipSuccess) {
        rc = set_error(HSFM_ENOMEM, "hsfm_revcomp_batch: device allocation failed");
    } else if (hipMemcpyAsync(dq, qflat, total, hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(doff, qoff, ((size_t)nq + 1) * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = set_error(HSFM_EDEVICE, "hsfm_revcomp_batch: copy failed");
    } else if ((rc = hsfm_revcomp_device(dev, dq, doff, nq, dout)) == HSFM_OK) {
        if (hipMemcpyAsync(out, dout, total, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = set_error(HSFM_EDEVICE, "hsfm_revcomp_batch: copy back failed");
    }
    (void)hipStreamSynchronize(s);
    if (dq) (void)hipFree(dq);
    if (dout) (void)hipFree(dout);
    if (doff) (void)hipFree(doff);
    return rc;
}

