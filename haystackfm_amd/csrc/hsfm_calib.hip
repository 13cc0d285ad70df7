// hsfm_calib.hip -- measurement aid, not on the query path: the rate at which
// this GPU serves random dependent loads of 64-byte records, the access
// pattern of every backward-search step (one Occ record per probe point, its
// address known only after the previous step's load).  It bounds the query
// kernels from above more tightly than the HBM streaming peak does.
//
// Every lane runs `iters` loads in a chain: the next record index mixes a
// per-lane generator with the value just loaded, so a load cannot issue before
// the previous one returns.  The table is 64-byte records; indexes are masked
// to the table (or to a window of it), so every access stays in bounds.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// mode 0: a whole 64-byte record (four 16-byte loads); mode 1: 16 bytes of a
// record; mode 2: two records a few positions apart (the lo/hi ends of an
// interval), both whole; mode 3: a whole record, loaded by the four lanes of a
// quad together and exchanged by shuffles; mode 4: every sixth load is 16
// bytes anywhere in the whole table (`mask_all`), the others whole records in
// the window (the mix of a seed-table lookup and the probes that follow it).
template <int MODE>
__global__ __launch_bounds__(256) void k_calib_chase(const uint4 *__restrict__ tab, uint64_t mask, uint64_t mask_all,
                                                     uint32_t iters, uint32_t *__restrict__ sink) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x = (t + 1) * 0x9E3779B97F4A7C15ull;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < iters; i++) {
        const uint64_t r = ((x >> 20) ^ acc) & mask;
        const uint4 *p = tab + r * 4;
        uint32_t a = 0;
        if (MODE == 3) {
            // the four lanes of a quad fetch each other's records together: in
            // round i, lane l loads part (l & 3) of the record of quad lane i,
            // so one instruction touches 16 records, each line once
            const uint32_t ql = threadIdx.x & 3u, qb = threadIdx.x & ~3u;
            const uint64_t rr = r;
            uint4 mine[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint64_t ri = (uint64_t)__shfl((long long)rr, (int)(qb + i), 64) & mask;
                const uint4 v = tab[ri * 4 + ql];
                mine[i] = v; // part ql of quad lane i's record
            }
            // lane l gathers the four parts of its own record from its quad: in
            // round k it reads lane s = (l + k) & 3, which offers the part it
            // holds of record (s - k) & 3, i.e. part s of lane l's record
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t o = (ql - (uint32_t)k) & 3u;
                const uint4 off = o == 0 ? mine[0] : o == 1 ? mine[1] : o == 2 ? mine[2] : mine[3];
                const int src = (int)(qb + ((ql + (uint32_t)k) & 3u));
                uint4 v;
                v.x = __shfl(off.x, src, 64);
                v.y = __shfl(off.y, src, 64);
                v.z = __shfl(off.z, src, 64);
                v.w = __shfl(off.w, src, 64);
                a += v.x ^ v.w;
            }
        } else if (MODE == 1) {
            const uint4 v = p[0];
            a = v.x ^ v.w;
        } else if (MODE == 4 && i % 6 == 5) {
            const uint4 v = tab[(((x >> 20) ^ acc) & mask_all) * 4];
            a = v.x ^ v.w;
        } else {
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const uint4 v = p[w];
                a += v.x ^ v.w;
            }
        }
        if (MODE == 2) {
            const uint4 *p2 = tab + ((r + 1 + ((x >> 40) & 3)) & mask) * 4;
#pragma unroll
            for (int w = 0; w < 4; w++) a += p2[w].y;
        }
        acc += a;
        x = x * 6364136223846793005ull + 1442695040888963407ull;
    }
    if (acc == 0x7fffffffu) sink[0] = acc; // keeps the chain live
}

__global__ void k_calib_fill(uint4 *tab, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t h = (uint32_t)(i * 2654435761u);
        tab[i] = make_uint4(h, h >> 3, h >> 7, h >> 11);
    }
}

} // namespace

// Times `reps` launches of `lanes` chains of `iters` loads over a table of
// `table_bytes` (a power of two, >= 64 KiB) of which the chains touch the first
// `window_bytes` (a power of two, <= table_bytes).  ms_out: the fastest launch;
// records_out: records loaded per launch.  Returns 0, or a hipError_t value.
extern "C" int hsfm_calib_random_records(int ordinal, uint64_t table_bytes, uint64_t window_bytes, uint32_t lanes,
                                         uint32_t iters, int mode, int reps, double *ms_out, double *records_out) {
    if (table_bytes < 65536 || (table_bytes & (table_bytes - 1)) || window_bytes < 64 ||
        (window_bytes & (window_bytes - 1)) || window_bytes > table_bytes || lanes == 0 || lanes % 256 ||
        mode < 0 || mode > 4 || reps < 1 || !ms_out || !records_out)
        return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(ordinal);
    if (e != hipSuccess) return (int)e;
    uint4 *tab = nullptr;
    uint32_t *sink = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    float best = 1e30f;
    const uint64_t mask = window_bytes / 64 - 1, mask_all = table_bytes / 64 - 1;
    const dim3 grid(lanes / 256), block(256);
    if ((e = hipMalloc((void **)&tab, table_bytes)) != hipSuccess) return (int)e;
    if ((e = hipMalloc((void **)&sink, 64)) != hipSuccess) goto done;
    if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess) goto done;
    hipLaunchKernelGGL(k_calib_fill, dim3(4096), dim3(256), 0, 0, tab, table_bytes / 16);
    for (int rep = 0; rep <= reps && e == hipSuccess; rep++) { // launch 0 warms up
        e = hipEventRecord(a, 0);
        if (mode == 0) hipLaunchKernelGGL(k_calib_chase<0>, grid, block, 0, 0, tab, mask, mask_all, iters, sink);
        else if (mode == 1) hipLaunchKernelGGL(k_calib_chase<1>, grid, block, 0, 0, tab, mask, mask_all, iters, sink);
        else if (mode == 2) hipLaunchKernelGGL(k_calib_chase<2>, grid, block, 0, 0, tab, mask, mask_all, iters, sink);
        else if (mode == 3) hipLaunchKernelGGL(k_calib_chase<3>, grid, block, 0, 0, tab, mask, mask_all, iters, sink);
        else hipLaunchKernelGGL(k_calib_chase<4>, grid, block, 0, 0, tab, mask, mask_all, iters, sink);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(b, 0);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess && rep > 0) best = ms < best ? ms : best;
    }
    if (e == hipSuccess) {
        *ms_out = best;
        *records_out = (double)lanes * iters * (mode == 2 ? 2.0 : 1.0);
    }
done:
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (sink) (void)hipFree(sink);
    (void)hipFree(tab);
    return (int)e;
}
