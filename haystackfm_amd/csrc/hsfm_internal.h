// hsfm_internal.h -- device layout and rank primitives of the gfx950 query path.
//
// Device Occ layout (one per index direction), built once at upload from the
// reference's interleaved OccTable records (src/occ/mod.rs:114-183):
//
//   record b (64 bytes, 16-byte aligned) covers BWT positions [64b, 64b+64):
//     bytes  0..31  four u64 bitplanes; bit j of plane p = bit p of the lane
//                   index of position 64b+j (unused planes are zero)
//     bytes 32..63  sixteen u16 "below" counts: slot k (1..15) = number of
//                   positions in [superblock start, 64b) whose lane is < k;
//                   slot 0 is always 0
//   superblock s (65536 positions = 1024 records): sixteen u32, slot k = number
//                   of positions in [0, 65536 s) whose lane is < k (slot 0 = 0)
//
//   There are (n >> 6) + 1 records, so every count point x in [0, n] reads the
//   single record x >> 6:  #lane-k in [0, x) = sb[k+1]-sb[k] + cum[k+1]-cum[k]
//   + popcount(eq_k & ((1 << (x & 63)) - 1)), with slot 16 implied
//   (sb: 65536 s, cum: 64 (b mod 1024)).  One 64-byte HBM sector per rank probe;
//   the superblock table (64 B per 65536 positions) stays cache resident.
//
// Sampled-SA markers: one uint4 per 64 BWT rows {bits lo, bits hi, rank base, 0}
// (from src/suffix_array/mod.rs:94-110); sa_vals copied unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hsfm.h"

namespace hsfm {

constexpr uint32_t kRecShift = 6;   // 64 positions per record
constexpr uint32_t kSbRecShift = 10; // 1024 records per superblock
constexpr uint32_t kSbShift = 16;
constexpr uint32_t kLocateRegCap = 4; // intervals held in registers per read

// Per-index constants, resident in device memory and staged into LDS.
struct DevConst {
    uint16_t compat_lanes[256]; // bit k: lane k's symbol is compatible with query byte
    int8_t lookup_digit[256];   // digit of a byte among the lookup core symbols, -1 if none
    uint32_t c_lane[16];        // C[symbol of lane]
    uint32_t lane_sym[16];
    uint32_t num_lanes;
    uint32_t lookup_radix;
    uint32_t lookup_depth;      // 0 = no lookup table
    uint32_t pad;
    // Seed-table expansion of a query byte: bit 4 set = exact base with digit in
    // bits 0-1; else bits 0-3 = set of bases whose compat sets union to this
    // byte's compat set (IupacDna ambiguity codes); 0 = no expansion.
    uint8_t seed_expand[256];
    // Base set of each text lane for matching (bit x: the lane is compatible
    // with query base x); nib_text = 1 when, for every query byte b, lane l is
    // compatible with b exactly when nibble(b) & text_nib[l] != 0, nibble(b)
    // being b's base-set nibble of the query nibble stream (true for IupacDna
    // and ExactDna), so texts and queries can be compared as nibbles.
    uint8_t text_nib[16];
    uint32_t nib_text;
};

// Device view of one index direction (passed by value to kernels).
struct DevDir {
    const uint4 *occ;     // 4 uint4 per record
    const uint32_t *sb;   // 16 u32 per superblock
    const uint4 *samp;    // per 64 rows
    const uint32_t *sa_vals;
    const uint2 *lookup;  // (lo, hi) per k-mer, or null
    uint32_t n;           // text length (including sentinels)
    // SA value of every row, built on the device from the samples (resolve
    // becomes one load), or null
    const uint32_t *sa_full;
    // lane of every text position (built with sa_full; padded by 16 bytes), or null
    const uint8_t *text;
    // row of every text position (the inverse of sa_full), or null
    const uint32_t *isa;
    // text_nib of every text position's lane, eight per word (position i in
    // bits 4 (i & 7) of word i >> 3; zero past the end, padded by 16 bytes),
    // or null when the alphabet does not allow nibble comparison
    const uint32_t *text4;
    // base-set nibble of query bytes 0..15 (nibble b = byte b's set) when no
    // other byte has one, for comparing query bytes with text4 in registers;
    // 0 otherwise
    unsigned long long qnib16;
    uint32_t num_samples;
    uint32_t sample_rate;
    uint32_t num_lanes;
    uint64_t lookup_len;
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t popc64(uint64_t v) { return (uint32_t)__popcll(v); }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 ld4(const uint4 *p) {
    const uint4 v = *p;
    return u32x4{v.x, v.y, v.z, v.w};
}

// One loaded record plus its superblock row, kept as scalar/vector values (no
// arrays) so that per-lane dynamic selection compiles to v_cndmask chains.
struct RecView {
    uint64_t p0, p1, p2, p3;  // bitplanes
    uint64_t c0, c1, c2, c3;  // cum slots 0..15, four u16 per word
    u32x4 s0, s1, s2, s3;     // superblock slots 0..15
    uint32_t sb16, cum16;     // implied slot-16 values
};

__device__ __forceinline__ void load_rec(const DevDir &d, uint32_t x, RecView &r) {
    x = x < d.n ? x : d.n;
    const uint32_t blk = x >> kRecShift;
    const uint4 *rp = d.occ + ((size_t)blk << 2);
    const uint4 *sp = reinterpret_cast<const uint4 *>(d.sb + ((size_t)(blk >> kSbRecShift) << 4));
    const uint4 a = rp[0], b = rp[1], c = rp[2], e = rp[3];
    r.s0 = ld4(sp);
    r.s1 = ld4(sp + 1);
    r.s2 = ld4(sp + 2);
    r.s3 = ld4(sp + 3);
    r.p0 = (uint64_t)a.x | ((uint64_t)a.y << 32);
    r.p1 = (uint64_t)a.z | ((uint64_t)a.w << 32);
    r.p2 = (uint64_t)b.x | ((uint64_t)b.y << 32);
    r.p3 = (uint64_t)b.z | ((uint64_t)b.w << 32);
    r.c0 = (uint64_t)c.x | ((uint64_t)c.y << 32);
    r.c1 = (uint64_t)c.z | ((uint64_t)c.w << 32);
    r.c2 = (uint64_t)e.x | ((uint64_t)e.y << 32);
    r.c3 = (uint64_t)e.z | ((uint64_t)e.w << 32);
    r.sb16 = (blk >> kSbRecShift) << kSbShift;
    r.cum16 = (blk & ((1u << kSbRecShift) - 1)) << kRecShift;
}

// slot k (0..16) of the cum / superblock counts, k a runtime value
__device__ __forceinline__ uint32_t rec_cum(const RecView &r, uint32_t k) {
    const uint64_t w = (k & 8) ? ((k & 4) ? r.c3 : r.c2) : ((k & 4) ? r.c1 : r.c0);
    const uint32_t v = (uint32_t)(w >> ((k & 3) * 16)) & 0xffffu;
    return k >= 16 ? r.cum16 : (k == 0 ? 0u : v);
}
__device__ __forceinline__ uint32_t rec_sb(const RecView &r, uint32_t k) {
    const u32x4 q = (k & 8) ? ((k & 4) ? r.s3 : r.s2) : ((k & 4) ? r.s1 : r.s0);
    const uint32_t v = (k & 2) ? ((k & 1) ? q.w : q.z) : ((k & 1) ? q.y : q.x);
    return k >= 16 ? r.sb16 : (k == 0 ? 0u : v);
}
__device__ __forceinline__ uint64_t rec_plane(const RecView &r, int j) {
    return j == 0 ? r.p0 : j == 1 ? r.p1 : j == 2 ? r.p2 : r.p3;
}

// Counts of every lane k < L in [0, x): out[k].  x must be in [0, n].
__device__ __forceinline__ void lane_counts(const RecView &r, uint32_t x, uint32_t L, uint32_t (&out)[16]) {
    const uint64_t low = (1ull << (x & 63)) - 1;
    uint32_t below_prev = 0; // sb[k] + cum[k]
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t eq = ((k & 1) ? r.p0 : ~r.p0) & ((k & 2) ? r.p1 : ~r.p1) &
                            ((k & 4) ? r.p2 : ~r.p2) & ((k & 8) ? r.p3 : ~r.p3);
        const uint32_t below_next = rec_sb(r, (uint32_t)k + 1) + rec_cum(r, (uint32_t)k + 1);
        const uint32_t v = (below_next - below_prev) + popc64(eq & low);
        below_prev = below_next;
        out[k] = ((uint32_t)k < L) ? v : 0u;
    }
}

// Rows of every lane among the rows selected by m, from the four bitplanes of a
// 64-row block: a 4-level split tree (lane k = 8*p3 + 4*p2 + 2*p1 + p0), 30
// masks instead of 16 independent four-way ANDs.
__device__ __forceinline__ void tree_popc16(uint64_t m, uint64_t p0, uint64_t p1, uint64_t p2, uint64_t p3,
                                            uint32_t (&pc)[16]) {
#pragma unroll
    for (int h3 = 0; h3 < 2; h3++) {
        const uint64_t a = h3 ? (m & p3) : (m & ~p3);
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            const uint64_t b = h2 ? (a & p2) : (a & ~p2);
#pragma unroll
            for (int h1 = 0; h1 < 2; h1++) {
                const uint64_t c = h1 ? (b & p1) : (b & ~p1);
                pc[8 * h3 + 4 * h2 + 2 * h1 + 0] = popc64(c & ~p0);
                pc[8 * h3 + 4 * h2 + 2 * h1 + 1] = popc64(c & p0);
            }
        }
    }
}

// The loads of one probe point pair: the records of xlo and xhi and the
// superblock row of xlo (issued by pair_issue, consumed by pair_eval, so a
// caller can overlap them with other work).  The fields double as generic
// load slots for the small control-state loads of the pipelined kernels.
struct PairRaw {
    uint4 l0, l1, l2, l3, h0, h1, h2, h3, s0, s1, s2, s3;
};
__device__ __forceinline__ void pair_issue(const DevDir &d, uint32_t xlo, uint32_t xhi, PairRaw &w) {
    const uint32_t bl = xlo >> kRecShift, bh = xhi >> kRecShift;
    const uint4 *rpl = d.occ + ((size_t)bl << 2);
    const uint4 *rph = d.occ + ((size_t)bh << 2);
    const uint4 *spl = reinterpret_cast<const uint4 *>(d.sb + ((size_t)(bl >> kSbRecShift) << 4));
    w.l0 = rpl[0]; w.l1 = rpl[1]; w.l2 = rpl[2]; w.l3 = rpl[3];
    w.s0 = spl[0]; w.s1 = spl[1]; w.s2 = spl[2]; w.s3 = spl[3];
    // the record of xhi only when it is another one (nearly every narrow
    // interval has both ends in one record): lanes with one record make no
    // request, and the consumers take the xlo record's words (pair_hi)
    w.h0 = w.h1 = w.h2 = w.h3 = make_uint4(0u, 0u, 0u, 0u);
    if (bh != bl) {
        w.h0 = rph[0]; w.h1 = rph[1]; w.h2 = rph[2]; w.h3 = rph[3];
    }
}
// The words of the record of xhi issued by pair_issue: its own, or those of
// the record of xlo when both ends lie in one record.
__device__ __forceinline__ void pair_hi(const PairRaw &w, uint32_t xlo, uint32_t xhi, uint4 &h0, uint4 &h1,
                                        uint4 &h2, uint4 &h3) {
    const bool same = (xlo >> kRecShift) == (xhi >> kRecShift);
    h0 = same ? w.l0 : w.h0;
    h1 = same ? w.l1 : w.h1;
    h2 = same ? w.l2 : w.h2;
    h3 = same ? w.l3 : w.h3;
}
// Counts of every lane in [0, xlo) and [0, xhi) (xlo <= xhi <= n) from the two
// records and the superblock row of xlo only; the rare pair that spans two
// superblocks reloads the row of xhi.  Lanes the index does not use count 0
// because no position holds their code.
__device__ __forceinline__ void pair_eval(const DevDir &d, const PairRaw &w, uint32_t xlo, uint32_t xhi,
                                          uint32_t (&cl)[16], uint32_t (&ch)[16]) {
    uint4 h0, h1, h2, h3;
    pair_hi(w, xlo, xhi, h0, h1, h2, h3);
    const uint32_t bl = xlo >> kRecShift, bh = xhi >> kRecShift;
    const uint64_t lp0 = (uint64_t)w.l0.x | ((uint64_t)w.l0.y << 32), lp1 = (uint64_t)w.l0.z | ((uint64_t)w.l0.w << 32);
    const uint64_t lp2 = (uint64_t)w.l1.x | ((uint64_t)w.l1.y << 32), lp3 = (uint64_t)w.l1.z | ((uint64_t)w.l1.w << 32);
    const uint64_t hp0 = (uint64_t)h0.x | ((uint64_t)h0.y << 32), hp1 = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
    const uint64_t hp2 = (uint64_t)h1.x | ((uint64_t)h1.y << 32), hp3 = (uint64_t)h1.z | ((uint64_t)h1.w << 32);
    const uint32_t lcw[8] = {w.l2.x, w.l2.y, w.l2.z, w.l2.w, w.l3.x, w.l3.y, w.l3.z, w.l3.w};
    const uint32_t hcw[8] = {h2.x, h2.y, h2.z, h2.w, h3.x, h3.y, h3.z, h3.w};
    const uint32_t sbw[16] = {w.s0.x, w.s0.y, w.s0.z, w.s0.w, w.s1.x, w.s1.y, w.s1.z, w.s1.w,
                              w.s2.x, w.s2.y, w.s2.z, w.s2.w, w.s3.x, w.s3.y, w.s3.z, w.s3.w};
    const uint64_t lowl = (1ull << (xlo & 63)) - 1, lowh = (1ull << (xhi & 63)) - 1;
    const uint32_t sb16 = (bl >> kSbRecShift) << kSbShift;
    const uint32_t lcum16 = (bl & ((1u << kSbRecShift) - 1)) << kRecShift;
    const uint32_t hcum16 = (bh & ((1u << kSbRecShift) - 1)) << kRecShift;
    tree_popc16(lowl, lp0, lp1, lp2, lp3, cl);
    tree_popc16(lowh, hp0, hp1, hp2, hp3, ch);
    uint32_t pl = 0, ph = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        // slot k+1 of the cumulative counts (slot 16 = positions before the record)
        const uint32_t sbn = k == 15 ? sb16 : sbw[k + 1];
        const uint32_t cln = k == 15 ? lcum16 : (lcw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu;
        const uint32_t chn = k == 15 ? hcum16 : (hcw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu;
        const uint32_t nl = sbn + cln, nh = sbn + chn;
        cl[k] += nl - pl;
        ch[k] += nh - ph;
        pl = nl;
        ph = nh;
    }
    if ((bl >> kSbRecShift) != (bh >> kSbRecShift)) {
        RecView rh;
        load_rec(d, xhi, rh);
        lane_counts(rh, xhi, 16, ch);
    }
}
// pair_eval for callers that only walk the lanes in order: the counts at xlo
// first (16 values), then the counts at xhi one lane at a time, handed to
// visit(k, count at xlo, count at xhi).  The superblock row of xhi replaces
// that of xlo (reloaded, rare) between the two halves.
template <typename V>
__device__ __forceinline__ void pair_visit(const DevDir &d, PairRaw &w, uint32_t xlo, uint32_t xhi, V &&visit) {
    uint4 h0, h1, h2, h3;
    pair_hi(w, xlo, xhi, h0, h1, h2, h3);
    const uint32_t bl = xlo >> kRecShift, bh = xhi >> kRecShift;
    uint32_t cl[16];
    {
        const uint64_t p0 = (uint64_t)w.l0.x | ((uint64_t)w.l0.y << 32), p1 = (uint64_t)w.l0.z | ((uint64_t)w.l0.w << 32);
        const uint64_t p2 = (uint64_t)w.l1.x | ((uint64_t)w.l1.y << 32), p3 = (uint64_t)w.l1.z | ((uint64_t)w.l1.w << 32);
        const uint32_t cw[8] = {w.l2.x, w.l2.y, w.l2.z, w.l2.w, w.l3.x, w.l3.y, w.l3.z, w.l3.w};
        const uint32_t sbw[16] = {w.s0.x, w.s0.y, w.s0.z, w.s0.w, w.s1.x, w.s1.y, w.s1.z, w.s1.w,
                                  w.s2.x, w.s2.y, w.s2.z, w.s2.w, w.s3.x, w.s3.y, w.s3.z, w.s3.w};
        const uint64_t low = (1ull << (xlo & 63)) - 1;
        const uint32_t sb16 = (bl >> kSbRecShift) << kSbShift;
        const uint32_t cum16 = (bl & ((1u << kSbRecShift) - 1)) << kRecShift;
        uint32_t prev = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t e = ((k & 1) ? p0 : ~p0) & ((k & 2) ? p1 : ~p1) & ((k & 4) ? p2 : ~p2) & ((k & 8) ? p3 : ~p3);
            const uint32_t nx = (k == 15 ? sb16 : sbw[k + 1]) +
                                (k == 15 ? cum16 : (cw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu);
            cl[k] = (nx - prev) + popc64(e & low);
            prev = nx;
        }
    }
    if ((bl >> kSbRecShift) != (bh >> kSbRecShift)) {
        const uint4 *sph = reinterpret_cast<const uint4 *>(d.sb + ((size_t)(bh >> kSbRecShift) << 4));
        w.s0 = sph[0]; w.s1 = sph[1]; w.s2 = sph[2]; w.s3 = sph[3];
    }
    const uint64_t p0 = (uint64_t)h0.x | ((uint64_t)h0.y << 32), p1 = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
    const uint64_t p2 = (uint64_t)h1.x | ((uint64_t)h1.y << 32), p3 = (uint64_t)h1.z | ((uint64_t)h1.w << 32);
    const uint32_t cw[8] = {h2.x, h2.y, h2.z, h2.w, h3.x, h3.y, h3.z, h3.w};
    const uint32_t sbw[16] = {w.s0.x, w.s0.y, w.s0.z, w.s0.w, w.s1.x, w.s1.y, w.s1.z, w.s1.w,
                              w.s2.x, w.s2.y, w.s2.z, w.s2.w, w.s3.x, w.s3.y, w.s3.z, w.s3.w};
    const uint64_t low = (1ull << (xhi & 63)) - 1;
    const uint32_t sb16 = (bh >> kSbRecShift) << kSbShift;
    const uint32_t cum16 = (bh & ((1u << kSbRecShift) - 1)) << kRecShift;
    uint32_t prev = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t e = ((k & 1) ? p0 : ~p0) & ((k & 2) ? p1 : ~p1) & ((k & 4) ? p2 : ~p2) & ((k & 8) ? p3 : ~p3);
        const uint32_t nx = (k == 15 ? sb16 : sbw[k + 1]) +
                            (k == 15 ? cum16 : (cw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu);
        visit(k, cl[k], (nx - prev) + popc64(e & low));
        prev = nx;
    }
}

// pair_visit with both points evaluated lane by lane in one loop (no per-lane
// arrays); the pair that spans two superblocks takes the reload path.
template <typename V>
__device__ __forceinline__ void pair_visit_fused(const DevDir &d, PairRaw &w, uint32_t xlo, uint32_t xhi,
                                                 V &&visit) {
    uint4 h0, h1, h2, h3;
    pair_hi(w, xlo, xhi, h0, h1, h2, h3);
    const uint32_t bl = xlo >> kRecShift, bh = xhi >> kRecShift;
    if ((bl >> kSbRecShift) != (bh >> kSbRecShift)) {
        pair_visit(d, w, xlo, xhi, visit);
        return;
    }
    const uint64_t lp0 = (uint64_t)w.l0.x | ((uint64_t)w.l0.y << 32), lp1 = (uint64_t)w.l0.z | ((uint64_t)w.l0.w << 32);
    const uint64_t lp2 = (uint64_t)w.l1.x | ((uint64_t)w.l1.y << 32), lp3 = (uint64_t)w.l1.z | ((uint64_t)w.l1.w << 32);
    const uint64_t hp0 = (uint64_t)h0.x | ((uint64_t)h0.y << 32), hp1 = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
    const uint64_t hp2 = (uint64_t)h1.x | ((uint64_t)h1.y << 32), hp3 = (uint64_t)h1.z | ((uint64_t)h1.w << 32);
    const uint32_t lcw[8] = {w.l2.x, w.l2.y, w.l2.z, w.l2.w, w.l3.x, w.l3.y, w.l3.z, w.l3.w};
    const uint32_t hcw[8] = {h2.x, h2.y, h2.z, h2.w, h3.x, h3.y, h3.z, h3.w};
    const uint32_t sbw[16] = {w.s0.x, w.s0.y, w.s0.z, w.s0.w, w.s1.x, w.s1.y, w.s1.z, w.s1.w,
                              w.s2.x, w.s2.y, w.s2.z, w.s2.w, w.s3.x, w.s3.y, w.s3.z, w.s3.w};
    const uint64_t lowl = (1ull << (xlo & 63)) - 1, lowh = (1ull << (xhi & 63)) - 1;
    const uint32_t sb16 = (bl >> kSbRecShift) << kSbShift;
    const uint32_t lcum16 = (bl & ((1u << kSbRecShift) - 1)) << kRecShift;
    const uint32_t hcum16 = (bh & ((1u << kSbRecShift) - 1)) << kRecShift;
    uint32_t pl = 0, ph = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t el = ((k & 1) ? lp0 : ~lp0) & ((k & 2) ? lp1 : ~lp1) & ((k & 4) ? lp2 : ~lp2) &
                            ((k & 8) ? lp3 : ~lp3);
        const uint64_t eh = ((k & 1) ? hp0 : ~hp0) & ((k & 2) ? hp1 : ~hp1) & ((k & 4) ? hp2 : ~hp2) &
                            ((k & 8) ? hp3 : ~hp3);
        const uint32_t sbn = k == 15 ? sb16 : sbw[k + 1];
        const uint32_t nl = sbn + (k == 15 ? lcum16 : (lcw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu);
        const uint32_t nh = sbn + (k == 15 ? hcum16 : (hcw[(k + 1) >> 1] >> (((k + 1) & 1) * 16)) & 0xffffu);
        visit(k, (nl - pl) + popc64(el & lowl), (nh - ph) + popc64(eh & lowh));
        pl = nl;
        ph = nh;
    }
}

__device__ __forceinline__ void pair_counts(const DevDir &d, uint32_t xlo, uint32_t xhi, uint32_t (&cl)[16],
                                            uint32_t (&ch)[16]) {
    PairRaw w;
    pair_issue(d, xlo, xhi, w);
    pair_eval(d, w, xlo, xhi, cl, ch);
}

// Lane of BWT row `row` and #same-lane rows in [0, row): the LF step of
// src/occ/mod.rs:456-480 on the device layout.
__device__ __forceinline__ void lf_lane_rank(const RecView &r, uint32_t row, uint32_t &lane, uint32_t &rank) {
    const uint32_t off = row & 63;
    const uint32_t v = (uint32_t)((r.p0 >> off) & 1ull) | ((uint32_t)((r.p1 >> off) & 1ull) << 1) |
                       ((uint32_t)((r.p2 >> off) & 1ull) << 2) | ((uint32_t)((r.p3 >> off) & 1ull) << 3);
    uint64_t eq = ~0ull;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint64_t m = 0ull - (uint64_t)((v >> j) & 1u);
        eq &= ~(rec_plane(r, j) ^ m);
    }
    const uint64_t low = (1ull << off) - 1;
    lane = v;
    // below(v+1) - below(v) as a masked sum over the 16 static slots: keeps the
    // selection in registers (a select chain over the row gets rewritten into
    // an indexed scratch access by the optimizer).
    uint32_t acc = 0, prev = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint32_t next = rec_sb(r, (uint32_t)k + 1) + rec_cum(r, (uint32_t)k + 1);
        acc += (next - prev) & (0u - (uint32_t)(v == (uint32_t)k));
        prev = next;
    }
    rank = acc + popc64(eq & low);
}

// Sampled-SA lookup of row (src/suffix_array/mod.rs:185-201).
__device__ __forceinline__ bool sa_sampled(const DevDir &d, uint32_t row, uint32_t &val) {
    const uint4 w = d.samp[row >> 6];
    const uint64_t bits = (uint64_t)w.x | ((uint64_t)w.y << 32);
    const uint32_t off = row & 63;
    if (!((bits >> off) & 1ull)) return false;
    uint32_t rk = w.z + popc64(bits & ((1ull << off) - 1));
    rk = rk < d.num_samples ? rk : d.num_samples - 1;
    val = d.sa_vals[rk];
    return true;
}

// Counts of every lane at the two probe points of a bidirectional interval
// {fwd_lo, fwd_hi, rev_lo, -}: fwd points for a left extension (P -> cP on the
// forward index), rev points for a right extension (P -> Pc on the reverse
// index), as in bidir.rs:80-130.
template <bool LEFT>
__device__ __forceinline__ void interval_counts(const DevDir &D, uint32_t L, const uint4 &iv, uint32_t (&cl)[16],
                                                uint32_t (&ch)[16]) {
    const uint32_t size = iv.y - iv.x;
    const uint32_t xlo = LEFT ? iv.x : iv.z;
    const uint32_t xhi = LEFT ? iv.y : iv.z + size;
    RecView rl, rh;
    load_rec(D, xlo, rl);
    load_rec(D, xhi, rh);
    lane_counts(rl, xlo, L, cl);
    lane_counts(rh, xhi, L, ch);
}
// Child of iv for lane k: clane = C of the lane, (a, b) = counts at the two
// points, acc = sum over lanes below k of (b - a) (count_smaller_than).
template <bool LEFT>
__device__ __forceinline__ uint4 bidir_child(const uint4 &iv, uint32_t clane, uint32_t a, uint32_t b, uint32_t acc) {
    return LEFT ? make_uint4(clane + a, clane + b, iv.z + acc, 0u)
                : make_uint4(iv.x + acc, iv.x + acc + (b - a), clane + a, 0u);
}

// Seed table view (built by build_seed_table): per ACGT F-mer index, the
// offsets of its interval list, the bidirectional intervals, and lsl.
struct SeedView {
    const uint64_t *off;
    const uint4 *iv;
    const uint8_t *lsl;
    uint32_t F;
    // packed records (build_seed_packed): {fwd_lo, fwd_hi, rev_lo, meta} for an
    // F-mer with one interval, {iv offset lo, hi, 0, meta} otherwise;
    // meta = lsl | extension mask << 5 | count << 9.  Bit b of the mask is
    // set when the set of the F-mer has a left child under query base b
    // (all four bits set when not known: several intervals)
    const uint4 *rec;
};
constexpr uint32_t kSeedLslBits = 5;
constexpr uint32_t kSeedCntShift = 9;
__host__ __device__ __forceinline__ uint32_t seed_rec_lsl(uint32_t meta) { return meta & 31u; }
__host__ __device__ __forceinline__ uint32_t seed_rec_mask(uint32_t meta) { return (meta >> kSeedLslBits) & 15u; }
__host__ __device__ __forceinline__ uint32_t seed_rec_cnt(uint32_t meta) { return meta >> kSeedCntShift; }

// A seed window q[j-F..j) as table indices: ACGT codes give fixed digits, up
// to kSeedAmb ambiguity codes expand to every base they contain (the union of
// the bases' compatible sets is the code's own).  pos holds 8 bits, mask 4
// bits (the bases) per code; unused codes keep mask 1, digit 0.
constexpr uint32_t kSeedAmb = 3;
constexpr uint32_t kSeedCombos = 1u << (2 * kSeedAmb);
struct SeedWin {
    uint32_t base, pos, mask, namb;
};
__device__ __forceinline__ bool seed_window(const uint8_t *w, uint32_t F, const uint8_t *s_expand, SeedWin &sw,
                                            uint32_t max_amb = kSeedAmb) {
    sw.base = 0; sw.pos = 0; sw.mask = 0x111u; sw.namb = 0;
    bool ok = true;
    for (uint32_t u = 0; u < F; u++) {
        const uint32_t x = s_expand[w[u]];
        if (x & 0x10u) sw.base |= (x & 3u) << (2 * u);
        else if (x && sw.namb < max_amb) {
            sw.pos |= u << (8 * sw.namb);
            sw.mask = (sw.mask & ~(0xfu << (4 * sw.namb))) | (x << (4 * sw.namb));
            sw.namb++;
        } else ok = false;
    }
    return ok;
}
// seed_window of q[0..F) (12 <= F <= 16) from aligned dwords: seed_window_load
// issues the loads (only dwords holding bytes of the window, so every load
// stays inside memory the window itself occupies), seed_window_decode reads
// the codes out of them.
__device__ __forceinline__ void seed_window_load(const uint8_t *q, uint32_t F, uint4 &d, uint32_t &d4) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(q);
    const uint32_t *base = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    const uint32_t last = ((uint32_t)(a & 3) + F - 1) >> 2; // 2..4
    d.x = base[0];
    d.y = base[1];
    d.z = base[2];
    d.w = base[last < 3 ? last : 3];
    d4 = base[last];
}
__device__ __forceinline__ bool seed_window_decode(const uint8_t *q, uint32_t F, const uint4 &d, uint32_t d4,
                                                   const uint8_t *s_expand, SeedWin &sw, uint32_t max_amb) {
    const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(q) & 3);
    const uint32_t dd[5] = {d.x, d.y, d.z, d.w, d4};
    uint32_t b[4]; // the 16 bytes from the window start, four per word
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) b[i] = sh ? (dd[i] >> (8 * sh)) | (dd[i + 1] << (32 - 8 * sh)) : dd[i];
    // all sixteen table lookups first (independent LDS reads), then a
    // branch-free pass: exact bases give digits, ambiguity codes are counted
    // and the first max_amb (<= 3) of them recorded in order
    uint32_t xs[16];
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) xs[u] = s_expand[(b[u >> 2] >> (8 * (u & 3))) & 0xffu];
    uint32_t base = 0, pos = 0, mask = 0x111u, na = 0;
    bool bad = false;
#pragma unroll
    for (uint32_t u = 0; u < 16; u++) {
        const uint32_t x = xs[u];
        const bool in = u < F;
        const bool ex = in && (x & 0x10u);
        const bool am = in && !(x & 0x10u) && x != 0u;
        bad |= in && x == 0u;
        base |= ex ? (x & 3u) << (2 * u) : 0u;
        const bool rec = am && na < max_amb;
        pos |= rec ? u << (8 * na) : 0u;
        mask = rec ? (mask & ~(0xfu << (4 * na))) | (x << (4 * na)) : mask;
        na += am ? 1u : 0u;
    }
    sw.base = base;
    sw.pos = pos;
    sw.mask = mask;
    sw.namb = na < max_amb ? na : max_amb;
    return !bad && na <= max_amb;
}
// Queries as base-set nibbles (k_pack_bases): 4 bits per code, eight codes per
// word, code i of the flat query array in bits 4(i & 7) of word i >> 3.  The
// nibble of an exact base (codes 1..4) is that base alone; of another code,
// the bases of its seed expansion (0 when it has none).  seed_window_nib
// decodes the window of F (<= 14) codes starting at flat position g: digits
// of the one-base nibbles, ambiguity codes (several bases) in order.  It
// gives the same table indices as seed_window (a code whose expansion is a
// single base reads as that base, the one index its expansion would give).
__device__ __forceinline__ uint32_t base_nibble(uint32_t c, const uint8_t *s_expand) {
    return (c >= 1 && c <= 4) ? 1u << (c - 1) : (uint32_t)(s_expand[c] & 15u);
}
__device__ __forceinline__ void seed_window_nib_load(const uint32_t *nib, uint64_t g, uint32_t &w0, uint32_t &w1,
                                                     uint32_t &w2) {
    const uint32_t *b = nib + (g >> 3);
    w0 = b[0];
    w1 = b[1];
    w2 = b[2];
}
__device__ __forceinline__ bool seed_window_nib(uint64_t g, uint32_t F, uint32_t w0, uint32_t w1, uint32_t w2,
                                                SeedWin &sw, uint32_t max_amb) {
    const uint32_t sh = 4 * (uint32_t)(g & 7);
    uint64_t x = ((uint64_t)w0 | ((uint64_t)w1 << 32)) >> sh;
    if (sh) x |= (uint64_t)w2 << (64 - sh);
    const uint64_t ones = 0x1111111111111111ull >> (64 - 4 * F); // bit 0 of each window nibble
    x &= ones * 15u;
    // a nibble with no base: no expansion; with two or more: ambiguity code
    const uint64_t any = (x | (x >> 1) | (x >> 2) | (x >> 3)) & ones;
    const uint64_t multi = ((x & (x >> 1)) | (x & (x >> 2)) | (x & (x >> 3)) | ((x >> 1) & (x >> 2)) |
                            ((x >> 1) & (x >> 3)) | ((x >> 2) & (x >> 3))) & ones;
    // digit of a one-base nibble n (1, 2, 4, 8 -> 0..3): (n >> 1) - (n >> 3)
    uint64_t d = ((x >> 1) & (ones * 7u)) - ((x >> 3) & ones);
    d &= ~(multi * 15u);
    // four bits per digit -> two bits per digit, in window order
    d = (d | (d >> 2)) & 0x0f0f0f0f0f0f0f0full;
    d = (d | (d >> 4)) & 0x00ff00ff00ff00ffull;
    d = (d | (d >> 8)) & 0x0000ffff0000ffffull;
    d = (d | (d >> 16)) & 0xffffffffull;
    sw.base = (uint32_t)d;
    sw.pos = 0;
    sw.mask = 0x111u;
    const uint32_t na = (uint32_t)__popcll((unsigned long long)multi);
    sw.namb = na < max_amb ? na : max_amb;
    uint64_t mm = multi;
#pragma unroll
    for (uint32_t i = 0; i < kSeedAmb; i++) {
        if (i < sw.namb) {
            const uint32_t bit = (uint32_t)__ffsll((unsigned long long)mm) - 1u; // 4 * position
            mm &= mm - 1;
            sw.pos |= (bit >> 2) << (8 * i);
            sw.mask = (sw.mask & ~(0xfu << (4 * i))) | ((uint32_t)((x >> bit) & 15u) << (4 * i));
        }
    }
    return any == ones && na <= max_amb;
}
// seed_window over the last F (<= 16) of 16 codes held one per register.
__device__ __forceinline__ bool seed_window16(const uint32_t (&wb)[16], uint32_t F, const uint8_t *s_expand,
                                              SeedWin &sw) {
    sw.base = 0; sw.pos = 0; sw.mask = 0x111u; sw.namb = 0;
    bool ok = true;
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
        if (k + F < 16) continue;
        const uint32_t u = k + F - 16; // position in the window
        const uint32_t x = s_expand[wb[k] & 0xffu];
        if (x & 0x10u) sw.base |= (x & 3u) << (2 * u);
        else if (x && sw.namb < kSeedAmb) {
            sw.pos |= u << (8 * sw.namb);
            sw.mask = (sw.mask & ~(0xfu << (4 * sw.namb))) | (x << (4 * sw.namb));
            sw.namb++;
        } else ok = false;
    }
    return ok;
}
// Expansions of a window: xi in [0, 4^namb) (digit i of xi = the base taken
// for ambiguity code i).
__device__ __forceinline__ uint32_t seed_ncombo(const SeedWin &sw) { return 1u << (2 * sw.namb); }
// Table index of expansion xi (base digit i = bits 2i..2i+1 of xi); false if
// a digit is not a base of its code.
__device__ __forceinline__ bool seed_combo(const SeedWin &sw, uint32_t xi, uint32_t &idx) {
    idx = sw.base;
    bool ok = true;
#pragma unroll
    for (uint32_t i = 0; i < kSeedAmb; i++) {
        const uint32_t d = (xi >> (2 * i)) & 3u;
        ok &= ((sw.mask >> (4 * i + d)) & 1u) != 0;
        if (i < sw.namb) idx |= d << (2 * ((sw.pos >> (8 * i)) & 0xffu));
    }
    return ok;
}

// A seed window with up to kWideAmb ambiguity codes (expanded cooperatively
// by a workgroup): 4-bit positions and base masks per code.
constexpr uint32_t kWideAmb = 5;
struct WideWin {
    uint32_t base, pos, mask, namb;
};
__device__ __forceinline__ bool wide_window(const uint8_t *w, uint32_t F, const uint8_t *s_expand, WideWin &ww) {
    ww.base = 0; ww.pos = 0; ww.mask = 0; ww.namb = 0;
    bool ok = true;
    for (uint32_t u = 0; u < F; u++) {
        const uint32_t x = s_expand[w[u]];
        if (x & 0x10u) ww.base |= (x & 3u) << (2 * u);
        else if (x && ww.namb < kWideAmb) {
            ww.pos |= u << (4 * ww.namb);
            ww.mask |= x << (4 * ww.namb);
            ww.namb++;
        } else ok = false;
    }
    return ok;
}
// Table index of expansion xi in [0, 4^namb); false if a digit is not a base
// of its code.
__device__ __forceinline__ bool wide_combo(const WideWin &ww, uint32_t xi, uint32_t &idx) {
    idx = ww.base;
    bool ok = true;
    for (uint32_t i = 0; i < ww.namb; i++) {
        const uint32_t d = (xi >> (2 * i)) & 3u;
        ok &= ((ww.mask >> (4 * i + d)) & 1u) != 0;
        idx |= d << (2 * ((ww.pos >> (4 * i)) & 0xfu));
    }
    return ok;
}

// wave64 helpers
__device__ __forceinline__ uint32_t wave_lane() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return __shfl(wave_incl_scan(v), 63, 64); }

} // namespace hsfm
