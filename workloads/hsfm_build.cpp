// hsfm_build.cpp -- host-side FmIndex construction (the reference's build_cpu +
// to_bytes) and seeded workload synthesis for the benchmark and the tests.
// Not part of libhsfm.so (the device builder hsfm_build_index_device is).
//
// Construction follows FmIndex::build_cpu_inner (src/fm_index/mod.rs:117-195):
//   C array from the text (c_array.rs:14-31); suffix array in plain
//   lexicographic order with the 0 sentinels compared as ordinary symbols
//   (suffix_array/cpu.rs:20-25); BWT[i] = T[SA[i]-1 mod n] (bwt/cpu.rs:14-35);
//   sampled SA with forced sequence starts (suffix_array/mod.rs:71-118);
//   compact-lane Occ with interleaved records (occ/cpu.rs:23-119,
//   occ/mod.rs:114-183); lookup table BFS (lookup.rs:37-91).
// Serialization is bincode 1 in SerializableFmIndex field order
// (serialize.rs:55-67).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "hsfm_build.h"

namespace {

// Stage lines on stderr while a large text is sorted (HSFM_BUILD_PROGRESS=1):
// a 1 Gbp suffix sort runs for minutes, and a caller watching its output can
// tell a slow build from a stalled one.
bool g_progress = false;
void progress(const char *what, int64_t n, int level) {
    if (!g_progress || n < (1 << 24)) return;
    fprintf(stderr, "[build] %lld symbols, recursion level %d: %s\n", (long long)n, level, what);
    fflush(stderr);
}

// ---------------------------------------------------------------------------
// suffix sorting: induced sorting (SA-IS) with an implicit end marker that is
// smaller than every symbol
// ---------------------------------------------------------------------------
template <class CharT>
void induce(const CharT *T, int32_t *SA, int32_t n, const std::vector<uint8_t> &stype,
            const std::vector<int64_t> &cnt, std::vector<int64_t> &bkt) {
    const size_t K = cnt.size();
    int64_t s = 0;
    for (size_t k = 0; k < K; k++) { bkt[k] = s; s += cnt[k]; }
    SA[bkt[T[n - 1]]++] = n - 1;
    for (int32_t i = 0; i < n; i++) {
        const int32_t j = SA[i] - 1;
        if (SA[i] > 0 && !stype[j]) SA[bkt[T[j]]++] = j;
    }
    s = 0;
    for (size_t k = 0; k < K; k++) { s += cnt[k]; bkt[k] = s; }
    for (int32_t i = n - 1; i >= 0; i--) {
        const int32_t j = SA[i] - 1;
        if (SA[i] > 0 && stype[j]) SA[--bkt[T[j]]] = j;
    }
}

template <class CharT>
void sais(const CharT *T, int32_t *SA, int32_t n, int32_t K, int level = 0) {
    if (n <= 0) return;
    if (n == 1) { SA[0] = 0; return; }
    std::vector<uint8_t> st(n);
    st[n - 1] = 0;
    for (int32_t i = n - 2; i >= 0; i--) st[i] = (T[i] < T[i + 1] || (T[i] == T[i + 1] && st[i + 1])) ? 1 : 0;
    auto lms = [&](int32_t i) { return i > 0 && st[i] && !st[i - 1]; };
    std::vector<int64_t> cnt(K, 0), bkt(K, 0);
    for (int32_t i = 0; i < n; i++) cnt[T[i]]++;
    int64_t s = 0;
    for (int32_t k = 0; k < K; k++) { s += cnt[k]; bkt[k] = s; }
    std::fill(SA, SA + n, -1);
    for (int32_t i = 1; i < n; i++)
        if (lms(i)) SA[--bkt[T[i]]] = i;
    induce(T, SA, n, st, cnt, bkt);
    progress("LMS substrings sorted", n, level);
    int32_t n1 = 0;
    for (int32_t i = 0; i < n; i++)
        if (lms(SA[i])) SA[n1++] = SA[i];
    std::fill(SA + n1, SA + n, -1);
    int32_t names = 0, prev = -1;
    for (int32_t i = 0; i < n1; i++) {
        const int32_t p = SA[i];
        bool diff = prev < 0;
        for (int32_t d = 0; !diff; d++) {
            if (p + d == n || prev + d == n || T[p + d] != T[prev + d] || st[p + d] != st[prev + d]) { diff = true; break; }
            if (d > 0 && (lms(p + d) || lms(prev + d))) { diff = !(lms(p + d) && lms(prev + d)); break; }
        }
        if (diff) { names++; prev = p; }
        SA[n1 + p / 2] = names - 1;
    }
    for (int32_t i = n - 1, j = n - 1; i >= n1; i--)
        if (SA[i] >= 0) SA[j--] = SA[i];
    int32_t *s1 = SA + n - n1;
    progress("LMS substrings named", n, level);
    if (names < n1) sais<int32_t>(s1, SA, n1, names, level + 1);
    else for (int32_t i = 0; i < n1; i++) SA[s1[i]] = i;
    for (int32_t i = 1, j = 0; i < n; i++)
        if (lms(i)) s1[j++] = i;
    for (int32_t i = 0; i < n1; i++) SA[i] = s1[SA[i]];
    std::fill(SA + n1, SA + n, -1);
    s = 0;
    for (int32_t k = 0; k < K; k++) { s += cnt[k]; bkt[k] = s; }
    for (int32_t i = n1 - 1; i >= 0; i--) {
        const int32_t j = SA[i];
        SA[i] = -1;
        SA[--bkt[T[j]]] = j;
    }
    induce(T, SA, n, st, cnt, bkt);
    progress("suffixes induced", n, level);
}

// ---------------------------------------------------------------------------
// index pieces in the reference layout
// ---------------------------------------------------------------------------
struct Occ {
    uint8_t L = 0, P = 0, s2l[16], l2s[16];
    std::vector<uint32_t> sb;
    std::vector<uint8_t> blk;
    uint64_t stride = 0;
    uint32_t enc = 0, n = 0;
};
struct Ssa {
    std::vector<uint8_t> words;
    std::vector<uint32_t> vals;
    uint32_t rate = 0, n = 0;
};
struct Index {
    uint32_t c[16];
    Occ occ;
    Ssa sa;
    uint32_t n = 0, nseq = 0;
    std::vector<uint32_t> bounds;
    std::vector<std::string> headers;
    bool has_lookup = false;
    uint32_t depth = 0;
    std::vector<uint8_t> core;
    std::vector<uint32_t> lk; // pairs
    uint8_t tag = 0;
};

template <class F>
void parallel_for(uint64_t n, int nthreads, F f) {
    if (nthreads <= 1 || n < 4096) { f(0, n); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const uint64_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto &x : th) x.join();
}

void build_occ(const uint8_t *bwt, uint32_t n, uint32_t enc, Occ &o) {
    const uint64_t nb = (n + 63ull) / 64;
    uint64_t tot[16] = {0};
    for (uint32_t i = 0; i < n; i++) tot[bwt[i] & 15]++;
    memset(o.s2l, 0xff, 16);
    memset(o.l2s, 0, 16);
    o.L = 0;
    for (int c = 0; c < 16; c++)
        if (tot[c]) { o.s2l[c] = o.L; o.l2s[o.L] = (uint8_t)c; o.L++; }
    o.P = 0;
    if (o.L > 1) while ((1u << o.P) < o.L) o.P++;
    const uint32_t W = enc == 0 ? o.P : o.L;
    o.stride = 6ull * o.L + 8ull * W;
    o.enc = enc;
    o.n = n;
    o.blk.assign(nb * o.stride, 0);
    o.sb.assign(((nb + 7) / 8) * o.L, 0);
    uint32_t cum[16] = {0}, base[16] = {0};
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t bits[16] = {0};
        uint32_t cnt[16] = {0};
        const uint32_t st = (uint32_t)(b * 64), en = std::min<uint32_t>(st + 64, n);
        for (uint32_t p = st; p < en; p++) { bits[bwt[p]] |= 1ull << (p - st); cnt[bwt[p]]++; }
        if (b % 8 == 0)
            for (int l = 0; l < o.L; l++) { o.sb[(b / 8) * o.L + l] = cum[l]; base[l] = cum[l]; }
        uint8_t *rec = o.blk.data() + b * o.stride;
        uint64_t ld[16] = {0};
        for (int c = 0; c < 16; c++) {
            const uint8_t lane = o.s2l[c];
            if (lane == 0xff) continue;
            const uint16_t d = (uint16_t)(cum[lane] - base[lane]);
            memcpy(rec + 4ull * o.L + 2ull * lane, &d, 2);
            if (enc == 0) { for (uint32_t p = 0; p < W; p++) if ((lane >> p) & 1) ld[p] |= bits[c]; }
            else ld[lane] = bits[c];
            cum[lane] += cnt[c];
        }
        for (int l = 0; l < o.L; l++) memcpy(rec + 4ull * l, &o.sb[(b / 8) * o.L + l], 4);
        for (uint32_t w = 0; w < W; w++) memcpy(rec + 6ull * o.L + 8ull * w, &ld[w], 8);
    }
}

inline uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
inline uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
inline uint16_t rd16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }

// rank(c, i) on the reference record layout (occ/mod.rs:330-363)
uint32_t occ_rank(const Occ &o, uint8_t c, uint32_t i) {
    if (i == 0 || c >= 16 || o.s2l[c] == 0xff) return 0;
    const uint32_t lane = o.s2l[c], pos = i - 1, off = pos % 64;
    const uint8_t *rec = o.blk.data() + (uint64_t)(pos / 64) * o.stride;
    uint64_t m;
    if (o.enc == 1) m = rd64(rec + 6ull * o.L + 8ull * lane);
    else if (o.P == 0) m = ~0ull;
    else {
        m = ~0ull;
        for (uint32_t p = 0; p < o.P; p++) {
            const uint64_t pv = rd64(rec + 6ull * o.L + 8ull * p);
            m &= ((lane >> p) & 1) ? pv : ~pv;
        }
    }
    const uint64_t mask = off == 63 ? ~0ull : ((1ull << (off + 1)) - 1);
    return rd32(rec + 4ull * lane) + rd16(rec + 4ull * o.L + 2ull * lane) + (uint32_t)__builtin_popcountll(m & mask);
}

void build_lookup(Index &ix, uint32_t depth, int nthreads) {
    ix.has_lookup = true;
    ix.depth = depth;
    ix.core = {1, 2, 3, 4};
    uint64_t size = 1;
    for (uint32_t i = 0; i < depth; i++) size *= 4;
    ix.lk.assign(size * 2, 0);
    struct Node { uint32_t idx, lo, hi; };
    std::vector<Node> cur{{0, 0, ix.n}};
    for (uint32_t level = 1; level <= depth; level++) {
        const bool last = level == depth;
        const int nt = cur.size() < 4096 ? 1 : nthreads;
        std::vector<std::vector<Node>> outs(nt);
        std::vector<std::thread> th;
        auto work = [&](int t) {
            const uint64_t b = cur.size() * t / nt, e = cur.size() * (t + 1) / nt;
            for (uint64_t k = b; k < e; k++) {
                const Node p = cur[k];
                for (uint32_t d = 0; d < 4; d++) {
                    const uint8_t sym = ix.core[d];
                    const uint32_t child = p.idx * 4 + d;
                    const uint32_t up = sym + 1 < 16 ? ix.c[sym + 1] : ix.n;
                    if (up - ix.c[sym] == 0) continue;
                    const uint32_t nl = ix.c[sym] + occ_rank(ix.occ, sym, p.lo), nh = ix.c[sym] + occ_rank(ix.occ, sym, p.hi);
                    if (last) { ix.lk[2ull * child] = nl; ix.lk[2ull * child + 1] = nh; }
                    else if (nl < nh) outs[t].push_back({child, nl, nh});
                }
            }
        };
        for (int t = 1; t < nt; t++) th.emplace_back(work, t);
        work(0);
        for (auto &x : th) x.join();
        if (!last) {
            std::vector<Node> nxt;
            for (auto &o : outs) nxt.insert(nxt.end(), o.begin(), o.end());
            cur.swap(nxt);
        }
    }
}

bool build(const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq, const char *headers,
           uint32_t rate, uint32_t depth, uint32_t enc, uint8_t tag, int nthreads, Index &ix) {
    if (n == 0 || n >= 0x7fffffffull || rate == 0 || nseq == 0 || depth > 15 || enc > 1) return false;
    ix.n = (uint32_t)n;
    ix.nseq = nseq;
    ix.tag = tag;
    ix.bounds.assign(bounds, bounds + nseq);
    const char *hp = headers;
    for (uint32_t i = 0; i < nseq; i++) {
        if (hp && *hp) { ix.headers.emplace_back(hp); hp += strlen(hp) + 1; }
        else { ix.headers.push_back("seq_" + std::to_string(i)); if (hp) hp++; }
    }
    uint64_t f[16] = {0};
    for (uint64_t i = 0; i < n; i++) f[text[i] & 15]++;
    uint32_t s = 0;
    for (int i = 0; i < 16; i++) { ix.c[i] = s; s += (uint32_t)f[i]; }
    std::vector<int32_t> sa(n);
    sais<uint8_t>(text, sa.data(), (int32_t)n, 16);
    std::vector<uint8_t> bwt(n);
    parallel_for(n, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) bwt[i] = text[sa[i] == 0 ? n - 1 : sa[i] - 1];
    });
    // sampled SA: SA % rate == 0 or a sequence start
    std::vector<uint8_t> forced((n + 7) / 8, 0);
    forced[0] |= 1;
    for (uint32_t i = 0; i + 1 < nseq; i++) forced[bounds[i] / 8] |= (uint8_t)(1u << (bounds[i] % 8));
    const uint64_t nw = (n + 63) / 64;
    std::vector<uint64_t> bv(nw, 0);
    for (uint64_t i = 0; i < n; i++) {
        const uint32_t v = (uint32_t)sa[i];
        if (v % rate == 0 || ((forced[v / 8] >> (v % 8)) & 1)) {
            bv[i / 64] |= 1ull << (i % 64);
            ix.sa.vals.push_back(v);
        }
    }
    ix.sa.words.assign(nw * 14, 0);
    uint32_t cum = 0, base = 0;
    for (uint64_t w = 0; w < nw; w++) {
        if (w % 8 == 0) base = cum;
        const uint16_t d = (uint16_t)(cum - base);
        uint8_t *rec = ix.sa.words.data() + w * 14;
        memcpy(rec, &bv[w], 8);
        memcpy(rec + 8, &base, 4);
        memcpy(rec + 12, &d, 2);
        cum += (uint32_t)__builtin_popcountll(bv[w]);
    }
    ix.sa.rate = rate;
    ix.sa.n = (uint32_t)n;
    sa.clear();
    sa.shrink_to_fit();
    progress("BWT and SA samples written", (int64_t)n, 0);
    build_occ(bwt.data(), (uint32_t)n, enc, ix.occ);
    if (depth) build_lookup(ix, depth, nthreads);
    return true;
}

struct Out {
    std::vector<uint8_t> b;
    void put(const void *p, size_t k) { const uint8_t *q = (const uint8_t *)p; b.insert(b.end(), q, q + k); }
    void u8(uint8_t v) { put(&v, 1); }
    void u32(uint32_t v) { put(&v, 4); }
    void u64(uint64_t v) { put(&v, 8); }
};

void serialize(const Index &ix, Out &o) {
    o.put(ix.c, 64);
    o.u8(ix.occ.L);
    o.u8(ix.occ.P);
    o.put(ix.occ.s2l, 16);
    o.put(ix.occ.l2s, 16);
    o.u64(ix.occ.sb.size());
    o.put(ix.occ.sb.data(), ix.occ.sb.size() * 4);
    o.u64(ix.occ.blk.size());
    o.put(ix.occ.blk.data(), ix.occ.blk.size());
    o.u64(ix.occ.stride);
    o.u32(ix.occ.enc);
    o.u32(ix.occ.n);
    o.u64(ix.sa.words.size());
    o.put(ix.sa.words.data(), ix.sa.words.size());
    o.u64(ix.sa.vals.size());
    o.put(ix.sa.vals.data(), ix.sa.vals.size() * 4);
    o.u32(ix.sa.rate);
    o.u32(ix.sa.n);
    o.u32(ix.n);
    o.u32(ix.nseq);
    o.u64(ix.bounds.size());
    o.put(ix.bounds.data(), ix.bounds.size() * 4);
    o.u64(ix.headers.size());
    for (const auto &h : ix.headers) { o.u64(h.size()); o.put(h.data(), h.size()); }
    if (ix.has_lookup) {
        o.u8(1);
        o.u32(ix.depth);
        o.u64(ix.core.size());
        o.put(ix.core.data(), ix.core.size());
        o.u64(ix.lk.size() / 2);
        o.put(ix.lk.data(), ix.lk.size() * 4);
    } else {
        o.u8(0);
    }
    o.u8(ix.tag);
}

uint8_t *to_malloc(const std::vector<uint8_t> &v) {
    uint8_t *p = (uint8_t *)malloc(v.size() ? v.size() : 1);
    if (p && !v.empty()) memcpy(p, v.data(), v.size());
    return p;
}

// ---------------------------------------------------------------------------
// workload synthesis (SplitMix64 streams)
// ---------------------------------------------------------------------------
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t k) { return k ? next() % k : 0; }
};
inline uint8_t gc_base(Rng &r, double gc) {
    const double u = r.unit();
    if (u < gc / 2) return 2;
    if (u < gc) return 3;
    if (u < gc + (1 - gc) / 2) return 1;
    return 4;
}
inline uint8_t other_base(Rng &r, uint8_t b) {
    if (b < 1 || b > 4) return (uint8_t)(1 + r.below(4));
    uint8_t o = (uint8_t)(1 + r.below(3));
    return o >= b ? o + 1 : o;
}
// IUPAC complement: A<->T, C<->G, R<->Y, K<->M, B<->V, D<->H; N, S, W fixed
const uint8_t kComp[16] = {0, 4, 3, 2, 1, 5, 7, 6, 8, 9, 11, 10, 15, 14, 13, 12};

} // namespace

extern "C" int hsfm_build_index(const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq,
                                const char *headers, uint32_t sample_rate, uint32_t lookup_depth, uint32_t encoding,
                                uint8_t tag, int nthreads, uint8_t **out, uint64_t *out_len) {
    const char *pv = getenv("HSFM_BUILD_PROGRESS");
    g_progress = pv && *pv && *pv != '0';
    if (!text || !bounds || !out || !out_len) return -1;
    Index ix;
    if (!build(text, n, bounds, nseq, headers, sample_rate, lookup_depth, encoding, tag, nthreads, ix)) return -1;
    Out o;
    serialize(ix, o);
    *out = to_malloc(o.b);
    *out_len = o.b.size();
    return *out ? 0 : -4;
}

extern "C" int hsfm_build_bidir(const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq,
                                const char *headers, uint32_t sample_rate, uint32_t lookup_depth, uint32_t encoding,
                                uint8_t tag, int nthreads, uint8_t **out, uint64_t *out_len) {
    const char *pv = getenv("HSFM_BUILD_PROGRESS");
    g_progress = pv && *pv && *pv != '0';
    if (!text || !bounds || !out || !out_len || n == 0) return -1;
    // reverse text: strip the trailing sentinel, reverse, one sentinel at the end
    const uint64_t m = text[n - 1] == 0 ? n - 1 : n;
    std::vector<uint8_t> rt(m + 1);
    for (uint64_t i = 0; i < m; i++) rt[i] = text[m - 1 - i];
    rt[m] = 0;
    const uint32_t rb = (uint32_t)(m + 1);
    Index fwd, rev;
    bool ok_f = false, ok_r = false;
    const int half = nthreads > 1 ? nthreads / 2 : 1;
    std::thread t([&] { ok_r = build(rt.data(), m + 1, &rb, 1, nullptr, sample_rate, 0, encoding, tag, half, rev); });
    ok_f = build(text, n, bounds, nseq, headers, sample_rate, lookup_depth, encoding, tag, nthreads - half, fwd);
    t.join();
    if (!ok_f || !ok_r) return -1;
    Out of, orv;
    serialize(fwd, of);
    serialize(rev, orv);
    std::vector<uint8_t> all;
    all.reserve(4 + of.b.size() + orv.b.size());
    const uint32_t fl = (uint32_t)of.b.size();
    all.insert(all.end(), (const uint8_t *)&fl, (const uint8_t *)&fl + 4);
    all.insert(all.end(), of.b.begin(), of.b.end());
    all.insert(all.end(), orv.b.begin(), orv.b.end());
    *out = to_malloc(all);
    *out_len = all.size();
    return *out ? 0 : -4;
}

extern "C" void hsfm_build_free(void *p) { free(p); }

extern "C" int hsfm_synth_reference(uint64_t seed, uint64_t total, uint32_t ncontig, double gc, double repeat_frac,
                                    uint32_t n_units, uint32_t unit_min, uint32_t unit_max, double div_min,
                                    double div_max, double iupac_frac, const uint8_t *iupac_codes, uint32_t n_iupac,
                                    uint8_t **text_out, uint64_t *n_out, uint32_t **bounds_out) {
    if (ncontig == 0 || total < 2ull * ncontig || !text_out || !n_out || !bounds_out) return -1;
    Rng r{seed};
    std::vector<uint64_t> cuts;
    for (;;) {
        cuts.clear();
        for (uint32_t i = 0; i + 1 < ncontig; i++) cuts.push_back(1 + r.below(total - 1));
        std::sort(cuts.begin(), cuts.end());
        cuts.push_back(total);
        bool ok = true;
        for (size_t i = 0; i < cuts.size(); i++) ok &= cuts[i] > (i ? cuts[i - 1] : 0);
        if (ok) break;
    }
    std::vector<uint8_t> seq(total);
    for (auto &b : seq) b = gc_base(r, gc);
    if (repeat_frac > 0 && n_units) {
        std::vector<std::vector<uint8_t>> units(n_units);
        for (auto &u : units) {
            u.resize(unit_min + r.below(unit_max - unit_min + 1));
            for (auto &b : u) b = gc_base(r, gc);
        }
        const uint64_t target = (uint64_t)(repeat_frac * (double)total);
        for (uint64_t covered = 0; covered < target;) {
            const auto &u = units[r.below(n_units)];
            if (u.size() >= total) break;
            const uint64_t at = r.below(total - u.size());
            const double d = div_min + (div_max - div_min) * r.unit();
            for (size_t i = 0; i < u.size(); i++) seq[at + i] = r.unit() < d ? other_base(r, u[i]) : u[i];
            covered += u.size();
        }
    }
    if (iupac_frac > 0 && n_iupac)
        for (auto &b : seq)
            if (r.unit() < iupac_frac) b = iupac_codes[r.below(n_iupac)];
    uint8_t *text = (uint8_t *)malloc(total + ncontig);
    uint32_t *bnd = (uint32_t *)malloc(ncontig * 4ull);
    uint64_t w = 0, prev = 0;
    for (uint32_t c = 0; c < ncontig; c++) {
        memcpy(text + w, seq.data() + prev, cuts[c] - prev);
        w += cuts[c] - prev;
        text[w++] = 0;
        bnd[c] = (uint32_t)w;
        prev = cuts[c];
    }
    *text_out = text;
    *n_out = w;
    *bounds_out = bnd;
    return 0;
}

extern "C" int hsfm_synth_reads(uint64_t seed, const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq,
                                uint64_t nreads, uint32_t len, double sigma, uint32_t len_min, uint32_t len_max,
                                double sub, double indel, int both_strands, uint64_t n_random, double gc, int shuffle,
                                uint8_t **flat_out, uint64_t **off_out) {
    (void)n;
    if (!text || !bounds || nseq == 0 || !flat_out || !off_out || len == 0) return -1;
    Rng r{seed};
    const uint64_t total = nreads + n_random;
    std::vector<uint32_t> lens(total);
    for (auto &L : lens) {
        L = len;
        if (sigma > 0) {
            double u1 = std::max(r.unit(), 1e-300), u2 = r.unit();
            double v = std::exp(std::log((double)len) + sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            v = std::min(std::max(v, (double)len_min), (double)len_max);
            L = (uint32_t)(v + 0.5);
        }
    }
    std::vector<uint64_t> order(total);
    std::iota(order.begin(), order.end(), 0);
    if (shuffle)
        for (uint64_t i = total; i > 1; i--) std::swap(order[i - 1], order[r.below(i)]);
    uint64_t *off = (uint64_t *)malloc((total + 1) * 8);
    off[0] = 0;
    for (uint64_t i = 0; i < total; i++) off[i + 1] = off[i] + lens[order[i]];
    uint8_t *flat = (uint8_t *)malloc(off[total] ? off[total] : 1);
    std::vector<uint64_t> cs(nseq + 1, 0);
    for (uint32_t c = 0; c < nseq; c++) cs[c + 1] = bounds[c];
    const uint64_t span = bounds[nseq - 1];
    std::vector<uint8_t> win;
    for (uint64_t i = 0; i < total; i++) {
        const uint64_t src = order[i];
        const uint32_t L = lens[src];
        uint8_t *dst = flat + off[i];
        if (src >= nreads) {
            for (uint32_t k = 0; k < L; k++) dst[k] = gc_base(r, gc);
            continue;
        }
        const uint64_t need = (uint64_t)L + (indel > 0 ? 8 : 0);
        uint32_t c = 0;
        for (int tries = 0; tries < 1000; tries++) {
            const uint64_t p = r.below(span);
            c = (uint32_t)(std::upper_bound(cs.begin() + 1, cs.end(), p) - cs.begin() - 1);
            if (cs[c + 1] - cs[c] - 1 >= need) break;
        }
        const uint64_t clen = cs[c + 1] - cs[c] - 1, lim = cs[c] + clen;
        const uint64_t at = cs[c] + (clen > need ? r.below(clen - need + 1) : 0);
        win.assign(L, 0);
        uint32_t k = 0;
        for (uint64_t p = at; k < L && p < lim;) {
            if (indel > 0 && r.unit() < indel) {
                const uint32_t il = 1 + (uint32_t)r.below(3);
                if (r.next() & 1) for (uint32_t t = 0; t < il && k < L; t++) win[k++] = gc_base(r, gc);
                else p += il;
                continue;
            }
            win[k++] = text[p++];
        }
        while (k < L) win[k++] = gc_base(r, gc);
        const bool rc = both_strands && (r.next() & 1);
        for (uint32_t j = 0; j < L; j++) {
            uint8_t b = rc ? kComp[win[L - 1 - j] & 15] : win[j];
            if (sub > 0 && r.unit() < sub) b = other_base(r, b);
            dst[j] = b;
        }
    }
    *flat_out = flat;
    *off_out = off;
    return 0;
}
