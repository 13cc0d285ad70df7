/*
 * hsfm_oracle.c -- CPU restatement of haystackfm's batched query path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the haystackfm crate root).  Construction is restated too, so the oracle can
 * emit FmIndex::to_bytes streams for the product's from_bytes boundary.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   compat table ........ golden 16x16 table (alphabet.rs:546-610) + base-set closed form
 *   C array ............. golden arrays (c_array.rs:76-119)
 *   SA .................. SA-IS == comparison sort (suffix_array/cpu.rs:20-25)
 *   rank / lf_step ...... naive_rank over every (c,i) (occ/cpu.rs:143-145)
 *   count / locate ...... known answers (tests/common/mod.rs:47-66, query.rs:505-646)
 *                         + IUPAC brute force + sampling-rate independence
 *   lookup seeding ...... closed form: exact ACGT on the last k, IUPAC on the rest
 *   find_mems/smems ..... known answers (smem.rs:405-728) + IUPAC brute force
 *                         + smems == containment-maximal(mems) (smem.rs:805-870)
 *   to_bytes ............ round trip + field-order layout + tag-is-last-byte
 *                         (byte layout vs. real bincode: parity unpinned)
 */
#define _GNU_SOURCE
#include "hsfm_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ALPHA 16
#define BLOCK 64u
#define SUPERBLOCK 512u
#define SA_REC 14u
#define SA_SB_WORDS 8u

/* ------------------------------------------------------------------------ */
/* small growable arrays                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t *a; size_t n, cap; } vu32;
static void vu32_reserve(vu32 *v, size_t cap) {
    if (cap > v->cap) {
        size_t c = v->cap ? v->cap : 16;
        while (c < cap) c *= 2;
        v->a = (uint32_t *)realloc(v->a, c * sizeof(uint32_t));
        v->cap = c;
    }
}
static void vu32_push(vu32 *v, uint32_t x) { vu32_reserve(v, v->n + 1); v->a[v->n++] = x; }

typedef struct { or_biv *a; size_t n, cap; } vbiv;
static void vbiv_push(vbiv *v, or_biv x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 8;
        v->a = (or_biv *)realloc(v->a, v->cap * sizeof(or_biv));
    }
    v->a[v->n++] = x;
}
static vbiv vbiv_clone(const vbiv *s) {
    vbiv d = {0, 0, 0};
    if (s->n) {
        d.a = (or_biv *)malloc(s->n * sizeof(or_biv));
        memcpy(d.a, s->a, s->n * sizeof(or_biv));
        d.n = d.cap = s->n;
    }
    return d;
}

/* ------------------------------------------------------------------------ */
/* instrumentation                                                           */
/* ------------------------------------------------------------------------ */
static __thread or_probe_counts g_probe;
static int g_probe_on = 0;
void or_probe_reset(void) { memset(&g_probe, 0, sizeof g_probe); }
void or_probe_get(or_probe_counts *out) { *out = g_probe; }
void or_probe_enable(int on) { g_probe_on = on; }
static inline void probe_pair(uint32_t lo, uint32_t hi) {
    if (!g_probe_on) return;
    uint32_t bl = lo ? (lo - 1) / BLOCK : 0xffffffffu;
    uint32_t bh = hi ? (hi - 1) / BLOCK : 0xffffffffu;
    g_probe.occ_probes += (lo != 0) + (hi != 0 && bh != bl);
    g_probe.steps++;
}

/* ------------------------------------------------------------------------ */
/* alphabet (src/alphabet.rs)                                                */
/* ------------------------------------------------------------------------ */
enum { SENT = 0, A_ = 1, C_ = 2, G_ = 3, T_ = 4, N_ = 5, R_ = 6, Y_ = 7, S_ = 8,
       W_ = 9, K_ = 10, M_ = 11, B_ = 12, D_ = 13, H_ = 14, V_ = 15 };

/* alphabet.rs:54-109: ASCII -> code, case-insensitive, U->T */
int or_encode_byte(uint8_t b) {
    static const char up[] = "$ACGTNRYSWKMBDHV";
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    if (b == 'U') return T_;
    for (int i = 0; i < 16; i++)
        if ((uint8_t)up[i] == b) return i;
    return -1;
}

/* alphabet.rs:174-193, lists in the reference's order */
static const uint8_t COMPAT[16][16] = {
    {0},
    {A_, N_, R_, W_, M_, D_, H_, V_},
    {C_, N_, Y_, S_, M_, B_, H_, V_},
    {G_, N_, R_, S_, K_, B_, D_, V_},
    {T_, N_, Y_, W_, K_, B_, D_, H_},
    {A_, C_, G_, T_, N_, R_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
    {A_, G_, N_, R_, S_, W_, K_, M_, B_, D_, H_, V_},
    {C_, T_, N_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
    {C_, G_, N_, R_, Y_, S_, K_, M_, B_, D_, H_, V_},
    {A_, T_, N_, R_, Y_, W_, K_, M_, B_, D_, H_, V_},
    {G_, T_, N_, R_, Y_, S_, W_, K_, B_, D_, H_, V_},
    {A_, C_, N_, R_, Y_, S_, W_, M_, B_, D_, H_, V_},
    {C_, G_, T_, N_, R_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
    {A_, G_, T_, N_, R_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
    {A_, C_, T_, N_, R_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
    {A_, C_, G_, N_, R_, Y_, S_, W_, K_, M_, B_, D_, H_, V_},
};
static const uint8_t COMPAT_LEN[16] = {0, 8, 8, 8, 8, 15, 12, 12, 12, 12, 12, 12, 14, 14, 14, 14};

/* compatible_fn of the alphabet with tag (alphabet_fns_from_tag,
 * alphabet.rs:240-254): tag 0 = IupacDna, tag 1 = ExactDna (:286-294). */
static inline uint32_t compat_list(uint8_t tag, uint8_t code, const uint8_t **out) {
    static const uint8_t EX[5][1] = {{0}, {A_}, {C_}, {G_}, {T_}};
    if (tag == 1) {
        if (code >= A_ && code <= T_) { *out = EX[code]; return 1; }
        *out = EX[0];
        return 0;
    }
    if (code >= 16) { *out = COMPAT[0]; return 0; }
    *out = COMPAT[code];
    return COMPAT_LEN[code];
}
uint32_t or_compatible(uint8_t tag, uint8_t code, uint8_t out[16]) {
    const uint8_t *l;
    uint32_t k = compat_list(tag, code, &l);
    memcpy(out, l, k);
    return k;
}
uint16_t or_compat_mask(uint8_t tag, uint8_t code) {
    const uint8_t *l;
    uint32_t k = compat_list(tag, code, &l);
    uint16_t m = 0;
    for (uint32_t i = 0; i < k; i++) m |= (uint16_t)(1u << l[i]);
    return m;
}
/* alphabet.rs:149-168 */
uint32_t or_iupac_bases(uint8_t code, uint8_t out[4]) {
    static const uint8_t BS[16] = {0, 1, 2, 4, 8, 15, 5, 10, 6, 9, 12, 3, 14, 13, 11, 7};
    /* bit 0=A 1=C 2=G 3=T */
    if (code >= 16) return 0;
    uint32_t k = 0;
    for (int b = 0; b < 4; b++)
        if (BS[code] >> b & 1) out[k++] = (uint8_t)(A_ + b);
    return k;
}

/* ------------------------------------------------------------------------ */
/* suffix array: SA-IS (Nong, Zhang & Chan 2009) with a virtual terminator   */
/* smaller than every symbol, which yields the plain lexicographic order of  */
/* suffix_array/cpu.rs:20-25 (internal 0 sentinels compare as ordinary 0s).  */
/* ------------------------------------------------------------------------ */
static void sais_i32(const int32_t *T, int32_t *SA, int32_t n, int32_t K);

#define SAIS_BODY(CT)                                                                     \
    if (n <= 0) return;                                                                   \
    if (n == 1) { SA[0] = 0; return; }                                                     \
    uint8_t *t = (uint8_t *)malloc((size_t)n);                                            \
    int32_t *Cn = (int32_t *)calloc((size_t)K, sizeof(int32_t));                           \
    int32_t *Bk = (int32_t *)malloc((size_t)K * sizeof(int32_t));                          \
    t[n - 1] = 0;                                                                         \
    for (int32_t i = n - 2; i >= 0; i--)                                                  \
        t[i] = (T[i] < T[i + 1] || (T[i] == T[i + 1] && t[i + 1])) ? 1 : 0;               \
    for (int32_t i = 0; i < n; i++) Cn[T[i]]++;                                           \
    /* bucket ends */                                                                     \
    {                                                                                     \
        int64_t s = 0;                                                                    \
        for (int32_t k = 0; k < K; k++) { s += Cn[k]; Bk[k] = (int32_t)s; }               \
    }                                                                                     \
    for (int32_t i = 0; i < n; i++) SA[i] = -1;                                           \
    for (int32_t i = 1; i < n; i++)                                                       \
        if (t[i] && !t[i - 1]) SA[--Bk[T[i]]] = i;                                        \
    INDUCE(CT)                                                                            \
    int32_t n1 = 0;                                                                       \
    for (int32_t i = 0; i < n; i++) {                                                     \
        int32_t p = SA[i];                                                                \
        if (p > 0 && t[p] && !t[p - 1]) SA[n1++] = p;                                     \
    }                                                                                     \
    for (int32_t i = n1; i < n; i++) SA[i] = -1;                                          \
    int32_t name = 0, prev = -1;                                                          \
    for (int32_t i = 0; i < n1; i++) {                                                    \
        int32_t pos = SA[i];                                                              \
        int diff = 0;                                                                     \
        if (prev < 0) diff = 1;                                                           \
        else {                                                                            \
            for (int32_t d = 0;; d++) {                                                   \
                if (pos + d == n || prev + d == n || T[pos + d] != T[prev + d] ||         \
                    t[pos + d] != t[prev + d]) { diff = 1; break; }                       \
                if (d > 0) {                                                              \
                    int la = t[pos + d] && !t[pos + d - 1];                               \
                    int lb = t[prev + d] && !t[prev + d - 1];                             \
                    if (la || lb) { diff = !(la && lb); break; }                          \
                }                                                                         \
            }                                                                             \
        }                                                                                 \
        if (diff) { name++; prev = pos; }                                                 \
        SA[n1 + pos / 2] = name - 1;                                                      \
    }                                                                                     \
    for (int32_t i = n - 1, j = n - 1; i >= n1; i--)                                      \
        if (SA[i] >= 0) SA[j--] = SA[i];                                                  \
    int32_t *s1 = SA + n - n1;                                                            \
    if (name < n1) sais_i32(s1, SA, n1, name);                                            \
    else for (int32_t i = 0; i < n1; i++) SA[s1[i]] = i;                                  \
    {                                                                                     \
        int32_t j = 0;                                                                    \
        for (int32_t i = 1; i < n; i++)                                                   \
            if (t[i] && !t[i - 1]) s1[j++] = i;                                           \
    }                                                                                     \
    for (int32_t i = 0; i < n1; i++) SA[i] = s1[SA[i]];                                   \
    for (int32_t i = n1; i < n; i++) SA[i] = -1;                                          \
    {                                                                                     \
        int64_t s = 0;                                                                    \
        for (int32_t k = 0; k < K; k++) { s += Cn[k]; Bk[k] = (int32_t)s; }               \
    }                                                                                     \
    for (int32_t i = n1 - 1; i >= 0; i--) {                                               \
        int32_t j = SA[i];                                                                \
        SA[i] = -1;                                                                       \
        SA[--Bk[T[j]]] = j;                                                               \
    }                                                                                     \
    INDUCE(CT)                                                                            \
    free(t); free(Cn); free(Bk);

#define INDUCE(CT)                                                                        \
    {                                                                                     \
        int64_t s = 0;                                                                    \
        for (int32_t k = 0; k < K; k++) { Bk[k] = (int32_t)s; s += Cn[k]; }               \
        SA[Bk[T[n - 1]]++] = n - 1;                                                       \
        for (int32_t i = 0; i < n; i++) {                                                 \
            int32_t j = SA[i] - 1;                                                        \
            if (SA[i] > 0 && !t[j]) SA[Bk[T[j]]++] = j;                                   \
        }                                                                                 \
        s = 0;                                                                            \
        for (int32_t k = 0; k < K; k++) { s += Cn[k]; Bk[k] = (int32_t)s; }               \
        for (int32_t i = n - 1; i >= 0; i--) {                                            \
            int32_t j = SA[i] - 1;                                                        \
            if (SA[i] > 0 && t[j]) SA[--Bk[T[j]]] = j;                                    \
        }                                                                                 \
    }

static void sais_i32(const int32_t *T, int32_t *SA, int32_t n, int32_t K) { SAIS_BODY(int32_t) }
static void sais_u8(const uint8_t *T, int32_t *SA, int32_t n, int32_t K) { SAIS_BODY(uint8_t) }

static const uint8_t *g_sort_text;
static uint64_t g_sort_n;
static int cmp_suffix(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    uint64_t lx = g_sort_n - x, ly = g_sort_n - y, l = lx < ly ? lx : ly;
    int c = memcmp(g_sort_text + x, g_sort_text + y, l);
    if (c) return c;
    return lx < ly ? -1 : (lx > ly ? 1 : 0);
}

/* naive (suffix_array/cpu.rs:20-25) or SA-IS; identical order */
int or_suffix_array(const uint8_t *text, uint64_t n, uint32_t *sa, int naive) {
    if (n == 0) return 0;
    if (n > 0x7fffffffULL) return -1;
    if (naive) {
        for (uint64_t i = 0; i < n; i++) sa[i] = (uint32_t)i;
        g_sort_text = text;
        g_sort_n = n;
        qsort(sa, n, sizeof(uint32_t), cmp_suffix);
        return 0;
    }
    sais_u8(text, (int32_t *)sa, (int32_t)n, ALPHA);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* index structures (reference layouts)                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint8_t num_lanes, num_planes;
    uint8_t s2l[ALPHA], l2s[ALPHA];
    uint32_t *sb; uint64_t sb_len;
    uint8_t *blk; uint64_t blk_len;
    uint64_t stride;
    uint32_t encoding;
    uint32_t text_len;
} occ_t;

typedef struct {
    uint8_t *words; uint64_t words_len;
    uint32_t *vals; uint64_t vals_len;
    uint32_t rate, text_len;
} ssa_t;

struct or_fmi {
    uint32_t c[ALPHA];
    occ_t occ;
    ssa_t sa;
    uint32_t text_len, num_sequences;
    uint32_t *bounds; uint64_t nbounds;
    char **headers; uint64_t nheaders;
    int has_lookup;
    uint32_t lk_depth;
    uint8_t *lk_core; uint64_t lk_ncore;
    uint32_t *lk_iv; uint64_t lk_n; /* pairs */
    uint8_t tag;
};

static inline uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint16_t rd16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }
static inline uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

/* c_array.rs:14-31 */
static void carray_from_text(const uint8_t *t, uint64_t n, uint32_t c[ALPHA]) {
    uint64_t f[ALPHA] = {0};
    for (uint64_t i = 0; i < n; i++)
        if (t[i] < ALPHA) f[t[i]]++;
    uint32_t s = 0;
    for (int i = 0; i < ALPHA; i++) { c[i] = s; s += (uint32_t)f[i]; }
}
/* c_array.rs:59-67 */
static inline uint32_t symbol_count(const or_fmi *f, uint8_t c) {
    uint32_t up = (c + 1 < ALPHA) ? f->c[c + 1] : f->text_len;
    return up - f->c[c];
}

/* occ/mod.rs:100-106 */
static uint8_t num_planes_for(uint8_t L) {
    if (L <= 1) return 0;
    uint8_t p = 0;
    while ((1u << p) < L) p++;
    return p;
}

/* occ/cpu.rs:23-119 + occ/mod.rs:114-183 (interleaved record layout) */
static void build_occ(const uint8_t *bwt, uint32_t n, uint32_t encoding, occ_t *o) {
    uint64_t nb = (n + BLOCK - 1) / BLOCK;
    uint64_t tot[ALPHA] = {0};
    for (uint32_t i = 0; i < n; i++)
        if (bwt[i] < ALPHA) tot[bwt[i]]++;
    memset(o, 0, sizeof *o);
    memset(o->s2l, 0xff, ALPHA);
    uint8_t L = 0;
    for (int c = 0; c < ALPHA; c++)
        if (tot[c]) o->s2l[c] = L++;
    for (int c = 0; c < ALPHA; c++)
        if (o->s2l[c] != 0xff) o->l2s[o->s2l[c]] = (uint8_t)c;
    o->num_lanes = L;
    o->num_planes = num_planes_for(L);
    uint32_t W = encoding == 0 ? o->num_planes : L;
    o->stride = (uint64_t)L * 4 + (uint64_t)L * 2 + (uint64_t)W * 8;
    o->encoding = encoding;
    o->text_len = n;
    o->blk_len = nb * o->stride;
    o->blk = (uint8_t *)calloc(o->blk_len ? o->blk_len : 1, 1);
    uint64_t nsb = (nb + 7) / 8;
    o->sb_len = nsb * L;
    o->sb = (uint32_t *)calloc(o->sb_len ? o->sb_len : 1, sizeof(uint32_t));
    uint32_t cum[ALPHA] = {0}, sbb[ALPHA] = {0};
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t bits[ALPHA] = {0};
        uint32_t cnt[ALPHA] = {0};
        uint32_t st = (uint32_t)(b * BLOCK), en = st + BLOCK < n ? st + BLOCK : n;
        for (uint32_t p = st; p < en; p++) {
            uint8_t ch = bwt[p];
            if (ch < ALPHA) { bits[ch] |= 1ULL << (p - st); cnt[ch]++; }
        }
        if (b % 8 == 0) {
            for (int l = 0; l < L; l++) { o->sb[(b / 8) * L + l] = cum[l]; sbb[l] = cum[l]; }
        }
        uint8_t *rec = o->blk + b * o->stride;
        uint64_t lane_data[ALPHA] = {0};
        for (int c = 0; c < ALPHA; c++) {
            uint8_t lane = o->s2l[c];
            if (lane == 0xff) continue;
            uint16_t d = (uint16_t)(cum[lane] - sbb[lane]);
            memcpy(rec + (uint64_t)L * 4 + lane * 2, &d, 2);
            if (encoding == 0) {
                for (uint32_t p = 0; p < W; p++)
                    if ((lane >> p) & 1) lane_data[p] |= bits[c];
            } else {
                lane_data[lane] = bits[c];
            }
            cum[lane] += cnt[c];
        }
        for (int l = 0; l < L; l++) memcpy(rec + l * 4, &o->sb[(b / 8) * L + l], 4);
        for (uint32_t w = 0; w < W; w++)
            memcpy(rec + (uint64_t)L * 6 + w * 8, &lane_data[w], 8);
    }
}

/* occ/mod.rs:276-298 */
static inline uint64_t lane_mask(const occ_t *o, const uint8_t *rec, uint32_t lane) {
    if (o->encoding == 1) {
        if (o->num_lanes == 0) return 0;
        return rd64(rec + (uint64_t)o->num_lanes * 6 + lane * 8);
    }
    if (o->num_planes == 0) return ~0ULL;
    uint64_t m = ~0ULL;
    for (uint32_t p = 0; p < o->num_planes; p++) {
        uint64_t pv = rd64(rec + (uint64_t)o->num_lanes * 6 + p * 8);
        m &= ((lane >> p) & 1) ? pv : ~pv;
    }
    return m;
}
/* occ/mod.rs:305-325 */
static inline uint32_t lane_at(const occ_t *o, const uint8_t *rec, uint32_t off) {
    if (o->encoding == 1) {
        for (uint32_t l = 0; l < o->num_lanes; l++)
            if ((rd64(rec + (uint64_t)o->num_lanes * 6 + l * 8) >> off) & 1) return l;
        return 0;
    }
    uint32_t lane = 0;
    for (uint32_t p = 0; p < o->num_planes; p++)
        lane |= (uint32_t)((rd64(rec + (uint64_t)o->num_lanes * 6 + p * 8) >> off) & 1) << p;
    return lane;
}
/* occ/mod.rs:346-363 */
static inline uint32_t rank_at_lane(const occ_t *o, uint32_t lane, uint32_t i) {
    uint32_t pos = i - 1, b = pos / BLOCK, off = pos % BLOCK;
    const uint8_t *rec = o->blk + (uint64_t)b * o->stride;
    uint32_t sbc = rd32(rec + lane * 4);
    uint32_t d = rd16(rec + (uint64_t)o->num_lanes * 4 + lane * 2);
    uint64_t bits = lane_mask(o, rec, lane);
    uint64_t mask = off == 63 ? ~0ULL : ((1ULL << (off + 1)) - 1);
    return sbc + d + (uint32_t)__builtin_popcountll(bits & mask);
}
/* occ/mod.rs:330-339 */
static inline uint32_t occ_rank(const occ_t *o, uint8_t c, uint32_t i) {
    if (i == 0) return 0;
    if (c >= ALPHA) return 0;
    uint8_t lane = o->s2l[c];
    if (lane == 0xff) return 0;
    return rank_at_lane(o, lane, i);
}
uint32_t or_rank(const or_fmi *f, uint8_t c, uint32_t i) { return occ_rank(&f->occ, c, i); }

/* occ/mod.rs:456-480 */
static inline void occ_lf_step(const occ_t *o, uint32_t pos, uint8_t *sym, uint32_t *rank) {
    uint32_t b = pos / BLOCK, off = pos % BLOCK;
    const uint8_t *rec = o->blk + (uint64_t)b * o->stride;
    uint32_t lane = lane_at(o, rec, off);
    uint64_t bits = lane_mask(o, rec, lane);
    uint32_t sbc = rd32(rec + lane * 4);
    uint32_t d = rd16(rec + (uint64_t)o->num_lanes * 4 + lane * 2);
    uint64_t mask = (1ULL << off) - 1;
    *sym = o->l2s[lane];
    *rank = sbc + d + (uint32_t)__builtin_popcountll(bits & mask);
}
void or_lf_step(const or_fmi *f, uint32_t pos, uint8_t *sym, uint32_t *rank) {
    occ_lf_step(&f->occ, pos, sym, rank);
}
uint8_t or_symbol_at(const or_fmi *f, uint32_t pos) {
    const occ_t *o = &f->occ;
    const uint8_t *rec = o->blk + (uint64_t)(pos / BLOCK) * o->stride;
    return o->l2s[lane_at(o, rec, pos % BLOCK)];
}

/* suffix_array/mod.rs:71-118 */
static void build_ssa(const uint32_t *sa, uint32_t n, uint32_t rate, const uint32_t *forced,
                      uint32_t nforced, ssa_t *s) {
    uint64_t nw = (n + 63) / 64;
    uint64_t *bv = (uint64_t *)calloc(nw ? nw : 1, 8);
    /* forced membership via a bitmap over text positions */
    uint64_t *fb = (uint64_t *)calloc((n + 63) / 64 + 1, 8);
    for (uint32_t k = 0; k < nforced; k++)
        if (forced[k] < n) fb[forced[k] / 64] |= 1ULL << (forced[k] % 64);
    uint64_t cnt = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t v = sa[i];
        if (v % rate == 0 || ((fb[v / 64] >> (v % 64)) & 1)) { bv[i / 64] |= 1ULL << (i % 64); cnt++; }
    }
    s->vals = (uint32_t *)malloc((cnt ? cnt : 1) * 4);
    s->vals_len = cnt;
    uint64_t k = 0;
    for (uint32_t i = 0; i < n; i++)
        if ((bv[i / 64] >> (i % 64)) & 1) s->vals[k++] = sa[i];
    s->words_len = nw * SA_REC;
    s->words = (uint8_t *)calloc(s->words_len ? s->words_len : 1, 1);
    uint32_t cum = 0, sbb = 0;
    for (uint64_t w = 0; w < nw; w++) {
        if (w % SA_SB_WORDS == 0) sbb = cum;
        uint16_t d = (uint16_t)(cum - sbb);
        uint8_t *rec = s->words + w * SA_REC;
        memcpy(rec, &bv[w], 8);
        memcpy(rec + 8, &sbb, 4);
        memcpy(rec + 12, &d, 2);
        cum += (uint32_t)__builtin_popcountll(bv[w]);
    }
    s->rate = rate;
    s->text_len = n;
    free(bv);
    free(fb);
}

/* suffix_array/mod.rs:185-201 */
static inline int ssa_get(const ssa_t *s, uint32_t i, uint32_t *val) {
    uint64_t w = i / 64;
    uint32_t bo = i % 64;
    const uint8_t *rec = s->words + w * SA_REC;
    uint64_t word = rd64(rec);
    if (g_probe_on) g_probe.sa_probes++;
    if (!((word >> bo) & 1)) return 0;
    uint64_t mask = bo == 0 ? 0 : ((1ULL << bo) - 1);
    uint32_t r = rd32(rec + 8) + rd16(rec + 12) + (uint32_t)__builtin_popcountll(word & mask);
    if (g_probe_on) g_probe.sa_val_reads++;
    *val = s->vals[r];
    return 1;
}
int or_sa_get(const or_fmi *f, uint32_t row, uint32_t *val) { return ssa_get(&f->sa, row, val); }

/* fm_index/mod.rs:288-291 */
static inline uint32_t lf_mapping(const or_fmi *f, uint32_t i) {
    uint8_t c;
    uint32_t r;
    occ_lf_step(&f->occ, i, &c, &r);
    if (g_probe_on) g_probe.occ_probes++;
    return f->c[c] + r;
}

/* query.rs:282-291 */
uint32_t or_resolve_sa(const or_fmi *f, uint32_t i) {
    uint32_t steps = 0, v;
    for (;;) {
        if (ssa_get(&f->sa, i, &v)) return v + steps;
        i = lf_mapping(f, i);
        steps++;
    }
}

/* query.rs:342-357 */
int or_map_position(const or_fmi *f, uint32_t pos, uint32_t *seq, uint32_t *off) {
    uint64_t lo = 0, hi = f->nbounds;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (f->bounds[mid] <= pos) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= f->nbounds) return 0;
    *seq = (uint32_t)lo;
    *off = pos - (lo == 0 ? 0 : f->bounds[lo - 1]);
    return 1;
}

/* lookup.rs:37-91 (BFS, leftmost code = least significant base-4 digit) */
typedef struct { uint32_t idx, lo, hi; } lkn;
typedef struct { const or_fmi *f; const lkn *cur; uint64_t b, e; int last; uint32_t radix;
                 lkn *outv; uint64_t outn; } lk_job;
static void *lk_worker(void *arg) {
    lk_job *j = (lk_job *)arg;
    const or_fmi *f = j->f;
    j->outn = 0;
    for (uint64_t k = j->b; k < j->e; k++) {
        lkn p = j->cur[k];
        for (uint32_t d = 0; d < j->radix; d++) {
            uint8_t sym = f->lk_core[d];
            uint32_t child = p.idx * j->radix + d;
            if (symbol_count(f, sym) == 0) continue;
            uint32_t cv = f->c[sym];
            uint32_t nl = cv + occ_rank(&f->occ, sym, p.lo), nh = cv + occ_rank(&f->occ, sym, p.hi);
            if (j->last) { f->lk_iv[2 * (uint64_t)child] = nl; f->lk_iv[2 * (uint64_t)child + 1] = nh; }
            else if (nl < nh) { lkn c2 = {child, nl, nh}; j->outv[j->outn++] = c2; }
        }
    }
    return NULL;
}
static void build_lookup(or_fmi *f, uint32_t depth, int nthreads) {
    uint32_t radix = 4;
    f->has_lookup = 1;
    f->lk_depth = depth;
    f->lk_ncore = 4;
    f->lk_core = (uint8_t *)malloc(4);
    for (int i = 0; i < 4; i++) f->lk_core[i] = (uint8_t)(A_ + i);
    uint64_t size = 1;
    for (uint32_t i = 0; i < depth; i++) size *= radix;
    f->lk_n = size;
    f->lk_iv = (uint32_t *)calloc(size * 2, 4);
    lkn *cur = (lkn *)malloc(sizeof(lkn));
    cur[0].idx = 0; cur[0].lo = 0; cur[0].hi = f->text_len;
    uint64_t ncur = 1;
    if (nthreads < 1) nthreads = 1;
    for (uint32_t level = 1; level <= depth; level++) {
        int last = level == depth;
        int nt = (ncur < 4096) ? 1 : nthreads;
        lk_job *jobs = (lk_job *)calloc(nt, sizeof(lk_job));
        pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
        for (int t = 0; t < nt; t++) {
            jobs[t].f = f; jobs[t].cur = cur; jobs[t].radix = radix; jobs[t].last = last;
            jobs[t].b = ncur * t / nt; jobs[t].e = ncur * (t + 1) / nt;
            jobs[t].outv = last ? NULL : (lkn *)malloc(((jobs[t].e - jobs[t].b) * radix + 1) * sizeof(lkn));
            if (nt > 1) pthread_create(&th[t], NULL, lk_worker, &jobs[t]);
            else lk_worker(&jobs[t]);
        }
        if (nt > 1) for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
        if (!last) {
            uint64_t tot = 0;
            for (int t = 0; t < nt; t++) tot += jobs[t].outn;
            lkn *nxt = (lkn *)malloc((tot ? tot : 1) * sizeof(lkn));
            uint64_t k = 0;
            for (int t = 0; t < nt; t++) {
                memcpy(nxt + k, jobs[t].outv, jobs[t].outn * sizeof(lkn));
                k += jobs[t].outn;
                free(jobs[t].outv);
            }
            free(cur);
            cur = nxt;
            ncur = tot;
        }
        free(jobs);
        free(th);
    }
    free(cur);
}

static or_fmi *build_from_text(const uint8_t *text, uint64_t n, const uint32_t *bounds,
                               uint32_t nseq, const char *headers, uint32_t rate,
                               uint32_t lookup_depth, uint32_t encoding, uint8_t tag,
                               int naive, int nthreads) {
    if (n == 0 || n > 0xffffffffULL || rate == 0) return NULL;
    or_fmi *f = (or_fmi *)calloc(1, sizeof(or_fmi));
    f->tag = tag;
    f->text_len = (uint32_t)n;
    f->num_sequences = nseq;
    f->nbounds = nseq;
    f->bounds = (uint32_t *)malloc((nseq ? nseq : 1) * 4);
    memcpy(f->bounds, bounds, nseq * 4);
    f->nheaders = nseq;
    f->headers = (char **)calloc(nseq ? nseq : 1, sizeof(char *));
    const char *hp = headers;
    for (uint32_t i = 0; i < nseq; i++) {
        /* fm_index/mod.rs:126-137: empty header -> "seq_{i}" */
        if (hp && *hp) {
            f->headers[i] = strdup(hp);
            hp += strlen(hp) + 1;
        } else {
            char buf[32];
            snprintf(buf, sizeof buf, "seq_%u", i);
            f->headers[i] = strdup(buf);
            if (hp) hp += 1;
        }
    }
    carray_from_text(text, n, f->c);
    uint32_t *sa = (uint32_t *)malloc(n * 4);
    or_suffix_array(text, n, sa, naive);
    uint8_t *bwt = (uint8_t *)malloc(n);
    for (uint64_t i = 0; i < n; i++) bwt[i] = text[sa[i] == 0 ? n - 1 : sa[i] - 1]; /* bwt/cpu.rs:14-35 */
    /* fm_index/mod.rs:154-162: sequence starts are force-sampled */
    uint32_t *starts = (uint32_t *)malloc((nseq ? nseq : 1) * 4);
    starts[0] = 0;
    for (uint32_t i = 0; i + 1 < nseq; i++) starts[i + 1] = bounds[i];
    build_ssa(sa, (uint32_t)n, rate, starts, nseq ? nseq : 1, &f->sa);
    free(starts);
    free(sa);
    build_occ(bwt, (uint32_t)n, encoding, &f->occ);
    free(bwt);
    if (lookup_depth > 0) build_lookup(f, lookup_depth, nthreads);
    return f;
}

or_fmi *or_build(const uint8_t *text, uint64_t n, const uint32_t *seq_bounds, uint32_t nseq,
                 const char *headers, uint32_t sample_rate, uint32_t lookup_depth,
                 uint32_t encoding, uint8_t tag, int naive_sa, int nthreads) {
    return build_from_text(text, n, seq_bounds, nseq, headers, sample_rate, lookup_depth,
                           encoding, tag, naive_sa, nthreads);
}

/* bidir_index.rs:61-77, 523-532 */
or_fmi *or_build_rev(const uint8_t *text, uint64_t n, uint32_t sample_rate, uint32_t encoding,
                     uint8_t tag, int naive_sa, int nthreads) {
    uint64_t m = (n && text[n - 1] == 0) ? n - 1 : n;
    uint8_t *r = (uint8_t *)malloc(m + 1);
    for (uint64_t i = 0; i < m; i++) r[i] = text[m - 1 - i];
    r[m] = 0;
    uint32_t b = (uint32_t)(m + 1);
    or_fmi *f = build_from_text(r, m + 1, &b, 1, NULL, sample_rate, 0, encoding, tag, naive_sa,
                                nthreads);
    free(r);
    return f;
}

void or_free(or_fmi *f) {
    if (!f) return;
    free(f->occ.sb); free(f->occ.blk);
    free(f->sa.words); free(f->sa.vals);
    free(f->bounds);
    for (uint64_t i = 0; i < f->nheaders; i++) free(f->headers[i]);
    free(f->headers);
    free(f->lk_core); free(f->lk_iv);
    free(f);
}
void or_free_buf(void *p) { free(p); }

uint32_t or_text_len(const or_fmi *f) { return f->text_len; }
uint32_t or_num_lanes(const or_fmi *f) { return f->occ.num_lanes; }
uint32_t or_sample_rate(const or_fmi *f) { return f->sa.rate; }
uint32_t or_num_sequences(const or_fmi *f) { return f->num_sequences; }
const uint32_t *or_seq_boundaries(const or_fmi *f) { return f->bounds; }
uint32_t or_lookup_depth(const or_fmi *f) { return f->has_lookup ? f->lk_depth : 0; }
void or_c_array(const or_fmi *f, uint32_t out[16]) { memcpy(out, f->c, 64); }

/* ------------------------------------------------------------------------ */
/* bincode-1 serialization (fm_index/serialize.rs:13-86): little endian,     */
/* fixed-width ints, u64 sequence lengths, Option = u8 tag, enum = u32.      */
/* ------------------------------------------------------------------------ */
typedef struct { uint8_t *p; uint64_t n, cap; } wbuf;
static void wb_put(wbuf *w, const void *d, uint64_t k) {
    if (w->n + k > w->cap) {
        uint64_t c = w->cap ? w->cap : 1024;
        while (c < w->n + k) c *= 2;
        w->p = (uint8_t *)realloc(w->p, c);
        w->cap = c;
    }
    memcpy(w->p + w->n, d, k);
    w->n += k;
}
static void wb_u8(wbuf *w, uint8_t v) { wb_put(w, &v, 1); }
static void wb_u32(wbuf *w, uint32_t v) { wb_put(w, &v, 4); }
static void wb_u64(wbuf *w, uint64_t v) { wb_put(w, &v, 8); }

int or_to_bytes(const or_fmi *f, uint8_t **out, uint64_t *len) {
    wbuf w = {0, 0, 0};
    for (int i = 0; i < ALPHA; i++) wb_u32(&w, f->c[i]);                 /* CArray{data:[u32;16]} */
    wb_u8(&w, f->occ.num_lanes);                                          /* OccTable */
    wb_u8(&w, f->occ.num_planes);
    wb_put(&w, f->occ.s2l, ALPHA);
    wb_put(&w, f->occ.l2s, ALPHA);
    wb_u64(&w, f->occ.sb_len);
    wb_put(&w, f->occ.sb, f->occ.sb_len * 4);
    wb_u64(&w, f->occ.blk_len);
    wb_put(&w, f->occ.blk, f->occ.blk_len);
    wb_u64(&w, f->occ.stride);
    wb_u32(&w, f->occ.encoding);
    wb_u32(&w, f->occ.text_len);
    wb_u64(&w, f->sa.words_len);                                          /* SampledSuffixArray */
    wb_put(&w, f->sa.words, f->sa.words_len);
    wb_u64(&w, f->sa.vals_len);
    wb_put(&w, f->sa.vals, f->sa.vals_len * 4);
    wb_u32(&w, f->sa.rate);
    wb_u32(&w, f->sa.text_len);
    wb_u32(&w, f->text_len);
    wb_u32(&w, f->num_sequences);
    wb_u64(&w, f->nbounds);
    wb_put(&w, f->bounds, f->nbounds * 4);
    wb_u64(&w, f->nheaders);
    for (uint64_t i = 0; i < f->nheaders; i++) {
        uint64_t l = strlen(f->headers[i]);
        wb_u64(&w, l);
        wb_put(&w, f->headers[i], l);
    }
    if (f->has_lookup) {                                                  /* Option<LookupTable> */
        wb_u8(&w, 1);
        wb_u32(&w, f->lk_depth);
        wb_u64(&w, f->lk_ncore);
        wb_put(&w, f->lk_core, f->lk_ncore);
        wb_u64(&w, f->lk_n);
        wb_put(&w, f->lk_iv, f->lk_n * 8);
    } else {
        wb_u8(&w, 0);
    }
    wb_u8(&w, f->tag);                                                    /* alphabet_tag (last) */
    *out = w.p;
    *len = w.n;
    return 0;
}

typedef struct { const uint8_t *p; uint64_t n, off; int bad; } rbuf;
static const uint8_t *rb_take(rbuf *r, uint64_t k) {
    if (r->bad || k > r->n - r->off) { r->bad = 1; return NULL; }
    const uint8_t *q = r->p + r->off;
    r->off += k;
    return q;
}
static uint8_t rb_u8(rbuf *r) { const uint8_t *q = rb_take(r, 1); return q ? q[0] : 0; }
static uint32_t rb_u32(rbuf *r) { const uint8_t *q = rb_take(r, 4); return q ? rd32(q) : 0; }
static uint64_t rb_u64(rbuf *r) { const uint8_t *q = rb_take(r, 8); return q ? rd64(q) : 0; }
static void *rb_vec(rbuf *r, uint64_t elem, uint64_t *count) {
    uint64_t k = rb_u64(r);
    if (r->bad || (elem && k > (r->n - r->off) / elem)) { r->bad = 1; *count = 0; return NULL; }
    const uint8_t *q = rb_take(r, k * elem);
    uint64_t bytes = k * elem;
    void *d = malloc(bytes ? bytes : 1);
    if (q) memcpy(d, q, k * elem);
    *count = k;
    return d;
}

or_fmi *or_from_bytes(const uint8_t *b, uint64_t len, char *err, int errlen) {
    rbuf r = {b, len, 0, 0};
    or_fmi *f = (or_fmi *)calloc(1, sizeof(or_fmi));
    for (int i = 0; i < ALPHA; i++) f->c[i] = rb_u32(&r);
    f->occ.num_lanes = rb_u8(&r);
    f->occ.num_planes = rb_u8(&r);
    const uint8_t *q = rb_take(&r, ALPHA);
    if (q) memcpy(f->occ.s2l, q, ALPHA);
    q = rb_take(&r, ALPHA);
    if (q) memcpy(f->occ.l2s, q, ALPHA);
    f->occ.sb = (uint32_t *)rb_vec(&r, 4, &f->occ.sb_len);
    f->occ.blk = (uint8_t *)rb_vec(&r, 1, &f->occ.blk_len);
    f->occ.stride = rb_u64(&r);
    f->occ.encoding = rb_u32(&r);
    f->occ.text_len = rb_u32(&r);
    f->sa.words = (uint8_t *)rb_vec(&r, 1, &f->sa.words_len);
    f->sa.vals = (uint32_t *)rb_vec(&r, 4, &f->sa.vals_len);
    f->sa.rate = rb_u32(&r);
    f->sa.text_len = rb_u32(&r);
    f->text_len = rb_u32(&r);
    f->num_sequences = rb_u32(&r);
    f->bounds = (uint32_t *)rb_vec(&r, 4, &f->nbounds);
    f->nheaders = rb_u64(&r);
    if (!r.bad && f->nheaders > r.n - r.off) r.bad = 1;
    if (!r.bad) {
        f->headers = (char **)calloc(f->nheaders ? f->nheaders : 1, sizeof(char *));
        for (uint64_t i = 0; i < f->nheaders && !r.bad; i++) {
            uint64_t l = rb_u64(&r);
            const uint8_t *s = rb_take(&r, l);
            f->headers[i] = (char *)calloc(l + 1, 1);
            if (s) memcpy(f->headers[i], s, l);
        }
    } else {
        f->nheaders = 0;
    }
    uint8_t opt = rb_u8(&r);
    if (opt == 1) {
        f->has_lookup = 1;
        f->lk_depth = rb_u32(&r);
        f->lk_core = (uint8_t *)rb_vec(&r, 1, &f->lk_ncore);
        f->lk_iv = (uint32_t *)rb_vec(&r, 8, &f->lk_n);
    } else if (opt != 0) {
        r.bad = 1;
    }
    f->tag = rb_u8(&r);
    const char *msg = NULL;
    if (r.bad) msg = "truncated or malformed FM-index byte stream";
    else if (f->tag > 1) msg = "unknown alphabet tag in serialized FM-index"; /* serialize.rs:35-40 */
    else if (f->occ.encoding > 1) msg = "unknown occ encoding";
    if (msg) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", msg);
        or_free(f);
        return NULL;
    }
    return f;
}

/* ------------------------------------------------------------------------ */
/* backward search (query.rs:219-277, 390-407)                               */
/* ------------------------------------------------------------------------ */
static int cmp_pair_lo(const void *a, const void *b) {
    uint32_t x = ((const uint32_t *)a)[0], y = ((const uint32_t *)b)[0];
    return x < y ? -1 : (x > y);
}
static void merge_intervals(vu32 *v) {
    size_t k = v->n / 2;
    qsort(v->a, k, 8, cmp_pair_lo);
    size_t w = 0;
    uint32_t cl = v->a[0], ch = v->a[1];
    for (size_t r = 1; r < k; r++) {
        uint32_t lo = v->a[2 * r], hi = v->a[2 * r + 1];
        if (lo <= ch) { if (hi > ch) ch = hi; }
        else { v->a[2 * w] = cl; v->a[2 * w + 1] = ch; w++; cl = lo; ch = hi; }
    }
    v->a[2 * w] = cl; v->a[2 * w + 1] = ch;
    v->n = 2 * (w + 1);
}

static void backward_search_v(const or_fmi *f, const uint8_t *p, uint32_t m, vu32 *cur, vu32 *nxt) {
    cur->n = 0;
    if (m == 0) { vu32_push(cur, 0); vu32_push(cur, f->text_len); return; }
    uint32_t start = m;
    vu32_push(cur, 0);
    vu32_push(cur, f->text_len);
    if (f->has_lookup && m >= f->lk_depth) {
        const uint8_t *suf = p + (m - f->lk_depth);
        uint64_t idx = 0;
        int ok = 1;
        for (int64_t k = (int64_t)f->lk_depth - 1; k >= 0; k--) {     /* lookup.rs:101-111 */
            int64_t d = -1;
            for (uint64_t s = 0; s < f->lk_ncore; s++)
                if (f->lk_core[s] == suf[k]) { d = (int64_t)s; break; }
            if (d < 0) { ok = 0; break; }
            idx = idx * f->lk_ncore + (uint64_t)d;
        }
        if (ok) {
            if (g_probe_on) g_probe.occ_probes++;
            uint32_t lo = f->lk_iv[2 * idx], hi = f->lk_iv[2 * idx + 1];
            if (lo >= hi) { cur->n = 0; return; }
            cur->a[0] = lo;
            cur->a[1] = hi;
            start = m - f->lk_depth;
        }
    }
    for (int64_t i = (int64_t)start - 1; i >= 0; i--) {
        const uint8_t *cl;
        uint32_t k = compat_list(f->tag, p[i], &cl);
        nxt->n = 0;
        for (size_t j = 0; j < cur->n / 2; j++) {
            uint32_t lo = cur->a[2 * j], hi = cur->a[2 * j + 1];
            probe_pair(lo, hi);
            for (uint32_t s = 0; s < k; s++) {
                uint8_t r = cl[s];
                if (symbol_count(f, r) == 0) continue;
                uint32_t cv = f->c[r];
                uint32_t nl = cv + occ_rank(&f->occ, r, lo), nh = cv + occ_rank(&f->occ, r, hi);
                if (nl < nh) { vu32_push(nxt, nl); vu32_push(nxt, nh); }
            }
        }
        if (nxt->n > 2) merge_intervals(nxt);
        if (nxt->n == 0) { cur->n = 0; return; }
        vu32 t = *cur; *cur = *nxt; *nxt = t;
    }
}

uint32_t or_backward_search(const or_fmi *f, const uint8_t *p, uint32_t m, uint32_t *ivs,
                            uint32_t cap) {
    vu32 a = {0, 0, 0}, b = {0, 0, 0};
    backward_search_v(f, p, m, &a, &b);
    uint32_t k = (uint32_t)(a.n / 2);
    for (uint32_t i = 0; i < k && i < cap; i++) { ivs[2 * i] = a.a[2 * i]; ivs[2 * i + 1] = a.a[2 * i + 1]; }
    free(a.a);
    free(b.a);
    return k;
}

uint32_t or_count(const or_fmi *f, const uint8_t *p, uint32_t m) {
    vu32 a = {0, 0, 0}, b = {0, 0, 0};
    backward_search_v(f, p, m, &a, &b);
    uint32_t s = 0;
    for (size_t i = 0; i < a.n / 2; i++) s += a.a[2 * i + 1] - a.a[2 * i];
    free(a.a);
    free(b.a);
    return s;
}

/* query.rs:305-339: lockstep walk; results in row order */
static void resolve_rows(const or_fmi *f, const uint32_t *rows, uint64_t nr, uint32_t *out) {
    uint32_t *cur = (uint32_t *)malloc((nr ? nr : 1) * 4), *steps = (uint32_t *)calloc(nr ? nr : 1, 4);
    uint32_t *act = (uint32_t *)malloc((nr ? nr : 1) * 4), *nact = (uint32_t *)malloc((nr ? nr : 1) * 4);
    memcpy(cur, rows, nr * 4);
    uint64_t na = nr;
    for (uint64_t i = 0; i < nr; i++) act[i] = (uint32_t)i;
    while (na) {
        for (uint64_t k = 0; k < na; k++) {
            __builtin_prefetch(f->sa.words + (uint64_t)(cur[act[k]] / 64) * SA_REC);
            __builtin_prefetch(f->occ.blk + (uint64_t)(cur[act[k]] / BLOCK) * f->occ.stride);
        }
        uint64_t nn = 0;
        for (uint64_t k = 0; k < na; k++) {
            uint32_t idx = act[k], v;
            if (ssa_get(&f->sa, cur[idx], &v)) out[idx] = v + steps[idx];
            else { cur[idx] = lf_mapping(f, cur[idx]); steps[idx]++; nact[nn++] = idx; }
        }
        uint32_t *t = act; act = nact; nact = t;
        na = nn;
    }
    free(cur); free(steps); free(act); free(nact);
}

/* ------------------------------------------------------------------------ */
/* thread pool helper                                                        */
/* ------------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int tid, uint64_t b, uint64_t e);
typedef struct { range_fn fn; void *ctx; int tid; uint64_t b, e; } rjob;
static void *rjob_run(void *a) { rjob *j = (rjob *)a; j->fn(j->ctx, j->tid, j->b, j->e); return NULL; }
static void parallel_range(uint64_t n, int nthreads, range_fn fn, void *ctx) {
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n) nthreads = n ? (int)n : 1;
    rjob *jobs = (rjob *)calloc(nthreads, sizeof(rjob));
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        jobs[t].fn = fn; jobs[t].ctx = ctx; jobs[t].tid = t;
        jobs[t].b = n * t / nthreads; jobs[t].e = n * (t + 1) / nthreads;
        if (nthreads > 1) pthread_create(&th[t], NULL, rjob_run, &jobs[t]);
        else rjob_run(&jobs[t]);
    }
    if (nthreads > 1) for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* batched locate (query.rs:22-39 per query)                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const or_fmi *f; const uint8_t *q; const uint64_t *qo; int counts_only;
    uint64_t *cnt;           /* per query */
    vu32 *hits;              /* per thread */
    or_probe_counts *probes; /* per thread */
} loc_ctx;
static void loc_worker(void *vc, int tid, uint64_t b, uint64_t e) {
    loc_ctx *c = (loc_ctx *)vc;
    vu32 cur = {0, 0, 0}, nxt = {0, 0, 0}, rows = {0, 0, 0};
    vu32 *out = &c->hits[tid];
    or_probe_reset();
    uint32_t *pos = NULL;
    size_t pcap = 0;
    for (uint64_t q = b; q < e; q++) {
        backward_search_v(c->f, c->q + c->qo[q], (uint32_t)(c->qo[q + 1] - c->qo[q]), &cur, &nxt);
        rows.n = 0;
        uint64_t tot = 0;
        for (size_t i = 0; i < cur.n / 2; i++) tot += cur.a[2 * i + 1] - cur.a[2 * i];
        c->cnt[q] = tot;
        if (c->counts_only || tot == 0) continue;
        for (size_t i = 0; i < cur.n / 2; i++)
            for (uint32_t r = cur.a[2 * i]; r < cur.a[2 * i + 1]; r++) vu32_push(&rows, r);
        if (rows.n > pcap) { pcap = rows.n; pos = (uint32_t *)realloc(pos, pcap * 4); }
        resolve_rows(c->f, rows.a, rows.n, pos);
        for (size_t i = 0; i < rows.n; i++) {
            uint32_t s = 0, o = 0;
            or_map_position(c->f, pos[i], &s, &o);
            vu32_push(out, s);
            vu32_push(out, o);
        }
    }
    or_probe_get(&c->probes[tid]);
    free(cur.a); free(nxt.a); free(rows.a); free(pos);
}

int or_locate_batch(const or_fmi *f, const uint8_t *qflat, const uint64_t *qoff, uint32_t nq,
                    int nthreads, int counts_only, or_locate_out *out) {
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > nq) nthreads = nq ? (int)nq : 1;
    loc_ctx c;
    c.f = f; c.q = qflat; c.qo = qoff; c.counts_only = counts_only;
    c.cnt = (uint64_t *)calloc(nq ? nq : 1, 8);
    c.hits = (vu32 *)calloc(nthreads, sizeof(vu32));
    c.probes = (or_probe_counts *)calloc(nthreads, sizeof(or_probe_counts));
    parallel_range(nq, nthreads, loc_worker, &c);
    out->offsets = (uint64_t *)malloc((nq + 1) * 8);
    out->offsets[0] = 0;
    for (uint32_t q = 0; q < nq; q++) out->offsets[q + 1] = out->offsets[q] + c.cnt[q];
    uint64_t tot = 0;
    for (int t = 0; t < nthreads; t++) tot += c.hits[t].n;
    out->hits = (uint32_t *)malloc(tot ? tot * 4 : 4);
    uint64_t k = 0;
    for (int t = 0; t < nthreads; t++) {
        memcpy(out->hits + k, c.hits[t].a, c.hits[t].n * 4);
        k += c.hits[t].n;
        free(c.hits[t].a);
    }
    memset(&g_probe, 0, sizeof g_probe);
    for (int t = 0; t < nthreads; t++) {
        g_probe.occ_probes += c.probes[t].occ_probes;
        g_probe.sa_probes += c.probes[t].sa_probes;
        g_probe.sa_val_reads += c.probes[t].sa_val_reads;
        g_probe.steps += c.probes[t].steps;
    }
    free(c.cnt); free(c.hits); free(c.probes);
    return 0;
}
void or_locate_out_free(or_locate_out *o) { free(o->offsets); free(o->hits); o->offsets = NULL; o->hits = NULL; }

typedef struct { const or_fmi *f; const uint32_t *rows; uint32_t *out; } rr_ctx;
static void rr_worker(void *vc, int tid, uint64_t b, uint64_t e) {
    (void)tid;
    rr_ctx *c = (rr_ctx *)vc;
    resolve_rows(c->f, c->rows + b, e - b, c->out + b);
}
int or_resolve_rows(const or_fmi *f, const uint32_t *rows, uint64_t nrows, uint32_t *pos_out,
                    int nthreads) {
    rr_ctx c = {f, rows, pos_out};
    parallel_range(nrows, nthreads, rr_worker, &c);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* bidirectional intervals (fm_index/bidir.rs)                               */
/* ------------------------------------------------------------------------ */
/* bidir.rs:136-143 */
static inline uint32_t count_smaller_than(uint8_t c, uint32_t lo, uint32_t hi, const or_fmi *f) {
    uint32_t s = 0, lim = c < ALPHA ? c : ALPHA;
    for (uint32_t b = 0; b < lim; b++) s += occ_rank(&f->occ, (uint8_t)b, hi) - occ_rank(&f->occ, (uint8_t)b, lo);
    return s;
}
/* bidir.rs:80-100 */
int or_extend_right(const or_fmi *rev, const or_biv *iv, uint8_t c, or_biv *out) {
    if (c >= ALPHA) return 0;
    uint32_t cv = rev->c[c];
    uint32_t nl = cv + occ_rank(&rev->occ, c, iv->rev_lo), nh = cv + occ_rank(&rev->occ, c, iv->rev_hi);
    if (nl >= nh) return 0;
    uint32_t off = count_smaller_than(c, iv->rev_lo, iv->rev_hi, rev);
    uint32_t sz = nh - nl;
    out->fwd_lo = iv->fwd_lo + off;
    out->fwd_hi = iv->fwd_lo + off + sz;
    out->rev_lo = nl;
    out->rev_hi = nh;
    return 1;
}
/* bidir.rs:112-130 */
int or_extend_left(const or_fmi *fwd, const or_biv *iv, uint8_t c, or_biv *out) {
    if (c >= ALPHA) return 0;
    uint32_t cv = fwd->c[c];
    uint32_t nl = cv + occ_rank(&fwd->occ, c, iv->fwd_lo), nh = cv + occ_rank(&fwd->occ, c, iv->fwd_hi);
    if (nl >= nh) return 0;
    uint32_t off = count_smaller_than(c, iv->fwd_lo, iv->fwd_hi, fwd);
    uint32_t sz = nh - nl;
    out->fwd_lo = nl;
    out->fwd_hi = nh;
    out->rev_lo = iv->rev_lo + off;
    out->rev_hi = iv->rev_lo + off + sz;
    return 1;
}

/* smem.rs:332-343: base-major, then intervals */
static void extend_multi_right(const vbiv *ivs, uint8_t c, const or_fmi *rev, vbiv *res) {
    const uint8_t *bl;
    uint32_t k = compat_list(rev->tag, c, &bl);
    res->n = 0;
    for (size_t i = 0; i < ivs->n; i++) probe_pair(ivs->a[i].rev_lo, ivs->a[i].rev_hi);
    for (uint32_t s = 0; s < k; s++)
        for (size_t i = 0; i < ivs->n; i++) {
            or_biv e;
            if (or_extend_right(rev, &ivs->a[i], bl[s], &e)) vbiv_push(res, e);
        }
}
/* smem.rs:346-357 */
static void extend_multi_left(const vbiv *ivs, uint8_t c, const or_fmi *fwd, vbiv *res) {
    const uint8_t *bl;
    uint32_t k = compat_list(fwd->tag, c, &bl);
    res->n = 0;
    for (size_t i = 0; i < ivs->n; i++) probe_pair(ivs->a[i].fwd_lo, ivs->a[i].fwd_hi);
    for (uint32_t s = 0; s < k; s++)
        for (size_t i = 0; i < ivs->n; i++) {
            or_biv e;
            if (or_extend_left(fwd, &ivs->a[i], bl[s], &e)) vbiv_push(res, e);
        }
}
static uint64_t coverage(const vbiv *v) {
    uint64_t s = 0;
    for (size_t i = 0; i < v->n; i++) s += v->a[i].fwd_hi - v->a[i].fwd_lo;
    return s;
}

typedef struct { uint32_t start, end, count; vbiv ivs; } rawmem;
typedef struct { rawmem *a; size_t n, cap; } vraw;
static void vraw_push(vraw *v, rawmem x) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 8; v->a = (rawmem *)realloc(v->a, v->cap * sizeof(rawmem)); }
    v->a[v->n++] = x;
}

/* smem.rs:250-295 */
static int raw_mem_from(const or_fmi *fwd, const or_fmi *rev, const uint8_t *q, uint32_t n,
                        uint32_t i, uint32_t min_len, rawmem *out) {
    vbiv ivs = {0, 0, 0}, nxt = {0, 0, 0}, tmp = {0, 0, 0};
    or_biv full = {0, fwd->text_len, 0, fwd->text_len};
    vbiv_push(&ivs, full);
    uint32_t j = i;
    int matched = 0;
    while (j < n) {
        extend_multi_right(&ivs, q[j], rev, &nxt);
        if (nxt.n == 0) break;
        vbiv t = ivs; ivs = nxt; nxt = t;
        j++;
        matched = 1;
    }
    int ok = 0;
    if (matched && j - i >= min_len) {
        int left_max = 1;
        if (i > 0) { extend_multi_left(&ivs, q[i - 1], fwd, &tmp); left_max = tmp.n == 0; }
        if (left_max) {
            out->start = i; out->end = j; out->count = (uint32_t)coverage(&ivs);
            out->ivs = ivs; ivs.a = NULL;
            ok = 1;
        }
    }
    free(ivs.a); free(nxt.a); free(tmp.a);
    return ok;
}

/* smem.rs:227-243 */
static uint32_t extend_left_maximally(const or_fmi *fwd, vbiv *ivs, const uint8_t *q, uint32_t from) {
    vbiv nxt = {0, 0, 0};
    uint32_t start = from;
    while (start > 0) {
        extend_multi_left(ivs, q[start - 1], fwd, &nxt);
        if (nxt.n == 0) break;
        vbiv t = *ivs; *ivs = nxt; nxt = t;
        start--;
    }
    free(nxt.a);
    return start;
}

/* smem.rs:154-222 */
static void collect_smem_candidates(const or_fmi *fwd, const or_fmi *rev, const uint8_t *q,
                                    uint32_t n, uint32_t min_len, vraw *out) {
    uint32_t pivot = 0;
    or_biv full = {0, fwd->text_len, 0, fwd->text_len};
    vbiv root = {0, 0, 0};
    vbiv_push(&root, full);
    vbiv curr = {0, 0, 0}, next = {0, 0, 0}, tmp = {0, 0, 0};
    typedef struct { vbiv ivs; uint32_t end; } pref;
    pref *pf = NULL;
    size_t npf = 0, cpf = 0;
    while (pivot < n) {
        extend_multi_right(&root, q[pivot], rev, &curr);
        if (curr.n == 0) { pivot++; continue; }
        uint64_t cov = coverage(&curr);
        npf = 0;
        uint32_t j = pivot + 1;
        for (;;) {
            if (npf == cpf) { cpf = cpf ? cpf * 2 : 16; pf = (pref *)realloc(pf, cpf * sizeof(pref)); }
            if (j == n) { pf[npf].ivs = vbiv_clone(&curr); pf[npf].end = n; npf++; break; }
            extend_multi_right(&curr, q[j], rev, &next);
            uint64_t ncov = coverage(&next);
            if (ncov != cov) { pf[npf].ivs = vbiv_clone(&curr); pf[npf].end = j; npf++; }
            if (next.n == 0) break;
            vbiv t = curr; curr = next; next = t;
            cov = ncov;
            j++;
        }
        uint32_t reach = npf ? pf[npf - 1].end : pivot + 1;
        for (size_t k = 0; k < npf; k++) {
            uint32_t end = pf[k].end;
            uint32_t start = extend_left_maximally(fwd, &pf[k].ivs, q, pivot);
            if (end - start < min_len) { free(pf[k].ivs.a); continue; }
            int right_max = 1;
            if (end != n) { extend_multi_right(&pf[k].ivs, q[end], rev, &tmp); right_max = tmp.n == 0; }
            if (!right_max) { free(pf[k].ivs.a); continue; }
            rawmem rm = {start, end, (uint32_t)coverage(&pf[k].ivs), pf[k].ivs};
            vraw_push(out, rm);
        }
        pivot = reach > pivot + 1 ? reach : pivot + 1;
    }
    free(root.a); free(curr.a); free(next.a); free(tmp.a); free(pf);
}

static int cmp_raw(const void *a, const void *b) {
    const rawmem *x = (const rawmem *)a, *y = (const rawmem *)b;
    if (x->start != y->start) return x->start < y->start ? -1 : 1;
    if (x->end != y->end) return x->end < y->end ? -1 : 1;
    return 0;
}
/* stable sort by (start,end) (Rust sort_by_key is stable): insertion-merge */
static void stable_sort_raw(vraw *v) {
    for (size_t i = 1; i < v->n; i++) {
        rawmem x = v->a[i];
        size_t j = i;
        while (j > 0 && cmp_raw(&v->a[j - 1], &x) > 0) { v->a[j] = v->a[j - 1]; j--; }
        v->a[j] = x;
    }
}

/* smem.rs:71-104 (mode 0) and 113-124 (mode 1) */
static void find_mems_one(const or_fmi *fwd, const or_fmi *rev, const uint8_t *q, uint32_t n,
                          uint32_t min_len, int mode, vraw *res) {
    res->n = 0;
    if (n == 0 || min_len == 0) return;
    vraw raws = {0, 0, 0};
    if (mode == 0) collect_smem_candidates(fwd, rev, q, n, min_len, &raws);
    else
        for (uint32_t i = 0; i < n; i++) {
            rawmem rm;
            if (raw_mem_from(fwd, rev, q, n, i, min_len, &rm)) vraw_push(&raws, rm);
        }
    stable_sort_raw(&raws);
    /* dedup_by_key: keep the first of each run */
    size_t w = 0;
    for (size_t i = 0; i < raws.n; i++) {
        if (w > 0 && raws.a[w - 1].start == raws.a[i].start && raws.a[w - 1].end == raws.a[i].end) {
            free(raws.a[i].ivs.a);
            continue;
        }
        raws.a[w++] = raws.a[i];
    }
    raws.n = w;
    if (mode == 0) {
        /* containment filter (smem.rs:87-103) */
        for (size_t i = 0; i < raws.n; i++) {
            int contained = 0;
            for (size_t j = 0; j < raws.n && !contained; j++)
                if (j != i && raws.a[j].start <= raws.a[i].start && raws.a[i].end <= raws.a[j].end &&
                    !(raws.a[j].start == raws.a[i].start && raws.a[j].end == raws.a[i].end))
                    contained = 1;
            if (!contained) vraw_push(res, raws.a[i]);
            else free(raws.a[i].ivs.a);
        }
        free(raws.a);
    } else {
        *res = raws;
    }
}

typedef struct {
    const or_fmi *fwd, *rev; const uint8_t *q; const uint64_t *qo; uint32_t min_len; int mode;
    vraw *per_q; or_probe_counts *probes;
} mem_ctx;
static void mem_worker(void *vc, int tid, uint64_t b, uint64_t e) {
    mem_ctx *c = (mem_ctx *)vc;
    or_probe_reset();
    for (uint64_t q = b; q < e; q++)
        find_mems_one(c->fwd, c->rev, c->q + c->qo[q], (uint32_t)(c->qo[q + 1] - c->qo[q]),
                      c->min_len, c->mode, &c->per_q[q]);
    or_probe_get(&c->probes[tid]);
}

int or_find_mems_batch(const or_fmi *fwd, const or_fmi *rev, const uint8_t *qflat,
                       const uint64_t *qoff, uint32_t nq, uint32_t min_len, int mode,
                       int nthreads, or_mem_out *out) {
    if (nthreads < 1) nthreads = 1;
    mem_ctx c = {fwd, rev, qflat, qoff, min_len, mode, NULL, NULL};
    c.per_q = (vraw *)calloc(nq ? nq : 1, sizeof(vraw));
    c.probes = (or_probe_counts *)calloc(nthreads, sizeof(or_probe_counts));
    parallel_range(nq, nthreads, mem_worker, &c);
    uint64_t nm = 0, ni = 0;
    for (uint32_t q = 0; q < nq; q++) {
        nm += c.per_q[q].n;
        for (size_t k = 0; k < c.per_q[q].n; k++) ni += c.per_q[q].a[k].ivs.n;
    }
    out->mem_off = (uint64_t *)malloc((nq + 1) * 8);
    out->mems = (uint32_t *)malloc((nm + 1) * 16);
    out->ivs = (uint32_t *)malloc((ni ? ni : 1) * 16);
    out->nmem = nm;
    out->niv = ni;
    uint64_t m = 0, iv = 0;
    out->mem_off[0] = 0;
    for (uint32_t q = 0; q < nq; q++) {
        for (size_t k = 0; k < c.per_q[q].n; k++) {
            rawmem *r = &c.per_q[q].a[k];
            out->mems[4 * m + 0] = r->start;
            out->mems[4 * m + 1] = r->end;
            out->mems[4 * m + 2] = r->count;
            out->mems[4 * m + 3] = (uint32_t)iv;
            for (size_t t = 0; t < r->ivs.n; t++) {
                out->ivs[4 * iv + 0] = r->ivs.a[t].fwd_lo;
                out->ivs[4 * iv + 1] = r->ivs.a[t].fwd_hi;
                out->ivs[4 * iv + 2] = r->ivs.a[t].rev_lo;
                out->ivs[4 * iv + 3] = r->ivs.a[t].rev_hi;
                iv++;
            }
            free(r->ivs.a);
            m++;
        }
        free(c.per_q[q].a);
        out->mem_off[q + 1] = m;
    }
    out->mems[4 * m + 0] = out->mems[4 * m + 1] = out->mems[4 * m + 2] = 0;
    out->mems[4 * m + 3] = (uint32_t)iv;
    memset(&g_probe, 0, sizeof g_probe);
    for (int t = 0; t < nthreads; t++) {
        g_probe.occ_probes += c.probes[t].occ_probes;
        g_probe.sa_probes += c.probes[t].sa_probes;
        g_probe.sa_val_reads += c.probes[t].sa_val_reads;
        g_probe.steps += c.probes[t].steps;
    }
    free(c.per_q);
    free(c.probes);
    return 0;
}
void or_mem_out_free(or_mem_out *o) {
    free(o->mem_off); free(o->mems); free(o->ivs);
    o->mem_off = NULL; o->mems = NULL; o->ivs = NULL;
}

/* ------------------------------------------------------------------------ */
/* brute force (tests/common/mod.rs:29-44 and smem.rs:380-403, generalised   */
/* to the IUPAC predicate: text[p+k] in compatible(q[k]) for all k)          */
/* ------------------------------------------------------------------------ */
static int match_at(const uint8_t *text, uint64_t n, uint64_t p, const uint8_t *q, uint32_t m,
                    const uint16_t *mask) {
    if (p + m > n) return 0;
    for (uint32_t k = 0; k < m; k++) {
        uint8_t t = text[p + k];
        if (t >= 16 || !((mask[q[k]] >> t) & 1)) return 0;
    }
    return 1;
}
uint64_t or_naive_locate(const uint8_t *text, uint64_t n, const uint8_t *q, uint32_t m,
                         uint8_t tag, uint32_t *pos_out, uint64_t cap) {
    uint16_t mask[256];
    for (int b = 0; b < 256; b++) mask[b] = or_compat_mask(tag, (uint8_t)b);
    uint64_t k = 0;
    if (m == 0) return 0;
    for (uint64_t p = 0; p + m <= n; p++)
        if (match_at(text, n, p, q, m, mask)) { if (k < cap) pos_out[k] = (uint32_t)p; k++; }
    return k;
}
static uint64_t occ_count(const uint8_t *text, uint64_t n, const uint8_t *q, uint32_t m,
                          const uint16_t *mask) {
    uint64_t k = 0;
    for (uint64_t p = 0; p + m <= n; p++) k += match_at(text, n, p, q, m, mask);
    return k;
}
uint64_t or_brute_mems(const uint8_t *text, uint64_t n, const uint8_t *q, uint32_t m,
                       uint32_t min_len, uint8_t tag, uint32_t *out, uint64_t cap) {
    uint16_t mask[256];
    for (int b = 0; b < 256; b++) mask[b] = or_compat_mask(tag, (uint8_t)b);
    uint64_t k = 0;
    for (uint32_t s = 0; s < m; s++)
        for (uint32_t e = s + (min_len ? min_len : 1); e <= m; e++) {
            uint64_t c = occ_count(text, n, q + s, e - s, mask);
            if (!c) continue;
            int rmax = e == m || !occ_count(text, n, q + s, e - s + 1, mask);
            int lmax = s == 0 || !occ_count(text, n, q + s - 1, e - s + 1, mask);
            if (rmax && lmax) {
                if (k < cap) { out[3 * k] = s; out[3 * k + 1] = e; out[3 * k + 2] = (uint32_t)c; }
                k++;
            }
        }
    return k;
}

/* ------------------------------------------------------------------------ */
/* synthetic workloads                                                       */
/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t *s) { return (or_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint8_t gc_base(uint64_t *s, double gc) {
    double u = u01(s);
    if (u < gc / 2) return C_;
    if (u < gc) return G_;
    if (u < gc + (1 - gc) / 2) return A_;
    return T_;
}
static inline uint8_t other_base(uint64_t *s, uint8_t b) {
    if (b < A_ || b > T_) return (uint8_t)(A_ + (or_splitmix64(s) % 4));
    uint8_t o = (uint8_t)(A_ + (or_splitmix64(s) % 3));
    if (o >= b) o++;
    return o;
}
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

int or_gen_reference(uint64_t seed, uint64_t total, uint32_t ncontig, double gc,
                     double repeat_frac, uint32_t n_units, uint32_t unit_min, uint32_t unit_max,
                     double div_min, double div_max, double iupac_frac, const uint8_t *iupac_codes,
                     uint32_t n_iupac, uint8_t **text_out, uint64_t *n_out, uint32_t **bounds_out) {
    if (ncontig == 0 || total < ncontig) return -1;
    uint64_t st = seed;
    uint64_t *cuts = (uint64_t *)malloc((ncontig + 1) * 8);
    for (;;) {
        for (uint32_t i = 0; i + 1 < ncontig; i++) cuts[i] = 1 + or_splitmix64(&st) % (total - 1);
        cuts[ncontig - 1] = total;
        qsort(cuts, ncontig - 1, 8, cmp_u64);
        int ok = 1;
        uint64_t prev = 0;
        for (uint32_t i = 0; i < ncontig; i++) { if (cuts[i] <= prev) ok = 0; prev = cuts[i]; }
        if (ok) break;
    }
    uint8_t *seq = (uint8_t *)malloc(total);
    for (uint64_t i = 0; i < total; i++) seq[i] = gc_base(&st, gc);
    /* repeat families: copies of units at per-copy divergence */
    if (repeat_frac > 0 && n_units > 0) {
        uint8_t **units = (uint8_t **)malloc(n_units * sizeof(uint8_t *));
        uint32_t *ulen = (uint32_t *)malloc(n_units * 4);
        for (uint32_t u = 0; u < n_units; u++) {
            ulen[u] = unit_min + (uint32_t)(or_splitmix64(&st) % (unit_max - unit_min + 1));
            units[u] = (uint8_t *)malloc(ulen[u]);
            for (uint32_t i = 0; i < ulen[u]; i++) units[u][i] = gc_base(&st, gc);
        }
        uint64_t covered = 0, target = (uint64_t)(repeat_frac * (double)total);
        while (covered < target) {
            uint32_t u = (uint32_t)(or_splitmix64(&st) % n_units);
            if (ulen[u] >= total) break;
            uint64_t at = or_splitmix64(&st) % (total - ulen[u]);
            double d = div_min + (div_max - div_min) * u01(&st);
            for (uint32_t i = 0; i < ulen[u]; i++) {
                uint8_t b = units[u][i];
                if (u01(&st) < d) b = other_base(&st, b);
                seq[at + i] = b;
            }
            covered += ulen[u];
        }
        for (uint32_t u = 0; u < n_units; u++) free(units[u]);
        free(units);
        free(ulen);
    }
    if (iupac_frac > 0 && n_iupac > 0)
        for (uint64_t i = 0; i < total; i++)
            if (u01(&st) < iupac_frac) seq[i] = iupac_codes[or_splitmix64(&st) % n_iupac];
    uint8_t *text = (uint8_t *)malloc(total + ncontig);
    uint32_t *bounds = (uint32_t *)malloc(ncontig * 4);
    uint64_t w = 0, prev = 0;
    for (uint32_t c = 0; c < ncontig; c++) {
        memcpy(text + w, seq + prev, cuts[c] - prev);
        w += cuts[c] - prev;
        text[w++] = 0;
        bounds[c] = (uint32_t)w;
        prev = cuts[c];
    }
    free(seq);
    free(cuts);
    *text_out = text;
    *n_out = w;
    *bounds_out = bounds;
    return 0;
}

/* IUPAC complement: A<->T, C<->G, R<->Y, K<->M, B<->V, D<->H; N, S, W fixed */
static const uint8_t COMP[16] = {0, T_, G_, C_, A_, N_, Y_, R_, S_, W_, M_, K_, V_, H_, D_, B_};

int or_gen_reads(uint64_t seed, const uint8_t *text, uint64_t n, const uint32_t *bounds,
                 uint32_t nseq, uint64_t nreads, uint32_t len, double sigma, uint32_t len_min,
                 uint32_t len_max, double sub, double indel, int both_strands, uint64_t n_random,
                 double gc, uint8_t **flat_out, uint64_t **off_out) {
    (void)n;
    uint64_t st = seed;
    uint64_t total = nreads + n_random;
    uint64_t *off = (uint64_t *)malloc((total + 1) * 8);
    uint32_t *lens = (uint32_t *)malloc((total ? total : 1) * 4);
    uint64_t sum = 0;
    for (uint64_t r = 0; r < total; r++) {
        uint32_t L = len;
        if (sigma > 0) {
            double u1 = u01(&st), u2 = u01(&st);
            if (u1 < 1e-300) u1 = 1e-300;
            double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            double v = exp(log((double)len) + sigma * z);
            if (v < len_min) v = len_min;
            if (v > len_max) v = len_max;
            L = (uint32_t)(v + 0.5);
        }
        lens[r] = L;
        off[r] = sum;
        sum += L;
    }
    off[total] = sum;
    uint8_t *flat = (uint8_t *)malloc(sum ? sum : 1);
    uint64_t *cstart = (uint64_t *)malloc((nseq + 1) * 8);
    cstart[0] = 0;
    for (uint32_t c = 0; c < nseq; c++) cstart[c + 1] = bounds[c];
    uint64_t span = bounds[nseq - 1];
    uint8_t *win = (uint8_t *)malloc((size_t)len_max * 2 + 64 + (size_t)len * 2);
    for (uint64_t r = 0; r < total; r++) {
        uint32_t L = lens[r];
        uint8_t *dst = flat + off[r];
        if (r >= nreads) {
            for (uint32_t i = 0; i < L; i++) dst[i] = gc_base(&st, gc);
            continue;
        }
        /* contig weighted by length (uniform position over the text), window fits */
        uint32_t c = 0;
        uint64_t need = (uint64_t)L + (indel > 0 ? 8 : 0);
        for (int tries = 0;; tries++) {
            uint64_t p = or_splitmix64(&st) % span;
            c = 0;
            while (cstart[c + 1] <= p) c++;
            uint64_t clen = cstart[c + 1] - cstart[c] - 1;
            if (clen >= need || tries > 1000) break;
        }
        uint64_t cs = cstart[c], clen = cstart[c + 1] - cstart[c] - 1;
        if (clen < need) need = clen;
        uint64_t at = cs + (clen > need ? or_splitmix64(&st) % (clen - need + 1) : 0);
        /* copy with indels until L bases are produced */
        uint32_t k = 0;
        uint64_t p = at, lim = cs + clen;
        while (k < L && p < lim) {
            if (indel > 0 && u01(&st) < indel) {
                uint32_t il = 1 + (uint32_t)(or_splitmix64(&st) % 3);
                if (or_splitmix64(&st) & 1) {
                    for (uint32_t t = 0; t < il && k < L; t++) win[k++] = gc_base(&st, gc);
                } else {
                    p += il;
                }
                continue;
            }
            win[k++] = text[p++];
        }
        while (k < L) win[k++] = gc_base(&st, gc);
        int rc = both_strands && (or_splitmix64(&st) & 1);
        for (uint32_t i = 0; i < L; i++) {
            uint8_t b = rc ? COMP[win[L - 1 - i] & 15] : win[i];
            if (sub > 0 && u01(&st) < sub) b = other_base(&st, b);
            dst[i] = b;
        }
    }
    free(win);
    free(cstart);
    free(lens);
    *flat_out = flat;
    *off_out = off;
    return 0;
}
