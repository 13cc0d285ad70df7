/*
 * hsfm_oracle.h -- CPU restatement of haystackfm's query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (haystackfm_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm at the cited file:line of
 * sriram98v/haystackfm (paths relative to the reference crate root).  Index
 * construction (SA, BWT, Occ, sampled SA, lookup table) is restated so that the
 * oracle can emit the exact byte stream of FmIndex::to_bytes for the product's
 * FmIndex::from_bytes boundary.
 */
#ifndef HSFM_ORACLE_H
#define HSFM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- alphabet (src/alphabet.rs) ---------------------------------------- */
int      or_encode_byte(uint8_t b);                       /* alphabet.rs:102, -1 = invalid */
uint32_t or_compatible(uint8_t tag, uint8_t code, uint8_t out[16]); /* alphabet.rs:174-305 */
uint16_t or_compat_mask(uint8_t tag, uint8_t code);
uint32_t or_iupac_bases(uint8_t code, uint8_t out[4]);    /* alphabet.rs:149-168 */

/* ---- index ------------------------------------------------------------- */
typedef struct or_fmi or_fmi;

/* Build an FmIndex from the concatenated text (each sequence followed by a
 * 0 sentinel, alphabet.rs:375-393).  Restates FmIndex::build_cpu_inner
 * (fm_index/mod.rs:117-195).  headers: '\0'-separated blob or NULL for the
 * default "seq_{i}" names.  naive_sa != 0 uses the comparison-sort SA of
 * suffix_array/cpu.rs:20-25 instead of SA-IS (identical order). */
or_fmi *or_build(const uint8_t *text, uint64_t n, const uint32_t *seq_bounds,
                 uint32_t nseq, const char *headers, uint32_t sample_rate,
                 uint32_t lookup_depth, uint32_t encoding, uint8_t tag,
                 int naive_sa, int nthreads);
/* Reverse-text index of BidirFmIndex::build_cpu_with (bidir_index.rs:53-87,
 * 523-532): strip the trailing sentinel, reverse, re-append one sentinel,
 * lookup_depth 0. */
or_fmi *or_build_rev(const uint8_t *text, uint64_t n, uint32_t sample_rate,
                     uint32_t encoding, uint8_t tag, int naive_sa, int nthreads);
void or_free(or_fmi *f);

/* bincode-1 stream of FmIndex::to_bytes (fm_index/serialize.rs:13-86). */
int     or_to_bytes(const or_fmi *f, uint8_t **out, uint64_t *len);
or_fmi *or_from_bytes(const uint8_t *b, uint64_t len, char *err, int errlen);
void    or_free_buf(void *p);

uint32_t or_text_len(const or_fmi *f);
uint32_t or_num_lanes(const or_fmi *f);
uint32_t or_sample_rate(const or_fmi *f);
uint32_t or_num_sequences(const or_fmi *f);
const uint32_t *or_seq_boundaries(const or_fmi *f);
uint32_t or_lookup_depth(const or_fmi *f);
void or_c_array(const or_fmi *f, uint32_t out[16]);

/* ---- primitives -------------------------------------------------------- */
uint32_t or_rank(const or_fmi *f, uint8_t c, uint32_t i);             /* occ/mod.rs:330-363 */
void     or_lf_step(const or_fmi *f, uint32_t pos, uint8_t *sym, uint32_t *rank); /* occ/mod.rs:456-480 */
uint8_t  or_symbol_at(const or_fmi *f, uint32_t pos);                 /* occ/mod.rs:418-423 */
int      or_sa_get(const or_fmi *f, uint32_t row, uint32_t *val);      /* suffix_array/mod.rs:185-201 */
uint32_t or_resolve_sa(const or_fmi *f, uint32_t row);                /* query.rs:282-291 */
int      or_map_position(const or_fmi *f, uint32_t pos, uint32_t *seq, uint32_t *off); /* query.rs:342-357 */

/* backward_search (query.rs:219-277): writes up to cap (lo,hi) pairs, returns
 * the number of intervals (call again with a larger cap if it exceeds cap). */
uint32_t or_backward_search(const or_fmi *f, const uint8_t *p, uint32_t m,
                            uint32_t *ivs, uint32_t cap);
uint32_t or_count(const or_fmi *f, const uint8_t *p, uint32_t m);      /* query.rs:9-14 */

/* Batched locate over nq queries (CSR input).  Output CSR: per query the
 * (seq_idx, pos) pairs in the order FmIndex::locate returns them
 * (query.rs:22-39).  counts_only != 0 skips the resolve (count only). */
typedef struct {
    uint64_t *offsets; /* nq + 1 */
    uint32_t *hits;    /* 2 * offsets[nq]: (seq_idx, pos) */
} or_locate_out;
int  or_locate_batch(const or_fmi *f, const uint8_t *qflat, const uint64_t *qoff,
                     uint32_t nq, int nthreads, int counts_only, or_locate_out *out);
void or_locate_out_free(or_locate_out *o);

/* ---- bidirectional / MEM ----------------------------------------------- */
typedef struct { uint32_t fwd_lo, fwd_hi, rev_lo, rev_hi; } or_biv;
int or_extend_right(const or_fmi *rev, const or_biv *iv, uint8_t c, or_biv *out); /* bidir.rs:80-100 */
int or_extend_left(const or_fmi *fwd, const or_biv *iv, uint8_t c, or_biv *out);  /* bidir.rs:112-130 */

/* find_smems (mode 0, smem.rs:71-104) or find_mems (mode 1, smem.rs:113-124)
 * over a batch.  Output per query: Mems sorted by (start,end) with their
 * accepted interval sets in the order the algorithm produced them.
 *   mem_off[q]..mem_off[q+1]  : Mems of query q
 *   mems[k*4 + {0,1,2,3}]      : query_start, query_end, match_count, iv_off
 *   iv_off indexes ivs (4 u32 per interval: fwd_lo fwd_hi rev_lo rev_hi);
 *   interval count of Mem k = iv_off[k+1] - iv_off[k] (mems has nmem+1 rows). */
typedef struct {
    uint64_t *mem_off; /* nq + 1 */
    uint32_t *mems;    /* 4 * (nmem + 1) */
    uint32_t *ivs;     /* 4 * niv */
    uint64_t  nmem, niv;
} or_mem_out;
int  or_find_mems_batch(const or_fmi *fwd, const or_fmi *rev, const uint8_t *qflat,
                        const uint64_t *qoff, uint32_t nq, uint32_t min_len, int mode,
                        int nthreads, or_mem_out *out);
void or_mem_out_free(or_mem_out *o);

/* Resolve SA rows to text positions (resolve_sa_batch, query.rs:305-339). */
int or_resolve_rows(const or_fmi *f, const uint32_t *rows, uint64_t nrows,
                    uint32_t *pos_out, int nthreads);

/* ---- instrumentation ---------------------------------------------------- */
/* Per-call probe counters of the query algorithms: distinct Occ blocks per
 * backward/extension step, SA marker records per LF step, sa_vals reads. */
typedef struct { uint64_t occ_probes, sa_probes, sa_val_reads, steps; } or_probe_counts;
void or_probe_reset(void);
void or_probe_get(or_probe_counts *out);
void or_probe_enable(int on);

/* ---- brute force checkers (tests/common/mod.rs:29-44, smem.rs:380-403) -- */
/* positions p with compat(q[k]) containing text[p+k] for all k */
uint64_t or_naive_locate(const uint8_t *text, uint64_t n, const uint8_t *q, uint32_t m,
                         uint8_t tag, uint32_t *pos_out, uint64_t cap);
uint64_t or_brute_mems(const uint8_t *text, uint64_t n, const uint8_t *q, uint32_t m,
                       uint32_t min_len, uint8_t tag, uint32_t *out, uint64_t cap);

/* ---- suffix array ------------------------------------------------------- */
int or_suffix_array(const uint8_t *text, uint64_t n, uint32_t *sa, int naive);

/* ---- seeded synthetic workloads ---------------------------------------- */
/* SplitMix64-driven generators for the reference and read sets. */
uint64_t or_splitmix64(uint64_t *state);
/* Reference: ncontig contigs of uniform random lengths summing to total;
 * background i.i.d. bases with GC fraction gc; repeat_frac of the sequence
 * covered by copies of n_units units (len in [unit_min,unit_max]) at divergence
 * in [div_min,div_max]; iupac_frac of positions replaced by a code drawn from
 * iupac_codes.  Writes concatenated text (with sentinels) and boundaries. */
int or_gen_reference(uint64_t seed, uint64_t total, uint32_t ncontig, double gc,
                     double repeat_frac, uint32_t n_units, uint32_t unit_min,
                     uint32_t unit_max, double div_min, double div_max,
                     double iupac_frac, const uint8_t *iupac_codes, uint32_t n_iupac,
                     uint8_t **text_out, uint64_t *n_out, uint32_t **bounds_out);
/* Reads sampled from the reference (both strands when both_strands), with
 * substitution rate sub, indel event rate indel (1-3 bp), plus n_random
 * random-sequence reads.  Lengths: fixed len, or lognormal(ln median, sigma)
 * clipped to [len_min,len_max] when sigma > 0.  Output CSR (flat codes). */
int or_gen_reads(uint64_t seed, const uint8_t *text, uint64_t n, const uint32_t *bounds,
                 uint32_t nseq, uint64_t nreads, uint32_t len, double sigma,
                 uint32_t len_min, uint32_t len_max, double sub, double indel,
                 int both_strands, uint64_t n_random, double gc,
                 uint8_t **flat_out, uint64_t **off_out);

#ifdef __cplusplus
}
#endif
#endif
