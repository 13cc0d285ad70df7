/*
 * hsfm_build.h -- host-side index construction and seeded workload synthesis.
 *
 * Index construction is not part of the GPU query path; these entry points
 * produce the byte streams that the reference's FmIndex::build_cpu +
 * FmIndex::to_bytes (src/fm_index/mod.rs:117-195, src/fm_index/serialize.rs:13-26)
 * and BidirFmIndex::build_cpu + to_bytes (src/fm_index/bidir_index.rs:53-80,
 * 183-192) produce, so that hsfm_index_view_from_bytes / hsfm_index_upload can be
 * fed without the Rust crate.  Host code only (libhsfm_build.so).
 */
#ifndef HSFM_BUILD_H
#define HSFM_BUILD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* text: concatenated codes with a 0 sentinel after each sequence
 * (concatenate_sequences, src/alphabet.rs:375-393); bounds: exclusive ends.
 * headers: '\0'-separated names or NULL for "seq_{i}".  encoding 0 = Bitplane,
 * 1 = OneHot; tag 0 = IupacDna, 1 = ExactDna.  Returns 0 on success; *out is
 * malloc'd (free with hsfm_build_free). */
int hsfm_build_index(const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq,
                     const char *headers, uint32_t sample_rate, uint32_t lookup_depth,
                     uint32_t encoding, uint8_t tag, int nthreads, uint8_t **out, uint64_t *out_len);
/* [u32 LE fwd_len][fwd][rev]; the reverse index has lookup_depth 0. */
int hsfm_build_bidir(const uint8_t *text, uint64_t n, const uint32_t *bounds, uint32_t nseq,
                     const char *headers, uint32_t sample_rate, uint32_t lookup_depth,
                     uint32_t encoding, uint8_t tag, int nthreads, uint8_t **out, uint64_t *out_len);
void hsfm_build_free(void *p);

/* Seeded synthetic references: ncontig contigs of uniform random lengths
 * summing to total, i.i.d. background at GC fraction gc, repeat_frac of the
 * sequence covered by diverged copies of n_units random units, iupac_frac of
 * positions replaced by codes drawn from iupac_codes. */
int hsfm_synth_reference(uint64_t seed, uint64_t total, uint32_t ncontig, double gc,
                         double repeat_frac, uint32_t n_units, uint32_t unit_min, uint32_t unit_max,
                         double div_min, double div_max, double iupac_frac,
                         const uint8_t *iupac_codes, uint32_t n_iupac, uint8_t **text_out,
                         uint64_t *n_out, uint32_t **bounds_out);
/* Reads sampled from the reference (optionally both strands) with
 * substitutions and 1-3 bp indels, plus n_random random reads; fixed length or
 * lognormal(ln len, sigma) clipped to [len_min, len_max]; shuffled when
 * shuffle != 0.  CSR output (malloc'd). */
int hsfm_synth_reads(uint64_t seed, const uint8_t *text, uint64_t n, const uint32_t *bounds,
                     uint32_t nseq, uint64_t nreads, uint32_t len, double sigma, uint32_t len_min,
                     uint32_t len_max, double sub, double indel, int both_strands,
                     uint64_t n_random, double gc, int shuffle, uint8_t **flat_out,
                     uint64_t **off_out);

#ifdef __cplusplus
}
#endif
#endif
