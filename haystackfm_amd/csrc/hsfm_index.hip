// hsfm_index.hip -- FmIndex byte-stream view, validation, upload and the
// conversion of the reference's Occ / sampled-SA records into the device layout.
#include <string.h>

#include "hsfm_host.h"

using namespace hsfm;

namespace {

// ---------------------------------------------------------------------------
// bincode-1 reader (little endian, fixed-width ints, u64 sequence lengths,
// Option = u8 tag, unit enum = u32), field order of SerializableFmIndex
// (src/fm_index/serialize.rs:55-67).
// ---------------------------------------------------------------------------
struct Reader {
    const uint8_t *p;
    uint64_t n, off;
    bool bad;
    const uint8_t *take(uint64_t k) {
        if (bad || k > n - off) {
            bad = true;
            return nullptr;
        }
        const uint8_t *q = p + off;
        off += k;
        return q;
    }
    uint8_t u8() { const uint8_t *q = take(1); return q ? q[0] : 0; }
    uint32_t u32() { uint32_t v = 0; const uint8_t *q = take(4); if (q) memcpy(&v, q, 4); return v; }
    uint64_t u64() { uint64_t v = 0; const uint8_t *q = take(8); if (q) memcpy(&v, q, 8); return v; }
    const uint8_t *vec(uint64_t elem, uint64_t *count) {
        uint64_t k = u64();
        if (bad || (elem && k > (n - off) / elem)) {
            bad = true;
            *count = 0;
            return nullptr;
        }
        *count = k;
        return take(k * elem);
    }
};

// alphabet.rs:174-193 (IupacDna) and :286-294 (ExactDna) as symbol bitmasks.
const uint16_t kIupacCompat[16] = {
    0x0000, 0xEA62, 0xD9A4, 0xB568, 0x76B0, 0xFFFE, 0xFF6A, 0xFFB4,
    0xFDEC, 0xFEF2, 0xF7F8, 0xFBE6, 0xFFFC, 0xFFFA, 0xFFF6, 0xFFEE,
};

uint8_t planes_for(uint8_t L) {
    if (L <= 1) return 0;
    uint8_t p = 0;
    while ((1u << p) < L) p++;
    return p;
}

int validate(const hsfm_index_view *v, const char *which) {
    const uint64_t n = v->text_len;
    if (n == 0 || n >= 0xffffffffull) return set_error(HSFM_EINVAL, "%s: text_len %llu out of range", which, (unsigned long long)n);
    const uint8_t L = v->occ_num_lanes;
    if (L == 0 || L > 16) return set_error(HSFM_EINVAL, "%s: num_lanes %u", which, L);
    if (v->occ_encoding > 1) return set_error(HSFM_EINVAL, "%s: occ encoding %u", which, v->occ_encoding);
    if (v->occ_num_planes != planes_for(L)) return set_error(HSFM_EINVAL, "%s: num_planes %u for %u lanes", which, v->occ_num_planes, L);
    const uint64_t W = v->occ_encoding == 0 ? v->occ_num_planes : L;
    if (v->occ_block_stride != 6ull * L + 8ull * W) return set_error(HSFM_EINVAL, "%s: block stride %llu", which, (unsigned long long)v->occ_block_stride);
    if (v->occ_text_len != n || v->sa_text_len != n) return set_error(HSFM_EINVAL, "%s: text_len mismatch", which);
    const uint64_t nb = (n + 63) / 64;
    if (v->occ_block_data_len != nb * v->occ_block_stride || !v->occ_block_data)
        return set_error(HSFM_EINVAL, "%s: block_data length %llu != %llu", which, (unsigned long long)v->occ_block_data_len, (unsigned long long)(nb * v->occ_block_stride));
    for (uint32_t l = 0; l < L; l++) {
        const uint8_t s = v->occ_lane_to_symbol[l];
        if (s >= 16 || v->occ_symbol_to_lane[s] != l) return set_error(HSFM_EINVAL, "%s: lane map inconsistent at lane %u", which, l);
        if (l && v->occ_lane_to_symbol[l - 1] >= s) return set_error(HSFM_EINVAL, "%s: lanes not in symbol order", which);
    }
    if (v->sa_word_data_len != nb * 14 || !v->sa_word_data) return set_error(HSFM_EINVAL, "%s: SA marker length", which);
    if (v->sa_vals_len == 0 || v->sa_vals_len > n || !v->sa_vals) return set_error(HSFM_EINVAL, "%s: sa_vals length", which);
    if (v->sa_sample_rate == 0) return set_error(HSFM_EINVAL, "%s: sample rate 0", which);
    if (v->num_sequences == 0 || v->seq_boundaries_len != v->num_sequences || !v->seq_boundaries)
        return set_error(HSFM_EINVAL, "%s: seq_boundaries", which);
    uint32_t last;
    memcpy(&last, (const uint8_t *)v->seq_boundaries + 4 * (v->seq_boundaries_len - 1), 4);
    if (last != n) return set_error(HSFM_EINVAL, "%s: last sequence boundary %u != text_len", which, last);
    if (v->has_lookup) {
        if (v->lookup_depth == 0 || v->lookup_core_len == 0 || v->lookup_core_len > 16 || !v->lookup_core_symbols)
            return set_error(HSFM_EINVAL, "%s: lookup table header", which);
        uint64_t sz = 1;
        for (uint32_t i = 0; i < v->lookup_depth; i++) {
            sz *= v->lookup_core_len;
            if (sz > (1ull << 34)) return set_error(HSFM_EINVAL, "%s: lookup table too large", which);
        }
        if (v->lookup_intervals_len != sz || !v->lookup_intervals) return set_error(HSFM_EINVAL, "%s: lookup intervals length", which);
    }
    return HSFM_OK;
}

// ---------------------------------------------------------------------------
// conversion kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rd32u(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint64_t rd64u(const uint8_t *p) {
    return (uint64_t)rd32u(p) | ((uint64_t)rd32u(p + 4) << 32);
}

struct ConvParams {
    uint64_t stride;
    uint32_t L, P, enc, nb_ref, nrec;
    uint32_t totals[16]; // per lane
};

// Absolute per-lane counts at the start of reference block bb (record fields
// sb_count + delta, src/occ/mod.rs:201-231), or the totals past the end.
__device__ void ref_abs(const uint8_t *ref, const ConvParams &c, uint32_t bb, uint32_t (&out)[16]) {
    if (bb < c.nb_ref) {
        const uint8_t *rec = ref + (uint64_t)bb * c.stride;
        for (uint32_t l = 0; l < 16; l++)
            out[l] = l < c.L ? rd32u(rec + 4 * l) + ((uint32_t)rec[4 * c.L + 2 * l] | ((uint32_t)rec[4 * c.L + 2 * l + 1] << 8)) : 0u;
    } else {
        for (uint32_t l = 0; l < 16; l++) out[l] = c.totals[l];
    }
}

__global__ void k_convert_occ(const uint8_t *__restrict__ ref, ConvParams c, uint4 *__restrict__ occ,
                              uint32_t *__restrict__ sb) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= c.nrec) return;
    uint32_t ab[16], as[16];
    ref_abs(ref, c, b, ab);
    const uint32_t s0 = b & ~((1u << kSbRecShift) - 1);
    ref_abs(ref, c, s0, as);
    uint32_t below_b[16], below_s[16];
    uint32_t accb = 0, accs = 0;
    for (int k = 0; k < 16; k++) {
        below_b[k] = accb;
        below_s[k] = accs;
        accb += ab[k];
        accs += as[k];
    }
    uint64_t pl[4] = {0, 0, 0, 0};
    if (b < c.nb_ref) {
        const uint8_t *rec = ref + (uint64_t)b * c.stride + 6ull * c.L;
        if (c.enc == 0) {
            for (uint32_t p = 0; p < c.P && p < 4; p++) pl[p] = rd64u(rec + 8 * p);
        } else {
            for (uint32_t l = 0; l < c.L; l++) {
                const uint64_t w = rd64u(rec + 8 * l);
                for (int p = 0; p < 4; p++)
                    if ((l >> p) & 1) pl[p] |= w;
            }
        }
    }
    uint32_t cum[16];
    for (int k = 0; k < 16; k++) cum[k] = k ? (below_b[k] - below_s[k]) & 0xffffu : 0u;
    uint4 *r = occ + ((size_t)b << 2);
    r[0] = make_uint4((uint32_t)pl[0], (uint32_t)(pl[0] >> 32), (uint32_t)pl[1], (uint32_t)(pl[1] >> 32));
    r[1] = make_uint4((uint32_t)pl[2], (uint32_t)(pl[2] >> 32), (uint32_t)pl[3], (uint32_t)(pl[3] >> 32));
    r[2] = make_uint4(cum[0] | (cum[1] << 16), cum[2] | (cum[3] << 16), cum[4] | (cum[5] << 16), cum[6] | (cum[7] << 16));
    r[3] = make_uint4(cum[8] | (cum[9] << 16), cum[10] | (cum[11] << 16), cum[12] | (cum[13] << 16), cum[14] | (cum[15] << 16));
    if ((b & ((1u << kSbRecShift) - 1)) == 0) {
        uint32_t *srow = sb + ((size_t)(b >> kSbRecShift) << 4);
        for (int k = 0; k < 16; k++) srow[k] = k ? below_b[k] : 0u;
    }
}

// src/suffix_array/mod.rs:97-110 record [word u64][sb u32][delta u16] -> uint4
__global__ void k_convert_samp(const uint8_t *__restrict__ words, uint32_t nw, uint4 *__restrict__ out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const uint8_t *rec = words + (uint64_t)w * 14;
    const uint64_t bits = rd64u(rec);
    const uint32_t base = rd32u(rec + 8) + ((uint32_t)rec[12] | ((uint32_t)rec[13] << 8));
    out[w] = make_uint4((uint32_t)bits, (uint32_t)(bits >> 32), base, 0u);
}

int upload_dir(hsfm_index *idx, const hsfm_index_view *v, DirBuffers &db, bool want_lookup) {
    hipStream_t s = idx->dev->stream;
    const uint32_t n = v->text_len;
    const uint32_t nb_ref = (uint32_t)((n + 63ull) / 64);
    const uint32_t nrec = (n >> kRecShift) + 1;
    const uint32_t nsb = ((nrec - 1) >> kSbRecShift) + 1;
    int rc;
    if ((rc = db.occ.ensure((size_t)nrec * 64))) return rc;
    if ((rc = db.sb.ensure((size_t)nsb * 64))) return rc;
    if ((rc = db.samp.ensure((size_t)nb_ref * 16))) return rc;
    if ((rc = db.sa_vals.ensure(v->sa_vals_len * 4))) return rc;

    DevBuf staging;
    if ((rc = staging.ensure(v->occ_block_data_len))) return rc;
    HSFM_HIP(hipMemcpyAsync(staging.p, v->occ_block_data, v->occ_block_data_len, hipMemcpyHostToDevice, s));
    ConvParams c{};
    c.stride = v->occ_block_stride;
    c.L = v->occ_num_lanes;
    c.P = v->occ_num_planes;
    c.enc = v->occ_encoding;
    c.nb_ref = nb_ref;
    c.nrec = nrec;
    for (uint32_t l = 0; l < c.L; l++) {
        const uint8_t sym = v->occ_lane_to_symbol[l];
        const uint32_t up = sym + 1 < 16 ? v->c_array[sym + 1] : n;
        c.totals[l] = up - v->c_array[sym];
    }
    hipLaunchKernelGGL(k_convert_occ, dim3((nrec + 255) / 256), dim3(256), 0, s, staging.as<uint8_t>(), c,
                       db.occ.as<uint4>(), db.sb.as<uint32_t>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipStreamSynchronize(s));
    staging.release();

    if ((rc = staging.ensure(v->sa_word_data_len))) return rc;
    HSFM_HIP(hipMemcpyAsync(staging.p, v->sa_word_data, v->sa_word_data_len, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_convert_samp, dim3((nb_ref + 255) / 256), dim3(256), 0, s, staging.as<uint8_t>(), nb_ref,
                       db.samp.as<uint4>());
    HSFM_HIP(hipGetLastError());
    HSFM_HIP(hipMemcpyAsync(db.sa_vals.p, v->sa_vals, v->sa_vals_len * 4, hipMemcpyHostToDevice, s));
    HSFM_HIP(hipStreamSynchronize(s));
    staging.release();

    db.view = DevDir{};
    if (want_lookup && v->has_lookup) {
        if ((rc = db.lookup.ensure(v->lookup_intervals_len * 8))) return rc;
        HSFM_HIP(hipMemcpyAsync(db.lookup.p, v->lookup_intervals, v->lookup_intervals_len * 8, hipMemcpyHostToDevice, s));
        HSFM_HIP(hipStreamSynchronize(s));
        db.view.lookup = db.lookup.as<uint2>();
        db.view.lookup_len = v->lookup_intervals_len;
    }
    db.view.occ = db.occ.as<uint4>();
    db.view.sb = db.sb.as<uint32_t>();
    db.view.samp = db.samp.as<uint4>();
    db.view.sa_vals = db.sa_vals.as<uint32_t>();
    db.view.n = n;
    db.view.num_samples = (uint32_t)v->sa_vals_len;
    db.view.sample_rate = v->sa_sample_rate;
    db.view.num_lanes = v->occ_num_lanes;
    return HSFM_OK;
}

void free_dir(DirBuffers &d) {
    d.occ.release();
    d.sb.release();
    d.samp.release();
    d.sa_vals.release();
    d.lookup.release();
    d.sa_full.release();
    d.text.release();
    d.text4.release();
    d.isa.release();
    d.sa_full_tried = false;
    d.view.sa_full = nullptr;
    d.view.text = nullptr;
    d.view.text4 = nullptr;
    d.view.isa = nullptr;
}

void free_workspace(Workspace &w) {
    DevBuf *all[] = {&w.qcopy, &w.qoffcopy, &w.refb, &w.cnt, &w.niv, &w.ivs, &w.status, &w.ovf_list,
                     &w.counters, &w.fb_sets, &w.fb_ivs, &w.fb_meta, &w.hit_off, &w.rows, &w.hits,
                     &w.scan_tmp, &w.mem_recs, &w.mem_cnt, &w.mem_off, &w.mem_sorted, &w.mem_hits,
                     &w.pos_off, &w.positions, &w.mem_flags, &w.read_list, &w.wave_sets,
                     &w.mem_slots, &w.mem_status, &w.items, &w.item_cnt, &w.item_first, &w.read_off,
                     &w.qnib, &w.defer_ids, &w.part_cnt, &w.item_tot};
    for (DevBuf *b : all) b->release();
    if (w.h_pinned) (void)hipHostFree(w.h_pinned);
    w.h_pinned = nullptr;
    for (hipEvent_t &e : w.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

} // namespace

extern "C" int hsfm_index_view_from_bytes(const uint8_t *bytes, uint64_t len, hsfm_index_view *out) {
    if (!bytes || !out) return set_error(HSFM_EINVAL, "hsfm_index_view_from_bytes: null argument");
    Reader r{bytes, len, 0, false};
    hsfm_index_view v;
    memset(&v, 0, sizeof v);
    for (int i = 0; i < 16; i++) v.c_array[i] = r.u32();
    v.occ_num_lanes = r.u8();
    v.occ_num_planes = r.u8();
    const uint8_t *q = r.take(16);
    if (q) memcpy(v.occ_symbol_to_lane, q, 16);
    q = r.take(16);
    if (q) memcpy(v.occ_lane_to_symbol, q, 16);
    v.occ_superblock_checkpoints = (const uint32_t *)r.vec(4, &v.occ_superblock_len);
    v.occ_block_data = r.vec(1, &v.occ_block_data_len);
    v.occ_block_stride = r.u64();
    v.occ_encoding = r.u32();
    v.occ_text_len = r.u32();
    v.sa_word_data = r.vec(1, &v.sa_word_data_len);
    v.sa_vals = (const uint32_t *)r.vec(4, &v.sa_vals_len);
    v.sa_sample_rate = r.u32();
    v.sa_text_len = r.u32();
    v.text_len = r.u32();
    v.num_sequences = r.u32();
    v.seq_boundaries = (const uint32_t *)r.vec(4, &v.seq_boundaries_len);
    uint64_t nh = r.u64();
    if (!r.bad && nh > (r.n - r.off) / 8) r.bad = true;
    for (uint64_t i = 0; i < nh && !r.bad; i++) {
        uint64_t l = r.u64();
        r.take(l);
    }
    const uint8_t opt = r.u8();
    if (opt == 1) {
        v.has_lookup = 1;
        v.lookup_depth = r.u32();
        v.lookup_core_symbols = r.vec(1, &v.lookup_core_len);
        v.lookup_intervals = (const uint32_t *)r.vec(8, &v.lookup_intervals_len);
    } else if (opt != 0) {
        r.bad = true;
    }
    v.alphabet_tag = r.u8();
    if (r.bad) return set_error(HSFM_EDESER, "deserialization failed: truncated or malformed FM-index byte stream");
    if (v.alphabet_tag > 1)
        return set_error(HSFM_EDESER, "deserialization failed: unknown alphabet tag %u in serialized FM-index", v.alphabet_tag);
    for (int b = 0; b < 256; b++) {
        uint16_t m = 0;
        if (v.alphabet_tag == 0) m = b < 16 ? kIupacCompat[b] : 0;
        else m = (b >= 1 && b <= 4) ? (uint16_t)(1u << b) : 0;
        v.compat_mask[b] = m;
    }
    *out = v;
    return HSFM_OK;
}

extern "C" int hsfm_index_upload(hsfm_device *dev, const hsfm_index_view *fwd, const hsfm_index_view *rev,
                                 hsfm_index **out) {
    if (!dev || !fwd || !out) return set_error(HSFM_EINVAL, "hsfm_index_upload: null argument");
    int rc = validate(fwd, "forward index");
    if (rc) return rc;
    if (rev) {
        if ((rc = validate(rev, "reverse index"))) return rc;
        if (rev->text_len != fwd->text_len || rev->occ_num_lanes != fwd->occ_num_lanes ||
            memcmp(rev->occ_lane_to_symbol, fwd->occ_lane_to_symbol, 16) != 0 ||
            memcmp(rev->c_array, fwd->c_array, 64) != 0)
            return set_error(HSFM_EINVAL, "reverse index does not match the forward index alphabet/length");
        if (memcmp(rev->compat_mask, fwd->compat_mask, sizeof fwd->compat_mask) != 0)
            return set_error(HSFM_EINVAL, "forward and reverse index use different alphabets");
    }
    HSFM_HIP(hipSetDevice(dev->ordinal));
    hsfm_index *idx = new hsfm_index();
    idx->dev = dev;
    rc = upload_dir(idx, fwd, idx->fwd, true);
    if (!rc && rev) {
        rc = upload_dir(idx, rev, idx->rev, false);
        idx->has_rev = true;
    }
    if (!rc) {
        DevConst &c = idx->h_const;
        memset(&c, 0, sizeof c);
        const uint32_t L = fwd->occ_num_lanes;
        c.num_lanes = L;
        for (uint32_t l = 0; l < L; l++) {
            c.lane_sym[l] = fwd->occ_lane_to_symbol[l];
            c.c_lane[l] = fwd->c_array[c.lane_sym[l]];
        }
        for (int b = 0; b < 256; b++) {
            uint16_t m = 0;
            for (uint32_t l = 0; l < L; l++)
                if ((fwd->compat_mask[b] >> c.lane_sym[l]) & 1) m |= (uint16_t)(1u << l);
            c.compat_lanes[b] = m;
            c.lookup_digit[b] = -1;
        }
        // seed-table expansion: ACGT map to digits; another byte expands to the
        // bases whose compat sets are contained in its own, when their union is
        // exactly its compat set (true for every IupacDna ambiguity code)
        for (int b = 0; b < 256; b++) {
            uint8_t e = 0;
            if (b >= 1 && b <= 4) {
                e = (uint8_t)(0x10 | (b - 1));
            } else {
                const uint16_t mb = fwd->compat_mask[b];
                uint16_t uni = 0;
                uint8_t bases = 0;
                for (int x = 0; x < 4; x++) {
                    const uint16_t mx = fwd->compat_mask[x + 1];
                    if (mx && (mx & ~mb) == 0) {
                        bases |= (uint8_t)(1u << x);
                        uni |= mx;
                    }
                }
                if (bases && uni == mb) e = bases;
            }
            c.seed_expand[b] = e;
        }
        // text lanes as base sets, and whether nibble comparison equals the
        // compat table for every query byte
        for (uint32_t l = 0; l < L && l < 16; l++) {
            uint8_t tn = 0;
            for (int x = 0; x < 4; x++)
                if ((c.compat_lanes[x + 1] >> l) & 1u) tn |= (uint8_t)(1u << x);
            c.text_nib[l] = tn;
        }
        c.nib_text = 1;
        for (int b = 0; b < 256 && c.nib_text; b++) {
            const uint32_t qn = (b >= 1 && b <= 4) ? 1u << (b - 1) : (uint32_t)(c.seed_expand[b] & 15u);
            for (uint32_t l = 0; l < L && l < 16; l++)
                if ((((c.compat_lanes[b] >> l) & 1u) != 0) != ((qn & c.text_nib[l]) != 0)) c.nib_text = 0;
        }
        if (fwd->has_lookup) {
            c.lookup_depth = fwd->lookup_depth;
            c.lookup_radix = (uint32_t)fwd->lookup_core_len;
            for (uint64_t i = fwd->lookup_core_len; i-- > 0;) c.lookup_digit[fwd->lookup_core_symbols[i]] = (int8_t)i;
        }
        rc = idx->d_const.ensure(sizeof(DevConst));
        if (!rc) {
            hipError_t e = hipMemcpy(idx->d_const.p, &c, sizeof c, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = set_error(HSFM_EDEVICE, "upload constants: %s", hipGetErrorString(e));
        }
        idx->nseq = (uint32_t)fwd->seq_boundaries_len;
        if (!rc) rc = idx->seq_bounds.ensure((size_t)idx->nseq * 4);
        if (!rc) {
            hipError_t e = hipMemcpy(idx->seq_bounds.p, fwd->seq_boundaries, (size_t)idx->nseq * 4, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = set_error(HSFM_EDEVICE, "upload boundaries: %s", hipGetErrorString(e));
        }
    }
    if (rc) {
        free_dir(idx->fwd);
        free_dir(idx->rev);
        idx->d_const.release();
        idx->seq_bounds.release();
        delete idx;
        return rc;
    }
    *out = idx;
    return HSFM_OK;
}

extern "C" int hsfm_index_free(hsfm_index *idx) {
    if (!idx) return HSFM_OK;
    if (tl_callback_index == idx)
        return set_error(HSFM_EINVAL, "hsfm_index_free: called from a batch callback on the same index");
    (void)hipSetDevice(idx->dev->ordinal);
    (void)hipStreamSynchronize(idx->dev->stream);
    for (hipStream_t bs : idx->bstream)
        if (bs) (void)hipStreamSynchronize(bs);
    free_dir(idx->fwd);
    free_dir(idx->rev);
    idx->d_const.release();
    idx->seq_bounds.release();
    free_workspace(idx->ws);
    idx->seed_off.release();
    idx->seed_iv.release();
    idx->seed_lsl.release();
    idx->seed_rec12.release();
    idx->seed_rec14.release();
    idx->seed_iv14.release();
    idx->seed_rec15.release();
    idx->seed_iv15.release();
    for (int k = 0; k < hsfm_index::kMaxBatchSets; k++) {
        free_batch_bufs(idx->batch[k]);
        if (idx->bstream[k]) (void)hipStreamDestroy(idx->bstream[k]);
    }
    delete idx;
    return HSFM_OK;
}
