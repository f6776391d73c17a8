// capi_gf256.cpp - C ABI: profiles, GF(2^8) batch encoder and batch decoder.
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/scrs.h"
#include "device.hpp"
#include "kernels.hpp"
#include "matrices.hpp"
#include "status.hpp"

using namespace scrs;

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------

struct scrs_encode_scratch {
    int32_t engine;
    uint64_t k, m, L;
    DevBuf d_data, d_rep;
    HostBuf h_data, h_rep;
};

struct scrs_batch_encoder {
    scrs_profile prof{};
    int device = 0;
    uint32_t groups = 0;
    DevBuf d_tables;
    hipStream_t stream = nullptr;
    std::mutex mu;            // guards `own` for scratch-less host-slice calls
    scrs_encode_scratch own{};
};

struct DecodeWorkspace {
    DevBuf d_idx, d_tables, d_rows, d_src, d_dst, d_copy, d_flag, d_payload_ptrs;
    HostBuf h_idx;
    hipEvent_t idx_copied = nullptr;
    uint64_t cap = 0;  // codewords the buffers hold
    ~DecodeWorkspace() {
        if (idx_copied) (void)hipEventDestroy(idx_copied);
    }
};

struct scrs_decode_scratch {
    int32_t engine;
    uint64_t k, m, L;
    DevBuf d_payload, d_out;
    HostBuf h_payload, h_out;
    DecodeWorkspace ws;
};

struct scrs_batch_decoder {
    scrs_profile prof{};
    int device = 0;
    uint32_t rmax = 0, groups = 0;
    DevBuf d_coeffs;
    hipStream_t stream = nullptr;
    DecodeWorkspace ws;          // batched device calls
    scrs_decode_scratch own{};   // scratch-less host-slice calls
};

namespace {

bool is_gf256(int32_t engine) {
    return engine == SCRS_ENGINE_STANDARD_CAUCHY || engine == SCRS_ENGINE_GOOD_CAUCHY;
}

size_t engine_capacity(int32_t engine) {
    switch (engine) {
        case SCRS_ENGINE_STANDARD_CAUCHY: return kCapStandard;
        case SCRS_ENGINE_GOOD_CAUCHY: return kCapGood;
        case SCRS_ENGINE_TOWER: return kCapTower;
        default: return kCapAfft;
    }
}

std::vector<uint8_t> gf256_matrix(const scrs_profile &p) {
    return p.engine == SCRS_ENGINE_GOOD_CAUCHY ? good_cauchy_matrix(p.k, p.m)
                                               : standard_cauchy_matrix(p.k, p.m);
}

// Device OOM while building a codec maps onto TooManySymbols{cap}, the
// reference's allocation-failure precedent (tower/encoder.rs:37-44).
scrs_status construction_error(hipError_t err, int32_t engine) {
    if (err == hipErrorOutOfMemory) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(engine));
    return st(SCRS_ERR_DEVICE, static_cast<uint64_t>(err));
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

hipError_t ensure_workspace(DecodeWorkspace &ws, uint64_t count, uint32_t k, uint32_t rmax,
                            uint32_t groups) {
    hipError_t err = hipSuccess;
    if (!ws.idx_copied) {
        err = hipEventCreateWithFlags(&ws.idx_copied, hipEventDisableTiming);
        if (err != hipSuccess) return err;
    }
    if (count <= ws.cap) return hipSuccess;
    const size_t kk = k, rr = rmax ? rmax : 1;
    if ((err = ws.d_idx.ensure(count * kk * sizeof(uint16_t))) != hipSuccess) return err;
    if ((err = ws.d_tables.ensure(count * kk * groups * 128 * sizeof(uint32_t))) != hipSuccess) return err;
    if ((err = ws.d_rows.ensure(count * sizeof(uint32_t))) != hipSuccess) return err;
    if ((err = ws.d_src.ensure(count * kk * sizeof(void *))) != hipSuccess) return err;
    if ((err = ws.d_dst.ensure(count * rr * sizeof(void *))) != hipSuccess) return err;
    if ((err = ws.d_copy.ensure(count * kk * sizeof(void *))) != hipSuccess) return err;
    if ((err = ws.d_flag.ensure(sizeof(uint32_t))) != hipSuccess) return err;
    if ((err = ws.h_idx.ensure(count * kk * sizeof(uint16_t))) != hipSuccess) return err;
    ws.cap = count;
    return hipSuccess;
}

// Validate one codeword's received indices with the reference's order
// (batch/codec.rs:304-317) minus the payload-length check.
scrs_status check_indices_u16(const uint16_t *idx, uint64_t k, uint64_t n) {
    unsigned char seen[65536 / 8];
    std::memset(seen, 0, (n + 7) / 8);
    for (uint64_t t = 0; t < k; t++) {
        const uint64_t i = idx[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i >> 3] & (1u << (i & 7))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
        seen[i >> 3] |= static_cast<unsigned char>(1u << (i & 7));
    }
    return ok();
}

// Enqueue solve + multiply-XOR for `count` codewords whose indices are
// already validated.  Either `payload_ptrs` (device array) or the data /
// repair layout bases name the received payloads.
scrs_status enqueue_gf256_decode(const scrs_batch_decoder *dec, DecodeWorkspace &ws,
                                 const uint16_t *indices, const uint8_t *const *payload_ptrs,
                                 const uint8_t *data_base, const uint8_t *repair_base,
                                 uint8_t *out, uint64_t count, hipStream_t s, bool vec16) {
    const uint32_t k = static_cast<uint32_t>(dec->prof.k), m = static_cast<uint32_t>(dec->prof.m);
    const uint64_t L = dec->prof.symbol_len;
    const uint64_t chunk_max = 65535;
    hipError_t err = ensure_workspace(ws, count < chunk_max ? count : chunk_max, k, dec->rmax, dec->groups);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    for (uint64_t c0 = 0; c0 < count; c0 += chunk_max) {
        const uint64_t cn = count - c0 < chunk_max ? count - c0 : chunk_max;
        // The pinned index buffer is reused: wait for the previous upload.
        if ((err = hipEventSynchronize(ws.idx_copied)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        std::memcpy(ws.h_idx.ptr, indices + c0 * k, cn * k * sizeof(uint16_t));
        if ((err = hipMemcpyAsync(ws.d_idx.ptr, ws.h_idx.ptr, cn * k * sizeof(uint16_t),
                                  hipMemcpyHostToDevice, s)) != hipSuccess)
            return st(SCRS_ERR_DEVICE, err);
        if ((err = hipEventRecord(ws.idx_copied, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        DecodeSolveArgs sa;
        sa.indices = ws.d_idx.as<uint16_t>();
        sa.coeffs = dec->d_coeffs.as<uint8_t>();
        sa.payload_ptrs = payload_ptrs ? payload_ptrs + c0 * k : nullptr;
        sa.data_base = data_base ? data_base + c0 * k * L : nullptr;
        sa.repair_base = repair_base ? repair_base + c0 * m * L : nullptr;
        sa.out_base = out + c0 * k * L;
        sa.tables = ws.d_tables.as<uint32_t>();
        sa.rows = ws.d_rows.as<uint32_t>();
        sa.src_ptrs = ws.d_src.as<const uint8_t *>();
        sa.dst_ptrs = ws.d_dst.as<uint8_t *>();
        sa.copy_ptrs = ws.d_copy.as<uint8_t *>();
        sa.error_flag = ws.d_flag.as<uint32_t>();
        sa.k = k;
        sa.m = m;
        sa.rmax = dec->rmax;
        sa.groups = dec->groups;
        sa.L = L;
        sa.count = cn;
        if ((err = launch_decode_solve_gf256(sa, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        RowsTermsArgs ra;
        ra.src_ptrs = ws.d_src.as<const uint8_t *const>();
        ra.dst_ptrs = ws.d_dst.as<uint8_t *const>();
        ra.copy_ptrs = ws.d_copy.as<uint8_t *const>();
        ra.rows = ws.d_rows.as<uint32_t>();
        ra.tables = ws.d_tables.as<uint32_t>();
        ra.table_cw_stride = static_cast<uint64_t>(k) * dec->groups * 128;
        ra.T = k;
        ra.rmax = dec->rmax;
        ra.groups = dec->groups;
        ra.L = L;
        ra.count = cn;
        if ((err = launch_rows_terms_lds(ra, vec16, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    }
    return ok();
}

}  // namespace

// ---------------------------------------------------------------------------
// Library / profiles
// ---------------------------------------------------------------------------

extern "C" uint32_t scrs_abi_version(void) { return 1; }

extern "C" int32_t scrs_engine_field(int32_t engine) {
    switch (engine) {
        case SCRS_ENGINE_STANDARD_CAUCHY:
        case SCRS_ENGINE_GOOD_CAUCHY: return SCRS_FIELD_GF256;
        case SCRS_ENGINE_TOWER:
        case SCRS_ENGINE_AFFT: return SCRS_FIELD_GF65536;
        default: return -1;
    }
}

extern "C" int32_t scrs_recommended_gf16_engine(uint64_t k, uint64_t m) {
    const bool large_block = (k + m < k) || k + m > 256;  // saturating add
    const bool high_redundancy = (m > UINT64_MAX / 3) || m * 3 >= k;
    return (large_block || high_redundancy) ? SCRS_ENGINE_AFFT : SCRS_ENGINE_TOWER;
}

extern "C" scrs_status scrs_profile_resolve(int32_t engine, uint64_t k, uint64_t m,
                                            uint64_t symbol_len, scrs_profile *out) {
    const int32_t field = scrs_engine_field(engine);
    if (field < 0) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 5);
    if (k == 0 || m == 0) return st(SCRS_CONFIG_ZERO_DIMENSION);
    if (symbol_len == 0) return st(SCRS_CONFIG_ZERO_SYMBOL_LEN);
    if (field == SCRS_FIELD_GF65536 && symbol_len % 2 != 0) return st(SCRS_CONFIG_ODD_SYMBOL_LEN);
    const uint64_t cap = engine_capacity(engine);
    if (k + m < k || k + m > cap) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, cap);
    *out = scrs_profile{engine, k, m, symbol_len};
    return ok();
}

extern "C" scrs_status scrs_profile_recommended(int32_t field, uint64_t k, uint64_t m,
                                                uint64_t symbol_len, scrs_profile *out) {
    int32_t engine;
    if (field == SCRS_FIELD_GF256)
        engine = (k + m >= k && k + m <= 255) ? SCRS_ENGINE_GOOD_CAUCHY : SCRS_ENGINE_STANDARD_CAUCHY;
    else if (field == SCRS_FIELD_GF65536)
        engine = scrs_recommended_gf16_engine(k, m);
    else
        return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    return scrs_profile_resolve(engine, k, m, symbol_len, out);
}

extern "C" scrs_status scrs_fill_random_dev(uint8_t *dst_dev, uint64_t len, uint64_t seed,
                                            uint64_t byte_offset, void *stream) {
    if (!dst_dev && len) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (byte_offset % 8) return st(SCRS_ERR_INVALID_ARGUMENT, 4);
    hipError_t err = launch_fill_splitmix(dst_dev, len, seed, byte_offset, static_cast<hipStream_t>(stream));
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

// ---------------------------------------------------------------------------
// Batch encoder
// ---------------------------------------------------------------------------

scrs_status gf256_encoder_init(scrs_batch_encoder *enc) {
    const scrs_profile &p = enc->prof;
    const std::vector<uint8_t> coeffs = gf256_matrix(p);  // k x m, term-major already
    enc->groups = static_cast<uint32_t>((p.m + 15) / 16);
    std::vector<uint32_t> tables(p.k * enc->groups * 128);
    build_nibble_tables(coeffs.data(), p.k, p.m, enc->groups, tables.data());
    hipError_t err = enc->d_tables.ensure(tables.size() * sizeof(uint32_t));
    if (err != hipSuccess) return construction_error(err, p.engine);
    err = hipMemcpy(enc->d_tables.ptr, tables.data(), tables.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (err != hipSuccess) return construction_error(err, p.engine);
    return ok();
}

extern "C" scrs_status scrs_batch_encoder_create(const scrs_profile *profile, scrs_batch_encoder **out) {
    if (!profile) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(profile->engine, profile->k, profile->m, profile->symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    if (p.engine == SCRS_ENGINE_TOWER) return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    if (!is_gf256(p.engine)) return scrs_afft_encoder_create_impl(&p, out);
    auto *enc = new scrs_batch_encoder();
    enc->prof = p;
    (void)hipGetDevice(&enc->device);
    hipError_t err = hipStreamCreateWithFlags(&enc->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete enc;
        return construction_error(err, p.engine);
    }
    enc->own.engine = p.engine;
    enc->own.k = p.k;
    enc->own.m = p.m;
    enc->own.L = p.symbol_len;
    s = gf256_encoder_init(enc);
    if (s.code != SCRS_OK) {
        scrs_batch_encoder_destroy(enc);
        return s;
    }
    *out = enc;
    return ok();
}

extern "C" void scrs_batch_encoder_destroy(scrs_batch_encoder *enc) {
    if (!enc) return;
    if (enc->prof.engine == SCRS_ENGINE_AFFT) {
        scrs_afft_encoder_destroy_impl(enc);
        return;
    }
    DeviceGuard g(enc->device);
    if (enc->stream) (void)hipStreamSynchronize(enc->stream);
    if (enc->stream) (void)hipStreamDestroy(enc->stream);
    delete enc;
}

extern "C" scrs_status scrs_batch_encoder_profile(const scrs_batch_encoder *enc, scrs_profile *out) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = enc->prof;
    return ok();
}

extern "C" scrs_status scrs_encode_scratch_create(const scrs_batch_encoder *enc, scrs_encode_scratch **out) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    auto *s = new scrs_encode_scratch();
    s->engine = enc->prof.engine;
    s->k = enc->prof.k;
    s->m = enc->prof.m;
    s->L = enc->prof.symbol_len;
    *out = s;
    return ok();
}

extern "C" void scrs_encode_scratch_destroy(scrs_encode_scratch *scratch) { delete scratch; }

static scrs_status gf256_encode_host(const scrs_batch_encoder *enc, const uint8_t *data,
                                     uint8_t *repairs, scrs_encode_scratch &sc) {
    const uint64_t k = enc->prof.k, m = enc->prof.m, L = enc->prof.symbol_len;
    hipError_t err;
    if ((err = sc.h_data.ensure(k * L)) != hipSuccess || (err = sc.d_data.ensure(k * L)) != hipSuccess ||
        (err = sc.h_rep.ensure(m * L)) != hipSuccess || (err = sc.d_rep.ensure(m * L)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    std::memcpy(sc.h_data.ptr, data, k * L);
    hipStream_t s = enc->stream;
    if ((err = hipMemcpyAsync(sc.d_data.ptr, sc.h_data.ptr, k * L, hipMemcpyHostToDevice, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    scrs_status r = scrs_encode_batch_dev(enc, sc.d_data.as<uint8_t>(), sc.d_rep.as<uint8_t>(), 1, s);
    if (r.code != SCRS_OK) return r;
    if ((err = hipMemcpyAsync(sc.h_rep.ptr, sc.d_rep.ptr, m * L, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if ((err = hipStreamSynchronize(s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    std::memcpy(repairs, sc.h_rep.ptr, m * L);
    return ok();
}

static scrs_status check_encode_lens(const scrs_profile &p, uint64_t data_len, uint64_t repairs_len) {
    const uint64_t din = p.k * p.symbol_len, dout = p.m * p.symbol_len;
    if (data_len != din) return st(SCRS_ENCODE_WRONG_INPUT_LEN, din, data_len);
    if (repairs_len != dout) return st(SCRS_ENCODE_WRONG_OUTPUT_LEN, dout, repairs_len);
    return ok();
}

extern "C" scrs_status scrs_encode_into(const scrs_batch_encoder *enc, const uint8_t *data,
                                        uint64_t data_len, uint8_t *repairs, uint64_t repairs_len) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (enc->prof.engine == SCRS_ENGINE_AFFT)
        return scrs_afft_encode_into_impl(enc, data, data_len, repairs, repairs_len, nullptr);
    scrs_status s = check_encode_lens(enc->prof, data_len, repairs_len);
    if (s.code != SCRS_OK) return s;
    if ((!data && data_len) || (!repairs && repairs_len)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    auto *mut = const_cast<scrs_batch_encoder *>(enc);
    std::lock_guard<std::mutex> lock(mut->mu);
    DeviceGuard g(enc->device);
    return gf256_encode_host(enc, data, repairs, mut->own);
}

extern "C" scrs_status scrs_encode_into_with(const scrs_batch_encoder *enc, const uint8_t *data,
                                             uint64_t data_len, uint8_t *repairs, uint64_t repairs_len,
                                             scrs_encode_scratch *scratch) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 6);
    // The type-erased layer matches the scratch variant first
    // (selector.rs:516-527): GF(256) codecs take the unit scratch of any
    // GF(256) geometry, AFFT takes only an AFFT scratch.
    const bool scratch_gf256 = is_gf256(scratch->engine);
    if (is_gf256(enc->prof.engine) != scratch_gf256) return st(SCRS_ENCODE_SCRATCH_MISMATCH);
    if (enc->prof.engine == SCRS_ENGINE_AFFT)
        return scrs_afft_encode_into_impl(enc, data, data_len, repairs, repairs_len, scratch);
    scrs_status s = check_encode_lens(enc->prof, data_len, repairs_len);
    if (s.code != SCRS_OK) return s;
    if ((!data && data_len) || (!repairs && repairs_len)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    DeviceGuard g(enc->device);
    return gf256_encode_host(enc, data, repairs, *scratch);
}

extern "C" scrs_status scrs_encode_batch_dev(const scrs_batch_encoder *enc, const uint8_t *data_dev,
                                             uint8_t *repairs_dev, uint64_t count, void *stream) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count == 0) return ok();
    if (!data_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!repairs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    if (enc->prof.engine == SCRS_ENGINE_AFFT)
        return scrs_afft_encode_batch_dev_impl(enc, data_dev, repairs_dev, count, stream);
    const uint64_t k = enc->prof.k, m = enc->prof.m, L = enc->prof.symbol_len;
    const bool vec16 = aligned16(data_dev) && aligned16(repairs_dev) && L % 16 == 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (uint64_t c0 = 0; c0 < count; c0 += 65535) {
        RowsTermsArgs a;
        a.src_base = data_dev + c0 * k * L;
        a.src_cw_stride = k * L;
        a.dst_base = repairs_dev + c0 * m * L;
        a.dst_cw_stride = m * L;
        a.tables = enc->d_tables.as<uint32_t>();
        a.table_cw_stride = 0;
        a.T = static_cast<uint32_t>(k);
        a.rmax = static_cast<uint32_t>(m);
        a.groups = enc->groups;
        a.L = L;
        a.count = count - c0 < 65535 ? count - c0 : 65535;
        hipError_t err = launch_rows_terms_lds(a, vec16, s);
        if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    }
    return ok();
}

// ---------------------------------------------------------------------------
// Batch decoder
// ---------------------------------------------------------------------------

extern "C" scrs_status scrs_batch_decoder_create(const scrs_profile *profile, scrs_batch_decoder **out) {
    if (!profile) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(profile->engine, profile->k, profile->m, profile->symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    if (!is_gf256(p.engine)) return scrs_gf16_batch_decoder_create_impl(&p, out);
    auto *dec = new scrs_batch_decoder();
    dec->prof = p;
    (void)hipGetDevice(&dec->device);
    const uint64_t rmax = p.k < p.m ? p.k : p.m;
    dec->rmax = static_cast<uint32_t>(rmax);
    dec->groups = static_cast<uint32_t>((rmax + 15) / 16);
    dec->own.engine = p.engine;
    dec->own.k = p.k;
    dec->own.m = p.m;
    dec->own.L = p.symbol_len;
    hipError_t err = hipStreamCreateWithFlags(&dec->stream, hipStreamNonBlocking);
    if (err == hipSuccess) {
        const std::vector<uint8_t> coeffs = gf256_matrix(p);
        err = dec->d_coeffs.ensure(coeffs.size());
        if (err == hipSuccess)
            err = hipMemcpy(dec->d_coeffs.ptr, coeffs.data(), coeffs.size(), hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        scrs_batch_decoder_destroy(dec);
        return construction_error(err, p.engine);
    }
    *out = dec;
    return ok();
}

extern "C" void scrs_batch_decoder_destroy(scrs_batch_decoder *dec) {
    if (!dec) return;
    if (!is_gf256(dec->prof.engine)) {
        scrs_gf16_batch_decoder_destroy_impl(dec);
        return;
    }
    DeviceGuard g(dec->device);
    if (dec->stream) {
        (void)hipStreamSynchronize(dec->stream);
        (void)hipStreamDestroy(dec->stream);
    }
    delete dec;
}

extern "C" scrs_status scrs_batch_decoder_profile(const scrs_batch_decoder *dec, scrs_profile *out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = dec->prof;
    return ok();
}

extern "C" scrs_status scrs_decode_scratch_create(const scrs_batch_decoder *dec, scrs_decode_scratch **out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    auto *s = new scrs_decode_scratch();
    s->engine = dec->prof.engine;
    s->k = dec->prof.k;
    s->m = dec->prof.m;
    s->L = dec->prof.symbol_len;
    *out = s;
    return ok();
}

extern "C" void scrs_decode_scratch_destroy(scrs_decode_scratch *scratch) { delete scratch; }

static scrs_status gf256_decode_host(scrs_batch_decoder *dec, const uint64_t *indices,
                                     const uint8_t *const *payloads, const uint64_t *payload_lens,
                                     uint64_t count, uint8_t *out, uint64_t out_len,
                                     scrs_decode_scratch &sc) {
    const uint64_t k = dec->prof.k, m = dec->prof.m, n = k + m, L = dec->prof.symbol_len;
    if (out_len != k * L) return st(SCRS_DECODE_WRONG_OUTPUT_LEN, k * L, out_len);
    if (count != k) return st(SCRS_DECODE_WRONG_COUNT, k, count);
    if (!indices || !payloads || !payload_lens) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    bool seen[256] = {false};
    uint64_t e = 0;
    std::vector<uint16_t> idx16(k);
    for (uint64_t t = 0; t < count; t++) {
        const uint64_t i = indices[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i]) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
        seen[i] = true;
        if (payload_lens[t] != L) return st(SCRS_DECODE_WRONG_PAYLOAD_LEN, L, payload_lens[t]);
        if (!payloads[t]) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
        idx16[t] = static_cast<uint16_t>(i);
        e += i >= k;
    }
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 6);
    DeviceGuard g(dec->device);
    hipError_t err;
    if ((err = sc.h_payload.ensure(k * L)) != hipSuccess || (err = sc.d_payload.ensure(k * L)) != hipSuccess ||
        (err = sc.h_out.ensure(k * L)) != hipSuccess || (err = sc.d_out.ensure(k * L)) != hipSuccess ||
        (err = sc.ws.d_payload_ptrs.ensure(k * sizeof(void *))) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    hipStream_t s = dec->stream;
    // Received payloads packed back to back: payload t at d_payload + t*L.
    for (uint64_t t = 0; t < k; t++) std::memcpy(sc.h_payload.as<uint8_t>() + t * L, payloads[t], L);
    std::vector<const uint8_t *> ptrs(k);
    for (uint64_t t = 0; t < k; t++) ptrs[t] = sc.d_payload.as<uint8_t>() + t * L;
    if ((err = hipMemcpyAsync(sc.d_payload.ptr, sc.h_payload.ptr, k * L, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (err = hipMemcpyAsync(sc.ws.d_payload_ptrs.ptr, ptrs.data(), k * sizeof(void *), hipMemcpyHostToDevice, s)) !=
            hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if ((err = ensure_workspace(sc.ws, 1, static_cast<uint32_t>(k), dec->rmax, dec->groups)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if ((err = hipMemsetAsync(sc.ws.d_flag.ptr, 0, sizeof(uint32_t), s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    const bool vec16 = L % 16 == 0;
    scrs_status r = enqueue_gf256_decode(dec, sc.ws, idx16.data(), sc.ws.d_payload_ptrs.as<const uint8_t *const>(),
                                         nullptr, nullptr, sc.d_out.as<uint8_t>(), 1, s, vec16);
    if (r.code != SCRS_OK) return r;
    uint32_t flag = 0;
    if ((err = hipMemcpyAsync(sc.h_out.ptr, sc.d_out.ptr, k * L, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (err = hipMemcpyAsync(&flag, sc.ws.d_flag.ptr, sizeof(flag), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (err = hipStreamSynchronize(s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if (flag) return st(SCRS_DECODE_INSUFFICIENT_RANK, e - 1, e);
    std::memcpy(out, sc.h_out.ptr, k * L);
    return ok();
}

extern "C" scrs_status scrs_decode_into(scrs_batch_decoder *dec, const uint64_t *indices,
                                        const uint8_t *const *payloads, const uint64_t *payload_lens,
                                        uint64_t count, uint8_t *out, uint64_t out_len) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!is_gf256(dec->prof.engine))
        return scrs_gf16_decode_into_impl(dec, indices, payloads, payload_lens, count, out, out_len, nullptr);
    return gf256_decode_host(dec, indices, payloads, payload_lens, count, out, out_len, dec->own);
}

extern "C" scrs_status scrs_decode_into_with(scrs_batch_decoder *dec, const uint64_t *indices,
                                             const uint8_t *const *payloads, const uint64_t *payload_lens,
                                             uint64_t count, uint8_t *out, uint64_t out_len,
                                             scrs_decode_scratch *scratch) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 8);
    if (!is_gf256(dec->prof.engine))
        return scrs_gf16_decode_into_impl(dec, indices, payloads, payload_lens, count, out, out_len, scratch);
    // ScratchMismatch is the first check (batch/codec.rs:289-291; selector.rs:330-357).
    if (!is_gf256(scratch->engine) || scratch->k != dec->prof.k || scratch->m != dec->prof.m ||
        scratch->L != dec->prof.symbol_len)
        return st(SCRS_DECODE_SCRATCH_MISMATCH);
    return gf256_decode_host(dec, indices, payloads, payload_lens, count, out, out_len, *scratch);
}

static scrs_status gf256_decode_batch(scrs_batch_decoder *dec, const uint16_t *indices,
                                      const uint8_t *const *payload_ptrs_dev, const uint8_t *data_dev,
                                      const uint8_t *repairs_dev, uint8_t *out_dev, uint64_t count,
                                      void *stream) {
    const uint64_t k = dec->prof.k, n = k + dec->prof.m, L = dec->prof.symbol_len;
    if (count == 0) return ok();
    if (!indices) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!out_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 5);
    for (uint64_t c = 0; c < count; c++) {
        scrs_status s = check_indices_u16(indices + c * k, k, n);
        if (s.code != SCRS_OK) return s;
    }
    DeviceGuard g(dec->device);
    const bool vec16 = L % 16 == 0 && aligned16(out_dev) &&
                       (payload_ptrs_dev || (aligned16(data_dev) && aligned16(repairs_dev)));
    return enqueue_gf256_decode(dec, dec->ws, indices, payload_ptrs_dev, data_dev, repairs_dev, out_dev, count,
                                static_cast<hipStream_t>(stream), vec16 && !payload_ptrs_dev);
}

extern "C" scrs_status scrs_decode_batch_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                             const uint8_t *data_dev, const uint8_t *repairs_dev,
                                             uint8_t *out_dev, uint64_t count, void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!is_gf256(dec->prof.engine))
        return scrs_gf16_decode_batch_impl(dec, indices, nullptr, data_dev, repairs_dev, out_dev, count, stream);
    if (count && (!data_dev || !repairs_dev)) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    return gf256_decode_batch(dec, indices, nullptr, data_dev, repairs_dev, out_dev, count, stream);
}

extern "C" scrs_status scrs_decode_batch_ptrs_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                                  const uint8_t *const *payload_ptrs_dev, uint8_t *out_dev,
                                                  uint64_t count, void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count && !payload_ptrs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    if (!is_gf256(dec->prof.engine))
        return scrs_gf16_decode_batch_impl(dec, indices, payload_ptrs_dev, nullptr, nullptr, out_dev, count, stream);
    return gf256_decode_batch(dec, indices, payload_ptrs_dev, nullptr, nullptr, out_dev, count, stream);
}
