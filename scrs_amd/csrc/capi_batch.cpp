// capi_batch.cpp - C ABI: profile selection, BatchEncoder, BatchDecoder.
#include <cstring>
#include <vector>

#include "handles.hpp"
#include "matrices.hpp"
#include "status.hpp"

using namespace scrs;

namespace {

scrs_status construction_error(hipError_t err, int32_t engine) {
    // Device OOM while building a codec maps onto TooManySymbols{cap}, the
    // reference's allocation-failure precedent (tower/encoder.rs:37-44).
    if (err == hipErrorOutOfMemory) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(engine));
    return st(SCRS_ERR_DEVICE, static_cast<uint64_t>(err));
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
static bool aligned4(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; }

// Received indices of one codeword: IndexOutOfRange -> DuplicateIndex, in
// order (batch/codec.rs:304-317; the blanket BatchDecoder's push order,
// codec.rs:228-236).
scrs_status check_indices_u16(const uint16_t *idx, uint64_t k, uint64_t n, std::vector<uint64_t> &seen) {
    seen.assign((n + 63) / 64, 0);
    for (uint64_t t = 0; t < k; t++) {
        const uint64_t i = idx[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
        seen[i >> 6] |= 1ull << (i & 63);
    }
    return ok();
}

// Scratch kinds of the type-erased layer (selector.rs:248-258, 446-455).
int32_t scratch_kind(int32_t engine) { return engine_is_gf256(engine) ? 0 : engine; }

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

extern "C" scrs_status scrs_profile_resolve(int32_t engine, uint64_t k, uint64_t m, uint64_t symbol_len,
                                            scrs_profile *out) {
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

extern "C" scrs_status scrs_profile_recommended(int32_t field, uint64_t k, uint64_t m, uint64_t symbol_len,
                                                scrs_profile *out) {
    int32_t engine;
    if (field == SCRS_FIELD_GF256)
        engine = (k + m >= k && k + m <= 255) ? SCRS_ENGINE_GOOD_CAUCHY : SCRS_ENGINE_STANDARD_CAUCHY;
    else if (field == SCRS_FIELD_GF65536)
        engine = scrs_recommended_gf16_engine(k, m);
    else
        return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    return scrs_profile_resolve(engine, k, m, symbol_len, out);
}

extern "C" scrs_status scrs_fill_random_dev(uint8_t *dst_dev, uint64_t len, uint64_t seed, uint64_t byte_offset,
                                            void *stream) {
    if (!dst_dev && len) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (byte_offset % 8) return st(SCRS_ERR_INVALID_ARGUMENT, 4);
    hipError_t err = launch_fill_splitmix(dst_dev, len, seed, byte_offset, static_cast<hipStream_t>(stream));
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

// ---------------------------------------------------------------------------
// Batch encoder
// ---------------------------------------------------------------------------

extern "C" scrs_status scrs_batch_encoder_create(const scrs_profile *profile, scrs_batch_encoder **out) {
    if (!profile) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(profile->engine, profile->k, profile->m, profile->symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    // Tower is incremental-only (selector.rs:457-467).
    if (p.engine == SCRS_ENGINE_TOWER) return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    auto *enc = new scrs_batch_encoder();
    enc->prof = p;
    (void)hipGetDevice(&enc->device);
    enc->own.engine = p.engine;
    enc->own.k = p.k;
    enc->own.m = p.m;
    enc->own.L = p.symbol_len;
    hipError_t err = hipSuccess;
    if (p.engine == SCRS_ENGINE_AFFT) {
        s = scrs_afft_engine_create(p, false, &enc->afft);
        if (s.code != SCRS_OK) {
            scrs_batch_encoder_destroy(enc);
            return s;
        }
    } else if (err == hipSuccess) {
        std::vector<uint32_t> terms(p.k);
        for (uint32_t i = 0; i < p.k; i++) terms[i] = i;
        const std::vector<uint32_t> tables = coefficient_tables(p, terms, &enc->groups);
        err = enc->d_tables.ensure(tables.size() * 4);
        if (err == hipSuccess) err = hipMemcpy(enc->d_tables.ptr, tables.data(), tables.size() * 4, hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        scrs_batch_encoder_destroy(enc);
        return construction_error(err, p.engine);
    }
    *out = enc;
    return ok();
}

extern "C" void scrs_batch_encoder_destroy(scrs_batch_encoder *enc) {
    if (!enc) return;
    DeviceGuard g(enc->device);
    if (enc->stream) {
        (void)hipStreamSynchronize(enc->stream);
        (void)hipStreamDestroy(enc->stream);
    }
    if (enc->afft) scrs_afft_engine_destroy(enc->afft);
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

static scrs_status check_encode_lens(const scrs_profile &p, uint64_t data_len, uint64_t repairs_len) {
    const uint64_t din = p.k * p.symbol_len, dout = p.m * p.symbol_len;
    if (data_len != din) return st(SCRS_ENCODE_WRONG_INPUT_LEN, din, data_len);
    if (repairs_len != dout) return st(SCRS_ENCODE_WRONG_OUTPUT_LEN, dout, repairs_len);
    return ok();
}

static scrs_status encode_host(const scrs_batch_encoder *enc, const uint8_t *data, uint8_t *repairs,
                               scrs_encode_scratch &sc) {
    const uint64_t k = enc->prof.k, m = enc->prof.m, L = enc->prof.symbol_len;
    hipError_t err;
    if ((err = sc.h_data.ensure(k * L)) != hipSuccess || (err = sc.d_data.ensure(k * L)) != hipSuccess ||
        (err = sc.h_rep.ensure(m * L)) != hipSuccess || (err = sc.d_rep.ensure(m * L)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    std::memcpy(sc.h_data.ptr, data, k * L);
    auto *mut = const_cast<scrs_batch_encoder *>(enc);
    std::call_once(mut->stream_once,
                   [mut] { mut->stream_err = hipStreamCreateWithFlags(&mut->stream, hipStreamNonBlocking); });
    if (mut->stream_err != hipSuccess) return st(SCRS_ERR_DEVICE, mut->stream_err);
    hipStream_t s = enc->stream;
    if ((err = hipMemcpyAsync(sc.d_data.ptr, sc.h_data.ptr, k * L, hipMemcpyHostToDevice, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    scrs_status r = scrs_encode_batch_dev(enc, sc.d_data.as<uint8_t>(), sc.d_rep.as<uint8_t>(), 1, s);
    if (r.code != SCRS_OK) return r;
    if ((err = hipMemcpyAsync(sc.h_rep.ptr, sc.d_rep.ptr, m * L, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (err = hipStreamSynchronize(s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    std::memcpy(repairs, sc.h_rep.ptr, m * L);
    return ok();
}

extern "C" scrs_status scrs_encode_into(const scrs_batch_encoder *enc, const uint8_t *data, uint64_t data_len,
                                        uint8_t *repairs, uint64_t repairs_len) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = check_encode_lens(enc->prof, data_len, repairs_len);
    if (s.code != SCRS_OK) return s;
    if ((!data && data_len) || (!repairs && repairs_len)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    auto *mut = const_cast<scrs_batch_encoder *>(enc);
    std::lock_guard<std::mutex> lock(mut->mu);
    DeviceGuard g(enc->device);
    return encode_host(enc, data, repairs, mut->own);
}

extern "C" scrs_status scrs_encode_into_with(const scrs_batch_encoder *enc, const uint8_t *data, uint64_t data_len,
                                             uint8_t *repairs, uint64_t repairs_len, scrs_encode_scratch *scratch) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 6);
    // The type-erased layer matches the scratch variant first
    // (selector.rs:516-527): GF(256) codecs take the unit scratch of any
    // GF(256) geometry; AFFT takes an AFFT scratch.
    if (scratch_kind(enc->prof.engine) != scratch_kind(scratch->engine)) return st(SCRS_ENCODE_SCRATCH_MISMATCH);
    // An AFFT scratch is a workspace that is resized on use, so one from
    // another AFFT geometry is accepted (afft/encoder.rs:250-272).
    scrs_status s = check_encode_lens(enc->prof, data_len, repairs_len);
    if (s.code != SCRS_OK) return s;
    if ((!data && data_len) || (!repairs && repairs_len)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    DeviceGuard g(enc->device);
    return encode_host(enc, data, repairs, *scratch);
}

extern "C" scrs_status scrs_encode_batch_dev(const scrs_batch_encoder *enc, const uint8_t *data_dev,
                                             uint8_t *repairs_dev, uint64_t count, void *stream) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count == 0) return ok();
    if (!data_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!repairs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    if (enc->prof.engine == SCRS_ENGINE_AFFT)
        return scrs_afft_encode(enc->afft, data_dev, repairs_dev, count, static_cast<hipStream_t>(stream));
    const uint64_t k = enc->prof.k, m = enc->prof.m, L = enc->prof.symbol_len;
    const bool vec16 = aligned16(data_dev) && aligned16(repairs_dev) && L % 16 == 0;
    RowsTermsArgs a;
    a.src_base = data_dev;
    a.src_cw_stride = k * L;
    a.dst_base = repairs_dev;
    a.dst_cw_stride = m * L;
    a.tables = enc->d_tables.as<uint32_t>();
    a.table_cw_stride = 0;
    a.T = static_cast<uint32_t>(k);
    a.rmax = static_cast<uint32_t>(m);
    a.groups = enc->groups;
    a.L = L;
    a.count = count;
    a.dst_aligned4 = aligned4(repairs_dev) && L % 4 == 0;
    hipError_t err;
    if (enc->kernel != SCRS_KERNEL_TABLE && aligned4(data_dev) && aligned4(repairs_dev)) {
        a.tables = enc->d_bm_tables.as<uint32_t>();
        err = launch_rows_bitmatrix(a, static_cast<hipStream_t>(stream), static_cast<uint32_t>(enc->kernel));
    } else {
        hipStream_t s = static_cast<hipStream_t>(stream);
        auto *mut = const_cast<scrs_batch_encoder *>(enc);
        const bool attached = mut->split.acquire(a, static_cast<uint32_t>(m), s, true);
        err = launch_rows_terms_lds(a, vec16, s);
        if (attached) mut->split.release(s);
    }
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

// Validation shared by both set_kernel calls.
static scrs_status check_kernel(const scrs_profile &p, int32_t kernel) {
    if (kernel != SCRS_KERNEL_TABLE && kernel != SCRS_KERNEL_BITMATRIX && kernel != SCRS_KERNEL_BITMATRIX_FP4)
        return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!engine_is_gf256(p.engine)) return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    if (kernel != SCRS_KERNEL_TABLE) {
        const uint32_t cap = bitmatrix_max_terms(static_cast<uint32_t>(kernel));
        if (p.k > cap) return st(SCRS_ERR_CAPACITY, cap);
        if (p.symbol_len % 4) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    }
    return ok();
}

extern "C" scrs_status scrs_batch_encoder_set_kernel(scrs_batch_encoder *enc, int32_t kernel) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = check_kernel(enc->prof, kernel);
    if (s.code != SCRS_OK) return s;
    if (kernel != SCRS_KERNEL_TABLE && enc->bm_kind != kernel) {
        const uint32_t k = static_cast<uint32_t>(enc->prof.k), m = static_cast<uint32_t>(enc->prof.m);
        const std::vector<uint8_t> c = enc->prof.engine == SCRS_ENGINE_GOOD_CAUCHY ? good_cauchy_matrix(k, m)
                                                                                   : standard_cauchy_matrix(k, m);
        std::vector<uint8_t> tab(bitmatrix_table_bytes(k, enc->groups, static_cast<uint32_t>(kernel)));
        build_bitmatrix_tables(c.data(), k, m, enc->groups, tab.data(), static_cast<uint32_t>(kernel));
        DeviceGuard g(enc->device);
        hipError_t err = enc->d_bm_tables.ensure(tab.size());
        if (err == hipSuccess) err = hipMemcpy(enc->d_bm_tables.ptr, tab.data(), tab.size(), hipMemcpyHostToDevice);
        if (err != hipSuccess) return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
        enc->bm_kind = kernel;
    }
    enc->kernel = kernel;
    return ok();
}

extern "C" scrs_status scrs_batch_decoder_set_kernel(scrs_batch_decoder *dec, int32_t kernel) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = check_kernel(dec->prof, kernel);
    if (s.code != SCRS_OK) return s;
    dec->kernel = kernel;
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
    auto *dec = new scrs_batch_decoder();
    dec->prof = p;
    (void)hipGetDevice(&dec->device);
    dec->own.engine = p.engine;
    dec->own.k = p.k;
    dec->own.m = p.m;
    dec->own.L = p.symbol_len;
    s = p.engine == SCRS_ENGINE_AFFT ? scrs_afft_engine_create(p, true, &dec->afft) : dec->eng.init(p);
    if (s.code != SCRS_OK) {
        scrs_batch_decoder_destroy(dec);
        return s;
    }
    *out = dec;
    return ok();
}

extern "C" void scrs_batch_decoder_destroy(scrs_batch_decoder *dec) {
    if (!dec) return;
    DeviceGuard g(dec->device);
    if (dec->stream) {
        (void)hipStreamSynchronize(dec->stream);
        (void)hipStreamDestroy(dec->stream);
    }
    if (dec->afft) scrs_afft_engine_destroy(dec->afft);
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

// Host-slice decode of one codeword: validation in the engine's order, then
// the received payloads are packed into pinned memory, uploaded once, and
// decoded on the device.
static scrs_status decode_host(scrs_batch_decoder *dec, const uint64_t *indices, const uint8_t *const *payloads,
                               const uint64_t *payload_lens, uint64_t count, uint8_t *out, uint64_t out_len,
                               scrs_decode_scratch &sc, bool with_scratch) {
    const scrs_profile &p = dec->prof;
    const uint64_t k = p.k, n = p.k + p.m, L = p.symbol_len;
    const bool gf256 = engine_is_gf256(p.engine);
    const bool geometry_ok = sc.k == p.k && sc.m == p.m && sc.L == L;
    if (gf256) {
        // BatchCodec::decode_into_with (batch/codec.rs:289-331).
        if (with_scratch && !geometry_ok) return st(SCRS_DECODE_SCRATCH_MISMATCH);
        if (out_len != k * L) return st(SCRS_DECODE_WRONG_OUTPUT_LEN, k * L, out_len);
        if (count != k) return st(SCRS_DECODE_WRONG_COUNT, k, count);
    } else if (count != k) {
        // Blanket BatchDecoder over Decoder (codec.rs:219-239).
        return st(SCRS_DECODE_WRONG_COUNT, k, count);
    }
    if (!indices || !payloads || !payload_lens) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    std::vector<uint64_t> &seen = sc.seen;
    std::vector<uint16_t> &idx16 = sc.idx16;
    seen.assign((n + 63) / 64, 0);
    idx16.resize(k);
    for (uint64_t t = 0; t < count; t++) {
        const uint64_t i = indices[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (gf256) {
            if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
            seen[i >> 6] |= 1ull << (i & 63);
            if (payload_lens[t] != L) return st(SCRS_DECODE_WRONG_PAYLOAD_LEN, L, payload_lens[t]);
        } else {
            // push: IndexOutOfRange -> WrongPayloadLen -> Dependent
            // (tower/decoder.rs:372-405), Dependent -> DuplicateIndex.
            if (payload_lens[t] != L) return st(SCRS_DECODE_WRONG_PAYLOAD_LEN, L, payload_lens[t]);
            if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
            seen[i >> 6] |= 1ull << (i & 63);
        }
        if (!payloads[t]) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
        idx16[t] = static_cast<uint16_t>(i);
    }
    if (!gf256) {
        // finalize_into_with: InsufficientRank -> WrongOutputLen -> ScratchMismatch.
        if (out_len != k * L) return st(SCRS_DECODE_WRONG_OUTPUT_LEN, k * L, out_len);
        if (with_scratch && !geometry_ok) return st(SCRS_DECODE_SCRATCH_MISMATCH);
    }
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 6);
    DeviceGuard g(dec->device);
    hipError_t err;
    if (!dec->stream && (err = hipStreamCreateWithFlags(&dec->stream, hipStreamNonBlocking)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if ((err = sc.h_payload.ensure(k * L)) != hipSuccess || (err = sc.d_payload.ensure(k * L)) != hipSuccess ||
        (err = sc.h_out.ensure(k * L)) != hipSuccess || (err = sc.d_out.ensure(k * L)) != hipSuccess ||
        (err = sc.d_ptrs.ensure(k * sizeof(void *))) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    hipStream_t s = dec->stream;
    for (uint64_t t = 0; t < k; t++) std::memcpy(sc.h_payload.as<uint8_t>() + t * L, payloads[t], L);
    std::vector<const uint8_t *> &ptrs = sc.ptrs;
    ptrs.resize(k);
    for (uint64_t t = 0; t < k; t++) ptrs[t] = sc.d_payload.as<uint8_t>() + t * L;
    if ((err = hipMemcpyAsync(sc.d_payload.ptr, sc.h_payload.ptr, k * L, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (err = hipMemcpyAsync(sc.d_ptrs.ptr, ptrs.data(), k * sizeof(void *), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (!dec->afft && (err = dec->eng.clear_error(s)) != hipSuccess))
        return st(SCRS_ERR_DEVICE, err);
    scrs_status r = dec->afft ? scrs_afft_decode(dec->afft, idx16.data(), sc.d_ptrs.as<const uint8_t *const>(),
                                                 nullptr, nullptr, sc.d_out.as<uint8_t>(), 1, s)
                              : dec->eng.enqueue(idx16.data(), sc.d_ptrs.as<const uint8_t *const>(), nullptr,
                                                 nullptr, sc.d_out.as<uint8_t>(), 1, s, L % 16 == 0,
                                                 static_cast<uint32_t>(dec->kernel));
    if (r.code != SCRS_OK) return r;
    uint32_t flag = 0;
    if ((err = hipMemcpyAsync(sc.h_out.ptr, sc.d_out.ptr, k * L, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (err = dec->afft ? hipStreamSynchronize(s) : dec->eng.read_error(&flag, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if (flag) return st(SCRS_DECODE_INSUFFICIENT_RANK, k - 1, k);
    std::memcpy(out, sc.h_out.ptr, k * L);
    return ok();
}

extern "C" scrs_status scrs_decode_into(scrs_batch_decoder *dec, const uint64_t *indices,
                                        const uint8_t *const *payloads, const uint64_t *payload_lens, uint64_t count,
                                        uint8_t *out, uint64_t out_len) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    return decode_host(dec, indices, payloads, payload_lens, count, out, out_len, dec->own, false);
}

extern "C" scrs_status scrs_decode_into_with(scrs_batch_decoder *dec, const uint64_t *indices,
                                             const uint8_t *const *payloads, const uint64_t *payload_lens,
                                             uint64_t count, uint8_t *out, uint64_t out_len,
                                             scrs_decode_scratch *scratch) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 8);
    // Variant check of the type-erased layer comes first (selector.rs:330-357).
    if (scratch_kind(dec->prof.engine) != scratch_kind(scratch->engine)) return st(SCRS_DECODE_SCRATCH_MISMATCH);
    return decode_host(dec, indices, payloads, payload_lens, count, out, out_len, *scratch, true);
}

static scrs_status decode_batch(scrs_batch_decoder *dec, const uint16_t *indices,
                                const uint8_t *const *payload_ptrs_dev, const uint8_t *data_dev,
                                const uint8_t *repairs_dev, uint8_t *out_dev, uint64_t count, void *stream) {
    const uint64_t k = dec->prof.k, n = k + dec->prof.m, L = dec->prof.symbol_len;
    if (count == 0) return ok();
    if (!indices) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!out_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 5);
    for (uint64_t c = 0; c < count; c++) {
        scrs_status s = check_indices_u16(indices + c * k, k, n, dec->seen);
        if (s.code != SCRS_OK) return s;
    }
    if (dec->prof.engine == SCRS_ENGINE_AFFT) {
        DeviceGuard g(dec->device);
        // With one codeword, the validation's bitmap is its received set.
        return scrs_afft_decode(dec->afft, indices, payload_ptrs_dev, data_dev, repairs_dev, out_dev, count,
                                static_cast<hipStream_t>(stream), count == 1 ? dec->seen.data() : nullptr);
    }
    DeviceGuard g(dec->device);
    // Pointer-table payloads may be unaligned: the 16-byte path is taken
    // only for the layout form, whose alignment is known here.
    const bool vec16 = !payload_ptrs_dev && L % 16 == 0 && aligned16(out_dev) && aligned16(data_dev) &&
                       aligned16(repairs_dev);
    const uint32_t bitmatrix = (!payload_ptrs_dev && aligned4(out_dev) && aligned4(data_dev) && aligned4(repairs_dev))
                                   ? static_cast<uint32_t>(dec->kernel)
                                   : 0u;
    return dec->eng.enqueue(indices, payload_ptrs_dev, data_dev, repairs_dev, out_dev, count,
                            static_cast<hipStream_t>(stream), vec16, bitmatrix);
}

extern "C" scrs_status scrs_decode_batch_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                             const uint8_t *data_dev, const uint8_t *repairs_dev, uint8_t *out_dev,
                                             uint64_t count, void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count && (!data_dev || !repairs_dev)) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    return decode_batch(dec, indices, nullptr, data_dev, repairs_dev, out_dev, count, stream);
}

extern "C" scrs_status scrs_decode_batch_ptrs_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                                  const uint8_t *const *payload_ptrs_dev, uint8_t *out_dev,
                                                  uint64_t count, void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count && !payload_ptrs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    return decode_batch(dec, indices, payload_ptrs_dev, nullptr, nullptr, out_dev, count, stream);
}
