// capi_stream.cpp - C ABI: streaming Decoder and IncrementalEncoder.
#include <algorithm>
#include <cstring>
#include <vector>

#include "handles.hpp"
#include "matrices.hpp"
#include "status.hpp"

using namespace scrs;

namespace {

constexpr uint32_t kPending = 32;  // fed symbols staged before one flush

bool bit(const std::vector<uint64_t> &b, uint64_t i) { return (b[i >> 6] >> (i & 63)) & 1u; }

scrs_push_outcome outcome(int32_t kind, uint64_t rank = 0, uint64_t received = 0) {
    scrs_push_outcome o;
    o.kind = kind;
    o.rank = rank;
    o.received = received;
    return o;
}

}  // namespace

// ---------------------------------------------------------------------------
// Decoder (codec.rs:180-201, transport/symbol_sink.rs:36-54)
// ---------------------------------------------------------------------------

extern "C" scrs_status scrs_decoder_create(const scrs_profile *profile, scrs_decoder **out) {
    if (!profile) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(profile->engine, profile->k, profile->m, profile->symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    auto *dec = new scrs_decoder();
    dec->prof = p;
    (void)hipGetDevice(&dec->device);
    const uint64_t n = p.k + p.m, L = p.symbol_len;
    dec->bits.assign((n + 63) / 64, 0);
    dec->idx.reserve(p.k);
    dec->key.reserve(dec->bits.size() + 3);
    hipError_t err = hipStreamCreateWithFlags(&dec->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&dec->order_ev, hipEventDisableTiming);
    if (err == hipSuccess) err = dec->d_slots.ensure(n * L);
    if (err == hipSuccess) err = dec->h_slots.ensure(n * L);
    if (err != hipSuccess) {
        scrs_decoder_destroy(dec);
        // Allocation failure at construction is TooManySymbols{cap}
        // (tower/decoder.rs:98-108).
        if (err == hipErrorOutOfMemory) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(p.engine));
        return st(SCRS_ERR_DEVICE, err);
    }
    s = p.engine == SCRS_ENGINE_AFFT ? scrs_afft_engine_create(p, true, &dec->afft) : dec->eng.init(p);
    if (s.code != SCRS_OK) {
        scrs_decoder_destroy(dec);
        return s;
    }
    *out = dec;
    return ok();
}

extern "C" void scrs_decoder_destroy(scrs_decoder *dec) {
    if (!dec) return;
    DeviceGuard g(dec->device);
    if (dec->stream) {
        (void)hipStreamSynchronize(dec->stream);
        (void)hipStreamDestroy(dec->stream);
    }
    if (dec->order_ev) (void)hipEventDestroy(dec->order_ev);
    if (dec->afft) scrs_afft_engine_destroy(dec->afft);
    delete dec;
}

extern "C" scrs_status scrs_decoder_profile(const scrs_decoder *dec, scrs_profile *out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = dec->prof;
    return ok();
}

// SymbolSink::push check order: IndexOutOfRange -> WrongPayloadLen ->
// TooManySymbols -> received += 1 -> Dependent (complete or duplicate) ->
// store -> Complete / Advanced (decoder/streaming.rs:388-429,
// tower/decoder.rs:372-405).
static scrs_status push_common(scrs_decoder *dec, uint64_t index, uint64_t len, scrs_push_outcome *out,
                               bool *store) {
    const uint64_t n = dec->prof.k + dec->prof.m, k = dec->prof.k;
    *store = false;
    if (index >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, index, n);
    if (len != dec->prof.symbol_len) return st(SCRS_DECODE_WRONG_PAYLOAD_LEN, dec->prof.symbol_len, len);
    if (dec->received >= n) return st(SCRS_DECODE_TOO_MANY_SYMBOLS, n, dec->received);
    dec->received += 1;
    if (dec->distinct >= k || bit(dec->bits, index)) {
        if (out) *out = outcome(SCRS_PUSH_DEPENDENT);
        return ok();
    }
    dec->bits[index >> 6] |= 1ull << (index & 63);
    dec->distinct += 1;
    *store = true;
    if (out)
        *out = dec->distinct >= k ? outcome(SCRS_PUSH_COMPLETE)
                                  : outcome(SCRS_PUSH_ADVANCED, dec->distinct, dec->received);
    return ok();
}

extern "C" scrs_status scrs_decoder_push(scrs_decoder *dec, uint64_t index, const uint8_t *payload, uint64_t len,
                                         scrs_push_outcome *out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!payload && len) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    bool store = false;
    scrs_status s = push_common(dec, index, len, out, &store);
    if (s.code != SCRS_OK || !store) return s;
    DeviceGuard g(dec->device);
    const uint64_t L = dec->prof.symbol_len;
    // Each slot is written once per block (reset synchronises the stream
    // before slots are reused), so the pinned copy stays valid until the
    // asynchronous upload has read it.
    std::memcpy(dec->h_slots.as<uint8_t>() + index * L, payload, L);
    hipError_t err = hipMemcpyAsync(dec->d_slots.as<uint8_t>() + index * L, dec->h_slots.as<uint8_t>() + index * L,
                                    L, hipMemcpyHostToDevice, dec->stream);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

extern "C" scrs_status scrs_decoder_push_dev(scrs_decoder *dec, uint64_t index, const uint8_t *payload_dev,
                                             uint64_t len, void *stream, scrs_push_outcome *out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!payload_dev && len) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    bool store = false;
    scrs_status s = push_common(dec, index, len, out, &store);
    if (s.code != SCRS_OK || !store) return s;
    DeviceGuard g(dec->device);
    const uint64_t L = dec->prof.symbol_len;
    hipStream_t cs = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(dec->d_slots.as<uint8_t>() + index * L, payload_dev, L, hipMemcpyDeviceToDevice, cs);
    // Finalize orders itself after the decoder's stream: chain the copy there.
    if (err == hipSuccess) err = hipEventRecord(dec->order_ev, cs);
    if (err == hipSuccess) err = hipStreamWaitEvent(dec->stream, dec->order_ev, 0);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

extern "C" uint64_t scrs_decoder_rank(const scrs_decoder *dec) { return dec ? dec->distinct : 0; }
extern "C" uint64_t scrs_decoder_received(const scrs_decoder *dec) { return dec ? dec->received : 0; }
extern "C" int32_t scrs_decoder_is_complete(const scrs_decoder *dec) {
    return dec && dec->distinct >= dec->prof.k ? 1 : 0;
}

extern "C" scrs_status scrs_decoder_reset(scrs_decoder *dec) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    DeviceGuard g(dec->device);
    // In-flight uploads still read the pinned slots: drain them first.
    hipError_t err = hipStreamSynchronize(dec->stream);
    std::fill(dec->bits.begin(), dec->bits.end(), 0);
    dec->distinct = 0;
    dec->received = 0;
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

extern "C" scrs_status scrs_decoder_scratch_create(const scrs_decoder *dec, scrs_decoder_scratch **out) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    auto *s = new scrs_decoder_scratch();
    s->engine = dec->prof.engine;
    s->k = dec->prof.k;
    s->m = dec->prof.m;
    s->L = dec->prof.symbol_len;
    s->capacity = 16;  // Decoder::scratch() = RecipeCache::new(16) (decoder/streaming.rs:461-463)
    *out = s;
    return ok();
}

extern "C" void scrs_decoder_scratch_destroy(scrs_decoder_scratch *scratch) { delete scratch; }

extern "C" scrs_status scrs_decoder_scratch_stats(const scrs_decoder_scratch *scratch, uint64_t *hits,
                                                  uint64_t *misses, uint64_t *len, uint64_t *recipe_bytes) {
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    uint64_t bytes = 0;
    for (const RecipeEntry &e : scratch->lru) bytes += e.bytes;
    if (hits) *hits = scratch->hits;
    if (misses) *misses = scratch->misses;
    if (len) *len = scratch->lru.size();
    if (recipe_bytes) *recipe_bytes = bytes;
    return ok();
}

// Reconstruct into device memory on `s`.  With a recipe cache (GF(2^8)
// engines), a hit replays the cached product tables; a miss solves and then
// stores the new recipe (decoder/streaming.rs:134-161, decoder/cache.rs).
static scrs_status finalize_dev(scrs_decoder *dec, uint8_t *out_dev, hipStream_t s, scrs_decoder_scratch *cache) {
    const uint64_t k = dec->prof.k, n = k + dec->prof.m, L = dec->prof.symbol_len;
    std::vector<uint16_t> &idx = dec->idx;  // capacity k from creation
    idx.clear();
    for (uint64_t i = 0; i < n && idx.size() < k; i++)
        if (bit(dec->bits, i)) idx.push_back(static_cast<uint16_t>(i));
    // Slots are laid out as one codeword: data rows then repair rows.
    const uint8_t *data_base = dec->d_slots.as<uint8_t>();
    const uint8_t *repair_base = data_base + k * L;
    // Host pushes were queued on the decoder's stream: order them first.
    hipError_t err = hipEventRecord(dec->order_ev, dec->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(s, dec->order_ev, 0);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    if (dec->afft) return scrs_afft_decode(dec->afft, idx.data(), nullptr, data_base, repair_base, out_dev, 1, s);
    const bool vec16 = L % 16 == 0 && (reinterpret_cast<uintptr_t>(out_dev) & 15u) == 0;
    const bool cacheable = cache && engine_is_gf256(dec->prof.engine);
    if (cacheable) {
        // Key: engine, k, m and the receipt pattern (decoder/recipe.rs:9-23).
        std::vector<uint64_t> &key = dec->key;  // capacity reserved at creation
        key.assign(dec->bits.begin(), dec->bits.end());
        key.push_back(static_cast<uint64_t>(dec->prof.engine));
        key.push_back(dec->prof.k);
        key.push_back(dec->prof.m);
        for (auto it = cache->lru.begin(); it != cache->lru.end(); ++it) {
            if (it->pattern != key) continue;
            cache->hits++;
            cache->lru.splice(cache->lru.begin(), cache->lru, it);
            RecipeEntry &e = cache->lru.front();
            return dec->eng.enqueue_recipe(idx.data(), e.d_tables.as<uint32_t>(), e.d_missing.as<uint16_t>(),
                                           e.d_rows.as<uint32_t>(), data_base, repair_base, out_dev, s, vec16);
        }
        cache->misses++;
        dec->eng.write_tables = true;  // the recipe keeps the product tables
        scrs_status r = dec->eng.enqueue(idx.data(), nullptr, data_base, repair_base, out_dev, 1, s, vec16);
        dec->eng.write_tables = false;
        if (r.code != SCRS_OK) return r;
        cache->lru.emplace_front();
        RecipeEntry &e = cache->lru.front();
        e.pattern = key;
        const uint64_t tb = dec->eng.recipe_table_bytes(), mb = (dec->eng.rmax ? dec->eng.rmax : 1) * 2ull;
        if ((err = e.d_tables.ensure(tb)) != hipSuccess || (err = e.d_missing.ensure(mb)) != hipSuccess ||
            (err = e.d_rows.ensure(4)) != hipSuccess ||
            (err = hipMemcpyAsync(e.d_tables.ptr, dec->eng.d_tables.ptr, tb, hipMemcpyDeviceToDevice, s)) !=
                hipSuccess ||
            (err = hipMemcpyAsync(e.d_missing.ptr, dec->eng.d_missing.ptr, mb, hipMemcpyDeviceToDevice, s)) !=
                hipSuccess ||
            (err = hipMemcpyAsync(e.d_rows.ptr, dec->eng.d_rows.ptr, 4, hipMemcpyDeviceToDevice, s)) != hipSuccess) {
            (void)hipStreamSynchronize(s);
            cache->lru.pop_front();
            return st(SCRS_ERR_DEVICE, err);
        }
        e.bytes = tb + mb + 4;
        while (cache->lru.size() > cache->capacity) {
            // The evicted recipe may still be read by queued work.
            (void)hipStreamSynchronize(s);
            cache->lru.pop_back();
        }
        return ok();
    }
    return dec->eng.enqueue(idx.data(), nullptr, data_base, repair_base, out_dev, 1, s, vec16);
}

// Decoder::finalize_into(_with) checks: InsufficientRank -> WrongOutputLen
// (decoder/streaming.rs:472-507) -> ScratchMismatch for the geometry-tagged
// Tower scratch (tower/decoder.rs:205-221).
static scrs_status finalize_checks(const scrs_decoder *dec, uint64_t out_len, const scrs_decoder_scratch *scratch) {
    const uint64_t k = dec->prof.k, L = dec->prof.symbol_len;
    if (dec->distinct < k) return st(SCRS_DECODE_INSUFFICIENT_RANK, dec->distinct, k);
    if (out_len != k * L) return st(SCRS_DECODE_WRONG_OUTPUT_LEN, k * L, out_len);
    if (scratch && !engine_is_gf256(dec->prof.engine) &&
        (scratch->k != dec->prof.k || scratch->m != dec->prof.m || scratch->L != L))
        return st(SCRS_DECODE_SCRATCH_MISMATCH);
    return ok();
}

static scrs_status finalize_host(scrs_decoder *dec, uint8_t *out, uint64_t out_len, scrs_decoder_scratch *scratch) {
    scrs_status s = finalize_checks(dec, out_len, scratch);
    if (s.code != SCRS_OK) return s;
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    DeviceGuard g(dec->device);
    const uint64_t bytes = dec->prof.k * dec->prof.symbol_len;
    hipError_t err;
    if ((err = dec->d_out.ensure(bytes)) != hipSuccess || (err = dec->h_out.ensure(bytes)) != hipSuccess ||
        (!dec->afft && (err = dec->eng.clear_error(dec->stream)) != hipSuccess))
        return st(SCRS_ERR_DEVICE, err);
    s = finalize_dev(dec, dec->d_out.as<uint8_t>(), dec->stream, scratch);
    if (s.code != SCRS_OK) return s;
    uint32_t flag = 0;
    if ((err = hipMemcpyAsync(dec->h_out.ptr, dec->d_out.ptr, bytes, hipMemcpyDeviceToHost, dec->stream)) !=
            hipSuccess ||
        (err = dec->afft ? hipStreamSynchronize(dec->stream) : dec->eng.read_error(&flag, dec->stream)) !=
            hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if (flag) return st(SCRS_DECODE_INSUFFICIENT_RANK, dec->distinct - 1, dec->prof.k);
    std::memcpy(out, dec->h_out.ptr, bytes);
    return ok();
}

extern "C" scrs_status scrs_decoder_finalize_into(scrs_decoder *dec, uint8_t *out, uint64_t out_len) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    return finalize_host(dec, out, out_len, nullptr);
}

extern "C" scrs_status scrs_decoder_finalize_into_with(scrs_decoder *dec, uint8_t *out, uint64_t out_len,
                                                       scrs_decoder_scratch *scratch) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!scratch) return st(SCRS_ERR_INVALID_ARGUMENT, 4);
    // Scratch variant first (selector.rs:224-240): Cauchy / Tower / AFFT.
    const bool want_gf256 = engine_is_gf256(dec->prof.engine);
    if (want_gf256 != engine_is_gf256(scratch->engine) ||
        (!want_gf256 && scratch->engine != dec->prof.engine))
        return st(SCRS_DECODE_SCRATCH_MISMATCH);
    return finalize_host(dec, out, out_len, scratch);
}

extern "C" scrs_status scrs_decoder_finalize_into_dev(scrs_decoder *dec, uint8_t *out_dev, uint64_t out_len,
                                                      void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = finalize_checks(dec, out_len, nullptr);
    if (s.code != SCRS_OK) return s;
    if (!out_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    DeviceGuard g(dec->device);
    return finalize_dev(dec, out_dev, static_cast<hipStream_t>(stream), nullptr);
}

// ---------------------------------------------------------------------------
// IncrementalEncoder (codec.rs:122-131)
// ---------------------------------------------------------------------------

extern "C" scrs_status scrs_incr_encoder_create(const scrs_profile *profile, scrs_incr_encoder **out) {
    if (!profile) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(profile->engine, profile->k, profile->m, profile->symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    // Block-final engines have no incremental encoder (selector.rs:370-380).
    if (p.engine != SCRS_ENGINE_GOOD_CAUCHY && p.engine != SCRS_ENGINE_TOWER)
        return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    auto *enc = new scrs_incr_encoder();
    enc->prof = p;
    enc->gf16 = p.engine == SCRS_ENGINE_TOWER;
    (void)hipGetDevice(&enc->device);
    enc->fed.assign(p.k, 0);
    std::vector<uint32_t> terms(p.k);
    for (uint32_t i = 0; i < p.k; i++) terms[i] = i;
    const std::vector<uint32_t> tables = coefficient_tables(p, terms, &enc->groups);
    enc->group_bytes = table_group_bytes(enc->gf16);
    hipError_t err = enc->d_tables.ensure(tables.size() * 4);
    if (err == hipSuccess) err = hipMemcpy(enc->d_tables.ptr, tables.data(), tables.size() * 4, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        scrs_incr_encoder_destroy(enc);
        if (err == hipErrorOutOfMemory) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(p.engine));
        return st(SCRS_ERR_DEVICE, err);
    }
    *out = enc;
    return ok();
}

extern "C" void scrs_incr_encoder_destroy(scrs_incr_encoder *enc) {
    if (!enc) return;
    DeviceGuard g(enc->device);
    if (enc->stream) {
        (void)hipStreamSynchronize(enc->stream);
        (void)hipStreamDestroy(enc->stream);
    }
    if (enc->order_ev) (void)hipEventDestroy(enc->order_ev);
    delete enc;
}

extern "C" scrs_status scrs_incr_encoder_profile(const scrs_incr_encoder *enc, scrs_profile *out) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = enc->prof;
    return ok();
}

// The feed path's stream, accumulators and staging are made on first use, so
// an encoder used only through scrs_incr_encode_batch_dev holds its tables
// alone.
static scrs_status ensure_feed_state(scrs_incr_encoder *enc) {
    if (enc->feed_ready) return ok();
    const uint64_t m = enc->prof.m, L = enc->prof.symbol_len;
    hipError_t err = hipStreamCreateWithFlags(&enc->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = enc->d_repairs.ensure(m * L);
    if (err == hipSuccess) err = hipMemsetAsync(enc->d_repairs.ptr, 0, m * L, enc->stream);
    if (err == hipSuccess) err = enc->d_pending.ensure(kPending * L);
    if (err == hipSuccess) err = enc->h_pending.ensure(kPending * L);
    if (err == hipSuccess) err = enc->d_ptab.ensure(kPending * 4);
    if (err == hipSuccess) err = enc->h_repairs.ensure(m * L);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&enc->order_ev, hipEventDisableTiming);
    if (err == hipSuccess) err = hipStreamSynchronize(enc->stream);
    if (err != hipSuccess) return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
    enc->pending.reserve(kPending);
    enc->feed_ready = true;
    return ok();
}

// Fold every staged symbol into the repair accumulators in one multiply-XOR
// launch.  Feeding is lazy: the repairs always equal the partial sums of the
// symbols fed so far once flushed (encoder/streaming.rs:192-197).
static scrs_status flush(scrs_incr_encoder *enc) {
    if (enc->pending.empty()) return ok();
    const uint64_t L = enc->prof.symbol_len;
    hipError_t err = hipMemcpyAsync(enc->d_ptab.ptr, enc->pending.data(), enc->pending.size() * 4,
                                    hipMemcpyHostToDevice, enc->stream);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    RowsTermsArgs a;
    a.src_base = enc->d_pending.as<uint8_t>();
    a.src_cw_stride = 0;
    a.dst_base = enc->d_repairs.as<uint8_t>();
    a.dst_cw_stride = 0;
    a.tables = enc->d_tables.as<uint32_t>();
    a.term_index = enc->d_ptab.as<uint32_t>();
    a.T = static_cast<uint32_t>(enc->pending.size());
    a.rmax = static_cast<uint32_t>(enc->prof.m);
    a.groups = enc->groups;
    a.accumulate = 1;
    a.L = L;
    a.count = 1;
    err = launch_rows_terms_field(enc->gf16, a, L % 16 == 0, enc->stream);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    // The staging buffers are reused by the next feeds: wait for this flush.
    err = hipStreamSynchronize(enc->stream);
    enc->pending.clear();
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

// feed check order: IndexOutOfRange{index, n} -> WrongPayloadLen ->
// DuplicateData (encoder/streaming.rs:162-181, tower/encoder.rs:89-116).
static scrs_status feed_checks(scrs_incr_encoder *enc, uint64_t index, uint64_t len) {
    const uint64_t k = enc->prof.k, n = k + enc->prof.m;
    if (index >= k) return st(SCRS_ENCODE_INDEX_OUT_OF_RANGE, index, n);
    if (len != enc->prof.symbol_len) return st(SCRS_ENCODE_WRONG_PAYLOAD_LEN, enc->prof.symbol_len, len);
    if (enc->fed[index]) return st(SCRS_ENCODE_DUPLICATE_DATA, index);
    return ok();
}

extern "C" scrs_status scrs_incr_feed(scrs_incr_encoder *enc, uint64_t index, const uint8_t *data, uint64_t len) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = feed_checks(enc, index, len);
    if (s.code != SCRS_OK) return s;
    if (!data) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    DeviceGuard g(enc->device);
    if ((s = ensure_feed_state(enc)).code != SCRS_OK) return s;
    const uint64_t L = enc->prof.symbol_len, slot = enc->pending.size();
    std::memcpy(enc->h_pending.as<uint8_t>() + slot * L, data, L);
    hipError_t err = hipMemcpyAsync(enc->d_pending.as<uint8_t>() + slot * L, enc->h_pending.as<uint8_t>() + slot * L,
                                    L, hipMemcpyHostToDevice, enc->stream);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    enc->fed[index] = 1;
    enc->fed_count++;
    enc->mirror_valid = false;
    enc->pending.push_back(static_cast<uint32_t>(index));
    return enc->pending.size() == kPending ? flush(enc) : ok();
}

extern "C" scrs_status scrs_incr_feed_dev(scrs_incr_encoder *enc, uint64_t index, const uint8_t *data_dev,
                                          uint64_t len, void *stream) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    scrs_status s = feed_checks(enc, index, len);
    if (s.code != SCRS_OK) return s;
    if (!data_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    DeviceGuard g(enc->device);
    if ((s = ensure_feed_state(enc)).code != SCRS_OK) return s;
    const uint64_t L = enc->prof.symbol_len, slot = enc->pending.size();
    hipStream_t cs = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(enc->d_pending.as<uint8_t>() + slot * L, data_dev, L, hipMemcpyDeviceToDevice, cs);
    if (err == hipSuccess) err = hipEventRecord(enc->order_ev, cs);
    if (err == hipSuccess) err = hipStreamWaitEvent(enc->stream, enc->order_ev, 0);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    enc->fed[index] = 1;
    enc->fed_count++;
    enc->mirror_valid = false;
    enc->pending.push_back(static_cast<uint32_t>(index));
    return enc->pending.size() == kPending ? flush(enc) : ok();
}

extern "C" scrs_status scrs_incr_repair(scrs_incr_encoder *enc, uint64_t index, const uint8_t **ptr, uint64_t *len) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    const uint64_t m = enc->prof.m, n = enc->prof.k + m, L = enc->prof.symbol_len;
    if (index >= m) return st(SCRS_ENCODE_INDEX_OUT_OF_RANGE, index, n);
    if (!ptr) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    DeviceGuard g(enc->device);
    scrs_status fs = ensure_feed_state(enc);
    if (fs.code != SCRS_OK) return fs;
    if (!enc->mirror_valid) {
        scrs_status s = flush(enc);
        if (s.code != SCRS_OK) return s;
        hipError_t err = hipMemcpyAsync(enc->h_repairs.ptr, enc->d_repairs.ptr, m * L, hipMemcpyDeviceToHost,
                                        enc->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(enc->stream);
        if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        enc->mirror_valid = true;
    }
    // Borrow valid until the next feed or reset (codec.rs:126-127).
    *ptr = enc->h_repairs.as<uint8_t>() + index * L;
    if (len) *len = L;
    return ok();
}

extern "C" scrs_status scrs_incr_repairs_dev(scrs_incr_encoder *enc, uint8_t **repairs_dev, void *stream) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!repairs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    DeviceGuard g(enc->device);
    scrs_status s = ensure_feed_state(enc);
    if (s.code != SCRS_OK) return s;
    s = flush(enc);
    if (s.code != SCRS_OK) return s;
    hipError_t err = hipEventRecord(enc->order_ev, enc->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(static_cast<hipStream_t>(stream), enc->order_ev, 0);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    *repairs_dev = enc->d_repairs.as<uint8_t>();
    return ok();
}

extern "C" uint64_t scrs_incr_fed_count(const scrs_incr_encoder *enc) { return enc ? enc->fed_count : 0; }

extern "C" scrs_status scrs_incr_reset(scrs_incr_encoder *enc) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    DeviceGuard g(enc->device);
    scrs_status fs = ensure_feed_state(enc);
    if (fs.code != SCRS_OK) return fs;
    hipError_t err = hipStreamSynchronize(enc->stream);
    if (err == hipSuccess)
        err = hipMemsetAsync(enc->d_repairs.ptr, 0, enc->prof.m * enc->prof.symbol_len, enc->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(enc->stream);
    std::fill(enc->fed.begin(), enc->fed.end(), 0);
    enc->fed_count = 0;
    enc->pending.clear();
    enc->mirror_valid = false;
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

extern "C" scrs_status scrs_incr_encode_batch_dev(const scrs_incr_encoder *enc, const uint8_t *data_dev,
                                                  uint8_t *repairs_dev, uint64_t count, void *stream) {
    if (!enc) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count == 0) return ok();
    if (!data_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!repairs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    const uint64_t k = enc->prof.k, m = enc->prof.m, L = enc->prof.symbol_len;
    const bool vec16 = L % 16 == 0 && (reinterpret_cast<uintptr_t>(data_dev) & 15u) == 0 &&
                       (reinterpret_cast<uintptr_t>(repairs_dev) & 15u) == 0;
    RowsTermsArgs a;
    a.src_base = data_dev;
    a.src_cw_stride = k * L;
    a.dst_base = repairs_dev;
    a.dst_cw_stride = m * L;
    a.tables = enc->d_tables.as<uint32_t>();
    a.T = static_cast<uint32_t>(k);
    a.rmax = static_cast<uint32_t>(m);
    a.groups = enc->groups;
    a.L = L;
    a.count = count;
    a.dst_aligned4 = (reinterpret_cast<uintptr_t>(repairs_dev) & 3u) == 0 && L % 4 == 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *mut = const_cast<scrs_incr_encoder *>(enc);
    const bool attached = mut->split.acquire(a, static_cast<uint32_t>(m), s, true);
    hipError_t err = launch_rows_terms_field(enc->gf16, a, vec16, s);
    if (attached) mut->split.release(s);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}
