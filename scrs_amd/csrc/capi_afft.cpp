// capi_afft.cpp - the additive-FFT engine behind the shared batch-encoder,
// batch-decoder and decoder handles.
#include "afft.hpp"
#include "handles.hpp"
#include "status.hpp"

using namespace scrs;

namespace {

AfftEngine *engine_of(void *p) { return static_cast<AfftEngine *>(p); }

}  // namespace

scrs_status scrs_afft_engine_create(const scrs_profile &p, bool for_decode, void **out) {
    auto *e = new AfftEngine();
    scrs_status s = e->init(p, for_decode);
    if (s.code != SCRS_OK) {
        delete e;
        return s;
    }
    *out = e;
    return ok();
}

void scrs_afft_engine_destroy(void *engine) { delete engine_of(engine); }

scrs_status scrs_afft_encode(void *engine, const uint8_t *data_dev, uint8_t *repairs_dev, uint64_t count,
                             hipStream_t s) {
    AfftEngine *e = engine_of(engine);
    std::lock_guard<std::mutex> lock(e->mu);
    hipError_t err = e->begin(s);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    scrs_status r = e->encode(data_dev, repairs_dev, count, s);
    err = e->finish(s);
    if (r.code == SCRS_OK && err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    return r;
}

scrs_status scrs_afft_decode(void *engine, const uint16_t *indices, const uint8_t *const *ptrs_dev,
                             const uint8_t *data_base, const uint8_t *repair_base, uint8_t *out_dev, uint64_t count,
                             hipStream_t s, const uint64_t *received64) {
    AfftEngine *e = engine_of(engine);
    std::lock_guard<std::mutex> lock(e->mu);
    hipError_t err = e->begin(s);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    const uint64_t k = e->k, m = e->m, L = e->L;
    scrs_status r = ok();
    for (uint64_t c = 0; c < count && r.code == SCRS_OK; c++)
        r = e->decode(indices + c * k, ptrs_dev ? ptrs_dev + c * k : nullptr,
                      data_base ? data_base + c * k * L : nullptr, repair_base ? repair_base + c * m * L : nullptr,
                      out_dev + c * k * L, s, count == 1 ? received64 : nullptr);
    err = e->finish(s);
    if (r.code == SCRS_OK && err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    return r;
}
