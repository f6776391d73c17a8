// status.hpp - status construction and the engine hooks shared between the
// C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/scrs.h"

inline scrs_status st(int32_t code, uint64_t a = 0, uint64_t b = 0) {
    scrs_status s;
    s.code = code;
    s.a = a;
    s.b = b;
    return s;
}

inline scrs_status ok() { return st(SCRS_OK); }

// Additive-FFT engine behind the shared batch-encoder / batch-decoder /
// decoder handles (capi_afft.cpp).  Engines are opaque; calls on one engine
// are serialised.
scrs_status scrs_afft_engine_create(const scrs_profile &p, bool for_decode, void **out);
void scrs_afft_engine_destroy(void *engine);
scrs_status scrs_afft_encode(void *engine, const uint8_t *data_dev, uint8_t *repairs_dev, uint64_t count,
                             hipStream_t s);
// `count` codewords; codeword c's k received indices at indices + c*k (host,
// validated), payloads from ptrs_dev + c*k or the data / repair layout.
// received64 (optional, count == 1): the validation's bitmap of the received
// indices, bit i of word i / 64, over the n indices.
scrs_status scrs_afft_decode(void *engine, const uint16_t *indices, const uint8_t *const *ptrs_dev,
                             const uint8_t *data_base, const uint8_t *repair_base, uint8_t *out_dev, uint64_t count,
                             hipStream_t s, const uint64_t *received64 = nullptr);
