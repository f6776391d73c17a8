// status.hpp - status construction and the engine hooks shared between the
// C-ABI translation units.
#pragma once

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
// decoder handles (capi_afft.cpp).
scrs_status scrs_afft_encoder_init_impl(scrs_batch_encoder *enc);
void scrs_afft_encoder_release_impl(scrs_batch_encoder *enc);
scrs_status scrs_afft_encode_batch_dev_impl(const scrs_batch_encoder *enc, const uint8_t *data_dev,
                                            uint8_t *repairs_dev, uint64_t count, void *stream);
scrs_status scrs_afft_batch_decoder_init_impl(scrs_batch_decoder *dec);
void scrs_afft_batch_decoder_release_impl(scrs_batch_decoder *dec);
scrs_status scrs_afft_decode_into_impl(scrs_batch_decoder *dec, const uint64_t *indices,
                                       const uint8_t *const *payloads, const uint64_t *payload_lens, uint64_t count,
                                       uint8_t *out, uint64_t out_len, scrs_decode_scratch *scratch);
scrs_status scrs_afft_decode_batch_impl(scrs_batch_decoder *dec, const uint16_t *indices,
                                        const uint8_t *const *payload_ptrs_dev, const uint8_t *data_dev,
                                        const uint8_t *repairs_dev, uint8_t *out_dev, uint64_t count, void *stream);
