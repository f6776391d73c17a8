// status.hpp - status construction and the engine entry points shared
// between the C-ABI translation units.
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

// GF(2^16) engines behind the shared batch-encoder / batch-decoder handles.
scrs_status scrs_afft_encoder_create_impl(const scrs_profile *p, scrs_batch_encoder **out);
void scrs_afft_encoder_destroy_impl(scrs_batch_encoder *enc);
scrs_status scrs_afft_encode_into_impl(const scrs_batch_encoder *enc, const uint8_t *data, uint64_t data_len,
                                       uint8_t *repairs, uint64_t repairs_len, scrs_encode_scratch *scratch);
scrs_status scrs_afft_encode_batch_dev_impl(const scrs_batch_encoder *enc, const uint8_t *data_dev,
                                            uint8_t *repairs_dev, uint64_t count, void *stream);

scrs_status scrs_gf16_batch_decoder_create_impl(const scrs_profile *p, scrs_batch_decoder **out);
void scrs_gf16_batch_decoder_destroy_impl(scrs_batch_decoder *dec);
scrs_status scrs_gf16_decode_into_impl(scrs_batch_decoder *dec, const uint64_t *indices,
                                       const uint8_t *const *payloads, const uint64_t *payload_lens,
                                       uint64_t count, uint8_t *out, uint64_t out_len,
                                       scrs_decode_scratch *scratch);
scrs_status scrs_gf16_decode_batch_impl(scrs_batch_decoder *dec, const uint16_t *indices,
                                        const uint8_t *const *payload_ptrs_dev, const uint8_t *data_dev,
                                        const uint8_t *repairs_dev, uint8_t *out_dev, uint64_t count,
                                        void *stream);
