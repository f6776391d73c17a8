// capi_afft.cpp - additive-FFT engine behind the shared handles.
//
// The additive-FFT device engine is not built yet: constructing an AFFT
// codec reports UnsupportedMode for the engine rather than computing
// anything.
#include "handles.hpp"
#include "status.hpp"

scrs_status scrs_afft_encoder_init_impl(scrs_batch_encoder *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}
void scrs_afft_encoder_release_impl(scrs_batch_encoder *) {}
scrs_status scrs_afft_encode_batch_dev_impl(const scrs_batch_encoder *, const uint8_t *, uint8_t *, uint64_t, void *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}
scrs_status scrs_afft_batch_decoder_init_impl(scrs_batch_decoder *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}
void scrs_afft_batch_decoder_release_impl(scrs_batch_decoder *) {}
scrs_status scrs_afft_decode_into_impl(scrs_batch_decoder *, const uint64_t *, const uint8_t *const *,
                                       const uint64_t *, uint64_t, uint8_t *, uint64_t, scrs_decode_scratch *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}
scrs_status scrs_afft_decode_batch_impl(scrs_batch_decoder *, const uint16_t *, const uint8_t *const *,
                                        const uint8_t *, const uint8_t *, uint8_t *, uint64_t, void *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}
