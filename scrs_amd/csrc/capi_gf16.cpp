// capi_gf16.cpp - GF(2^16) engines (Tower, AFFT) behind the shared handles.
//
// The GF(2^16) device engines are not built yet: every entry point reports
// UnsupportedMode for the requested engine rather than computing anything.
#include "status.hpp"

scrs_status scrs_afft_encoder_create_impl(const scrs_profile *p, scrs_batch_encoder **out) {
    *out = nullptr;
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, static_cast<uint64_t>(p->engine));
}

void scrs_afft_encoder_destroy_impl(scrs_batch_encoder *) {}

scrs_status scrs_afft_encode_into_impl(const scrs_batch_encoder *, const uint8_t *, uint64_t, uint8_t *, uint64_t,
                                       scrs_encode_scratch *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}

scrs_status scrs_afft_encode_batch_dev_impl(const scrs_batch_encoder *, const uint8_t *, uint8_t *, uint64_t,
                                            void *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_AFFT);
}

scrs_status scrs_gf16_batch_decoder_create_impl(const scrs_profile *p, scrs_batch_decoder **out) {
    *out = nullptr;
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, static_cast<uint64_t>(p->engine));
}

void scrs_gf16_batch_decoder_destroy_impl(scrs_batch_decoder *) {}

scrs_status scrs_gf16_decode_into_impl(scrs_batch_decoder *, const uint64_t *, const uint8_t *const *,
                                       const uint64_t *, uint64_t, uint8_t *, uint64_t, scrs_decode_scratch *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_TOWER);
}

scrs_status scrs_gf16_decode_batch_impl(scrs_batch_decoder *, const uint16_t *, const uint8_t *const *,
                                        const uint8_t *, const uint8_t *, uint8_t *, uint64_t, void *) {
    return st(SCRS_CONFIG_UNSUPPORTED_MODE, SCRS_ENGINE_TOWER);
}

extern "C" const char *scrs_status_name(int32_t code) {
    switch (code) {
        case SCRS_OK: return "Ok";
        case SCRS_CONFIG_ZERO_DIMENSION: return "ConfigError::ZeroDimension";
        case SCRS_CONFIG_TOO_MANY_SYMBOLS: return "ConfigError::TooManySymbols";
        case SCRS_CONFIG_ZERO_SYMBOL_LEN: return "ConfigError::ZeroSymbolLen";
        case SCRS_CONFIG_ODD_SYMBOL_LEN: return "ConfigError::OddSymbolLen";
        case SCRS_CONFIG_UNSUPPORTED_MODE: return "ConfigError::UnsupportedMode";
        case SCRS_ENCODE_WRONG_INPUT_LEN: return "EncodeError::WrongInputLen";
        case SCRS_ENCODE_INDEX_OUT_OF_RANGE: return "EncodeError::IndexOutOfRange";
        case SCRS_ENCODE_WRONG_PAYLOAD_LEN: return "EncodeError::WrongPayloadLen";
        case SCRS_ENCODE_DUPLICATE_DATA: return "EncodeError::DuplicateData";
        case SCRS_ENCODE_WRONG_OUTPUT_LEN: return "EncodeError::WrongOutputLen";
        case SCRS_ENCODE_SCRATCH_MISMATCH: return "EncodeError::ScratchMismatch";
        case SCRS_DECODE_WRONG_COUNT: return "DecodeError::WrongCount";
        case SCRS_DECODE_INDEX_OUT_OF_RANGE: return "DecodeError::IndexOutOfRange";
        case SCRS_DECODE_DUPLICATE_INDEX: return "DecodeError::DuplicateIndex";
        case SCRS_DECODE_WRONG_PAYLOAD_LEN: return "DecodeError::WrongPayloadLen";
        case SCRS_DECODE_TOO_MANY_SYMBOLS: return "DecodeError::TooManySymbols";
        case SCRS_DECODE_INSUFFICIENT_RANK: return "DecodeError::InsufficientRank";
        case SCRS_DECODE_WRONG_OUTPUT_LEN: return "DecodeError::WrongOutputLen";
        case SCRS_DECODE_SCRATCH_MISMATCH: return "DecodeError::ScratchMismatch";
        case SCRS_ERR_INVALID_ARGUMENT: return "InvalidArgument";
        case SCRS_ERR_DEVICE: return "DeviceError";
        case SCRS_ERR_OUT_OF_MEMORY: return "OutOfMemory";
        case SCRS_ERR_CAPACITY: return "CapacityExceeded";
        default: return "Unknown";
    }
}
