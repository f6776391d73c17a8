// capi_common.cpp - status names and the additive-FFT hooks.
#include <atomic>

#include "afft.hpp"
#include "handles.hpp"
#include "status.hpp"

extern "C" uint64_t scrs_device_alloc_count(void) { return scrs::alloc_counter().load(std::memory_order_relaxed); }

namespace scrs {
namespace {
std::atomic<int32_t> g_term_split{1};
std::atomic<int32_t> g_afft_targeted_max{static_cast<int32_t>(kAfftTargetedDefault)};
}  // namespace
bool term_split_enabled() { return g_term_split.load(std::memory_order_relaxed) != 0; }
uint32_t afft_targeted_max() { return static_cast<uint32_t>(g_afft_targeted_max.load(std::memory_order_relaxed)); }
}  // namespace scrs

extern "C" int32_t scrs_set_term_split(int32_t enable) {
    return scrs::g_term_split.exchange(enable ? 1 : 0, std::memory_order_relaxed);
}

extern "C" int32_t scrs_afft_set_targeted_max(int32_t max_missing) {
    const int32_t cap = static_cast<int32_t>(scrs::kAfftDirectCap);
    const int32_t v = max_missing < 0 ? 0 : (max_missing > cap ? cap : max_missing);
    return scrs::g_afft_targeted_max.exchange(v, std::memory_order_relaxed);
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
        case SCRS_TRANSFORM_WRONG_LEN: return "TransformLengthError";
        case SCRS_ERR_INVALID_ARGUMENT: return "InvalidArgument";
        case SCRS_ERR_DEVICE: return "DeviceError";
        case SCRS_ERR_OUT_OF_MEMORY: return "OutOfMemory";
        case SCRS_ERR_CAPACITY: return "CapacityExceeded";
        default: return "Unknown";
    }
}
