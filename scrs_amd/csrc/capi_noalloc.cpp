// capi_noalloc.cpp - batched GF(2^8) device decode that allocates nothing.
//
// For callers that cannot allocate memory during a call (code under HIP
// graph capture, tight serving loops): the per-launch workspace of
// scrs_decode_batch_dev (received indices, the closed-form solve's outputs
// and work, the pointer tables) lives in one file-scope device array instead
// of buffers the decoder grows on demand.  Each device holds its own copy of
// the array for the life of the process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "handles.hpp"
#include "status.hpp"

using namespace scrs;

namespace {

constexpr uint64_t kScratchBytes = 32ull << 20;  // per device, held until the process ends
constexpr uint64_t kMaxCount = 65535;             // codewords per call (one launch grid row each)
constexpr int kMaxDevices = 64;

// Device-side workspace.  Host code never passes `scratch` itself to a
// kernel or a copy: the name is a host-side stand-in, and the device address
// comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The array's address on the current device (looked up on the first call
// on that device).
hipError_t scratch_address(int device, uint8_t **out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<uint8_t *>(g_scratch_ptr[device]);
    return hipSuccess;
}

// decode_into's index checks (IndexOutOfRange -> DuplicateIndex per
// codeword, batch/codec.rs:304-317) with a bitmap on the stack: GF(2^8)
// codewords have at most 256 indices.
scrs_status check_indices(const uint16_t *idx, uint64_t count, uint64_t k, uint64_t n) {
    for (uint64_t c = 0; c < count; c++) {
        uint64_t seen[4] = {0, 0, 0, 0};
        for (uint64_t t = 0; t < k; t++) {
            const uint64_t i = idx[c * k + t];
            if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
            if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
            seen[i >> 6] |= 1ull << (i & 63);
        }
    }
    return ok();
}

}  // namespace

extern "C" uint64_t scrs_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" uint64_t scrs_decode_batch_dev_noalloc_bytes(const scrs_batch_decoder *dec, uint64_t count) {
    if (!dec || !engine_is_gf256(dec->prof.engine)) return 0;
    return dec->eng.workspace_bytes(count);
}

extern "C" scrs_status scrs_decode_batch_dev_noalloc(scrs_batch_decoder *dec, const uint16_t *indices,
                                                     const uint8_t *data_dev, const uint8_t *repairs_dev,
                                                     uint8_t *out_dev, uint64_t count, void *stream) {
    if (!dec) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!engine_is_gf256(dec->prof.engine)) return st(SCRS_CONFIG_UNSUPPORTED_MODE, dec->prof.engine);
    if (count == 0) return ok();
    if (!indices) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!data_dev || !repairs_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    if (!out_dev) return st(SCRS_ERR_INVALID_ARGUMENT, 5);
    const uint64_t k = dec->prof.k, n = k + dec->prof.m, L = dec->prof.symbol_len;
    scrs_status v = check_indices(indices, count, k, n);
    if (v.code != SCRS_OK) return v;
    // The whole launch must fit the fixed array: count * per-codeword bytes
    // (rounded per array) at most kScratchBytes.
    if (count > kMaxCount) return st(SCRS_ERR_CAPACITY, kMaxCount, count);
    const uint64_t need = dec->eng.workspace_bytes(count);
    if (need > kScratchBytes) return st(SCRS_ERR_CAPACITY, kScratchBytes, need);
    DeviceGuard g(dec->device);
    uint8_t *base = nullptr;
    hipError_t err = scratch_address(dec->device, &base);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    const DecodeWorkspace ws = dec->eng.carve_workspace(base, count);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // The indices go straight from the caller's memory: a copy from
    // pageable memory has consumed its source when it returns.
    if ((err = hipMemcpyAsync(ws.idx, indices, count * k * sizeof(uint16_t), hipMemcpyHostToDevice, s)) !=
        hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    const bool vec16 = L % 16 == 0 && (reinterpret_cast<uintptr_t>(out_dev) & 15u) == 0 &&
                       (reinterpret_cast<uintptr_t>(data_dev) & 15u) == 0 &&
                       (reinterpret_cast<uintptr_t>(repairs_dev) & 15u) == 0;
    // A split launch reuses the decoder's partial-row buffer only if it
    // already holds it, else XORs atomically.
    return dec->eng.run_chunk(ws, nullptr, data_dev, repairs_dev, out_dev, count, s, vec16, false,
                              /*may_allocate=*/false);
}
