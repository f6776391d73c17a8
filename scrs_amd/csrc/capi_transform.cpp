// capi_transform.cpp - C ABI: the additive-FFT transform plan on its own
// (TransformPlan, scrs/src/afft/transform.rs:24-95).
//
// The transform runs on the codecs' butterfly passes: a block of `size` rows
// is one codeword of the pass kernel, every element column an independent
// transform.  Forward passes go from the top bits down (the root's level
// first), inverse passes from the bottom bits up, each pass at most 8 levels
// (afft/transform.rs:301-331).
#include <cstring>
#include <mutex>

#include "afft.hpp"
#include "afft_math.hpp"
#include "status.hpp"

using namespace scrs;

struct scrs_transform_plan {
    uint64_t size = 0;
    uint32_t log_size = 0;
    int device = 0;
    const AfftDeviceTables *dev = nullptr;
    // Host-slice calls: a stream and staging buffers, serialised by `mu`.
    mutable std::mutex mu;
    mutable hipStream_t stream = nullptr;
    mutable DevBuf d_values;
    mutable HostBuf h_values;
};

namespace {

constexpr uint32_t kTransformPassLog = 8;

hipError_t run_passes(const scrs_transform_plan *p, uint8_t *rows, uint64_t L, uint64_t count, bool forward,
                      hipStream_t s) {
    const uint32_t D = p->log_size;
    if (D == 0) return hipSuccess;  // one point: the transform is the identity
    const uint32_t parts = (D + kTransformPassLog - 1) / kTransformPassLog;
    uint32_t lo[kAfftMaxLog], g[kAfftMaxLog];
    uint32_t b = 0;
    for (uint32_t i = 0; i < parts; i++) {
        g[i] = D / parts + (i < D % parts ? 1u : 0u);
        lo[i] = b;
        b += g[i];
    }
    AfftPassArgs base;
    base.node_fac = p->dev->node_fac.as<uint16_t>();
    base.deriv_tabs = p->dev->deriv_tabs.as<uint16_t>();
    base.L = L;
    base.in_base = rows;
    base.in_limit = p->size;
    base.in_cw_stride = p->size * L;
    base.out_base = rows;
    base.out_first = 0;
    base.out_limit = p->size;
    base.out_cw_stride = p->size * L;
    base.inv_base = 0;
    base.fwd_base = 0;
    for (uint32_t j = 0; j < parts; j++) {
        const uint32_t i = forward ? parts - 1 - j : j;
        AfftPassArgs a = base;
        a.b_lo = lo[i];
        a.g = g[i];
        a.flags = forward ? kAfftFwd : kAfftInv;
        hipError_t err = launch_afft_pass(a, D, static_cast<uint32_t>(count), s);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

scrs_status transform_dev(const scrs_transform_plan *p, uint8_t *rows, uint64_t L, uint64_t count, bool forward,
                          hipStream_t s) {
    if (!p) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count == 0) return ok();
    if (!rows) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (L == 0) return st(SCRS_CONFIG_ZERO_SYMBOL_LEN);
    if (L % 2) return st(SCRS_CONFIG_ODD_SYMBOL_LEN);
    if (count > 65535) return st(SCRS_ERR_CAPACITY, 65535);
    DeviceGuard dg(p->device);
    hipError_t err = run_passes(p, rows, L, count, forward, s);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

scrs_status transform_host(const scrs_transform_plan *p, uint16_t *values, uint64_t len, bool forward) {
    if (!p) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (len != p->size) return st(SCRS_TRANSFORM_WRONG_LEN, p->size, len);
    if (!values) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard dg(p->device);
    const uint64_t bytes = len * 2;
    hipError_t err = hipSuccess;
    if (!p->stream) err = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = p->d_values.ensure(bytes);
    if (err == hipSuccess) err = p->h_values.ensure(bytes);
    if (err != hipSuccess) return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
    std::memcpy(p->h_values.ptr, values, bytes);
    hipStream_t s = p->stream;
    if ((err = hipMemcpyAsync(p->d_values.ptr, p->h_values.ptr, bytes, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (err = run_passes(p, p->d_values.as<uint8_t>(), 2, 1, forward, s)) != hipSuccess ||
        (err = hipMemcpyAsync(p->h_values.ptr, p->d_values.ptr, bytes, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (err = hipStreamSynchronize(s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    std::memcpy(values, p->h_values.ptr, bytes);
    return ok();
}

}  // namespace

extern "C" scrs_status scrs_transform_plan_create(uint64_t size, scrs_transform_plan **out) {
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    if (size == 0 || size > 65536 || (size & (size - 1))) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    auto *p = new scrs_transform_plan();
    p->size = size;
    while ((uint64_t{1} << p->log_size) < size) p->log_size++;
    (void)hipGetDevice(&p->device);
    hipError_t err = afft_device_tables(&p->dev);
    if (err != hipSuccess) {
        delete p;
        return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
    }
    *out = p;
    return ok();
}

extern "C" void scrs_transform_plan_destroy(scrs_transform_plan *plan) {
    if (!plan) return;
    DeviceGuard dg(plan->device);
    if (plan->stream) {
        (void)hipStreamSynchronize(plan->stream);
        (void)hipStreamDestroy(plan->stream);
    }
    delete plan;
}

extern "C" uint64_t scrs_transform_plan_size(const scrs_transform_plan *plan) { return plan ? plan->size : 0; }
extern "C" uint32_t scrs_transform_plan_log_size(const scrs_transform_plan *plan) {
    return plan ? plan->log_size : 0;
}

extern "C" scrs_status scrs_transform_forward(const scrs_transform_plan *plan, uint16_t *values, uint64_t len) {
    return transform_host(plan, values, len, true);
}

extern "C" scrs_status scrs_transform_inverse(const scrs_transform_plan *plan, uint16_t *values, uint64_t len) {
    return transform_host(plan, values, len, false);
}

extern "C" scrs_status scrs_transform_forward_dev(const scrs_transform_plan *plan, uint8_t *rows_dev,
                                                  uint64_t symbol_len, uint64_t count, void *stream) {
    return transform_dev(plan, rows_dev, symbol_len, count, true, static_cast<hipStream_t>(stream));
}

extern "C" scrs_status scrs_transform_inverse_dev(const scrs_transform_plan *plan, uint8_t *rows_dev,
                                                  uint64_t symbol_len, uint64_t count, void *stream) {
    return transform_dev(plan, rows_dev, symbol_len, count, false, static_cast<hipStream_t>(stream));
}
