// engine.cpp - device decode engine and coefficient plans.
#include "engine.hpp"

#include <algorithm>
#include <cstring>

#include "matrices.hpp"
#include "status.hpp"

namespace scrs {

size_t engine_capacity(int32_t engine) {
    switch (engine) {
        case SCRS_ENGINE_STANDARD_CAUCHY: return kCapStandard;
        case SCRS_ENGINE_GOOD_CAUCHY: return kCapGood;
        case SCRS_ENGINE_TOWER: return kCapTower;
        default: return kCapAfft;
    }
}

std::vector<uint16_t> cauchy_coordinates(const scrs_profile &p) {
    const size_t n = p.k + p.m;
    std::vector<uint16_t> c(n);
    if (p.engine == SCRS_ENGINE_STANDARD_CAUCHY) {
        for (size_t i = 0; i < n; i++) c[i] = static_cast<uint16_t>(i);
    } else if (p.engine == SCRS_ENGINE_GOOD_CAUCHY) {
        for (size_t i = 0; i < n; i++) c[i] = gf256().pow3(i);
    } else {
        uint16_t v = 1;
        for (size_t i = 0; i < n; i++) {
            c[i] = v;
            v = Gf16::mul(v, Gf16::kGenerator);
        }
    }
    return c;
}

std::vector<uint32_t> coefficient_tables(const scrs_profile &p, const std::vector<uint32_t> &terms,
                                         uint32_t *groups_out) {
    const size_t k = p.k, m = p.m, T = terms.size();
    const uint32_t groups = static_cast<uint32_t>((m + 15) / 16);
    *groups_out = groups;
    if (engine_is_gf256(p.engine)) {
        const std::vector<uint8_t> c = p.engine == SCRS_ENGINE_GOOD_CAUCHY ? good_cauchy_matrix(k, m)
                                                                           : standard_cauchy_matrix(k, m);
        std::vector<uint8_t> coef(T * m);
        for (size_t t = 0; t < T; t++)
            for (size_t j = 0; j < m; j++) coef[t * m + j] = c[terms[t] * m + j];
        std::vector<uint32_t> tab(T * groups * 128);
        build_nibble_tables(coef.data(), T, m, groups, tab.data());
        return tab;
    }
    const std::vector<uint16_t> c = tower_cauchy_matrix(k, m);
    std::vector<uint16_t> coef(T * m);
    for (size_t t = 0; t < T; t++)
        for (size_t j = 0; j < m; j++) coef[t * m + j] = c[terms[t] * m + j];
    std::vector<uint32_t> tab(T * groups * 512);
    build_nibble_tables16(coef.data(), T, m, groups, tab.data());
    return tab;
}

hipError_t launch_rows_terms_field(bool gf16, const RowsTermsArgs &a, bool vec16, hipStream_t s) {
    return gf16 ? launch_rows_terms16(a, vec16, s) : launch_rows_terms_lds(a, vec16, s);
}

SplitBuffer::~SplitBuffer() {
    if (done) (void)hipEventDestroy(done);
}

bool SplitBuffer::acquire(RowsTermsArgs &a, uint32_t rows, hipStream_t s, bool may_allocate) {
    if (rows == 0 || rows > a.rmax || !a.dst_aligned4 || a.L % 4) return false;
    const uint32_t segs = rows_terms_segments(a);
    if (segs < 2) return false;
    const uint64_t need = a.count * rows * segs * a.L;
    if (need > kMaxBytes) return false;
    mu.lock();
    hipError_t err = hipSuccess;
    if (need > buf.size) err = may_allocate ? buf.ensure(need) : hipErrorOutOfMemory;
    if (err == hipSuccess && !done) err = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (err == hipSuccess) err = hipStreamWaitEvent(s, done, 0);
    if (err != hipSuccess) {
        (void)hipGetLastError();  // no buffer only means atomics
        mu.unlock();
        return false;
    }
    a.partials = buf.as<uint8_t>();
    a.partial_rows = rows;
    a.partials_bytes = need;
    return true;
}

void SplitBuffer::release(hipStream_t s) {
    (void)hipEventRecord(done, s);
    mu.unlock();
}

DecodeEngine::~DecodeEngine() {
    if (idx_copied) (void)hipEventDestroy(idx_copied);
}

scrs_status DecodeEngine::init(const scrs_profile &p) {
    prof = p;
    gf16 = !engine_is_gf256(p.engine);
    k = static_cast<uint32_t>(p.k);
    m = static_cast<uint32_t>(p.m);
    L = p.symbol_len;
    rmax = k < m ? k : m;
    groups = (rmax + 15) / 16;
    return init_device(cauchy_coordinates(p));
}

scrs_status DecodeEngine::init_afft_direct(const scrs_profile &p, uint32_t max_missing,
                                           const std::vector<uint16_t> &wlog) {
    prof = p;
    gf16 = true;
    k = static_cast<uint32_t>(p.k);
    m = static_cast<uint32_t>(p.m);
    L = p.symbol_len;
    rmax = std::min({k, m, max_missing});
    groups = (rmax + 15) / 16;
    weighted = true;
    force_split = true;
    const uint64_t n = uint64_t{k} + m;
    if (wlog.size() != n) return st(SCRS_ERR_INVALID_ARGUMENT, 3);
    std::vector<uint16_t> coords(n);
    for (uint64_t i = 0; i < n; i++) coords[i] = static_cast<uint16_t>(i);
    hipError_t err = d_wlog.ensure(n * sizeof(uint16_t));
    if (err == hipSuccess) err = hipMemcpy(d_wlog.ptr, wlog.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (err != hipSuccess)
        return err == hipErrorOutOfMemory ? st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(p.engine))
                                          : st(SCRS_ERR_DEVICE, err);
    return init_device(coords);
}

scrs_status DecodeEngine::init_device(const std::vector<uint16_t> &coords) {
    const scrs_profile &p = prof;
    hipError_t err = d_coords.ensure(coords.size() * sizeof(uint16_t));
    if (err == hipSuccess)
        err = hipMemcpy(d_coords.ptr, coords.data(), coords.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&idx_copied, hipEventDisableTiming);
    if (err == hipSuccess) err = d_flag.ensure(sizeof(uint32_t));
    if (err == hipSuccess) err = hipMemset(d_flag.ptr, 0, sizeof(uint32_t));
    if (err != hipSuccess) {
        if (err == hipErrorOutOfMemory) return st(SCRS_CONFIG_TOO_MANY_SYMBOLS, engine_capacity(p.engine));
        return st(SCRS_ERR_DEVICE, err);
    }
    // Bound the per-launch workspace: product tables dominate (the larger
    // of the two GF(2^8) kernels' layouts).
    uint64_t table_bytes = cw_table_bytes(0);
    for (uint32_t bm : {kBitmatrixInt8, kBitmatrixFp4})
        if (!gf16 && k <= bitmatrix_max_terms(bm) && cw_table_bytes(bm) > table_bytes) table_bytes = cw_table_bytes(bm);
    const uint64_t budget = 1ull << 30;
    chunk = table_bytes ? budget / table_bytes : 65535;
    if (chunk > 65535) chunk = 65535;
    return ok();
}

scrs_status DecodeEngine::enqueue(const uint16_t *indices, const uint8_t *const *payload_ptrs,
                                  const uint8_t *data_base, const uint8_t *repair_base, uint8_t *out,
                                  uint64_t count, hipStream_t s, bool vec16, uint32_t bitmatrix) {
    if (count == 0) return ok();
    if (bitmatrix && (gf16 || k > bitmatrix_max_terms(bitmatrix) || L % 4)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    const uint64_t tb = cw_table_bytes(bitmatrix);
    if (chunk == 0) return st(SCRS_ERR_CAPACITY, tb);
    const uint64_t want = count < chunk ? count : chunk;
    hipError_t err;
    if (want > cap) {
        const uint64_t kk = k, rr = rmax ? rmax : 1;
        const uint64_t work = cauchy_solve_work_elems(k, m, rmax);
        if ((err = d_idx.ensure(want * kk * 2)) != hipSuccess || (err = h_idx.ensure(want * kk * 2)) != hipSuccess ||
            (err = d_fused.ensure(want * rr * kk * 2)) != hipSuccess ||
            (err = d_missing.ensure(want * rr * 2)) != hipSuccess || (err = d_rows.ensure(want * 4)) != hipSuccess ||
            (err = d_work.ensure(want * work * 2)) != hipSuccess ||
            (err = d_src.ensure(want * kk * 8)) != hipSuccess || (err = d_dst.ensure(want * rr * 8)) != hipSuccess ||
            (err = d_copy.ensure(want * kk * 8)) != hipSuccess)
            return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
        cap = want;
    }
    // HBM product tables only for the kernels that read them.
    const bool hbm_tables = bitmatrix != 0 || write_tables;
    if (hbm_tables && (err = d_tables.ensure(cap * tb)) != hipSuccess)
        return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
    DecodeWorkspace ws;
    ws.idx = d_idx.as<uint16_t>();
    ws.fused = d_fused.as<uint16_t>();
    ws.missing = d_missing.as<uint16_t>();
    ws.rows = d_rows.as<uint32_t>();
    ws.work = d_work.as<uint16_t>();
    ws.src = d_src.as<const uint8_t *>();
    ws.dst = d_dst.as<uint8_t *>();
    ws.copy = d_copy.as<uint8_t *>();
    ws.tables = hbm_tables ? d_tables.as<uint32_t>() : nullptr;
    ws.cap = cap;
    for (uint64_t c0 = 0; c0 < count; c0 += cap) {
        const uint64_t cn = count - c0 < cap ? count - c0 : cap;
        // The pinned index buffer is reused: wait for the previous upload.
        if ((err = hipEventSynchronize(idx_copied)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        std::memcpy(h_idx.ptr, indices + c0 * k, cn * k * sizeof(uint16_t));
        if ((err = hipMemcpyAsync(d_idx.ptr, h_idx.ptr, cn * k * sizeof(uint16_t), hipMemcpyHostToDevice, s)) !=
                hipSuccess ||
            (err = hipEventRecord(idx_copied, s)) != hipSuccess)
            return st(SCRS_ERR_DEVICE, err);
        scrs_status r = run_chunk(ws, payload_ptrs ? payload_ptrs + c0 * k : nullptr,
                                  data_base ? data_base + c0 * k * L : nullptr,
                                  repair_base ? repair_base + c0 * m * L : nullptr, out + c0 * k * L, cn, s, vec16,
                                  bitmatrix);
        if (r.code != SCRS_OK) return r;
    }
    return ok();
}

static uint64_t round256(uint64_t x) { return (x + 255) & ~uint64_t{255}; }

uint64_t DecodeEngine::workspace_bytes(uint64_t count) const {
    const uint64_t kk = k, rr = rmax ? rmax : 1;
    return round256(count * kk * 2) + round256(count * rr * kk * 2) + round256(count * rr * 2) +
           round256(count * 4) + round256(count * cauchy_solve_work_elems(k, m, rmax) * 2) +
           round256(count * kk * 8) + round256(count * rr * 8) + round256(count * kk * 8);
}

DecodeWorkspace DecodeEngine::carve_workspace(uint8_t *base, uint64_t count) const {
    const uint64_t kk = k, rr = rmax ? rmax : 1;
    DecodeWorkspace ws;
    uint8_t *p = base;
    auto take = [&p](uint64_t bytes) {
        uint8_t *q = p;
        p += round256(bytes);
        return q;
    };
    ws.idx = reinterpret_cast<uint16_t *>(take(count * kk * 2));
    ws.fused = reinterpret_cast<uint16_t *>(take(count * rr * kk * 2));
    ws.missing = reinterpret_cast<uint16_t *>(take(count * rr * 2));
    ws.rows = reinterpret_cast<uint32_t *>(take(count * 4));
    ws.work = reinterpret_cast<uint16_t *>(take(count * cauchy_solve_work_elems(k, m, rmax) * 2));
    ws.src = reinterpret_cast<const uint8_t **>(take(count * kk * 8));
    ws.dst = reinterpret_cast<uint8_t **>(take(count * rr * 8));
    ws.copy = reinterpret_cast<uint8_t **>(take(count * kk * 8));
    ws.tables = nullptr;
    ws.cap = count;
    return ws;
}

scrs_status DecodeEngine::run_chunk(const DecodeWorkspace &ws, const uint8_t *const *payload_ptrs,
                                    const uint8_t *data_base, const uint8_t *repair_base, uint8_t *out, uint64_t cn,
                                    hipStream_t s, bool vec16, uint32_t bitmatrix, bool may_allocate) {
    const uint64_t tb = cw_table_bytes(bitmatrix);
    // Table path (either field): the multiply-XOR builds its product tables
    // in LDS from the fused coefficients, so none are written to HBM, unless
    // a recipe cache asked for them.
    const bool fused_in_rows = !bitmatrix && !write_tables;
    if (!fused_in_rows && !ws.tables) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (cn > ws.cap) return st(SCRS_ERR_CAPACITY, ws.cap);
    hipError_t err;
    CauchySolveArgs sa;
    sa.indices = ws.idx;
    sa.coords = d_coords.as<uint16_t>();
    sa.fused = ws.fused;
    sa.missing = ws.missing;
    sa.rows = ws.rows;
    sa.work = ws.work;
    sa.work_stride = cauchy_solve_work_elems(k, m, rmax);
    sa.error_flag = d_flag.as<uint32_t>();
    sa.k = k;
    sa.m = m;
    sa.rmax = rmax;
    sa.count = cn;
    if (weighted) sa.point_wlog = d_wlog.as<uint16_t>();
    if ((err = launch_cauchy_solve(sa, gf16, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    DecodePrepArgs pa;
    pa.indices = ws.idx;
    pa.fused = ws.fused;
    pa.missing = ws.missing;
    pa.rows = ws.rows;
    pa.payload_ptrs = payload_ptrs;
    pa.data_base = data_base;
    pa.repair_base = repair_base;
    pa.out_base = out;
    pa.tables = ws.tables;
    pa.src_ptrs = ws.src;
    pa.dst_ptrs = ws.dst;
    pa.copy_ptrs = ws.copy;
    pa.k = k;
    pa.m = m;
    pa.rmax = rmax;
    pa.groups = groups;
    pa.L = L;
    pa.count = cn;
    if (bitmatrix) {
        // Bit-matrix operands from the solve; the prepare step then
        // writes only the pointer tables.
        if ((err = launch_bitmatrix_tables(pa, s, bitmatrix)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        pa.fused = nullptr;
    } else if (fused_in_rows) {
        pa.fused = nullptr;
    }
    if ((err = launch_decode_prepare(pa, gf16, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    RowsTermsArgs ra;
    ra.src_ptrs = ws.src;
    ra.dst_ptrs = ws.dst;
    ra.copy_ptrs = ws.copy;
    ra.rows = ws.rows;
    ra.tables = ws.tables;
    ra.table_cw_stride = tb / 4;
    ra.T = k;
    ra.rmax = rmax;
    ra.groups = groups;
    ra.L = L;
    ra.count = cn;
    if (fused_in_rows) ra.fused = ws.fused;
    ra.force_split = force_split;
    // Missing rows are out + index * L.
    ra.dst_aligned4 = (reinterpret_cast<uintptr_t>(out) & 3u) == 0 && L % 4 == 0;
    const bool attached = !bitmatrix && split.acquire(ra, force_split && rows_hint ? rows_hint : rmax, s, may_allocate);
    err = bitmatrix ? launch_rows_bitmatrix(ra, s, bitmatrix) : launch_rows_terms_field(gf16, ra, vec16, s);
    if (attached) split.release(s);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

scrs_status DecodeEngine::enqueue_recipe(const uint16_t *indices, const uint32_t *tables, const uint16_t *missing,
                                         const uint32_t *rows, const uint8_t *data_base,
                                         const uint8_t *repair_base, uint8_t *out, hipStream_t s, bool vec16) {
    hipError_t err;
    if (cap == 0) {
        // Pointer tables and indices for one codeword.
        if ((err = d_idx.ensure(k * 2ull)) != hipSuccess || (err = h_idx.ensure(k * 2ull)) != hipSuccess ||
            (err = d_src.ensure(k * 8ull)) != hipSuccess || (err = d_dst.ensure((rmax ? rmax : 1) * 8ull)) != hipSuccess ||
            (err = d_copy.ensure(k * 8ull)) != hipSuccess)
            return st(SCRS_ERR_DEVICE, err);
    }
    if ((err = hipEventSynchronize(idx_copied)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    std::memcpy(h_idx.ptr, indices, k * sizeof(uint16_t));
    if ((err = hipMemcpyAsync(d_idx.ptr, h_idx.ptr, k * sizeof(uint16_t), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (err = hipEventRecord(idx_copied, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    DecodePrepArgs pa;
    pa.indices = d_idx.as<uint16_t>();
    pa.missing = missing;
    pa.rows = rows;
    pa.data_base = data_base;
    pa.repair_base = repair_base;
    pa.out_base = out;
    pa.src_ptrs = d_src.as<const uint8_t *>();
    pa.dst_ptrs = d_dst.as<uint8_t *>();
    pa.copy_ptrs = d_copy.as<uint8_t *>();
    pa.k = k;
    pa.m = m;
    pa.rmax = rmax;
    pa.groups = groups;
    pa.L = L;
    pa.count = 1;
    if ((err = launch_decode_prepare(pa, gf16, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    RowsTermsArgs ra;
    ra.src_ptrs = d_src.as<const uint8_t *const>();
    ra.dst_ptrs = d_dst.as<uint8_t *const>();
    ra.copy_ptrs = d_copy.as<uint8_t *const>();
    ra.rows = rows;
    ra.tables = tables;
    ra.table_cw_stride = 0;
    ra.T = k;
    ra.rmax = rmax;
    ra.groups = groups;
    ra.L = L;
    ra.count = 1;
    ra.dst_aligned4 = (reinterpret_cast<uintptr_t>(out) & 3u) == 0 && L % 4 == 0;
    const bool attached = split.acquire(ra, rmax, s, true);
    err = launch_rows_terms_field(gf16, ra, vec16, s);
    if (attached) split.release(s);
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

hipError_t DecodeEngine::read_error(uint32_t *flag, hipStream_t s) {
    hipError_t err = hipMemcpyAsync(flag, d_flag.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (err != hipSuccess) return err;
    return hipStreamSynchronize(s);
}

hipError_t DecodeEngine::clear_error(hipStream_t s) { return hipMemsetAsync(d_flag.ptr, 0, sizeof(uint32_t), s); }

}  // namespace scrs
