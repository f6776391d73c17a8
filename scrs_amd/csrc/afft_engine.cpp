// afft_engine.cpp - additive-FFT Reed-Solomon codec on the device.
//
// Encode, k a power of two and m <= k (afft/encoder.rs:112-126): inverse
// transform of the k data rows over the points [0, k), then the forward
// transform of those coefficients over the coset [k, 2k); the first m rows
// are the repairs.
//
// Everything else is erasure decoding over the domain [0, N), N the power
// of two >= k + m (afft/decoder.rs:269-350): with Lambda the product of
// (x + e) over the points e without a value, the rows Lambda(x) P(x) of the
// known points go through the inverse transform, the formal derivative and
// the forward transform, and P(e) = (Lambda P)'(e) / Lambda'(e).  The encode
// of other shapes is the same computation with the data points known and
// the repair points as targets.
//
// Above one pass the derivative runs with unit factors: the last inverse
// pass scales coefficient row q by T(q), the product of the derivative
// factors d_b over the bits b of q, the derivative passes only XOR rows, and
// the pass that completes it scales row p by 1/T(p) before the forward
// levels.  Since T(p | 2^b) = T(p) d_b for a bit b clear in p, that equals
// the reference's sum of d_b C[p | 2^b] (transform.rs:188-213) with two
// products per row instead of one per term.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "afft.hpp"
#include "afft_math.hpp"
#include "matrices.hpp"
#include "status.hpp"

namespace scrs {

namespace {

uint32_t log2_ceil(uint64_t x) {
    uint32_t d = 0;
    while ((uint64_t{1} << d) < x) d++;
    return d;
}

// Consecutive bit ranges (first bit, count) of a transform's passes, held
// inline: a codec call plans its passes without touching the heap.
struct BitParts {
    std::pair<uint32_t, uint32_t> p[kAfftMaxLog];
    uint32_t n = 0;
    uint32_t size() const { return n; }
    const std::pair<uint32_t, uint32_t> &operator[](uint32_t i) const { return p[i]; }
};

// Split bits [0, D) into ceil(D / cap) nearly equal consecutive parts.
BitParts plan_bits(uint32_t D, uint32_t cap) {
    BitParts parts;
    const uint32_t s = (D + cap - 1) / cap;
    uint32_t lo = 0;
    for (uint32_t i = 0; i < s && i < kAfftMaxLog; i++) {
        const uint32_t g = D / s + (i < D % s ? 1 : 0);
        parts.p[parts.n++] = {lo, g};
        lo += g;
    }
    return parts;
}

// Largest pass whose LDS still holds every table a single-pass decode builds.
constexpr uint32_t kSinglePassLog = 7;
#ifndef SCRS_AFFT_PASS_LOG
#define SCRS_AFFT_PASS_LOG 8
#endif
// Levels per butterfly pass at most.
constexpr uint32_t kPassLog = SCRS_AFFT_PASS_LOG;
constexpr uint64_t kWorkBudget = 1ull << 30;

}  // namespace

hipError_t afft_device_tables(const AfftDeviceTables **out) {
    static std::mutex mu;
    static std::map<int, AfftDeviceTables *> cache;
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) {
        *out = it->second;
        return hipSuccess;
    }
    const AfftMath &m = afft_math();
    auto *t = new AfftDeviceTables();
    DevBuf consts;
    std::vector<uint16_t> d(m.deriv, m.deriv + kAfftMaxLog);
    // T(q) and 1/T(q) through logarithms; every derivative factor is
    // nonzero (the linear coefficient of a normalised subspace polynomial).
    std::vector<uint16_t> dscale(65536), dunscale(65536);
    for (uint32_t q = 0; q < 65536; q++) {
        uint32_t lg = 0;
        for (uint32_t b = 0; b < kAfftMaxLog; b++)
            if ((q >> b) & 1u) lg = (lg + m.log[m.deriv[b]]) % kAfftOrder;
        dscale[q] = m.exp[lg];
        dunscale[q] = m.exp[(kAfftOrder - lg) % kAfftOrder];
    }
    auto build = [&](const std::vector<uint16_t> &c, DevBuf &fac) -> hipError_t {
        hipError_t e;
        if ((e = fac.ensure(c.size() * 2)) != hipSuccess) return e;
        return hipMemcpy(fac.ptr, c.data(), c.size() * 2, hipMemcpyHostToDevice);
    };
    if ((err = consts.ensure(65536 * 2)) == hipSuccess &&
        (err = t->deriv_tabs.ensure(kAfftMaxLog * 128)) == hipSuccess &&
        (err = t->exp.ensure(kAfftOrder * 2)) == hipSuccess &&
        (err = hipMemcpy(t->exp.ptr, m.exp.data(), kAfftOrder * 2, hipMemcpyHostToDevice)) == hipSuccess &&
        (err = t->node_fac.ensure(65536 * 2)) == hipSuccess &&
        (err = hipMemcpy(t->node_fac.ptr, m.factors.data(), 65536 * 2, hipMemcpyHostToDevice)) == hipSuccess &&
        (err = hipMemcpy(consts.ptr, d.data(), kAfftMaxLog * 2, hipMemcpyHostToDevice)) == hipSuccess &&
        (err = launch_afft_tables(consts.as<uint16_t>(), kAfftMaxLog, t->deriv_tabs.as<uint16_t>(), nullptr)) ==
            hipSuccess &&
        (err = hipStreamSynchronize(nullptr)) == hipSuccess && (err = build(dscale, t->deriv_scale)) == hipSuccess)
        err = build(dunscale, t->deriv_unscale);
    if (err != hipSuccess) {
        delete t;
        return err;
    }
    cache[dev] = t;
    *out = t;
    return hipSuccess;
}

AfftEngine::~AfftEngine() {
    if (bits_copied) (void)hipEventDestroy(bits_copied);
    if (last) (void)hipEventDestroy(last);
}

scrs_status AfftEngine::init(const scrs_profile &p, bool for_decode) {
    prof = p;
    k = static_cast<uint32_t>(p.k);
    m = static_cast<uint32_t>(p.m);
    n = k + m;
    L = p.symbol_len;
    D = log2_ceil(n);
    Dk = log2_ceil(k);
    N = uint64_t{1} << D;
    fused = (uint64_t{1} << Dk) == k && D == Dk + 1 && k >= 2;
    hipError_t err = afft_device_tables(&dev);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&last, hipEventDisableTiming);
    auto fail = [&](hipError_t e) {
        return e == hipErrorOutOfMemory ? st(SCRS_CONFIG_TOO_MANY_SYMBOLS, kCapAfft) : st(SCRS_ERR_DEVICE, e);
    };
    if (err != hipSuccess) return fail(err);
    if (!for_decode && !fused) {
        std::vector<uint8_t> known(N, 0);
        for (uint32_t x = 0; x < k; x++) known[x] = 1;
        const std::vector<uint16_t> c = afft_locator_scales(known, D);
        if ((err = enc_scale.ensure(N * 2)) != hipSuccess ||
            (err = hipMemcpy(enc_scale.ptr, c.data(), N * 2, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(err);
    }
    if (for_decode) {
        const std::vector<uint32_t> lw = afft_log_wht(D);
        // Whole 64-bit words, so a validation bitmap (64-bit words) copies in.
        const uint64_t words = (N + 63) / 64 * 2;
        h_bits.assign(words, 0);
        if ((err = log_wht.ensure(N * 4)) != hipSuccess ||
            (err = hipMemcpy(log_wht.ptr, lw.data(), N * 4, hipMemcpyHostToDevice)) != hipSuccess ||
            (err = bits.ensure(words * 4)) != hipSuccess || (err = h_bits_pinned.ensure(words * 4)) != hipSuccess ||
            (err = loc_work.ensure(N * 4)) != hipSuccess || (err = loc_out.ensure(N * 2)) != hipSuccess ||
            (err = in_rows.ensure(N * 8)) != hipSuccess ||
            (err = copy_rows.ensure(N * 8)) != hipSuccess || (err = out_rows.ensure(N * 8)) != hipSuccess ||
            (err = work.ensure(N * L)) != hipSuccess || (err = work2.ensure(N * L)) != hipSuccess ||
            (err = hipEventCreateWithFlags(&bits_copied, hipEventDisableTiming)) != hipSuccess)
            return fail(err);
        // Direct decode weights from the locator of the data points
        // (afft/decoder.rs:256-267): with the points x >= k known, the
        // locator constants are P(x) there and 1 / P'(d) at the data
        // points d < k.
        const AfftMath &am = afft_math();
        std::vector<uint8_t> known(N, 0);
        for (uint64_t x = k; x < N; x++) known[x] = 1;
        const std::vector<uint16_t> c = afft_locator_scales(known, D);
        std::vector<uint16_t> wlog(n);
        for (uint32_t x = 0; x < n; x++) {
            const uint32_t lg = am.log[c[x]];
            wlog[x] = static_cast<uint16_t>(x < k ? lg : (kAfftOrder - lg) % kAfftOrder);
        }
        scrs_status r = direct.init_afft_direct(p, kAfftDirectCap, wlog);
        if (r.code != SCRS_OK) return r;
    }
    return ok();
}

scrs_status AfftEngine::encode(const uint8_t *data, uint8_t *repairs, uint64_t count, hipStream_t s) {
    if (count == 0) return ok();
    const uint64_t rows = fused ? k : N;  // work rows per codeword
    const uint64_t per_cw = rows * L * (fused ? 1 : 2);
    uint64_t chunk = per_cw ? kWorkBudget / per_cw : count;
    if (chunk == 0) chunk = 1;
    if (chunk > count) chunk = count;
    if (chunk > 65535) chunk = 65535;
    hipError_t err;
    if (chunk > enc_cap) {
        if ((err = work.ensure(chunk * rows * L)) != hipSuccess ||
            (!fused && (err = work2.ensure(chunk * rows * L)) != hipSuccess))
            return st(err == hipErrorOutOfMemory ? SCRS_ERR_OUT_OF_MEMORY : SCRS_ERR_DEVICE, err);
        enc_cap = chunk;
    }
    AfftPassArgs base;
    base.node_fac = dev->node_fac.as<uint16_t>();
    base.deriv_tabs = dev->deriv_tabs.as<uint16_t>();
    base.L = L;
    uint8_t *w1 = work.as<uint8_t>(), *w2 = work2.as<uint8_t>();
    const uint64_t wcw = rows * L;
    for (uint64_t c0 = 0; c0 < count; c0 += chunk) {
        const uint32_t cn = static_cast<uint32_t>(count - c0 < chunk ? count - c0 : chunk);
        const uint8_t *din = data + c0 * k * L;
        uint8_t *rout = repairs + c0 * m * L;
        auto set_in_data = [&](AfftPassArgs &a) {
            a.in_base = din;
            a.in_limit = k;
            a.in_cw_stride = k * L;
        };
        auto set_in_work = [&](AfftPassArgs &a, const uint8_t *w) {
            a.in_base = w;
            a.in_limit = rows;
            a.in_cw_stride = wcw;
        };
        auto set_out_work = [&](AfftPassArgs &a, uint8_t *w) {
            a.out_base = w;
            a.out_first = 0;
            a.out_limit = rows;
            a.out_cw_stride = wcw;
        };
        auto set_out_repairs = [&](AfftPassArgs &a, uint64_t first) {
            a.out_base = rout;
            a.out_first = first;
            a.out_limit = first + m;
            a.out_cw_stride = m * L;
        };
        if (fused) {
            const auto parts = plan_bits(Dk, kPassLog);
            const uint32_t ns = static_cast<uint32_t>(parts.size());
            for (uint32_t i = 0; i < ns; i++) {
                AfftPassArgs a = base;
                a.b_lo = parts[i].first;
                a.g = parts[i].second;
                a.flags = kAfftInv | (i + 1 == ns ? kAfftFwd : 0u);
                a.inv_base = 0;
                a.fwd_base = k;
                if (i == 0) set_in_data(a);
                else set_in_work(a, w1);
                if (i + 1 == ns) a.row_limit = m;
                if (ns == 1) set_out_repairs(a, 0);
                else set_out_work(a, w1);
                if ((err = launch_afft_pass(a, Dk, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
            }
            for (int i = static_cast<int>(ns) - 2; i >= 0; i--) {
                AfftPassArgs a = base;
                a.b_lo = parts[i].first;
                a.g = parts[i].second;
                a.flags = kAfftFwd;
                a.fwd_base = k;
                a.row_limit = m;
                set_in_work(a, w1);
                if (i == 0) set_out_repairs(a, 0);
                else set_out_work(a, w1);
                if ((err = launch_afft_pass(a, Dk, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
            }
            continue;
        }
        // Erasure decoding of the repair points from the data points.
        const uint16_t *sc = enc_scale.as<uint16_t>();
        if (D <= kSinglePassLog) {
            AfftPassArgs a = base;
            a.b_lo = 0;
            a.g = D;
            a.flags = kAfftInv | kAfftDeriv | kAfftFwd;
            set_in_data(a);
            a.scale_in = sc;
            set_out_repairs(a, k);
            a.scale_out = sc;
            a.row_limit = n;
            if ((err = launch_afft_pass(a, D, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
            continue;
        }
        const auto parts = plan_bits(D, kPassLog);
        const uint32_t ns = static_cast<uint32_t>(parts.size());
        for (uint32_t i = 0; i < ns; i++) {
            AfftPassArgs a = base;
            a.b_lo = parts[i].first;
            a.g = parts[i].second;
            a.flags = kAfftInv;
            if (i == 0) {
                set_in_data(a);
                a.scale_in = sc;
            } else {
                set_in_work(a, w1);
            }
            if (i + 1 == ns) a.scale_out = dev->deriv_scale.as<uint16_t>();  // unit-factor derivative
            set_out_work(a, w1);
            if ((err = launch_afft_pass(a, D, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        }
        for (uint32_t i = 0; i + 1 < ns; i++) {
            AfftPassArgs a = base;
            a.b_lo = parts[i].first;
            a.g = parts[i].second;
            a.flags = kAfftDerivUnit;
            set_in_work(a, w1);
            if (i > 0) {
                a.aux_base = w2;
                a.aux_cw_stride = wcw;
            }
            set_out_work(a, w2);
            if ((err = launch_afft_pass(a, D, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        }
        for (int i = static_cast<int>(ns) - 1; i >= 0; i--) {
            AfftPassArgs a = base;
            a.b_lo = parts[i].first;
            a.g = parts[i].second;
            a.row_limit = n;
            if (i + 1 == static_cast<int>(ns)) {
                a.flags = kAfftDerivUnit | kAfftFwd;
                a.scale_der = dev->deriv_unscale.as<uint16_t>();
                set_in_work(a, w1);
                if (ns > 1) {  // lower parts' derivative contributions
                    a.aux_base = w2;
                    a.aux_cw_stride = wcw;
                }
            } else {
                a.flags = kAfftFwd;
                set_in_work(a, w2);
            }
            if (i == 0) {
                set_out_repairs(a, k);
                a.scale_out = sc;
            } else {
                set_out_work(a, w2);
            }
            if ((err = launch_afft_pass(a, D, cn, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
        }
    }
    return ok();
}

scrs_status AfftEngine::decode(const uint16_t *indices, const uint8_t *const *ptrs, const uint8_t *data_base,
                               const uint8_t *repair_base, uint8_t *out, hipStream_t s, const uint64_t *received64) {
    hipError_t err;
    // Known-point bitmap (through a pinned staging buffer reused per call):
    // the caller's validation bitmap when given (the same bits in 64-bit
    // words, little-endian), else built from the indices.  The missing data
    // symbols are k minus the known points below k.
    const uint64_t n_words64 = (uint64_t{n} + 63) / 64;
    if (received64) {
        std::memcpy(h_bits.data(), received64, n_words64 * 8);
        std::fill(reinterpret_cast<uint64_t *>(h_bits.data()) + n_words64,
                  reinterpret_cast<uint64_t *>(h_bits.data() + h_bits.size()), uint64_t{0});
    } else {
        std::fill(h_bits.begin(), h_bits.end(), 0u);
        for (uint32_t t = 0; t < k; t++) h_bits[indices[t] >> 5] |= 1u << (indices[t] & 31u);
    }
    uint32_t known_data = 0;
    for (uint32_t w = 0; w < k / 32; w++) known_data += static_cast<uint32_t>(__builtin_popcount(h_bits[w]));
    if (k % 32) known_data += static_cast<uint32_t>(__builtin_popcount(h_bits[k / 32] & ((1u << (k % 32)) - 1u)));
    const uint32_t missing = k - known_data;
    if (missing > 0 && missing <= std::min(afft_targeted_max(), direct.rmax)) {
        auto al16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
        const bool vec16 = !ptrs && L % 16 == 0 && al16(data_base) && al16(repair_base) && al16(out);
        direct.rows_hint = missing;  // the solve's row count is exactly this
        const scrs_status r = direct.enqueue(indices, ptrs, data_base, repair_base, out, 1, s, vec16);
        direct.rows_hint = 0;
        return r;
    }
    const uint64_t words = h_bits.size();
    if ((err = hipEventSynchronize(bits_copied)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    std::memcpy(h_bits_pinned.ptr, h_bits.data(), words * 4);
    if ((err = hipMemcpyAsync(bits.ptr, h_bits_pinned.ptr, words * 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (err = hipEventRecord(bits_copied, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    AfftRowsArgs ra;
    ra.known_bits = bits.as<uint32_t>();
    ra.data_base = ptrs ? nullptr : data_base;
    ra.repair_base = ptrs ? nullptr : repair_base;
    ra.out = out;
    ra.in_rows = in_rows.as<const uint8_t *>();
    ra.copy_rows = copy_rows.as<uint8_t *>();
    ra.out_rows = out_rows.as<uint8_t *>();
    ra.k = k;
    ra.n = n;
    ra.D = D;
    ra.L = L;
    if ((err = launch_afft_rows(ra, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    if (ptrs && (err = hipMemcpyAsync(loc_work.ptr, indices, k * 2, hipMemcpyHostToDevice, s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if (ptrs && (err = launch_afft_scatter_ptrs(loc_work.as<uint16_t>(), ptrs, k, in_rows.as<const uint8_t *>(),
                                                s)) != hipSuccess)
        return st(SCRS_ERR_DEVICE, err);
    if (ptrs && (err = hipStreamSynchronize(s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    AfftPassArgs base;
    base.node_fac = dev->node_fac.as<uint16_t>();
    base.deriv_tabs = dev->deriv_tabs.as<uint16_t>();
    base.L = L;
    // Row tables point into the data, repair and output buffers at multiples
    // of L (caller pointers from `ptrs` are not inspected).
    auto al4 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; };
    base.rows_aligned4 = !ptrs && L % 4 == 0 && al4(data_base) && al4(repair_base) && al4(out);
    const uint8_t *const *rows_in = in_rows.as<const uint8_t *const>();
    uint8_t *const *rows_copy = copy_rows.as<uint8_t *const>();
    uint8_t *const *rows_out = out_rows.as<uint8_t *const>();
    if (missing == 0) {
        // Every data symbol arrived: the first pass's copies are the output.
        AfftPassArgs a = base;
        a.b_lo = 0;
        a.g = D < kPassLog ? D : kPassLog;
        a.flags = kAfftInv;
        a.in_rows = rows_in;
        a.copy_rows = rows_copy;
        a.row_limit = k;
        err = launch_afft_pass(a, D, 1, s);
        return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
    }
    AfftLocatorArgs la;
    la.known_bits = bits.as<uint32_t>();
    la.log_wht = log_wht.as<uint32_t>();
    la.exp = dev->exp.as<uint16_t>();
    la.work = loc_work.as<uint32_t>();
    la.out = loc_out.as<uint16_t>();
    la.D = D;
    if ((err = launch_afft_locator(la, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    const uint16_t *sc = loc_out.as<uint16_t>();
    uint8_t *w1 = work.as<uint8_t>(), *w2 = work2.as<uint8_t>();
    auto in_work = [&](AfftPassArgs &a, const uint8_t *w) {
        a.in_base = w;
        a.in_limit = N;
    };
    auto out_work = [&](AfftPassArgs &a, uint8_t *w) {
        a.out_base = w;
        a.out_first = 0;
        a.out_limit = N;
    };
    if (D <= kSinglePassLog) {
        AfftPassArgs a = base;
        a.b_lo = 0;
        a.g = D;
        a.flags = kAfftInv | kAfftDeriv | kAfftFwd;
        a.in_rows = rows_in;
        a.copy_rows = rows_copy;
        a.scale_in = sc;
        a.out_rows = rows_out;
        a.scale_out = sc;
        err = launch_afft_pass(a, D, 1, s);
        return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
    }
    const auto parts = plan_bits(D, kPassLog);
    const uint32_t ns = static_cast<uint32_t>(parts.size());
    for (uint32_t i = 0; i < ns; i++) {
        AfftPassArgs a = base;
        a.b_lo = parts[i].first;
        a.g = parts[i].second;
        a.flags = kAfftInv;
        if (i == 0) {
            a.in_rows = rows_in;
            a.copy_rows = rows_copy;
            a.scale_in = sc;
        } else {
            in_work(a, w1);
        }
        if (i + 1 == ns) a.scale_out = dev->deriv_scale.as<uint16_t>();  // unit-factor derivative
        out_work(a, w1);
        if ((err = launch_afft_pass(a, D, 1, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    }
    for (uint32_t i = 0; i + 1 < ns; i++) {
        AfftPassArgs a = base;
        a.b_lo = parts[i].first;
        a.g = parts[i].second;
        a.flags = kAfftDerivUnit;
        in_work(a, w1);
        if (i > 0) a.aux_base = w2;
        out_work(a, w2);
        if ((err = launch_afft_pass(a, D, 1, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    }
    for (int i = static_cast<int>(ns) - 1; i >= 0; i--) {
        AfftPassArgs a = base;
        a.b_lo = parts[i].first;
        a.g = parts[i].second;
        a.row_limit = k;
        if (i + 1 == static_cast<int>(ns)) {
            a.flags = kAfftDerivUnit | kAfftFwd;
            a.scale_der = dev->deriv_unscale.as<uint16_t>();
            in_work(a, w1);
            if (ns > 1) a.aux_base = w2;  // lower parts' derivative contributions
        } else {
            a.flags = kAfftFwd;
            in_work(a, w2);
        }
        if (i == 0) {
            a.out_rows = rows_out;
            a.scale_out = sc;
        } else {
            out_work(a, w2);
        }
        if ((err = launch_afft_pass(a, D, 1, s)) != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    }
    return ok();
}

}  // namespace scrs
