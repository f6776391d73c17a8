// batch_engine.cpp - planner, resident caches, graph replay and autotuner
// of the fused uneven batch, and its C ABI.
#include "batch_engine.hpp"

#include <algorithm>
#include <cstring>

#include "engine.hpp"
#include "gf_device.hpp"
#include "matrices.hpp"
#include "status.hpp"

using namespace scrs;

struct scrs_batch_engine {
    BatchEngine eng;
};

namespace scrs {

namespace {

constexpr uint32_t kSolveGrid = 256;
constexpr uint16_t kDataSlot = 0xFFFF;
constexpr uint64_t kTuneMinJobs = 32;
#ifndef SCRS_FILL_UNITS
#define SCRS_FILL_UNITS 512
#endif
#ifndef SCRS_MIN_SEG_TERMS
#define SCRS_MIN_SEG_TERMS 8
#endif
constexpr uint64_t kFillUnits = SCRS_FILL_UNITS;     // units a batch should offer the device
constexpr uint32_t kMinSegTerms = SCRS_MIN_SEG_TERMS;  // shortest term segment of a split range

uint64_t align16(uint64_t x) { return (x + 15) & ~uint64_t{15}; }
bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
bool aligned4(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; }
uint32_t log2_floor(uint64_t x) {
    uint32_t d = 0;
    while (x > 1) {
        x >>= 1;
        d++;
    }
    return d;
}
int coord_slot(int32_t engine) {
    return engine == SCRS_ENGINE_STANDARD_CAUCHY ? 0 : (engine == SCRS_ENGINE_GOOD_CAUCHY ? 1 : 2);
}

}  // namespace

struct BatchEngine::Plan {
    std::vector<BatchRange> ranges;
    std::vector<BatchSolve> solves;
    std::vector<uint64_t> ptrs;   // pointer arrays, device addresses as integers
    std::vector<uint16_t> u16s;   // received / slot / missing arrays of the solves
    // Offsets of each range's and solve's arrays, patched into device
    // addresses at upload.
    struct RangeRef {
        uint64_t src = 0, dst = 0, copy = 0;
        uint64_t part = 0;  // byte offset of a term-split range's partial rows
        bool has_copy = false;
    };
    struct SolveRef {
        uint64_t idx = 0, slot = 0, missing = 0, work = 0, fused = 0;
    };
    std::vector<RangeRef> rref;
    std::vector<SolveRef> sref;
    std::vector<BatchReduce> red;    // term-split ranges (dst and partials patched at upload)
    std::vector<uint32_t> red_range; // the range of each
    uint64_t part_bytes = 0, red_items = 0;
    uint64_t work_u32 = 0, fused_u16 = 0;
    uint32_t units = 0;
};

BatchEngine::~BatchEngine() {
    DeviceGuard g(device_);
    for (Slot &s : slots_) {
        if (s.done) {
            (void)hipEventSynchronize(s.done);
            (void)hipEventDestroy(s.done);
        }
        for (auto &kv : s.graphs) (void)hipGraphExecDestroy(kv.second.second);
    }
    if (capture_) (void)hipStreamDestroy(capture_);
    if (last_solved_) (void)hipEventDestroy(last_solved_);
    for (Sample &x : samples_) {
        (void)hipEventSynchronize(x.e1);
        event_pool_.push_back(x.e0);
        event_pool_.push_back(x.e1);
    }
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    for (CacheEntry *e : lru_) delete e;
    for (CacheEntry *e : retired_) delete e;
}

scrs_status BatchEngine::init(const scrs_batch_config &cfg) {
    cfg_ = cfg;
    if (cfg_.cache_bytes == 0) cfg_.cache_bytes = 1ull << 30;
    if (cfg_.tile_bytes != 0 && cfg_.tile_bytes != 1024 && cfg_.tile_bytes != 2048 && cfg_.tile_bytes != 4096)
        return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (cfg_.cw_per_unit != 0 && cfg_.cw_per_unit != 1 && cfg_.cw_per_unit != 2 && cfg_.cw_per_unit != 4 &&
        cfg_.cw_per_unit != 8)
        return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    (void)hipGetDevice(&device_);
    hipError_t err = hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking);
    // Coordinates of every wire index for each Cauchy engine, at capacity.
    const int32_t engines[3] = {SCRS_ENGINE_STANDARD_CAUCHY, SCRS_ENGINE_GOOD_CAUCHY, SCRS_ENGINE_TOWER};
    for (int i = 0; i < 3 && err == hipSuccess; i++) {
        const scrs_profile p{engines[i], 1, engine_capacity(engines[i]) - 1, 2};
        const std::vector<uint16_t> c = cauchy_coordinates(p);
        if ((err = coords_[i].ensure(c.size() * 2)) == hipSuccess)
            err = hipMemcpy(coords_[i].ptr, c.data(), c.size() * 2, hipMemcpyHostToDevice);
    }
    uint8_t t[768];
    gf8_exp_log(t, t + 512);
    if (err == hipSuccess && (err = gf8_.ensure(768)) == hipSuccess)
        err = hipMemcpy(gf8_.ptr, t, 768, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = gf16_log_exp_device(&gf16_log_exp_);
    if (err == hipSuccess) err = arena_.init(cfg_.cache_bytes);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&last_solved_, hipEventDisableTiming);
    for (Slot &s : slots_) {
        if (err == hipSuccess) err = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (err == hipSuccess) err = s.counter.ensure(64);
        if (err == hipSuccess) err = hipMemset(s.counter.ptr, 0, 64);  // launches count on from claims = 0
    }
    return err == hipSuccess ? ok() : st(SCRS_ERR_DEVICE, err);
}

scrs_status BatchEngine::validate(const scrs_job &j) const {
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(j.profile.engine, j.profile.k, j.profile.m, j.profile.symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    if (p.engine == SCRS_ENGINE_AFFT) return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    if (j.op == SCRS_JOB_ENCODE) {
        if (!j.src || !j.dst) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
        return ok();
    }
    if (j.op != SCRS_JOB_DECODE) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!j.indices || !j.payloads || !j.dst) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    const uint64_t k = p.k, n = p.k + p.m;
    std::vector<uint64_t> &seen = seen_;
    seen.assign((n + 63) / 64, 0);
    for (uint64_t t = 0; t < k; t++) {
        const uint64_t i = j.indices[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
        seen[i >> 6] |= 1ull << (i & 63);
        if (!j.payloads[t]) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    }
    return ok();
}

hipError_t DeviceArena::init(uint64_t bytes) {
    cap_ = bytes & ~uint64_t{255};
    hipError_t err = buf_.ensure(cap_);
    if (err == hipSuccess) free_[0] = cap_;
    return err;
}

bool DeviceArena::alloc(uint64_t bytes, uint64_t *off) {
    bytes = (bytes + 255) & ~uint64_t{255};
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        const uint64_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
    }
    return false;
}

void DeviceArena::release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + 255) & ~uint64_t{255};
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
        bytes += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += bytes;
            return;
        }
    }
    free_[off] = bytes;
}

void BatchEngine::touch(CacheEntry *e) {
    e->last_use = seq_;
    lru_.splice(lru_.end(), lru_, e->lru);
}

void BatchEngine::retire_completed() {
    // Batches may finish out of order on several streams: every batch up to
    // the oldest one still running is complete (a slot is reused only once
    // its previous batch has finished, so older batches are done).
    uint64_t oldest_running = 0, newest = 0;
    for (Slot &s : slots_) {
        if (!s.seq) continue;
        newest = std::max(newest, s.seq);
        if (hipEventQuery(s.done) != hipSuccess && (!oldest_running || s.seq < oldest_running))
            oldest_running = s.seq;
    }
    completed_ = std::max(completed_, oldest_running ? oldest_running - 1 : newest);
    auto it = std::remove_if(retired_.begin(), retired_.end(), [&](CacheEntry *e) {
        if (e->last_use > completed_) return false;
        arena_.release(e->off, e->bytes);
        delete e;
        return true;
    });
    retired_.erase(it, retired_.end());
}

// Free arena space: evict the least recently used entry the batch being
// planned does not read, or wait for earlier batches to release evicted
// entries.  False when nothing can be freed.
bool BatchEngine::make_room() {
    for (auto it = lru_.begin(); it != lru_.end(); ++it) {
        CacheEntry *e = *it;
        if (e->last_use == seq_) continue;
        if (e->recipe) recipe_cache_.erase(e->recipe_key);
        else enc_cache_.erase(e->enc_key);
        cache_bytes_ -= e->bytes;
        lru_.erase(it);
        retired_.push_back(e);
        retire_completed();
        return true;
    }
    if (retired_.empty()) return false;
    for (Slot &s : slots_) (void)hipEventSynchronize(s.done);
    retire_completed();
    return true;
}

CacheEntry *BatchEngine::new_entry(uint64_t bytes, scrs_status *err) {
    if (bytes > arena_.capacity()) {
        *err = st(SCRS_ERR_CAPACITY, arena_.capacity());
        return nullptr;
    }
    uint64_t off = 0;
    while (!arena_.alloc(bytes, &off)) {
        if (!make_room()) {
            *err = st(SCRS_ERR_CAPACITY, arena_.capacity());
            return nullptr;
        }
    }
    auto *e = new CacheEntry();
    e->off = off;
    e->bytes = bytes;
    e->last_use = seq_;
    cache_bytes_ += bytes;
    e->lru = lru_.insert(lru_.end(), e);
    return e;
}

// Encode product tables of one geometry; a new geometry's tables are
// computed on the device by a table job of the batch being planned.
CacheEntry *BatchEngine::encoder_tables(const scrs_profile &p, Plan &plan, scrs_status *err) {
    const auto key = std::make_tuple(p.engine, p.k, p.m);
    auto it = enc_cache_.find(key);
    if (it != enc_cache_.end()) {
        st_.table_hits++;
        touch(it->second);
        return it->second;
    }
    st_.table_misses++;
    const bool gf16 = !engine_is_gf256(p.engine);
    const uint32_t groups = static_cast<uint32_t>((p.m + 15) / 16);
    CacheEntry *e = new_entry(p.k * groups * table_group_bytes(gf16), err);
    if (!e) return nullptr;
    e->rows = static_cast<uint32_t>(p.m);
    e->groups = groups;
    e->enc_key = key;
    enc_cache_[key] = e;
    BatchSolve sv;
    sv.coords = coords_[coord_slot(p.engine)].as<uint16_t>();
    sv.tables = entry_ptr(e);
    sv.k = static_cast<uint32_t>(p.k);
    sv.r = static_cast<uint32_t>(p.m);
    sv.groups = groups;
    sv.gf16 = gf16;
    sv.encode = 1;
    plan.solves.push_back(sv);
    plan.sref.emplace_back();
    return e;
}

scrs_status BatchEngine::plan(const scrs_job *jobs, uint64_t count, uint32_t tile, uint32_t cw, Plan &p) {
    p = Plan();
    // Encode jobs of one geometry, length and alignment class share a range
    // (class 2: rows 16-byte aligned; 1: destination rows 4-byte aligned):
    // sorted by that key, each run of equal keys is one range, in job order.
    enc_keys_.clear();
    decodes_.clear();
    uint64_t nptrs = 0;
    for (uint64_t i = 0; i < count; i++) {
        const scrs_job &j = jobs[i];
        const uint64_t L = j.profile.symbol_len;
        if (j.op == SCRS_JOB_ENCODE) {
            const int v = (L % 16 == 0 && aligned16(j.src) && aligned16(j.dst)) ? 2
                          : (L % 4 == 0 && aligned4(j.dst))                     ? 1
                                                                                : 0;
            enc_keys_.push_back({j.profile.engine, v, j.profile.k, j.profile.m, L, i});
            nptrs += j.profile.k + j.profile.m;
        } else {
            decodes_.push_back(i);
            nptrs += 3 * j.profile.k;
        }
    }
    std::sort(enc_keys_.begin(), enc_keys_.end());
    p.ptrs.reserve(nptrs);
    uint64_t units = 0;
    auto add_range = [&](BatchRange r, const Plan::RangeRef &ref) -> bool {
        r.segs = 1;
        r.seg_terms = r.T;
        r.unit_base = static_cast<uint32_t>(units);
        units += static_cast<uint64_t>((r.cw_count + r.ncw - 1) / r.ncw) * r.groups * r.tiles;
        p.ranges.push_back(r);
        p.rref.push_back(ref);
        return units <= 0xffffffffull;
    };
    auto same_range = [](const EncKey &a, const EncKey &b) {
        return a.engine == b.engine && a.k == b.k && a.m == b.m && a.L == b.L && a.cls == b.cls;
    };
    for (size_t g0 = 0; g0 < enc_keys_.size();) {
        size_t g1 = g0 + 1;
        while (g1 < enc_keys_.size() && same_range(enc_keys_[g0], enc_keys_[g1])) g1++;
        const scrs_job &j0 = jobs[enc_keys_[g0].job];
        const scrs_profile &pr = j0.profile;
        scrs_status es = ok();
        CacheEntry *e = encoder_tables(pr, p, &es);
        if (!e) return es;
        const uint64_t k = pr.k, m = pr.m, L = pr.symbol_len;
        BatchRange r;
        r.L = L;
        r.tables = entry_ptr(e);
        r.T = static_cast<uint32_t>(k);
        r.rows = static_cast<uint32_t>(m);
        r.groups = e->groups;
        r.tiles = static_cast<uint32_t>((L + tile - 1) / tile);
        r.cw_count = static_cast<uint32_t>(g1 - g0);
        r.ncw = std::min<uint32_t>(cw, r.cw_count);
        const int cls = enc_keys_[g0].cls;
        r.flags = (cls == 2 ? kRangeVec16 | kRangeDst4 : cls == 1 ? kRangeDst4 : 0u) |
                  (engine_is_gf256(pr.engine) ? 0u : kRangeGf16);
        Plan::RangeRef ref;
        ref.src = p.ptrs.size();
        for (size_t g = g0; g < g1; g++) {
            const uint8_t *src = jobs[enc_keys_[g].job].src;
            for (uint64_t t = 0; t < k; t++) p.ptrs.push_back(reinterpret_cast<uint64_t>(src + t * L));
        }
        ref.dst = p.ptrs.size();
        for (size_t g = g0; g < g1; g++) {
            uint8_t *dst = jobs[enc_keys_[g].job].dst;
            for (uint64_t t = 0; t < m; t++) p.ptrs.push_back(reinterpret_cast<uint64_t>(dst + t * L));
        }
        if (!add_range(r, ref)) return st(SCRS_ERR_CAPACITY, 0xffffffffull);
        g0 = g1;
    }
    for (uint64_t ji : decodes_) {
        const scrs_job &j = jobs[ji];
        const scrs_profile &pr = j.profile;
        const uint64_t k = pr.k, m = pr.m, n = k + m, L = pr.symbol_len;
        const bool gf16 = !engine_is_gf256(pr.engine);
        // Terms in ascending index order (the recipe depends on the set
        // only): mark the received indices, then read the bitmap in order.
        // pos_ is written before it is read for every marked index.
        std::vector<uint64_t> &key = key_;
        key.assign((n + 63) / 64 + 3, 0);
        if (pos_.size() < n) pos_.resize(n);
        for (uint64_t t = 0; t < k; t++) {
            const uint16_t x = j.indices[t];
            key[x >> 6] |= 1ull << (x & 63);
            pos_[x] = static_cast<uint32_t>(t);
        }
        order_.clear();
        missing_.clear();
        bool vec = L % 16 == 0 && aligned16(j.dst);
        for (uint64_t w = 0; w < (n + 63) / 64; w++)
            for (uint64_t bits = key[w]; bits; bits &= bits - 1) {
                const uint16_t x = static_cast<uint16_t>(64 * w + __builtin_ctzll(bits));
                order_.push_back(x);
                vec = vec && aligned16(j.payloads[pos_[x]]);
            }
        for (uint64_t i = 0; i < k; i++)
            if (!((key[i >> 6] >> (i & 63)) & 1)) missing_.push_back(static_cast<uint16_t>(i));
        const std::vector<uint16_t> &missing = missing_;
        const uint32_t r = static_cast<uint32_t>(missing.size());
        key[key.size() - 3] = static_cast<uint64_t>(pr.engine);
        key[key.size() - 2] = k;
        key[key.size() - 1] = m;
        BatchRange br;
        br.L = L;
        br.T = static_cast<uint32_t>(k);
        br.rows = r;
        br.groups = r ? (r + 15) / 16 : 1;
        br.tiles = static_cast<uint32_t>((L + tile - 1) / tile);
        br.cw_count = 1;
        br.ncw = 1;
        const bool dst4 = L % 4 == 0 && aligned4(j.dst);
        br.flags = kRangeDecode | (vec ? kRangeVec16 : 0u) | (dst4 ? kRangeDst4 : 0u) | (gf16 ? kRangeGf16 : 0u);
        if (r) {
            CacheEntry *e;
            auto it = recipe_cache_.find(key);
            if (it != recipe_cache_.end()) {
                st_.recipe_hits++;
                e = it->second;
                touch(e);
            } else {
                st_.recipe_misses++;
                scrs_status es = ok();
                e = new_entry(k * br.groups * table_group_bytes(gf16), &es);
                if (!e) return es;
                e->rows = r;
                e->groups = br.groups;
                e->recipe = true;
                e->recipe_key = key;
                recipe_cache_[key] = e;
                BatchSolve sv;
                sv.tables = entry_ptr(e);
                sv.k = static_cast<uint32_t>(k);
                sv.r = r;
                sv.groups = br.groups;
                sv.gf16 = gf16;
                sv.coords = coords_[coord_slot(pr.engine)].as<uint16_t>();
                Plan::SolveRef sr;
                sr.idx = p.u16s.size();
                for (uint16_t x : order_) p.u16s.push_back(x);
                sr.slot = p.u16s.size();
                uint16_t rank = 0;
                for (uint16_t x : order_) p.u16s.push_back(x < k ? kDataSlot : rank++);
                sr.missing = p.u16s.size();
                for (uint16_t v : missing) p.u16s.push_back(v);
                sr.work = p.work_u32;
                p.work_u32 += 4ull * r + k;
                sr.fused = p.fused_u16;
                p.fused_u16 += static_cast<uint64_t>(r) * k;
                p.solves.push_back(sv);
                p.sref.push_back(sr);
            }
            br.tables = entry_ptr(e);
        }
        Plan::RangeRef ref;
        ref.src = p.ptrs.size();
        for (uint16_t x : order_) p.ptrs.push_back(reinterpret_cast<uint64_t>(j.payloads[pos_[x]]));
        ref.dst = p.ptrs.size();
        for (uint16_t v : missing) p.ptrs.push_back(reinterpret_cast<uint64_t>(j.dst + v * L));
        ref.copy = p.ptrs.size();
        ref.has_copy = true;
        for (uint16_t x : order_) p.ptrs.push_back(x < k ? reinterpret_cast<uint64_t>(j.dst + x * L) : 0);
        if (!add_range(br, ref)) return st(SCRS_ERR_CAPACITY, 0xffffffffull);
    }
    split_terms(p, units);
    p.units = static_cast<uint32_t>(units);
    return ok();
}

// A batch too small to give every CU work splits the term loop of its
// ranges with dword-aligned destination rows: each segment's unit stores its
// partial rows, and a reduce launch XORs them into the destinations.  A
// unit's term loop waits out one memory latency per term, so the latency of
// a lone codeword drops with the segment length (kMinSegTerms at the
// shortest); a full batch is left as it is.
void BatchEngine::split_terms(Plan &p, uint64_t &units) {
    if (units >= kFillUnits || units == 0) return;
    const uint64_t factor = (kFillUnits + units - 1) / units;
    bool any = false;
    for (size_t i = 0; i < p.ranges.size(); i++) {
        BatchRange &r = p.ranges[i];
        if (!(r.flags & kRangeDst4) || r.rows == 0 || r.T < 2 * kMinSegTerms) continue;
        const uint32_t max_segs = r.T / kMinSegTerms;
        const uint32_t want = static_cast<uint32_t>(std::min<uint64_t>(max_segs, factor));
        if (want < 2) continue;
        r.seg_terms = (r.T + want - 1) / want;
        r.segs = (r.T + r.seg_terms - 1) / r.seg_terms;
        if (r.segs < 2) {
            r.segs = 1;
            r.seg_terms = r.T;
            continue;
        }
        any = true;
        r.part_stride = align16(r.L);
        p.rref[i].part = p.part_bytes;
        const uint64_t out_rows = static_cast<uint64_t>(r.cw_count) * r.rows;
        p.part_bytes += out_rows * r.segs * r.part_stride;
        BatchReduce rd;
        rd.L = r.L;
        rd.stride = r.part_stride;
        rd.item_base = p.red_items;
        rd.rows = r.rows;
        rd.segs = r.segs;
        p.red.push_back(rd);
        p.red_range.push_back(static_cast<uint32_t>(i));
        p.red_items += out_rows * (r.L / 4);
    }
    if (!any) return;
    units = 0;
    for (BatchRange &r : p.ranges) {
        r.unit_base = static_cast<uint32_t>(units);
        units += static_cast<uint64_t>((r.cw_count + r.ncw - 1) / r.ncw) * r.groups * r.tiles * r.segs;
    }
}

hipError_t BatchEngine::run_graph(Slot &slot, uint32_t tile, bool reduce, hipStream_t s) {
    const uint32_t key = tile * 2 + (reduce ? 1 : 0);
    auto it = slot.graphs.find(key);
    if (it == slot.graphs.end() || it->second.first != slot.generation) {
        if (it != slot.graphs.end()) {
            (void)hipGraphExecDestroy(it->second.second);
            slot.graphs.erase(it);
        }
        const BatchHeader *hdr = slot.dev.as<BatchHeader>();
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipError_t err = hipStreamBeginCapture(capture_, hipStreamCaptureModeThreadLocal);
        if (err != hipSuccess) return err;
        // The unit counter needs no reset between launches: no memset node.
        // The reduce launch is sized for the largest reduce (grid-stride).
        hipError_t e3 = launch_batch_rows(hdr, slot.counter.as<uint32_t>(), tile, batch_rows_grid(tile), capture_);
        if (e3 == hipSuccess && reduce) e3 = launch_batch_reduce(hdr, kBatchReduceGrid, capture_);
        err = hipStreamEndCapture(capture_, &graph);
        if (err == hipSuccess) err = e3;
        if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (err != hipSuccess) return err;
        it = slot.graphs.emplace(key, std::make_pair(slot.generation, exec)).first;
    }
    return hipGraphLaunch(it->second.second, s);
}

hipError_t BatchEngine::launch(Slot &slot, const Plan &p, uint32_t tile, hipStream_t s) {
    const uint64_t off_r = align16(sizeof(BatchHeader));
    const uint64_t off_s = align16(off_r + p.ranges.size() * sizeof(BatchRange));
    const uint64_t off_p = align16(off_s + p.solves.size() * sizeof(BatchSolve));
    const uint64_t off_u = align16(off_p + p.ptrs.size() * 8);
    const uint64_t off_red = align16(off_u + p.u16s.size() * 2);
    const uint64_t total = align16(off_red + p.red.size() * sizeof(BatchReduce));
    const uint64_t sc_work = 0, sc_fused = align16(p.work_u32 * 4);
    const uint64_t sc_part = align16(sc_fused + p.fused_u16 * 2);
    const uint64_t sc_total = std::max<uint64_t>(16, align16(sc_part + p.part_bytes));
    hipError_t err;
    // The slot's previous batch may still read its staging and arrays.
    if ((err = hipEventSynchronize(slot.done)) != hipSuccess) return err;
    if (total > slot.dev.size || sc_total > slot.scratch.size) {
        const uint64_t want = std::max<uint64_t>(total * 2, 1 << 20), want_sc = std::max<uint64_t>(sc_total * 2, 1 << 20);
        if (total > slot.dev.size && (err = slot.dev.ensure(want)) != hipSuccess) return err;
        if (sc_total > slot.scratch.size && (err = slot.scratch.ensure(want_sc)) != hipSuccess) return err;
        slot.generation++;
    }
    if ((err = slot.host.ensure(slot.dev.size)) != hipSuccess) return err;
    uint8_t *h = slot.host.as<uint8_t>();
    uint8_t *d = slot.dev.as<uint8_t>();
    uint8_t *sc = slot.scratch.as<uint8_t>();
    BatchHeader hdr;
    hdr.nranges = static_cast<uint32_t>(p.ranges.size());
    hdr.nunits = p.units;
    hdr.nsolve = static_cast<uint32_t>(p.solves.size());
    hdr.ranges = reinterpret_cast<const BatchRange *>(d + off_r);
    hdr.solves = reinterpret_cast<const BatchSolve *>(d + off_s);
    hdr.nred = static_cast<uint32_t>(p.red.size());
    hdr.red_items = p.red_items;
    hdr.red = reinterpret_cast<const BatchReduce *>(d + off_red);
    hdr.claim_base = slot.claims;
    std::memcpy(h, &hdr, sizeof(hdr));
    auto *hr = reinterpret_cast<BatchRange *>(h + off_r);
    for (size_t i = 0; i < p.ranges.size(); i++) {
        BatchRange r = p.ranges[i];
        const Plan::RangeRef &ref = p.rref[i];
        r.src = reinterpret_cast<const uint8_t *const *>(d + off_p + ref.src * 8);
        r.dst = reinterpret_cast<uint8_t *const *>(d + off_p + ref.dst * 8);
        r.copy = ref.has_copy ? reinterpret_cast<uint8_t *const *>(d + off_p + ref.copy * 8) : nullptr;
        r.partials = r.segs > 1 ? sc + sc_part + ref.part : nullptr;
        hr[i] = r;
    }
    auto *hred = reinterpret_cast<BatchReduce *>(h + off_red);
    for (size_t i = 0; i < p.red.size(); i++) {
        BatchReduce rd = p.red[i];
        rd.dst = hr[p.red_range[i]].dst;
        rd.partials = hr[p.red_range[i]].partials;
        hred[i] = rd;
    }
    auto *hs = reinterpret_cast<BatchSolve *>(h + off_s);
    for (size_t i = 0; i < p.solves.size(); i++) {
        BatchSolve v = p.solves[i];
        const Plan::SolveRef &ref = p.sref[i];
        if (v.encode) {
            hs[i] = v;
            continue;
        }
        v.idx = reinterpret_cast<const uint16_t *>(d + off_u + ref.idx * 2);
        v.slot = reinterpret_cast<const uint16_t *>(d + off_u + ref.slot * 2);
        v.missing = reinterpret_cast<const uint16_t *>(d + off_u + ref.missing * 2);
        v.work = reinterpret_cast<uint32_t *>(sc + sc_work + ref.work * 4);
        v.fused = reinterpret_cast<uint16_t *>(sc + sc_fused + ref.fused * 2);
        hs[i] = v;
    }
    if (!p.ptrs.empty()) std::memcpy(h + off_p, p.ptrs.data(), p.ptrs.size() * 8);
    if (!p.u16s.empty()) std::memcpy(h + off_u, p.u16s.data(), p.u16s.size() * 2);
    // Tables and recipes written by an earlier batch's solve may be read
    // here, and it may have run on another stream.  Only the latest solve is
    // waited for: the multiply-XORs may still run beside this batch.
    if (solves_recorded_ && (err = hipStreamWaitEvent(s, last_solved_, 0)) != hipSuccess) return err;
    if ((err = hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s)) != hipSuccess) return err;
    const BatchHeader *dh = slot.dev.as<BatchHeader>();
    // Nothing to solve (every table and recipe resident): the multiply-XOR
    // runs alone, and the latest solve stays the one to wait for.
    if (!p.solves.empty()) {
        if ((err = launch_batch_solve(dh, gf8_.as<uint8_t>(), gf16_log_exp_, kSolveGrid, s)) != hipSuccess)
            return err;
        if ((err = hipEventRecord(last_solved_, s)) != hipSuccess) return err;
        solves_recorded_ = true;
    }
    const bool reduce = !p.red.empty();
    if (cfg_.use_graph) {
        err = run_graph(slot, tile, reduce, s);
        if (err != hipSuccess) return err;
        slot.claims += p.units;
    } else {
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(batch_rows_grid(tile), p.units));
        if ((err = launch_batch_rows(dh, slot.counter.as<uint32_t>(), tile, grid, s)) != hipSuccess) return err;
        slot.claims += p.units;  // the launch is enqueued: it will draw nunits times
        if (reduce && (err = launch_batch_reduce(dh, batch_reduce_grid(p.red_items), s)) != hipSuccess) return err;
    }
    slot.seq = seq_;
    return hipEventRecord(slot.done, s);
}

scrs_status BatchEngine::submit(const scrs_job *jobs, uint64_t count, hipStream_t s, uint64_t *failed) {
    std::lock_guard<std::mutex> lock(mu_);
    for (uint64_t i = 0; i < count; i++) {
        scrs_status v = validate(jobs[i]);
        if (v.code != SCRS_OK) {
            if (failed) *failed = i;
            return v;
        }
    }
    if (count == 0) return ok();
    DeviceGuard g(device_);
    retire_completed();
    // Autotuner signature: job-count and mean-length buckets, field mix and
    // whether any encode geometry repeats (codewords per unit matters).
    uint64_t lsum = 0, enc = 0, gf16 = 0;
    sig_keys_.clear();
    for (uint64_t i = 0; i < count; i++) {
        const scrs_job &j = jobs[i];
        lsum += j.profile.symbol_len;
        gf16 += engine_is_gf256(j.profile.engine) ? 0 : 1;
        if (j.op == SCRS_JOB_ENCODE) {
            enc++;
            sig_keys_.emplace_back(j.profile.engine, j.profile.k, j.profile.m, j.profile.symbol_len);
        }
    }
    std::sort(sig_keys_.begin(), sig_keys_.end());
    const uint32_t max_rep =
        std::adjacent_find(sig_keys_.begin(), sig_keys_.end()) != sig_keys_.end() ? 2u : (enc ? 1u : 0u);
    const Sig sig = {log2_floor(lsum / count), gf16 ? 1u : 0u, static_cast<uint32_t>(4 * enc / (count + 1)),
                     max_rep > 1 ? 1u : 0u};
    uint32_t tile = cfg_.tile_bytes, cw = cfg_.cw_per_unit;
    const bool tune = (tile == 0 || cw == 0);
    harvest_samples();
    auto found = tuned_.find(sig);
    if (tune && found != tuned_.end()) {
        if (!tile) tile = found->second.first;
        if (!cw) cw = found->second.second;
    }
    // A signature not tuned yet runs each large batch on its next candidate
    // configuration and times it on the device; small batches are not
    // representative and run with the defaults until then.  Nothing waits:
    // the timings are collected by later submissions.
    Sample sample;
    bool timed = false;
    if (tune && found == tuned_.end() && count >= kTuneMinJobs) {
        Tuning &tg = tuning_[sig];
        if (tg.cands.empty()) {
            const std::vector<uint32_t> tiles =
                cfg_.tile_bytes ? std::vector<uint32_t>{cfg_.tile_bytes} : std::vector<uint32_t>{1024, 2048, 4096};
            const std::vector<uint32_t> cws = cfg_.cw_per_unit ? std::vector<uint32_t>{cfg_.cw_per_unit}
                                              : (max_rep > 1 ? std::vector<uint32_t>{1, 2, 4, 8}
                                                             : std::vector<uint32_t>{1});
            for (uint32_t tt : tiles)
                for (uint32_t cc : cws) tg.cands.push_back({tt, cc, -1.0});
        }
        const uint32_t ci = tg.next % static_cast<uint32_t>(tg.cands.size());
        tg.next++;
        tile = tg.cands[ci].tile;
        cw = tg.cands[ci].cw;
        sample.sig = sig;
        sample.cand = ci;
        for (uint64_t i = 0; i < count; i++) {
            const scrs_job &j = jobs[i];
            const uint64_t k = j.profile.k, m = j.profile.m, L = j.profile.symbol_len;
            sample.bytes += j.op == SCRS_JOB_ENCODE ? (k + m) * L : 2 * k * L;
        }
        for (hipEvent_t *e : {&sample.e0, &sample.e1}) {
            if (!event_pool_.empty()) {
                *e = event_pool_.back();
                event_pool_.pop_back();
            } else if (hipEventCreate(e) != hipSuccess) {
                *e = nullptr;
            }
        }
        timed = sample.e0 && sample.e1 && hipEventRecord(sample.e0, s) == hipSuccess;
    }
    if (!tile) tile = 4096;
    if (!cw) cw = 1;
    seq_++;
    Plan p;
    scrs_status r = plan(jobs, count, tile, cw, p);
    if (r.code != SCRS_OK) return r;
    Slot &slot = slots_[next_slot_];
    next_slot_ = (next_slot_ + 1) % kSlots;
    hipError_t err = launch(slot, p, tile, s);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    if (timed && hipEventRecord(sample.e1, s) == hipSuccess) {
        samples_.push_back(std::move(sample));
    } else {
        if (sample.e0) event_pool_.push_back(sample.e0);
        if (sample.e1) event_pool_.push_back(sample.e1);
    }
    st_.tile_bytes = tile;
    st_.cw_per_unit = cw;
    st_.batches++;
    st_.jobs += count;
    return ok();
}

// Timings of finished tuning runs; a signature whose candidates have all
// run keeps the one with the least device time per byte.
void BatchEngine::harvest_samples() {
    while (!samples_.empty() && hipEventQuery(samples_.front().e1) == hipSuccess) {
        Sample x = std::move(samples_.front());
        samples_.pop_front();
        float ms = 0;
        if (hipEventElapsedTime(&ms, x.e0, x.e1) == hipSuccess && x.bytes) {
            auto it = tuning_.find(x.sig);
            if (it != tuning_.end() && !tuned_.count(x.sig)) {
                Candidate &c = it->second.cands[x.cand];
                const double v = ms * 1e6 / static_cast<double>(x.bytes);
                if (c.best_ns_per_byte < 0 || v < c.best_ns_per_byte) c.best_ns_per_byte = v;
                c.runs++;
                const Candidate *best = nullptr;
                bool all = true;
                for (const Candidate &k : it->second.cands) {
                    if (k.runs < kTuneRuns) all = false;
                    if (k.best_ns_per_byte >= 0 && (!best || k.best_ns_per_byte < best->best_ns_per_byte)) best = &k;
                }
                if (all && best) {
                    tuned_[x.sig] = {best->tile, best->cw};
                    st_.tuned_signatures = static_cast<uint32_t>(tuned_.size());
                    tuning_.erase(it);
                }
            }
        }
        event_pool_.push_back(x.e0);
        event_pool_.push_back(x.e1);
    }
}

scrs_batch_stats BatchEngine::stats() const {
    std::lock_guard<std::mutex> lock(mu_);
    scrs_batch_stats s = st_;
    s.cache_bytes = cache_bytes_;
    return s;
}

}  // namespace scrs

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" scrs_status scrs_batch_engine_create(const scrs_batch_config *cfg, scrs_batch_engine **out) {
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    scrs_batch_config c{};
    if (cfg) c = *cfg;
    auto *e = new scrs_batch_engine();
    scrs_status s = e->eng.init(c);
    if (s.code != SCRS_OK) {
        delete e;
        return s;
    }
    *out = e;
    return ok();
}

extern "C" void scrs_batch_engine_destroy(scrs_batch_engine *eng) { delete eng; }

extern "C" scrs_status scrs_batch_engine_submit(scrs_batch_engine *eng, const scrs_job *jobs, uint64_t count,
                                                void *stream, uint64_t *failed_job) {
    if (!eng) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!jobs && count) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    return eng->eng.submit(jobs, count, static_cast<hipStream_t>(stream), failed_job);
}

extern "C" scrs_status scrs_batch_engine_stats(const scrs_batch_engine *eng, scrs_batch_stats *out) {
    if (!eng) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = eng->eng.stats();
    return ok();
}
