// batch_engine.hpp - planner, resident caches, graph replay and autotuner
// of the fused uneven batch.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/scrs.h"
#include "batch.hpp"
#include "device.hpp"

namespace scrs {

// Resident tables live in one device arena: first-fit over a free list with
// coalescing, so filling a cache entry costs no allocation call.
class DeviceArena {
  public:
    hipError_t init(uint64_t bytes);
    bool alloc(uint64_t bytes, uint64_t *off);
    void release(uint64_t off, uint64_t bytes);
    uint8_t *base() const { return buf_.as<uint8_t>(); }
    uint64_t capacity() const { return cap_; }

  private:
    DevBuf buf_;
    uint64_t cap_ = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> size
};

// A cache entry is released only once the last batch that read it has
// completed.
struct CacheEntry {
    uint64_t off = 0, bytes = 0;
    uint64_t last_use = 0;  // batch sequence number
    uint32_t rows = 0, groups = 0;
    std::list<CacheEntry *>::iterator lru;
    bool recipe = false;
    std::tuple<int32_t, uint64_t, uint64_t> enc_key;
    std::vector<uint64_t> recipe_key;
};

// Recipe keys: received-index bitmap, then engine, k and m.
struct RecipeKeyHash {
    size_t operator()(const std::vector<uint64_t> &v) const {
        uint64_t h = 0x9E3779B97F4A7C15ull ^ v.size();
        for (uint64_t w : v) {
            h ^= w + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
            h = (h ^ (h >> 31)) * 0xBF58476D1CE4E5B9ull;
        }
        return static_cast<size_t>(h ^ (h >> 29));
    }
};

class BatchEngine {
  public:
    BatchEngine() = default;
    BatchEngine(const BatchEngine &) = delete;
    BatchEngine &operator=(const BatchEngine &) = delete;
    ~BatchEngine();

    scrs_status init(const scrs_batch_config &cfg);
    // Validate every job, then enqueue the whole batch on `s`.  Nothing is
    // enqueued when a job fails validation; *failed gets its index.
    scrs_status submit(const scrs_job *jobs, uint64_t count, hipStream_t s, uint64_t *failed);
    scrs_batch_stats stats() const;

  private:
    struct Slot {
        HostBuf host;        // staged header, ranges, solves, pointer and index arrays
        DevBuf dev;          // device copy of the staged bytes
        DevBuf scratch;      // solve work and fused coefficients
        DevBuf counter;      // unit queue: draws only ever add to it
        uint32_t claims = 0; // its value when the next launch starts
        hipEvent_t done = nullptr;
        uint64_t seq = 0;
        uint32_t generation = 0;  // bumps when `dev` moves
        // Multiply-XOR launches captured against this slot's buffers, per
        // tile width and with or without the reduce launch.
        std::map<uint32_t, std::pair<uint32_t, hipGraphExec_t>> graphs;
    };
    // Staging slots: a batch reuses the slot of the batch four back, so
    // batches in flight on several streams rarely wait for a slot.
    static constexpr uint32_t kSlots = 4;
    // Online tuning of one batch signature: every candidate (tile bytes,
    // codewords per unit) runs on real batches of that signature, timed with
    // events, and the fastest per byte is kept once each has run.
    struct Candidate {
        uint32_t tile = 0, cw = 0;
        double best_ns_per_byte = -1;
        uint32_t runs = 0;
    };
    // Timed runs per candidate before the signature is decided (the best of
    // them counts, so one run slowed by a neighbour does not decide).
    static constexpr uint32_t kTuneRuns = 2;
    struct Tuning {
        std::vector<Candidate> cands;
        uint32_t next = 0;
    };
    // Batch signature: log2 of the mean symbol length, GF(2^16) present,
    // encode share in quarters, and whether an encode geometry repeats.
    using Sig = std::array<uint32_t, 4>;
    struct Sample {
        Sig sig{};
        uint32_t cand = 0;
        uint64_t bytes = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
    };
    void harvest_samples();
    struct Plan;

    scrs_status validate(const scrs_job &j) const;
    scrs_status plan(const scrs_job *jobs, uint64_t count, uint32_t tile, uint32_t cw, Plan &p);
    void split_terms(Plan &p, uint64_t &units);
    hipError_t launch(Slot &slot, const Plan &p, uint32_t tile, hipStream_t s);
    hipError_t run_graph(Slot &slot, uint32_t tile, bool reduce, hipStream_t s);
    CacheEntry *encoder_tables(const scrs_profile &p, Plan &plan, scrs_status *err);
    CacheEntry *new_entry(uint64_t bytes, scrs_status *err);
    bool make_room();
    void retire_completed();
    void touch(CacheEntry *e);
    uint32_t *entry_ptr(const CacheEntry *e) const { return reinterpret_cast<uint32_t *>(arena_.base() + e->off); }

    scrs_batch_config cfg_{};
    int device_ = 0;
    hipStream_t capture_ = nullptr;
    DevBuf coords_[3];   // Standard, Good, Tower coordinates of every wire index
    DevBuf gf8_;         // exp (512) + log (256)
    const uint16_t *gf16_log_exp_ = nullptr;  // GF(2^16) log / exp (shared per device)
    Slot slots_[kSlots];
    uint32_t next_slot_ = 0;
    uint64_t seq_ = 0;

    DeviceArena arena_;
    // The previous batch's tables and recipes are written (its solve has
    // run), on whatever stream: a batch's multiply-XOR may read them.
    hipEvent_t last_solved_ = nullptr;
    bool solves_recorded_ = false;
    // Resident caches, least recently used first.
    std::map<std::tuple<int32_t, uint64_t, uint64_t>, CacheEntry *> enc_cache_;
    std::unordered_map<std::vector<uint64_t>, CacheEntry *, RecipeKeyHash> recipe_cache_;
    std::list<CacheEntry *> lru_;
    std::vector<CacheEntry *> retired_;
    uint64_t cache_bytes_ = 0;
    std::deque<std::pair<uint64_t, hipEvent_t>> inflight_;
    uint64_t completed_ = 0;

    // Planner scratch reused across jobs and batches (no per-job allocation).
    struct EncKey {
        int32_t engine;
        int cls;
        uint64_t k, m, L, job;
        bool operator<(const EncKey &o) const {
            return std::tie(engine, k, m, L, cls, job) < std::tie(o.engine, o.k, o.m, o.L, o.cls, o.job);
        }
    };
    std::vector<EncKey> enc_keys_;
    std::vector<uint64_t> decodes_, key_;
    mutable std::vector<uint64_t> seen_;  // validate(), under mu_
    std::vector<uint32_t> pos_;     // wire index -> position in the job's index list
    std::vector<uint16_t> order_, missing_;
    std::vector<std::tuple<int32_t, uint64_t, uint64_t, uint64_t>> sig_keys_;

    // Autotuner: batch signature -> (tile bytes, codewords per unit).
    std::map<Sig, std::pair<uint32_t, uint32_t>> tuned_;
    std::map<Sig, Tuning> tuning_;
    std::deque<Sample> samples_;
    std::vector<hipEvent_t> event_pool_;
    mutable std::mutex mu_;
    scrs_batch_stats st_{};
};

}  // namespace scrs
