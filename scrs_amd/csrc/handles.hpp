// handles.hpp - the opaque handle types behind the C ABI.
#pragma once

#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/scrs.h"
#include "device.hpp"
#include "engine.hpp"

// Caller-owned encode workspace (BatchEncoder::Scratch, codec.rs:138).
struct scrs_encode_scratch {
    int32_t engine = 0;
    uint64_t k = 0, m = 0, L = 0;
    scrs::DevBuf d_data, d_rep, d_work;
    scrs::HostBuf h_data, h_rep;
};

struct scrs_batch_encoder {
    scrs_profile prof{};
    int device = 0;
    uint32_t groups = 0;
    scrs::DevBuf d_tables;       // GF(2^8) product tables of all k terms
    int32_t kernel = SCRS_KERNEL_TABLE;
    scrs::DevBuf d_bm_tables;    // bit-matrix operands of the kernel bm_kind
    int32_t bm_kind = SCRS_KERNEL_TABLE;  // none built yet
    scrs::SplitBuffer split;     // partial rows of launches too small to fill the device
    hipStream_t stream = nullptr;  // host-slice calls only, created on first use
    std::once_flag stream_once;
    hipError_t stream_err = hipSuccess;
    std::mutex mu;               // guards `own` for scratch-less host-slice calls
    scrs_encode_scratch own{};
    void *afft = nullptr;        // AFFT engine state (afft engine only)
};

// Caller-owned batch-decode workspace (BatchDecoder::Scratch, codec.rs:159).
struct scrs_decode_scratch {
    int32_t engine = 0;
    uint64_t k = 0, m = 0, L = 0;
    scrs::DevBuf d_payload, d_out, d_ptrs;
    scrs::HostBuf h_payload, h_out;
    // Per-call host state, sized on the first call and reused after.
    std::vector<uint64_t> seen;
    std::vector<uint16_t> idx16;
    std::vector<const uint8_t *> ptrs;
};

struct scrs_batch_decoder {
    scrs_profile prof{};
    int device = 0;
    hipStream_t stream = nullptr;  // host-slice calls only, created on first use
    scrs::DecodeEngine eng;
    scrs_decode_scratch own{};
    void *afft = nullptr;        // AFFT engine state (afft engine only)
    int32_t kernel = SCRS_KERNEL_TABLE;
    std::vector<uint64_t> seen;  // index bitmap of the device batch calls, reused
};

// One cached reconstruction recipe: the device product tables, row count
// and missing rows of one receipt pattern (decoder/recipe.rs:9-60).
struct RecipeEntry {
    std::vector<uint64_t> pattern;  // receipt bitmap words
    scrs::DevBuf d_tables, d_missing, d_rows;
    uint64_t bytes = 0;
};

// Streaming-decoder scratch (Decoder::Scratch): a recipe cache for the
// GF(2^8) engines (decoder/cache.rs:15-100), a geometry-tagged workspace for
// Tower and AFFT.
struct scrs_decoder_scratch {
    int32_t engine = 0;
    uint64_t k = 0, m = 0, L = 0;
    uint64_t capacity = 16;
    std::list<RecipeEntry> lru;     // most recent first
    uint64_t hits = 0, misses = 0;
};

struct scrs_decoder {
    scrs_profile prof{};
    int device = 0;
    hipStream_t stream = nullptr;
    scrs::DecodeEngine eng;
    scrs::DevBuf d_slots;           // n x L received payloads
    scrs::HostBuf h_slots;          // pinned staging of host pushes
    scrs::DevBuf d_out;
    scrs::HostBuf h_out;
    std::vector<uint64_t> bits;     // receipt bitmap
    uint64_t distinct = 0, received = 0;
    void *afft = nullptr;
    // Per-finalize host state (reserved at creation) and the event that
    // orders device pushes and finalizes against the decoder's stream.
    std::vector<uint16_t> idx;
    std::vector<uint64_t> key;
    hipEvent_t order_ev = nullptr;
};

struct scrs_incr_encoder {
    scrs_profile prof{};
    int device = 0;
    bool gf16 = false;
    hipStream_t stream = nullptr;
    uint32_t groups = 0;
    uint64_t group_bytes = 0;       // table bytes per (term, group)
    scrs::DevBuf d_tables;          // product tables of all k terms
    scrs::DevBuf d_repairs;         // m x L accumulators
    scrs::DevBuf d_pending;         // staged fed symbols, kPending x L
    scrs::DevBuf d_ptab;            // their gathered tables
    scrs::HostBuf h_pending;        // pinned staging of host feeds
    scrs::HostBuf h_repairs;        // host mirror returned by repair()
    scrs::SplitBuffer split;        // partial rows of lone-codeword batch encodes
    std::vector<uint32_t> pending;  // data indices staged, in feed order
    std::vector<uint8_t> fed;
    uint64_t fed_count = 0;
    bool mirror_valid = false;
    bool feed_ready = false;        // stream and accumulators exist (first feed / repair)
    hipEvent_t order_ev = nullptr;  // orders device feeds and repairs_dev readers (made with the stream)
};
