/*
 * scrs.h - C ABI of the SCRS payload-arithmetic engine for AMD MI355X (gfx950).
 *
 * The reference exposes its codecs as Rust traits.  Each entry point below is
 * the C-ABI form of one trait method or constructor; the comment names the
 * reference item it replaces (paths relative to the reference repository).
 * Handles are opaque.  Every call returns an scrs_status whose code and
 * fields mirror the reference's error variants (scrs_status.h), and the
 * order of checks matches the reference, so the same bad input reports the
 * same error.
 *
 * Buffers:
 *   - Host-slice calls (`*_into`, `*_into_with`) take ordinary host memory,
 *     as the reference's `&[u8]` / `&mut [u8]` arguments do; they stage the
 *     bytes through device memory and return when the result is in host
 *     memory.
 *   - Device calls (`*_dev`) take device pointers and a hipStream_t (passed as
 *     void*, null = the null stream) and return once the work is enqueued.
 *
 * Threading: encoders may be used from several threads at once (the
 * reference's encoders take &self); decoders and scratches need external
 * synchronisation (the reference's decoders take &mut self).
 */
#ifndef SCRS_H
#define SCRS_H

#include <stddef.h>
#include <stdint.h>

#include "scrs_status.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Validated geometry (codec.rs:54-103). */
typedef struct scrs_profile {
    int32_t engine; /* scrs_engine */
    uint64_t k;
    uint64_t m;
    uint64_t symbol_len;
} scrs_profile;

typedef struct scrs_batch_encoder scrs_batch_encoder;
typedef struct scrs_batch_decoder scrs_batch_decoder;
typedef struct scrs_encode_scratch scrs_encode_scratch;
typedef struct scrs_decode_scratch scrs_decode_scratch;
typedef struct scrs_decoder scrs_decoder;
typedef struct scrs_decoder_scratch scrs_decoder_scratch;
typedef struct scrs_incr_encoder scrs_incr_encoder;
typedef struct scrs_batch_engine scrs_batch_engine;
typedef struct scrs_server scrs_server;

/* ---- Library ---------------------------------------------------------- */

/* Version of this ABI (no reference counterpart). */
uint32_t scrs_abi_version(void);

/* Human-readable name of a status code (no reference counterpart). */
const char *scrs_status_name(int32_t code);

/* Device and pinned-host allocations the library has made since it was
 * loaded (no reference counterpart).  The observable of the zero-allocation
 * steady state the reference tests with a counting allocator
 * (tests/zero_alloc.rs:48-146): once a codec and its scratch have run one
 * block, further blocks of the same geometry and receipt pattern leave it
 * unchanged. */
uint64_t scrs_device_alloc_count(void);

/* Heap allocations (operator new, of any size) made by the library's own
 * code since it was loaded (no reference counterpart).  The host side of the
 * same steady state: the reference's counting global allocator
 * (tests/zero_alloc.rs:13-46) sees no allocation in the second block, and
 * neither does this count.  Allocations inside the HIP runtime are not the
 * library's and are not counted. */
uint64_t scrs_host_alloc_count(void);

/* Term split of small codec launches (no reference counterpart).  When
 * enabled (the default), a codec launch too small to fill the device (a lone
 * codeword, say) splits each codeword's terms across workgroups: each
 * segment stores its partial rows to a buffer the codec keeps, and a second
 * kernel XORs them into the destinations (atomic XORs into rows cleared
 * first when no buffer fits).  It applies to launches whose destination
 * rows and symbol_len are multiples of 4 bytes and that overwrite their
 * destinations.  The bytes are the same either way.  Disabled, a lone
 * codeword runs on one workgroup per row group and column tile, each
 * folding every term.  Process-wide; returns the previous setting. */
int32_t scrs_set_term_split(int32_t enable);

/* Largest number of missing data symbols for which an additive-FFT decode
 * takes the targeted path: a closed-form solve over the systematic locator
 * and one multiply-XOR of the k received symbols into the missing ones,
 * instead of the transforms over the whole domain.  Replaces the constant
 * TARGETED_MAX_MISSING = 5 of scrs/src/afft/decoder.rs:578 (the same
 * default here); the decoded bytes are the same on either path, so the
 * setting only moves the crossover.  Clamped to [0, 64]; 0 sends every
 * decode through the transforms.  Process-wide; returns the previous
 * setting. */
int32_t scrs_afft_set_targeted_max(int32_t max_missing);

/* ---- Profile selection ------------------------------------------------ */

/* Engine::field (codec.rs:42-50).  Returns -1 for an unknown engine. */
int32_t scrs_engine_field(int32_t engine);

/* recommended_gf16_engine (lib.rs:85-93): AFFT when k+m > 256 or 3m >= k,
 * else Tower.  Returns an scrs_engine. */
int32_t scrs_recommended_gf16_engine(uint64_t k, uint64_t m);

/* Profile::resolve (selector.rs:40-61).  Check order: ZeroDimension ->
 * ZeroSymbolLen -> OddSymbolLen -> TooManySymbols{cap}. */
scrs_status scrs_profile_resolve(int32_t engine, uint64_t k, uint64_t m, uint64_t symbol_len,
                                 scrs_profile *out);

/* Profile::recommended (selector.rs:67-87): GF(256) picks Good Cauchy when
 * k+m <= 255, else Standard; GF(65536) follows scrs_recommended_gf16_engine. */
scrs_status scrs_profile_recommended(int32_t field, uint64_t k, uint64_t m, uint64_t symbol_len,
                                     scrs_profile *out);

/* ---- BatchEncoder (codec.rs:136-151) ------------------------------------ */

/* batch_encoder(&profile) (selector.rs:457-467).  Tower is incremental-only:
 * UnsupportedMode{engine}. */
scrs_status scrs_batch_encoder_create(const scrs_profile *profile, scrs_batch_encoder **out);
void scrs_batch_encoder_destroy(scrs_batch_encoder *enc);
/* Coded::{k, m, symbol_len} (codec.rs:106-117). */
scrs_status scrs_batch_encoder_profile(const scrs_batch_encoder *enc, scrs_profile *out);

/* BatchEncoder::scratch (codec.rs:140). */
scrs_status scrs_encode_scratch_create(const scrs_batch_encoder *enc, scrs_encode_scratch **out);
void scrs_encode_scratch_destroy(scrs_encode_scratch *scratch);

/* BatchEncoder::encode_into (codec.rs:143; batch/codec.rs:156-193).
 * data = k * symbol_len bytes, repairs = m * symbol_len bytes, overwritten.
 * Check order: WrongInputLen -> WrongOutputLen. */
scrs_status scrs_encode_into(const scrs_batch_encoder *enc, const uint8_t *data, uint64_t data_len,
                             uint8_t *repairs, uint64_t repairs_len);

/* BatchEncoder::encode_into_with (codec.rs:145-150; selector.rs:510-528).
 * A scratch of the wrong engine kind: ScratchMismatch (checked first). */
scrs_status scrs_encode_into_with(const scrs_batch_encoder *enc, const uint8_t *data,
                                  uint64_t data_len, uint8_t *repairs, uint64_t repairs_len,
                                  scrs_encode_scratch *scratch);

/* Batched device encode (no reference counterpart; the batch form of
 * encode_into): `count` codewords back to back, codeword c's data at
 * data_dev + c*k*L and its repairs written to repairs_dev + c*m*L. */
scrs_status scrs_encode_batch_dev(const scrs_batch_encoder *enc, const uint8_t *data_dev,
                                  uint8_t *repairs_dev, uint64_t count, void *stream);

/* ---- Multiply-XOR kernel of the GF(256) codecs (no reference counterpart) */

typedef enum scrs_kernel {
    SCRS_KERNEL_TABLE = 0,         /* split-nibble product tables in LDS (default) */
    SCRS_KERNEL_BITMATRIX = 1,     /* GF(2) bit-matrix product on int8 MFMA */
    SCRS_KERNEL_BITMATRIX_FP4 = 2  /* the same product on e2m1 (fp4) MFMA, fp32 accumulation */
} scrs_kernel;

/* Select the multiply-XOR kernel of a GF(256) batch encoder or decoder.  All
 * kernels produce the same bytes.  BITMATRIX expresses the code as an
 * (8m) x (8k) 0/1 matrix product with int32 accumulation and parity
 * extraction; it requires k <= 64 (else SCRS_ERR_CAPACITY{a = 64}).
 * BITMATRIX_FP4 computes that product with 0/1 products on 4-bit operands
 * and exact fp32 sums (k <= 256, so every GF(256) geometry).  Both require
 * symbol_len % 4 == 0 (else SCRS_ERR_INVALID_ARGUMENT{a = 2}).  Calls whose
 * buffers are not 4-byte aligned, scrs_decode_batch_ptrs_dev and the
 * streaming codecs use the table kernel.  Tower and AFFT handles:
 * UnsupportedMode{engine}.  Not to be called while other work uses the
 * handle. */
scrs_status scrs_batch_encoder_set_kernel(scrs_batch_encoder *enc, int32_t kernel);
scrs_status scrs_batch_decoder_set_kernel(scrs_batch_decoder *dec, int32_t kernel);

/* ---- BatchDecoder (codec.rs:157-175) ------------------------------------ */

/* batch_decoder(&profile) (selector.rs:270-279). */
scrs_status scrs_batch_decoder_create(const scrs_profile *profile, scrs_batch_decoder **out);
void scrs_batch_decoder_destroy(scrs_batch_decoder *dec);
scrs_status scrs_batch_decoder_profile(const scrs_batch_decoder *dec, scrs_profile *out);

/* BatchDecoder::scratch (codec.rs:161; batch/codec.rs:245-258). */
scrs_status scrs_decode_scratch_create(const scrs_batch_decoder *dec, scrs_decode_scratch **out);
void scrs_decode_scratch_destroy(scrs_decode_scratch *scratch);

/* BatchDecoder::decode_into (codec.rs:163-167; batch/codec.rs:265-271).
 * `count` symbols: indices[i] in 0..n with payloads[i] of payload_lens[i]
 * bytes.  out = k * symbol_len bytes, every byte overwritten.
 * GF(256) check order: WrongOutputLen -> WrongCount -> per symbol
 * IndexOutOfRange -> DuplicateIndex -> WrongPayloadLen
 * (batch/codec.rs:293-331). */
scrs_status scrs_decode_into(scrs_batch_decoder *dec, const uint64_t *indices,
                             const uint8_t *const *payloads, const uint64_t *payload_lens,
                             uint64_t count, uint8_t *out, uint64_t out_len);

/* BatchDecoder::decode_into_with (codec.rs:169-174; batch/codec.rs:278-406).
 * A scratch from another geometry: ScratchMismatch (checked first). */
scrs_status scrs_decode_into_with(scrs_batch_decoder *dec, const uint64_t *indices,
                                  const uint8_t *const *payloads, const uint64_t *payload_lens,
                                  uint64_t count, uint8_t *out, uint64_t out_len,
                                  scrs_decode_scratch *scratch);

/* Batched device decode (no reference counterpart; the batch form of
 * decode_into).  For codeword c, indices[c*k .. c*k+k) (host memory) name
 * the k received symbols; symbol i < k is read from data_dev + (c*k+i)*L and
 * symbol k+j from repairs_dev + (c*m+j)*L.  The k data symbols are written
 * to out_dev + c*k*L.  Indices are validated on the host with the
 * decode_into check order before anything is enqueued. */
scrs_status scrs_decode_batch_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                  const uint8_t *data_dev, const uint8_t *repairs_dev,
                                  uint8_t *out_dev, uint64_t count, void *stream);

/* Same, with the received payloads given as a device array of `count * k`
 * device pointers (payload_ptrs_dev[c*k + t] is the payload of indices[c*k+t]). */
scrs_status scrs_decode_batch_ptrs_dev(scrs_batch_decoder *dec, const uint16_t *indices,
                                       const uint8_t *const *payload_ptrs_dev, uint8_t *out_dev,
                                       uint64_t count, void *stream);

/* Same as scrs_decode_batch_dev, allocating nothing (no reference
 * counterpart: for callers that cannot allocate during a call, such as code
 * under HIP graph capture or a tight serving loop).  The launch workspace
 * (indices, solve outputs, pointer tables) is a fixed per-device array of
 * scrs_noalloc_scratch_bytes() bytes (32 MiB), held until the process ends;
 * scrs_decode_batch_dev_noalloc_bytes() gives the bytes a call of `count`
 * codewords needs.  GF(256) engines only (else UnsupportedMode{engine}), on
 * the split-nibble table kernel.  Same index checks as scrs_decode_batch_dev;
 * then count > 65535 or a workspace above the array's size is
 * SCRS_ERR_CAPACITY{a = limit, b = requested}.  The indices are copied from
 * host memory before the call returns.  Every call on one device shares
 * that device's array: calls that overlap on the device (different streams,
 * or threads) race and must be serialised by the caller. */
scrs_status scrs_decode_batch_dev_noalloc(scrs_batch_decoder *dec, const uint16_t *indices,
                                          const uint8_t *data_dev, const uint8_t *repairs_dev, uint8_t *out_dev,
                                          uint64_t count, void *stream);
uint64_t scrs_noalloc_scratch_bytes(void);
/* Workspace bytes of a scrs_decode_batch_dev_noalloc call of `count`
 * codewords (0 for a null or non-GF(256) decoder). */
uint64_t scrs_decode_batch_dev_noalloc_bytes(const scrs_batch_decoder *dec, uint64_t count);

/* ---- Decoder: streaming receive (codec.rs:180-201) ---------------------- */

/* decoder(&profile) (selector.rs:117-125).  Device buffers for n received
 * symbols are reserved up front; allocation failure is TooManySymbols{cap}
 * (tower/decoder.rs:98-108). */
scrs_status scrs_decoder_create(const scrs_profile *profile, scrs_decoder **out);
void scrs_decoder_destroy(scrs_decoder *dec);
scrs_status scrs_decoder_profile(const scrs_decoder *dec, scrs_profile *out);

/* SymbolSink::push (transport/symbol_sink.rs:36-54; decoder/streaming.rs:387-429;
 * tower/decoder.rs:371-405).  Check order: IndexOutOfRange{index,n} ->
 * WrongPayloadLen -> TooManySymbols{cap=n, received}; then received += 1;
 * a duplicate index or a push after completion is Dependent; otherwise the
 * payload is stored and the outcome is Complete or Advanced{rank, received}.
 * The payload is copied before the call returns. */
scrs_status scrs_decoder_push(scrs_decoder *dec, uint64_t index, const uint8_t *payload,
                              uint64_t len, scrs_push_outcome *out);

/* push with a device-resident payload, copied on `stream` (no reference
 * counterpart; same checks and outcomes as scrs_decoder_push). */
scrs_status scrs_decoder_push_dev(scrs_decoder *dec, uint64_t index, const uint8_t *payload_dev,
                                  uint64_t len, void *stream, scrs_push_outcome *out);

/* Decoder::{rank, received} (codec.rs:187-189), SymbolSink::is_complete. */
uint64_t scrs_decoder_rank(const scrs_decoder *dec);
uint64_t scrs_decoder_received(const scrs_decoder *dec);
int32_t scrs_decoder_is_complete(const scrs_decoder *dec);

/* Decoder::reset (codec.rs:191): clears the receipt state, keeps buffers. */
scrs_status scrs_decoder_reset(scrs_decoder *dec);

/* Decoder::scratch (codec.rs:185; selector.rs:188-196): a 16-entry recipe
 * cache for the GF(256) engines (decoder/streaming.rs:461-463,
 * decoder/cache.rs), a geometry-tagged workspace for Tower. */
scrs_status scrs_decoder_scratch_create(const scrs_decoder *dec, scrs_decoder_scratch **out);
void scrs_decoder_scratch_destroy(scrs_decoder_scratch *scratch);
/* RecipeCache::{hits, misses, len} and the bytes of the cached recipes
 * (decoder/cache.rs:60-100). */
scrs_status scrs_decoder_scratch_stats(const scrs_decoder_scratch *scratch, uint64_t *hits,
                                       uint64_t *misses, uint64_t *len, uint64_t *recipe_bytes);

/* Decoder::finalize_into (codec.rs:194; decoder/streaming.rs:476-488).
 * Check order: InsufficientRank{rank,k} -> WrongOutputLen. */
scrs_status scrs_decoder_finalize_into(scrs_decoder *dec, uint8_t *out, uint64_t out_len);

/* Decoder::finalize_into_with (codec.rs:196-200; selector.rs:224-240).  A
 * scratch of another engine kind is ScratchMismatch (checked first); then
 * InsufficientRank -> WrongOutputLen -> ScratchMismatch for a Tower scratch
 * of another geometry (tower/decoder.rs:200-221). */
scrs_status scrs_decoder_finalize_into_with(scrs_decoder *dec, uint8_t *out, uint64_t out_len,
                                            scrs_decoder_scratch *scratch);

/* finalize_into writing device memory on `stream` (no reference
 * counterpart; same checks as scrs_decoder_finalize_into). */
scrs_status scrs_decoder_finalize_into_dev(scrs_decoder *dec, uint8_t *out_dev, uint64_t out_len,
                                           void *stream);

/* ---- IncrementalEncoder (codec.rs:122-131) ------------------------------ */

/* incremental_encoder(&profile) (selector.rs:370-381): Good Cauchy and
 * Tower; any other engine is UnsupportedMode{engine}. */
scrs_status scrs_incr_encoder_create(const scrs_profile *profile, scrs_incr_encoder **out);
void scrs_incr_encoder_destroy(scrs_incr_encoder *enc);
scrs_status scrs_incr_encoder_profile(const scrs_incr_encoder *enc, scrs_profile *out);

/* IncrementalEncoder::feed (encoder/streaming.rs:162-190; tower/encoder.rs:89-116).
 * Check order: IndexOutOfRange{index, n=k+m} for index >= k ->
 * WrongPayloadLen -> DuplicateData{index}.  The data is copied before the
 * call returns; contributions are folded in batches on the device. */
scrs_status scrs_incr_feed(scrs_incr_encoder *enc, uint64_t index, const uint8_t *data,
                           uint64_t len);

/* feed with device-resident data read on `stream` (no reference counterpart;
 * same checks as scrs_incr_feed). */
scrs_status scrs_incr_feed_dev(scrs_incr_encoder *enc, uint64_t index, const uint8_t *data_dev,
                               uint64_t len, void *stream);

/* IncrementalEncoder::repair (encoder/streaming.rs:197-210; tower/encoder.rs:122-131):
 * IndexOutOfRange{index, n} for index >= m.  *ptr receives a host view of
 * repair `index` (the partial sum of the data fed so far), valid until the
 * next feed, reset or destroy. */
scrs_status scrs_incr_repair(scrs_incr_encoder *enc, uint64_t index, const uint8_t **ptr,
                             uint64_t *len);

/* Device view of all m repair accumulators (m * symbol_len bytes, repair j
 * at j * symbol_len), current for work enqueued on `stream` after the call
 * (no reference counterpart). */
scrs_status scrs_incr_repairs_dev(scrs_incr_encoder *enc, uint8_t **repairs_dev, void *stream);

/* IncrementalEncoder::{fed_count, reset} (codec.rs:128-130). */
uint64_t scrs_incr_fed_count(const scrs_incr_encoder *enc);
scrs_status scrs_incr_reset(scrs_incr_encoder *enc);

/* Batched device encode with the incremental encoder's coefficients, equal
 * to feeding all k data symbols of each of `count` codewords (no reference
 * counterpart).  Layout as scrs_encode_batch_dev.  The only batched encoder
 * of the Tower engine. */
scrs_status scrs_incr_encode_batch_dev(const scrs_incr_encoder *enc, const uint8_t *data_dev,
                                       uint8_t *repairs_dev, uint64_t count, void *stream);

/* ---- Additive-FFT transform plan (afft/transform.rs:24-95) --------------- */

typedef struct scrs_transform_plan scrs_transform_plan;

/* TransformPlan::new(size) (afft/transform.rs:41-60): the LCH novel-basis
 * transform over the additive points 0..size-1 of GF(2^16).  size must be a
 * power of two in 1..=65536; any other size is SCRS_ERR_INVALID_ARGUMENT
 * {a = 1} (the reference returns None). */
scrs_status scrs_transform_plan_create(uint64_t size, scrs_transform_plan **out);
void scrs_transform_plan_destroy(scrs_transform_plan *plan);
/* TransformPlan::{size, log_size} (afft/transform.rs:63-71). */
uint64_t scrs_transform_plan_size(const scrs_transform_plan *plan);
uint32_t scrs_transform_plan_log_size(const scrs_transform_plan *plan);

/* TransformPlan::forward (afft/transform.rs:73-82): evaluate `len`
 * novel-basis coefficients at the plan's points, in place.
 * TransformPlan::inverse (:84-91): evaluations back to coefficients.
 * len != size is TransformLengthError{a = expected, b = got}.  Host
 * memory; returns once the result is back. */
scrs_status scrs_transform_forward(const scrs_transform_plan *plan, uint16_t *values, uint64_t len);
scrs_status scrs_transform_inverse(const scrs_transform_plan *plan, uint16_t *values, uint64_t len);

/* Device form (no reference counterpart; the byte form the codecs use,
 * TransformPlan::forward_bytes / inverse_bytes, afft/transform.rs:93-99,
 * 167-172): `count` blocks back to back, each of `size` rows of symbol_len
 * bytes (interleaved little-endian elements, symbol_len even), every element
 * column transformed independently, in place on `stream`.  Element slices
 * are symbol_len = 2.  An odd symbol_len is SCRS_CONFIG_ODD_SYMBOL_LEN; count
 * above 65535 is SCRS_ERR_CAPACITY{a = 65535}. */
scrs_status scrs_transform_forward_dev(const scrs_transform_plan *plan, uint8_t *rows_dev, uint64_t symbol_len,
                                       uint64_t count, void *stream);
scrs_status scrs_transform_inverse_dev(const scrs_transform_plan *plan, uint8_t *rows_dev, uint64_t symbol_len,
                                       uint64_t count, void *stream);

/* ---- Fused batches of uneven codewords (no reference counterpart) ------- */

typedef enum scrs_job_op { SCRS_JOB_ENCODE = 0, SCRS_JOB_DECODE = 1 } scrs_job_op;

/* One codeword's work.  Payload buffers are device memory owned by the
 * caller.  Encode: src = k*symbol_len data bytes, dst = m*symbol_len repairs
 * (Tower encodes as feeding all k data symbols).  Decode: indices = k
 * received wire indices and payloads = their k device pointers (both host
 * arrays, read during the call), dst = k*symbol_len data bytes. */
typedef struct scrs_job {
    int32_t op; /* scrs_job_op */
    scrs_profile profile;
    const uint8_t *src;
    uint8_t *dst;
    const uint16_t *indices;
    const uint8_t *const *payloads;
} scrs_job;

typedef struct scrs_batch_config {
    uint32_t tile_bytes;  /* column tile per work unit: 1024, 2048, 4096; 0 = autotune */
    uint32_t cw_per_unit; /* codewords of one encode geometry per unit: 1, 2, 4, 8; 0 = autotune */
    uint32_t use_graph;   /* replay the batch launches from a captured hipGraph */
    uint64_t cache_bytes; /* device bytes for resident tables and recipes; 0 = 1 GiB */
} scrs_batch_config;

typedef struct scrs_batch_stats {
    uint64_t batches, jobs;
    uint64_t recipe_hits, recipe_misses; /* decode reconstruction recipes */
    uint64_t table_hits, table_misses;   /* encode product tables */
    uint64_t cache_bytes;
    uint32_t tile_bytes, cw_per_unit;    /* configuration of the last batch */
    uint32_t tuned_signatures;
} scrs_batch_stats;

/* A batch engine keeps generator tables and decode recipes resident on the
 * device and runs a whole batch - GF(256) and Tower jobs of any k, m and
 * symbol_len - as one persistent launch chain (optionally a hipGraph).
 * Checks per job follow the codec calls: Profile::resolve, then for decode
 * IndexOutOfRange -> DuplicateIndex per index; AFFT is UnsupportedMode. */
scrs_status scrs_batch_engine_create(const scrs_batch_config *cfg, scrs_batch_engine **out);
void scrs_batch_engine_destroy(scrs_batch_engine *eng);
/* Validates every job first (on failure *failed_job = its index and nothing
 * is enqueued), then enqueues the batch on `stream`.  With autotuning, the
 * first batch of a new shape signature runs once per candidate
 * configuration, synchronously, and the fastest is kept. */
scrs_status scrs_batch_engine_submit(scrs_batch_engine *eng, const scrs_job *jobs, uint64_t count,
                                     void *stream, uint64_t *failed_job);
scrs_status scrs_batch_engine_stats(const scrs_batch_engine *eng, scrs_batch_stats *out);

/* ---- Serving loop (no reference counterpart) ---------------------------- */

typedef struct scrs_server_config {
    int32_t batched;         /* 1: continuous batches through a batch engine; 0: one codec call per job */
    uint32_t max_batch_jobs; /* jobs per batch at most; 0 = 4096 */
    uint64_t hbm_budget;     /* payload bytes admitted at once; 0 = 8 GiB */
    scrs_batch_config batch; /* batch engine configuration (batched mode) */
} scrs_server_config;

typedef struct scrs_completion {
    uint64_t tag;
    scrs_status status;
    uint64_t t_submit_ns, t_start_ns, t_done_ns; /* scrs_now_ns() clock */
} scrs_completion;

/* A server owns a stream and a dispatcher thread.  Jobs are admitted in
 * arrival order while their payload footprint fits the budget; in batched
 * mode everything admitted since the last launch forms the next batch,
 * queued behind the one on the device, so arrivals join while earlier work
 * runs.  Completion times are taken when the dispatcher observes the
 * batch's (or job's) event. */
scrs_status scrs_server_create(const scrs_server_config *cfg, scrs_server **out);
void scrs_server_destroy(scrs_server *srv);
/* Queue one job (copied, including its index and payload arrays). */
scrs_status scrs_server_submit(scrs_server *srv, const scrs_job *job, uint64_t tag);
/* Move up to `cap` completions into `out`; wait != 0 blocks until at least
 * one is available or nothing is outstanding. */
scrs_status scrs_server_poll(scrs_server *srv, scrs_completion *out, uint64_t cap, uint64_t *n, int32_t wait);
/* Submit jobs[i] at arrival_ns[i] after the call starts (a native thread
 * paces the arrivals) and wait for all of them; completions[i] belongs to
 * jobs[i] (tag = i). */
scrs_status scrs_server_replay(scrs_server *srv, const scrs_job *jobs, const uint64_t *arrival_ns,
                               uint64_t count, scrs_completion *completions);
uint64_t scrs_now_ns(void);

typedef struct scrs_server_stats {
    uint64_t batches, jobs, max_batch_jobs;
    uint64_t host_ns;           /* dispatcher time spent enqueueing batches */
    uint64_t gpu_ns, max_gpu_ns; /* device time of the batches (events around them) */
} scrs_server_stats;
/* Counters since creation; read after the dispatcher has seen the work. */
scrs_status scrs_server_get_stats(scrs_server *srv, scrs_server_stats *out);

/* ---- Utilities (no reference counterpart) ----------------------------- */

/* Fill len bytes of device memory with the seeded splitmix64 counter stream
 * (byte i = byte i%8 of splitmix64 word (byte_offset+i)/8 of state seed);
 * byte_offset must be a multiple of 8. */
scrs_status scrs_fill_random_dev(uint8_t *dst_dev, uint64_t len, uint64_t seed,
                                 uint64_t byte_offset, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SCRS_H */
