/*
 * scrs_status.h - status codes shared by the SCRS HIP library's C ABI.
 *
 * Every entry point returns an `scrs_status`: `code` names the error
 * variant, and `a` / `b` carry the variant's fields in the order the
 * reference declares them.  The variants mirror the reference's three error
 * enums one for one:
 *   ConfigError   scrs/src/error.rs:12-31
 *   EncodeError   scrs/src/error.rs:35-71
 *   DecodeError   scrs/src/error.rs:75-126
 * and the transform's length error (scrs/src/afft/transform.rs:16-22).
 * Codes 90+ have no reference counterpart: they report misuse of the C ABI
 * itself (null handles) or a failure of the HIP runtime.
 */
#ifndef SCRS_STATUS_H
#define SCRS_STATUS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum scrs_code {
    SCRS_OK = 0,

    /* ConfigError (error.rs:12) */
    SCRS_CONFIG_ZERO_DIMENSION = 10,
    SCRS_CONFIG_TOO_MANY_SYMBOLS = 11, /* a = cap */
    SCRS_CONFIG_ZERO_SYMBOL_LEN = 12,
    SCRS_CONFIG_ODD_SYMBOL_LEN = 13,
    SCRS_CONFIG_UNSUPPORTED_MODE = 14, /* a = engine */

    /* EncodeError (error.rs:35) */
    SCRS_ENCODE_WRONG_INPUT_LEN = 20,    /* a = expected, b = got */
    SCRS_ENCODE_INDEX_OUT_OF_RANGE = 21, /* a = index, b = n */
    SCRS_ENCODE_WRONG_PAYLOAD_LEN = 22,  /* a = expected, b = got */
    SCRS_ENCODE_DUPLICATE_DATA = 23,     /* a = index */
    SCRS_ENCODE_WRONG_OUTPUT_LEN = 24,   /* a = expected, b = got */
    SCRS_ENCODE_SCRATCH_MISMATCH = 25,

    /* DecodeError (error.rs:75) */
    SCRS_DECODE_WRONG_COUNT = 30,        /* a = expected, b = got */
    SCRS_DECODE_INDEX_OUT_OF_RANGE = 31, /* a = index, b = n */
    SCRS_DECODE_DUPLICATE_INDEX = 32,    /* a = index */
    SCRS_DECODE_WRONG_PAYLOAD_LEN = 33,  /* a = expected, b = got */
    SCRS_DECODE_TOO_MANY_SYMBOLS = 34,   /* a = cap, b = received */
    SCRS_DECODE_INSUFFICIENT_RANK = 35,  /* a = rank, b = k */
    SCRS_DECODE_WRONG_OUTPUT_LEN = 36,   /* a = expected, b = got */
    SCRS_DECODE_SCRATCH_MISMATCH = 37,

    /* TransformLengthError (afft/transform.rs:16-22) */
    SCRS_TRANSFORM_WRONG_LEN = 40,       /* a = expected, b = got */

    /* C-ABI / runtime failures (no reference counterpart) */
    SCRS_ERR_INVALID_ARGUMENT = 90, /* null handle or pointer; a = argument position */
    SCRS_ERR_DEVICE = 91,           /* a = hipError_t value */
    SCRS_ERR_OUT_OF_MEMORY = 92,    /* a = bytes requested */
    SCRS_ERR_CAPACITY = 93          /* a request exceeds a fixed library limit; a = limit */
} scrs_code;

typedef struct scrs_status {
    int32_t code;
    uint64_t a;
    uint64_t b;
} scrs_status;

/* Engines, numbered as the reference's `Engine` enum (codec.rs:30-40). */
typedef enum scrs_engine {
    SCRS_ENGINE_STANDARD_CAUCHY = 0,
    SCRS_ENGINE_GOOD_CAUCHY = 1,
    SCRS_ENGINE_TOWER = 2,
    SCRS_ENGINE_AFFT = 3
} scrs_engine;

/* Fields, as the reference's `Field` enum (codec.rs:20-26). */
typedef enum scrs_field {
    SCRS_FIELD_GF256 = 0,
    SCRS_FIELD_GF65536 = 1
} scrs_field;

/* PushOutcome (transport/symbol_sink.rs:11-27). */
typedef enum scrs_push_kind {
    SCRS_PUSH_ADVANCED = 0,
    SCRS_PUSH_COMPLETE = 1,
    SCRS_PUSH_DEPENDENT = 2
} scrs_push_kind;

typedef struct scrs_push_outcome {
    int32_t kind;      /* scrs_push_kind */
    uint64_t rank;     /* valid for SCRS_PUSH_ADVANCED */
    uint64_t received; /* valid for SCRS_PUSH_ADVANCED */
} scrs_push_outcome;

#ifdef __cplusplus
}
#endif

#endif /* SCRS_STATUS_H */
