"""SCRS payload arithmetic on AMD MI355X (gfx950).

The codecs live in the HIP library ``libscrs_hip.so`` behind the C ABI of
``include/scrs.h``; this package is its Python binding, mirroring the
reference's ``Profile`` selection and codec traits.
"""
from .codec import (  # noqa: F401
    BatchDecoder,
    BatchEngine,
    Job,
    Server,
    BatchEncoder,
    Decoder,
    DecoderScratch,
    IncrementalEncoder,
    PushOutcome,
    ConfigError,
    DecodeError,
    DeviceError,
    EncodeError,
    Engine,
    Field,
    Kernel,
    TransformLengthError,
    TransformPlan,
    Profile,
    ScrsError,
    batch_decoder,
    batch_encoder,
    decoder,
    incremental_encoder,
    fill_random_dev,
    now_ns,
    device_alloc_count,
    host_alloc_count,
    set_term_split,
    afft_set_targeted_max,
    recommended_gf16_engine,
)
