"""Zero-allocation steady state, after tests/zero_alloc.rs:48-155 of the
reference: once a codec and its scratch have run one block, another block of
the same geometry and receipt pattern through encode_into_with, the
incremental encoder, push + finalize_into_with and decode_into_with makes no
allocation of any kind: no device or pinned-host allocation (hipMalloc /
hipHostMalloc) and no heap allocation (operator new) by the library's code,
both counted by the library.  The reference's five cases, plus additive-FFT
decodes of a few erasures at k >= 16, which take the targeted path with its
split multiply-XOR and partial rows."""
import numpy as np
import pytest

from scrs_amd.codec import (ConfigError, Engine, Profile, batch_decoder, batch_encoder, decoder,
                            device_alloc_count, host_alloc_count, incremental_encoder)

pytestmark = pytest.mark.gpu


def maybe(factory, profile):
    try:
        return factory(profile)
    except ConfigError:
        return None


@pytest.mark.parametrize("engine,k,m,L,missing", [
    (Engine.STANDARD_CAUCHY, 8, 4, 64, 2),
    (Engine.GOOD_CAUCHY, 8, 4, 64, 2),
    (Engine.TOWER, 8, 4, 64, 2),
    (Engine.AFFT, 8, 4, 64, 2),
    (Engine.AFFT, 16, 8, 64, 6),
    (Engine.AFFT, 64, 32, 256, 4),
    (Engine.AFFT, 1024, 512, 128, 5),
])
def test_reusable_codecs_allocate_nothing(torch_cuda, engine, k, m, L, missing):
    p = Profile.resolve(engine, k, m, L)
    constructed = host_alloc_count()
    data = ((np.arange(k * L) * 131 + 7) & 0xFF).astype(np.uint8)
    repairs = np.zeros(m * L, np.uint8)
    enc, inc = maybe(batch_encoder, p), maybe(incremental_encoder, p)
    es = enc.scratch() if enc else None
    # First block: every codec runs once (the incremental encoder included:
    # it sets up its device state on the first feed).
    if enc:
        enc.encode_into_with(data, repairs, es)
    if inc:
        for i in range(k):
            inc.feed(i, data[i * L:(i + 1) * L])
        fed = np.frombuffer(b"".join(inc.repair(j) for j in range(m)), np.uint8)
        if enc:
            assert np.array_equal(fed, repairs)
        repairs = fed.copy()
        inc.reset()
    word = [data[i * L:(i + 1) * L] for i in range(k)] + [repairs[j * L:(j + 1) * L] for j in range(m)]
    received = [(i, word[i]) for i in list(range(missing, k)) + list(range(k, k + missing))]

    streaming = decoder(p)
    ss = streaming.scratch()
    stream_out = np.zeros(k * L, np.uint8)
    for i, payload in received:
        streaming.push(i, payload)
    streaming.finalize_into_with(stream_out, ss)
    streaming.reset()
    batch = batch_decoder(p)
    bs = batch.scratch()
    batch_out = np.zeros(k * L, np.uint8)
    batch.decode_into_with(received, batch_out, bs)

    # The counter sees the library's own allocations: building the codecs
    # (handles, coefficient tables, caches) made some.
    assert host_alloc_count() > constructed
    before, host_before = device_alloc_count(), host_alloc_count()
    if enc:
        enc.encode_into_with(data, repairs, es)
    if inc:
        inc.reset()
        for i in range(k):
            inc.feed(i, data[i * L:(i + 1) * L])
    for i, payload in received:
        streaming.push(i, payload)
    streaming.finalize_into_with(stream_out, ss)
    batch.decode_into_with(received, batch_out, bs)
    assert device_alloc_count() == before, engine
    assert host_alloc_count() == host_before, engine
    assert np.array_equal(stream_out, data)
    assert np.array_equal(batch_out, data)
