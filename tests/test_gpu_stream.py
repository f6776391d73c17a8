"""GF(2^8) streaming decoder and incremental encoder against the oracle.

Follows the reference's own tests of these types: decoder/streaming.rs:545-760
(recipe cache hits and misses, engine separation, every k-subset, duplicates,
finalize_into vs finalize) and encoder/streaming.rs:259-360.
"""
import itertools

import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchEncoder, ConfigError, DecodeError, Decoder, EncodeError, Engine,
                            IncrementalEncoder, Profile, PushOutcome)

pytestmark = pytest.mark.gpu

ENGINES = [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY]


def encoded(engine, k, m, L, data):
    rep = BatchEncoder(Profile.resolve(engine, k, m, L)).encode(data)
    return [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]


def test_recipe_cache_records_hits_and_misses(torch_cuda):
    # decoder/streaming.rs:551-576
    k, m, L = 4, 3, 8
    data = (np.arange(k * L) % 256).astype(np.uint8)
    sym = encoded(Engine.STANDARD_CAUCHY, k, m, L, data)
    arrival = list(range(k, k + m)) + list(range(0, k - m))
    dec = Decoder(Profile.resolve(Engine.STANDARD_CAUCHY, k, m, L))
    cache = dec.scratch()
    for it in range(2):
        d = Decoder(Profile.resolve(Engine.STANDARD_CAUCHY, k, m, L))
        for i in arrival:
            d.push(i, sym[i])
        out = np.zeros(k * L, np.uint8)
        d.finalize_into_with(out, cache)
        assert np.array_equal(out, data)
        s = cache.stats()
        assert (s["misses"], s["hits"]) == ((1, 0) if it == 0 else (1, 1))


def test_recipe_cache_separates_engines(torch_cuda):
    # decoder/streaming.rs:578-614
    k, m, L = 4, 3, 32
    data = (np.arange(k * L) % 256).astype(np.uint8)
    arrival = [k, k + 1, 2, 3]
    good = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    std = Decoder(Profile.resolve(Engine.STANDARD_CAUCHY, k, m, L))
    cache = good.scratch()
    for dec, engine in ((good, Engine.GOOD_CAUCHY), (std, Engine.STANDARD_CAUCHY)):
        sym = encoded(engine, k, m, L, data)
        for i in arrival:
            dec.push(i, sym[i])
        out = np.zeros(k * L, np.uint8)
        dec.finalize_into_with(out, cache)
        assert np.array_equal(out, data)
    s = cache.stats()
    assert (s["misses"], s["hits"], s["len"]) == (2, 0, 2)


def test_recipe_cache_evicts_least_recent(torch_cuda):
    # RecipeCache capacity 16 (decoder/streaming.rs:461-463, decoder/cache.rs)
    k, m, L = 6, 6, 64
    rng = np.random.default_rng(11)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    sym = encoded(Engine.GOOD_CAUCHY, k, m, L, data)
    dec = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    cache = dec.scratch()
    subsets = list(itertools.combinations(range(k + m), k))[:20]
    for sub in subsets + subsets[-3:]:
        dec.reset()
        for i in sub:
            dec.push(i, sym[i])
        out = np.zeros(k * L, np.uint8)
        dec.finalize_into_with(out, cache)
        assert np.array_equal(out, data), sub
    s = cache.stats()
    assert s["len"] == 16
    assert s["misses"] == 20 and s["hits"] == 3
    assert s["recipe_bytes"] > 0


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", [(3, 2, 4), (4, 3, 8), (5, 4, 33)])
def test_roundtrip_all_subsets_small(torch_cuda, engine, k, m, L):
    # decoder/streaming.rs:625-664, pushed in reverse arrival order
    rng = np.random.default_rng(k * 13 + m)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    sym = encoded(engine, k, m, L, data)
    dec = Decoder(Profile.resolve(engine, k, m, L))
    cache = dec.scratch()
    for sub in itertools.combinations(range(k + m), k):
        dec.reset()
        outcomes = [dec.push(i, sym[i]) for i in sub[::-1]]
        assert outcomes[-1] == PushOutcome.COMPLETE
        assert outcomes[:-1] == [PushOutcome.advanced(r + 1, r + 1) for r in range(k - 1)]
        assert np.array_equal(dec.finalize(), data), sub
        out = np.zeros(k * L, np.uint8)
        dec.finalize_into_with(out, cache)
        assert np.array_equal(out, data), sub


def test_streaming_matches_oracle_at_larger_shapes(torch_cuda):
    for engine in ENGINES:
        for (k, m, L) in [(32, 16, 65536), (100, 28, 1400), (16, 200, 256)]:
            if engine == Engine.GOOD_CAUCHY and k + m > 255:
                continue
            rng = np.random.default_rng(k + m + L)
            data = rng.integers(0, 256, k * L, dtype=np.uint8)
            sym = encoded(engine, k, m, L, data)
            c = oracle.gf256_matrix(int(engine), k, m)
            dec = Decoder(Profile.resolve(engine, k, m, L))
            order = rng.permutation(k + m).tolist()
            for i in order:
                dec.push(i, sym[i])
            assert dec.rank() == k and dec.received() == k + m
            received = [i for i in order][:k]
            code, _, _, want = oracle.gf256_decode(c, k, m, L, [(i, sym[i]) for i in received])
            assert code == 0
            assert np.array_equal(dec.finalize(), want)


def test_duplicate_before_complete_is_dependent(torch_cuda):
    # decoder/streaming.rs:722-739
    k, m, L = 3, 2, 4
    sym = encoded(Engine.STANDARD_CAUCHY, k, m, L, np.full(k * L, 0x42, np.uint8))
    dec = Decoder(Profile.resolve(Engine.STANDARD_CAUCHY, k, m, L))
    assert dec.push(0, sym[0]).kind == "Advanced"
    assert dec.push(0, sym[0]) == PushOutcome.DEPENDENT
    assert dec.rank() == 1 and dec.received() == 2


def test_push_dev_and_finalize_into_dev(torch_cuda):
    torch = torch_cuda
    k, m, L = 12, 6, 4096
    rng = np.random.default_rng(21)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    sym = encoded(Engine.GOOD_CAUCHY, k, m, L, data)
    dev = torch.from_numpy(np.concatenate(sym)).cuda()
    dec = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    for i in rng.permutation(k + m)[:k].tolist():
        dec.push_dev(i, dev.data_ptr() + i * L, L)
    out = torch.zeros(k * L, dtype=torch.uint8, device="cuda")
    dec.finalize_into_dev(out, k * L)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), data)


def test_decoder_errors(torch_cuda):
    k, m, L = 3, 2, 4
    dec = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    with pytest.raises(DecodeError) as e:
        dec.push(5, bytes(1))
    assert e.value == DecodeError("IndexOutOfRange", index=5, n=5)
    with pytest.raises(DecodeError) as e:
        dec.push(1, bytes(1))
    assert e.value == DecodeError("WrongPayloadLen", expected=4, got=1)
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(12, np.uint8))
    assert e.value == DecodeError("InsufficientRank", rank=0, k=3)
    for i in range(5):
        dec.push(i, bytes(4))
    with pytest.raises(DecodeError) as e:
        dec.push(0, bytes(4))
    assert e.value == DecodeError("TooManySymbols", cap=5, received=5)
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(13, np.uint8))
    assert e.value == DecodeError("WrongOutputLen", expected=12, got=13)
    # a recipe cache of any GF(2^8) geometry is the right scratch kind
    other = Decoder(Profile.resolve(Engine.STANDARD_CAUCHY, 9, 9, 9)).scratch()
    out = np.full(12, 0xAB, np.uint8)
    dec.finalize_into_with(out, other)
    assert np.array_equal(out, np.zeros(12, np.uint8))
    tower = Decoder(Profile.resolve(Engine.TOWER, 3, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(out, tower)
    assert e.value == DecodeError("ScratchMismatch")


def test_decoder_blanket_batch_decode(torch_cuda):
    # codec.rs:203-240 through the streaming decoder
    k, m, L = 5, 3, 40
    rng = np.random.default_rng(4)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    sym = encoded(Engine.GOOD_CAUCHY, k, m, L, data)
    dec = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    out = np.zeros(k * L, np.uint8)
    dec.decode_into([(i, sym[i]) for i in (7, 1, 5, 3, 0)], out)
    assert np.array_equal(out, data)
    with pytest.raises(DecodeError) as e:
        dec.decode_into([(i, sym[i]) for i in (7, 1, 7, 3, 0)], out)
    assert e.value == DecodeError("DuplicateIndex", index=7)


# ---- incremental encoder (Good Cauchy) -------------------------------------

def test_incremental_encoder_engines(torch_cuda):
    # selector.rs:370-381
    IncrementalEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 254, 1, 1))
    with pytest.raises(ConfigError) as e:
        IncrementalEncoder(Profile.resolve(Engine.STANDARD_CAUCHY, 4, 2, 8))
    assert e.value == ConfigError("UnsupportedMode", engine=Engine.STANDARD_CAUCHY)


@pytest.mark.parametrize("k,m,L", [(1, 1, 1), (4, 3, 17), (10, 4, 4096), (32, 16, 65536), (100, 60, 1400),
                                   (40, 33, 4097)])
def test_streaming_matches_batch_repairs(torch_cuda, k, m, L):
    # encoder/streaming.rs:279-300
    rng = np.random.default_rng(k * 101 + m + L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    enc = IncrementalEncoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    for i in rng.permutation(k).tolist():
        enc.feed(i, data[i * L:(i + 1) * L])
    got = np.frombuffer(b"".join(enc.repair(j) for j in range(m)), np.uint8)
    want = oracle.gf256_encode(oracle.good_cauchy_matrix(k, m), k, m, L, data)
    assert np.array_equal(got, want)


def test_reset_reuses_encoder_for_second_block(torch_cuda):
    # encoder/streaming.rs:302-341
    k, m, L = 5, 3, 48
    rng = np.random.default_rng(8)
    enc = IncrementalEncoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    c = oracle.good_cauchy_matrix(k, m)
    for block in range(2):
        data = rng.integers(0, 256, k * L, dtype=np.uint8)
        for i in range(k):
            enc.feed(i, data[i * L:(i + 1) * L])
            partial = data.copy()
            partial[(i + 1) * L:] = 0
            assert enc.repair(m - 1) == oracle.gf256_encode(c, k, m, L, partial)[(m - 1) * L:].tobytes()
        assert enc.fed_count() == k
        enc.reset()
        assert enc.fed_count() == 0
        assert enc.repair(0) == bytes(L)


def test_incremental_feed_errors(torch_cuda):
    # encoder/streaming.rs:162-210
    enc = IncrementalEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 3, 2, 4))
    with pytest.raises(EncodeError) as e:
        enc.feed(4, bytes(1))
    assert e.value == EncodeError("IndexOutOfRange", index=4, n=5)
    with pytest.raises(EncodeError) as e:
        enc.feed(2, bytes(5))
    assert e.value == EncodeError("WrongPayloadLen", expected=4, got=5)
    enc.feed(2, bytes(4))
    with pytest.raises(EncodeError) as e:
        enc.feed(2, bytes(4))
    assert e.value == EncodeError("DuplicateData", index=2)
    with pytest.raises(EncodeError) as e:
        enc.repair(2)
    assert e.value == EncodeError("IndexOutOfRange", index=2, n=5)


def test_feed_dev_and_repairs_dev(torch_cuda):
    torch = torch_cuda
    k, m, L = 40, 12, 8192
    rng = np.random.default_rng(13)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    dev = torch.from_numpy(data).cuda()
    enc = IncrementalEncoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    for i in rng.permutation(k).tolist():
        enc.feed_dev(i, dev.data_ptr() + i * L, L)
    want = oracle.gf256_encode(oracle.good_cauchy_matrix(k, m), k, m, L, data)
    assert b"".join(enc.repair(j) for j in range(m)) == want.tobytes()
    # The device view feeds a streaming decode with the first m data erased.
    addr = enc.repairs_dev()
    dec = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    for i in range(m, k):
        dec.push_dev(i, dev.data_ptr() + i * L, L)
    for j in range(m):
        dec.push_dev(k + j, addr + j * L, L)
    out = torch.zeros(k * L, dtype=torch.uint8, device="cuda")
    dec.finalize_into_dev(out, k * L)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), data)
