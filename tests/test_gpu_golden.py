"""The HIP library against the committed codec vectors
(tests/golden/codec_vectors.json) through every codec surface of the C ABI."""
import json
from pathlib import Path

import numpy as np
import pytest

from scrs_amd.codec import (BatchDecoder, BatchEncoder, Engine, IncrementalEncoder, Kernel, Profile, decoder)

pytestmark = pytest.mark.gpu

VECTORS = json.loads((Path(__file__).resolve().parent / "golden" / "codec_vectors.json").read_text())["cases"]
ENGINE = {"StandardCauchy": Engine.STANDARD_CAUCHY, "GoodCauchy": Engine.GOOD_CAUCHY, "Tower": Engine.TOWER,
          "Afft": Engine.AFFT}


def ident(case):
    return f"{case['engine']}-{case['k']}-{case['m']}-{case['symbol_len']}"


@pytest.mark.parametrize("case", VECTORS, ids=ident)
def test_library_reproduces_vectors(torch_cuda, case):
    e, k, m, L = ENGINE[case["engine"]], case["k"], case["m"], case["symbol_len"]
    p = Profile.resolve(e, k, m, L)
    data = np.frombuffer(bytes.fromhex(case["data"]), np.uint8)
    rep = np.frombuffer(bytes.fromhex(case["repairs"]), np.uint8)
    syms = [(i, data[i * L:(i + 1) * L] if i < k else rep[(i - k) * L:(i - k + 1) * L]) for i in case["received"]]
    if e != Engine.TOWER:
        enc = BatchEncoder(p)
        assert np.array_equal(enc.encode(data), rep)
        if e != Engine.AFFT and L % 4 == 0 and k <= 64:
            enc.set_kernel(Kernel.BITMATRIX)
            assert np.array_equal(enc.encode(data), rep)
    if e in (Engine.GOOD_CAUCHY, Engine.TOWER):
        inc = IncrementalEncoder(p)
        for i in range(k):
            inc.feed(i, data[i * L:(i + 1) * L])
        assert b"".join(inc.repair(j) for j in range(m)) == rep.tobytes()
    dec = BatchDecoder(p)
    assert np.array_equal(dec.decode(syms), data)
    sd = decoder(p)
    for i, payload in syms:
        sd.push(i, payload)
    assert sd.is_complete()
    assert np.array_equal(sd.finalize(), data)
