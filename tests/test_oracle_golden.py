"""The oracle against the committed codec vectors (tests/golden/codec_vectors.json):
its encode, decode and input stream must reproduce them byte for byte."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

VECTORS = json.loads((Path(__file__).resolve().parent / "golden" / "codec_vectors.json").read_text())["cases"]
ENGINE = {"StandardCauchy": 0, "GoodCauchy": 1, "Tower": 2, "Afft": 3}


def ident(case):
    return f"{case['engine']}-{case['k']}-{case['m']}-{case['symbol_len']}"


def symbols(case, data, rep):
    k, L = case["k"], case["symbol_len"]
    out = []
    for i in case["received"]:
        out.append((i, data[i * L:(i + 1) * L] if i < k else rep[(i - k) * L:(i - k + 1) * L]))
    return out


@pytest.mark.parametrize("case", VECTORS, ids=ident)
def test_oracle_reproduces_vectors(case):
    e, k, m, L = ENGINE[case["engine"]], case["k"], case["m"], case["symbol_len"]
    data = np.frombuffer(bytes.fromhex(case["data"]), np.uint8)
    rep = np.frombuffer(bytes.fromhex(case["repairs"]), np.uint8)
    assert np.array_equal(oracle.fill_bytes(case["seed"], 0, k * L), data)
    if e in (0, 1):
        c = oracle.gf256_matrix(e, k, m)
        assert np.array_equal(oracle.gf256_encode(c, k, m, L, data), rep)
        code, _, _, out = oracle.gf256_decode(c, k, m, L, symbols(case, data, rep))
        assert code == 0 and np.array_equal(out, data)
    elif e == 2:
        c = oracle.tower_cauchy_matrix(k, m)
        assert np.array_equal(oracle.tower_encode(c, k, m, L, data), rep)
        assert np.array_equal(oracle.tower_decode_rref(c, k, m, L, symbols(case, data, rep)), data)
    else:
        assert np.array_equal(oracle.afft_encode(k, m, L, data), rep)
        assert np.array_equal(oracle.afft_decode(k, L, symbols(case, data, rep)), data)


def test_vectors_cover_every_engine_and_erase_data():
    assert {c["engine"] for c in VECTORS} == set(ENGINE)
    for c in VECTORS:
        assert len(c["received"]) == c["k"] and len(set(c["received"])) == c["k"]
        assert not set(c["received"]) & set(c["erased"])
    # Most patterns must lose data symbols, so decode does real work.
    assert sum(any(i < c["k"] for i in c["erased"]) for c in VECTORS) >= len(VECTORS) // 2
