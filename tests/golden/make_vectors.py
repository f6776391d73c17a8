"""Write codec_vectors.json: small seeded encode/decode vectors for every
engine, computed by the oracle.

The vectors freeze the oracle's outputs on fixed inputs: the CPU suite checks
that the oracle still reproduces them, the GPU suite runs the same inputs
through the C ABI.  Inputs are the seeded splitmix64 byte stream; erasure
patterns are seeded permutations.  Run from the repository root:

    python tests/golden/make_vectors.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402

ENGINES = {0: "StandardCauchy", 1: "GoodCauchy", 2: "Tower", 3: "Afft"}
SHAPES = {
    0: [(3, 2, 8), (10, 4, 64), (1, 5, 16), (20, 12, 33)],
    1: [(3, 2, 8), (10, 4, 64), (6, 6, 17), (32, 16, 48)],
    2: [(3, 2, 8), (10, 4, 64), (7, 9, 18), (40, 10, 32)],
    3: [(3, 2, 8), (8, 8, 64), (5, 11, 18), (32, 32, 32)],
}


def encode(engine: int, k: int, m: int, L: int, data: np.ndarray) -> np.ndarray:
    if engine in (0, 1):
        return oracle.gf256_encode(oracle.gf256_matrix(engine, k, m), k, m, L, data)
    if engine == 2:
        return oracle.tower_encode(oracle.tower_cauchy_matrix(k, m), k, m, L, data)
    return oracle.afft_encode(k, m, L, data)


def build() -> dict:
    cases = []
    for engine, shapes in SHAPES.items():
        for n, (k, m, L) in enumerate(shapes):
            seed = 0x601D_0000 + 16 * engine + n
            data = oracle.fill_bytes(seed, 0, k * L)
            rep = encode(engine, k, m, L, data)
            rng = np.random.default_rng(seed)
            r = int(rng.integers(1, m + 1))
            erased = sorted(int(x) for x in rng.permutation(k + m)[:r])
            keep = [i for i in range(k + m) if i not in erased]
            received = [int(x) for x in rng.permutation(keep)[:k]]
            cases.append({
                "engine": ENGINES[engine], "k": k, "m": m, "symbol_len": L, "seed": seed,
                "data": data.tobytes().hex(), "repairs": rep.tobytes().hex(),
                "erased": erased, "received": received,
            })
    return {
        "note": "data = splitmix64 byte stream of `seed`; repairs = the engine's systematic encode; "
                "decoding the `received` indices (in that order) must return `data`",
        "cases": cases,
    }


if __name__ == "__main__":
    out = HERE / "codec_vectors.json"
    out.write_text(json.dumps(build(), indent=1) + "\n")
    print(f"wrote {out}")
