"""Pin the CPU oracle's GF((2^8)^2) arithmetic and Tower codec."""
import itertools

import numpy as np

import oracle


def test_known_answers_pin_basis_and_wire_order():
    # algebra/gf65536.rs:221-228
    assert oracle.gf16_mul(0x1234, 0xABCD) == 0x8EE3
    assert oracle.gf16_square(0x0100) == 0x0120
    assert oracle.gf16_inv(0x1234) == 0xBEA3
    assert (0x1234).to_bytes(2, "little") == bytes([0x34, 0x12])


def test_delta_has_absolute_trace_one():
    # algebra/gf65536.rs:211-219
    tr, p = 0, 0x20
    for _ in range(8):
        tr ^= p
        p = oracle.gf256_mul(p, p)
    assert tr == 1


def test_generator_is_primitive():
    # algebra/gf65536.rs:231-236
    for q in (3, 5, 17, 257):
        assert oracle.gf16_pow(0x0108, 65535 // q) != 1
    assert oracle.gf16_pow(0x0108, 65535) == 1


def test_karatsuba_matches_schoolbook_and_axioms():
    rng = np.random.default_rng(11)
    for a, b, c in rng.integers(0, 65536, size=(3000, 3)):
        a, b, c = int(a), int(b), int(c)
        assert oracle.gf16_mul(a, b) == oracle.gf16_mul_schoolbook(a, b)
        assert oracle.gf16_mul(a, b) == oracle.gf16_mul(b, a)
        assert oracle.gf16_mul(oracle.gf16_mul(a, b), c) == oracle.gf16_mul(a, oracle.gf16_mul(b, c))
        assert oracle.gf16_mul(a, b ^ c) == oracle.gf16_mul(a, b) ^ oracle.gf16_mul(a, c)
        if a:
            assert oracle.gf16_mul(a, oracle.gf16_inv(a)) == 1
        assert oracle.gf16_square(a) == oracle.gf16_mul(a, a)


def test_tower_matrix_entries_are_inverse_sums():
    k, m = 7, 5
    c = oracle.tower_cauchy_matrix(k, m)
    for i in range(k):
        for j in range(m):
            s = oracle.gf16_pow(0x0108, i) ^ oracle.gf16_pow(0x0108, k + j)
            assert oracle.gf16_mul(int(c[i, j]), s) == 1


def test_tower_encode_closed_form_and_round_trip():
    rng = np.random.default_rng(2)
    k, m, L = 4, 3, 6
    c = oracle.tower_cauchy_matrix(k, m)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    rep = oracle.tower_encode(c, k, m, L, data)
    d16 = data.view("<u2").reshape(k, L // 2)
    r16 = rep.view("<u2").reshape(m, L // 2)
    for j in range(m):
        for p in range(L // 2):
            acc = 0
            for i in range(k):
                acc ^= oracle.gf16_mul(int(c[i, j]), int(d16[i, p]))
            assert r16[j, p] == acc
    sym = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    for subset in itertools.combinations(range(k + m), k):
        out = oracle.tower_decode_rref(c, k, m, L, [(i, sym[i]) for i in subset])
        assert np.array_equal(out, data), subset
