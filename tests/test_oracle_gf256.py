"""Pin the CPU oracle's GF(2^8) arithmetic and Cauchy codec.

Each check is independent of the formula it pins: known answers from the
reference's own tests, a hardware multiplier (GF2P8MULB), closed forms brute
forced on tiny inputs, and decode(encode(x)) == x over every k-subset.
"""
import itertools

import numpy as np
import pytest

import oracle


# ---- field -----------------------------------------------------------------

def test_known_small_products():
    # algebra/gf256.rs:430-439
    assert oracle.gf256_mul(0x02, 0x02) == 0x04
    assert oracle.gf256_mul(0x02, 0x1B) == 0x36
    assert oracle.gf256_mul(0x80, 0x02) == 0x1B
    assert oracle.gf256_mul(0x03, 0x03) == 0x05


def test_inverse_known_values():
    # algebra/gf256.rs:442-456
    assert oracle.gf256_inv(0x01) == 0x01
    assert oracle.gf256_inv(0x02) == 0x8D
    assert oracle.gf256_inv(0x03) == 0xF6
    assert oracle.gf256_inv(0x00) == 0x00


def test_generator_cycles_through_all_nonzero_elements():
    # algebra/gf256.rs:413-427
    seen = set()
    g = 1
    for _ in range(255):
        seen.add(g)
        g = oracle.gf256_mul(g, 0x03)
    assert g == 1
    assert seen == set(range(1, 256))


def test_exp_log_tables_are_mutual_inverses():
    # algebra/gf256.rs:458-465
    exp, log = oracle.gf256_exp_log()
    for i in range(255):
        assert log[exp[i]] == i
    assert exp[0] == 1 and exp[1] == 3


def test_multiplication_matches_gfni_exhaustively():
    # GF2P8MULB is a hardware multiplier for the same 0x11B field
    # (simd/x86.rs:1402-1425).
    if oracle.gf256_mul_gfni(1, 1) < 0:
        pytest.skip("host CPU lacks GFNI")
    table = oracle.gf256_mul_table()
    for a in range(256):
        for b in range(256):
            assert table[a, b] == oracle.gf256_mul_gfni(a, b)


def test_field_axioms_sampled():
    rng = np.random.default_rng(1)
    t = oracle.gf256_mul_table()
    for a, b, c in rng.integers(0, 256, size=(2000, 3)):
        assert t[a, b] == t[b, a]
        assert t[t[a, b], c] == t[a, t[b, c]]
        assert t[a, b ^ c] == t[a, b] ^ t[a, c]
    for a in range(1, 256):
        assert t[a, oracle.gf256_inv(a)] == 1


# ---- matrices --------------------------------------------------------------

def test_good_cauchy_factorisation_matches_direct_form():
    # matrices/good_cauchy.rs:125-151 vs :94-101
    for k, m in [(1, 1), (3, 2), (10, 4), (32, 16), (128, 64), (254, 1), (1, 254)]:
        assert np.array_equal(oracle.good_cauchy_matrix(k, m), oracle.good_cauchy_matrix(k, m, factored=True))


def test_standard_cauchy_entries_are_inverse_sums():
    c = oracle.std_cauchy_matrix(10, 4)
    for i in range(10):
        for j in range(4):
            assert oracle.gf256_mul(c[i, j], i ^ (10 + j)) == 1


def test_capacities():
    # batch/codec.rs:645-651
    oracle.std_cauchy_matrix(255, 1)
    with pytest.raises(ValueError):
        oracle.std_cauchy_matrix(255, 2)
    oracle.good_cauchy_matrix(254, 1)
    with pytest.raises(ValueError):
        oracle.good_cauchy_matrix(255, 1)


def _det(mat):
    t = oracle.gf256_mul_table()
    a = [list(map(int, r)) for r in mat]
    n = len(a)
    det = 1
    for col in range(n):
        piv = next((r for r in range(col, n) if a[r][col]), None)
        if piv is None:
            return 0
        a[col], a[piv] = a[piv], a[col]
        det = t[det, a[col][col]]
        inv = oracle.gf256_inv(a[col][col])
        for r in range(col + 1, n):
            f = t[a[r][col], inv]
            for c in range(col, n):
                a[r][c] ^= t[f, a[col][c]]
    return det


@pytest.mark.parametrize("engine", [0, 1])
def test_every_square_minor_is_nonsingular(engine):
    # MDS: every square submatrix of a Cauchy matrix is nonsingular
    # (matrices/cauchy.rs:163-200).
    k, m = 5, 4
    c = oracle.gf256_matrix(engine, k, m)
    for r in range(1, min(k, m) + 1):
        for rows in itertools.combinations(range(k), r):
            for cols in itertools.combinations(range(m), r):
                assert _det(c[np.ix_(rows, cols)]) != 0


# ---- encode ----------------------------------------------------------------

def test_reference_repair_vector():
    # batch/codec.rs:631-642: k=2, m=1, L=2, data [1,2,3,4]
    c = oracle.std_cauchy_matrix(2, 1)
    rep = oracle.gf256_encode(c, 2, 1, 2, np.array([1, 2, 3, 4], np.uint8))
    a00, a10 = c[0, 0], c[1, 0]
    assert rep[0] == oracle.gf256_mul(1, a00) ^ oracle.gf256_mul(3, a10)
    assert rep[1] == oracle.gf256_mul(2, a00) ^ oracle.gf256_mul(4, a10)
    # a00 = 1/(0^2), a10 = 1/(1^2)
    assert a00 == oracle.gf256_inv(2) and a10 == oracle.gf256_inv(3)


@pytest.mark.parametrize("engine", [0, 1])
def test_encode_matches_closed_form_per_byte(engine):
    rng = np.random.default_rng(7)
    t = oracle.gf256_mul_table()
    for k, m, L in [(1, 1, 3), (3, 2, 5), (5, 5, 8), (10, 4, 17)]:
        c = oracle.gf256_matrix(engine, k, m)
        data = rng.integers(0, 256, k * L, dtype=np.uint8)
        rep = oracle.gf256_encode(c, k, m, L, data).reshape(m, L)
        d = data.reshape(k, L)
        for j in range(m):
            for p in range(L):
                acc = 0
                for i in range(k):
                    acc ^= int(t[c[i, j], d[i, p]])
                assert rep[j, p] == acc


def test_fast_rows_terms_matches_scalar():
    # The CPU-baseline kernel against the scalar loop at vector-boundary
    # lengths (simd/rows.rs:447).
    rng = np.random.default_rng(3)
    for L in [0, 1, 7, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 1400, 4099]:
        for rows, terms in [(1, 3), (4, 5), (7, 2)]:
            coef = rng.integers(0, 256, (terms, rows), dtype=np.uint8)
            coef[0, :] = [0, 1, 2, 0x53, 0xFF, 7, 9][:rows]
            srcs = [rng.integers(0, 256, L, dtype=np.uint8) for _ in range(terms)]
            base = rng.integers(0, 256, (rows, L), dtype=np.uint8)
            a, b = base.copy(), base.copy()
            oracle.gf256_rows_terms(a, coef, srcs)
            oracle.gf256_rows_terms(b, coef, srcs, fast=True)
            assert np.array_equal(a, b), (L, rows, terms)


def test_encode_many_matches_reference_encode():
    rng = np.random.default_rng(5)
    k, m, L, count = 32, 16, 4096 + 64, 3
    c = oracle.good_cauchy_matrix(k, m)
    data = rng.integers(0, 256, count * k * L, dtype=np.uint8)
    many = oracle.gf256_encode_many(c, k, m, L, data, count, threads=2).reshape(count, m * L)
    for cw in range(count):
        ref = oracle.gf256_encode(c, k, m, L, data[cw * k * L:(cw + 1) * k * L])
        assert np.array_equal(many[cw], ref)


# ---- decode ----------------------------------------------------------------

@pytest.mark.parametrize("engine", [0, 1])
@pytest.mark.parametrize("k,m,L", [(3, 2, 4), (4, 3, 8), (3, 3, 4), (5, 4, 3)])
def test_decode_every_subset_round_trips(engine, k, m, L):
    # batch/codec.rs:709-803
    rng = np.random.default_rng(k * 100 + m)
    c = oracle.gf256_matrix(engine, k, m)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    rep = oracle.gf256_encode(c, k, m, L, data)
    sym = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    for subset in itertools.combinations(range(k + m), k):
        symbols = [(i, sym[i]) for i in subset]
        code, _, _, out = oracle.gf256_decode(c, k, m, L, symbols)
        assert code == 0
        assert np.array_equal(out, data), subset
        rref = oracle.gf256_decode_rref(c, k, m, L, symbols)
        assert np.array_equal(rref, data)
        # Received order does not change the bytes.
        code, _, _, out2 = oracle.gf256_decode(c, k, m, L, symbols[::-1])
        assert code == 0 and np.array_equal(out2, data)


def test_decode_error_order():
    # batch/codec.rs:293-331 and the reference's error tests :805-866
    k, m, L = 3, 2, 4
    c = oracle.std_cauchy_matrix(k, m)
    ok = np.zeros(L, np.uint8)
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok)])[:3] == (30, 3, 2)
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok), (5, ok)])[:3] == (31, 5, 5)
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok), (0, ok)])[:2] == (32, 0)
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok), (2, ok[:3])])[:3] == (33, 4, 3)
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok), (2, ok)], out_len=11)[:3] == (36, 12, 11)
    # Wrong output length is reported before a wrong count.
    assert oracle.gf256_decode(c, k, m, L, [(0, ok)], out_len=11)[0] == 36
    # Index range is checked before the payload length of the same symbol.
    assert oracle.gf256_decode(c, k, m, L, [(0, ok), (1, ok), (9, ok[:1])])[0] == 31


def test_seeded_bytes_numpy_matches_c():
    for seed, off, n in [(1, 0, 100), (0xDEADBEEF, 64, 37), (2**63 + 5, 8 * 1000, 1000)]:
        assert np.array_equal(oracle.fill_bytes(seed, off, n), oracle.splitmix_bytes_np(seed, off, n))


def test_splitmix64_known_answer():
    # splitmix64 from state 0: first output 0xE220A8397B1DCDAF (published
    # reference sequence of Vigna's splitmix64.c).
    b = oracle.fill_bytes(0, 0, 8)
    assert int.from_bytes(bytes(b), "little") == 0xE220A8397B1DCDAF
