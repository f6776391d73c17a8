"""The mixed serving stream at full size through the serving loop.

The seeded trace is the benchmark's: 20,000 requests, 70% encode / 30%
decode, engine uniform over {Standard Cauchy, Good Cauchy, Tower}, k ~
U[4,200], m ~ U[1,32], symbol_len from {512, 1500, 4096, 65536}, r ~ U[1,m]
erasures per decode.  Every decode reads a real codeword.  The same trace
runs through the server in batched mode (continuous batches, Poisson
arrivals) and in plain mode (one codec call per request), each into its own
output set; then

  * every decode output equals its codeword's data (both modes),
  * every output of the batched run equals the plain run's, byte for byte,
  * a seeded sample of encode jobs of every engine (at least 210) and of the
    pre-encoded decode codewords equals the CPU oracle.
"""
import numpy as np
import pytest

import bench
from scrs_amd import Server

pytestmark = pytest.mark.gpu

N_REQ = 20000


@pytest.fixture(scope="module")
def serving(torch_cuda):
    specs = bench.serving_trace(N_REQ)
    setup = bench.ServingSetup(specs, sets=2)
    setup.encode_codewords()
    yield setup
    del setup
    torch_cuda.cuda.empty_cache()


def test_trace_is_the_full_mix(serving):
    specs = serving.specs
    assert len(specs) == N_REQ
    enc = sum(s[0] for s in specs)
    assert 0.68 * N_REQ < enc < 0.72 * N_REQ
    for eng in range(3):
        assert sum(1 for s in specs if s[1] == eng and s[0]) > 4000
        assert sum(1 for s in specs if s[1] == eng and not s[0]) > 1500
    for sl in bench.SERVE_LENS:
        assert sum(1 for s in specs if s[4] == sl and not s[0]) > 1000, sl
    assert max(s[2] for s in specs) == 200 and min(s[2] for s in specs) == 4
    assert max(s[3] for s in specs) == 32 and min(s[3] for s in specs) == 1


def _replay(srv, jobs, rate=None):
    n = len(jobs)
    if rate is None:
        arr = np.zeros(n, np.uint64)
    else:
        arr = np.cumsum(np.random.default_rng(0xA55).exponential(1e9 / rate, size=n)).astype(np.uint64)
    res = srv.replay(Server.prepare(jobs), arr)
    bad = np.flatnonzero(res["code"] != 0)
    assert bad.size == 0, f"{bad.size} requests failed, first {int(bad[0])}: status {int(res['code'][bad[0]])}"
    return res


@pytest.fixture(scope="module")
def plain_outputs(serving, torch_cuda):
    """Output set 1: the trace through the server in plain mode, one codec
    call per request, every request alone."""
    plain = Server(batched=False)
    try:
        _replay(plain, serving.jobs(1))
        torch_cuda.cuda.synchronize()
    finally:
        plain.close()
    return serving


@pytest.mark.timeout(900)
def test_plain_mode_decodes_and_encodes_match_the_oracle(plain_outputs):
    serving = plain_outputs
    assert serving.decodes_ok(1), "a plain decode did not reconstruct its codeword's data"
    assert serving.sample_encodes(1, per_engine=70) >= 210


@pytest.mark.timeout(600)
def test_batched_mode_equals_plain_mode(plain_outputs, torch_cuda):
    """Poisson arrivals: batches form while earlier ones run on the device;
    every output equals the same request run alone."""
    serving, torch = plain_outputs, torch_cuda
    serving.clear(0)
    batched = Server(batched=True, max_batch_jobs=1024, use_graph=False, cache_bytes=16 << 30)
    try:
        _replay(batched, serving.jobs(0), rate=200_000)
        torch.cuda.synchronize()
        stats = batched.stats()
        assert stats["batches"] > 1 and stats["jobs"] == N_REQ
    finally:
        batched.close()
    assert serving.decodes_ok(0), "a batched decode did not reconstruct its codeword's data"
    assert serving.sets_equal(0, 1), "batched and plain outputs differ"


@pytest.mark.timeout(600)
def test_saturated_batches_equal_plain_mode(plain_outputs, torch_cuda):
    """All requests at once (largest batches, autotuned tiles), replayed
    twice on one server so the second pass reuses cached tables and recipes."""
    serving, torch = plain_outputs, torch_cuda
    serving.clear(0)
    batched = Server(batched=True, max_batch_jobs=4096, use_graph=True, cache_bytes=16 << 30)
    try:
        for _ in range(2):
            _replay(batched, serving.jobs(0))
            torch.cuda.synchronize()
            assert serving.decodes_ok(0)
            assert serving.sets_equal(0, 1)
            serving.clear(0)
    finally:
        batched.close()
