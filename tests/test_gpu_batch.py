"""Fused uneven batches and the serving loop, against the per-codeword path.

Every job's output from a mixed batch (GF(256) Standard and Good Cauchy and
GF(2^16) Tower, uneven k, m and symbol_len, encode and decode together) must
equal the output of the same job run alone through the codec calls, which
are pinned to the oracle by the other GPU suites; a sample is also checked
against the oracle directly.
"""
import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, BatchEngine, ConfigError, DecodeError, Engine,
                            IncrementalEncoder, Job, Profile, Server, fill_random_dev)

pytestmark = pytest.mark.gpu

ENGINES = [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY, Engine.TOWER]


def rand_geometry(rng, engine, lens=(512, 1500, 4096, 65536), kmax=200, mmax=32):
    k = int(rng.integers(4, kmax + 1))
    m = int(rng.integers(1, mmax + 1))
    if engine == Engine.GOOD_CAUCHY:
        k = min(k, 255 - m)
    L = int(rng.choice(lens))
    return Profile.resolve(engine, k, m, L)


def encode_alone(torch, p, data):
    rep = torch.empty(p.m * p.symbol_len, dtype=torch.uint8, device="cuda")
    if p.engine == Engine.TOWER:
        IncrementalEncoder(p).encode_batch_dev(data, rep, 1)
    else:
        BatchEncoder(p).encode_batch_dev(data, rep, 1)
    torch.cuda.synchronize()
    return rep


class Workload:
    """A seeded mix of encode and decode jobs with their inputs on the device."""

    def __init__(self, torch, seed, count, lens=(512, 1500, 4096, 65536), decode_frac=0.3):
        rng = np.random.default_rng(seed)
        self.torch = torch
        self.items = []
        for i in range(count):
            engine = ENGINES[int(rng.integers(0, 3))]
            p = rand_geometry(rng, engine, lens)
            k, m, L = p.k, p.m, p.symbol_len
            data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
            fill_random_dev(data, data.numel(), seed * 1000 + i)
            if rng.random() < decode_frac:
                rep = encode_alone(torch, p, data)
                r = int(rng.integers(1, m + 1))
                erased = set(rng.choice(k + m, size=r, replace=False).tolist())
                survivors = [x for x in range(k + m) if x not in erased]
                idx = [int(x) for x in rng.permutation(survivors)[:k]]
                ptrs = [data.data_ptr() + x * L if x < k else rep.data_ptr() + (x - k) * L for x in idx]
                out = torch.full((k * L,), 0xAB, dtype=torch.uint8, device="cuda")
                self.items.append(dict(job=Job.decode(p, idx, ptrs, out), data=data, rep=rep, out=out, p=p))
            else:
                out = torch.full((m * L,), 0xAB, dtype=torch.uint8, device="cuda")
                self.items.append(dict(job=Job.encode(p, data, out), data=data, out=out, p=p))
        torch.cuda.synchronize()
        # The per-codeword answers.
        self.want = []
        for it in self.items:
            if it["job"].op == "encode":
                self.want.append(encode_alone(torch, it["p"], it["data"]).clone())
            else:
                self.want.append(it["data"].clone())

    @property
    def jobs(self):
        return [it["job"] for it in self.items]

    def clear(self):
        for it in self.items:
            it["out"].fill_(0xAB)
        self.torch.cuda.synchronize()

    def check(self):
        for i, it in enumerate(self.items):
            assert self.torch.equal(it["out"], self.want[i]), (i, it["job"].op, it["p"])


def test_mixed_batch_matches_alone(torch_cuda):
    torch = torch_cuda
    w = Workload(torch, 1, 60)
    eng = BatchEngine(tile_bytes=4096, cw_per_unit=1, use_graph=False)
    eng.submit(w.jobs)
    torch.cuda.synchronize()
    w.check()
    st = eng.stats()
    assert st["batches"] == 1 and st["jobs"] == 60
    # A sample against the oracle directly.
    for it in w.items[:12]:
        p = it["p"]
        if it["job"].op != "encode" or p.symbol_len > 4096:
            continue
        d = it["data"].cpu().numpy()
        if p.engine == Engine.TOWER:
            want = oracle.tower_encode(oracle.tower_cauchy_matrix(p.k, p.m), p.k, p.m, p.symbol_len, d)
        else:
            want = oracle.gf256_encode(oracle.gf256_matrix(int(p.engine), p.k, p.m), p.k, p.m, p.symbol_len, d)
        assert np.array_equal(it["out"].cpu().numpy(), want)


@pytest.mark.parametrize("tile,cw", [(1024, 1), (2048, 2), (4096, 8)])
def test_tile_shapes_and_codewords_per_unit(torch_cuda, tile, cw):
    torch = torch_cuda
    # Repeated encode geometries so several codewords share one unit.
    p = Profile.resolve(Engine.GOOD_CAUCHY, 20, 6, 1500)
    q = Profile.resolve(Engine.TOWER, 12, 5, 4096)
    jobs, outs, wants = [], [], []
    for i in range(12):
        prof = p if i % 3 else q
        data = torch.empty(prof.k * prof.symbol_len, dtype=torch.uint8, device="cuda")
        fill_random_dev(data, data.numel(), 77 + i)
        out = torch.zeros(prof.m * prof.symbol_len, dtype=torch.uint8, device="cuda")
        jobs.append(Job.encode(prof, data, out))
        outs.append(out)
        wants.append(encode_alone(torch, prof, data))
    w = Workload(torch, 3, 20, lens=(512, 1500, 4096))
    eng = BatchEngine(tile_bytes=tile, cw_per_unit=cw, use_graph=True)
    eng.submit(jobs + w.jobs)
    torch.cuda.synchronize()
    for o, wv in zip(outs, wants):
        assert torch.equal(o, wv)
    w.check()


def test_graph_replays_match_eager(torch_cuda):
    torch = torch_cuda
    w = Workload(torch, 5, 40, lens=(512, 1500, 4096))
    eager = BatchEngine(tile_bytes=2048, cw_per_unit=1, use_graph=False)
    eager.submit(w.jobs)
    torch.cuda.synchronize()
    w.check()
    eng = BatchEngine(tile_bytes=2048, cw_per_unit=1, use_graph=True)
    scratch = torch.zeros(1 << 20, dtype=torch.float32, device="cuda")
    for rep in range(4):
        w.clear()
        eng.submit(w.jobs)
        scratch.mul_(0.5).add_(1.0)  # the program's own eager work between replays
        torch.cuda.synchronize()
        w.check()
    st = eng.stats()
    assert st["batches"] == 4
    # Recipes solved in the first replay are reused by the later ones.
    assert st["recipe_misses"] > 0 and st["recipe_hits"] == 3 * st["recipe_misses"]


def test_autotuner_picks_a_configuration(torch_cuda):
    """Online tuning: successive batches of one signature each run the next
    candidate configuration (every output correct), and once every candidate
    has been timed the fastest is kept for the later batches."""
    torch = torch_cuda
    w = Workload(torch, 7, 40, lens=(512, 4096))
    eng = BatchEngine(tile_bytes=0, cw_per_unit=0)
    seen = set()
    for _ in range(30):  # at most 3 tiles x 4 codewords per unit, two runs each
        w.clear()
        eng.submit(w.jobs)
        torch.cuda.synchronize()
        w.check()
        st = eng.stats()
        seen.add((st["tile_bytes"], st["cw_per_unit"]))
        if st["tuned_signatures"] == 1:
            break
    assert st["tuned_signatures"] == 1 and len(seen) > 1
    chosen = []
    for _ in range(3):  # the same signature: the tuned configuration, no more tuning
        w.clear()
        eng.submit(w.jobs)
        torch.cuda.synchronize()
        w.check()
        st = eng.stats()
        chosen.append((st["tile_bytes"], st["cw_per_unit"]))
    assert st["tuned_signatures"] == 1 and len(set(chosen)) == 1
    assert chosen[0][0] in (1024, 2048, 4096)


def test_validation_rejects_before_enqueue(torch_cuda):
    torch = torch_cuda
    w = Workload(torch, 9, 6, lens=(512,), decode_frac=0.0)
    p = Profile.resolve(Engine.STANDARD_CAUCHY, 3, 2, 512)
    out = torch.full((3 * 512,), 0xAB, dtype=torch.uint8, device="cuda")
    src = torch.zeros(5 * 512, dtype=torch.uint8, device="cuda")
    bad = Job.decode(p, [0, 1, 1], [src.data_ptr()] * 3, out)
    eng = BatchEngine(tile_bytes=1024, cw_per_unit=1)
    with pytest.raises(DecodeError) as e:
        eng.submit(w.jobs + [bad])
    assert e.value == DecodeError("DuplicateIndex", index=1)
    assert e.value.job == 6
    torch.cuda.synchronize()
    for it in w.items:  # nothing ran
        assert bool((it["out"] == 0xAB).all())
    afft = Job.encode(Profile.resolve(Engine.AFFT, 4, 2, 512), src, out)
    with pytest.raises(ConfigError) as e:
        eng.submit([afft])
    assert e.value == ConfigError("UnsupportedMode", engine=Engine.AFFT)


@pytest.mark.parametrize("batched", [True, False])
def test_server_results_equal_alone(torch_cuda, batched):
    torch = torch_cuda
    w = Workload(torch, 11, 80, lens=(512, 1500, 4096))
    srv = Server(batched=batched, max_batch_jobs=32, tile_bytes=2048, cw_per_unit=1)
    arrivals = np.cumsum(np.random.default_rng(11).exponential(20_000, size=80)).astype(np.uint64)
    res = srv.replay(w.jobs, arrivals)
    torch.cuda.synchronize()
    assert (res["code"] == 0).all()
    assert (res["t_done_ns"] >= res["t_start_ns"]).all() and (res["t_start_ns"] >= res["t_submit_ns"]).all()
    w.check()
    srv.close()


def test_server_admission_budget(torch_cuda):
    torch = torch_cuda
    # A budget smaller than two jobs' footprint still completes every job.
    w = Workload(torch, 13, 16, lens=(4096,), decode_frac=0.5)
    srv = Server(batched=True, hbm_budget=1, tile_bytes=4096, cw_per_unit=1)
    for i, j in enumerate(w.jobs):
        srv.submit(j, i)
    got = []
    while len(got) < 16:
        got += srv.poll(wait=True)
    assert sorted(c["tag"] for c in got) == list(range(16))
    assert all(c["code"] == 0 for c in got)
    torch.cuda.synchronize()
    w.check()
    srv.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_lone_large_jobs_split_terms(torch_cuda, engine):
    # A batch far too small to fill the device splits each job's term loop
    # across units that store partial rows; a reduce launch combines them.
    torch = torch_cuda
    rng = np.random.default_rng(int(engine) + 40)
    m = 32 if engine != Engine.GOOD_CAUCHY else 30
    p = Profile.resolve(engine, 200 if engine != Engine.GOOD_CAUCHY else 220, m, 65536)
    k, L = p.k, p.symbol_len
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 4242 + int(engine))
    want_rep = encode_alone(torch, p, data)
    eng = BatchEngine(tile_bytes=4096, cw_per_unit=1)
    rep = torch.full((m * L,), 0xAB, dtype=torch.uint8, device="cuda")
    eng.submit([Job.encode(p, data, rep)])
    torch.cuda.synchronize()
    assert torch.equal(rep, want_rep)
    erased = set(rng.choice(k, size=m, replace=False).tolist())
    idx = [int(x) for x in rng.permutation([x for x in range(k + m) if x not in erased])[:k]]
    ptrs = [data.data_ptr() + x * L if x < k else rep.data_ptr() + (x - k) * L for x in idx]
    out = torch.full((k * L,), 0xAB, dtype=torch.uint8, device="cuda")
    for _ in range(2):  # second pass reuses the recipe
        out.fill_(0xAB)
        eng.submit([Job.decode(p, idx, ptrs, out)])
        torch.cuda.synchronize()
        assert torch.equal(out, data)


@pytest.mark.parametrize("use_graph", [False, True])
def test_unit_counter_across_many_batches(torch_cuda, use_graph):
    """The multiply-XOR draws its units from a counter that only grows: each
    launch starts from the value the host recorded after the previous one.
    Thirty batches of 1, 5 and 40 jobs (split and unsplit ranges) through one
    engine, every output checked after each batch."""
    torch = torch_cuda
    loads = [Workload(torch, 50 + i, n, lens=(512, 1500, 4096)) for i, n in enumerate((1, 5, 40))]
    eng = BatchEngine(tile_bytes=4096, cw_per_unit=1, use_graph=use_graph)
    for b in range(30):
        w = loads[b % 3]
        w.clear()
        eng.submit(w.jobs)
        torch.cuda.synchronize()
        w.check()
    assert eng.stats()["batches"] == 30
