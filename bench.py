"""Benchmark: GF(256) Good Cauchy block encode + erasure decode on one MI355X.

Workload (one step): 4096 independent codewords, k=32, m=16, symbol_len=65536
(8 GiB of data resident in HBM).  Encode all repairs, then decode every
codeword from the k lowest surviving indices after erasing a seeded uniform
subset of 1..16 of its 48 indices.  Throughput counts algorithmic bytes:
(k+m)*L per encoded codeword plus 2*k*L per decoded codeword.

Prints one JSON line: whole-job throughput, per-phase kernel times measured
with HIP events on the launch stream, the encode launch against the HBM
roofline, and the CPU port of the reference algorithm timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

K, M, L, COUNT = 32, 16, 65536, 4096
SEED_DATA = 0x5C25_0002
SEED_ERASE = 0x5C25_E7A5
HBM_PEAK_GBS = 8000.0
INT8_PEAK_OPS = 5.0e15  # MI355X dense int8 MFMA
FP4_PEAK_FLOPS = 10.1e15  # MI355X dense fp4 / fp6 MFMA
# LDS read bandwidth: 256 CUs x 232 B/clk (ds_read_b128, measured) x 2.4 GHz.
LDS_PEAK_TBS = 256 * 232 * 2.4e9 / 1e12
# Scattered 2-byte table reads: an LDS unit serves 32 lanes per clock
# without bank conflicts (SQ_LDS_IDX_ACTIVE = 2 cycles per wave64 read),
# 256 CUs x 32 x 2.4 GHz.
LDS_LOOKUP_PEAK_TPS = 256 * 32 * 2.4e9 / 1e12


def afft_table_reads(k: int, m: int, sl: int) -> tuple[int, int]:
    """Table reads (one 2-byte LDS read per lane) of the additive-FFT encode
    and erasure decode at k = m = 2^D/2: a butterfly multiplies each 16-bit
    element of its high row by the node constant, 3 reads with the 6+5+5-bit
    tables; the 7-level encode pass (inverse and forward tables) keeps nibble
    tables, 4 reads, on its first level so that it fits two workgroups per
    CU; a per-row scale is 4 reads per element.  Encode: inverse over the
    2^Dk data rows and forward over the coset, in passes of (8, Dk-8)
    levels.  Decode over N = k + m rows: inverse in two 8-level passes
    (input scale, then the T(q) scale), the XOR-only derivative pass, the
    forward over the high bits (1/T scale) and the forward over the low bits
    of the rows below k (output scale).
    The first node of every level has factor 0: in the top stage of the
    top pass (its upper 4 levels, 3 in a 7-level pass) of a transform that
    starts at point 0, the 2^B butterflies of level B on that node skip
    their product.  The encode's inverse and the decode's inverse and
    forward start at 0; the encode's forward starts at k.  The decode's
    forward over the high bits needs only the rows below k = N/2: below
    its top level it runs on that half alone, where levels 15 and 14 lie on
    node 0 and levels 13 and 12 on nodes 0-1 and 0-3."""
    assert k == m and k & (k - 1) == 0 and 2 * k == 1 << 16
    el = sl // 2
    per = 3
    dk = k.bit_length() - 1
    skipped = lambda top, levels: sum(1 << b for b in range(top - levels, top))  # noqa: E731
    mid = 4 + (dk - 9) * per  # the 7-level pass: nibble first level, wide above
    enc = (8 * per + 2 * mid + 8 * per) * (k // 2) * el
    enc -= skipped(dk, 3) * per * el
    n = 2 * k
    inv = 2 * 8 * per * (n // 2) * el + 2 * n * el * 4 - skipped(16, 4) * per * el
    q = n // 4  # butterflies per level among the rows below k
    fwd_hi = (q // 2 + 3 * q // 4 + 4 * q) * per * el + n * el * 4
    fwd_lo = 8 * per * (k // 2) * el + k * el * 4
    dec = inv + fwd_hi + fwd_lo
    return enc, dec


def erasure_indices(count: int, k: int, m: int, max_r: int, seed: int) -> np.ndarray:
    """Per codeword: erase a uniform subset of r ~ U{1..max_r} indices of n,
    keep the k lowest survivors (ascending)."""
    rng = np.random.default_rng(seed)
    n = k + m
    idx = np.empty((count, k), np.uint16)
    for c in range(count):
        r = int(rng.integers(1, max_r + 1))
        erased = rng.permutation(n)[:r]
        keep = np.ones(n, bool)
        keep[erased] = False
        idx[c] = np.flatnonzero(keep)[:k]
    return idx


def cpu_threads() -> tuple[int, dict]:
    """Threads for the CPU baselines: the CPU share this process may use
    (its affinity set, capped by OMP_NUM_THREADS where the host sets one, as
    a GPU box does per GPU), and a description of the host."""
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(share, int(omp)) if omp.isdigit() and int(omp) > 0 else share
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, {"cpu_model": model, "host_cpus": os.cpu_count(), "affinity_cpus": share,
                     "omp_num_threads": omp or None}


def _timed_reps(fn, target_s: float, min_reps: int = 5) -> tuple[float, float, int]:
    """Run fn() once to warm up, then at least min_reps times and until
    target_s has passed; returns (min, median, reps) seconds per call."""
    fn()
    times = []
    t_end = time.perf_counter() + target_s
    while len(times) < min_reps or time.perf_counter() < t_end:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return min(times), float(np.median(times)), len(times)


def cpu_baseline(idx: np.ndarray, rep_gpu_sample: np.ndarray, target_s: float = 10.0) -> dict:
    """The oracle's register-blocked GFNI port, timed on this host on a
    bounded sample of the same workload; its outputs are checked against the
    GPU's for the sampled codewords."""
    import oracle

    threads, host = cpu_threads()
    sample = max(16, threads)
    coeffs = oracle.good_cauchy_matrix(K, M)
    data = oracle.splitmix_bytes_np(SEED_DATA, 0, sample * K * L)
    rep = oracle.gf256_encode_many(coeffs, K, M, L, data, sample, threads)
    out = oracle.gf256_decode_many(coeffs, K, M, L, idx[:sample], data, rep, sample, threads)
    if not np.array_equal(rep[:rep_gpu_sample.size], rep_gpu_sample):
        raise SystemExit("CPU and GPU repairs differ on the sampled codewords")
    if not np.array_equal(out, data):
        raise SystemExit("CPU decode did not round-trip")

    def step():
        r = oracle.gf256_encode_many(coeffs, K, M, L, data, sample, threads)
        oracle.gf256_decode_many(coeffs, K, M, L, idx[:sample], data, r, sample, threads)

    t_min, t_med, reps = _timed_reps(step, target_s)
    nbytes = sample * ((K + M) * L + 2 * K * L)
    # The reference is single-threaded: one core on a smaller sample too.
    t1, _, _ = _timed_reps(lambda: (oracle.gf256_decode_many(
        coeffs, K, M, L, idx[:2], data[: 2 * K * L],
        oracle.gf256_encode_many(coeffs, K, M, L, data[: 2 * K * L], 2, 1), 2, 1)), 1.0, 3)
    plan = {2: "GFNI+AVX2", 1: "AVX2 nibble", 0: "scalar"}[oracle.cpu_plan()]
    return {
        "value": round(nbytes / t_min / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} codewords of the workload, encode + decode, one codeword per thread ({plan}); "
                  f"best of {reps} runs (median {round(nbytes / t_med / 1e9, 3)} GB/s)",
        "single_core_value": round(2 * ((K + M) * L + 2 * K * L) / t1 / 1e9, 3),
        **host,
    }


def timed_phases(steps, stream, phases):
    """Run `phases` (list of callables) `steps` times with HIP events between
    them on `stream`; returns (wall seconds, per-phase mean ms)."""
    import torch

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(phases) + 1)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for row in ev:
        row[0].record(stream)
        for i, ph in enumerate(phases):
            ph()
            row[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [sum(r[i].elapsed_time(r[i + 1]) for r in ev) / steps for i in range(len(phases))]
    return wall, per


def tower_cpu_baseline(k, m, sl, idx, data_gpu, rep_gpu, out_gpu, target_s: float = 10.0) -> dict:
    """The reference's Tower CPU path, ported (oracle/scrs_cpu16.c): the
    GFNI byte-swap rows kernel for the incremental feed and the
    rational-Lagrange finalize, one codeword per thread, timed on a bounded
    sample of the workload's codewords; the sample's repairs and decoded data
    are checked against the GPU's."""
    import oracle

    threads, host = cpu_threads()
    sample = max(8, threads)
    data = data_gpu[: sample * k * sl].cpu().numpy()
    coeffs = oracle.tower_cauchy_matrix(k, m)
    rep = oracle.tower_encode_port(coeffs, k, m, sl, data, count=sample, threads=threads)
    if not np.array_equal(rep, rep_gpu[: sample * m * sl].cpu().numpy()):
        raise SystemExit("CPU and GPU Tower repairs differ on the sampled codewords")
    out = oracle.tower_finalize_port(k, m, sl, idx[:sample], data, rep, count=sample, threads=threads)
    if out is None or not np.array_equal(out, out_gpu[: sample * k * sl].cpu().numpy()):
        raise SystemExit("CPU and GPU Tower finalize outputs differ on the sampled codewords")

    def step():
        r = oracle.tower_encode_port(coeffs, k, m, sl, data, count=sample, threads=threads)
        oracle.tower_finalize_port(k, m, sl, idx[:sample], data, r, count=sample, threads=threads)

    t_min, t_med, reps = _timed_reps(step, target_s)
    nbytes = sample * ((k + m) * sl + 2 * k * sl)
    plan = "GFNI+AVX2" if oracle.cpu16_plan() == 2 else "scalar"
    return {"value": round(nbytes / t_min / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{sample} of the workload's codewords, incremental feed + finalize, one codeword per "
                      f"thread ({plan}); best of {reps} runs (median {round(nbytes / t_med / 1e9, 3)} GB/s)",
            **host}


def afft_cpu_baseline(k, m, sl, idx, data_gpu, rep_gpu, target_s: float = 10.0) -> dict:
    """The reference's additive-FFT CPU path, ported (oracle/scrs_cpu16.c):
    the strip-blocked fused encode (strips shared out over the threads) and
    the complete erasure-locator decode (column slices over the threads), on
    the whole workload; outputs are checked against the GPU's."""
    import oracle

    threads, host = cpu_threads()
    data = data_gpu.cpu().numpy()
    rep_g = rep_gpu.cpu().numpy()
    rep, width = oracle.afft_encode_port(k, m, sl, data, threads=threads)
    if not np.array_equal(rep, rep_g):
        raise SystemExit("CPU and GPU additive-FFT repairs differ")
    sym = [(i, data[i * sl:(i + 1) * sl] if i < k else rep_g[(i - k) * sl:(i - k + 1) * sl])
           for i in (int(x) for x in idx[0])]
    out = oracle.afft_decode_port(k, m, sl, sym, threads=threads)
    if not np.array_equal(out, data):
        raise SystemExit("CPU additive-FFT decode did not reconstruct the data")
    e_min, e_med, e_reps = _timed_reps(lambda: oracle.afft_encode_port(k, m, sl, data, threads=threads, out=rep),
                                       target_s / 2)
    d_min, d_med, d_reps = _timed_reps(lambda: oracle.afft_decode_port(k, m, sl, sym, threads=threads),
                                       target_s / 2)
    nbytes = (k + m) * sl + 2 * k * sl
    plan = "GFNI+AVX2" if oracle.cpu16_plan() == 2 else "scalar"
    return {"value": round(nbytes / (e_min + d_min) / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload: fused strip-blocked encode ({width}-byte strips, strips over the "
                      f"threads) + complete decode (column slices over the threads), {plan}; best of "
                      f"{e_reps}/{d_reps} runs",
            "encode_ms": round(e_min * 1e3, 2), "decode_ms": round(d_min * 1e3, 2),
            "median_encode_ms": round(e_med * 1e3, 2), "median_decode_ms": round(d_med * 1e3, 2), **host}


def gf16_workload(args, which: str | None = None) -> dict:
    """GF(2^16) lines: Tower k=200 m=50 symbol_len=8192 x 256 codewords
    (incremental feed of all k, then finalize with exactly 50 data symbols
    erased), or AFFT k=m=32768 symbol_len=2048 (one 128 MiB codeword, decode
    from a seeded random half of the indices)."""
    import torch

    from scrs_amd import Engine, Profile, batch_decoder, batch_encoder, fill_random_dev, incremental_encoder

    which = which or args.workload
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    if which == "tower":
        k, m, sl, count = 200, 50, 8192, 256
        prof = Profile.resolve(Engine.TOWER, k, m, sl)
        enc = incremental_encoder(prof)
        rng = np.random.default_rng(SEED_ERASE)
        idx = np.empty((count, k), np.uint16)
        for c in range(count):
            erased = set(rng.choice(k, size=m, replace=False).tolist())
            idx[c] = [i for i in range(k + m) if i not in erased]
        desc = ("GF(65536) Tower k=200 m=50 symbol_len=8192, 256 codewords: incremental feed of all k, "
                "finalize with exactly 50 data symbols erased (50x50 inverse per codeword on device)")
    else:
        k, m, sl, count = 32768, 32768, 2048, 1
        prof = Profile.resolve(Engine.AFFT, k, m, sl)
        enc = batch_encoder(prof)
        rng = np.random.default_rng(SEED_ERASE)
        idx = np.sort(rng.choice(k + m, size=k, replace=False)).astype(np.uint16)[None, :]
        desc = ("GF(65536) additive FFT k=32768 m=32768 symbol_len=2048, one 128 MiB codeword: block encode, "
                "decode from a seeded random half of the 65536 indices")
    dec = batch_decoder(prof)
    data = torch.empty(count * k * sl, dtype=torch.uint8, device="cuda")
    rep = torch.empty(count * m * sl, dtype=torch.uint8, device="cuda")
    out = torch.empty_like(data)
    fill_random_dev(data, data.numel(), SEED_DATA, 0, stream)
    phases = [lambda: enc.encode_batch_dev(data, rep, count, stream),
              lambda: dec.decode_batch_dev(idx, data, rep, out, count, stream)]
    for _ in range(args.warmup):
        for ph in phases:
            ph()
    wall, (enc_ms, dec_ms) = timed_phases(args.steps, stream, phases)
    if not torch.equal(out, data):
        raise SystemExit("decode output differs from the original data")
    enc_bytes, dec_bytes = count * (k + m) * sl, count * 2 * k * sl
    targeted = None
    if which == "afft":
        targeted = afft_targeted_line(args, dec, data, rep, out, stream, k, m)
    lds_roof = None
    if which == "afft":
        enc_reads, dec_reads = afft_table_reads(k, m, sl)
        enc_tps = enc_reads / (enc_ms / 1e3) / 1e12
        dec_tps = dec_reads / (dec_ms / 1e3) / 1e12
        lds_roof = {"kernel": "afft_pass_kernel, encode phase", "achieved": round(enc_tps, 3),
                    "peak": round(LDS_LOOKUP_PEAK_TPS, 2), "unit": "T table reads/s",
                    "frac": round(enc_tps / LDS_LOOKUP_PEAK_TPS, 4), "reads_per_launch": enc_reads,
                    "decode": {"achieved": round(dec_tps, 3), "frac": round(dec_tps / LDS_LOOKUP_PEAK_TPS, 4),
                               "reads_per_launch": dec_reads}}
    if which == "tower":
        # The feed reads, per element and term, four nibble-table entries of
        # each 16-row group: 2 bytes per row of the group, 8 bytes for groups
        # of at most 4 rows and 16 for at most 8 (the staged entry sizes).
        group_rows = [min(16, m - g) for g in range(0, m, 16)]
        entry = sum(8 if r <= 4 else (16 if r <= 8 else 32) for r in group_rows)
        lds_bytes = count * k * (sl // 2) * 4 * entry
        lds_tbs = lds_bytes / (enc_ms / 1e3) / 1e12
        lds_roof = {"kernel": "rows_terms16_kernel, feed", "achieved": round(lds_tbs, 2),
                    "peak": LDS_PEAK_TBS, "unit": "TB/s", "frac": round(lds_tbs / LDS_PEAK_TBS, 4),
                    "lds_bytes_per_launch": lds_bytes}
    cpu = None
    if not args.no_cpu:
        torch.cuda.synchronize()
        cpu = (tower_cpu_baseline(k, m, sl, idx, data, rep, out) if which == "tower"
               else afft_cpu_baseline(k, m, sl, idx, data, rep))
    value = (enc_bytes + dec_bytes) / (wall / args.steps) / 1e9
    enc_gbs = enc_bytes / (enc_ms / 1e3) / 1e9
    dec_gbs = dec_bytes / (dec_ms / 1e3) / 1e9
    return {
        "metric": "encode/decode GB/s (bit-exact), % HBM roofline; batched vs plain GPU; p99 latency",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic (seeded splitmix64 bytes in HBM; seeded erasures)",
        "config": {"workload": desc, "codewords": count, "k": k, "m": m, "symbol_len": sl},
        "encode_gbs": round(enc_gbs, 2), "decode_gbs": round(dec_gbs, 2),
        "encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
        "roofline": {"bound": "hbm", "kernel": "encode phase", "achieved": round(enc_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(enc_gbs / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     # GF(2^16) work is bound by LDS table lookups, not HBM: the
                     # fraction is of the HBM peak for comparison only.
                     "limiter": "lds"},
        "lds_roofline": lds_roof,
        "targeted_decode": targeted,
        "cpu_baseline": cpu,
    }


def afft_targeted_line(args, dec, data, rep, out, stream, k, m) -> dict:
    """AFFT decodes of 5 missing data symbols (the reference's
    TARGETED_MAX_MISSING) through the targeted path and, for the same
    pattern, through the transforms; both outputs are checked."""
    import torch

    from scrs_amd.codec import afft_set_targeted_max

    rng = np.random.default_rng(SEED_ERASE + 5)
    r = 5
    missing = set(rng.choice(k, size=r, replace=False).tolist())
    repairs = sorted(rng.choice(m, size=r, replace=False).tolist())
    idx = np.array([i for i in range(k) if i not in missing] + [k + j for j in repairs], np.uint16)[None, :]
    res = {}
    prev = afft_set_targeted_max(5)
    try:
        for name, limit in (("targeted_ms", 5), ("transforms_ms", 0)):
            afft_set_targeted_max(limit)
            # The buffer already holds the data from the main decode: poison
            # it so that this variant's own output is what gets compared.
            out.fill_(0xAB)
            phase = [lambda: dec.decode_batch_dev(idx, data, rep, out, 1, stream)]
            for _ in range(args.warmup):
                phase[0]()
            _, (ms,) = timed_phases(args.steps, stream, phase)
            if not torch.equal(out, data):
                raise SystemExit("targeted-path comparison: decode output differs from the original data")
            res[name] = round(ms, 4)
    finally:
        afft_set_targeted_max(prev)
    # Algorithmic bytes as for every decode: k received symbols read, k
    # data symbols written.
    gbs = 2 * k * data.numel() // k / (res["targeted_ms"] / 1e3) / 1e9
    return {"missing": r, **res, "speedup": round(res["transforms_ms"] / res["targeted_ms"], 2),
            "targeted_gbs": round(gbs, 2), "targeted_hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


SEED_SERVE = 0x5C25_0005
SERVE_LENS = (512, 1500, 4096, 65536)


def serving_trace(n_req: int, seed: int = SEED_SERVE) -> list[tuple]:
    """The mixed serving stream, from one seed: per request
    (is_encode, engine, k, m, symbol_len, received indices or None).
    70% encode / 30% decode, engine uniform over {Standard Cauchy, Good
    Cauchy, Tower}, k ~ U[4,200], m ~ U[1,32], symbol_len from
    {512, 1500, 4096, 65536} (all even, as GF(2^16) needs), and a decode
    erases r ~ U[1,m] of its k+m indices and receives k survivors in a
    random order."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_req):
        enc = bool(rng.random() < 0.7)
        eng = int(rng.integers(0, 3))
        k, m = int(rng.integers(4, 201)), int(rng.integers(1, 33))
        sl = int(rng.choice(SERVE_LENS))
        idx = None
        if not enc:
            r = int(rng.integers(1, m + 1))
            keep = np.ones(k + m, bool)
            keep[rng.choice(k + m, size=r, replace=False)] = False
            idx = rng.permutation(np.flatnonzero(keep))[:k].astype(np.uint16)
        specs.append((enc, eng, k, m, sl, idx))
    return specs


def _align256(x: int) -> int:
    return (x + 255) // 256 * 256


class ServingSetup:
    """Device buffers and jobs for a serving trace.

    Encode jobs read k*L seeded bytes at a seeded offset of a shared pool.
    Every decode job reads a real codeword: its data symbols sit in
    ``ddata`` and its repairs in ``drep``, encoded beforehand.  Each job
    writes its own 256-byte-aligned output region of one of ``sets`` output
    sets; decode outputs are laid out exactly like the codewords' data in
    ``ddata`` (zero padding in both), so one comparison checks every decode.
    """

    POOL = 1 << 30

    def __init__(self, specs, sets: int = 2, seed: int = SEED_DATA):
        import torch

        from scrs_amd import Engine, Profile

        self.specs = specs
        engines = [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY, Engine.TOWER]
        rng = np.random.default_rng(seed ^ 0x0FF5E7)
        self.profiles, self.src_off, self.out_off, self.rep_off = [], [], [], []
        eo = do = ro = 0
        self.enc_bytes = self.dec_bytes = 0
        for enc, eng, k, m, sl, idx in specs:
            self.profiles.append(Profile.resolve(engines[eng], k, m, sl))
            if enc:
                self.src_off.append(int(rng.integers(0, (self.POOL - k * sl) // 256)) * 256)
                self.out_off.append(eo)
                self.rep_off.append(None)
                eo += _align256(m * sl)
                self.enc_bytes += (k + m) * sl
            else:
                self.src_off.append(None)
                self.out_off.append(do)
                self.rep_off.append(ro)
                do += _align256(k * sl)
                ro += _align256(m * sl)
                self.dec_bytes += 2 * k * sl
        self.algorithmic_bytes = self.enc_bytes + self.dec_bytes
        dev = "cuda"
        self.pool = torch.empty(self.POOL, dtype=torch.uint8, device=dev)
        self.ddata = torch.zeros(max(do, 256), dtype=torch.uint8, device=dev)
        self.drep = torch.zeros(max(ro, 256), dtype=torch.uint8, device=dev)
        self.out_e = [torch.zeros(max(eo, 256), dtype=torch.uint8, device=dev) for _ in range(sets)]
        self.out_d = [torch.zeros(max(do, 256), dtype=torch.uint8, device=dev) for _ in range(sets)]
        from scrs_amd import fill_random_dev

        fill_random_dev(self.pool, self.POOL, seed)
        # Seeded bytes in each decode codeword's data rows, zeros between.
        for i, (enc, _e, k, _m, sl, _idx) in enumerate(specs):
            if not enc:
                o = self.out_off[i]
                fill_random_dev(self.ddata[o:], k * sl, seed + 1 + i)
        torch.cuda.synchronize()

    def _encode_job(self, i, out_set):
        from scrs_amd import Job

        return Job.encode(self.profiles[i], self.pool.data_ptr() + self.src_off[i],
                          self.out_e[out_set].data_ptr() + self.out_off[i])

    def encode_codewords(self) -> None:
        """Repairs of every decode job's codeword, by one batch-engine
        submission (checked against the oracle by sample_encodes)."""
        import torch

        from scrs_amd import BatchEngine, Job

        jobs = []
        for i, (enc, _e, k, _m, sl, _idx) in enumerate(self.specs):
            if not enc:
                jobs.append(Job.encode(self.profiles[i], self.ddata.data_ptr() + self.out_off[i],
                                       self.drep.data_ptr() + self.rep_off[i]))
        if jobs:
            eng = BatchEngine(tile_bytes=4096, cw_per_unit=1, use_graph=False)
            eng.submit(jobs)
            torch.cuda.synchronize()
            eng.close()

    def jobs(self, out_set: int) -> list:
        from scrs_amd import Job

        out = []
        for i, (enc, _e, k, _m, sl, idx) in enumerate(self.specs):
            if enc:
                out.append(self._encode_job(i, out_set))
                continue
            base_d = self.ddata.data_ptr() + self.out_off[i]
            base_r = self.drep.data_ptr() + self.rep_off[i]
            ptrs = [base_d + int(x) * sl if x < k else base_r + (int(x) - k) * sl for x in idx]
            out.append(Job.decode(self.profiles[i], idx, ptrs, self.out_d[out_set].data_ptr() + self.out_off[i]))
        return out

    def decodes_ok(self, out_set: int) -> bool:
        """Every decode output equals its codeword's data (one comparison)."""
        import torch

        return bool(torch.equal(self.out_d[out_set], self.ddata))

    def sets_equal(self, a: int, b: int) -> bool:
        import torch

        return bool(torch.equal(self.out_e[a], self.out_e[b]) and torch.equal(self.out_d[a], self.out_d[b]))

    def clear(self, out_set: int) -> None:
        self.out_e[out_set].zero_()
        self.out_d[out_set].zero_()

    def sample_encodes(self, out_set: int, per_engine: int, seed: int = 0x5A3) -> int:
        """Compare `per_engine` seeded encode jobs of each engine, and as
        many pre-encoded decode codewords, with the oracle (GF(256) Cauchy
        product or GF(2^16) Tower product over the same input bytes).
        Raises SystemExit on a mismatch; returns the jobs checked."""
        from concurrent.futures import ThreadPoolExecutor

        import oracle

        rng = np.random.default_rng(seed)
        picks = []
        for eng in range(3):
            enc_i = [i for i, s in enumerate(self.specs) if s[0] and s[1] == eng]
            dec_i = [i for i, s in enumerate(self.specs) if not s[0] and s[1] == eng]
            for group in (enc_i, dec_i):
                if group:
                    n = min(per_engine if group is enc_i else max(1, per_engine // 4), len(group))
                    picks += [int(x) for x in rng.choice(group, size=n, replace=False)]
        work = []
        for i in picks:
            enc, eng, k, m, sl, _idx = self.specs[i]
            if enc:
                src = self.pool[self.src_off[i]:self.src_off[i] + k * sl]
                rep = self.out_e[out_set][self.out_off[i]:self.out_off[i] + m * sl]
            else:
                src = self.ddata[self.out_off[i]:self.out_off[i] + k * sl]
                rep = self.drep[self.rep_off[i]:self.rep_off[i] + m * sl]
            work.append((i, eng, k, m, sl, src.cpu().numpy(), rep.cpu().numpy()))

        def check(w):
            i, eng, k, m, sl, src, rep = w
            if eng == 2:
                want = oracle.tower_encode(oracle.tower_cauchy_matrix(k, m), k, m, sl, src)
            else:
                want = oracle.gf256_encode(oracle.gf256_matrix(eng, k, m), k, m, sl, src)
            return i if not np.array_equal(want, rep) else None

        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            bad = [i for i in ex.map(check, work) if i is not None]
        if bad:
            raise SystemExit(f"serving outputs differ from the oracle for requests {bad[:8]}")
        return len(work)


def serve_cpu_baseline(setup: "ServingSetup", target_s: float = 10.0, sample: int = 512) -> dict:
    """The reference's CPU paths on a seeded sample of the trace's requests
    (GF(256): the GFNI register-blocked multiply-XOR encode and recipe
    decode; Tower: the GFNI byte-swap feed and rational-Lagrange finalize),
    one request per thread; every sampled output is checked against the
    GPU's batched output."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle

    threads, host = cpu_threads()
    rng = np.random.default_rng(0xC9B)
    picks = sorted(int(i) for i in rng.choice(len(setup.specs), size=min(sample, len(setup.specs)), replace=False))
    work, nbytes = [], 0
    for i in picks:
        enc, eng, k, m, sl, idx = setup.specs[i]
        if enc:
            src = setup.pool[setup.src_off[i]:setup.src_off[i] + k * sl].cpu().numpy()
            want = setup.out_e[0][setup.out_off[i]:setup.out_off[i] + m * sl].cpu().numpy()
            work.append((True, eng, k, m, sl, None, src, None, want))
            nbytes += (k + m) * sl
        else:
            d = setup.ddata[setup.out_off[i]:setup.out_off[i] + k * sl].cpu().numpy()
            r = setup.drep[setup.rep_off[i]:setup.rep_off[i] + m * sl].cpu().numpy()
            want = setup.out_d[0][setup.out_off[i]:setup.out_off[i] + k * sl].cpu().numpy()
            work.append((False, eng, k, m, sl, np.ascontiguousarray(idx, np.uint16)[None, :], d, r, want))
            nbytes += 2 * k * sl
    mats = {}

    def matrix(eng, k, m):
        key = (eng, k, m)
        if key not in mats:
            mats[key] = (oracle.tower_cauchy_matrix(k, m) if eng == 2 else oracle.gf256_matrix(eng, k, m))
        return mats[key]

    for w in work:
        matrix(w[1], w[2], w[3])

    def one(w):
        enc, eng, k, m, sl, idx, a, b, _want = w
        c = mats[(eng, k, m)]
        if eng == 2:
            return (oracle.tower_encode_port(c, k, m, sl, a) if enc
                    else oracle.tower_finalize_port(k, m, sl, idx, a, b))
        return (oracle.gf256_encode_many(c, k, m, sl, a, 1, 1) if enc
                else oracle.gf256_decode_many(c, k, m, sl, idx, a, b, 1, 1))

    with ThreadPoolExecutor(max_workers=threads) as ex:
        got = list(ex.map(one, work))
        bad = [picks[i] for i, (g, w) in enumerate(zip(got, work)) if g is None or not np.array_equal(g, w[8])]
        if bad:
            raise SystemExit(f"CPU port and GPU serving outputs differ for requests {bad[:8]}")
        t_min, t_med, reps = _timed_reps(lambda: list(ex.map(one, work)), target_s)
    return {"value": round(nbytes / t_min / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{len(work)} seeded requests of the trace, one request per thread (GF(256) GFNI "
                      f"multiply-XOR / recipe decode, Tower GFNI feed / rational-Lagrange finalize); "
                      f"best of {reps} passes (median {round(nbytes / t_med / 1e9, 3)} GB/s)",
            "req_per_s": round(len(work) / t_min, 1), **host}


def serve_workload(args, bounded: bool = False) -> dict:
    """Mixed serving stream: 20000 requests, 70% encode / 30% decode, engine
    uniform over {Standard Cauchy, Good Cauchy, Tower}, k ~ U[4,200],
    m ~ U[1,32], symbol_len from {512, 1500, 4096, 65536}, erasures ~ U[1,m],
    one seed.  The same trace runs through the serving loop in plain mode
    (one codec call per request) and in batched mode (continuous batches
    through the fused engine, hipGraph replay, autotuned tiles).  Every
    decode output is compared with its codeword's data, every batched output
    with the plain one, and a seeded sample of encodes with the oracle.
    `bounded` trims the latency sweep (the sub-object of the default line)."""
    import torch

    from scrs_amd import Server, set_term_split

    torch.cuda.set_device(0)
    n_req = args.requests
    specs = serving_trace(n_req)
    setup = ServingSetup(specs, sets=2)
    setup.encode_codewords()
    total_alg = setup.algorithmic_bytes
    prepared = [Server.prepare(setup.jobs(0)), Server.prepare(setup.jobs(1))]

    def run(srv, out_set: int, rate: float | None, label: str):
        if rate is None:
            arr = np.zeros(n_req, np.uint64)
        else:
            gaps = np.random.default_rng(0xA55).exponential(1e9 / rate, size=n_req)
            arr = np.cumsum(gaps).astype(np.uint64)
        before = srv.stats()
        res = srv.replay(prepared[out_set], arr)
        after = srv.stats()
        if (res["code"] != 0).any():
            bad = np.flatnonzero(res["code"] != 0)
            raise SystemExit(f"{bad.size} serving requests failed, first {int(bad[0])} with status "
                             f"{int(res['code'][bad[0]])}")
        lat = (res["t_done_ns"] - res["t_submit_ns"]) / 1e6
        span = (res["t_done_ns"].max() - res["t_submit_ns"].min()) / 1e9
        out = {"gbs": round(total_alg / span / 1e9, 2), "req_per_s": round(n_req / span, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3)}
        nb = max(1, after["batches"] - before["batches"])
        out["batches"] = after["batches"] - before["batches"]
        out["mean_batch_jobs"] = round((after["jobs"] - before["jobs"]) / nb, 1)
        out["host_us_per_batch"] = round((after["host_ns"] - before["host_ns"]) / nb / 1e3, 1)
        out["gpu_us_per_batch"] = round((after["gpu_ns"] - before["gpu_ns"]) / nb / 1e3, 1)
        print(f"serve {label} rate={rate}: {out}", file=sys.stderr, flush=True)
        return out

    def verify(out_set: int, what: str):
        torch.cuda.synchronize()
        if not setup.decodes_ok(out_set):
            raise SystemExit(f"{what}: a decode did not reconstruct its codeword's data")

    # Each server first replays the whole trace once untimed: the plain path's
    # codecs and the batched path's resident tables, recipes and tuned
    # configurations are then warm, as in a long-running server.
    plain = Server(batched=False)
    run(plain, 1, None, "plain warm-up")
    verify(1, "plain")
    checked = setup.sample_encodes(1, per_engine=24 if bounded else 70)
    setup.clear(1)
    plain_sat = run(plain, 1, None, "plain")
    verify(1, "plain")
    batched = Server(batched=True, max_batch_jobs=args.max_batch, tile_bytes=0, cw_per_unit=0, use_graph=False,
                     cache_bytes=16 << 30)
    run(batched, 0, None, "batched warm-up")
    setup.clear(0)
    batched_sat = run(batched, 0, None, "batched")
    verify(0, "batched")
    if not setup.sets_equal(0, 1):
        raise SystemExit("batched and per-request outputs differ")
    cpu = None if args.no_cpu else serve_cpu_baseline(setup)

    def sweep(plain_srv, plain_capacity, mults, label):
        """Plain at half its capacity; then the batched path at the highest
        load whose p99 stays within the plain path's p99 there.  A load whose
        p99 lands above it is measured once more before the sweep ends (the
        p99 of a one-second replay moves with single stalls of the host);
        the kept point says whether it was the second measurement."""
        load = 0.5 * plain_capacity
        p_half = run(plain_srv, 1, load, f"{label} half load")
        best = None
        for mult in mults:
            rate = load * mult
            if rate > 1.2 * batched_sat["req_per_s"]:
                break
            rr = run(batched, 0, rate, "batched")
            retried = False
            if rr["p99_ms"] > p_half["p99_ms"]:
                rr = run(batched, 0, rate, "batched (again)")
                retried = True
                if rr["p99_ms"] > p_half["p99_ms"]:
                    break
            best = dict(rr, offered_req_per_s=round(rate, 1), retried=retried)
        ratio = round(best["req_per_s"] / p_half["req_per_s"], 2) if best else None
        return p_half, best, ratio

    mults = (1.5, 2, 3, 4, 6, 8, 12, 16) if bounded else (1.25, 1.5, 2, 2.5, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32)
    plain_half, best, ratio = sweep(plain, plain_sat["req_per_s"], mults, "plain")
    batched_half = run(batched, 0, 0.5 * plain_sat["req_per_s"], "batched at plain half load")
    # For reference, the plain path with the codecs' term split switched off:
    # a lone codeword then runs on groups x tiles workgroups, each folding
    # every term (the weaker baseline).
    prev_split = set_term_split(False)
    try:
        plain_ts = Server(batched=False)
        run(plain_ts, 1, None, "plain unsplit warm-up")
        setup.clear(1)
        plain_ts_sat = run(plain_ts, 1, None, "plain unsplit")
        verify(1, "plain unsplit")
        if not setup.sets_equal(0, 1):
            raise SystemExit("unsplit plain and batched outputs differ")
        ts_mults = (4, 8, 16, 32, 64) if bounded else (2, 4, 8, 12, 16, 24, 32, 48, 64, 96)
        plain_ts_half, best_ts, ratio_ts = sweep(plain_ts, plain_ts_sat["req_per_s"], ts_mults, "plain unsplit")
        plain_ts.close()
    finally:
        set_term_split(prev_split)
    plain.close()
    batched.close()
    value = batched_sat["gbs"]
    return {
        "metric": "encode/decode GB/s (bit-exact), % HBM roofline; batched vs plain GPU; p99 latency",
        "value": value, "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(total_alg / (value * 1e9) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/u16",
        "data": "synthetic (seeded splitmix64 bytes in HBM; seeded request trace; decodes read real codewords)",
        "config": {"workload": "serving stream: %d requests, Poisson arrivals, 70%% encode / 30%% decode, "
                               "engine uniform {Standard, Good, Tower}, k~U[4,200], m~U[1,32], "
                               "symbol_len in {512,1500,4096,65536}, erasures~U[1,m]" % n_req,
                   "requests": n_req, "max_batch_jobs": args.max_batch,
                   "algorithmic_bytes": total_alg},
        "checks": {"decodes_vs_data": n_req - sum(s[0] for s in specs), "batched_vs_plain": n_req,
                   "encodes_vs_oracle": checked},
        "plain_saturated": plain_sat, "batched_saturated": batched_sat,
        "equal_load": {"offered_req_per_s": round(0.5 * plain_sat["req_per_s"], 1), "plain": plain_half,
                       "batched": batched_half},
        "batched_at_plain_p99": best, "throughput_ratio_at_equal_p99": ratio,
        "plain_unsplit": {"saturated": plain_ts_sat, "equal_load": plain_ts_half,
                          "batched_at_its_p99": best_ts, "throughput_ratio_at_equal_p99": ratio_ts},
        "roofline": {"bound": "hbm", "kernel": "batched serving (whole stream)", "achieved": value,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(value / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }


def host_slice_rates(enc, dec, data, rep, idx, sample: int = 32) -> dict:
    """The drop-in host-slice calls (encode_into_with / decode_into_with) on
    ordinary host buffers: every call stages its bytes over PCIe and returns
    with the result in host memory.  Reported beside the device-resident
    figures, never as the headline value."""
    h_data = data[: sample * K * L].cpu().numpy().reshape(sample, K * L)
    h_rep = rep[: sample * M * L].cpu().numpy().reshape(sample, M * L)
    got = np.empty(M * L, np.uint8)
    out = np.empty(K * L, np.uint8)
    es, ds = enc.scratch(), dec.scratch()
    for c in range(2):  # warm the staging buffers
        enc.encode_into_with(h_data[c], got, es)
    t0 = time.perf_counter()
    for c in range(sample):
        enc.encode_into_with(h_data[c], got, es)
    t_enc = time.perf_counter() - t0
    if not np.array_equal(got, h_rep[sample - 1]):
        raise SystemExit("host-slice encode differs from the device batch encode")
    syms = []
    for c in range(sample):
        row = []
        for t in idx[c]:
            t = int(t)
            row.append((t, h_data[c, t * L:(t + 1) * L] if t < K else h_rep[c, (t - K) * L:(t - K + 1) * L]))
        syms.append(row)
    dec.decode_into_with(syms[0], out, ds)
    t0 = time.perf_counter()
    for c in range(sample):
        dec.decode_into_with(syms[c], out, ds)
    t_dec = time.perf_counter() - t0
    if not np.array_equal(out, h_data[sample - 1]):
        raise SystemExit("host-slice decode differs from the original data")
    return {
        "encode_gbs": round(sample * (K + M) * L / t_enc / 1e9, 2),
        "decode_gbs": round(sample * 2 * K * L / t_dec / 1e9, 2),
        "sample": f"{sample} codewords, one encode_into_with / decode_into_with call each, pageable host buffers",
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--count", type=int, default=COUNT, help="codewords per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-bitmatrix", action="store_true",
                    help="skip the bit-matrix (int8 MFMA) kernel leg of the GF(256) workload")
    ap.add_argument("--no-host-slice", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) encode_into / decode_into leg")
    ap.add_argument("--workload", choices=["gf256", "tower", "afft", "serve"], default="gf256",
                    help="gf256: the headline GF(256) batch line; tower / afft: the GF(2^16) engines; "
                         "serve: the mixed request stream, plain vs batched")
    ap.add_argument("--requests", type=int, default=20000, help="serve: requests in the trace")
    ap.add_argument("--max-batch", type=int, default=1024, help="serve: jobs per batch at most")
    ap.add_argument("--headline-only", action="store_true",
                    help="default workload: print only the GF(256) line, without the Tower, additive-FFT "
                         "and serving sub-objects")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before the timed steps, run a plain 1 GiB device copy (torch copy_) so a "
                         "counter profile can scale FETCH_SIZE / WRITE_SIZE by a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the workload shards by "
                          "codeword with no exchange, so there is no multi-GPU step to time"}))
        raise SystemExit(2)
    if args.workload == "serve":
        print(json.dumps(serve_workload(args)))
        return
    if args.workload != "gf256":
        print(json.dumps(gf16_workload(args)))
        return
    line = gf256_line(args)
    if not args.headline_only:
        # The other configurations, each with its own roofline and CPU
        # baseline, as sub-objects of the one line.
        import torch

        for name in ("tower", "afft"):
            torch.cuda.empty_cache()
            print(f"bench: {name} workload", file=sys.stderr, flush=True)
            line[name] = gf16_workload(args, name)
        torch.cuda.empty_cache()
        print("bench: serve workload", file=sys.stderr, flush=True)
        line["serve"] = serve_workload(args, bounded=True)
    print(json.dumps(line))


def gf256_line(args) -> dict:
    """The headline line: GF(256) Good Cauchy k=32 m=16 L=65536 x 4096."""
    import torch

    from scrs_amd import Engine, Profile, batch_decoder, batch_encoder, fill_random_dev

    torch.cuda.set_device(0)
    count = args.count
    prof = Profile.resolve(Engine.GOOD_CAUCHY, K, M, L)
    enc, dec = batch_encoder(prof), batch_decoder(prof)
    data = torch.empty(count * K * L, dtype=torch.uint8, device="cuda")
    rep = torch.empty(count * M * L, dtype=torch.uint8, device="cuda")
    out = torch.empty(count * K * L, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    fill_random_dev(data, data.numel(), SEED_DATA, 0, stream)
    idx = erasure_indices(count, K, M, 16, SEED_ERASE)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        gib = 1 << 30
        src_c = torch.ones(gib, dtype=torch.uint8, device="cuda")
        dst_c = torch.empty_like(src_c)
        for _ in range(2):
            dst_c.copy_(src_c)
        torch.cuda.synchronize()
        del src_c, dst_c

    def step():
        enc.encode_batch_dev(data, rep, count, stream)
        dec.decode_batch_dev(idx, data, rep, out, count, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        enc.encode_batch_dev(data, rep, count, stream)
        e1.record(stream)
        dec.decode_batch_dev(idx, data, rep, out, count, stream)
        e2.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    enc_ms = sum(a.elapsed_time(b) for a, b, _ in ev) / args.steps
    dec_ms = sum(b.elapsed_time(c) for _, b, c in ev) / args.steps

    # Bit-exactness: every codeword round-trips; sampled repairs vs the oracle.
    if not torch.equal(out, data):
        raise SystemExit("decode output differs from the original data")
    enc_bytes = count * (K + M) * L
    dec_bytes = count * 2 * K * L
    ms_per_step = wall * 1e3 / args.steps
    value = (enc_bytes + dec_bytes) / (wall / args.steps) / 1e9

    # The other GF(256) kernels on the same workload: the bit-matrix product
    # on int8 MFMA and on e2m1 (fp4) MFMA, timed the same way and checked
    # against the table path.
    def mfma_leg(kind, chunk, peak, label):
        benc, bdec = batch_encoder(prof), batch_decoder(prof)
        benc.set_kernel(kind)
        bdec.set_kernel(kind)
        rep_b = torch.empty_like(rep)
        out_b = torch.empty_like(out)
        for _ in range(max(1, args.warmup)):
            benc.encode_batch_dev(data, rep_b, count, stream)
            bdec.decode_batch_dev(idx, data, rep_b, out_b, count, stream)
        evb = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
                torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        torch.cuda.synchronize()
        for e0, e1, e2 in evb:
            e0.record(stream)
            benc.encode_batch_dev(data, rep_b, count, stream)
            e1.record(stream)
            bdec.decode_batch_dev(idx, data, rep_b, out_b, count, stream)
            e2.record(stream)
        torch.cuda.synchronize()
        if not torch.equal(rep_b, rep) or not torch.equal(out_b, data):
            raise SystemExit(f"{kind.name} kernel output differs from the table kernel")
        b_enc = sum(a.elapsed_time(b) for a, b, _ in evb) / args.steps
        b_dec = sum(b.elapsed_time(c) for _, b, c in evb) / args.steps
        # MFMA work per launch: 2 * (8 row bits) * (8 source bits) per byte
        # column, rows padded to 16 per group and sources to whole chunks.
        groups = (M + 15) // 16
        k_pad = (K + chunk - 1) // chunk * chunk
        ops = 2.0 * (8 * 16 * groups) * (8 * k_pad) * L * count
        del rep_b, out_b
        return {
            "encode_ms": round(b_enc, 4), "decode_ms": round(b_dec, 4),
            "encode_gbs": round(enc_bytes / (b_enc / 1e3) / 1e9, 2),
            "decode_gbs": round(dec_bytes / (b_dec / 1e3) / 1e9, 2),
            f"encode_{label}_tops": round(ops / (b_enc / 1e3) / 1e12, 1),
            f"mfma_frac_of_{label}_peak": round(ops / (b_enc / 1e3) / peak, 4),
        }

    bitmatrix = bitmatrix_fp4 = None
    if not args.no_bitmatrix:
        from scrs_amd import Kernel
        bitmatrix = mfma_leg(Kernel.BITMATRIX, 8, INT8_PEAK_OPS, "int8")
        bitmatrix_fp4 = mfma_leg(Kernel.BITMATRIX_FP4, 16, FP4_PEAK_FLOPS, "fp4")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(idx, rep[: 16 * M * L].cpu().numpy())
    host = host_slice_rates(enc, dec, data, rep, idx) if not args.no_host_slice else None

    enc_gbs = enc_bytes / (enc_ms / 1e3) / 1e9
    enc_lds_bytes = count * K * L * 32  # two 16-byte entries per source byte, M = 16 rows
    dec_gbs = dec_bytes / (dec_ms / 1e3) / 1e9
    traffic = traffic_enc = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        measured = json.loads(tf.read_text())
        traffic = measured.get("gf256_decode_rows_terms_lds")
        traffic_enc = measured.get("gf256_encode_rows_terms_lds")
    line = {
        "metric": "encode/decode GB/s (bit-exact), % HBM roofline; batched vs plain GPU; p99 latency",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded splitmix64 bytes generated in HBM; seeded erasure subsets)",
        "config": {
            "workload": "GF(256) Good Cauchy block encode + batch decode, k=32 m=16 symbol_len=65536, "
                        "4096 codewords (8 GiB data), erase U{1..16} of 48 per codeword",
            "codewords": count, "k": K, "m": M, "symbol_len": L,
            "kernel": "lds-nibble-tables",
        },
        "encode_gbs": round(enc_gbs, 2),
        "decode_gbs": round(dec_gbs, 2),
        "encode_ms": round(enc_ms, 4),
        "decode_ms": round(dec_ms, 4),
        # The dominant kernel is the decode's multiply-XOR (the longest launch
        # of the step); its rate is taken over the whole decode phase (solve and
        # pointer kernels included), timed with events on the
        # launch stream.  The encode launch is reported beside it.
        "roofline": {
            "bound": "hbm",
            "kernel": "rows_terms_lds_kernel, decode (whole decode phase timed)",
            "achieved": round(dec_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dec_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "roofline_encode": {
            "bound": "hbm",
            "kernel": "rows_terms_lds_kernel, encode",
            "achieved": round(enc_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(enc_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_enc,
            # Each source byte reads two 16-byte table entries (all 16 rows):
            # the encode is bound by LDS reads as much as by HBM.
            "lds_achieved_tbs": round(enc_lds_bytes / (enc_ms / 1e3) / 1e12, 2),
            "lds_peak_tbs": round(LDS_PEAK_TBS, 2),
            "lds_frac": round(enc_lds_bytes / (enc_ms / 1e3) / 1e12 / LDS_PEAK_TBS, 4),
        },
        "cpu_baseline": cpu,
        "bitmatrix_kernel": bitmatrix,
        "bitmatrix_fp4_kernel": bitmatrix_fp4,
        "host_slice": host,
    }
    del data, rep, out
    return line


if __name__ == "__main__":
    main()
