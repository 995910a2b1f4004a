"""Benchmark: single-proof Virgo prover latency on one MI355X.

Workload (the headline configuration): a random layered circuit with 16
compute layers of 2^20 gates (add/mult with probability 1/2, wiring uniform
over the previous layer, splitmix64 seed 1) and 2^20 inputs uniform in
[0, 2^31), proved with the reference's ZK-GKR sumcheck and the Virgo
polynomial commitment of the input layer (rate 1/32, 64+1 slices, FRI down to
32, 33 queries).  A step is one complete proof: inputs and gates are resident
in HBM when the timed region starts; the step ends with the full transcript
assembled in host memory.

Prints one JSON line.  `value` is milliseconds per proof (lower is better).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# HIP's default of 4 hardware queues per process: more queues measured slower
# for the batched workload (675 vs 596 and 536 proofs/s with 4, 24 and 16
# queues at 12-16 workers), set before the HIP runtime starts
os.environ["GPU_MAX_HW_QUEUES"] = "4"

PEAK_HBM_GBS = 8000.0
# MI355X VALU issue at the published 2.4 GHz: 256 CUs x 4 SIMDs, each
# issuing one wave64 three-input bitwise operation (v_bitop3_b32, v_xor_b32)
# every 2 cycles -- 32 lanes per cycle, the issue width behind the published
# 157.3 TFLOP/s of packed FP32 -- and one funnel shift (v_alignbit_b32) every
# 4 cycles (16 lanes per cycle); vg_valu_rates measures both on the device
# (about 0.40 and 0.23 wave instructions per SIMD per cycle at 8 waves per
# SIMD).
VALU_SIMDS = 256 * 4
CLOCK_GHZ = 2.4
# one op per lane per cycle at 16 lanes: the single-rate figure
PEAK_VALU_GOPS = VALU_SIMDS * 16 * CLOCK_GHZ


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-width", type=int, default=20)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log-width", type=int, default=0,
                    help="oracle baseline at this width, scaled to --log-width (default: the full width)")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-workers", type=int, default=16)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-rate", type=float, default=100.0, help="offered requests per second (Poisson)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-budget-gb", type=int, default=96)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-microset", action="store_true")
    ap.add_argument("--microset-sizes", default="18,20,22,24")
    ap.add_argument("--part-only", choices=["batch", "serve", "microset"],
                    help="run only the batched or the serving part and print its object (for profiling)")
    return ap.parse_args()


def leaf_traffic(n):
    """HBM bytes per k_leaf_multi launch, from rocprofv3 FETCH_SIZE and
    WRITE_SIZE of the 2^20-gate proof scaled by the same counters for a 1 GiB
    copy (FETCH_SIZE x 2048 B/KiB, WRITE_SIZE x 1024 B/KiB): 798999.8 KiB
    read, 24575.5 KiB written.  Other widths: null (not measured)."""
    if n != 20:
        return None
    return int(798999.8 * 2048 + 24575.5 * 1024)


def c1(args):
    """The reference's own CPU-runnable case: prove + verify of the 32 x 32
    matrix-product circuit (inputs padded to 2^15), on the GPU (best of 5
    after a warm-up; the device verifier) and, unless --no-cpu-baseline, with
    the CPU oracle (1 core); transcripts compared byte for byte."""
    import virgo_amd as V
    from virgo_amd import workloads as W
    circ = W.matmul_circuit(32, seed=1)
    p = V.Prover(circ)
    tr = p.prove()
    tp, tv = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        tr = p.prove()
        tp.append(time.perf_counter() - t0)
        ok, _, vms = p.verify(tr)
        tv.append(vms)
    p.close()
    out = {"workload": "32x32 matmul, inputs uniform in [0, 2^31), seed 1", "gpu_prove_ms": round(min(tp) * 1e3, 3),
           "gpu_verify_ms": round(min(tv), 3), "accepted": bool(ok)}
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        t0 = time.perf_counter()
        ref, _ = O.prove(circ)
        out["cpu_prove_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        t0 = time.perf_counter()
        ok_o, _ = O.verify(circ, tr)
        out["cpu_verify_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["bit_exact_vs_cpu"] = ref == tr and bool(ok_o)
    return out


MICRO_RATE_LOG = 3      # the kernel micro-set's extension rate 1/8
# Algorithmic 32-bit VALU operations of one Keccak-f[1600] as SHA3-256 of a
# 64-byte block uses it (leaf chains, Merkle nodes, FRI leaves): the FIPS 202
# round on 32-bit halves (theta 80, rho 48, chi 50, iota 1-2 per round) with
# the constant lanes of the padded block folded and the work that never
# reaches the four digest lanes dropped; derived symbolically in
# tests/test_keccak_opcount.py (a general state costs 24 x 180 - 11 = 4309).
KECCAK_F_OPS = 4174
# the same split by issue class (tests/test_keccak_opcount.py derives both)
KECCAK_F_BIT_OPS = 2822
KECCAK_F_SHIFT_OPS = 1352


def keccak_peak_gops(bit_cpi=2.0, shift_cpi=4.0, ghz=CLOCK_GHZ):
    """Algorithmic ops/s of SHA3-256 (64-byte block) permutations at full
    VALU issue: a SIMD spends bit_cpi cycles per wave64 bitwise operation and
    shift_cpi per funnel shift, 64 lanes each."""
    cycles = KECCAK_F_BIT_OPS * bit_cpi + KECCAK_F_SHIFT_OPS * shift_cpi   # per 64 permutations on one SIMD
    return KECCAK_F_OPS * 64 * VALU_SIMDS * ghz / cycles


def _rand_fe(rng, n):
    import numpy as np
    import virgo_amd as V
    a = np.zeros(n, dtype=V.FE)
    a["real"] = rng.integers(0, (1 << 61) - 1, n, dtype=np.uint64)
    a["img"] = rng.integers(0, (1 << 61) - 1, n, dtype=np.uint64)
    return a


def microset(args):
    """The kernel micro-set: one polynomial of 2^k uniform F_{p^2}
    coefficients (k in --microset-sizes), forward and inverse transform,
    rate-1/8 extension, Merkle tree over the extension's 32-byte pairs, FRI
    commit (fold + tree per level) down to 32 values; each stage timed on the
    device (mean of 5 after one warm-up, vg_microset), with SHA3-256 (the
    reference's hash) and SHA-256 trees.  Rates: transform butterflies/s,
    tree hashes/s, and for SHA3 trees the VALU fraction at KECCAK_F_OPS ops per
    Keccak-f.  CPU: the oracle's stages (the reference's FFT, SHA3 tree and
    fold, 1 core) at 2^18 and 2^20."""
    import numpy as np
    import virgo_amd as V
    out = {"workload": "one polynomial of 2^k uniform F_{p^2} coefficients, rate 1/8 extension, Merkle tree of "
                       "32-byte pairs, FRI commit to 32 values, seed 400+k", "sizes": {}}
    for k in [int(x) for x in args.microset_sizes.split(",") if x]:
        rng = np.random.default_rng(400 + k)
        L = k + MICRO_RATE_LOG
        c, rs = _rand_fe(rng, 1 << k), _rand_fe(rng, L - 5)
        row = {}
        for h in ("sha3", "sha256"):
            roots, _, ok, ms = V.microset(c, MICRO_RATE_LOG, rs, hash=h, reps=5)
            M = 1 << L
            fly = (1 << k) // 2 * k
            hashes = M // 2 - 1
            e = {s + "_ms": round(v, 4) for s, v in ms.items()}
            e["total_ms"] = round(sum(ms.values()), 3)
            e["roundtrip_exact"] = ok
            e["fft_gbfly_s"] = round(fly / (ms["fft"] * 1e-3) / 1e9, 2)
            e["extension_gbfly_s"] = round((fly << MICRO_RATE_LOG) / (ms["extension"] * 1e-3) / 1e9, 2)
            e["merkle_mhash_s"] = round(hashes / (ms["merkle"] * 1e-3) / 1e6, 1)
            if h == "sha3":
                gops = hashes * KECCAK_F_OPS / (ms["merkle"] * 1e-3) / 1e9
                e["merkle_valu_frac"] = round(gops / keccak_peak_gops(), 3)
            e["root"] = bytes(roots[0]).hex()[:16]
            row[h] = e
        out["sizes"]["2^%d" % k] = row
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cpu = {}
        for k in (18, 20):
            rng = np.random.default_rng(400 + k)
            L = k + MICRO_RATE_LOG
            N, M = 1 << k, 1 << L
            c, rs = _rand_fe(rng, N), _rand_fe(rng, L - 5)
            w = O.root_of_unity(k)
            t = [time.perf_counter()]
            f = O.fft(c, N, w)
            t.append(time.perf_counter())
            O.ifft(f, N, w)
            t.append(time.perf_counter())
            E = O.fft(c, M, O.root_of_unity(L))
            t.append(time.perf_counter())
            lv = np.ascontiguousarray(np.stack([E[:M // 2], E[M // 2:]], axis=1)).view(np.uint8).reshape(-1, 32)
            root0 = bytes(O.merkle_tree(lv)[1]).hex()[:16]
            t.append(time.perf_counter())
            cw = E
            for j in range(L - 5):
                cw = O.fri_fold1(cw, rs[j:j + 1])
                if j + 1 < L - 5:
                    h2 = len(cw) // 2
                    O.merkle_tree(np.ascontiguousarray(np.stack([cw[:h2], cw[h2:]], axis=1)).view(np.uint8).reshape(-1, 32))
            t.append(time.perf_counter())
            names = ("fft", "ifft", "extension", "merkle", "fri")
            e = {n + "_ms": round((t[i + 1] - t[i]) * 1e3, 1) for i, n in enumerate(names)}
            e["total_ms"] = round((t[-1] - t[0]) * 1e3, 1)
            e["same_root_as_gpu"] = root0 == out["sizes"].get("2^%d" % k, {}).get("sha3", {}).get("root", root0)
            cpu["2^%d" % k] = e
        out["cpu_oracle_1core"] = cpu
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def keccak_rates():
    """Keccak-f[1600] permutations per second on this core: the oracle's
    SHA3-256 of 64-byte blocks (one permutation each) against OpenSSL's
    (hashlib.sha3_256 over one long message: one permutation per 136-byte
    block, so no per-call Python overhead)."""
    import hashlib
    import numpy as np
    from oracle import oracle as O
    n = 1 << 18
    blocks = np.random.default_rng(3).integers(0, 256, (n, 64), dtype=np.uint8)
    O.sha3_64(blocks[:1024])
    t0 = time.perf_counter()
    O.sha3_64(blocks)
    t_o = time.perf_counter() - t0
    msg = bytes(136 * n)
    t0 = time.perf_counter()
    hashlib.sha3_256(msg).digest()
    t_h = time.perf_counter() - t0
    return {"oracle": round(n / t_o), "openssl_hashlib": round(n / t_h), "ratio": round(t_h / t_o, 3)}


def cpu_baseline(args):
    """Times the CPU oracle (single-threaded C restatement of the reference
    prover, -O3 with AVX2 and mulx as the reference builds) proving the same
    circuit as the GPU (2^log_width gates x depth layers, seed 1), once, at
    full size, pinned to one host core, before this process initialises the
    GPU.  Also reports the host CPU model and the oracle's Keccak-f rate
    against OpenSSL's on the same core."""
    from oracle import oracle as O
    from virgo_amd import workloads as W
    lw = args.cpu_sample_log_width or args.log_width
    circ = W.random_circuit(lw, args.depth, seed=1)
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[0]})
    try:
        rates = keccak_rates()
        t0 = time.perf_counter()
        tr, parts = O.prove(circ)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, set(cpus))
    scale = 2 ** (args.log_width - lw)
    out = {"value": round(dt * 1e3 * scale, 1), "unit": "ms per proof", "cores": 1, "kind": "port",
           "sample": "one oracle proof of the random circuit 2^%d x %d layers, seed 1, on core %d, before GPU init"
                     % (lw, args.depth, cpus[0]) + ("" if scale == 1 else "; x%d width scaling" % scale),
           "cpu_model": _cpu_model(),
           "stage_ms": {k: round(float(v) * 1e3, 1) for k, v in
                        zip(("evaluate", "gkr", "commit", "queries", "fft_gkr"), parts)},
           "keccak_f_per_s": rates}
    return out, (tr if scale == 1 else None)


def _cpu_prove_spec(spec):
    """One oracle proof of a batch job (runs in a worker process)."""
    from oracle import oracle as O
    from virgo_amd import workloads as W
    circ = W.make_job(spec)
    t0 = time.perf_counter()
    O.prove(circ)
    return time.perf_counter() - t0


def _noop(_):
    return 0


def cpu_batch(args, procs=16, stride=8):
    """CPU throughput baseline of the batched workload: the oracle
    (single-threaded C restatement of the reference prover) proving every
    `stride`-th job of workloads.batch_specs in `procs` worker processes at
    once (the host cores of one GPU's share).  proofs/s = sample size over
    (summed proof seconds / procs): the whole 256-job batch would keep the
    processes evenly busy, which a sample's wall time, set by its longest
    proof, does not show (reported beside it)."""
    import multiprocessing as mp
    from virgo_amd import workloads as W
    sample = W.batch_specs(args.batch_jobs)[::stride]
    with mp.get_context("spawn").Pool(procs) as pool:
        pool.map(_noop, range(procs))                   # workers started and imported
        t0 = time.perf_counter()
        times = pool.map(_cpu_prove_spec, sample, chunksize=1)
        wall = time.perf_counter() - t0
    return {"value": round(len(sample) / (sum(times) / procs), 2), "unit": "proofs/s", "cores": procs, "kind": "port",
            "sample": "oracle proofs of every %dth job (%d of %d), %d processes: %.1f s wall, %.1f s CPU"
                      % (stride, len(sample), args.batch_jobs, procs, wall, sum(times))}


def batched(args):
    """Batched throughput on the uneven workload (workloads.batch_jobs).

    Plain loop: one proof at a time, each on a context in its lowest-latency
    mode (own streams, stages overlapped); a step is the second proof on a
    fresh context, summed over the jobs.  Batched: every circuit resident in
    HBM at once, contexts created round-robin on `batch_workers` streams, and
    vg_prove_batch proving them from that many library threads (each proof in
    order on its stream, the streams overlapping on the device); the median
    of five timed batches after three warm-up batches.  Both legs time the
    library calls (vg_prove, vg_prove_batch: inputs resident in HBM to every
    transcript in the library's host buffers); the copies into Python bytes
    after them are not counted (for the batch they took 10-50 ms, most of it
    first-touch page faults of 256 fresh objects).  Every batched transcript
    must equal the plain one byte for byte."""
    import virgo_amd as V
    from virgo_amd import workloads as W
    jobs = W.batch_jobs(args.batch_jobs)
    plain, clock = [], []
    for _, c in jobs:
        p = V.Prover(c)
        p.prove()
        plain.append(p.prove(clock))
        p.close()
    t_plain = sum(clock)
    t0 = time.perf_counter()
    streams = [V.Stream() for _ in range(args.batch_workers)]
    provers = [V.Prover(c, stream=streams[k % args.batch_workers]) for k, (_, c) in enumerate(jobs)]
    setup = time.perf_counter() - t0
    dev_bytes = sum(p.device_bytes() for p in provers)
    for _ in range(3):                                  # warm-up: a process's first batches run slower
        V.prove_batch(provers, args.batch_workers)
    runs = []
    for _ in range(5):
        out = V.prove_batch(provers, args.batch_workers, runs)
    t_batch = sorted(runs)[2]                           # median of five
    same = all(a == b for a, b in zip(plain, out))
    for p in provers:
        p.close()
    for st in streams:
        st.close()
    n = len(provers)
    return {"workload": "%d proofs: matmul n log-uniform in {8..256}, random layered log2 width U[10,18] x depth "
                        "U[4,20], seed 7" % n,
            "proofs_per_s": round(n / t_batch, 1), "plain_loop_proofs_per_s": round(n / t_plain, 1),
            "speedup_vs_plain": round(t_plain / t_batch, 2), "workers": args.batch_workers,
            "bit_exact_vs_plain": bool(same), "device_gb": round(dev_bytes / 1e9, 1),
            "setup_s": round(setup, 1)}


def serving(args):
    """Continuous batching under a Poisson stream of requests
    (workloads.serve_jobs: log2 width U[10,20], depth U[8,24]) at a fixed
    offered rate, through the library's proof server: each request is
    admitted as a slot and its lease of the HBM budget free up, with every
    admitted proof's state resident on the device until its transcript is on
    the host.  Latency = submit to transcript, measured in the library.  Then
    every request again through the same server, one at a time: its alone
    latency (the floor of its latency under this server) and, per request,
    the slowdown latency / alone latency.  Each transcript is compared with
    the same circuit alone (through the server and on a context of its own)
    and checked by the device verifier."""
    import numpy as np
    import virgo_amd as V
    from virgo_amd import workloads as W
    jobs = W.serve_jobs(args.serve_requests)
    srv = V.Server(device_budget=args.serve_budget_gb << 30, slots=args.serve_slots)
    for _, c in jobs[:4]:                                # warm-up
        srv.wait(srv.submit(c))
    gaps = np.random.default_rng(13).exponential(1.0 / args.serve_rate, len(jobs))
    t0 = time.perf_counter()
    arrive = t0 + np.cumsum(gaps)
    ids, t_sub = [], []
    for k, (_, c) in enumerate(jobs):
        delay = arrive[k] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        t_sub.append(time.perf_counter())
        ids.append(srv.submit(c))
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    # the same requests through the same server one at a time: each one's
    # latency with nothing else on the device (its floor under this server)
    alone = []
    for _, c in jobs:
        tr, t = srv.wait(srv.submit(c))
        alone.append((tr, t["latency"]))
    srv.close()
    lat = np.array([t["latency"] for _, t in res])
    lat_alone = np.array([a for _, a in alone])
    slow = lat / lat_alone
    t_end = max(ts + t["latency"] * 1e-3 for ts, (_, t) in zip(t_sub, res))
    same, accepted = all(a == tr for (a, _), (tr, _) in zip(alone, res)), True
    for (_, c), (tr, _) in zip(jobs, res):
        p = V.Prover(c)
        same &= p.prove() == tr
        accepted &= p.verify(tr)[0]
        p.close()
    mean = lambda k: round(float(np.mean([t[k] for _, t in res])), 2)  # noqa: E731
    return {"workload": "%d requests: random layered log2 width U[10,20] x depth U[8,24], seed 11; Poisson "
                        "arrivals, seed 13" % len(jobs),
            "offered_per_s": args.serve_rate, "proofs_per_s": round(len(jobs) / (t_end - t0), 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_ms": round(float(lat.mean()), 2),
            "alone_p50_ms": round(float(np.percentile(lat_alone, 50)), 2),
            "alone_p99_ms": round(float(np.percentile(lat_alone, 99)), 2),
            "slowdown_p50": round(float(np.percentile(slow, 50)), 2),
            "slowdown_p99": round(float(np.percentile(slow, 99)), 2),
            "mean_queued_ms": mean("queued"), "mean_plan_ms": mean("plan"),
            "mean_prove_ms": mean("prove"), "slots": args.serve_slots, "budget_gb": args.serve_budget_gb,
            "peak_admitted": st["peak_admitted"], "peak_device_gb": round(st["device_peak"] / 1e9, 1),
            "bit_exact_vs_alone": bool(same), "all_accepted": bool(accepted)}


_T0 = time.perf_counter()


def progress(what):
    """One line on stderr as each part of the run finishes."""
    print("bench: %s done at %.1f s" % (what, time.perf_counter() - _T0), file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import virgo_amd as V
    from virgo_amd import workloads as W
    if args.part_only:
        part = {"batch": batched, "serve": serving, "microset": microset}[args.part_only]
        print(json.dumps({args.part_only: part(args)}))
        return
    # the CPU throughput baseline of the batch first: its worker processes
    # start before this process initialises the GPU
    cpu_b = cpu_batch(args) if not args.no_cpu_baseline and not args.no_batch else None
    if cpu_b:
        progress("cpu batch baseline")
    cpu_1, cpu_tr = cpu_baseline(args) if not args.no_cpu_baseline else (None, None)
    if cpu_1:
        progress("cpu single-proof baseline")
    circ = W.random_circuit(args.log_width, args.depth, seed=1)
    p = V.Prover(circ)
    for _ in range(args.warmup):
        p.prove()
    stage, clock = [], []
    for _ in range(args.steps):
        tr = p.prove(clock)
        stage.append(p.stage_ms.copy())
    ms = sum(clock) * 1e3 / args.steps
    progress("timed proofs")
    import numpy as np
    st = np.mean(np.array(stage), axis=0)
    names = V.STAGES
    n = args.log_width
    # Dominant kernel: k_leaf_multi, the SHA3-256 leaf chains of every
    # commitment (initial L and H trees and every FRI level) in one launch:
    # 65 permutations per chain, KECCAK_F_OPS algorithmic operations each.
    chains = 2 * 2 ** (n - 2) + sum(2 ** (n - 3 - k) for k in range(n - 6))
    ops = chains * 65 * KECCAK_F_OPS
    leaf_ms = float(st[7])
    achieved = ops / (leaf_ms * 1e-3) / 1e9
    peak = keccak_peak_gops()
    line = {
        "metric": "prover ms @2^20-gate circuit; batched proofs/s vs plain per-proof GPU loop",
        "value": round(ms, 3),
        "unit": "ms per proof",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 F_{p^2} (p=2^61-1), exact",
        "data": "synthetic",
        "config": {"workload": "random layered circuit 2^%d gates x %d layers, 2^%d inputs, seed 1"
                               % (args.log_width, args.depth, args.log_width),
                   "transcript_bytes": len(tr),
                   "stage_ms": {k: round(float(v), 3) for k, v in zip(names, st)}},
        "roofline": {"bound": "valu", "achieved": round(achieved, 1), "peak": round(peak, 1), "unit": "Gop/s",
                     "frac": round(achieved / peak, 4), "traffic": leaf_traffic(n),
                     "kernel": "k_leaf_multi (SHA3-256 leaf chains, %d chains x 65 Keccak-f)" % chains,
                     "kernel_ms": round(leaf_ms, 3),
                     "peak_model": "2.4 GHz x 1024 SIMDs; per Keccak-f %d bitwise ops at 2 cycles and %d funnel "
                                   "shifts at 4 cycles per wave64" % (KECCAK_F_BIT_OPS, KECCAK_F_SHIFT_OPS),
                     "single_rate_peak": PEAK_VALU_GOPS},
    }
    # the issue rates the peak assumes, measured on this device (vg_valu_rates)
    try:
        r = V.valu_rates()
        mp = keccak_peak_gops(1.0 / r["bitwise_per_cycle"], 1.0 / r["shift_per_cycle"], CLOCK_GHZ)
        line["roofline"]["measured_issue"] = {k: round(v, 3) for k, v in r.items()}
        line["roofline"]["frac_at_measured_issue"] = round(achieved / mp, 4)
        # the same at the shader clock measured under load (below the published
        # 2.4 GHz): what is left is issue stalls, not clock
        mc = keccak_peak_gops(1.0 / r["bitwise_per_cycle"], 1.0 / r["shift_per_cycle"], r["clock_mhz"] / 1e3)
        line["roofline"]["frac_at_measured_issue_and_clock"] = round(achieved / mc, 4)
    except V.VirgoError:
        pass
    # the reference's own accounting (zk_verifier.cpp:1232 adds the prover's
    # and the commitment prover's total_time, vpd_verifier.cpp:74 and
    # fft_circuit_GKR.cpp:812-823 the fft_gkr prover time; evaluate is not
    # counted): the same parts from the device stages.  The stages overlap on
    # the device, so the sum exceeds the proof's wall time.
    sd = dict(zip(names, (float(x) for x in st)))
    commit = sum(sd[k] for k in ("lde_l", "lde_qh", "fri_fold", "hash", "queries"))
    line["config"]["reference_accounting_ms"] = {
        "zk_prover (V_res + sumcheck)": round(sd["gkr"], 3), "poly_commit_prover (commit + FRI + queries)": round(commit, 3),
        "fft_gkr prover": round(sd["fft_gkr"], 3), "sum": round(sd["gkr"] + commit + sd["fft_gkr"], 3)}
    # the NTT autotuner's choice for this proof's transform sizes
    line["config"]["ntt_tiles"] = {"2^%d" % k: V.ntt_plan(k)[0] for k in (n - 6, n - 5)}
    line["config"]["ntt_tile_us"] = {"2^%d" % k: V.ntt_plan(k)[1] for k in (n - 6, n - 5)}
    # the device verifier on the same transcript (the sumcheck round checks
    # on the host, predicates, public array, query evaluations and paths on
    # the device)
    if not args.no_verify:
        p.verify(tr)
        vt = []
        for _ in range(3):
            ok, why, vms = p.verify(tr)
            vt.append(vms)
        line["verify"] = {"accepted": bool(ok), "device_verifier_ms": round(min(vt), 2)}
    p.close()
    if not args.no_c1:
        line["c1"] = c1(args)
        progress("c1")
    if not args.no_batch:
        line["batched"] = batched(args)
        progress("batched")
    if not args.no_serve:
        line["serving"] = serving(args)
        progress("serving")
    if not args.no_microset:
        line["microset"] = microset(args)
        progress("microset")
    if cpu_1:
        line["cpu_baseline"] = cpu_1
        if cpu_tr is not None:
            line["cpu_baseline"]["bit_exact_vs_gpu"] = cpu_tr == tr
        if cpu_b:
            line["batched"]["cpu_baseline"] = cpu_b
            line["batched"]["speedup_vs_cpu"] = round(line["batched"]["proofs_per_s"] / cpu_b["value"], 1)
        if not args.no_verify:
            from oracle import oracle as O
            t0 = time.perf_counter()
            ok_o, _ = O.verify(circ, tr)
            line["verify"]["oracle_verifier_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            line["verify"]["oracle_accepted"] = bool(ok_o)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
