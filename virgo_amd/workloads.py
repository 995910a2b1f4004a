"""Synthetic circuits in the reference's layered-circuit file layout.

A circuit is a dict {"n": int64[d], "ty": int32[sum n], "u": int64[...],
"v": int64[...]} holding the d file layers of the reference format
(zk_verifier.cpp:15-217): file layer 1 is the input layer (type 3 gates whose
u is the value), each later layer lists its gates in gate order.  The same
arrays feed the HIP prover's C ABI and the text writer below.

Workload randomness is splitmix64 from a stated seed (counter form:
x_i = mix(seed + (i+1) * 0x9E3779B97F4A7C15)).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, stream=0):
    """n outputs of splitmix64 started at seed ^ (stream * golden^2)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03))
        z = np.arange(1, n + 1, dtype=np.uint64)
        z *= GAMMA
        z += base
        t = np.empty_like(z)
        for sh, mul in ((30, 0xBF58476D1CE4E5B9), (27, 0x94D049BB133111EB)):
            np.right_shift(z, np.uint64(sh), out=t)
            z ^= t
            z *= np.uint64(mul)
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z


def _pack(layers):
    n = np.array([len(t) for (t, _, _) in layers], dtype=np.int64)
    ty = np.concatenate([t for (t, _, _) in layers]).astype(np.int32)
    u = np.concatenate([a for (_, a, _) in layers]).astype(np.int64)
    v = np.concatenate([b for (_, _, b) in layers]).astype(np.int64)
    return {"n": n, "ty": ty, "u": u, "v": v}


def matmul_circuit(mat, seed=1):
    """C = A*B over F_p with true matmul wiring: mult gate i*n^2+j*n+k takes
    (A[i][k], B[k][j]); sum gate i*n+j adds [i*n^2+j*n, +n).  Inputs uniform
    in [0, 2^31) (the reference truncates input values through int,
    zk_prover.cpp:60)."""
    n = int(mat)
    assert n & (n - 1) == 0
    vals = (splitmix64(seed, 2 * n * n) >> np.uint64(33)).astype(np.int64)
    inp = (np.full(2 * n * n, 3, np.int32), vals, np.zeros(2 * n * n, np.int64))
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    mult = (np.ones(n ** 3, np.int32), (i * n + k).astype(np.int64), (n * n + k * n + j).astype(np.int64))
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    start = (ii * n * n + jj * n).astype(np.int64)
    summ = (np.full(n * n, 5, np.int32), start, start + n)
    return _pack([inp, mult, summ])


def _mod(r, m):
    """r mod m as int64 (a mask when m is a power of two: the same values)."""
    if m & (m - 1) == 0:
        return (r & np.uint64(m - 1)).astype(np.int64)
    return (r % np.uint64(m)).astype(np.int64)


def wide_inputs(seed, n):
    """n input values uniform in [2^40, 2^63): the reference reads them as
    long long and keeps only the low 32 bits through int u
    (zk_prover.cpp:57-66), so about half become negative ints."""
    return ((splitmix64(seed, n, 1) >> np.uint64(1)) | np.uint64(1 << 40)).astype(np.int64)


def random_circuit(log_width, depth, seed=1, log_inputs=None, v_bits=None, wide=False):
    """Random layered circuit: `depth` compute layers of 2^log_width gates,
    each gate add or mult with probability 1/2, u and v uniform over the
    previous layer; 2^log_inputs inputs uniform in [0, 2^31) (wide: in
    [2^40, 2^63), see wide_inputs).  v_bits: v uniform over the first
    2^v_bits entries only (skewed fan-out)."""
    if log_inputs is None:
        log_inputs = log_width
    W, NI = 1 << log_width, 1 << log_inputs
    if wide:
        vals = wide_inputs(seed, NI)
    else:
        vals = (splitmix64(seed, NI, 1) >> np.uint64(33)).astype(np.int64)
    layers = [(np.full(NI, 3, np.int32), vals, np.zeros(NI, np.int64))]
    prev = NI
    for layer in range(depth):
        r = splitmix64(seed, 3 * W, 2 + layer)
        ty = (r[:W] & np.uint64(1)).astype(np.int32)
        vr = prev if v_bits is None else min(prev, 1 << v_bits)
        u = _mod(r[W:2 * W], prev)
        v = _mod(r[2 * W:], vr)
        layers.append((ty, u, v))
        prev = W
    return _pack(layers)


def write_reference_files(circ, circuit_path, meta_path):
    """Writes the circuit in the reference's text format (circuit + meta)."""
    n = circ["n"]
    off = 0
    with open(circuit_path, "w") as f:
        f.write("%d\n" % len(n))
        for cnt in n:
            f.write("%d " % cnt)
            parts = []
            for g in range(int(cnt)):
                parts.append("%d %d %d %d" % (circ["ty"][off + g], g, circ["u"][off + g], circ["v"][off + g]))
            f.write(" ".join(parts) + "\n")
            off += int(cnt)
    with open(meta_path, "w") as f:
        for _ in n:
            f.write("0 0 1 0 0\n")


def read_reference_files(circuit_path):
    """Parses the reference's circuit text format (meta is not needed: the
    parallel-layer flags only change how the verifier evaluates predicates)."""
    toks = open(circuit_path).read().split()
    pos = 0
    d = int(toks[pos]); pos += 1
    ns, ty, u, v = [], [], [], []
    for _ in range(d):
        cnt = int(toks[pos]); pos += 1
        ns.append(cnt)
        arr = np.array(toks[pos:pos + 4 * cnt], dtype=np.int64).reshape(cnt, 4)
        pos += 4 * cnt
        if not np.array_equal(arr[:, 1], np.arange(cnt)):
            raise ValueError("gates must be in sorted order, and full [0, 2^n - 1]")
        ty.append(arr[:, 0]); u.append(arr[:, 2]); v.append(arr[:, 3])
    return {"n": np.array(ns, np.int64), "ty": np.concatenate(ty).astype(np.int32),
            "u": np.concatenate(u), "v": np.concatenate(v)}


def batch_specs(count=256, seed=7):
    """Sizes and seeds of the batched workload (see batch_jobs), without
    building the circuits: a list of (name, kind, a, b, circuit seed)."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        if rng.integers(0, 2) == 0:
            n = 1 << int(rng.integers(3, 9))
            specs.append(("matmul%d_s%d" % (n, k), "matmul", n, 0, 1000 + k))
        else:
            lw, d = int(rng.integers(10, 19)), int(rng.integers(4, 21))
            specs.append(("random_w%d_d%d_s%d" % (lw, d, k), "random", lw, d, 1000 + k))
    return specs


def make_job(spec):
    """The circuit of one batch_specs entry."""
    _, kind, a, b, seed = spec
    return matmul_circuit(a, seed=seed) if kind == "matmul" else random_circuit(a, b, seed=seed)


def batch_jobs(count=256, seed=7):
    """The batched workload: `count` independent proofs of uneven size, half
    n x n matmul circuits with n log-uniform in {8, ..., 256}, half random
    layered circuits with log2 width uniform in [10, 18] and depth uniform in
    [4, 20] (add/mult 50/50, uniform wiring, inputs in [0, 2^31)).  Returns
    a list of (name, circuit)."""
    return [(sp[0], make_job(sp)) for sp in batch_specs(count, seed)]


def serve_jobs(count=96, seed=11):
    """The serving workload: `count` proof requests, each a random layered
    circuit with log2 width uniform in [10, 20] and depth uniform in [8, 24]
    (add/mult 50/50, uniform wiring, inputs in [0, 2^31)).  Returns a list
    of (name, circuit) in arrival order."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        lw, d = int(rng.integers(10, 21)), int(rng.integers(8, 25))
        jobs.append(("random_w%d_d%d_s%d" % (lw, d, k), random_circuit(lw, d, seed=5000 + k)))
    return jobs


# gate types a circuit description may use (zk_prover.cpp:66-155): 0 add,
# 1 mult, 2 dummy (zero), 5 sum of a range, 6 not, 7 minus, 8 xor, 9 naab,
# 10 relay, 12 sum of a range weighted by powers of 2, 13 bit test (u == v).
# Type 4 (direct relay) is the layer the loader inserts above the inputs
# (verifier_fast_track.cpp:102-106), never written in a description.
MIXED_TYPES = (0, 1, 2, 5, 6, 7, 8, 9, 10, 12, 13)


def mixed_circuit(log_width, depth, seed=1, log_inputs=None):
    """Random layered circuit using every gate type: each gate's type is
    uniform over MIXED_TYPES, u and v uniform over the previous layer within
    each type's rules (sums and power-weighted sums span 1..8 entries, bit
    tests have u == v, not, relay and dummy gates carry v = 0); inputs
    uniform in [0, 2^31)."""
    if log_inputs is None:
        log_inputs = log_width
    W, NI = 1 << log_width, 1 << log_inputs
    vals = (splitmix64(seed, NI, 1) >> np.uint64(33)).astype(np.int64)
    layers = [(np.full(NI, 3, np.int32), vals, np.zeros(NI, np.int64))]
    prev = NI
    types = np.array(MIXED_TYPES, dtype=np.int32)
    for layer in range(depth):
        r = splitmix64(seed, 4 * W, 2 + layer)
        ty = types[(r[:W] % np.uint64(len(types))).astype(np.int64)]
        u = _mod(r[W:2 * W], prev)
        v = _mod(r[2 * W:3 * W], prev)
        span = (r[3 * W:] % np.uint64(8)).astype(np.int64) + 1
        rng_sum = (ty == 5)
        u[rng_sum] = np.minimum(u[rng_sum], prev - span[rng_sum])
        v[rng_sum] = u[rng_sum] + span[rng_sum]                  # [u, v), v <= prev
        rng_exp = (ty == 12)
        u[rng_exp] = np.minimum(u[rng_exp], prev - span[rng_exp])
        v[rng_exp] = u[rng_exp] + span[rng_exp] - 1              # [u, v], v < prev
        v[ty == 13] = u[ty == 13]
        v[(ty == 6) | (ty == 10) | (ty == 2)] = 0
        u[ty == 2] = 0
        layers.append((ty, u, v))
        prev = W
    return _pack(layers)
