"""MI355X-native Virgo prover.

Python front end over libvirgo_hip.so (the C ABI declared in
include/virgo.h).  The proving work runs in hand-written HIP kernels for
gfx950; this module only marshals arrays.  There is no CPU fallback: if the
HIP library or the GPU is missing, every entry point raises.
"""
import ctypes
import os
import time

import numpy as np

from . import workloads  # noqa: F401


HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvirgo_hip.so")
FE = np.dtype([("img", "<u8"), ("real", "<u8")])     # prime_field.h:93 layout

EXPORTS = ["vg_init", "vg_last_error", "vg_free", "vg_ctx_create", "vg_ctx_destroy", "vg_ctx_device_bytes",
           "vg_prove", "vg_prove_async", "vg_finish", "vg_output", "vg_fft", "vg_ifft", "vg_sha3_64",
           "vg_merkle_tree", "vg_fri_fold", "vg_fe_mul", "vg_fft_gkr", "vg_stream_create", "vg_stream_destroy", "vg_prove_batch",
           "vg_verify", "vg_server_create", "vg_server_submit", "vg_server_wait", "vg_server_stats",
           "vg_server_destroy", "vg_ntt_plan", "vg_sha256_64", "vg_merkle_tree_sha256", "vg_microset", "vg_fft_noalloc", "vg_layer_values",
           "vg_proof_values", "vg_query_positions", "vg_plan", "vg_valu_rates"]


class VirgoError(RuntimeError):
    pass


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "-j4"], check=True)


def lib():
    """Loads libvirgo_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VirgoError("libvirgo_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.vg_init.argtypes = [ctypes.c_int]
        L.vg_last_error.restype = ctypes.c_char_p
        L.vg_free.argtypes = [vp]
        L.vg_ctx_create.argtypes = [ctypes.POINTER(vp), vp, vp]
        L.vg_ctx_destroy.argtypes = [vp]
        L.vg_ctx_device_bytes.argtypes = [vp]
        L.vg_ctx_device_bytes.restype = ctypes.c_int64
        L.vg_plan.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.vg_valu_rates.argtypes = [vp]
        L.vg_prove.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t), vp]
        L.vg_prove_async.argtypes = [vp]
        L.vg_finish.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t), vp]
        L.vg_output.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
        L.vg_layer_values.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
        L.vg_proof_values.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.vg_query_positions.argtypes = [vp, vp]
        L.vg_verify.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_double)]
        L.vg_fft.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.vg_ifft.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.vg_sha3_64.argtypes = [vp, vp, ctypes.c_size_t]
        L.vg_merkle_tree.argtypes = [vp, ctypes.c_int, vp]
        L.vg_sha256_64.argtypes = [vp, vp, ctypes.c_size_t]
        L.vg_merkle_tree_sha256.argtypes = [vp, ctypes.c_int, vp]
        L.vg_microset.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp, vp,
                                  ctypes.POINTER(ctypes.c_int), vp]
        L.vg_fe_mul.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.vg_fri_fold.argtypes = [vp, ctypes.c_int, vp, vp]
        L.vg_fft_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.vg_ntt_plan.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp]
        L.vg_prove_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.vg_stream_create.argtypes = [ctypes.POINTER(vp)]
        L.vg_stream_destroy.argtypes = [vp]
        L.vg_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        L.vg_server_submit.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64)]
        L.vg_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t), vp]
        L.vg_server_stats.argtypes = [vp, vp]
        L.vg_server_destroy.argtypes = [vp]
        L.vg_fft_gkr.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t),
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise VirgoError("virgo error %d: %s" % (rc, lib().vg_last_error().decode()))


_inited = False


def init(device=0):
    global _inited
    if not _inited:
        _check(lib().vg_init(device))
        _inited = True


class _Desc(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int), ("n", ctypes.c_void_p), ("ty", ctypes.c_void_p),
                ("u", ctypes.c_void_p), ("v", ctypes.c_void_p)]


STAGES = ["evaluate", "gkr", "lde_l", "lde_qh", "fri_fold", "hash", "queries", "leaf_kernel", "device_total",
          "fft_gkr", "host_prepare", "host_enqueue", "host_assemble"]


class Stream:
    """A HIP stream owned by the library (for Prover(circ, stream=...))."""

    def __init__(self):
        init()
        self.handle = ctypes.c_void_p()
        _check(lib().vg_stream_create(ctypes.byref(self.handle)))

    def close(self):
        if self.handle:
            lib().vg_stream_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class Prover:
    """One proof context: the circuit resident in HBM plus every buffer of a
    proof.  prove() runs the whole Virgo prover on the device and returns the
    transcript bytes (the prover messages in the reference verifier's order)."""

    def __init__(self, circ, stream=None):
        """stream: None (the context owns streams and overlaps its stages) or
        a Stream (the whole proof runs in order on it; contexts on different
        streams overlap -- batched proving)."""
        init()
        if isinstance(stream, Stream):
            stream = stream.handle
        self._keep = tuple(np.ascontiguousarray(circ[k], dtype=t) for k, t in
                           (("n", np.int64), ("ty", np.int32), ("u", np.int64), ("v", np.int64)))
        n, ty, u, v = self._keep
        d = _Desc(len(n), _p(n), _p(ty), _p(u), _p(v))
        self._ctx = ctypes.c_void_p()
        _check(lib().vg_ctx_create(ctypes.byref(self._ctx), ctypes.byref(d), stream))
        self.stage_ms = np.zeros(len(STAGES), dtype=np.float64)

    def device_bytes(self):
        return int(lib().vg_ctx_device_bytes(self._ctx))

    def prove(self, clock=None):
        """Proves the context's circuit; returns the transcript as bytes.
        clock (a list): the seconds vg_prove took (inputs resident on the
        device to the transcript in the library's host buffer) are appended;
        the copy into a Python object after it is not counted."""
        out = ctypes.c_void_p()
        ln = ctypes.c_size_t()
        t0 = time.perf_counter()
        rc = lib().vg_prove(self._ctx, ctypes.byref(out), ctypes.byref(ln), _p(self.stage_ms))
        if clock is not None:
            clock.append(time.perf_counter() - t0)
        _check(rc)
        data = ctypes.string_at(out, ln.value)
        lib().vg_free(out)
        return data

    def prove_async(self):
        _check(lib().vg_prove_async(self._ctx))

    def finish(self):
        out = ctypes.c_void_p()
        ln = ctypes.c_size_t()
        _check(lib().vg_finish(self._ctx, ctypes.byref(out), ctypes.byref(ln), _p(self.stage_ms)))
        data = ctypes.string_at(out, ln.value)
        lib().vg_free(out)
        return data

    def verify(self, transcript):
        """Checks a transcript of this context's circuit with the reference
        verifier's checks (hot loops on the device).  Returns (accepted,
        reason, ms): reason names the failed check when not accepted."""
        ok = ctypes.c_int()
        ms = ctypes.c_double()
        _check(lib().vg_verify(self._ctx, bytes(transcript), len(transcript), ctypes.byref(ok), ctypes.byref(ms)))
        return bool(ok.value), ("" if ok.value else lib().vg_last_error().decode()), ms.value

    def output(self):
        return self.layer(-1)

    def layer(self, idx):
        """Values of layer idx of the padded circuit as last evaluated (0:
        the input layer; -1: the output layer)."""
        L, bits = lib(), ctypes.c_int()
        if idx < 0:
            def fn(out, cap):
                return L.vg_output(self._ctx, out, cap, ctypes.byref(bits))
        else:
            def fn(out, cap):
                return L.vg_layer_values(self._ctx, idx, out, cap, ctypes.byref(bits))
        _check(fn(None, 0))
        buf = np.zeros(1 << bits.value, dtype=FE)
        _check(fn(_p(buf), len(buf)))
        return buf

    VALUES = {"draws": 0, "general": 1, "pub_mask": 2, "pri_mask": 3}

    def proof_values(self, what):
        """vg_proof_values of the last proof: "draws", "general", "pub_mask"
        or "pri_mask"."""
        n = ctypes.c_size_t()
        _check(lib().vg_proof_values(self._ctx, self.VALUES[what], None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, dtype=FE)
        _check(lib().vg_proof_values(self._ctx, self.VALUES[what], _p(buf), n.value, ctypes.byref(n)))
        return buf

    def query_positions(self):
        pos = np.zeros(33, dtype=np.int64)
        _check(lib().vg_query_positions(self._ctx, _p(pos)))
        return pos

    def close(self):
        if self._ctx:
            lib().vg_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "failed", "queued", "running",
                                               "device_in_use", "device_peak", "peak_admitted", "device_budget")]


class Server:
    """Continuous-batching proof server (vg_server_*): requests of any size
    are submitted at any time; each is admitted when a slot and its lease of
    the device budget are free, and proves beside the others on the device.
    submit() returns an id; wait(id) returns (transcript, times) with times
    in ms: latency (submit to transcript), queued (for a slot), plan (host
    planning + waiting for memory), prove (upload + device proof), and the
    device bytes of the request's lease."""

    TIMES = ("latency", "queued", "plan", "prove", "device_bytes")

    def __init__(self, device_budget, slots=16, pinned_budget=0):
        init()
        self._s = ctypes.c_void_p()
        self._keep = {}
        _check(lib().vg_server_create(ctypes.byref(self._s), int(device_budget), int(pinned_budget), int(slots)))

    def submit(self, circ):
        keep = tuple(np.ascontiguousarray(circ[k], dtype=t) for k, t in
                     (("n", np.int64), ("ty", np.int32), ("u", np.int64), ("v", np.int64)))
        n, ty, u, v = keep
        d = _Desc(len(n), _p(n), _p(ty), _p(u), _p(v))
        jid = ctypes.c_int64()
        _check(lib().vg_server_submit(self._s, ctypes.byref(d), ctypes.byref(jid)))
        self._keep[jid.value] = keep        # the arrays stay alive until wait()
        return jid.value

    def wait(self, jid):
        out = ctypes.c_void_p()
        ln = ctypes.c_size_t()
        t = np.zeros(5, dtype=np.float64)
        rc = lib().vg_server_wait(self._s, jid, ctypes.byref(out), ctypes.byref(ln), _p(t))
        self._keep.pop(jid, None)
        _check(rc)
        data = ctypes.string_at(out, ln.value)
        lib().vg_free(out)
        return data, dict(zip(self.TIMES, t.tolist()))

    def stats(self):
        st = _ServerStats()
        _check(lib().vg_server_stats(self._s, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._s:
            lib().vg_server_destroy(self._s)
            self._s = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _desc(circ):
    keep = tuple(np.ascontiguousarray(circ[k], dtype=t) for k, t in
                 (("n", np.int64), ("ty", np.int32), ("u", np.int64), ("v", np.int64)))
    n, ty, u, v = keep
    return _Desc(len(n), _p(n), _p(ty), _p(u), _p(v)), keep


def plan(circ):
    """Checks the circuit as a context would and plans one proof on the host
    (vg_plan; no GPU needed): returns (device bytes, page-locked host bytes)
    of a context of it on a caller's stream.  Raises VirgoError with the
    loader's message for a bad circuit."""
    d, keep = _desc(circ)
    dev, pin = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().vg_plan(ctypes.byref(d), ctypes.byref(dev), ctypes.byref(pin)))
    del keep
    return dev.value, pin.value


def valu_rates():
    """Measured VALU issue rates (vg_valu_rates): wave instructions per SIMD
    per cycle for three-input bitwise operations and for funnel shifts, and
    the shader clock in MHz during the measurement."""
    init()
    r = np.zeros(3, dtype=np.float64)
    _check(lib().vg_valu_rates(_p(r)))
    return {"bitwise_per_cycle": float(r[0]), "shift_per_cycle": float(r[1]), "clock_mhz": float(r[2])}


def prove(circ):
    p = Prover(circ)
    try:
        return p.prove()
    finally:
        p.close()


def prove_batch(provers, workers, clock=None):
    """Proves every Prover's circuit with `workers` host threads in the
    library (vg_prove_batch); returns the transcripts in order.  clock (a
    list): the seconds vg_prove_batch took are appended (the copies into
    Python objects after it are not counted)."""
    n = len(provers)
    ctxs = (ctypes.c_void_p * n)(*[p._ctx.value for p in provers])
    outs = (ctypes.c_void_p * n)()
    lens = (ctypes.c_size_t * n)()
    t0 = time.perf_counter()
    rc = lib().vg_prove_batch(ctxs, n, workers, outs, lens)
    if clock is not None:
        clock.append(time.perf_counter() - t0)
    res = []
    for k in range(n):
        if outs[k]:
            res.append(ctypes.string_at(outs[k], lens[k]))
            lib().vg_free(outs[k])
        else:
            res.append(None)
    _check(rc)
    return res


def fft_gkr(lg):
    """fft_gkr proof of the FFT/evaluation circuit with 2^lg coefficients:
    returns (transcript section, accepted, proof_size, device ms)."""
    init()
    out = ctypes.c_void_p()
    ln = ctypes.c_size_t()
    acc, ps, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().vg_fft_gkr(lg, ctypes.byref(out), ctypes.byref(ln), ctypes.byref(acc), ctypes.byref(ps),
                            ctypes.byref(ms)))
    data = ctypes.string_at(out, ln.value)
    lib().vg_free(out)
    return data, bool(acc.value), ps.value, ms.value


def fe_mul(a, b):
    init()
    a = np.ascontiguousarray(a, dtype=FE)
    b = np.ascontiguousarray(b, dtype=FE)
    o = np.zeros_like(a)
    _check(lib().vg_fe_mul(_p(a), _p(b), _p(o), len(a)))
    return o


def fft(coef, order, rou):
    init()
    coef = np.ascontiguousarray(coef, dtype=FE)
    o = np.zeros(order, dtype=FE)
    _check(lib().vg_fft(_p(coef), len(coef), order, _p(np.ascontiguousarray(rou, dtype=FE)), _p(o)))
    return o


def ifft(evals, coef_len, rou):
    init()
    evals = np.ascontiguousarray(evals, dtype=FE)
    o = np.zeros(min(coef_len, len(evals)), dtype=FE)
    _check(lib().vg_ifft(_p(evals), coef_len, len(evals), _p(np.ascontiguousarray(rou, dtype=FE)), _p(o)))
    return o


def sha3_64(blocks):
    init()
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
    o = np.zeros((len(blocks), 32), dtype=np.uint8)
    _check(lib().vg_sha3_64(_p(blocks), _p(o), len(blocks)))
    return o


def ntt_plan(log_n):
    """(tile_log, {tile_log: us}) of the NTT autotuner for 2^log_n
    transforms (vg_ntt_plan); tile_log 1 is the one-block transform of
    2^13 and 2^14 values."""
    t = ctypes.c_int()
    us = np.zeros(4, dtype=np.float32)
    _check(lib().vg_ntt_plan(log_n, ctypes.byref(t), _p(us)))
    names = [10, 11, 12, 1]                    # tiles 2^10..2^12; 1: the one-block transform
    return t.value, {names[k]: round(float(us[k]), 1) for k in range(4) if us[k] > 0}


def fri_fold(cw, r):
    """One FRI fold of a natural-order codeword (len a power of two) with
    challenge r: the codeword over the squared domain (vg_fri_fold)."""
    init()
    cw = np.ascontiguousarray(cw, dtype=FE)
    n = len(cw)
    assert n >= 2 and n & (n - 1) == 0
    o = np.zeros(n // 2, dtype=FE)
    _check(lib().vg_fri_fold(_p(cw), n.bit_length() - 1, _p(np.ascontiguousarray(r, dtype=FE).reshape(1)), _p(o)))
    return o


def sha256_64(blocks):
    """SHA-256 of each 64-byte row (vg_sha256_64)."""
    init()
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
    o = np.zeros((len(blocks), 32), dtype=np.uint8)
    _check(lib().vg_sha256_64(_p(blocks), _p(o), len(blocks)))
    return o


HASHES = {"sha3": 0, "sha256": 1}


def merkle_tree(leaves, hash="sha3"):
    """Heap-layout Merkle tree of 32-byte leaves (row 1 = root), node hash
    SHA3-256 (the proofs') or SHA-256."""
    init()
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8).reshape(-1, 32)
    pad = 1
    while pad < len(leaves):
        pad *= 2
    tree = np.zeros((2 * pad, 32), dtype=np.uint8)
    fn = lib().vg_merkle_tree if HASHES[hash] == 0 else lib().vg_merkle_tree_sha256
    _check(fn(_p(leaves), len(leaves), _p(tree)))
    return tree


def microset(coef, rate_log, fri_r, hash="sha3", reps=1):
    """Kernel micro-set on one polynomial (vg_microset): transform, inverse,
    rate 2^-rate_log extension, Merkle tree of the extension's pairs, FRI
    commit down to 32 values.  Returns (roots [L-5, 32] uint8, final 32
    values, inverse returned coef exactly, {stage: mean device ms})."""
    init()
    coef = np.ascontiguousarray(coef, dtype=FE)
    n = len(coef)
    assert n >= 32 and n & (n - 1) == 0
    L = n.bit_length() - 1 + rate_log
    fri_r = np.ascontiguousarray(fri_r, dtype=FE)
    assert len(fri_r) >= L - 5
    roots = np.zeros((L - 5, 32), dtype=np.uint8)
    final = np.zeros(32, dtype=FE)
    ok = ctypes.c_int()
    ms = np.zeros(5, dtype=np.float64)
    _check(lib().vg_microset(n.bit_length() - 1, rate_log, HASHES[hash], _p(coef), _p(fri_r), reps, _p(roots),
                             _p(final), ctypes.byref(ok), _p(ms)))
    return roots, final, bool(ok.value), dict(zip(("fft", "ifft", "extension", "merkle", "fri"), ms.tolist()))
