"""GPU: the allocation-free transform vg_fft_noalloc (device inputs and
outputs, working memory in the library's static per-GPU array) against
vg_fft and the oracle, bit for bit, on calls in a row whose sizes shrink and
grow again, so that a result can only be right if each call clears and
rebuilds what the last one left in the shared array."""
import ctypes

import numpy as np
import pytest

import virgo_amd as V
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = (1 << 61) - 1
H2D, D2H = 1, 2


def hip():
    V.init()
    h = ctypes.CDLL("libamdhip64.so.7")          # the runtime libvirgo_hip.so is linked against
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.p = h, ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.h.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


def rand_fe(rng, n):
    a = np.zeros(n, dtype=V.FE)
    a["real"] = rng.integers(0, P, n, dtype=np.uint64)
    a["img"] = rng.integers(0, P, n, dtype=np.uint64)
    return a


def test_fft_noalloc_matches_vg_fft():
    h = hip()
    L = V.lib()
    st = V.Stream()
    big = 1 << 20
    cin, cout = DevBuf(h, 16 * big), DevBuf(h, 16 * big)
    rng = np.random.default_rng(31)
    # (coef_len, order, stream): a large call, a smaller one, the largest allowed, then small again
    for k, (coef_len, order, stream) in enumerate([(1 << 12, 1 << 16, None), (1 << 5, 1 << 9, st.handle),
                                                   (1 << 20, 1 << 20, st.handle), (4, 16, None)]):
        print("call", k, "coef_len", coef_len, "order", order, flush=True)
        c = rand_fe(rng, coef_len)
        w = O.root_of_unity(order.bit_length() - 1)
        cin.put(c)
        rc = L.vg_fft_noalloc(cin.p, coef_len, order, w.ctypes.data, cout.p, stream)
        assert rc == 0, L.vg_last_error().decode()
        assert h.hipStreamSynchronize(stream) == 0
        got = cout.get(np.zeros(order, dtype=V.FE))
        assert np.array_equal(got, V.fft(c, order, w)), k
        if order <= 1 << 16:
            assert np.array_equal(got, O.fft(c, order, w)), k
    w = O.root_of_unity(21)
    assert L.vg_fft_noalloc(cin.p, 1 << 10, 1 << 21, w.ctypes.data, cout.p, None) != 0
    assert "scratch limit" in L.vg_last_error().decode()
    assert L.vg_fft_noalloc(cin.p, 1 << 10, 1 << 9, w.ctypes.data, cout.p, None) != 0
    cin.free()
    cout.free()
    st.close()
