"""ibem3d_gnn_predict_noalloc (the prediction that allocates nothing) against ibem3d_gnn_predict, through
the C ABI: device inputs and outputs (allocated here through the HIP runtime the library links), the
activations in the library's file-scope scratch array. Two calls in a row, the second with fewer
spheres, in FP32 and BF16: bit-identical to the allocating entry point; the spheres past the second
call's count keep the first call's values; invalid arguments are refused."""
import ctypes

import numpy as np
import pytest

from ibemproject_amd import engine as E
from ibemproject_amd import learned as Lr

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2
EINVAL = -1


def case(msg):
    print(f"[case] {msg}", flush=True)


def inputs(rng, S):
    edge = rng.standard_normal((S, 16, 8)).astype(np.float32)
    edge[:, :, 6] = (rng.uniform(size=(S, 16)) > 0.2).astype(np.float32)
    edge[:, :, 7] = 0.0
    node = rng.standard_normal((S, 4)).astype(np.float32)
    node[:, 3] = 0.0
    return edge, node


class Hip:
    """The few HIP runtime calls the test needs (the runtime already loaded by the library)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.h.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.check(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def get(self, p, shape, dtype=np.float32):
        a = np.empty(shape, dtype=dtype)
        self.check(self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H), "hipMemcpy D2H")
        return a

    def sync(self):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_all(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_noalloc_matches_predict():
    L = E.load()
    g = Lr.Gnn(seed=11)
    hip = Hip()
    rng = np.random.default_rng(8)
    p = g.params()
    g.set_params(p + 0.05 * rng.standard_normal(p.size).astype(np.float32))
    sizes = (4000, 700)  # the first past two 1664-sphere chunks, the second inside one
    edge0, node0 = inputs(rng, sizes[0])
    try:
        for bf16 in (0, 1):
            d_y = hip.put(np.full((sizes[0], 360), np.nan, dtype=np.float32))
            prev = None
            for S in sizes:
                case(f"predict_noalloc S={S} bf16={bf16}")
                edge, node = (edge0, node0) if S == sizes[0] else inputs(rng, S)
                d_e, d_n = hip.put(edge), hip.put(node)
                rc = L.ibem3d_gnn_predict_noalloc(g.g, d_e, d_n, S, bf16, d_y, None)
                assert rc == 0, L.ibem3d_gnn_last_error().decode()
                hip.sync()
                got = hip.get(d_y, (sizes[0], 360))
                want = g.predict(edge, node, bf16=bool(bf16)) / g.norm
                assert np.array_equal(got[:S], want), (S, bf16, float(np.abs(got[:S] - want).max()))
                if prev is not None:  # the rows past S are the first call's
                    assert np.array_equal(got[S:], prev[S:])
                prev = got
        # refused: no spheres, missing pointers
        assert L.ibem3d_gnn_predict_noalloc(g.g, d_e, d_n, 0, 0, d_y, None) == EINVAL
        assert L.ibem3d_gnn_predict_noalloc(g.g, None, d_n, 10, 0, d_y, None) == EINVAL
    finally:
        hip.sync()
        hip.free_all()
        g.close()
