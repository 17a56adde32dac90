"""gsap_mdl_noalloc_i32 (the description length on device buffers, no device allocation)
against gsap_mdl_i32 on the same graph and labels, bit for bit, on two calls in a row (the
second on a smaller graph and fewer blocks), and its size limit."""
import numpy as np
import pytest

import gsap_amd
from gsap_amd import Options, workloads

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls the test needs for device buffers of its own (ctypes on
    libamdhip64)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed with {rc}"

    def upload(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        self.check(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1),
                   "hipMemcpy H2D")
        return p.value

    def download(self, ptr, dtype, n):
        C = self.C
        out = np.empty(n, dtype=dtype)
        self.check(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                    C.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def noalloc_S(ctx, g, lab, B, model=0):
    h = Hip()
    try:
        optr, oadj, ow = h.upload(g.out_ptr), h.upload(g.out_adj), h.upload(g.out_w)
        dl = h.upload(np.asarray(lab, np.int32))
        dS = h.upload(np.array([-1.0]))
        h.sync()
        ctx.mdl_noalloc(optr, oadj, ow, dl, g.N, g.E, B, dS, 0, model)
        h.sync()
        return float(h.download(dS, np.float64, 1)[0])
    finally:
        h.sync()
        h.free()


def test_noalloc_mdl_equals_mdl():
    cases = [(workloads.graph("c2"), 128, 5, 0), (workloads.graph("c1"), 11, 6, 1),
             (gsap_amd.generate_dcsbm(300, 4, 1.0, 2.0, 8.0, seed=7), 1, 7, 0)]
    with gsap_amd.Context(0) as ctx:
        for g, B, seed, model in cases:
            lab = np.random.default_rng(seed).integers(0, B, g.N).astype(np.int32)
            print(f"noalloc mdl N={g.N} E={g.E} B={B} model={model}", flush=True)
            want = ctx.mdl(g, lab, B, Options(model=model))
            got = noalloc_S(ctx, g, lab, B, model)
            assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got, want)


def test_noalloc_mdl_rejects_beyond_limit():
    g = workloads.graph("c1")
    lab = np.zeros(g.N, np.int32)
    with gsap_amd.Context(0) as ctx:
        with pytest.raises(gsap_amd.GsapError, match="1..128"):
            noalloc_S(ctx, g, lab, 129)
        with pytest.raises(gsap_amd.GsapError, match="1..128"):
            noalloc_S(ctx, g, lab, 0)
