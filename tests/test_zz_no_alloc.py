"""pf_generate_noalloc (the generation entry point that allocates no device
memory) against pf_generate, through the C ABI: bit-identical fields on two
calls in a row, the second with a smaller field, so the second call reuses
the static buffer the first one filled."""
import ctypes

import numpy as np
import pytest

import pf_binding as pfb

pytestmark = pytest.mark.gpu


def desc(dim, cells, **kw):
    kw.setdefault("extensions", [1.0] * dim)
    return pfb.make_desc(dim, cells, **kw)


def _noalloc(lib, plan, seed, first, batch, transform, param, out):
    fn = lib.pf_generate_noalloc
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32,
                   ctypes.c_int32, ctypes.c_double, ctypes.c_void_p]
    return fn(plan.h, seed, first, batch, pfb.TRANSFORMS[transform], param, out.data_ptr())


def test_noalloc_matches_pf_generate_twice_second_smaller():
    import torch
    lib = pfb.load()
    cases = [
        # 3D 64^3 cells (128^3 embedding) FP32, batch 5: the buffer holds several
        # realizations per group
        (desc(3, [64, 64, 64], covariance="matern32", variance=2.0, anisotropy="axiparallel",
              corr_length=[0.1, 0.1, 0.05], approximate=True, precision="fp32"), 5, "logNormal"),
        # then a smaller 2D FP64 field, batch 3
        (desc(2, [96, 80], covariance="exponential", corr_length=0.1, precision="fp64"), 3,
         "none"),
    ]
    for i, (d, batch, tr) in enumerate(cases):
        print(f"no-alloc case {i}: dim {d.dim} batch {batch} {tr}", flush=True)
        p = pfb.Plan(d)
        p.eigs()
        dt = torch.float32 if d.precision == 1 else torch.float64
        got = torch.full((batch, p.domain), float("nan"), dtype=dt, device="cuda")
        torch.cuda.synchronize()
        pfb.check(_noalloc(lib, p, 31, 200 + i, batch, tr, 0.0, got))
        want = p.generate(batch, seed=31, first=200 + i, transform=tr)
        assert np.array_equal(got.cpu().numpy(), want), i
        p.close()


def test_noalloc_rejects_unprepared_plans_and_host_output():
    import torch
    lib = pfb.load()
    p = pfb.Plan(desc(2, [32, 32], covariance="exponential", corr_length=0.1, precision="fp32"))
    out = torch.empty((1, p.domain), dtype=torch.float32, device="cuda")
    # no pf_eigs yet: the call would have to allocate
    assert _noalloc(lib, p, 1, 0, 1, "none", 0.0, out) == pfb.PF_ESTATE
    p.eigs()
    assert _noalloc(lib, p, 1, 0, 1, "none", 0.0, out) == pfb.PF_OK

    class HostBuf:
        def __init__(self, a):
            self.a = a

        def data_ptr(self):
            return self.a.ctypes.data

    host = np.empty((1, p.domain), dtype=np.float32)
    assert _noalloc(lib, p, 1, 0, 1, "none", 0.0, HostBuf(host)) == pfb.PF_EINVAL
    # the plan still generates after the rejected call
    assert _noalloc(lib, p, 1, 0, 1, "none", 0.0, out) == pfb.PF_OK
    p.close()
