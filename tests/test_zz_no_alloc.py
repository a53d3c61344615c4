"""msw_dot_noalloc (the allocation-free inner product, partials in the
library's file-scope device array) against msw_dot on a context of the same
mesh, batch and precision: bitwise equal, on two calls in a row, the second
with a smaller mesh and batch (the array keeps the first call's partials)."""

import ctypes as C

import numpy as np
import pytest
import torch

from helpers import case
from mpas_sw_amd import _lib
from mpas_sw_amd.core import SWModel

pytestmark = pytest.mark.gpu


def _noalloc(lib, m, xh, xu, yh, yu, wh, wu):
    out = torch.full((m.batch,), float("nan"), dtype=torch.float64, device=m.device)
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _lib.check(lib.msw_dot_noalloc(m.nC, m.nE, m.batch, m._prec, p(xh), p(xu), p(yh), p(yu), p(wh), p(wu),
                                   p(out), C.c_void_p(m.stream.cuda_stream)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_dot_equals_context_dot_on_two_calls(prec):
    lib = _lib.load()
    for level, batch in ((4, 64), (3, 16)):   # the second call is smaller
        print(f"msw_dot_noalloc L{level} B={batch} {prec}", flush=True)
        mesh, phys = case(level, 5)
        m = SWModel(mesh, phys, 300.0, test_case=5, batch=batch, precision=prec)
        g = torch.Generator("cpu").manual_seed(level * 100 + batch)
        dt = m.dtype
        xh = torch.randn((m.nC, batch), generator=g, dtype=torch.float64).to(m.device, dt)
        xu = torch.randn((m.nE, batch), generator=g, dtype=torch.float64).to(m.device, dt)
        yh = torch.randn((m.nC, batch), generator=g, dtype=torch.float64).to(m.device, dt)
        yu = torch.randn((m.nE, batch), generator=g, dtype=torch.float64).to(m.device, dt)
        wh = m.to_internal(np.asarray(phys["areaCell"]), "c", member=False)
        wu = m.to_internal(np.asarray(phys["dcEdge"]) * np.asarray(phys["dvEdge"]), "e", member=False)
        for args in ((xh, xu, yh, yu, wh, wu), (xh, xu, None, None, None, None), (xh, None, yh, None, wh, None),
                     (None, xu, None, yu, None, wu)):
            ref = m.dot(*args)
            got = _noalloc(lib, m, *args)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (level, batch, prec, (got - ref).abs().max().item())
        m.close()


def test_noalloc_dot_rejects_what_the_scratch_cannot_hold():
    lib = _lib.load()
    out = torch.zeros(512, dtype=torch.float64, device="cuda")
    x = torch.zeros((10, 512), dtype=torch.float64, device="cuda")
    rc = lib.msw_dot_noalloc(10, 10, 512, _lib.MSW_FP64, C.c_void_p(x.data_ptr()), None, None, None, None,
                             None, C.c_void_p(out.data_ptr()), None)
    assert rc == _lib.MSW_EINVAL and b"256" in lib.msw_last_error()
    rc = lib.msw_dot_noalloc(10, 10, 3, _lib.MSW_FP64, C.c_void_p(x.data_ptr()), None, None, None, None,
                             None, C.c_void_p(out.data_ptr()), None)
    assert rc == _lib.MSW_EINVAL
