"""sw_integrals_noalloc against sw_integrals (GPU).

The no-allocation variant runs the same kernels with the block partials in
the library's static device array and writes the result to device memory.
Two calls in a row (a C24 grid, then a smaller C12 grid, with no host
synchronisation between them) must give the bits sw_integrals gives for each,
in both precisions: the second call must not see the first call's partials.
"""

from __future__ import annotations

import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _case(torch, nx, dtype, seed):
    from licom4py_gpu_amd import barotr, duogrid, initial
    g = duogrid.grid(nx)
    h, u, v, _ = initial.seeded_balanced(g, seed=seed)
    ctx = barotr.SWContext(duogrid.barotr_consts(g, 5000.0), dtype)
    fields = [torch.as_tensor(x).to(ctx.device, dtype) for x in (h, u, v)]
    return ctx, fields


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_integrals_noalloc_matches_integrals(dtype):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import _hip
    tdt = torch.float64 if dtype == "f64" else torch.float32
    lib = _hip.lib()
    cases = [_case(torch, 24, tdt, 3), _case(torch, 12, tdt, 4)]
    want = [ctx.integrals(*f) for ctx, f in cases]  # sw_integrals: waits for its result
    outs = [torch.full((2,), float("nan"), dtype=torch.float64, device=cases[0][0].device) for _ in cases]
    torch.cuda.synchronize()
    for (ctx, f), out in zip(cases, outs):  # back to back, no synchronisation
        ctx._bind_stream()
        rc = lib.sw_integrals_noalloc(ctx._h, *(ctypes.c_void_p(x.data_ptr()) for x in f),
                                      ctypes.c_void_p(out.data_ptr()))
        _hip.check(rc, "sw_integrals_noalloc")
    torch.cuda.synchronize()
    for w, out in zip(want, outs):
        got = out.cpu().tolist()
        assert got == list(w), (got, w)
    # a second round on the same contexts (the array keeps the last call's partials)
    for (ctx, f), out in zip(reversed(cases), reversed(outs)):
        out.fill_(float("nan"))
        _hip.check(lib.sw_integrals_noalloc(ctx._h, *(ctypes.c_void_p(x.data_ptr()) for x in f),
                                            ctypes.c_void_p(out.data_ptr())), "sw_integrals_noalloc")
    torch.cuda.synchronize()
    for w, out in zip(want, outs):
        assert out.cpu().tolist() == list(w)


def test_integrals_noalloc_rejects_null():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import _hip
    lib = _hip.lib()
    assert lib.sw_integrals_noalloc(None, None, None, None, None) == -1
    assert b"null" in lib.sw_last_error()
