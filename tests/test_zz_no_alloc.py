"""impes_amg_cycles_noalloc (V-cycle iteration in a static device workspace, no
device allocation) against impes_amg_cycles, both through the C ABI.

The two entry points run the same kernels in the same order on the same
hierarchy, so their results must be identical bits: solution, relative residual
and cycle count. Checked on two no-allocation calls in a row, the second on a
smaller hierarchy (the workspace keeps the first call's contents), for both
hierarchy precisions, with and without an initial guess; then at the workspace's
row limit and one row above it (a clear error, nothing launched)."""
import ctypes as C

import numpy as np
import pytest
import torch

from matrices import impes_like

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def hierarchy(dims, seed, dev):
    import impes_hip as H

    _, d, ox, oy, oz = impes_like(*dims, seed=seed)
    h = H.AmgHierarchy(max_levels=25)
    t = [torch.tensor(a, dtype=torch.float64, device=dev) for a in (d, ox, oy, oz)]
    h.setup_st7(dims, *t)
    return h


def cycles(name, h, precision, b, x0):
    """One call of `name`; x0 None: no initial guess (the output buffer starts as NaN,
    which neither path may read)."""
    from impes_hip._lib import call, ptr, stream_ptr

    x = x0.clone() if x0 is not None else torch.full_like(b, float("nan"))
    rel, done = C.c_double(), C.c_int()
    call(name, h.handle, int(precision), ptr(b), ptr(x), 0 if x0 is None else 1, 2, 1e-9, 6, C.byref(rel),
         C.byref(done), stream_ptr())
    torch.cuda.synchronize()
    return x, rel.value, done.value


@pytest.mark.parametrize("use_x0", [False, True])
@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_cycles_equal_allocating_cycles(dev, precision, use_x0):
    rng = np.random.default_rng(7)
    big = hierarchy((40, 40, 20), 1, dev)
    small = hierarchy((20, 16, 9), 2, dev)
    assert small.n < big.n
    inputs = []
    for h in (big, small):
        b = torch.tensor(rng.standard_normal(h.n), dtype=torch.float64, device=dev)
        x0 = torch.tensor(rng.standard_normal(h.n), dtype=torch.float64, device=dev) if use_x0 else None
        inputs.append((h, b, x0))
    # two no-allocation calls in a row (big, then small), then the allocating path
    got = [cycles("impes_amg_cycles_noalloc", h, precision, b, x0) for h, b, x0 in inputs]
    ref = [cycles("impes_amg_cycles", h, precision, b, x0) for h, b, x0 in inputs]
    for (xg, rg, cg), (xr, rr, cr) in zip(got, ref):
        assert torch.isfinite(xr).all()
        assert torch.equal(xg, xr)
        assert rg == rr and cg == cr
    # and once more on the big hierarchy after the small one
    xg, rg, cg = cycles("impes_amg_cycles_noalloc", *inputs[0][:1], precision, *inputs[0][1:])
    assert torch.equal(xg, ref[0][0]) and rg == ref[0][1] and cg == ref[0][2]


def test_noalloc_cycles_row_limit(dev):
    import impes_hip as H

    at = hierarchy((1024, 1024, 1), 3, dev)
    assert at.n == MAX_ROWS
    b = torch.ones(at.n, dtype=torch.float64, device=dev)
    xg, rg, cg = cycles("impes_amg_cycles_noalloc", at, 32, b, None)
    xr, rr, cr = cycles("impes_amg_cycles", at, 32, b, None)
    assert torch.equal(xg, xr) and rg == rr and cg == cr
    del at
    above = hierarchy((1025, 1024, 1), 3, dev)
    assert above.n == MAX_ROWS + 1024
    out = torch.full_like(torch.ones(above.n, dtype=torch.float64, device=dev), 5.0)
    with pytest.raises(H.ImpesError, match="above the static workspace's maximum"):
        from impes_hip._lib import call, ptr, stream_ptr

        call("impes_amg_cycles_noalloc", above.handle, 32, ptr(torch.ones_like(out)), ptr(out), 0, 1, 1e-9, 2,
             None, None, stream_ptr())
    torch.cuda.synchronize()
    assert bool((out == 5.0).all())
