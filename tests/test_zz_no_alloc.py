"""The no-allocation point sample (pss_ksample_noalloc) against pss_ksample, through the C ABI.

Both launch the same kernels on the same operands, so the samples must agree bit for bit.
Two calls in a row on one GPU, the second on a smaller grid: the library's device array of
partial sums keeps the first call's contents, and the second call must not read them.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(shape, seed):
    import torch

    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.rhoPJ import trilinear
    from pulsesuite_amd.typespace import GetXArray, GetYArray, GetZArray, ss

    sp = ss(3, *shape, 20e-9, 20e-9, 20e-9, 9.1)
    ctx = GridContext(sp, 1e-17, with_cpml=False)
    rng = np.random.default_rng(seed)
    F = [torch.as_tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex128,
                         device="cuda") for _ in range(3)]
    rx, ry, rz = GetXArray(sp), GetYArray(sp), GetZArray(sp)
    x0, y0, z0 = 0.37 * rx[-1], -0.21 * ry[-1] if len(ry) > 1 else 0.0, 0.11 * rz[-1] if len(rz) > 1 else 0.0
    idx, w = trilinear(shape, rx, ry, rz, x0, y0, z0)
    return ctx, F, torch.as_tensor(idx, device="cuda"), torch.as_tensor(w, device="cuda")


def _sample(ctx, F, idx, w, noalloc):
    import torch

    from pulsesuite_amd import _lib

    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    P = _lib.ptr
    st = _lib.stream_ptr()
    if noalloc:
        _lib.check(_lib.lib.pss_ksample_noalloc(ctx.gp, P(F[0]), P(F[1]), P(F[2]), P(idx), P(w), P(out), st),
                   "pss_ksample_noalloc")
    else:
        part = torch.zeros(6144, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib.pss_ksample(ctx.gp, P(F[0]), P(F[1]), P(F[2]), P(idx), P(w), P(part), P(out), st),
                   "pss_ksample")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_sample_equals_allocating_sample():
    first = _case((64, 64, 64), 1)
    second = _case((32, 16, 8), 2)  # smaller grid: fewer partial slots than the first call wrote
    for ctx, F, idx, w in (first, second):
        want = _sample(ctx, F, idx, w, noalloc=False)
        got = _sample(ctx, F, idx, w, noalloc=True)
        assert np.any(want != 0.0)
        assert np.array_equal(got, want), (got, want)


def test_noalloc_sample_refuses_unsupported_grids():
    """Axis lengths beyond what the kernels (and the 6144-double scratch) are laid out for are
    refused on the host before any launch."""
    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import ss

    ctx = GridContext(ss(3, 64, 8, 8, 1.0, 1.0, 1.0, 1.0), 1.0, with_cpml=False)
    ctx.g.Nx = 4096  # host-side struct only: refused before any device work
    assert _lib.lib.pss_ksample_noalloc(ctx.gp, None, None, None, None, None, None, None) != 0
    assert b"powers of two" in _lib.lib.pss_last_error()
    ctx.g.Nx = 96
    assert _lib.lib.pss_ksample_noalloc(ctx.gp, None, None, None, None, None, None, None) != 0
