"""Classic RK4 on device arrays with fused stage combines.

Same call surface as libpulsesuite/integrator.py:143-176 (rk4_dp, rk4_dpc,
rk4_3D_dpc): ``t_new = rk4_dpc(y, dydt, t, h, D)`` advances ``y`` in place with
the Numerical Recipes stage order, where ``D(t, y)`` returns dy/dt.  Here y and
dydt are contiguous CUDA float64 / complex128 tensors and D returns a tensor of
the same shape; the four stage combines each run as one kernel pass
(pss_rk4_stage), the last one reading y, dydt, dyt, dym and writing y.
"""

from . import _lib


def _flat_len(t):
    import torch

    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("rk4: arrays must be contiguous CUDA tensors")
    if t.dtype == torch.complex128:
        return 2 * t.numel()
    if t.dtype == torch.float64:
        return t.numel()
    raise TypeError(f"rk4: unsupported dtype {t.dtype} (float64 or complex128)")


def _stage(stage, n, h, y, dydt=None, dyt=None, dym=None, ytmp=None):
    _lib.check(_lib.lib.pss_rk4_stage(stage, n, float(h), _lib.ptr(y), _lib.ptr(dydt), _lib.ptr(dyt),
                                      _lib.ptr(dym), _lib.ptr(ytmp), _lib.stream_ptr()), "pss_rk4_stage")


def _derivative(D, t, y, like):
    out = D(t, y)
    if out.shape != like.shape or out.dtype != like.dtype:
        raise ValueError("rk4: D(t, y) must return an array shaped and typed like y")
    return out.contiguous()


def _rk4(y, dydt, t, h, D):
    import torch

    n = _flat_len(y)
    if _flat_len(dydt) != n or dydt.dtype != y.dtype:
        raise ValueError("rk4: y and dydt differ in shape or dtype")
    ytmp = torch.empty_like(y)
    _stage(0, n, h, y, dydt=dydt, ytmp=ytmp)
    dyt = _derivative(D, t + h / 2.0, ytmp, y)
    _stage(1, n, h, y, dyt=dyt, ytmp=ytmp)
    dym = _derivative(D, t + h / 2.0, ytmp, y)
    _stage(2, n, h, y, dyt=dyt, dym=dym, ytmp=ytmp)
    dyt = _derivative(D, t + h, ytmp, y)
    _stage(3, n, h, y, dydt=dydt, dyt=dyt, dym=dym)
    return t + h


def rk4_dp(y, dydt, t, h, D):
    """Classic RK4 for real ODEs (libpulsesuite/integrator.py:160-162).  y in place; returns t + h."""
    return _rk4(y, dydt, t, h, D)


def rk4_dpc(y, dydt, t, h, D):
    """Classic RK4 for complex ODEs (libpulsesuite/integrator.py:165-167).  y in place; returns t + h."""
    return _rk4(y, dydt, t, h, D)


def rk4_3D_dpc(y, dydt, t, h, D):
    """Classic RK4 for 3-D complex arrays (libpulsesuite/integrator.py:170-172).  y in place; returns t + h."""
    return _rk4(y, dydt, t, h, D)
