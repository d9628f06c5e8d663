"""Fused RK4 stage combines (pss_rk4_stage) against the stage expressions of
libpulsesuite/integrator.py:143-157 (bit-exact), and the device RK4 step against
the oracle's rk4 and the closed-form amplification factor of linear ODEs."""

import numpy as np
import pytest

from oracle.helpers import rk4 as ork4

pytestmark = pytest.mark.gpu


def same(a, b):
    return bool(np.array_equal(a, b))


def cu(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_stage_combines_bit_exact(n):
    from pulsesuite_amd import _lib

    rng = np.random.default_rng(n)
    y, dydt, dyt, dym = (rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(4))
    h = 3.7e-3
    Y, K1, K2, KM = cu(y), cu(dydt), cu(dyt), cu(dym)
    T = cu(np.zeros(n, complex))
    s = _lib.stream_ptr()
    P = _lib.ptr
    _lib.check(_lib.lib.pss_rk4_stage(0, 2 * n, h, P(Y), P(K1), None, None, P(T), s), "s0")
    ok = same(T.cpu().numpy(), y + (h / 2.0) * dydt)
    assert ok, "stage 0"
    _lib.check(_lib.lib.pss_rk4_stage(1, 2 * n, h, P(Y), None, P(K2), None, P(T), s), "s1")
    ok = same(T.cpu().numpy(), y + (h / 2.0) * dyt)
    assert ok, "stage 1"
    _lib.check(_lib.lib.pss_rk4_stage(2, 2 * n, h, P(Y), None, P(K2), P(KM), P(T), s), "s2")
    ok = same(T.cpu().numpy(), y + h * dym)
    assert ok, "stage 2 ytmp"
    dym2 = dym + dyt
    ok = same(KM.cpu().numpy(), dym2)
    assert ok, "stage 2 dym"
    _lib.check(_lib.lib.pss_rk4_stage(3, 2 * n, h, P(Y), P(K1), P(K2), P(KM), None, s), "s3")
    ok = same(Y.cpu().numpy(), y + (h / 6.0) * (dydt + dyt + 2.0 * dym2))
    assert ok, "stage 3"


def test_null_operand_rejected():
    from pulsesuite_amd import _lib

    assert _lib.lib.pss_rk4_stage(3, 4, 1.0, None, None, None, None, None, None) != 0
    assert b"null" in _lib.lib.pss_last_error()


@pytest.mark.parametrize("dtype", ["complex", "real"])
def test_rk4_step_matches_oracle_and_closed_form(dtype):
    import torch

    from pulsesuite_amd.integrator import rk4_dp, rk4_dpc

    rng = np.random.default_rng(11)
    n = 4096
    if dtype == "complex":
        lam = -rng.uniform(0, 2, n) + 1j * rng.uniform(-5, 5, n)
        y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        step = rk4_dpc
    else:
        lam = -rng.uniform(0, 2, n)
        y0 = rng.standard_normal(n)
        step = rk4_dp
    h = 0.05
    L = cu(lam)
    Y = cu(y0)
    t = 0.0
    yo = y0.copy()
    to = 0.0
    for _ in range(20):
        t = step(Y, L * Y, t, h, lambda tt, yy: L * yy)
        to = ork4(yo, lam * yo, to, h, lambda tt, yy: lam * yy)
    torch.cuda.synchronize()
    assert t == pytest.approx(to)
    got = Y.cpu().numpy()
    e1 = float(np.linalg.norm(got - yo) / np.linalg.norm(yo))
    assert e1 <= 1e-14, e1
    z = lam * h
    R = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    e2 = float(np.linalg.norm(got - y0 * R**20) / np.linalg.norm(yo))
    assert e2 <= 1e-13, e2
