"""Pins for the PSTD oracle: the reference's own single-mode known answers
(T/test_PSTD3D.py:146-246), k.B = 0, explicit DFT, CPML closed forms, and the
golden fixtures written from the reference's field-path modules
(tests/golden/make_golden.py)."""

import os

import numpy as np
import pytest
from scipy.constants import c as c0
from scipy.constants import mu_0 as mu0

from oracle import pstd as P
from oracle.helpers import rescale_1D_dpc, rk4

GOLD = os.path.join(os.path.dirname(__file__), "golden", "field_path.npz")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def small_space():
    return P.Space(3, 16, 12, 8, 1e-7, 1e-7, 1e-7, 4.0)


def test_updateE_single_mode():
    sp = small_space()
    tm = P.Time(0.0, 1e-12, 1e-17, 0)
    sh = (sp.Nx, sp.Ny, sp.Nz)
    z = lambda: np.zeros(sh, complex)  # noqa: E731
    Bz = z()
    Bz[1, 0, 0] = 1.0
    Ex, Ey, Ez = z(), z(), z()
    P.UpdateE3D(sp, tm, z(), z(), Bz, z(), z(), z(), Ex, Ey, Ez)
    kx = P.GetKxArray(sp)[:, None, None]
    assert np.allclose(Ey, 1j * (-kx * Bz) * (c0**2 / 4.0) * tm.dt, rtol=1e-12, atol=1e-12)
    Jx = np.ones(sh, complex)
    Ex = z()
    P.UpdateE3D(sp, tm, z(), z(), z(), Jx, z(), z(), Ex, z(), z())
    assert np.allclose(Ex, -mu0 * (c0**2 / 4.0) * tm.dt * Jx, rtol=1e-12)


def test_updateB_single_mode_and_divB():
    sp = small_space()
    tm = P.Time(0.0, 1e-12, 1e-17, 0)
    sh = (sp.Nx, sp.Ny, sp.Nz)
    z = lambda: np.zeros(sh, complex)  # noqa: E731
    Ey = z()
    Ey[1, 0, 0] = 1.0
    Bx, By, Bz = z(), z(), z()
    P.UpdateB3D(sp, tm, z(), Ey, z(), Bx, By, Bz)
    kx = P.GetKxArray(sp)[:, None, None]
    assert np.allclose(Bz, -1j * kx * Ey * tm.dt, rtol=1e-12, atol=1e-12)
    rng = np.random.default_rng(42)
    Bx, By, Bz = z(), z(), z()
    for _ in range(3):
        E = [rng.standard_normal(sh) + 1j * rng.standard_normal(sh) for _ in range(3)]
        P.UpdateB3D(sp, tm, *E, Bx, By, Bz)
    kx, ky, kz = P.GetKxArray(sp)[:, None, None], P.GetKyArray(sp)[None, :, None], P.GetKzArray(sp)[None, None, :]
    div = kx * Bx + ky * By + kz * Bz
    assert np.abs(div).max() < 1e-10 * np.abs(Bx).max() * np.abs(kx).max()


def test_fft3d_is_dft():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((4, 6, 5)) + 1j * rng.standard_normal((4, 6, 5))
    B = A.copy()
    P.fft_3D(B)
    ref = np.zeros_like(A)
    for (i, j, k), _ in np.ndenumerate(A):
        ph = np.exp(-2j * np.pi * (i * np.arange(4)[:, None, None] / 4 + j * np.arange(6)[None, :, None] / 6
                                   + k * np.arange(5)[None, None, :] / 5))
        ref[i, j, k] = np.sum(A * ph)
    assert np.allclose(B, ref, rtol=1e-12, atol=1e-12)
    P.ifft_3D(B)
    assert np.allclose(B, A, rtol=1e-12, atol=1e-13)


def test_cpml_closed_forms():
    cp = P.CPML(64, 64, 1, 1e-8, 1e-8, 1e-8, 1e-17, npml=12)
    assert (cp.npml_x, cp.npml_y, cp.npml_z) == (12, 12, 0)
    inner = slice(12, 52)
    # interior: sigma = 0, kappa = 1, alpha = alpha_max  ->  b = exp(-alpha_max dt/eps0), c = 0
    from scipy.constants import epsilon_0 as eps0
    assert np.allclose(cp.bx[inner], np.exp(-0.05 * 1e-17 / eps0), rtol=1e-15)
    assert np.all(cp.cx[inner] == 0.0) and np.all(cp.kappa_x[inner] == 1.0)
    assert np.all(cp.cz == 0.0)
    assert P.CalcNPML(128) == 6 and P.CalcNPML(31) == 0 and P.CalcNPML(1024) == 51
    # graded profile: sigma_max (pos)^4 with pos = (npml-1-i+0.5)/npml on the left
    pos = (12 - 1 - 0 + 0.5) / 12
    assert np.isclose(cp.sigma_x[0], P.CalcSigmaMax(12e-8) * pos**4, rtol=1e-14)
    assert np.allclose(cp.sigma_x, cp.sigma_x[::-1])


def test_cpml_golden(gold):
    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    Nx, Ny, Nz = int(Nx), int(Ny), int(Nz)
    sp = P.Space(3, Nx, Ny, Nz, dx, dy, dz, epsr)
    cp = P.CPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr)
    assert (cp.npml_x, cp.npml_y, cp.npml_z) == tuple(gold["cpml_npml"])
    E = [gold["cpml_in_E" + n].copy() for n in "xyz"]
    B = [gold["cpml_in_B" + n].copy() for n in "xyz"]
    kx, ky, kz = P.GetKxArray(sp), P.GetKyArray(sp), P.GetKzArray(sp)
    v2 = c0**2 / epsr
    cp.ApplyCPML_E(*E, *B, kx, ky, kz, v2, dt)
    cp.ApplyCPML_B(*B, *E, kx, ky, kz, dt)
    cp.ApplyCPML_E(*E, *B, kx, ky, kz, v2, dt)
    for i, n in enumerate("xyz"):
        assert rel(E[i], gold["cpml_out_E" + n]) < 1e-12
        assert rel(B[i], gold["cpml_out_B" + n]) < 1e-12


def test_tfsf_ic_golden(gold):
    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    sp = P.Space(3, int(Nx), int(Ny), int(Nz), dx, dy, dz, epsr)
    lam, amp, tw, tp, ch, t = gold["tfsf_pulse"]
    pu = P.Pulse(lam, amp, tw, tp, ch)
    prof = P.InitializeTFSF(sp, pu)
    assert rel(prof, gold["tfsf_profile"]) < 1e-14
    E = gold["tfsf_in"].copy()
    P.UpdateTFSC(E, prof, sp, P.Time(t, 1e-12, dt, 0), pu, amp)
    assert rel(E, gold["tfsf_out"]) < 1e-13
    lam, amp, tw, tp, ch, npx = gold["ic_pulse"]
    Ey = np.zeros(gold["ic_Ey"].shape, complex)
    Bz = np.zeros_like(Ey)
    P.SeedInitialCondition(sp, P.Time(0.0, 1e-12, dt, 0), P.Pulse(lam, amp, tw, tp, ch), Ey, Bz, int(npx))
    assert rel(Ey, gold["ic_Ey"]) < 1e-13 and rel(Bz, gold["ic_Bz"]) < 1e-13


def test_propagator_golden(gold):
    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    sp = P.Space(3, int(Nx), int(Ny), int(Nz), dx, dy, dz, epsr)
    lam, amp, tw, tp, ch, _ = gold["tfsf_pulse"]
    pu = P.Pulse(lam, amp, tw, tp, ch)
    prop = P.Propagator(sp, P.Time(0.0, 1.0, dt, 0), pu, "soft", "cpml")
    prop.run(int(gold["run_nsteps"][0]))
    for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        assert rel(getattr(prop, n), gold["run_" + n]) < 1e-11, n


def test_spline_rk4_golden(gold):
    z1 = rescale_1D_dpc(gold["spl_x0"], gold["spl_z0"], gold["spl_x1"])
    assert rel(z1, gold["spl_z1"]) < 1e-13
    A = gold["rk4_A"]
    y = gold["rk4_y0"].copy()
    D = lambda t, yy: A @ yy + np.sin(t)  # noqa: E731
    rk4(y, D(0.0, y), 0.0, 0.1, D)
    assert rel(y, gold["rk4_y1"]) < 1e-14


def test_absorber_mask_closed_forms():
    """The masking absorber's profile (PSTD3D/absorber.py:62-122): 1 - U0 = 0.5 at the outer
    left cell, exp(-gmax rho^3 dt) with gmax = ln 2 / dt on both layers (the right one's grading
    reaching (npml - 1)/npml at the last cell), 1 inside; thickness rule of :39-59; the 3-D mask
    is the product of the three profiles.  The device module builds the same host profiles."""
    dt = 3e-17
    for N, want in ((4, 0), (32, 3), (64, 6), (128, 6), (256, 12), (1024, 51)):
        assert P.CalcNPML_mask(N) == want, N
    N, p = 128, 12
    m = P.build_1d_mask(N, p, dt)
    assert m[0] == 0.5
    assert np.all(m[p:N - p] == 1.0)
    for i in range(p):
        assert np.isclose(m[i], np.exp(-np.log(2.0) * ((p - 1 - i) / (p - 1)) ** 3), rtol=1e-15)
    for i in range(N - p, N):
        assert np.isclose(m[i], np.exp(-np.log(2.0) * ((i - (N - p)) / p) ** 3), rtol=1e-15)
    assert m[N - 1] > 0.5 and m[N - p] == 1.0  # the right layer is graded one cell short
    a = P.Absorber(32, 8, 64, dt)
    mx, my, mz = (P.build_1d_mask(n, P.CalcNPML_mask(n), dt) for n in (32, 8, 64))
    assert np.array_equal(a.mask, mx[:, None, None] * my[None, :, None] * mz[None, None, :])
    import importlib.util

    if importlib.util.find_spec("torch") is not None:
        from pulsesuite_amd import absorber as A  # host-side profiles only (no device work)

        # the same formula element by element (scalar exp as the reference's loop; the oracle's
        # vectorised exp may differ in the last bit)
        assert np.max(np.abs(A.build_1d_mask(N, p, dt) - m) / m) <= 4e-16
        assert all(A.CalcNPML(n) == P.CalcNPML_mask(n) for n in (4, 5, 31, 64, 100, 300, 2048))
