"""Masking absorber on the GPU: drop-in for PSTD3D/absorber.py (the reference's default
boundary of PSTD3DPropagator).

Multiplicative mask with polynomial grading (m = 3), 1 - U0 = 0.5 at the outermost left
cell, separable over the axes: mask(i, j, k) = mask_x(i) mask_y(j) mask_z(k).  The 1-D
profiles follow PSTD3D/absorber.py:62-122 (including its right-hand layer, whose grading
variable reaches (npml - 1)/npml at the last cell, not 1).  ``ApplyAbsorber_E`` /
``ApplyAbsorber_B`` take k-space complex128 CUDA tensors and apply the mask in real space
(pss_mask_half: inverse 3-D FFT, mask, forward 3-D FFT), in place.
"""

import numpy as np
import torch

from . import _lib
from ._grid import GridContext
from .typespace import ss

U0_PER_STEP = 0.5  # PSTD3D/absorber.py:18
M_GRADE = 3
FRAC_PML = 0.05


def CalcNPML(N):
    """Absorbing layer thickness (PSTD3D/absorber.py:39-59)."""
    if N <= 4:
        return 0
    npml = max(6, int(FRAC_PML * N))
    if npml > N // 10:
        npml = N // 10
    if 2 * npml >= N:
        npml = max(1, N // 4)
    return npml


def build_1d_mask(N, npml, dt):
    """1-D mask profile of one axis (PSTD3D/absorber.py:62-122)."""
    mask = np.ones(N, dtype=np.float64)
    if npml <= 0:
        return mask
    gamma_max = -np.log(1.0 - U0_PER_STEP) / dt
    for i in range(npml):
        rho = (npml - 1 - i) / max(npml - 1, 1)
        mask[i] = np.exp(-gamma_max * rho**M_GRADE * dt)
    for i in range(N - npml, N):
        rho = (i - (N - npml)) / max(npml, 1)
        mask[i] = np.exp(-gamma_max * rho**M_GRADE * dt)
    return mask


class Absorber:
    """Device masks of one grid (InitAbsorber, PSTD3D/absorber.py:125-158)."""

    def __init__(self, Nx, Ny, Nz, dt, device=None, ctx=None):
        self.ctx = ctx if ctx is not None else GridContext(ss(3, Nx, Ny, Nz, 1.0, 1.0, 1.0, 1.0), dt,
                                                           with_cpml=False, device=device)
        dev = self.ctx.device
        self.npml_x, self.npml_y, self.npml_z = CalcNPML(Nx), CalcNPML(Ny), CalcNPML(Nz)
        self.mask_x = build_1d_mask(Nx, self.npml_x, dt)
        self.mask_y = build_1d_mask(Ny, self.npml_y, dt)
        self.mask_z = build_1d_mask(Nz, self.npml_z, dt)
        self._m = [torch.as_tensor(m, device=dev) for m in (self.mask_x, self.mask_y, self.mask_z)]

    @property
    def mask(self):
        """The 3-D mask (host), as the reference stores it."""
        return self.mask_x[:, None, None] * self.mask_y[None, :, None] * self.mask_z[None, None, :]

    def GetAbsorberInfo(self):
        return self.npml_x, self.npml_y, self.npml_z

    def apply(self, Fx, Fy, Fz, src=None, stream=None):
        """Mask the three k-space fields in place; ``src`` = (component, parameters, profile,
        rx, ry, rz) adds the soft source in the same real-space pass."""
        comp, par, srcx, rx, ry, rz = src if src is not None else (-1, None, None, None, None, None)
        P = _lib.ptr
        _lib.check(_lib.lib.pss_mask_half(self.ctx.gp, P(Fx), P(Fy), P(Fz), *(P(m) for m in self._m), comp, par,
                                          P(srcx), P(rx), P(ry), P(rz), _lib.stream_ptr(stream)), "pss_mask_half")


_state = {"absorber": None}  # module-level state, as the reference keeps it (not thread-safe)


def InitAbsorber(Nx, Ny, Nz, dt):
    """PSTD3D/absorber.py:125-158 (no console report)."""
    _state["absorber"] = Absorber(Nx, Ny, Nz, dt)


def _get():
    a = _state["absorber"]
    if a is None:
        raise RuntimeError("Call InitAbsorber before applying the absorber.")
    return a


def ApplyAbsorber_E(Ex, Ey, Ez):
    """PSTD3D/absorber.py:160-175: spectral -> real -> mask -> spectral, in place."""
    _get().apply(Ex, Ey, Ez)


def ApplyAbsorber_B(Bx, By, Bz):
    """PSTD3D/absorber.py:178-191."""
    _get().apply(Bx, By, Bz)


def GetAbsorberInfo():
    return _get().GetAbsorberInfo()
