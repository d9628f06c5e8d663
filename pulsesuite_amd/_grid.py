"""Device-side grid context: k arrays, FFT twiddles and CPML coefficient arrays
for one (Nx, Ny, Nz) grid, packed in the ``pss_grid`` struct of include/pss.h."""

import ctypes as C

import numpy as np
import torch
from scipy.constants import c as c0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import mu_0 as mu0

from . import _lib
from .typespace import GetDx, GetDy, GetDz, GetEpsr, GetKxArray, GetKyArray, GetKzArray

M_PROFILE = 4       # PSTD3D/cpml.py:56-59
KAPPA_MAX = 8.0
ALPHA_MAX = 0.05
R_TARGET = 1.0e-8


def CalcNPML(N):
    """PSTD3D/cpml.py:67-90"""
    if N < 32:
        return 0
    npml = max(6, int(0.05 * N))
    if 2 * npml >= N:
        npml = max(6, N // 4)
    return npml


def CalcSigmaMax(L):
    return -(M_PROFILE + 1) * eps0 * c0 * np.log(R_TARGET) / (2.0 * L)


def CalcCoefficients1D(sigma, kappa, alpha, dt):
    """PSTD3D/cpml.py:109-140"""
    b = np.exp(-(sigma / kappa + alpha) * (dt / eps0))
    denom = sigma * kappa + alpha * kappa**2
    safe = np.abs(denom) > 1.0e-20
    c = np.where(safe, sigma / np.where(safe, denom, 1.0) * (b - 1.0), 0.0)
    return b.astype(np.float64), c.astype(np.float64)


def build_profile_1d(N, h, npml):
    """PSTD3D/cpml.py:455-488"""
    if npml == 0:
        return np.zeros(N), np.ones(N), np.zeros(N)
    L = npml * h
    smax = CalcSigmaMax(L)
    i = np.arange(N, dtype=np.float64)
    pos = np.zeros(N)
    left = i < npml
    right = i >= (N - npml)
    pos[left] = (npml - 1.0 - i[left] + 0.5) / npml
    pos[right] = (i[right] - (N - npml) + 0.5) / npml
    return smax * pos**M_PROFILE, 1.0 + (KAPPA_MAX - 1.0) * pos**M_PROFILE, ALPHA_MAX * (1.0 - pos)


def _pow2(n):
    return n >= 1 and (n & (n - 1)) == 0 and n <= 2048


class GridContext:
    """Everything the field kernels need for one grid and time step."""

    def __init__(self, space, dt, npml=None, with_cpml=True, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        Nx, Ny, Nz = space.Nx, space.Ny, space.Nz
        for n in (Nx, Ny, Nz):
            if not _pow2(n):
                raise ValueError(f"grid axis {n}: the device FFT needs powers of two <= 2048")
        self.shape = (Nx, Ny, Nz)
        self.n = Nx * Ny * Nz
        self.dt = dt
        self.epsr = GetEpsr(space)
        self.v2 = c0**2 / self.epsr
        dev = self.device
        self.k_host = [GetKxArray(space), GetKyArray(space), GetKzArray(space)]
        nps = [CalcNPML(Nx), CalcNPML(Ny), CalcNPML(Nz)]
        if npml is not None:
            nps = [int(npml) if d > 0 else 0 for d in nps]
        if not with_cpml:
            nps = [0, 0, 0]
        self.npml = nps
        hs = [GetDx(space), GetDy(space), GetDz(space)]
        self.prof = [build_profile_1d(N, h, p) for N, h, p in zip(self.shape, hs, nps)]
        coef = [CalcCoefficients1D(s, k, a, dt) for (s, k, a) in self.prof]
        self._keep = {}

        def up(name, arr, dtype=torch.float64):
            t = torch.as_tensor(np.ascontiguousarray(arr), dtype=dtype).to(dev)
            self._keep[name] = t
            return t.data_ptr()

        g = _lib.Grid()
        g.Nx, g.Ny, g.Nz = Nx, Ny, Nz
        g.npml_x, g.npml_y, g.npml_z = nps
        g.dt, g.v2, g.mu0 = dt, self.v2, mu0
        g.kx, g.ky, g.kz = (up("k" + a, k) for a, k in zip("xyz", self.k_host))
        for a, (b, c), (s, kap, al) in zip("xyz", coef, self.prof):
            setattr(g, "b" + a, up("b" + a, b))
            setattr(g, "c" + a, up("c" + a, c))
            setattr(g, "ik" + a, up("ik" + a, 1.0 / kap - 1.0))
        for a, N in zip("xyz", self.shape):
            m = np.arange(max(N // 2, 1))
            tw = np.exp(-2j * np.pi * m / N)
            setattr(g, "tw" + a, up("tw" + a, tw.view(np.float64)))
        self.g = g
        self.gp = C.pointer(g)

    def field(self):
        return torch.zeros(self.shape, dtype=torch.complex128, device=self.device)
