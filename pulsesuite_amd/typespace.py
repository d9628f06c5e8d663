"""Spatial grid record ``ss`` and getters (PSTD3D/typespace.py:63-465), host side."""

from dataclasses import dataclass

import numpy as np

twopi = 2.0 * np.pi


@dataclass
class ss:
    """Spatial grid.  Mirrors PSTD3D/typespace.py:63-73."""
    Dims: int
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    epsr: float


def GetNx(s):
    return s.Nx


def GetNy(s):
    return s.Ny


def GetNz(s):
    return s.Nz


def GetDx(s):
    """1.0 on a collapsed axis (typespace.py:153-178)."""
    return 1.0 if s.Nx == 1 else s.dx


def GetDy(s):
    return 1.0 if s.Ny == 1 else s.dy


def GetDz(s):
    return 1.0 if s.Nz == 1 else s.dz


def GetEpsr(s):
    """1.0 when Nz == 1, as the reference does (typespace.py:237-263)."""
    return 1.0 if s.Nz == 1 else s.epsr


def GetXWidth(s):
    return s.dx * (s.Nx - 1)


def GetYWidth(s):
    return s.dy * (s.Ny - 1)


def GetZWidth(s):
    return s.dz * (s.Nz - 1)


def GetSpaceArray(N, L):
    """N points over [-L/2, L/2] (typespace.py:352-359)."""
    return np.array([0.0]) if N == 1 else np.linspace(-L / 2.0, L / 2.0, N)


def GetKArray(N, L):
    """FFT-order wavenumbers with +N/2 at index N/2 (typespace.py:362-381)."""
    if N == 1:
        return np.array([0.0])
    dl = twopi / L if L > 0 else 1.0
    i = np.arange(N)
    return np.where(i <= N // 2, i, i - N).astype(np.float64) * dl


def GetXArray(s):
    return np.array([0.0]) if s.Nx == 1 else GetSpaceArray(s.Nx, GetXWidth(s))


def GetYArray(s):
    return np.array([0.0]) if s.Ny == 1 else GetSpaceArray(s.Ny, GetYWidth(s))


def GetZArray(s):
    return np.array([0.0]) if s.Nz == 1 else GetSpaceArray(s.Nz, GetZWidth(s))


def GetKxArray(s):
    return np.array([0.0]) if s.Nx == 1 else GetKArray(s.Nx, GetXWidth(s))


def GetKyArray(s):
    return np.array([0.0]) if s.Ny == 1 else GetKArray(s.Ny, GetYWidth(s))


def GetKzArray(s):
    return np.array([0.0]) if s.Nz == 1 else GetKArray(s.Nz, GetZWidth(s))


def GetDVol(s):
    return GetDx(s) * GetDy(s) * GetDz(s)
