"""CPML for the PSTD solver on the GPU (drop-in for PSTD3D/cpml.py:340-818).

``CPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr, npml=None)`` keeps the 12 auxiliary psi
fields (each on the PML slabs of its derivative axis, where it can be nonzero) and a
6-field scratch on the device; ``ApplyCPML_E`` /
``ApplyCPML_B`` take k-space complex128 CUDA tensors and correct them in place
(spectral derivatives, real-space recursion, FFT back), as the reference does.
``npml`` overrides the reference's 5 %-of-the-grid rule on every axis that has
a PML (N >= 32).
"""

import torch

from . import _lib
from ._grid import GridContext
from .typespace import ss


class CPML:
    def __init__(self, Nx, Ny, Nz, dx, dy, dz, dt, epsr=1.0, npml=None, device=None):
        space = ss(3, Nx, Ny, Nz, dx, dy, dz, epsr)
        self.ctx = GridContext(space, dt, npml=npml, with_cpml=True, device=device)
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.npml_x, self.npml_y, self.npml_z = self.ctx.npml
        n = self.ctx.n
        dev = self.ctx.device
        # psi on the PML slabs of their derivative axes only (zero everywhere else)
        npsi = max(1, int(_lib.lib.pss_cpml_psi_len(self.ctx.gp)))
        self.psi_B = torch.zeros(npsi, dtype=torch.float64, device=dev)
        self.psi_E = torch.zeros(npsi, dtype=torch.float64, device=dev)
        self.scratch = torch.empty(6 * n, dtype=torch.complex128, device=dev)

    def GetCPMLInfo(self):
        return self.npml_x, self.npml_y, self.npml_z

    def _half(self, which, F, G, dt, v2=None, src=None, sample=None, stream=None):
        g = self.ctx.g
        g.dt = dt
        if v2 is not None:
            g.v2 = v2
        comp, par, srcx, rx, ry, rz = src if src is not None else (-1, None, None, None, None, None)
        sidx, sw, sout = sample if sample is not None else (None, None, None)
        psi = self.psi_E if which == 0 else self.psi_B
        _lib.check(_lib.lib.pss_cpml_half(self.ctx.gp, which, *(_lib.ptr(t) for t in F), *(_lib.ptr(t) for t in G),
                                          _lib.ptr(psi), _lib.ptr(self.scratch), comp, par, _lib.ptr(srcx),
                                          _lib.ptr(rx), _lib.ptr(ry), _lib.ptr(rz), _lib.ptr(sidx), _lib.ptr(sw),
                                          _lib.ptr(sout), _lib.stream_ptr(stream)), "pss_cpml_half")

    def UpdateB_ApplyCPML_B(self, Bx, By, Bz, Ex, Ey, Ez, dt, src=None, stream=None):
        """B -= i dt (k x E) and ApplyCPML_B in one call (pss_b_half: the curl term rides on
        the correction's last forward pass)."""
        g = self.ctx.g
        g.dt = dt
        comp, par, srcx, rx, ry, rz = src if src is not None else (-1, None, None, None, None, None)
        P = _lib.ptr
        _lib.check(_lib.lib.pss_b_half(self.ctx.gp, P(Bx), P(By), P(Bz), P(Ex), P(Ey), P(Ez), P(self.psi_B),
                                       P(self.scratch), comp, par, P(srcx), P(rx), P(ry), P(rz),
                                       _lib.stream_ptr(stream)), "pss_b_half")

    def CurlE_ApplyCPML_E(self, Ex, Ey, Ez, Bx, By, Bz, dt, v2, src=None, stream=None):
        """E += i v2 dt (k x B) and ApplyCPML_E in one call (pss_cpml_e_curl: the curl term
        rides on the correction's last forward pass); the current term of the E update is
        left to the caller."""
        g = self.ctx.g
        g.dt = dt
        g.v2 = v2
        comp, par, srcx, rx, ry, rz = src if src is not None else (-1, None, None, None, None, None)
        P = _lib.ptr
        _lib.check(_lib.lib.pss_cpml_e_curl(self.ctx.gp, P(Ex), P(Ey), P(Ez), P(Bx), P(By), P(Bz), P(self.psi_E),
                                            P(self.scratch), comp, par, P(srcx), P(rx), P(ry), P(rz),
                                            _lib.stream_ptr(stream)), "pss_cpml_e_curl")

    def ApplyCPML_E(self, Ex, Ey, Ez, Bx, By, Bz, kx=None, ky=None, kz=None, v2=None, dt=None, **kw):
        """PSTD3D/cpml.py:640-702 (kx, ky, kz are implied by the grid)."""
        self._half(0, (Ex, Ey, Ez), (Bx, By, Bz), self.ctx.dt if dt is None else dt, v2, **kw)

    def ApplyCPML_B(self, Bx, By, Bz, Ex, Ey, Ez, kx=None, ky=None, kz=None, dt=None, **kw):
        """PSTD3D/cpml.py:575-638"""
        self._half(1, (Bx, By, Bz), (Ex, Ey, Ez), self.ctx.dt if dt is None else dt, None, **kw)


_default = None


def InitCPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr=1.0, npml=None):
    global _default
    _default = CPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr, npml)


def ApplyCPML_E(Ex, Ey, Ez, Bx, By, Bz, kx, ky, kz, v2, dt):
    if _default is None:
        raise RuntimeError("Call InitCPML before ApplyCPML_E.")
    _default.ApplyCPML_E(Ex, Ey, Ez, Bx, By, Bz, kx, ky, kz, v2, dt)


def ApplyCPML_B(Bx, By, Bz, Ex, Ey, Ez, kx, ky, kz, dt):
    if _default is None:
        raise RuntimeError("Call InitCPML before ApplyCPML_B.")
    _default.ApplyCPML_B(Bx, By, Bz, Ex, Ey, Ez, kx, ky, kz, dt)


def GetCPMLInfo():
    if _default is None:
        raise RuntimeError("Call InitCPML before GetCPMLInfo.")
    return _default.GetCPMLInfo()
