"""3-D FFTs on the GPU with the numpy/pyFFTW convention of core/fftw.py:217-226:
``fft_3D`` unnormalised forward, ``ifft_3D`` scaled by 1/N, both in place on a
contiguous complex128 CUDA tensor.  Runs the LDS-staged radix-2 passes of
libpss (powers of two up to 2048 per axis)."""

import torch

from . import _lib
from ._grid import GridContext

_ctx = {}


def _context(shape, device):
    key = (tuple(shape), str(device))
    if key not in _ctx:
        from .typespace import ss
        Nx, Ny, Nz = shape
        _ctx[key] = GridContext(ss(3, Nx, Ny, Nz, 1.0, 1.0, 1.0, 1.0), 1.0, with_cpml=False, device=device)
    return _ctx[key]


def _check(Z):
    if not (isinstance(Z, torch.Tensor) and Z.is_cuda and Z.dtype == torch.complex128 and Z.is_contiguous()
            and Z.dim() == 3):
        raise TypeError("fft_3D/ifft_3D take a contiguous 3-D complex128 CUDA tensor")


def fft_3D(Z):
    _check(Z)
    ctx = _context(Z.shape, Z.device)
    _lib.check(_lib.lib.pss_fft3d(ctx.gp, _lib.ptr(Z), -1, _lib.stream_ptr()), "pss_fft3d")


def ifft_3D(Z):
    _check(Z)
    ctx = _context(Z.shape, Z.device)
    _lib.check(_lib.lib.pss_fft3d(ctx.gp, _lib.ptr(Z), +1, _lib.stream_ptr()), "pss_fft3d")
