"""Phonon scattering on the GPU: drop-in for PulseSuite's PSTD3D/phonons.py module surface.

``InitializePhonons(ky, Ee, Eh, L, epsr, Gph, Oph)`` (PSTD3D/phonons.py:55-136) sets the
module state: the LO-phonon Bose number N0 = 1/(exp(hbar Oph / kB T) - 1) at T = 77 K and the
Coulomb scale Vscale = hbar Oph epsr (1/epsr_inf - 1/epsr_0), with a one-wire device context
for the rates.  ``MBPE(ne, VC, E1D, Win, Wout)`` / ``MBPH(nh, VC, E1D, Win, Wout)``
(:140-268) add the electron / hole in and out rates to Win, Wout in place:

  Win[k]  += sum_i Vep[k, i] n[i] EP[k, i],   Wout[k] += sum_i Vep[k, i] (1 - n[i]) EP[i, k],
  EP[a, b] = 2 Gph (N0 / ((E[a] - E[b] - hbar Oph)^2 + (hbar Gph)^2)
                    + (N0 + 1) / ((E[a] - E[b] + hbar Oph)^2 + (hbar Gph)^2)),  a != b,
  Vep = VC[:, :, 1] / E1D Vscale  (holes: VC[:, :, 2]),

on the step's own kernel (k_wsum through pss_coll_rates); EP and HP are evaluated where they
are read, not stored.  ``EP``/``HP``/``EPT``/``HPT``/``NO``/``Vscale`` module attributes
give the reference's tables on request (host, for inspection).  VC and E1D must be Toeplitz in
k1 - k2, as the reference's are.
"""

import numpy as np
import torch
from scipy.constants import hbar
from scipy.constants import k as kB

from .SBEs import DeviceWires
from .wire_model import EPSR0_HOST, EPSRINF_HOST, TEMP, MBFlags, explicit_model

_state = {"dev": None, "model": None, "NO": 0.0, "Vscale": 0.0, "Ee": None, "Eh": None, "Gph": 0.0,
          "Oph": 0.0}


def _np(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _add(dst, val):
    if isinstance(dst, torch.Tensor):
        dst += torch.as_tensor(val, dtype=dst.dtype, device=dst.device)
    else:
        dst += val


def InitializePhonons(ky, Ee, Eh, L, epsr, Gph, Oph, device=None):
    """PSTD3D/phonons.py:55-136 (Gph, Oph in rad/s)."""
    mb = MBFlags(Optics=False, Excitons=False, EHs=False, Screened=False, Phonon=True, LF=False, DiagDph=False,
                 OffDiagDph=False)
    m = explicit_model(_np(ky), _np(Ee), _np(Eh), mb, L=L, er=epsr, Oph=float(Oph), Gph=float(Gph), epsr_ph=epsr)
    dev = DeviceWires(m, 1, 0.0, 0.0, "fp64", device)
    dev.force_on()
    _state.update(dev=dev, model=m, NO=float(m.scalars["N0"]), Vscale=float(m.scalars["Vscale"]),
                  Ee=m.arrays["Ee"], Eh=m.arrays["Eh"], Gph=float(Gph), Oph=float(Oph))


def _rates(n, VC, E1D, holes):
    dev = _state["dev"]
    if dev is None:
        raise ValueError("Phonons not initialized. Call InitializePhonons first.")
    Nk = dev.models[0].Nk
    vc = _np(VC)[:, 0, :].T.astype(np.float64)
    dev.workspace_view(0, "vc").copy_(torch.as_tensor(vc.reshape(-1)))
    dev.workspace_view(0, "e1d").copy_(torch.as_tensor(np.real(_np(E1D))[:, 0].astype(np.float64)))
    name = "D" if holes else "C"
    dev.buffer(0, name, "next").diagonal().copy_(torch.as_tensor(np.real(_np(n)).astype(complex)))
    dev.kernel("pss_coll_rates", 2)
    W = dev.workspace_view(0, "W").cpu().numpy().reshape(4, Nk)
    return (W[2], W[3]) if holes else (W[0], W[1])


def MBPE(ne, VC, E1D, Win, Wout):
    """Electron-phonon in / out rates added to Win, Wout (PSTD3D/phonons.py:157-205)."""
    wi, wo = _rates(ne, VC, E1D, False)
    _add(Win, wi)
    _add(Wout, wo)


def MBPH(nh, VC, E1D, Win, Wout):
    """Hole-phonon in / out rates added to Win, Wout (PSTD3D/phonons.py:223-268)."""
    wi, wo = _rates(nh, VC, E1D, True)
    _add(Win, wi)
    _add(Wout, wo)


def FermiDistr(En):
    """PSTD3D/phonons.py:317-340"""
    return 1.0 / (np.exp(np.asarray(En) / kB / TEMP) + 1.0)


def BoseDistr(En):
    """PSTD3D/phonons.py:343-364"""
    return 1.0 / (np.exp(np.asarray(En) / kB / TEMP) - 1.0)


def N00():
    """The LO-phonon Bose number of the initialised module (PSTD3D/phonons.py:367-371)."""
    return _state["NO"]


def _table(E):
    N0, hO, hG, G2 = _state["NO"], hbar * _state["Oph"], hbar * _state["Gph"], 2.0 * _state["Gph"]
    dE = E[:, None] - E[None, :]
    T = (N0 / ((dE - hO) ** 2 + hG**2) + (N0 + 1.0) / ((dE + hO) ** 2 + hG**2)) * G2
    np.fill_diagonal(T, 0.0)
    return T


def __getattr__(name):
    """The reference's module tables (PSTD3D/phonons.py:21-33), built on the host on request."""
    if name in ("_EP", "_EPT", "_HP", "_HPT"):
        if _state["Ee"] is None:
            return None
        T = _table(_state["Ee"] if name[1] == "E" else _state["Eh"])
        return T.T.copy() if name.endswith("T") else T
    if name == "_NO":
        return _state["NO"]
    if name == "_Vscale":
        return _state["Vscale"]
    if name in ("_Temp", "_epsr0", "_epsrINF"):
        return {"_Temp": TEMP, "_epsr0": EPSR0_HOST, "_epsrINF": EPSRINF_HOST}[name]
    raise AttributeError(f"module 'phonons' has no attribute {name!r}")
