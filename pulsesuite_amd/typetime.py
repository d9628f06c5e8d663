"""Time grid record ``ts`` and getters (PSTD3D/typetime.py:64-300), host side."""

from dataclasses import dataclass

import numpy as np
from scipy.constants import c as _c0


@dataclass
class ts:
    """Time grid.  Mirrors PSTD3D/typetime.py:64-134."""
    t: float
    tf: float
    dt: float
    n: int

    def CalcNt(self):
        return int((self.tf - self.t) / self.dt)

    def UpdateT(self, dt):
        self.t += dt

    def UpdateN(self, dn):
        self.n += dn


def GetT(t):
    return t.t


def GetTf(t):
    return t.tf


def GetDt(t):
    return t.dt


def GetN(t):
    return t.n


def CalcNt(t):
    return t.CalcNt()


def UpdateT(t, dt):
    t.UpdateT(dt)


def CalculateOptimalDt(dx, dy, dz, eps_r, safety=0.9):
    """safety * min(dx) / (v sqrt 3)  (PSTD3D/typetime.py:254-275)."""
    v = _c0 / np.sqrt(eps_r)
    return safety * min(dx, dy, dz) / (v * np.sqrt(3.0))
