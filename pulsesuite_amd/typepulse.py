"""Laser pulse record ``ps`` (PSTD3D/typepulse.py:47-190), host side."""

from dataclasses import dataclass

import numpy as np
from scipy.constants import c as _c0

_twopi = 2.0 * np.pi


@dataclass
class ps:
    lambda_: float
    Amp: float
    Tw: float
    Tp: float
    chirp: float
    pol: int = 0
    w0: float = float("inf")

    def CalcK0(self):
        return _twopi / self.lambda_

    def CalcOmega0(self):
        return _twopi * _c0 / self.lambda_

    def CalcTau(self):
        """Tw / sqrt(2 ln 2)  (typepulse.py:104-106)."""
        return self.Tw / np.sqrt(2.0 * np.log(2.0))


def GetLambda(p):
    return p.lambda_


def GetAmp(p):
    return p.Amp


def GetTw(p):
    return p.Tw


def GetTp(p):
    return p.Tp


def GetChirp(p):
    return p.chirp


def GetW0(p):
    return p.w0
