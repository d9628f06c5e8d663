"""Positional parameter files of the field path: space, time and pulse records.

Readers and writers of PSTD3D/typespace.py:76-119, typetime.py:310-346 and
typepulse.py:315-355, host side.  One value per line, the first token of the line; blank
lines and lines starting with '#' or '!' are skipped; a trailing ':' is dropped and Fortran
exponents ('1.5d-9', '2D3') are read as 'e' (GetFileParam, PSTD3D/typetime.py:31-50).
Writers use the E25.15E3-like layout of the reference with its comments.
"""

import os

from .typepulse import ps
from .typespace import ss
from .typetime import ts

_PFRMTA = "{:25.15E}"


def GetFileParam(f):
    """Next numeric parameter of an open text file (PSTD3D/typetime.py:31-50)."""
    while True:
        line = f.readline()
        if not line:
            raise ValueError("Unexpected end of file while reading parameter")
        stripped = line.strip()
        if not stripped or stripped[0] in ("#", "!"):
            continue
        token = stripped.split()[0].rstrip(":")
        token = token.replace("d", "e").replace("D", "E")
        try:
            return float(token)
        except ValueError as exc:
            raise ValueError(f"Could not parse parameter from line: {line!r}") from exc


def readspaceparams_sub(f, space):
    """Dims, Nx, Ny, Nz, dx, dy, dz, epsr (PSTD3D/typespace.py:76-85)."""
    space.Dims = int(GetFileParam(f))
    space.Nx = int(GetFileParam(f))
    space.Ny = int(GetFileParam(f))
    space.Nz = int(GetFileParam(f))
    space.dx = GetFileParam(f)
    space.dy = GetFileParam(f)
    space.dz = GetFileParam(f)
    space.epsr = GetFileParam(f)


def ReadSpaceParams(filename, space=None):
    """Read a space parameter file into ``space`` (a new ss when None); returns it."""
    space = space if space is not None else ss(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0)
    with open(filename, "r", encoding="utf-8") as f:
        readspaceparams_sub(f, space)
    return space


def WriteSpaceParams_sub(f, space):
    """PSTD3D/typespace.py:95-108"""
    f.write(f"{space.Dims:25d} : Number of dimensions.\n")
    f.write(f"{space.Nx:25d} : Number of space X points.\n")
    f.write(f"{space.Ny:25d} : Number of space Y points.\n")
    f.write(f"{space.Nz:25d} : Number of space Z points.\n")
    f.write(f"{space.dx:25.15E} : The Width of the X pixel. (m)\n")
    f.write(f"{space.dy:25.15E} : The Width of the Y pixel. (m)\n")
    f.write(f"{space.dz:25.15E} : The Width of the Z pixel. (m)\n")
    f.write(f"{space.epsr:25.15E} : The relative background dielectric constant\n")


def writespaceparams(filename, space):
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "w", encoding="utf-8") as f:
        WriteSpaceParams_sub(f, space)


def readtimeparams_sub(f, time):
    """t, tf, dt, n (PSTD3D/typetime.py:310-318)."""
    time.t = float(GetFileParam(f))
    time.tf = float(GetFileParam(f))
    time.dt = float(GetFileParam(f))
    time.n = int(GetFileParam(f))


def ReadTimeParams(cmd, time=None):
    time = time if time is not None else ts(0.0, 0.0, 0.0, 0)
    with open(cmd, "r", encoding="utf-8") as f:
        readtimeparams_sub(f, time)
    return time


def WriteTimeParams_sub(f, time):
    """PSTD3D/typetime.py:331-336"""
    f.write(f"{_PFRMTA.format(time.t)}  : Current time of simulation. (s)\n")
    f.write(f"{_PFRMTA.format(time.tf)} : Final time of simulation. (s)\n")
    f.write(f"{_PFRMTA.format(time.dt)} : Time pixel size [dt]. (s)\n")
    f.write(f"{time.n:25d} : Current time index.\n")


def writetimeparams(cmd, time):
    with open(cmd, "w", encoding="utf-8") as f:
        WriteTimeParams_sub(f, time)


def readpulseparams_sub(f, pulse):
    """lambda, Amp, Tw, Tp, chirp (PSTD3D/typepulse.py:315-325)."""
    pulse.lambda_ = float(GetFileParam(f))
    pulse.Amp = float(GetFileParam(f))
    pulse.Tw = float(GetFileParam(f))
    pulse.Tp = float(GetFileParam(f))
    pulse.chirp = float(GetFileParam(f))


def ReadPulseParams(cmd, pulse=None):
    pulse = pulse if pulse is not None else ps(0.0, 0.0, 0.0, 0.0, 0.0)
    with open(cmd, "r", encoding="utf-8") as f:
        readpulseparams_sub(f, pulse)
    return pulse


def writepulseparams_sub(f, pulse):
    """PSTD3D/typepulse.py:338-346"""
    f.write(f"{_PFRMTA.format(pulse.lambda_)} : The pulse wavelength. (m)\n")
    f.write(f"{_PFRMTA.format(pulse.Amp)}     : The pulse amplitude. (V/m)\n")
    f.write(f"{_PFRMTA.format(pulse.Tw)}      : The pulsewidth. (s)\n")
    f.write(f"{_PFRMTA.format(pulse.Tp)}      : The time the pulse crosses the origin. (s)\n")
    f.write(f"{_PFRMTA.format(pulse.chirp)}   : The pulse chirp constant. (rad/s^2)\n")


def WritePulseParams(cmd, pulse):
    with open(cmd, "w", encoding="utf-8") as f:
        writepulseparams_sub(f, pulse)
