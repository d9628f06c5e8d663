"""Oracle: wire <-> grid coupling glue (the ``calc_j`` hook of the coupled step).

Test infrastructure only (see ``oracle/__init__.py``).  Follows
PSTD3D/rhoPJ.py:1060-1329 (QWArray.QuantumWire) for the current-density path:
point-sample E at the wire centre (EAtXYZ, PSTD3D/usefulsubs.py:2016-2085),
run every sub-band wire's QWCalculator on the x line, J = sum_w (P_w - P_old_w)/dt / Nw
+ CalcJfx (PSTD3D/rhoPJ.py:561-592, 1183-1194), deposit J(x) g_y(y) g_z(z)
(QWPlacement, PSTD3D/rhoPJ.py:736-781) and FFT it to k-space.

Current semantics (``current=``): "reference" follows the reference's code, which
re-creates PxxOld/PyyOld/PzzOld/RhoOld as zeros on every call (rhoPJ.py:1138-1147)
before forming J, so J = sum_w P_w/dt/Nw; "delta" keeps the previous call's P and
rho (J = dP/dt, as the reference's comment at rhoPJ.py:1182 describes).  The
longitudinal-field diagnostics (ElongfromRho, RhoB*) are not on the J path and are
omitted.

Pinning: EAtXYZ, CalcJfx, QWPlacement pinned (closed forms); the composition is
build-defined (no reference driver ships) and parity unpinned.
"""

import numpy as np

from .helpers import locator
from .pstd import GetDx, GetXArray, GetYArray, GetZArray, GetKxArray, fft_3D, ifft_3D
from .sbe import OracleWire


def EAtXYZ(f, x, y, z, x0, y0, z0):
    """Trilinear sample of a 3-D complex field.  PSTD3D/usefulsubs.py:2016-2085"""
    def axis(a, a0):
        if len(a) == 1:
            return 0, np.array([1.0, 0.0])
        i = locator(a, a0)
        al = (a0 - a[i]) / (a[i + 1] - a[i])
        return i, np.array([1.0 - al, al])
    i, wx = axis(x, x0)
    j, wy = axis(y, y0)
    k, wz = axis(z, z0)
    f0 = 0.0 + 0.0j
    for r in range(2):
        for q in range(2):
            for p in range(2):
                if wx[p] == 0.0 and (i + p) >= f.shape[0]:
                    continue
                if wy[q] == 0.0 and (j + q) >= f.shape[1]:
                    continue
                if wz[r] == 0.0 and (k + r) >= f.shape[2]:
                    continue
                f0 = f0 + wx[p] * wy[q] * wz[r] * f[i + p, j + q, k + r]
    return f0


def CalcJfx(RhoNew, RhoPrev, dt, dx):
    """Free current from -int d(rho)/dt dx.  PSTD3D/rhoPJ.py:561-592"""
    drhodt = (RhoNew - RhoPrev) / dt
    out = np.zeros(len(RhoNew), dtype=np.complex128)
    out[1:] = -np.cumsum(drhodt[1:] * dx)
    return out


class OracleCoupling:
    """Nw sub-band wires along x through (x0, y0, z0) feeding J into the grid."""

    def __init__(self, space, lam, Emax0, Nw=1, qw=None, mb=None, y0=0.0, z0=0.0, ay=5e-9, az=5e-9, x0=0.0,
                 current="reference", integrator="leapfrog"):
        self.current = current
        self.space = space
        self.rx = GetXArray(space)
        self.ry = GetYArray(space)
        self.rz = GetZArray(space)
        self.qx = GetKxArray(space)
        self.Nw = Nw
        self.y0, self.z0, self.x0 = y0, z0, x0
        self.wire = OracleWire(self.rx, 0.0, Emax0, lam, Nw, qw, mb, integrator=integrator)
        Nx = space.Nx
        self.Pold = np.zeros((3, Nx, Nw), dtype=np.complex128)
        self.RhoOld = np.zeros((Nx, Nw), dtype=np.complex128)
        self.gy = np.exp(-((self.ry - y0) ** 2) / (2.0 * ay**2))
        self.gz = np.exp(-((self.rz - z0) ** 2) / (2.0 * az**2))
        self.last_line = None

    def calc_j(self, space, time, Ex, Ey, Ez, Jx, Jy, Jz):
        dt = time.dt
        if self.current == "reference":  # rhoPJ.py:1138-1147: the old arrays start from zero each call
            self.Pold = np.zeros_like(self.Pold)
            self.RhoOld = np.zeros_like(self.RhoOld)
        Er = []
        for F in (Ex, Ey, Ez):
            G = F.copy()
            ifft_3D(G)
            Er.append(G)
        Nx = space.Nx
        P = np.zeros((3, Nx, self.Nw), dtype=np.complex128)
        Rho = np.zeros((Nx, self.Nw), dtype=np.complex128)
        for w in range(self.Nw):
            e = [EAtXYZ(G, self.rx, self.ry, self.rz, self.x0, self.y0, self.z0) for G in Er]
            Exx = np.full(Nx, e[0], dtype=np.complex128)
            Eyy = np.full(Nx, e[1], dtype=np.complex128)
            Ezz = np.full(Nx, e[2], dtype=np.complex128)
            Pxx, Pyy, Pzz, R = self.wire.QWCalculator(Exx, Eyy, Ezz, self.rx, dt, w + 1)
            P[0, :, w], P[1, :, w], P[2, :, w] = Pxx, Pyy, Pzz
            Rho[:, w] = R
        J = np.sum(P - self.Pold, axis=2) / dt / float(self.Nw)
        J[0] = J[0] + CalcJfx(np.sum(Rho, axis=1), np.sum(self.RhoOld, axis=1), dt, GetDx(space)) / float(self.Nw)
        self.Pold = P
        self.RhoOld = Rho
        self.last_line = J
        g = self.gy[None, :, None] * self.gz[None, None, :]
        for comp, Jg in enumerate((Jx, Jy, Jz)):
            Jg[...] = J[comp][:, None, None] * g
            fft_3D(Jg)
