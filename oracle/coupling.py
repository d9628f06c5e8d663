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
rho (J = dP/dt, as the reference's comment at rhoPJ.py:1182 describes).

Longitudinal-field diagnostics of QuantumWire (``longitudinal=True``): ElongfromRho
(PSTD3D/rhoPJ.py:784-969, k-space loop :596-700) on the placed rho and P, and the
bound-charge diagnostics RhoBAnalytic, RhoBSmooth, RhoBAnSmooth (rhoPJ.py:1235-1285).

Pinning: EAtXYZ, CalcJfx, QWPlacement pinned (closed forms); ElongfromRho pinned by
closed forms (one Fourier mode of rho or P: amplitude, sign and axis of each output,
Gauss's law for RhoB, the q = 0 mode, longitudinality); the RhoB diagnostics pinned by
the product rule on a quadratic line (exact central difference) and by the filter of
one Fourier mode; the composition is build-defined (no reference driver ships) and
parity unpinned.
"""

import numpy as np

from .helpers import locator
from scipy.constants import epsilon_0 as eps0

from .pstd import (GetDx, GetDy, GetDz, GetEpsr, GetKxArray, GetKyArray, GetKzArray, GetXArray, GetYArray,
                   GetZArray, fft_3D, ifft_3D)
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


def QWPlacement(Fwire, gx, gy, gz):
    """Fgrid[i, j, k] = gx[i] gy[j] gz[k] Fwire[i].  PSTD3D/rhoPJ.py:736-781"""
    return (gx * Fwire)[:, None, None] * gy[None, :, None] * gz[None, None, :]


def ElongfromRho(space, Rho, Px, Py, Pz):
    """Longitudinal fields from P and the free charge.  PSTD3D/rhoPJ.py:784-969 (k-space
    loop :596-700).  Px, Py, Pz are transformed to k-space in place, as the reference does.
    Returns (E from P + rho, E from P, E from rho) as three (x, y, z) triples of real-space
    grids, and RhoBSpec (real space)."""
    qx, qy, qz = GetKxArray(space), GetKyArray(space), GetKzArray(space)
    small = (qx[1] - qx[0]) / 1e4
    inv_epsr = 1.0 / (eps0 * GetEpsr(space))
    Rk = np.fft.fftn(Rho)
    for P in (Px, Py, Pz):
        P[...] = np.fft.fftn(P)
    QX, QY, QZ = qx[:, None, None], qy[None, :, None], qz[None, None, :]
    qsq = QX**2 + QY**2 + QZ**2 + small
    qdotP = QX * Px + QY * Py + QZ * Pz
    rho_term = -1j * Rk * inv_epsr / qsq
    P_term = -qdotP * inv_epsr / qsq
    ER = [rho_term * Q for Q in (QX, QY, QZ)]
    EP = [P_term * Q for Q in (QX, QY, QZ)]
    EPR = [(rho_term + P_term) * Q for Q in (QX, QY, QZ)]
    qdotE = QX * EPR[0] + QY * EPR[1] + QZ * EPR[2]
    RhoB = 1j * qdotE / inv_epsr - Rk
    back = lambda F: np.fft.ifftn(F)  # noqa: E731
    return [back(F) for F in EPR], [back(F) for F in EP], [back(F) for F in ER], back(RhoB)


def rhob_diagnostics(space, Pavg, gy, gz, dgy, dgz, RhoBSpec):
    """RhoBAnalytic (product rule, -div P of the placed P), RhoBSmooth, RhoBAnSmooth
    (Gaussian k-space filter of widths 2 dx, 2 dy, 2 dz).  PSTD3D/rhoPJ.py:1235-1285"""
    Nx = space.Nx
    gx = np.ones(Nx)
    dx = GetDx(space)
    dPxdx = (np.roll(Pavg[0], -1) - np.roll(Pavg[0], 1)) / (2.0 * dx)
    An = -QWPlacement(dPxdx, gx, gy, gz)
    An = An - QWPlacement(Pavg[1], gx, dgy, gz)
    An = An - QWPlacement(Pavg[2], gx, gy, dgz)
    s = (2.0 * GetDx(space), 2.0 * GetDy(space), 2.0 * GetDz(space))
    qx, qy, qz = GetKxArray(space), GetKyArray(space), GetKzArray(space)
    filt = np.exp(-0.5 * (s[0]**2 * qx[:, None, None]**2 + s[1]**2 * qy[None, :, None]**2
                          + s[2]**2 * qz[None, None, :]**2))
    smooth = np.fft.ifftn(np.fft.fftn(RhoBSpec) * filt)
    an_smooth = np.fft.ifftn(np.fft.fftn(An) * filt)
    return An, smooth, an_smooth


class OracleCoupling:
    """Nw sub-band wires along x through (x0, y0, z0) feeding J into the grid."""

    def __init__(self, space, lam, Emax0, Nw=1, qw=None, mb=None, y0=0.0, z0=0.0, ay=5e-9, az=5e-9, x0=0.0,
                 current="reference", integrator="leapfrog", longitudinal=False):
        self.current = current
        self.longitudinal = longitudinal
        self.diag = None
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
        self.dgy = -(self.ry - y0) / ay**2 * self.gy
        self.dgz = -(self.rz - z0) / az**2 * self.gz
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
        if self.longitudinal:
            self._diagnostics(space, P, Rho)
        self.Pold = P
        self.RhoOld = Rho
        self.last_line = J
        g = self.gy[None, :, None] * self.gz[None, None, :]
        for comp, Jg in enumerate((Jx, Jy, Jz)):
            Jg[...] = J[comp][:, None, None] * g
            fft_3D(Jg)

    def _diagnostics(self, space, P, Rho):
        """The longitudinal fields and bound-charge diagnostics of QuantumWire
        (PSTD3D/rhoPJ.py:1206-1285) from this call's wire outputs."""
        gx = np.ones(space.Nx)
        Pavg = np.sum(P, axis=2) / float(self.Nw)
        Rgrid = QWPlacement(np.sum(Rho, axis=1) / float(self.Nw), gx, self.gy, self.gz)
        Pg = [QWPlacement(Pavg[c], gx, self.gy, self.gz) for c in range(3)]
        EPR, EP, ER, RhoB = ElongfromRho(space, Rgrid, *Pg)
        An, smooth, an_smooth = rhob_diagnostics(space, Pavg, self.gy, self.gz, self.dgy, self.dgz, RhoB)
        self.diag = {"Rho": Rgrid, "Px": Pg[0], "Py": Pg[1], "Pz": Pg[2],
                     "ExlfromPRho": EPR[0], "EylfromPRho": EPR[1], "EzlfromPRho": EPR[2],
                     "ExlfromP": EP[0], "EylfromP": EP[1], "EzlfromP": EP[2],
                     "ExlfromRho": ER[0], "EylfromRho": ER[1], "EzlfromRho": ER[2],
                     "RhoBSpec": RhoB, "RhoBAnalytic": An, "RhoBSmooth": smooth, "RhoBAnSmooth": an_smooth}
