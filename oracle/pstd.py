"""Oracle: 3-D PSTD Maxwell propagator with CPML and soft source (FP64 NumPy).

Test infrastructure only (see ``oracle/__init__.py``).  Restates
PSTD3D/PSTD3D.py, PSTD3D/cpml.py, PSTD3D/tfsf.py, core/fftw.py (3-D FFTs) and
the grid getters of PSTD3D/typespace.py, typetime.py, typepulse.py.

Pinning:
  UpdateE3D / UpdateB3D            pinned (single-mode known answers of the
                                   reference's own tests, k.B = 0)
  CPML profiles / coefficients     pinned (closed forms: b=1,c=0 outside the PML)
  absorber mask                    pinned (closed forms: 1 - U0 at the outer left cell,
                                   exp(-gmax rho^3 dt), 1 inside, separability)
  ApplyCPML_E/B, UpdateTFSC,
  SeedInitialCondition, run loop   golden (fixtures from the reference modules)
  fft_3D / ifft_3D                 pinned (explicit DFT at small N)
"""

from dataclasses import dataclass

import numpy as np
from scipy.constants import c as c0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import mu_0 as mu0

twopi = 2.0 * np.pi

M_PROFILE = 4  # PSTD3D/cpml.py:56-59
KAPPA_MAX = 8.0
ALPHA_MAX = 0.05
R_TARGET = 1.0e-8


# --- grid records  PSTD3D/typespace.py:63-73, typetime.py:64-134, typepulse.py:47-190
@dataclass
class Space:
    Dims: int
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    epsr: float


@dataclass
class Time:
    t: float
    tf: float
    dt: float
    n: int


@dataclass
class Pulse:
    lambda_: float
    Amp: float
    Tw: float
    Tp: float
    chirp: float
    pol: int = 0
    w0: float = float("inf")

    def CalcOmega0(self):
        return twopi * c0 / self.lambda_

    def CalcTau(self):
        return self.Tw / np.sqrt(2.0 * np.log(2.0))


def GetDx(s):
    return 1.0 if s.Nx == 1 else s.dx  # typespace.py:153-178


def GetDy(s):
    return 1.0 if s.Ny == 1 else s.dy


def GetDz(s):
    return 1.0 if s.Nz == 1 else s.dz


def GetEpsr(s):
    return 1.0 if s.Nz == 1 else s.epsr  # typespace.py:237-263 (Nz==1 -> 1.0)


def _space_array(N, L):
    return np.array([0.0]) if N == 1 else np.linspace(-L / 2.0, L / 2.0, N)


def _k_array(N, L):
    if N == 1:
        return np.array([0.0])
    dl = twopi / L if L > 0 else 1.0
    i = np.arange(N)
    return np.where(i <= N // 2, i, i - N).astype(np.float64) * dl


def GetXArray(s):
    return _space_array(s.Nx, s.dx * (s.Nx - 1))  # typespace.py:297-302,384-389


def GetYArray(s):
    return _space_array(s.Ny, s.dy * (s.Ny - 1))


def GetZArray(s):
    return _space_array(s.Nz, s.dz * (s.Nz - 1))


def GetKxArray(s):
    return np.array([0.0]) if s.Nx == 1 else _k_array(s.Nx, s.dx * (s.Nx - 1))  # :405-410


def GetKyArray(s):
    return np.array([0.0]) if s.Ny == 1 else _k_array(s.Ny, s.dy * (s.Ny - 1))


def GetKzArray(s):
    return np.array([0.0]) if s.Nz == 1 else _k_array(s.Nz, s.dz * (s.Nz - 1))


def fft_3D(Z):
    """Unnormalised forward 3-D FFT in place.  core/fftw.py:217-221"""
    Z[...] = np.fft.fftn(Z, axes=(0, 1, 2))


def ifft_3D(Z):
    """1/N-normalised inverse 3-D FFT in place.  core/fftw.py:222-226"""
    Z[...] = np.fft.ifftn(Z, axes=(0, 1, 2))


# --- spectral curl  PSTD3D/PSTD3D.py:29-113 ---------------------------------
def UpdateE3D(space, time, Bx, By, Bz, Jx, Jy, Jz, Ex, Ey, Ez):
    v2 = c0**2 / GetEpsr(space)
    dt = time.dt
    qx = GetKxArray(space)[:, None, None]
    qy = GetKyArray(space)[None, :, None]
    qz = GetKzArray(space)[None, None, :]
    coeff = v2 * dt
    Ex += 1j * (qy * Bz - qz * By) * coeff - mu0 * Jx * coeff
    Ey += 1j * (qz * Bx - qx * Bz) * coeff - mu0 * Jy * coeff
    Ez += 1j * (qx * By - qy * Bx) * coeff - mu0 * Jz * coeff


def UpdateB3D(space, time, Ex, Ey, Ez, Bx, By, Bz):
    dt = time.dt
    qx = GetKxArray(space)[:, None, None]
    qy = GetKyArray(space)[None, :, None]
    qz = GetKzArray(space)[None, None, :]
    Bx -= 1j * (qy * Ez - qz * Ey) * dt
    By -= 1j * (qz * Ex - qx * Ez) * dt
    Bz -= 1j * (qx * Ey - qy * Ex) * dt


# --- CPML  PSTD3D/cpml.py:67-140, 340-702 -----------------------------------
def CalcNPML(N):
    if N < 32:
        return 0
    npml = max(6, int(0.05 * N))
    if 2 * npml >= N:
        npml = max(6, N // 4)
    return npml


def CalcSigmaMax(L):
    return -(M_PROFILE + 1) * eps0 * c0 * np.log(R_TARGET) / (2.0 * L)


def CalcCoefficients1D(sigma, kappa, alpha, dt):
    b = np.exp(-(sigma / kappa + alpha) * (dt / eps0))
    denom = sigma * kappa + alpha * kappa**2
    safe = np.abs(denom) > 1.0e-20
    c = np.where(safe, sigma / np.where(safe, denom, 1.0) * (b - 1.0), 0.0)
    return b.astype(np.float64), c.astype(np.float64)


def build_profile_1d(N, h, npml):
    if npml == 0:
        return np.zeros(N), np.ones(N), np.zeros(N)
    L = npml * h
    sigma_max = CalcSigmaMax(L)
    m = float(M_PROFILE)
    i_arr = np.arange(N, dtype=np.float64)
    pos = np.zeros(N)
    left = i_arr < npml
    right = i_arr >= (N - npml)
    pos[left] = (npml - 1.0 - i_arr[left] + 0.5) / npml
    pos[right] = (i_arr[right] - (N - npml) + 0.5) / npml
    sigma = sigma_max * pos**m
    kappa = 1.0 + (KAPPA_MAX - 1.0) * pos**m
    alpha = ALPHA_MAX * (1.0 - pos)
    return sigma, kappa, alpha


class CPML:
    """CPML state; ``npml`` overrides CalcNPML per axis (None = reference rule)."""

    def __init__(self, Nx, Ny, Nz, dx, dy, dz, dt, epsr=1.0, npml=None):
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.dt = dt
        nps = [CalcNPML(Nx), CalcNPML(Ny), CalcNPML(Nz)]
        if npml is not None:
            # override only on axes the reference rule gives a PML (N >= 32)
            nps = [int(npml) if d > 0 else 0 for d in nps]
        self.npml_x, self.npml_y, self.npml_z = nps
        prof = [build_profile_1d(N, h, p) for N, h, p in zip((Nx, Ny, Nz), (dx, dy, dz), nps)]
        (self.sigma_x, self.kappa_x, self.alpha_x) = prof[0]
        (self.sigma_y, self.kappa_y, self.alpha_y) = prof[1]
        (self.sigma_z, self.kappa_z, self.alpha_z) = prof[2]
        self.bx, self.cx = CalcCoefficients1D(self.sigma_x, self.kappa_x, self.alpha_x, dt)
        self.by, self.cy = CalcCoefficients1D(self.sigma_y, self.kappa_y, self.alpha_y, dt)
        self.bz, self.cz = CalcCoefficients1D(self.sigma_z, self.kappa_z, self.alpha_z, dt)
        shape = (Nx, Ny, Nz)
        for name in ("Bxy", "Bxz", "Byx", "Byz", "Bzx", "Bzy", "Exy", "Exz", "Eyx", "Eyz", "Ezx", "Ezy"):
            setattr(self, "psi_" + name, np.zeros(shape))

    def mask(self):
        i = np.arange(self.Nx)
        j = np.arange(self.Ny)
        k = np.arange(self.Nz)
        inx = (i < self.npml_x) | (i >= self.Nx - self.npml_x)
        iny = (j < self.npml_y) | (j >= self.Ny - self.npml_y)
        inz = (k < self.npml_z) | (k >= self.Nz - self.npml_z)
        return inx[:, None, None] | iny[None, :, None] | inz[None, None, :]

    @staticmethod
    def _deriv(F, k1d, axis):
        shape = [1, 1, 1]
        shape[axis] = -1
        out = 1j * k1d.reshape(shape) * F
        ifft_3D(out)
        return out

    def _apply(self, F, G, kx, ky, kz, scale, sign, psis):
        """Shared body of ApplyCPML_E (F=E, G=B, sign=+1, scale=v2*dt) and
        ApplyCPML_B (F=B, G=E, sign=-1, scale=dt).  cpml.py:150-334, 575-702"""
        Fx, Fy, Fz = F
        Gx, Gy, Gz = G
        dGz_dy = self._deriv(Gz, ky, 1).real
        dGy_dz = self._deriv(Gy, kz, 2).real
        dGx_dz = self._deriv(Gx, kz, 2).real
        dGz_dx = self._deriv(Gz, kx, 0).real
        dGy_dx = self._deriv(Gy, kx, 0).real
        dGx_dy = self._deriv(Gx, ky, 1).real
        for A in (Fx, Fy, Fz):
            ifft_3D(A)
        m = self.mask()
        bx, by, bz = self.bx[:, None, None], self.by[None, :, None], self.bz[None, None, :]
        cx, cy, cz = self.cx[:, None, None], self.cy[None, :, None], self.cz[None, None, :]
        ikx = 1.0 / self.kappa_x[:, None, None] - 1.0
        iky = 1.0 / self.kappa_y[None, :, None] - 1.0
        ikz = 1.0 / self.kappa_z[None, None, :] - 1.0
        pxy, pxz, pyx, pyz, pzx, pzy = psis
        pxy[...] = np.where(m, by * pxy + cy * dGz_dy, pxy)
        pxz[...] = np.where(m, bz * pxz + cz * dGy_dz, pxz)
        pyx[...] = np.where(m, bx * pyx + cx * dGz_dx, pyx)
        pyz[...] = np.where(m, bz * pyz + cz * dGx_dz, pyz)
        pzx[...] = np.where(m, bx * pzx + cx * dGy_dx, pzx)
        pzy[...] = np.where(m, by * pzy + cy * dGx_dy, pzy)
        Fx += np.where(m, sign * scale * (iky * dGz_dy + pxy - ikz * dGy_dz - pxz), 0.0)
        Fy += np.where(m, sign * scale * (ikz * dGx_dz + pyz - ikx * dGz_dx - pyx), 0.0)
        Fz += np.where(m, sign * scale * (ikx * dGy_dx + pzx - iky * dGx_dy - pzy), 0.0)

    def ApplyCPML_E(self, Ex, Ey, Ez, Bx, By, Bz, kx, ky, kz, v2, dt, after_real=None):
        psis = (self.psi_Exy, self.psi_Exz, self.psi_Eyx, self.psi_Eyz, self.psi_Ezx, self.psi_Ezy)
        self._apply((Ex, Ey, Ez), (Bx, By, Bz), kx, ky, kz, v2 * dt, 1.0, psis)
        if after_real is not None:
            after_real(Ex, Ey, Ez)
        for A in (Ex, Ey, Ez):
            fft_3D(A)

    def ApplyCPML_B(self, Bx, By, Bz, Ex, Ey, Ez, kx, ky, kz, dt, after_real=None):
        psis = (self.psi_Bxy, self.psi_Bxz, self.psi_Byx, self.psi_Byz, self.psi_Bzx, self.psi_Bzy)
        self._apply((Bx, By, Bz), (Ex, Ey, Ez), kx, ky, kz, dt, -1.0, psis)
        if after_real is not None:
            after_real(Bx, By, Bz)
        for A in (Bx, By, Bz):
            fft_3D(A)


# --- masking absorber  PSTD3D/absorber.py:39-191 ------------------------------
def CalcNPML_mask(N):
    """PSTD3D/absorber.py:39-59"""
    if N <= 4:
        return 0
    npml = max(6, int(0.05 * N))
    if npml > N // 10:
        npml = N // 10
    if 2 * npml >= N:
        npml = max(1, N // 4)
    return npml


def build_1d_mask(N, npml, dt, U0=0.5, m=3):
    """PSTD3D/absorber.py:62-122 (the right layer's grading ends at (npml-1)/npml)."""
    mask = np.ones(N)
    if npml <= 0:
        return mask
    gmax = -np.log(1.0 - U0) / dt
    i = np.arange(N)
    left = i < npml
    right = i >= N - npml
    rho_l = (npml - 1 - i[left]) / max(npml - 1, 1)
    rho_r = (i[right] - (N - npml)) / max(npml, 1)
    mask[left] = np.exp(-gmax * rho_l**m * dt)
    mask[right] = np.exp(-gmax * rho_r**m * dt)
    return mask


class Absorber:
    """InitAbsorber / ApplyAbsorber_E/B (PSTD3D/absorber.py:125-191)."""

    def __init__(self, Nx, Ny, Nz, dt):
        self.npml = (CalcNPML_mask(Nx), CalcNPML_mask(Ny), CalcNPML_mask(Nz))
        mx, my, mz = (build_1d_mask(N, p, dt) for N, p in zip((Nx, Ny, Nz), self.npml))
        self.mask = mx[:, None, None] * my[None, :, None] * mz[None, None, :]

    def apply(self, *F):
        for A in F:
            ifft_3D(A)
            A *= self.mask
            fft_3D(A)


# --- soft source  PSTD3D/tfsf.py:26-142 --------------------------------------
def InitializeTFSF(space, pulse):
    Nx = space.Nx
    x = GetXArray(space)
    dx = GetDx(space)
    xp = x[int(round(Nx * 0.25))]
    sigma_src = 5.0 * dx
    return np.exp(-0.5 * ((x - xp) / sigma_src) ** 2) / (sigma_src * np.sqrt(twopi))


def tfsf_increment(tfsf, space, t_now, pulse, amp):
    """S(x)*E_inc(x,y,z,t)*dx, the additive term of UpdateTFSC (tfsf.py:106-136)."""
    omega0 = pulse.CalcOmega0()
    tau_G = pulse.CalcTau()
    v = c0 / np.sqrt(GetEpsr(space))
    dx = GetDx(space)
    x = GetXArray(space)
    yy = GetYArray(space)
    zz = GetZArray(space)
    tau = t_now - x / v - pulse.Tp
    tau_3 = tau[:, None, None]
    gauss_yz = np.exp(-(yy[None, :, None] ** 2 + zz[None, None, :] ** 2) / pulse.w0**2)
    E_inc = amp * gauss_yz * np.exp(-(tau_3**2) / tau_G**2) * np.cos(omega0 * tau_3 + pulse.chirp * tau_3**2)
    return tfsf[:, None, None] * E_inc * dx


def UpdateTFSC(E, tfsf, space, time, pulse, amp):
    E_real = E.copy()
    ifft_3D(E_real)
    E_real += tfsf_increment(tfsf, space, time.t, pulse, amp)
    fft_3D(E_real)
    E[...] = E_real


# --- IC seeding  PSTD3D/PSTD3D.py:134-226 ------------------------------------
def SeedInitialCondition(space, time, pulse, Ey, Bz, npml_x=0):
    Nx = Ey.shape[0]
    x = GetXArray(space)
    yy = GetYArray(space)
    zz = GetZArray(space)
    omega0 = pulse.CalcOmega0()
    tau_G = pulse.CalcTau()
    v = c0 / np.sqrt(GetEpsr(space))
    sigma_x = v * tau_G
    k_med = omega0 / v
    xi_3 = (x - x[Nx // 2])[:, None, None]
    gauss_yz = np.exp(-(yy[None, :, None] ** 2 + zz[None, None, :] ** 2) / pulse.w0**2)
    Ey_vals = (pulse.Amp * gauss_yz * np.exp(-(xi_3**2) / sigma_x**2)
               * np.exp(1j * (k_med * xi_3 + (pulse.chirp / v**2) * xi_3**2)))
    if npml_x > 0:
        taper = np.ones(Nx)
        for i in range(npml_x):
            taper[i] = 0.5 * (1.0 - np.cos(np.pi * i / npml_x))
        for i in range(Nx - npml_x, Nx):
            taper[i] = 0.5 * (1.0 - np.cos(np.pi * (Nx - 1 - i) / npml_x))
        Ey_vals = Ey_vals * taper[:, None, None]
    Ey[...] = Ey_vals
    Bz[...] = Ey_vals / v
    fft_3D(Ey)
    fft_3D(Bz)


# --- propagator  PSTD3D/PSTD3D.py:233-629 ------------------------------------
class Propagator:
    """Time loop of PSTD3DPropagator.run without file I/O or diagnostics IFFTs.

    ``calc_j(space, time, Ex, Ey, Ez, Jx, Jy, Jz)`` follows the reference hook
    (fields in k-space; J persistent and owned by the callback).
    """

    def __init__(self, space, time, pulse, source_type="ic", boundary_type="cpml", calc_j=None, npml=None,
                 polarization="y"):
        self.space, self.time, self.pulse = space, time, pulse
        # the source field: Ey with Bz = Ey/v (the reference), or Ez with By = -Ez/v
        self.polarization = polarization
        self.source_type, self.boundary_type = source_type, boundary_type
        self.calc_j = calc_j
        shape = (space.Nx, space.Ny, space.Nz)
        self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz, self.Jx, self.Jy, self.Jz = (
            np.zeros(shape, dtype=np.complex128) for _ in range(9))
        self.v = c0 / np.sqrt(GetEpsr(space))
        self.v2 = c0**2 / GetEpsr(space)
        self.kx, self.ky, self.kz = GetKxArray(space), GetKyArray(space), GetKzArray(space)
        self.cpml = None
        self.absorber = None
        if boundary_type == "cpml":
            self.cpml = CPML(space.Nx, space.Ny, space.Nz, GetDx(space), GetDy(space), GetDz(space),
                             time.dt, GetEpsr(space), npml=npml)
        elif boundary_type == "mask":
            self.absorber = Absorber(space.Nx, space.Ny, space.Nz, time.dt)
        self.tfsf = InitializeTFSF(space, pulse) if source_type == "soft" else None
        if source_type == "ic":
            npx = (self.cpml.npml_x if self.cpml is not None else
                   self.absorber.npml[0] if self.absorber is not None else 0)
            if polarization == "y":
                SeedInitialCondition(space, time, pulse, self.Ey, self.Bz, npx)
            else:
                SeedInitialCondition(space, time, pulse, self.Ez, self.By, npx)
                self.By *= -1.0

    def emax_real(self):
        E = [self.Ex.copy(), self.Ey.copy(), self.Ez.copy()]
        for A in E:
            ifft_3D(A)
        return np.sqrt(np.max(np.abs(E[0]) ** 2 + np.abs(E[1]) ** 2 + np.abs(E[2]) ** 2))

    def step(self):
        sp, tm, pu = self.space, self.time, self.pulse
        dt = tm.dt
        Amp = pu.Amp
        UpdateE3D(sp, tm, self.Bx, self.By, self.Bz, self.Jx, self.Jy, self.Jz, self.Ex, self.Ey, self.Ez)
        if self.cpml is not None:
            self.cpml.ApplyCPML_E(self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz,
                                  self.kx, self.ky, self.kz, self.v2, dt)
        elif self.absorber is not None:
            self.absorber.apply(self.Ex, self.Ey, self.Ez)
        if self.source_type == "soft":
            UpdateTFSC(self.Ey if self.polarization == "y" else self.Ez, self.tfsf, sp, tm, pu, Amp)
        tm.t += dt / 2.0
        if self.calc_j is not None:
            self.calc_j(sp, tm, self.Ex, self.Ey, self.Ez, self.Jx, self.Jy, self.Jz)
        UpdateB3D(sp, tm, self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz)
        if self.cpml is not None:
            self.cpml.ApplyCPML_B(self.Bx, self.By, self.Bz, self.Ex, self.Ey, self.Ez,
                                  self.kx, self.ky, self.kz, dt)
        elif self.absorber is not None:
            self.absorber.apply(self.Bx, self.By, self.Bz)
        if self.source_type == "soft":
            if self.polarization == "y":
                UpdateTFSC(self.Bz, self.tfsf, sp, tm, pu, Amp / self.v)
            else:
                UpdateTFSC(self.By, self.tfsf, sp, tm, pu, -Amp / self.v)
        tm.t += dt / 2.0
        tm.n += 1

    def run(self, nsteps, check_every=1):
        for n in range(nsteps):
            if check_every and n % check_every == 0:
                Emax = self.emax_real()
                if np.isnan(Emax) or Emax > 1.0e20:
                    raise RuntimeError(f"NaN or overflow at step {n}, Emax = {Emax}")
            self.step()
