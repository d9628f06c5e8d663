"""Write small seeded golden vectors from the reference's field-path modules.

Run in a container where the reference checkout exists (read-only):
    python tests/golden/make_golden.py /root/reference/src
It imports only modules that load without numba/pyfftw (typespace, typetime,
typepulse, cpml, tfsf, PSTD3D, core.fftw, libpulsesuite.spliner/integrator) and
saves inputs/outputs as ``.npz`` fixtures next to this script.  The fixtures
are data; the reference code itself never travels.

Functions covered (none of them is fixed by a closed form alone):
  cpml.CPML.ApplyCPML_E / ApplyCPML_B   (32x64x4 grid, seeded random k-space fields)
  tfsf.InitializeTFSF / UpdateTFSC      (same grid)
  PSTD3D.SeedInitialCondition           (same grid, npml_x from CPML)
  PSTD3D.PSTD3DPropagator (cpml + soft) 3 time steps, no-op calc_j
  spliner.rescale_1D_dpc                (non-uniform knots, complex data)
  integrator._rk4                       (linear ODE system)
"""

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    sys.path.insert(0, src)
    from pulsesuite.PSTD3D import PSTD3D as P
    from pulsesuite.PSTD3D import cpml as CP
    from pulsesuite.PSTD3D import tfsf as TF
    from pulsesuite.PSTD3D.typepulse import ps
    from pulsesuite.PSTD3D.typespace import ss
    from pulsesuite.PSTD3D.typetime import ts
    from pulsesuite.libpulsesuite import integrator, spliner

    rng = np.random.default_rng(1234)
    Nx, Ny, Nz = 32, 64, 4  # x, y >= 32 get a PML from the reference rule; z stays periodic
    dx, dy, dz = 25e-9, 30e-9, 35e-9
    epsr = 2.25
    dt = 0.5 * min(dx, dy, dz) / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    space = ss(3, Nx, Ny, Nz, dx, dy, dz, epsr)

    def rnd():
        return rng.standard_normal((Nx, Ny, Nz)) + 1j * rng.standard_normal((Nx, Ny, Nz))

    out = {}
    # ---- CPML one E and one B correction from seeded k-space fields
    CP.InitCPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr)
    from pulsesuite.PSTD3D.typespace import GetKxArray, GetKyArray, GetKzArray
    kx, ky, kz = GetKxArray(space), GetKyArray(space), GetKzArray(space)
    E = [rnd() for _ in range(3)]
    B = [rnd() * 1e-8 for _ in range(3)]
    for i, n in enumerate("xyz"):
        out["cpml_in_E" + n] = E[i].copy()
        out["cpml_in_B" + n] = B[i].copy()
    v2 = 2.99792458e8**2 / epsr
    CP.ApplyCPML_E(E[0], E[1], E[2], B[0], B[1], B[2], kx, ky, kz, v2, dt)
    CP.ApplyCPML_B(B[0], B[1], B[2], E[0], E[1], E[2], kx, ky, kz, dt)
    CP.ApplyCPML_E(E[0], E[1], E[2], B[0], B[1], B[2], kx, ky, kz, v2, dt)
    for i, n in enumerate("xyz"):
        out["cpml_out_E" + n] = E[i]
        out["cpml_out_B" + n] = B[i]
    out["cpml_npml"] = np.array(CP.GetCPMLInfo())
    out["cpml_grid"] = np.array([Nx, Ny, Nz, dx, dy, dz, epsr, dt])

    # ---- TFSF
    pulse = ps(800e-9, 1e8, 10e-15, 8e-15, 0.0)
    tfsf = TF.InitializeTFSF(space, pulse)
    out["tfsf_profile"] = tfsf
    Ey = rnd()
    out["tfsf_in"] = Ey.copy()
    time = ts(3.3e-15, 1e-12, dt, 0)
    TF.UpdateTFSC(Ey, tfsf, space, time, pulse, 1e8)
    out["tfsf_out"] = Ey
    out["tfsf_pulse"] = np.array([800e-9, 1e8, 10e-15, 8e-15, 0.0, 3.3e-15])

    # ---- IC seeding
    Ey = np.zeros((Nx, Ny, Nz), complex)
    Bz = np.zeros((Nx, Ny, Nz), complex)
    pulse_ic = ps(800e-9, 1e8, 5e-15, 0.0, 1e27)
    P.SeedInitialCondition(space, ts(0.0, 1e-12, dt, 0), pulse_ic, Ey, Bz, 6)
    out["ic_Ey"] = Ey
    out["ic_Bz"] = Bz
    out["ic_pulse"] = np.array([800e-9, 1e8, 5e-15, 0.0, 1e27, 6])

    # ---- 3 propagator steps (cpml + soft source), output dir in a temp folder
    with tempfile.TemporaryDirectory() as td:
        cwd = os.getcwd()
        os.chdir(td)
        try:
            time = ts(0.0, 3 * dt + 0.5 * dt, dt, 0)
            prop = P.PSTD3DPropagator(space, time, pulse, source_type="soft", boundary_type="cpml",
                                      output_dir=td, snapshot_interval=10**9)
            prop.run()
            time = prop.time
        finally:
            os.chdir(cwd)
    for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        out["run_" + n] = getattr(prop, n)
    out["run_nsteps"] = np.array([int(time.tf / dt) + 1])  # run() loops n = 0..CalcNt

    # ---- spline
    x0 = np.sort(rng.uniform(-1e-7, 1e-7, 23))
    z0 = np.sin(3e7 * x0) + 1j * np.cos(2e7 * x0) ** 2
    x1 = np.linspace(-1.2e-7, 1.2e-7, 57)
    z1 = np.zeros(57, complex)
    spliner.rescale_1D_dpc(x0, z0, x1, z1)
    out["spl_x0"], out["spl_z0"], out["spl_x1"], out["spl_z1"] = x0, z0, x1, z1

    # ---- rk4
    A = rng.standard_normal((4, 4))
    y = rng.standard_normal(4)
    out["rk4_A"], out["rk4_y0"] = A, y.copy()
    D = lambda t, yy: A @ yy + np.sin(t)  # noqa: E731
    integrator._rk4(y, D(0.0, y), 0.0, 0.1, D)
    out["rk4_y1"] = y

    np.savez_compressed(os.path.join(HERE, "field_path.npz"), **out)
    print("wrote", os.path.join(HERE, "field_path.npz"), len(out), "arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
