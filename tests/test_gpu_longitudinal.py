"""Longitudinal-field diagnostics of QWArray.QuantumWire on the device against the oracle:
ElongfromRho (PSTD3D/rhoPJ.py:784-969), QWPlacement (:736-781) and the bound-charge
diagnostics RhoBAnalytic / RhoBSmooth / RhoBAnSmooth (:1206-1285).  Relative L2 bars:
1e-12 on seeded synthetic inputs, 1e-10 inside coupled steps."""

import numpy as np
import pytest

from oracle import pstd as OP

pytestmark = pytest.mark.gpu

SHAPE, D, EPSR = (16, 8, 32), (20e-9, 30e-9, 10e-9), 9.1


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _spaces(shape=SHAPE, d=D):
    from pulsesuite_amd.typespace import ss

    return ss(3, *shape, *d, EPSR), OP.Space(3, *shape, *d, EPSR)


def _cgrid(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_elong_from_rho_against_oracle():
    import torch

    from oracle import coupling as OC
    from pulsesuite_amd.rhoPJ import ElongfromRho

    sp, osp = _spaces()
    rng = np.random.default_rng(7)
    Rho, Px, Py, Pz = (_cgrid(rng, SHAPE) for _ in range(4))
    dev = [torch.as_tensor(a, device="cuda") for a in (Rho, Px, Py, Pz)]
    outs = [torch.zeros(SHAPE, dtype=torch.complex128, device="cuda") for _ in range(10)]
    ElongfromRho(sp, *dev, *outs)
    torch.cuda.synchronize()
    oP = [Px.copy(), Py.copy(), Pz.copy()]
    EPR, EP, ER, RhoB = OC.ElongfromRho(osp, Rho, *oP)
    want = EPR + EP + ER + [RhoB]
    for n, (o, w) in enumerate(zip(outs, want)):
        assert rel(o, w) < 1e-12, (n, rel(o, w))
    for d, w in zip(dev[1:], oP):  # P transformed in place, as the reference does
        assert rel(d, w) < 1e-13
    assert rel(dev[0], Rho) == 0.0  # rho is not changed


def test_qwplacement_against_oracle():
    import torch

    from oracle import coupling as OC
    from pulsesuite_amd.rhoPJ import QWPlacement

    rng = np.random.default_rng(3)
    f, gx = _cgrid(rng, SHAPE[0]), rng.standard_normal(SHAPE[0])
    gy, gz = rng.standard_normal(SHAPE[1]), rng.standard_normal(SHAPE[2])
    out = torch.zeros(SHAPE, dtype=torch.complex128, device="cuda")
    QWPlacement(f, gx, gy, gz, out)
    assert rel(out, OC.QWPlacement(f, gx, gy, gz)) < 1e-15


def _coupled(shape, nk, x0, longitudinal=True, device_time=False, **qkw):
    from oracle.coupling import OracleCoupling
    from oracle.sbe import MBFlags as OMB
    from oracle.sbe import QWParams as OQW
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    dx = 20e-9
    sp, osp = _spaces(shape, (dx, dx, dx))
    dt = 0.9 * 2.0 * dx / (np.pi * 2.99792458e8 / np.sqrt(EPSR) * np.sqrt(3.0))
    pulse = dict(lambda_=800e-9, Amp=3e8, Tw=5e-15, Tp=4e-15, chirp=0.0)
    L = L_of_nk(nk)
    oc = OracleCoupling(osp, 800e-9, pulse["Amp"], Nw=2, qw=OQW(L=L), mb=OMB(), x0=x0, longitudinal=longitudinal)
    oprop = OP.Propagator(osp, OP.Time(0.0, 1.0, dt, 0), OP.Pulse(**pulse), "soft", "cpml", calc_j=oc.calc_j, npml=6)
    qa = QWArray(QWParams(L=L), MBFlags(), Nw=2, Emax0=pulse["Amp"], x0=x0, longitudinal=longitudinal, **qkw)
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(**pulse), source_type="soft", boundary_type="cpml",
                            calc_j=qa, npml=6, device_time=device_time)
    return sp, osp, oc, oprop, qa, prop


def test_diagnostics_of_seeded_wire_outputs():
    """QWArray._diagnostics on seeded P (Nw = 2 sub-band wires, 3 components) and rho lines
    against the oracle's, every grid of the 17."""
    import torch

    from pulsesuite_amd.rhoPJ import LONGITUDINAL

    sp, osp, oc, oprop, qa, prop = _coupled((32, 16, 16), 12, 0.0)
    Nx = sp.Nx
    rng = np.random.default_rng(11)
    P = _cgrid(rng, (2, 3, Nx))
    R = _cgrid(rng, (2, Nx))
    qa.P.copy_(torch.as_tensor(P))
    qa.Rho.copy_(torch.as_tensor(R))
    qa._diagnostics(sp)
    torch.cuda.synchronize()
    oc._diagnostics(osp, np.transpose(P, (1, 2, 0)), R.T)
    for n in LONGITUDINAL:
        assert rel(qa.longitudinal_grids[n], oc.diag[n]) < 1e-12, (n, rel(qa.longitudinal_grids[n], oc.diag[n]))


def test_diagnostics_in_coupled_steps():
    """Four coupled steps (wire on the soft source's plane, so it is driven from the first
    step) with the diagnostics on, against the oracle: the P-derived grids (rho is the
    difference of equal Fermi seeds, rounding noise at this point, so its grids are held to
    the seeded test above).  A pipelined graph run gives the same diagnostics."""
    import torch

    x0 = float(OP.GetXArray(OP.Space(3, 32, 16, 16, 20e-9, 20e-9, 20e-9, EPSR))[8])
    sp, osp, oc, oprop, qa, prop = _coupled((32, 16, 16), 12, x0)
    steps = 4
    for _ in range(steps):
        oprop.step()
        prop.step()
    torch.cuda.synchronize()
    names = ("Px", "Py", "Pz", "ExlfromP", "EylfromP", "EzlfromP", "RhoBAnalytic", "RhoBAnSmooth")
    assert np.linalg.norm(oc.diag["ExlfromP"]) > 0
    for n in names:
        assert rel(qa.longitudinal_grids[n], oc.diag[n]) < 1e-10, (n, rel(qa.longitudinal_grids[n], oc.diag[n]))
    _, _, _, _, qb, pb = _coupled((32, 16, 16), 12, x0, device_time=True)
    pb.run_pipelined(steps)
    torch.cuda.synchronize()
    for n in names:
        assert rel(qb.longitudinal_grids[n], qa.longitudinal_grids[n].cpu().numpy()) < 1e-10, n
