"""GPU parity of the PSTD field path (FFT, curl updates, CPML, soft source,
propagator) against numpy, the oracle and the reference's golden fixtures,
and of the coupled SBE-PSTD step against the oracle.  FP64 bar: relative L2
<= 1e-10 (1e-12 where the path is short)."""

import os

import numpy as np
import pytest

from oracle import pstd as OP

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "field_path.npz")


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def cu(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.complex128, device="cuda")


@pytest.mark.parametrize("shape", [(16, 8, 32), (128, 1, 1), (1, 1, 1024), (32, 64, 4), (4, 2, 2),
                                   (256, 2, 2), (2, 512, 2), (8, 8, 256), (2048, 1, 1)])
def test_fft3d_matches_numpy(shape):
    from pulsesuite_amd.fftw import fft_3D, ifft_3D

    rng = np.random.default_rng(1)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    T = cu(A)
    fft_3D(T)
    assert rel(T, np.fft.fftn(A)) < 1e-14
    ifft_3D(T)
    assert rel(T, A) < 1e-14


@pytest.mark.parametrize("shape,axis", [((128, 1, 1), 0), ((8, 64, 4), 1), ((2, 4, 1024), 2), ((4, 16, 2), 1),
                                        ((2048, 1, 1), 0)])
@pytest.mark.parametrize("nb", [1, 2, 3])
def test_fft_line_batch_matches_numpy(shape, axis, nb):
    """One launch over nb arrays (in place and out of place) = numpy's FFT along the axis;
    lengths 16 and 2048 take the one-launch-per-array path."""
    import ctypes as C

    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import ss

    ctx = GridContext(ss(3, *shape, 1.0, 1.0, 1.0, 1.0), 1.0, with_cpml=False)
    rng = np.random.default_rng(nb)
    A = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(nb)]
    for inplace in (True, False):
        T = [cu(a) for a in A]
        O = T if inplace else [torch.empty_like(t) for t in T]
        ins = (C.c_void_p * nb)(*(t.data_ptr() for t in T))
        outs = (C.c_void_p * nb)(*(t.data_ptr() for t in O))
        _lib.check(_lib.lib.pss_fft_line_batch(ctx.gp, ins, outs, nb, axis, -1, _lib.stream_ptr()), "batch")
        for b in range(nb):
            assert rel(O[b], np.fft.fft(A[b], axis=axis)) < 1e-14
        ins = (C.c_void_p * nb)(*(t.data_ptr() for t in O))
        _lib.check(_lib.lib.pss_fft_line_batch(ctx.gp, ins, ins, nb, axis, +1, _lib.stream_ptr()), "batch")
        for b in range(nb):
            assert rel(O[b], A[b]) < 1e-14
    bad = (C.c_void_p * 1)(T[0].data_ptr())
    with pytest.raises(RuntimeError):
        _lib.check(_lib.lib.pss_fft_line_batch(ctx.gp, bad, bad, 4, axis, -1, _lib.stream_ptr()), "batch")


@pytest.mark.parametrize("shape", [(32, 64, 16), (128, 1, 1), (16, 16, 256)])
def test_update_e_sample_and_ksample(shape):
    """pss_update_e3d_sample: E as pss_update_e3d (the same per-cell update, to the last
    bits), and a point sample equal to pss_ksample of the new E and to numpy's IFFT of it at the
    eight trilinear corners (weights of a point inside the grid; -1 marks unused corners
    on length-1 axes)."""
    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import ss

    Nx, Ny, Nz = shape
    ctx = GridContext(ss(3, Nx, Ny, Nz, 2e-8, 2e-8, 2e-8, 9.1), 1e-17, with_cpml=False)
    ctx.g.dt, ctx.g.v2 = 1e-17, 2.99792458e8**2 / 9.1
    rng = np.random.default_rng(3)
    E = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    jl = rng.standard_normal((3, Nx)) + 1j * rng.standard_normal((3, Nx))
    gy, gz = rng.standard_normal(Ny) + 0j, rng.standard_normal(Nz) + 0j
    # trilinear corners of a point (x, y, z) = (i0 + fx, j0 + fy, k0 + fz)
    f = [0.3, 0.6, 0.8]
    base = [n // 3 for n in shape]
    idx, wts = [], []
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                ok = all(shape[d] > 1 or s == 0 for d, s in enumerate((a, b, c)))
                w = np.prod([(f[d] if s else 1 - f[d]) if shape[d] > 1 else 1.0 for d, s in enumerate((a, b, c))])
                if ok:
                    i, j, k = base[0] + a, base[1] + b, base[2] + c
                    idx.append((i * Ny + j) * Nz + k)
                    wts.append(w)
                else:
                    idx.append(-1)
                    wts.append(0.0)
    dev = torch.device("cuda")
    t_idx = torch.tensor(idx, dtype=torch.int32, device=dev)
    t_w = torch.tensor(wts, dtype=torch.float64, device=dev)
    part = torch.zeros(6144, dtype=torch.float64, device=dev)
    clock = torch.full((1,), 2.0, dtype=torch.float64, device=dev)
    lines = torch.zeros((2, 3, Nx), dtype=torch.complex128, device=dev)
    out1 = torch.zeros(3, dtype=torch.complex128, device=dev)
    out2 = torch.zeros(3, dtype=torch.complex128, device=dev)
    dB, djl, dgy, dgz = [cu(a) for a in B], cu(jl), cu(gy), cu(gz)
    E1, E2 = [cu(a) for a in E], [cu(a) for a in E]
    P, st = _lib.ptr, _lib.stream_ptr()
    _lib.check(_lib.lib.pss_update_e3d(ctx.gp, *(P(t) for t in dB), None, None, None, P(djl), P(dgy), P(dgz),
                                       *(P(t) for t in E1), st), "e")
    _lib.check(_lib.lib.pss_update_e3d_sample(ctx.gp, *(P(t) for t in dB), None, None, None, P(djl), P(dgy), P(dgz),
                                              *(P(t) for t in E2), P(t_idx), P(t_w), P(part), P(out1), P(lines), Nx, 2,
                                              P(clock), 0.25, 0, st), "es")
    _lib.check(_lib.lib.pss_ksample(ctx.gp, *(P(t) for t in E1), P(t_idx), P(t_w), P(part), P(out2), st), "ks")
    torch.cuda.synchronize()
    for a, b in zip(E1, E2):
        assert rel(b, a.cpu().numpy()) < 1e-15
    assert rel(out1, out2.cpu().numpy()) < 1e-14
    assert float(clock.item()) == 2.25  # the clock advanced once by the given tick
    # both wire lines hold the sample along their whole length
    ln = lines.cpu().numpy()
    assert np.array_equal(ln, np.broadcast_to(out1.cpu().numpy()[None, :, None], ln.shape))
    real = [np.fft.ifftn(t.cpu().numpy()) for t in E1]
    ref = [sum(w * r.reshape(-1)[o] for o, w in zip(idx, wts) if o >= 0) for r in real]
    assert rel(out1, np.array(ref)) < 1e-12
    # update_b: B -= i dt k x E_new in the same pass, as pss_update_b3d after the E update
    E7, B7, B8 = [cu(a) for a in E], [cu(a) for a in B], [cu(a) for a in B]
    _lib.check(_lib.lib.pss_update_e3d_sample(ctx.gp, *(P(t) for t in B7), None, None, None, P(djl), P(dgy), P(dgz),
                                              *(P(t) for t in E7), P(t_idx), P(t_w), P(part), P(out2), None, 0, 0,
                                              None, 0.0, 1, st), "es7")
    _lib.check(_lib.lib.pss_update_b3d(ctx.gp, *(P(t) for t in E1), *(P(t) for t in B8), st), "b8")
    torch.cuda.synchronize()
    for a, b in zip(E7, E1):
        assert rel(a, b.cpu().numpy()) < 1e-15
    for a, b in zip(B7, B8):
        assert rel(a, b.cpu().numpy()) < 1e-15
    assert rel(out2, out1.cpu().numpy()) < 1e-14
    # null B: the current term only, the same as a full update with B = 0
    zB = [torch.zeros_like(t) for t in dB]
    E3, E4 = [cu(a) for a in E], [cu(a) for a in E]
    _lib.check(_lib.lib.pss_update_e3d(ctx.gp, *(P(t) for t in zB), None, None, None, P(djl), P(dgy), P(dgz),
                                       *(P(t) for t in E3), st), "e0")
    _lib.check(_lib.lib.pss_update_e3d_sample(ctx.gp, None, None, None, None, None, None, P(djl), P(dgy), P(dgz),
                                              *(P(t) for t in E4), P(t_idx), P(t_w), P(part), P(out1), None, 0, 0,
                                              None, 0.0, 0, st), "es0")
    _lib.check(_lib.lib.pss_ksample(ctx.gp, *(P(t) for t in E3), P(t_idx), P(t_w), P(part), P(out2), st), "ks0")
    torch.cuda.synchronize()
    for a, b in zip(E3, E4):
        assert rel(b, a.cpu().numpy()) < 1e-15
    assert rel(out1, out2.cpu().numpy()) < 1e-14
    assert _lib.lib.pss_update_e3d_sample(ctx.gp, P(dB[0]), None, None, None, None, None, P(djl), P(dgy), P(dgz),
                                          *(P(t) for t in E4), P(t_idx), P(t_w), P(part), P(out1), None, 0, 0,
                                          None, 0.0, 0, st) != 0  # B in part is refused


def test_update_e_b_match_oracle():
    from pulsesuite_amd import PSTD3D as G
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts

    sp = ss(3, 16, 32, 8, 2e-8, 3e-8, 4e-8, 2.5)
    tm = ts(0.0, 1e-12, 3e-17, 0)
    osp = OP.Space(3, 16, 32, 8, 2e-8, 3e-8, 4e-8, 2.5)
    otm = OP.Time(0.0, 1e-12, 3e-17, 0)
    rng = np.random.default_rng(2)
    F = [rng.standard_normal((16, 32, 8)) + 1j * rng.standard_normal((16, 32, 8)) for _ in range(9)]
    E = [f.copy() for f in F[:3]]
    B = [1e-8 * f for f in F[3:6]]
    J = [1e3 * f for f in F[6:]]
    dE, dB, dJ = [cu(a) for a in E], [cu(a) for a in B], [cu(a) for a in J]
    G.UpdateE3D(sp, tm, *dB, *dJ, *dE)
    OP.UpdateE3D(osp, otm, *B, *J, *E)
    G.UpdateB3D(sp, tm, *dE, *dB)
    OP.UpdateB3D(osp, otm, *E, *B)
    for a, b in zip(dE + dB, E + B):
        assert rel(a, b) < 1e-13


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_cpml_golden(gold):
    from pulsesuite_amd.cpml import CPML

    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    Nx, Ny, Nz = int(Nx), int(Ny), int(Nz)
    cp = CPML(Nx, Ny, Nz, dx, dy, dz, dt, epsr)
    assert cp.GetCPMLInfo() == tuple(gold["cpml_npml"])
    E = [cu(gold["cpml_in_E" + n]) for n in "xyz"]
    B = [cu(gold["cpml_in_B" + n]) for n in "xyz"]
    v2 = 2.99792458e8**2 / epsr
    cp.ApplyCPML_E(*E, *B, None, None, None, v2, dt)
    cp.ApplyCPML_B(*B, *E, None, None, None, dt)
    cp.ApplyCPML_E(*E, *B, None, None, None, v2, dt)
    for i, n in enumerate("xyz"):
        assert rel(E[i], gold["cpml_out_E" + n]) < 1e-12
        assert rel(B[i], gold["cpml_out_B" + n]) < 1e-12


def test_propagator_golden(gold):
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts

    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    sp = ss(3, int(Nx), int(Ny), int(Nz), dx, dy, dz, epsr)
    lam, amp, tw, tp, ch, _ = gold["tfsf_pulse"]
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(lam, amp, tw, tp, ch), source_type="soft",
                            boundary_type="cpml")
    prop.run(int(gold["run_nsteps"][0]))
    for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        assert rel(getattr(prop, n), gold["run_" + n]) < 1e-11, n


def test_ic_seed_golden(gold):
    import torch

    from pulsesuite_amd.PSTD3D import SeedInitialCondition
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts

    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    sp = ss(3, int(Nx), int(Ny), int(Nz), dx, dy, dz, epsr)
    lam, amp, tw, tp, ch, npx = gold["ic_pulse"]
    Ey = torch.zeros((int(Nx), int(Ny), int(Nz)), dtype=torch.complex128, device="cuda")
    Bz = torch.zeros_like(Ey)
    SeedInitialCondition(sp, ts(0.0, 1e-12, dt, 0), ps(lam, amp, tw, tp, ch), Ey, Bz, int(npx))
    assert rel(Ey, gold["ic_Ey"]) < 1e-13 and rel(Bz, gold["ic_Bz"]) < 1e-13


@pytest.mark.parametrize("current", ["reference", "delta"])
@pytest.mark.parametrize("grid,nk,steps", [((32, 32, 32), 14, 6), ((64, 1, 1), 12, 8)])
def test_coupled_step_matches_oracle(grid, nk, steps, current):
    """Soft-source PSTD + CPML with one wire (Nw = 2 sub-bands) driven by the
    point field, current fed back through the separable deposit; both current
    semantics (the reference's J = P/dt and J = dP/dt) against the oracle's."""
    from oracle.coupling import OracleCoupling
    from oracle.sbe import MBFlags as OMB
    from oracle.sbe import QWParams as OQW
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    Nx, Ny, Nz = grid
    dx = 20e-9
    epsr = 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr if Nz > 1 else 1.0) * np.sqrt(3.0))
    L = L_of_nk(nk)
    pulse = dict(lambda_=800e-9, Amp=3e8, Tw=5e-15, Tp=4e-15, chirp=0.0)
    npml = 6 if Nx >= 32 else None
    # oracle
    osp = OP.Space(3, Nx, Ny, Nz, dx, dx, dx, epsr)
    oc = OracleCoupling(osp, 800e-9, pulse["Amp"], Nw=2, qw=OQW(L=L), mb=OMB(), current=current)
    oprop = OP.Propagator(osp, OP.Time(0.0, 1.0, dt, 0), OP.Pulse(**pulse), "soft", "cpml", calc_j=oc.calc_j,
                          npml=npml)
    # device
    sp = ss(3, Nx, Ny, Nz, dx, dx, dx, epsr)
    qa = QWArray(QWParams(L=L), MBFlags(), Nw=2, Emax0=pulse["Amp"], current=current)
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(**pulse), source_type="soft", boundary_type="cpml",
                            calc_j=qa, npml=npml)
    for _ in range(steps):
        oprop.step()
        prop.step()
    # relative L2 of the vector fields E and B: the undriven components (Ex, Ez, Bx, By)
    # are ~1e-9 of the field and carry CalcJfx of Rho = RhoH - RhoE, which is rounding
    # noise while both densities are the equal Fermi seeds, so they are held to the
    # field's norm, not their own
    for F in (("Ex", "Ey", "Ez"), ("Bx", "By", "Bz")):
        num = sum(np.linalg.norm(getattr(prop, n).cpu().numpy() - getattr(oprop, n)) ** 2 for n in F)
        den = sum(np.linalg.norm(getattr(oprop, n)) ** 2 for n in F)
        assert np.sqrt(num / den) < 1e-10, F
    for w in range(2):
        st = qa.wires.get_state(w)
        so = oc.wire.state[w]
        assert rel(st["p2"], so.p2) < 1e-10
        assert rel(st["C2"], so.C2) < 1e-10
    jl = oc.last_line
    assert np.linalg.norm(jl) > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_overlapped_step_is_identical(precision):
    """The wire step on a side stream beside the B half step gives bit-identical
    fields, wire state and line current to the serial order (same launches, same
    operands).  The wire's current is fed back (the point-current feedback grows by
    ~1.5e16 per second, so 15 steps keep every value finite)."""
    import torch

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    N, nk, dx, epsr = 64, 64, 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    out = []
    for overlap in (False, True):
        sp = ss(3, N, N, N, dx, dx, dx, epsr)
        qa = QWArray(QWParams(L=L_of_nk(nk)), MBFlags(Screened=False), Nw=1, Emax0=0.0, precision=precision)
        prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(800e-9, 3e8, 5e-15, 4e-15, 0.0), source_type="soft",
                                boundary_type="cpml", calc_j=qa, npml=8, precision=precision, overlap=overlap)
        assert prop.overlap == overlap
        for _ in range(15):
            prop.step()
        torch.cuda.synchronize()
        st = qa.wires.get_state(0)
        out.append([getattr(prop, n).cpu().numpy() for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")] +
                   [st["p2"], st["C2"], st["D2"], qa.jl.cpu().numpy()])
    assert all(bool(np.isfinite(a).all()) for a in out[0])
    assert np.linalg.norm(out[0][1]) > 0 and np.linalg.norm(out[0][6]) > 0 and np.linalg.norm(out[0][9]) > 0
    for i, (a, b) in enumerate(zip(*out)):
        ok = bool(np.array_equal(a, b))
        assert ok, f"array {i} differs between serial and overlapped steps"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_captured_step_matches_eager(precision):
    """Replays of the step recorded as a HIP graph (device-side source time, wire step
    on the side stream) give bit-identical fields, wire state and current line to eager
    steps; every replay is checked."""
    import torch

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    N, nk, dx, epsr = 64, 64, 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))

    def make():
        sp = ss(3, N, N, N, dx, dx, dx, epsr)
        qa = QWArray(QWParams(L=L_of_nk(nk)), MBFlags(Screened=False), Nw=1, Emax0=0.0, precision=precision)
        prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(800e-9, 3e8, 5e-15, 4e-15, 0.0), source_type="soft",
                                boundary_type="cpml", calc_j=qa, npml=8, precision=precision, device_time=True)
        return prop, qa

    def snap(prop, qa):
        torch.cuda.synchronize()
        st = qa.wires.get_state(0)
        return [getattr(prop, n).cpu().numpy() for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")] + \
               [st["p2"], st["C2"], qa.jl.cpu().numpy()]

    pe, qe = make()
    pg, qg = make()
    for _ in range(2):
        pe.step()
        pg.step()
    pg.capture_step()
    for k in range(10):
        pe.step()
        pg.run_graph(1)
        a, b = snap(pe, qe), snap(pg, qg)
        assert all(bool(np.isfinite(x).all()) for x in a)
        for i, (x, y) in enumerate(zip(a, b)):
            ok = bool(np.array_equal(x, y))
            assert ok, f"replay {k}: array {i} differs from the eager step"
    assert pe.time.n == pg.time.n and pe.time.t == pg.time.t
    assert np.linalg.norm(a[1]) > 0 and np.linalg.norm(a[6]) > 0


@pytest.mark.parametrize("precision,feedback,steps,tol,shape,spg", [("fp64", False, 12, 1e-12, (64, 64, 64), 1),
                                                                    ("mixed", False, 12, 1e-6, (64, 64, 64), 3),
                                                                    ("fp64", True, 8, 1e-12, (64, 64, 64), 2),
                                                                    ("mixed", False, 6, 1e-6, (128, 64, 64), 1)])
def test_pipelined_run_matches_eager(precision, feedback, steps, tol, shape, spg):
    """run_pipelined (the next step's CPML-E correction beside the wire step, in mixed
    precision with the E update's curl term, graph replays of spg steps, of one step and
    of the closing step) against as many eager steps, in chunks of 1, 2 and the rest (one
    grid whose axes differ: the slab-derivative stages then take one launch per axis).  The same
    updates are summed in another order, so agreement is to rounding, measured on each
    field as a vector (its three components together: with the current fed back, Ey
    grows to ~1e12 while Ex stays ~1e4, and a rounding difference of the order of Ey's
    last bits is a large part of Ex).  FP64 1e-12; mixed: the point sample is taken from
    the field after the FP32 correction instead of before it plus the correction, 1e-6."""
    import torch

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    nk, dx, epsr = 64, 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))

    def make():
        sp = ss(3, *shape, dx, dx, dx, epsr)
        # the wire through the source plane (x = -Nx dx / 4): it samples the driven field from
        # the first step (farther away it sees a precursor many orders below the field, whose
        # relative rounding error says nothing)
        qa = QWArray(QWParams(L=L_of_nk(nk)), MBFlags(Screened=False, ignorewire=not feedback), Nw=1, Emax0=0.0,
                     precision=precision, x0=-0.25 * shape[0] * dx)
        prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(800e-9, 3e8, 5e-15, 4e-15, 0.0), source_type="soft",
                                boundary_type="cpml", calc_j=qa, npml=8, precision=precision, device_time=True)
        return prop, qa

    def snap(prop, qa):
        torch.cuda.synchronize()
        st = qa.wires.get_state(0)
        return [np.stack([getattr(prop, F + c).cpu().numpy() for c in "xyz"]) for F in "EB"] + \
               [st["p2"], st["C2"], qa.jl.cpu().numpy(), qa.sample.cpu().numpy()]

    pe, qe = make()
    pp, qp = make()
    pp.steps_per_graph = spg
    assert pp.pipelined()
    for _ in range(2):
        pe.step()
        pp.step()
    done = 0
    for chunk in (1, 2, steps - 3):
        for _ in range(chunk):
            pe.step()
        pp.run_pipelined(chunk)
        done += chunk
        a, b = snap(pe, qe), snap(pp, qp)
        assert all(bool(np.isfinite(x).all()) for x in b)
        for i, (x, y) in enumerate(zip(a, b)):
            assert rel(y, x) <= tol, f"after {done} steps: array {i} differs by {rel(y, x):.3e}"
    assert pe.time.n == pp.time.n and abs(pe.time.t - pp.time.t) <= 1e-12 * abs(pe.time.t)
    assert np.linalg.norm(a[0]) > 0 and np.linalg.norm(a[3]) > 0 and np.linalg.norm(a[5]) > 0
    if feedback:
        assert np.linalg.norm(a[4]) > 0


def test_cpml_128_matches_oracle():
    """One CPML E and one CPML B half step on a 128^3 grid (the bench grid) with
    12-cell layers: every FFT kernel path the bench uses, with the i*k premultiply."""
    from pulsesuite_amd.cpml import CPML

    N, dx, epsr = 128, 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    rng = np.random.default_rng(5)
    E = [rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))) for _ in range(3)]
    dE, dB = [cu(a) for a in E], [cu(a) for a in B]
    sp = OP.Space(3, N, N, N, dx, dx, dx, epsr)
    ocp = OP.CPML(N, N, N, dx, dx, dx, dt, epsr, npml=12)
    kx, ky, kz = OP.GetKxArray(sp), OP.GetKyArray(sp), OP.GetKzArray(sp)
    v2 = 2.99792458e8**2 / epsr
    ocp.ApplyCPML_E(*E, *B, kx, ky, kz, v2, dt)
    ocp.ApplyCPML_B(*B, *E, kx, ky, kz, dt)
    cp = CPML(N, N, N, dx, dx, dx, dt, epsr, npml=12)
    cp.ApplyCPML_E(*dE, *dB, None, None, None, v2, dt)
    cp.ApplyCPML_B(*dB, *dE, None, None, None, dt)
    for a, b in zip(dE + dB, E + B):
        assert rel(a, b) < 1e-13


@pytest.mark.parametrize("shape", [(64, 128, 256), (128, 64, 1)])
def test_cpml_anisotropic_matches_oracle(shape):
    """Two CPML E/B half-step pairs on grids whose axes differ (layers 6/6/12 cells by
    the default rule; one grid with a length-1 axis): derivatives and psi on the PML
    slabs of each axis, psi carried from one step to the next."""
    from pulsesuite_amd.cpml import CPML

    Nx, Ny, Nz = shape
    dx, epsr = 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    rng = np.random.default_rng(7)
    E = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    dE, dB = [cu(a) for a in E], [cu(a) for a in B]
    sp = OP.Space(3, Nx, Ny, Nz, dx, dx, dx, epsr)
    ocp = OP.CPML(Nx, Ny, Nz, dx, dx, dx, dt, epsr)
    cp = CPML(Nx, Ny, Nz, dx, dx, dx, dt, epsr)
    assert cp.GetCPMLInfo() == (ocp.npml_x, ocp.npml_y, ocp.npml_z)
    assert len(set(cp.GetCPMLInfo())) > 1  # the layers differ between axes
    kx, ky, kz = OP.GetKxArray(sp), OP.GetKyArray(sp), OP.GetKzArray(sp)
    v2 = 2.99792458e8**2 / epsr
    for _ in range(2):
        ocp.ApplyCPML_E(*E, *B, kx, ky, kz, v2, dt)
        ocp.ApplyCPML_B(*B, *E, kx, ky, kz, dt)
        cp.ApplyCPML_E(*dE, *dB, None, None, None, v2, dt)
        cp.ApplyCPML_B(*dB, *dE, None, None, None, dt)
    for a, b in zip(dE + dB, E + B):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("shape,fused", [((64, 64, 128), True), ((128, 64, 1), True), ((1024, 1, 1), True),
                                         ((64, 64, 64), True), ((128, 128, 128), True), ((256, 64, 64), True),
                                         ((64, 64, 32), False), ((2048, 1, 1), False)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cpml_e_curl_matches_update_then_cpml(shape, fused, precision):
    """pss_cpml_e_curl (E += i v2 dt k x B added in the correction's last forward pass
    when that axis is 64 ... 1024 long) against pss_update_e3d without current followed by
    pss_cpml_half (which = 0), soft source on Ey, two half steps (psi carried): the same
    sums in another order, FP64 rounding (1e-13 on E as a vector).  Mixed: FP32 passes that
    round differently in the two forms (with Ny = Nz = 64 or 128 the fused form transforms
    each x-plane along z, then y, in LDS; the split form along y, then z): 3e-7, a few units
    of FP32 roundoff (6e-8) through the three passes.  Other lengths take the two calls."""
    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd.cpml import CPML
    from pulsesuite_amd.PSTD3D import InitializeTFSF
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import GetXArray, GetYArray, GetZArray, ss

    Nx, Ny, Nz = shape
    dx, epsr = 20e-9, 9.1
    v2 = (2.99792458e8 / np.sqrt(epsr)) ** 2
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    rng = np.random.default_rng(12)
    E = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    sp = ss(3, Nx, Ny, Nz, dx, dx, dx, epsr)
    pulse = ps(800e-9, 3e8, 5e-15, 4e-15, 0.0)
    dev = torch.device("cuda")
    srcx = torch.as_tensor(InitializeTFSF(sp, pulse), device=dev)
    r = [torch.as_tensor(f(sp), device=dev) for f in (GetXArray, GetYArray, GetZArray)]
    par = (_lib.C.c_double * 9)(2e-15, 3e8 / 1e8, pulse.CalcOmega0(), pulse.CalcTau(), 0.0, 1e8, pulse.Tp, 0.0, dx)
    out = []
    for mode in ("split", "fused"):
        cp = CPML(Nx, Ny, Nz, dx, dx, dx, dt, epsr)
        cp.ctx.g.precision = _lib.MIXED if precision == "mixed" else _lib.FP64
        dE, dB = [cu(a) for a in E], [cu(a) for a in B]
        src = (1, par, srcx, *r)
        for _ in range(2):
            if mode == "split":
                cp.ctx.g.dt, cp.ctx.g.v2 = dt, v2
                _lib.check(_lib.lib.pss_update_e3d(cp.ctx.gp, *(_lib.ptr(t) for t in dB), None, None, None, None,
                                                   None, None, *(_lib.ptr(t) for t in dE), _lib.stream_ptr()), "e")
                cp.ApplyCPML_E(*dE, *dB, dt=dt, v2=v2, src=src)
            else:
                cp.CurlE_ApplyCPML_E(*dE, *dB, dt, v2, src=src)
        torch.cuda.synchronize()
        out.append(np.stack([t.cpu().numpy() for t in dE]))
    assert np.isfinite(out[1]).all()
    assert rel(out[1], out[0]) < (1e-13 if precision == "fp64" else 3e-7)


@pytest.mark.parametrize("shape,fused", [((64, 64, 128), True), ((128, 64, 1), True), ((1024, 1, 1), True),
                                         ((64, 64, 64), True), ((128, 128, 128), True), ((256, 64, 64), True),
                                         ((64, 64, 32), False), ((2048, 1, 1), False)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_b_half_matches_update_then_cpml(shape, fused, precision):
    """pss_b_half (B -= i dt k x E added in the correction's last forward pass when that
    axis is 64 ... 1024 long, in mixed precision; FP64 runs the two updates) against pss_update_b3d followed by pss_cpml_half, with the
    soft source on Bz, two half steps (psi carried): the same sums in another order,
    FP64 rounding (1e-13 on B as a vector).  Mixed: the correction's last pass is an FP32
    transform in both, but in two different kernels (plain and curl-adding), whose FP32
    butterflies the compiler may contract differently: FP32 rounding of the correction,
    3e-7 on B (with Ny = Nz = 64 or 128 the fused form transforms each x-plane along z,
    then y, in LDS; the split form along y, then z).  The non-fused lengths take the two
    calls."""
    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd.cpml import CPML
    from pulsesuite_amd.PSTD3D import InitializeTFSF
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import GetXArray, GetYArray, GetZArray, ss

    Nx, Ny, Nz = shape
    dx, epsr = 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    rng = np.random.default_rng(11)
    E = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    sp = ss(3, Nx, Ny, Nz, dx, dx, dx, epsr)
    pulse = ps(800e-9, 3e8, 5e-15, 4e-15, 0.0)
    dev = torch.device("cuda")
    srcx = torch.as_tensor(InitializeTFSF(sp, pulse), device=dev)
    r = [torch.as_tensor(f(sp), device=dev) for f in (GetXArray, GetYArray, GetZArray)]
    par = (_lib.C.c_double * 9)(2e-15, 3e8 / 1e8, pulse.CalcOmega0(), pulse.CalcTau(), 0.0, 1e8, pulse.Tp, 0.0, dx)
    out = []
    for mode in ("split", "fused"):
        cp = CPML(Nx, Ny, Nz, dx, dx, dx, dt, epsr)
        cp.ctx.g.precision = _lib.MIXED if precision == "mixed" else _lib.FP64
        dE, dB = [cu(a) for a in E], [cu(a) for a in B]
        src = (2, par, srcx, *r)
        for _ in range(2):
            if mode == "split":
                _lib.check(_lib.lib.pss_update_b3d(cp.ctx.gp, *(_lib.ptr(t) for t in dE + dB), _lib.stream_ptr()), "b")
                cp.ApplyCPML_B(*dB, *dE, dt=dt, src=src)
            else:
                cp.UpdateB_ApplyCPML_B(*dB, *dE, dt, src=src)
        torch.cuda.synchronize()
        out.append(np.stack([t.cpu().numpy() for t in dB]))
    assert np.isfinite(out[1]).all()
    assert rel(out[1], out[0]) < (1e-13 if precision == "fp64" else 3e-7)
    if not fused:  # the same two calls: bit-identical
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("N,npml", [(128, 12), (64, 8)])
def test_cpml_mixed_precision_derivatives(N, npml):
    """Mixed precision: the six spectral derivatives of a CPML half step in FP32 (real
    parts, used only in PML cells; at 64 and 128 points per axis the passes after the slab
    pass run as 2-D plane transforms in LDS), fields FP64.  Bar 1e-6 against the FP64
    oracle on one E and one B half step."""
    from pulsesuite_amd import _lib
    from pulsesuite_amd.cpml import CPML

    dx, epsr = 20e-9, 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    rng = np.random.default_rng(6)
    E = [rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)) for _ in range(3)]
    B = [1e-8 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))) for _ in range(3)]
    dE, dB = [cu(a) for a in E], [cu(a) for a in B]
    sp = OP.Space(3, N, N, N, dx, dx, dx, epsr)
    ocp = OP.CPML(N, N, N, dx, dx, dx, dt, epsr, npml=npml)
    kx, ky, kz = OP.GetKxArray(sp), OP.GetKyArray(sp), OP.GetKzArray(sp)
    v2 = 2.99792458e8**2 / epsr
    ocp.ApplyCPML_E(*E, *B, kx, ky, kz, v2, dt)
    ocp.ApplyCPML_B(*B, *E, kx, ky, kz, dt)
    cp = CPML(N, N, N, dx, dx, dx, dt, epsr, npml=npml)
    cp.ctx.g.precision = _lib.MIXED
    cp.ApplyCPML_E(*dE, *dB, None, None, None, v2, dt)
    cp.ApplyCPML_B(*dB, *dE, None, None, None, dt)
    errs = [float(rel(a, b)) for a, b in zip(dE + dB, E + B)]
    print("mixed CPML rel errors", ["%.1e" % e for e in errs], flush=True)
    assert max(errs) < 1e-6
    assert max(errs) > 0.0  # the FP32 path ran


def test_propagator_golden_mixed(gold):
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts

    Nx, Ny, Nz, dx, dy, dz, epsr, dt = gold["cpml_grid"]
    sp = ss(3, int(Nx), int(Ny), int(Nz), dx, dy, dz, epsr)
    lam, amp, tw, tp, ch, _ = gold["tfsf_pulse"]
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), ps(lam, amp, tw, tp, ch), source_type="soft",
                            boundary_type="cpml", precision="mixed")
    prop.run(int(gold["run_nsteps"][0]))
    for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        assert rel(getattr(prop, n), gold["run_" + n]) < 1e-6, n


@pytest.mark.gpu
@pytest.mark.parametrize("keep_old", [0, 1])
@pytest.mark.parametrize("Nx,Nw", [(1, 1), (64, 2), (128, 1), (2048, 3)])
def test_jline_fft_in_kernel(Nx, Nw, keep_old):
    """pss_jline with fft = 1: the current line and then its forward FFT along x in the
    same kernel (radix-2 in LDS), against the line without FFT and numpy's FFT of it
    (FP64 rounding of an N log N transform: 1e-14), and the same Pold/RhoOld update.
    The line itself against the oracle's CalcJfx and the sums of PSTD3D/rhoPJ.py:1183-1194,
    with the old arrays kept (keep_old = 1) or zero (keep_old = 0, the reference's
    re-created PxxOld/RhoOld, PSTD3D/rhoPJ.py:1138-1147)."""
    import torch

    from oracle.coupling import CalcJfx
    from pulsesuite_amd import _lib

    rng = np.random.default_rng(Nx + Nw)
    c = lambda *sh: rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    P, Rho, Pold, RhoOld = c(Nw, 3, Nx), c(Nw, Nx), c(Nw, 3, Nx), c(Nw, Nx)
    dt, dx = 1e-17, 2e-8
    out = []
    for fft in (0, 1):
        t = [cu(a) for a in (P, Rho, Pold, RhoOld)]
        jl = torch.zeros((3, Nx), dtype=torch.complex128, device="cuda")
        _lib.check(_lib.lib.pss_jline(*(_lib.ptr(a) for a in t), Nw, Nx, dt, dx, _lib.ptr(jl), fft, keep_old,
                                      _lib.stream_ptr()), "jline")
        torch.cuda.synchronize()
        out.append((jl.cpu().numpy(), t[2].cpu().numpy(), t[3].cpu().numpy()))
    (j0, po0, ro0), (j1, po1, ro1) = out
    assert rel(j1, np.fft.fft(j0, axis=1)) < 1e-14
    assert np.array_equal(po0, P) and np.array_equal(po1, P)
    assert np.array_equal(ro0, Rho) and np.array_equal(ro1, Rho)
    po, ro = (Pold, RhoOld) if keep_old else (0 * Pold, 0 * RhoOld)
    want = np.sum(P - po, axis=0) / dt / Nw
    want[0] += CalcJfx(np.sum(Rho, axis=0), np.sum(ro, axis=0), dt, dx) / Nw
    assert rel(j0, want) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("shape,axis,in32,curl", [((16, 8, 128), 2, 1, True), ((16, 8, 128), 2, 0, True),
                                                  ((64, 4, 2), 0, 1, False), ((4, 256, 2), 1, 0, False),
                                                  ((8, 2, 32), 2, 0, False), ((256, 4, 32), 0, 1, True),
                                                  ((4, 256, 8), 1, 1, False), ((1024, 2, 8), 0, 1, False)])
def test_fft_pass_acc_matches_numpy(shape, axis, in32, curl):
    """pss_fft_pass_acc: out_b += FFT_axis(in_b) (complex64 input computed in FP32: 1e-6 of the
    transform; complex128: 1e-14), plus coef i (k x G)_b when G is given (the curl term of the
    E update, PSTD3D/PSTD3D.py:29-76, on the accumulating last pass); in is left as is."""
    import ctypes as C

    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import GetKxArray, GetKyArray, GetKzArray, ss

    sp = ss(3, *shape, 20e-9, 20e-9, 20e-9, 9.1)
    ctx = GridContext(sp, 1e-17, with_cpml=False)
    rng = np.random.default_rng(sum(shape) + axis)
    c = lambda: rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    A = [c() for _ in range(3)]
    F = [c() * 1e3 for _ in range(3)]
    G = [c() * 1e-5 for _ in range(3)]
    dt = torch.complex64 if in32 else torch.complex128
    tin = [torch.as_tensor(a, dtype=dt, device="cuda") for a in A]
    tout = [cu(f) for f in F]
    tG = [cu(g) for g in G] if curl else [None] * 3
    coef = 1.3e-2
    ins = (C.c_void_p * 3)(*(t.data_ptr() for t in tin))
    outs = (C.c_void_p * 3)(*(t.data_ptr() for t in tout))
    _lib.check(_lib.lib.pss_fft_pass_acc(ctx.gp, ins, outs, 3, axis, -1, in32, *(_lib.ptr(t) for t in tG), coef,
                                         _lib.stream_ptr()), "pss_fft_pass_acc")
    torch.cuda.synchronize()
    Ain = [a.astype(np.complex64).astype(np.complex128) if in32 else a for a in A]
    want = [f + np.fft.fft(a, axis=axis) for f, a in zip(F, Ain)]
    if curl:
        kx, ky, kz = GetKxArray(sp)[:, None, None], GetKyArray(sp)[None, :, None], GetKzArray(sp)[None, None, :]
        cx = ky * G[2] - kz * G[1]
        cy = kz * G[0] - kx * G[2]
        cz = kx * G[1] - ky * G[0]
        want = [w + coef * 1j * cc for w, cc in zip(want, (cx, cy, cz))]
    tol = 1e-6 if in32 else 1e-14
    for t, w, f, a in zip(tout, want, F, Ain):
        got = t.cpu().numpy()
        scale = np.linalg.norm(np.fft.fft(a, axis=axis))
        assert np.linalg.norm(got - w) / scale < tol
    for t, a in zip(tin, A):  # input untouched
        assert np.array_equal(t.cpu().numpy(), a.astype(np.complex64) if in32 else a)


@pytest.mark.gpu
@pytest.mark.parametrize("src_comp", [-1, 1, 2])
def test_mask_half_matches_oracle(src_comp):
    """pss_mask_half: the masking absorber (PSTD3D/absorber.py:160-191) of three k-space
    fields and, on src_comp, the soft-source add of the same real-space pass, against the
    oracle's ApplyAbsorber + UpdateTFSC (which round-trips the field once more)."""
    import torch

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts

    shape, dx, dt = (32, 16, 64), 20e-9, 5e-17
    pulse = (800e-9, 3e8, 5e-15, 4e-15, 0.0)
    prop = PSTD3DPropagator(ss(3, *shape, dx, dx, dx, 9.1), ts(2e-15, 1.0, dt, 0), ps(*pulse), source_type="soft",
                            boundary_type="mask")
    rng = np.random.default_rng(4 + src_comp)
    F = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    T = [cu(f) for f in F]
    src = prop._src(src_comp, 3e8, 2e-15) if src_comp >= 0 else None
    prop.absorber.apply(*T, src=src)
    torch.cuda.synchronize()
    osp = OP.Space(3, *shape, dx, dx, dx, 9.1)
    a = OP.Absorber(*shape, dt)
    W = [f.copy() for f in F]
    a.apply(*W)
    if src_comp >= 0:
        OP.UpdateTFSC(W[src_comp], OP.InitializeTFSF(osp, OP.Pulse(*pulse)), osp, OP.Time(2e-15, 1.0, dt, 0),
                      OP.Pulse(*pulse), 3e8)
    for t, w in zip(T, W):
        assert rel(t, w) < 1e-13
    assert not np.allclose(W[0], F[0])  # the mask acted


@pytest.mark.gpu
@pytest.mark.parametrize("polarization", ["y", "z"])
def test_mask_boundary_coupled_matches_oracle(polarization):
    """The reference's default boundary (mask) in the coupled step: soft source of either
    polarisation, one wire (Nw = 2) with the current fed back, against the oracle (E and B as
    vector fields, the wire's p and C)."""
    from oracle.coupling import OracleCoupling
    from oracle.sbe import MBFlags as OMB
    from oracle.sbe import QWParams as OQW
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    N, nk, dx, epsr, steps = 32, 14, 20e-9, 9.1, 10
    dt = 0.9 * 2 * dx / (np.pi * 2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    pulse = dict(lambda_=800e-9, Amp=3e8, Tw=5e-15, Tp=4e-15, chirp=0.0)
    x0 = -0.25 * N * dx  # near the source plane: the wire is driven from the start
    osp = OP.Space(3, N, N, N, dx, dx, dx, epsr)
    oc = OracleCoupling(osp, 800e-9, 0.0, Nw=2, qw=OQW(L=L_of_nk(nk)), mb=OMB(), x0=x0)
    oprop = OP.Propagator(osp, OP.Time(0.0, 1.0, dt, 0), OP.Pulse(**pulse), "soft", "mask", calc_j=oc.calc_j,
                          polarization=polarization)
    qa = QWArray(QWParams(L=L_of_nk(nk)), MBFlags(), Nw=2, Emax0=0.0, x0=x0)
    prop = PSTD3DPropagator(ss(3, N, N, N, dx, dx, dx, epsr), ts(0.0, 1.0, dt, 0), ps(**pulse), source_type="soft",
                            calc_j=qa, polarization=polarization)
    assert prop.boundary_type == "mask"  # the reference's default
    for _ in range(steps):
        oprop.step()
        prop.step()
    for names in (("Ex", "Ey", "Ez"), ("Bx", "By", "Bz")):
        num = sum(np.linalg.norm(getattr(prop, n).cpu().numpy() - getattr(oprop, n)) ** 2 for n in names)
        den = sum(np.linalg.norm(getattr(oprop, n)) ** 2 for n in names)
        assert np.sqrt(num / den) < 1e-10, names
    big = "Ey" if polarization == "y" else "Ez"
    assert np.linalg.norm(getattr(oprop, big)) > 10 * np.linalg.norm(getattr(oprop, "Ez" if big == "Ey" else "Ey"))
    for w in range(2):
        st, so = qa.wires.get_state(w), oc.wire.state[w]
        assert rel(st["p2"], so.p2) < 1e-10 and rel(st["C2"], so.C2) < 1e-10
    assert np.abs(oc.wire.state[0].p2).max() > 0


@pytest.mark.gpu
def test_mask_closed_loop_stays_bounded():
    """128^3 with an Nk = 256 wire on the source plane and its current fed back: with the
    CPML the undriven components grow from the point current (~1e4 per 50 steps), with the
    mask they track the driven field over 600 steps: the wire's own polarisation has x and z
    parts (its dipole elements Xcv, Zcv), so Ex and Ez are not small, but their ratio to Ey
    settles instead of growing (measured: Ex / Ey 0.098 -> 0.107, Ez / Ey 0.005 -> 0.008)."""
    import torch

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    N, dx, epsr = 128, 20e-9, 9.1
    dt = 0.9 * 2 * dx / (np.pi * 2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    qa = QWArray(QWParams(L=L_of_nk(256)), MBFlags(Screened=False), Nw=1, Emax0=0.0, x0=-0.25 * N * dx)
    prop = PSTD3DPropagator(ss(3, N, N, N, dx, dx, dx, epsr), ts(0.0, 1.0, dt, 0), ps(800e-9, 1e7, 100e-15, 50e-15, 0.0),
                            source_type="soft", boundary_type="mask", calc_j=qa, check_every=50, device_time=True)
    prop.capture_step()
    norms = []
    for _ in range(6):
        prop.run_graph(100)
        torch.cuda.synchronize()
        norms.append([float(torch.linalg.vector_norm(getattr(prop, n))) for n in ("Ex", "Ey", "Ez")])
        print("mask closed loop", prop.time.n, ["%.3e" % v for v in norms[-1]], flush=True)
    st = qa.wires.get_state(0)
    assert np.abs(np.diag(st["C2"])).max() > 1e-8  # driven
    assert np.isfinite(norms).all()
    r0 = (norms[0][0] / norms[0][1], norms[0][2] / norms[0][1])
    for ex, ey, ez in norms[1:]:
        assert ex / ey < 2.0 * r0[0] and ez / ey < 3.0 * r0[1]
        assert ey < 10.0 * norms[0][1]


@pytest.mark.parametrize("shape", [(128, 128, 128), (64, 64, 64), (256, 64, 64), (32, 128, 128)])
def test_fft_planes_acc_matches_numpy(shape):
    """pss_fft_planes_acc (the y and z forward passes of the CPML correction's dF, one 2-D
    LDS transform per x-plane, accumulated into the field with the curl term): dF is real,
    so the kernel reads the x pass's output on the x-planes 0 ... Nx/2 only and takes the
    others as their conjugate partners.  Against numpy in FP64 on a random real dF (x pass
    done on the host, rounded to complex64): F_c + FFT3(dF_c) + coef i (k x G)_c, relative
    error of the FFT part within 1e-6 (FP32 transforms), the rest exact to FP64 rounding.
    Planes Nx/2+1 ... Nx-1 of the input are filled with garbage to show they are not read."""
    import ctypes as C

    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import GetKxArray, GetKyArray, GetKzArray, ss

    Nx, Ny, Nz = shape
    sp = ss(3, Nx, Ny, Nz, 2e-8, 2e-8, 2e-8, 9.1)
    ctx = GridContext(sp, 1e-17, with_cpml=False)
    rng = np.random.default_rng(21)
    dF = [rng.standard_normal(shape) for _ in range(3)]
    F = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    G = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    coef = 0.37
    tx = [np.fft.fft(d, axis=0).astype(np.complex64) for d in dF]
    for t in tx:
        t[Nx // 2 + 1:] = np.nan  # not read
    dev = torch.device("cuda")
    din = [torch.as_tensor(t, device=dev) for t in tx]
    dout = [cu(a) for a in F]
    dG = [cu(a) for a in G]
    ins = (C.c_void_p * 3)(*(t.data_ptr() for t in din))
    outs = (C.c_void_p * 3)(*(t.data_ptr() for t in dout))
    _lib.check(_lib.lib.pss_fft_planes_acc(ctx.gp, ins, outs, *(_lib.ptr(t) for t in dG), coef, _lib.stream_ptr()),
               "pss_fft_planes_acc")
    torch.cuda.synchronize()
    k = np.meshgrid(GetKxArray(sp), GetKyArray(sp), GetKzArray(sp), indexing="ij")
    for c in range(3):
        fft = np.fft.fftn(dF[c])
        curl = k[(c + 1) % 3] * G[(c + 2) % 3] - k[(c + 2) % 3] * G[(c + 1) % 3]
        want = F[c] + fft + coef * 1j * curl
        got = dout[c].cpu().numpy()
        assert np.isfinite(got).all()
        assert rel(got - F[c] - coef * 1j * curl, fft) < 1e-6, c
        assert rel(got, want) < 1e-6, c
