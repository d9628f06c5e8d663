"""The five benchmark workloads at their full sizes, as parity cases.

Shared by the fixture generator (tests/golden/make_configs.py, which runs the FP64 oracle
on the CPU) and the GPU tests (tests/test_gpu_configs.py, which run the HIP path on the
same inputs and compare with the fixtures).  Each builder returns the oracle's objects or
the device's objects for exactly the same configuration.

  C1  1 wire Nk = 128 on a 1-D line of 1024 cells (dx 10 nm), 16-cell CPML, soft source,
      current fed back, 2000 steps.  On an Nx x 1 x 1 grid the reference takes eps_r = 1
      (GetEpsr, PSTD3D/typespace.py:237-263), so the leapfrog's spectral step limit is
      2 dx / (pi c) = 21.2 as; dt = 0.9 of it (the 50 as of the configuration list diverges
      from rounding noise within ~20 steps).  Source amplitude 10^U(6, 9) V/m, seed 1.
      The current is J = dP/dt (the "delta" coupling): with the reference's own J = P / dt
      (its QuantumWire rebuilds P_old as zeros on every call) this driven 1-D line grows
      without bound from step ~400 and is NaN by step ~1250, in the oracle and on the GPU
      alike, so the 2000-step comparison runs the finite form.
  C2  128^3 grid (dx 20 nm, eps_r 9.1), 12-cell CPML, soft source, one wire Nk = 256 whose
      sample point sits on the source plane (driven from the first step), current fed back;
      dt = 0.9 x 2 dx / (pi v sqrt 3) = 66.6 as.
  C3  256^3 grid, 12-cell CPML, one wire Nk = 1024 on the source plane, the same dt.
  C4  512 wires, Nk uniform over the even values in [64, 1024], E0 log-uniform
      [1e6, 1e9] V/m, detuning U(-50, 50) meV, screened, dt 10 as, from 2 fs before the pulse
      peak (seed 0); eight wires spanning the Nk range are the oracle cases.
  C5  served requests (wire + 1-D line of 1024 cells, 16-cell CPML); the oracle case is one
      request of Nk = 64 run for 500 steps.

Large arrays are compared through a Rademacher sketch: S x with a fixed random +-1 matrix S
of SKETCH_ROWS rows, regenerated from a seed on both sides.  |S (a - b)| / |S b| estimates
the relative L2 |a - b| / |b| (unbiased in the square, relative spread sqrt(2 / rows) = 0.25):
an error anywhere in the array shows in it.
"""

import math

import numpy as np

SKETCH_ROWS = 32
_CHUNK = 1 << 22
C0 = 2.99792458e8


def sketch(arrays, seed, rows=SKETCH_ROWS):
    """Rademacher sketches [len(arrays)][rows] (complex) of equally sized complex arrays.
    The sign stream depends only on the seed, the size and the fixed chunking, so both sides
    regenerate the same matrix."""
    flats = [np.ascontiguousarray(a, dtype=np.complex128).reshape(-1) for a in arrays]
    n = flats[0].size
    assert all(f.size == n for f in flats)
    views = [f.view(np.float64).reshape(-1, 2) for f in flats]
    out = np.zeros((len(flats), rows), dtype=np.complex128)
    rng = np.random.default_rng(seed)
    for j in range(rows):
        for o in range(0, n, _CHUNK):
            m = min(_CHUNK, n - o)
            s = rng.integers(0, 2, size=m, dtype=np.int8).astype(np.float64) * 2.0 - 1.0
            for a, v in enumerate(views):
                re, im = s @ v[o:o + m]
                out[a, j] += re + 1j * im
    return out


def sketch_rel(sa, sb):
    """Relative L2 estimate |a - b| / |b| from the sketches of a and b."""
    return float(np.linalg.norm(sa - sb) / max(np.linalg.norm(sb), 1e-300))


# ---------------------------------------------------------------- field + wire configurations
def spectral_dt(dx, v, dims):
    """0.9 of the leapfrog spectral limit 2 dx / (pi v sqrt(dims))."""
    return 0.9 * 2.0 * dx / (math.pi * v * math.sqrt(dims))


COUPLED = {
    "C1": dict(shape=(1024, 1, 1), dx=10e-9, epsr=9.1, npml=16, nk=128, steps=2000, checkpoints=(1000, 2000),
               amp_seed=1, src_plane=False, screened=False, current="delta"),
    "C2": dict(shape=(128, 128, 128), dx=20e-9, epsr=9.1, npml=12, nk=256, steps=10, checkpoints=(5, 10),
               amp=1e7, src_plane=True, screened=False),
    "C3": dict(shape=(256, 256, 256), dx=20e-9, epsr=9.1, npml=12, nk=1024, steps=2, checkpoints=(2,),
               amp=1e7, src_plane=True, screened=False),
}


def coupled_params(name):
    c = dict(COUPLED[name])
    Nx, Ny, Nz = c["shape"]
    epsr_eff = 1.0 if Nz == 1 else c["epsr"]  # GetEpsr
    v = C0 / math.sqrt(epsr_eff)
    dims = sum(1 for n in c["shape"] if n > 1)
    c["dt"] = spectral_dt(c["dx"], v, dims)
    if "amp" not in c:
        c["amp"] = float(10.0 ** np.random.default_rng(c["amp_seed"]).uniform(6.0, 9.0))
    # the wire's sample point: grid centre, or the source plane (x index round(0.25 Nx))
    x = np.linspace(-c["dx"] * (Nx - 1) / 2.0, c["dx"] * (Nx - 1) / 2.0, Nx)
    c["x0"] = float(x[int(round(0.25 * Nx))]) if c["src_plane"] else 0.0
    c["pulse"] = (800e-9, c["amp"], 100e-15, 50e-15, 0.0)
    c.setdefault("current", "reference")
    return c


def oracle_coupled(name, current=None):
    """(propagator, coupling) of the FP64 oracle for config `name`."""
    from oracle import pstd as OP
    from oracle.coupling import OracleCoupling
    from oracle.sbe import L_for_nk, MBFlags, QWParams

    c = coupled_params(name)
    sp = OP.Space(3, *c["shape"], c["dx"], c["dx"], c["dx"], c["epsr"])
    oc = OracleCoupling(sp, 800e-9, 0.0, Nw=1, qw=QWParams(L=L_for_nk(c["nk"])), mb=MBFlags(Screened=c["screened"]),
                        x0=c["x0"], current=current or c["current"])
    prop = OP.Propagator(sp, OP.Time(0.0, 1.0, c["dt"], 0), OP.Pulse(*c["pulse"]), "soft", "cpml", calc_j=oc.calc_j,
                         npml=c["npml"])
    assert oc.wire.Nk == c["nk"]
    return prop, oc


def device_coupled(name, precision="fp64", feedback=True, current=None, boundary="cpml"):
    """(propagator, QWArray) of the HIP path for config `name` (device time; pipelined-capable
    with the CPML; boundary "mask" swaps in the reference's default absorber)."""
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    c = coupled_params(name)
    sp = ss(3, *c["shape"], c["dx"], c["dx"], c["dx"], c["epsr"])
    qa = QWArray(QWParams(L=L_of_nk(c["nk"])), MBFlags(Screened=c["screened"], ignorewire=not feedback), Nw=1,
                 Emax0=0.0, x0=c["x0"], precision=precision, current=current or c["current"])
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, c["dt"], 0), ps(*c["pulse"]), source_type="soft",
                            boundary_type=boundary, calc_j=qa, npml=c["npml"], check_every=0, precision=precision,
                            device_time=True)
    assert qa.model.Nk == c["nk"]
    return prop, qa


FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def coupled_record(fields, wire_state, P, Rho, seed):
    """What a checkpoint keeps: field sketches and norms, the wire's populations, the
    sketches of p, C, D, and the wire's polarisation and charge lines (on the x line)."""
    rec = {}
    sk = sketch([fields[n] for n in FIELDS], seed)
    for i, n in enumerate(FIELDS):
        rec["sk_" + n] = sk[i]
        rec["norm_" + n] = np.linalg.norm(fields[n])
    if fields["Ex"].size <= 1 << 16:  # 1-D lines: the fields themselves (copies: the run goes on)
        for n in FIELDS:
            rec["f_" + n] = np.array(fields[n], copy=True).reshape(-1)
    sk = sketch([wire_state["p2"], wire_state["C2"], wire_state["D2"]], seed + 1)
    for i, n in enumerate(("p2", "C2", "D2")):
        rec["sk_" + n] = sk[i]
        rec["norm_" + n] = np.linalg.norm(wire_state[n])
    rec["ne"] = np.diag(wire_state["C2"]).real.copy()
    rec["nh"] = np.diag(wire_state["D2"]).real.copy()
    rec["P"] = np.asarray(P)
    rec["Rho"] = np.asarray(Rho)
    return rec


# ---------------------------------------------------------------- batched wires (C4)
C4 = dict(nw=512, seed=0, nk_lo=64, nk_hi=1024, e_lo=1e6, e_hi=1e9, dt=10e-18, steps=2, nrr=16, screened=True)


def c4_specs():
    from pulsesuite_amd.batch import sample_specs

    specs = sample_specs(C4["nw"], seed=C4["seed"], nk_lo=C4["nk_lo"], nk_hi=C4["nk_hi"], e_lo=C4["e_lo"],
                         e_hi=C4["e_hi"])
    for sp in specs:  # 2 fs before the pulse peak: the steps are driven
        sp.t0 = sp.tp - 2e-15
    return specs


def c4_oracle_picks(specs):
    """Eight spec indices spanning the Nk range (sorted-Nk quantiles)."""
    order = sorted(range(len(specs)), key=lambda i: (specs[i].Nk, i))
    return [order[int(round(q * (len(order) - 1)))] for q in np.linspace(0.0, 1.0, 8)]


def c4_drive(sp, w0, t):
    """Uniform Gaussian drive of a batched wire (docs/source/examples/sbe_wire_1d.py:45-54)."""
    u = w0 * (t - sp.tp)
    w0tw = w0 * sp.tw
    return math.exp(-(u**2) / w0tw**2) * math.cos(u) * math.exp(-(u**20) / (2 * w0tw) ** 20)


def c4_oracle_wire(sp):
    """Oracle steps of one batched wire: returns (state after C4['steps'] steps, last outputs)."""
    from oracle import sbe as O

    HBAR, E = 1.0545718176461565e-34, 1.602176634e-19
    L = O.L_for_nk(sp.Nk)
    rr = np.linspace(-1.2 * L, 1.2 * L, C4["nrr"])
    ref = O.OracleWire(rr, 0.0, 0.0, 800e-9, 1, O.QWParams(L=L), O.MBFlags(Screened=C4["screened"]))
    assert ref.Nk == sp.Nk
    w0 = (ref.gap + sp.detuning_eV * E) / HBAR
    r = None
    for n in range(C4["steps"]):
        f = c4_drive(sp, w0, sp.t0 + n * C4["dt"])
        Ex, Ey, Ez = (np.full(C4["nrr"], sp.E0 * c * f, dtype=complex) for c in sp.pol)
        r = ref.QWCalculator(Ex, Ey, Ez, rr, C4["dt"], 1)
    return ref.state[0], r


# ---------------------------------------------------------------- serving (C5)
C5_ORACLE = dict(nk=64, steps=500, E0=4e7)
# (Nk, steps, E0) served with the wire on the source plane, so every step is driven
C5_MIX = ((2048, 12, 2e7), (700, 30, 1e7), (130, 40, 5e7), (64, 50, 3e7))
