/*
 * pss.h -- C ABI of the MI355X-native coupled SBE-PSTD step (libpss.so).
 *
 * All pointers are device pointers (HIP, gfx950) unless a comment says "host".
 * Complex arrays are interleaved (re, im) float64 pairs, i.e. the layout of
 * numpy/torch complex128.  Every call returns 0 on success or a nonzero
 * status; pss_last_error() gives the message.  No call allocates memory: the
 * caller owns every buffer (state, tables, workspace).  A descriptor array is
 * used by one host thread at a time; calls are asynchronous on `stream`.
 *
 * Citations are paths relative to the reference's src/pulsesuite.
 */
#ifndef PSS_H
#define PSS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Many-body flags, bit positions follow mb.params order (PSTD3D/SBEs.py:5016-5034). */
enum {
  PSS_OPTICS = 1 << 0,
  PSS_EXCITONS = 1 << 1,
  PSS_EHS = 1 << 2,
  PSS_SCREENED = 1 << 3,
  PSS_PHONON = 1 << 4,
  PSS_DCTRANS = 1 << 5,
  PSS_LF = 1 << 6,
  PSS_DIAGDPH = 1 << 8,
  PSS_OFFDIAGDPH = 1 << 9,
  PSS_RECOMB = 1 << 10,
  PSS_IGNOREWIRE = 1 << 12,
  PSS_LORENTZDELTA = 1 << 14,
};

/* Precision of the contractions. */
enum { PSS_FP64 = 0, PSS_MIXED = 1 };

/*
 * One quantum wire (one sub-band index w of an SBESolver) resident on the GPU.
 * Filled by the host shim; copied once to device memory; the step kernels read
 * it from there (one descriptor per wire in a batch).
 */
typedef struct pss_wire_desc {
  int32_t Nk, Nr, nrr, w;        /* k points, wire grid points, propagation points, 1-based wire */
  int32_t flags, precision, pad0, pad1;
  double dt, L, gap, ehint, area, dcv, dkr, dr_r, drr, yw, zsign;
  double gam_e, gam_h, gam_eh;
  double Vscale, N0, hOph, hGph, Gph;
  double density_max, gate;      /* gate = 1e-3 * Emax0 (wireoff threshold) */
  double Edc;                    /* background DC field (V/m) of the transport shift (PSS_DCTRANS) */
  double rsp[4];                 /* PSS_RECOMB: Rsp(E) = rsp[0] + rsp[1] E + rsp[2] E^2 + rsp[3] E^3 */
  double pad2[3];
  /* tables (read only) */
  const double *kr, *Ee, *Eh;    /* [Nk] */
  const double *win;             /* [Nr] QW window on r */
  const double *r;               /* [Nr] */
  const double *V0;              /* [3][Nk] unscreened V_eh, V_ee, V_hh by |k1-k2| */
  const double *chie, *chih, *qe, *qh; /* [Nk] Toeplitz rows of Chi1De/Chi1Dh/qe/qh */
  const int32_t *ktab;           /* [6][Nk*Nk] k_p_q, k_m_q, k1p, k1p_m_q, k1, k1_m_q (row k, col q) */
  const double *xe, *xh;         /* [Nk] */
  const int32_t *iqv;            /* [Nk] Vxx2 index table */
  const double *twid;            /* [2*Nr] exp(-2 pi i m/Nr), interleaved */
  /* spline of the propagation grid rr (Prop2QW source) and its evaluation at r */
  const double *sa_x, *sa_t, *sa_b, *sa_h; /* [nrr] knots, elimination factors, pivots, spacings */
  const double *sa_end;          /* [8] end-condition constants */
  const int32_t *sa_idx;         /* [Nr] interval index of r_j (-1 = outside) */
  const double *sa_dx;           /* [Nr] r_j - rr[idx] */
  /* spline of the wire grid r (QW2Prop source) and its evaluation at rr */
  const double *sb_t, *sb_b, *sb_h, *sb_end;
  const int32_t *sb_idx;         /* [nrr] */
  const double *sb_dx;           /* [nrr] */
  /* state (read/write) */
  double *p[3], *C[3], *D[3];    /* complex [Nk*Nk], rotating prev/cur/next */
  int32_t *ctl;                  /* [4] step counter of this wire, -, -, - */
  int32_t *on;                   /* wireoff latch, shared by the wires of one solver */
  /* per-step input / output (complex [nrr]) */
  const double *Exx, *Eyy, *Ezz;
  double *Pxx, *Pyy, *Pzz, *Rho;
  /* workspace: pss_sbe_workspace_bytes(Nk, Nr, nrr, precision) bytes, 256-byte aligned */
  double *ws;
} pss_wire_desc;

/* Bytes of workspace one wire needs. */
int64_t pss_sbe_workspace_bytes(int32_t Nk, int32_t Nr, int32_t nrr, int32_t precision);

/* Split-K factor of the contractions (power of two <= 16; 0 = each wire's own choice from
 * its Nk, see pss_sbe_step_p).  Process-wide; a tuning knob with no reference counterpart.
 * A wire's summation order follows it, so results under different settings differ by
 * rounding. */
int pss_sbe_tune(int32_t ksplit);

/*
 * One SBE time step of `nwires` wires whose descriptors sit in device memory
 * at `d_descs`.  Replaces, per wire, SBESolver.QWCalculator minus file output
 * (PSTD3D/SBEs.py:5838-5997): Prop2QW, _sbe_calculator, QW2Prop.
 * min_Nk / max_* bound the per-wire sizes in the batch (grid sizing); min_Nk must not
 * exceed the smallest wire's Nk.  Every wire's arithmetic depends on its own sizes only
 * (its split-K factor is chosen from its own Nk), so a wire gives the same result bit for
 * bit whichever wires share the launch.
 */
int pss_sbe_step(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                 int32_t max_nrr, void* stream);
/* The same step for a batch whose wires hold only the kinds in the mask `precisions`
 * (1: PSS_FP64 wires, 2: PSS_MIXED wires, at least one of them; 4: wires with PSS_DCTRANS,
 * 8: wires with PSS_RECOMB): the kernels of an absent kind are not launched (pss_sbe_step
 * launches all of them; each returns at once for wires it does not apply to).  A wire whose
 * flags ask for DCTrans or Recomb in a launch without that bit gets NaN outputs (Pxx, Pyy,
 * Pzz, Rho) rather than a step without the term.  ksplit_cap > 0 (a power of two <= 16)
 * caps every wire's split-K factor, which is otherwise chosen to fill the chip as if the
 * wire ran alone: a step that runs beside other work (the wire step of the coupled step)
 * does better with fewer workgroups.  pss_sbe_step = precisions 15, no cap.
 *
 * DCTrans (PSTD3D/dcfield.py:1213-1247, called at PSTD3D/SBEs.py:5698-5702): after the
 * relaxation, C3, D3 and p3 are shifted in momentum by dk = -e (Edc + Re Ex(q=0)) dt / hbar,
 * both indices when PSS_LF is set, else the diagonals only.  The FFT shift of the reference
 * is the circulant matrix S = F^-1 diag(exp(-i Y dk)) F (X -> S X S^T); its first column
 * has the closed form of a Dirichlet kernel, and the two products run as FP64 tiles.
 * Recomb (PSTD3D/emission.py:94-131, SBEs.py:5335-5337): Rsp from the cur populations and
 * screened V, then ne3 *= exp(-Rsp nh3 dt), nh3 *= exp(-Rsp ne3 dt) inside the relaxation. */
int pss_sbe_step_p(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                   int32_t max_nrr, int32_t precisions, int32_t ksplit_cap, void* stream);

/*
 * Kernel-level entries: parts of pss_sbe_step on the wires' own buffers, with the same kernels.
 * The state buffers rotate with the wire's step counter ctl[0]: "cur" is buffer 1 + (ctl[0] & 1),
 * "next" buffer 2 - (ctl[0] & 1), "prev" buffer 0.
 *
 * pss_calc_h: CalcH (PSTD3D/SBEs.py:820-935, _calc_Meh :5165-5175) of the cur state and the
 *   propagation lines Exx/Eyy/Ezz, with everything it needs before it (Prop2QW
 *   PSTD3D/qwoptics.py:673-721, populations and CalcScreenedArrays PSTD3D/coulomb.py:1195-1230,
 *   the diagonal traces of the direct terms; CalcGammaE/H PSTD3D/dephasing.py:540-624 as a
 *   by-product).  Results in the workspace: H = (Heh, Hee, Hhh) complex [3][Nk][Nk], the
 *   screened V(|q|) rows [3][Nk], E1D [Nk], GamE/GamH [2][Nk].  Does not touch the state.
 * pss_sbe_rhs: dpdt, dCdt, dDdt (PSTD3D/SBEs.py:1953-1976, 2320-2338, 2618-2636) of the cur
 *   state with the Hamiltonian, dephasing rates and off-diagonal dephasing
 *   (PSTD3D/dephasing.py:713-742) the step would use, written to the next buffers (dp/dt,
 *   dC/dt, dD/dt); prev and cur untouched.
 * pss_coll_rates: MBCE/MBCH (PSTD3D/coulomb.py:957-1071, which & 1) and MBPE/MBPH
 *   (PSTD3D/phonons.py:140-268, which & 2) of the populations on the next buffers' diagonals,
 *   with the screened V(|q|) rows and E1D in the workspace: Win_e, Wout_e, Win_h, Wout_h
 *   [4][Nk] in the workspace (the wire's flags must include the terms asked for).
 * pss_qwpol: QWPolarization3 and the k sums of QWRho5 (PSTD3D/qwoptics.py:815-946, 948-1077
 *   before their centred FFTs) of the next buffers p, C, D: P(r) [3][Nr] and rho_e, rho_h
 *   complex [2][Nr] in the workspace.  Populations are not relaxed or renormalised.
 * Offsets of the workspace segments: ws_layout in csrc/sbe.hip (mirrored by
 * pulsesuite_amd/SBEs.py ws_layout).
 */
/* Any subset of the step's stages, in order (bits of `stages`): 1 FRONT (wireoff gate,
 * Prop2QW, populations, screening, E(q)), 2 TRACES (direct-term traces, CalcGammaE/H),
 * 4 OFFDIAG (De, Dh), 8 CALCH (Hamiltonians), 16 RHS (contractions + epilogue), 32 COLL
 * (Win/Wout), 64 TAIL (relaxation, renormalisation, P(r), rho(r), QW2Prop).  opt: bit 0 =
 * RHS writes dX/dt only; bit 1 = TAIL without relaxation / renormalisation and QW2Prop;
 * bits 8.. = a mask on the wires' flags for COLL.  pss_sbe_step = all stages, opt 0. */
int pss_sbe_stages(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                   int32_t max_nrr, int32_t precisions, int32_t stages, int32_t opt, void* stream);
int pss_calc_h(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
               int32_t max_nrr, int32_t precisions, void* stream);
int pss_sbe_rhs(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                int32_t max_nrr, int32_t precisions, void* stream);
int pss_coll_rates(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                   int32_t max_nrr, int32_t which, void* stream);
int pss_qwpol(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
              int32_t max_nrr, void* stream);

/* dpdt (which 0), dCdt (1), dDdt (2) of explicit complex [Nk][Nk] matrices, the reference's
 * module functions (PSTD3D/SBEs.py:1953-1976, 2320-2338, 2618-2636) with the same argument
 * meaning: dpdt uses GamE, GamH [Nk] and Off (OffP, may be null = 0); out [Nk][Nk]. */
int pss_rhs_dense(int32_t Nk, int32_t which, const double* C, const double* D, const double* p, const double* Heh,
                  const double* Hee, const double* Hhh, const double* GamE, const double* GamH, const double* Off,
                  double* out, void* stream);

/* State of one wire (descriptor at d_desc, device) to / from a packed device buffer
 * [6][Nk][Nk] complex = (p1, p2, C1, C2, D1, D2) (the reference's YY1/YY2/CC1/CC2/DD1/DD2,
 * PSTD3D/SBEs.py:4782-4790, as its checkpoint writes them, :5536-5581).  The prev / cur
 * buffers are resolved on the device from ctl[0], so the copies can be captured. */
int pss_sbe_get_state(const pss_wire_desc* d_desc, int32_t Nk, double* packed, void* stream);
int pss_sbe_set_state(const pss_wire_desc* d_desc, int32_t Nk, const double* packed, void* stream);

/*
 * Uniform drive of batched wires: Exx/Eyy/Ezz of every wire are set along the wire to
 * E0_c exp(-u^2/(w0 tw)^2) cos(u) exp(-u^20/(2 w0 tw)^20), u = w0 (t - tp), with
 * t = t0 + ctl[1] dt and ctl[1] advanced by one (the Gaussian pulse of
 * docs/source/examples/sbe_wire_1d.py:45-54 applied as at :99-100).
 * d_pulses: device [nwires][8] = {E0x, E0y, E0z, w0, tw, tp, t0, unused}.
 */
int pss_drive_uniform(const pss_wire_desc* d_descs, const double* d_pulses, int32_t nwires, void* stream);

/*
 * Coulomb integral V(q) = prefactor * Vint(q) for a list of q (PSTD3D/coulomb.py:70-104,
 * 502-577), with the reference's series/asymptotic K0 (PSTD3D/coulomb.py:39-66).
 * y: [Ny] wire grid; q: [nq]; out: [nq].
 */
int pss_vint(const double* y, int32_t Ny, const double* q, int32_t nq, double alpha1, double alpha2,
             double Delta0, double prefactor, double* out, void* stream);

/* ------------------------------------------------------------------ PSTD */

/* 3-D grid (C order: x slowest, z fastest), complex128 fields. */
typedef struct pss_grid {
  int32_t Nx, Ny, Nz, npml_x, npml_y, npml_z;
  int32_t precision; /* PSS_FP64, or PSS_MIXED: the CPML's spectral derivatives in FP32 */
  int32_t pad1;
  double dt, v2, mu0;
  const double *kx, *ky, *kz;            /* [N] FFT-order wavenumbers */
  const double *bx, *cx, *ikx;           /* [Nx] CPML b, c, 1/kappa - 1 */
  const double *by, *cy, *iky;
  const double *bz, *cz, *ikz;
  const double *twx, *twy, *twz;         /* [N] FFT twiddles exp(-2 pi i m/N), interleaved, N/2 used */
  const double* tdev;                    /* optional device time: the source time is *tdev + src_par[0] */
} pss_grid;

/*
 * E += i v2 dt (k x B) - mu0 v2 dt J  (PSTD3D/PSTD3D.py:29-76).  J is either a dense k-space
 * field triple (Jx,Jy,Jz non-null) or separable: J_c(k) = jline_c(kx) gy(ky) gz(kz) given by
 * jline [3][Nx], gy [Ny], gz [Nz] (complex), as deposited by the wire coupling.
 */
int pss_update_e3d(const pss_grid* g, const double* Bx, const double* By, const double* Bz,
                   const double* Jx, const double* Jy, const double* Jz, const double* jline,
                   const double* gy, const double* gz, double* Ex, double* Ey, double* Ez, void* stream);
/* pss_update_e3d followed by the point sample of the new E (as pss_ksample), in one
 * pass over E: partial needs 6144 doubles of device scratch, sample_out[6] = (Re, Im)
 * of each component at the wire point.  clock (may be null): a device scalar advanced by
 * tick after the sample (the soft source's device time, as pss_add_scalar would).
 * Bx, By, Bz all null: the current term only (E -= mu0 v2 dt J), for a step whose curl
 * term was applied before by pss_cpml_e_curl.  lines (may be null): nlines wire lines
 * [nlines][3][line_n] complex, each filled with the sample as pss_fill_line would.
 * update_b != 0 (B given): also B -= i dt (k x E) from the new E in the same pass (the
 * curl term of UpdateB3D, PSTD3D/PSTD3D.py:79-113), so B is written too. */
int pss_update_e3d_sample(const pss_grid* g, const double* Bx, const double* By, const double* Bz,
                          const double* Jx, const double* Jy, const double* Jz, const double* jline,
                          const double* gy, const double* gz, double* Ex, double* Ey, double* Ez,
                          const int32_t* sample_idx, const double* sample_w, double* partial, double* sample_out,
                          double* lines, int32_t line_n, int32_t nlines, double* clock, double tick,
                          int32_t update_b, void* stream);
/* B -= i dt (k x E)  (PSTD3D/PSTD3D.py:79-113). */
int pss_update_b3d(const pss_grid* g, const double* Ex, const double* Ey, const double* Ez,
                   double* Bx, double* By, double* Bz, void* stream);
/* In-place 3-D FFT, numpy convention: forward unnormalised, inverse scaled 1/N
 * (core/fftw.py:217-226).  dir = -1 forward, +1 inverse. */
int pss_fft3d(const pss_grid* g, double* F, int32_t dir, void* stream);

/*
 * One CPML half step on fields F (k-space in, k-space out) driven by G:
 * which = 0: F = E, G = B (ApplyCPML_E, PSTD3D/cpml.py:640-702, scale v2*dt, sign +1);
 * which = 1: F = B, G = E (ApplyCPML_B, PSTD3D/cpml.py:575-638, scale dt, sign -1).
 * psi: the 6 auxiliary arrays (order xy, xz, yx, yz, zx, zy), each stored on the PML
 * slab of its derivative axis b (shape with axis b of length 2 npml_b), back to back:
 * pss_cpml_psi_len(g) doubles.  scratch: 6 complex [N].
 * src_comp >= 0 fuses UpdateTFSC (PSTD3D/tfsf.py:63-142) on that component: the
 * additive S(x) E_inc(x,y,z,t) dx is applied in real space before the forward FFT;
 * src_par (host, 9 doubles) = {t, amp, omega0, tau_G, chirp, v, Tp, 1/w0^2, dx};
 * srcx: [Nx] profile; rx/ry/rz: grid coordinates.
 * sample_idx (may be null): 8 flat cell indices (-1 = unused) and sample_w 8 weights
 * of a trilinear sample (EAtXYZ, PSTD3D/usefulsubs.py:2016-2085) of the corrected
 * real-space F, written to sample_out (complex [3]) -- the wire's point field.
 */
int pss_cpml_half(const pss_grid* g, int32_t which, double* Fx, double* Fy, double* Fz,
                  const double* Gx, const double* Gy, const double* Gz, double* psi, double* scratch,
                  int32_t src_comp, const double* src_par, const double* srcx, const double* rx,
                  const double* ry, const double* rz, const int32_t* sample_idx, const double* sample_w,
                  double* sample_out, void* stream);

/* The B half step of a leapfrog step, UpdateB3D followed by ApplyCPML_B
 * (PSTD3D/PSTD3D.py:79-113 and PSTD3D/cpml.py:575-638, as PSTD3D/PSTD3D.py:523-627 calls
 * them): B -= i dt (k x E), then the CPML-B correction with the soft source on src_comp.
 * Both updates read only E, so in mixed precision the curl term is added in the last
 * pass of the correction's forward transform when that axis is 64 ... 1024 long (one
 * read-modify-write of B instead of two); FP64 grids and other lengths run the two
 * updates one after the other (with the FP64 correction the fused pass was slower).
 * Arguments as for pss_update_b3d and pss_cpml_half (which = 1, no sample). */
int pss_b_half(const pss_grid* g, double* Bx, double* By, double* Bz, const double* Ex, const double* Ey,
               const double* Ez, double* psi, double* scratch, int32_t src_comp, const double* src_par,
               const double* srcx, const double* rx, const double* ry, const double* rz, void* stream);

/* The CPML-E correction (PSTD3D/cpml.py:640-702) together with the curl term of the E
 * update, E += i v2 dt (k x B) (PSTD3D/PSTD3D.py:29-76 without its current term): both
 * read only B, so the curl term is added in the last forward pass of the correction when
 * that axis is 64 ... 1024 long, else the two updates run one after the other.  The
 * current term then follows with pss_update_e3d_sample and null B.  Arguments as for
 * pss_update_e3d and pss_cpml_half (which = 0, no sample). */
int pss_cpml_e_curl(const pss_grid* g, double* Ex, double* Ey, double* Ez, const double* Bx, const double* By,
                    const double* Bz, double* psi, double* scratch, int32_t src_comp, const double* src_par,
                    const double* srcx, const double* rx, const double* ry, const double* rz, void* stream);

/* The masking absorber of one half step, ApplyAbsorber_E / ApplyAbsorber_B
 * (PSTD3D/absorber.py:160-191), followed by the soft source UpdateTFSC on src_comp
 * (PSTD3D/tfsf.py:63-142, as PSTD3D/PSTD3D.py:523-627 calls them): the three k-space
 * fields go to real space, are multiplied by the separable mask mx[i] my[j] mz[k]
 * (InitAbsorber, :125-158) and, on src_comp >= 0, get S(x) E_inc dx added (src_par, srcx,
 * rx/ry/rz as for pss_cpml_half), then go back to k-space.  Axis lengths: powers of two
 * <= 2048. */
int pss_mask_half(const pss_grid* g, double* Fx, double* Fy, double* Fz, const double* mx, const double* my,
                  const double* mz, int32_t src_comp, const double* src_par, const double* srcx, const double* rx,
                  const double* ry, const double* rz, void* stream);

/* Launch timing of the accumulating last pass with the curl term (pss_b_half and
 * pss_cpml_e_curl in mixed precision): no counterpart in the reference; for measurement.
 * pss_pass_timing_arm(n) arms n event pairs (created on the first call that needs them, so
 * call it outside any stream capture); each following such launch on a stream that is not
 * capturing records one pair around itself, until n are taken.  n = 0 disarms.
 * pss_pass_timing_read, after the work has completed, writes the durations (ms) of the
 * first min(n, taken) launches and the count taken. */
int pss_pass_timing_arm(int32_t n);
int pss_pass_timing_read(float* ms, int32_t n, int32_t* got);

/* The trilinear point sample of pss_cpml_half taken on its own, from k-space fields that
 * already hold the whole half step (EAtXYZ of the IFFT, PSTD3D/usefulsubs.py:2016-2085):
 * sample_out[6] = (Re, Im) of each component.  partial: 6144 doubles of device scratch.
 * Axis lengths: powers of two <= 2048. */
int pss_ksample(const pss_grid* g, const double* Fx, const double* Fy, const double* Fz, const int32_t* sample_idx,
                const double* sample_w, double* partial, double* sample_out, void* stream);

/* pss_ksample without the partial buffer: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc): the per-block partial sums go to a 6144-double (48 KiB) array in the
 * library's own device memory, one per GPU, held until the process ends.  The call
 * allocates nothing.  Same kernels and arithmetic as pss_ksample, so the same result bit
 * for bit.  Grids: axis lengths powers of two <= 2048 (the partial sums never exceed 6144
 * doubles).  Calls that overlap on one GPU (two streams, two threads) share the array and
 * race; serialise them. */
int pss_ksample_noalloc(const pss_grid* g, const double* Fx, const double* Fy, const double* Fz,
                        const int32_t* sample_idx, const double* sample_w, double* sample_out, void* stream);

/* Length in doubles of the psi buffer of pss_cpml_half for this grid (sizing helper of
 * the CPML state of PSTD3D/cpml.py:340-450, which keeps full-grid psi arrays). */
int64_t pss_cpml_psi_len(const pss_grid* g);

/* One FFT pass along `axis` of the grid's shape (the other axes batch), in -> out
 * (may alias).  Used for 1-D transforms of the separable current factors. */
int pss_fft_line(const pss_grid* g, const double* in, double* out, int32_t axis, int32_t dir, void* stream);
/* The same pass over nb <= 3 arrays of the grid's shape in one launch: in[b] -> out[b]
 * (host arrays of device pointers; each pair may alias).  Used for the three components
 * of the separable current line J(x), whose product with the transformed transverse
 * gates is the 3-D FFT of the deposited current (QWPlacement, PSTD3D/rhoPJ.py:736-781). */
int pss_fft_line_batch(const pss_grid* g, double* const* in, double* const* out, int32_t nb, int32_t axis,
                       int32_t dir, void* stream);

/* One accumulating pass of a 3-D transform along `axis`, as the last forward pass of the
 * CPML correction's dF transform runs it (PSTD3D/cpml.py:575-702 add FFT(dF) to the
 * k-space field): out_b += pass(in_b) for b < nb, in read as complex64 (in32 = 1, axes of
 * 64 ... 1024) or complex128, out complex128; in is not modified.  With Gx, Gy, Gz non-null
 * (nb = 3) the curl term coef i (k x G)_b of the E or B update (PSTD3D/PSTD3D.py:29-113) is
 * added in the same pass (coef = v2 dt for E from B, -dt for B from E).  A kernel-level entry
 * for tests and for timing that pass alone. */
int pss_fft_pass_acc(const pss_grid* g, double* const* in, double* const* out, int32_t nb, int32_t axis, int32_t dir,
                     int32_t in32, const double* Gx, const double* Gy, const double* Gz, double coef, void* stream);
/* The last two forward passes (along y and z) of the CPML correction's dF transform with
 * the accumulation and the curl term, as the mixed-precision correction runs them when
 * Ny = Nz = 64 or 128: one 2-D transform per x-plane in LDS.  in: three complex64 arrays
 * (dF after its forward pass along x; dF is real, so only the x-planes 0 ... Nx/2 are read
 * and plane Nx - g of the result is taken as conj of plane g with (ky, kz) reversed), out:
 * complex128 fields, out_c += FFT_yz(in_c) + coef i (k x G)_c
 * (PSTD3D/cpml.py:575-702 and the curl term of PSTD3D/PSTD3D.py:29-113).  A kernel-level
 * entry for tests and for timing that kernel alone; no counterpart entry point in the
 * reference. */
int pss_fft_planes_acc(const pss_grid* g, const double* const* in, double* const* out, const double* Gx,
                       const double* Gy, const double* Gz, double coef, void* stream);

/* Fill three uniform wire lines [3][n] with the point sample (complex [3]): the
 * wire sees E at one point (QWArray.QuantumWire, PSTD3D/rhoPJ.py:1160-1164). */
int pss_fill_line(const double* sample, double* lines, int32_t n, void* stream);

/* Current line of the wires (PSTD3D/rhoPJ.py:1183-1194 with CalcJfx :561-592):
 * J_c = sum_w (P_cw - Pold_cw)/dt/Nw, J_x += CalcJfx(sum Rho, sum RhoOld, dt, dx)/Nw;
 * then Pold <- P, RhoOld <- Rho.  P: [Nw][3][Nx], Rho: [Nw][Nx], jline: [3][Nx].
 * keep_old = 0 is the reference's behaviour: QWArray.QuantumWire re-creates PxxOld and
 * RhoOld as zeros on every call (PSTD3D/rhoPJ.py:1138-1147), so J_c = sum_w P_cw/dt/Nw
 * and CalcJfx sees RhoOld = 0; keep_old = 1 keeps the previous call's values
 * (J = dP/dt).  Pold and RhoOld are written in both cases.
 * fft != 0: each line then takes its forward FFT along x (numpy convention, as
 * pss_fft_line_batch with dir = -1) in the same kernel; Nx a power of two. */
int pss_jline(const double* P, const double* Rho, double* Pold, double* RhoOld, int32_t Nw, int32_t Nx,
              double dt, double dx, double* jline, int32_t fft, int32_t keep_old, void* stream);

/*
 * Longitudinal-field diagnostics of QWArray.QuantumWire (no counterpart entry points in the
 * reference; they carry its ElongfromRho and RhoB steps, PSTD3D/rhoPJ.py:784-969, 1206-1285).
 * pss_elong_k: the k-space loop of ElongfromRho (rhoPJ.py:596-700) on the transforms of rho
 * and P (all complex [Nx][Ny][Nz], the grid's FFT-order kx, ky, kz): out[10] = E from P + rho
 * (x, y, z), E from P (x, y, z), E from rho (x, y, z), RhoB, still in k-space; small and
 * inv_eps as the reference's (dq_x / 1e4, 1 / (eps0 eps_r)).
 */
int pss_elong_k(const pss_grid* g, const double* Rho_k, const double* Px, const double* Py, const double* Pz,
                double* const* out, double small, double inv_eps, void* stream);
/* QWPlacement (rhoPJ.py:736-781) of nterms <= 3 placed terms summed in order into out
 * [Nx][Ny][Nz] complex: term t is sign[t] (gx[t][i] gy[t][j]) gz[t][k] F_t(i), F_t the mean of
 * nw complex lines (lines[t] + w lstride[t] complex), or for mode[t] = 1 its periodic central
 * difference (F(i+1) - F(i-1)) / (2 dx); gx null (or gx[t] null) means gx = 1. */
int pss_qw_place(int32_t Nx, int32_t Ny, int32_t Nz, int32_t nw, int32_t nterms, const double* const* lines,
                 const int64_t* lstride, const double* const* gx, const double* const* gy, const double* const* gz,
                 const double* sign, const int32_t* mode, double dx, double* out, void* stream);
/* F *= exp(-(sx^2 kx^2 + sy^2 ky^2 + sz^2 kz^2) / 2) on a complex k-space grid (the
 * smoothing filter of RhoBSmooth / RhoBAnSmooth, rhoPJ.py:1263-1285). */
int pss_kfilter(const pss_grid* g, double* F, double sx, double sy, double sz, void* stream);

/*
 * Fused stage combine of the classic RK4 step (libpulsesuite/integrator.py:143-157,
 * Numerical Recipes stage order) on n doubles (a complex array is 2n doubles):
 *   stage 0: ytmp = y + (h/2) dydt
 *   stage 1: ytmp = y + (h/2) dyt
 *   stage 2: ytmp = y + h dym;  dym = dym + dyt
 *   stage 3: y = y + (h/6) (dydt + dyt + 2 dym)
 * Separately rounded FP64 operations in that order (bit-identical to the expressions).
 */
int pss_rk4_stage(int32_t stage, int64_t n, double h, double* y, const double* dydt, const double* dyt,
                  double* dym, double* ytmp, void* stream);

/* x[0] += v on the device (advances a device-side time counter inside captured steps). */
int pss_add_scalar(double* x, double v, void* stream);

/* Message of the last failing call (host string, valid until the next call). */
const char* pss_last_error(void);
/* Library version string (host); starts with "pss 0.2" for this ABI. */
const char* pss_version(void);

#ifdef __cplusplus
}
#endif
#endif
