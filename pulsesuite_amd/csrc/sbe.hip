// SBE wire step on MI355X (gfx950): FP64 path.
//
// One call advances every wire of a batch by one leapfrog step, following
// SBESolver.QWCalculator -> Prop2QW -> _sbe_calculator -> QW2Prop
// (PSTD3D/SBEs.py:5838-5997, 5624-5832; PSTD3D/qwoptics.py:673-811).
// Wires are independent; every kernel takes the batch index from blockIdx.z
// (or blockIdx.y) and reads that wire's descriptor from device memory, so wires
// of different Nk share one launch (blocks past a wire's extent exit).
//
// State rotation: prev (X1) lives in buffer 0 and is averaged in place; cur (X2)
// and next (X3) swap between buffers 1 and 2 by the parity of the wire's step
// counter ctl[0]; *on is the solver-wide latch of the reference's wireoff gate.
#include <algorithm>
#include <stdio.h>

#include "common.h"
#include "../../include/pss.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kHbar = 1.0545718176461565e-34;  // scipy.constants.hbar (h / 2 pi)
constexpr double kE0 = 1.602176634e-19;      // scipy.constants.e

// n / d via the hardware reciprocal and Newton steps (within 1 ulp of IEEE division).
__device__ inline double ddiv(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  double q = n * r;
  return fma(fma(-d, q, n), r, q);
}

// ---------------------------------------------------------------- workspace
struct WS {
  int64_t er, spa, eq, vc, e1d, ne2, nh2, tr, gam, De, Dh, H, W, n3, red, ds, pr, rho, spb, outr, cp, dcc, rsp, sp,
      pk, total_fp64, total;
};

__host__ __device__ inline int64_t al(int64_t n) { return (n + 31) / 32 * 32; }

__host__ __device__ inline WS ws_layout(int Nk, int Nr, int nrr) {
  WS w;
  int64_t o = 0;
  int64_t nd = 2 * Nk - 1;
  w.er = o;   o += al(3 * (int64_t)Nr);
  w.spa = o;  o += al(3 * (int64_t)nrr);
  w.eq = o;   o += al(3 * nd * 2);
  w.vc = o;   o += al(3 * (int64_t)Nk);
  w.e1d = o;  o += al(Nk);
  w.ne2 = o;  o += al(2 * (int64_t)Nk);
  w.nh2 = o;  o += al(2 * (int64_t)Nk);
  w.tr = o;   o += al(4 * nd * 2);
  w.gam = o;  o += al(2 * (int64_t)Nk);
  w.De = o;   o += al((2 * (int64_t)Nk + 1) * Nk);
  w.Dh = o;   o += al((2 * (int64_t)Nk + 1) * Nk);
  w.H = o;    o += al(3 * (int64_t)Nk * Nk * 2);
  w.W = o;    o += al(4 * (int64_t)Nk);
  w.n3 = o;   o += al(4 * (int64_t)Nk);
  w.red = o;  o += al(32);
  w.ds = o;   o += al(4 * nd * 2);
  w.pr = o;   o += al(3 * (int64_t)Nr);
  w.rho = o;  o += al(2 * (int64_t)Nr * 2);
  w.spb = o;  o += al(5 * (int64_t)Nr);
  w.outr = o; o += al(5 * (int64_t)nrr);
  w.cp = o;   o += al(2 * (int64_t)Nk * Nk);
  w.dcc = o;  o += al(8 * (int64_t)Nk);   // DCTrans: circulant column (complex [Nk]), shifted diagonals [3][Nk]
  w.rsp = o;  o += al(Nk);                // Recomb: Rsp [Nk]
  const int64_t np = (Nk + 63) / 64 * 64;
  const int ksm = np <= 128 ? 16 : np <= 256 ? 8 : np <= 512 ? 4 : 1;
  w.pk = o;   o += al(ksm > 1 ? 6 * (int64_t)ksm * Nk * Nk : 0);  // split-K partial tiles
  w.total_fp64 = o;
  w.sp = o;   o += al(16 * np * np);  // mixed precision only: FP16 operand planes
  w.total = o;
  return w;
}

struct Bufs {
  const cplx *p1, *p2, *C1, *C2, *D1, *D2;  // prev / cur
  cplx *pp, *Cp, *Dp;                       // prev (written in place)
  cplx *pn, *Cn, *Dn;                       // next
};

__device__ inline Bufs bufs(const pss_wire_desc& d) {
  const int s = d.ctl[0] & 1;
  const int cur = 1 + s, nxt = 2 - s;
  Bufs b;
  b.p1 = (const cplx*)d.p[0]; b.C1 = (const cplx*)d.C[0]; b.D1 = (const cplx*)d.D[0];
  b.pp = (cplx*)d.p[0];       b.Cp = (cplx*)d.C[0];       b.Dp = (cplx*)d.D[0];
  b.p2 = (const cplx*)d.p[cur]; b.C2 = (const cplx*)d.C[cur]; b.D2 = (const cplx*)d.D[cur];
  b.pn = (cplx*)d.p[nxt]; b.Cn = (cplx*)d.C[nxt]; b.Dn = (cplx*)d.D[nxt];
  return b;
}

__device__ inline bool wire_on(const pss_wire_desc& d) { return *d.on != 0; }

// ------------------------------------------------------------ FMM spline
// Second-derivative-like coefficients c[] of libpulsesuite/spliner.py:1287-1363
// for real data y on knots with spacings h, given the y-independent elimination
// factors t[] and pivots b[] (precomputed on the host with the same operations).
// n >= 4.  Run by one full wavefront: the two first-order recurrences (forward
// elimination, back substitution) are split into per-lane chunks joined by an
// affine-map scan across the 64 lanes; inside a chunk each value is computed with
// the reference's own expression from its predecessor.
__device__ inline double sdiff(const double* __restrict__ y, int ys, const double* __restrict__ h, int i) {
  return (y[(int64_t)(i + 1) * ys] - y[(int64_t)i * ys]) / h[i];
}

// The loops below take a lane's elements in chunks of SCH whose loads are all issued before
// the chunk's arithmetic: the recurrences are serial per lane, so one memory round trip per
// element would set their latency (the wire kernels run beside HBM-bound field passes,
// where a round trip costs microseconds).  The arithmetic is that of the element loop.
constexpr int SCH = 4;

__device__ void spline_solve_wave(const double* __restrict__ y, int ys, int n, const double* __restrict__ h,
                                  const double* __restrict__ t, const double* __restrict__ b,
                                  const double* __restrict__ e, double* __restrict__ c) {
  const int lane = threadIdx.x & 63;
  const int m = (n + 63) / 64;
  const int lo = min(n, lane * m), hi = min(n, lo + m);
  // right-hand side: interior second differences, FMM end conditions
  // e = {x3-x1, x2-x0, h0^2, x3-x0, x[n-1]-x[n-3], x[n-2]-x[n-4], h[n-2]^2, x[n-1]-x[n-4]}
  for (int i0 = lo; i0 < hi; i0 += SCH) {
    double yy[SCH + 2], hq[SCH + 1];  // y[i0-1 .. i0+SCH], h[i0-1 .. i0+SCH-1]
#pragma unroll
    for (int u = 0; u < SCH + 2; ++u) {
      const int i = i0 - 1 + u;
      yy[u] = (i >= 0 && i < n) ? y[(int64_t)i * ys] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SCH + 1; ++u) {
      const int i = i0 - 1 + u;
      hq[u] = (i >= 0 && i < n - 1) ? h[i] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u)  // sdiff(i) - sdiff(i - 1); the two ends are set below
      if (i0 + u < hi) c[i0 + u] = (yy[u + 2] - yy[u + 1]) / hq[u + 1] - (yy[u + 1] - yy[u]) / hq[u];
  }
  if (lo == 0 && hi > 0) {
    const double c1 = sdiff(y, ys, h, 1) - sdiff(y, ys, h, 0);
    const double c2 = sdiff(y, ys, h, 2) - sdiff(y, ys, h, 1);
    double r = c2 / e[0] - c1 / e[1];
    c[0] = r * e[2] / e[3];
  }
  if (lo <= n - 1 && n - 1 < hi) {
    const double cm2 = sdiff(y, ys, h, n - 2) - sdiff(y, ys, h, n - 3);
    const double cm3 = sdiff(y, ys, h, n - 3) - sdiff(y, ys, h, n - 4);
    double r = cm2 / e[4] - cm3 / e[5];
    c[n - 1] = -r * e[6] / e[7];
  }
  // forward elimination x_i = r_i - t_i x_{i-1} (x_0 = r_0): the lane's affine map first
  double A = 1.0, B = 0.0;
  for (int i0 = lo; i0 < hi; i0 += SCH) {
    double cc[SCH], tt[SCH];
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i0 + u;
      cc[u] = i < hi ? c[i] : 0.0;
      tt[u] = i < hi ? t[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i0 + u;
      if (i < hi) {
        const double a = i == 0 ? 0.0 : -tt[u];
        B = a * B + cc[u];
        A = a * A;
      }
    }
  }
  for (int off = 1; off < 64; off <<= 1) {
    const double Ap = __shfl_up(A, off, 64), Bp = __shfl_up(B, off, 64);
    if (lane >= off) {
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  double x = __shfl_up(B, 1, 64);
  if (lane == 0) x = 0.0;
  for (int i0 = lo; i0 < hi; i0 += SCH) {
    double cc[SCH], tt[SCH];
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i0 + u;
      cc[u] = i < hi ? c[i] : 0.0;
      tt[u] = i < hi ? t[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i0 + u;
      if (i < hi) {
        x = i == 0 ? cc[u] : cc[u] - tt[u] * x;
        cc[u] = x;
      }
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u)
      if (i0 + u < hi) c[i0 + u] = cc[u];
  }
  // back substitution x_i = (x_i - h_i x_{i+1}) / b_i (x_{n-1} = x_{n-1} / b_{n-1})
  A = 1.0;
  B = 0.0;
  for (int i1 = hi - 1; i1 >= lo; i1 -= SCH) {
    double cc[SCH], hh[SCH], bb[SCH];
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i1 - u;
      cc[u] = i >= lo ? c[i] : 0.0;
      hh[u] = i >= lo ? h[i] : 0.0;
      bb[u] = i >= lo ? b[i] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i1 - u;
      if (i >= lo) {
        const double a = i == n - 1 ? 0.0 : -hh[u] / bb[u];
        B = a * B + cc[u] / bb[u];
        A = a * A;
      }
    }
  }
  for (int off = 1; off < 64; off <<= 1) {
    const double Ap = __shfl_down(A, off, 64), Bp = __shfl_down(B, off, 64);
    if (lane + off < 64) {
      B = A * Bp + B;
      A = A * Ap;
    }
  }
  x = __shfl_down(B, 1, 64);
  if (lane == 63) x = 0.0;
  for (int i1 = hi - 1; i1 >= lo; i1 -= SCH) {
    double cc[SCH], hh[SCH], bb[SCH];
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i1 - u;
      cc[u] = i >= lo ? c[i] : 0.0;
      hh[u] = i >= lo ? h[i] : 0.0;
      bb[u] = i >= lo ? b[i] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u) {
      const int i = i1 - u;
      if (i >= lo) {
        x = i == n - 1 ? cc[u] / bb[u] : (cc[u] - hh[u] * x) / bb[u];
        cc[u] = x;
      }
    }
#pragma unroll
    for (int u = 0; u < SCH; ++u)
      if (i1 - u >= lo) c[i1 - u] = cc[u];
  }
}

// Evaluate the spline of (y, c) at interval i with offset dx (seval_dpc, spliner.py:1367-1396).
__device__ inline double spline_eval(const double* y, int ystride, const double* c, const double* h, int n,
                                     int i, double dx) {
  double bi, ci, di;
  const double yi = y[(int64_t)i * ystride];
  if (i < n - 1) {
    const double Di = (y[(int64_t)(i + 1) * ystride] - yi) / h[i];
    bi = Di - h[i] * (c[i + 1] + 2.0 * c[i]);
    di = (c[i + 1] - c[i]) / h[i];
    ci = 3.0 * c[i];
  } else {
    const double Dl = (yi - y[(int64_t)(n - 2) * ystride]) / h[n - 2];
    bi = Dl + h[n - 2] * (c[n - 2] + 2.0 * c[n - 1]);
    di = (c[n - 1] - c[n - 2]) / h[n - 2];
    ci = 3.0 * c[n - 1];
  }
  return yi + dx * (bi + dx * (ci + dx * di));
}

// ---------------------------------- K2: populations, screening (CalcScreenedArrays)
__device__ void prep_body(const pss_wire_desc& d, double* sh) {
  const WS L = ws_layout(d.Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const int Nk = d.Nk;
  double s = 0.0;
  for (int k = threadIdx.x; k < Nk; k += 256) {
    const cplx c = B.C2[(int64_t)k * Nk + k], e = B.D2[(int64_t)k * Nk + k];
    ws[L.ne2 + 2 * k] = c.x; ws[L.ne2 + 2 * k + 1] = c.y;
    ws[L.nh2 + 2 * k] = e.x; ws[L.nh2 + 2 * k + 1] = e.y;
    s += c.x + e.x;
  }
  s = block_sum<256>(s, sh);
  // PSTD3D/coulomb.py:1218-1230
  double n1 = s / 2.0 / d.L;
  n1 = fmin(n1, d.density_max);
  const bool scr = d.flags & PSS_SCREENED;
  for (int q = threadIdx.x; q < Nk; q += 256) {
    double e1 = 1.0;
    if (scr) {
      e1 = 1.0 - d.chie[q] * 2 * log(fabs((d.qe[q] - kPi * n1) / (d.qe[q] + n1)))
               - d.chih[q] * 2 * log(fabs((d.qh[q] - kPi * n1) / (d.qh[q] + n1)));
    }
    ws[L.e1d + q] = e1;
    for (int n = 0; n < 3; ++n) {
      const double v = scr ? d.V0[n * Nk + q] / e1 : d.V0[n * Nk + q];
      ws[L.vc + n * Nk + q] = v;
    }
  }
  if (threadIdx.x == 0) ws[L.red + 0] = n1;
  if (threadIdx.x < 8) ws[L.red + 8 + threadIdx.x] = 0.0;  // max-abs slots of the mixed path
}

// -------------------------------------------------- K1: Prop2QW (rr -> r)
// Two workgroups per wire (blockIdx.y), both taking the wireoff gate (the same sum of the
// same lines, so the latch gets one value): x = 0 solves the splines on rr of the three
// propagation lines and evaluates them on the wire grid r_j (window applied, real part
// kept: Prop2QW takes the real part after windowing, and the spline is real-linear, so Re
// commutes with it); x = 1 runs K2 (populations, screening) beside it.
__global__ void __launch_bounds__(256) k_prop_front(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  __shared__ double sh[16];
  double* ws = d.ws;
  const WS L = ws_layout(d.Nk, d.Nr, d.nrr);
  // wireoff gate (PSTD3D/SBEs.py:5861-5870): max_r sqrt(|Ex|^2+|Ey|^2+|Ez|^2) vs 1e-3 Emax0
  double m = 0.0;
  for (int i = threadIdx.x; i < d.nrr; i += 256) {
    const double a = d.Exx[2 * i], b = d.Exx[2 * i + 1];
    const double c = d.Eyy[2 * i], e = d.Eyy[2 * i + 1];
    const double f = d.Ezz[2 * i], g = d.Ezz[2 * i + 1];
    m = fmax(m, sqrt((a * a + b * b) + (c * c + e * e) + (f * f + g * g)));
  }
  m = block_max<256>(m, sh);
  if (threadIdx.x == 0) {
    if (!(*d.on == 0 && m < d.gate)) *d.on = 1;
  }
  __syncthreads();
  if (!wire_on(d)) return;
  if (blockIdx.x == 1) {
    prep_body(d, sh);
    return;
  }
  const int wv = threadIdx.x >> 6;
  if (wv < 3) {
    const double* y = (wv == 0 ? d.Exx : wv == 1 ? d.Eyy : d.Ezz);
    spline_solve_wave(y, 2, d.nrr, d.sa_h, d.sa_t, d.sa_b, d.sa_end, ws + L.spa + wv * d.nrr);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < d.Nr; j += 256) {
    const int i = d.sa_idx[j];
    double v[3];  // the three evaluations' loads before the stores
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      v[f] = 0.0;
      if (i >= 0) {
        const double* y = (f == 0 ? d.Exx : f == 1 ? d.Eyy : d.Ezz);
        v[f] = spline_eval(y, 2, ws + L.spa + f * d.nrr, d.sa_h, d.nrr, i, d.sa_dx[j]);
      }
    }
    const double wj = d.win[j];
#pragma unroll
    for (int f = 0; f < 3; ++f) ws[L.er + f * d.Nr + j] = v[f] * wj;
  }
}

// Centred DFT (FFTG, PSTD3D/usefulsubs.py:2278-2304) of the three real fields at
// the 2Nk-1 wavenumber indices m = NQ0 + dd that kkp reaches (NQ0 = Nk+1).
__global__ void __launch_bounds__(256) k_prop_dft(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d)) return;
  const int nd = 2 * d.Nk - 1;
  const int od = blockIdx.x;
  if (od >= nd) return;
  __shared__ double sh[6 * 4 + 6];
  const WS L = ws_layout(d.Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const int Nr = d.Nr;
  const int m = (d.Nk + 1) + (od - (d.Nk - 1));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < Nr; j += 256) {
    const int64_t tix = ((int64_t)j * m) % Nr;
    const double wr = d.twid[2 * tix], wi = d.twid[2 * tix + 1];
    const double s = (j & 1) ? -1.0 : 1.0;
    for (int f = 0; f < 3; ++f) {
      const double v = s * ws[L.er + f * Nr + j];
      acc[2 * f] += v * wr;
      acc[2 * f + 1] += v * wi;
    }
  }
  block_sum_n<256, 6>(acc, sh);
  if (threadIdx.x == 0) {
    const double sm = (m & 1) ? -1.0 : 1.0;
    for (int f = 0; f < 3; ++f) {
      ws[L.eq + (f * nd + od) * 2] = -(sm * acc[2 * f]) / Nr;
      ws[L.eq + (f * nd + od) * 2 + 1] = -(sm * acc[2 * f + 1]) / Nr;
    }
  }
}

__device__ void maxabs_body(const pss_wire_desc& d, const int m, const int bx, const int nbx, double* sh);

// ---------------------- K3: diagonal traces (direct Hartree terms) + CalcGammaE/H; the
// workgroups from mx_base on take max |p|, |C|, |D| of mixed-precision wires (64 each)
__global__ void __launch_bounds__(256) k_traces_gamma(const pss_wire_desc* __restrict__ descs, const int mx_base) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d)) return;
  __shared__ double sh[8 * 4 + 8];
  if ((int)blockIdx.x >= mx_base) {
    if (d.precision != PSS_MIXED) return;
    const int r = (int)blockIdx.x - mx_base;
    maxabs_body(d, r >> 6, r & 63, 64, sh);
    return;
  }
  const int Nk = d.Nk, nd = 2 * Nk - 1;
  if ((int)blockIdx.x >= nd + Nk) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  if ((int)blockIdx.x < nd) {
    const int qd = (int)blockIdx.x - (Nk - 1);
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int k0 = max(0, -qd), k1 = min(Nk - 1, Nk - 1 - qd);
    for (int k = k0 + threadIdx.x; k <= k1; k += 256) {
      const cplx c = B.C2[(int64_t)k * Nk + k + qd];   // C[k, k+qd]
      const cplx e = B.D2[(int64_t)(k + qd) * Nk + k]; // D[k+qd, k]
      const cplx f = B.D2[(int64_t)k * Nk + k + qd];   // D[k, k+qd]
      const cplx g = B.C2[(int64_t)(k + qd) * Nk + k]; // C[k+qd, k]
      a[0] += c.x; a[1] += c.y; a[2] += e.x; a[3] += e.y; a[4] += f.x; a[5] += f.y; a[6] += g.x; a[7] += g.y;
    }
    block_sum_n<256, 8>(a, sh);
    if (threadIdx.x == 0)
      for (int q = 0; q < 4; ++q) {
        ws[L.tr + (q * nd + blockIdx.x) * 2] = a[2 * q];
        ws[L.tr + (q * nd + blockIdx.x) * 2 + 1] = a[2 * q + 1];
      }
    return;
  }
  // CalcGammaE/H  PSTD3D/dephasing.py:29-103 with ne = Re(C2_kk), nh = Re(D2_kk)
  const int k = (int)blockIdx.x - nd;
  const double* vc = ws + L.vc;
  const double* ne = ws + L.ne2;  // interleaved complex; real part at 2*i
  const double* nh = ws + L.nh2;
  const int NN = Nk * Nk;
  const int32_t* kpq = d.ktab + 0 * NN;
  const int32_t* kmq = d.ktab + 1 * NN;
  const int32_t* k1p = d.ktab + 2 * NN;
  const int32_t* k1pmq = d.ktab + 3 * NN;
  const int32_t* k1t = d.ktab + 4 * NN;
  const int32_t* k1mq = d.ktab + 5 * NN;
  const double f = kPi / kHbar;
  double ge = 0.0, gh = 0.0;
  if (d.flags & PSS_DIAGDPH) {
    for (int q = threadIdx.x; q < Nk; q += 256) {
      const int iq = d.iqv[q];
      const double veh = vc[0 * Nk + iq], vee = vc[1 * Nk + iq], vhh = vc[2 * Nk + iq];
      const double Veh2 = veh * veh, Vee2 = vee * vee, Vhh2 = vhh * vhh;
      const double axe = fabs(d.xe[q]), axh = fabs(d.xh[q]);
      const int64_t kq = (int64_t)k * Nk + q;
      int a = kpq[kq];
      if (a >= 0 && a < Nk) {
        const double na = ne[2 * a], ha = nh[2 * a];
        ge += f * Vee2 * na * (1.0 - na) * axe;
        gh += f * Vhh2 * ha * (1.0 - ha) * axh;
      }
      const int c = kmq[kq];
      a = k1pmq[kq];
      int b = k1p[kq];
      if (a >= 0 && a < Nk && b >= 0 && b < Nk && c >= 0 && c < Nk) {
        const double nha = nh[2 * a], nhb = nh[2 * b], nec = ne[2 * c];
        ge += f * Veh2 * (nha * (1.0 - nhb) * nec + (1.0 - nha) * nhb * (1.0 - nec)) * axh;
      }
      a = k1mq[kq];
      b = k1t[kq];
      if (a >= 0 && a < Nk && b >= 0 && b < Nk && c >= 0 && c < Nk) {
        const double nea = ne[2 * a], neb = ne[2 * b], nhc = nh[2 * c];
        gh += f * Veh2 * (nea * (1.0 - neb) * nhc + (1.0 - nea) * neb * (1.0 - nhc)) * axe;
      }
    }
    double gg[2] = {ge, gh};
    block_sum_n<256, 2>(gg, sh);
    ge = gg[0];
    gh = gg[1];
  } else {
    ge = d.gam_e;
    gh = d.gam_h;
  }
  if (threadIdx.x == 0) {
    ws[L.gam + k] = ge;
    ws[L.gam + Nk + k] = gh;
  }
}

// ------------------------------- K4: off-diagonal dephasing tables De, Dh
// PSTD3D/dephasing.py:227-317 (CalcOffDiagDeph_E2/H2) with n = |n2|:
//   D[q, k] = pi/hbar sum_p V^2 L(Ed) (n_pm (1-n_p)(1-n_k) + (1-n_pm) n_p n_k),
//   Ed = E[k+q] - E[k] + E[p-q] - E[p]   (same species, p - q in [1, Nk-1]),
//   Ed = E[k+q] - E[k] + E'[p+q] - E'[p] (other species E', n'),
// L the Lorentzian of width b.  Per q the p-dependent parts G_p = E[p-q] - E[p],
// a_p = n_pm (1-n_p), b_p = (1-n_pm) n_p are shared by every k, so a workgroup
// owns one q and 256 k, stages (G, a, b) of a p chunk in LDS (broadcast reads)
// and accumulates Sa = sum L(F + G_p) a_p, Sb = sum L(F + G_p) b_p with F =
// E[k+q] - E[k]; D = pi/hbar V^2 ((1-n_k) Sa + n_k Sb).  For the same-species term
// the sums read a per-workgroup table of the line shape over m = k - p + q (see below).
// grid (ceil(Nk/256), 2Nk+1 (q + Nk), 2 nwires (species)), dynamic LDS (256 + max Nk)
// doubles;  electrons -> De.
__device__ inline void offdiag_sum64(const double* __restrict__ G, const double* __restrict__ A,
                                     const double* __restrict__ Bv, int n, double F, double pre, double b2,
                                     double& Sa, double& Sb) {
  for (int j = 0; j < n; ++j) {
    const double Ed = F + G[j];
    const double l = ddiv(pre, Ed * Ed + b2);
    Sa = fma(l, A[j], Sa);
    Sb = fma(l, Bv[j], Sb);
  }
}

// mixed precision: P[j] = (G_j / b, a_j, b_j), z = F / b + G_j / b and 1 / (1 + z^2)
// in FP32, partial sums of 64 terms in FP32 folded into FP64; the caller applies
// the factor 1 / (pi b) of (b/pi)/(E^2+b^2) = (1/(pi b)) / (1 + (E/b)^2).
__device__ inline void offdiag_sum32(const float4* __restrict__ P, int n, float Fs, double& Sa, double& Sb) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j1 = min(n, j0 + 64);
    float sa = 0.f, sb = 0.f;
    for (int j = j0; j < j1; ++j) {
      const float4 v = P[j];
      const float z = Fs + v.x;
      // the hardware reciprocal (v_rcp_f32, 1 ulp): one instruction instead of the
      // correctly rounded division sequence; z^2 + 1 >= 1, so no denormal or range case
      const float l = __builtin_amdgcn_rcpf(fmaf(z, z, 1.0f));
      sa = fmaf(l, v.y, sa);
      sb = fmaf(l, v.z, sb);
    }
    Sa += (double)sa;
    Sb += (double)sb;
  }
}

// table forms: T[-j] is the line shape of partner j (T points at the entry of j = 0)
__device__ inline void offdiag_tab64(const double* __restrict__ T, const double* __restrict__ A,
                                     const double* __restrict__ Bv, int n, double& Sa, double& Sb) {
  for (int j = 0; j < n; ++j) {
    const double l = T[-j];
    Sa = fma(l, A[j], Sa);
    Sb = fma(l, Bv[j], Sb);
  }
}

__device__ inline void offdiag_tab32(const float* __restrict__ T, const float4* __restrict__ P, int n, double& Sa,
                                     double& Sb) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j1 = min(n, j0 + 64);
    float sa = 0.f, sb = 0.f;
    for (int j = j0; j < j1; ++j) {
      const float l = T[-j];
      const float4 v = P[j];
      sa = fmaf(l, v.y, sa);
      sb = fmaf(l, v.z, sb);
    }
    Sa += (double)sa;
    Sb += (double)sb;
  }
}

__global__ void __launch_bounds__(256) k_offdiag(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.z >> 1];
  if (!wire_on(d) || !(d.flags & PSS_OFFDIAGDPH)) return;
  const int Nk = d.Nk;
  const int q = (int)blockIdx.y - Nk;
  if (q > Nk) return;
  if (blockIdx.x * 256 >= Nk) return;
  const bool holes = blockIdx.z & 1;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  // outputs with k + q outside [0, Nk) are zero; a block with none inside only writes zeros
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int klo = max(0, -q), khi = min(Nk, Nk - q);  // valid k in [klo, khi)
  const int kb = blockIdx.x * 256;
  if (kb + 256 <= klo || kb >= khi) {
    if (k < Nk) ws[(holes ? L.Dh : L.De) + (int64_t)(q + Nk) * Nk + k] = 0.0;
    return;
  }
  const double* vc = ws + L.vc;
  const double* nA = ws + (holes ? L.nh2 : L.ne2);
  const double* nB = ws + (holes ? L.ne2 : L.nh2);
  const double* EA = holes ? d.Eh : d.Ee;
  const double* EB = holes ? d.Ee : d.Eh;
  constexpr int CH = 512;
  __shared__ double sG[CH], sa[CH], sb[CH];
  __shared__ float4 sP[CH];
  extern __shared__ double sT[];  // 256 + Nk - 1 line-shape values (FP64, or FP32 in mixed)
  float* sTf = (float*)sT;
  // g = [gam_eh, gam_e, gam_h]; E2 gets (g0, g2), H2 gets (g1, g2)  (dephasing.py:735-736)
  const double bAA = kHbar * (holes ? d.gam_e : d.gam_eh);
  const double bAB = kHbar * d.gam_h;
  const bool fast = d.precision == PSS_MIXED;
  const int aq = q < 0 ? -q : q;
  double vAA = 0.0, vAB = 0.0;
  if (aq >= 1 && aq <= Nk - 1) {
    const double x = vc[(holes ? 2 : 1) * Nk + aq], y = vc[aq];
    vAA = x * x;
    vAB = y * y;
  }
  const int kq = k + q;
  const bool valid = k >= klo && k < khi;
  const double F = valid ? EA[kq] - EA[k] : 0.0;
  double acc = 0.0;
  for (int term = 0; term < 2; ++term) {
    const double v2 = term == 0 ? vAA : vAB;
    if (v2 == 0.0) continue;  // uniform over the block
    const double bw = term == 0 ? bAA : bAB;
    const double pre = bw / kPi, b2 = bw * bw, ib = 1.0 / bw;
    const double* E = term == 0 ? EA : EB;
    const double* nn = term == 0 ? nA : nB;
    // same species: partner p - q in [1, Nk-1]; other species: p + q in [0, Nk-1]
    const int s = term == 0 ? -q : q;
    const int p0 = term == 0 ? max(q + 1, 0) : max(-q, 0);
    const int p1 = term == 0 ? min(q + Nk - 1, Nk - 1) : min(Nk - 1 - q, Nk - 1);
    double Sa = 0.0, Sb = 0.0;
    // same species: the bands are parabolic on a uniform grid (E[j] = c kr[j]^2, kr[j] =
    // dkr (j - j0)), so F_k + G_p = D (k - p + q) with D = 2 c dkr^2 q and the line shape
    // depends on m = k - p + q only: tabulated once per workgroup for the m it meets,
    // m in [mlo, mlo + nm)
    const bool tab = term == 0;
    const int mlo = kb - p1 + q, nm = 256 + (p1 - p0);
    if (tab) {
      const double D = 2.0 * (E[0] / (d.kr[0] * d.kr[0])) * d.dkr * d.dkr * (double)q;
      __syncthreads();
      for (int i = threadIdx.x; i < nm; i += 256) {
        const double Ed = D * (double)(mlo + i);
        if (fast) {
          const double z = Ed * ib;
          sTf[i] = (float)(1.0 / (1.0 + z * z));
        } else {
          sT[i] = ddiv(pre, Ed * Ed + b2);
        }
      }
    }
    for (int c0 = p0; c0 <= p1; c0 += CH) {
      const int n = min(CH, p1 - c0 + 1);
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += 256) {
        const int p = c0 + j, pr = p + s;
        const double np_ = cabs(mk(nn[2 * p], nn[2 * p + 1]));
        const double npr = cabs(mk(nn[2 * pr], nn[2 * pr + 1]));
        const double G = E[pr] - E[p], av = npr * (1.0 - np_), bv = (1.0 - npr) * np_;
        if (fast) {
          sP[j] = make_float4((float)(G * ib), (float)av, (float)bv, 0.f);
        } else {
          sG[j] = G;
          sa[j] = av;
          sb[j] = bv;
        }
      }
      __syncthreads();
      if (valid) {
        if (tab) {
          const int base = k - c0 + q - mlo;  // table index of p = c0 + j is base - j
          if (fast) offdiag_tab32(sTf + base, sP, n, Sa, Sb);
          else offdiag_tab64(sT + base, sa, sb, n, Sa, Sb);
        } else if (fast) {
          offdiag_sum32(sP, n, (float)(F * ib), Sa, Sb);
        } else {
          offdiag_sum64(sG, sa, sb, n, F, pre, b2, Sa, Sb);
        }
      }
    }
    if (valid) {
      const double nk = cabs(mk(nA[2 * k], nA[2 * k + 1]));
      const double f = fast ? 1.0 / (kPi * bw) : 1.0;
      acc += v2 * f * ((1.0 - nk) * Sa + nk * Sb);
    }
  }
  if (k < Nk) ws[(holes ? L.Dh : L.De) + (int64_t)(q + Nk) * Nk + k] = acc * kPi / kHbar;
}

typedef double v4d __attribute__((ext_vector_type(4)));

// ------------------------------------------------- K5: Hamiltonians (CalcH)
// PSTD3D/SBEs.py:820-935 and _calc_Meh :5165-5175.  Exchange terms are Toeplitz
// products along diagonals; direct terms use the diagonal traces of K3.
__global__ void __launch_bounds__(256) k_calc_h(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.z];
  if (!wire_on(d)) return;
  const int Nk = d.Nk;
  const int k1 = blockIdx.x * 16 + (threadIdx.x & 15);
  const int k2 = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (k1 >= Nk || k2 >= Nk) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const double* vc = ws + L.vc;
  const int nd = 2 * Nk - 1;
  const int64_t o = (int64_t)k1 * Nk + k2;
  cplx heh = mk(0, 0), hee = mk(0, 0), hhh = mk(0, 0);
  if (d.flags & PSS_OPTICS) {
    // Meh[ke=k1, kh=k2] = -ehint (Xcv Ex[q] + Ycv Ey[q] yw + Zcv Ez[q] (-1)^w), q = kkp
    const int od = (k1 - k2) + (Nk - 1);
    const cplx ex = *(const cplx*)(ws + L.eq + (0 * nd + od) * 2);
    const cplx ey = *(const cplx*)(ws + L.eq + (1 * nd + od) * 2);
    const cplx ez = *(const cplx*)(ws + L.eq + (2 * nd + od) * 2);
    const double xcv = d.dcv * (((k2 + 1) & 1) ? -1.0 : 1.0);
    heh = (-d.ehint * xcv) * ex - (d.ehint * d.dcv) * ey * d.yw - (d.ehint * -d.dcv) * ez * d.zsign;
  }
  if (k1 == k2) {
    hee = mk(d.Ee[k1] + d.gap, 0.0);
    hhh = mk(d.Eh[k1], 0.0);
  }
  if (d.flags & PSS_EXCITONS) {
    const bool lf = d.flags & PSS_LF;
    // the exchange sums sum_{q != 0} V(|q|) X^T[k1+q, k2+q] of Heh (and of Hee, Hhh with
    // LF) are added by k_exch_fft; without LF Hee / Hhh keep only their diagonal, summed here
    if (!lf && k1 == k2) {
      cplx b = mk(0, 0), c = mk(0, 0);
      for (int q = -k1; q <= Nk - 1 - k1; ++q) {
        if (q == 0) continue;
        const int aq = q < 0 ? -q : q;
        const int64_t t = (int64_t)(k2 + q) * Nk + (k1 + q);
        b = b + vc[Nk + aq] * B.C2[t];
        c = c + vc[2 * Nk + aq] * B.D2[t];
      }
      hee = hee - b;
      hhh = hhh - c;
    }
    if (lf && k1 != k2) {
      const int qd = k1 - k2;
      const int aq = qd < 0 ? -qd : qd;
      const int od = qd + (Nk - 1);
      const cplx tC = *(const cplx*)(ws + L.tr + (0 * nd + od) * 2);
      const cplx tD = *(const cplx*)(ws + L.tr + (1 * nd + od) * 2);
      const cplx tD2 = *(const cplx*)(ws + L.tr + (2 * nd + od) * 2);
      const cplx tC2 = *(const cplx*)(ws + L.tr + (3 * nd + od) * 2);
      hee = hee + (vc[Nk + aq] * tC - vc[aq] * tD);
      hhh = hhh + (vc[2 * Nk + aq] * tD2 - vc[aq] * tC2);
    }
  }
  cplx* H = (cplx*)(ws + L.H);
  const int64_t NN = (int64_t)Nk * Nk;
  H[o] = heh;
  H[NN + o] = hee;
  H[2 * NN + o] = hhh;
}

__host__ __device__ inline int np64(int Nk) { return (Nk + 63) / 64 * 64; }

__host__ __device__ inline int ksplit_max(int Nk) {
  const int np = np64(Nk);
  return np <= 128 ? 16 : np <= 256 ? 8 : np <= 512 ? 4 : 1;
}

// Split-K factor of one wire's contractions, a function of that wire alone (so a wire's
// summation order, and its result, do not depend on which wires share the launch):
// knob > 0 forces it (pss_sbe_tune), else enough workgroups to fill the chip were the wire
// alone (3 t64^2 ks >= 1024); capped by its partial buffer (ksplit_max), by cap > 0 (a step
// that runs beside other work) and by the launch's slice count KS.  Non-increasing in Nk.
__host__ __device__ inline int wire_ks(int Nk, int knob, int cap, int KS) {
  int ks = knob;
  if (ks <= 0) {
    const int t64 = np64(Nk) / 64;
    const int64_t tiles = (int64_t)t64 * t64 * 3;
    ks = 1;
    while (ks < ksplit_max(Nk) && tiles * ks < 1024) ks *= 2;
  }
  ks = ks < ksplit_max(Nk) ? ks : ksplit_max(Nk);
  if (cap > 0 && ks > cap) ks = cap;
  return ks < KS ? ks : KS;
}

// ------------------------------- K5b: exchange sums as FFT convolutions along diagonals
// For a matrix X (p, C or D) and potential row V, the exchange term of CalcH
// (PSTD3D/SBEs.py:820-907) is  E[k1][k2] = sum_{q != 0} V(|q|) X[k2+q][k1+q].
// Along a diagonal d = k1 - k2 it is a symmetric Toeplitz product: with
// x_d[s] = X[r][r+d], r = s + max(0, -d), s in [0, n), n = Nk - |d|,
//   E[t + max(0, d)][t + max(0, -d)] = sum_{s != t} V(|s - t|) x_d[s],
// the linear convolution of x_d with c[j] = V(|j|) (0 < |j| < Nk, c[0] = 0).  On a
// line of M >= 2 Nk - 1 points (a power of two) with x_d zero-padded and c wrapped
// (c[M - j] = c[j]) the circular convolution equals it on [0, n), so
//   E_d = IFFT(FFT(c) FFT(x_d)),  FFT(c) real (c is real and even):
// O(M log M) per diagonal instead of O(Nk^2), the whole exchange O(Nk^2 log Nk).
// A workgroup holds LN lines of M complex points in LDS: line 0 the kernel c, lines
// 1 .. LN-1 consecutive diagonals.  The transforms are in-place Stockham autosort
// steps of radix 8 (the last step radix 4 or 2; Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC 2008), each thread holding 8
// points per step; the inverse is conj(FFT(conj(.))) / M.  X is read and H updated row
// by row over the block's diagonals (contiguous elements).  m = 0: Heh += E (p, V_eh);
// with LF m = 1: Hee -= E (C, V_ee), m = 2: Hhh -= E (D, V_hh).  FP64 in both
// precisions.  Grid (ceil((2 Nk - 1) / (LN - 1)), 3, nwires), LN M / 8 threads for the
// launch's largest M, dynamic LDS (LN (9 M / 8 + 1) + exch_tw(M)) complex values (at most
// 153 KiB).
__host__ __device__ inline int exch_m(int Nk) {
  int M = 8;
  while (M < 2 * Nk - 1) M <<= 1;
  return M;
}
__host__ __device__ inline int exch_lines(int M) { return M <= 128 ? 16 : M <= 256 ? 8 : M <= 2048 ? 4 : 2; }
// Lines per workgroup for a wire of line size M in a launch sized for Mmax with LN lines:
// the launch's threads (LN Mmax / 8) all busy, at most 64 lines.
__host__ __device__ inline int exch_lines_w(int LN, int Mmax, int M) {
  const int n = LN * (Mmax / M);
  return n < 64 ? n : 64;
}
// Twiddles: a step with span Ns = 8^a > 1 and radix R reads w = W_{Ns R}^k, k < Ns, from
// its own table at offset (Ns - 8) / 7 (consecutive butterflies read consecutive entries);
// the tables of all steps hold exch_tw(M) entries.
__host__ __device__ inline int exch_tw(int M) {
  int n = 0;
  for (int Ns = 8; Ns < M; Ns *= 8) n += Ns;
  return n;
}

// forward DFTs of 2, 4 and 8 points in registers (W = exp(-2 pi i / R))
__device__ __forceinline__ void dft2(cplx* v) {
  const cplx a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}
__device__ __forceinline__ void dft4(cplx& a0, cplx& a1, cplx& a2, cplx& a3) {
  const cplx t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, u = a1 - a3;
  const cplx t3 = mk(u.y, -u.x);  // (a1 - a3) (-i)
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}
__device__ __forceinline__ void dft8(cplx* v) {
  constexpr double h = 0.70710678118654752440;  // 1 / sqrt(2)
  cplx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cplx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  const cplx w1 = mk(h * (o1.x + o1.y), h * (o1.y - o1.x));   // W8 o1
  const cplx w2 = mk(o2.y, -o2.x);                             // W8^2 o2 = -i o2
  const cplx w3 = mk(h * (o3.y - o3.x), -h * (o3.x + o3.y));  // W8^3 o3
  v[0] = e0 + o0; v[4] = e0 - o0;
  v[1] = e1 + w1; v[5] = e1 - w1;
  v[2] = e2 + w2; v[6] = e2 - w2;
  v[3] = e3 + w3; v[7] = e3 - w3;
}

// LDS position of point i of a line: one pad slot per 8 points, so that the stride-R
// stores of a step spread over the banks
__device__ __forceinline__ int xpad(int i) { return i + (i >> 3); }

// LDS points are 16-byte aligned (ds_read_b128 / ds_write_b128: one bank group per lane)
struct __align__(16) lcplx {
  double x, y;
};
__device__ __forceinline__ cplx lget(const lcplx& a) { return mk(a.x, a.y); }
__device__ __forceinline__ void lput(lcplx& a, cplx v) { a = lcplx{v.x, v.y}; }

// One Stockham step of radix R on a line of M = 2^lm points: butterfly j in [0, M/R)
// takes x[j + r M/R], multiplies by w^r, w = W_M^{(j mod Ns) M/(Ns R)}, transforms, and
// stores at (j / Ns) Ns R + (j mod Ns) + r Ns.  Thread t < M/8 owns butterflies
// t + u M/8, u < 8/R.  One table read per butterfly; w^2 .. w^7 by products.
template <int R>
__device__ __forceinline__ void st_load(const lcplx* ln, int M, int t, cplx* v) {
#pragma unroll
  for (int u = 0; u < 8 / R; ++u)
#pragma unroll
    for (int r = 0; r < R; ++r) v[u * R + r] = lget(ln[xpad(t + u * (M / 8) + r * (M / R))]);
}
template <int R>
__device__ __forceinline__ void st_store(lcplx* ln, int lm, int Ns, int t, cplx* v, const lcplx* tq) {
  constexpr int LR = R == 8 ? 3 : R == 4 ? 2 : 1;
#pragma unroll
  for (int u = 0; u < 8 / R; ++u) {
    const int j = t + u * (1 << (lm - 3)), k = j & (Ns - 1);
    cplx* w = v + u * R;
    if (Ns > 1) {
      const cplx w1 = lget(tq[(Ns - 8) / 7 + k]);  // W_{Ns R}^k
      w[1] = w[1] * w1;
      if (R >= 4) {
        const cplx w2 = w1 * w1, w3 = w2 * w1;
        w[2] = w[2] * w2;
        w[3] = w[3] * w3;
        if (R == 8) {
          const cplx w4 = w2 * w2;
          w[4] = w[4] * w4;
          w[5] = w[5] * (w4 * w1);
          w[6] = w[6] * (w3 * w3);
          w[7] = w[7] * (w4 * w3);
        }
      }
    }
    if (R == 8) dft8(w);
    else if (R == 4) dft4(w[0], w[1], w[2], w[3]);
    else dft2(w);
    const int base = ((j - k) << LR) + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lput(ln[xpad(base + r * Ns)], w[r]);
  }
}

// Ordering of a step's LDS reads before its writes (and of the writes before the next
// step's reads): a block barrier, or, when each line lies in one wavefront (M <= 512), a
// wavefront-scope fence (the wave's LDS operations complete in order).
template <bool WAVE>
__device__ __forceinline__ void lines_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// forward transform of the lines [l0, LN) in place (line l at sm + l MP); every thread of
// the block calls it
template <bool WAVE>
__device__ __forceinline__ void lines_fft(lcplx* sm, const lcplx* tq, int lm, int MP, int LN, int l0) {
  const int M = 1 << lm;
  const int l = (int)threadIdx.x >> (lm - 3), t = (int)threadIdx.x & ((M >> 3) - 1);
  const bool act = l >= l0 && l < LN;
  lcplx* ln = sm + (int64_t)l * MP;
  cplx v[8];
  for (int lns = 0; lns < lm;) {
    const int rem = lm - lns;
    const int R = rem >= 3 ? 8 : rem == 2 ? 4 : 2;
    const int Ns = 1 << lns;
    if (act) {
      if (R == 8) st_load<8>(ln, M, t, v);
      else if (R == 4) st_load<4>(ln, M, t, v);
      else st_load<2>(ln, M, t, v);
    }
    lines_sync<WAVE>();
    if (act) {
      if (R == 8) st_store<8>(ln, lm, Ns, t, v, tq);
      else if (R == 4) st_store<4>(ln, lm, Ns, t, v, tq);
      else st_store<2>(ln, lm, Ns, t, v, tq);
    }
    lines_sync<WAVE>();
    lns += R == 8 ? 3 : R == 4 ? 2 : 1;
  }
}

__global__ void __launch_bounds__(1024) k_exch_fft(const pss_wire_desc* __restrict__ descs, int LN0, int Mmax) {
  const int wire = blockIdx.z, m = blockIdx.y;
  const pss_wire_desc& d = descs[wire];
  if (!wire_on(d) || !(d.flags & PSS_EXCITONS)) return;
  if (m > 0 && !(d.flags & PSS_LF)) return;
  const int Nk = d.Nk, nd = 2 * Nk - 1, M = exch_m(Nk), LN = exch_lines_w(LN0, Mmax, M), D = LN - 1;
  const int lm = __builtin_ctz(M), MP = M + (M >> 3) + 1;  // odd stride: lines start on different banks
  const int od0 = blockIdx.x * D;
  if (od0 >= nd) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  const double* ws = d.ws;
  const Bufs B = bufs(d);
  const cplx* X = m == 0 ? B.p2 : m == 1 ? B.C2 : B.D2;
  const double* V = ws + L.vc + (int64_t)m * Nk;
  extern __shared__ lcplx sm[];  // LN lines of MP, then the steps' twiddle tables
  lcplx* tq = sm + (int64_t)LN * MP;
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int Ns = 8; Ns < M; Ns *= 8) {
    const int R = M / Ns >= 8 ? 8 : M / Ns;
    for (int k = tid; k < Ns; k += nt) {
      double sn, cs;
      sincospi(2.0 * k / (Ns * R), &sn, &cs);
      lput(tq[(Ns - 8) / 7 + k], mk(cs, -sn));
    }
  }
  // line 0: c; lines 1..D: zero, then the diagonals' elements row by row
  for (int e = tid; e < LN * M; e += nt) {
    const int l = e >> lm, s = e & (M - 1);
    cplx v = mk(0, 0);
    if (l == 0) {
      const int j = s <= M / 2 ? s : M - s;
      if (j >= 1 && j < Nk) v = mk(V[j], 0.0);
    }
    lput(sm[(int64_t)l * MP + xpad(s)], v);
  }
  __syncthreads();
  // thread -> (diagonal l, rows r0, r0 + rs, ...): consecutive threads read consecutive
  // columns of one row
  const int ld = tid % D, r0 = tid / D, rs = nt / D;
  const int od = od0 + ld, dg = od - (Nk - 1);
  const bool lon = r0 < rs && od < nd;
  lcplx* dl = sm + (int64_t)(ld + 1) * MP;
  // four rows' loads in flight together, then their LDS stores
  constexpr int U = 4;
  if (lon)
    for (int rb = r0; rb < Nk; rb += U * rs) {
      cplx v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = rb + u * rs, col = r + dg;
        v[u] = (r < Nk && col >= 0 && col < Nk) ? X[(int64_t)r * Nk + col] : mk(0, 0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = rb + u * rs, col = r + dg;
        if (r < Nk && col >= 0 && col < Nk) lput(dl[xpad(r < col ? r : col)], v[u]);
      }
    }
  __syncthreads();
  if (M <= 512) lines_fft<true>(sm, tq, lm, MP, LN, 0);
  else lines_fft<false>(sm, tq, lm, MP, LN, 0);
  __syncthreads();
  // conj(x^ c^) / M: the inverse as a forward transform of the conjugate
  const double iM = 1.0 / M;
  for (int e = M + tid; e < LN * M; e += nt) {
    const int l = e >> lm, s = e & (M - 1);
    const double c = sm[xpad(s)].x * iM;
    lcplx& x = sm[(int64_t)l * MP + xpad(s)];
    x = lcplx{x.x * c, -x.y * c};
  }
  __syncthreads();
  if (M <= 512) lines_fft<true>(sm, tq, lm, MP, LN, 1);
  else lines_fft<false>(sm, tq, lm, MP, LN, 1);
  __syncthreads();
  cplx* H = (cplx*)(d.ws + L.H) + (int64_t)m * Nk * Nk;
  const double sg = m == 0 ? 1.0 : -1.0;
  if (lon)
    for (int kb = r0; kb < Nk; kb += U * rs) {
      cplx hv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k1 = kb + u * rs, k2 = k1 - dg;
        hv[u] = (k1 < Nk && k2 >= 0 && k2 < Nk) ? H[(int64_t)k1 * Nk + k2] : mk(0, 0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k1 = kb + u * rs, k2 = k1 - dg;
        if (k1 < Nk && k2 >= 0 && k2 < Nk) {
          const cplx y = lget(dl[xpad(k1 < k2 ? k1 : k2)]);
          H[(int64_t)k1 * Nk + k2] = hv[u] + mk(sg * y.x, -sg * y.y);  // conj of the forward transform
        }
      }
    }
}

// ------------------------------------------- K6: RHS GEMMs + leapfrog epilogue
// Output X in {p, C, D} = sum of complex products op(A) op(B) (dpdt/dCdt/dDdt,
// PSTD3D/SBEs.py:1953-1976, 2320-2338, 2618-2636), then
//   X3 = X1 + 2 dt R, R = S/(i hbar) (+ elementwise terms for p),
//   next <- X3, prev <- (X1 + X3)/2   (PSTD3D/SBEs.py:5685-5718).
// The off-diagonal dephasing OffP = i hbar (Mh p + p Me^T) (dephasing.py:321-356)
// enters as two more products with operands pre-scaled by i*hbar.
// Tile 32x32 per wave (64 threads), v_mfma_f64_16x16x4f64, K block 16.
enum Op { OP_N = 0, OP_T = 1, OP_C = 2, OP_H = 3, OP_MH = 4, OP_MET = 5 };
enum Src { S_HHH, S_HEE, S_HEH, S_C, S_D, S_P, S_NONE };

struct Term {
  int8_t a, aop, b, bop;
  int8_t neg;
};

// terms per output: p: 6, C: 4, D: 4
__constant__ Term kTerms[3][6] = {
    {{S_HHH, OP_N, S_P, OP_N, 0}, {S_P, OP_N, S_HEE, OP_T, 0}, {S_HEH, OP_T, S_C, OP_N, 1},
     {S_D, OP_T, S_HEH, OP_T, 1}, {S_NONE, OP_MH, S_P, OP_N, 0}, {S_P, OP_N, S_NONE, OP_MET, 0}},
    {{S_C, OP_N, S_HEE, OP_T, 0}, {S_HEE, OP_T, S_C, OP_N, 1}, {S_P, OP_H, S_HEH, OP_T, 0},
     {S_HEH, OP_C, S_P, OP_N, 1}, {S_NONE, 0, S_NONE, 0, 0}, {S_NONE, 0, S_NONE, 0, 0}},
    {{S_D, OP_N, S_HHH, OP_T, 0}, {S_HHH, OP_T, S_D, OP_N, 1}, {S_P, OP_C, S_HEH, OP_N, 0},
     {S_HEH, OP_H, S_P, OP_T, 1}, {S_NONE, 0, S_NONE, 0, 0}, {S_NONE, 0, S_NONE, 0, 0}},
};


__device__ inline const cplx* src_ptr(int s, const Bufs& B, const cplx* H, int64_t NN) {
  switch (s) {
    case S_HEH: return H;
    case S_HEE: return H + NN;
    case S_HHH: return H + 2 * NN;
    case S_C: return B.C2;
    case S_D: return B.D2;
    default: return B.p2;
  }
}

// element (i, k) of op(X) where X is Nk x Nk; for OP_MH / OP_MET the real
// matrices built from the dephasing tables, times i*hbar.
__device__ inline cplx load_op(const cplx* X, int op, int i, int k, int Nk, const double* Dh,
                               const double* De) {
  if (i >= Nk || k >= Nk) return mk(0, 0);
  switch (op) {
    case OP_N: return X[(int64_t)i * Nk + k];
    case OP_T: return X[(int64_t)k * Nk + i];
    case OP_C: return conj(X[(int64_t)i * Nk + k]);
    case OP_H: return conj(X[(int64_t)k * Nk + i]);
    case OP_MH: {  // Mh[kh=i, k'=k] = Dh[k'-kh+Nk, kh]
      if (i == k) return mk(0, 0);
      return mk(0.0, kHbar * Dh[(int64_t)(k - i + Nk) * Nk + i]);
    }
    default: {  // OP_MET: Me^T[k'=i, ke=k] = Me[ke, k'] = De[k'-ke+Nk, ke]
      if (i == k) return mk(0, 0);
      return mk(0.0, kHbar * De[(int64_t)(i - k + Nk) * Nk + k]);
    }
  }
}

// FP64 leapfrog epilogue of one element with the products' sum S (the Hee / Hhh
// diagonals are inside the products): + Heh^T - i hbar gamma p for dp.
// Whether the RHS epilogue writes the averaged prev entry (X1 + X3) / 2: not on the C, D
// diagonals (written after the relaxation), and with DCTrans not where the transport shift
// changes X3 after the epilogue (every entry with PSS_LF, else the diagonal of p).
__device__ inline bool avg_in_epilogue(const pss_wire_desc& d, int out, int i, int j) {
  if (i == j && out != 0) return false;
  if (d.flags & PSS_DCTRANS) return !(d.flags & PSS_LF) && i != j;
  return true;
}

__device__ inline void fp64_epilogue(const pss_wire_desc& d, const Bufs& B, const cplx* H, const double* gam,
                                     int out, int i, int j, cplx S, int deriv) {
  const int Nk = d.Nk;
  const int64_t o = (int64_t)i * Nk + j;
  const cplx* X1 = out == 0 ? B.p1 : out == 1 ? B.C1 : B.D1;
  cplx* Xp = out == 0 ? B.pp : out == 1 ? B.Cp : B.Dp;
  cplx* Xn = out == 0 ? B.pn : out == 1 ? B.Cn : B.Dn;
  if (out == 0) {
    // + Heh^T[kh=i, ke=j] - i hbar (GamE[ke] + GamH[kh]) p[kh, ke]
    const cplx hT = H[(int64_t)j * Nk + i];
    const cplx pv = B.p2[o];
    const double g = kHbar * (gam[j] + gam[Nk + i]);
    S = S + hT - mk(-g * pv.y, g * pv.x);
  }
  const cplx R = div_ih(S, kHbar);
  if (deriv) {  // kernel-level entry: dX/dt into the next buffer, nothing integrated
    Xn[o] = R;
    return;
  }
  const cplx x1 = X1[o];
  const cplx x3 = x1 + (R * d.dt) * 2.0;
  Xn[o] = x3;
  if (avg_in_epilogue(d, out, i, j)) Xp[o] = (x1 + x3) * 0.5;
}

// One 64 x 64 output tile per workgroup, 4 waves (32 x 32 quadrants, 2 x 2 blocks of
// 16 x 16, v_mfma_f64_16x16x4f64); A (64 x 16) and B (16 x 64) complex operands of one
// k chunk staged in LDS, shared by the waves; coalesced loads along the contiguous
// index of op(X).  The (term, chunk) list is cut into `ks` slices (blockIdx.z); with
// ks == 1 the epilogue runs here, else FP64 partial tiles go to the pk buffer.
__global__ void __launch_bounds__(256) k_rhs_gemm(const pss_wire_desc* __restrict__ descs, int KS, int knob,
                                                   int cap, int deriv) {
  const int zz = blockIdx.z;
  const int slice = zz % KS, out = (zz / KS) % 3, wire = zz / (3 * KS);
  const pss_wire_desc& d = descs[wire];
  if (!wire_on(d) || d.precision == PSS_MIXED) return;
  const int Nk = d.Nk;
  const int ks = wire_ks(Nk, knob, cap, KS);  // this wire's own split-K factor
  if (slice >= ks) return;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  if (m0 >= Nk || n0 >= Nk) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const int64_t NN = (int64_t)Nk * Nk;
  const cplx* H = (const cplx*)(ws + L.H);
  const double* Dh = ws + L.Dh;
  const double* De = ws + L.De;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv >> 1, wc = wv & 1;

  __shared__ double Ar[16][64], Ai[16][64], Br[16][64], Bi[16][64];
  v4d accr[2][2], acci[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = v4d{0, 0, 0, 0};
      acci[a][b] = v4d{0, 0, 0, 0};
    }
  const int nterms = out == 0 ? ((d.flags & PSS_OFFDIAGDPH) ? 6 : 4) : 4;
  const int nkc = (Nk + 15) / 16;
  const int nitems = nterms * nkc;
  const int it0 = (int)((int64_t)nitems * slice / ks), it1 = (int)((int64_t)nitems * (slice + 1) / ks);
  for (int it = it0; it < it1; ++it) {
    const int t = it / nkc, k0 = (it % nkc) * 16;
    const Term T = kTerms[out][t];
    const cplx* Ap = src_ptr(T.a, B, H, NN);
    const cplx* Bp = src_ptr(T.b, B, H, NN);
    const double sg = T.neg ? -1.0 : 1.0;
    const bool aT = (T.aop == OP_T || T.aop == OP_H || T.aop == OP_MET);
    const bool bT = (T.bop == OP_T || T.bop == OP_H || T.bop == OP_MET);
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += 256) {
      int mm, kk;
      if (!aT) { mm = e >> 4; kk = e & 15; } else { kk = e >> 6; mm = e & 63; }
      const cplx v = load_op(Ap, T.aop, m0 + mm, k0 + kk, Nk, Dh, De);
      Ar[kk][mm] = sg * v.x;
      Ai[kk][mm] = sg * v.y;
    }
    for (int e = threadIdx.x; e < 1024; e += 256) {
      int kk, nn;
      if (!bT) { kk = e >> 6; nn = e & 63; } else { nn = e >> 4; kk = e & 15; }
      // op(B)[k, n]: for transposed ops load_op's (i, k) = (k, n) of op(B)
      const cplx v = load_op(Bp, T.bop, k0 + kk, n0 + nn, Nk, Dh, De);
      Br[kk][nn] = v.x;
      Bi[kk][nn] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {
      const int kr = kq * 4 + (lane >> 4);
      const int li = lane & 15;
      double ar[2], ai[2], br[2], bi[2];
      for (int x = 0; x < 2; ++x) {
        ar[x] = Ar[kr][wr * 32 + x * 16 + li];
        ai[x] = Ai[kr][wr * 32 + x * 16 + li];
        br[x] = Br[kr][wc * 32 + x * 16 + li];
        bi[x] = Bi[kr][wc * 32 + x * 16 + li];
      }
      for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
          accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], accr[x][y], 0, 0, 0);
          accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[x], bi[y], accr[x][y], 0, 0, 0);
          acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], acci[x][y], 0, 0, 0);
          acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], br[y], acci[x][y], 0, 0, 0);
        }
    }
  }
  // v_mfma_f64_16x16x4f64 result: lane l, element v holds row (l/16) + 4v, column l%16
  const double* gam = ws + L.gam;
  cplx* pk = (cplx*)(ws + L.pk) + ((int64_t)slice * 3 + out) * NN;
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      for (int v = 0; v < 4; ++v) {
        const int i = m0 + wr * 32 + x * 16 + (lane >> 4) + 4 * v;
        const int j = n0 + wc * 32 + y * 16 + (lane & 15);
        if (i >= Nk || j >= Nk) continue;
        const cplx S = mk(accr[x][y][v], acci[x][y][v]);
        if (ks == 1) fp64_epilogue(d, B, H, gam, out, i, j, S, deriv);
        else pk[(int64_t)i * Nk + j] = S;
      }
}

// ------------------------------- K6m: mixed-precision RHS contractions
// Same products as k_rhs_gemm, with each complex operand split per real
// component into FP16 hi + lo (x s = hi + lo, s a power of two per matrix from
// its max |x|), three MFMA products hi*hi + hi*lo + lo*hi accumulated in FP32,
// folded into FP64 per term.  The diagonals of Hee and Hhh (single-particle
// energies + exchange diagonal, which set the optical phase) are kept out of the
// split products and applied exactly in FP64 in the epilogue, as is the
// leapfrog accumulation.  Max-abs slots: red[8 + m], m = p, C, D, Heh, Hee, Hhh,
// hbar Dh, hbar De.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline void atomic_max_pos(double* addr, double v) {
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

// max |x| of matrix m over the elements bx, bx + nbx*256, ... of one wire (workgroup bx of nbx)
__device__ void maxabs_body(const pss_wire_desc& d, const int m, const int bx, const int nbx, double* sh) {
  const int Nk = d.Nk;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const int64_t NN = (int64_t)Nk * Nk;
  const cplx* H = (const cplx*)(ws + L.H);
  double v = 0.0;
  if (m < 6) {
    const cplx* X = m == 0 ? B.p2 : m == 1 ? B.C2 : m == 2 ? B.D2 : H + (m - 3) * NN;
    for (int64_t o = (int64_t)bx * 256 + threadIdx.x; o < NN; o += (int64_t)nbx * 256) {
      if ((m == 4 || m == 5) && (o / Nk == o % Nk)) continue;  // diagonal handled in FP64
      v = fmax(v, fmax(fabs(X[o].x), fabs(X[o].y)));
    }
  } else if (d.flags & PSS_OFFDIAGDPH) {
    const double* D = ws + (m == 6 ? L.Dh : L.De);
    const int64_t n = (2 * (int64_t)Nk + 1) * Nk;
    for (int64_t o = (int64_t)bx * 256 + threadIdx.x; o < n; o += (int64_t)nbx * 256)
      v = fmax(v, fabs(kHbar * D[o]));
  }
  v = block_max<256>(v, sh);
  if (threadIdx.x == 0 && v > 0.0) atomic_max_pos(ws + L.red + 8 + m, v);
}

__global__ void __launch_bounds__(256) k_maxabs(const pss_wire_desc* __restrict__ descs, int m0) {
  const pss_wire_desc& d = descs[blockIdx.z];
  if (!wire_on(d) || d.precision != PSS_MIXED) return;
  __shared__ double sh[16];
  maxabs_body(d, m0 + (int)blockIdx.y, (int)blockIdx.x, (int)gridDim.x, sh);
}

__device__ inline double pow2_scale(double mx) {
  if (!(mx > 0.0)) return 1.0;
  int e;
  frexp(mx, &e);
  return ldexp(1.0, 14 - e);
}

// matrix index (red slot) of a term operand
__device__ inline int slot_of(int src, int op) {
  if (op == OP_MH) return 6;
  if (op == OP_MET) return 7;
  switch (src) {
    case S_P: return 0;
    case S_C: return 1;
    case S_D: return 2;
    case S_HEH: return 3;
    case S_HEE: return 4;
    default: return 5;  // S_HHH
  }
}

__device__ inline void split16(double x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (double)hi);
}

// ---- operand planes.  k_split writes, for every source matrix (slot m: p, C, D,
// Heh, Hee with zero diagonal, Hhh with zero diagonal, Mh = i hbar Dh-matrix,
// Q = Me^T = i hbar De-matrix), the scaled FP16 hi/lo planes (re hi, re lo, im hi,
// im lo) in two orientations: o = 0 row-major X[i][k], o = 1 transposed X[k][i].
// Planes are NP x NP (NP = Nk rounded up to 64) with zeros outside Nk, so the
// contraction kernel loads whole 16-wide k segments without bounds checks.

__device__ inline _Float16* plane(double* ws, const WS& L, int NP, int m, int o, int q) {
  return (_Float16*)(ws + L.sp) + ((int64_t)((m * 2 + o) * 4 + q)) * NP * NP;
}

// element of source slot m at (i, k), i, k < Nk, as a complex value (before scaling)
__device__ inline cplx slot_value(int m, int i, int k, int Nk, const Bufs& B, const cplx* H, const double* Dh,
                                  const double* De) {
  const int64_t NN = (int64_t)Nk * Nk;
  const int64_t o = (int64_t)i * Nk + k;
  switch (m) {
    case 0: return B.p2[o];
    case 1: return B.C2[o];
    case 2: return B.D2[o];
    case 3: return H[o];
    case 4: return i == k ? mk(0, 0) : H[NN + o];
    case 5: return i == k ? mk(0, 0) : H[2 * NN + o];
    case 6: return i == k ? mk(0, 0) : mk(0.0, kHbar * Dh[(int64_t)(k - i + Nk) * Nk + i]);  // Mh[i][k]
    default: return i == k ? mk(0, 0) : mk(0.0, kHbar * De[(int64_t)(i - k + Nk) * Nk + k]);  // Me^T[i][k]
  }
}

// grid (NP/32, NP/32, 8 * nwires): one 32 x 32 tile of one slot, both orientations
__global__ void __launch_bounds__(256) k_split(const pss_wire_desc* __restrict__ descs) {
  const int wire = blockIdx.z >> 3, m = blockIdx.z & 7;
  const pss_wire_desc& d = descs[wire];
  if (!wire_on(d) || d.precision != PSS_MIXED) return;
  if (m >= 6 && !(d.flags & PSS_OFFDIAGDPH)) return;
  const int Nk = d.Nk, NP = np64(Nk);
  const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  if (k0 >= NP || i0 >= NP) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const cplx* H = (const cplx*)(ws + L.H);
  const double sc = pow2_scale(ws[L.red + 8 + m]);
  __shared__ _Float16 T[4][32][33];
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int ii = e >> 5, kk = e & 31;
    const int i = i0 + ii, k = k0 + kk;
    cplx v = mk(0, 0);
    if (i < Nk && k < Nk) v = slot_value(m, i, k, Nk, B, H, ws + L.Dh, ws + L.De);
    _Float16 h[4];
    split16(v.x * sc, h[0], h[1]);
    split16(v.y * sc, h[2], h[3]);
    for (int q = 0; q < 4; ++q) {
      plane(ws, L, NP, m, 0, q)[(int64_t)i * NP + k] = h[q];
      T[q][kk][ii] = h[q];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int kk = e >> 5, ii = e & 31;
    for (int q = 0; q < 4; ++q) plane(ws, L, NP, m, 1, q)[(int64_t)(k0 + kk) * NP + i0 + ii] = T[q][kk][ii];
  }
}

// operand mapping of a term: A side needs op(A)[i][k] with k contiguous, B side
// op(B)[k][j] with k contiguous for fixed j
__device__ inline void a_side(int src, int op, int& m, int& o, bool& cj) {
  m = slot_of(src, op);
  o = (op == OP_T || op == OP_H) ? 1 : 0;
  cj = (op == OP_C || op == OP_H);
}
__device__ inline void b_side(int src, int op, int& m, int& o, bool& cj) {
  m = slot_of(src, op);
  o = (op == OP_N || op == OP_C || op == OP_MET) ? 1 : 0;
  cj = (op == OP_C || op == OP_H);
}

// leapfrog epilogue of one element with the products' sum S (mixed path): exact
// diagonal parts Hhh[kh,kh] p + p Hee[ke,ke] (+ Heh^T - i hbar gamma p) for dp,
// X (Hd[j,j] - Hd[i,i]) for dC / dD; R = S / (i hbar); X3 = X1 + 2 dt R.
__device__ inline void mixed_epilogue(const pss_wire_desc& d, const Bufs& B, const cplx* H, const double* gam,
                                      int out, int i, int j, cplx S, int deriv) {
  const int Nk = d.Nk;
  const int64_t NN = (int64_t)Nk * Nk;
  const int64_t o = (int64_t)i * Nk + j;
  const cplx* X1 = out == 0 ? B.p1 : out == 1 ? B.C1 : B.D1;
  const cplx* X2 = out == 0 ? B.p2 : out == 1 ? B.C2 : B.D2;
  cplx* Xp = out == 0 ? B.pp : out == 1 ? B.Cp : B.Dp;
  cplx* Xn = out == 0 ? B.pn : out == 1 ? B.Cn : B.Dn;
  const cplx* Hee = H + NN;
  const cplx* Hhh = H + 2 * NN;
  const cplx x2 = X2[o];
  if (out == 0) {
    const cplx hT = H[(int64_t)j * Nk + i];
    const double g = kHbar * (gam[j] + gam[Nk + i]);
    S = S + Hhh[(int64_t)i * Nk + i] * x2 + x2 * Hee[(int64_t)j * Nk + j] + hT - mk(-g * x2.y, g * x2.x);
  } else {
    const cplx* Hd = out == 1 ? Hee : Hhh;
    S = S + x2 * (Hd[(int64_t)j * Nk + j] - Hd[(int64_t)i * Nk + i]);
  }
  const cplx R = div_ih(S, kHbar);
  if (deriv) {  // kernel-level entry: dX/dt into the next buffer, nothing integrated
    Xn[o] = R;
    return;
  }
  const cplx x1 = X1[o];
  const cplx x3 = x1 + (R * d.dt) * 2.0;
  Xn[o] = x3;
  if (avg_in_epilogue(d, out, i, j)) Xp[o] = (x1 + x3) * 0.5;
}

// ---- contraction: one 64 x 64 output tile per workgroup, 4 waves (32 x 32 each,
// 2 x 2 blocks of 16 x 16).  A stage is 32 k values: the FP16 planes of A (64
// rows) and B (64 rows) of the next stage are loaded into registers while the
// current one computes, then staged in LDS and shared by the waves.  12
// v_mfma_f32_16x16x32_f16 per block and stage: re += ar br - ai bi, im += ar bi +
// ai br, each product hi*hi + hi*lo + lo*hi.  FP32 partials are folded into FP64
// (times 1/(sa sb)) every 4 stages (128 k values) and at term ends.  The (term,
// stage) list is cut into `ks` contiguous slices (blockIdx.z); with ks == 1 the
// epilogue runs here, else FP64 partial tiles go to the pk buffer.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
constexpr int kKs = 32;        // k values per stage
constexpr int kLdk = kKs + 8;  // padded stride of a staged row (halves), a 16-byte multiple
constexpr int kFold = 4;       // stages per FP32 partial

__device__ inline f4 mfma32(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

__global__ void __launch_bounds__(256) k_rhs_mx(const pss_wire_desc* __restrict__ descs, int KS, int knob, int cap,
                                                 int deriv) {
  const int zz = blockIdx.z;
  const int slice = zz % KS, out = (zz / KS) % 3, wire = zz / (3 * KS);
  const pss_wire_desc& d = descs[wire];
  if (!wire_on(d) || d.precision != PSS_MIXED) return;
  const int Nk = d.Nk, NP = np64(Nk);
  const int ks = wire_ks(Nk, knob, cap, KS);  // this wire's own split-K factor
  if (slice >= ks) return;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  if (m0 >= NP || n0 >= NP) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const double* mx = ws + L.red + 8;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = wv >> 1, wc = wv & 1;

  __shared__ _Float16 As[4][64][kLdk];
  __shared__ _Float16 Bs[4][64][kLdk];

  double accr[2][2][4], acci[2][2][4];
  f4 fr[2][2], fi[2][2];
  #pragma unroll
  for (int a = 0; a < 2; ++a)
    #pragma unroll
    for (int b = 0; b < 2; ++b) {
      fr[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      fi[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      #pragma unroll
      for (int v = 0; v < 4; ++v) { accr[a][b][v] = 0.0; acci[a][b][v] = 0.0; }
    }
  const int nterms = out == 0 ? ((d.flags & PSS_OFFDIAGDPH) ? 6 : 4) : 4;
  const int nkc = NP / kKs;
  const int nitems = nterms * nkc;
  const int it0 = (int)((int64_t)nitems * slice / ks), it1 = (int)((int64_t)nitems * (slice + 1) / ks);
  int cur_t = -1, nacc = 0;
  double cur_f = 0.0;
  bool acj = false, bcj = false;
  auto fold = [&]() {
    #pragma unroll
    for (int a = 0; a < 2; ++a)
      #pragma unroll
      for (int b = 0; b < 2; ++b)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
          accr[a][b][v] += (double)fr[a][b][v] * cur_f;
          acci[a][b][v] += (double)fi[a][b][v] * cur_f;
          fr[a][b][v] = 0.f;
          fi[a][b][v] = 0.f;
        }
    nacc = 0;
  };
  // loader: a stage is 2 operands x 4 planes x 64 rows x 4 sixteen-byte quarters;
  // thread t loads quarter t & 3 of row t >> 2 of every plane of A and of B, so 4
  // consecutive threads read one row's 64 contiguous bytes
  const int lrow = threadIdx.x >> 2, lqt = threadIdx.x & 3;
  const int64_t PL = (int64_t)NP * NP;
  const _Float16* spl = (const _Float16*)(ws + L.sp);
  // named registers, not an array: the prefetched segments must stay in VGPRs
  uint4 pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3;
  auto prefetch = [&](int it) {
    const int t = it / nkc, k0 = (it - t * nkc) * kKs;
    const Term T = kTerms[out][t];
    int ma, oa, mb, ob;
    bool c;
    a_side(T.a, T.aop, ma, oa, c);
    b_side(T.b, T.bop, mb, ob, c);
    const _Float16* ga = spl + (ma * 2 + oa) * 4 * PL + (int64_t)(m0 + lrow) * NP + k0 + 8 * lqt;
    const _Float16* gb = spl + (mb * 2 + ob) * 4 * PL + (int64_t)(n0 + lrow) * NP + k0 + 8 * lqt;
    pa0 = *(const uint4*)ga;
    pa1 = *(const uint4*)(ga + PL);
    pa2 = *(const uint4*)(ga + 2 * PL);
    pa3 = *(const uint4*)(ga + 3 * PL);
    pb0 = *(const uint4*)gb;
    pb1 = *(const uint4*)(gb + PL);
    pb2 = *(const uint4*)(gb + 2 * PL);
    pb3 = *(const uint4*)(gb + 3 * PL);
  };
  if (it0 < it1) prefetch(it0);
  for (int it = it0; it < it1; ++it) {
    const int t = it / nkc;
    if (t != cur_t) {
      if (cur_t >= 0) fold();
      const Term T = kTerms[out][t];
      int ma, oa, mb, ob;
      a_side(T.a, T.aop, ma, oa, acj);
      b_side(T.b, T.bop, mb, ob, bcj);
      const double sa = pow2_scale(mx[ma]), sb = pow2_scale(mx[mb]);
      cur_f = (T.neg ? -1.0 : 1.0) / (sa * sb);
      cur_t = t;
    }
    *(uint4*)&As[0][lrow][8 * lqt] = pa0;
    *(uint4*)&As[1][lrow][8 * lqt] = pa1;
    *(uint4*)&As[2][lrow][8 * lqt] = pa2;
    *(uint4*)&As[3][lrow][8 * lqt] = pa3;
    *(uint4*)&Bs[0][lrow][8 * lqt] = pb0;
    *(uint4*)&Bs[1][lrow][8 * lqt] = pb1;
    *(uint4*)&Bs[2][lrow][8 * lqt] = pb2;
    *(uint4*)&Bs[3][lrow][8 * lqt] = pb3;
    __syncthreads();
    if (it + 1 < it1) prefetch(it + 1);
    {
      // fragments: lane l holds row l % 16, k values 8 (l / 16) .. + 7 of the stage
      const int li = lane & 15, kq = (lane >> 4) * 8;
      h8 a[2][4], bq[2][4];
      #pragma unroll
      for (int x = 0; x < 2; ++x)
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[x][q] = *(const h8*)&As[q][wr * 32 + x * 16 + li][kq];
          bq[x][q] = *(const h8*)&Bs[q][wc * 32 + x * 16 + li][kq];
        }
      if (acj)
        #pragma unroll
        for (int x = 0; x < 2; ++x) { a[x][2] = -a[x][2]; a[x][3] = -a[x][3]; }
      if (bcj)
        #pragma unroll
        for (int x = 0; x < 2; ++x) { bq[x][2] = -bq[x][2]; bq[x][3] = -bq[x][3]; }
      #pragma unroll
      for (int x = 0; x < 2; ++x) {
        const h8 nih = -a[x][2], nil = -a[x][3];
        #pragma unroll
        for (int y = 0; y < 2; ++y) {
          f4 r = fr[x][y], im = fi[x][y];
          // re: ar br - ai bi
          r = mfma32(a[x][0], bq[y][0], r);
          r = mfma32(a[x][0], bq[y][1], r);
          r = mfma32(a[x][1], bq[y][0], r);
          r = mfma32(nih, bq[y][2], r);
          r = mfma32(nih, bq[y][3], r);
          r = mfma32(nil, bq[y][2], r);
          // im: ar bi + ai br
          im = mfma32(a[x][0], bq[y][2], im);
          im = mfma32(a[x][0], bq[y][3], im);
          im = mfma32(a[x][1], bq[y][2], im);
          im = mfma32(a[x][2], bq[y][0], im);
          im = mfma32(a[x][2], bq[y][1], im);
          im = mfma32(a[x][3], bq[y][0], im);
          fr[x][y] = r;
          fi[x][y] = im;
        }
      }
    }
    if (++nacc == kFold) fold();
    __syncthreads();
  }
  if (cur_t >= 0) fold();
  // result element (x, y, v): row m0 + 32 wr + 16 x + 4 (lane / 16) + v, column n0 + 32 wc + 16 y + lane % 16
  const Bufs B = bufs(d);
  const cplx* H = (const cplx*)(ws + L.H);
  const double* gam = ws + L.gam;
  cplx* pk = (cplx*)(ws + L.pk) + ((int64_t)slice * 3 + out) * Nk * Nk;
  #pragma unroll
  for (int x = 0; x < 2; ++x)
    #pragma unroll
    for (int y = 0; y < 2; ++y)
      #pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = m0 + wr * 32 + x * 16 + 4 * (lane >> 4) + v;
        const int j = n0 + wc * 32 + y * 16 + (lane & 15);
        if (i >= Nk || j >= Nk) continue;
        const cplx S = mk(accr[x][y][v], acci[x][y][v]);
        if (ks == 1) mixed_epilogue(d, B, H, gam, out, i, j, S, deriv);
        else pk[(int64_t)i * Nk + j] = S;
      }
}

// sum of the wire's ks partial tiles (fixed order) + epilogue; grid (ceil(NN/256), 3, nwires);
// a wire with ks = 1 had its epilogue in the contraction kernel
__global__ void __launch_bounds__(256) k_rhs_epi(const pss_wire_desc* __restrict__ descs, int KS, int knob, int cap,
                                                  int deriv) {
  const pss_wire_desc& d = descs[blockIdx.z];
  if (!wire_on(d)) return;
  const int Nk = d.Nk, out = blockIdx.y;
  const int ks = wire_ks(Nk, knob, cap, KS);
  if (ks == 1) return;
  const int64_t NN = (int64_t)Nk * Nk;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= NN) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const cplx* pk = (const cplx*)(ws + L.pk);
  cplx S = mk(0, 0);
  for (int s = 0; s < ks; ++s) S = S + pk[((int64_t)s * 3 + out) * NN + o];
  if (d.precision == PSS_MIXED)
    mixed_epilogue(d, bufs(d), (const cplx*)(ws + L.H), ws + L.gam, out, (int)(o / Nk), (int)(o % Nk), S, deriv);
  else
    fp64_epilogue(d, bufs(d), (const cplx*)(ws + L.H), ws + L.gam, out, (int)(o / Nk), (int)(o % Nk), S, deriv);
}

// --------------------------- K7: carrier-carrier + phonon collision rates
// MBCE/MBCH (PSTD3D/coulomb.py:957-1071) with the Ceh/Cee/Chh tables generated
// on the fly and read with the reference's one-slot index shift (slot (a,b,c)
// holds the entry built for (a-1,b-1,c-1); slot 0 is zero); MBPE/MBPH
// (PSTD3D/phonons.py:140-268).  n = |Re n3| for MBC*, Re n3 for MBP*.
//
// Regrouping of the eight reference sums (1-based momenta, n padded with n[0]=0).
// Every term is X(a,b,c) * (population factors) with, for e-h scattering,
// X = V_eh(|a-c|)^2 C_eh[a,b,c] and the fourth momentum d = b + c - a, and the
// shifted table read C_eh[a,b,c] = delta(Ee[a-2] + Eh[b-2] - Eh[d-2] - Ee[c-2]),
// nonzero only for a,b,c,d >= 2 and a != c:
//   Win_e[a]  += X (1-nh[b]) nh[d] ne[c]      (MBCE in,  coulomb.py:975)
//   Wout_e[c] += X (1-ne[a]) (1-nh[b]) nh[d]  (MBCE out, coulomb.py:986)
//   Win_h[b]  += X (1-ne[a]) nh[d] ne[c]      (MBCH in)
//   Wout_h[d] += X (1-ne[a]) (1-nh[b]) ne[c]  (MBCH out)
// and for e-e (h-h alike) with d = a + b - c and C_ee[a,b,c] =
// delta(Ee[a-2] + Ee[b-2] - Ee[d-2] - Ee[c-2]):
//   Win_e[a]  += X (1-ne[b]) ne[d] ne[c],   Wout_e[c] += X (1-ne[a]) (1-ne[b]) ne[d].
// A workgroup owns one output momentum o, a thread one partner momentum t:
//   electrons (a = o, c = t): S = sum_b delta(..) (1-nh[b]) nh[b+t-o]   (e-h)
//                                 + sum_b delta(..) (1-ne[b]) ne[o+b-t] (e-e, own V)
//     Win_e[o] += ne[t] V^2 S,  Wout_e[t] += (1-ne[o]) V^2 S
//   holes (e-h: b = o, d = t; h-h: a = o, c = t):
//     S = sum_a delta(..) (1-ne[a]) ne[t+a-o]   (e-h),  sum_b delta(..) (1-nh[b]) nh[o+b-t]  (h-h)
//     Win_h[o] += nh[t] V^2 S,  Wout_h[t] += (1-nh[o]) V^2 S
// so each line shape is evaluated once per (species, kind) instead of once per
// output term.  Wout partial sums go to a per-wire (species, t, o) array and are
// reduced over o by k_wsum, in a fixed order.
struct CollCtx {
  double hw2[3], ihw[3], pre[3];    // per kind (eh, ee, hh): (hbar g)^2, 1/(hbar g), prefactor
  bool lorentz;
  bool fast;                        // mixed precision: exp / reciprocal of the line shape in FP32
};

// GaussDelta: (2 pi/hbar) exp(-(E/hw)^2) / (sqrt(pi) hw)  (coulomb.py:245-263);
// Lorentz: 2 g / (E^2 + hw^2).  FAST (mixed precision): the exponential, or the
// reciprocal (2 g / hw^2) / (z^2 + 1) of the Lorentzian (hardware v_rcp_f32), in FP32.
template <bool LORENTZ, bool FAST>
__device__ inline double line_shape(const CollCtx& x, int kind, double Ed) {
  const double z = Ed * x.ihw[kind];
  if (LORENTZ) {
    if (FAST) return x.pre[kind] / x.hw2[kind] * (double)__builtin_amdgcn_rcpf((float)(z * z + 1.0));
    return ddiv(x.pre[kind], Ed * Ed + x.hw2[kind]);
  }
  if (FAST) return x.pre[kind] * (double)__expf(-(float)(z * z));
  return x.pre[kind] * exp(-(z * z));
}

__device__ inline CollCtx coll_ctx(const pss_wire_desc& d) {
  CollCtx x;
  x.lorentz = (d.flags & PSS_LORENTZDELTA) != 0;
  x.fast = d.precision == PSS_MIXED;
  const double g[3] = {d.gam_eh, d.gam_e, d.gam_h};
  for (int c = 0; c < 3; ++c) {
    const double hw = kHbar * g[c];
    x.hw2[c] = hw * hw;
    x.ihw[c] = 1.0 / hw;
    x.pre[c] = x.lorentz ? 2.0 * g[c] : (2.0 * kPi / kHbar) * (1.0 / (sqrt(kPi) * hw));
  }
  return x;
}

// One species' line data: energies E (global, E(k) = E[k-2] for 1-based k >= 2), and in
// LDS m = 1 - n and n indexed by 1-based momentum k = 0..Nk (n = |Re n3|).
struct LineSp {
  const double* E;
  const double* m;
  const double* n;
};
__device__ __forceinline__ double ek(const LineSp& S, int k) { return S.E[k - 2]; }

// Gaussian line (GaussDelta) by recurrence.  With parabolic bands on a uniform grid the
// energy mismatch of a line is linear in the summed momentum, z_j = (base + E(j) -
// E(j+s)) / hw = z_lo + (j - lo) beta, so e_j = exp(-z_j^2) obeys
//   e_{j+1} = e_j r_j,  r_j = exp(-2 z_j beta - beta^2),  r_{j+1} = r_j exp(-2 beta^2)
// (and the mirror image downwards).  Starting at the grid point nearest the line centre
// (|z| <= |beta| / 2) the terms only decrease, so the recurrence neither overflows nor
// loses the line to an underflowed start; 3 exp per line instead of one per term, FP64
// throughout (relative error ~1e-16 per step).
__device__ inline double coll_line_gauss(const CollCtx& x, int kind, double base, const LineSp& S, int s, int lo,
                                         int hi) {
  if (hi < lo) return 0.0;
  const double ih = x.ihw[kind];
  const double* __restrict__ m = S.m;
  const double* __restrict__ n = S.n;
  const double zlo = (base + ek(S, lo) - ek(S, lo + s)) * ih;
  if (hi == lo) return x.pre[kind] * exp(-(zlo * zlo)) * (m[lo] * n[lo + s]);
  const double zhi = (base + ek(S, hi) - ek(S, hi + s)) * ih;
  const double beta = (zhi - zlo) / (double)(hi - lo);
  double jc = beta != 0.0 ? (double)lo - zlo / beta : (double)lo;
  jc = fmin(fmax(jc, (double)lo), (double)hi);
  const int jp = (int)rint(jc);
  const double zp = (base + ek(S, jp) - ek(S, jp + s)) * ih;
  const double c = exp(-2.0 * beta * beta);
  const double e0 = exp(-(zp * zp));
  double acc = e0 * (m[jp] * n[jp + s]);
  // Walking away from the centre each weight is the last one times r < 1, r itself shrinking,
  // and (1 - n) n <= 1, so once e r / (1 - r) <= 2^-56 of the sum so far the rest of that side
  // cannot change it beyond rounding: stop there (the line is a few widths across, a few
  // grid points where the momentum transfer is large; a line whose centre lies far outside
  // the range has weights that underflow to 0 from the start and stops at once).
  // The test holds at any point of the walk; it is made once per 8 terms so that the loads
  // of a group issue together.
  constexpr double kTail = 0x1p-56;
  constexpr int G = 8;
  double e = e0, r = exp(-(2.0 * zp * beta + beta * beta));
  int j = jp + 1;
  bool done = false;
  for (; j + G - 1 <= hi; j += G) {
#pragma unroll
    for (int u = 0; u < G; ++u) {
      e *= r;
      r *= c;
      acc = fma(e, m[j + u] * n[j + u + s], acc);
    }
    if (r < 1.0 && e * r <= kTail * (1.0 - r) * acc) {
      done = true;
      break;
    }
  }
  if (!done)
    for (; j <= hi; ++j) {
      e *= r;
      r *= c;
      acc = fma(e, m[j] * n[j + s], acc);
    }
  e = e0;
  r = exp(2.0 * zp * beta - beta * beta);
  j = jp - 1;
  done = false;
  for (; j - G + 1 >= lo; j -= G) {
#pragma unroll
    for (int u = 0; u < G; ++u) {
      e *= r;
      r *= c;
      acc = fma(e, m[j - u] * n[j - u + s], acc);
    }
    if (r < 1.0 && e * r <= kTail * (1.0 - r) * acc) {
      done = true;
      break;
    }
  }
  if (!done)
    for (; j >= lo; --j) {
      e *= r;
      r *= c;
      acc = fma(e, m[j] * n[j + s], acc);
    }
  return x.pre[kind] * acc;
}

// sum over j in [lo, hi] of shape(base + E(j) - E(j+s)) (1 - n[j]) n[j+s] along one line
template <bool LORENTZ, bool FAST>
__device__ inline double coll_line(const CollCtx& x, int kind, double base, const LineSp& S, int s, int lo, int hi) {
  if (!LORENTZ) return coll_line_gauss(x, kind, base, S, s, lo, hi);
  double acc = 0.0;
  for (int j = lo; j <= hi; ++j)
    acc += line_shape<LORENTZ, FAST>(x, kind, base + ek(S, j) - ek(S, j + s)) * (S.m[j] * S.n[j + s]);
  return acc;
}

template <bool LORENTZ, bool FAST>
__device__ inline void coll_pair(const CollCtx& x, bool holes, int Nk, int o, int t, const LineSp& Se,
                                 const LineSp& Sh, double& Sx, double& Ss) {
  if (!holes) {
    const double base = ek(Se, o) - ek(Se, t);
    // e-h: a = o, c = t, loop b, d = b + t - o: Ed = Ee[o-2] - Ee[t-2] + Eh[b-2] - Eh[d-2]
    Sx = coll_line<LORENTZ, FAST>(x, 0, base, Sh, t - o, max(2, 2 + o - t), min(Nk, Nk + o - t));
    // e-e: a = o, c = t, loop b, d = o + b - t
    Ss = coll_line<LORENTZ, FAST>(x, 1, base, Se, o - t, max(2, 2 + t - o), min(Nk, Nk + t - o));
  } else {
    const double base = ek(Sh, o) - ek(Sh, t);
    // e-h: b = o, d = t, loop a, c = t + a - o: Ed = Ee[a-2] - Ee[c-2] + Eh[o-2] - Eh[t-2]
    Sx = coll_line<LORENTZ, FAST>(x, 0, base, Se, t - o, max(2, 2 + o - t), min(Nk, Nk + o - t));
    // h-h: a = o, c = t, loop b, d = o + b - t
    Ss = coll_line<LORENTZ, FAST>(x, 2, base, Sh, o - t, max(2, 2 + t - o), min(Nk, Nk + t - o));
  }
}

__global__ void __launch_bounds__(256) k_coll(const pss_wire_desc* __restrict__ descs, int fmask) {
  const pss_wire_desc& d = descs[blockIdx.z];
  if (!wire_on(d) || !(d.flags & fmask & PSS_EHS)) return;
  const int Nk = d.Nk;
  const int o = (int)blockIdx.x + 1;  // 1-based output momentum
  if (o > Nk) return;
  const bool holes = blockIdx.y == 1;
  __shared__ double sh[16];
  extern __shared__ double lsp[];  // m = 1 - n, n for k = 0..Nk: electrons (m, n), then holes (m, n)
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  double* mE = lsp;
  double* nE = lsp + (Nk + 1);
  double* mH = lsp + 2 * (Nk + 1);
  double* nH = lsp + 3 * (Nk + 1);
  for (int k = threadIdx.x; k <= Nk; k += 256) {
    const double ne = k == 0 ? 0.0 : fabs(B.Cn[(int64_t)(k - 1) * Nk + (k - 1)].x);
    const double nh = k == 0 ? 0.0 : fabs(B.Dn[(int64_t)(k - 1) * Nk + (k - 1)].x);
    mE[k] = 1.0 - ne;
    nE[k] = ne;
    mH[k] = 1.0 - nh;
    nH[k] = nh;
  }
  __syncthreads();
  const CollCtx x = coll_ctx(d);
  const double* vc = ws + L.vc;
  const LineSp Se{d.Ee, mE, nE}, Sh{d.Eh, mH, nH};
  const double* nout = holes ? nH : nE;  // output species
  const double no = nout[o];
  double win = 0.0;
  double* pout = ws + L.cp + (holes ? (int64_t)Nk * Nk : 0);
  for (int t = (int)threadIdx.x + 1; t <= Nk; t += 256) {
    double Sx = 0.0, Ss = 0.0;  // cross-species (e-h) and same-species (e-e / h-h) sums
    if (o >= 2 && t >= 2 && t != o) {
      if (x.lorentz) {
        if (x.fast) coll_pair<true, true>(x, holes, Nk, o, t, Se, Sh, Sx, Ss);
        else coll_pair<true, false>(x, holes, Nk, o, t, Se, Sh, Sx, Ss);
      } else {
        if (x.fast) coll_pair<false, true>(x, holes, Nk, o, t, Se, Sh, Sx, Ss);
        else coll_pair<false, false>(x, holes, Nk, o, t, Se, Sh, Sx, Ss);
      }
      const int q = abs(o - t);
      const double vx = vc[q], vs = vc[(holes ? 2 : 1) * Nk + q];
      Sx = vx * vx * Sx + vs * vs * Ss;
    }
    win += nout[t] * Sx;
    pout[(int64_t)(t - 1) * Nk + (o - 1)] = (1.0 - no) * Sx;
  }
  win = block_sum<256>(win, sh);
  if (threadIdx.x == 0) ws[L.W + (holes ? 2 : 0) * Nk + (o - 1)] = win;
}

// Win / Wout per momentum: collision in-rate from k_coll, out-rate reduced from
// its partials, plus MBPE / MBPH (phonons.py:140-268):
//   MBPE: Win[k] += sum_i Vep[k,i] n[i] EP[k,i];  Wout[k] += sum_i Vep[k,i] (1-n[i]) EP[i,k]
//   MBPH: Win[k] += sum_i Vhp[i,k] n[i] HP[k,i];  Wout[k] += sum_i Vhp[i,k] (1-n[i]) HP[i,k]
//   EP[a,b] = 2 Gph (N0/((E[a]-E[b]-hO)^2+hG^2) + (N0+1)/((E[a]-E[b]+hO)^2+hG^2))
__global__ void __launch_bounds__(256) k_wsum(const pss_wire_desc* __restrict__ descs, int fmask) {
  const pss_wire_desc& d = descs[blockIdx.z];
  const int flags = d.flags & fmask;
  if (!wire_on(d) || !(flags & (PSS_EHS | PSS_PHONON))) return;
  const int Nk = d.Nk;
  const int kk = blockIdx.x;
  if (kk >= Nk) return;
  const bool holes = blockIdx.y == 1;
  __shared__ double sh[16];
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  double win = 0.0, wout = 0.0;
  if (flags & PSS_EHS) {
    const double* pout = ws + L.cp + (holes ? (int64_t)Nk * Nk : 0) + (int64_t)kk * Nk;
    for (int i = threadIdx.x; i < Nk; i += 256) wout += pout[i];
  }
  if (flags & PSS_PHONON) {
    const double* E = holes ? d.Eh : d.Ee;
    const double* e1d = ws + L.e1d;
    const double* vrow = ws + L.vc + (holes ? 2 : 1) * Nk;
    const double N0 = d.N0, hO = d.hOph, hG2 = d.hGph * d.hGph, G2 = 2.0 * d.Gph;
    for (int i = threadIdx.x; i < Nk; i += 256) {
      if (i == kk) continue;
      const cplx ni_c = holes ? B.Dn[(int64_t)i * Nk + i] : B.Cn[(int64_t)i * Nk + i];
      const double ni = ni_c.x;
      const int dq = abs(kk - i);
      const double V = vrow[dq] / e1d[dq] * d.Vscale;
      const double dEab = E[kk] - E[i];  // EP[k, i]
      const double dEba = E[i] - E[kk];  // EP[i, k]
      const double Pki = (N0 / ((dEab - hO) * (dEab - hO) + hG2) + (N0 + 1.0) / ((dEab + hO) * (dEab + hO) + hG2)) * G2;
      const double Pik = (N0 / ((dEba - hO) * (dEba - hO) + hG2) + (N0 + 1.0) / ((dEba + hO) * (dEba + hO) + hG2)) * G2;
      win += V * ni * Pki;
      wout += V * (1.0 - ni) * Pik;
    }
  }
  double ww[2] = {win, wout};
  block_sum_n<256, 2>(ww, sh);
  win = ww[0];
  wout = ww[1];
  if (threadIdx.x == 0) {
    double* W = ws + L.W;
    const int iw = (holes ? 2 : 0) * Nk + kk, iwo = (holes ? 3 : 1) * Nk + kk;
    W[iw] = ((flags & PSS_EHS) ? W[iw] : 0.0) + win;
    W[iwo] = wout;
  }
}

// ------------------------------------------ Recomb: spontaneous emission rates
// Rsp[k] = SpontIntegral(gap + Ee[k] + Eh[k] + Ec[k]) (PSTD3D/emission.py:94-131) from the
// cur populations and the screened V of the step (the reference's _preparation order), with
//   Ec[k] = sum_i ne_i (Vee(0) - Vee|i-k|) + nh_i (Vhh(0) - Vhh|i-k|)
//           + sum_{i != k} (ne_i - nh_i) Veh(0) - Veh(0)           (emission.py:134-175)
// and the photon-energy sum sum_n (h_n + E)^3 w_n of SpontIntegral as the cubic
// rsp[0] + rsp[1] E + rsp[2] E^2 + rsp[3] E^3 of its moments (all terms positive, so the
// two forms agree to rounding).
__global__ void __launch_bounds__(256) k_rsp(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d) || !(d.flags & PSS_RECOMB)) return;
  const int Nk = d.Nk;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= Nk) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const double* vc = ws + L.vc;
  const double veh0 = vc[0], vee0 = vc[Nk], vhh0 = vc[2 * Nk];
  double s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < Nk; ++i) {
    const int q = abs(i - k);
    const double ne = ws[L.ne2 + 2 * i], nh = ws[L.nh2 + 2 * i];
    s1 += ne * (vee0 - vc[Nk + q]) + nh * (vhh0 - vc[2 * Nk + q]);
    if (i != k) s2 += ne * veh0 - nh * veh0;
  }
  const double E = d.gap + d.Ee[k] + d.Eh[k] + ((s1 + s2) - veh0);
  ws[L.rsp + k] = ((d.rsp[3] * E + d.rsp[2]) * E + d.rsp[1]) * E + d.rsp[0];
}

__device__ inline cplx cexp_(cplx z) {
  const double m = exp(z.x);
  double sn, cs;
  sincos(z.y, &sn, &cs);
  return mk(m * cs, m * sn);
}

// Exponential relaxation of one population pair (PSTD3D/SBEs.py:5321-5333) and, with
// Recomb, the recombination decay ne *= exp(-Rsp nh dt), nh *= exp(-Rsp ne dt) with the
// updated ne (:5335-5337).
__device__ inline void relax_pair(const pss_wire_desc& d, const double* ws, const WS& L, int k, cplx& ne, cplx& nh) {
  const int Nk = d.Nk;
  const double dt = d.dt;
  const double wi = ws[L.W + k], wo = ws[L.W + Nk + k];
  const double hi = ws[L.W + 2 * Nk + k], ho = ws[L.W + 3 * Nk + k];
  const double We = wi + wo, Wh = hi + ho;
  const double ee = exp(-We * dt), eh = exp(-Wh * dt);
  const double Ee1 = -1 + exp(We * dt), Eh1 = -1 + exp(Wh * dt);
  ne = ee * ((mk(Ee1, 0) + ne) * wi + ne * wo);
  ne = mk(ne.x / (We + 1e-200), ne.y / (We + 1e-200));
  nh = (eh * ((mk(Eh1, 0) + nh) * hi + nh * ho));
  nh = mk(nh.x / (Wh + 1e-200), nh.y / (Wh + 1e-200));
  if (d.flags & PSS_RECOMB) {
    const double r = ws[L.rsp + k];
    ne = ne * cexp_(mk(-r * nh.x * dt, -r * nh.y * dt));
    nh = nh * cexp_(mk(-r * ne.x * dt, -r * ne.y * dt));
  }
}

// ------------------------------------------------------------- DCTrans
// Transport (PSTD3D/dcfield.py:1169-1247) after the relaxation, before the renormalisation
// (PSTD3D/SBEs.py:5698-5702).  The reference multiplies the FFT of X by exp(-i Y_j dk),
// Y_j = s_j 2 pi / ((Nk - 1) dky) (s_j the FFT-order integer of index j), and transforms
// back: X -> S X S^T (both indices, ShiftN2D, when PSS_LF) or n -> S n on the diagonal
// (ShiftN1D), with the circulant S[a][m] = c[(a - m) mod Nk],
//   c[dd] = (1/Nk) sum_s exp(i s 2 pi u / Nk),  u = dd - sigma,  sigma = Nk dk / ((Nk-1) dky),
// s over the FFT-order integers: the Dirichlet kernel
//   c[dd] = (1/Nk) exp(i pi u h / Nk) sin(pi u) / sin(pi u / Nk),  h = 1 (Nk even), 0 (odd).
// sigma is the shift in grid points; an integer sigma is an exact circular roll.

// K8a (DCTrans wires): relaxation of the next diagonals, before their shift.
__global__ void __launch_bounds__(256) k_dc_relax(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.x];
  if (!wire_on(d) || !(d.flags & PSS_DCTRANS) || !(d.flags & (PSS_EHS | PSS_PHONON))) return;
  const int Nk = d.Nk;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  const Bufs B = bufs(d);
  for (int k = threadIdx.x; k < Nk; k += 256) {
    const int64_t o = (int64_t)k * Nk + k;
    cplx ne = B.Cn[o], nh = B.Dn[o];
    d.ws[L.n3 + 2 * Nk + k] = ne.x;  // populations before the relaxation (the rates' output files)
    d.ws[L.n3 + 3 * Nk + k] = nh.x;
    relax_pair(d, d.ws, L, k, ne, nh);
    B.Cn[o] = ne;
    B.Dn[o] = nh;
  }
}

// The circulant column c[dd] of the step's shift (Eac = Re Ex at q = 0 of the centred DFT,
// the field line's index NQ0 = od Nk - 1 of the workspace copy).
__global__ void __launch_bounds__(256) k_dc_coef(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d) || !(d.flags & PSS_DCTRANS)) return;
  const int Nk = d.Nk;
  const int dd = blockIdx.x * 256 + threadIdx.x;
  if (dd >= Nk) return;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const double Eac = ws[L.eq + (int64_t)(Nk - 1) * 2];
  const double dk = -kE0 * (d.Edc + Eac) / kHbar * d.dt;
  const double dky = Nk > 2 ? d.kr[2] - d.kr[1] : d.kr[1] - d.kr[0];
  const double sigma = (double)Nk * dk / ((double)(Nk - 1) * dky);
  const double u = (double)dd - sigma;
  const double v = u / (double)Nk;
  const double den = sinpi(v);
  double R;
  if (den != 0.0) {
    R = sinpi(u) / den;
  } else {  // u = m Nk: the kernel's limit Nk (-1)^(m (Nk - 1))
    const long long m = llrint(v);
    R = ((m * (Nk - 1)) & 1) ? -(double)Nk : (double)Nk;
  }
  const double ph = (Nk & 1) ? 0.0 : v;
  double sn, cs;
  sincospi(ph, &sn, &cs);
  ws[L.dcc + 2 * dd] = cs * R / (double)Nk;
  ws[L.dcc + 2 * dd + 1] = sn * R / (double)Nk;
}

// The two products of X3 -> S X3 S^T (PSS_LF wires), 64 x 64 FP64 complex output tiles of
// 256 threads (4 x 4 per thread), k chunks of 16 staged in LDS.  PASS 0: T = S X3 into the
// Hamiltonian segment of the workspace (free after the RHS); PASS 1: X3' = T S^T into the
// next buffer and the averaged prev (X1 + X3') / 2 (not on the C, D diagonals, which the
// renormalisation writes).
template <int PASS>
__global__ void __launch_bounds__(256) k_dc_gemm(const pss_wire_desc* __restrict__ descs) {
  const int mat = blockIdx.z % 3, wire = blockIdx.z / 3;
  const pss_wire_desc& d = descs[wire];
  if (!wire_on(d) || !(d.flags & PSS_DCTRANS) || !(d.flags & PSS_LF)) return;
  const int N = d.Nk;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  if (m0 >= N || n0 >= N) return;
  const WS L = ws_layout(N, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const cplx* c = (const cplx*)(ws + L.dcc);
  cplx* T = (cplx*)(ws + L.H) + (int64_t)mat * N * N;
  cplx* Xn = mat == 0 ? B.pn : mat == 1 ? B.Cn : B.Dn;
  __shared__ cplx As[64][17];
  __shared__ cplx Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  cplx acc[4][4];
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) acc[p][q] = mk(0.0, 0.0);
  for (int k0 = 0; k0 < N; k0 += 16) {
    for (int e = threadIdx.x; e < 1024; e += 256) {
      const int r = e >> 4, kk = e & 15;
      const int i = m0 + r, k = k0 + kk;
      cplx v = mk(0.0, 0.0);
      if (i < N && k < N) v = PASS == 0 ? c[i >= k ? i - k : i - k + N] : T[(int64_t)i * N + k];
      As[r][kk] = v;
    }
    for (int e = threadIdx.x; e < 1024; e += 256) {
      const int kk = e >> 6, col = e & 63;
      const int k = k0 + kk, j = n0 + col;
      cplx v = mk(0.0, 0.0);
      if (k < N && j < N) v = PASS == 0 ? Xn[(int64_t)k * N + j] : c[j >= k ? j - k : j - k + N];
      Bs[kk][col] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      cplx a[4], b[4];
      for (int p = 0; p < 4; ++p) a[p] = As[ty + 16 * p][kk];
      for (int q = 0; q < 4; ++q) b[q] = Bs[kk][tx + 16 * q];
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) {
          acc[p][q].x = fma(a[p].x, b[q].x, fma(-a[p].y, b[q].y, acc[p][q].x));
          acc[p][q].y = fma(a[p].x, b[q].y, fma(a[p].y, b[q].x, acc[p][q].y));
        }
    }
    __syncthreads();
  }
  const cplx* X1 = mat == 0 ? B.p1 : mat == 1 ? B.C1 : B.D1;
  cplx* Xp = mat == 0 ? B.pp : mat == 1 ? B.Cp : B.Dp;
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) {
      const int i = m0 + ty + 16 * p, j = n0 + tx + 16 * q;
      if (i >= N || j >= N) continue;
      const int64_t o = (int64_t)i * N + j;
      if (PASS == 0) {
        T[o] = acc[p][q];
      } else {
        Xn[o] = acc[p][q];
        if (!(mat != 0 && i == j)) Xp[o] = (X1[o] + acc[p][q]) * 0.5;
      }
    }
}

// Diagonal shift n -> S n of p3, C3, D3 (DCTrans wires without PSS_LF), one workgroup per
// matrix, the diagonal staged in LDS; p's averaged prev diagonal follows (the C, D ones come
// with the renormalisation).
__global__ void __launch_bounds__(256) k_dc_diag(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d) || !(d.flags & PSS_DCTRANS) || (d.flags & PSS_LF)) return;
  const int N = d.Nk, mat = blockIdx.x;
  const WS L = ws_layout(N, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const cplx* c = (const cplx*)(ws + L.dcc);
  cplx* Xn = mat == 0 ? B.pn : mat == 1 ? B.Cn : B.Dn;
  extern __shared__ double dsh[];
  cplx* n = (cplx*)dsh;
  for (int a = threadIdx.x; a < N; a += 256) n[a] = Xn[(int64_t)a * (N + 1)];
  __syncthreads();
  for (int a = threadIdx.x; a < N; a += 256) {
    cplx sacc = mk(0.0, 0.0);
    for (int m = 0; m < N; ++m) {
      const cplx cv = c[a >= m ? a - m : a - m + N];
      sacc.x = fma(cv.x, n[m].x, fma(-cv.y, n[m].y, sacc.x));
      sacc.y = fma(cv.x, n[m].y, fma(cv.y, n[m].x, sacc.y));
    }
    const int64_t o = (int64_t)a * (N + 1);
    Xn[o] = sacc;
    if (mat == 0) B.pp[o] = (B.p1[o] + sacc) * 0.5;
  }
}

// --------------------- K8: exponential relaxation + renormalisation of n3
// PSTD3D/SBEs.py:5321-5333 and 5704-5718 (diagonals of C3, D3 and of the averaged prev);
// DCTrans wires were relaxed before their shift (k_dc_relax).
__device__ void pops_body(const pss_wire_desc& d, double* sh) {
  const int Nk = d.Nk;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const bool relax = (d.flags & (PSS_EHS | PSS_PHONON)) && !(d.flags & PSS_DCTRANS);
  double se = 0.0, sh_ = 0.0;
  for (int k = threadIdx.x; k < Nk; k += 256) {
    const int64_t o = (int64_t)k * Nk + k;
    cplx ne = B.Cn[o], nh = B.Dn[o];
    if (relax) {
      ws[L.n3 + 2 * Nk + k] = ne.x;  // populations before the relaxation (the rates' output files)
      ws[L.n3 + 3 * Nk + k] = nh.x;
      relax_pair(d, ws, L, k, ne, nh);
    }
    const double ae = cabs(ne), ah = cabs(nh);
    ws[L.n3 + k] = ae;
    ws[L.n3 + Nk + k] = ah;
    se += ae;
    sh_ += ah;
  }
  double ss2[2] = {se, sh_};
  block_sum_n<256, 2>(ss2, sh);
  se = ss2[0];
  sh_ = ss2[1];
  const double total = (se + sh_) / 2.0;
  for (int k = threadIdx.x; k < Nk; k += 256) {
    const int64_t o = (int64_t)k * Nk + k;
    const double ne = ws[L.n3 + k] * total / (se + 1e-200);
    const double nh = ws[L.n3 + Nk + k] * total / (sh_ + 1e-200);
    ws[L.n3 + k] = ne;
    ws[L.n3 + Nk + k] = nh;
    B.Cn[o] = mk(ne, 0.0);
    B.Dn[o] = mk(nh, 0.0);
    const cplx c1 = B.C1[o], d1 = B.D1[o];
    B.Cp[o] = (c1 + mk(ne, 0.0)) * 0.5;
    B.Dp[o] = (d1 + mk(nh, 0.0)) * 0.5;
  }
}

// --------------------------- K9: diagonal sums of p3, C3, conj(D3); the workgroup of
// the main diagonal runs K8 before its sums
// QWPolarization3 / QWRho5 depend on k_e - k_h (resp. k2 - k1) only through
// exp(i (k_e - k_h) r) with k_e - k_h = dd * dkr, so the O(Nr Nk^2) sums of
// PSTD3D/qwoptics.py:56-82, 367-380 reduce to sums along diagonals dd.
__global__ void __launch_bounds__(256) k_diag_sums(const pss_wire_desc* __restrict__ descs, int pops) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d)) return;
  const int Nk = d.Nk, nd = 2 * Nk - 1;
  if ((int)blockIdx.x >= nd) return;
  __shared__ double sh[8 * 4 + 8];
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const Bufs B = bufs(d);
  const int dd = (int)blockIdx.x - (Nk - 1);
  if (dd == 0 && pops) {  // K8 on the main diagonal first: its sums read the new populations
    pops_body(d, sh);
    __syncthreads();
  }
  const int a0 = max(0, -dd), a1 = min(Nk - 1, Nk - 1 - dd);
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = a0 + threadIdx.x; a <= a1; a += 256) {
    const int64_t o = (int64_t)a * Nk + a + dd;
    const cplx pv = B.pn[o];
    const double sg = ((a + 1) & 1) ? -1.0 : 1.0;  // (-1)^(kh+1)
    s[0] += sg * pv.x; s[1] += sg * pv.y;
    s[2] += pv.x; s[3] += pv.y;
    const cplx cv = B.Cn[o], dv = B.Dn[o];
    s[4] += cv.x; s[5] += cv.y;
    s[6] += dv.x; s[7] -= dv.y;
  }
  block_sum_n<256, 8>(s, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) {
      ws[L.ds + (q * nd + blockIdx.x) * 2] = s[2 * q];
      ws[L.ds + (q * nd + blockIdx.x) * 2 + 1] = s[2 * q + 1];
    }
}

// P(r), rho(r) on the wire grid (before the centred FFT pair, which cancels).
// One wavefront per grid point r_j: lanes split the 2Nk-1 diagonals.
__global__ void __launch_bounds__(256) k_pol_rho(const pss_wire_desc* __restrict__ descs) {
  const pss_wire_desc& d = descs[blockIdx.y];
  if (!wire_on(d)) return;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= d.Nr) return;
  const int lane = threadIdx.x & 63;
  const int Nk = d.Nk, nd = 2 * Nk - 1;
  const WS L = ws_layout(Nk, d.Nr, d.nrr);
  double* ws = d.ws;
  const cplx* ds = (const cplx*)(ws + L.ds);
  const double a = d.dkr * d.r[j];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bool lf = d.flags & PSS_LF, opt = d.flags & PSS_OPTICS;
  for (int od = lane; od < nd; od += 64) {
    double sn, cs;
    sincos((double)(od - (Nk - 1)) * a, &sn, &cs);
    const cplx e = mk(cs, sn);
    if (opt) {
      const cplx x = ds[od] * e, p = ds[nd + od] * e;
      v[0] += x.x; v[1] += x.y; v[2] += p.x; v[3] += p.y;
    }
    if (lf) {
      const cplx c = ds[2 * nd + od] * e, q = ds[3 * nd + od] * e;
      v[4] += c.x; v[5] += c.y; v[6] += q.x; v[7] += q.y;
    }
  }
  for (int q = 0; q < 8; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  if (lane != 0) return;
  const double w = d.win[j];
  const double f = 2 * d.ehint / d.area / d.L;
  ws[L.pr + j] = opt ? (v[0] * d.dcv) * w * f : 0.0;
  ws[L.pr + d.Nr + j] = opt ? (v[2] * d.dcv) * w * d.yw * f : 0.0;
  ws[L.pr + 2 * d.Nr + j] = opt ? (v[2] * -d.dcv) * w * d.zsign * f : 0.0;
  cplx* rho = (cplx*)(ws + L.rho);
  rho[j] = lf ? mk(v[4], v[5]) * (w / (2.0 * d.L)) : mk(0, 0);
  rho[d.Nr + j] = lf ? mk(v[6], v[7]) * (w / (2.0 * d.L)) : mk(0, 0);
}

// QWRho5 edge removal + normalisation (qwoptics.py:1065-1077), the 2 e0 ehint/area
// scale (SBEs.py:5764-5765), QW2Prop plasmonic normalisation (qwoptics.py:784-789)
// and the five spline solves on r for QW2Prop.
__device__ void rho_spline_body(const pss_wire_desc& d, double* sh) {
  const int Nk = d.Nk, Nr = d.Nr;
  const WS L = ws_layout(Nk, Nr, d.nrr);
  double* ws = d.ws;
  cplx* rho = (cplx*)(ws + L.rho);
  const bool lf = d.flags & PSS_LF;
  double* y5 = ws + L.pr;  // Px, Py, Pz already here; re, rh appended below as real arrays
  if (lf) {
    const double dr = d.r[1] - d.r[0];
    cplx* R0 = rho;
    cplx* R1 = rho + Nr;
    // the two species side by side: edge values and population totals, then the sums
    // of |rho - edge mean|, then (after normalisation) the plasmonic totals -- three
    // reductions of several values instead of eight
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < 10; j += 320) {
      v[0] += R0[j].x + R0[Nr - 10 + j].x;
      v[1] += R1[j].x + R1[Nr - 10 + j].x;
    }
    for (int k = threadIdx.x; k < Nk; k += 320) {
      v[2] += fabs(ws[L.n3 + k]);
      v[3] += fabs(ws[L.n3 + Nk + k]);
    }
    block_sum_n<320, 4>(v, sh);
    const double sub0 = v[0] / 20.0, sub1 = v[1] / 20.0;
    const double nt0 = v[2] + 1e-50, nt1 = v[3] + 1e-50;
    double w[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < Nr; j += 320) {
      R0[j].x -= sub0;
      R1[j].x -= sub1;
      w[0] += cabs(R0[j]);
      w[1] += cabs(R1[j]);
    }
    block_sum_n<320, 2>(w, sh);
    const double sc0 = nt0 / (w[0] * dr + 1e-100), sc1 = nt1 / (w[1] * dr + 1e-100);
    // QW2Prop plasmonic step: re = |re|, normalise both to the common total
    double t[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < Nr; j += 320) {
      cplx v0 = R0[j] * sc0, v1 = R1[j] * sc1;
      v0 = mk(2.0 * kE0 * v0.x / d.area * d.ehint, 2.0 * kE0 * v0.y / d.area * d.ehint);
      v1 = mk(2.0 * kE0 * v1.x / d.area * d.ehint, 2.0 * kE0 * v1.y / d.area * d.ehint);
      const double a = cabs(v0), b = cabs(v1);
      ws[L.spb + 3 * Nr + j] = a;  // temporary
      ws[L.spb + 4 * Nr + j] = b;
      t[0] += a;
      t[1] += b;
    }
    block_sum_n<320, 2>(t, sh);
    const double s0 = t[0], s1 = t[1];
    const double total = (s0 * dr + s1 * dr + 1e-100) / 2.0;
    const double f0 = total / (s0 * dr + 1e-100), f1 = total / (s1 * dr + 1e-100);
    for (int j = threadIdx.x; j < Nr; j += 320) {
      rho[j] = mk(ws[L.spb + 3 * Nr + j] * f0, 0.0);
      rho[Nr + j] = mk(ws[L.spb + 4 * Nr + j] * f1, 0.0);
    }
    if (threadIdx.x == 0) ws[L.red + 1] = total;
  }
  __syncthreads();
  const int f = threadIdx.x >> 6;  // one wavefront per field
  if (f < 3 || lf) {
    const double* y = f < 3 ? (y5 + f * Nr) : (const double*)(rho + (f - 3) * Nr);
    const int ys = f < 3 ? 1 : 2;
    spline_solve_wave(y, ys, Nr, d.sb_h, d.sb_t, d.sb_b, d.sb_end, ws + L.spb + f * Nr);
  }
}

// Evaluate on rr, |.| and final normalisation of rho, Rho = RhoH - RhoE, zeros when
// gated off or ignorewire, NaN when the launch left out a term the wire's flags ask for;
// advance the step counter.
__device__ void qw2prop_body(const pss_wire_desc& d, const bool on, const bool nan_out, double* sh) {
  constexpr int NT = 320;
  const int n = d.nrr;
  const bool zero = !on || (d.flags & PSS_IGNOREWIRE);
  const int Nr = d.Nr;
  const WS L = ws_layout(d.Nk, Nr, n);
  double* ws = d.ws;
  const bool lf = d.flags & PSS_LF;
  double se = 0.0, sh_ = 0.0;
  if (on) {
    const double* y5 = ws + L.pr;
    const cplx* rho = (const cplx*)(ws + L.rho);
    for (int i = threadIdx.x; i < n; i += NT) {
      const int ix = d.sb_idx[i];
      double v[5];  // all five evaluations' loads before the stores
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        v[f] = 0.0;
        if (ix >= 0 && (f < 3 || lf)) {
          const double* y = f < 3 ? (y5 + f * Nr) : (const double*)(rho + (f - 3) * Nr);
          v[f] = spline_eval(y, f < 3 ? 1 : 2, ws + L.spb + f * Nr, d.sb_h, Nr, ix, d.sb_dx[i]);
        }
      }
#pragma unroll
      for (int f = 0; f < 5; ++f) ws[L.outr + f * n + i] = v[f];
      if (lf) {
        const double a = fabs(ws[L.outr + 3 * n + i]), b = fabs(ws[L.outr + 4 * n + i]);
        ws[L.outr + 3 * n + i] = a;
        ws[L.outr + 4 * n + i] = b;
        se += a;
        sh_ += b;
      }
    }
  }
  double ss2[2] = {se, sh_};
  block_sum_n<NT, 2>(ss2, sh);
  se = ss2[0];
  sh_ = ss2[1];
  const double total = ws[L.red + 1];
  const double fe = lf ? total / (se * d.drr + 1e-100) : 0.0;
  const double fh = lf ? total / (sh_ * d.drr + 1e-100) : 0.0;
  for (int i = threadIdx.x; i < n; i += NT) {
    if (nan_out) {
      const double q = __builtin_nan("");
      d.Pxx[2 * i] = q; d.Pxx[2 * i + 1] = q;
      d.Pyy[2 * i] = q; d.Pyy[2 * i + 1] = q;
      d.Pzz[2 * i] = q; d.Pzz[2 * i + 1] = q;
      d.Rho[2 * i] = q; d.Rho[2 * i + 1] = q;
      continue;
    }
    if (zero) {
      d.Pxx[2 * i] = 0; d.Pxx[2 * i + 1] = 0;
      d.Pyy[2 * i] = 0; d.Pyy[2 * i + 1] = 0;
      d.Pzz[2 * i] = 0; d.Pzz[2 * i + 1] = 0;
      d.Rho[2 * i] = 0; d.Rho[2 * i + 1] = 0;
      continue;
    }
    d.Pxx[2 * i] = ws[L.outr + i]; d.Pxx[2 * i + 1] = 0;
    d.Pyy[2 * i] = ws[L.outr + n + i]; d.Pyy[2 * i + 1] = 0;
    d.Pzz[2 * i] = ws[L.outr + 2 * n + i]; d.Pzz[2 * i + 1] = 0;
    const double re = lf ? ws[L.outr + 3 * n + i] * fe : 0.0;
    const double rh = lf ? ws[L.outr + 4 * n + i] * fh : 0.0;
    ws[L.outr + 3 * n + i] = re;  // RhoE, RhoH on the propagation grid (the output files)
    ws[L.outr + 4 * n + i] = rh;
    d.Rho[2 * i] = rh - re;
    d.Rho[2 * i + 1] = 0;
  }
  if (threadIdx.x == 0 && on) d.ctl[0] += 1;
}

// The rho spline (K10) and the QW2Prop evaluation on the propagation grid in one
// workgroup per wire; a wire that is off (or ignorewire) writes zeros.  feat: the optional
// terms the launch ran (4 DCTrans, 8 Recomb).
__global__ void __launch_bounds__(320) k_rho_qw2prop(const pss_wire_desc* __restrict__ descs, int feat) {
  const pss_wire_desc& d = descs[blockIdx.x];
  __shared__ double sh[4 * 5 + 4];
  const bool on = wire_on(d);
  const bool missing = ((d.flags & PSS_DCTRANS) && !(feat & 4)) || ((d.flags & PSS_RECOMB) && !(feat & 8));
  if (on) rho_spline_body(d, sh);
  __syncthreads();
  qw2prop_body(d, on, on && missing, sh);
}

// ------------------------------------------------------------ Vint (init)
__device__ inline double k03_series(double x) {
  // PSTD3D/coulomb.py:39-66
  if (x > 100.0) return 0.0;
  if (x < 1e-10) return 23.0;
  if (x < 2.0) {
    const double x2 = x * x;
    double result = -log(0.5 * x) - 0.5772156649015329;
    double term = 1.0;
    for (int k = 1; k < 20; ++k) {
      term = term * 0.25 * x2 / (double)(k * k);
      result = result + term * (1.0 / k - 0.5772156649015329);
      if (fabs(term) < 1e-15) break;
    }
    return result;
  }
  return sqrt(kPi / (2.0 * x)) * exp(-x);
}

__global__ void __launch_bounds__(256) k_vint(const double* __restrict__ y, int Ny, const double* __restrict__ q,
                                              double ae, double ah, double Delta0, double pref, double* out) {
  __shared__ double sh[16];
  const int iq = blockIdx.x;
  const int N1 = Ny / 4, N2 = 3 * Ny / 4;
  const int m = N2 - N1 + 1;
  const double kmin = (ae + ah) / 4.0;
  const double dk = fmax(fabs(q[iq]), kmin);
  const double mc = ae * ah / kPi * (y[1] - y[0]) * (y[1] - y[0]);
  double acc = 0.0;
  for (int64_t t = threadIdx.x; t < (int64_t)m * m; t += 256) {
    const int i = N1 + (int)(t / m), j = N1 + (int)(t % m);
    const double dy = y[i] - y[j];
    const double r = sqrt(dy * dy + Delta0 * Delta0);
    const double ex = exp(-(ae * y[i]) * (ae * y[i]) - (ah * y[j]) * (ah * y[j]));
    acc += ex * mc * k03_series(dk * r);
  }
  acc = block_sum<256>(acc, sh);
  if (threadIdx.x == 0) out[iq] = pref * acc;
}

// ------------------------------------------- uniform drive of batched wires
// Exx/Eyy/Ezz[:] = E0_c exp(-u^2/(w0 tw)^2) cos(u) exp(-u^20/(2 w0 tw)^20),
// u = w0 (t - tp), t = t0 + n dt with n the wire's drive counter ctl[1]: the
// Gaussian pulse with super-Gaussian cut-off of docs/source/examples/
// sbe_wire_1d.py:45-54, applied uniformly along the wire (:99-100).  The counter
// lives on the device, so a captured graph replays successive times.
// pulse record (8 doubles): E0x, E0y, E0z, w0, tw, tp, t0, unused.
__global__ void __launch_bounds__(256) k_drive(const pss_wire_desc* __restrict__ descs,
                                               const double* __restrict__ pulses) {
  const pss_wire_desc& d = descs[blockIdx.x];
  const double* pp = pulses + 8 * (int64_t)blockIdx.x;
  const int n = d.ctl[1];
  const double t = pp[6] + n * d.dt;
  const double w0 = pp[3], tw = pp[4], tp = pp[5];
  const double u = w0 * (t - tp);
  const double w0tw = w0 * tw;
  const double u2 = u * u;
  double u20 = u2 * u2;        // u^4
  u20 = u20 * u20 * u2 * u2 * u2 * u2 * u2 * u2;  // u^8 * u^12 = u^20
  double s = 2.0 * w0tw;
  double s2 = s * s, s4 = s2 * s2;
  const double s20 = s4 * s4 * s4 * s4 * s4;
  const double env = exp(-u2 / (w0tw * w0tw)) * cos(u) * exp(-u20 / s20);
  double* E[3] = {(double*)d.Exx, (double*)d.Eyy, (double*)d.Ezz};
  for (int i = threadIdx.x; i < 3 * d.nrr; i += 256) {
    const int c = i / d.nrr, j = i - c * d.nrr;
    E[c][2 * j] = pp[c] * env;
    E[c][2 * j + 1] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) d.ctl[1] = n + 1;
}

}  // namespace

// ================================================================== C ABI
extern "C" int64_t pss_sbe_workspace_bytes(int32_t Nk, int32_t Nr, int32_t nrr, int32_t precision) {
  const WS w = ws_layout(Nk, Nr, nrr);
  return (precision == PSS_MIXED ? w.total : w.total_fp64) * (int64_t)sizeof(double);
}

namespace {
int g_ksplit = 0;  // 0: each wire chooses from its own Nk (wire_ks)
}

extern "C" int pss_sbe_tune(int32_t ksplit) {
  if (ksplit < 0 || ksplit > 16 || (ksplit & (ksplit - 1)))
    return pss_set_msg(2, "pss_sbe_tune: ksplit must be 0 (auto) or a power of two <= 16");
  g_ksplit = ksplit;
  return 0;
}

// Stages of the step, in order (bits of a mask):
//   FRONT   K1: wireoff gate, Prop2QW splines onto r; K2: populations, screened V, E1D; K1b: E(q)
//   TRACES  K3: diagonal traces (direct terms), CalcGammaE/H, max |p|, |C|, |D| of mixed wires
//   OFFDIAG K4: off-diagonal dephasing tables De, Dh
//   CALCH   K5: Meh, band energies, direct terms; K5b: exchange (FFT convolutions along the diagonals)
//   RHS     K6: the dpdt/dCdt/dDdt contractions and their epilogue (leapfrog, or dX/dt only)
//   COLL    K7: collision and phonon rates Win/Wout from the populations of the next state
//   TAIL    K8 relaxation + renormalisation, K9 diagonal sums, P(r) and rho(r), QW2Prop
enum { ST_FRONT = 1, ST_TRACES = 2, ST_OFFDIAG = 4, ST_CALCH = 8, ST_RHS = 16, ST_COLL = 32, ST_TAIL = 64,
       ST_ALL = 127 };
// opt: bit 0 = dX/dt only in the RHS epilogue (no leapfrog); bit 1 = the tail's diagonal sums
// without K8 (populations left as they are); bits 8.. = mask on the descriptor flags of K7
static int sbe_stages(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk, int32_t max_Nr,
                      int32_t max_nrr, int32_t precisions, int32_t ksplit_cap, int stages, int opt,
                      hipStream_t s, const char* who) {
  if (nwires <= 0) return 0;
  char msg[160];
  if (min_Nk < 4 || min_Nk > max_Nk) {
    snprintf(msg, sizeof(msg), "%s: need 4 <= min_Nk <= max_Nk", who);
    return pss_set_msg(2, msg);
  }
  if (!(precisions & 3) || precisions > 15) {
    snprintf(msg, sizeof(msg),
             "%s: precisions is a mask of 1 (FP64 wires), 2 (mixed wires), at least one of them, "
             "4 (DCTrans wires), 8 (Recomb wires)", who);
    return pss_set_msg(2, msg);
  }
  if (ksplit_cap < 0 || ksplit_cap > 16 || (ksplit_cap & (ksplit_cap - 1))) {
    snprintf(msg, sizeof(msg), "%s: ksplit_cap must be 0 (none) or a power of two <= 16", who);
    return pss_set_msg(2, msg);
  }
  if (max_Nk > 4096 || max_Nr > 8194 || max_nrr < 4) {
    snprintf(msg, sizeof(msg), "%s: sizes out of range (4 <= Nk <= 4096, nrr >= 4)", who);
    return pss_set_msg(2, msg);
  }
  const bool any64 = precisions & 1, anymx = precisions & 2, anydc = precisions & 4, anyrc = precisions & 8;
  const int deriv = opt & 1, pops = (opt & 2) ? 0 : 1;
  const int fmask = (opt >> 8) ? (opt >> 8) : ~0;
  const int nd = 2 * max_Nk - 1;
  const size_t lds_rel = 2 * ((size_t)max_Nk + 1) * sizeof(double2);
  // gfx950 gives a workgroup up to 160 KB of LDS on request; raised once per process
  // for the largest size seen (outside any stream work, so a captured step is unaffected)
  static size_t lds_granted = 65536;
  if ((stages & ST_COLL) && lds_rel > lds_granted) {
    PSS_CHECK(hipFuncSetAttribute((const void*)k_coll, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rel));
    lds_granted = lds_rel;
  }
  if (stages & ST_FRONT) {
    hipLaunchKernelGGL(k_prop_front, dim3(2, nwires), dim3(256), 0, s, d_descs);
    hipLaunchKernelGGL(k_prop_dft, dim3(nd, nwires), dim3(256), 0, s, d_descs);
  }
  // max |p|, |C|, |D| of the mixed wires in 3 x 64 extra workgroups (after K2 cleared the slots)
  if (stages & ST_TRACES)
    hipLaunchKernelGGL(k_traces_gamma, dim3(nd + max_Nk + (anymx ? 3 * 64 : 0), nwires), dim3(256), 0, s, d_descs,
                       nd + max_Nk);
  if (stages & ST_OFFDIAG)
    hipLaunchKernelGGL(k_offdiag, dim3((max_Nk + 255) / 256, 2 * max_Nk + 1, 2 * nwires), dim3(256),
                       (size_t)(256 + max_Nk) * sizeof(double), s, d_descs);
  if (stages & ST_CALCH) {
    const int t16 = (max_Nk + 15) / 16;
    hipLaunchKernelGGL(k_calc_h, dim3(t16, t16, nwires), dim3(256), 0, s, d_descs);
    // exchange of both precisions: FFT convolutions along the diagonals (FP64)
    // LDS for the largest need of any line size in the launch (a wire of a smaller M packs
    // more lines into the same workgroup)
    const int M = exch_m(max_Nk), LN = exch_lines(M);
    size_t lds_x = 0;
    for (int Mw = 8; Mw <= M; Mw *= 2) {
      const size_t need = ((size_t)exch_lines_w(LN, M, Mw) * (Mw + Mw / 8 + 1) + exch_tw(Mw)) * 16;
      lds_x = need > lds_x ? need : lds_x;
    }
    static size_t lds_x_granted = 65536;
    if (lds_x > lds_x_granted) {
      PSS_CHECK(hipFuncSetAttribute((const void*)k_exch_fft, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_x));
      lds_x_granted = lds_x;
    }
    hipLaunchKernelGGL(k_exch_fft, dim3((2 * max_Nk - 1 + LN - 2) / (LN - 1), 3, nwires), dim3(LN * M / 8), lds_x, s,
                       d_descs, LN, M);
  }
  if (stages & ST_RHS) {
    const int np = np64(max_Nk), t64 = np / 64;
    // split-K: each wire's own factor (wire_ks), the launch sized for the largest of them,
    // which the smallest wire has
    const int KS = wire_ks(min_Nk, g_ksplit, ksplit_cap, 16);
    if (any64)
      hipLaunchKernelGGL(k_rhs_gemm, dim3(t64, t64, 3 * nwires * KS), dim3(256), 0, s, d_descs, KS, g_ksplit,
                         (int)ksplit_cap, deriv);
    if (anymx) {
      hipLaunchKernelGGL(k_maxabs, dim3(64, 5, nwires), dim3(256), 0, s, d_descs, 3);  // Heh .. hbar De
      hipLaunchKernelGGL(k_split, dim3(np / 32, np / 32, 8 * nwires), dim3(256), 0, s, d_descs);
      hipLaunchKernelGGL(k_rhs_mx, dim3(t64, t64, 3 * nwires * KS), dim3(256), 0, s, d_descs, KS, g_ksplit,
                         (int)ksplit_cap, deriv);
    }
    if (KS > 1) {
      const int64_t nn = (int64_t)max_Nk * max_Nk;
      hipLaunchKernelGGL(k_rhs_epi, dim3((unsigned)((nn + 255) / 256), 3, nwires), dim3(256), 0, s, d_descs, KS,
                         g_ksplit, (int)ksplit_cap, deriv);
    }
  }
  if (stages & ST_COLL) {
    hipLaunchKernelGGL(k_coll, dim3(max_Nk, 2, nwires), dim3(256), lds_rel, s, d_descs, fmask);
    hipLaunchKernelGGL(k_wsum, dim3(max_Nk, 2, nwires), dim3(256), 0, s, d_descs, fmask);
    if (anyrc) hipLaunchKernelGGL(k_rsp, dim3((max_Nk + 255) / 256, nwires), dim3(256), 0, s, d_descs);
  }
  if ((stages & ST_TAIL) && pops && anydc) {  // relaxation, then the transport shift of C3, D3, p3
    const int t64 = (max_Nk + 63) / 64;
    hipLaunchKernelGGL(k_dc_relax, dim3(nwires), dim3(256), 0, s, d_descs);
    hipLaunchKernelGGL(k_dc_coef, dim3((max_Nk + 255) / 256, nwires), dim3(256), 0, s, d_descs);
    hipLaunchKernelGGL(k_dc_gemm<0>, dim3(t64, t64, 3 * nwires), dim3(256), 0, s, d_descs);
    hipLaunchKernelGGL(k_dc_gemm<1>, dim3(t64, t64, 3 * nwires), dim3(256), 0, s, d_descs);
    hipLaunchKernelGGL(k_dc_diag, dim3(3, nwires), dim3(256), (size_t)max_Nk * sizeof(double2), s, d_descs);
  }
  if (stages & ST_TAIL) {
    hipLaunchKernelGGL(k_diag_sums, dim3(nd, nwires), dim3(256), 0, s, d_descs, pops);
    hipLaunchKernelGGL(k_pol_rho, dim3((max_Nr + 3) / 4, nwires), dim3(256), 0, s, d_descs);
    if (pops) hipLaunchKernelGGL(k_rho_qw2prop, dim3(nwires), dim3(320), 0, s, d_descs, precisions);
  }
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_sbe_step_p(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                              int32_t max_Nr, int32_t max_nrr, int32_t precisions, int32_t ksplit_cap,
                              void* stream) {
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, precisions, ksplit_cap, ST_ALL, 0,
                    (hipStream_t)stream, "pss_sbe_step_p");
}

extern "C" int pss_sbe_step(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                            int32_t max_Nr, int32_t max_nrr, void* stream) {
  return pss_sbe_step_p(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, 15, 0, stream);
}

extern "C" int pss_sbe_stages(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                              int32_t max_Nr, int32_t max_nrr, int32_t precisions, int32_t stages, int32_t opt,
                              void* stream) {
  if (stages < 1 || stages > ST_ALL) return pss_set_msg(2, "pss_sbe_stages: stages is a mask of 1 ... 64");
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, precisions, 0, stages, opt,
                    (hipStream_t)stream, "pss_sbe_stages");
}

// ---- kernel-level entries: parts of the step on the wires' own buffers (for tests and for
// callers of the reference's module functions); the same kernels as the step
extern "C" int pss_calc_h(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                          int32_t max_Nr, int32_t max_nrr, int32_t precisions, void* stream) {
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, precisions, 0, ST_FRONT | ST_TRACES | ST_CALCH,
                    0, (hipStream_t)stream, "pss_calc_h");
}

extern "C" int pss_sbe_rhs(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                           int32_t max_Nr, int32_t max_nrr, int32_t precisions, void* stream) {
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, precisions, 0,
                    ST_FRONT | ST_TRACES | ST_OFFDIAG | ST_CALCH | ST_RHS, 1, (hipStream_t)stream, "pss_sbe_rhs");
}

extern "C" int pss_coll_rates(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                              int32_t max_Nr, int32_t max_nrr, int32_t which, void* stream) {
  if (which < 1 || which > 3) return pss_set_msg(2, "pss_coll_rates: which is a mask of 1 (MBCE/MBCH), 2 (MBPE/MBPH)");
  const int fm = ((which & 1) ? PSS_EHS : 0) | ((which & 2) ? PSS_PHONON : 0);
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, 3, 0, ST_COLL, fm << 8, (hipStream_t)stream,
                    "pss_coll_rates");
}

extern "C" int pss_qwpol(const pss_wire_desc* d_descs, int32_t nwires, int32_t min_Nk, int32_t max_Nk,
                         int32_t max_Nr, int32_t max_nrr, void* stream) {
  return sbe_stages(d_descs, nwires, min_Nk, max_Nk, max_Nr, max_nrr, 3, 0, ST_TAIL, 2, (hipStream_t)stream,
                    "pss_qwpol");
}

// ---- dpdt / dCdt / dDdt of explicit matrices (the reference's module functions), one
// output element per thread, plain FP64 sums in the order of the reference's terms
__global__ void __launch_bounds__(256) k_rhs_dense(int Nk, int which, const cplx* __restrict__ C,
                                                   const cplx* __restrict__ D, const cplx* __restrict__ p,
                                                   const cplx* __restrict__ Heh, const cplx* __restrict__ Hee,
                                                   const cplx* __restrict__ Hhh, const double* __restrict__ GamE,
                                                   const double* __restrict__ GamH, const cplx* __restrict__ Off,
                                                   cplx* __restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= (int64_t)Nk * Nk) return;
  const int a = (int)(o / Nk), b = (int)(o % Nk);  // output [a][b]
  cplx t1 = mk(0, 0), t2 = mk(0, 0), t3 = mk(0, 0), t4 = mk(0, 0);
  cplx r;
  if (which == 0) {  // dpdt[kh = a][ke = b], PSTD3D/SBEs.py:1953-1976
    for (int j = 0; j < Nk; ++j) {
      t1 = t1 + Hhh[(int64_t)a * Nk + j] * p[(int64_t)j * Nk + b];
      t2 = t2 + Hee[(int64_t)b * Nk + j] * p[(int64_t)a * Nk + j];
      t3 = t3 + Heh[(int64_t)j * Nk + a] * C[(int64_t)j * Nk + b];
      t4 = t4 + Heh[(int64_t)b * Nk + j] * D[(int64_t)j * Nk + a];
    }
    const double g = kHbar * (GamE[b] + GamH[a]);
    const cplx pv = p[o];
    r = t1 + t2 - t3 - t4 + Heh[(int64_t)b * Nk + a] - mk(-g * pv.y, g * pv.x) + (Off ? Off[o] : mk(0, 0));
  } else if (which == 1) {  // dCdt[k1 = a][k2 = b], PSTD3D/SBEs.py:2320-2338
    for (int j = 0; j < Nk; ++j) {
      t1 = t1 + Hee[(int64_t)b * Nk + j] * C[(int64_t)a * Nk + j];
      t2 = t2 + Hee[(int64_t)j * Nk + a] * C[(int64_t)j * Nk + b];
      t3 = t3 + Heh[(int64_t)b * Nk + j] * conj(p[(int64_t)j * Nk + a]);
      t4 = t4 + conj(Heh[(int64_t)a * Nk + j]) * p[(int64_t)j * Nk + b];
    }
    r = t1 - t2 + t3 - t4;
  } else {  // dDdt[k1 = a][k2 = b], PSTD3D/SBEs.py:2618-2636
    for (int j = 0; j < Nk; ++j) {
      t1 = t1 + Hhh[(int64_t)b * Nk + j] * D[(int64_t)a * Nk + j];
      t2 = t2 + Hhh[(int64_t)j * Nk + a] * D[(int64_t)j * Nk + b];
      t3 = t3 + Heh[(int64_t)j * Nk + b] * conj(p[(int64_t)a * Nk + j]);
      t4 = t4 + conj(Heh[(int64_t)j * Nk + a]) * p[(int64_t)b * Nk + j];
    }
    r = t1 - t2 + t3 - t4;
  }
  out[o] = div_ih(r, kHbar);
}

extern "C" int pss_rhs_dense(int32_t Nk, int32_t which, const double* C, const double* D, const double* p,
                             const double* Heh, const double* Hee, const double* Hhh, const double* GamE,
                             const double* GamH, const double* Off, double* out, void* stream) {
  if (Nk < 1 || Nk > 8192 || which < 0 || which > 2) return pss_set_msg(2, "pss_rhs_dense: 1 <= Nk <= 8192, which 0..2");
  if (!C || !D || !p || !Heh || !Hee || !Hhh || !out || (which == 0 && (!GamE || !GamH)))
    return pss_set_msg(2, "pss_rhs_dense: null operand");
  const int64_t nn = (int64_t)Nk * Nk;
  hipLaunchKernelGGL(k_rhs_dense, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, (hipStream_t)stream, Nk, which,
                     (const cplx*)C, (const cplx*)D, (const cplx*)p, (const cplx*)Heh, (const cplx*)Hee,
                     (const cplx*)Hhh, GamE, GamH, (const cplx*)Off, (cplx*)out);
  PSS_CHECK(hipGetLastError());
  return 0;
}

// ---- state access: (p1, p2, C1, C2, D1, D2) of one wire <-> a packed [6][Nk][Nk] complex
// buffer; the prev / cur buffers are resolved on the device from the wire's step counter,
// so the copy can sit in a captured graph
__global__ void __launch_bounds__(256) k_state_copy(const pss_wire_desc* __restrict__ descs, double* packed,
                                                    int to_packed) {
  const pss_wire_desc& d = descs[0];
  const int Nk = d.Nk;
  const int64_t NN = (int64_t)Nk * Nk;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int m = blockIdx.y;  // 0 p1, 1 p2, 2 C1, 3 C2, 4 D1, 5 D2
  if (o >= NN) return;
  const int cur = 1 + (d.ctl[0] & 1);
  double* const* base = m < 2 ? d.p : m < 4 ? d.C : d.D;
  cplx* X = (cplx*)base[(m & 1) ? cur : 0];
  cplx* P = (cplx*)packed + (int64_t)m * NN;
  if (to_packed) P[o] = X[o];
  else X[o] = P[o];
}

static int state_copy(const pss_wire_desc* d_desc, int32_t Nk, double* packed, int to_packed, hipStream_t s,
                      const char* who) {
  if (!d_desc || !packed || Nk < 4 || Nk > 4096) {
    char msg[128];
    snprintf(msg, sizeof(msg), "%s: null pointer or Nk outside 4 ... 4096", who);
    return pss_set_msg(2, msg);
  }
  const int64_t NN = (int64_t)Nk * Nk;
  hipLaunchKernelGGL(k_state_copy, dim3((unsigned)((NN + 255) / 256), 6), dim3(256), 0, s, d_desc, packed, to_packed);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_sbe_get_state(const pss_wire_desc* d_desc, int32_t Nk, double* packed, void* stream) {
  return state_copy(d_desc, Nk, packed, 1, (hipStream_t)stream, "pss_sbe_get_state");
}

extern "C" int pss_sbe_set_state(const pss_wire_desc* d_desc, int32_t Nk, const double* packed, void* stream) {
  return state_copy(d_desc, Nk, (double*)packed, 0, (hipStream_t)stream, "pss_sbe_set_state");
}

extern "C" int pss_drive_uniform(const pss_wire_desc* d_descs, const double* d_pulses, int32_t nwires,
                                 void* stream) {
  if (nwires <= 0) return 0;
  if (!d_descs || !d_pulses) return pss_set_msg(2, "pss_drive_uniform: null pointer");
  hipLaunchKernelGGL(k_drive, dim3(nwires), dim3(256), 0, (hipStream_t)stream, d_descs, d_pulses);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_vint(const double* y, int32_t Ny, const double* q, int32_t nq, double alpha1, double alpha2,
                        double Delta0, double prefactor, double* out, void* stream) {
  if (nq <= 0) return 0;
  if (Ny < 4) return pss_set_msg(2, "pss_vint: Ny < 4");
  hipLaunchKernelGGL(k_vint, dim3(nq), dim3(256), 0, (hipStream_t)stream, y, Ny, q, alpha1, alpha2, Delta0,
                     prefactor, out);
  PSS_CHECK(hipGetLastError());
  return 0;
}
