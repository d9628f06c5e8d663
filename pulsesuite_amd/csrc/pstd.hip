// PSTD Maxwell step on MI355X (gfx950): spectral curl updates, LDS-staged
// radix-2 FFT passes, CPML recursion with the soft source fused in, and the
// wire<->grid coupling kernels (point sample, separable current).
//
// Fields are complex128, C order (x slowest, z fastest), k-space between
// half steps exactly as PSTD3D/PSTD3D.py keeps them.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "../../include/pss.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------- spectral curl updates
// PSTD3D/PSTD3D.py:29-76, at cell o = (ix, iy, iz); the new values are returned too
// CURL = false leaves out the curl term (B is not read): E -= mu0 v2 dt J only.
template <bool CURL = true>
__device__ __forceinline__ void update_e_cell(const pss_grid& g, int64_t o, int ix, int iy, int iz,
                                              const cplx* __restrict__ Bx, const cplx* __restrict__ By,
                                              const cplx* __restrict__ Bz, const cplx* __restrict__ Jx,
                                              const cplx* __restrict__ Jy, const cplx* __restrict__ Jz,
                                              const cplx* __restrict__ jl, const cplx* __restrict__ gy,
                                              const cplx* __restrict__ gz, cplx* Ex, cplx* Ey, cplx* Ez, cplx& ex,
                                              cplx& ey, cplx& ez) {
  const double qx = g.kx[ix], qy = g.ky[iy], qz = g.kz[iz];
  const double coeff = g.v2 * g.dt;
  cplx jx = mk(0, 0), jy = mk(0, 0), jz = mk(0, 0);
  if (Jx) {
    jx = Jx[o]; jy = Jy[o]; jz = Jz[o];
  } else if (jl) {
    const cplx gg = gy[iy] * gz[iz];
    jx = jl[ix] * gg;
    jy = jl[g.Nx + ix] * gg;
    jz = jl[2 * g.Nx + ix] * gg;
  }
  if (CURL) {
    const cplx bx = Bx[o], by = By[o], bz = Bz[o];
    const cplx cx = qy * bz - qz * by, cy = qz * bx - qx * bz, cz = qx * by - qy * bx;
    // i * c * coeff - mu0 * J * coeff
    ex = Ex[o] + (mk(-cx.y, cx.x) * coeff - (g.mu0 * jx) * coeff);
    ey = Ey[o] + (mk(-cy.y, cy.x) * coeff - (g.mu0 * jy) * coeff);
    ez = Ez[o] + (mk(-cz.y, cz.x) * coeff - (g.mu0 * jz) * coeff);
  } else {
    ex = Ex[o] - (g.mu0 * jx) * coeff;
    ey = Ey[o] - (g.mu0 * jy) * coeff;
    ez = Ez[o] - (g.mu0 * jz) * coeff;
  }
  Ex[o] = ex;
  Ey[o] = ey;
  Ez[o] = ez;
}

__global__ void __launch_bounds__(256) k_update_e(pss_grid g, const cplx* __restrict__ Bx, const cplx* __restrict__ By,
                                                  const cplx* __restrict__ Bz, const cplx* __restrict__ Jx,
                                                  const cplx* __restrict__ Jy, const cplx* __restrict__ Jz,
                                                  const cplx* __restrict__ jl, const cplx* __restrict__ gy,
                                                  const cplx* __restrict__ gz, cplx* Ex, cplx* Ey, cplx* Ez) {
  const int64_t n = (int64_t)g.Nx * g.Ny * g.Nz;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= n) return;
  const int lz = __builtin_ctz(g.Nz), ly = __builtin_ctz(g.Ny);  // axis lengths are powers of two
  const int iz = (int)(o & (g.Nz - 1));
  const int iy = (int)((o >> lz) & (g.Ny - 1));
  const int ix = (int)(o >> (lz + ly));
  cplx ex, ey, ez;
  update_e_cell(g, o, ix, iy, iz, Bx, By, Bz, Jx, Jy, Jz, jl, gy, gz, Ex, Ey, Ez, ex, ey, ez);
}

// PSTD3D/PSTD3D.py:79-113
__global__ void __launch_bounds__(256) k_update_b(pss_grid g, const cplx* __restrict__ Ex, const cplx* __restrict__ Ey,
                                                  const cplx* __restrict__ Ez, cplx* Bx, cplx* By, cplx* Bz) {
  const int64_t n = (int64_t)g.Nx * g.Ny * g.Nz;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= n) return;
  const int lz = __builtin_ctz(g.Nz), ly = __builtin_ctz(g.Ny);  // axis lengths are powers of two
  const int iz = (int)(o & (g.Nz - 1));
  const int iy = (int)((o >> lz) & (g.Ny - 1));
  const int ix = (int)(o >> (lz + ly));
  const double qx = g.kx[ix], qy = g.ky[iy], qz = g.kz[iz];
  const cplx ex = Ex[o], ey = Ey[o], ez = Ez[o];
  const cplx cx = qy * ez - qz * ey, cy = qz * ex - qx * ez, cz = qx * ey - qy * ex;
  Bx[o] = Bx[o] - mk(-cx.y, cx.x) * g.dt;
  By[o] = By[o] - mk(-cy.y, cy.x) * g.dt;
  Bz[o] = Bz[o] - mk(-cz.y, cz.x) * g.dt;
}

// ------------------------------------------------------------ FFT pass
// One radix-2 DIT transform of length N along `axis` for TB lines per block,
// staged in LDS (32 KB of data), loads/stores coalesced along z.  Optional
// pre-multiply by i*k_m along axis `mul_axis` (spectral derivative), scale on
// store (1/N per pass for the inverse).  In place is allowed (each block owns
// its lines).
// complex of either precision for the templated passes
template <class T>
struct cpx {
  T x, y;
};
template <class T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> __device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) { return {s * a.x, s * a.y}; }
struct cflt {
  float x, y;
};

struct PassArgs {
  const cplx* in;   // batch element 0 (blockIdx.y = 0)
  cplx* out;
  const cplx *in1, *in2;  // two-level passes: batch elements 1, 2 (same geometry, blockIdx.y = 1, 2)
  cplx *out1, *out2;
  int in32, out32;      // element types of in / out: 0 complex128 (real double), 1 complex64 (real float)
  const double* tw;     // exp(-2 pi i m/N), m < N/2, interleaved
  const double* kmul;   // wavenumbers along mul_axis (null = none)
  int Nx, Ny, Nz, axis, mul_axis, dir;  // dir -1 forward, +1 inverse
  double scale;
  int accum;  // 0: out = result, 1: out += result (F_k += FFT(dF)), 2: real part to ((double*)out)[i]
  int slab;   // > 0 (two-level passes only): write only positions k < slab or k >= N - slab along the
              // axis, into an array whose axis length is 2 slab (k -> k, or k - (N - 2 slab))
  int kmajor; // slab pass along z: the array is (2 slab, Nx, Ny), slab position outermost
  int half;   // store only the outputs k <= N / 2 (a real field's forward pass whose other half
              // the next stage takes from the conjugate pairs)
  // accum 3 (two-level passes, complex128 out): out_c += result + ccoef i (q x G)_c, the
  // spectral curl update of the other field G fused into the accumulating last pass
  // (batch element c = component); q from the grid's wavenumber tables
  const cplx *cg0, *cg1, *cg2;
  const double *qx, *qy, *qz;
  double ccoef;
};

template <int N, int MODE>
__global__ void __launch_bounds__(256) k_fft_pass(PassArgs a) {
  constexpr int TB = (2048 / N) > 0 ? (2048 / N) : 1;
  constexpr int LOG = __builtin_ctz(N);
  __shared__ cplx buf[TB * N];
  __shared__ cplx tws[N / 2 > 0 ? N / 2 : 1];
  const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
  // line geometry
  int64_t base0, sa, st;
  int nvalid;
  int cx = 0, cy = 0, cz0 = 0;  // coordinates of line t=0 (the pass axis coordinate is 0)
  if (a.axis == 2) {
    const int64_t nlines = (int64_t)Nx * Ny;
    const int64_t l0 = (int64_t)blockIdx.x * TB;
    if (l0 >= nlines) return;
    nvalid = (int)min((int64_t)TB, nlines - l0);
    base0 = l0 * N;
    sa = 1;
    st = N;
    cx = (int)(l0 / Ny);
    cy = (int)(l0 % Ny);
  } else {
    const int nch = (Nz + TB - 1) / TB;
    const int other = (int)(blockIdx.x / nch);
    cz0 = (int)(blockIdx.x % nch) * TB;
    if (a.axis == 1) {
      if (other >= Nx) return;
      cx = other;
      base0 = (int64_t)other * Ny * Nz + cz0;
      sa = Nz;
    } else {
      if (other >= Ny) return;
      cy = other;
      base0 = (int64_t)other * Nz + cz0;
      sa = (int64_t)Ny * Nz;
    }
    st = 1;
    nvalid = min(TB, Nz - cz0);
  }
  for (int m = threadIdx.x; m < N / 2; m += 256) {
    const double wr = a.tw[2 * m], wi = a.tw[2 * m + 1];
    tws[m] = mk(wr, a.dir > 0 ? -wi : wi);
  }
  // load (bit-reversed within each line)
  for (int e = threadIdx.x; e < TB * N; e += 256) {
    int t, i;
    if (a.axis == 2) { t = e / N; i = e % N; } else { i = e / TB; t = e % TB; }
    cplx v = mk(0, 0);
    if (t < nvalid) {
      v = a.in[base0 + (int64_t)t * st + (int64_t)i * sa];
      if (a.kmul) {
        int coord;
        if (a.mul_axis == a.axis) coord = i;
        else if (a.axis == 2) {
          // lines (x, y) consecutive in y
          const int64_t l = (int64_t)cx * Ny + cy + t;
          coord = a.mul_axis == 0 ? (int)(l / Ny) : (int)(l % Ny);
        } else if (a.mul_axis == 2) coord = cz0 + t;
        else coord = a.mul_axis == 0 ? cx : cy;
        const double k = a.kmul[coord];
        v = mk(-k * v.y, k * v.x);  // i k v
      }
    }
    const int r = __builtin_bitreverse32((unsigned)i) >> (32 - LOG);
    buf[t * N + r] = v;
  }
  __syncthreads();
  // radix-2 stages
#pragma unroll 1
  for (int s = 0; s < LOG; ++s) {
    const int half = 1 << s;
    for (int b = threadIdx.x; b < TB * (N / 2); b += 256) {
      const int t = b / (N / 2), bb = b % (N / 2);
      const int grp = bb >> s, j = bb & (half - 1);
      const int i0 = (grp << (s + 1)) + j, i1 = i0 + half;
      const cplx w = tws[j << (LOG - 1 - s)];
      const cplx x0 = buf[t * N + i0], x1 = buf[t * N + i1] * w;
      buf[t * N + i0] = x0 + x1;
      buf[t * N + i1] = x0 - x1;
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < TB * N; e += 256) {
    int t, i;
    if (a.axis == 2) { t = e / N; i = e % N; } else { i = e / TB; t = e % TB; }
    if (t < nvalid) {
      const int64_t o = base0 + (int64_t)t * st + (int64_t)i * sa;
      const cplx v = buf[t * N + i] * a.scale;
      if (MODE == 1) a.out[o] = a.out[o] + v;
      else if (MODE == 2) ((double*)a.out)[o] = v.x;
      else a.out[o] = v;
    }
  }
}

// In-register radix-2 DFT of R points (natural order in and out); twiddles
// W_{2h}^j = tw[j * (N / 2h)] from the length-N table in LDS.
template <int R, int N, class T>
__device__ __forceinline__ void dft_reg(cpx<T>* v, const cpx<T>* tw) {
  constexpr int LOG = __builtin_ctz(R);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int r = (int)(__builtin_bitreverse32((unsigned)i) >> (32 - LOG));
    if (r > i) {
      const cpx<T> tmp = v[i];
      v[i] = v[r];
      v[r] = tmp;
    }
  }
#pragma unroll
  for (int h = 1; h < R; h <<= 1) {
#pragma unroll
    for (int g0 = 0; g0 < R; g0 += 2 * h) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const cpx<T> w = tw[j * (N / (2 * h))];
        const cpx<T> x0 = v[g0 + j], x1 = v[g0 + j + h] * w;
        v[g0 + j] = x0 + x1;
        v[g0 + j + h] = x0 - x1;
      }
    }
  }
}

template <class T>
__device__ __forceinline__ cpx<T> load_in(const PassArgs& a, const cplx* in, int64_t o) {
  if (a.in32) {
    const cflt v = ((const cflt*)in)[o];
    return {(T)v.x, (T)v.y};
  }
  const cplx v = in[o];
  return {(T)v.x, (T)v.y};
}

// MODE 0: out = v, MODE 1: out += v (complex128), MODE 2: out = Re v (real array);
// MODE 3 (out += v + curl term) is stored by the pass itself
template <int MODE, class T>
__device__ __forceinline__ void store_out(const PassArgs& a, cplx* out, int64_t o, cpx<T> v) {
  if (MODE == 1) {
    out[o] = out[o] + mk((double)v.x, (double)v.y);
  } else if (MODE == 2) {
    if (a.out32) ((float*)out)[o] = (float)v.x;
    else ((double*)out)[o] = (double)v.x;
  } else {
    if (a.out32) ((cflt*)out)[o] = cflt{(float)v.x, (float)v.y};
    else out[o] = mk((double)v.x, (double)v.y);
  }
}

// Lines per workgroup of a two-level pass: 256 threads of N2 each, and at most 2048
// complex128 or 4096 complex64 elements (the LDS transpose tile's budget).  Along axes 0
// and 1 a workgroup's lines are consecutive in z, so TB elements of one line position are
// contiguous: 16 complex64 lines fill a 128-byte line where 8 would fill half of it.
template <int N1, int N2, bool F32>
constexpr int fft2_tb() {
  return (256 / N2) < ((F32 ? 4096 : 2048) / (N1 * N2)) ? (256 / N2) : ((F32 ? 4096 : 2048) / (N1 * N2));
}

// Two-level transform N = N1 x N2 along `axis`: each line is read straight from
// HBM into registers (N2 threads, N1 points each), N1-point DFTs, inter-level
// twiddles, one padded LDS transpose, N2-point DFTs, store.  One LDS round trip
// per element instead of log2(N) staged radix-2 sweeps.
// by value: a reference to the kernel's by-value argument would put it in scratch memory
template <int N1, int N2, int MODE, class T, bool SLAB>
__device__ __forceinline__ void fft2_body(const PassArgs a, const int bz, const int bx) {
  constexpr int N = N1 * N2;
  constexpr int G = N2;
  constexpr int TB = fft2_tb<N1, N2, sizeof(T) == 4>();
  __shared__ cpx<T> Y[TB * N1 * (N2 + 1)];
  __shared__ cpx<T> tws[N / 2];
  __shared__ T kl[N];  // wavenumbers along the pass axis (derivative passes)
  const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
  int64_t base0, sa, st;
  int nvalid;
  int cx = 0, cy = 0, cz0 = 0;
  // bz: batch element, a uniform choice among scalar arguments
  const cplx* in = bz == 0 ? a.in : bz == 1 ? a.in1 : a.in2;
  cplx* out = bz == 0 ? a.out : bz == 1 ? a.out1 : a.out2;
  const int P2 = SLAB ? 2 * a.slab : 0;
  int64_t obase0, ost, osa;  // output geometry (differs from the input's only for slab output)
  int64_t lb = 0;            // first line of the block (axis 2)
  if (a.axis == 2) {
    const int64_t nlines = (int64_t)Nx * Ny;
    const int64_t l0 = (int64_t)bx * TB;
    if (l0 >= nlines) return;
    lb = l0;
    nvalid = (int)min((int64_t)TB, nlines - l0);
    base0 = l0 * N;
    sa = 1;
    st = N;
    cx = (int)(l0 / Ny);
    cy = (int)(l0 % Ny);
    obase0 = SLAB ? (a.kmajor ? l0 : l0 * P2) : base0;
    ost = SLAB ? (a.kmajor ? 1 : P2) : st;
    osa = SLAB && a.kmajor ? nlines : 1;
  } else {
    const int nch = (Nz + TB - 1) / TB;
    const int other = (int)(bx / nch);
    cz0 = (int)(bx % nch) * TB;
    if (a.axis == 1) {
      if (other >= Nx) return;
      cx = other;
      base0 = (int64_t)other * Ny * Nz + cz0;
      sa = Nz;
      obase0 = SLAB ? (int64_t)other * P2 * Nz + cz0 : base0;
    } else {
      if (other >= Ny) return;
      cy = other;
      base0 = (int64_t)other * Nz + cz0;
      sa = (int64_t)Ny * Nz;
      obase0 = base0;
    }
    st = 1;
    nvalid = min(TB, Nz - cz0);
    ost = st;
    osa = sa;
  }
  for (int m = threadIdx.x; m < N / 2; m += TB * G) {
    const double wr = a.tw[2 * m], wi = a.tw[2 * m + 1];
    tws[m] = cpx<T>{(T)wr, (T)(a.dir > 0 ? -wi : wi)};
  }
  if (a.kmul && a.mul_axis == a.axis)
    for (int m = threadIdx.x; m < N; m += TB * G) kl[m] = (T)a.kmul[m];
  __syncthreads();
  const int tid = threadIdx.x;
  // ---- level 1: thread (t, n2) transforms x[n1 N2 + n2], n1 < N1
  {
    int t, n2;
    if (a.axis == 2) { n2 = tid % G; t = tid / G; } else { t = tid % TB; n2 = tid / TB; }
    cpx<T> v[N1];
    double kf = 0.0;
    bool kconst = false;
    if (a.kmul && a.mul_axis != a.axis) {
      int coord;
      if (a.axis == 2) {
        const int64_t l = (int64_t)cx * Ny + cy + t;
        coord = a.mul_axis == 0 ? (int)(l / Ny) : (int)(l % Ny);
      } else if (a.mul_axis == 2) coord = cz0 + t;
      else coord = a.mul_axis == 0 ? cx : cy;
      kf = a.kmul[min(coord, a.mul_axis == 0 ? Nx - 1 : a.mul_axis == 1 ? Ny - 1 : Nz - 1)];
      kconst = true;
    }
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      cpx<T> x = {(T)0, (T)0};
      if (t < nvalid) {
        const int idx = n1 * N2 + n2;
        x = load_in<T>(a, in, base0 + (int64_t)t * st + (int64_t)idx * sa);
        if (a.kmul) {
          const T k = kconst ? (T)kf : kl[idx];
          x = cpx<T>{-k * x.y, k * x.x};  // i k x
        }
      }
      v[n1] = x;
    }
    dft_reg<N1, N, T>(v, tws);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int m = (n2 * k1) % N;
      const cpx<T> w = m < N / 2 ? tws[m] : cpx<T>{-tws[m - N / 2].x, -tws[m - N / 2].y};
      Y[(t * N1 + k1) * (N2 + 1) + n2] = v[k1] * w;
    }
  }
  __syncthreads();
  // ---- level 2: thread (t, k1) transforms over n2 -> X[k1 + N1 k2]
  {
    int t, k1;
    if (a.axis == 2) { k1 = tid % G; t = tid / G; } else { t = tid % TB; k1 = tid / TB; }
    if (k1 < N1 && t < nvalid) {
      cpx<T> u[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) u[n2] = Y[(t * N1 + k1) * (N2 + 1) + n2];
      dft_reg<N2, N, T>(u, tws);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int k = k1 + N1 * k2;
        int kk = k;
        if (SLAB) {
          if (k >= a.slab && k < N - a.slab) continue;
          kk = k < a.slab ? k : k - (N - P2);
        }
        if (MODE == 0 && a.half && k > N / 2) continue;
        const int64_t o = obase0 + (int64_t)t * ost + (int64_t)kk * osa;
        if (MODE == 3) {
          int ix, iy, iz;
          if (a.axis == 2) {
            const int64_t l = lb + t;
            ix = (int)(l / Ny);
            iy = (int)(l % Ny);
            iz = k;
          } else if (a.axis == 1) {
            ix = cx; iy = k; iz = cz0 + t;
          } else {
            ix = k; iy = cy; iz = cz0 + t;
          }
          // component c = bz: (q x G)_c = q_{c+1} G_{c+2} - q_{c+2} G_{c+1}
          const double q0 = a.qx[ix], q1 = a.qy[iy], q2 = a.qz[iz];
          const double qa = bz == 0 ? q1 : bz == 1 ? q2 : q0, qb = bz == 0 ? q2 : bz == 1 ? q0 : q1;
          const cplx* ga = bz == 0 ? a.cg1 : bz == 1 ? a.cg2 : a.cg0;
          const cplx* gb = bz == 0 ? a.cg2 : bz == 1 ? a.cg0 : a.cg1;
          const cplx cr = qa * gb[o] - qb * ga[o];
          const cpx<T> v = u[k2] * (T)a.scale;
          out[o] = out[o] + (mk((double)v.x, (double)v.y) + mk(-cr.y, cr.x) * a.ccoef);
        } else {
          store_out<MODE, T>(a, out, o, u[k2] * (T)a.scale);
        }
      }
    }
  }
}

// MODE 3 runs on a 1-D grid that deals the three components of one line group to the same
// XCD back to back (workgroups go to the 8 XCDs in turn): line group g of component c is
// workgroup 24 (g / 8) + 8 c + g % 8, so the other field's lines that two components' curl
// terms both read are still in that XCD's L2 for the second.
template <int N1, int N2, int MODE, class T, bool SLAB = false>
__global__ void __launch_bounds__(256) k_fft2(PassArgs a) {
  if (MODE == 3) {
    const int lin = blockIdx.x, grp = lin / 24, r = lin - 24 * grp;
    fft2_body<N1, N2, MODE, T, SLAB>(a, r >> 3, grp * 8 + (r & 7));
  } else {
    fft2_body<N1, N2, MODE, T, SLAB>(a, blockIdx.y, blockIdx.x);
  }
}

// Up to three passes of one length, each over up to two arrays, possibly of different
// geometries and axes (one stage of the slab derivatives along the three axes) in one
// launch: blockIdx.y = sub-pass, blockIdx.x as for a single pass (a sub-pass with fewer
// workgroups than the grid's x extent lets the rest return).
struct MultiPass {
  PassArgs p[3];
  int q[6], b[6];  // blockIdx.y -> (pass, batch element)
};

template <int N1, int N2, int MODE, class T, bool SLAB>
__global__ void __launch_bounds__(256) k_fft2_multi(MultiPass m) {
  const int e = blockIdx.y;
  fft2_body<N1, N2, MODE, T, SLAB>(m.p[m.q[e]], m.b[e], blockIdx.x);
}

// workgroups of one two-level pass along `axis` of grid (Nx, Ny, Nz)
template <int N1, int N2, bool F32>
int64_t fft2_blocks(int axis, int Nx, int Ny, int Nz) {
  constexpr int TB = fft2_tb<N1, N2, F32>();
  if (axis == 2) return ((int64_t)Nx * Ny + TB - 1) / TB;
  return (int64_t)(axis == 1 ? Nx : Ny) * ((Nz + TB - 1) / TB);
}

template <int N1, int N2>
void launch_fft2_multi(const MultiPass& m, int np, int nsub, hipStream_t s) {
  constexpr int G = N2;
  const PassArgs& a = m.p[0];  // the stage's flags are common to the sub-passes
  // the kernel's arithmetic type: FP32 whenever either side is single precision (the slab
  // variant: its output side)
  const bool f32 = a.slab ? a.out32 != 0 : (a.in32 || a.out32);
  const int TB = f32 ? fft2_tb<N1, N2, true>() : fft2_tb<N1, N2, false>();
  int64_t blocks = 0;
  for (int q = 0; q < np; ++q)
    blocks = std::max(blocks, f32 ? fft2_blocks<N1, N2, true>(m.p[q].axis, m.p[q].Nx, m.p[q].Ny, m.p[q].Nz)
                                  : fft2_blocks<N1, N2, false>(m.p[q].axis, m.p[q].Nx, m.p[q].Ny, m.p[q].Nz));
  const dim3 gr((unsigned)blocks, (unsigned)nsub), bl(TB * G);
  if (a.slab) {
    if (a.out32) {
      if (a.accum == 2) hipLaunchKernelGGL((k_fft2_multi<N1, N2, 2, float, true>), gr, bl, 0, s, m);
      else hipLaunchKernelGGL((k_fft2_multi<N1, N2, 0, float, true>), gr, bl, 0, s, m);
    } else {
      if (a.accum == 2) hipLaunchKernelGGL((k_fft2_multi<N1, N2, 2, double, true>), gr, bl, 0, s, m);
      else hipLaunchKernelGGL((k_fft2_multi<N1, N2, 0, double, true>), gr, bl, 0, s, m);
    }
  } else if (a.in32 || a.out32) {
    if (a.accum == 2) hipLaunchKernelGGL((k_fft2_multi<N1, N2, 2, float, false>), gr, bl, 0, s, m);
    else hipLaunchKernelGGL((k_fft2_multi<N1, N2, 0, float, false>), gr, bl, 0, s, m);
  } else {
    if (a.accum == 2) hipLaunchKernelGGL((k_fft2_multi<N1, N2, 2, double, false>), gr, bl, 0, s, m);
    else hipLaunchKernelGGL((k_fft2_multi<N1, N2, 0, double, false>), gr, bl, 0, s, m);
  }
}

template <int N1, int N2>
void launch_fft2(const PassArgs& a, const pss_grid* g, hipStream_t s, int nb) {
  constexpr int G = N2;
  const bool f32 = a.slab ? a.out32 != 0 : (a.in32 || a.out32);  // as the launches below choose T
  const int TB = f32 ? fft2_tb<N1, N2, true>() : fft2_tb<N1, N2, false>();
  int64_t blocks;
  if (a.axis == 2) blocks = ((int64_t)g->Nx * g->Ny + TB - 1) / TB;
  else blocks = (int64_t)(a.axis == 1 ? g->Nx : g->Ny) * ((g->Nz + TB - 1) / TB);
  const dim3 gr(a.accum == 3 ? (unsigned)((blocks + 7) / 8 * 24) : (unsigned)blocks, a.accum == 3 ? 1u : (unsigned)nb),
      bl(TB * G);
  if (a.slab) {  // first pass of a slab derivative: store (complex) or real part
    if (a.out32) {
      if (a.accum == 2) hipLaunchKernelGGL((k_fft2<N1, N2, 2, float, true>), gr, bl, 0, s, a);
      else hipLaunchKernelGGL((k_fft2<N1, N2, 0, float, true>), gr, bl, 0, s, a);
    } else {
      if (a.accum == 2) hipLaunchKernelGGL((k_fft2<N1, N2, 2, double, true>), gr, bl, 0, s, a);
      else hipLaunchKernelGGL((k_fft2<N1, N2, 0, double, true>), gr, bl, 0, s, a);
    }
    return;
  }
  if (a.in32 || a.out32) {  // FP32 arithmetic whenever either side is single precision
    if (a.accum == 3) hipLaunchKernelGGL((k_fft2<N1, N2, 3, float>), gr, bl, 0, s, a);
    else if (a.accum == 1) hipLaunchKernelGGL((k_fft2<N1, N2, 1, float>), gr, bl, 0, s, a);
    else if (a.accum == 2) hipLaunchKernelGGL((k_fft2<N1, N2, 2, float>), gr, bl, 0, s, a);
    else hipLaunchKernelGGL((k_fft2<N1, N2, 0, float>), gr, bl, 0, s, a);
  } else if (a.accum == 3) hipLaunchKernelGGL((k_fft2<N1, N2, 3, double>), gr, bl, 0, s, a);
  else if (a.accum == 1) hipLaunchKernelGGL((k_fft2<N1, N2, 1, double>), gr, bl, 0, s, a);
  else if (a.accum == 2) hipLaunchKernelGGL((k_fft2<N1, N2, 2, double>), gr, bl, 0, s, a);
  else hipLaunchKernelGGL((k_fft2<N1, N2, 0, double>), gr, bl, 0, s, a);
}

// Curl update fused into an accumulating last pass (accum 3): out_c += ... + coef i (q x G)_c.
struct CurlArgs {
  const cplx* G[3];
  double coef;
};

// One pass along `axis` over nb <= 3 arrays of the same geometry (ins[b] -> outs[b]):
// one launch for the two-level sizes, one launch per array otherwise.
int launch_pass_b(const pss_grid* g, const cplx* const* ins, cplx* const* outs, int nb, int axis, int dir,
                  const double* kmul, int mul_axis, hipStream_t s, int accum = 0, int in32 = 0, int out32 = 0,
                  int slab = 0, const CurlArgs* cu = nullptr);

PassArgs make_pass_args(const pss_grid* g, const cplx* const* ins, cplx* const* outs, int nb, int axis, int dir,
                        const double* kmul, int mul_axis, int accum, int in32, int out32, int slab,
                        const CurlArgs* cu);
int launch_pass_radix2(const pss_grid* g, const PassArgs& a, int N, int axis, int accum, int in32, int out32,
                       int slab, hipStream_t s);

int launch_pass(const pss_grid* g, const cplx* in, cplx* out, int axis, int dir, const double* kmul, int mul_axis,
                hipStream_t s, int accum = 0, int in32 = 0, int out32 = 0, int slab = 0) {
  return launch_pass_b(g, &in, &out, 1, axis, dir, kmul, mul_axis, s, accum, in32, out32, slab);
}

int launch_pass_b(const pss_grid* g, const cplx* const* ins, cplx* const* outs, int nb, int axis, int dir,
                  const double* kmul, int mul_axis, hipStream_t s, int accum, int in32, int out32, int slab,
                  const CurlArgs* cu) {
  const int N = axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz;
  if (nb < 1 || nb > 3) return -1;
  if (accum == 3 && (!cu || nb != 3 || !(N == 64 || N == 128 || N == 256 || N == 512 || N == 1024)))
    return -4;  // the fused curl update: the three components, two-level lengths only
  if (N == 1) {
    for (int b = 0; b < nb; ++b)
      if (ins[b] != outs[b]) return -1;
    return 0;
  }
  const bool two_level = N == 64 || N == 128 || N == 256 || N == 512 || N == 1024;
  if (nb > 1 && !two_level) {
    for (int b = 0; b < nb; ++b) {
      const int rc = launch_pass_b(g, ins + b, outs + b, 1, axis, dir, kmul, mul_axis, s, accum, in32, out32, slab);
      if (rc) return rc;
    }
    return 0;
  }
  const PassArgs a = make_pass_args(g, ins, outs, nb, axis, dir, kmul, mul_axis, accum, in32, out32, slab, cu);
  switch (N) {
    case 64: launch_fft2<8, 8>(a, g, s, nb); return 0;
    case 128: launch_fft2<8, 16>(a, g, s, nb); return 0;
    case 256: launch_fft2<16, 16>(a, g, s, nb); return 0;
    case 512: launch_fft2<16, 32>(a, g, s, nb); return 0;
    case 1024: launch_fft2<32, 32>(a, g, s, nb); return 0;
    default: break;
  }
  return launch_pass_radix2(g, a, N, axis, accum, in32, out32, slab, s);
}

PassArgs make_pass_args(const pss_grid* g, const cplx* const* ins, cplx* const* outs, int nb, int axis, int dir,
                        const double* kmul, int mul_axis, int accum, int in32, int out32, int slab,
                        const CurlArgs* cu) {
  const int N = axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz;
  PassArgs a;
  a.in = ins[0];
  a.out = outs[0];
  a.in1 = nb > 1 ? ins[1] : nullptr;
  a.out1 = nb > 1 ? outs[1] : nullptr;
  a.in2 = nb > 2 ? ins[2] : nullptr;
  a.out2 = nb > 2 ? outs[2] : nullptr;
  a.tw = axis == 0 ? g->twx : axis == 1 ? g->twy : g->twz;
  a.kmul = kmul;
  a.Nx = g->Nx; a.Ny = g->Ny; a.Nz = g->Nz;
  a.axis = axis;
  a.mul_axis = mul_axis;
  a.dir = dir;
  a.scale = dir > 0 ? 1.0 / N : 1.0;
  a.accum = accum;
  a.in32 = in32;
  a.out32 = out32;
  a.slab = slab;
  a.kmajor = 0;
  a.half = 0;
  if (cu) {
    a.cg0 = cu->G[0]; a.cg1 = cu->G[1]; a.cg2 = cu->G[2];
    a.qx = g->kx; a.qy = g->ky; a.qz = g->kz;
    a.ccoef = cu->coef;
  } else {
    a.cg0 = a.cg1 = a.cg2 = nullptr;
    a.qx = a.qy = a.qz = nullptr;
    a.ccoef = 0.0;
  }
  return a;
}

// radix-2 LDS passes for the lengths without a two-level kernel (FP64 only)
int launch_pass_radix2(const pss_grid* g, const PassArgs& a, int N, int axis, int accum, int in32, int out32,
                       int slab, hipStream_t s) {
  if (in32 || out32 || slab) return -3;  // single-precision and slab passes exist for N = 64 ... 1024 only
  const int TB = (2048 / N) > 0 ? (2048 / N) : 1;
  int64_t blocks;
  if (axis == 2) blocks = ((int64_t)g->Nx * g->Ny + TB - 1) / TB;
  else blocks = (int64_t)(axis == 1 ? g->Nx : g->Ny) * ((g->Nz + TB - 1) / TB);
  dim3 grid((unsigned)blocks), blk(256);
  switch (N) {
#define PSS_CASE(n)                                                                   \
  case n:                                                                             \
    if (accum == 1) hipLaunchKernelGGL((k_fft_pass<n, 1>), grid, blk, 0, s, a);       \
    else if (accum == 2) hipLaunchKernelGGL((k_fft_pass<n, 2>), grid, blk, 0, s, a);  \
    else hipLaunchKernelGGL((k_fft_pass<n, 0>), grid, blk, 0, s, a);                  \
    break;
    PSS_CASE(2) PSS_CASE(4) PSS_CASE(8) PSS_CASE(16) PSS_CASE(32) PSS_CASE(64) PSS_CASE(128) PSS_CASE(256)
    PSS_CASE(512) PSS_CASE(1024) PSS_CASE(2048)
#undef PSS_CASE
    default: return -2;
  }
  return 0;
}

// Full 3-D transform from `in` to `out` (may alias), optional i*k premultiply.
int fft3(const pss_grid* g, const cplx* in, cplx* out, int dir, const double* kmul, int mul_axis, hipStream_t s) {
  const cplx* src = in;
  bool first = true;
  for (int axis = 0; axis < 3; ++axis) {
    const int N = axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz;
    if (N == 1) continue;
    int rc = launch_pass(g, src, out, axis, dir, first ? kmul : nullptr, mul_axis, s);
    if (rc) return rc;
    first = false;
    src = out;
  }
  if (first && in != out) return -1;  // all axes of length 1: nothing to do in place only
  return 0;
}

// 2-D transform of one N x N complex plane in LDS (N = N1 N2; buf: N rows of pitch N + 1,
// the odd pitch putting a column's points on different banks; tw: W^m, m < N/2, of the
// transform's direction): N-point transforms along the rows, then along the columns, each
// the two-level N1 x N2 transform of the pass kernels.  The rows' first level takes its
// points from load(r, c) and the columns' second level hands the results to
// store(r, c, X), so the plane makes 6 LDS sweeps.  NT threads; ends without a barrier.
// SYNC: store may write the plane buffer (every read of the last level precedes a barrier
// before the first store).
template <int N1, int N2, int NT, bool SYNC = false, class Load, class Store>
__device__ __forceinline__ void plane_fft2d(cpx<float>* buf, const cpx<float>* tw, Load load, Store store) {
  constexpr int N = N1 * N2, PITCH = N + 1;
  constexpr int R1 = N * N2 / NT, R2 = N * N1 / NT;  // level-1 / level-2 tasks per thread
  static_assert(R1 * NT == N * N2 && R2 * NT == N * N1, "tasks must divide evenly");
  const int tid = threadIdx.x;
  auto twid = [&](int mm) {  // W^mm, mm < N
    return mm < N / 2 ? tw[mm] : cpx<float>{-tw[mm - N / 2].x, -tw[mm - N / 2].y};
  };
  // rows, level 1: task (row l, n2) transforms x[n1 N2 + n2] over n1
#pragma unroll
  for (int q = 0; q < R1; ++q) {
    const int task = tid + NT * q, l = task / N2, n2 = task % N2;
    cpx<float> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = load(l, n1 * N2 + n2);
    dft_reg<N1, N, float>(v, tw);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[l * PITCH + k1 * N2 + n2] = v[k1] * twid((n2 * k1) % N);
  }
  __syncthreads();
  // rows, level 2: task (row l, k1) transforms y[k1 N2 + n2] over n2 -> X[k1 + N1 k2] (all
  // reads before the writes: the two index sets of a row overlap)
  {
    cpx<float> u[R2][N2];
#pragma unroll
    for (int q = 0; q < R2; ++q) {
      const int task = tid + NT * q, l = task / N1, k1 = task % N1;
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) u[q][n2] = buf[l * PITCH + k1 * N2 + n2];
      dft_reg<N2, N, float>(u[q], tw);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R2; ++q) {
      const int task = tid + NT * q, l = task / N1, k1 = task % N1;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) buf[l * PITCH + k1 + N1 * k2] = u[q][k2];
    }
  }
  __syncthreads();
  // columns, level 1 in place: task (column c, n2), consecutive threads on consecutive columns
#pragma unroll
  for (int q = 0; q < R1; ++q) {
    const int task = tid + NT * q, c = task % N, n2 = task / N;
    cpx<float> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = buf[(n1 * N2 + n2) * PITCH + c];
    dft_reg<N1, N, float>(v, tw);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) buf[(k1 * N2 + n2) * PITCH + c] = v[k1] * twid((n2 * k1) % N);
  }
  __syncthreads();
  // columns, level 2: results to store
  if (SYNC) {
    cpx<float> u[R2][N2];
#pragma unroll
    for (int q = 0; q < R2; ++q) {
      const int task = tid + NT * q, c = task % N, k1 = task / N;
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) u[q][n2] = buf[(k1 * N2 + n2) * PITCH + c];
      dft_reg<N2, N, float>(u[q], tw);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R2; ++q) {
      const int task = tid + NT * q, c = task % N, k1 = task / N;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) store(k1 + N1 * k2, c, u[q][k2]);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < R2; ++q) {
    const int task = tid + NT * q, c = task % N, k1 = task / N;
    cpx<float> u[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) u[n2] = buf[(k1 * N2 + n2) * PITCH + c];
    dft_reg<N2, N, float>(u, tw);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) store(k1 + N1 * k2, c, u[k2]);
  }
}

// LDS of a plane transform: the plane, then the twiddle table
template <int N>
constexpr size_t plane_lds() { return ((size_t)N * (N + 1) + N / 2) * sizeof(cpx<float>); }

// The passes after the slab pass of the slab derivatives, as one 2-D transform per slab
// plane in LDS (mixed precision, equal axis lengths N = 64 or 128: an N x N complex64 plane
// fits in LDS): N-point inverse transforms along the plane's rows, then along its columns,
// the real parts times 1/N^2 written.  Entry e: derivative array S[e] -> R[e], planes
// pstride apart, rows rstride apart.  Workgroup = one plane.
struct PlaneSet {
  const cflt* S[6];
  float* R[6];
  int64_t pstride[6], rstride[6];
  int first[7];  // first workgroup of entry e (prefix sums of the plane counts)
  int ne;
  const double* tw;  // exp(-2 pi i m/N), interleaved, m < N/2
};

template <int N1, int N2, int NT>
__global__ void __launch_bounds__(NT) k_slab_planes(PlaneSet ps) {
  constexpr int N = N1 * N2;
  extern __shared__ cpx<float> lds_pl[];
  cpx<float>* buf = lds_pl;
  cpx<float>* tw = lds_pl + N * (N + 1);
  const int b = blockIdx.x, tid = threadIdx.x;
  const int e = b >= ps.first[5] ? 5 : b >= ps.first[4] ? 4 : b >= ps.first[3] ? 3 : b >= ps.first[2] ? 2
              : b >= ps.first[1] ? 1 : 0;
  if (e >= ps.ne) return;
  const int pl = b - (e == 0 ? ps.first[0] : e == 1 ? ps.first[1] : e == 2 ? ps.first[2] : e == 3 ? ps.first[3]
                      : e == 4 ? ps.first[4] : ps.first[5]);
  const cflt* S = e == 0 ? ps.S[0] : e == 1 ? ps.S[1] : e == 2 ? ps.S[2] : e == 3 ? ps.S[3] : e == 4 ? ps.S[4] : ps.S[5];
  float* R = e == 0 ? ps.R[0] : e == 1 ? ps.R[1] : e == 2 ? ps.R[2] : e == 3 ? ps.R[3] : e == 4 ? ps.R[4] : ps.R[5];
  const int64_t pst = e == 0 ? ps.pstride[0] : e == 1 ? ps.pstride[1] : e == 2 ? ps.pstride[2]
                    : e == 3 ? ps.pstride[3] : e == 4 ? ps.pstride[4] : ps.pstride[5];
  const int64_t rst = e == 0 ? ps.rstride[0] : e == 1 ? ps.rstride[1] : e == 2 ? ps.rstride[2]
                    : e == 3 ? ps.rstride[3] : e == 4 ? ps.rstride[4] : ps.rstride[5];
  const int64_t base = (int64_t)pl * pst;
  for (int m = tid; m < N / 2; m += NT) tw[m] = cpx<float>{(float)ps.tw[2 * m], (float)-ps.tw[2 * m + 1]};  // inverse
  __syncthreads();
  constexpr float sc = 1.0f / (float)(N * N);
  plane_fft2d<N1, N2, NT>(
      buf, tw,
      [&](int r, int c) {
        const cflt x = S[base + (int64_t)r * rst + c];
        return cpx<float>{x.x, x.y};
      },
      [&](int r, int c, cpx<float> X) { R[base + (int64_t)r * rst + c] = X.x * sc; });
}

// The last two forward passes of dF (along y and z) as one 2-D transform per x-plane in
// LDS (mixed precision, Ny = Nz = N = 64 or 128), with the accumulation into the field
// and the spectral curl update of the other field fused in as the accumulating last pass
// does it (F_c += FFT(dF_c) + ccoef i (q x G)_c, term for term).  dF[c] holds the x pass's
// output (complex64) of a real field, so x-planes come in conjugate pairs:
// X(Nx - g, ky, kz) = conj X(g, -ky, -kz).  A workgroup transforms plane g <= Nx / 2 and
// updates planes g and Nx - g.  Workgroup 24 (g / 8) + 8 c + g % 8 takes component c of
// plane g, so the three components of a plane go to one XCD back to back and the other
// field's planes that two components' curl terms both read are still in that XCD's L2.
struct CurlPlanes {
  const cflt* dF[3];
  cplx* F[3];
  const cplx* G[3];
  const double *qx, *qy, *qz;
  double ccoef;
  const double* tw;  // exp(-2 pi i m/N), interleaved, m < N/2
  int Nx;
};

template <int N1, int N2, int NT>
__global__ void __launch_bounds__(NT) k_dF_planes(CurlPlanes cp) {
  constexpr int N = N1 * N2, PITCH = N + 1;
  extern __shared__ cpx<float> lds_pl[];
  cpx<float>* buf = lds_pl;
  cpx<float>* tw = lds_pl + N * PITCH;
  const int lin = blockIdx.x, grp = lin / 24, r8 = lin - 24 * grp;
  const int c = r8 >> 3, gx = grp * 8 + (r8 & 7);
  const int Nx = cp.Nx;
  if (gx > Nx / 2) return;
  const int tid = threadIdx.x;
  const cflt* dF = c == 0 ? cp.dF[0] : c == 1 ? cp.dF[1] : cp.dF[2];
  cplx* F = c == 0 ? cp.F[0] : c == 1 ? cp.F[1] : cp.F[2];
  // component c: (q x G)_c = q_{c+1} G_{c+2} - q_{c+2} G_{c+1}
  const cplx* ga = c == 0 ? cp.G[1] : c == 1 ? cp.G[2] : cp.G[0];
  const cplx* gb = c == 0 ? cp.G[2] : c == 1 ? cp.G[0] : cp.G[1];
  const int64_t base = (int64_t)gx * N * N;
  for (int m = tid; m < N / 2; m += NT) tw[m] = cpx<float>{(float)cp.tw[2 * m], (float)cp.tw[2 * m + 1]};  // forward
  __syncthreads();
  plane_fft2d<N1, N2, NT, true>(
      buf, tw,
      [&](int r, int col) {  // row r = y, column = z
        const cflt x = dF[base + (int64_t)r * N + col];
        return cpx<float>{x.x, x.y};
      },
      [&](int r, int col, cpx<float> X) { buf[r * PITCH + col] = X; });
  __syncthreads();
  // the field updates of plane gx and, when it is not its own partner, plane Nx - gx
  const int gp = (Nx - gx) & (Nx - 1);
  const bool pair = gp != gx;
  auto upd = [&](int g, int r, int col, cpx<float> X) {
    const int64_t o = ((int64_t)g * N + r) * N + col;
    const double q0 = cp.qx[g], q1 = cp.qy[r], q2 = cp.qz[col];
    const double qa = c == 0 ? q1 : c == 1 ? q2 : q0, qb = c == 0 ? q2 : c == 1 ? q0 : q1;
    const cplx cr = qa * gb[o] - qb * ga[o];
    F[o] = F[o] + (mk((double)X.x, (double)X.y) + mk(-cr.y, cr.x) * cp.ccoef);
  };
#pragma unroll 4
  for (int e = tid; e < N * N; e += NT) {
    const int r = e / N, col = e % N;
    upd(gx, r, col, buf[r * PITCH + col]);
    if (pair) {
      const cpx<float> Y = buf[((N - r) & (N - 1)) * PITCH + ((N - col) & (N - 1))];
      upd(gp, r, col, cpx<float>{Y.x, -Y.y});
    }
  }
}

template <int N1, int N2, int NT>
void launch_dF_planes(const CurlPlanes& cp, hipStream_t s) {
  constexpr size_t lds = plane_lds<N1 * N2>();
  static bool granted = false;  // once per process, outside any stream work
  if (!granted) {
    (void)hipFuncSetAttribute((const void*)k_dF_planes<N1, N2, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);  // a refusal shows as the launch's error
    granted = true;
  }
  hipLaunchKernelGGL((k_dF_planes<N1, N2, NT>), dim3((unsigned)((cp.Nx / 2 + 1 + 7) / 8 * 24)), dim3(NT), lds, s, cp);
}

template <int N1, int N2, int NT>
void launch_slab_planes(const PlaneSet& ps, int nblocks, hipStream_t s) {
  constexpr size_t lds = plane_lds<N1 * N2>();
  static bool granted = false;  // once per process, outside any stream work
  if (!granted) {
    (void)hipFuncSetAttribute((const void*)k_slab_planes<N1, N2, NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);  // a refusal shows as the launch's error
    granted = true;
  }
  hipLaunchKernelGGL((k_slab_planes<N1, N2, NT>), dim3(nblocks), dim3(NT), lds, s, ps);
}

// Launch timing of the accumulating last pass with the curl term (the largest kernel of a
// mixed-precision coupled step), for measurement: armed for n launches, each such launch on a
// stream that is not capturing records an event pair around itself.
std::vector<hipEvent_t> g_tev;
int g_tev_armed = 0, g_tev_next = 0;
std::mutex g_tev_mu;

int timed_slot(hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_tev_mu);
  if (g_tev_next >= g_tev_armed) return -1;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return -1;
  return g_tev_next++;
}

// out += FFT(tmp) along every axis of length > 1; tmp is overwritten (used in place
// for all passes but the last, which accumulates into out).  tmp32: tmp holds
// complex64 and the passes compute in FP32 (out stays complex128).
// nb arrays at once (tmps[b] -> outs[b]), one launch per axis for the two-level sizes.
int fft3_acc(const pss_grid* g, cplx* const* tmps, cplx* const* outs, int nb, int dir, hipStream_t s,
             int tmp32 = 0, const CurlArgs* cu = nullptr) {
  int last = -1;
  for (int axis = 0; axis < 3; ++axis)
    if ((axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz) > 1) last = axis;
  if (last < 0) return -1;
  // complex64 temporaries, curl fused, Ny = Nz = 64 or 128: the pass along x, then the
  // passes along y and z with the accumulation as one 2-D transform per x-plane
  if (cu && tmp32 && nb == 3 && dir < 0 && g->Ny == g->Nz && (g->Ny == 64 || g->Ny == 128)) {
    if (g->Nx > 1) {  // only the x-planes 0 ... Nx/2 are read (conjugate pairs)
      const cplx* ins[3] = {tmps[0], tmps[1], tmps[2]};
      const int two = g->Nx == 64 || g->Nx == 128 || g->Nx == 256 || g->Nx == 512 || g->Nx == 1024;
      if (!two) return -3;  // complex64 temporaries exist for two-level lengths only
      PassArgs a = make_pass_args(g, ins, tmps, 3, 0, dir, nullptr, -1, 0, tmp32, tmp32, 0, nullptr);
      a.half = 1;
      switch (g->Nx) {
        case 64: launch_fft2<8, 8>(a, g, s, 3); break;
        case 128: launch_fft2<8, 16>(a, g, s, 3); break;
        case 256: launch_fft2<16, 16>(a, g, s, 3); break;
        case 512: launch_fft2<16, 32>(a, g, s, 3); break;
        default: launch_fft2<32, 32>(a, g, s, 3); break;
      }
    }
    CurlPlanes cp;
    for (int c = 0; c < 3; ++c) {
      cp.dF[c] = (const cflt*)tmps[c];
      cp.F[c] = outs[c];
      cp.G[c] = cu->G[c];
    }
    cp.qx = g->kx;
    cp.qy = g->ky;
    cp.qz = g->kz;
    cp.ccoef = cu->coef;
    cp.tw = g->twy;
    cp.Nx = g->Nx;
    const int slot = timed_slot(s);
    if (slot >= 0) hipEventRecord(g_tev[2 * slot], s);
    if (g->Ny == 128) launch_dF_planes<8, 16, 1024>(cp, s);
    else launch_dF_planes<8, 8, 512>(cp, s);
    if (slot >= 0) hipEventRecord(g_tev[2 * slot + 1], s);
    return 0;
  }
  for (int axis = 0; axis < 3; ++axis) {
    const int N = axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz;
    if (N == 1) continue;
    const bool fin = axis == last;
    const cplx* ins[3] = {tmps[0], nb > 1 ? tmps[1] : nullptr, nb > 2 ? tmps[2] : nullptr};
    const int slot = fin && cu ? timed_slot(s) : -1;
    if (slot >= 0) hipEventRecord(g_tev[2 * slot], s);
    const int rc = launch_pass_b(g, ins, fin ? outs : tmps, nb, axis, dir, nullptr, -1, s, fin ? (cu ? 3 : 1) : 0,
                                 tmp32, fin ? 0 : tmp32, 0, fin ? cu : nullptr);
    if (slot >= 0) hipEventRecord(g_tev[2 * slot + 1], s);
    if (rc) return rc;
  }
  return 0;
}

// Real parts of the two spectral derivatives of one field G along axes a1 and a2
// (the axes other than c):  R_a = Re IFFT(i k_a G).  The plain inverse pass along c is
// shared: T = IFFT_c(G); R_a1 = Re IFFT_a1(i k_a1 IFFT_a2(T)); R_a2 likewise with the
// roles swapped (the factor i k_a is applied on the load of the pass along a, where
// it commutes with the passes along the other axes).  5 passes instead of 6, the last
// of each writing only the real part.  T, U: complex scratch.  Along an axis of
// length 1, k = 0 and the derivative is zero.
template <class T>
__global__ void k_zero(T* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = (T)0;
}

int deriv_pair(const pss_grid* g, const cplx* G, int c, int a1, int a2, void* R1, void* R2, cplx* T, cplx* U,
               int fp32, hipStream_t s) {
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  const int len[3] = {g->Nx, g->Ny, g->Nz};
  const double* ks[3] = {g->kx, g->ky, g->kz};
  const cplx* t = G;
  int t32 = 0;  // element type of t
  if (len[c] > 1) {
    int rc = launch_pass(g, G, T, c, +1, nullptr, -1, s, 0, 0, fp32);
    if (rc) return rc;
    t = T;
    t32 = fp32;
  }
  for (int h = 0; h < 2; ++h) {
    const int a = h == 0 ? a1 : a2, b = h == 0 ? a2 : a1;
    void* R = h == 0 ? R1 : R2;
    if (len[a] == 1) {
      // k = 0 along a length-1 axis: the derivative is zero
      if (fp32) hipLaunchKernelGGL((k_zero<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (float*)R, n);
      else hipLaunchKernelGGL((k_zero<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (double*)R, n);
      continue;
    }
    const cplx* u = t;
    int u32 = t32;
    if (len[b] > 1) {
      int rc = launch_pass(g, t, U, b, +1, nullptr, -1, s, 0, t32, fp32);
      if (rc) return rc;
      u = U;
      u32 = fp32;
    }
    int rc = launch_pass(g, u, (cplx*)R, a, +1, ks[a], a, s, 2, u32, fp32);
    if (rc) return rc;
  }
  return 0;
}

// The CPML correction reads dG/da only inside the PML slabs along a: outside them
// the profile gives sigma = 0 and kappa = 1, so c_a = 0 and 1/kappa_a - 1 = 0 multiply
// it (PSTD3D/cpml.py:109-140,455-488).  R_q = Re IFFT(i k_a G) is therefore computed
// only there: the pass along a (with the factor i k_a) runs first and writes only the
// 2 P_a slab planes, into an array whose a-axis is 2 P_a long; the passes along the
// other axes run on that array (2 P_a / N_a of the grid), the last writing the real
// part.  Derivative q: 0 dGz/dy, 1 dGy/dz, 2 dGx/dz, 3 dGz/dx, 4 dGy/dx, 5 dGx/dy.
struct SlabR {
  int64_t off[6];   // element offset of derivative q in R
  int on[6];        // q computed (else it is never read: axis of length 1 or no PML along a)
  int full;         // 1: full-grid arrays R[q n + o] (fallback path), 0: slab arrays
  int64_t poff[6];  // element offsets of psi_xy, psi_xz, psi_yx, psi_yz, psi_zx, psi_zy, each
                    // stored on the slab of its derivative axis (elsewhere c = 0 keeps it 0)
  int zmajor;       // the z derivatives (1, 2) are stored (2 Pz, Nx, Ny), slab position outermost
};

// derivative axis of psi_xy, psi_xz, psi_yx, psi_yz, psi_zx, psi_zy
constexpr int kPsiAxis[6] = {1, 2, 0, 2, 0, 1};

// offsets of the six slab-stored psi arrays; returns their total length in doubles
int64_t psi_layout(const pss_grid* g, int64_t poff[6]) {
  const int len[3] = {g->Nx, g->Ny, g->Nz};
  const int pml[3] = {g->npml_x, g->npml_y, g->npml_z};
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  int64_t off = 0;
  for (int j = 0; j < 6; ++j) {
    const int b = kPsiAxis[j];
    poff[j] = off;
    if (len[b] > 1 && pml[b] > 0) off += n / len[b] * 2 * pml[b];
  }
  return off;
}

constexpr int kDerivField[6] = {2, 1, 0, 2, 1, 0};
constexpr int kDerivAxis[6] = {1, 2, 2, 0, 0, 1};

// One stage (same dir, accum and element types) of up to three passes, each over two
// arrays of one geometry: one launch when every pass has the same two-level length,
// else one launch per pass.
struct PassSpec {
  const pss_grid* g;
  const cplx* in[2];
  cplx* out[2];
  int axis;
  const double* kmul;
  int mul_axis;
  int slab;
  int kmajor = 0;
};

int launch_pass_multi(const PassSpec* sp, int ns, int dir, hipStream_t s, int accum, int in32, int out32) {
  auto len = [&](int i) {
    const pss_grid* g = sp[i].g;
    return sp[i].axis == 0 ? g->Nx : sp[i].axis == 1 ? g->Ny : g->Nz;
  };
  const int N = len(0);
  bool one = ns > 1 && (N == 64 || N == 128 || N == 256 || N == 512 || N == 1024);
  for (int i = 1; i < ns; ++i)
    if (len(i) != N) one = false;
  if (!one) {
    for (int i = 0; i < ns; ++i) {
      const int rc = launch_pass_b(sp[i].g, sp[i].in, sp[i].out, 2, sp[i].axis, dir, sp[i].kmul, sp[i].mul_axis, s,
                                   accum, in32, out32, sp[i].slab);
      if (rc) return rc;
    }
    return 0;
  }
  MultiPass m{};
  int e = 0;
  for (int i = 0; i < ns; ++i) {
    m.p[i] = make_pass_args(sp[i].g, sp[i].in, sp[i].out, 2, sp[i].axis, dir, sp[i].kmul, sp[i].mul_axis, accum,
                            in32, out32, sp[i].slab, nullptr);
    m.p[i].kmajor = sp[i].kmajor;
    for (int b = 0; b < 2; ++b) {
      m.q[e] = i;
      m.b[e] = b;
      ++e;
    }
  }
  switch (N) {
    case 64: launch_fft2_multi<8, 8>(m, ns, e, s); break;
    case 128: launch_fft2_multi<8, 16>(m, ns, e, s); break;
    case 256: launch_fft2_multi<16, 16>(m, ns, e, s); break;
    case 512: launch_fft2_multi<16, 32>(m, ns, e, s); break;
    default: launch_fft2_multi<32, 32>(m, ns, e, s); break;
  }
  return 0;
}

// The two derivatives along one axis (of two fields) share every launch: batch of 2.
// When the temporaries of all axes fit in S (3 complex fields) at once, each stage of
// the three axes (the slab pass, then the passes on the compact arrays) is one launch;
// otherwise the axes run one after the other, reusing two slab arrays of at most n / 2
// complex each (2 P <= N).
int slab_derivs(const pss_grid* g, const cplx* const G[3], void* R, cplx* S, int fp32, SlabR& sr, hipStream_t s) {
  const int len[3] = {g->Nx, g->Ny, g->Nz};
  const int pml[3] = {g->npml_x, g->npml_y, g->npml_z};
  const double* ks[3] = {g->kx, g->ky, g->kz};
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  int64_t off = 0;
  sr.full = 0;
  for (int q = 0; q < 6; ++q) {
    const int a = kDerivAxis[q];
    sr.on[q] = 0;
    sr.off[q] = 0;
    if (len[a] == 1 || pml[a] == 0) continue;
    sr.on[q] = 1;
    sr.off[q] = off;
    off += n / len[a] * 2 * pml[a];
  }
  struct AxisPlan {
    int a, P, np, order[3];
    int64_t m;  // elements of one slab array
    cplx* Sq[2];
    cplx* Rq[2];
    const cplx* Gq[2];
    pss_grid gc;  // the slab array's geometry
  };
  AxisPlan ap[3];
  int na = 0;
  int64_t soff = 0;
  for (int a = 0; a < 3; ++a) {
    const int P = pml[a];
    if (len[a] == 1 || P == 0) continue;
    AxisPlan& x = ap[na++];
    x.a = a;
    x.P = P;
    int qs[2], nq = 0;
    for (int q = 0; q < 6; ++q)
      if (kDerivAxis[q] == a) qs[nq++] = q;
    x.m = n / len[a] * 2 * P;
    x.Sq[0] = S + soff;
    x.Sq[1] = S + soff + x.m;
    soff += 2 * x.m;
    for (int h = 0; h < 2; ++h) {
      x.Rq[h] = fp32 ? (cplx*)((float*)R + sr.off[qs[h]]) : (cplx*)((double*)R + sr.off[qs[h]]);
      x.Gq[h] = G[kDerivField[qs[h]]];
    }
    x.gc = *g;
    if (a == 0) x.gc.Nx = 2 * P;
    else if (a == 1) x.gc.Ny = 2 * P;
    else x.gc.Nz = 2 * P;
    x.np = 0;
    x.order[x.np++] = a;
    for (int b = 0; b < 3; ++b)
      if (b != a && len[b] > 1) x.order[x.np++] = b;
  }
  bool together = na > 1 && soff <= 3 * n;
  for (int k = 1; k < na; ++k)
    if (ap[k].np != ap[0].np) together = false;
  if (!together)
    for (int k = 0; k < na; ++k) {  // one axis after the other, two reused slab arrays
      ap[k].Sq[0] = S;
      ap[k].Sq[1] = S + ap[k].m;
    }
  const int np = na ? ap[0].np : 0;
  // plane mode: mixed precision, three axes of one length N = 64 or 128, at least two PML
  // axes run together: the slab pass, then one 2-D transform per slab plane (k_slab_planes);
  // the z derivatives then live (2 Pz, Nx, Ny), slab position outermost, so that a z-slab
  // plane is contiguous
  const int Nc = len[0];
  const bool planes = fp32 && together && na >= 2 && (Nc == 64 || Nc == 128) && len[1] == Nc && len[2] == Nc;
  sr.zmajor = 0;
  auto spec = [&](const AxisPlan& x, int i) {
    const bool last = i == x.np - 1;
    PassSpec p;
    p.g = i == 0 ? g : &x.gc;
    for (int h = 0; h < 2; ++h) {
      p.in[h] = i == 0 ? x.Gq[h] : x.Sq[h];
      p.out[h] = last ? x.Rq[h] : x.Sq[h];
    }
    p.axis = x.order[i];
    p.kmul = i == 0 ? ks[x.a] : nullptr;
    p.mul_axis = i == 0 ? x.a : -1;
    p.slab = i == 0 ? x.P : 0;
    return p;
  };
  if (planes) {
    PassSpec sp[3];
    for (int k = 0; k < na; ++k) {
      sp[k] = spec(ap[k], 0);
      sp[k].kmajor = ap[k].a == 2;
    }
    const int rc = launch_pass_multi(sp, na, +1, s, 0, 0, fp32);
    if (rc) return rc;
    PlaneSet ps{};
    int e = 0, nb = 0;
    for (int k = 0; k < na; ++k) {
      const AxisPlan& x = ap[k];
      for (int h = 0; h < 2; ++h, ++e) {
        ps.S[e] = (const cflt*)x.Sq[h];
        ps.R[e] = (float*)x.Rq[h];
        // slab arrays: x (2P, N, N) planes along x; y (N, 2P, N) planes at j'; z (2P, N, N)
        ps.pstride[e] = x.a == 1 ? Nc : (int64_t)Nc * Nc;
        ps.rstride[e] = x.a == 1 ? (int64_t)2 * x.P * Nc : Nc;
        ps.first[e] = nb;
        nb += 2 * x.P;
      }
      if (x.a == 2) sr.zmajor = 1;
    }
    for (int k = e; k <= 6; ++k) ps.first[k] = nb;
    ps.ne = e;
    ps.tw = g->twx;
    if (Nc == 128) launch_slab_planes<8, 16, 1024>(ps, nb, s);
    else launch_slab_planes<8, 8, 512>(ps, nb, s);
    return 0;
  }
  if (together) {
    for (int i = 0; i < np; ++i) {
      PassSpec sp[3];
      for (int k = 0; k < na; ++k) sp[k] = spec(ap[k], i);
      const int rc = launch_pass_multi(sp, na, +1, s, i == np - 1 ? 2 : 0, i == 0 ? 0 : fp32, fp32);
      if (rc) return rc;
    }
    return 0;
  }
  for (int k = 0; k < na; ++k)
    for (int i = 0; i < ap[k].np; ++i) {
      const PassSpec p = spec(ap[k], i);
      const int rc = launch_pass_b(p.g, p.in, p.out, 2, p.axis, +1, p.kmul, p.mul_axis, s, i == ap[k].np - 1 ? 2 : 0,
                                   i == 0 ? 0 : fp32, fp32, p.slab);
      if (rc) return rc;
    }
  return 0;
}

__device__ __forceinline__ int64_t slab_index(int a, int i, int j, int k, int Nx, int Ny, int Nz, int P) {
  if (a == 0) return ((int64_t)(i < P ? i : i - (Nx - 2 * P)) * Ny + j) * Nz + k;
  if (a == 1) return ((int64_t)i * (2 * P) + (j < P ? j : j - (Ny - 2 * P))) * Nz + k;
  return ((int64_t)i * Ny + j) * (2 * P) + (k < P ? k : k - (Nz - 2 * P));
}

// ------------------------------------------------------------ CPML
// Real-space correction of PSTD3D/cpml.py:150-334 (psi only in PML cells) and the
// soft-source add of PSTD3D/tfsf.py:106-136 on component src_comp, written as a
// real field dF: the reference adds them to Re F in real space between an IFFT and
// an FFT of F; by linearity F_k += FFT(dF) is the same update without taking F
// to real space and back.  R holds the real parts of the six derivatives; dF_x,
// dF_y, dF_z (imaginary part 0) are written for every cell.
struct SrcPar {
  double t, amp, omega0, tau_G, chirp, v, Tp, w0inv2, dx;
};

// ROWS (Nz >= 64): the 64 lanes of a wave lie in one (i, j) row, so i, j, the x- and
// y-slab tests and their coefficients are wave-uniform (scalar); only the z slab splits
// a wave.  Every load of a cell (derivatives, old psi, b/c and 1/kappa - 1 tables) is
// issued before the first use, so a wave waits on memory once.
template <class TR, class TF, bool ROWS>
__global__ void __launch_bounds__(256) k_cpml(pss_grid g, int which, const TR* __restrict__ R, SlabR sr,
                                              TF* __restrict__ dFx, TF* __restrict__ dFy, TF* __restrict__ dFz,
                                              double* __restrict__ psi, int src_comp, SrcPar sp,
                                              const double* __restrict__ srcx, const double* __restrict__ rx,
                                              const double* __restrict__ ry, const double* __restrict__ rz) {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)Nx * Ny * Nz;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= n) return;
  const int lz = __builtin_ctz(Nz), ly = __builtin_ctz(Ny);  // axis lengths are powers of two
  const int k = (int)(o & (Nz - 1));
  int j = (int)((o >> lz) & (Ny - 1));
  int i = (int)(o >> (lz + ly));
  if (ROWS) {
    i = __builtin_amdgcn_readfirstlane(i);
    j = __builtin_amdgcn_readfirstlane(j);
  }
  const int Px = g.npml_x, Py = g.npml_y, Pz = g.npml_z;  // 0 on axes without a PML (and length 1)
  const bool inx = Px > 0 && (i < Px || i >= Nx - Px);
  const bool iny = Py > 0 && (j < Py || j >= Ny - Py);
  const bool inz = Pz > 0 && (k < Pz || k >= Nz - Pz);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (inx || iny || inz) {
    // d0 = dGz/dy, d1 = dGy/dz, d2 = dGx/dz, d3 = dGz/dx, d4 = dGy/dx, d5 = dGx/dy (real parts)
    // and the old psi_xy, psi_xz, psi_yx, psi_yz, psi_zx, psi_zy, each read only on the slab
    // of its axis (elsewhere the coefficients that multiply it are 0, and psi stays 0)
    const double ikx = g.ikx[i], iky = g.iky[j], ikz = g.ikz[k];
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0, d5 = 0.0;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0, p5 = 0.0;
    double bX = 0.0, cX = 0.0, bY = 0.0, cY = 0.0, bZ = 0.0, cZ = 0.0;
    int64_t sx = 0, sy = 0, sz = 0;  // positions on the x, y, z slab arrays
    if (inx) {
      sx = ((int64_t)(i < Px ? i : i - (Nx - 2 * Px)) * Ny + j) * Nz + k;
      const int64_t r = sr.full ? o : sx;
      d3 = (double)R[sr.off[3] + r];
      d4 = (double)R[sr.off[4] + r];
      p2 = psi[sr.poff[2] + sx];
      p4 = psi[sr.poff[4] + sx];
      bX = g.bx[i];
      cX = g.cx[i];
    }
    if (iny) {
      sy = ((int64_t)i * (2 * Py) + (j < Py ? j : j - (Ny - 2 * Py))) * Nz + k;
      const int64_t r = sr.full ? o : sy;
      d0 = (double)R[sr.off[0] + r];
      d5 = (double)R[sr.off[5] + r];
      p0 = psi[sr.poff[0] + sy];
      p5 = psi[sr.poff[5] + sy];
      bY = g.by[j];
      cY = g.cy[j];
    }
    if (inz) {
      const int kz = k < Pz ? k : k - (Nz - 2 * Pz);
      sz = ((int64_t)i * Ny + j) * (2 * Pz) + kz;
      const int64_t r = sr.full ? o : sr.zmajor ? ((int64_t)kz * Nx + i) * Ny + j : sz;
      d1 = (double)R[sr.off[1] + r];
      d2 = (double)R[sr.off[2] + r];
      p1 = psi[sr.poff[1] + sz];
      p3 = psi[sr.poff[3] + sz];
      bZ = g.bz[k];
      cZ = g.cz[k];
    }
    // psi_ab = b_b psi_ab + c_b dG/db (0 outside the slab of b)
    const double vxy = bY * p0 + cY * d0, vxz = bZ * p1 + cZ * d1;
    const double vyx = bX * p2 + cX * d3, vyz = bZ * p3 + cZ * d2;
    const double vzx = bX * p4 + cX * d4, vzy = bY * p5 + cY * d5;
    if (inx) {
      psi[sr.poff[2] + sx] = vyx;
      psi[sr.poff[4] + sx] = vzx;
    }
    if (iny) {
      psi[sr.poff[0] + sy] = vxy;
      psi[sr.poff[5] + sy] = vzy;
    }
    if (inz) {
      psi[sr.poff[1] + sz] = vxz;
      psi[sr.poff[3] + sz] = vyz;
    }
    const double cx_ = iky * d0 + vxy - ikz * d1 - vxz;
    const double cy_ = ikz * d2 + vyz - ikx * d3 - vyx;
    const double cz_ = ikx * d4 + vzx - iky * d5 - vzy;
    const double sc = which == 0 ? g.v2 * g.dt : -g.dt;
    fx = sc * cx_;
    fy = sc * cy_;
    fz = sc * cz_;
  }
  if (src_comp >= 0) {
    const double t = g.tdev ? *g.tdev + sp.t : sp.t;
    const double tau = t - rx[i] / sp.v - sp.Tp;
    const double gyz = exp(-(ry[j] * ry[j] + rz[k] * rz[k]) * sp.w0inv2);
    const double einc = sp.amp * gyz * exp(-(tau * tau) / (sp.tau_G * sp.tau_G)) * cos(sp.omega0 * tau + sp.chirp * tau * tau);
    const double add = srcx[i] * einc * sp.dx;
    if (src_comp == 0) fx += add;
    else if (src_comp == 1) fy += add;
    else fz += add;
  }
  dFx[o] = TF{(decltype(TF::x))fx, 0};
  dFy[o] = TF{(decltype(TF::x))fy, 0};
  dFz[o] = TF{(decltype(TF::x))fz, 0};
}

// Masking absorber (PSTD3D/absorber.py:102-191) on real-space fields, with the soft source
// of the same half step (PSTD3D/tfsf.py:106-136) on src_comp: F *= (mx[i] my[j]) mz[k] (the
// complex field, as the reference multiplies it), then Re F_src += S(x) E_inc dx.
__global__ void __launch_bounds__(256) k_mask_src(pss_grid g, cplx* __restrict__ Fx, cplx* __restrict__ Fy,
                                                  cplx* __restrict__ Fz, const double* __restrict__ mx,
                                                  const double* __restrict__ my, const double* __restrict__ mz,
                                                  int src_comp, SrcPar sp, const double* __restrict__ srcx,
                                                  const double* __restrict__ rx, const double* __restrict__ ry,
                                                  const double* __restrict__ rz) {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)Nx * Ny * Nz;
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= n) return;
  const int lz = __builtin_ctz(Nz), ly = __builtin_ctz(Ny);
  const int k = (int)(o & (Nz - 1)), j = (int)((o >> lz) & (Ny - 1)), i = (int)(o >> (lz + ly));
  const double m = (mx[i] * my[j]) * mz[k];
  cplx fx = Fx[o] * m, fy = Fy[o] * m, fz = Fz[o] * m;
  if (src_comp >= 0) {
    const double t = g.tdev ? *g.tdev + sp.t : sp.t;
    const double tau = t - rx[i] / sp.v - sp.Tp;
    const double gyz = exp(-(ry[j] * ry[j] + rz[k] * rz[k]) * sp.w0inv2);
    const double einc = sp.amp * gyz * exp(-(tau * tau) / (sp.tau_G * sp.tau_G)) * cos(sp.omega0 * tau + sp.chirp * tau * tau);
    const double add = srcx[i] * einc * sp.dx;
    if (src_comp == 0) fx.x += add;
    else if (src_comp == 1) fy.x += add;
    else fz.x += add;
  }
  Fx[o] = fx;
  Fy[o] = fy;
  Fz[o] = fz;
}

// Trilinear sample (EAtXYZ, PSTD3D/usefulsubs.py:2016-2085) of the corrected
// real-space field F + dF at 8 corner cells, with F given in k-space:
//   F(x_c) = (1/N) sum_k F_k exp(+2 pi i (mx ix/Nx + my iy/Ny + mz iz/Nz)),
// so sum_c w_c F(x_c) = (1/N) sum_k F_k Phi(k), Phi(k) = sum_c w_c exp(...).
// Pass 1: per-block partial sums (grid-stride over k); pass 2: one block adds the
// partials in a fixed order, divides by N and adds sum_c w_c dF(x_c).
__device__ inline cplx twp(const cplx* tw, int N, int r) {  // exp(+2 pi i r / N), tw = exp(-2 pi i m/N), m < N/2
  r &= N - 1;
  return r < N / 2 ? conj(tw[r]) : mk(-tw[r - N / 2].x, tw[r - N / 2].y);
}

// Phi(k) = sum_c w_c exp(+2 pi i (mx cx/Nx + my cy/Ny + mz cz/Nz)).  The axis lengths are
// powers of two, so with L = max(Nx, Ny, Nz) the phase is exp(+2 pi i r/L) with the exact
// integer r = (mx cx L/Nx + my cy L/Ny + mz cz L/Nz) mod L: one twiddle of the longest
// axis per corner instead of a product of three.
struct SampleLDS {
  int cx[8], cy[8], cz[8];  // corner coordinates pre-scaled by L/N of their axis
  double cw[8];
  double red[6][4];
};

// tws: L/2 twiddles exp(-2 pi i m/L) (dynamic LDS); ends with a barrier
__device__ inline int sample_prep(const pss_grid& g, const int32_t* __restrict__ idx, const double* __restrict__ w,
                                  cplx* tws, SampleLDS& sl) {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int L = max(Nx, max(Ny, Nz));
  const double* twL = L == Nx ? g.twx : L == Ny ? g.twy : g.twz;
  for (int m = threadIdx.x; m < L / 2; m += 256) tws[m] = mk(twL[2 * m], twL[2 * m + 1]);
  if (threadIdx.x < 8) {
    const int64_t o = idx[threadIdx.x];
    const bool ok = o >= 0;
    sl.cx[threadIdx.x] = ok ? (int)(o / ((int64_t)Ny * Nz)) * (L / Nx) : 0;
    sl.cy[threadIdx.x] = ok ? (int)((o / Nz) % Ny) * (L / Ny) : 0;
    sl.cz[threadIdx.x] = ok ? (int)(o % Nz) * (L / Nz) : 0;
    sl.cw[threadIdx.x] = ok ? w[threadIdx.x] : 0.0;
  }
  __syncthreads();
  return L;
}

__device__ __forceinline__ cplx sample_phi(int L, const cplx* tws, const SampleLDS& sl, int mx, int my, int mz) {
  cplx phi = mk(0, 0);
  for (int c = 0; c < 8; ++c) {
    if (sl.cw[c] == 0.0) continue;
    if (L == 1) {
      phi.x += sl.cw[c];
      continue;
    }
    const int r = (mx * sl.cx[c] + my * sl.cy[c] + mz * sl.cz[c]) & (L - 1);
    phi = phi + sl.cw[c] * twp(tws, L, r);
  }
  return phi;
}

// the block's three complex sums -> partial[blockIdx.x * 6 + q]
__device__ inline void sample_reduce(cplx a0, cplx a1, cplx a2, SampleLDS& sl, double* partial) {
  double v[6] = {a0.x, a0.y, a1.x, a1.y, a2.x, a2.y};
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 6; ++q) {
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    if (lane == 0) sl.red[q][wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 6)
    partial[(int64_t)blockIdx.x * 6 + threadIdx.x] =
        ((sl.red[threadIdx.x][0] + sl.red[threadIdx.x][1]) + sl.red[threadIdx.x][2]) + sl.red[threadIdx.x][3];
}

__global__ void __launch_bounds__(256) k_ksample(pss_grid g, const cplx* __restrict__ Fx, const cplx* __restrict__ Fy,
                                                 const cplx* __restrict__ Fz, const int32_t* __restrict__ idx,
                                                 const double* __restrict__ w, double* partial) {
  extern __shared__ cplx tws[];
  __shared__ SampleLDS sl;
  const int L = sample_prep(g, idx, w, tws, sl);
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)Nx * Ny * Nz;
  const int lz = __builtin_ctz(Nz), ly = __builtin_ctz(Ny);
  cplx a0 = mk(0, 0), a1 = mk(0, 0), a2 = mk(0, 0);
  for (int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x; o < n; o += (int64_t)gridDim.x * 256) {
    const cplx f0 = Fx[o], f1 = Fy[o], f2 = Fz[o];
    const cplx phi = sample_phi(L, tws, sl, (int)(o >> (lz + ly)), (int)((o >> lz) & (Ny - 1)), (int)(o & (Nz - 1)));
    a0 = a0 + f0 * phi;
    a1 = a1 + f1 * phi;
    a2 = a2 + f2 * phi;
  }
  sample_reduce(a0, a1, a2, sl, partial);
}

// The E update with the point sample of the new E (partial sums as k_ksample): one pass
// over E instead of two.  CURL = false: the current term only (the curl term applied
// before, by pss_cpml_e_curl).
template <bool CURL, bool BUPD>
__global__ void __launch_bounds__(256) k_update_e_sample(pss_grid g, const cplx* Bx, const cplx* By, const cplx* Bz,
                                                         const cplx* __restrict__ Jx, const cplx* __restrict__ Jy,
                                                         const cplx* __restrict__ Jz, const cplx* __restrict__ jl,
                                                         const cplx* __restrict__ gy, const cplx* __restrict__ gz,
                                                         cplx* Ex, cplx* Ey, cplx* Ez, const int32_t* __restrict__ idx,
                                                         const double* __restrict__ w, double* partial, cplx* Bwx,
                                                         cplx* Bwy, cplx* Bwz) {
  extern __shared__ cplx tws[];
  __shared__ SampleLDS sl;
  const int L = sample_prep(g, idx, w, tws, sl);
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)Nx * Ny * Nz;
  const int lz = __builtin_ctz(Nz), ly = __builtin_ctz(Ny);
  cplx a0 = mk(0, 0), a1 = mk(0, 0), a2 = mk(0, 0);
  for (int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x; o < n; o += (int64_t)gridDim.x * 256) {
    const int iz = (int)(o & (Nz - 1)), iy = (int)((o >> lz) & (Ny - 1)), ix = (int)(o >> (lz + ly));
    cplx e0, e1, e2;
    update_e_cell<CURL>(g, o, ix, iy, iz, Bx, By, Bz, Jx, Jy, Jz, jl, gy, gz, Ex, Ey, Ez, e0, e1, e2);
    if (BUPD) {  // B -= i dt (k x E) from the new E (as k_update_b)
      const double qx = g.kx[ix], qy = g.ky[iy], qz = g.kz[iz];
      const cplx cx = qy * e2 - qz * e1, cy = qz * e0 - qx * e2, cz = qx * e1 - qy * e0;
      Bwx[o] = Bwx[o] - mk(-cx.y, cx.x) * g.dt;
      Bwy[o] = Bwy[o] - mk(-cy.y, cy.x) * g.dt;
      Bwz[o] = Bwz[o] - mk(-cz.y, cz.x) * g.dt;
    }
    const cplx phi = sample_phi(L, tws, sl, ix, iy, iz);
    a0 = a0 + e0 * phi;
    a1 = a1 + e1 * phi;
    a2 = a2 + e2 * phi;
  }
  sample_reduce(a0, a1, a2, sl, partial);
}

// One workgroup: the six sums over the per-block partials side by side (one reduction),
// the 24 corner reads of dF issued together (no real-space correction when dF is null).
template <class TF>
__global__ void __launch_bounds__(256) k_ksample_final(pss_grid g, const double* __restrict__ partial, int nblk,
                                                       const TF* __restrict__ dFx, const TF* __restrict__ dFy,
                                                       const TF* __restrict__ dFz, const int32_t* __restrict__ idx,
                                                       const double* __restrict__ w, double* out,
                                                       double* clock = nullptr, double tick = 0.0,
                                                       double* lines = nullptr, int line_n = 0, int nlines = 0) {
  __shared__ double red[6][4];
  __shared__ double sv[6];
  __shared__ double dfw[3][8];  // w_c dF_comp(x_c), 0 for unused corners
  const double invn = 1.0 / ((double)g.Nx * g.Ny * g.Nz);
  if (threadIdx.x < 24) {
    const int c = threadIdx.x & 7, comp = threadIdx.x >> 3;
    const TF* dF = comp == 0 ? dFx : comp == 1 ? dFy : dFz;
    const int32_t o = idx[c];
    dfw[comp][c] = o >= 0 && dF ? w[c] * (double)dF[o].x : 0.0;
  }
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += partial[(int64_t)b * 6 + q];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_down(acc[q], off, 64);
    if (lane == 0) red[q][wid] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    double r = (((red[q][0] + red[q][1]) + red[q][2]) + red[q][3]) * invn;
    if ((q & 1) == 0)  // dF is real: the sampled correction adds to the real parts
      for (int c = 0; c < 8; ++c) r += dfw[q >> 1][c];
    out[q] = r;
    sv[q] = r;
  }
  if (clock && threadIdx.x == 0) *clock += tick;  // device clock of the soft source, if given
  if (lines) {  // uniform wire lines [nlines][3][line_n] = the sample (as k_fill_line)
    __syncthreads();
    const int64_t m = (int64_t)nlines * 3 * line_n;
    for (int64_t i = threadIdx.x; i < m; i += 256) {
      const int f = (int)((i / line_n) % 3);
      lines[2 * i] = sv[2 * f];
      lines[2 * i + 1] = sv[2 * f + 1];
    }
  }
}

// Uniform wire line: Exx[f][x] = sample[f] for all x (QuantumWire, PSTD3D/rhoPJ.py:1160-1164).
__global__ void k_fill_line(const double* sample, double* lines, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int f = 0; f < 3; ++f) {
    lines[2 * ((int64_t)f * n + i)] = sample[2 * f];
    lines[2 * ((int64_t)f * n + i) + 1] = sample[2 * f + 1];
  }
}

// J line from the wires' polarisation and charge (PSTD3D/rhoPJ.py:1183-1194, CalcJfx :561-592):
// J_c(x) = sum_w (P_cw - Pold_cw)/dt / Nw, J_x += CalcJfx(sum_w Rho_w, sum_w RhoOld_w)/Nw;
// then Pold <- P, RhoOld <- Rho.  P: [Nw][3][Nx], Rho: [Nw][Nx] complex.  One block.
// keep_old = 0: Pold and RhoOld enter as zero (the reference re-creates them as zero
// arrays on every call, PSTD3D/rhoPJ.py:1138-1147, so its J is sum_w P_cw/dt/Nw); they
// are still written.  Subtracting an exact zero leaves the arithmetic of each term as is.
__global__ void __launch_bounds__(256) k_jline(const cplx* P, const cplx* Rho, cplx* Pold, cplx* RhoOld, int Nw,
                                               int Nx, double dt, double dx, cplx* jl, int fft, int keep_old) {
  __shared__ cplx drho[2048];
  for (int x = threadIdx.x; x < Nx; x += 256) {
    for (int c = 0; c < 3; ++c) {
      cplx s = mk(0, 0);
      for (int w = 0; w < Nw; ++w) {
        const int64_t o = ((int64_t)w * 3 + c) * Nx + x;
        s = s + (P[o] - (keep_old ? Pold[o] : mk(0, 0)));
        Pold[o] = P[o];
      }
      jl[c * Nx + x] = mk(s.x / dt / Nw, s.y / dt / Nw);
    }
    cplx rn = mk(0, 0), ro = mk(0, 0);
    for (int w = 0; w < Nw; ++w) {
      const int64_t o = (int64_t)w * Nx + x;
      rn = rn + Rho[o];
      if (keep_old) ro = ro + RhoOld[o];
      RhoOld[o] = Rho[o];
    }
    const cplx dr = rn - ro;
    drho[x] = mk(dr.x / dt, dr.y / dt);
  }
  __syncthreads();
  // CalcJfx: jl_x[x] += -dx sum_{x'=1..x} drho[x'] / Nw for x >= 1, as a block scan: each
  // thread sums a contiguous run, the runs' sums are scanned across the block (wave
  // shuffles, then the 4 wave totals), and each thread replays its run from its offset
  __shared__ cplx wsum[4];
  const int per = (Nx + 255) / 256;
  const int x0 = threadIdx.x * per, x1 = min(Nx, x0 + per);
  cplx loc = mk(0, 0);
  for (int x = max(x0, 1); x < x1; ++x) loc = loc + drho[x];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  cplx inc = loc;  // inclusive scan of the run sums within the wave
  for (int off = 1; off < 64; off <<= 1) {
    const double ur = __shfl_up(inc.x, off, 64), ui = __shfl_up(inc.y, off, 64);
    if (lane >= off) inc = inc + mk(ur, ui);
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  cplx pre = inc - loc;  // exclusive prefix within the wave
  for (int w = 0; w < wid; ++w) pre = pre + wsum[w];
  for (int x = max(x0, 1); x < x1; ++x) {
    pre = pre + drho[x];
    jl[x] = jl[x] + mk(-(pre.x * dx) / Nw, -(pre.y * dx) / Nw);
  }
  if (!fft || Nx == 1) return;
  // forward FFT of each line (numpy convention, unnormalised; Nx a power of two) in LDS:
  // bit-reversed load, then radix-2 stages, twiddles exp(-2 pi i k / len) = cospi/sinpi
  // of -k / (len / 2) (exact arguments)
  const int lg = __builtin_ctz(Nx);
  for (int c = 0; c < 3; ++c) {
    __syncthreads();  // the line is written and the LDS buffer is free
    for (int x = threadIdx.x; x < Nx; x += 256) drho[__brev((unsigned)x) >> (32 - lg)] = jl[c * Nx + x];
    __syncthreads();
    for (int h = 1; h < Nx; h <<= 1) {
      for (int t = threadIdx.x; t < Nx / 2; t += 256) {
        const int k = t & (h - 1), i = 2 * (t - k) + k;  // butterfly k of group t / h
        double sn, cs;
        sincospi(-(double)k / h, &sn, &cs);
        const cplx u = drho[i], v = drho[i + h] * mk(cs, sn);
        drho[i] = u + v;
        drho[i + h] = u - v;
      }
      __syncthreads();
    }
    for (int x = threadIdx.x; x < Nx; x += 256) jl[c * Nx + x] = drho[x];
  }
}

// ---- longitudinal fields and bound charge (diagnostics of QuantumWire)
// ElongfromRho's k-space loop (PSTD3D/rhoPJ.py:596-700), one cell per thread, on the
// transforms of the free charge and of P:  s = |q|^2 + small,
//   E from rho = -i rho ieps / s q,   E from P = -(q.P) ieps / s q,   E from both = their sum q,
//   RhoB = i (q.E_both) / ieps - rho   (Gauss's law),  ieps = 1 / (eps0 eps_r).
struct ElongOut {
  cplx* e[10];  // ExPR, EyPR, EzPR, ExP, EyP, EzP, ExR, EyR, EzR, RhoB
};

__global__ void __launch_bounds__(256) k_elong(pss_grid g, const cplx* __restrict__ Rk, const cplx* __restrict__ Px,
                                               const cplx* __restrict__ Py, const cplx* __restrict__ Pz, ElongOut o,
                                               double small, double ieps) {
  const int Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)g.Nx * Ny * Nz;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int k = (int)(c % Nz), j = (int)((c / Nz) % Ny), i = (int)(c / ((int64_t)Ny * Nz));
  const double qx = g.kx[i], qy = g.ky[j], qz = g.kz[k];
  const double qs = qx * qx + qy * qy + qz * qz + small;
  const cplx R = Rk[c], X = Px[c], Y = Py[c], Z = Pz[c];
  const cplx qP = qx * X + qy * Y + qz * Z;
  const cplx rt = mk(R.y * ieps / qs, -R.x * ieps / qs);     // -i rho ieps / s
  const cplx pt = mk(-qP.x * ieps / qs, -qP.y * ieps / qs);  // -(q.P) ieps / s
  const cplx sm = rt + pt;
  const cplx ex = sm * qx, ey = sm * qy, ez = sm * qz;
  o.e[0][c] = ex;
  o.e[1][c] = ey;
  o.e[2][c] = ez;
  o.e[3][c] = pt * qx;
  o.e[4][c] = pt * qy;
  o.e[5][c] = pt * qz;
  o.e[6][c] = rt * qx;
  o.e[7][c] = rt * qy;
  o.e[8][c] = rt * qz;
  const cplx qE = qx * ex + qy * ey + qz * ez;
  o.e[9][c] = mk(-qE.y / ieps - R.x, qE.x / ieps - R.y);  // i (q.E) / ieps - rho
}

// QWPlacement (PSTD3D/rhoPJ.py:736-781) of wire lines, up to three placed terms summed in
// order: out = (((s0 T0) + s1 T1) + s2 T2), T_t[i][j][k] = (gx[i] gy[j]) gz[k] F_t(i) with F_t
// the mean of nw lines (lines + w lstride) at i, or for mode 1 its periodic central
// difference (F(i+1) - F(i-1)) / (2 dx) (the reference's np.roll difference); gx null = 1.
struct PlaceTerm {
  const cplx* lines;
  const cplx *gx, *gy, *gz;
  int64_t lstride;
  double sign;
  int mode;
};
struct PlaceTerms {
  PlaceTerm t[3];
};

__device__ inline cplx line_mean(const PlaceTerm& t, int nw, int i) {
  cplx s = mk(0, 0);
  for (int w = 0; w < nw; ++w) s = s + t.lines[(int64_t)w * t.lstride + i];
  return mk(s.x / (double)nw, s.y / (double)nw);
}

__global__ void __launch_bounds__(256) k_qwplace(int Nx, int Ny, int Nz, int nw, PlaceTerms pt, int nterms, double dx,
                                                 cplx* __restrict__ out) {
  const int64_t n = (int64_t)Nx * Ny * Nz;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int k = (int)(c % Nz), j = (int)((c / Nz) % Ny), i = (int)(c / ((int64_t)Ny * Nz));
  cplx acc = mk(0, 0);
  for (int q = 0; q < nterms; ++q) {
    const PlaceTerm& t = pt.t[q];
    cplx F;
    if (t.mode == 1) {
      const cplx a = line_mean(t, nw, i + 1 < Nx ? i + 1 : 0), b = line_mean(t, nw, i > 0 ? i - 1 : Nx - 1);
      const cplx df = a - b;
      F = mk(df.x / (2.0 * dx), df.y / (2.0 * dx));
    } else {
      F = line_mean(t, nw, i);
    }
    const cplx gxy = t.gx ? t.gx[i] * t.gy[j] : t.gy[j];
    const cplx v = (gxy * t.gz[k]) * F;
    acc = q == 0 ? t.sign * v : acc + t.sign * v;
  }
  out[c] = acc;
}

// F *= exp(-(sx^2 qx^2 + sy^2 qy^2 + sz^2 qz^2) / 2): the Gaussian k-space filter of the
// smoothed bound-charge diagnostics (PSTD3D/rhoPJ.py:1263-1285).
__global__ void __launch_bounds__(256) k_kfilter(pss_grid g, cplx* __restrict__ F, double sx2, double sy2, double sz2) {
  const int Ny = g.Ny, Nz = g.Nz;
  const int64_t n = (int64_t)g.Nx * Ny * Nz;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int k = (int)(c % Nz), j = (int)((c / Nz) % Ny), i = (int)(c / ((int64_t)Ny * Nz));
  const double qx = g.kx[i], qy = g.ky[j], qz = g.kz[k];
  F[c] = F[c] * exp(-0.5 * (sx2 * (qx * qx) + sy2 * (qy * qy) + sz2 * (qz * qz)));
}

}  // namespace

// ================================================================== C ABI
static bool pow2_axes(const pss_grid* g) {
  for (int N : {g->Nx, g->Ny, g->Nz})
    if (N < 1 || (N & (N - 1))) return false;
  return true;
}

extern "C" int pss_update_e3d(const pss_grid* g, const double* Bx, const double* By, const double* Bz,
                              const double* Jx, const double* Jy, const double* Jz, const double* jline,
                              const double* gy, const double* gz, double* Ex, double* Ey, double* Ez, void* stream) {
  if (!pow2_axes(g)) return pss_set_msg(2, "pss_update_e3d: axis lengths must be powers of two");
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  hipLaunchKernelGGL(k_update_e, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *g,
                     (const cplx*)Bx, (const cplx*)By, (const cplx*)Bz, (const cplx*)Jx, (const cplx*)Jy,
                     (const cplx*)Jz, (const cplx*)jline, (const cplx*)gy, (const cplx*)gz, (cplx*)Ex, (cplx*)Ey,
                     (cplx*)Ez);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_update_e3d_sample(const pss_grid* g, const double* Bx, const double* By, const double* Bz,
                                     const double* Jx, const double* Jy, const double* Jz, const double* jline,
                                     const double* gy, const double* gz, double* Ex, double* Ey, double* Ez,
                                     const int32_t* sample_idx, const double* sample_w, double* partial,
                                     double* sample_out, double* lines, int32_t line_n, int32_t nlines,
                                     double* clock, double tick, int32_t update_b, void* stream) {
  if (!pow2_axes(g)) return pss_set_msg(2, "pss_update_e3d_sample: axis lengths must be powers of two");
  if (lines && (line_n < 1 || nlines < 1)) return pss_set_msg(2, "pss_update_e3d_sample: line_n, nlines >= 1");
  const bool curl = Bx != nullptr;
  if ((By != nullptr) != curl || (Bz != nullptr) != curl)
    return pss_set_msg(2, "pss_update_e3d_sample: B is given in full (curl term) or not at all (current only)");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  const int nblk = (int)std::min<int64_t>(1024, (n + 255) / 256);
  const size_t lds = (size_t)(std::max(g->Nx, std::max(g->Ny, g->Nz)) / 2 + 1) * sizeof(cplx);
  if (update_b && !curl) return pss_set_msg(2, "pss_update_e3d_sample: update_b needs B");
  auto kern = update_b ? k_update_e_sample<true, true> : curl ? k_update_e_sample<true, false>
                                                             : k_update_e_sample<false, false>;
  cplx* bw[3] = {nullptr, nullptr, nullptr};
  if (update_b) {
    bw[0] = (cplx*)Bx; bw[1] = (cplx*)By; bw[2] = (cplx*)Bz;
  }
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), lds, s, *g, (const cplx*)Bx, (const cplx*)By,
                     (const cplx*)Bz, (const cplx*)Jx, (const cplx*)Jy, (const cplx*)Jz, (const cplx*)jline,
                     (const cplx*)gy, (const cplx*)gz, (cplx*)Ex, (cplx*)Ey, (cplx*)Ez, sample_idx, sample_w, partial,
                     bw[0], bw[1], bw[2]);
  hipLaunchKernelGGL((k_ksample_final<cplx>), dim3(1), dim3(256), 0, s, *g, partial, nblk, (const cplx*)nullptr,
                     (const cplx*)nullptr, (const cplx*)nullptr, sample_idx, sample_w, sample_out, clock, tick,
                     lines, line_n, nlines);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_update_b3d(const pss_grid* g, const double* Ex, const double* Ey, const double* Ez, double* Bx,
                              double* By, double* Bz, void* stream) {
  if (!pow2_axes(g)) return pss_set_msg(2, "pss_update_b3d: axis lengths must be powers of two");
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  hipLaunchKernelGGL(k_update_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *g,
                     (const cplx*)Ex, (const cplx*)Ey, (const cplx*)Ez, (cplx*)Bx, (cplx*)By, (cplx*)Bz);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_fft3d(const pss_grid* g, double* F, int32_t dir, void* stream) {
  int rc = fft3(g, (const cplx*)F, (cplx*)F, dir, nullptr, -1, (hipStream_t)stream);
  if (rc) return pss_set_msg(2, "pss_fft3d: axis length must be a power of two <= 2048");
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_fft_line(const pss_grid* g, const double* in, double* out, int32_t axis, int32_t dir,
                            void* stream) {
  int rc = launch_pass(g, (const cplx*)in, (cplx*)out, axis, dir, nullptr, -1, (hipStream_t)stream);
  if (rc) return pss_set_msg(2, "pss_fft_line: bad axis length");
  PSS_CHECK(hipGetLastError());
  return 0;
}

static int ksample_launch(const pss_grid* g, const double* Fx, const double* Fy, const double* Fz,
                          const int32_t* sample_idx, const double* sample_w, double* partial, double* sample_out,
                          hipStream_t s) {
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  const int nblk = (int)std::min<int64_t>(1024, (n + 255) / 256);
  const size_t lds = (size_t)(std::max(g->Nx, std::max(g->Ny, g->Nz)) / 2 + 1) * sizeof(cplx);
  hipLaunchKernelGGL(k_ksample, dim3(nblk), dim3(256), lds, s, *g, (const cplx*)Fx, (const cplx*)Fy, (const cplx*)Fz,
                     sample_idx, sample_w, partial);
  hipLaunchKernelGGL((k_ksample_final<cplx>), dim3(1), dim3(256), 0, s, *g, partial, nblk, (const cplx*)nullptr,
                     (const cplx*)nullptr, (const cplx*)nullptr, sample_idx, sample_w, sample_out);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_ksample(const pss_grid* g, const double* Fx, const double* Fy, const double* Fz,
                           const int32_t* sample_idx, const double* sample_w, double* partial, double* sample_out,
                           void* stream) {
  if (!pow2_axes(g) || g->Nx > 2048 || g->Ny > 2048 || g->Nz > 2048)
    return pss_set_msg(2, "pss_ksample: axis lengths must be powers of two <= 2048");
  return ksample_launch(g, Fx, Fy, Fz, sample_idx, sample_w, partial, sample_out, (hipStream_t)stream);
}

// ---- the point sample without a caller-provided partial-sum buffer
// The per-block partial sums of k_ksample (at most 1024 blocks x 6 doubles, whatever the
// grid) live in this array: one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in symbol); it uses the
// device address that hipGetSymbolAddress returns for the current GPU, looked up once per GPU
// on that GPU's first call.  Every call on a GPU shares that GPU's copy: calls that overlap
// on one GPU (two streams, two threads) race on it.
#define PSS_KSAMPLE_SCRATCH_DOUBLES 6144
static __device__ __align__(256) double scratch[PSS_KSAMPLE_SCRATCH_DOUBLES];
static double* g_scratch_dev[64];
static std::mutex g_scratch_mu;

static int scratch_on_current_device(double** out) {
  int dev = -1;
  PSS_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return pss_set_msg(2, "pss_ksample_noalloc: device ordinal >= 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return pss_set_error(e, "pss_ksample_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) return pss_set_msg(2, "pss_ksample_noalloc: null scratch address");
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int pss_ksample_noalloc(const pss_grid* g, const double* Fx, const double* Fy, const double* Fz,
                                   const int32_t* sample_idx, const double* sample_w, double* sample_out,
                                   void* stream) {
  if (!pow2_axes(g) || g->Nx > 2048 || g->Ny > 2048 || g->Nz > 2048)
    return pss_set_msg(2, "pss_ksample_noalloc: axis lengths must be powers of two <= 2048");
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  const int64_t need = std::min<int64_t>(1024, (n + 255) / 256) * 6;
  if (need > PSS_KSAMPLE_SCRATCH_DOUBLES)
    return pss_set_msg(2, "pss_ksample_noalloc: partial sums exceed the 6144-double device scratch");
  double* part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  // every partial slot the final reduction reads is written by k_ksample first: nothing to clear
  return ksample_launch(g, Fx, Fy, Fz, sample_idx, sample_w, part, sample_out, (hipStream_t)stream);
}

extern "C" int pss_fft_line_batch(const pss_grid* g, double* const* in, double* const* out, int32_t nb,
                                  int32_t axis, int32_t dir, void* stream) {
  if (nb < 1 || nb > 3 || axis < 0 || axis > 2 || (dir != 1 && dir != -1))
    return pss_set_msg(2, "pss_fft_line_batch: 1 <= nb <= 3, axis 0..2, dir +-1");
  const cplx* ins[3] = {(const cplx*)in[0], nb > 1 ? (const cplx*)in[1] : nullptr, nb > 2 ? (const cplx*)in[2] : nullptr};
  cplx* outs[3] = {(cplx*)out[0], nb > 1 ? (cplx*)out[1] : nullptr, nb > 2 ? (cplx*)out[2] : nullptr};
  const int rc = launch_pass_b(g, ins, outs, nb, axis, dir, nullptr, -1, (hipStream_t)stream);
  if (rc) return pss_set_msg(2, "pss_fft_line_batch: bad axis length");
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_fft_pass_acc(const pss_grid* g, double* const* in, double* const* out, int32_t nb, int32_t axis,
                                int32_t dir, int32_t in32, const double* Gx, const double* Gy, const double* Gz,
                                double coef, void* stream) {
  if (nb < 1 || nb > 3 || axis < 0 || axis > 2 || (dir != 1 && dir != -1) || (in32 != 0 && in32 != 1))
    return pss_set_msg(2, "pss_fft_pass_acc: 1 <= nb <= 3, axis 0..2, dir +-1, in32 0/1");
  const int N = axis == 0 ? g->Nx : axis == 1 ? g->Ny : g->Nz;
  const bool two_level = N == 64 || N == 128 || N == 256 || N == 512 || N == 1024;
  const bool curl = Gx != nullptr;
  if (curl && (!Gy || !Gz || nb != 3))
    return pss_set_msg(2, "pss_fft_pass_acc: the curl term needs the three components of G and nb = 3");
  if ((curl || in32) && !two_level)
    return pss_set_msg(2, "pss_fft_pass_acc: complex64 input and the curl term need an axis of 64 ... 1024");
  const cplx* ins[3] = {(const cplx*)in[0], nb > 1 ? (const cplx*)in[1] : nullptr, nb > 2 ? (const cplx*)in[2] : nullptr};
  cplx* outs[3] = {(cplx*)out[0], nb > 1 ? (cplx*)out[1] : nullptr, nb > 2 ? (cplx*)out[2] : nullptr};
  const CurlArgs cu{{(const cplx*)Gx, (const cplx*)Gy, (const cplx*)Gz}, coef};
  const int rc = launch_pass_b(g, ins, outs, nb, axis, dir, nullptr, -1, (hipStream_t)stream, curl ? 3 : 1, in32, 0, 0,
                               curl ? &cu : nullptr);
  if (rc) return pss_set_msg(2, "pss_fft_pass_acc: bad axis length");
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_fft_planes_acc(const pss_grid* g, const double* const* in, double* const* out, const double* Gx,
                                  const double* Gy, const double* Gz, double coef, void* stream) {
  if (!(g->Ny == g->Nz && (g->Ny == 64 || g->Ny == 128)))
    return pss_set_msg(2, "pss_fft_planes_acc: needs Ny = Nz = 64 or 128");
  if (g->Nx < 1 || (g->Nx & (g->Nx - 1)))
    return pss_set_msg(2, "pss_fft_planes_acc: Nx must be a power of two (conjugate x-planes Nx - g)");
  if (!in || !out || !in[0] || !in[1] || !in[2] || !out[0] || !out[1] || !out[2] || !Gx || !Gy || !Gz)
    return pss_set_msg(2, "pss_fft_planes_acc: null array");
  CurlPlanes cp;
  const cplx* G[3] = {(const cplx*)Gx, (const cplx*)Gy, (const cplx*)Gz};
  for (int c = 0; c < 3; ++c) {
    cp.dF[c] = (const cflt*)in[c];
    cp.F[c] = (cplx*)out[c];
    cp.G[c] = G[c];
  }
  cp.qx = g->kx;
  cp.qy = g->ky;
  cp.qz = g->kz;
  cp.ccoef = coef;
  cp.tw = g->twy;
  cp.Nx = g->Nx;
  if (g->Ny == 128) launch_dF_planes<8, 16, 1024>(cp, (hipStream_t)stream);
  else launch_dF_planes<8, 8, 512>(cp, (hipStream_t)stream);
  PSS_CHECK(hipGetLastError());
  return 0;
}

// cu: the curl update of F from G, fused into the last forward pass (null: none)
static int cpml_half_impl(const pss_grid* g, int32_t which, double* Fx, double* Fy, double* Fz, const double* Gx,
                   const double* Gy, const double* Gz, double* psi, double* scratch, int32_t src_comp,
                   const double* src_par, const double* srcx, const double* rx, const double* ry, const double* rz,
                   const int32_t* sample_idx, const double* sample_w, double* sample_out, const CurlArgs* cu,
                   hipStream_t s) {
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  if (g->Nx > 2048 || g->Ny > 2048 || g->Nz > 2048)
    return pss_set_msg(2, "pss_cpml_half: axis length must be a power of two <= 2048");
  {
    const int len[3] = {g->Nx, g->Ny, g->Nz}, pml[3] = {g->npml_x, g->npml_y, g->npml_z};
    for (int a = 0; a < 3; ++a)
      if (pml[a] < 0 || (pml[a] > 0 && (len[a] == 1 || 2 * pml[a] > len[a])))
        return pss_set_msg(2, "pss_cpml_half: a PML needs an axis longer than 1 and 2 npml <= N");
  }
  // scratch: 6 complex fields.  Slots 0-2 hold the six real derivative arrays,
  // slots 3-4 the complex temporaries of the derivative passes, then slots 3-5 dF.
  cplx* sc = (cplx*)scratch;
  // mixed precision: derivative passes in FP32 when every axis uses the two-level kernels
  bool fp32 = g->precision == PSS_MIXED;
  for (int N : {g->Nx, g->Ny, g->Nz})
    if (N > 1 && (N < 64 || N > 1024)) fp32 = false;
  double* R = (double*)scratch;
  float* Rf = (float*)scratch;
  const cplx *gx = (const cplx*)Gx, *gy = (const cplx*)Gy, *gz = (const cplx*)Gz;
  // derivatives on the PML slabs when every axis has a two-level pass (64 ... 1024),
  // else full-grid derivative pairs.  R: 0 = dGz/dy, 1 = dGy/dz, 2 = dGx/dz, 3 = dGz/dx,
  // 4 = dGy/dx, 5 = dGx/dy
  bool slab = true;
  for (int N : {g->Nx, g->Ny, g->Nz})
    if (N > 1 && (N < 64 || N > 1024)) slab = false;
  SlabR sr{};
  psi_layout(g, sr.poff);
  int rc = 0;
  if (slab) {
    const cplx* G3[3] = {gx, gy, gz};
    rc = slab_derivs(g, G3, scratch, sc + 3 * n, fp32, sr, s);
  } else {
    sr.full = 1;
    for (int q = 0; q < 6; ++q) {
      sr.off[q] = q * n;
      sr.on[q] = 1;
    }
    auto rq = [&](int q) -> void* { return fp32 ? (void*)(Rf + q * n) : (void*)(R + q * n); };
    rc = deriv_pair(g, gz, 2, 1, 0, rq(0), rq(3), sc + 3 * n, sc + 4 * n, fp32, s);
    if (!rc) rc = deriv_pair(g, gy, 1, 2, 0, rq(1), rq(4), sc + 3 * n, sc + 4 * n, fp32, s);
    if (!rc) rc = deriv_pair(g, gx, 0, 2, 1, rq(2), rq(5), sc + 3 * n, sc + 4 * n, fp32, s);
  }
  if (rc) return pss_set_msg(2, "pss_cpml_half: axis length must be a power of two <= 2048");
  SrcPar sp{};
  if (src_comp >= 0) {
    sp.t = src_par[0]; sp.amp = src_par[1]; sp.omega0 = src_par[2]; sp.tau_G = src_par[3];
    sp.chirp = src_par[4]; sp.v = src_par[5]; sp.Tp = src_par[6]; sp.w0inv2 = src_par[7]; sp.dx = src_par[8];
  }
  cplx* dF = sc + 3 * n;       // complex128 dF (FP64), or
  cflt* dFf = (cflt*)dF;       // complex64 dF (mixed: transformed in FP32, added to the FP64 fields)
  {
    const dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    const bool rows = g->Nz >= 64;
    if (fp32) {
      if (rows)
        hipLaunchKernelGGL((k_cpml<float, cflt, true>), gr, bl, 0, s, *g, which, Rf, sr, dFf, dFf + n, dFf + 2 * n,
                           psi, src_comp, sp, srcx, rx, ry, rz);
      else
        hipLaunchKernelGGL((k_cpml<float, cflt, false>), gr, bl, 0, s, *g, which, Rf, sr, dFf, dFf + n, dFf + 2 * n,
                           psi, src_comp, sp, srcx, rx, ry, rz);
    } else {
      if (rows)
        hipLaunchKernelGGL((k_cpml<double, cplx, true>), gr, bl, 0, s, *g, which, R, sr, dF, dF + n, dF + 2 * n,
                           psi, src_comp, sp, srcx, rx, ry, rz);
      else
        hipLaunchKernelGGL((k_cpml<double, cplx, false>), gr, bl, 0, s, *g, which, R, sr, dF, dF + n, dF + 2 * n,
                           psi, src_comp, sp, srcx, rx, ry, rz);
    }
  }
  cplx* F[3] = {(cplx*)Fx, (cplx*)Fy, (cplx*)Fz};
  if (sample_idx) {
    // partial sums live in the (now free) derivative slots
    const int nblk = (int)std::min<int64_t>(1024, (n + 255) / 256);
    double* part = R;
    const size_t lds = (size_t)(std::max(g->Nx, std::max(g->Ny, g->Nz)) / 2 + 1) * sizeof(cplx);
    hipLaunchKernelGGL(k_ksample, dim3(nblk), dim3(256), lds, s, *g, F[0], F[1], F[2], sample_idx, sample_w, part);
    if (fp32)
      hipLaunchKernelGGL((k_ksample_final<cflt>), dim3(1), dim3(256), 0, s, *g, part, nblk, dFf, dFf + n,
                         dFf + 2 * n, sample_idx, sample_w, sample_out);
    else
      hipLaunchKernelGGL((k_ksample_final<cplx>), dim3(1), dim3(256), 0, s, *g, part, nblk, dF, dF + n,
                         dF + 2 * n, sample_idx, sample_w, sample_out);
  }
  {
    cplx* tmps[3];
    for (int c = 0; c < 3; ++c) tmps[c] = fp32 ? (cplx*)(dFf + c * n) : dF + c * n;
    rc = fft3_acc(g, tmps, F, 3, -1, s, fp32 ? 1 : 0, cu);
    if (rc) return pss_set_msg(2, "pss_cpml_half: fft");
  }
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_cpml_half(const pss_grid* g, int32_t which, double* Fx, double* Fy, double* Fz, const double* Gx,
                             const double* Gy, const double* Gz, double* psi, double* scratch, int32_t src_comp,
                             const double* src_par, const double* srcx, const double* rx, const double* ry,
                             const double* rz, const int32_t* sample_idx, const double* sample_w,
                             double* sample_out, void* stream) {
  return cpml_half_impl(g, which, Fx, Fy, Fz, Gx, Gy, Gz, psi, scratch, src_comp, src_par, srcx, rx, ry, rz,
                        sample_idx, sample_w, sample_out, nullptr, (hipStream_t)stream);
}

extern "C" int pss_mask_half(const pss_grid* g, double* Fx, double* Fy, double* Fz, const double* mx, const double* my,
                             const double* mz, int32_t src_comp, const double* src_par, const double* srcx,
                             const double* rx, const double* ry, const double* rz, void* stream) {
  if (!pow2_axes(g) || g->Nx > 2048 || g->Ny > 2048 || g->Nz > 2048)
    return pss_set_msg(2, "pss_mask_half: axis lengths must be powers of two <= 2048");
  if (!Fx || !Fy || !Fz || !mx || !my || !mz) return pss_set_msg(2, "pss_mask_half: null field or mask");
  if (src_comp > 2 || (src_comp >= 0 && (!src_par || !srcx || !rx || !ry || !rz)))
    return pss_set_msg(2, "pss_mask_half: src_comp 0..2 needs its parameters and coordinates");
  hipStream_t s = (hipStream_t)stream;
  cplx* F[3] = {(cplx*)Fx, (cplx*)Fy, (cplx*)Fz};
  const cplx* Fc[3] = {F[0], F[1], F[2]};
  // to real space (inverse passes, 1/N overall), mask and source, back to k-space
  for (int axis = 0; axis < 3; ++axis)
    if (launch_pass_b(g, Fc, F, 3, axis, +1, nullptr, -1, s)) return pss_set_msg(2, "pss_mask_half: fft");
  SrcPar sp{};
  if (src_comp >= 0) {
    sp.t = src_par[0]; sp.amp = src_par[1]; sp.omega0 = src_par[2]; sp.tau_G = src_par[3];
    sp.chirp = src_par[4]; sp.v = src_par[5]; sp.Tp = src_par[6]; sp.w0inv2 = src_par[7]; sp.dx = src_par[8];
  }
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  hipLaunchKernelGGL(k_mask_src, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, *g, F[0], F[1], F[2], mx, my, mz,
                     src_comp, sp, srcx, rx, ry, rz);
  for (int axis = 0; axis < 3; ++axis)
    if (launch_pass_b(g, Fc, F, 3, axis, -1, nullptr, -1, s)) return pss_set_msg(2, "pss_mask_half: fft");
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_b_half(const pss_grid* g, double* Bx, double* By, double* Bz, const double* Ex, const double* Ey,
                          const double* Ez, double* psi, double* scratch, int32_t src_comp, const double* src_par,
                          const double* srcx, const double* rx, const double* ry, const double* rz, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int last = 0;
  for (int a = 0; a < 3; ++a)
    if ((a == 0 ? g->Nx : a == 1 ? g->Ny : g->Nz) > 1) last = a;
  const int N = last == 0 ? g->Nx : last == 1 ? g->Ny : g->Nz;
  // fused in mixed precision only: with the FP64 correction (complex128 dF) the fused last
  // pass cost more than the separate B update (128^3 pipelined step: FP64 0.806 ms fused,
  // 0.792 ms separate; mixed 0.578 ms fused, 0.587 ms separate)
  const bool fuse = g->precision == PSS_MIXED && (N == 64 || N == 128 || N == 256 || N == 512 || N == 1024);
  if (!fuse) {
    const int rc = pss_update_b3d(g, Ex, Ey, Ez, Bx, By, Bz, stream);
    if (rc) return rc;
    return cpml_half_impl(g, 1, Bx, By, Bz, Ex, Ey, Ez, psi, scratch, src_comp, src_par, srcx, rx, ry, rz, nullptr,
                          nullptr, nullptr, nullptr, s);
  }
  const CurlArgs cu{{(const cplx*)Ex, (const cplx*)Ey, (const cplx*)Ez}, -g->dt};
  return cpml_half_impl(g, 1, Bx, By, Bz, Ex, Ey, Ez, psi, scratch, src_comp, src_par, srcx, rx, ry, rz, nullptr,
                        nullptr, nullptr, &cu, s);
}

extern "C" int pss_cpml_e_curl(const pss_grid* g, double* Ex, double* Ey, double* Ez, const double* Bx,
                               const double* By, const double* Bz, double* psi, double* scratch, int32_t src_comp,
                               const double* src_par, const double* srcx, const double* rx, const double* ry,
                               const double* rz, void* stream) {
  if (!Bx || !By || !Bz) return pss_set_msg(2, "pss_cpml_e_curl: null B");
  hipStream_t s = (hipStream_t)stream;
  int last = 0;
  for (int a = 0; a < 3; ++a)
    if ((a == 0 ? g->Nx : a == 1 ? g->Ny : g->Nz) > 1) last = a;
  const int N = last == 0 ? g->Nx : last == 1 ? g->Ny : g->Nz;
  const bool fuse = N == 64 || N == 128 || N == 256 || N == 512 || N == 1024;
  if (!fuse) {
    // the curl term alone (no current), then the correction
    const int rc = pss_update_e3d(g, Bx, By, Bz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, Ex, Ey, Ez,
                                  stream);
    if (rc) return rc;
    return cpml_half_impl(g, 0, Ex, Ey, Ez, Bx, By, Bz, psi, scratch, src_comp, src_par, srcx, rx, ry, rz, nullptr,
                          nullptr, nullptr, nullptr, s);
  }
  const CurlArgs cu{{(const cplx*)Bx, (const cplx*)By, (const cplx*)Bz}, g->v2 * g->dt};
  return cpml_half_impl(g, 0, Ex, Ey, Ez, Bx, By, Bz, psi, scratch, src_comp, src_par, srcx, rx, ry, rz, nullptr,
                        nullptr, nullptr, &cu, s);
}

__global__ void k_add_scalar(double* x, double v) {
  if (threadIdx.x == 0) x[0] += v;
}

extern "C" int64_t pss_cpml_psi_len(const pss_grid* g) {
  int64_t poff[6];
  return psi_layout(g, poff);
}

extern "C" int pss_add_scalar(double* x, double v, void* stream) {
  if (!x) return pss_set_msg(2, "pss_add_scalar: null pointer");
  hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(64), 0, (hipStream_t)stream, x, v);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_fill_line(const double* sample, double* lines, int32_t n, void* stream) {
  hipLaunchKernelGGL(k_fill_line, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, sample, lines, n);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_jline(const double* P, const double* Rho, double* Pold, double* RhoOld, int32_t Nw, int32_t Nx,
                         double dt, double dx, double* jline, int32_t fft, int32_t keep_old, void* stream) {
  if (Nx < 1 || Nx > 2048) return pss_set_msg(2, "pss_jline: 1 <= Nx <= 2048");
  if (Nw < 1) return pss_set_msg(2, "pss_jline: Nw >= 1");
  if (fft && (Nx & (Nx - 1))) return pss_set_msg(2, "pss_jline: the FFT of the line needs a power-of-two Nx");
  hipLaunchKernelGGL(k_jline, dim3(1), dim3(256), 0, (hipStream_t)stream, (const cplx*)P, (const cplx*)Rho,
                     (cplx*)Pold, (cplx*)RhoOld, Nw, Nx, dt, dx, (cplx*)jline, (int)(fft != 0), (int)(keep_old != 0));
  PSS_CHECK(hipGetLastError());
  return 0;
}

// ---- launch timing of the fused last pass (diagnostics; no counterpart in the reference)
extern "C" int pss_pass_timing_arm(int32_t n) {
  if (n < 0 || n > 4096) return pss_set_msg(2, "pss_pass_timing_arm: 0 <= n <= 4096");
  std::lock_guard<std::mutex> lock(g_tev_mu);
  while ((int)g_tev.size() < 2 * n) {
    hipEvent_t e;
    PSS_CHECK(hipEventCreate(&e));
    g_tev.push_back(e);
  }
  g_tev_armed = n;
  g_tev_next = 0;
  return 0;
}

extern "C" int pss_pass_timing_read(float* ms, int32_t n, int32_t* got) {
  if (!ms || !got || n < 0) return pss_set_msg(2, "pss_pass_timing_read: null output or n < 0");
  std::lock_guard<std::mutex> lock(g_tev_mu);
  const int k = std::min(n, g_tev_next);
  for (int i = 0; i < k; ++i) PSS_CHECK(hipEventElapsedTime(&ms[i], g_tev[2 * i], g_tev[2 * i + 1]));
  *got = k;
  return 0;
}


// ---- longitudinal-field diagnostics (no counterpart entry in the reference; ElongfromRho's loop)
extern "C" int pss_elong_k(const pss_grid* g, const double* Rho_k, const double* Px, const double* Py,
                           const double* Pz, double* const* out, double small, double inv_eps, void* stream) {
  if (!g || !Rho_k || !Px || !Py || !Pz || !out) return pss_set_msg(2, "pss_elong_k: null argument");
  ElongOut o;
  for (int q = 0; q < 10; ++q) {
    if (!out[q]) return pss_set_msg(2, "pss_elong_k: null output");
    o.e[q] = (cplx*)out[q];
  }
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  hipLaunchKernelGGL(k_elong, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *g,
                     (const cplx*)Rho_k, (const cplx*)Px, (const cplx*)Py, (const cplx*)Pz, o, small, inv_eps);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_qw_place(int32_t Nx, int32_t Ny, int32_t Nz, int32_t nw, int32_t nterms, const double* const* lines,
                            const int64_t* lstride, const double* const* gx, const double* const* gy,
                            const double* const* gz, const double* sign, const int32_t* mode, double dx, double* out,
                            void* stream) {
  if (Nx < 1 || Ny < 1 || Nz < 1 || nw < 1 || nterms < 1 || nterms > 3 || !lines || !lstride || !gy || !gz ||
      !sign || !mode || !out)
    return pss_set_msg(2, "pss_qw_place: bad sizes or null argument (1 <= nterms <= 3)");
  PlaceTerms pt{};
  for (int q = 0; q < nterms; ++q) {
    if (!lines[q] || !gy[q] || !gz[q] || (mode[q] != 0 && mode[q] != 1))
      return pss_set_msg(2, "pss_qw_place: null line or gate, or mode not 0/1");
    pt.t[q] = PlaceTerm{(const cplx*)lines[q], gx ? (const cplx*)gx[q] : nullptr, (const cplx*)gy[q],
                        (const cplx*)gz[q], lstride[q], sign[q], mode[q]};
  }
  const int64_t n = (int64_t)Nx * Ny * Nz;
  hipLaunchKernelGGL(k_qwplace, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, Nx, Ny, Nz, nw,
                     pt, nterms, dx, (cplx*)out);
  PSS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int pss_kfilter(const pss_grid* g, double* F, double sx, double sy, double sz, void* stream) {
  if (!g || !F) return pss_set_msg(2, "pss_kfilter: null argument");
  const int64_t n = (int64_t)g->Nx * g->Ny * g->Nz;
  hipLaunchKernelGGL(k_kfilter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *g, (cplx*)F,
                     sx * sx, sy * sy, sz * sz);
  PSS_CHECK(hipGetLastError());
  return 0;
}
