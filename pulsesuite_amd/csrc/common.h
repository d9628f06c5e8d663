// Shared device helpers for the pss HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PSS_INLINE __device__ __forceinline__

struct cplx {
  double x, y;
};

PSS_INLINE cplx mk(double a, double b) { return cplx{a, b}; }
PSS_INLINE cplx operator+(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
PSS_INLINE cplx operator-(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
PSS_INLINE cplx operator*(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
PSS_INLINE cplx operator*(double s, cplx a) { return {s * a.x, s * a.y}; }
PSS_INLINE cplx operator*(cplx a, double s) { return {s * a.x, s * a.y}; }
PSS_INLINE cplx conj(cplx a) { return {a.x, -a.y}; }
PSS_INLINE double cabs(cplx a) { return hypot(a.x, a.y); }
// a / (i*h) for real h
PSS_INLINE cplx div_ih(cplx a, double h) { return {a.y / h, -a.x / h}; }

// Block-wide sum of doubles (blockDim.x a multiple of 64, <= 1024).
template <int NT>
PSS_INLINE double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; ++i) r += sh[i];
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// K sums at once, each by the same tree as block_sum (bit-identical results), one set of
// barriers; sh holds K * NT / 64 doubles.
template <int NT, int K>
PSS_INLINE void block_sum_n(double* v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) sh[q * (NT / 64) + wid] = v[q];
  __syncthreads();
  if (threadIdx.x < K) {
    double r = 0.0;
    for (int i = 0; i < NT / 64; ++i) r += sh[threadIdx.x * (NT / 64) + i];
    sh[K * (NT / 64) + threadIdx.x] = r;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = sh[K * (NT / 64) + q];
  __syncthreads();
}

template <int NT>
PSS_INLINE double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh[i]);
    sh[0] = r;
  }
  __syncthreads();
  double r = sh[0];
  __syncthreads();
  return r;
}

// K maxima at once (max is exact in any order), one set of barriers; sh holds
// K * NT / 64 + K doubles.
template <int NT, int K>
PSS_INLINE void block_max_n(double* v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] = fmax(v[q], __shfl_down(v[q], o, 64));
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) sh[q * (NT / 64) + wid] = v[q];
  __syncthreads();
  if (threadIdx.x < K) {
    double r = sh[threadIdx.x * (NT / 64)];
    for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh[threadIdx.x * (NT / 64) + i]);
    sh[K * (NT / 64) + threadIdx.x] = r;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = sh[K * (NT / 64) + q];
  __syncthreads();
}

#define PSS_CHECK(expr)                                      \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return pss_set_error(_e, #expr);   \
  } while (0)

int pss_set_error(hipError_t e, const char* what);
int pss_set_msg(int code, const char* msg);
