// Classic RK4 stage combines, one fused HBM pass per stage.
//
// libpulsesuite/integrator.py:143-157 (_rk4, Numerical Recipes stage order):
//   ytmp = y + (h/2) dydt ; dyt = D(t + h/2, ytmp)
//   ytmp = y + (h/2) dyt  ; dym = D(t + h/2, ytmp)
//   ytmp = y + h dym ; dym = dym + dyt ; dyt = D(t + h, ytmp)
//   y    = y + (h/6) (dydt + dyt + 2 dym)
// The coefficients are real, so a complex array is processed as 2n doubles.
// Every operation is a separately rounded IEEE multiply or add in the order of
// the expressions above (FMA contraction switched off for this kernel), so the result is bit-identical
// to the same expressions evaluated elementwise in FP64.
#include <algorithm>

#include "common.h"
#include "../../include/pss.h"

namespace {

__global__ void __launch_bounds__(256) k_rk4_stage(int stage, int64_t n, double h, double* __restrict__ y,
                                                   const double* __restrict__ dydt, const double* __restrict__ dyt,
                                                   double* __restrict__ dym, double* __restrict__ ytmp) {
#pragma clang fp contract(off)
  const double h2 = h / 2.0, h6 = h / 6.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    switch (stage) {
      case 0: ytmp[i] = y[i] + h2 * dydt[i]; break;
      case 1: ytmp[i] = y[i] + h2 * dyt[i]; break;
      case 2: {
        const double m = dym[i];
        ytmp[i] = y[i] + h * m;
        dym[i] = m + dyt[i];
        break;
      }
      default: {
        const double s = (dydt[i] + dyt[i]) + 2.0 * dym[i];
        y[i] = y[i] + h6 * s;
      }
    }
  }
}

}  // namespace

extern "C" int pss_rk4_stage(int32_t stage, int64_t n, double h, double* y, const double* dydt, const double* dyt,
                             double* dym, double* ytmp, void* stream) {
  if (n <= 0) return 0;
  if (stage < 0 || stage > 3) return pss_set_msg(2, "pss_rk4_stage: stage must be 0..3");
  if (!y || (stage == 0 && (!dydt || !ytmp)) || (stage == 1 && (!dyt || !ytmp)) ||
      (stage == 2 && (!dym || !dyt || !ytmp)) || (stage == 3 && (!dydt || !dyt || !dym)))
    return pss_set_msg(2, "pss_rk4_stage: null operand for this stage");
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_rk4_stage, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (int)stage, n, h, y,
                     dydt, dyt, dym, ytmp);
  PSS_CHECK(hipGetLastError());
  return 0;
}
