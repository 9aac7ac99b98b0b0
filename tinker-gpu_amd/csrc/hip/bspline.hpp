// Order-5 cardinal B-splines of the PME grid, shared by the spreading and
// gathering kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace t9 {

// Order-5 cardinal B-spline weights and derivatives for a point at fractional
// offset w in [0,1): th[lv][m] = d^lv/dx^lv M5(x) at x = w + 4 - m,
// m = 0..4 (grid point base + m).
template <int LV, class real>
__device__ __forceinline__ void bspline5(real w, real th[LV][5]) {
  real A2[2] = {w, 1 - w};
  real A3[3], A4[4], A5[5];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    real lo = k >= 1 ? A2[k - 1] : 0, hi = k <= 1 ? A2[k] : 0;
    A3[k] = ((w + k) * hi + (3 - w - k) * lo) * (real)0.5;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    real lo = k >= 1 ? A3[k - 1] : 0, hi = k <= 2 ? A3[k] : 0;
    A4[k] = ((w + k) * hi + (4 - w - k) * lo) * (real)(1.0 / 3.0);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    real lo = k >= 1 ? A4[k - 1] : 0, hi = k <= 3 ? A4[k] : 0;
    A5[k] = ((w + k) * hi + (5 - w - k) * lo) * (real)0.25;
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const int k = 4 - m;
    th[0][m] = A5[k];
    if (LV > 1) {
      real a = k <= 3 ? A4[k] : 0, b = k >= 1 ? A4[k - 1] : 0;
      th[LV > 1 ? 1 : 0][m] = a - b;
    }
    if (LV > 2) {
      real a = k <= 2 ? A3[k] : 0, b = (k >= 1 && k - 1 <= 2) ? A3[k - 1] : 0, c = k >= 2 ? A3[k - 2] : 0;
      th[LV > 2 ? 2 : 0][m] = a - 2 * b + c;
    }
    if (LV > 3) {
      real a = k <= 1 ? A2[k] : 0, b = (k >= 1 && k - 1 <= 1) ? A2[k - 1] : 0;
      real c = (k >= 2 && k - 2 <= 1) ? A2[k - 2] : 0, d = k >= 3 ? A2[k - 3] : 0;
      th[LV > 3 ? 3 : 0][m] = a - 3 * b + 3 * c - d;
    }
  }
}

}  // namespace t9
