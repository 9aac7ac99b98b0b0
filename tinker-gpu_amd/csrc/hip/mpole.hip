// Local-frame -> global multipoles (mpoleInit: chkpole + rotpole).
// One thread per sorted atom, arithmetic in FP64 from the FP64 caller-order
// coordinates; the result is stored in the working precision, sorted order,
// and the frame itself (rows = local x, y, z axes) in FP32 for the learned
// initial guess.
// Follows include/seq/rotpole.h:8-224 (frame rules) and
// src/cu/amoeba/rotpole.cu:7-42.
#include "ctx.hpp"

namespace t9 {
namespace {

__device__ __forceinline__ void unit_vec(const double* xyz, int i, int j, double lx, double ly, double lz,
                                         int periodic, double* v) {
  double x = xyz[3 * j] - xyz[3 * i], y = xyz[3 * j + 1] - xyz[3 * i + 1], z = xyz[3 * j + 2] - xyz[3 * i + 2];
  if (periodic) {
    x -= lx * rint(x / lx);
    y -= ly * rint(y / ly);
    z -= lz * rint(z / lz);
  }
  double r = rsqrt(x * x + y * y + z * z);
  v[0] = x * r;
  v[1] = y * r;
  v[2] = z * r;
}

__device__ __forceinline__ void normalize(double* v) {
  double r = rsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= r;
  v[1] *= r;
  v[2] *= r;
}

template <class real>
__global__ void k_rotpole(int n, const int* __restrict__ perm, const double* __restrict__ xyz, double lx, double ly,
                          double lz, int periodic, const double* __restrict__ pole, const int* __restrict__ frame,
                          const int* __restrict__ zaxis, const int* __restrict__ xaxis, const int* __restrict__ yaxis,
                          const int* __restrict__ ysign, real* __restrict__ rpole, float* __restrict__ frot) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  double p[10];
  for (int k = 0; k < 10; ++k) p[k] = pole[10 * i + k];
  const int fr = frame[i];
  // chirality: Z-then-X frames with a y-defining atom flip the y components
  if (fr == 2 && yaxis[i] >= 0 && ysign[i] != 0) {
    const int ia = i, ib = zaxis[i], ic = xaxis[i], id = yaxis[i];
    double xad = xyz[3 * ia] - xyz[3 * id], yad = xyz[3 * ia + 1] - xyz[3 * id + 1], zad = xyz[3 * ia + 2] - xyz[3 * id + 2];
    double xbd = xyz[3 * ib] - xyz[3 * id], ybd = xyz[3 * ib + 1] - xyz[3 * id + 1], zbd = xyz[3 * ib + 2] - xyz[3 * id + 2];
    double xcd = xyz[3 * ic] - xyz[3 * id], ycd = xyz[3 * ic + 1] - xyz[3 * id + 1], zcd = xyz[3 * ic + 2] - xyz[3 * id + 2];
    double c1 = ybd * zcd - zbd * ycd, c2 = ycd * zad - zcd * yad, c3 = yad * zbd - zad * ybd;
    double vol = xad * c1 + xbd * c2 + xcd * c3;
    int k = ysign[i];
    if ((k < 0 && vol > 0) || (k > 0 && vol < 0)) {
      p[2] = -p[2];
      p[7] = -p[7];
      p[9] = -p[9];
    }
  }
  double a[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (fr != 0) {
    double* xx = a[0];
    double* zz = a[2];
    double yy[3] = {0, 0, 0};
    unit_vec(xyz, i, zaxis[i], lx, ly, lz, periodic, zz);
    if (fr == 1) {
      bool ok = !(fabs(zz[0]) > 0.866);
      xx[0] = ok ? 1 : 0;
      xx[1] = ok ? 0 : 1;
      xx[2] = 0;
    } else {
      unit_vec(xyz, i, xaxis[i], lx, ly, lz, periodic, xx);
    }
    if (fr == 4 || fr == 5) unit_vec(xyz, i, yaxis[i], lx, ly, lz, periodic, yy);
    if (fr == 3) {
      for (int d = 0; d < 3; ++d) zz[d] += xx[d];
      normalize(zz);
    } else if (fr == 4) {
      for (int d = 0; d < 3; ++d) xx[d] += yy[d];
      normalize(xx);
    } else if (fr == 5) {
      for (int d = 0; d < 3; ++d) zz[d] += xx[d] + yy[d];
      normalize(zz);
    }
    double dot = xx[0] * zz[0] + xx[1] * zz[1] + xx[2] * zz[2];
    for (int d = 0; d < 3; ++d) xx[d] -= dot * zz[d];
    normalize(xx);
    a[1][0] = zz[1] * xx[2] - zz[2] * xx[1];
    a[1][1] = zz[2] * xx[0] - zz[0] * xx[2];
    a[1][2] = zz[0] * xx[1] - zz[1] * xx[0];
  }
  for (int k = 0; k < 3; ++k)
    for (int g = 0; g < 3; ++g) frot[9 * (size_t)s + 3 * k + g] = (float)a[k][g];
  real* out = rpole + 10 * (size_t)s;
  out[0] = (real)p[0];
  for (int g = 0; g < 3; ++g) out[1 + g] = (real)(a[0][g] * p[1] + a[1][g] * p[2] + a[2][g] * p[3]);
  const double m[3][3] = {{p[4], p[7], p[8]}, {p[7], p[5], p[9]}, {p[8], p[9], p[6]}};
  double r[3][3];
  for (int g = 0; g < 3; ++g)
    for (int h = g; h < 3; ++h) {
      double t = 0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) t += a[k][g] * a[l][h] * m[k][l];
      r[g][h] = t;
    }
  out[4] = (real)r[0][0];
  out[5] = (real)r[1][1];
  out[6] = (real)r[2][2];
  out[7] = (real)r[0][1];
  out[8] = (real)r[0][2];
  out[9] = (real)r[1][2];
}

template <class real>
void mpole_init_t(Ctx& c) {
  const double lx = c.periodic ? c.box[0] : 1.0, ly = c.periodic ? c.box[1] : 1.0, lz = c.periodic ? c.box[2] : 1.0;
  c.frot.alloc(sizeof(float) * 9 * (size_t)c.n);
  hipLaunchKernelGGL(k_rotpole<real>, dim3(ceil_div(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.as<int>(),
                     c.xyz.as<double>(), lx, ly, lz, c.periodic ? 1 : 0, c.pole.as<double>(), c.frame.as<int>(),
                     c.zaxis.as<int>(), c.xaxis.as<int>(), c.yaxis.as<int>(), c.ysign.as<int>(), c.rpole.as<real>(),
                     c.frot.as<float>());
  T9_LAUNCH_CHECK();
  c.rpole_valid = true;
  ++c.rpole_stamp;
}

}  // namespace

void mpole_init(Ctx& c) {
  if (c.prec == 32) mpole_init_t<float>(c);
  else mpole_init_t<double>(c);
}

}  // namespace t9
