// dof.h -- rotation helpers of the DOF kernels and the trilinear interpolation, shared
// by the batched kernels, the persistent minimiser and the Monte Carlo chains.
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"
#include "kernels.h"

namespace mmo {

struct Rod { double m[9]; };

// Rot.of_axis_angle, src/rot.ml:121-146, from the sine and cosine of the angle
__device__ __forceinline__ Rod rodrigues(double x, double y, double z, double s, double c) {
  const double o = 1.0 - c;
  Rod r;
  r.m[0] = c + x * x * o;     r.m[1] = x * y * o - z * s; r.m[2] = x * z * o + y * s;
  r.m[3] = x * y * o + z * s; r.m[4] = c + y * y * o;     r.m[5] = y * z * o - x * s;
  r.m[6] = x * z * o - y * s; r.m[7] = y * z * o + x * s; r.m[8] = c + z * z * o;
  return r;
}

// Rot.r_xyz, src/rot.ml:52-66, and its partial derivatives; sn/cs hold the sines and
// cosines of (alpha, beta, gamma)
__device__ __forceinline__ void rot_xyz(const double* sn, const double* cs, double* m, double* da,
                                        double* db, double* dg) {
  const double as = sn[0], ac = cs[0], bs = sn[1], bc = cs[1], gs = sn[2], gc = cs[2];
  m[0] = bc * gc;  m[1] = gc * as * bs - ac * gs;  m[2] = as * gs + ac * gc * bs;
  m[3] = bc * gs;  m[4] = ac * gc + as * bs * gs;  m[5] = ac * bs * gs - gc * as;
  m[6] = -bs;      m[7] = bc * as;                 m[8] = ac * bc;
  if (da) {
    da[0] = 0.0;      da[1] = gc * ac * bs + as * gs; da[2] = ac * gs - as * gc * bs;
    da[3] = 0.0;      da[4] = -as * gc + ac * bs * gs; da[5] = -as * bs * gs - gc * ac;
    da[6] = 0.0;      da[7] = bc * ac;               da[8] = -as * bc;
    db[0] = -bs * gc; db[1] = gc * as * bc;          db[2] = ac * gc * bc;
    db[3] = -bs * gs; db[4] = as * bc * gs;          db[5] = ac * bc * gs;
    db[6] = -bc;      db[7] = -bs * as;              db[8] = -ac * bs;
    dg[0] = -bc * gs; dg[1] = -gs * as * bs - ac * gc; dg[2] = as * gc - ac * gs * bs;
    dg[3] = bc * gc;  dg[4] = -ac * gs + as * bs * gc; dg[5] = ac * bs * gc + gs * as;
    dg[6] = 0.0;      dg[7] = 0.0;                   dg[8] = 0.0;
  }
}

// The 8 corner values of cell (i0, j0, k0) of type ty's grid in the corner order a..h of
// src/G3D.ml:144-151: two 16-byte loads from the cell-packed copy when there is one,
// else 8 gathers from the node array.
// CELLS = false: the node arrays only (callers whose lanes share nodes).
template <bool CELLS = true>
__device__ __forceinline__ void grid_corners(const GridDev& g, int ty, int i0, int j0, int k0,
                                             float (&c)[8]) {
  if (CELLS && g.cells) {
    const size_t cell = (((size_t)ty * (g.nz - 1) + k0) * (g.ny - 1) + j0) * (g.nx - 1) + i0;
    const float4* q = reinterpret_cast<const float4*>(g.cells + 8 * cell);
    const float4 lo = q[0], hi = q[1];
    c[0] = lo.x; c[1] = lo.y; c[2] = lo.z; c[3] = lo.w;
    c[4] = hi.x; c[5] = hi.y; c[6] = hi.z; c[7] = hi.w;
    return;
  }
  const float* arr = g.data + (size_t)ty * g.nvox;
  const size_t nx = g.nx, nxy = (size_t)g.nx * g.ny;
  const size_t b = i0 + j0 * nx + k0 * nxy;
  c[0] = arr[b]; c[1] = arr[b + 1]; c[2] = arr[b + 1 + nx]; c[3] = arr[b + nx];
  c[4] = arr[b + nxy]; c[5] = arr[b + 1 + nxy]; c[6] = arr[b + 1 + nx + nxy];
  c[7] = arr[b + nx + nxy];
}

// G3D.trilin (src/G3D.ml:97-157) of type ty's grid at p - origin = (px, py, pz), FP64
// weights, corner order a..h; with GRAD also its gradient dg (A^-1).  ok = false (and
// 0) when the cell lies outside the grid.
template <bool GRAD, bool CELLS = true>
__device__ __forceinline__ double trilin(const GridDev& g, int ty, double px, double py,
                                         double pz, bool& ok, double* dg) {
  // src/G3D.ml:97-157: i0 = int_of_float (x * one_div_step) (truncation)
  const double fx = px * g.ods, fy = py * g.ods, fz = pz * g.ods;
  // cell test in floating point before the conversion (also rejects NaN and values
  // that would saturate the integer conversion); equals 0 <= i0 < nx - 1 with i0 the
  // truncation of a non-negative fx
  if (!(px >= 0.0 && py >= 0.0 && pz >= 0.0 && fx < (double)(g.nx - 1) &&
        fy < (double)(g.ny - 1) && fz < (double)(g.nz - 1))) {
    ok = false;
    return 0.0;
  }
  const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
  float cv[8];
  grid_corners<CELLS>(g, ty, i0, j0, k0, cv);
  const double A = cv[0], B = cv[1], C = cv[2], D = cv[3];
  const double E = cv[4], F = cv[5], G = cv[6], H = cv[7];
  const double wlx = (px - i0 * g.step) * g.ods, wly = (py - j0 * g.step) * g.ods,
               wlz = (pz - k0 * g.step) * g.ods;
  const double whx = 1.0 - wlx, why = 1.0 - wly, whz = 1.0 - wlz;
  if (GRAD) {
    dg[0] = g.ods * ((B - A) * why * whz + (C - D) * wly * whz + (F - E) * why * wlz + (G - H) * wly * wlz);
    dg[1] = g.ods * ((D - A) * whx * whz + (C - B) * wlx * whz + (H - E) * whx * wlz + (G - F) * wlx * wlz);
    dg[2] = g.ods * ((E - A) * whx * why + (F - B) * wlx * why + (G - C) * wlx * wly + (H - D) * whx * wly);
  }
  // corner order a..h, src/G3D.ml:144-151
  return A * (whx * why * whz) + B * (wlx * why * whz) + C * (wlx * wly * whz) +
         D * (whx * wly * whz) + E * (whx * why * wlz) + F * (wlx * why * wlz) +
         G * (wlx * wly * wlz) + H * (whx * wly * wlz);
}

}  // namespace mmo
