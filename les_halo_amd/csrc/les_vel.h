// vel2.f95:48-123: the interpolated velocities nou, velocity gradients diu and their products
// cov at one point, each with the reference's FP32 operation order.
#pragma once
#include "les_kernels.h"

namespace les {

template <bool P2>
struct VelCtx {
  Dims d;
  Grid1D g;
  Buf u, v, w;  // raw-buffer views of u, v (u's bounds) and w
  __device__ VelCtx(Dims d_, Grid1D g_, const float* u_, const float* v_, const float* w_)
      : d(d_), g(g_), u(u_, d_.nU()), v(v_, d_.nU()), w(w_, d_.nW()) {}
  __device__ float U(int i, int j, int k) const { return u[d.U(i, j, k)]; }
  __device__ float V(int i, int j, int k) const { return v[d.U(i, j, k)]; }
  __device__ float W(int i, int j, int k) const { return w[d.W(i, j, k)]; }
  __device__ float dx1(int i) const { return g.DX1(i); }
  __device__ float dy1(int j) const { return g.DY1(j); }
  __device__ float dzn(int k) const { return g.DZN(k); }
  __device__ float dzs(int k) const { return g.DZS(k); }
  __device__ double2 rzn(int k) const { return g.RZN(k); }
  __device__ double2 rzs(int k) const { return g.RZS(k); }
  __device__ double2 rzp(int k) const { return g.RZP(k); }

  // vel2.f95:51-60
  __device__ float diu1(int i, int j, int k) const { return xdiv<P2>(-U(i - 1, j, k) + U(i, j, k), dx1(i)); }
  __device__ float cov1(int i, int j, int k) const {
    return ((U(i - 1, j, k) + U(i, j, k)) / 2.0f) * diu1(i, j, k);
  }
  __device__ float diu5(int i, int j, int k) const { return xdiv<P2>(-V(i, j - 1, k) + V(i, j, k), dy1(j)); }
  __device__ float cov5(int i, int j, int k) const {
    return ((V(i, j - 1, k) + V(i, j, k)) / 2.0f) * diu5(i, j, k);
  }
  __device__ float diu9(int i, int j, int k) const { return zdiv(-W(i, j, k - 1) + W(i, j, k), dzn(k), rzn(k)); }
  __device__ float cov9(int i, int j, int k) const {
    return ((W(i, j, k - 1) + W(i, j, k)) / 2.0f) * diu9(i, j, k);
  }
  // vel2.f95:68-70
  __device__ float diu2(int i, int j, int k) const {
    return xdiv<P2>(2.0f * (-U(i, j - 1, k) + U(i, j, k)), dy1(j - 1) + dy1(j));
  }
  __device__ float cov2(int i, int j, int k) const {
    float nou = xdiv<P2>(dx1(i + 1) * V(i, j - 1, k) + dx1(i) * V(i + 1, j - 1, k), dx1(i) + dx1(i + 1));
    return nou * diu2(i, j, k);
  }
  // vel2.f95:78-80
  __device__ float diu3(int i, int j, int k) const {
    return zdiv(-U(i, j, k - 1) + U(i, j, k), dzs(k - 1), rzs(k - 1));
  }
  __device__ float cov3(int i, int j, int k) const {
    float nou = xdiv<P2>(dx1(i + 1) * W(i, j, k - 1) + dx1(i) * W(i + 1, j, k - 1), dx1(i) + dx1(i + 1));
    return nou * diu3(i, j, k);
  }
  // vel2.f95:88-90 (the "- u0" with u0 = 0 is exact)
  __device__ float diu4(int i, int j, int k) const {
    return xdiv<P2>(2.0f * (-V(i - 1, j, k) + V(i, j, k)), dx1(i - 1) + dx1(i));
  }
  __device__ float cov4(int i, int j, int k) const {
    float nou = xdiv<P2>(dy1(j + 1) * U(i - 1, j, k) + dy1(j) * U(i - 1, j + 1, k), dy1(j) + dy1(j + 1));
    return nou * diu4(i, j, k);
  }
  // vel2.f95:98-100
  __device__ float diu6(int i, int j, int k) const {
    return zdiv(-V(i, j, k - 1) + V(i, j, k), dzs(k - 1), rzs(k - 1));
  }
  __device__ float cov6(int i, int j, int k) const {
    float nou = xdiv<P2>(dy1(j + 1) * W(i, j, k - 1) + dy1(j) * W(i, j + 1, k - 1), dy1(j) + dy1(j + 1));
    return nou * diu6(i, j, k);
  }
  // vel2.f95:108-110
  __device__ float diu7(int i, int j, int k) const {
    return xdiv<P2>(2.0f * (-W(i - 1, j, k) + W(i, j, k)), dx1(i - 1) + dx1(i));
  }
  __device__ float cov7(int i, int j, int k) const {
    float nou = zdiv(dzn(k + 1) * U(i - 1, j, k) + dzn(k) * U(i - 1, j, k + 1), dzn(k) + dzn(k + 1),
                     rzp(k));
    return nou * diu7(i, j, k);
  }
  // vel2.f95:118-120
  __device__ float diu8(int i, int j, int k) const {
    return xdiv<P2>(2.0f * (-W(i, j - 1, k) + W(i, j, k)), dy1(j - 1) + dy1(j));
  }
  __device__ float cov8(int i, int j, int k) const {
    float nou = zdiv(dzn(k + 1) * V(i, j - 1, k) + dzn(k) * V(i, j - 1, k + 1), dzn(k) + dzn(k + 1),
                     rzp(k));
    return nou * diu8(i, j, k);
  }
};


// Stores diu1..9 of interior cell (i,j,k) with every halo copy vel2.f95:125-188 derives from it
// (i = im+1 column copies, periodic j rows, the "--les" copies for k up to km+1).  d3top/d6top
// are diu3/diu6 at km+1 (used when k == km).  Halo cells whose source vel2 never computes are
// copied from memory; the kernel never writes those sources, so the copies are race-free.
// A tile (e.tiled) makes the k copies and, at the domain's east face, the i = im+1 copies;
// every copy involving the periodic j rows comes from the exchange after the stage instead.
__device__ inline void store_diu(const Dims& d, const Fields& F, int i, int j, int k, float d1,
                                 float d2, float d3, float d4, float d5, float d6, float d7,
                                 float d8, float d9, float d3top, float d6top,
                                 Edges e = EDGES_ALL) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  const bool hk = k <= km - 1;
  const bool ie = (i == im) && e.e;  // i = im+1 copies
  const bool pj = !e.tiled;          // periodic j copies
  // the interior stores through buffer views; the halo copies (few threads) through pointers
  Buf(F.diu1, d.nD15()).st(d.D15(i, j, k), d1);
  if (ie) F.diu1[d.D15(im + 1, j, k)] = d1;
  Buf(F.diu5, d.nD15()).st(d.D15(i, j, k), d5);
  if (pj && j == jm) F.diu5[d.D15(i, 0, k)] = d5;
  if (pj && j == 1) F.diu5[d.D15(i, jm + 1, k)] = d5;
  Buf(F.diu9, d.nD()).st(d.D(i, j, k), d9);
  Buf(F.diu2, d.nD()).st(d.D(i, j, k), d2);
  if (pj && j == jm) F.diu2[d.D(i, 0, k)] = d2;
  if (pj && j == 1) F.diu2[d.D(i, jm + 1, k)] = d2;
  if (ie) {
    F.diu2[d.D(im + 1, j, k)] = d2;
    if (pj && j == 1) F.diu2[d.D(im + 1, jm + 1, k)] = d2;
    if (k == km) {
      F.diu2[d.D(im + 1, j, km + 1)] = F.diu2[d.D(im, j, km + 1)];
      if (pj && j == jm) F.diu2[d.D(im + 1, jm + 1, km + 1)] = F.diu2[d.D(im, jm + 1, km + 1)];
    }
  }
  Buf(F.diu3, d.nD()).st(d.D(i, j, k), d3);
  if (k == km) F.diu3[d.D(i, j, km + 1)] = d3top;
  if (ie) {
    F.diu3[d.D(im + 1, j, k)] = d3;
    if (k == km) F.diu3[d.D(im + 1, j, km + 1)] = d3top;
    if (pj && j == jm) {
      F.diu3[d.D(im + 1, jm + 1, k)] = F.diu3[d.D(im, jm + 1, k)];
      if (k == km) F.diu3[d.D(im + 1, jm + 1, km + 1)] = F.diu3[d.D(im, jm + 1, km + 1)];
    }
  }
  Buf(F.diu4, d.nD()).st(d.D(i, j, k), d4);
  if (ie) F.diu4[d.D(im + 1, j, k)] = d4;
  if (pj && j == jm) {
    F.diu4[d.D(i, 0, k)] = d4;
    if (ie) F.diu4[d.D(im + 1, 0, k)] = d4;
    if (k == km) {
      F.diu4[d.D(i, 0, km + 1)] = F.diu4[d.D(i, jm, km + 1)];
      if (ie) F.diu4[d.D(im + 1, 0, km + 1)] = F.diu4[d.D(im + 1, jm, km + 1)];
    }
  }
  Buf(F.diu6, d.nD()).st(d.D(i, j, k), d6);
  if (k == km) F.diu6[d.D(i, j, km + 1)] = d6top;
  if (pj && j == jm) {
    F.diu6[d.D(i, 0, k)] = d6;
    if (k == km) F.diu6[d.D(i, 0, km + 1)] = d6top;
    if (ie) {
      F.diu6[d.D(im + 1, 0, k)] = F.diu6[d.D(im + 1, jm, k)];
      if (k == km) F.diu6[d.D(im + 1, 0, km + 1)] = F.diu6[d.D(im + 1, jm, km + 1)];
    }
  }
  if (hk) {
    Buf(F.diu7, d.nD()).st(d.D(i, j, k), d7);
    if (ie) F.diu7[d.D(im + 1, j, k)] = d7;
    Buf(F.diu8, d.nD()).st(d.D(i, j, k), d8);
    if (pj && j == jm) F.diu8[d.D(i, 0, k)] = d8;
    if (pj && j == 1) F.diu8[d.D(i, jm + 1, k)] = d8;
  }
}

}  // namespace les
