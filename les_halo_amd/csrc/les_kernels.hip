// HIP kernels (gfx950) for the LES_Halo time step.  Each kernel restates one reference stage
// (file:line cited per kernel) in the reference's Fortran array layouts, with the reference's
// FP32 operation order, compiled without FP contraction, so that every element is bit-identical
// to the sequential reference.  Threads run along i (the contiguous axis) so every wavefront
// reads and writes 64 consecutive floats.  Boundary ("halo") values are written by the thread
// that owns their source, or by a dedicated boundary kernel whose every read is of a cell that
// the kernel never writes; no kernel relies on inter-workgroup ordering.
#include "les_kernels.h"
#include "les_vel.h"

namespace les {

namespace {

constexpr int BX = 64;  // one wavefront along i
constexpr int BY = 4;   // rows of j per workgroup

inline dim3 interior_grid(Dims d, int kcount) {
  return dim3((d.ip + BX - 1) / BX, (d.jp + BY - 1) / BY, kcount);
}

__device__ __forceinline__ float mbit(uint8_t m, int b) { return (m >> b) & 1u ? 1.0f : 0.0f; }

}  // namespace

// ==========================================================================================
// velnw.f95:27-59 fused with the inflow of bondv1_ext_wind_profile.f95:27-50
// (state 1 of run_LES_kernel).  Threads cover i = 0..imax, j = 1..jm, k = 1..km.
// ==========================================================================================
// A thread takes VN_KPT consecutive levels of one (i,j) column: all their operands are loaded
// before the first store (the stores could alias them for the compiler, which would otherwise
// wait for each load group in turn), so more bytes are in flight per wavefront.  p(i,j,kp+1), h
// and w at k = kp exist; the w update is skipped there.
#ifndef LES_VELNW_KPT
#define LES_VELNW_KPT 2
#endif
constexpr int VN_KPT = LES_VELNW_KPT;
template <bool P2>
__global__ __launch_bounds__(256) void k_velnw_inflow(Dims d, Grid1D g, Fields F, Physics ph,
                                                      int do_velnw, int inflow_n, int imax,
                                                      StripGrid sg) {
  int bx, by, bz, L;
  if (!strip_block(sg, bx, by, bz, L)) return;
  const int i = bx * (int)blockDim.x + threadIdx.x;
  const int j = 1 + by * (int)blockDim.y + threadIdx.y;
  const int k0 = 1 + bz * VN_KPT;
  if (i > imax || j > d.jp) return;
  if (inflow_n > 0 && (i <= 1 || inflow_n == 1)) {
    // inflow planes i = 0,1 (and, at n == 1, the fill u(i,:,:) = u(1,:,:) for i = 2..im)
#pragma unroll
    for (int q = 0; q < VN_KPT; ++q) {
      const int k = k0 + q;
      if (k > d.kp) break;
      F.u[d.U(i, j, k)] = F.wind[(j - 1) + d.jp * (k - 1)];
      F.v[d.U(i, j, k)] = 0.0f;
      F.w[d.W(i, j, k)] = 0.0f;
    }
    return;
  }
  if (!do_velnw || i < 1) return;
  const float ro = ph.ro, dt = ph.dt;
  const Buf bp(F.p, d.nP()), bf(F.f, d.nF()), bg(F.g, d.nF()), bh(F.h, d.nF());
  const Buf bu(F.u, d.nU()), bv(F.v, d.nU()), bw(F.w, d.nW());
  float p0[VN_KPT], pe[VN_KPT], pn[VN_KPT], pt[VN_KPT];
  float fv[VN_KPT], gv[VN_KPT], hv[VN_KPT], uv[VN_KPT], vv[VN_KPT], wv[VN_KPT];
#pragma unroll
  for (int q = 0; q < VN_KPT; ++q) {
    const int k = min(k0 + q, d.kp);  // levels past kp: loaded, not stored
    const long ifg = d.F(i, j, k), iu = d.U(i, j, k);
    p0[q] = bp[d.P(i, j, k)];
    pe[q] = bp[d.P(i + 1, j, k)];
    pn[q] = bp[d.P(i, j + 1, k)];
    pt[q] = bp[d.P(i, j, k + 1)];
    fv[q] = bf[ifg];
    gv[q] = bg[ifg];
    hv[q] = bh[ifg];
    uv[q] = bu[iu];
    vv[q] = bv[iu];
    wv[q] = bw[d.W(i, j, k)];
  }
#pragma unroll
  for (int q = 0; q < VN_KPT; ++q) {
    const int k = k0 + q;
    if (k > d.kp) break;
    const long iu = d.U(i, j, k);
    float pz = xdiv<P2>((-p0[q] + pe[q]) / ro, g.DXS(i));
    bu.st(iu, uv[q] + dt * (fv[q] - pz));
    pz = xdiv<P2>((-p0[q] + pn[q]) / ro, g.DYS(j));
    bv.st(iu, vv[q] + dt * (gv[q] - pz));
    if (k <= d.kp - 1) {
      pz = ((-p0[q] + pt[q]) / ro) / g.DZS(k);
      bw.st(d.W(i, j, k), wv[q] + dt * (hv[q] - pz));
    }
  }
}

void launch_velnw_inflow(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velnw,
                         int inflow_n, int shape) {
  int imax = (do_velnw || inflow_n == 1) ? d.ip : 1;
  const int sx = stencil_shape_x(shape), sy = stencil_shape_y(shape);
  const StripGrid sg = strip_grid((imax + 1 + sx - 1) / sx, (d.jp + sy - 1) / sy,
                                  (d.kp + VN_KPT - 1) / VN_KPT);
  if (g.p2xy)
    hipLaunchKernelGGL(k_velnw_inflow<true>, strip_dim(sg), dim3(sx, sy), 0, st, d, g, F, ph,
                       do_velnw, inflow_n, imax, sg);
  else
    hipLaunchKernelGGL(k_velnw_inflow<false>, strip_dim(sg), dim3(sx, sy), 0, st, d, g, F, ph,
                       do_velnw, inflow_n, imax, sg);
}

// ==========================================================================================
// bondv1.f95:67-75: uout = (max(0, max u(im,:,:)) + min(0, min u(im,:,:))) / 2.
// One workgroup per level k; max/min are order-independent, so the last workgroup to finish
// folds the per-level partials (seeded with 0 as amax1/amin1 are) into the exact result and
// re-arms the counter for the next launch.
// ==========================================================================================
namespace {
__device__ __forceinline__ void block_maxmin(float& mx, float& mn, float* smax, float* smin) {
  for (int o = 32; o > 0; o >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, o));
    mn = fminf(mn, __shfl_xor(mn, o));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smax[wv] = mx;
    smin[wv] = mn;
  }
  __syncthreads();
  mx = smax[0];
  mn = smin[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
    mx = fmaxf(mx, smax[q]);
    mn = fminf(mn, smin[q]);
  }
  __syncthreads();
}
}  // namespace

__global__ __launch_bounds__(256) void k_uout(Dims d, Fields F, DevScalars* sc, float* part) {
  __shared__ float smax[4], smin[4];
  __shared__ int last;
  const int k = 1 + blockIdx.x;
  float mx = 0.0f, mn = 0.0f;
  for (int j = 1 + threadIdx.x; j <= d.jp; j += 256) {
    const float x = F.u[d.U(d.ip, j, k)];
    mx = fmaxf(mx, x);
    mn = fminf(mn, x);
  }
  block_maxmin(mx, mn, smax, smin);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mx;
    part[2 * blockIdx.x + 1] = mn;
    __threadfence();
    const unsigned prev = atomicAdd(&sc->uout_count, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  mx = 0.0f;
  mn = 0.0f;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) {
    mx = fmaxf(mx, __hip_atomic_load(&part[2 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    mn = fminf(mn, __hip_atomic_load(&part[2 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  block_maxmin(mx, mn, smax, smin);
  if (threadIdx.x == 0) {
    sc->uout_max = mx;
    sc->uout_min = mn;
    sc->uout = (mx + mn) / 2.0f;
    sc->uout_count = 0u;
  }
}

void launch_uout(hipStream_t st, Dims d, Fields F, DevScalars* sc, float* part) {
  hipLaunchKernelGGL(k_uout, dim3(d.kp), dim3(256), 0, st, d, F, sc, part);
}

// ==========================================================================================
// bondv1.f95:81-97: advective outflow at i = im (u) and i = im+1 (v, w)
// ==========================================================================================
__global__ __launch_bounds__(256) void k_outflow(Dims d, Grid1D g, Fields F, Physics ph,
                                                 const DevScalars* sc) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= d.jp * d.kp) return;
  const int j = 1 + t % d.jp, k = 1 + t / d.jp;
  const int im = d.ip;
  const float dt = ph.dt, uout = sc->uout, dxs = g.DXS(im);
  long a = d.U(im, j, k), b = d.U(im - 1, j, k);
  F.u[a] = F.u[a] - dt * uout * (F.u[a] - F.u[b]) / dxs;
  a = d.U(im + 1, j, k);
  b = d.U(im, j, k);
  F.v[a] = F.v[a] - dt * uout * (F.v[a] - F.v[b]) / dxs;
  a = d.W(im + 1, j, k);
  b = d.W(im, j, k);
  F.w[a] = F.w[a] - dt * uout * (F.w[a] - F.w[b]) / dxs;
}

// bondv1.f95:99-137: periodic j, then top/bottom; each halo cell gets its final value from
// interior rows (which this kernel never writes).  The reference's w loop runs i = -1..im+1
// although w starts at i = 0; in the column-major layout w(-1,j,k) is w(ip+1,j-1,k), so the
// only cells it reaches beyond the i = im+1 column are w(ip+1,jp+1,-1) and w(ip+1,jp+1,km-1),
// both zeroed here as well.
// A tile (e.tiled) has its j faces and internal i faces from the exchange before this kernel:
// it makes no periodic copies, takes the k planes from its own (exchanged) rows, and writes the
// w(-1,-1,...) cells only where they lie in the whole domain (the tile at its east and north
// faces).
__global__ __launch_bounds__(256) void k_bondv1_halo(Dims d, Fields F, Edges e) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  const int ni = im + 2;
  const long n1 = e.tiled ? 0 : (long)ni * km * 2;  // j faces, k = 1..km, two faces
  const long n2 = (long)ni * (jm + 3) * 2;         // k faces, j = -1..jm+1, two planes
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n1) {
    const int i = (int)(t % ni);
    const int k = 1 + (int)((t / ni) % km);
    const int face = (int)(t / ((long)ni * km));
    const int jdst = face ? jm + 1 : 0, jsrc = face ? 1 : jm;
    F.u[d.U(i, jdst, k)] = F.u[d.U(i, jsrc, k)];
    F.v[d.U(i, jdst, k)] = F.v[d.U(i, jsrc, k)];
    if (k <= km - 1) {
      float wv = F.w[d.W(i, jsrc, k)];
      if (i == im + 1 && face == 1 && k == km - 1) wv = 0.0f;  // the w(-1,-1,km) write
      F.w[d.W(i, jdst, k)] = wv;
    }
    return;
  }
  t -= n1;
  if (t >= n2) return;
  const int i = (int)(t % ni);
  const int j = -1 + (int)((t / ni) % (jm + 3));
  const int top = (int)(t / ((long)ni * (jm + 3)));
  // w: zero planes k = 0 and k = km (j = -1..jm+1)
  F.w[d.W(i, j, top ? km : 0)] = 0.0f;
  const bool corner = i == im + 1 && j == jm + 1 && (!e.tiled || (e.e && e.n));
  if (!top && corner) F.w[d.W(im + 1, jm + 1, -1)] = 0.0f;  // w(-1,-1,0)
  if (top && corner && e.tiled) F.w[d.W(im + 1, jm + 1, km - 1)] = 0.0f;  // w(-1,-1,km)
  if (j < 0) return;
  // u, v: k = 0 antisymmetric, k = km+1 copy, from the periodic-completed rows
  const int js = e.tiled ? j : ((j == 0) ? jm : ((j == jm + 1) ? 1 : j));
  if (top) {
    F.u[d.U(i, j, km + 1)] = F.u[d.U(i, js, km)];
    F.v[d.U(i, j, km + 1)] = F.v[d.U(i, js, km)];
  } else {
    F.u[d.U(i, j, 0)] = -F.u[d.U(i, js, 1)];
    F.v[d.U(i, j, 0)] = -F.v[d.U(i, js, 1)];
  }
}

// do_outflow = 1: outflow then the halo kernel; 0: the halo kernel only; -1: outflow only
void launch_bondv1_uvw(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, DevScalars* sc,
                       Edges e, int do_outflow) {
  int n = d.jp * d.kp;
  if (do_outflow != 0)
    hipLaunchKernelGGL(k_outflow, dim3((n + 255) / 256), dim3(256), 0, st, d, g, F, ph, sc);
  if (do_outflow < 0) return;
  long nh = (e.tiled ? 0L : (long)(d.ip + 2) * d.kp * 2) + (long)(d.ip + 2) * (d.jp + 3) * 2;
  hipLaunchKernelGGL(k_bondv1_halo, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, d, F,
                     e);
}

// ==========================================================================================
// vel2.f95:48-123 formulas at one point; velFG.f95:59-105; feedbf.f95:38-67.
// ==========================================================================================

// One thread per interior cell.  Neighbour values of cov/diu at i+1 = im+1 use the vel2 halo
// copy (value at im), at j+1 = jm+1 the periodic copy (value at j = 1); k+1 values are
// computed (vel2 computes them up to km+1).  diu1..9 are stored with every halo copy of
// vel2.f95:125-188, each written by the thread owning its source.
template <bool P2>
__global__ __launch_bounds__(256) void k_velfg_feedbf(Dims d, Grid1D g, Fields F, Physics ph,
                                                      int do_velfg, int do_feedbf) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int im = d.ip, jm = d.jp, km = d.kp;
  if (i > im || j > jm) return;
  VelCtx<P2> c{d, g, F.u, F.v, F.w};
  const int ip1 = (i == im) ? im : i + 1;  // i+1 with the im+1 -> im copy
  const int jp1 = (j == jm) ? 1 : j + 1;   // j+1 with the jm+1 -> 1 copy
  const long ifg = d.F(i, j, k);
  float fv = 0.0f, gv = 0.0f, hv = 0.0f;
  const float two = 2.0f;

  // ---- diu values at this cell (also stored) ----
  const float d1 = c.diu1(i, j, k), d2 = c.diu2(i, j, k), d3 = c.diu3(i, j, k);
  const float d4 = c.diu4(i, j, k), d5 = c.diu5(i, j, k), d6 = c.diu6(i, j, k);
  const float d9 = c.diu9(i, j, k);
  const bool hk = k <= km - 1;
  const float d7 = hk ? c.diu7(i, j, k) : 0.0f, d8 = hk ? c.diu8(i, j, k) : 0.0f;

  if (do_velfg) {
    const float dx1i = c.dx1(i), dx1ip = c.dx1(i + 1), dy1j = c.dy1(j), dy1jp = c.dy1(j + 1);
    const float dznk = c.dzn(k), dznkp = c.dzn(k + 1);
    const double2 rznk = c.rzn(k);
    // u momentum (velFG.f95:59-72)
    {
      float covx1 = xdiv<P2>(dx1ip * c.cov1(i, j, k) + dx1i * c.cov1(ip1, j, k), dx1i + dx1ip);
      float covy1 = (c.cov2(i, j, k) + c.cov2(i, jp1, k)) / two;
      float covz1 = (c.cov3(i, j, k) + c.cov3(i, j, k + 1)) / two;
      float covc = covx1 + covy1 + covz1;
      float dfu1 = xdiv<P2>(two * (-d1 + c.diu1(ip1, j, k)), dx1i + dx1ip) +
                   xdiv<P2>(-d2 + c.diu2(i, jp1, k), dy1j) + zdiv(-d3 + c.diu3(i, j, k + 1), dznk, rznk);
      float df = ph.vn * dfu1;
      fv = (-covc + df);
    }
    // v momentum (velFG.f95:75-88)
    {
      float covx1 = (c.cov4(i, j, k) + c.cov4(ip1, j, k)) / two;
      float covy1 = xdiv<P2>(dy1jp * c.cov5(i, j, k) + dy1j * c.cov5(i, jp1, k), dy1j + dy1jp);
      float covz1 = (c.cov6(i, j, k) + c.cov6(i, j, k + 1)) / two;
      float covc = covx1 + covy1 + covz1;
      float dfv1 = xdiv<P2>(-d4 + c.diu4(ip1, j, k), dx1i) +
                   xdiv<P2>(two * (-d5 + c.diu5(i, jp1, k)), dy1j + dy1jp) +
                   zdiv(-d6 + c.diu6(i, j, k + 1), dznk, rznk);
      float df = ph.vn * dfv1;
      gv = (-covc + df);
    }
    // w momentum (velFG.f95:92-105), k <= km-1
    if (hk) {
      float covx1 = (c.cov7(i, j, k) + c.cov7(ip1, j, k)) / two;
      float covy1 = (c.cov8(i, j, k) + c.cov8(i, jp1, k)) / two;
      float covz1 = zdiv(dznkp * c.cov9(i, j, k) + dznk * c.cov9(i, j, k + 1), dznk + dznkp, c.rzp(k));
      float covc = covx1 + covy1 + covz1;
      float dfw1 = xdiv<P2>(-d7 + c.diu7(ip1, j, k), dx1i) + xdiv<P2>(-d8 + c.diu8(i, jp1, k), dy1j) +
                   zdiv(-d9 + c.diu9(i, j, k + 1), c.dzs(k), c.rzs(k));
      float df = ph.vn * dfw1;
      hv = (-covc + df);
    } else {
      hv = F.h[ifg];  // velfg leaves h(:,:,km) untouched
    }
  } else {
    fv = F.f[ifg];
    gv = F.g[ifg];
    hv = F.h[ifg];
  }

  if (do_feedbf) {  // feedbf.f95:41-65
    const uint8_t m = F.mask[d.M(i, j, k)];
    const float bm = mbit(m, 0), cm = mbit(m, 1), dm = mbit(m, 2);
    const float uu = c.U(i, j, k), vv = c.V(i, j, k), ww = c.W(i, j, k);
    const float us = (F.usum[ifg] + uu) * bm;
    const float vs = (F.vsum[ifg] + vv) * cm;
    const float ws = (F.wsum[ifg] + ww) * dm;
    F.usum[ifg] = us;
    F.vsum[ifg] = vs;
    F.wsum[ifg] = ws;
    const float f1x = ph.alpha * us * ph.dt, f1y = ph.alpha * vs * ph.dt, f1z = ph.alpha * ws * ph.dt;
    const float bc = F.cd ? ph.beta * F.cd[(i - 1) + d.ip * ((j - 1) + d.jp * (k - 1))] : ph.beta;
    const float f2x = bc * uu * bm, f2y = bc * vv * cm, f2z = bc * ww * dm;
    fv = fv + (f1x + f2x);
    gv = gv + (f1y + f2y);
    hv = hv + (f1z + f2z);
  }
  F.f[ifg] = fv;
  F.g[ifg] = gv;
  F.h[ifg] = hv;

  if (!do_velfg) return;
  const float d3top = (k == km) ? c.diu3(i, j, km + 1) : 0.0f;
  const float d6top = (k == km) ? c.diu6(i, j, km + 1) : 0.0f;
  store_diu(d, F, i, j, k, d1, d2, d3, d4, d5, d6, d7, d8, d9, d3top, d6top);
}

// velfg + feedbf with every i-face and j-face quantity computed once per point.  A 64 x 8
// workgroup covers 63 x 7 cells: each lane computes the i-face quantities (diu1/cov1, diu4/cov4,
// diu7/cov7) at its own i and hands them to the lane before it (its i+1), each row the j-face
// quantities (diu2/cov2, diu5/cov5, diu8/cov8) at its own j through LDS to the row before it;
// lane 63 and row 7 only produce.  The k-face quantities at k and k+1 are computed in the
// thread.  Each value comes from the same formula at the same point as in k_velfg_feedbf, so
// the results are bit-identical; the i-face and j-face work is halved.
constexpr int VX = 64, VY = 8;
// The body of one workgroup with logical index L (k_velfg_xy, and the ensemble's batched launch)
template <bool P2>
__device__ __forceinline__ void velfg_xy_body(const Dims& d, const Grid1D& g, const Fields& F,
                                              const Physics& ph, int do_feedbf,
                                              const StripGrid& sg, const Edges& e, int L) {
  __shared__ float sy[6][VY][VX];  // cov2 diu2 cov5 diu5 cov8 diu8 of each row
  int bx, by, bz;
  if (!strip_map(sg, L, bx, by, bz)) return;
  const int x = threadIdx.x, y = threadIdx.y;
  const int im = d.ip, jm = d.jp, km = d.kp;
  const int i = 1 + bx * (VX - 1) + x;
  const int j = 1 + by * (VY - 1) + y;
  const int k = 1 + bz;
  const bool hk = k <= km - 1;
  VelCtx<P2> c{d, g, F.u, F.v, F.w};
  // i-face quantities at (i, j, k); a tile whose east face is internal computes them at im+1
  // too, from its exchanged halo (the whole domain's values there)
  const int imx = e.e ? im : im + 1;
  float x1 = 0.0f, xc1 = 0.0f, x4 = 0.0f, xc4 = 0.0f, x7 = 0.0f, xc7 = 0.0f;
  if (y < VY - 1 && i <= imx && j <= jm) {
    x1 = c.diu1(i, j, k);
    xc1 = c.cov1(i, j, k);
    x4 = c.diu4(i, j, k);
    xc4 = c.cov4(i, j, k);
    if (hk) {
      x7 = c.diu7(i, j, k);
      xc7 = c.cov7(i, j, k);
    }
  }
  // j-face quantities at (i, jq, k); row jm+1 is the periodic copy of row 1 (a tile computes
  // it from its exchanged halo rows)
  if (x < VX - 1 && i <= im && j <= jm + 1) {
    const int jq = (j == jm + 1 && !e.tiled) ? 1 : j;
    sy[0][y][x] = c.cov2(i, jq, k);
    sy[1][y][x] = c.diu2(i, jq, k);
    sy[2][y][x] = c.cov5(i, jq, k);
    sy[3][y][x] = c.diu5(i, jq, k);
    if (hk) {
      sy[4][y][x] = c.cov8(i, jq, k);
      sy[5][y][x] = c.diu8(i, jq, k);
    }
  }
  float x1n = __shfl_down(x1, 1), xc1n = __shfl_down(xc1, 1);
  float x4n = __shfl_down(x4, 1), xc4n = __shfl_down(xc4, 1);
  float x7n = __shfl_down(x7, 1), xc7n = __shfl_down(xc7, 1);
  __syncthreads();
  if (x >= VX - 1 || y >= VY - 1 || i > im || j > jm) return;
  if (i == im && e.e) {  // vel2.f95:125-165: the i-face quantities at im+1 are copies of im
    x1n = x1;
    xc1n = xc1;
    x4n = x4;
    xc4n = xc4;
    x7n = x7;
    xc7n = xc7;
  }
  const float y2 = sy[1][y][x], y5 = sy[3][y][x];
  const float yc2 = sy[0][y][x], yc2n = sy[0][y + 1][x], y2n = sy[1][y + 1][x];
  const float yc5 = sy[2][y][x], yc5n = sy[2][y + 1][x], y5n = sy[3][y + 1][x];
  // k-face quantities at k and k+1
  const float z3 = c.diu3(i, j, k), zc3 = c.cov3(i, j, k);
  const float z3n = c.diu3(i, j, k + 1), zc3n = c.cov3(i, j, k + 1);
  const float z6 = c.diu6(i, j, k), zc6 = c.cov6(i, j, k);
  const float z6n = c.diu6(i, j, k + 1), zc6n = c.cov6(i, j, k + 1);
  const float z9 = c.diu9(i, j, k);
  const long ifg = d.F(i, j, k);
  const float two = 2.0f;
  const float dx1i = c.dx1(i), dx1ip = c.dx1(i + 1), dy1j = c.dy1(j), dy1jp = c.dy1(j + 1);
  const float dznk = c.dzn(k), dznkp = c.dzn(k + 1);
  const double2 rznk = c.rzn(k);
  float fv, gv, hv;
  {  // velFG.f95:59-72
    float covx1 = xdiv<P2>(dx1ip * xc1 + dx1i * xc1n, dx1i + dx1ip);
    float covy1 = (yc2 + yc2n) / two;
    float covz1 = (zc3 + zc3n) / two;
    float covc = covx1 + covy1 + covz1;
    float dfu1 = xdiv<P2>(two * (-x1 + x1n), dx1i + dx1ip) + xdiv<P2>(-y2 + y2n, dy1j) +
                 zdiv(-z3 + z3n, dznk, rznk);
    float df = ph.vn * dfu1;
    fv = (-covc + df);
  }
  {  // velFG.f95:75-88
    float covx1 = (xc4 + xc4n) / two;
    float covy1 = xdiv<P2>(dy1jp * yc5 + dy1j * yc5n, dy1j + dy1jp);
    float covz1 = (zc6 + zc6n) / two;
    float covc = covx1 + covy1 + covz1;
    float dfv1 = xdiv<P2>(-x4 + x4n, dx1i) + xdiv<P2>(two * (-y5 + y5n), dy1j + dy1jp) +
                 zdiv(-z6 + z6n, dznk, rznk);
    float df = ph.vn * dfv1;
    gv = (-covc + df);
  }
  float y8 = 0.0f;
  if (hk) {  // velFG.f95:92-105
    const float yc8 = sy[4][y][x], yc8n = sy[4][y + 1][x];
    y8 = sy[5][y][x];
    const float y8n = sy[5][y + 1][x];
    const float z9n = c.diu9(i, j, k + 1), zc9 = c.cov9(i, j, k), zc9n = c.cov9(i, j, k + 1);
    float covx1 = (xc7 + xc7n) / two;
    float covy1 = (yc8 + yc8n) / two;
    float covz1 = zdiv(dznkp * zc9 + dznk * zc9n, dznk + dznkp, c.rzp(k));
    float covc = covx1 + covy1 + covz1;
    float dfw1 = xdiv<P2>(-x7 + x7n, dx1i) + xdiv<P2>(-y8 + y8n, dy1j) +
                 zdiv(-z9 + z9n, c.dzs(k), c.rzs(k));
    float df = ph.vn * dfw1;
    hv = (-covc + df);
  } else {
    hv = Buf(F.h, d.nF())[ifg];  // velfg leaves h(:,:,km) untouched
  }
  if (do_feedbf) {  // feedbf.f95:41-65
    const uint8_t m = F.mask[d.M(i, j, k)];
    const float bm = mbit(m, 0), cm = mbit(m, 1), dm = mbit(m, 2);
    const float uu = c.U(i, j, k), vv = c.V(i, j, k), ww = c.W(i, j, k);
    const Buf bus(F.usum, d.nF()), bvs(F.vsum, d.nF()), bws(F.wsum, d.nF());
    const float us = (bus[ifg] + uu) * bm;
    const float vs = (bvs[ifg] + vv) * cm;
    const float ws = (bws[ifg] + ww) * dm;
    bus.st(ifg, us);
    bvs.st(ifg, vs);
    bws.st(ifg, ws);
    const float f1x = ph.alpha * us * ph.dt, f1y = ph.alpha * vs * ph.dt, f1z = ph.alpha * ws * ph.dt;
    const float bc = F.cd ? ph.beta * F.cd[(i - 1) + d.ip * ((j - 1) + d.jp * (k - 1))] : ph.beta;
    const float f2x = bc * uu * bm, f2y = bc * vv * cm, f2z = bc * ww * dm;
    fv = fv + (f1x + f2x);
    gv = gv + (f1y + f2y);
    hv = hv + (f1z + f2z);
  }
  Buf(F.f, d.nF()).st(ifg, fv);
  Buf(F.g, d.nF()).st(ifg, gv);
  Buf(F.h, d.nF()).st(ifg, hv);
  store_diu(d, F, i, j, k, x1, y2, z3, x4, y5, z6, x7, y8, z9, z3n, z6n, e);
}

template <bool P2>
__global__ __launch_bounds__(512) void k_velfg_xy(Dims d, Grid1D g, Fields F, Physics ph,
                                                  int do_feedbf, StripGrid sg, Edges e) {
  velfg_xy_body<P2>(d, g, F, ph, do_feedbf, sg, e, xcd_tile((int)blockIdx.x, (int)gridDim.x));
}

StripGrid velfg_strip(Dims d) {
  return strip_grid((d.ip + VX - 2) / (VX - 1), (d.jp + VY - 2) / (VY - 1), d.kp);
}

void launch_velfg(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velfg,
                  int do_feedbf, Edges e) {
  if (do_velfg) {
    // (the last column of blocks always has a lane at i = im+1: lane 63 of a block sits one
    // cell past its last output)
    const StripGrid sg = strip_grid((d.ip + VX - 2) / (VX - 1), (d.jp + VY - 2) / (VY - 1), d.kp);
    if (g.p2xy)
      hipLaunchKernelGGL(k_velfg_xy<true>, strip_dim(sg), dim3(VX, VY), 0, st, d, g, F, ph,
                         do_feedbf, sg, e);
    else
      hipLaunchKernelGGL(k_velfg_xy<false>, strip_dim(sg), dim3(VX, VY), 0, st, d, g, F, ph,
                         do_feedbf, sg, e);
    return;
  }
  if (g.p2xy)
    hipLaunchKernelGGL(k_velfg_feedbf<true>, interior_grid(d, d.kp), dim3(BX, BY), 0, st, d, g,
                       F, ph, do_velfg, do_feedbf);
  else
    hipLaunchKernelGGL(k_velfg_feedbf<false>, interior_grid(d, d.kp), dim3(BX, BY), 0, st, d, g,
                       F, ph, do_velfg, do_feedbf);
}

// ==========================================================================================
// les.f95:40-60: Smagorinsky eddy viscosity sm from the stored diu fields
// ==========================================================================================
// les.f95:40-56 at one interior point: the diu values sm(i,j,k) reads, loaded by sm_load
// before any arithmetic, and the formula (sm_value) in the reference's operation order
struct SmIn {
  float d1, d5, d9, d2[4], d3[4], d4[4], d6[4], d7[4], d8[4];
};
__device__ __forceinline__ SmIn sm_load(const Dims& d, const Fields& F, int i, int j, int k) {
  const Buf diu2(F.diu2, d.nD()), diu3(F.diu3, d.nD()), diu4(F.diu4, d.nD());
  const Buf diu6(F.diu6, d.nD()), diu7(F.diu7, d.nD()), diu8(F.diu8, d.nD());
  SmIn v;
  v.d1 = Buf(F.diu1, d.nD15())[d.D15(i, j, k)];
  v.d5 = Buf(F.diu5, d.nD15())[d.D15(i, j, k)];
  v.d9 = Buf(F.diu9, d.nD())[d.D(i, j, k)];
  v.d2[0] = diu2[d.D(i - 1, j, k)];
  v.d2[1] = diu2[d.D(i - 1, j + 1, k)];
  v.d2[2] = diu2[d.D(i, j, k)];
  v.d2[3] = diu2[d.D(i, j + 1, k)];
  v.d3[0] = diu3[d.D(i - 1, j, k)];
  v.d3[1] = diu3[d.D(i - 1, j, k + 1)];
  v.d3[2] = diu3[d.D(i, j, k)];
  v.d3[3] = diu3[d.D(i, j, k + 1)];
  v.d4[0] = diu4[d.D(i, j, k)];
  v.d4[1] = diu4[d.D(i, j - 1, k)];
  v.d4[2] = diu4[d.D(i + 1, j, k)];
  v.d4[3] = diu4[d.D(i + 1, j - 1, k)];
  v.d6[0] = diu6[d.D(i, j - 1, k)];
  v.d6[1] = diu6[d.D(i, j - 1, k + 1)];
  v.d6[2] = diu6[d.D(i, j, k)];
  v.d6[3] = diu6[d.D(i, j, k + 1)];
  v.d7[0] = diu7[d.D(i, j, k)];
  v.d7[1] = diu7[d.D(i, j, k - 1)];
  v.d7[2] = diu7[d.D(i + 1, j, k)];
  v.d7[3] = diu7[d.D(i + 1, j, k - 1)];
  v.d8[0] = diu8[d.D(i, j, k)];
  v.d8[1] = diu8[d.D(i, j, k - 1)];
  v.d8[2] = diu8[d.D(i, j + 1, k)];
  v.d8[3] = diu8[d.D(i, j + 1, k - 1)];
  return v;
}
__device__ __forceinline__ float sm_value(const SmIn& v, float delx1k, float cs0) {
  const float q = 0.25f, two = 2.0f;
  float dudxx1 = v.d1;
  float dudyx1 = (v.d2[0] + v.d2[1] + v.d2[2] + v.d2[3]) * q;
  float dudzx1 = (v.d3[0] + v.d3[1] + v.d3[2] + v.d3[3]) * q;
  float dvdxx1 = (v.d4[0] + v.d4[1] + v.d4[2] + v.d4[3]) * q;
  float dvdyx1 = v.d5;
  float dvdzx1 = (v.d6[0] + v.d6[1] + v.d6[2] + v.d6[3]) * q;
  float dwdxx1 = (v.d7[0] + v.d7[1] + v.d7[2] + v.d7[3]) * q;
  float dwdyx1 = (v.d8[0] + v.d8[1] + v.d8[2] + v.d8[3]) * q;
  float dwdzx1 = v.d9;
  float cd = cs0 * delx1k;
  float s = two * (dudxx1 * dudxx1 + dvdyx1 * dvdyx1 + dwdzx1 * dwdzx1) +
            (dudyx1 + dvdxx1) * (dudyx1 + dvdxx1) + (dwdyx1 + dvdzx1) * (dwdyx1 + dvdzx1) +
            (dudzx1 + dwdxx1) * (dudzx1 + dwdxx1);
  return (cd * cd) * sqrtf(s);
}

// A thread takes SM_KPT consecutive levels of one column, every load before the first store
// (which could alias them for the compiler), so the loads two levels share (diu3/6 at k+1 of
// one level are its neighbour's k, diu7/8 at k-1 likewise) are made once.  The levels are
// k0 + q unclamped: past kp the buffer views read zero and nothing is stored.  One level
// measured fastest: two (23 loads a level instead of 27) took 92.8 us against 90.5 at
// 300x300x90.
#ifndef LES_SM_KPT
#define LES_SM_KPT 1
#endif
constexpr int SM_KPT = LES_SM_KPT;
int calc_sm_kblocks(int kp) { return (kp + SM_KPT - 1) / SM_KPT; }
// boundsm.f95:13-32 composed with the interior value val = sm(i,j,k): every halo cell whose
// final value boundsm takes from this cell, i.e. (a,b,c) in A x B x C with A = {i, 0 if i = 1,
// im+1 if i = im}, B likewise in j, C = {k, 0 if k = 1, km+1 if k = km} (sm(.,.,0) = -sm(.,.,1)),
// and, for the row j = 1, the halo cells of row j = -1, whose sources sm(x,-1,z) no kernel
// writes.  The three loops of boundsm leave exactly these values (see k_boundsm).
__device__ inline void boundsm_from(const Dims& d, const Fields& F, int i, int j, int k,
                                    float val) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  const int na = (i == 1) + (i == im), nb = (j == 1) + (j == jm), nc = (k == 1) + (k == km);
  if (na + nb + nc == 0) return;
  int A[3] = {i, i == 1 ? 0 : im + 1, im + 1};
  int B[3] = {j, j == 1 ? 0 : jm + 1, jm + 1};
  int C[3] = {k, k == 1 ? 0 : km + 1, km + 1};
  for (int qa = 0; qa <= na; ++qa)
    for (int qb = 0; qb <= nb; ++qb)
      for (int qc = 0; qc <= nc; ++qc) {
        if (qa + qb + qc == 0) continue;
        const int c = C[qc];
        F.sm[d.S(A[qa], B[qb], c)] = (c == 0) ? -val : val;
      }
  if (j == 1) {  // row j = -1: sm(a,-1,c) for a in A, c in C, from sm(i,-1,.)
    const float s1 = F.sm[d.S(i, -1, k)];
    for (int qa = 0; qa <= na; ++qa)
      for (int qc = 0; qc <= nc; ++qc) {
        if (qa + qc == 0) continue;
        const int c = C[qc];
        F.sm[d.S(A[qa], -1, c)] = (c == 0) ? -s1 : s1;
      }
  }
}

// with_bound: also boundsm's halo cells (boundsm_from), instead of a k_boundsm launch
__device__ __forceinline__ void calc_sm_body(const Dims& d, const Grid1D& g, const Fields& F,
                                             const Physics& ph, const StripGrid& sg,
                                             int with_bound, int L) {
  int bx, by, bz;
  if (!strip_map(sg, L, bx, by, bz)) return;
  const int i = 1 + bx * (int)blockDim.x + threadIdx.x;
  const int j = 1 + by * (int)blockDim.y + threadIdx.y;
  const int k0 = 1 + bz * SM_KPT;
  if (i > d.ip || j > d.jp) return;
  SmIn v[SM_KPT];
#pragma unroll
  for (int q = 0; q < SM_KPT; ++q) v[q] = sm_load(d, F, i, j, k0 + q);
  const Buf bsm(F.sm, d.nS());
#pragma unroll
  for (int q = 0; q < SM_KPT; ++q) {
    const int k = k0 + q;
    if (k > d.kp) break;
    const float val = sm_value(v[q], g.delx1[k - 1], ph.cs0);
    bsm.st(d.S(i, j, k), val);
    if (with_bound) boundsm_from(d, F, i, j, k, val);
  }
}

__global__ __launch_bounds__(256) void k_calc_sm(Dims d, Grid1D g, Fields F, Physics ph,
                                                 StripGrid sg, int with_bound) {
  calc_sm_body(d, g, F, ph, sg, with_bound, xcd_tile((int)blockIdx.x, (int)gridDim.x));
}

void launch_calc_sm(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int with_bound,
                    int shape) {
  const int sx = stencil_shape_x(shape), sy = stencil_shape_y(shape);
  const StripGrid sg = strip_grid((d.ip + sx - 1) / sx, (d.jp + sy - 1) / sy, calc_sm_kblocks(d.kp));
  hipLaunchKernelGGL(k_calc_sm, strip_dim(sg), dim3(sx, sy), 0, st, d, g, F, ph, sg, with_bound);
}

// ==========================================================================================
// boundsm.f95:13-32.  The three loops (i faces, j faces, k planes) are composed: every halo
// cell (a,b,c) takes its final value from cells the kernel never writes.
// ==========================================================================================
__global__ __launch_bounds__(256) void k_boundsm(Dims d, Fields F) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  // the written cells (a = 0..im+1, b = -1..jm+1, c = 0..km+1 with a, b or c on a face) as
  // three disjoint sets: the k planes, the j faces between them, the i faces of the rest
  const long n1 = 2L * (im + 2) * (jm + 3);
  const long n2 = 2L * (im + 2) * km;
  const long n3 = 2L * (jm + 1) * km;
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  int a, b, c;
  if (t < n1) {
    a = (int)(t % (im + 2));
    b = -1 + (int)((t / (im + 2)) % (jm + 3));
    c = (t / ((long)(im + 2) * (jm + 3))) ? km + 1 : 0;
  } else if ((t -= n1) < n2) {
    a = (int)(t % (im + 2));
    c = 1 + (int)((t / (im + 2)) % km);
    b = (t / ((long)(im + 2) * km)) ? jm + 1 : 0;
  } else if ((t -= n2) < n3) {
    const int bb = (int)(t % (jm + 1));  // b in {-1} U [1, jm]
    b = (bb == 0) ? -1 : bb;
    c = 1 + (int)((t / (jm + 1)) % km);
    a = (t / ((long)(jm + 1) * km)) ? im + 1 : 0;
  } else {
    return;
  }
  // V1: after the i-face loop; V2: after the j-face loop
  auto V1 = [&](int x, int y, int z) -> float {
    int xs = (x == 0) ? 1 : ((x == im + 1) ? im : x);
    return F.sm[d.S(xs, y, z)];
  };
  auto V2 = [&](int x, int y, int z) -> float {
    if (y == 0) return V1(x, 1, z);
    if (y == jm + 1) return V1(x, jm, z);
    return V1(x, y, z);
  };
  float val;
  if (c == 0)
    val = -V2(a, b, 1);
  else if (c == km + 1)
    val = V2(a, b, km);
  else
    val = V2(a, b, c);
  F.sm[d.S(a, b, c)] = val;
}

void launch_boundsm(hipStream_t st, Dims d, Fields F) {
  long n = 2L * (d.ip + 2) * (d.jp + 3) + 2L * (d.ip + 2) * d.kp + 2L * (d.jp + 1) * d.kp;
  hipLaunchKernelGGL(k_boundsm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, F);
}

// ==========================================================================================
// les.f95:78-182 (viscous terms) fused with adam.f95:32-46 (Adams-Bashforth)
// ==========================================================================================
// les.f95:78-180: the viscous terms added to f, g, h at interior point (i,j,k); sm(x,y,z)
// returns the eddy viscosity (boundsm applied) at the 15 points it is called with
template <bool P2, class SM>
__device__ __forceinline__ void visc_terms(const Dims& d, const Grid1D& g, const Fields& F, int i,
                                           int j, int k, SM sm, float& fv, float& gv, float& hv) {
  const float two = 2.0f;
  const Buf b1(F.diu1, d.nD15()), b5(F.diu5, d.nD15());
  const Buf b2(F.diu2, d.nD()), b3(F.diu3, d.nD()), b4(F.diu4, d.nD()), b6(F.diu6, d.nD());
  const Buf b7(F.diu7, d.nD()), b8(F.diu8, d.nD()), b9(F.diu9, d.nD());
  auto D1 = [&](int x, int y, int z) -> float { return b1[d.D15(x, y, z)]; };
  auto D5 = [&](int x, int y, int z) -> float { return b5[d.D15(x, y, z)]; };
  auto D2 = [&](int x, int y, int z) -> float { return b2[d.D(x, y, z)]; };
  auto D3 = [&](int x, int y, int z) -> float { return b3[d.D(x, y, z)]; };
  auto D4 = [&](int x, int y, int z) -> float { return b4[d.D(x, y, z)]; };
  auto D6 = [&](int x, int y, int z) -> float { return b6[d.D(x, y, z)]; };
  auto D7 = [&](int x, int y, int z) -> float { return b7[d.D(x, y, z)]; };
  auto D8 = [&](int x, int y, int z) -> float { return b8[d.D(x, y, z)]; };
  auto D9 = [&](int x, int y, int z) -> float { return b9[d.D(x, y, z)]; };
  const float dx1i = g.DX1(i), dx1ip = g.DX1(i + 1), dx1im = g.DX1(i - 1);
  const float dy1j = g.DY1(j), dy1jp = g.DY1(j + 1), dy1jm = g.DY1(j - 1);
  const float dznk = g.DZN(k), dznkp = g.DZN(k + 1), dznkm = g.DZN(k - 1);
  const double2 rznk = g.RZN(k), rzpk = g.RZP(k), rzpkm = g.RZP(k - 1);
  const float s000 = sm(i, j, k), s100 = sm(i + 1, j, k);
  // x-momentum (les.f95:78-109)
  {
    float evsx2 = s100;
    float evsx1 = s000;
    float evsy2 = xdiv<P2>(dy1jp * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)) +
                       dy1j * (xdiv<P2>(dx1ip * sm(i, j + 1, k) + dx1i * sm(i + 1, j + 1, k), dx1i + dx1ip)),
                       dy1j + dy1jp);
    float evsy1 = xdiv<P2>(dy1jp * (xdiv<P2>(dx1ip * sm(i, j - 1, k) + dx1i * sm(i + 1, j - 1, k), dx1i + dx1ip)) +
                       dy1j * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)),
                       dy1j + dy1jp);
    float evsz2 = zdiv(dznkp * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)) +
                       dznk * (xdiv<P2>(dx1ip * sm(i, j, k + 1) + dx1i * sm(i + 1, j, k + 1), dx1i + dx1ip)), dznk + dznkp, rzpk);
    float evsz1 = zdiv(dznk * (xdiv<P2>(dx1ip * sm(i, j, k - 1) + dx1i * sm(i + 1, j, k - 1), dx1i + dx1ip)) +
                       dznkm * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)), dznkm + dznk, rzpkm);
    float visux2 = (evsx2) * two * D1(i + 1, j, k);
    float visux1 = (evsx1) * two * D1(i, j, k);
    float visuy2 = (evsy2) * (D2(i, j + 1, k) + D4(i + 1, j, k));
    float visuy1 = (evsy1) * (D2(i, j, k) + D4(i + 1, j - 1, k));
    float visuz2 = (evsz2) * (D3(i, j, k + 1) + D7(i + 1, j, k));
    float visuz1 = (evsz1) * (D3(i, j, k) + D7(i + 1, j, k - 1));
    float vfu = xdiv<P2>(visux2 - visux1, dx1i) +
                xdiv<P2>(visuy2 - visuy1, dy1j) +
                zdiv(visuz2 - visuz1, dznk, rznk);
    fv = (fv + vfu);
  }
  // y-momentum (les.f95:112-143)
  {
    float evsy2 = sm(i, j + 1, k);
    float evsy1 = s000;
    float evsx2 = xdiv<P2>(dy1jp * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)) +
                       dy1j * (xdiv<P2>(dx1ip * sm(i, j + 1, k) + dx1i * sm(i + 1, j + 1, k), dx1i + dx1ip)),
                       dy1j + dy1jp);
    float evsx1 = xdiv<P2>(dy1jp * (xdiv<P2>(dx1i * sm(i - 1, j, k) + dx1im * s000, dx1im + dx1i)) +
                       dy1j * (xdiv<P2>(dx1i * sm(i - 1, j + 1, k) + dx1im * sm(i, j + 1, k), dx1im + dx1i)),
                       dy1j + dy1jp);
    float evsz2 = zdiv(dznkp * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)) +
                       dznk * (xdiv<P2>(dx1ip * sm(i, j, k + 1) + dx1i * sm(i + 1, j, k + 1), dx1i + dx1ip)), dznk + dznkp, rzpk);
    float evsz1 = zdiv(dznk * (xdiv<P2>(dx1ip * sm(i, j, k - 1) + dx1i * sm(i + 1, j, k - 1), dx1i + dx1ip)) +
                       dznkm * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)), dznkm + dznk, rzpkm);
    float visvx2 = (evsx2) * (D2(i, j + 1, k) + D4(i + 1, j, k));
    float visvx1 = (evsx1) * (D2(i - 1, j + 1, k) + D4(i, j, k));
    float visvy2 = (evsy2) * two * D5(i, j + 1, k);
    float visvy1 = (evsy1) * two * D5(i, j, k);
    float visvz2 = (evsz2) * (D6(i, j, k + 1) + D8(i, j + 1, k));
    float visvz1 = (evsz1) * (D6(i, j, k) + D8(i, j + 1, k - 1));
    float vfv = xdiv<P2>(visvx2 - visvx1, dx1i) +
                xdiv<P2>(visvy2 - visvy1, dy1j) +
                zdiv(visvz2 - visvz1, dznk, rznk);
    gv = (gv + vfv);
  }
  // z-momentum (les.f95:146-180)
  {
    float evsz2 = sm(i, j, k + 1);
    float evsz1 = s000;
    float evsx2 = zdiv(dznkp * (xdiv<P2>(dx1ip * s000 + dx1i * s100, dx1i + dx1ip)) +
                       dznk * (xdiv<P2>(dx1ip * sm(i, j, k + 1) + dx1i * sm(i + 1, j, k + 1), dx1i + dx1ip)), dznk + dznkp, rzpk);
    float evsx1 = zdiv(dznkp * (xdiv<P2>(dx1i * sm(i - 1, j, k) + dx1im * s000, dx1im + dx1i)) +
                       dznk * (xdiv<P2>(dx1i * sm(i - 1, j, k + 1) + dx1im * sm(i, j, k + 1), dx1im + dx1i)), dznk + dznkp, rzpk);
    float evsy2 = zdiv(dznkp * (xdiv<P2>(dy1jp * s000 + dy1j * sm(i, j + 1, k), dy1j + dy1jp)) +
                       dznk * (xdiv<P2>(dy1jp * sm(i, j, k + 1) + dy1j * sm(i, j + 1, k + 1), dy1j + dy1jp)), dznk + dznkp, rzpk);
    float evsy1 = zdiv(dznkp * (xdiv<P2>(dy1j * sm(i, j - 1, k) + dy1jm * s000, dy1jm + dy1j)) +
                       dznk * (xdiv<P2>(dy1j * sm(i, j - 1, k + 1) + dy1jm * sm(i, j, k + 1), dy1jm + dy1j)), dznk + dznkp, rzpk);
    float viswx2 = (evsx2) * (D3(i, j, k + 1) + D7(i + 1, j, k));
    float viswx1 = (evsx1) * (D3(i - 1, j, k + 1) + D7(i, j, k));
    float viswy2 = (evsy2) * (D6(i, j, k + 1) + D8(i, j + 1, k));
    float viswy1 = (evsy1) * (D6(i, j - 1, k + 1) + D8(i, j, k));
    float viswz2 = (evsz2) * two * D9(i, j, k + 1);
    float viswz1 = (evsz1) * two * D9(i, j, k);
    float vfw = xdiv<P2>(viswx2 - viswx1, dx1i) +
                xdiv<P2>(viswy2 - viswy1, dy1j) +
                zdiv(viswz2 - viswz1, dznk, rznk);
    hv = (hv + vfw);
  }
}

// visc (les.f95:78-180) and adam (adam.f95:35-43) on the two cells (i,j,k) and (i,j,k+1) of one
// thread (k = 1 + 2 bz) of the workgroup with logical index L (k_visc_adam, and the ensemble's
// batched launch): every load of both cells comes before any store, so the loads the two share
// (sm at k+1, diu3/6/9 at k+1, ...) are made once (39 loads a cell instead of 45).  With an odd
// kp the last thread's second cell is level kp+1: its loads past the arrays' ends read zero
// (buffer views), and nothing of it is stored.
int visc_kblocks(int kp) { return (kp + 1) / 2; }

template <bool P2>
__device__ __forceinline__ void visc_adam_body(const Dims& d, const Grid1D& g, const Fields& F,
                                               int do_visc, int do_adam, const StripGrid& sg,
                                               int L) {
  int bx, by, bz;
  if (!strip_map(sg, L, bx, by, bz)) return;
  const int i = 1 + bx * (int)blockDim.x + threadIdx.x;
  const int j = 1 + by * (int)blockDim.y + threadIdx.y;
  const int k = 1 + 2 * bz;
  if (i > d.ip || j > d.jp) return;
  const bool two = k + 1 <= d.kp;
  const int k2 = k + 1;
  const long i1 = d.F(i, j, k), i2 = d.F(i, j, k2);
  const Buf bf(F.f, d.nF()), bg(F.g, d.nF()), bh(F.h, d.nF());
  float f1 = bf[i1], g1 = bg[i1], h1 = bh[i1];
  float f2 = bf[i2], g2 = bg[i2], h2 = bh[i2];
  if (do_visc) {
    const Buf bsm(F.sm, d.nS());
    auto sm = [&](int x, int y, int z) -> float { return bsm[d.S(x, y, z)]; };
    visc_terms<P2>(d, g, F, i, j, k, sm, f1, g1, h1);
    visc_terms<P2>(d, g, F, i, j, k2, sm, f2, g2, h2);
  }
  if (do_adam) {  // adam.f95:35-43: fold <- f, f <- 1.5 f - 0.5 fold (g, h alike)
    const long o1 = d.O(i, j, k), o2 = d.O(i, j, k2);
    const Buf bfo(F.fold, d.nO()), bgo(F.gold, d.nO()), bho(F.hold, d.nO());
    const float fo1 = bfo[o1], go1 = bgo[o1], ho1 = bho[o1];
    const float fo2 = bfo[o2], go2 = bgo[o2], ho2 = bho[o2];
    bfo.st(o1, f1);
    bgo.st(o1, g1);
    bho.st(o1, h1);
    if (two) {
      bfo.st(o2, f2);
      bgo.st(o2, g2);
      bho.st(o2, h2);
    }
    f1 = 1.5f * f1 - 0.5f * fo1;
    g1 = 1.5f * g1 - 0.5f * go1;
    h1 = 1.5f * h1 - 0.5f * ho1;
    f2 = 1.5f * f2 - 0.5f * fo2;
    g2 = 1.5f * g2 - 0.5f * go2;
    h2 = 1.5f * h2 - 0.5f * ho2;
  }
  bf.st(i1, f1);
  bg.st(i1, g1);
  bh.st(i1, h1);
  if (two) {
    bf.st(i2, f2);
    bg.st(i2, g2);
    bh.st(i2, h2);
  }
}

template <bool P2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_visc_adam(Dims d, Grid1D g, Fields F, int do_visc,
                                                   int do_adam, StripGrid sg) {
  visc_adam_body<P2>(d, g, F, do_visc, do_adam, sg, xcd_tile((int)blockIdx.x, (int)gridDim.x));
}

void launch_calc_visc_adam(hipStream_t st, Dims d, Grid1D g, Fields F, int do_visc, int do_adam,
                           int shape) {
  const int sx = stencil_shape_x(shape), sy = stencil_shape_y(shape);
  const StripGrid sg = strip_grid((d.ip + sx - 1) / sx, (d.jp + sy - 1) / sy, visc_kblocks(d.kp));
  if (g.p2xy)
    hipLaunchKernelGGL(k_visc_adam<true>, strip_dim(sg), dim3(sx, sy), 0, st, d, g, F, do_visc,
                       do_adam, sg);
  else
    hipLaunchKernelGGL(k_visc_adam<false>, strip_dim(sg), dim3(sx, sy), 0, st, d, g, F, do_visc,
                       do_adam, sg);
}

// ==========================================================================================
// Reductions: blocks accumulate exact limbs in LDS and write them to partials[block]; one
// finalisation workgroup adds all partials (integer, exact) and rounds once.
// ==========================================================================================

int reduction_blocks(Dims d) {
  return ((d.ip + BX - 1) / BX) * ((d.jp + BY - 1) / BY) * RED_KC;
}

// ---- bondFG.f95:15-32 on its own ----
__global__ __launch_bounds__(256) void k_bondfg(Dims d, Fields F) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  const long n1 = (long)jm * km, n2 = (long)im * km, n3 = (long)im * jm;
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n1) {
    int j = 1 + (int)(t % jm), k = 1 + (int)(t / jm);
    F.f[d.F(0, j, k)] = F.f[d.F(1, j, k)];
    return;
  }
  t -= n1;
  if (t < n2) {
    int i = 1 + (int)(t % im), k = 1 + (int)(t / im);
    F.g[d.F(i, 0, k)] = F.g[d.F(i, jm, k)];
    return;
  }
  t -= n2;
  if (t < n3) {
    int i = 1 + (int)(t % im), j = 1 + (int)(t / im);
    F.h[d.F(i, j, 0)] = 0.0f;
    F.h[d.F(i, j, km)] = 0.0f;
  }
}

void launch_bondfg(hipStream_t st, Dims d, Fields F) {
  long n = (long)d.jp * d.kp + (long)d.ip * d.kp + (long)d.ip * d.jp;
  hipLaunchKernelGGL(k_bondfg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, F);
}

// ---- press.f95:49-75: bondfg + rhs + exact partial sums of dx1*dy1*dzn*rhs ----
template <bool P2>
// A tile reads f(0) and g(i,0) from its exchanged halo unless its west face is the domain's,
// and makes bondFG's f and g copies only at the domain's faces (the exchange made the rest).
// rsplit[0] != nullptr: rhs(i,j,k) goes to rsplit[(i+j+k)&1] at ((i-1)>>1) + nr*((j-1) +
// jp*(k-1)) (the split copy's layout) instead of F.rhs.
__global__ __launch_bounds__(256) void k_rhs(Dims d, Grid1D g, Fields F, Physics ph,
                                             int64_t* partials, StripGrid sg, Edges e,
                                             float* rs0, float* rs1, int nr, int64_t* slots) {
  __shared__ unsigned long long lds[NLIMB];
  int bx, by, bz, L;
  if (!strip_block(sg, bx, by, bz, L)) return;  // padding: writes no partial
  block_limbs_init(lds);
  const int i = 1 + bx * BX + threadIdx.x;
  const int j = 1 + by * BY + threadIdx.y;
  const int im = d.ip, jm = d.jp, km = d.kp;
  const int kchunk = (km + RED_KC - 1) / RED_KC;
  const int k0 = 1 + bz * kchunk, k1 = min(km, k0 + kchunk - 1);
  ThreadAcc acc;
  acc.init();
  if (i <= im && j <= jm) {
    const float dx1i = g.DX1(i), dy1j = g.DY1(j), dt = ph.dt;
    const int jm1 = (j == 1 && !e.tiled) ? jm : j - 1;  // g(i,0,k) = g(i,jm,k) (bondFG)
    const Buf bu(F.u, d.nU()), bv(F.v, d.nU()), bw(F.w, d.nW());
    const Buf bf(F.f, d.nF()), bg(F.g, d.nF()), bh(F.h, d.nF()), brh(F.rhs, d.nR());
    // Operands of level k with bondFG's values (f(0) = f(1), g(0) = g(jm), h(0) = h(km) = 0)
    // taken in place of the halo cells, so nothing read here is written by this kernel.
    struct In {
      float um, uc, vm, vc, wm, wc, fc, fm, gc, gm, hc, hm;
    };
    auto load = [&](int k) {
      In t;
      t.um = bu[d.U(i - 1, j, k)];
      t.uc = bu[d.U(i, j, k)];
      t.vm = bv[d.U(i, j - 1, k)];
      t.vc = bv[d.U(i, j, k)];
      t.wm = bw[d.W(i, j, k - 1)];
      t.wc = bw[d.W(i, j, k)];
      const long ifg = d.F(i, j, k);
      t.fc = bf[ifg];
      const float f0 = bf[d.F(i - 1, j, k)];
      t.fm = (i == 1 && e.w) ? t.fc : f0;
      t.gc = bg[ifg];
      t.gm = bg[d.F(i, jm1, k)];
      const float h0 = bh[ifg], h1 = bh[d.F(i, j, k - 1)];
      t.hc = (k == km) ? 0.0f : h0;
      t.hm = (k == 1) ? 0.0f : h1;
      return t;
    };
    // the loads of level k+1 are issued before the stores of level k
    In nxt = load(k0);
    for (int k = k0; k <= k1; ++k) {
      const In c = nxt;
      if (k < k1) nxt = load(k + 1);
      const float dznk = g.DZN(k);
      const double2 rznk = g.RZN(k);
      float r = xdiv<P2>(-c.um + c.uc, dx1i) + xdiv<P2>(-c.vm + c.vc, dy1j) +
                zdiv(-c.wm + c.wc, dznk, rznk);
      r = xdiv<P2>(c.fc - c.fm, dx1i) + xdiv<P2>(c.gc - c.gm, dy1j) +
          zdiv(c.hc - c.hm, dznk, rznk) + (r) / dt;
      if (rs0)
        (((i + j + k) & 1) ? rs1 : rs0)[((i - 1) >> 1) + nr * ((j - 1) + jm * (k - 1))] = r;
      else
        brh.st(d.R(i, j, k), r);
      acc.add(dx1i * dy1j * dznk * r, lds);
    }
    // bondFG: f(0,j,k) = f(1,j,k); g(i,0,k) = g(i,jm,k); h(i,j,0) = h(i,j,km) = 0
    for (int k = k0; k <= k1; ++k) {
      if (i == 1 && e.w) F.f[d.F(0, j, k)] = F.f[d.F(1, j, k)];
      if (j == jm && !e.tiled) F.g[d.F(i, 0, k)] = F.g[d.F(i, jm, k)];
      if (k == 1) F.h[d.F(i, j, 0)] = 0.0f;
      if (k == km) F.h[d.F(i, j, km)] = 0.0f;
    }
  }
  acc.flush(lds);
  if (slots)
    block_limbs_atomic(lds, slots, L);
  else
    block_limbs_store(lds, partials + (long)L * NLIMB);
}

// Also re-arms the SOR scalars for the solve that follows (the work of k_sor_reset).
__global__ __launch_bounds__(FIN_THREADS) void k_rhsav_finalize(const int64_t* partials, int nb,
                                                                DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) {
    sc->rhsav = limbs_to_float(limbs) / sc->area;
    sc->sor_done = 0;
    sc->sor_iters = 0;
    sc->sor = 0.0f;
  }
}

__global__ __launch_bounds__(RED_SLOTS * NLIMB) void k_rhsav_finalize_slots(int64_t* slots,
                                                                           DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_slots(slots, limbs);
  if (threadIdx.x == 0) {
    sc->rhsav = limbs_to_float(limbs) / sc->area;
    sc->sor_done = 0;
    sc->sor_iters = 0;
    sc->sor = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_rhs_adjust(Dims d, Fields F, const DevScalars* sc) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const long ir = d.R(i, j, k);
  F.rhs[ir] = F.rhs[ir] - sc->rhsav;
}

void launch_rhs(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, RedBuf rb,
                DevScalars* sc, Edges e, int finalize, const SplitSOR* split) {
  const StripGrid sg = strip_grid((d.ip + BX - 1) / BX, (d.jp + BY - 1) / BY, RED_KC);
  const int nb = sg.nbx * sg.nby * sg.nbz;
  float* rs0 = split ? split->r[0] : nullptr;
  float* rs1 = split ? split->r[1] : nullptr;
  const int nr = split ? split->nr : 0;
  // a finalised sum goes through the slots; partials stay for callers that combine them
  int64_t* slots = finalize ? rb.slots : nullptr;
  if (g.p2xy)
    hipLaunchKernelGGL(k_rhs<true>, strip_dim(sg), dim3(BX, BY), 0, st, d, g, F, ph, rb.partials,
                       sg, e, rs0, rs1, nr, slots);
  else
    hipLaunchKernelGGL(k_rhs<false>, strip_dim(sg), dim3(BX, BY), 0, st, d, g, F, ph,
                       rb.partials, sg, e, rs0, rs1, nr, slots);
  if (!finalize) return;
  if (slots)
    hipLaunchKernelGGL(k_rhsav_finalize_slots, dim3(1), dim3(RED_SLOTS * NLIMB), 0, st, slots, sc);
  else
    hipLaunchKernelGGL(k_rhsav_finalize, dim3(1), dim3(FIN_THREADS), 0, st, rb.partials, nb, sc);
}

void launch_rhsav_finalize(hipStream_t st, const int64_t* partials, int nb, DevScalars* sc) {
  hipLaunchKernelGGL(k_rhsav_finalize, dim3(1), dim3(FIN_THREADS), 0, st, partials, nb, sc);
}

void launch_rhs_adjust(hipStream_t st, Dims d, Fields F, DevScalars* sc) {
  hipLaunchKernelGGL(k_rhs_adjust, interior_grid(d, d.kp), dim3(BX, BY), 0, st, d, F, sc);
}

// ==========================================================================================
// press.f95:90-100: one red-black colour sweep (in place; a colour reads only the other
// colour and its own cell, so the parallel sweep equals the sequential one bit for bit).
// Threads: x -> every second i of the row's colour, y -> j, z -> k.
// ==========================================================================================
__global__ __launch_bounds__(256) void k_sor_colour(Dims d, Grid1D g, Fields F, Physics ph,
                                                    int nrd, int want_resid, int64_t* partials,
                                                    const DevScalars* sc, SeqTerms tm) {
  __shared__ unsigned long long lds[NLIMB];
  if (sc->sor_done) return;
  if (want_resid) block_limbs_init(lds);
  const int t = blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int i = 1 + ((k + j + nrd) & 1) + 2 * t;
  ThreadAcc acc;
  acc.init();
  if (i <= d.ip && j <= d.jp) {
    const long ip = d.P(i, j, k);
    const long sj = d.ip + 3, sk = (long)(d.ip + 3) * (d.jp + 3);
    const float pc = F.p[ip];
    const float reltmp =
        ph.omega * (F.cn1[d.O(i, j, k)] *
                        (g.cn2l[i - 1] * F.p[ip + 1] + g.cn2s[i - 1] * F.p[ip - 1] +
                         g.cn3l[j - 1] * F.p[ip + sj] + g.cn3s[j - 1] * F.p[ip - sj] +
                         g.cn4l[k - 1] * F.p[ip + sk] + g.cn4s[k - 1] * F.p[ip - sk] -
                         F.rhs[d.R(i, j, k)]) -
                    pc);
    F.p[ip] = pc + reltmp;
    if (want_resid) {
      if (tm.t[0])  // the reference-order sum adds the terms afterwards
        tm.t[(1 + nrd) & 1][((i - 1) >> 1) + tm.nrt * ((j - 1) + d.jp * (k - 1))] = reltmp * reltmp;
      else
        acc.add(reltmp * reltmp, lds);
    }
  }
  if (want_resid) {  // into the accumulation slots (partials: RedBuf::slots)
    acc.flush(lds);
    const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_limbs_atomic(lds, partials, blk);
  }
}

int sor_resid_blocks(Dims d) {
  int half = (d.ip + 1) / 2;
  return ((half + BX - 1) / BX) * ((d.jp + BY - 1) / BY) * d.kp;
}

void launch_sor_colour(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int nrd,
                       int want_resid, RedBuf rb, DevScalars* sc, SeqTerms tm) {
  int half = (d.ip + 1) / 2;
  dim3 grid((half + BX - 1) / BX, (d.jp + BY - 1) / BY, d.kp);
  hipLaunchKernelGGL(k_sor_colour, grid, dim3(BX, BY), 0, st, d, g, F, ph, nrd, want_resid,
                     rb.slots, sc, tm);
}

// residual of one SOR iteration (both colours) and the early exit of press.f95:114
// the residual of the iteration's two colour sweeps from the accumulation slots (cleared as
// they are read), and the early exit of press.f95:114
__global__ __launch_bounds__(FIN_THREADS) void k_sor_resid_finalize(int64_t* slots, Physics ph,
                                                                    DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  if (sc->sor_done) return;  // (the sweeps returned too: the slots are zero)
  sum_slots(slots, limbs);
  if (threadIdx.x == 0) {
    float s = limbs_to_float(limbs);
    sc->sor = s;
    sc->sor_iters += 1;
    if (s < ph.pjuge) sc->sor_done = 1;
  }
}

void launch_sor_resid_finalize(hipStream_t st, RedBuf rb, Physics ph, DevScalars* sc) {
  hipLaunchKernelGGL(k_sor_resid_finalize, dim3(1), dim3(FIN_THREADS), 0, st, rb.slots, ph, sc);
}

__global__ void k_sor_reset(DevScalars* sc) {
  sc->sor_done = 0;
  sc->sor_iters = 0;
  sc->sor = 0.0f;
}

void launch_sor_reset(hipStream_t st, DevScalars* sc) {
  hipLaunchKernelGGL(k_sor_reset, dim3(1), dim3(1), 0, st, sc);
}

// boundp.f95:23-44 (boundp1): i faces for j = 0..jm+1 then j faces for i = 0..im+1, at every
// k = 0..km+1; composed so that every source is an interior (i,j) column.
// A tile copies only at the domain's i faces and makes no periodic j copies (the exchange
// after this kernel fills its internal faces and j faces).
__global__ __launch_bounds__(256) void k_boundp1(Dims d, Fields F, const DevScalars* sc,
                                                 int check_done, Edges e) {
  if (check_done && sc->sor_done) return;
  const int im = d.ip, jm = d.jp, km = d.kp;
  const long nface_i = (long)jm * (km + 2) * 2;       // a in {0,im+1}, b = 1..jm
  const long nface_j = (long)(im + 2) * (km + 2) * 2; // b in {0,jm+1}, a = 0..im+1
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < nface_i) {
    const int b = 1 + (int)(t % jm);
    const int c = (int)((t / jm) % (km + 2));
    const int side = (int)(t / ((long)jm * (km + 2)));
    const int a = side ? im + 1 : 0, as = side ? im : 1;
    if (side ? e.e : e.w) F.p[d.P(a, b, c)] = F.p[d.P(as, b, c)];
    return;
  }
  t -= nface_i;
  if (t >= nface_j || e.tiled) return;
  const int a = (int)(t % (im + 2));
  const int c = (int)((t / (im + 2)) % (km + 2));
  const int side = (int)(t / ((long)(im + 2) * (km + 2)));
  const int b = side ? jm + 1 : 0, bs = side ? 1 : jm;
  const int as = (a == 0) ? 1 : ((a == im + 1) ? im : a);
  F.p[d.P(a, b, c)] = F.p[d.P(as, bs, c)];
}

// boundp.f95:5-21 (boundp2): k planes from k = 1 and k = km for i = 0..im+1, j = 0..jm+1
__global__ __launch_bounds__(256) void k_boundp2(Dims d, Fields F, const DevScalars* sc,
                                                 int check_done) {
  if (check_done && sc->sor_done) return;
  const int im = d.ip, jm = d.jp, km = d.kp;
  const long n = (long)(im + 2) * (jm + 2);
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * n) return;
  const int top = (int)(t / n);
  t %= n;
  const int a = (int)(t % (im + 2)), b = (int)(t / (im + 2));
  if (top)
    F.p[d.P(a, b, km + 1)] = F.p[d.P(a, b, km)];
  else
    F.p[d.P(a, b, 0)] = F.p[d.P(a, b, 1)];
}

void launch_boundp1(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done,
                    Edges e) {
  long n = (long)d.jp * (d.kp + 2) * 2 + (long)(d.ip + 2) * (d.kp + 2) * 2;
  hipLaunchKernelGGL(k_boundp1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, F, sc,
                     check_done, e);
}

void launch_boundp2(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done) {
  long n = 2L * (d.ip + 2) * (d.jp + 2);
  hipLaunchKernelGGL(k_boundp2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, F, sc,
                     check_done);
}

// ==========================================================================================
// press.f95:123-141: pav = sum(p*dx1*dy1*dzn)/pco; p -= pav; then boundp1 + boundp2 which,
// composed, set every halo cell of p to its mirror interior cell.
// ==========================================================================================
__global__ __launch_bounds__(256) void k_pav(Dims d, Grid1D g, Fields F, int64_t* partials) {
  __shared__ unsigned long long lds[NLIMB];
  block_limbs_init(lds);
  const int i = 1 + blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int km = d.kp;
  const int kchunk = (km + RED_KC - 1) / RED_KC;
  const int k0 = 1 + blockIdx.z * kchunk, k1 = min(km, k0 + kchunk - 1);
  ThreadAcc acc;
  acc.init();
  if (i <= d.ip && j <= d.jp) {
    const float dx1i = g.DX1(i), dy1j = g.DY1(j);
    for (int k = k0; k <= k1; ++k) acc.add(F.p[d.P(i, j, k)] * dx1i * dy1j * g.DZN(k), lds);
  }
  acc.flush(lds);
  const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_limbs_store(lds, partials + (long)blk * NLIMB);
}

__global__ __launch_bounds__(FIN_THREADS) void k_pav_finalize(const int64_t* partials, int nb,
                                                      DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) sc->pav = limbs_to_float(limbs) / sc->pco;
}

void launch_pav(hipStream_t st, Dims d, Grid1D g, Fields F, RedBuf rb, DevScalars* sc,
                int finalize) {
  dim3 grid = interior_grid(d, RED_KC);
  int nb = (int)(grid.x * grid.y * grid.z);
  hipLaunchKernelGGL(k_pav, grid, dim3(BX, BY), 0, st, d, g, F, rb.partials);
  if (finalize)
    hipLaunchKernelGGL(k_pav_finalize, dim3(1), dim3(FIN_THREADS), 0, st, rb.partials, nb, sc);
}

void launch_pav_finalize(hipStream_t st, const int64_t* partials, int nb, DevScalars* sc) {
  hipLaunchKernelGGL(k_pav_finalize, dim3(1), dim3(FIN_THREADS), 0, st, partials, nb, sc);
}

// p -= pav on the interior; with_boundp also writes every mirror image of the cell
// (boundp1 then boundp2 leave p(a,b,c) = p(M(a),M(b),M(c)) with M(0)=1, M(im+1)=im in i,
// M(0)=jm, M(jm+1)=1 in j, M(0)=1, M(km+1)=km in k).
__global__ __launch_bounds__(256) void k_adj(Dims d, Fields F, const DevScalars* sc,
                                             int with_boundp) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int im = d.ip, jm = d.jp, km = d.kp;
  if (i > im || j > jm) return;
  const long ip = d.P(i, j, k);
  const float val = F.p[ip] - sc->pav;
  F.p[ip] = val;
  if (!with_boundp) return;
  int ai[2], bj[2], ck[2];
  int na = 0, nb = 0, nc = 0;
  ai[na++] = i;
  if (i == 1) ai[na++] = 0;
  if (i == im) ai[na++] = im + 1;  // (im >= 2: at most one of the two)
  bj[nb++] = j;
  if (j == jm) bj[nb++] = 0;
  if (j == 1) bj[nb++] = jm + 1;
  ck[nc++] = k;
  if (k == 1) ck[nc++] = 0;
  if (k == km) ck[nc++] = km + 1;
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y)
      for (int z = 0; z < nc; ++z)
        if (x | y | z) F.p[d.P(ai[x], bj[y], ck[z])] = val;
}

void launch_adj(hipStream_t st, Dims d, Fields F, DevScalars* sc, int with_boundp) {
  hipLaunchKernelGGL(k_adj, interior_grid(d, d.kp), dim3(BX, BY), 0, st, d, F, sc, with_boundp);
}

// ==========================================================================================
// Exact sum of a device vector (the reduction on its own, for testing)
// ==========================================================================================
__global__ __launch_bounds__(256) void k_exact_sum_part(const float* x, long n, int64_t* partials) {
  __shared__ unsigned long long lds[NLIMB];
  block_limbs_init(lds);
  ThreadAcc acc;
  acc.init();
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256)
    acc.add(x[t], lds);
  acc.flush(lds);
  block_limbs_store(lds, partials + (long)blockIdx.x * NLIMB);
}

__global__ __launch_bounds__(FIN_THREADS) void k_exact_sum_fin(const int64_t* partials, int nb, float* out) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) *out = limbs_to_float(limbs);
}

void launch_exact_sum(hipStream_t st, const float* x, long n, RedBuf rb, float* out) {
  int nb = (int)min((long)rb.max_blocks, max(1L, (n + 255) / 256));
  hipLaunchKernelGGL(k_exact_sum_part, dim3(nb), dim3(256), 0, st, x, n, rb.partials);
  hipLaunchKernelGGL(k_exact_sum_fin, dim3(1), dim3(FIN_THREADS), 0, st, rb.partials, nb, out);
}

// ==========================================================================================
// feedbfm.f95:42-63 on the device: amask1 = 0 where zbm(i,j) > z2(k)+0.5*dzn(k), else 1;
// bmask1 = cmask1 = dmask1 = 1 - amask1 over the interior (halos are cleared by the caller),
// and the interior bit mask the stencils read.  zbm(-1:ip+1,-1:jp+1), z2(1:kp+2).
// ==========================================================================================
__global__ __launch_bounds__(256) void k_feedbfm(Dims d, Grid1D g, const float* z2,
                                                 const float* zbm, float* am, float* bm, float* cm,
                                                 float* dm, uint8_t* bits) {
  const int i = 1 + blockIdx.x * BX + threadIdx.x;
  const int j = 1 + blockIdx.y * BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int ip = d.ip, jp = d.jp;
  const float hb = zbm[(i + 1) + (long)(ip + 3) * (j + 1)];
  const bool solid = hb > z2[k - 1] + 0.5f * g.DZN(k);
  const float a = solid ? 0.0f : 1.0f, o = solid ? 1.0f : 0.0f;
  const long ia = i + (long)(ip + 2) * (j + (long)(jp + 2) * k);
  am[ia] = a;
  dm[ia] = o;
  bm[(i + 1) + (long)(ip + 3) * (j + (long)(jp + 2) * k)] = o;
  cm[i + (long)(ip + 2) * ((j + 1) + (long)(jp + 3) * k)] = o;
  bits[d.M(i, j, k)] = solid ? 7 : 8;  // b|c|d or a
}

void launch_feedbfm(hipStream_t st, Dims d, Grid1D g, const float* z2, const float* zbm,
                    float* am, float* bm, float* cm, float* dm, uint8_t* bits) {
  hipLaunchKernelGGL(k_feedbfm, interior_grid(d, d.kp), dim3(BX, BY), 0, st, d, g, z2, zbm, am,
                     bm, cm, dm, bits);
}

// ==========================================================================================
// Halo exchange pack/unpack: halo_exchange.cl:3-51 (write), :156-185 (read) and
// halo_exchange_correct.cl:156-185 (read, inner ring without corners), with the OpenCL
// vector components replaced by separate fields (buffer block v = field v).
// Thread t -> (x, k) with x along the row/column, i.e. the reference's gl_id mapping.
// ==========================================================================================
struct HaloPtrs {
  float* a[16];
};

__global__ __launch_bounds__(256) void k_halo(HaloPtrs arr, int v_dim, int im, int jm, int km,
                                              int hw, int correct, int write, float* buffer) {
  const long L = (long)max(im, jm);
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= L * km) return;
  const int k = (int)(t / L);
  const int x = (int)(t % L);
  const long plane = (long)im * jm;
  if (!correct) {
    const long v_sz = 2L * km * im * hw + 2L * km * (jm - 2 * hw) * hw;
    for (int v = 0; v < v_dim; ++v) {
      float* A = arr.a[v];
      float* B = buffer + v * v_sz;
      for (int w = 0; w < hw; ++w) {
        if (x < im) {
          long an = x + (long)im * w + k * plane, bn = x + (long)im * w + (long)k * im * hw;
          long as = an + (long)im * (jm - hw), bs = (long)km * im * hw + bn;
          if (write) {
            A[an] = B[bn];
            A[as] = B[bs];
          } else {
            B[bn] = A[an];
            B[bs] = A[as];
          }
        }
        if (x < jm - hw && x > hw - 1) {
          long aw = (long)x * im + w + k * plane;
          long bw = 2L * km * im * hw + (long)(x - hw) * hw + w + (long)k * (jm - 2 * hw) * hw;
          long ae = aw + im - hw;
          long be = bw + (long)km * (jm - 2 * hw) * hw;
          if (write) {
            A[aw] = B[bw];
            A[ae] = B[be];
          } else {
            B[bw] = A[aw];
            B[be] = A[ae];
          }
        }
      }
    }
  } else {  // read only: inner ring, corners excluded
    const long v_sz = 2L * km * (im + jm - 4 * hw) * hw;
    for (int v = 0; v < v_dim; ++v) {
      const float* A = arr.a[v];
      float* B = buffer + v * v_sz;
      for (int w = 0; w < hw; ++w) {
        if (x < im - hw && x > hw - 1) {
          long bn = (x - hw) + (long)(im - 2 * hw) * w + (long)k * (im - 2 * hw) * hw;
          B[bn] = A[(long)im * hw + x + (long)im * w + k * plane];
          B[(long)km * (im - 2 * hw) * hw + bn] = A[x + k * plane + (long)im * (jm - 2 * hw) + (long)im * w];
        }
        if (x < jm - hw && x > hw - 1) {
          long bw = 2L * km * (im - 2 * hw) * hw + (long)(x - hw) * hw + w + (long)k * (jm - 2 * hw) * hw;
          B[bw] = A[(long)x * im + w + hw + k * plane];
          B[bw + (long)km * (jm - 2 * hw) * hw] = A[(long)x * im + w + k * plane + im - 2 * hw];
        }
      }
    }
  }
}

void launch_halo_read(hipStream_t st, const float* const* arrays, int v_dim, int im, int jm,
                      int km, int h_w, int correct, float* buffer) {
  HaloPtrs hp;
  for (int v = 0; v < v_dim && v < 16; ++v) hp.a[v] = const_cast<float*>(arrays[v]);
  long n = (long)max(im, jm) * km;
  hipLaunchKernelGGL(k_halo, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hp, v_dim, im,
                     jm, km, h_w, correct, 0, buffer);
}

void launch_halo_write(hipStream_t st, float* const* arrays, int v_dim, int im, int jm, int km,
                       int h_w, const float* buffer) {
  HaloPtrs hp;
  for (int v = 0; v < v_dim && v < 16; ++v) hp.a[v] = arrays[v];
  long n = (long)max(im, jm) * km;
  hipLaunchKernelGGL(k_halo, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hp, v_dim, im,
                     jm, km, h_w, 0, 1, const_cast<float*>(buffer));
}

// ==========================================================================================
// Tiles of one domain on one device (les_tiled): face exchange and scalar combination
// ==========================================================================================

// Every box of the plan in one launch: blockIdx.y selects the box, threads run over its cells
// (i fastest).  Source and destination boxes never overlap a box written in the same launch.
__global__ __launch_bounds__(256) void k_exchange(ExchangePlan plan) {
  const BoxCopy& b = plan.box[blockIdx.y];
  const long n = (long)b.ni * b.nj * b.nk;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int a = (int)(t % b.ni);
    const int c = (int)((t / b.ni) % b.nj);
    const int z = b.k0 + (int)(t / ((long)b.ni * b.nj));
    const ArrView& S = b.src;
    const ArrView& D = b.dst;
    const float v = S.p[(b.si + a - S.lo0) + (long)S.n0 * ((b.sj + c - S.lo1) + (long)S.n1 * (z - S.lo2))];
    D.p[(b.di + a - D.lo0) + (long)D.n0 * ((b.dj + c - D.lo1) + (long)D.n1 * (z - D.lo2))] = v;
  }
}

void launch_exchange(hipStream_t st, const ExchangePlan& plan) {
  if (plan.nbox <= 0) return;
  long most = 0;
  for (int q = 0; q < plan.nbox; ++q)
    most = std::max(most, (long)plan.box[q].ni * plan.box[q].nj * plan.box[q].nk);
  const unsigned gx = (unsigned)std::min<long>((most + 255) / 256, 1024);
  hipLaunchKernelGGL(k_exchange, dim3(gx, plan.nbox), dim3(256), 0, st, plan);
}

struct ScalarDst {
  DevScalars* p[8];
};

__global__ void k_scalar_bcast(const DevScalars* src, ScalarDst dst, int ndst, int off) {
  const int q = threadIdx.x;
  if (q >= ndst) return;
  const float v = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(src) + off);
  *reinterpret_cast<float*>(reinterpret_cast<char*>(dst.p[q]) + off) = v;
}

void launch_scalar_bcast(hipStream_t st, const DevScalars* src, DevScalars* const* dst, int ndst,
                         int off) {
  ScalarDst d{};
  for (int q = 0; q < ndst && q < 8; ++q) d.p[q] = dst[q];
  hipLaunchKernelGGL(k_scalar_bcast, dim3(1), dim3(64), 0, st, src, d, std::min(ndst, 8), off);
}

// bondv1.f95:67-75 over the union of the east tiles' planes: max and min are exact and order
// free, and every tile's max/min already includes the 0 seed of amax1/amin1
__global__ void k_uout_combine(ScalarDst east, int neast, ScalarDst all, int nall) {
  if (threadIdx.x != 0) return;
  float mx = east.p[0]->uout_max, mn = east.p[0]->uout_min;
  for (int q = 1; q < neast; ++q) {
    mx = fmaxf(mx, east.p[q]->uout_max);
    mn = fminf(mn, east.p[q]->uout_min);
  }
  const float u = (mx + mn) / 2.0f;
  for (int q = 0; q < nall; ++q) all.p[q]->uout = u;
}

void launch_uout_combine(hipStream_t st, DevScalars* const* east, int neast,
                         DevScalars* const* all, int nall) {
  ScalarDst e{}, a{};
  for (int q = 0; q < neast && q < 8; ++q) e.p[q] = east[q];
  for (int q = 0; q < nall && q < 8; ++q) a.p[q] = all[q];
  hipLaunchKernelGGL(k_uout_combine, dim3(1), dim3(64), 0, st, e, std::min(neast, 8), a,
                     std::min(nall, 8));
}

// boundsm.f95:13-32 on a tile, in the order the reference composes it, with the exchanges
// between the phases: phase 0 copies the i faces that are the domain's (rows j = 1..jm),
// phase 1 the j faces that are the domain's (columns i = 0..im+1, after the i faces have been
// exchanged), phase 2 the k planes of every column (a, b) of 0..im+1 x 0..jm+1 from the then
// complete levels 1 and km.
__global__ __launch_bounds__(256) void k_boundsm_tiled(Dims d, Fields F, Edges e, int phase) {
  const int im = d.ip, jm = d.jp, km = d.kp;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (phase == 0) {
    if (t >= 2L * jm * km) return;
    const int b = 1 + (int)(t % jm), c = 1 + (int)((t / jm) % km), side = (int)(t / ((long)jm * km));
    if (side ? e.e : e.w) F.sm[d.S(side ? im + 1 : 0, b, c)] = F.sm[d.S(side ? im : 1, b, c)];
  } else if (phase == 1) {
    if (t >= 2L * (im + 2) * km) return;
    const int a = (int)(t % (im + 2)), c = 1 + (int)((t / (im + 2)) % km);
    const int side = (int)(t / ((long)(im + 2) * km));
    if (side ? e.n : e.s) F.sm[d.S(a, side ? jm + 1 : 0, c)] = F.sm[d.S(a, side ? jm : 1, c)];
  } else {
    if (t >= 2L * (im + 2) * (jm + 2)) return;
    const int a = (int)(t % (im + 2)), b = (int)((t / (im + 2)) % (jm + 2));
    const int top = (int)(t / ((long)(im + 2) * (jm + 2)));
    if (top)
      F.sm[d.S(a, b, km + 1)] = F.sm[d.S(a, b, km)];
    else
      F.sm[d.S(a, b, 0)] = -F.sm[d.S(a, b, 1)];
  }
}

void launch_boundsm_tiled(hipStream_t st, Dims d, Fields F, Edges e, int phase) {
  const long n = phase == 0 ? 2L * d.jp * d.kp
                            : (phase == 1 ? 2L * (d.ip + 2) * d.kp : 2L * (d.ip + 2) * (d.jp + 2));
  hipLaunchKernelGGL(k_boundsm_tiled, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, F,
                     e, phase);
}

// ---- batched launches: workgroup w of the grid (after the XCD remap) belongs to member
// wgm[w] and is that member's logical workgroup w - base ----
template <bool P2>
__global__ __launch_bounds__(512) void k_velfg_batch(const BatchStencil* __restrict__ tab,
                                                     const uint8_t* __restrict__ wgm, int total) {
  const int w = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (w >= total) return;
  // the member's operands into registers once (the body's stores cannot alias them there)
  const BatchStencil& t = tab[wgm[w]];
  const Dims d = t.d;
  const Grid1D g = t.g;
  const Fields F = t.F;
  const Physics ph = t.ph;
  const StripGrid sg = t.sg_velfg;
  velfg_xy_body<P2>(d, g, F, ph, 1, sg, EDGES_ALL, w - t.base_velfg);
}

__global__ __launch_bounds__(256) void k_calc_sm_batch(const BatchStencil* __restrict__ tab,
                                                       const uint8_t* __restrict__ wgm, int total) {
  const int w = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (w >= total) return;
  const BatchStencil& t = tab[wgm[w]];
  const Dims d = t.d;
  const Grid1D g = t.g;
  const Fields F = t.F;
  const Physics ph = t.ph;
  const StripGrid sg = t.sg_sm;
  calc_sm_body(d, g, F, ph, sg, 1, w - t.base_sm);
}

template <bool P2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_visc_adam_batch(const BatchStencil* __restrict__ tab,
                                                         const uint8_t* __restrict__ wgm,
                                                         int total) {
  const int w = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (w >= total) return;
  const BatchStencil& t = tab[wgm[w]];
  const Dims d = t.d;
  const Grid1D g = t.g;
  const Fields F = t.F;
  const StripGrid sg = t.sg_visc;
  visc_adam_body<P2>(d, g, F, 1, 1, sg, w - t.base_visc);
}

void launch_velfg_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm, int total,
                        int p2xy) {
  const dim3 grid((unsigned)((total + 7) / 8 * 8));
  if (p2xy)
    hipLaunchKernelGGL(k_velfg_batch<true>, grid, dim3(VX, VY), 0, st, tab, wgm, total);
  else
    hipLaunchKernelGGL(k_velfg_batch<false>, grid, dim3(VX, VY), 0, st, tab, wgm, total);
}

void launch_calc_sm_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm, int total,
                          int shape) {
  const dim3 grid((unsigned)((total + 7) / 8 * 8));
  hipLaunchKernelGGL(k_calc_sm_batch, grid, dim3(stencil_shape_x(shape), stencil_shape_y(shape)),
                     0, st, tab, wgm, total);
}

void launch_visc_adam_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm,
                            int total, int shape, int p2xy) {
  const dim3 grid((unsigned)((total + 7) / 8 * 8));
  const dim3 blk(stencil_shape_x(shape), stencil_shape_y(shape));
  if (p2xy)
    hipLaunchKernelGGL(k_visc_adam_batch<true>, grid, blk, 0, st, tab, wgm, total);
  else
    hipLaunchKernelGGL(k_visc_adam_batch<false>, grid, blk, 0, st, tab, wgm, total);
}

}  // namespace les
