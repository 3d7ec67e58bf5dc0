// Mixed-precision pressure solve (LES_PREC_MIXED): the stencils and the SOR sweeps stay FP32;
// the rhs mean, the residual norm, the pressure iterate and its mean run in FP64.
//
// press.f95:49-146 restated as iterative refinement of the discrete Poisson problem that the
// SOR of press.f95:91-100 relaxes.  Its fixed point p = cn1*(sum_nb c_nb*p_nb - rhs) is
// L p = rhs with L p = sum_nb c_nb*p_nb - p/cn1 (cn2l/cn2s along i, cn3l/cn3s along j,
// cn4l/cn4s along k, init.f95:82-107), mirror images across the i and k faces and periodic j
// (boundp.f95).  The solve:
//   rhs64 = rhs - rhsav64                 (FP64 mean of vol*rhs, exact sum rounded to FP64)
//   rhs64 = rhs64 - sum(w rhs64)/sum(w)   (the solvable part; w below, vol())
//   p32   = SOR_nmaxp(p, rhs64)           (the FP32 colour-split sweeps)
//   p64   = p32
//   repeat: r64 = rhs64 - L p64 (FP64);  stop when |r64|_2 <= tol * |rhs64|_2
//           e32 = SOR_nmaxp(0, r64)       (FP32 sweeps on the correction)
//           p64 = p64 + e32
//   p = p64 - pav64                        (FP64 mean of vol*p64), with every boundary image.
// In exact arithmetic SOR from p on rhs equals p plus SOR from 0 on rhs - L p, so each pass
// continues the reference iteration while FP64 carries the accumulated solution.  All state is
// in the colour-split layout of les_sor.hip (interior cells, colour (i+j+k)&1).
#include "les_mixed.h"

#include <algorithm>
#include <utility>

namespace les {

namespace {

__device__ __forceinline__ int split_index(const SplitSOR& sp, int i, int j, int k) {
  return ((i - 1) >> 1) + sp.nr * ((j - 1) + sp.nj * (k - 1));
}

__device__ __forceinline__ double p64_at(const SplitSOR& sp, const Mixed& mx, int i, int j, int k) {
  return mx.p64[(i + j + k) & 1][split_index(sp, i, j, k)];
}

__device__ __forceinline__ int tid() { return threadIdx.x + blockDim.x * threadIdx.y; }

// Block sum of one double per thread (256 threads, 4 wavefronts), in a fixed order; the total
// is valid in thread 0.  Every thread of the block must call it.
__device__ double block_sum(double v, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = tid() & 63, wave = tid() >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (tid() == 0) s = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  return s;
}

constexpr int MX_BX = 64, MX_BY = 4;

inline dim3 mx_grid(Dims d) { return dim3((d.ip + MX_BX - 1) / MX_BX, (d.jp + MX_BY - 1) / MX_BY, d.kp); }
inline int mx_blocks(Dims d) {
  dim3 g = mx_grid(d);
  return (int)(g.x * g.y * g.z);
}

__device__ __forceinline__ int flat_block() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// The weight w = (dxs(i-1)+dxs(i)) (dys(j-1)+dys(j)) (dzs(k-1)+dzs(k)) in which L (boundary
// images folded) is symmetric: w(i) cn2l(i) = w(i+1) cn2s(i+1) = 2/dxs(i) with init.f95:82-107's
// coefficients, alike along j and k.  The constants are L's null space and w its left null
// vector, so L p = rhs has a solution only if sum w rhs = 0.
__device__ __forceinline__ double vol(const Grid1D& g, int i, int j, int k) {
  return ((double)g.DXS(i - 1) + (double)g.DXS(i)) * ((double)g.DYS(j - 1) + (double)g.DYS(j)) *
         ((double)g.DZS(k - 1) + (double)g.DZS(k));
}

}  // namespace

// ---- rhsav in FP64 from the exact partial sums of k_rhs (press.f95:65-75) ----
__global__ __launch_bounds__(256) void k_mx_rhsav(const int64_t* partials, int nb, DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) {
    sc->rhsav64 = limbs_to_double(limbs) / sc->area64;
    sc->rhsav = (float)sc->rhsav64;
  }
}

// ---- p -> split FP32; rhs64 = rhs - rhsav64 in FP64; block sums of w*rhs64 and w ----
__global__ __launch_bounds__(256) void k_mx_scatter(Dims d, Grid1D g, Fields F, SplitSOR sp,
                                                    Mixed mx, const DevScalars* sc) {
  __shared__ double lds[4];
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double a = 0.0, b = 0.0;
  if (i <= d.ip && j <= d.jp) {
    const int par = (i + j + k) & 1;
    const int is = split_index(sp, i, j, k);
    sp.p[par][is] = F.p[d.P(i, j, k)];
    const double r = (double)F.rhs[d.R(i, j, k)] - sc->rhsav64;
    mx.r64[par][is] = r;
    b = vol(g, i, j, k);
    a = b * r;
  }
  const double sa = block_sum(a, lds);
  if (tid() == 0) mx.part[flat_block()] = sa;
  __syncthreads();
  const double sb = block_sum(b, lds);
  if (tid() == 0) mx.part[mx.nblk + flat_block()] = sb;
}

// ---- the solvable rhs: rhs64 -= sum(w rhs64) / sum(w) (cg[7]); the sweeps get its FP32
// rounding, and so does rhs (the right-hand side the step solved) ----
__global__ __launch_bounds__(256) void k_mx_wmean(Mixed mx) {
  __shared__ double lds[4];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < mx.nblk; q += 256) {
    a += mx.part[q];
    b += mx.part[mx.nblk + q];
  }
  const double sa = block_sum(a, lds);
  __syncthreads();
  const double sb = block_sum(b, lds);
  if (threadIdx.x == 0) mx.cg[7] = (sb > 0.0) ? sa / sb : 0.0;
}

__global__ __launch_bounds__(256) void k_mx_project_rhs(Dims d, Fields F, SplitSOR sp, Mixed mx) {
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
  const double r = mx.r64[par][is] - mx.cg[7];
  mx.r64[par][is] = r;
  sp.r[par][is] = (float)r;
  F.rhs[d.R(i, j, k)] = (float)r;
}

// ---- p64 = p32 ----
__global__ __launch_bounds__(256) void k_mx_promote(Dims d, SplitSOR sp, Mixed mx) {
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
  mx.p64[par][is] = (double)sp.p[par][is];
}

// ---- r = rhs64 - L p64 in FP64; sweeps get (float)r and a zero correction; block sums of
// r^2 (and of rhs64^2 when first != 0).  L in FP64 in flux form, L p = sum over the real
// neighbours of c_nb (p_nb - p_c), with init.f95:82-107's coefficients evaluated in FP64 from the
// grid spacings: in exact arithmetic 1/cn1 is the sum of the c_nb, so this is the operator whose
// fixed point the sweeps relax, with the constants exactly in its null space and w exactly its
// left null vector (the FP32 coefficients and cn1 break both at rounding level, which left a
// residual of (constant part of p) x (FP32 row sums) no correction could remove). ----
__global__ __launch_bounds__(256) void k_mx_resid(Dims d, Grid1D g, SplitSOR sp, Mixed mx,
                                                  int first) {
  __shared__ double lds[4];
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double rr = 0.0, bb = 0.0;
  if (i <= d.ip && j <= d.jp) {
    const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
    const double pc = mx.p64[par][is];
    const double xm = g.DXS(i - 1), xp = g.DXS(i), ym = g.DYS(j - 1), yp = g.DYS(j);
    const double zm = g.DZS(k - 1), zp = g.DZS(k);
    // mirror images across the i and k faces contribute p_c - p_c = 0
    double lp = 0.0;
    if (i < d.ip) lp += 2.0 / (xp * (xm + xp)) * (p64_at(sp, mx, i + 1, j, k) - pc);
    if (i > 1) lp += 2.0 / (xm * (xm + xp)) * (p64_at(sp, mx, i - 1, j, k) - pc);
    lp += 2.0 / (yp * (ym + yp)) * (p64_at(sp, mx, i, j == d.jp ? 1 : j + 1, k) - pc);
    lp += 2.0 / (ym * (ym + yp)) * (p64_at(sp, mx, i, j == 1 ? d.jp : j - 1, k) - pc);
    if (k < d.kp) lp += 2.0 / (zp * (zm + zp)) * (p64_at(sp, mx, i, j, k + 1) - pc);
    if (k > 1) lp += 2.0 / (zm * (zm + zp)) * (p64_at(sp, mx, i, j, k - 1) - pc);
    const double b = mx.r64[par][is];
    const double r = b - lp;
    sp.r[par][is] = (float)r;
    sp.p[par][is] = 0.0f;
    rr = r * r;
    bb = b * b;
  }
  const double s = block_sum(rr, lds);
  if (tid() == 0) mx.part[flat_block()] = s;
  if (first) {
    __syncthreads();
    const double t = block_sum(bb, lds);
    if (tid() == 0) mx.part[mx.nblk + flat_block()] = t;
  }
}

// ---- fixed-order sum of the block partials; relative residual and the stop flag ----
__global__ __launch_bounds__(256) void k_mx_resid_finalize(Mixed mx, int first, double tol,
                                                           DevScalars* sc) {
  __shared__ double lds[4];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < mx.nblk; q += 256) {
    a += mx.part[q];
    if (first) b += mx.part[mx.nblk + q];
  }
  const double sa = block_sum(a, lds);
  __syncthreads();
  const double sb = block_sum(b, lds);
  if (threadIdx.x == 0) {
    if (first) {
      sc->rhs_l2 = sqrt(sb);
      mx.cg[6] = 0.0;  // the conjugate-gradient iteration count of this press
    }
    const double rn = sqrt(sa);
    sc->resid_l2 = (sc->rhs_l2 > 0.0) ? rn / sc->rhs_l2 : rn;
    sc->refine_done = (sc->resid_l2 <= tol) ? 1 : 0;
  }
}

// ---- p64 += e32 ----
__global__ __launch_bounds__(256) void k_mx_accumulate(Dims d, SplitSOR sp, Mixed mx) {
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
  mx.p64[par][is] = mx.p64[par][is] + (double)sp.p[par][is];
}

// ---- pav64 = sum(vol * p64) / sum(vol) (press.f95:123-134 in FP64) ----
__global__ __launch_bounds__(256) void k_mx_pav(Dims d, Grid1D g, SplitSOR sp, Mixed mx) {
  __shared__ double lds[4];
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double v = 0.0;
  if (i <= d.ip && j <= d.jp)
    v = p64_at(sp, mx, i, j, k) * ((double)g.DX1(i) * (double)g.DY1(j) * (double)g.DZN(k));
  const double s = block_sum(v, lds);
  if (tid() == 0) mx.part[flat_block()] = s;
}

__global__ __launch_bounds__(256) void k_mx_pav_finalize(Mixed mx, DevScalars* sc) {
  __shared__ double lds[4];
  double a = 0.0;
  for (int q = threadIdx.x; q < mx.nblk; q += 256) a += mx.part[q];
  const double s = block_sum(a, lds);
  if (threadIdx.x == 0) {
    sc->pav64 = s / sc->area64;
    sc->pav = (float)sc->pav64;
  }
}

// ---- p(0:ip+1,0:jp+1,0:kp+1) = p64(mirror cell) - pav64, rounded once to FP32 ----
__global__ __launch_bounds__(256) void k_mx_gather(Dims d, Fields F, SplitSOR sp, Mixed mx,
                                                   const DevScalars* sc) {
  const int a = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int c = blockIdx.z;
  if (a > d.ip + 1 || b > d.jp + 1) return;
  const int i = (a == 0) ? 1 : ((a == d.ip + 1) ? d.ip : a);
  const int j = (b == 0) ? d.jp : ((b == d.jp + 1) ? 1 : b);
  const int k = (c == 0) ? 1 : ((c == d.kp + 1) ? d.kp : c);
  F.p[d.P(a, b, c)] = (float)(p64_at(sp, mx, i, j, k) - sc->pav64);
}

// ---- conjugate gradients for the correction (refine_solver 1) ----
// A = -L with the boundary images folded into the matrix (a face neighbour across an i or k face
// is the cell itself, so it drops out): (A x)_c = sum over the real neighbours of
// c_nb * (x_c - x_nb), and the preconditioner's diagonal dg_c is the sum of those c_nb.  With
// cn2l(i) = 2/(dxs(i)(dxs(i-1)+dxs(i))) and cn2s(i) = 2/(dxs(i-1)(dxs(i-1)+dxs(i)))
// (init.f95:82-107, and alike along j and k), w*cn2l(i) = w(i+1)*cn2s(i+1) for the weight
// w = (dxs(i-1)+dxs(i)) (dys(j-1)+dys(j)) (dzs(k-1)+dzs(k)): A is symmetric in the inner product
// weighted by w (the periodic j coupling as well when dys(0) = dys(jp), as on a uniform
// horizontal grid).  CG runs in that inner product, preconditioned by symmetric Gauss-Seidel
// (colour 0, colour 1, colour 0, from zero) on the same matrix.  Vectors FP32, inner products
// FP64.
//
// A work item is group grp (entries 4 grp .. 4 grp + 3) of colour row j over a chunk of CG_KC
// levels, marching up in k, with the items dealt over the workgroups in the XCD-aware order of
// the SOR sweeps: every load and store is a float4 of one colour row, no lane idles on the other
// colour, and the other colour's levels k - 1, k, k + 1 stay in registers, so each array is read
// once per chunk (plus two levels).  Cell i of colour c sits at entry (i - 1) / 2 of its row;
// its j and k neighbours at the same entry of the other colour's rows j +- 1 and levels k +- 1,
// its i neighbours at entries e - 1, e (the row's first cell of colour c at i = 1) or e, e + 1
// (at i = 2).
namespace {

// levels per work item: at 600x600x90, 10 levels gave 509 us per CG iteration, 5 levels 510
// and 20 levels 520
#ifndef LES_CG_KC
#define LES_CG_KC 10
#endif
constexpr int CG_KC = LES_CG_KC;

struct CgDir {  // the search direction of the last iteration and the one being formed
  float* old_[2];
  float* new_[2];
};

struct Col {
  int j, grp, row, k0, k1;  // row: the group's offset in level 1
  bool ok;
};

__device__ __forceinline__ Col cg_col(Dims d, const SplitSOR& sp, int nthreads) {
  const int t = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  Col it;
  it.ok = t < nthreads;
  const int tc = min(t, nthreads - 1);  // past the end: the last item's loads, nothing stored
  const int r = (int)__umulhi((unsigned)tc, sp.m_ng);  // chunk * jp + j - 1
  it.grp = tc - r * (sp.nr >> 2);
  const int ch = (int)__umulhi((unsigned)r, sp.m_jp);
  it.j = 1 + r - ch * d.jp;
  it.k0 = 1 + ch * CG_KC;
  it.k1 = min(d.kp, it.k0 + CG_KC - 1);
  it.row = sp.nr * (it.j - 1) + 4 * it.grp;
  return it;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, const float v[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void unpack(float4 a, float v[4]) {
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}
__device__ __forceinline__ void copy4(float* a, const float* b) {
#pragma unroll
  for (int q = 0; q < 4; ++q) a[q] = b[q];
}

// The weights of the item's cells without their k factor: (dxs(i-1)+dxs(i)) (dys(j-1)+dys(j))
// for the four cells of a row whose first cell is i = 1 (f = 1) or i = 2 (f = 2); the padding
// entries' index is clamped (their terms are never added)
struct Wxy {
  double a[2][4];
};

__device__ __forceinline__ Wxy cg_wxy(Dims d, const Grid1D& g, int j, int grp) {
  Wxy w;
  const double wy = (double)g.DYS(j - 1) + (double)g.DYS(j);
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = min(1 + f + 2 * (4 * grp + q), d.ip);
      w.a[f][q] = ((double)g.DXS(i - 1) + (double)g.DXS(i)) * wy;
    }
  return w;
}

__device__ __forceinline__ double cg_wz(const Grid1D& g, int k) {
  return (double)g.DZS(k - 1) + (double)g.DZS(k);
}

// The i coefficients of the item's cells, held for all its levels: cn2l and cn2s of the four
// cells of a row whose first cell is i = 1 (index 0) or i = 2 (index 1), zero where the
// neighbour is absent (i = ip, i = 1, padding), and the preconditioner's diagonal less its k
// terms, c3l + c3s + l + s.  A zero term adds exactly nothing: every value it multiplies is
// finite (the padding entries of z, dold and dnew stay zero).
struct Cx {
  float l[2][4], s[2][4], dxy[2][4];
};

__device__ __forceinline__ Cx cg_cx(Dims d, const Grid1D& g, int grp, float c3l, float c3s) {
  Cx x;
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = 1 + f + 2 * (4 * grp + q);
      const int ic = min(i, d.ip);
      x.l[f][q] = i < d.ip ? g.cn2l[ic - 1] : 0.0f;
      x.s[f][q] = (i > 1 && i <= d.ip) ? g.cn2s[ic - 1] : 0.0f;
      x.dxy[f][q] = ((c3l + c3s) + x.l[f][q]) + x.s[f][q];
    }
  return x;
}

// the row's first cell of colour c at level k is i = 1 (0) or i = 2 (1)
__device__ __forceinline__ int cg_f(int c, int j, int k) { return ((1 + j + k) & 1) == c ? 0 : 1; }

}  // namespace

// workgroups of the CG kernels (a multiple of the 8 XCDs), each writing one partial
int cg_blocks(Dims d, int nr) {
  const int nt = nr / 4 * d.jp * ((d.kp + CG_KC - 1) / CG_KC);
  return ((nt + 255) / 256 + 7) / 8 * 8;
}

// q = A dnew over both colours (dnew = z + beta dold, formed for every value read), stores dnew
// and q, block partials of <dnew, q>.  No thread reads what another writes (dnew and dold are
// different arrays).
template <bool FIRST>
__global__ __launch_bounds__(256) void k_cg_apply(Dims d, Grid1D g, SplitSOR sp, Mixed mx,
                                                  CgDir dd, int nthreads) {
  __shared__ double lds[4];
  if (mx.cg[4] != 0.0) return;  // converged (uniform: every block returns)
  const Col it = cg_col(d, sp, nthreads);
  const int j = it.j, grp = it.grp, nr = sp.nr;
  const int lev = nr * sp.nj;
  const float beta = FIRST ? 0.0f : (float)mx.cg[3];
  const int dn_ = (j == d.jp) ? -nr * (d.jp - 1) : nr;  // row j+1 (periodic)
  const int ds_ = (j == 1) ? nr * (d.jp - 1) : -nr;     // row j-1
  auto dnew4 = [&](int c, int o, float v[4]) {  // dnew of colour c at entries o .. o+3
    unpack(ld4(mx.z[c] + o), v);
    if (!FIRST) {
      float b[4];
      unpack(ld4(dd.old_[c] + o), b);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = v[q] + beta * b[q];
    }
  };
  auto dnew1 = [&](int c, int o) {
    const float a = mx.z[c][o];
    return FIRST ? a : a + beta * dd.old_[c][o];
  };
  const Wxy wxy = cg_wxy(d, g, j, grp);
  const float c3l = g.cn3l[j - 1], c3s = g.cn3s[j - 1];
  const Cx cx = cg_cx(d, g, grp, c3l, c3s);
  // dnew of both colours at levels k-1, k, k+1
  float xm[2][4] = {}, xc[2][4], xp[2][4] = {};  // zero where a level is absent
  const int b0 = it.row + lev * (it.k0 - 1);
  if (it.k0 > 1) {
    dnew4(0, b0 - lev, xm[0]);
    dnew4(1, b0 - lev, xm[1]);
  }
  dnew4(0, b0, xc[0]);
  dnew4(1, b0, xc[1]);
  double acc = 0.0;
  for (int k = it.k0; k <= it.k1; ++k) {
    const int b = it.row + lev * (k - 1);
    if (k < d.kp) {
      dnew4(0, b + lev, xp[0]);
      dnew4(1, b + lev, xp[1]);
    }
    const float c4l = k < d.kp ? g.cn4l[k - 1] : 0.0f, c4s = k > 1 ? g.cn4s[k - 1] : 0.0f;
    const double wz = cg_wz(g, k);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int o = c ^ 1;
      const int f = cg_f(c, j, k);
      float xn[4], xs[4];
      dnew4(o, b + dn_, xn);
      dnew4(o, b + ds_, xs);
      // the other colour's entry before the group (f == 0) or after it (f == 1)
      float edge = 0.0f;
      if (f == 0 && grp > 0) edge = dnew1(o, b - 1);
      if (f == 1 && 4 * grp + 4 < nr) edge = dnew1(o, b + 4);
      const float m[6] = {edge, xc[o][0], xc[o][1], xc[o][2], xc[o][3], edge};
      float qv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float x = xc[c][q];
        const float pw = f == 0 ? m[q] : m[1 + q];      // other colour at i - 1
        const float pe = f == 0 ? m[1 + q] : m[2 + q];  // other colour at i + 1
        const float l = f == 0 ? cx.l[0][q] : cx.l[1][q];
        const float sw = f == 0 ? cx.s[0][q] : cx.s[1][q];
        // flux form, sum c_nb (x_c - x_nb): A annihilates the constants exactly, and a smooth x
        // loses no digits to the cancellation of a large diagonal against its neighbours
        float qq = l * (x - pe);
        qq += sw * (x - pw);
        qq += c3l * (x - xn[q]);
        qq += c3s * (x - xs[q]);
        qq += c4l * (x - xp[o][q]);
        qq += c4s * (x - xm[o][q]);
        const bool real = f + 1 + 2 * (4 * grp + q) <= d.ip;
        if (!real) qq = 0.0f;
        const double w = f == 0 ? wxy.a[0][q] : wxy.a[1][q];
        if (real) acc += w * wz * (double)x * (double)qq;
        qv[q] = qq;
      }
      if (it.ok) {
        st4(dd.new_[c] + b, xc[c]);
        st4(mx.q[c] + b, qv);
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      copy4(xm[c], xc[c]);
      copy4(xc[c], xp[c]);
    }
  }
  if (!it.ok) acc = 0.0;
  const double t = block_sum(acc, lds);
  if (tid() == 0) mx.part[blockIdx.x] = t;
}

// The preconditioner z = M^-1 rr (rr in sp.r, negated: the system is -L e = -r): symmetric SOR
// from z = 0 with relaxation om, z_c <- (1 - om) z_c + om (-rr_c + sum c_nb z_nb) / dg_c, over
// colour 0, colour 1, colour 1, colour 0.  Colour 1's two passes see the same colour-0 values,
// so they are one: z_1 = om (2 - om) gs_1.  Three passes:
//   PASS 0: colour 0 from zero, z_0 = om (-rr_0) / dg (no neighbours); UPDATE first takes the
//           iteration's e += alpha dnew, rr += alpha q on both colours (q = A dnew = -L dnew);
//   PASS 1: colour 1 from the colour-0 neighbours, z_1 = om (2 - om) gs;
//   PASS 2: colour 0, z_0 = (1 - om) z_0 + om gs, and block partials of <rr, z>.
template <int PASS, bool UPDATE>
__global__ __launch_bounds__(256) void k_cg_sgs(Dims d, Grid1D g, SplitSOR sp, Mixed mx, float om,
                                                CgDir dd, int nthreads) {
  constexpr int C = PASS == 1 ? 1 : 0, O = C ^ 1;
  __shared__ double lds[4];
  if (mx.cg[4] != 0.0) return;
  const Col it = cg_col(d, sp, nthreads);
  const int j = it.j, grp = it.grp, nr = sp.nr;
  const int lev = nr * sp.nj;
  const int dn_ = (j == d.jp) ? -nr * (d.jp - 1) : nr;
  const int ds_ = (j == 1) ? nr * (d.jp - 1) : -nr;
  const float c3l = g.cn3l[j - 1], c3s = g.cn3s[j - 1];
  const float om1 = PASS == 1 ? om * (2.0f - om) : om;
  const Cx cx = cg_cx(d, g, grp, c3l, c3s);
  Wxy wxy;
  if constexpr (PASS == 2) wxy = cg_wxy(d, g, j, grp);
  // the other colour's z at levels k-1, k, k+1 (passes 1 and 2), zero where a level is absent
  float zm[4] = {0, 0, 0, 0}, zc[4] = {0, 0, 0, 0}, zp[4] = {0, 0, 0, 0};
  const int b0 = it.row + lev * (it.k0 - 1);
  if constexpr (PASS != 0) {
    if (it.k0 > 1) unpack(ld4(mx.z[O] + b0 - lev), zm);
    unpack(ld4(mx.z[O] + b0), zc);
  }
  double acc = 0.0;
  for (int k = it.k0; k <= it.k1; ++k) {
    const int b = it.row + lev * (k - 1);
    float rr[2][4];
    unpack(ld4(sp.r[C] + b), rr[C]);
    if constexpr (PASS == 2) unpack(ld4(sp.r[O] + b), rr[O]);
    if constexpr (UPDATE) {
      const float a = (float)mx.cg[2];
      unpack(ld4(sp.r[O] + b), rr[O]);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        float e[4], dn4[4], q4[4];
        unpack(ld4(sp.p[cc] + b), e);
        unpack(ld4(dd.new_[cc] + b), dn4);
        unpack(ld4(mx.q[cc] + b), q4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          e[q] = e[q] + a * dn4[q];
          rr[cc][q] = rr[cc][q] + a * q4[q];
        }
        if (it.ok) {
          st4(sp.p[cc] + b, e);
          st4(sp.r[cc] + b, rr[cc]);
        }
      }
    }
    const float c4l = k < d.kp ? g.cn4l[k - 1] : 0.0f, c4s = k > 1 ? g.cn4s[k - 1] : 0.0f;
    const int f = cg_f(C, j, k);
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (PASS != 0) {
      if (k < d.kp) unpack(ld4(mx.z[O] + b + lev), zp);
      float zn[4], zs[4];
      unpack(ld4(mx.z[O] + b + dn_), zn);
      unpack(ld4(mx.z[O] + b + ds_), zs);
      float edge = 0.0f;
      if (f == 0 && grp > 0) edge = mx.z[O][b - 1];
      if (f == 1 && 4 * grp + 4 < nr) edge = mx.z[O][b + 4];
      const float m[6] = {edge, zc[0], zc[1], zc[2], zc[3], edge};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float pw = f == 0 ? m[q] : m[1 + q];
        const float pe = f == 0 ? m[1 + q] : m[2 + q];
        const float l = f == 0 ? cx.l[0][q] : cx.l[1][q];
        const float sw = f == 0 ? cx.s[0][q] : cx.s[1][q];
        float t = c3l * zn[q] + c3s * zs[q];
        t += l * pe;
        t += sw * pw;
        t += c4l * zp[q];
        t += c4s * zm[q];
        s[q] = t;
      }
    }
    float zo[4] = {0, 0, 0, 0};
    if constexpr (PASS == 2) unpack(ld4(mx.z[C] + b), zo);
    float zn4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float dg = ((f == 0 ? cx.dxy[0][q] : cx.dxy[1][q]) + c4l) + c4s;
      const float gs = (-rr[C][q] + s[q]) / dg;
      zn4[q] = PASS == 2 ? (1.0f - om) * zo[q] + om * gs : om1 * gs;
      if (f + 1 + 2 * (4 * grp + q) > d.ip) zn4[q] = 0.0f;  // padding entries stay zero
    }
    if (it.ok) st4(mx.z[C] + b, zn4);
    if constexpr (PASS == 2) {
      const double wz = cg_wz(g, k);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const int fc = cg_f(cc, j, k);
        const float* zz = cc == C ? zn4 : zc;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double w = fc == 0 ? wxy.a[0][q] : wxy.a[1][q];
          if (fc + 1 + 2 * (4 * grp + q) <= d.ip) acc += w * wz * (-(double)rr[cc][q] * (double)zz[q]);
        }
      }
    }
    if constexpr (PASS != 0) {
      copy4(zm, zc);
      copy4(zc, zp);
    }
  }
  if constexpr (PASS == 2) {
    if (!it.ok) acc = 0.0;
    const double t = block_sum(acc, lds);
    if (tid() == 0) mx.part[blockIdx.x] = t;
  }
}

// which 0: rz = rz0 = sum (start); 1: alpha = rz / sum(pq); 2: beta = sum / rz, rz = sum, and
// stop (cg[4] = 1) once rz <= stop * rz0.  cg: rz, pq, alpha, beta, stopped, rz0
__global__ __launch_bounds__(256) void k_cg_scalar(Mixed mx, int which, double stop, int nb) {
  __shared__ double lds[4];
  if (mx.cg[4] != 0.0) return;
  double a = 0.0;
  for (int q = threadIdx.x; q < nb; q += 256) a += mx.part[q];
  const double s = block_sum(a, lds);
  if (threadIdx.x == 0) {
    if (which == 0) {
      mx.cg[0] = s;
      mx.cg[5] = s;
    } else if (which == 1) {
      mx.cg[1] = s;
      mx.cg[2] = (s > 0.0) ? mx.cg[0] / s : 0.0;
      if (!(s > 0.0)) mx.cg[4] = 1.0;  // direction in the null space or lost to rounding: stop
    } else {
      mx.cg[3] = (mx.cg[0] != 0.0) ? s / mx.cg[0] : 0.0;
      mx.cg[0] = s;
      mx.cg[6] += 1.0;  // iterations of this press
      if (!(s > stop * mx.cg[5])) mx.cg[4] = 1.0;  // also stops on 0 or NaN
    }
  }
}

__global__ void k_cg_reset(Mixed mx) {
  if (threadIdx.x < 6) mx.cg[threadIdx.x] = 0.0;
}

// A is singular (constants); its range is the w-orthogonal complement of the constants.  The
// FP32 right-hand side is projected onto it before the iteration (its w-mean removed), or CG
// would chase the inconsistent part into the null space.
__global__ __launch_bounds__(256) void k_cg_mean_part(Dims d, Grid1D g, SplitSOR sp, Mixed mx) {
  __shared__ double lds[4];
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double a = 0.0, b = 0.0;
  if (i <= d.ip && j <= d.jp) {
    const double w = vol(g, i, j, k);
    a = w * (double)sp.r[(i + j + k) & 1][split_index(sp, i, j, k)];
    b = w;
  }
  const double sa = block_sum(a, lds);
  if (tid() == 0) mx.part[flat_block()] = sa;
  __syncthreads();
  const double sb = block_sum(b, lds);
  if (tid() == 0) mx.part[mx.nblk + flat_block()] = sb;
}

__global__ __launch_bounds__(256) void k_cg_mean(Mixed mx) {
  __shared__ double lds[4];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < mx.nblk; q += 256) {
    a += mx.part[q];
    b += mx.part[mx.nblk + q];
  }
  const double sa = block_sum(a, lds);
  __syncthreads();
  const double sb = block_sum(b, lds);
  if (threadIdx.x == 0) mx.cg[1] = (sb > 0.0) ? sa / sb : 0.0;
}

__global__ __launch_bounds__(256) void k_cg_project(Dims d, SplitSOR sp, Mixed mx) {
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
  sp.r[par][is] = (float)((double)sp.r[par][is] - mx.cg[1]);
}

void launch_mx_cg(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Mixed mx, int iters, float om,
                  double stop) {
  const dim3 grid = mx_grid(d), blk(MX_BX, MX_BY);
  CgDir dd{{mx.dir[0], mx.dir[1]}, {mx.dir2[0], mx.dir2[1]}};
  // four entries of a colour row over CG_KC levels per thread
  const int nt = sp.nr / 4 * d.jp * ((d.kp + CG_KC - 1) / CG_KC), nwg = cg_blocks(d, sp.nr);
  const dim3 cg(nwg), cb(256);
  // the last pass adds up <rr, z>
  auto precondition = [&](bool update) {
    if (update)
      hipLaunchKernelGGL((k_cg_sgs<0, true>), cg, cb, 0, st, d, g, sp, mx, om, dd, nt);
    else
      hipLaunchKernelGGL((k_cg_sgs<0, false>), cg, cb, 0, st, d, g, sp, mx, om, dd, nt);
    hipLaunchKernelGGL((k_cg_sgs<1, false>), cg, cb, 0, st, d, g, sp, mx, om, dd, nt);
    hipLaunchKernelGGL((k_cg_sgs<2, false>), cg, cb, 0, st, d, g, sp, mx, om, dd, nt);
  };
  hipLaunchKernelGGL(k_cg_reset, dim3(1), dim3(64), 0, st, mx);
  hipLaunchKernelGGL(k_cg_mean_part, grid, blk, 0, st, d, g, sp, mx);
  hipLaunchKernelGGL(k_cg_mean, dim3(1), dim3(256), 0, st, mx);
  hipLaunchKernelGGL(k_cg_project, grid, blk, 0, st, d, sp, mx);
  precondition(false);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 0, stop, nwg);
  for (int it = 0; it < iters; ++it) {
    // dnew = z + beta dold into the other buffer; q = A dnew; alpha
    if (it == 0)
      hipLaunchKernelGGL(k_cg_apply<true>, cg, cb, 0, st, d, g, sp, mx, dd, nt);
    else
      hipLaunchKernelGGL(k_cg_apply<false>, cg, cb, 0, st, d, g, sp, mx, dd, nt);
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 1, stop, nwg);
    precondition(true);  // e += alpha dnew, rr += alpha q, z = M^-1 rr, <rr, z>
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 2, stop, nwg);
    std::swap(dd.old_[0], dd.new_[0]);
    std::swap(dd.old_[1], dd.new_[1]);
  }
}

int mixed_blocks(Dims d) { return mx_blocks(d); }

int mixed_part_len(Dims d, int nr) { return std::max(2 * mx_blocks(d), cg_blocks(d, nr)); }

void launch_mx_rhsav(hipStream_t st, Dims d, RedBuf rb, DevScalars* sc) {
  hipLaunchKernelGGL(k_mx_rhsav, dim3(1), dim3(256), 0, st, rb.partials, reduction_blocks(d), sc);
}

void launch_mx_scatter(hipStream_t st, Dims d, Grid1D g, Fields F, SplitSOR sp, Mixed mx,
                       DevScalars* sc) {
  hipLaunchKernelGGL(k_mx_scatter, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, g, F, sp, mx, sc);
  hipLaunchKernelGGL(k_mx_wmean, dim3(1), dim3(256), 0, st, mx);
  hipLaunchKernelGGL(k_mx_project_rhs, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, F, sp, mx);
}

void launch_mx_promote(hipStream_t st, Dims d, SplitSOR sp, Mixed mx) {
  hipLaunchKernelGGL(k_mx_promote, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, sp, mx);
}

void launch_mx_resid(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Mixed mx, int first,
                     double tol, DevScalars* sc) {
  hipLaunchKernelGGL(k_mx_resid, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, g, sp, mx, first);
  hipLaunchKernelGGL(k_mx_resid_finalize, dim3(1), dim3(256), 0, st, mx, first, tol, sc);
}

void launch_mx_accumulate(hipStream_t st, Dims d, SplitSOR sp, Mixed mx) {
  hipLaunchKernelGGL(k_mx_accumulate, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, sp, mx);
}

void launch_mx_pav_gather(hipStream_t st, Dims d, Grid1D g, Fields F, SplitSOR sp, Mixed mx,
                          DevScalars* sc) {
  hipLaunchKernelGGL(k_mx_pav, mx_grid(d), dim3(MX_BX, MX_BY), 0, st, d, g, sp, mx);
  hipLaunchKernelGGL(k_mx_pav_finalize, dim3(1), dim3(256), 0, st, mx, sc);
  dim3 grid((d.ip + 2 + 63) / 64, (d.jp + 2 + 3) / 4, d.kp + 2);
  hipLaunchKernelGGL(k_mx_gather, grid, dim3(64, 4), 0, st, d, F, sp, mx, sc);
}

}  // namespace les
