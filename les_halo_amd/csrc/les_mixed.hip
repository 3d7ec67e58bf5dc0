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
// is the cell itself, so its coefficient joins the diagonal): (A x)_c = sum over the real
// neighbours of c_nb * (x_c - x_nb); the preconditioner's diagonal is dg_c = 1/cn1 - (image
// coefficients), the same sum up to rounding.  With cn2l(i) =
// 2/(dxs(i)(dxs(i-1)+dxs(i))) and cn2s(i) = 2/(dxs(i-1)(dxs(i-1)+dxs(i))) (init.f95:82-107, and
// alike along j and k), w*cn2l(i) = w(i+1)*cn2s(i+1) for the weight w = (dxs(i-1)+dxs(i)) *
// (dys(j-1)+dys(j)) * (dzs(k-1)+dzs(k)): A is symmetric in the inner product weighted by w (the
// periodic j coupling as well when dys(0) = dys(jp), as on a uniform horizontal grid).  CG runs
// in that inner product, preconditioned by symmetric Gauss-Seidel (colour 0, colour 1, colour 0,
// from zero) on the same matrix.  Vectors FP32, inner products FP64.
namespace {

struct CgDir {  // the search direction of the last iteration and the one being formed
  float* old_[2];
  float* new_[2];
};

__device__ __forceinline__ float xs(const float* const* v, const SplitSOR& sp, int i, int j, int k) {
  return v[(i + j + k) & 1][split_index(sp, i, j, k)];
}

// the matrix row of cell (i,j,k) applied to v (its own value excluded)
__device__ __forceinline__ float offdiag_sum(Dims d, const Grid1D& g, const SplitSOR& sp,
                                             const float* const* v, int i, int j, int k,
                                             float& dg, float cn1) {
  float s = 0.0f;
  dg = 1.0f / cn1;
  const float c2l = g.cn2l[i - 1], c2s = g.cn2s[i - 1], c3l = g.cn3l[j - 1], c3s = g.cn3s[j - 1];
  const float c4l = g.cn4l[k - 1], c4s = g.cn4s[k - 1];
  if (i < d.ip) s += c2l * xs(v, sp, i + 1, j, k); else dg -= c2l;
  if (i > 1) s += c2s * xs(v, sp, i - 1, j, k); else dg -= c2s;
  s += c3l * xs(v, sp, i, j == d.jp ? 1 : j + 1, k);
  s += c3s * xs(v, sp, i, j == 1 ? d.jp : j - 1, k);
  if (k < d.kp) s += c4l * xs(v, sp, i, j, k + 1); else dg -= c4l;
  if (k > 1) s += c4s * xs(v, sp, i, j, k - 1); else dg -= c4s;
  return s;
}

}  // namespace

// dnew = z + beta dold (beta = 0 on the first iteration: dnew = z), q = A dnew, block partials of
// <dnew, q>.  The neighbours' dnew values are formed from their z and dold the same way, so no
// thread reads what another writes (dnew and dold are different arrays).
__global__ __launch_bounds__(256) void k_cg_apply(Dims d, Grid1D g, SplitSOR sp, Mixed mx,
                                                  CgDir dd, int first) {
  __shared__ double lds[4];
  if (mx.cg[4] != 0.0) return;  // converged (uniform: every block returns)
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double acc = 0.0;
  if (i <= d.ip && j <= d.jp) {
    const float beta = first ? 0.0f : (float)mx.cg[3];
    auto dn = [&](int a, int b, int c) {
      const int par = (a + b + c) & 1, is = split_index(sp, a, b, c);
      return first ? mx.z[par][is] : mx.z[par][is] + beta * dd.old_[par][is];
    };
    const int par = (i + j + k) & 1, is = split_index(sp, i, j, k);
    const float c2l = g.cn2l[i - 1], c2s = g.cn2s[i - 1], c3l = g.cn3l[j - 1], c3s = g.cn3s[j - 1];
    const float c4l = g.cn4l[k - 1], c4s = g.cn4s[k - 1];
    const float xv = dn(i, j, k);
    // flux form, sum c_nb (x_c - x_nb): A annihilates the constants exactly, and a smooth x
    // loses no digits to the cancellation of a large diagonal against its neighbours
    float qv = 0.0f;
    if (i < d.ip) qv += c2l * (xv - dn(i + 1, j, k));
    if (i > 1) qv += c2s * (xv - dn(i - 1, j, k));
    qv += c3l * (xv - dn(i, j == d.jp ? 1 : j + 1, k));
    qv += c3s * (xv - dn(i, j == 1 ? d.jp : j - 1, k));
    if (k < d.kp) qv += c4l * (xv - dn(i, j, k + 1));
    if (k > 1) qv += c4s * (xv - dn(i, j, k - 1));
    dd.new_[par][is] = xv;
    mx.q[par][is] = qv;
    acc = vol(g, i, j, k) * (double)xv * (double)qv;
  }
  const double t = block_sum(acc, lds);
  if (tid() == 0) mx.part[flat_block()] = t;
}

__device__ __forceinline__ void cg_relax_body(Dims d, const Grid1D& g, SplitSOR sp, Mixed mx,
                                              int i, int j, int k, int is, int colour, int first,
                                              int fresh, float om) {
  float dg;
  const float s = first ? 0.0f : offdiag_sum(d, g, sp, mx.z, i, j, k, dg, mx.cn1[d.O(i, j, k)]);
  if (first) {
    dg = 1.0f / mx.cn1[d.O(i, j, k)];
    if (i == d.ip) dg -= g.cn2l[i - 1];
    if (i == 1) dg -= g.cn2s[i - 1];
    if (k == d.kp) dg -= g.cn4l[k - 1];
    if (k == 1) dg -= g.cn4s[k - 1];
  }
  const float gs = (-sp.r[colour][is] + s) / dg;
  mx.z[colour][is] = (first || fresh) ? om * gs : (1.0f - om) * mx.z[colour][is] + om * gs;
}

// one colour of the symmetric SOR preconditioner z = M^-1 rr (rr in sp.r, negated: the system
// is -L e = -r), relaxation om: z_c <- (1 - om) z_c + om (rr_c + sum c_nb z_nb) / dg_c.  The
// application starts from z = 0: first (colour 0 forward) ignores the neighbours and its own
// value, fresh (colour 1 forward) its own value.
// update (first pass of an iteration): every cell first takes e += alpha dnew and rr -= alpha q.
// rz (last pass): block partials of <rr, z> over both colours (colour 1 is final by then).
__global__ __launch_bounds__(256) void k_cg_sgs(Dims d, Grid1D g, SplitSOR sp, Mixed mx, int colour,
                                                int first, int fresh, float om, CgDir dd,
                                                int update, int rz) {
  __shared__ double lds[4];
  const int i = 1 + blockIdx.x * MX_BX + threadIdx.x;
  const int j = 1 + blockIdx.y * MX_BY + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (mx.cg[4] != 0.0) return;  // converged (uniform: every block returns)
  const bool in = i <= d.ip && j <= d.jp;
  const int par = (i + j + k) & 1;
  const int is = in ? split_index(sp, i, j, k) : 0;
  if (in && update) {
    const float a = (float)mx.cg[2];
    sp.p[par][is] = sp.p[par][is] + a * dd.new_[par][is];
    sp.r[par][is] = sp.r[par][is] + a * mx.q[par][is];
  }
  if (in && par == colour) cg_relax_body(d, g, sp, mx, i, j, k, is, colour, first, fresh, om);
  if (rz) {
    double acc = 0.0;
    if (in) acc = vol(g, i, j, k) * -(double)sp.r[par][is] * (double)mx.z[par][is];
    const double t = block_sum(acc, lds);
    if (tid() == 0) mx.part[flat_block()] = t;
  }
}

// which 0: rz = rz0 = sum (start); 1: alpha = rz / sum(pq); 2: beta = sum / rz, rz = sum, and
// stop (cg[4] = 1) once rz <= stop * rz0.  cg: rz, pq, alpha, beta, stopped, rz0
__global__ __launch_bounds__(256) void k_cg_scalar(Mixed mx, int which, double stop) {
  __shared__ double lds[4];
  if (mx.cg[4] != 0.0) return;
  double a = 0.0;
  for (int q = threadIdx.x; q < mx.nblk; q += 256) a += mx.part[q];
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
  // symmetric SOR from zero: colour 0, colour 1 forward, colour 1, colour 0 backward; the first
  // pass may take the iteration's e and rr update, the last adds up <rr, z>
  auto precondition = [&](int update) {
    hipLaunchKernelGGL(k_cg_sgs, grid, blk, 0, st, d, g, sp, mx, 0, 1, 1, om, dd, update, 0);
    hipLaunchKernelGGL(k_cg_sgs, grid, blk, 0, st, d, g, sp, mx, 1, 0, 1, om, dd, 0, 0);
    hipLaunchKernelGGL(k_cg_sgs, grid, blk, 0, st, d, g, sp, mx, 1, 0, 0, om, dd, 0, 0);
    hipLaunchKernelGGL(k_cg_sgs, grid, blk, 0, st, d, g, sp, mx, 0, 0, 0, om, dd, 0, 1);
  };
  hipLaunchKernelGGL(k_cg_reset, dim3(1), dim3(64), 0, st, mx);
  hipLaunchKernelGGL(k_cg_mean_part, grid, blk, 0, st, d, g, sp, mx);
  hipLaunchKernelGGL(k_cg_mean, dim3(1), dim3(256), 0, st, mx);
  hipLaunchKernelGGL(k_cg_project, grid, blk, 0, st, d, sp, mx);
  precondition(0);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 0, stop);
  for (int it = 0; it < iters; ++it) {
    // dnew = z + beta dold into the other buffer; q = A dnew; alpha
    hipLaunchKernelGGL(k_cg_apply, grid, blk, 0, st, d, g, sp, mx, dd, it == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 1, stop);
    precondition(1);  // e += alpha dnew, rr -= alpha q, z = M^-1 rr, <rr, z>
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, st, mx, 2, stop);
    std::swap(dd.old_[0], dd.new_[0]);
    std::swap(dd.old_[1], dd.new_[1]);
  }
}

int mixed_blocks(Dims d) { return mx_blocks(d); }

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
