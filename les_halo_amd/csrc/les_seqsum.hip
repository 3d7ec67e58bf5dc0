// press.f95's sums in the reference's own order (les_config::sum_order = LES_SUM_REFERENCE).
//
// The reference accumulates rhsav (press.f95:65-75), pav (press.f95:123-134) and the SOR
// residual (press.f95:89-97) one term at a time in FP32, in loop order: k, then j, then i (for
// the residual: colour nrd = 0 over the whole grid, then nrd = 1).  Each partial sum is rounded,
// so the result depends on the order and no parallel reduction reproduces it.  These kernels
// run the additions in exactly that order on one wavefront: the lanes form 64 consecutive
// terms at a time (with the reference's FP32 operation order, e.g.
// ((dx1(i)*dy1(j))*dzn(k))*rhs(i,j,k)), stage them in LDS and fold them into the running sum
// one by one, while the next 64 terms are already being loaded.  A sum over N cells costs
// about N dependent FP32 additions (tens of milliseconds at 300x300x90): a parity mode that
// reproduces the reference's results bit for bit, not the benchmark path, whose exact sums
// (les_common.h) are order-independent and correctly rounded.
#include "les_kernels.h"

namespace les {

namespace {

// Sequential FP32 sum over the terms of src, in the order rows 0..nrows-1, entries
// 0..len(r)-1 of each row.  Called by one wavefront of 64 threads; every lane returns the sum.
template <class Src>
__device__ float seq_sum(const Src& src, float* stage) {
  const int lane = threadIdx.x;
  const int nrows = src.nrows();
  float s = 0.0f;
  int r = 0, e = 0;
  while (r < nrows && src.len(r) == 0) ++r;
  if (r >= nrows) return s;
  int n_next = min(64, src.len(r) - e);
  float t_next = lane < n_next ? src.term(r, e + lane) : 0.0f;
  while (r < nrows) {
    const float t = t_next;
    const int n = n_next;
    e += 64;
    if (e >= src.len(r)) {
      e = 0;
      ++r;
      while (r < nrows && src.len(r) == 0) ++r;
    }
    if (r < nrows) {  // next 64 terms in flight while these are added
      n_next = min(64, src.len(r) - e);
      t_next = lane < n_next ? src.term(r, e + lane) : 0.0f;
    }
    stage[lane] = t;
    __syncthreads();
    float v[64];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const float4 x = reinterpret_cast<const float4*>(stage)[q];
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
    if (n == 64) {  // a whole chunk: 64 dependent additions and nothing else
#pragma unroll
      for (int l = 0; l < 64; ++l) s = s + v[l];
    } else {
#pragma unroll
      for (int l = 0; l < 64; ++l)
        if (l < n) s = s + v[l];
    }
    __syncthreads();
  }
  return s;
}

// press.f95:67-73: terms dx1(i)*dy1(j)*dzn(k)*rhs(i,j,k), rows (j,k) of the Fortran rhs
struct RhsTerms {
  Dims d;
  Grid1D g;
  const float* rhs;
  __device__ int nrows() const { return d.jp * d.kp; }
  __device__ int len(int) const { return d.ip; }
  __device__ float term(int r, int e) const {
    const int k = 1 + r / d.jp, j = 1 + r - (k - 1) * d.jp, i = 1 + e;
    return g.DX1(i) * g.DY1(j) * g.DZN(k) * rhs[d.R(i, j, k)];
  }
};

// press.f95:125-131: terms p(i,j,k)*dx1(i)*dy1(j)*dzn(k), p in the Fortran array or in the
// colour-split copy (cell (i,j,k) in array (i+j+k)&1 at ((i-1)>>1) of its row)
struct PavTerms {
  Dims d;
  Grid1D g;
  const float* p;  // Fortran p, or null for the split copy
  SplitSOR sp;
  __device__ int nrows() const { return d.jp * d.kp; }
  __device__ int len(int) const { return d.ip; }
  __device__ float term(int r, int e) const {
    const int k = 1 + r / d.jp, j = 1 + r - (k - 1) * d.jp, i = 1 + e;
    const float pv = p ? p[d.P(i, j, k)]
                       : sp.p[(i + j + k) & 1][((i - 1) >> 1) + sp.nr * ((j - 1) + sp.nj * (k - 1))];
    return pv * g.DX1(i) * g.DY1(j) * g.DZN(k);
  }
};

// press.f95:91-97: the residual terms reltmp**2 of one iteration as the sweeps stored them,
// colour c of row (j,k) at ((i-1)>>1) + nrt*((j-1) + jp*(k-1)) of terms[c]; the sweep nrd = 0
// (colour 1: i = 1 + mod(k+j,2)) comes first, then nrd = 1 (colour 0)
struct ResidTerms {
  Dims d;
  const float* t1;
  const float* t0;
  int nrt;
  __device__ int nrows() const { return 2 * d.jp * d.kp; }
  __device__ void rc(int r, int& c, int& j, int& k) const {
    const int plane = d.jp * d.kp;
    c = r < plane ? 1 : 0;
    const int rr = r < plane ? r : r - plane;
    k = 1 + rr / d.jp;
    j = 1 + rr - (k - 1) * d.jp;
  }
  __device__ int len(int r) const {
    int c, j, k;
    rc(r, c, j, k);
    const int i0 = ((1 + j + k) & 1) == c ? 1 : 2;
    return i0 <= d.ip ? (d.ip - i0) / 2 + 1 : 0;
  }
  __device__ float term(int r, int e) const {
    int c, j, k;
    rc(r, c, j, k);
    return (c ? t1 : t0)[e + nrt * ((j - 1) + d.jp * (k - 1))];
  }
};

}  // namespace

// rhsav = sum / area; also re-arms the SOR scalars (as k_rhsav_finalize does)
__global__ __launch_bounds__(64) void k_seq_rhsav(Dims d, Grid1D g, const float* rhs,
                                                  DevScalars* sc) {
  __shared__ __align__(16) float stage[64];
  const float s = seq_sum(RhsTerms{d, g, rhs}, stage);
  if (threadIdx.x == 0) {
    sc->rhsav = s / sc->area;
    sc->sor_done = 0;
    sc->sor_iters = 0;
    sc->sor = 0.0f;
  }
}

__global__ __launch_bounds__(64) void k_seq_pav(Dims d, Grid1D g, const float* p, SplitSOR sp,
                                                DevScalars* sc) {
  __shared__ __align__(16) float stage[64];
  const float s = seq_sum(PavTerms{d, g, p, sp}, stage);
  if (threadIdx.x == 0) sc->pav = s / sc->pco;
}

// residual of one iteration and the early exit of press.f95:114
__global__ __launch_bounds__(64) void k_seq_resid(Dims d, SeqTerms tm, Physics ph,
                                                  DevScalars* sc) {
  __shared__ __align__(16) float stage[64];
  if (sc->sor_done) return;
  const float s = seq_sum(ResidTerms{d, tm.t[1], tm.t[0], tm.nrt}, stage);
  if (threadIdx.x == 0) {
    sc->sor = s;
    sc->sor_iters += 1;
    if (s < ph.pjuge) sc->sor_done = 1;
  }
}

void launch_seq_rhsav(hipStream_t st, Dims d, Grid1D g, const float* rhs, DevScalars* sc) {
  hipLaunchKernelGGL(k_seq_rhsav, dim3(1), dim3(64), 0, st, d, g, rhs, sc);
}

void launch_seq_pav(hipStream_t st, Dims d, Grid1D g, const float* p, const SplitSOR* sp,
                    DevScalars* sc) {
  SplitSOR s{};
  if (sp) s = *sp;
  hipLaunchKernelGGL(k_seq_pav, dim3(1), dim3(64), 0, st, d, g, sp ? nullptr : p, s, sc);
}

void launch_seq_resid(hipStream_t st, Dims d, SeqTerms tm, Physics ph, DevScalars* sc) {
  hipLaunchKernelGGL(k_seq_resid, dim3(1), dim3(64), 0, st, d, tm, ph, sc);
}

}  // namespace les
