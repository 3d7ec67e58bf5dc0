// Red-black SOR of press.f95:87-115 on a colour-split copy of p and rhs.
//
// The reference sweeps the Fortran array in place with stride 2 along i.  Here, for the
// duration of the solve, cell (i,j,k) of the box [0,ip+1]x[0,jp+1]x[0,kp+1] is stored in array
// (i+j+k)&1 at position (i>>1, j, k).  In every (j,k) row the cells of one colour are then
// contiguous, and their six neighbours sit in the other array at (i>>1 or i>>1 +- 1, j, k),
// (i>>1, j+-1, k), (i>>1, j, k+-1): a sweep reads one colour's p, rhs and the other colour's
// p and writes one colour's p - 8 bytes per cell of the grid, all in whole cache lines.
//
// cn1 is recomputed from its 1-D factors (init.f95:98-104: 2/(dxs(i-1)dxs(i)) +
// 2/(dys(j-1)dys(j)) + 2/(dzs(k-1)dzs(k)), then the reciprocal) with the reference's order of
// operations, so it is bit-identical to the stored array and costs no memory traffic.
//
// boundp1 (after each colour) and boundp2 (after each iteration) are fused: the thread that
// updates a cell also writes every mirror image of it (i faces, periodic j faces, k faces).
// Each mirror of a cell is read only by that cell, except the periodic j images, which are
// read by the cell across the seam; for even jp that cell has the other colour and is read in
// the other sweep.  So, for even jp and mirror-consistent halos on entry (the state every
// press leaves), every read returns the value the reference's sequential sweep reads, and the
// state after each full iteration is the reference's state after boundp2.
#include "les_kernels.h"

namespace les {

namespace {
constexpr int SX = 32;  // colour cells along i per workgroup row (a wavefront = 2 rows)
constexpr int SY = 8;   // rows of j per workgroup

__device__ __forceinline__ void put_split(const SplitSOR& sp, int a, int b, int c, float v) {
  sp.p[(a + b + c) & 1][sp.S(a >> 1, b, c)] = v;
}
}  // namespace

// ---- p (Fortran layout) -> split copy; rhs -= rhsav (press.f95:76-82) into both layouts ----
__global__ __launch_bounds__(256) void k_split_scatter(Dims d, Fields F, SplitSOR sp,
                                                       const DevScalars* sc, int adjust_rhs) {
  const int i = blockIdx.x * 64 + threadIdx.x;  // 0..ip+1
  const int j = blockIdx.y * 4 + threadIdx.y;   // 0..jp+1
  const int k = blockIdx.z;                     // 0..kp+1
  if (i > d.ip + 1 || j > d.jp + 1) return;
  const int par = (i + j + k) & 1;
  const long is = sp.S(i >> 1, j, k);
  sp.p[par][is] = F.p[d.P(i, j, k)];
  if (i >= 1 && i <= d.ip && j >= 1 && j <= d.jp && k >= 1 && k <= d.kp) {
    const long ir = d.R(i, j, k);
    float r = F.rhs[ir];
    if (adjust_rhs) {
      r = r - sc->rhsav;
      F.rhs[ir] = r;
    }
    sp.r[par][is] = r;
  }
}

void launch_split_scatter(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                          int adjust_rhs) {
  dim3 grid((d.ip + 2 + 63) / 64, (d.jp + 2 + 3) / 4, d.kp + 2);
  hipLaunchKernelGGL(k_split_scatter, grid, dim3(64, 4), 0, st, d, F, sp, sc, adjust_rhs);
}

// ---- one colour sweep ----
__global__ __launch_bounds__(256) void k_sor_split(Dims d, Grid1D g, SplitSOR sp, Physics ph,
                                                   int nrd, int want_resid, int64_t* partials,
                                                   const DevScalars* sc, int nbx, int nby,
                                                   int ntiles) {
  __shared__ unsigned long long lds[NLIMB];
  if (sc->sor_done) return;
  if (want_resid) block_limbs_init(lds);
  ThreadAcc acc;
  acc.init();
  const int L = xcd_tile(blockIdx.x, gridDim.x);
  if (L < ntiles) {
    const int bx = L % nbx;
    const int by = (L / nbx) % nby;
    const int k = 1 + L / (nbx * nby);
    const int j = 1 + by * SY + threadIdx.y;
    const int i = 1 + ((k + j + nrd) & 1) + 2 * (bx * SX + threadIdx.x);
    if (i <= d.ip && j <= d.jp) {
      const int c = (i + j + k) & 1, o = c ^ 1;
      const float* po = sp.p[o];
      const int ii = i >> 1;
      const long s0 = sp.S(ii, j, k);
      const long sj = sp.nh, sk = (long)sp.nh * sp.nj;
      const float pc = sp.p[c][s0];
      const float pe = po[sp.S((i + 1) >> 1, j, k)];
      const float pw = po[sp.S((i - 1) >> 1, j, k)];
      const float cn1 = 1.0f / (sp.ax[i] + sp.by[j] + sp.cz[k]);
      const float reltmp =
          ph.omega * (cn1 * (g.cn2l[i - 1] * pe + g.cn2s[i - 1] * pw + g.cn3l[j - 1] * po[s0 + sj] +
                             g.cn3s[j - 1] * po[s0 - sj] + g.cn4l[k - 1] * po[s0 + sk] +
                             g.cn4s[k - 1] * po[s0 - sk] - sp.r[c][s0]) -
                      pc);
      const float pn = pc + reltmp;
      sp.p[c][s0] = pn;
      if (want_resid) acc.add(reltmp * reltmp, &lds[0]);
      // mirror images (boundp1 i/j faces, boundp2 k faces)
      const bool mi = (i == 1) || (i == d.ip), mj = (j == 1) || (j == d.jp),
                 mk = (k == 1) || (k == d.kp);
      if (mi || mj || mk) {
        const int ia = (i == 1) ? 0 : d.ip + 1;
        const int jb = (j == d.jp) ? 0 : d.jp + 1;
        const int kc = (k == 1) ? 0 : d.kp + 1;
        if (mi) put_split(sp, ia, j, k, pn);
        if (mj) put_split(sp, i, jb, k, pn);
        if (mk) put_split(sp, i, j, kc, pn);
        if (mi && mj) put_split(sp, ia, jb, k, pn);
        if (mi && mk) put_split(sp, ia, j, kc, pn);
        if (mj && mk) put_split(sp, i, jb, kc, pn);
        if (mi && mj && mk) put_split(sp, ia, jb, kc, pn);
      }
    }
  }
  if (want_resid) {
    acc.flush(lds);
    block_limbs_store(lds, partials + (long)blockIdx.x * NLIMB);
  }
}

namespace {
void split_tiles(Dims d, int& nbx, int& nby, int& ntiles, int& nwg) {
  const int half = (d.ip + 1) / 2;  // colour cells per row (upper bound)
  nbx = (half + SX - 1) / SX;
  nby = (d.jp + SY - 1) / SY;
  ntiles = nbx * nby * d.kp;
  nwg = (ntiles + 7) / 8 * 8;
}
}  // namespace

int sor_split_blocks(Dims d) {
  int nbx, nby, nt, nwg;
  split_tiles(d, nbx, nby, nt, nwg);
  return nwg;
}

void launch_sor_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, int nrd,
                      int want_resid, RedBuf rb, int blk_offset, DevScalars* sc) {
  int nbx, nby, nt, nwg;
  split_tiles(d, nbx, nby, nt, nwg);
  hipLaunchKernelGGL(k_sor_split, dim3(nwg), dim3(SX, SY), 0, st, d, g, sp, ph, nrd, want_resid,
                     rb.partials + (long)blk_offset * NLIMB, sc, nbx, nby, nt);
}

// ---- pav partial sums (press.f95:125-132) from the split copy ----
__global__ __launch_bounds__(256) void k_pav_split(Dims d, Grid1D g, SplitSOR sp,
                                                   int64_t* partials) {
  __shared__ unsigned long long lds[NLIMB];
  block_limbs_init(lds);
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int kchunk = (d.kp + RED_KC - 1) / RED_KC;
  const int k0 = 1 + blockIdx.z * kchunk, k1 = min(d.kp, k0 + kchunk - 1);
  ThreadAcc acc;
  acc.init();
  if (i <= d.ip && j <= d.jp) {
    const float dx1i = g.DX1(i), dy1j = g.DY1(j);
    for (int k = k0; k <= k1; ++k)
      acc.add(sp.p[(i + j + k) & 1][sp.S(i >> 1, j, k)] * dx1i * dy1j * g.DZN(k), lds);
  }
  acc.flush(lds);
  const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_limbs_store(lds, partials + (long)blk * NLIMB);
}

__global__ __launch_bounds__(256) void k_pav_split_finalize(const int64_t* partials, int nb,
                                                            DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) sc->pav = limbs_to_float(limbs) / sc->pco;
}

void launch_pav_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, RedBuf rb, DevScalars* sc) {
  dim3 grid((d.ip + 63) / 64, (d.jp + 3) / 4, RED_KC);
  int nb = (int)(grid.x * grid.y * grid.z);
  hipLaunchKernelGGL(k_pav_split, grid, dim3(64, 4), 0, st, d, g, sp, rb.partials);
  hipLaunchKernelGGL(k_pav_split_finalize, dim3(1), dim3(256), 0, st, rb.partials, nb, sc);
}

// ---- split copy -> p, minus pav (press.f95:135-146): every box cell, the halos being the
// mirror images of the adjusted interior exactly as boundp1 + boundp2 leave them ----
__global__ __launch_bounds__(256) void k_split_gather_adj(Dims d, Fields F, SplitSOR sp,
                                                          const DevScalars* sc, int subtract_pav) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i > d.ip + 1 || j > d.jp + 1) return;
  const float v = sp.p[(i + j + k) & 1][sp.S(i >> 1, j, k)];
  F.p[d.P(i, j, k)] = subtract_pav ? v - sc->pav : v;
}

void launch_split_gather_adj(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                             int subtract_pav) {
  dim3 grid((d.ip + 2 + 63) / 64, (d.jp + 2 + 3) / 4, d.kp + 2);
  hipLaunchKernelGGL(k_split_gather_adj, grid, dim3(64, 4), 0, st, d, F, sp, sc, subtract_pav);
}

}  // namespace les
