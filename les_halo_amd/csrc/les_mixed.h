// Device state and launchers of the mixed-precision pressure solve (les_mixed.hip).
#pragma once
#include "les_kernels.h"

namespace les {

// FP64 companions of the colour-split SOR copy (same indexing as SplitSOR::p / SplitSOR::r)
struct Mixed {
  double* p64[2];     // pressure iterate
  double* r64[2];     // rhs - rhsav64
  double* part;       // per-block partial sums, mixed_part_len entries
  const float* cn1;   // cn1(1:ip,1:jp,1:kp)
  int nblk;
  // conjugate-gradient correction solve (refine_solver 1): FP32 split vectors and FP64 scalars
  float* z[2];        // preconditioned residual
  float* dir[2];      // search direction (two buffers: the next is formed from the last)
  float* dir2[2];
  float* q[2];        // A * dir
  double* cg;         // rz, pq, alpha, beta, stopped flag, rz at the start, iterations, w-mean
                      // of the rhs
};

int mixed_blocks(Dims d);
// entries of Mixed::part: two per block of the cell kernels, or one per conjugate-gradient block
int mixed_part_len(Dims d, int nr);
void launch_mx_rhsav(hipStream_t st, Dims d, RedBuf rb, DevScalars* sc);
// p -> split copy; rhs64 = rhs - rhsav64 minus its w-mean (the solvable right-hand side)
void launch_mx_scatter(hipStream_t st, Dims d, Grid1D g, Fields F, SplitSOR sp, Mixed mx,
                       DevScalars* sc);
void launch_mx_promote(hipStream_t st, Dims d, SplitSOR sp, Mixed mx);
void launch_mx_resid(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Mixed mx, int first,
                     double tol, DevScalars* sc);
void launch_mx_accumulate(hipStream_t st, Dims d, SplitSOR sp, Mixed mx);
// the correction solve as conjugate gradients on -L e = -r (sp.r: r, sp.p: e from zero),
// preconditioned by symmetric red-black SOR with relaxation om; at most iters iterations,
// stopping on the device once the preconditioned residual norm squared falls to stop times its
// start value
void launch_mx_cg(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Mixed mx, int iters, float om,
                  double stop);
void launch_mx_pav_gather(hipStream_t st, Dims d, Grid1D g, Fields F, SplitSOR sp, Mixed mx,
                          DevScalars* sc);

}  // namespace les
