// Device state shared by the kernels and the host API.
#pragma once
#include "les_common.h"

namespace les {

// Device pointers of every field, in the reference's Fortran layouts.
struct Fields {
  float *p, *u, *v, *w, *usum, *vsum, *wsum, *f, *g, *h, *fold, *gold, *hold, *rhs, *sm;
  float *diu1, *diu2, *diu3, *diu4, *diu5, *diu6, *diu7, *diu8, *diu9;
  const uint8_t* mask;  // interior (1:ip,1:jp,1:kp): bit0 bmask1, bit1 cmask1, bit2 dmask1
  const float* cn1;     // cn1(1:ip,1:jp,1:kp)
  const float* wind;    // wind_profile(1:jp,1:kp)
};

// Scalars kept on the device between stages (graph-capturable: no host round trip).
struct DevScalars {
  float uout, rhsav, area, pav, pco, sor;
  int sor_iters, sor_done;
  double resid_l2;
  int refine_passes;
  int pad;
};

struct Physics {
  float dt, ro, vn, alpha, beta, cs0, omega, pjuge;
};

// Reduction scratch: per-block carry-save limbs.
struct RedBuf {
  int64_t* partials;  // [max_blocks][NLIMB]
  int max_blocks;
};

// ---- launchers (les_kernels.hip); all asynchronous on 'st' ----
void launch_velnw_inflow(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velnw,
                         int inflow_n);
void launch_uout(hipStream_t st, Dims d, Fields F, DevScalars* sc);
void launch_bondv1_uvw(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, DevScalars* sc);
void launch_velfg(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velfg,
                  int do_feedbf);
void launch_calc_sm(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph);
void launch_boundsm(hipStream_t st, Dims d, Fields F);
void launch_calc_visc_adam(hipStream_t st, Dims d, Grid1D g, Fields F, int do_visc,
                           int do_adam);
void launch_rhs(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, RedBuf rb,
                DevScalars* sc);
void launch_bondfg(hipStream_t st, Dims d, Fields F);
void launch_rhs_adjust(hipStream_t st, Dims d, Fields F, DevScalars* sc);
void launch_sor_colour(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int nrd,
                       int want_resid, RedBuf rb, int blk_offset, DevScalars* sc);
void launch_sor_resid_finalize(hipStream_t st, RedBuf rb, int nblocks, Physics ph,
                               DevScalars* sc);
void launch_boundp1(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done);
void launch_boundp2(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done);
void launch_sor_reset(hipStream_t st, DevScalars* sc);
void launch_pav(hipStream_t st, Dims d, Grid1D g, Fields F, RedBuf rb, DevScalars* sc);
void launch_adj(hipStream_t st, Dims d, Fields F, DevScalars* sc, int with_boundp);
void launch_boundp_mirror(hipStream_t st, Dims d, Fields F);
int sor_resid_blocks(Dims d);
int reduction_blocks(Dims d);

// exact sum of n device floats into *out (device), for tests of the reduction itself
void launch_exact_sum(hipStream_t st, const float* x, long n, RedBuf rb, float* out);

// halo exchange pack/unpack (halo_exchange.cl semantics, SoA over the fields)
void launch_halo_read(hipStream_t st, const float* const* arrays, int v_dim, int im, int jm,
                      int km, int h_w, int correct, float* buffer);
void launch_halo_write(hipStream_t st, float* const* arrays, int v_dim, int im, int jm, int km,
                       int h_w, const float* buffer);

}  // namespace les
