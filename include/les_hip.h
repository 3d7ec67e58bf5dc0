/*
 * les_hip.h - C ABI of the MI355X (gfx950) LES_Halo timestep library, libles_hip.so.
 *
 * It replaces the OpenCLIntegration layer under the reference's GPU host module
 * (src/OpenCL/module_LES_combined_ocl_TEMPL_V2.f95, generated into
 * src/OpenCL/Wrappers/module_LES_combined_ocl.f95) and exposes one entry point per
 * kernel-level stage of the reference time step, with the reference's names.
 *
 * Conventions
 *  - Every array keeps the reference's Fortran bounds and column-major layout (i fastest),
 *    exactly as declared in src/Fortran/main_ocl.f95:40-167; les_field_bounds() gives them.
 *    Uploads and downloads are plain copies of those arrays (FP32).
 *  - Every call returns LES_OK (0) or an LES_E_* code; les_last_error() gives the message.
 *    No call exits the process.
 *  - One HIP stream per handle; calls are asynchronous on it except les_download,
 *    les_get_scalars, les_sync and the les_exchange_halo_* host-buffer forms, which
 *    synchronise.  A handle must not be used from two threads at once.
 *  - The GPU code path is the only path: without a usable HIP device les_create fails with
 *    LES_E_NODEV.
 */
#ifndef LES_HIP_H
#define LES_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
enum {
  LES_OK = 0,
  LES_E_INVALID = 1, /* bad argument (size, field id, value) */
  LES_E_HIP = 2,     /* a HIP runtime call failed */
  LES_E_NOMEM = 3,   /* device allocation failed */
  LES_E_STATE = 4,   /* call not valid in the handle's current state */
  LES_E_NODEV = 5    /* no HIP device */
};

/* ---- fields: the arrays of main_ocl.f95:40-167 --------------------------------------- */
typedef enum les_field {
  LES_P = 0,   /* p(0:ip+2,0:jp+2,0:kp+1) */
  LES_U,       /* u(0:ip+1,-1:jp+1,0:kp+1) */
  LES_V,       /* v(0:ip+1,-1:jp+1,0:kp+1) */
  LES_W,       /* w(0:ip+1,-1:jp+1,-1:kp+1) */
  LES_USUM,    /* usum(0:ip,0:jp,0:kp) */
  LES_VSUM,
  LES_WSUM,
  LES_F,       /* f(0:ip,0:jp,0:kp) */
  LES_G,
  LES_H,
  LES_FOLD,    /* fold(ip,jp,kp) */
  LES_GOLD,
  LES_HOLD,
  LES_RHS,     /* rhs(0:ip+1,0:jp+1,0:kp+1) */
  LES_SM,      /* sm(-1:ip+1,-1:jp+1,0:kp+1) */
  LES_DIU1,    /* diu1(-1:ip+2,0:jp+2,0:kp+2) */
  LES_DIU2,    /* diu2..4,6..9(0:ip+2,0:jp+2,0:kp+2) */
  LES_DIU3,
  LES_DIU4,
  LES_DIU5,    /* diu5(-1:ip+2,0:jp+2,0:kp+2) */
  LES_DIU6,
  LES_DIU7,
  LES_DIU8,
  LES_DIU9,
  LES_AMASK1,  /* amask1(0:ip+1,0:jp+1,0:kp+1), values 0/1 */
  LES_BMASK1,  /* bmask1(-1:ip+1,0:jp+1,0:kp+1), values 0/1 */
  LES_CMASK1,  /* cmask1(0:ip+1,-1:jp+1,0:kp+1), values 0/1 */
  LES_DMASK1,  /* dmask1(0:ip+1,0:jp+1,0:kp+1), values 0/1 */
  LES_CN1,     /* cn1(ip,jp,kp) */
  LES_CN2L,    /* cn2l(ip) */
  LES_CN2S,
  LES_CN3L,    /* cn3l(jp) */
  LES_CN3S,
  LES_CN4L,    /* cn4l(kp) */
  LES_CN4S,
  LES_DX1,     /* dx1(-1:ip+1) */
  LES_DXS,     /* dxs(0:ip) */
  LES_DY1,     /* dy1(0:jp+1) */
  LES_DYS,     /* dys(0:jp) */
  LES_DZN,     /* dzn(-1:kp+2) */
  LES_DZS,     /* dzs(-1:kp+2) */
  LES_Z2,      /* z2(kp+2) */
  LES_DELX1,   /* delx1(kp) */
  LES_WIND_PROFILE, /* wind_profile(jp,kp) (main_ocl.f95:162) */
  LES_NFIELDS
} les_field;

/* Fortran bounds lo[3]..hi[3] of a field for interior sizes ip, jp, kp (1-D fields use only
 * lo[0]/hi[0], 2-D wind_profile lo[0..1]/hi[0..1]); returns the element count, or -1. */
int64_t les_field_bounds(int field, int ip, int jp, int kp, int lo[3], int hi[3]);

/* ---- configuration ------------------------------------------------------------------- */
enum { LES_PREC_FP32 = 0, LES_PREC_MIXED = 1 };
/* Order of press.f95's sums (rhsav, pav, the SOR residual).  LES_SUM_EXACT: exact and
 * order-independent, rounded once to FP32 (the default; at least as accurate as any order).
 * LES_SUM_REFERENCE: the reference's own sequential FP32 accumulation in its loop order
 * (press.f95:65-75,89-97,123-134), which makes the results bit-identical to the reference's
 * CPU code; a parity mode, far slower (each sum is one chain of N dependent additions). */
enum { LES_SUM_EXACT = 0, LES_SUM_REFERENCE = 1 };

typedef struct les_config {
  int ip, jp, kp;   /* interior cells (params_common_sn.f95:3) */
  float dt;         /* time step (set.f95:76) */
  float ro;         /* density (set.f95:78) */
  float vn;         /* kinematic viscosity (set.f95:79) */
  float alpha;      /* feedback forcing gains (set.f95:81-82) */
  float beta;
  float cs0;        /* Smagorinsky constant (les.f95:31) */
  float omega;      /* SOR relaxation (press.f95:46) */
  float pjuge;      /* SOR early-exit threshold on the residual; 0 disables (press.f95:44,114) */
  int nmaxp;        /* SOR iterations per step (press.f95:45) */
  int precision;    /* LES_PREC_FP32: reference arithmetic; LES_PREC_MIXED: FP64 reductions and
                       outer refinement of the pressure solve */
  double refine_tol;  /* LES_PREC_MIXED: target relative L2 norm of the pressure residual,
                         |rhs - L p|_2 / |rhs|_2 (FP64) */
  int refine_max;     /* LES_PREC_MIXED: maximum outer refinement passes */
  int device;         /* HIP device ordinal */
  int sub_i, sub_j;   /* subdomains along i and j on the one device (1,1 = single domain) */
  float refine_omega; /* LES_PREC_MIXED: relaxation of the correction solves' SOR sweeps; 0 =
                         omega (SOR corrections) or 1 (the CG preconditioner) */
  int refine_solver;  /* LES_PREC_MIXED: 0 = each correction by nmaxp SOR iterations from zero;
                         1 = by nmaxp conjugate-gradient iterations preconditioned by symmetric
                         red-black Gauss-Seidel (the FP32 colour sweeps), FP64 inner products */
  int sum_order;      /* LES_SUM_EXACT or LES_SUM_REFERENCE (FP32 single-domain handles only) */
} les_config;

/* Fills the reference's defaults (set.f95, les.f95, press.f95) for an ip x jp x kp grid. */
int les_default_config(les_config* cfg, int ip, int jp, int kp);

typedef struct les_handle les_handle;

/* Replaces oclInit + the oclMake*Buffer calls of initialise_LES_kernel
 * (Wrappers/module_LES_combined_ocl.f95:224-306): allocates every device array, zeroed. */
int les_create(const les_config* cfg, les_handle** out);
int les_destroy(les_handle* h);
const char* les_last_error(const les_handle* h);
/* Interior sizes of the handle's grid. */
int les_get_dims(const les_handle* h, int* ip, int* jp, int* kp);
/* Waits for all work queued on the handle's stream. */
int les_sync(les_handle* h);

/* Replace the oclWrite* and oclRead* calls of one array (Wrappers/module_LES_combined_ocl.f95:359-418,
 * TEMPL:1035-1056).  host points at the full Fortran-bounded array.  Mask fields must hold
 * only 0 and 1.  Uploading a grid field (dx1..dzs) does not recompute coefficients: call
 * les_init_coefficients afterwards, or upload cn1..cn4s and delx1 as well. */
int les_upload(les_handle* h, int field, const float* host);
int les_download(les_handle* h, int field, float* host);

/* Device address and element count of a field (for callers that keep data on the device).
 * The grid (dx1..dzs) is read through tables that les_upload derives from it (power-of-two
 * flags, the FP64 reciprocals of the vertical spacings): change grid fields with les_upload,
 * not through this address. */
int les_device_ptr(les_handle* h, int field, void** dptr, int64_t* count);

/* init.f95:82-107 (cn1, cn2l/s, cn3l/s, cn4l/s) and les.f95:33-37 (delx1) from the uploaded
 * grid arrays, evaluated on the host in FP32 exactly as the reference does, then uploaded. */
int les_init_coefficients(les_handle* h);

/* ---- the kernel-level stages (reference names) ---------------------------------------- */
/* velnw.f95:27-59 */
int les_velnw(les_handle* h);
/* bondv1_ext_wind_profile.f95:27-50: inflow planes i=0,1 from wind_profile, n==1 fill */
int les_bondv1_init_uvw(les_handle* h, int n);
/* state 1 of run_LES_kernel (TEMPL:459-497): velnw followed by the inflow/n==1 fill */
int les_velnw__bondv1_init_uvw(les_handle* h, int n);
/* bondv1.f95:67-75: uout = (max(0,max u(im,:,:)) + min(0,min u(im,:,:)))/2, kept on device */
int les_bondv1_calc_uout(les_handle* h);
/* bondv1.f95:81-137: outflow, periodic j, top/bottom of u, v, w */
int les_bondv1_calc_uvw(les_handle* h);
/* velFG.f95:56-105 with vel2.f95:48-188 (diu1..9 and their halo copies, f, g, h) */
int les_velfg(les_handle* h);
/* feedbf.f95:38-67 */
int les_feedbf(les_handle* h);
/* les.f95:33-60: Smagorinsky sm */
int les_les_calc_sm(les_handle* h);
/* state 4 (TEMPL:585-627): velfg + feedbf + les_calc_sm */
int les_velfg__feedbf__les_calc_sm(les_handle* h);
/* boundsm.f95:13-32 (state 5) */
int les_les_bound_sm(les_handle* h);
/* les.f95:78-182 */
int les_les_calc_visc(les_handle* h);
/* adam.f95:32-46 */
int les_adam(les_handle* h);
/* state 6 (TEMPL:668-706): les_calc_visc + adam */
int les_les_calc_visc__adam(les_handle* h);
/* les.f95 as a whole: calc_sm, boundsm, calc_visc */
int les_les(les_handle* h);
/* state 7: bondFG.f95:15-32, press.f95:51-82 (rhs, rhsav = sum(vol*rhs)/sum(vol), rhs -= rhsav);
 * the volume sum is exact and order-independent, rounded once to FP32 */
int les_press_rhsav(les_handle* h);
/* state 8 (TEMPL:768-868): nrd = 0 or 1 -> one red-black colour sweep of press.f95:90-100
 * followed by boundp1; nrd = 2 -> boundp2 */
int les_press_sor(les_handle* h, int nrd);
/* press.f95:87-115: the whole SOR loop, nmaxp iterations (early exit when pjuge > 0) */
int les_press_sor_solve(les_handle* h);
/* state 9: pav = sum(vol*p)/sum(vol) (press.f95:123-134), kept on device */
int les_press_pav(les_handle* h);
/* state 10: p -= pav (press.f95:135-141) */
int les_press_adj(les_handle* h);
/* state 11: boundp1 then boundp2 (press.f95:144-146, boundp.f95) */
int les_press_boundp(les_handle* h);
/* press.f95:49-146 as a whole */
int les_press(les_handle* h);
/* bondFG.f95, boundp.f95 on their own */
int les_bondfg(les_handle* h);
int les_boundp1(les_handle* h);
int les_boundp2(les_handle* h);

/* One whole time step n (main_ocl.f95:264-299 / run_LES_kernel states 1-11). */
int les_step(les_handle* h, int n);
/* Steps n0 .. n0+nsteps-1, replayed from a captured HIP graph of one step. */
int les_run(les_handle* h, int n0, int nsteps);
/* Ensemble of independent members (handles of any sizes, each with its own state and
 * stream), advanced together: every step enqueues one graph launch of each member, step-major,
 * so the members' kernels run concurrently on the device.  No counterpart in the reference
 * (one domain per run); les_ensemble_run(e, n0, k) leaves every member as les_run(member, n0, k)
 * would.  Members keep their ownership (les_ensemble_destroy does not destroy them); mixed
 * precision members are not supported. */
typedef struct les_ensemble les_ensemble;
int les_ensemble_create(les_handle* const* members, int n, les_ensemble** out);
int les_ensemble_run(les_ensemble* e, int n0, int nsteps);
int les_ensemble_sync(les_ensemble* e);
int les_ensemble_destroy(les_ensemble* e);
/* LES_ENSEMBLE_BATCHED (the default): when every member allows it (FP32, exact sums, no SOR
 * early exit, an even nmaxp, a tuned fused SOR tile, even jp, the same kind of horizontal grid,
 * at most 256 members, p and rhs never handed out by address), each les_ensemble_run step is
 * one captured graph on the ensemble's own stream in which velfg + feedbf, calc_sm + boundsm,
 * the viscous terms + adam and every fused SOR iteration are one launch each over the
 * workgroups of all members (a member table per stage deals the launch's workgroups to the
 * members; the fused SOR runs 4-row tiles with each member's tuned depth; calc_sm and visc use
 * the workgroup shape most members' autotuning chose), and the other stages are launched per
 * member inside the same graph.  Members must then outlive the ensemble.  Otherwise, and with
 * LES_ENSEMBLE_CONCURRENT, every member replays its own graph on its own stream.  The results
 * are the same bit for bit.  les_ensemble_batched returns 1 when the last run was batched. */
enum { LES_ENSEMBLE_CONCURRENT = 0, LES_ENSEMBLE_BATCHED = 1 };
int les_ensemble_set_mode(les_ensemble* e, int mode);
int les_ensemble_batched(const les_ensemble* e);
/* ifdata.f95:258-273 on the device: bondv1 (inflow from wind_profile, n=1), boundp1,
 * boundp2, velfg, feedbf, les.  Fields must be zero (les_create) or uploaded. */
int les_ifdata(les_handle* h);

/* ---- aveflow statistics (aveflow.f95:42-157) -------------------------------------------
 * les_aveflow(h, n, n1, nmax) is the reference's call after step n: zeroes the statistics at
 * n == n1, accumulates for n >= n1, and at n == nmax turns them into means, RMS values and
 * spanwise means.  les_aveflow_download copies one array: 'which' in the order below;
 * (ip,jp,kp) arrays for 0..10, (ip,kp) arrays for 11..17, flat with i fastest.  The uw-flux
 * correction at k = 1 reads aveu/avew at level 0, before the arrays; those values read as 0. */
enum {
  LES_AVE_AVEU = 0, LES_AVE_AVEV, LES_AVE_AVEW, LES_AVE_AVEP, LES_AVE_AVEL, LES_AVE_AVEUU,
  LES_AVE_AVEVV, LES_AVE_AVEWW, LES_AVE_AVESM, LES_AVE_AVESMSM, LES_AVE_UWFX,
  LES_AVE_AVESU, LES_AVE_AVESV, LES_AVE_AVESW, LES_AVE_AVESUU, LES_AVE_AVESVV, LES_AVE_AVESWW,
  LES_AVE_UWFXS, LES_AVE_COUNT
};
int les_aveflow(les_handle* h, int n, int n1, int nmax);
int les_aveflow_download(les_handle* h, int which, float* host);

/* ---- restart and output files (Fortran unformatted sequential records) ---------------
 * les_write_restart writes data20 (press.f95:192-199: u, v, w, p, usum, vsum, wsum over the
 * interior, one record each) and data21 (adam.f95:24-29: fold, gold, hold, fghold = 0); with
 * with_header != 0 data20 starts with the record (n, time), which makes it the data30 file of
 * the ICAL = 1 start.  les_read_restart reads data30/data31 as ifdata.f95:227-247 does into the
 * interiors (halos untouched; run les_ifdata afterwards, as the reference's init does) and
 * returns n and time.  Both synchronise. */
int les_write_restart(les_handle* h, const char* data20, const char* data21, int n, float time,
                      int with_header);
int les_read_restart(les_handle* h, const char* data30, const char* data31, int* n, float* time);

/* ---- one domain as sub_i x sub_j tiles on one device ---------------------------------
 * les_tiled_create takes the whole domain's configuration with cfg->sub_i, cfg->sub_j >= 1
 * (at most 8 tiles; ip, jp divisible; the tile sizes even along a cut direction, so that the
 * red-black colours agree; FP32, no SOR early exit).  Each tile is a handle of
 * (ip/sub_i, jp/sub_j, kp) cells.  les_tiled_upload / les_tiled_download take whole-domain
 * arrays in the reference's Fortran bounds (main_ocl.f95:40-167) and slice them into / gather
 * them from the tiles.  les_tiled_ifdata and les_tiled_step are les_ifdata and les_step of
 * the whole domain: the stages run tile by tile with the faces the tiles share, and the
 * periodic j faces, exchanged between stages (the halo exchange of halo_exchange.cl's stages
 * ST_HALO_*, module_LES_combined_ocl_TEMPL_V2.f95:356-363, as device copies since every tile is
 * on this GPU), the domain's own faces under the reference's boundary rules, and the sums
 * combined exactly: the results are bitwise those of one handle on the whole domain. */
typedef struct les_tiled les_tiled;
int les_tiled_create(const les_config* cfg, les_tiled** out);
int les_tiled_destroy(les_tiled* t);
const char* les_tiled_last_error(const les_tiled* t);
int les_tiled_sync(les_tiled* t);
int les_tiled_upload(les_tiled* t, int field, const float* host);
int les_tiled_download(les_tiled* t, int field, float* host);
/* init.f95:82-107 coefficients and delx1 on every tile, from the uploaded grid */
int les_tiled_init_coefficients(les_tiled* t);
int les_tiled_ifdata(les_tiled* t);
int les_tiled_step(les_tiled* t, int n);
/* les_tiled_step for n0 .. n0+nsteps-1, the steps n >= 2 as replays of one captured HIP graph */
int les_tiled_run(les_tiled* t, int n0, int nsteps);

/* NetCDF snapshots (module_LES_write_netcdf.f95, called from press.f95:189): les_netcdf_create
 * is init_netcdf_file (:111-300): it creates dir/LES_output.nc, LES_output_p.nc,
 * LES_output_u/v/w.nc and LES_output_uvwsum.nc in the netCDF classic format with ntimes time
 * levels (NTIMESTEPS = 20 in the reference), every data value set to the float fill value.
 * les_netcdf_write is write_to_netcdf_file(p,u,v,w,usum,vsum,wsum,t) (:302-346): time level t
 * (1-based; the reference passes n/1000) of every variable from the handle's fields;
 * les_netcdf_write_arrays takes host arrays in the Fortran bounds instead.  les_netcdf_close is
 * close_netcdf_file (:348-386). */
typedef struct les_netcdf les_netcdf;
int les_netcdf_create(const char* dir, int ip, int jp, int kp, int ntimes, les_netcdf** out);
int les_netcdf_write(les_netcdf* nc, les_handle* h, int t);
int les_netcdf_write_arrays(les_netcdf* nc, int t, const float* p, const float* u, const float* v,
                            const float* w, const float* usum, const float* vsum,
                            const float* wsum);
int les_netcdf_close(les_netcdf* nc);

/* anime(n,n0,nmax,...,data22,data23,u,w,v,amask1) (anime.f95:5-49, main.f95:224): when n == n0,
 * n == nmax or n % 1000 == 0, writes the grid file data22 and the cell-centred velocity and
 * mask file data23 (Fortran unformatted records); otherwise does nothing.  les_anime_arrays
 * takes host arrays in the Fortran bounds (dx1(-1:ip+1), dy1(0:jp+1), z2(kp+2), u, v, w,
 * amask1) instead of a handle. */
int les_anime(les_handle* h, int n, int n0, int nmax, const char* data22, const char* data23);
int les_anime_arrays(int n, int n0, int nmax, int ip, int jp, int kp, const float* dx1,
                     const float* dy1, const float* z2, const float* u, const float* v,
                     const float* w, const float* amask1, const char* data22,
                     const char* data23);

/* Scalars of the last stages (synchronises). */
typedef struct les_scalars {
  float uout;      /* bondv1.f95:75 */
  float rhsav;     /* press.f95:75 (after the division) */
  float area;      /* press.f95:71 */
  float pav;       /* press.f95:134 (after the division) */
  float pco;       /* press.f95:129 */
  float sor;       /* press.f95:97 residual of the last SOR iteration (when computed) */
  int sor_iters;   /* SOR iterations run in the last solve */
  double resid_l2; /* LES_PREC_MIXED: final relative L2 norm of the pressure residual,
                      |rhs - L p|_2 / |rhs|_2 */
  int refine_passes;
  int cg_iters;    /* LES_PREC_MIXED, refine_solver 1: conjugate-gradient iterations of the last
                      solve, all passes */
} les_scalars;
int les_get_scalars(les_handle* h, les_scalars* out);

/* ---- halo exchange (src/OpenCL/Kernels/halo_exchange.cl, halo_exchange_correct.cl) --------
 * Pack ("read") or unpack ("write") the h_w-wide N/S/W/E ring of one or more fields into a
 * linear buffer.  The buffer holds v_dim consecutive blocks (one per field, in the order
 * given), each laid out north|south|west|east as in halo_exchange.cl:3-51 and :156-185;
 * "extent" arrays are the allocated (im, jm, km) of the fields, which must agree.
 * correct = 0: halo_exchange.cl (ring rows/columns on the array edge, corners in N/S);
 * correct = 1: halo_exchange_correct.cl read (inner ring, corners excluded). */
int64_t les_halo_size(int v_dim, int im, int jm, int km, int h_w, int correct);
int les_exchange_halo_read(les_handle* h, const int* fields, int v_dim, int h_w, int correct,
                           float* dev_buffer);
int les_exchange_halo_write(les_handle* h, const int* fields, int v_dim, int h_w,
                            const float* dev_buffer);
/* The same pack/unpack on caller-owned device arrays of allocated extent (im, jm, km), one per
 * vector component: exchange_{1,2,4,16}_halo_read(array, buffer, im, jm, km, h_w)
 * (halo_exchange.cl:156-290, halo_exchange_correct.cl:156-290) and
 * exchange_{1,2,4,16}_halo_write (halo_exchange.cl:3-155), with the OpenCL vector's
 * components as separate arrays (v_dim = 1..16).  Enqueued on the handle's stream. */
int les_halo_read_arrays(les_handle* h, float* const* dev_arrays, int v_dim, int im, int jm,
                         int km, int h_w, int correct, float* dev_buffer);
int les_halo_write_arrays(les_handle* h, float* const* dev_arrays, int v_dim, int im, int jm,
                          int km, int h_w, const float* dev_buffer);

/* Exact, order-independent sum of n device floats, rounded once to FP32 (the reduction that
 * rhsav, pav and the SOR residual use), result written to *out on the host. */
int les_reduce_sum_exact(les_handle* h, const float* dev_x, int64_t n, float* out);

/* les_reduce_sum_exact without a handle and without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same exact sum of n device floats
 * (bit-identical result), written to the device float *dev_out, enqueued on 'stream' (a
 * hipStream_t; NULL = the null stream) of the current device; returns without waiting.  The
 * workgroup partials live in a fixed device array of the library (160 KiB per device, held
 * until the process ends) that every call on a device shares: calls that overlap on one device
 * race, so serialise them (one stream).  n is at most LES_NOALLOC_SUM_MAX (else
 * LES_E_INVALID); LES_E_HIP when the array's device address cannot be obtained or the launch
 * fails.  Calls no hipMalloc/hipFree. */
#define LES_NOALLOC_SUM_MAX 2147483647LL
int les_reduce_sum_exact_noalloc(const float* dev_x, int64_t n, float* dev_out, void* stream);

/* ---- measurement (replaces the per-launch exectime of runOcl, TEMPL:366-370,1061-1071) ----
 * Launches one kernel of the step 'reps' times on the handle's stream between two HIP events
 * and returns the average duration per launch in milliseconds.  The kernel runs on the
 * current device state (it changes it, as the step would). */
enum {
  LES_KERNEL_SOR_SWEEP = 0,  /* one red-black colour sweep (the dominant kernel) */
  LES_KERNEL_VELFG = 1,      /* velfg + feedbf */
  LES_KERNEL_VISC_ADAM = 2,  /* les viscous terms + adam */
  LES_KERNEL_CALC_SM = 3,
  LES_KERNEL_RHS = 4,
  LES_KERNEL_VELNW = 5,
  LES_KERNEL_STEP = 6,       /* one whole step replayed from the captured graph */
  LES_KERNEL_SOR_ITER = 7    /* one fused SOR iteration (both colours; uniform grids, ip <= 320) */
};
int les_time_kernel(les_handle* h, int which, int reps, float* avg_ms);

/* Autotuner (no counterpart in the reference, whose OpenCL launch shapes are fixed): times one
 * SOR iteration without residual as two colour sweeps and as the fused launch with every
 * workgroup tile (rows of j x levels of k), 'reps' iterations each, on this handle's grid, and
 * keeps the fastest for later solves (untuned handles use the two sweeps); then times the
 * velnw, calc_sm and viscous-term kernels with each of their workgroup shapes (64x4, 128x2,
 * 32x8, 256x1) and keeps the fastest of each.  Changes no simulation state (the arrays the
 * stencil trials update are saved and restored).  *out (may be NULL) receives the choice. */
typedef struct les_tuning {
  int sor_tile_j, sor_tile_k;  /* the fused tile, or 0, 0 for the two colour sweeps */
  float sor_iter_ms;           /* average time of one iteration with that choice, or -1 when
                                  the grid allows neither (odd jp, ip > 320 or non-uniform) */
  /* workgroup shapes (lanes along i, rows of j) chosen for the point-wise stencil kernels
   * [0] velnw, [1] les calc_sm (with boundsm), [2] les viscous terms + adam, and the average
   * time of one launch with that shape */
  int stencil_bx[3], stencil_by[3];
  float stencil_ms[3];
} les_tuning;
int les_autotune(les_handle* h, int reps, les_tuning* out);
/* Sets the SOR iteration form by hand: tile_j = 0 for two colour sweeps, else the fused launch
 * with tile_j (4, 8 or 16) rows of j and tile_k levels of k per workgroup. */
int les_set_sor_tile(les_handle* h, int tile_j, int tile_k);

/* ---- the reference's GPU host API (module_LES_combined_ocl) as single-handle drop-ins ---- */
/* initialise_LES_kernel (Wrappers/module_LES_combined_ocl.f95:15-441), external wind profile
 * form: creates the library's default handle and uploads every array. */
int les_initialise_LES_kernel(
    const float* p, const float* u, const float* v, const float* w, const float* usum,
    const float* vsum, const float* wsum, const float* f, const float* g, const float* h,
    const float* fold, const float* gold, const float* hold, const float* diu1,
    const float* diu2, const float* diu3, const float* diu4, const float* diu5,
    const float* diu6, const float* diu7, const float* diu8, const float* diu9,
    const float* amask1, const float* bmask1, const float* cmask1, const float* dmask1,
    const float* cn1, const float* cn2l, const float* cn2s, const float* cn3l,
    const float* cn3s, const float* cn4l, const float* cn4s, const float* rhs,
    const float* sm, const float* dxs, const float* dys, const float* dzs, const float* dx1,
    const float* dy1, const float* dzn, const float* wind_profile, float dt, int im, int jm,
    int km);
/* run_LES_kernel(n, nmax, wind_profile) (TEMPL:219-1072): one step on the default handle. */
int les_run_LES_kernel(int n, int nmax, const float* wind_profile);
/* The default handle (for downloads/checkpoints), or NULL. */
les_handle* les_default_handle(void);
/* Releases the default handle. */
int les_finalise_LES_kernel(void);

/* ---- synthetic inputs for the benchmark configurations (host only, deterministic) ------- */
/* Grid: dx1 = dx, dy1 = dy everywhere; dzn = dz0 while the cell top is <= z_flat, then each
 * level is the previous one times ratio; halos, z2, dxs, dys, dzs as grid.f95:41-82. */
int les_gen_grid(int ip, int jp, int kp, float dx, float dy, float dz0, float z_flat,
                 float ratio, float* dx1, float* dxs, float* dy1, float* dys, float* dzn,
                 float* dzs, float* z2);
/* Terrain: zbm(-1:ip+1,-1:jp+1) of seeded rectangular blocks (widths in [wmin,wmax] m,
 * heights in [0,hmax] m), max-overlap, until >= coverage of the (i,j) columns are covered. */
int les_gen_terrain(int ip, int jp, float dx, float dy, uint64_t seed, float coverage,
                    float wmin, float wmax, float hmax, float* zbm);
/* feedbfm.f95:42-63: masks from zbm (amask1 = 0 where zbm(i,j) > z2(k)+0.5*dzn(k), then
 * bmask1 = cmask1 = dmask1 = 1 there); halos are left 0. */
int les_gen_masks(int ip, int jp, int kp, const float* zbm, const float* z2, const float* dzn,
                  float* amask1, float* bmask1, float* cmask1, float* dmask1);
/* feedbfm.f95:31-37: reads the GIS building-height file (the format of GIS/Tokyo_20mgrid.txt:
 * 100 x 100 heights, one per line, rows j = 100..1) into zbm(-1:ipmax+1,-1:jpmax+1) at
 * (i+25, j+25); other entries are left as they are.  Needs ipmax, jpmax >= 125. */
int les_gis_read_zbm(const char* path, int ipmax, int jpmax, float* zbm);
/* feedbfm.f95:42-63 on the device: the four masks (interiors; halos cleared) and the stencils'
 * bit mask from zbm(-1:ip+1,-1:jp+1); z2 and dzn must be uploaded. */
int les_set_terrain(les_handle* h, const float* zbm);
/* Log-law inflow u(z) = (ustar/0.4) ln(z/z0) (1 + noise*xi_jk), z = z2(k)+0.5*dzn(k)
 * (the inflow height of bondv1.f95:29), xi ~ N(0,1) seeded. */
int les_gen_inflow(int jp, int kp, const float* z2, const float* dzn, float ustar, float z0,
                   float noise, uint64_t seed, float* wind_profile);
/* initial_wind_profile.f95:22-32 (the reference's power-law profile, kp >= 78). */
int les_initial_wind_profile(int jp, int kp, const float* z2, const float* dzn,
                             float* wind_profile);

/* The adjoint form.  No entry point of the reference corresponds: the reference has no
 * derivative.  One handle holds the step's arrays in FP32 (fp64 = 0) or FP64 and runs the step
 * of main_ocl.f95:264-299 for n >= 2 (EXTERNAL_WIND_PROFILE) as generated statement kernels,
 * forward (les_adj_step) and transposed in reverse order (les_adj_reverse_step: one step's
 * vector-Jacobian product, the step recomputed from the current state first).  An FP32 handle
 * keeps an internal plain handle on the same grid whose kernels run the SOR (forward and
 * transposed) and the window's first forward pass, with the same values.  feedbf's beta
 * term is beta * cd(i,j,k) * u * mask with a drag coefficient field cd (field LES_ADJ_CD, 1 =
 * the reference).  Fields go in and out as doubles in the reference's Fortran bounds;
 * les_adj_{upload,download}_adjoint address the adjoints of the carried state (u, v, w, p, f, g,
 * h, fold, gold, hold, usum, vsum, wsum) and of the controls (LES_WIND_PROFILE, LES_ADJ_CD).
 * les_adj_set_sums gives press.f95's volume sums (area, pco).  les_adj_window advances nsteps
 * steps from the current state with a checkpoint every ckpt steps, returns the objective
 * J = mean over the steps and the nprobes (i,j,k) probes of u after each step, and leaves the
 * gradient of J in the adjoints of the controls and of the initial state (recomputing each
 * segment from its checkpoint into per-step copies of every array the step writes, which the
 * transposed steps then read); the handle's state is the state after the window.
 * les_adj_set_slots(h, 0) makes the window keep no per-step copies and recompute each step
 * again from its state before its transpose instead (less memory, one forward pass more, the
 * same gradient); 1 (the default) keeps them when they fit in device memory.  The window keeps
 * p after every step of its first forward pass, so the recomputed steps do not solve again.
 * les_adj_set_transposes(h, 1) (the default) transposes the blocks of statements that have
 * hand-written fused kernels (vel2 + velFG; les: calc_sm, boundsm, visc) with those; 0 runs every statement's generated
 * transpose (the same values in exact arithmetic; the check of the hand-written kernels). */
#define LES_ADJ_CD 100
typedef struct les_adj les_adj;
int les_adj_create(const les_config* cfg, int fp64, les_adj** out);
int les_adj_destroy(les_adj* h);
const char* les_adj_last_error(const les_adj* h);
int les_adj_upload(les_adj* h, int field, const double* host);
int les_adj_download(les_adj* h, int field, double* host);
int les_adj_upload_adjoint(les_adj* h, int field, const double* host);
int les_adj_download_adjoint(les_adj* h, int field, double* host);
int les_adj_set_sums(les_adj* h, double area, double pco);
int les_adj_zero_adjoints(les_adj* h);
int les_adj_step(les_adj* h, int nsteps);
int les_adj_reverse_step(les_adj* h);
int les_adj_window(les_adj* h, int nsteps, int ckpt, const int* probes, int nprobes,
                   double* objective);
int les_adj_set_slots(les_adj* h, int enable);
int les_adj_set_transposes(les_adj* h, int hand);

#ifdef __cplusplus
}
#endif
#endif /* LES_HIP_H */
