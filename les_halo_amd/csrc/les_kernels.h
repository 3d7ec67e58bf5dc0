// Device state shared by the kernels and the host API.
#pragma once
#include "les_common.h"

struct les_handle;

namespace les {

// Device pointers of every field, in the reference's Fortran layouts.
struct Fields {
  float *p, *u, *v, *w, *usum, *vsum, *wsum, *f, *g, *h, *fold, *gold, *hold, *rhs, *sm;
  float *diu1, *diu2, *diu3, *diu4, *diu5, *diu6, *diu7, *diu8, *diu9;
  const uint8_t* mask;  // interior (1:ip,1:jp,1:kp): bit0 bmask1, bit1 cmask1, bit2 dmask1
  const float* cn1;     // cn1(1:ip,1:jp,1:kp)
  const float* wind;    // wind_profile(1:jp,1:kp)
  // drag coefficient of feedbf's beta term per interior cell (i fastest), beta*cd*u*mask; the
  // reference's term (cd = 1) when null.  Set by the adjoint form's forward only.
  const float* cd = nullptr;
};

// Scalars kept on the device between stages (graph-capturable: no host round trip).
struct DevScalars {
  float uout, rhsav, area, pav, pco, sor;
  int sor_iters, sor_done;
  double resid_l2;          // mixed precision: |rhs - L p|_2 / |rhs|_2 of the last check
  int refine_passes;
  unsigned int uout_count;  // workgroups of k_uout done (re-armed to 0 by the last one)
  // mixed precision (FP64 reductions and iterate)
  double rhsav64, area64, pav64, rhs_l2;
  int refine_done;
  float uout_max, uout_min;  // bondv1.f95:71-72 before the average (max seeded and min seeded 0)
  float pav_carry;  // the pav that k_split_gather_adj took off p while the split copy kept p
};

struct Physics {
  float dt, ro, vn, alpha, beta, cs0, omega, pjuge;
};

// Red/black split copy of p and rhs used inside the SOR solve: interior cell (i,j,k) lives in
// array (i+j+k)&1 at ((i-1)>>1) + nr*((j-1) + nj*(k-1)); each (j,k) row of one colour is
// contiguous from element 0, rows padded to nr (a multiple of 4) floats.
// cn1(i,j,k) = 1/((ax(i)+by(j))+cz(k)) (init.f95:98-104); on a uniform horizontal grid
// (uniform != 0) the sweep uses cn2l(1), cn2s(1), cn3l(1), cn3s(1) and cn1k(k) instead.
struct SplitSOR {
  float* p[2];
  float* r[2];
  const float *ax, *by, *cz, *cn1k;
  int nr, nj;  // padded colour row length, rows per level (jp, or jp + 2 with jhalo)
  int uniform;
  uint32_t m_ng, m_jp;  // multiply-high reciprocals of nr/4 and jp (sor_split_divisors)
  // A tile of a cut domain (k_sor_split only): jhalo stores rows j = jp+1 and j = 0 of each
  // colour after the interior rows (row indices jp and jp+1) instead of wrapping around in j;
  // hw / he, when set, hold p(0,j,k) / p(ip+1,j,k) at (j-1) + jp*(k-1) for a west / east face
  // shared with a neighbour (else the cell's own value is its image, boundp1).
  int jhalo;
  const float *hw, *he;
};
// Per-cell residual terms reltmp**2 of one SOR iteration, for the reference-order sum
// (les_seqsum.hip): colour c's cell (i,j,k) at ((i-1)>>1) + nrt*((j-1) + jp*(k-1)) of t[c],
// nrt = (ip+1)/2.  t[0] == nullptr: the sweeps accumulate exact partials instead.
struct SeqTerms {
  float* t[2];
  int nrt;
};
// The split copies of the tiles of one cut domain and their neighbours, for the face exchange
// of the split layout (k_split_exchange) after the scatter
constexpr int SPLIT_TILES_MAX = 8;
struct SplitTiles {
  float* p[SPLIT_TILES_MAX][2];
  float* hw[SPLIT_TILES_MAX];  // nullptr on the domain's west face
  float* he[SPLIT_TILES_MAX];  // nullptr on the domain's east face
  int n[SPLIT_TILES_MAX], s[SPLIT_TILES_MAX], w[SPLIT_TILES_MAX], e[SPLIT_TILES_MAX];
  int ntiles, nr, nj;
};

// Fused SOR iteration (k_sor_iter): previous and next iterate, colour 0 and colour 1 each
struct SorPair {
  const float* pin[2];
  float* pout[2];
  // The first iteration of a solve on the resident split copy (k_sor_iter<.., FIRST>): it reads
  // the raw rhs that k_rhs stored in the split layout and subtracts rhsav (press.f95:76-82)
  // from every rhs value it reads, writing the result for the cells it owns to rout; with
  // p_shift it also subtracts DevScalars::pav_carry from every p value it reads (the split
  // copy then holds the previous solve's iterate before its pav was taken off).
  float* rout[2];
  int p_shift;
};
// workgroup tile of k_sor_iter: tj rows of j (4, 8 or 16), kc levels of k
struct SorIterShape {
  int tj, kc;
};
constexpr int SOR_ITER_NRMAX = 160;  // longest colour row (ip <= 320) the fused kernel stages in LDS
constexpr int SOR_ITER_KCMAX = 128;  // most levels per workgroup of the fused kernel

// Reduction scratch: per-block carry-save limbs.
struct RedBuf {
  int64_t* partials;  // [max_blocks][NLIMB]
  int max_blocks;
  int64_t* slots;     // [RED_SLOTS][NLIMB], zero between uses (nullptr: partials only)
};

// Which faces of a (sub)domain are faces of the whole domain.  A tile of a domain cut into
// subdomains on one device (les_tiled) takes its internal faces, and the periodic j faces, from
// its neighbours by copies after the stage instead of applying the single-domain rules there.
struct Edges {
  int w, e;   // the west (i = 0) / east (i = ip+1) face is the domain's inflow / outflow face
  int s, n;   // the south (j = 0) / north (j = jp+1) face is the domain's j = 0 / jm+1 face
  int tiled;  // 0: one domain (periodic j copies made in place); 1: faces come by exchange
};
constexpr Edges EDGES_ALL{1, 1, 1, 1, 0};

// box copy between two arrays of the same field (the halo exchange of les_tiled): element
// (i,j,k) of an array is at p[(i-lo0) + n0*((j-lo1) + n1*(k-lo2))]
struct ArrView {
  float* p;
  int lo0, lo1, lo2, n0, n1;
};
struct BoxCopy {
  ArrView src, dst;
  int si, sj, di, dj, k0;  // first source / destination (i, j), first level
  int ni, nj, nk;          // box extents
};
constexpr int EXCH_MAX = 32;  // boxes per exchange launch
struct ExchangePlan {
  int nbox;
  BoxCopy box[EXCH_MAX];
};
void launch_exchange(hipStream_t st, const ExchangePlan& plan);
// copies the float at byte offset off of src into the same place of up to 8 DevScalars
void launch_scalar_bcast(hipStream_t st, const DevScalars* src, DevScalars* const* dst, int ndst,
                         int off);
// uout of les_tiled: (max(0, max of the uout_max) + min(0, min of the uout_min)) / 2 over the
// east tiles, written to every tile
void launch_uout_combine(hipStream_t st, DevScalars* const* east, int neast,
                         DevScalars* const* all, int nall);
void launch_rhsav_finalize(hipStream_t st, const int64_t* partials, int nb, DevScalars* sc);
void launch_pav_finalize(hipStream_t st, const int64_t* partials, int nb, DevScalars* sc);
// boundsm of a tile in three phases: 0 physical i faces, 1 physical j faces, 2 k planes
void launch_boundsm_tiled(hipStream_t st, Dims d, Fields F, Edges e, int phase);

// ---- launchers (les_kernels.hip); all asynchronous on 'st' ----
// Workgroup shapes (lanes along i x rows of j, 256 threads) the autotuner chooses from for the
// point-wise stencil kernels (velnw, calc_sm, visc+adam); shape 0 is the untuned default.
constexpr int STENCIL_SHAPES = 4;
__host__ __device__ inline int stencil_shape_x(int s) {
  return s == 1 ? 128 : (s == 2 ? 32 : (s == 3 ? 256 : 64));
}
__host__ __device__ inline int stencil_shape_y(int s) { return 256 / stencil_shape_x(s); }
void launch_velnw_inflow(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velnw,
                         int inflow_n, int shape = 0);
void launch_uout(hipStream_t st, Dims d, Fields F, DevScalars* sc, float* part);
void launch_bondv1_uvw(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, DevScalars* sc,
                       Edges e = EDGES_ALL, int do_outflow = 1);
void launch_velfg(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int do_velfg,
                  int do_feedbf, Edges e = EDGES_ALL);
// with_bound: also boundsm.f95's halo cells, as k_boundsm would set them afterwards
void launch_calc_sm(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int with_bound = 0,
                    int shape = 0);
// workgroup rows along k of k_calc_sm (a thread takes consecutive levels)
int calc_sm_kblocks(int kp);
void launch_boundsm(hipStream_t st, Dims d, Fields F);
// workgroup rows along k of k_visc_adam (a thread takes two levels)
int visc_kblocks(int kp);
void launch_calc_visc_adam(hipStream_t st, Dims d, Grid1D g, Fields F, int do_visc,
                           int do_adam, int shape = 0);
// split != nullptr: the raw rhs goes to the split copy's rhs arrays, not to F.rhs
void launch_rhs(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, RedBuf rb,
                DevScalars* sc, Edges e = EDGES_ALL, int finalize = 1,
                const SplitSOR* split = nullptr);
void launch_bondfg(hipStream_t st, Dims d, Fields F);
void launch_rhs_adjust(hipStream_t st, Dims d, Fields F, DevScalars* sc);
// want_resid: the sweep adds its residual terms' exact sum into the slots rb.slots (or, with
// tm.t[0] set, stores each term for the reference-order sum)
void launch_sor_colour(hipStream_t st, Dims d, Grid1D g, Fields F, Physics ph, int nrd,
                       int want_resid, RedBuf rb, DevScalars* sc, SeqTerms tm = SeqTerms{});
// the iteration's residual from rb.slots (clearing them) and the early exit
void launch_sor_resid_finalize(hipStream_t st, RedBuf rb, Physics ph, DevScalars* sc);
void launch_boundp1(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done,
                    Edges e = EDGES_ALL);
void launch_boundp2(hipStream_t st, Dims d, Fields F, DevScalars* sc, int check_done);
void launch_sor_reset(hipStream_t st, DevScalars* sc);
void launch_pav(hipStream_t st, Dims d, Grid1D g, Fields F, RedBuf rb, DevScalars* sc,
                int finalize = 1);
void launch_adj(hipStream_t st, Dims d, Fields F, DevScalars* sc, int with_boundp);
int sor_resid_blocks(Dims d);

// split-layout SOR (jp even; p halos mirror-consistent on entry)
// adjust_rhs: rhs -= rhsav on the way (also stored back to F.rhs); with_rhs = 0: p only
void launch_split_scatter(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                          int adjust_rhs, int with_rhs = 1);
void launch_sor_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, int nrd,
                      int want_resid, RedBuf rb, DevScalars* sc, SeqTerms tm = SeqTerms{});
// press.f95's sums in the reference's sequential order (les_seqsum.hip): rhsav over the Fortran
// rhs; pav over the Fortran p (sp == nullptr) or the split copy; the residual of one iteration
// over the terms the sweeps stored, with the early exit
void launch_seq_rhsav(hipStream_t st, Dims d, Grid1D g, const float* rhs, DevScalars* sc);
void launch_seq_pav(hipStream_t st, Dims d, Grid1D g, const float* p, const SplitSOR* sp,
                    DevScalars* sc);
void launch_seq_resid(hipStream_t st, Dims d, SeqTerms tm, Physics ph, DevScalars* sc);
// sets SplitSOR::m_ng, m_jp; false when the grid is too large for the multiply-high quotients
bool sor_split_divisors(Dims d, SplitSOR& sp);
int sor_iter_supported(Dims d, const SplitSOR& sp);
// first != 0: the first iteration of a resident solve (see SorPair)
void launch_sor_iter(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, SorPair io,
                     DevScalars* sc, SorIterShape shape, int first = 0);
void launch_pav_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, RedBuf rb, DevScalars* sc);
// Where a tile's sweep stores the cells it updates that are a neighbour's halo: its row j = 1
// goes to row index jp (j = jp+1) of the south tile's s[c], its row j = jp to row index jp+1
// (j = 0) of the north tile's n[c], its p(ip,j,k) to the east tile's hw, its p(1,j,k) to the
// west tile's he (nullptr: no neighbour on that face).  During a sweep of colour c a tile reads
// only colour 1-c halo entries, so these stores never meet a read of the same sweep.
struct SplitPush {
  float* n[2];
  float* s[2];
  float *e, *w;
};
// the sweep operands of every tile of a cut domain (k_sor_split_tiles)
struct SorTiles {
  SplitSOR sp[SPLIT_TILES_MAX];
  Grid1D g[SPLIT_TILES_MAX];
  const DevScalars* sc[SPLIT_TILES_MAX];
  SplitPush push[SPLIT_TILES_MAX];
  int ntiles;
};
// one colour sweep (no residual) of every tile, all tiles of the same size d; the sweep stores
// the tiles' updated face cells into their neighbours' halos (SplitPush), so one exchange after
// the scatter keeps every halo current for the whole solve
void launch_sor_split_tiles(hipStream_t st, Dims d, const SorTiles& T, Physics ph, int nrd);
// fills every tile's halo rows (both colours) and i-halo columns from its neighbours' interior
void launch_split_exchange(hipStream_t st, Dims d, const SplitTiles& T);
// set_carry: also records pav in DevScalars::pav_carry (the split copy keeps p before adj)
void launch_split_gather_adj(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                             int subtract_pav, int set_carry = 0);
// F.rhs over the interior from the split copy's rhs arrays ra (the resident solve's
// adjusted rhs): the Fortran array materialised on demand
void launch_split_gather_rhs(hipStream_t st, Dims d, Fields F, SplitSOR sp);
// Transposes of the colour-split solve (scatter, iterations, gather) for adjoints pb of p and
// rb of rhs in their Fortran layouts; the split adjoints live in sp.p (of p) and sp.r (of rhs).
// gather_t: sp.p = the sum of pb over the box cells mirroring each interior cell, those pb
// zeroed, sp.r zeroed.  sweep_t on the colour-c cells (uniform horizontal grid), mode 0: add
// the gather from the other colour; 1: scale by the own coefficient, accumulate the rhs
// adjoint, add the gather; 2: scale and accumulate only.  scatter_t: pb, rb += sp.p, sp.r.
void launch_split_gather_t(hipStream_t st, Dims d, float* pb, SplitSOR sp);
void launch_split_sweep_t(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, float omega, int c,
                          int mode);
void launch_split_scatter_t(hipStream_t st, Dims d, float* pb, float* rb, SplitSOR sp);

// ---- the batched ensemble (one launch per stage over every member's workgroups) ----
// Per-member operands of the batched stencil launches, and each member's first logical
// workgroup in the launch (members in order, their workgroups in each kernel's own order).
struct BatchStencil {
  Dims d;
  Grid1D g;
  Fields F;
  Physics ph;
  StripGrid sg_velfg, sg_sm, sg_visc;
  int base_velfg, base_sm, base_visc;
};
// Per-member operands of the batched fused SOR iteration (TJ = 4, 256 threads): the first
// iteration of the resident solve (reads sp's raw rhs) and the following ones (sra: the
// adjusted rhs), in the two ping-pong directions.
struct BatchSor {
  Dims d;
  Grid1D g;
  SplitSOR sp, sra;
  Physics ph;
  SorPair first, fwd, back;
  const DevScalars* sc;
  int kc, njt, base;
};
// wgm[w]: the member of logical workgroup w (at most 256 members per batch)
void launch_velfg_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm, int total,
                        int p2xy);
void launch_calc_sm_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm, int total,
                          int shape);
void launch_visc_adam_batch(hipStream_t st, const BatchStencil* tab, const uint8_t* wgm,
                            int total, int shape, int p2xy);
// which: 0 = first iteration, 1 = fwd, 2 = back
void launch_sor_iter_batch(hipStream_t st, const BatchSor* tab, const uint8_t* wgm, int total,
                           int which);
// velfg's workgroup grid for a member (the strip order k_velfg_xy uses) and its block shape
StripGrid velfg_strip(Dims d);
constexpr int VELFG_THREADS = 512;
int reduction_blocks(Dims d);

// exact sum of n device floats into *out (device), for tests of the reduction itself
void launch_exact_sum(hipStream_t st, const float* x, long n, RedBuf rb, float* out);

// aveflow.f95 statistics: (ip,jp,kp) arrays a3[AVE_*] and (ip,kp) arrays a2 (avesu, avesv,
// avesw, avesuu, avesvv, avesww, uwfxs), flat with i fastest
enum { AVE_U, AVE_V, AVE_W, AVE_P, AVE_L, AVE_UU, AVE_VV, AVE_WW, AVE_SM, AVE_SMSM, AVE_UWFX,
       AVE_N3 };
constexpr int AVE_N2 = 7;
struct AveStats {
  float* a3[AVE_N3];
  float* a2[AVE_N2];
};
void launch_aveflow(hipStream_t st, Dims d, Fields F, AveStats s, int n, int n1, int nmax);

// feedbfm.f95:42-63 masks from building heights
void launch_feedbfm(hipStream_t st, Dims d, Grid1D g, const float* z2, const float* zbm,
                    float* am, float* bm, float* cm, float* dm, uint8_t* bits);

// halo exchange pack/unpack (halo_exchange.cl semantics, SoA over the fields)
void launch_halo_read(hipStream_t st, const float* const* arrays, int v_dim, int im, int jm,
                      int km, int h_w, int correct, float* buffer);
void launch_halo_write(hipStream_t st, float* const* arrays, int v_dim, int im, int jm, int km,
                       int h_w, const float* buffer);

// press.f95:87-115 by a plain handle's colour-split solve on p and rhs held elsewhere, on
// stream st (les_api.hip; the adjoint form's FP32 forward)
int sor_solve_external(::les_handle* h, hipStream_t st, float* p, float* rhs);
// its transpose (uniform horizontal grid): adjoints pb of p and rb of rhs
int sor_adjoint_external(::les_handle* h, hipStream_t st, float* pb, float* rb);
// a plain handle running the adjoint form's FP32 forward window (les_api.hip): its field
// arrays, the drag field of its feedbf, whether p's halos are mirror images, one eager step of
// main_ocl.f95:264-299 on stream st
float* field_ptr(::les_handle* h, int field);
void set_drag(::les_handle* h, const float* cd);
void set_p_consistent(::les_handle* h, bool consistent);
int step_external(::les_handle* h, hipStream_t st, int n);
// the stages of that step before the pressure solve (velnw to visc/adam), for the adjoint
// form's recomputation: u, v, w after velnw and the inflow (before the outflow) are copied to
// u_pre, v_pre, w_pre, and uout to uout_dst (device pointers)
int stages_external(::les_handle* h, hipStream_t st, int n, float* u_pre, float* v_pre,
                    float* w_pre, float* uout_dst);

}  // namespace les
