// C ABI of libles_hip.so (include/les_hip.h): handle, device arrays, uploads/downloads and
// the stage orchestration that replaces run_LES_kernel's state machine
// (src/OpenCL/module_LES_combined_ocl_TEMPL_V2.f95:219-1072).  Everything between the stages
// stays on the device: scalars (uout, rhsav, pav, residual) live in device memory and the
// SOR early exit is a device flag, so a whole step can be captured in one HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/les_hip.h"
#include "les_kernels.h"
#include "les_mixed.h"

using namespace les;

struct les_handle {
  les_config cfg;
  Dims d;
  Physics ph;
  hipStream_t st = nullptr;
  float* fld[LES_NFIELDS] = {};
  int64_t cnt[LES_NFIELDS] = {};
  uint8_t* mask = nullptr;  // interior bits: 1 b, 2 c, 4 d, 8 a
  double2* rz = nullptr;     // Grid1D::rzn, rzs, rzp, each cnt[LES_DZN] long (upload_zrecip)
  DevScalars* sc = nullptr;
  RedBuf rb{};
  std::vector<float> hgrid[LES_NFIELDS];  // host copies of the 1-D grid/coefficient arrays
  std::string err;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  float* scratch_f = nullptr;  // single-float device scratch
  // colour-split SOR copy and the 1-D factors of cn1
  SplitSOR sp{};
  float* split_mem = nullptr;
  std::vector<float> hcn1;     // host copy of cn1 (to know whether it is the separable form)
  bool cn1_separable = false;  // cn1 == 1/((ax(i)+by(j))+cz(k)) bit for bit
  bool p_consistent = true;    // every halo of p equals its mirror interior cell
  // mixed precision: FP64 iterate and rhs in the split layout, block partials
  Mixed mx{};
  double* mx_mem = nullptr;
  float* cg_mem = nullptr;  // mixed precision: the conjugate-gradient vectors
  // second iterate of the fused SOR iteration (ping-pong with sp.p) and its workgroup tile
  float* palt[2] = {nullptr, nullptr};
  SorIterShape iter_shape{0, 0};  // tj == 0: two colour sweeps per iteration (untuned default)
  const float* drag = nullptr;    // feedbf's drag coefficient field (Fields::cd), null: 1
  bool tuned = false;
  bool p2xy = false;  // see grid_p2xy
  bool split_ok = false;  // the colour-split sweep's index arithmetic covers this grid
  // aveflow statistics, allocated by the first les_aveflow call
  float* ave_mem = nullptr;
  AveStats ave{};
  // LES_SUM_REFERENCE: per-cell residual terms of one SOR iteration (both colours)
  float* seq_mem = nullptr;
  SeqTerms seq_terms{};
  // Resident split solve (enqueue_press): between steps the split copy keeps the last
  // iterate, p = iterate - pav_carry, so the next solve needs no scatter of p; k_rhs stores
  // the raw rhs in the split layout and the first fused iteration writes the adjusted rhs to
  // ra; F.rhs is filled from ra only when something reads it (sync_rhs).
  float* ra[2] = {nullptr, nullptr};
  bool resident_ok = true;    // off once p or rhs are handed out by address, in mixed
                              // precision and with reference-order sums
  bool sp_p_current = false;  // the split copy holds the carry state of F.p's interior
  bool rhs_stale = false;     // F.rhs's interior is behind ra
  bool graph_resident = false;  // the captured step is a resident solve (its replays leave
                                // sp_p_current and rhs_stale set)
  // workgroup shapes of the point-wise stencil kernels (les_autotune; 0 = 64 x 4)
  int shp_velnw = 0, shp_sm = 0, shp_visc = 0;
};

namespace {

les_handle* g_default = nullptr;

int fail(les_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((h), LES_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Bounds table: lo/hi offsets relative to (0, ip, jp, kp) per dimension.
struct BoundSpec {
  int ndim;
  int lo[3];
  int hi_base[3];  // 0 = ip, 1 = jp, 2 = kp
  int hi_off[3];
};

bool spec_of(int f, BoundSpec& s) {
  auto three = [&](int l0, int h0, int l1, int h1, int l2, int h2) {
    s.ndim = 3;
    s.lo[0] = l0; s.hi_base[0] = 0; s.hi_off[0] = h0;
    s.lo[1] = l1; s.hi_base[1] = 1; s.hi_off[1] = h1;
    s.lo[2] = l2; s.hi_base[2] = 2; s.hi_off[2] = h2;
    return true;
  };
  auto one = [&](int base, int l0, int h0) {
    s.ndim = 1;
    s.lo[0] = l0; s.hi_base[0] = base; s.hi_off[0] = h0;
    return true;
  };
  switch (f) {
    case LES_P: return three(0, 2, 0, 2, 0, 1);
    case LES_U: case LES_V: return three(0, 1, -1, 1, 0, 1);
    case LES_W: return three(0, 1, -1, 1, -1, 1);
    case LES_USUM: case LES_VSUM: case LES_WSUM: case LES_F: case LES_G: case LES_H:
      return three(0, 0, 0, 0, 0, 0);
    case LES_FOLD: case LES_GOLD: case LES_HOLD: case LES_CN1: return three(1, 0, 1, 0, 1, 0);
    case LES_RHS: return three(0, 1, 0, 1, 0, 1);
    case LES_SM: return three(-1, 1, -1, 1, 0, 1);
    case LES_DIU1: case LES_DIU5: return three(-1, 2, 0, 2, 0, 2);
    case LES_DIU2: case LES_DIU3: case LES_DIU4: case LES_DIU6: case LES_DIU7: case LES_DIU8:
    case LES_DIU9: return three(0, 2, 0, 2, 0, 2);
    case LES_AMASK1: case LES_DMASK1: return three(0, 1, 0, 1, 0, 1);
    case LES_BMASK1: return three(-1, 1, 0, 1, 0, 1);
    case LES_CMASK1: return three(0, 1, -1, 1, 0, 1);
    case LES_CN2L: case LES_CN2S: return one(0, 1, 0);
    case LES_CN3L: case LES_CN3S: return one(1, 1, 0);
    case LES_CN4L: case LES_CN4S: case LES_DELX1: return one(2, 1, 0);
    case LES_DX1: return one(0, -1, 1);
    case LES_DXS: return one(0, 0, 0);
    case LES_DY1: return one(1, 0, 1);
    case LES_DYS: return one(1, 0, 0);
    case LES_DZN: case LES_DZS: return one(2, -1, 2);
    case LES_Z2: return one(2, 1, 2);
    case LES_WIND_PROFILE:
      s.ndim = 2;
      s.lo[0] = 1; s.hi_base[0] = 1; s.hi_off[0] = 0;
      s.lo[1] = 1; s.hi_base[1] = 2; s.hi_off[1] = 0;
      return true;
    default: return false;
  }
}

bool is_mask(int f) { return f >= LES_AMASK1 && f <= LES_DMASK1; }
bool is_1d_host(int f) { return f >= LES_CN2L && f <= LES_WIND_PROFILE; }

// Every per-lane denominator of the stencils along i and j is a positive power of two with a
// normal reciprocal: dx1, dy1, dxs, dys and the neighbour sums dx1(i)+dx1(i+1), dy1(j)+dy1(j+1)
// (as FP32 sums).  Then the kernels divide by multiplying with the exact reciprocal.
bool pow2_normal_host(float b) {
  uint32_t bb;
  std::memcpy(&bb, &b, 4);
  return (bb & 0x807fffffu) == 0u && (bb & 0x7f800000u) - 0x00800000u <= 0x7e000000u;
}

bool grid_p2xy(const les_handle* h) {
  const auto& dx1 = h->hgrid[LES_DX1];
  const auto& dy1 = h->hgrid[LES_DY1];
  const auto& dxs = h->hgrid[LES_DXS];
  const auto& dys = h->hgrid[LES_DYS];
  if (dx1.empty() || dy1.empty() || dxs.empty() || dys.empty()) return false;
  for (const auto* a : {&dx1, &dy1, &dxs, &dys})
    for (float x : *a)
      if (!pow2_normal_host(x)) return false;
  for (size_t q = 0; q + 1 < dx1.size(); ++q)
    if (!pow2_normal_host(dx1[q] + dx1[q + 1])) return false;
  for (size_t q = 0; q + 1 < dy1.size(); ++q)
    if (!pow2_normal_host(dy1[q] + dy1[q + 1])) return false;
  return true;
}

Grid1D grid_of(les_handle* h) {
  Grid1D g;
  g.p2xy = h->p2xy ? 1 : 0;
  g.dx1 = h->fld[LES_DX1];
  g.dxs = h->fld[LES_DXS];
  g.dy1 = h->fld[LES_DY1];
  g.dys = h->fld[LES_DYS];
  g.dzn = h->fld[LES_DZN];
  g.dzs = h->fld[LES_DZS];
  g.cn2l = h->fld[LES_CN2L];
  g.cn2s = h->fld[LES_CN2S];
  g.cn3l = h->fld[LES_CN3L];
  g.cn3s = h->fld[LES_CN3S];
  g.cn4l = h->fld[LES_CN4L];
  g.cn4s = h->fld[LES_CN4S];
  g.delx1 = h->fld[LES_DELX1];
  const int64_t nz = h->cnt[LES_DZN];
  g.rzn = h->rz;
  g.rzs = h->rz + nz;
  g.rzp = h->rz + 2 * nz;
  return g;
}

Fields fields_of(les_handle* h) {
  Fields F;
  F.cd = h->drag;
  F.p = h->fld[LES_P];
  F.u = h->fld[LES_U];
  F.v = h->fld[LES_V];
  F.w = h->fld[LES_W];
  F.usum = h->fld[LES_USUM];
  F.vsum = h->fld[LES_VSUM];
  F.wsum = h->fld[LES_WSUM];
  F.f = h->fld[LES_F];
  F.g = h->fld[LES_G];
  F.h = h->fld[LES_H];
  F.fold = h->fld[LES_FOLD];
  F.gold = h->fld[LES_GOLD];
  F.hold = h->fld[LES_HOLD];
  F.rhs = h->fld[LES_RHS];
  F.sm = h->fld[LES_SM];
  F.diu1 = h->fld[LES_DIU1];
  F.diu2 = h->fld[LES_DIU2];
  F.diu3 = h->fld[LES_DIU3];
  F.diu4 = h->fld[LES_DIU4];
  F.diu5 = h->fld[LES_DIU5];
  F.diu6 = h->fld[LES_DIU6];
  F.diu7 = h->fld[LES_DIU7];
  F.diu8 = h->fld[LES_DIU8];
  F.diu9 = h->fld[LES_DIU9];
  F.mask = h->mask;
  F.cn1 = h->fld[LES_CN1];
  F.wind = h->fld[LES_WIND_PROFILE];
  return F;
}

// interior mask bits from the four float masks (host arrays in their Fortran bounds)
void build_mask_bits(const les_handle* h, const std::vector<float>* hm, std::vector<uint8_t>& out) {
  const int ip = h->d.ip, jp = h->d.jp, kp = h->d.kp;
  out.assign((size_t)ip * jp * kp, 0);
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i) {
        long ia = i + (long)(ip + 2) * (j + (long)(jp + 2) * k);        // a,d: (0:ip+1,0:jp+1,0:kp+1)
        long ib = (i + 1) + (long)(ip + 3) * (j + (long)(jp + 2) * k);  // b: (-1:ip+1,0:jp+1,0:kp+1)
        long ic = i + (long)(ip + 2) * ((j + 1) + (long)(jp + 3) * k);  // c: (0:ip+1,-1:jp+1,0:kp+1)
        uint8_t m = 0;
        if (hm[0].size() && hm[0][ia] != 0.0f) m |= 8;
        if (hm[1].size() && hm[1][ib] != 0.0f) m |= 1;
        if (hm[2].size() && hm[2][ic] != 0.0f) m |= 2;
        if (hm[3].size() && hm[3][ia] != 0.0f) m |= 4;
        out[(size_t)(i - 1) + (size_t)ip * ((j - 1) + (size_t)jp * (k - 1))] = m;
      }
}

// The vertical divisors of velfg, visc and rhs with their FP64 reciprocals (Grid1D::rzn, rzs,
// rzp; zdiv): (b, RN64(1/b)) for b = dzn(k), dzs(k) and the FP32 sum dzn(k) + dzn(k+1), from
// the host copies of the grid (a zero spacing gives an infinite reciprocal, which zdiv's fixup
// turns into the division's result).
int upload_zrecip(les_handle* h) {
  const auto& dzn = h->hgrid[LES_DZN];
  const auto& dzs = h->hgrid[LES_DZS];
  const size_t n = (size_t)h->cnt[LES_DZN];
  std::vector<double2> t(3 * n);
  auto pair = [](float b) { return make_double2((double)b, 1.0 / (double)b); };
  for (size_t q = 0; q < n; ++q) {
    t[q] = pair(dzn[q]);
    t[n + q] = pair(dzs[q]);
    t[2 * n + q] = pair(q + 1 < n ? dzn[q] + dzn[q + 1] : dzn[q]);  // last entry: never read
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(h->rz, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
  return LES_OK;
}

int upload_scalar_area(les_handle* h) {
  // press.f95:67-74/125-133: sum of dx1(i)*dy1(j)*dzn(k) in the reference's sequential FP32
  // order; a grid constant, evaluated once on the host.
  const auto& dx1 = h->hgrid[LES_DX1];
  const auto& dy1 = h->hgrid[LES_DY1];
  const auto& dzn = h->hgrid[LES_DZN];
  if (dx1.empty() || dy1.empty() || dzn.empty()) return LES_OK;
  float area = 0.0f;
  for (int k = 1; k <= h->d.kp; ++k)
    for (int j = 1; j <= h->d.jp; ++j)
      for (int i = 1; i <= h->d.ip; ++i) area = area + dx1[i + 1] * dy1[j] * dzn[k + 1];
  // the same volumes summed in FP64 for the mixed-precision means
  double area64 = 0.0;
  for (int k = 1; k <= h->d.kp; ++k)
    for (int j = 1; j <= h->d.jp; ++j)
      for (int i = 1; i <= h->d.ip; ++i) area64 += (double)(dx1[i + 1] * dy1[j] * dzn[k + 1]);
  HIPCHK(h, hipMemcpyAsync(&h->sc->area, &area, sizeof(float), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(&h->sc->pco, &area, sizeof(float), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipMemcpyAsync(&h->sc->area64, &area64, sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return LES_OK;
}

// The 1-D factors of cn1 (init.f95:98-103) for the split SOR, and whether the stored cn1 is
// exactly 1/((ax(i)+by(j))+cz(k)) (it is whenever it came from init.f95's formula).
int refresh_split_factors(les_handle* h) {
  const int ip = h->d.ip, jp = h->d.jp, kp = h->d.kp;
  const auto& dxs = h->hgrid[LES_DXS];
  const auto& dys = h->hgrid[LES_DYS];
  const auto& dzs = h->hgrid[LES_DZS];
  std::vector<float> ax(ip + 2, 0.0f), by(jp + 2, 0.0f), cz(kp + 2, 0.0f);
  for (int i = 1; i <= ip; ++i) ax[i] = 2.0f / (dxs[i - 1] * dxs[i]);
  for (int j = 1; j <= jp; ++j) by[j] = 2.0f / (dys[j - 1] * dys[j]);
  for (int k = 1; k <= kp; ++k) cz[k] = 2.0f / (dzs[k] * dzs[k + 1]);  // dzs(k-1)*dzs(k)
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(const_cast<float*>(h->sp.ax), ax.data(), ax.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(const_cast<float*>(h->sp.by), by.data(), by.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(const_cast<float*>(h->sp.cz), cz.data(), cz.size() * sizeof(float), hipMemcpyHostToDevice));
  bool sep = true;
  for (int k = 1; k <= kp && sep; ++k)
    for (int j = 1; j <= jp && sep; ++j)
      for (int i = 1; i <= ip; ++i) {
        const float c = 1.0f / (ax[i] + by[j] + cz[k]);
        const float s = h->hcn1[(size_t)(i - 1) + (size_t)ip * ((j - 1) + (size_t)jp * (k - 1))];
        uint32_t a, b;
        std::memcpy(&a, &c, 4);
        std::memcpy(&b, &s, 4);
        if (a != b) {
          sep = false;
          break;
        }
      }
  h->cn1_separable = sep;
  // uniform horizontal grid: every ax, by, cn2l, cn2s, cn3l, cn3s equal (bitwise); the sweep
  // then takes them as scalars and cn1 from a per-k table, the same values bit for bit
  auto all_equal = [](const float* a, int n) {
    for (int q = 1; q < n; ++q)
      if (std::memcmp(&a[q], &a[0], 4) != 0) return false;
    return true;
  };
  const bool uni = all_equal(ax.data() + 1, ip) && all_equal(by.data() + 1, jp) &&
                   all_equal(h->hgrid[LES_CN2L].data(), ip) &&
                   all_equal(h->hgrid[LES_CN2S].data(), ip) &&
                   all_equal(h->hgrid[LES_CN3L].data(), jp) &&
                   all_equal(h->hgrid[LES_CN3S].data(), jp);
  std::vector<float> cn1k(kp + 2, 0.0f);
  for (int k = 1; k <= kp; ++k) cn1k[k] = 1.0f / (ax[1] + by[1] + cz[k]);
  HIPCHK(h, hipMemcpy(const_cast<float*>(h->sp.cn1k), cn1k.data(), cn1k.size() * sizeof(float),
                      hipMemcpyHostToDevice));
  h->sp.uniform = uni ? 1 : 0;
  h->graph_valid = false;
  return LES_OK;
}

int check_launch(les_handle* h) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(h, LES_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return LES_OK;
}

// F.rhs's interior from the resident solve's adjusted rhs (ra), when it is behind
void sync_rhs(les_handle* h) {
  if (!h->rhs_stale) return;
  SplitSOR s = h->sp;
  s.r[0] = h->ra[0];
  s.r[1] = h->ra[1];
  launch_split_gather_rhs(h->st, h->d, fields_of(h), s);
  h->rhs_stale = false;
}

// p or rhs handed out by address: the caller may read or write them at any time, so the
// resident solve (which keeps them in the split copy between steps) stops for good
void expose_field(les_handle* h, int field) {
  if (field != LES_P && field != LES_RHS) return;
  sync_rhs(h);
  h->resident_ok = false;
  h->sp_p_current = false;
  h->graph_valid = false;
}

}  // namespace

// ------------------------------------------------------------------------------------------
extern "C" int64_t les_field_bounds(int field, int ip, int jp, int kp, int lo[3], int hi[3]) {
  BoundSpec s;
  if (!spec_of(field, s)) return -1;
  const int base[3] = {ip, jp, kp};
  int64_t n = 1;
  for (int q = 0; q < 3; ++q) {
    if (q < s.ndim) {
      lo[q] = s.lo[q];
      hi[q] = base[s.hi_base[q]] + s.hi_off[q];
      n *= (int64_t)(hi[q] - lo[q] + 1);
    } else {
      lo[q] = 0;
      hi[q] = 0;
    }
  }
  return n;
}

extern "C" int les_default_config(les_config* c, int ip, int jp, int kp) {
  if (!c) return LES_E_INVALID;
  std::memset(c, 0, sizeof(*c));
  c->ip = ip;
  c->jp = jp;
  c->kp = kp;
  c->dt = 0.2f;           // set.f95:76
  c->ro = 1.1763f;        // set.f95:78
  c->vn = 15.83f * powf(10.0f, -6.0f);  // set.f95:79
  c->alpha = -10.0f;      // set.f95:81
  c->beta = -1.0f;        // set.f95:82
  c->cs0 = 0.1f;          // les.f95:31
  c->omega = 1.0f;        // press.f95:46
  c->pjuge = 0.0001f;     // press.f95:44
  c->nmaxp = 50;          // press.f95:45
  c->precision = LES_PREC_FP32;
  c->refine_tol = 1e-7;
  c->refine_max = 8;
  c->device = 0;
  c->sub_i = 1;
  c->sub_j = 1;
  c->refine_solver = 0;
  return LES_OK;
}

extern "C" int les_create(const les_config* cfg, les_handle** out) {
  if (!cfg || !out) return LES_E_INVALID;
  *out = nullptr;
  if (cfg->ip < 2 || cfg->jp < 2 || cfg->kp < 2 || cfg->ip > 16384 || cfg->jp > 16384 ||
      cfg->kp > 4096 || cfg->nmaxp < 0)
    return LES_E_INVALID;
  // 32-bit device indexing and buffer byte offsets: the largest array,
  // diu1(-1:ip+2,0:jp+2,0:kp+2), below 2^30 elements
  if ((int64_t)(cfg->ip + 4) * (cfg->jp + 3) * (cfg->kp + 3) >= (int64_t(1) << 30))
    return LES_E_INVALID;
  if (cfg->sub_i != 1 || cfg->sub_j != 1) return LES_E_INVALID;  // one domain per handle
  if (cfg->precision != LES_PREC_FP32 && cfg->precision != LES_PREC_MIXED) return LES_E_INVALID;
  if (cfg->sum_order != LES_SUM_EXACT && cfg->sum_order != LES_SUM_REFERENCE) return LES_E_INVALID;
  if (cfg->sum_order == LES_SUM_REFERENCE && cfg->precision != LES_PREC_FP32) return LES_E_INVALID;
  if (cfg->precision == LES_PREC_MIXED && (!(cfg->refine_tol > 0.0) || cfg->refine_max < 0 ||
                                            cfg->refine_omega < 0.0f || cfg->refine_omega >= 2.0f ||
                                            cfg->refine_solver < 0 || cfg->refine_solver > 1))
    return LES_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LES_E_NODEV;
  if (cfg->device < 0 || cfg->device >= ndev) return LES_E_INVALID;
  les_handle* h = new les_handle();
  h->cfg = *cfg;
  h->d = Dims{cfg->ip, cfg->jp, cfg->kp};
  h->ph = Physics{cfg->dt, cfg->ro, cfg->vn, cfg->alpha, cfg->beta, cfg->cs0, cfg->omega, cfg->pjuge};
  auto bail = [&](int code) {
    les_destroy(h);
    return code;
  };
  if (hipSetDevice(cfg->device) != hipSuccess) return bail(LES_E_HIP);
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return bail(LES_E_HIP);
  for (int f = 0; f < LES_NFIELDS; ++f) {
    int lo[3], hi[3];
    int64_t n = les_field_bounds(f, cfg->ip, cfg->jp, cfg->kp, lo, hi);
    h->cnt[f] = n;
    if (hipMalloc(&h->fld[f], (size_t)n * sizeof(float)) != hipSuccess) return bail(LES_E_NOMEM);
    if (hipMemsetAsync(h->fld[f], 0, (size_t)n * sizeof(float), h->st) != hipSuccess)
      return bail(LES_E_HIP);
    if (is_1d_host(f) || f == LES_DX1 || f == LES_DXS || f == LES_DY1 || f == LES_DYS ||
        f == LES_DZN || f == LES_DZS || f == LES_Z2)
      h->hgrid[f].assign((size_t)n, 0.0f);
  }
  size_t nmask = (size_t)cfg->ip * cfg->jp * cfg->kp;
  if (hipMalloc(&h->mask, nmask) != hipSuccess) return bail(LES_E_NOMEM);
  if (hipMemsetAsync(h->mask, 0, nmask, h->st) != hipSuccess) return bail(LES_E_HIP);
  if (hipMalloc(&h->sc, sizeof(DevScalars)) != hipSuccess) return bail(LES_E_NOMEM);
  if (hipMemsetAsync(h->sc, 0, sizeof(DevScalars), h->st) != hipSuccess) return bail(LES_E_HIP);
  h->rb.max_blocks = std::max(std::max(reduction_blocks(h->d), 2 * sor_resid_blocks(h->d)), 4096);
  if (hipMalloc(&h->rb.partials, (size_t)h->rb.max_blocks * NLIMB * sizeof(int64_t)) != hipSuccess)
    return bail(LES_E_NOMEM);
  if (hipMalloc(&h->rb.slots, (size_t)RED_SLOTS * NLIMB * sizeof(int64_t)) != hipSuccess)
    return bail(LES_E_NOMEM);
  if (hipMemsetAsync(h->rb.slots, 0, (size_t)RED_SLOTS * NLIMB * sizeof(int64_t), h->st) != hipSuccess)
    return bail(LES_E_HIP);
  if (hipMalloc(&h->scratch_f, 256) != hipSuccess) return bail(LES_E_NOMEM);
  if (hipMalloc(&h->rz, 3 * (size_t)h->cnt[LES_DZN] * sizeof(double2)) != hipSuccess)
    return bail(LES_E_NOMEM);
  if (upload_zrecip(h) != LES_OK) return bail(LES_E_HIP);
  {  // split SOR copy: 2 colours x (p, rhs) of interior rows padded to nr floats, plus the
     // cn1 factors ax(0:ip+1), by(0:jp+1), cz(0:kp+1) and the table cn1k(0:kp+1)
    SplitSOR& sp = h->sp;
    sp.nr = ((cfg->ip + 1) / 2 + 3) / 4 * 4;
    sp.nj = cfg->jp;
    const size_t ns = (size_t)sp.nr * cfg->jp * cfg->kp;
    const size_t nf = 8 * ns + (size_t)(cfg->ip + 2) + (cfg->jp + 2) + 2 * (cfg->kp + 2);
    if (hipMalloc(&h->split_mem, nf * sizeof(float)) != hipSuccess) return bail(LES_E_NOMEM);
    if (hipMemsetAsync(h->split_mem, 0, nf * sizeof(float), h->st) != hipSuccess) return bail(LES_E_HIP);
    sp.p[0] = h->split_mem;
    sp.p[1] = sp.p[0] + ns;
    sp.r[0] = sp.p[1] + ns;
    sp.r[1] = sp.r[0] + ns;
    h->palt[0] = sp.r[1] + ns;  // float4-aligned like the four arrays before it
    h->palt[1] = h->palt[0] + ns;
    h->ra[0] = h->palt[1] + ns;
    h->ra[1] = h->ra[0] + ns;
    float* fac = h->ra[1] + ns;
    sp.ax = fac;
    sp.by = fac + (cfg->ip + 2);
    sp.cz = fac + (cfg->ip + 2) + (cfg->jp + 2);
    sp.cn1k = fac + (cfg->ip + 2) + (cfg->jp + 2) + (cfg->kp + 2);
    sp.uniform = 0;
    h->split_ok = sor_split_divisors(h->d, sp);
    h->hcn1.assign((size_t)cfg->ip * cfg->jp * cfg->kp, 0.0f);
    if (cfg->precision == LES_PREC_MIXED) {
      Mixed& mx = h->mx;
      mx.nblk = mixed_blocks(h->d);
      const size_t np = (size_t)mixed_part_len(h->d, sp.nr);
      const size_t nd = 4 * ns + np + 8;
      if (hipMalloc(&h->mx_mem, nd * sizeof(double)) != hipSuccess) return bail(LES_E_NOMEM);
      if (hipMemsetAsync(h->mx_mem, 0, nd * sizeof(double), h->st) != hipSuccess)
        return bail(LES_E_HIP);
      mx.p64[0] = h->mx_mem;
      mx.p64[1] = mx.p64[0] + ns;
      mx.r64[0] = mx.p64[1] + ns;
      mx.r64[1] = mx.r64[0] + ns;
      mx.part = mx.r64[1] + ns;
      mx.cg = mx.part + np;
      mx.cn1 = h->fld[LES_CN1];
      // FP32 vectors of the conjugate-gradient correction solve
      if (hipMalloc(&h->cg_mem, 8 * ns * sizeof(float)) != hipSuccess) return bail(LES_E_NOMEM);
      if (hipMemsetAsync(h->cg_mem, 0, 8 * ns * sizeof(float), h->st) != hipSuccess)
        return bail(LES_E_HIP);
      for (int c = 0; c < 2; ++c) {
        mx.z[c] = h->cg_mem + c * ns;
        mx.dir[c] = h->cg_mem + (2 + c) * ns;
        mx.q[c] = h->cg_mem + (4 + c) * ns;
        mx.dir2[c] = h->cg_mem + (6 + c) * ns;
      }
    }
  }
  h->resident_ok = cfg->precision == LES_PREC_FP32 && cfg->sum_order == LES_SUM_EXACT;
  if (cfg->sum_order == LES_SUM_REFERENCE) {
    const int nrt = (cfg->ip + 1) / 2;
    const size_t nt = (size_t)nrt * cfg->jp * cfg->kp;
    if (hipMalloc(&h->seq_mem, 2 * nt * sizeof(float)) != hipSuccess) return bail(LES_E_NOMEM);
    h->seq_terms = SeqTerms{{h->seq_mem, h->seq_mem + nt}, nrt};
  }
  if (hipStreamSynchronize(h->st) != hipSuccess) return bail(LES_E_HIP);
  *out = h;
  return LES_OK;
}

extern "C" int les_destroy(les_handle* h) {
  if (!h) return LES_OK;
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  if (h->st) hipStreamSynchronize(h->st);
  for (int f = 0; f < LES_NFIELDS; ++f)
    if (h->fld[f]) hipFree(h->fld[f]);
  if (h->mask) hipFree(h->mask);
  if (h->rz) hipFree(h->rz);
  if (h->sc) hipFree(h->sc);
  if (h->rb.partials) hipFree(h->rb.partials);
  if (h->rb.slots) hipFree(h->rb.slots);
  if (h->scratch_f) hipFree(h->scratch_f);
  if (h->split_mem) hipFree(h->split_mem);
  if (h->mx_mem) hipFree(h->mx_mem);
  if (h->cg_mem) hipFree(h->cg_mem);
  if (h->ave_mem) hipFree(h->ave_mem);
  if (h->seq_mem) hipFree(h->seq_mem);
  if (h->st) hipStreamDestroy(h->st);
  if (h == g_default) g_default = nullptr;
  delete h;
  return LES_OK;
}

extern "C" const char* les_last_error(const les_handle* h) { return h ? h->err.c_str() : ""; }

extern "C" int les_get_dims(const les_handle* h, int* ip, int* jp, int* kp) {
  if (!h || !ip || !jp || !kp) return LES_E_INVALID;
  *ip = h->d.ip;
  *jp = h->d.jp;
  *kp = h->d.kp;
  return LES_OK;
}

extern "C" int les_sync(les_handle* h) {
  if (!h) return LES_E_INVALID;
  HIPCHK(h, hipStreamSynchronize(h->st));
  return LES_OK;
}

extern "C" int les_upload(les_handle* h, int field, const float* host) {
  if (!h || !host || field < 0 || field >= LES_NFIELDS) return fail(h, LES_E_INVALID, "upload: bad argument");
  const size_t n = (size_t)h->cnt[field];
  if (is_mask(field)) {
    for (size_t q = 0; q < n; ++q)
      if (host[q] != 0.0f && host[q] != 1.0f)
        return fail(h, LES_E_INVALID, "upload: mask values must be 0 or 1");
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(h->fld[field], host, n * sizeof(float), hipMemcpyHostToDevice));
  if (!h->hgrid[field].empty()) std::memcpy(h->hgrid[field].data(), host, n * sizeof(float));
  if (is_mask(field)) {
    // rebuild the interior bit mask from the four device masks
    std::vector<float> hm[4];
    for (int q = 0; q < 4; ++q) {
      hm[q].resize((size_t)h->cnt[LES_AMASK1 + q]);
      HIPCHK(h, hipMemcpy(hm[q].data(), h->fld[LES_AMASK1 + q], hm[q].size() * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
    std::vector<uint8_t> bits;
    build_mask_bits(h, hm, bits);
    HIPCHK(h, hipMemcpy(h->mask, bits.data(), bits.size(), hipMemcpyHostToDevice));
  }
  if (field == LES_DX1 || field == LES_DY1 || field == LES_DZN) {
    int r = upload_scalar_area(h);
    if (r) return r;
  }
  if (field == LES_DZN || field == LES_DZS) {
    int r = upload_zrecip(h);
    if (r) return r;
  }
  if (field == LES_DX1 || field == LES_DY1 || field == LES_DXS || field == LES_DYS) {
    h->p2xy = grid_p2xy(h);
    h->graph_valid = false;
  }
  if (field == LES_P) {
    h->p_consistent = false;
    h->sp_p_current = false;
  }
  if (field == LES_RHS) h->rhs_stale = false;  // the upload replaced every value
  if (field == LES_CN1) std::memcpy(h->hcn1.data(), host, n * sizeof(float));
  if (field == LES_CN1 || field == LES_DXS || field == LES_DYS || field == LES_DZS ||
      field == LES_CN2L || field == LES_CN2S || field == LES_CN3L || field == LES_CN3S) {
    int r = refresh_split_factors(h);
    if (r) return r;
  }
  return LES_OK;
}

// aveflow.f95:42-157 after step n (statistics from n1 to nmax), on the device
extern "C" int les_aveflow(les_handle* h, int n, int n1, int nmax) {
  if (!h) return LES_E_INVALID;
  const Dims d = h->d;
  const Fields F = fields_of(h);
  if (n1 > nmax) return fail(h, LES_E_INVALID, "aveflow: n1 > nmax");
  const size_t n3 = (size_t)d.ip * d.jp * d.kp, n2 = (size_t)d.ip * d.kp;
  if (!h->ave_mem) {
    const size_t nf = AVE_N3 * n3 + AVE_N2 * n2;
    HIPCHK(h, hipMalloc(&h->ave_mem, nf * sizeof(float)));
    HIPCHK(h, hipMemsetAsync(h->ave_mem, 0, nf * sizeof(float), h->st));
    for (int q = 0; q < AVE_N3; ++q) h->ave.a3[q] = h->ave_mem + q * n3;
    for (int q = 0; q < AVE_N2; ++q) h->ave.a2[q] = h->ave_mem + AVE_N3 * n3 + q * n2;
  }
  launch_aveflow(h->st, d, F, h->ave, n, n1, nmax);
  return check_launch(h);
}

extern "C" int les_aveflow_download(les_handle* h, int which, float* host) {
  if (!h || !host || which < 0 || which >= AVE_N3 + AVE_N2) return LES_E_INVALID;
  if (!h->ave_mem) return fail(h, LES_E_STATE, "aveflow_download: no statistics yet");
  const size_t n3 = (size_t)h->d.ip * h->d.jp * h->d.kp, n2 = (size_t)h->d.ip * h->d.kp;
  const float* src = which < AVE_N3 ? h->ave.a3[which] : h->ave.a2[which - AVE_N3];
  const size_t cnt = which < AVE_N3 ? n3 : n2;
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(host, src, cnt * sizeof(float), hipMemcpyDeviceToHost));
  return LES_OK;
}

// feedbfm.f95:42-63 on the device from a host zbm(-1:ip+1,-1:jp+1) (ipmax = ip, jpmax = jp);
// needs z2 and dzn uploaded
extern "C" int les_set_terrain(les_handle* h, const float* zbm) {
  if (!h || !zbm) return LES_E_INVALID;
  const size_t nz = (size_t)(h->d.ip + 3) * (h->d.jp + 3);
  float* dz = nullptr;
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMalloc(&dz, nz * sizeof(float)));
  hipError_t e = hipMemcpy(dz, zbm, nz * sizeof(float), hipMemcpyHostToDevice);
  for (int f = LES_AMASK1; f <= LES_DMASK1 && e == hipSuccess; ++f)
    e = hipMemsetAsync(h->fld[f], 0, (size_t)h->cnt[f] * sizeof(float), h->st);
  if (e == hipSuccess) {
    launch_feedbfm(h->st, h->d, grid_of(h), h->fld[LES_Z2], dz, h->fld[LES_AMASK1],
                   h->fld[LES_BMASK1], h->fld[LES_CMASK1], h->fld[LES_DMASK1], h->mask);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  hipFree(dz);
  if (e != hipSuccess) return fail(h, LES_E_HIP, std::string("set_terrain: ") + hipGetErrorString(e));
  h->graph_valid = false;
  return LES_OK;
}

extern "C" int les_download(les_handle* h, int field, float* host) {
  if (!h || !host || field < 0 || field >= LES_NFIELDS) return fail(h, LES_E_INVALID, "download: bad argument");
  if (field == LES_RHS) sync_rhs(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(host, h->fld[field], (size_t)h->cnt[field] * sizeof(float), hipMemcpyDeviceToHost));
  return LES_OK;
}

extern "C" int les_device_ptr(les_handle* h, int field, void** dptr, int64_t* count) {
  if (!h || field < 0 || field >= LES_NFIELDS) return LES_E_INVALID;
  expose_field(h, field);
  if (dptr) *dptr = h->fld[field];
  if (count) *count = h->cnt[field];
  return LES_OK;
}

// init.f95:82-107 and les.f95:33-37 on the host, FP32, reference operation order
extern "C" int les_init_coefficients(les_handle* h) {
  if (!h) return LES_E_INVALID;
  const int ip = h->d.ip, jp = h->d.jp, kp = h->d.kp;
  auto& dxs = h->hgrid[LES_DXS];
  auto& dys = h->hgrid[LES_DYS];
  auto& dzs = h->hgrid[LES_DZS];
  auto& dx1 = h->hgrid[LES_DX1];
  auto& dy1 = h->hgrid[LES_DY1];
  auto& dzn = h->hgrid[LES_DZN];
  std::vector<float> cn2s(ip), cn2l(ip), cn3s(jp), cn3l(jp), cn4s(kp), cn4l(kp), delx1(kp);
  const float two = 2.0f, one = 1.0f;
  for (int i = 1; i <= ip; ++i) {
    cn2s[i - 1] = two / (dxs[i - 1] * (dxs[i - 1] + dxs[i]));
    cn2l[i - 1] = two / (dxs[i] * (dxs[i - 1] + dxs[i]));
  }
  for (int j = 1; j <= jp; ++j) {
    cn3s[j - 1] = two / (dys[j - 1] * (dys[j - 1] + dys[j]));
    cn3l[j - 1] = two / (dys[j] * (dys[j - 1] + dys[j]));
  }
  std::vector<float> cn1((size_t)ip * jp * kp);
  for (int k = 1; k <= kp; ++k) {
    const float dz1 = dzs[k], dz2 = dzs[k + 1];  // dzs(k-1), dzs(k) with offset -1
    cn4s[k - 1] = two / (dz1 * (dz1 + dz2));
    cn4l[k - 1] = two / (dz2 * (dz1 + dz2));
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i) {
        float c = two / (dxs[i - 1] * dxs[i]) + two / (dys[j - 1] * dys[j]) + two / (dz1 * dz2);
        cn1[(size_t)(i - 1) + (size_t)ip * ((j - 1) + (size_t)jp * (k - 1))] = one / c;
      }
    // les.f95:36 (dx1(0)*dy1(0)*dzn(k))**(1./3.)
    delx1[k - 1] = powf(dx1[1] * dy1[0] * dzn[k + 1], 1.0f / 3.0f);
  }
  int r;
  if ((r = les_upload(h, LES_CN2S, cn2s.data()))) return r;
  if ((r = les_upload(h, LES_CN2L, cn2l.data()))) return r;
  if ((r = les_upload(h, LES_CN3S, cn3s.data()))) return r;
  if ((r = les_upload(h, LES_CN3L, cn3l.data()))) return r;
  if ((r = les_upload(h, LES_CN4S, cn4s.data()))) return r;
  if ((r = les_upload(h, LES_CN4L, cn4l.data()))) return r;
  if ((r = les_upload(h, LES_CN1, cn1.data()))) return r;
  if ((r = les_upload(h, LES_DELX1, delx1.data()))) return r;
  return upload_scalar_area(h);
}

// ------------------------------------------------------------------------------------------
// stages
// ------------------------------------------------------------------------------------------
#define STAGE_BEGIN(h)                 \
  if (!(h)) return LES_E_INVALID;      \
  const Dims d = (h)->d;               \
  const Grid1D g = grid_of(h);         \
  const Fields F = fields_of(h);       \
  (void)d; (void)g; (void)F;

namespace {
// press.f95's sums in the reference's sequential order (les_config::sum_order)
bool seq_sums(const les_handle* h) { return h->cfg.sum_order == LES_SUM_REFERENCE; }

// rhs and rhsav (press.f95:51-75), the SOR scalars re-armed
void enqueue_rhs_rhsav(les_handle* h, const Dims& d, const Grid1D& g, const Fields& F) {
  launch_rhs(h->st, d, g, F, h->ph, h->rb, h->sc, EDGES_ALL, seq_sums(h) ? 0 : 1);
  if (seq_sums(h)) launch_seq_rhsav(h->st, d, g, F.rhs, h->sc);
}

// the residual of one iteration from the sweeps' partials or terms, and the early exit
void enqueue_resid(les_handle* h, const Dims& d, const Physics& ph) {
  if (seq_sums(h))
    launch_seq_resid(h->st, d, h->seq_terms, ph, h->sc);
  else
    launch_sor_resid_finalize(h->st, h->rb, ph, h->sc);
}
}  // namespace

extern "C" int les_velnw(les_handle* h) {
  STAGE_BEGIN(h);
  launch_velnw_inflow(h->st, d, g, F, h->ph, 1, 0, h->shp_velnw);
  return check_launch(h);
}

extern "C" int les_bondv1_init_uvw(les_handle* h, int n) {
  STAGE_BEGIN(h);
  launch_velnw_inflow(h->st, d, g, F, h->ph, 0, n < 1 ? 2 : n, h->shp_velnw);
  return check_launch(h);
}

extern "C" int les_velnw__bondv1_init_uvw(les_handle* h, int n) {
  STAGE_BEGIN(h);
  launch_velnw_inflow(h->st, d, g, F, h->ph, 1, n < 1 ? 2 : n, h->shp_velnw);
  return check_launch(h);
}

extern "C" int les_bondv1_calc_uout(les_handle* h) {
  STAGE_BEGIN(h);
  launch_uout(h->st, d, F, h->sc, reinterpret_cast<float*>(h->rb.partials));
  return check_launch(h);
}

extern "C" int les_bondv1_calc_uvw(les_handle* h) {
  STAGE_BEGIN(h);
  launch_bondv1_uvw(h->st, d, g, F, h->ph, h->sc);
  return check_launch(h);
}

extern "C" int les_velfg(les_handle* h) {
  STAGE_BEGIN(h);
  launch_velfg(h->st, d, g, F, h->ph, 1, 0);
  return check_launch(h);
}

extern "C" int les_feedbf(les_handle* h) {
  STAGE_BEGIN(h);
  launch_velfg(h->st, d, g, F, h->ph, 0, 1);
  return check_launch(h);
}

extern "C" int les_les_calc_sm(les_handle* h) {
  STAGE_BEGIN(h);
  launch_calc_sm(h->st, d, g, F, h->ph, 0, h->shp_sm);
  return check_launch(h);
}

extern "C" int les_velfg__feedbf__les_calc_sm(les_handle* h) {
  STAGE_BEGIN(h);
  launch_velfg(h->st, d, g, F, h->ph, 1, 1);
  launch_calc_sm(h->st, d, g, F, h->ph, 0, h->shp_sm);
  return check_launch(h);
}

extern "C" int les_les_bound_sm(les_handle* h) {
  STAGE_BEGIN(h);
  launch_boundsm(h->st, d, F);
  return check_launch(h);
}

extern "C" int les_les_calc_visc(les_handle* h) {
  STAGE_BEGIN(h);
  launch_calc_visc_adam(h->st, d, g, F, 1, 0, h->shp_visc);
  return check_launch(h);
}

extern "C" int les_adam(les_handle* h) {
  STAGE_BEGIN(h);
  launch_calc_visc_adam(h->st, d, g, F, 0, 1, h->shp_visc);
  return check_launch(h);
}

extern "C" int les_les_calc_visc__adam(les_handle* h) {
  STAGE_BEGIN(h);
  launch_calc_visc_adam(h->st, d, g, F, 1, 1, h->shp_visc);
  return check_launch(h);
}

extern "C" int les_les(les_handle* h) {
  STAGE_BEGIN(h);
  launch_calc_sm(h->st, d, g, F, h->ph, 1, h->shp_sm);  // boundsm's halo cells included
  launch_calc_visc_adam(h->st, d, g, F, 1, 0, h->shp_visc);
  return check_launch(h);
}

extern "C" int les_bondfg(les_handle* h) {
  STAGE_BEGIN(h);
  launch_bondfg(h->st, d, F);
  return check_launch(h);
}

extern "C" int les_press_rhsav(les_handle* h) {
  STAGE_BEGIN(h);
  enqueue_rhs_rhsav(h, d, g, F);
  launch_rhs_adjust(h->st, d, F, h->sc);
  h->rhs_stale = false;  // F.rhs holds the adjusted rhs of this call
  return check_launch(h);
}

extern "C" int les_boundp1(les_handle* h) {
  STAGE_BEGIN(h);
  launch_boundp1(h->st, d, F, h->sc, 0);
  return check_launch(h);
}

extern "C" int les_boundp2(les_handle* h) {
  STAGE_BEGIN(h);
  launch_boundp2(h->st, d, F, h->sc, 0);
  return check_launch(h);
}

extern "C" int les_press_sor(les_handle* h, int nrd) {
  STAGE_BEGIN(h);
  if (nrd < 0 || nrd > 2) return fail(h, LES_E_INVALID, "press_sor: nrd must be 0, 1 or 2");
  sync_rhs(h);
  h->sp_p_current = false;
  if (nrd == 2) {
    launch_boundp2(h->st, d, F, h->sc, 0);
  } else {
    launch_sor_colour(h->st, d, g, F, h->ph, nrd, 0, h->rb, h->sc);
    launch_boundp1(h->st, d, F, h->sc, 0);
  }
  h->p_consistent = false;  // a lone colour or boundp2 leaves an intermediate halo state
  return check_launch(h);
}

namespace {
// The colour-split solve (les_sor.hip) applies when its equivalence to the reference's
// in-place sweep holds: even jp, mirror-consistent p halos, cn1 of the separable form.
bool use_split(const les_handle* h) {
  return (h->d.jp % 2 == 0) && h->p_consistent && h->cn1_separable && h->split_ok;
}

// press.f95:87-115 on the Fortran arrays: colour sweep, boundp1 per colour, boundp2 per
// iteration (exact for any input state)
int enqueue_sor_inplace(les_handle* h) {
  STAGE_BEGIN(h);
  launch_sor_reset(h->st, h->sc);
  const bool resid = h->ph.pjuge > 0.0f;
  const SeqTerms tm = seq_sums(h) ? h->seq_terms : SeqTerms{};
  for (int l = 1; l <= h->cfg.nmaxp; ++l) {
    for (int nrd = 0; nrd <= 1; ++nrd) {
      launch_sor_colour(h->st, d, g, F, h->ph, nrd, resid ? 1 : 0, h->rb, h->sc, tm);
      launch_boundp1(h->st, d, F, h->sc, 1);
    }
    launch_boundp2(h->st, d, F, h->sc, 1);
    if (resid) enqueue_resid(h, d, h->ph);
  }
  return check_launch(h);
}

// the same iterations on the colour-split copy (already scattered); early exit only when
// pjuge > 0 and allowed.  Without a residual, pairs of iterations run as fused launches
// (k_sor_iter) from sp.p to the second iterate and back, so the result ends in sp.p.
// reset = false when the SOR scalars were just re-armed (k_rhsav_finalize does it).
int enqueue_sor_split_loop(les_handle* h, bool allow_exit = true, const Physics* phys = nullptr,
                           bool reset = true) {
  STAGE_BEGIN(h);
  const Physics ph = phys ? *phys : h->ph;
  if (reset) launch_sor_reset(h->st, h->sc);
  const bool resid = allow_exit && h->ph.pjuge > 0.0f;
  int l0 = 1;
  if (!resid && h->iter_shape.tj > 0 && sor_iter_supported(d, h->sp)) {
    const SorPair fwd{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}};
    const SorPair back{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}};
    for (; l0 + 1 <= h->cfg.nmaxp; l0 += 2) {
      launch_sor_iter(h->st, d, g, h->sp, ph, fwd, h->sc, h->iter_shape);
      launch_sor_iter(h->st, d, g, h->sp, ph, back, h->sc, h->iter_shape);
    }
  }
  const SeqTerms tm = seq_sums(h) ? h->seq_terms : SeqTerms{};
  for (int l = l0; l <= h->cfg.nmaxp; ++l) {
    for (int nrd = 0; nrd <= 1; ++nrd)
      launch_sor_split(h->st, d, g, h->sp, ph, nrd, resid ? 1 : 0, h->rb, h->sc, tm);
    if (resid) enqueue_resid(h, d, ph);
  }
  return check_launch(h);
}

int enqueue_sor_solve(les_handle* h) {
  STAGE_BEGIN(h);
  int r;
  if (use_split(h)) {
    launch_split_scatter(h->st, d, F, h->sp, h->sc, 0);
    if ((r = enqueue_sor_split_loop(h))) return r;
    launch_split_gather_adj(h->st, d, F, h->sp, h->sc, 0);
  } else if ((r = enqueue_sor_inplace(h))) {
    return r;
  }
  if (h->cfg.nmaxp > 0) h->p_consistent = true;
  return check_launch(h);
}

}  // namespace

namespace les {
// press.f95:87-115 (nmaxp iterations, no early exit) on p and rhs arrays of the handle's
// Fortran layouts held elsewhere, enqueued on stream st: the colour-split solve of the handle
// (its split copy, cn1 factors and tuned iteration), which leaves p with the boundary images
// boundp1/boundp2 leave.  Same conditions as the handle's own split solve (even jp, cn1 of the
// init.f95 form) and, up to the caller, mirror-consistent halos of p on entry.
int sor_solve_external(les_handle* h, hipStream_t st, float* p, float* rhs) {
  if (!h || !p || !rhs) return LES_E_INVALID;
  h->sp_p_current = false;  // the split copy takes this solve's arrays
  if (!(h->d.jp % 2 == 0 && h->cn1_separable && h->split_ok))
    return fail(h, LES_E_STATE, "split solve does not apply to this grid");
  const hipStream_t own = h->st;
  h->st = st;
  Fields F = fields_of(h);
  F.p = p;
  F.rhs = rhs;
  launch_split_scatter(st, h->d, F, h->sp, h->sc, 0);
  int r = enqueue_sor_split_loop(h, false);
  if (!r) launch_split_gather_adj(st, h->d, F, h->sp, h->sc, 0);
  h->st = own;
  return r ? r : check_launch(h);
}

// The transpose of sor_solve_external on a uniform horizontal grid: given the adjoints pb of p
// (after the solve) and rb of rhs, pb becomes the adjoint of p before it and rb gains the
// rhs part.  The split copy of the handle holds the split adjoints meanwhile.
int sor_adjoint_external(les_handle* h, hipStream_t st, float* pb, float* rb) {
  if (!h || !pb || !rb) return LES_E_INVALID;
  h->sp_p_current = false;  // the split copy holds the adjoints meanwhile
  if (!(h->d.jp % 2 == 0 && h->cn1_separable && h->split_ok && h->sp.uniform))
    return fail(h, LES_E_STATE, "split solve transpose needs a uniform horizontal grid");
  const Dims d = h->d;
  const Grid1D g = grid_of(h);
  const float om = h->ph.omega;
  const int n = h->cfg.nmaxp;
  launch_split_gather_t(st, d, pb, h->sp);
  if (n > 0) {
    // the iterations swept colour 1 then colour 0: their transposes run in reverse, colour 0
    // first; the other-colour part of each transposed sweep joins the own part of the next
    launch_split_sweep_t(st, d, g, h->sp, om, 1, 0);
    for (int l = n; l >= 1; --l) {
      launch_split_sweep_t(st, d, g, h->sp, om, 0, 1);
      launch_split_sweep_t(st, d, g, h->sp, om, 1, l > 1 ? 1 : 2);
    }
  }
  launch_split_scatter_t(st, d, pb, rb, h->sp);
  return check_launch(h);
}
}  // namespace les

namespace {
// press.f95:49-146 with FP64 means and outer iterative refinement (les_mixed.hip).  Runs
// eagerly: the host reads the residual after each pass.
int enqueue_press_mixed(les_handle* h) {
  STAGE_BEGIN(h);
  if (!(h->d.jp % 2 == 0 && h->cn1_separable && h->split_ok))
    return fail(h, LES_E_STATE, "mixed precision needs an even jp and cn1 of the init.f95 form");
  int r;
  if (!h->p_consistent) {  // the refinement works on mirror-consistent boundary images
    launch_boundp1(h->st, d, F, h->sc, 0);
    launch_boundp2(h->st, d, F, h->sc, 0);
  }
  launch_rhs(h->st, d, g, F, h->ph, h->rb, h->sc, EDGES_ALL, 0);  // partials for the FP64 sum
  launch_mx_rhsav(h->st, d, h->rb, h->sc);
  launch_mx_scatter(h->st, d, g, F, h->sp, h->mx, h->sc);
  if ((r = enqueue_sor_split_loop(h, false))) return r;
  launch_mx_promote(h->st, d, h->sp, h->mx);
  Physics ph_corr = h->ph;  // the correction solves may relax with their own omega
  if (h->cfg.refine_omega > 0.0f) ph_corr.omega = h->cfg.refine_omega;
  int passes = 0;
  for (;;) {
    launch_mx_resid(h->st, d, g, h->sp, h->mx, passes == 0 ? 1 : 0, h->cfg.refine_tol, h->sc);
    int done = 0;
    double resid = 0.0;
    HIPCHK(h, hipMemcpyAsync(&done, &h->sc->refine_done, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipMemcpyAsync(&resid, &h->sc->resid_l2, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    if (done || passes >= h->cfg.refine_max) break;
    if (h->cfg.refine_solver == 1) {
      // CG on the correction until its preconditioned residual norm has fallen by what this pass
      // needs to reach a third of the tolerance (at most 1e-6: FP32 vectors), up to 20 x nmaxp
      // iterations
      const double need = resid > 0.0 ? 0.3 * h->cfg.refine_tol / resid : 1.0;
      const double stop = std::max(1e-12, std::min(1e-2, need * need));
      // preconditioner relaxation: refine_omega, else 1 (symmetric Gauss-Seidel; at 600x600x90
      // 1420 iterations per step against 1600 for 1.5 and 2330 for 1.8)
      const float om = h->cfg.refine_omega > 0.0f ? h->cfg.refine_omega : 1.0f;
      launch_mx_cg(h->st, d, g, h->sp, h->mx, 20 * h->cfg.nmaxp, om, stop);
    }
    if (h->cfg.refine_solver != 1 && (r = enqueue_sor_split_loop(h, false, &ph_corr))) return r;
    launch_mx_accumulate(h->st, d, h->sp, h->mx);
    ++passes;
  }
  HIPCHK(h, hipMemcpyAsync(&h->sc->refine_passes, &passes, sizeof(int), hipMemcpyHostToDevice, h->st));
  launch_mx_pav_gather(h->st, d, g, F, h->sp, h->mx, h->sc);
  HIPCHK(h, hipStreamSynchronize(h->st));  // 'passes' lives on this stack frame
  h->p_consistent = true;
  return check_launch(h);
}

// The resident split solve applies: fused iterations in pairs, no residual, FP32 with exact
// sums, p and rhs not handed out by address.
bool resident_applies(const les_handle* h) {
  return h->resident_ok && use_split(h) && !(h->ph.pjuge > 0.0f) && h->iter_shape.tj > 0 &&
         sor_iter_supported(h->d, h->sp) && h->cfg.nmaxp >= 2 && h->cfg.nmaxp % 2 == 0;
}

// press.f95:49-146 on the resident split copy.  The split copy of p carries over from the last
// solve (its iterate; F.p = iterate - pav_carry), or is scattered from F.p; k_rhs stores the
// raw rhs in the split layout; the first fused iteration subtracts rhsav from every rhs value
// (and pav_carry from every p value) it reads, as the scatter and the gather would, and keeps
// the adjusted rhs in ra for the other iterations; pav and the gather close the solve without
// touching the split iterate.  F.rhs is filled from ra on demand (sync_rhs).  Every value is
// the one the scatter / gather path computes.
int enqueue_press_resident(les_handle* h) {
  STAGE_BEGIN(h);
  const bool carry = h->sp_p_current;
  if (!carry) launch_split_scatter(h->st, d, F, h->sp, h->sc, 0, 0);  // p only
  launch_rhs(h->st, d, g, F, h->ph, h->rb, h->sc, EDGES_ALL, 1, &h->sp);  // also re-arms SOR
  SplitSOR sra = h->sp;
  sra.r[0] = h->ra[0];
  sra.r[1] = h->ra[1];
  const SorPair first{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}, {h->ra[0], h->ra[1]},
                      carry ? 1 : 0};
  const SorPair fwd{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}, {nullptr, nullptr}, 0};
  const SorPair back{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}, {nullptr, nullptr}, 0};
  launch_sor_iter(h->st, d, g, h->sp, h->ph, first, h->sc, h->iter_shape, 1);
  launch_sor_iter(h->st, d, g, sra, h->ph, back, h->sc, h->iter_shape);
  for (int l = 3; l + 1 <= h->cfg.nmaxp; l += 2) {
    launch_sor_iter(h->st, d, g, sra, h->ph, fwd, h->sc, h->iter_shape);
    launch_sor_iter(h->st, d, g, sra, h->ph, back, h->sc, h->iter_shape);
  }
  launch_pav_split(h->st, d, g, h->sp, h->rb, h->sc);
  launch_split_gather_adj(h->st, d, F, h->sp, h->sc, 1, 1);  // p - pav, halos; pav_carry
  h->sp_p_current = true;
  h->rhs_stale = true;
  h->p_consistent = true;
  return check_launch(h);
}

// press.f95:49-146
int enqueue_press(les_handle* h) {
  STAGE_BEGIN(h);
  if (h->cfg.precision == LES_PREC_MIXED) return enqueue_press_mixed(h);
  if (resident_applies(h)) return enqueue_press_resident(h);
  h->sp_p_current = false;
  h->rhs_stale = false;  // F.rhs is written below
  int r;
  enqueue_rhs_rhsav(h, d, g, F);
  if (use_split(h)) {
    launch_split_scatter(h->st, d, F, h->sp, h->sc, 1);  // rhs -= rhsav on the way
    if ((r = enqueue_sor_split_loop(h, true, nullptr, false))) return r;
    if (seq_sums(h))
      launch_seq_pav(h->st, d, g, nullptr, &h->sp, h->sc);
    else
      launch_pav_split(h->st, d, g, h->sp, h->rb, h->sc);
    launch_split_gather_adj(h->st, d, F, h->sp, h->sc, 1);  // p - pav, halos included
  } else {
    launch_rhs_adjust(h->st, d, F, h->sc);
    if ((r = enqueue_sor_inplace(h))) return r;
    if (seq_sums(h))
      launch_seq_pav(h->st, d, g, F.p, nullptr, h->sc);
    else
      launch_pav(h->st, d, g, F, h->rb, h->sc);
    launch_adj(h->st, d, F, h->sc, 1);
  }
  h->p_consistent = true;
  return check_launch(h);
}

int enqueue_step(les_handle* h, int n) {
  STAGE_BEGIN(h);
  launch_velnw_inflow(h->st, d, g, F, h->ph, 1, n < 1 ? 2 : n, h->shp_velnw);
  launch_uout(h->st, d, F, h->sc, reinterpret_cast<float*>(h->rb.partials));
  launch_bondv1_uvw(h->st, d, g, F, h->ph, h->sc);
  launch_velfg(h->st, d, g, F, h->ph, 1, 1);
  launch_calc_sm(h->st, d, g, F, h->ph, 1, h->shp_sm);  // boundsm's halo cells included
  launch_calc_visc_adam(h->st, d, g, F, 1, 1, h->shp_visc);
  return enqueue_press(h);
}
}  // namespace

namespace les {
float* field_ptr(les_handle* h, int field) {
  if (!h || field < 0 || field >= LES_NFIELDS) return nullptr;
  expose_field(h, field);
  return h->fld[field];
}
void set_drag(les_handle* h, const float* cd) {
  h->drag = cd;
  h->graph_valid = false;
}
void set_p_consistent(les_handle* h, bool consistent) { h->p_consistent = consistent; }
int step_external(les_handle* h, hipStream_t st, int n) {
  if (!h) return LES_E_INVALID;
  const hipStream_t own = h->st;
  h->st = st;
  const int r = enqueue_step(h, n);
  h->st = own;
  return r;
}
int stages_external(les_handle* h, hipStream_t st, int n, float* u_pre, float* v_pre, float* w_pre,
                    float* uout_dst) {
  if (!h || !u_pre || !v_pre || !w_pre || !uout_dst) return LES_E_INVALID;
  const hipStream_t own = h->st;
  h->st = st;
  const Dims d = h->d;
  const Grid1D g = grid_of(h);
  const Fields F = fields_of(h);
  launch_velnw_inflow(st, d, g, F, h->ph, 1, n < 1 ? 2 : n, h->shp_velnw);
  hipError_t e = hipMemcpyAsync(u_pre, F.u, (size_t)h->cnt[LES_U] * sizeof(float),
                                hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(v_pre, F.v, (size_t)h->cnt[LES_V] * sizeof(float), hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(w_pre, F.w, (size_t)h->cnt[LES_W] * sizeof(float), hipMemcpyDeviceToDevice, st);
  launch_uout(st, d, F, h->sc, reinterpret_cast<float*>(h->rb.partials));
  if (e == hipSuccess)
    e = hipMemcpyAsync(uout_dst, &h->sc->uout, sizeof(float), hipMemcpyDeviceToDevice, st);
  launch_bondv1_uvw(st, d, g, F, h->ph, h->sc);
  launch_velfg(st, d, g, F, h->ph, 1, 1);
  launch_calc_sm(st, d, g, F, h->ph, 1, h->shp_sm);
  launch_calc_visc_adam(st, d, g, F, 1, 1, h->shp_visc);
  h->st = own;
  if (e != hipSuccess) return fail(h, LES_E_HIP, std::string("stages: ") + hipGetErrorString(e));
  return check_launch(h);
}
}  // namespace les

extern "C" int les_press_sor_solve(les_handle* h) {
  if (!h) return LES_E_INVALID;
  sync_rhs(h);
  h->sp_p_current = false;
  return enqueue_sor_solve(h);
}

extern "C" int les_press_pav(les_handle* h) {
  STAGE_BEGIN(h);
  if (seq_sums(h))
    launch_seq_pav(h->st, d, g, F.p, nullptr, h->sc);
  else
    launch_pav(h->st, d, g, F, h->rb, h->sc);
  return check_launch(h);
}

extern "C" int les_press_adj(les_handle* h) {
  STAGE_BEGIN(h);
  h->sp_p_current = false;
  launch_adj(h->st, d, F, h->sc, 0);
  h->p_consistent = false;  // interior shifted, halos not yet
  return check_launch(h);
}

extern "C" int les_press_boundp(les_handle* h) {
  STAGE_BEGIN(h);
  launch_boundp1(h->st, d, F, h->sc, 0);
  launch_boundp2(h->st, d, F, h->sc, 0);
  h->p_consistent = true;
  return check_launch(h);
}

extern "C" int les_press(les_handle* h) {
  if (!h) return LES_E_INVALID;
  return enqueue_press(h);
}

extern "C" int les_step(les_handle* h, int n) {
  if (!h) return LES_E_INVALID;
  return enqueue_step(h, n);
}

namespace {
// Steps that cannot be replayed from the graph (n <= 1, a state with arbitrary p halos, mixed
// precision) run eagerly; returns the first step number left for the graph.
int run_eager_prefix(les_handle* h, int n, int& nsteps, int& r) {
  r = LES_OK;
  // the n == 1 step differs (interior fill)
  while (nsteps > 0 && n <= 1) {
    if ((r = enqueue_step(h, n))) return n;
    ++n;
    --nsteps;
  }
  // a step from a state with arbitrary p halos takes the in-place solve once; every step leaves
  // the halos mirror-consistent, the state the captured graph assumes.  Mixed precision reads
  // the residual on the host between passes, so its steps are not captured.
  while (nsteps > 0 && (!h->p_consistent || h->cfg.precision == LES_PREC_MIXED)) {
    if ((r = enqueue_step(h, n))) return n;
    ++n;
    --nsteps;
  }
  // the captured step of the resident solve starts from a carried split copy
  while (nsteps > 0 && resident_applies(h) && !h->sp_p_current) {
    if ((r = enqueue_step(h, n))) return n;
    ++n;
    --nsteps;
  }
  return n;
}

// captures one step (n >= 2) of the handle into its graph, if not valid yet
int ensure_graph(les_handle* h) {
  if (h->graph_valid) return LES_OK;
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  int r = enqueue_step(h, 2);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(h->st, &gr);
  if (r) {
    if (gr) hipGraphDestroy(gr);
    return r;
  }
  if (e != hipSuccess) return fail(h, LES_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  h->graph = gr;
  HIPCHK(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  h->graph_valid = true;
  h->graph_resident = h->sp_p_current;  // set by the captured step iff it is resident
  return LES_OK;
}

// the host-side state a replay of the captured step leaves
void after_replays(les_handle* h) {
  h->p_consistent = true;
  if (h->graph_resident) {
    h->sp_p_current = true;
    h->rhs_stale = true;
  }
}
}  // namespace

extern "C" int les_run(les_handle* h, int n0, int nsteps) {
  if (!h || nsteps < 0) return LES_E_INVALID;
  int r;
  run_eager_prefix(h, n0, nsteps, r);
  if (r) return r;
  if (nsteps == 0) return LES_OK;
  if ((r = ensure_graph(h))) return r;
  for (int s = 0; s < nsteps; ++s) HIPCHK(h, hipGraphLaunch(h->gexec, h->st));
  after_replays(h);
  return LES_OK;
}

// ------------------------------------------------------------------------------------------
// ensemble: independent members of any sizes advanced together
// ------------------------------------------------------------------------------------------
struct les_ensemble {
  std::vector<les_handle*> m;
  std::string err;
  int mode = LES_ENSEMBLE_BATCHED;  // les_ensemble_set_mode
  int last_batched = 0;             // the last les_ensemble_run used the batched launches
  // the batched form: one stream and one captured step of every member
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  static constexpr int NAUX = 8;  // streams of the per-member segments inside the graph
  hipStream_t aux[NAUX] = {};
  hipEvent_t fork_ev = nullptr, join_ev[NAUX] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<int> sig;  // the members' launch settings the graph was captured with
  BatchStencil* tab_st = nullptr;
  BatchSor* tab_sor = nullptr;
  uint8_t* wgm = nullptr;  // member of each logical workgroup: velfg | calc_sm | visc | sor
  int tot[4] = {0, 0, 0, 0};
  int shp_sm = 0, shp_visc = 0;
};

namespace {
void batch_release(les_ensemble* e) {
  if (e->gexec) hipGraphExecDestroy(e->gexec);
  if (e->graph) hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  if (e->st) hipStreamSynchronize(e->st);
  if (e->tab_st) hipFree(e->tab_st);
  if (e->tab_sor) hipFree(e->tab_sor);
  if (e->wgm) hipFree(e->wgm);
  e->tab_st = nullptr;
  e->tab_sor = nullptr;
  e->wgm = nullptr;
  e->sig.clear();
}

// The batched step applies when every member takes the resident solve on the same kind of
// horizontal grid (at most 256 members: the workgroup-to-member table holds bytes).
bool batch_eligible(const les_ensemble* e) {
  if (e->mode != LES_ENSEMBLE_BATCHED || e->m.size() > 256) return false;
  const les_handle* h0 = e->m[0];
  for (const les_handle* h : e->m) {
    if (h->cfg.precision != LES_PREC_FP32 || h->cfg.sum_order != LES_SUM_EXACT) return false;
    if (!h->resident_ok || h->drag || h->p2xy != h0->p2xy || h->d.jp % 2 || !h->cn1_separable ||
        !h->split_ok || h->ph.pjuge > 0.0f || !sor_iter_supported(h->d, h->sp) ||
        h->cfg.nmaxp < 2 || h->cfg.nmaxp % 2)
      return false;
    if (h->cfg.nmaxp != h0->cfg.nmaxp) return false;
  }
  return true;
}

std::vector<int> batch_signature(const les_ensemble* e) {
  std::vector<int> v;
  for (const les_handle* h : e->m) {
    v.push_back(h->iter_shape.kc);
    v.push_back(h->shp_velnw);
    v.push_back(h->shp_sm);
    v.push_back(h->shp_visc);
    v.push_back(h->p2xy ? 1 : 0);
  }
  return v;
}

// the workgroup shape most members' autotuning chose for a kernel
int majority_shape(const les_ensemble* e, int les_handle::*field) {
  int cnt[STENCIL_SHAPES] = {};
  for (const les_handle* h : e->m) cnt[h->*field] += 1;
  int best = 0;
  for (int q = 1; q < STENCIL_SHAPES; ++q)
    if (cnt[q] > cnt[best]) best = q;
  return best;
}

// Captures one step (n >= 2) of every member on the ensemble's stream: per member velnw and
// bondv1; velfg + feedbf, calc_sm + boundsm and visc + adam as one launch each over all
// members' workgroups; per member rhs and rhsav; the fused SOR iterations as one launch each
// over all members' tiles (the resident solve); per member pav and the gather.
int batch_capture(les_ensemble* e) {
  batch_release(e);
  const int nm = (int)e->m.size();
  e->shp_sm = majority_shape(e, &les_handle::shp_sm);
  e->shp_visc = majority_shape(e, &les_handle::shp_visc);
  std::vector<BatchStencil> ts(nm);
  std::vector<BatchSor> tq(nm);
  std::vector<uint8_t> w[4];
  for (int q = 0; q < nm; ++q) {
    les_handle* h = e->m[q];
    BatchStencil& a = ts[q];
    a.d = h->d;
    a.g = grid_of(h);
    a.F = fields_of(h);
    a.ph = h->ph;
    a.sg_velfg = velfg_strip(h->d);
    const int sx = stencil_shape_x(e->shp_sm), sy = stencil_shape_y(e->shp_sm);
    a.sg_sm = strip_grid((h->d.ip + sx - 1) / sx, (h->d.jp + sy - 1) / sy, calc_sm_kblocks(h->d.kp));
    const int vx = stencil_shape_x(e->shp_visc), vy = stencil_shape_y(e->shp_visc);
    a.sg_visc = strip_grid((h->d.ip + vx - 1) / vx, (h->d.jp + vy - 1) / vy, visc_kblocks(h->d.kp));
    const StripGrid* sgs[3] = {&a.sg_velfg, &a.sg_sm, &a.sg_visc};
    int* bases[3] = {&a.base_velfg, &a.base_sm, &a.base_visc};
    for (int k = 0; k < 3; ++k) {
      *bases[k] = (int)w[k].size();
      w[k].insert(w[k].end(), (size_t)sgs[k]->nbx * sgs[k]->nby * sgs[k]->nbz, (uint8_t)q);
    }
    BatchSor& b = tq[q];
    b.d = h->d;
    b.g = grid_of(h);
    b.sp = h->sp;
    b.sra = h->sp;
    b.sra.r[0] = h->ra[0];
    b.sra.r[1] = h->ra[1];
    b.ph = h->ph;
    b.first = SorPair{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}, {h->ra[0], h->ra[1]}, 1};
    b.fwd = SorPair{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}, {nullptr, nullptr}, 0};
    b.back = SorPair{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}, {nullptr, nullptr}, 0};
    b.sc = h->sc;
    b.kc = std::min(h->iter_shape.kc > 0 ? h->iter_shape.kc : 10, SOR_ITER_KCMAX);
    b.njt = (h->d.jp + 3) / 4;
    b.base = (int)w[3].size();
    w[3].insert(w[3].end(), (size_t)b.njt * ((h->d.kp + b.kc - 1) / b.kc), (uint8_t)q);
  }
  size_t nw = 0;
  for (int k = 0; k < 4; ++k) {
    e->tot[k] = (int)w[k].size();
    nw += w[k].size();
  }
  if (hipMalloc(&e->tab_st, nm * sizeof(BatchStencil)) != hipSuccess ||
      hipMalloc(&e->tab_sor, nm * sizeof(BatchSor)) != hipSuccess ||
      hipMalloc(&e->wgm, nw) != hipSuccess) {
    batch_release(e);
    return LES_E_NOMEM;
  }
  std::vector<uint8_t> all;
  for (int k = 0; k < 4; ++k) all.insert(all.end(), w[k].begin(), w[k].end());
  if (hipMemcpy(e->tab_st, ts.data(), nm * sizeof(BatchStencil), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->tab_sor, tq.data(), nm * sizeof(BatchSor), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->wgm, all.data(), nw, hipMemcpyHostToDevice) != hipSuccess) {
    batch_release(e);
    return LES_E_HIP;
  }
  const uint8_t* wg[4];
  size_t off = 0;
  for (int k = 0; k < 4; ++k) {
    wg[k] = e->wgm + off;
    off += w[k].size();
  }
  const int p2 = e->m[0]->p2xy ? 1 : 0;
  for (int a = 0; a < les_ensemble::NAUX; ++a) {
    if (!e->aux[a] && hipStreamCreateWithFlags(&e->aux[a], hipStreamNonBlocking) != hipSuccess)
      return LES_E_HIP;
    if (!e->join_ev[a] && hipEventCreateWithFlags(&e->join_ev[a], hipEventDisableTiming) != hipSuccess)
      return LES_E_HIP;
  }
  if (!e->fork_ev && hipEventCreateWithFlags(&e->fork_ev, hipEventDisableTiming) != hipSuccess)
    return LES_E_HIP;
  if (hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    batch_release(e);
    return LES_E_HIP;
  }
  // The members' own (small) stages run as NAUX parallel chains of the graph: forked from the
  // ensemble's stream, member q on chain q % NAUX, joined before the next batched launch.
  auto per_member = [&](auto&& body) {
    hipEventRecord(e->fork_ev, e->st);
    for (int a = 0; a < les_ensemble::NAUX; ++a) hipStreamWaitEvent(e->aux[a], e->fork_ev, 0);
    for (size_t q = 0; q < e->m.size(); ++q) body(e->m[q], e->aux[q % les_ensemble::NAUX]);
    for (int a = 0; a < les_ensemble::NAUX; ++a) {
      hipEventRecord(e->join_ev[a], e->aux[a]);
      hipStreamWaitEvent(e->st, e->join_ev[a], 0);
    }
  };
  per_member([&](les_handle* h, hipStream_t st) {
    const Dims d = h->d;
    const Grid1D g = grid_of(h);
    const Fields F = fields_of(h);
    launch_velnw_inflow(st, d, g, F, h->ph, 1, 2, h->shp_velnw);
    launch_uout(st, d, F, h->sc, reinterpret_cast<float*>(h->rb.partials));
    launch_bondv1_uvw(st, d, g, F, h->ph, h->sc);
  });
  launch_velfg_batch(e->st, e->tab_st, wg[0], e->tot[0], p2);
  launch_calc_sm_batch(e->st, e->tab_st, wg[1], e->tot[1], e->shp_sm);
  launch_visc_adam_batch(e->st, e->tab_st, wg[2], e->tot[2], e->shp_visc, p2);
  per_member([&](les_handle* h, hipStream_t st) {
    launch_rhs(st, h->d, grid_of(h), fields_of(h), h->ph, h->rb, h->sc, EDGES_ALL, 1, &h->sp);
  });
  launch_sor_iter_batch(e->st, e->tab_sor, wg[3], e->tot[3], 0);
  launch_sor_iter_batch(e->st, e->tab_sor, wg[3], e->tot[3], 2);
  for (int l = 3; l + 1 <= e->m[0]->cfg.nmaxp; l += 2) {
    launch_sor_iter_batch(e->st, e->tab_sor, wg[3], e->tot[3], 1);
    launch_sor_iter_batch(e->st, e->tab_sor, wg[3], e->tot[3], 2);
  }
  per_member([&](les_handle* h, hipStream_t st) {
    launch_pav_split(st, h->d, grid_of(h), h->sp, h->rb, h->sc);
    launch_split_gather_adj(st, h->d, fields_of(h), h->sp, h->sc, 1, 1);
  });
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->st, &gr);
  if (ec != hipSuccess || !gr) {
    if (gr) hipGraphDestroy(gr);
    batch_release(e);
    e->err = std::string("batched capture: ") + hipGetErrorString(ec);
    return LES_E_HIP;
  }
  e->graph = gr;
  if (hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0) != hipSuccess) {
    batch_release(e);
    return LES_E_HIP;
  }
  e->sig = batch_signature(e);
  return LES_OK;
}

// The batched form of les_ensemble_run: eager steps of every member in lockstep until each can
// take the captured step (n >= 2, mirror-consistent p halos, a carried split copy), then the
// ensemble's graph, ordered after the members' streams and before them again.
int batch_run(les_ensemble* e, int n0, int nsteps) {
  int r;
  // the batched launches run every member's SOR as fused iterations: a member tuned to the
  // two colour sweeps takes the fused form with the default depth (the same values)
  for (les_handle* h : e->m)
    if (h->iter_shape.tj <= 0) {
      h->iter_shape = SorIterShape{4, 10};
      h->graph_valid = false;
    }
  auto ready = [&](const les_handle* h) { return h->p_consistent && h->sp_p_current; };
  while (nsteps > 0) {
    bool all = n0 >= 2;
    for (const les_handle* h : e->m) all = all && ready(h);
    if (all) break;
    for (les_handle* h : e->m)
      if ((r = enqueue_step(h, n0))) return r;
    ++n0;
    --nsteps;
  }
  if (nsteps == 0) return LES_OK;
  if (!e->st && hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) return LES_E_HIP;
  if (!e->ev && hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess) return LES_E_HIP;
  for (les_handle* h : e->m) {  // the ensemble's stream after every member's queued work
    if (hipEventRecord(e->ev, h->st) != hipSuccess || hipStreamWaitEvent(e->st, e->ev, 0) != hipSuccess)
      return LES_E_HIP;
  }
  if (!e->gexec || e->sig != batch_signature(e)) {
    if ((r = batch_capture(e))) return r;
  }
  for (int s = 0; s < nsteps; ++s)
    if (hipGraphLaunch(e->gexec, e->st) != hipSuccess) return LES_E_HIP;
  // the members' state as the captured steps leave it; their streams after the ensemble's
  if (hipEventRecord(e->ev, e->st) != hipSuccess) return LES_E_HIP;
  for (les_handle* h : e->m) {
    h->p_consistent = true;
    h->sp_p_current = true;
    h->rhs_stale = true;
    if (hipStreamWaitEvent(h->st, e->ev, 0) != hipSuccess) return LES_E_HIP;
  }
  return LES_OK;
}
}  // namespace

extern "C" int les_ensemble_create(les_handle* const* members, int n, les_ensemble** out) {
  if (!members || n < 1 || !out) return LES_E_INVALID;
  for (int q = 0; q < n; ++q)
    if (!members[q]) return LES_E_INVALID;
  les_ensemble* e = new les_ensemble();
  e->m.assign(members, members + n);
  *out = e;
  return LES_OK;
}

extern "C" int les_ensemble_set_mode(les_ensemble* e, int mode) {
  if (!e || (mode != LES_ENSEMBLE_CONCURRENT && mode != LES_ENSEMBLE_BATCHED)) return LES_E_INVALID;
  e->mode = mode;
  return LES_OK;
}

extern "C" int les_ensemble_batched(const les_ensemble* e) { return e ? e->last_batched : 0; }

extern "C" int les_ensemble_destroy(les_ensemble* e) {
  if (!e) return LES_OK;
  batch_release(e);
  if (e->ev) hipEventDestroy(e->ev);
  if (e->fork_ev) hipEventDestroy(e->fork_ev);
  for (int a = 0; a < les_ensemble::NAUX; ++a) {
    if (e->join_ev[a]) hipEventDestroy(e->join_ev[a]);
    if (e->aux[a]) hipStreamDestroy(e->aux[a]);
  }
  if (e->st) hipStreamDestroy(e->st);
  delete e;
  return LES_OK;
}

// Batched (when every member allows it): one launch per stage over all members.  Otherwise
// every member's stream gets one graph launch per step, step-major, so the members' kernels
// are in flight together and small members fill the device between the large ones.
extern "C" int les_ensemble_run(les_ensemble* e, int n0, int nsteps) {
  if (!e || nsteps < 0) return LES_E_INVALID;
  if (batch_eligible(e)) {
    e->last_batched = 1;
    return batch_run(e, n0, nsteps);
  }
  e->last_batched = 0;
  const int nm = (int)e->m.size();
  std::vector<int> left(nm, nsteps);
  for (int q = 0; q < nm; ++q) {
    les_handle* h = e->m[q];
    int r;
    run_eager_prefix(h, n0, left[q], r);
    if (r) return r;
    if (left[q] > 0 && h->cfg.precision == LES_PREC_MIXED) return LES_E_STATE;
    if (left[q] > 0 && (r = ensure_graph(h))) return r;
  }
  for (int s = 0; s < nsteps; ++s)
    for (int q = 0; q < nm; ++q) {
      les_handle* h = e->m[q];
      if (s < nsteps - left[q]) continue;  // this member ran its first steps eagerly
      HIPCHK(h, hipGraphLaunch(h->gexec, h->st));
    }
  for (int q = 0; q < nm; ++q)
    if (left[q] > 0) after_replays(e->m[q]);
  return LES_OK;
}

extern "C" int les_ensemble_sync(les_ensemble* e) {
  if (!e) return LES_E_INVALID;
  if (e->st && hipStreamSynchronize(e->st) != hipSuccess) return LES_E_HIP;
  for (les_handle* h : e->m) {
    int r = les_sync(h);
    if (r) return r;
  }
  return LES_OK;
}

extern "C" int les_ifdata(les_handle* h) {
  STAGE_BEGIN(h);
  launch_velnw_inflow(h->st, d, g, F, h->ph, 0, 1, h->shp_velnw);
  launch_uout(h->st, d, F, h->sc, reinterpret_cast<float*>(h->rb.partials));
  launch_bondv1_uvw(h->st, d, g, F, h->ph, h->sc);
  launch_boundp1(h->st, d, F, h->sc, 0);
  launch_boundp2(h->st, d, F, h->sc, 0);
  h->p_consistent = true;
  launch_velfg(h->st, d, g, F, h->ph, 1, 1);
  launch_calc_sm(h->st, d, g, F, h->ph, 1, h->shp_sm);  // boundsm's halo cells included
  launch_calc_visc_adam(h->st, d, g, F, 1, 0, h->shp_visc);
  return check_launch(h);
}

extern "C" int les_get_scalars(les_handle* h, les_scalars* out) {
  if (!h || !out) return LES_E_INVALID;
  DevScalars s;
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost));
  out->uout = s.uout;
  out->rhsav = s.rhsav;
  out->area = s.area;
  out->pav = s.pav;
  out->pco = s.pco;
  out->sor = s.sor;
  out->sor_iters = (h->ph.pjuge > 0.0f) ? s.sor_iters : h->cfg.nmaxp;
  out->resid_l2 = s.resid_l2;
  out->refine_passes = s.refine_passes;
  out->cg_iters = 0;
  if (h->cfg.precision == LES_PREC_MIXED) {
    double it = 0.0;
    HIPCHK(h, hipMemcpy(&it, h->mx.cg + 6, sizeof(double), hipMemcpyDeviceToHost));
    out->cg_iters = (int)it;
  }
  return LES_OK;
}

// ------------------------------------------------------------------------------------------
// halo exchange
// ------------------------------------------------------------------------------------------
namespace {
bool extent_of(const les_handle* h, int field, int& im, int& jm, int& km) {
  int lo[3], hi[3];
  if (les_field_bounds(field, h->d.ip, h->d.jp, h->d.kp, lo, hi) < 0) return false;
  BoundSpec s;
  spec_of(field, s);
  if (s.ndim != 3) return false;
  im = hi[0] - lo[0] + 1;
  jm = hi[1] - lo[1] + 1;
  km = hi[2] - lo[2] + 1;
  return true;
}
}  // namespace

extern "C" int64_t les_halo_size(int v_dim, int im, int jm, int km, int h_w, int correct) {
  if (v_dim < 1 || im < 1 || jm < 1 || km < 1 || h_w < 1) return -1;
  if (correct) return (int64_t)v_dim * 2 * km * (int64_t)(im + jm - 4 * h_w) * h_w;
  return (int64_t)v_dim * (2LL * km * im * h_w + 2LL * km * (jm - 2 * h_w) * h_w);
}

static int halo_common(les_handle* h, const int* fields, int v_dim, int h_w, float** ptrs, int& im,
                       int& jm, int& km) {
  if (!h || !fields || v_dim < 1 || v_dim > 16 || h_w < 1) return LES_E_INVALID;
  for (int v = 0; v < v_dim; ++v) {
    int a, b, c;
    if (fields[v] < 0 || fields[v] >= LES_NFIELDS || !extent_of(h, fields[v], a, b, c))
      return fail(h, LES_E_INVALID, "halo: not a 3-D field");
    if (v == 0) {
      im = a;
      jm = b;
      km = c;
    } else if (a != im || b != jm || c != km) {
      return fail(h, LES_E_INVALID, "halo: fields of one exchange must have equal extents");
    }
    ptrs[v] = h->fld[fields[v]];
  }
  if (jm < 2 * h_w + 1 || im < 2 * h_w + 1) return fail(h, LES_E_INVALID, "halo: width too large");
  return LES_OK;
}

extern "C" int les_exchange_halo_read(les_handle* h, const int* fields, int v_dim, int h_w,
                                      int correct, float* dev_buffer) {
  float* ptrs[16];
  int im, jm, km;
  int r = halo_common(h, fields, v_dim, h_w, ptrs, im, jm, km);
  if (r) return r;
  if (!dev_buffer) return fail(h, LES_E_INVALID, "halo: null buffer");
  sync_rhs(h);
  launch_halo_read(h->st, ptrs, v_dim, im, jm, km, h_w, correct ? 1 : 0, dev_buffer);
  return check_launch(h);
}

extern "C" int les_exchange_halo_write(les_handle* h, const int* fields, int v_dim, int h_w,
                                       const float* dev_buffer) {
  float* ptrs[16];
  int im, jm, km;
  int r = halo_common(h, fields, v_dim, h_w, ptrs, im, jm, km);
  if (r) return r;
  if (!dev_buffer) return fail(h, LES_E_INVALID, "halo: null buffer");
  sync_rhs(h);
  for (int q = 0; q < v_dim; ++q)
    if (fields[q] == LES_P) h->sp_p_current = false;
  launch_halo_write(h->st, ptrs, v_dim, im, jm, km, h_w, dev_buffer);
  return check_launch(h);
}

static int halo_arrays_check(les_handle* h, float* const* dev_arrays, int v_dim, int im, int jm,
                             int km, int h_w, const float* dev_buffer) {
  if (!h) return LES_E_INVALID;
  if (!dev_arrays || !dev_buffer || v_dim < 1 || v_dim > 16 || km < 1 || h_w < 1 ||
      im < 2 * h_w + 1 || jm < 2 * h_w + 1)
    return fail(h, LES_E_INVALID, "halo: bad arrays, sizes or width");
  for (int v = 0; v < v_dim; ++v)
    if (!dev_arrays[v]) return fail(h, LES_E_INVALID, "halo: null array");
  return LES_OK;
}

extern "C" int les_halo_read_arrays(les_handle* h, float* const* dev_arrays, int v_dim, int im,
                                    int jm, int km, int h_w, int correct, float* dev_buffer) {
  int r = halo_arrays_check(h, dev_arrays, v_dim, im, jm, km, h_w, dev_buffer);
  if (r) return r;
  launch_halo_read(h->st, dev_arrays, v_dim, im, jm, km, h_w, correct ? 1 : 0, dev_buffer);
  return check_launch(h);
}

extern "C" int les_halo_write_arrays(les_handle* h, float* const* dev_arrays, int v_dim, int im,
                                     int jm, int km, int h_w, const float* dev_buffer) {
  int r = halo_arrays_check(h, dev_arrays, v_dim, im, jm, km, h_w, dev_buffer);
  if (r) return r;
  launch_halo_write(h->st, dev_arrays, v_dim, im, jm, km, h_w, dev_buffer);
  return check_launch(h);
}

// exact-sum utility on a device vector; result to host
extern "C" int les_reduce_sum_exact(les_handle* h, const float* dev_x, int64_t n, float* out) {
  if (!h || !dev_x || !out || n < 0) return LES_E_INVALID;
  launch_exact_sum(h->st, dev_x, (long)n, h->rb, h->scratch_f);
  HIPCHK(h, hipMemcpyAsync(out, h->scratch_f, sizeof(float), hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return LES_OK;
}

extern "C" int les_time_kernel(les_handle* h, int which, int reps, float* avg_ms) {
  if (!h || !avg_ms || reps < 1) return LES_E_INVALID;
  STAGE_BEGIN(h);
  sync_rhs(h);
  h->sp_p_current = false;  // the timed kernels overwrite the split copy
  if (which == LES_KERNEL_STEP && !h->graph_valid) {
    int r = les_run(h, 2, 1);  // captures the graph
    if (r) return r;
  }
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  HIPCHK(h, hipEventRecord(e0, h->st));
  for (int q = 0; q < reps; ++q) {
    switch (which) {
      case LES_KERNEL_SOR_SWEEP:
        if (use_split(h))
          launch_sor_split(h->st, d, g, h->sp, h->ph, q & 1, 0, h->rb, h->sc);
        else
          launch_sor_colour(h->st, d, g, F, h->ph, q & 1, 0, h->rb, h->sc);
        break;
      case LES_KERNEL_VELFG: launch_velfg(h->st, d, g, F, h->ph, 1, 1); break;
      case LES_KERNEL_VISC_ADAM: launch_calc_visc_adam(h->st, d, g, F, 1, 1, h->shp_visc); break;
      case LES_KERNEL_CALC_SM: launch_calc_sm(h->st, d, g, F, h->ph, 0, h->shp_sm); break;
      case LES_KERNEL_RHS: launch_rhs(h->st, d, g, F, h->ph, h->rb, h->sc); break;
      case LES_KERNEL_VELNW: launch_velnw_inflow(h->st, d, g, F, h->ph, 1, 2, h->shp_velnw); break;
      case LES_KERNEL_STEP: HIPCHK(h, hipGraphLaunch(h->gexec, h->st)); break;
      case LES_KERNEL_SOR_ITER: {
        if (!sor_iter_supported(d, h->sp)) {
          hipEventDestroy(e0);
          hipEventDestroy(e1);
          return fail(h, LES_E_STATE, "time_kernel: the fused SOR iteration does not apply");
        }
        const SorPair io = (q & 1) ? SorPair{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}}
                                   : SorPair{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}};
        launch_sor_iter(h->st, d, g, h->sp, h->ph, io, h->sc, h->iter_shape);
        break;
      }
      default:
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return fail(h, LES_E_INVALID, "time_kernel: unknown kernel");
    }
  }
  HIPCHK(h, hipEventRecord(e1, h->st));
  HIPCHK(h, hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *avg_ms = ms / (float)reps;
  return check_launch(h);
}

extern "C" int les_set_sor_tile(les_handle* h, int tile_j, int tile_k) {
  if (!h) return LES_E_INVALID;
  if (tile_j == 0) {
    h->iter_shape = SorIterShape{0, 0};
  } else {
    if ((tile_j != 4 && tile_j != 8 && tile_j != 16) || tile_k < 1 || tile_k > SOR_ITER_KCMAX)
      return fail(h, LES_E_INVALID,
                  "set_sor_tile: tile_j must be 0, 4, 8 or 16 and tile_k in 1..128");
    if (!sor_iter_supported(h->d, h->sp))
      return fail(h, LES_E_STATE, "set_sor_tile: the fused SOR iteration does not apply to this grid");
    h->iter_shape = SorIterShape{tile_j, tile_k};
  }
  h->graph_valid = false;
  return LES_OK;
}

// Autotuner: times one SOR iteration as two colour sweeps and as the fused launch with every
// workgroup tile on this handle's grid (the split copy is scratch between solves, so the trial
// launches change no simulation state) and keeps the fastest; the captured step graph is
// rebuilt with it.
namespace {
// The workgroup shape of each point-wise stencil kernel: every shape timed over 'reps' launches
// from the same state (the arrays the kernels update are saved first and put back after each
// trial and at the end).  Without memory for the copy the shapes stay as they are.
int tune_stencils(les_handle* h, int reps, les_tuning& best) {
  STAGE_BEGIN(h);
  const int saved[] = {LES_U, LES_V, LES_W, LES_F, LES_G, LES_H, LES_FOLD, LES_GOLD, LES_HOLD,
                       LES_SM};
  size_t tot = 0;
  for (int f : saved) tot += (size_t)h->cnt[f];
  int* shp[3] = {&h->shp_velnw, &h->shp_sm, &h->shp_visc};
  for (int q = 0; q < 3; ++q) {
    best.stencil_bx[q] = stencil_shape_x(*shp[q]);
    best.stencil_by[q] = stencil_shape_y(*shp[q]);
    best.stencil_ms[q] = -1.0f;
  }
  float* keep = nullptr;
  if (hipMalloc(&keep, tot * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    return LES_OK;
  }
  auto copy_all = [&](bool to_keep) {
    size_t off = 0;
    for (int f : saved) {
      const size_t n = (size_t)h->cnt[f] * sizeof(float);
      if (to_keep)
        (void)hipMemcpyAsync(keep + off, h->fld[f], n, hipMemcpyDeviceToDevice, h->st);
      else
        (void)hipMemcpyAsync(h->fld[f], keep + off, n, hipMemcpyDeviceToDevice, h->st);
      off += (size_t)h->cnt[f];
    }
  };
  auto launch = [&](int q, int shape) {
    if (q == 0) launch_velnw_inflow(h->st, d, g, F, h->ph, 1, 2, shape);
    if (q == 1) launch_calc_sm(h->st, d, g, F, h->ph, 1, shape);
    if (q == 2) launch_calc_visc_adam(h->st, d, g, F, 1, 1, shape);
  };
  hipEvent_t e0, e1;
  int r = LES_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    r = fail(h, LES_E_HIP, "autotune: event");
  } else {
    copy_all(true);
    for (int q = 0; q < 3 && !r; ++q) {
      float best_ms = -1.0f;
      int best_shape = *shp[q];
      for (int shape = 0; shape < STENCIL_SHAPES; ++shape) {
        copy_all(false);
        launch(q, shape);  // warm-up
        (void)hipEventRecord(e0, h->st);
        for (int t = 0; t < reps; ++t) launch(q, shape);
        (void)hipEventRecord(e1, h->st);
        if (hipEventSynchronize(e1) != hipSuccess) {
          r = fail(h, LES_E_HIP, "autotune: stencil trial");
          break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        if (best_ms < 0.0f || ms < best_ms) {
          best_ms = ms;
          best_shape = shape;
        }
      }
      *shp[q] = best_shape;
      best.stencil_bx[q] = stencil_shape_x(best_shape);
      best.stencil_by[q] = stencil_shape_y(best_shape);
      best.stencil_ms[q] = best_ms;
    }
    copy_all(false);
    (void)hipStreamSynchronize(h->st);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(keep);
  h->graph_valid = false;
  return r ? r : check_launch(h);
}
}  // namespace

extern "C" int les_autotune(les_handle* h, int reps, les_tuning* out) {
  if (!h || reps < 1) return LES_E_INVALID;
  STAGE_BEGIN(h);
  sync_rhs(h);
  h->sp_p_current = false;  // the trials overwrite the split copy
  les_tuning best{};
  best.sor_iter_ms = -1.0f;
  if (h->split_ok && sor_iter_supported(d, h->sp)) {
    const int tjs[] = {4, 8, 16};
    const int kcs[] = {5, 10, 15, 23, 30, 45};
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    {  // the two colour sweeps of one iteration
      for (int q = 0; q < 4; ++q) launch_sor_split(h->st, d, g, h->sp, h->ph, q & 1, 0, h->rb, h->sc);
      HIPCHK(h, hipEventRecord(e0, h->st));
      for (int q = 0; q < 2 * reps; ++q)
        launch_sor_split(h->st, d, g, h->sp, h->ph, q & 1, 0, h->rb, h->sc);
      HIPCHK(h, hipEventRecord(e1, h->st));
      HIPCHK(h, hipEventSynchronize(e1));
      float ms = 0.0f;
      HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
      best.sor_iter_ms = ms / (float)reps;
      best.sor_tile_j = 0;
      best.sor_tile_k = 0;
    }
    for (int tj : tjs)
      for (int kc : kcs) {
        if (kc > h->d.kp) continue;
        const SorIterShape shp{tj, kc};
        for (int q = 0; q < 4; ++q) {  // warm-up
          const SorPair io = (q & 1) ? SorPair{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}}
                                     : SorPair{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}};
          launch_sor_iter(h->st, d, g, h->sp, h->ph, io, h->sc, shp);
        }
        HIPCHK(h, hipEventRecord(e0, h->st));
        for (int q = 0; q < reps; ++q) {
          const SorPair io = (q & 1) ? SorPair{{h->palt[0], h->palt[1]}, {h->sp.p[0], h->sp.p[1]}}
                                     : SorPair{{h->sp.p[0], h->sp.p[1]}, {h->palt[0], h->palt[1]}};
          launch_sor_iter(h->st, d, g, h->sp, h->ph, io, h->sc, shp);
        }
        HIPCHK(h, hipEventRecord(e1, h->st));
        HIPCHK(h, hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        if (ms < best.sor_iter_ms) {
          best.sor_iter_ms = ms;
          best.sor_tile_j = tj;
          best.sor_tile_k = kc;
        }
      }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    h->iter_shape = SorIterShape{best.sor_tile_j, best.sor_tile_k};
    h->graph_valid = false;
  }
  {
    int r = tune_stencils(h, reps, best);
    if (r) return r;
  }
  h->tuned = true;
  if (out) *out = best;
  return check_launch(h);
}

// ------------------------------------------------------------------------------------------
// the reference's GPU host API on a default handle
// ------------------------------------------------------------------------------------------
extern "C" int les_initialise_LES_kernel(
    const float* p, const float* u, const float* v, const float* w, const float* usum,
    const float* vsum, const float* wsum, const float* f, const float* g, const float* hh,
    const float* fold, const float* gold, const float* hold, const float* diu1,
    const float* diu2, const float* diu3, const float* diu4, const float* diu5,
    const float* diu6, const float* diu7, const float* diu8, const float* diu9,
    const float* amask1, const float* bmask1, const float* cmask1, const float* dmask1,
    const float* cn1, const float* cn2l, const float* cn2s, const float* cn3l,
    const float* cn3s, const float* cn4l, const float* cn4s, const float* rhs,
    const float* sm, const float* dxs, const float* dys, const float* dzs, const float* dx1,
    const float* dy1, const float* dzn, const float* wind_profile, float dt, int im, int jm,
    int km) {
  if (g_default) les_destroy(g_default);  // les_destroy clears g_default
  g_default = nullptr;
  les_config cfg;
  les_default_config(&cfg, im, jm, km);
  cfg.dt = dt;
  les_handle* h = nullptr;
  int r = les_create(&cfg, &h);
  if (r) return r;
  g_default = h;
  const float* arrs[] = {p, u, v, w, usum, vsum, wsum, f, g, hh, fold, gold, hold, rhs, sm,
                         diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9,
                         amask1, bmask1, cmask1, dmask1, cn1, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s,
                         dx1, dxs, dy1, dys, dzn, dzs};
  const int ids[] = {LES_P, LES_U, LES_V, LES_W, LES_USUM, LES_VSUM, LES_WSUM, LES_F, LES_G,
                     LES_H, LES_FOLD, LES_GOLD, LES_HOLD, LES_RHS, LES_SM, LES_DIU1, LES_DIU2,
                     LES_DIU3, LES_DIU4, LES_DIU5, LES_DIU6, LES_DIU7, LES_DIU8, LES_DIU9,
                     LES_AMASK1, LES_BMASK1, LES_CMASK1, LES_DMASK1, LES_CN1, LES_CN2L, LES_CN2S,
                     LES_CN3L, LES_CN3S, LES_CN4L, LES_CN4S, LES_DX1, LES_DXS, LES_DY1, LES_DYS,
                     LES_DZN, LES_DZS};
  for (size_t q = 0; q < sizeof(ids) / sizeof(ids[0]); ++q) {
    if (!arrs[q]) continue;
    if ((r = les_upload(h, ids[q], arrs[q]))) return r;
  }
  if (wind_profile && (r = les_upload(h, LES_WIND_PROFILE, wind_profile))) return r;
  // delx1 (les.f95:36) from the uploaded grid
  {
    std::vector<float> delx1(km);
    const auto& hdx1 = h->hgrid[LES_DX1];
    const auto& hdy1 = h->hgrid[LES_DY1];
    const auto& hdzn = h->hgrid[LES_DZN];
    for (int k = 1; k <= km; ++k) delx1[k - 1] = powf(hdx1[1] * hdy1[0] * hdzn[k + 1], 1.0f / 3.0f);
    if ((r = les_upload(h, LES_DELX1, delx1.data()))) return r;
  }
  return les_sync(h);
}

extern "C" int les_run_LES_kernel(int n, int nmax, const float* wind_profile) {
  (void)nmax;
  if (!g_default) return LES_E_STATE;
  int r;
  if (wind_profile && (r = les_upload(g_default, LES_WIND_PROFILE, wind_profile))) return r;
  return les_step(g_default, n);
}

extern "C" les_handle* les_default_handle(void) { return g_default; }

extern "C" int les_finalise_LES_kernel(void) {
  if (g_default) les_destroy(g_default);
  g_default = nullptr;
  return LES_OK;
}

// ------------------------------------------------------------------------------------------
// One domain cut into sub_i x sub_j tiles on one device.  Each tile is an ordinary handle of
// (ip/sub_i, jp/sub_j, kp) cells with the whole domain's arrays sliced into it (halos
// included).  A step runs the stages tile by tile on one stream; after a stage, the faces a
// tile shares with a neighbour (and the periodic j faces) are copied between the tiles'
// arrays (the halo exchange, here a device copy since every tile is on the same GPU), the
// domain's own faces keep the reference's boundary rules (Edges), and the sums (uout, rhsav,
// pav) combine the tiles' exact partials, so the tiled step is bitwise the single-domain step.
// Pressure sweeps each tile's colour-split copy (k_sor_split_tiles with halo rows and
// columns, which also stores each tile's updated face cells into its neighbours' halos); tiles
// whose grid the split sweep does not cover use the in-place colour sweep (k_sor_colour) with
// Fortran-layout exchanges.
// ------------------------------------------------------------------------------------------
struct les_tiled {
  les_config cfg;  // the whole domain
  int si = 1, sj = 1, ipl = 0, jpl = 0;
  std::vector<les_handle*> t;  // tile (a, b) at t[a + si * b]
  std::vector<Edges> e;
  int64_t* partials = nullptr;
  std::vector<int> red_off;  // first partial of each tile
  int red_total = 0;
  std::vector<float> gdx1, gdy1, gdzn;  // the whole domain's grid, for area = pco
  std::string err;
  hipGraph_t graph = nullptr;  // one captured step (n >= 2) for les_tiled_run
  hipGraphExec_t gexec = nullptr;
  // colour-split copies of p and rhs with halo rows and i-halo columns, per tile (nullptr:
  // the tiles sweep in place)
  float* split_mem = nullptr;
  long split_ns = 0, split_per_tile = 0;
  // the boundary images of p are mirror images of the interior (after tiled_boundp; not after
  // an upload of p), which the split sweep assumes at the domain's faces
  bool p_consistent = false;
};

namespace {

int tfail(les_tiled* T, int code, const std::string& msg) {
  if (T) T->err = msg;
  return code;
}

// axis of a 1-D or 2-D field along which tiles cut it (0 i, 1 j, 2 k)
int tiled_axis(int f) {
  switch (f) {
    case LES_DX1: case LES_DXS: case LES_CN2L: case LES_CN2S: return 0;
    case LES_DY1: case LES_DYS: case LES_CN3L: case LES_CN3S: case LES_WIND_PROFILE: return 1;
    default: return 2;
  }
}

bool is_3d(int f) { return f <= LES_DMASK1 || f == LES_CN1; }

ArrView view_of(les_handle* h, int f) {
  int lo[3], hi[3];
  les_field_bounds(f, h->d.ip, h->d.jp, h->d.kp, lo, hi);
  return ArrView{h->fld[f], lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1};
}

// exchange of the faces of `fields`: dir 0 copies the i faces (rows 1..jpl) between
// neighbours along i; dir 1 the j faces (columns 0..ipl+1, so corners follow the i faces),
// wrapping around in j when `periodic`.
void tiled_exchange(les_tiled* T, hipStream_t st, std::initializer_list<int> fields, int dir,
                    bool periodic = true) {
  ExchangePlan plan{};
  auto flush = [&]() {
    launch_exchange(st, plan);
    plan.nbox = 0;
  };
  auto add = [&](const BoxCopy& b) {
    plan.box[plan.nbox++] = b;
    if (plan.nbox == EXCH_MAX) flush();
  };
  for (int f : fields) {
    for (int b = 0; b < T->sj; ++b)
      for (int a = 0; a < T->si; ++a) {
        les_handle* s = T->t[a + T->si * b];
        const ArrView vs = view_of(s, f);
        const int klo = vs.lo2;
        int lo[3], hi[3];
        les_field_bounds(f, T->ipl, T->jpl, T->cfg.kp, lo, hi);
        const int nk = hi[2] - lo[2] + 1;
        // only the halo columns/rows the array has (f, g, h end at ip, jp)
        const bool has_e = hi[0] >= T->ipl + 1, has_w = lo[0] <= 0;
        const bool has_n = hi[1] >= T->jpl + 1, has_s = lo[1] <= 0;
        const int i0 = std::max(0, lo[0]), i1 = std::min(T->ipl + 1, hi[0]);
        if (dir == 0) {
          if (a + 1 >= T->si) continue;
          les_handle* r = T->t[a + 1 + T->si * b];
          const ArrView vr = view_of(r, f);
          // right's column 1 -> my column ipl+1; my column ipl -> right's column 0
          if (has_e) add(BoxCopy{vr, vs, 1, 1, T->ipl + 1, 1, klo, 1, T->jpl, nk});
          if (has_w) add(BoxCopy{vs, vr, T->ipl, 1, 0, 1, klo, 1, T->jpl, nk});
        } else {
          const int bn = (b + 1) % T->sj;
          if (b + 1 == T->sj && !periodic) continue;
          les_handle* r = T->t[a + T->si * bn];
          const ArrView vr = view_of(r, f);
          // north's row 1 -> my row jpl+1; my row jpl -> north's row 0 (columns i0..i1)
          if (has_n) add(BoxCopy{vr, vs, i0, 1, i0, T->jpl + 1, klo, i1 - i0 + 1, 1, nk});
          if (has_s) add(BoxCopy{vs, vr, i0, T->jpl, i0, 0, klo, i1 - i0 + 1, 1, nk});
        }
      }
  }
  flush();
}

int tiled_write_area(les_tiled* T) {
  if (T->gdx1.empty() || T->gdy1.empty() || T->gdzn.empty()) return LES_OK;
  // press.f95:67-74: the whole domain's sequential FP32 sum (the single-domain handle's value)
  float area = 0.0f;
  double area64 = 0.0;
  for (int k = 1; k <= T->cfg.kp; ++k)
    for (int j = 1; j <= T->cfg.jp; ++j)
      for (int i = 1; i <= T->cfg.ip; ++i) {
        const float v = T->gdx1[i + 1] * T->gdy1[j] * T->gdzn[k + 1];
        area = area + v;
        area64 += (double)v;
      }
  les_handle* h0 = T->t[0];
  HIPCHK(h0, hipStreamSynchronize(h0->st));
  HIPCHK(h0, hipMemcpy(&h0->sc->area, &area, sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h0, hipMemcpy(&h0->sc->pco, &area, sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(h0, hipMemcpy(&h0->sc->area64, &area64, sizeof(double), hipMemcpyHostToDevice));
  return LES_OK;
}

std::vector<DevScalars*> tile_scalars(les_tiled* T, bool east_only) {
  std::vector<DevScalars*> v;
  for (int b = 0; b < T->sj; ++b)
    for (int a = 0; a < T->si; ++a)
      if (!east_only || a == T->si - 1) v.push_back(T->t[a + T->si * b]->sc);
  return v;
}

void tiled_bcast(les_tiled* T, hipStream_t st, size_t off) {
  std::vector<DevScalars*> all = tile_scalars(T, false);
  launch_scalar_bcast(st, all[0], all.data() + 1, (int)all.size() - 1, (int)off);
}

// velnw + bondv1 (inflow, uout, outflow, periodic and top/bottom faces) on every tile
void tiled_bondv1(les_tiled* T, hipStream_t st, int do_velnw, int inflow_n) {
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    // n == 1 fills u with the profile everywhere; later steps set the inflow planes only
    const int in = (inflow_n == 1 || T->e[q].w) ? inflow_n : 0;
    launch_velnw_inflow(st, h->d, grid_of(h), fields_of(h), h->ph, do_velnw, in);
  }
  std::vector<DevScalars*> east = tile_scalars(T, true), all = tile_scalars(T, false);
  for (size_t q = 0; q < T->t.size(); ++q)
    if (T->e[q].e) {
      les_handle* h = T->t[q];
      launch_uout(st, h->d, fields_of(h), h->sc, reinterpret_cast<float*>(h->rb.partials));
    }
  launch_uout_combine(st, east.data(), (int)east.size(), all.data(), (int)all.size());
  for (size_t q = 0; q < T->t.size(); ++q)
    if (T->e[q].e) {
      les_handle* h = T->t[q];
      launch_bondv1_uvw(st, h->d, grid_of(h), fields_of(h), h->ph, h->sc, T->e[q], -1);
    }
  tiled_exchange(T, st, {LES_U, LES_V, LES_W}, 0);
  tiled_exchange(T, st, {LES_U, LES_V, LES_W}, 1);
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    launch_bondv1_uvw(st, h->d, grid_of(h), fields_of(h), h->ph, h->sc, T->e[q], 0);
  }
}

// boundp1 (the domain's i faces), the exchange, boundp2
void tiled_boundp(les_tiled* T, hipStream_t st, int check_done) {
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    launch_boundp1(st, h->d, fields_of(h), h->sc, check_done, T->e[q]);
  }
  tiled_exchange(T, st, {LES_P}, 0);
  tiled_exchange(T, st, {LES_P}, 1);
  for (les_handle* h : T->t) launch_boundp2(st, h->d, fields_of(h), h->sc, check_done);
  T->p_consistent = true;
}

// velfg + feedbf, les (calc_sm, boundsm), visc (+ adam)
void tiled_velfg_les(les_tiled* T, hipStream_t st, int do_adam) {
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    launch_velfg(st, h->d, grid_of(h), fields_of(h), h->ph, 1, 1, T->e[q]);
  }
  const std::initializer_list<int> diu = {LES_DIU1, LES_DIU2, LES_DIU3, LES_DIU4, LES_DIU5,
                                          LES_DIU6, LES_DIU7, LES_DIU8, LES_DIU9};
  tiled_exchange(T, st, diu, 0);
  tiled_exchange(T, st, diu, 1);
  for (les_handle* h : T->t) launch_calc_sm(st, h->d, grid_of(h), fields_of(h), h->ph);
  for (size_t q = 0; q < T->t.size(); ++q)
    launch_boundsm_tiled(st, T->t[q]->d, fields_of(T->t[q]), T->e[q], 0);
  tiled_exchange(T, st, {LES_SM}, 0);
  for (size_t q = 0; q < T->t.size(); ++q)
    launch_boundsm_tiled(st, T->t[q]->d, fields_of(T->t[q]), T->e[q], 1);
  tiled_exchange(T, st, {LES_SM}, 1, false);  // boundsm's j faces are copies, not periodic
  for (size_t q = 0; q < T->t.size(); ++q)
    launch_boundsm_tiled(st, T->t[q]->d, fields_of(T->t[q]), T->e[q], 2);
  for (les_handle* h : T->t) launch_calc_visc_adam(st, h->d, grid_of(h), fields_of(h), 1, do_adam);
}

// tile q's colour-split copy (its handle's coefficients, the tiled arrays)
SplitSOR tile_split(les_tiled* T, int q) {
  les_handle* h = T->t[q];
  SplitSOR sp = h->sp;
  float* base = T->split_mem + (long)q * T->split_per_tile;
  const long ns = T->split_ns;
  sp.p[0] = base;
  sp.p[1] = base + ns;
  sp.r[0] = base + 2 * ns;
  sp.r[1] = base + 3 * ns;
  sp.nj = T->jpl + 2;
  sp.jhalo = 1;
  sp.hw = T->e[q].w ? nullptr : base + 4 * ns;
  sp.he = T->e[q].e ? nullptr : base + 4 * ns + (long)T->jpl * T->cfg.kp;
  return sp;
}

SplitTiles split_tiles(les_tiled* T) {
  SplitTiles x{};
  x.ntiles = (int)T->t.size();
  for (int b = 0; b < T->sj; ++b)
    for (int a = 0; a < T->si; ++a) {
      const int q = a + T->si * b;
      const SplitSOR sp = tile_split(T, q);
      x.p[q][0] = sp.p[0];
      x.p[q][1] = sp.p[1];
      x.hw[q] = const_cast<float*>(sp.hw);
      x.he[q] = const_cast<float*>(sp.he);
      x.n[q] = a + T->si * ((b + 1) % T->sj);
      x.s[q] = a + T->si * ((b + T->sj - 1) % T->sj);
      x.w[q] = a > 0 ? q - 1 : -1;
      x.e[q] = a + 1 < T->si ? q + 1 : -1;
      x.nr = sp.nr;
      x.nj = sp.nj;
    }
  return x;
}

// press.f95:87-115 on the tiles' split copies: one face exchange of the split layout after the
// scatter, then the colour sweeps of every tile, each storing its updated face cells into the
// neighbours' halos as well; the k faces need no images (a cell reads itself there), and the
// gather leaves the interior of p for pav, adj and the boundary images
void tiled_sor_split(les_tiled* T, hipStream_t st) {
  const SplitTiles x = split_tiles(T);
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    launch_split_scatter(st, h->d, fields_of(h), tile_split(T, (int)q), h->sc, 0);
  }
  launch_split_exchange(st, T->t[0]->d, x);
  SorTiles sw{};
  sw.ntiles = (int)T->t.size();
  for (int q = 0; q < sw.ntiles; ++q) {
    sw.sp[q] = tile_split(T, q);
    sw.g[q] = grid_of(T->t[q]);
    sw.sc[q] = T->t[q]->sc;
    SplitPush& pu = sw.push[q];
    for (int c = 0; c < 2; ++c) {
      pu.n[c] = x.p[x.n[q]][c];
      pu.s[c] = x.p[x.s[q]][c];
    }
    pu.e = x.e[q] >= 0 ? x.hw[x.e[q]] : nullptr;
    pu.w = x.w[q] >= 0 ? x.he[x.w[q]] : nullptr;
  }
  for (int l = 1; l <= T->cfg.nmaxp; ++l)
    for (int nrd = 0; nrd <= 1; ++nrd) launch_sor_split_tiles(st, T->t[0]->d, sw, T->t[0]->ph, nrd);
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    launch_split_gather_adj(st, h->d, fields_of(h), tile_split(T, (int)q), h->sc, 0);
  }
}

void tiled_press(les_tiled* T, hipStream_t st) {
  tiled_exchange(T, st, {LES_F, LES_G}, 0);
  tiled_exchange(T, st, {LES_F, LES_G}, 1);
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    RedBuf rb{T->partials + (long)T->red_off[q] * NLIMB, 0};
    launch_rhs(st, h->d, grid_of(h), fields_of(h), h->ph, rb, h->sc, T->e[q], 0);
  }
  launch_rhsav_finalize(st, T->partials, T->red_total, T->t[0]->sc);
  tiled_bcast(T, st, offsetof(DevScalars, rhsav));
  for (size_t q = 1; q < T->t.size(); ++q) launch_sor_reset(st, T->t[q]->sc);
  for (les_handle* h : T->t) launch_rhs_adjust(st, h->d, fields_of(h), h->sc);
  bool split = T->split_mem && T->p_consistent;
  for (les_handle* h : T->t) split = split && h->cn1_separable;
  if (split) tiled_sor_split(T, st);
  for (int l = 1; !split && l <= T->cfg.nmaxp; ++l) {
    for (int nrd = 0; nrd <= 1; ++nrd) {
      for (les_handle* h : T->t)
        launch_sor_colour(st, h->d, grid_of(h), fields_of(h), h->ph, nrd, 0, h->rb, h->sc);
      for (size_t q = 0; q < T->t.size(); ++q) {
        les_handle* h = T->t[q];
        launch_boundp1(st, h->d, fields_of(h), h->sc, 1, T->e[q]);
      }
      tiled_exchange(T, st, {LES_P}, 0);
      tiled_exchange(T, st, {LES_P}, 1);
    }
    for (les_handle* h : T->t) launch_boundp2(st, h->d, fields_of(h), h->sc, 1);
  }
  for (size_t q = 0; q < T->t.size(); ++q) {
    les_handle* h = T->t[q];
    RedBuf rb{T->partials + (long)T->red_off[q] * NLIMB, 0};
    launch_pav(st, h->d, grid_of(h), fields_of(h), rb, h->sc, 0);
  }
  launch_pav_finalize(st, T->partials, T->red_total, T->t[0]->sc);
  tiled_bcast(T, st, offsetof(DevScalars, pav));
  for (les_handle* h : T->t) launch_adj(st, h->d, fields_of(h), h->sc, 0);
  tiled_boundp(T, st, 0);
}

int tiled_check(les_tiled* T) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tfail(T, LES_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return LES_OK;
}

}  // namespace

extern "C" int les_tiled_create(const les_config* cfg, les_tiled** out) {
  if (!cfg || !out) return LES_E_INVALID;
  *out = nullptr;
  const int si = cfg->sub_i, sj = cfg->sub_j;
  if (si < 1 || sj < 1 || si * sj > 8 || cfg->ip % si || cfg->jp % sj) return LES_E_INVALID;
  const int ipl = cfg->ip / si, jpl = cfg->jp / sj;
  // red-black colours: a tile's local (i+j+k) parity must be the domain's
  if ((si > 1 && ipl % 2) || (sj > 1 && jpl % 2) || ipl < 2 || jpl < 2) return LES_E_INVALID;
  if (cfg->precision != LES_PREC_FP32 || cfg->pjuge > 0.0f) return LES_E_INVALID;
  if (cfg->sum_order != LES_SUM_EXACT) return LES_E_INVALID;  // tiles combine exact partials
  les_tiled* T = new les_tiled();
  T->cfg = *cfg;
  T->si = si;
  T->sj = sj;
  T->ipl = ipl;
  T->jpl = jpl;
  les_config c = *cfg;
  c.ip = ipl;
  c.jp = jpl;
  c.sub_i = 1;
  c.sub_j = 1;
  for (int b = 0; b < sj; ++b)
    for (int a = 0; a < si; ++a) {
      les_handle* h = nullptr;
      const int r = les_create(&c, &h);
      if (r) {
        les_tiled_destroy(T);
        return r;
      }
      T->t.push_back(h);
      T->e.push_back(Edges{a == 0, a == si - 1, b == 0, b == sj - 1, 1});
      T->red_off.push_back(T->red_total);
      T->red_total += reduction_blocks(h->d);
    }
  if (hipMalloc(&T->partials, (size_t)T->red_total * NLIMB * sizeof(int64_t)) != hipSuccess) {
    les_tiled_destroy(T);
    return LES_E_NOMEM;
  }
  // the split sweep needs an even jp per tile (the colour across the periodic seam) and its
  // index arithmetic on the tile grid
  bool split = jpl % 2 == 0 && (int)T->t.size() <= SPLIT_TILES_MAX;
  for (les_handle* h : T->t) split = split && h->split_ok;
  if (split) {
    T->split_ns = (long)T->t[0]->sp.nr * (jpl + 2) * cfg->kp;
    T->split_per_tile = 4 * T->split_ns + 2L * jpl * cfg->kp;
    const size_t bytes = (size_t)T->split_per_tile * T->t.size() * sizeof(float);
    if (hipMalloc(&T->split_mem, bytes) != hipSuccess) {
      les_tiled_destroy(T);
      return LES_E_NOMEM;
    }
    // padding entries are read (never used) by whole-float4 loads
    if (hipMemset(T->split_mem, 0, bytes) != hipSuccess) {
      les_tiled_destroy(T);
      return LES_E_HIP;
    }
  }
  *out = T;
  return LES_OK;
}

extern "C" int les_tiled_destroy(les_tiled* T) {
  if (!T) return LES_E_INVALID;
  if (T->gexec) hipGraphExecDestroy(T->gexec);
  if (T->graph) hipGraphDestroy(T->graph);
  for (les_handle* h : T->t) les_destroy(h);
  if (T->partials) hipFree(T->partials);
  if (T->split_mem) hipFree(T->split_mem);
  delete T;
  return LES_OK;
}

extern "C" const char* les_tiled_last_error(const les_tiled* T) {
  if (!T) return "";
  if (!T->err.empty()) return T->err.c_str();
  for (les_handle* h : T->t)
    if (!h->err.empty()) return h->err.c_str();
  return "";
}

extern "C" int les_tiled_sync(les_tiled* T) {
  if (!T) return LES_E_INVALID;
  for (les_handle* h : T->t) {
    const int r = les_sync(h);
    if (r) return r;
  }
  return LES_OK;
}

// whole-domain array -> every tile's slice (halos included)
extern "C" int les_tiled_upload(les_tiled* T, int field, const float* host) {
  if (!T || !host || field < 0 || field >= LES_NFIELDS) return LES_E_INVALID;
  int glo[3], ghi[3], llo[3], lhi[3];
  const int64_t gn = les_field_bounds(field, T->cfg.ip, T->cfg.jp, T->cfg.kp, glo, ghi);
  les_field_bounds(field, T->ipl, T->jpl, T->cfg.kp, llo, lhi);
  if (gn <= 0) return LES_E_INVALID;
  const long g0 = ghi[0] - glo[0] + 1, g1 = ghi[1] - glo[1] + 1;
  for (int b = 0; b < T->sj; ++b)
    for (int a = 0; a < T->si; ++a) {
      les_handle* h = T->t[a + T->si * b];
      const int io = a * T->ipl, jo = b * T->jpl;
      std::vector<float> loc((size_t)h->cnt[field]);
      if (is_3d(field)) {
        const long l0 = lhi[0] - llo[0] + 1, l1 = lhi[1] - llo[1] + 1;
        for (int k = llo[2]; k <= lhi[2]; ++k)
          for (int j = llo[1]; j <= lhi[1]; ++j)
            for (int i = llo[0]; i <= lhi[0]; ++i)
              loc[(i - llo[0]) + l0 * ((j - llo[1]) + l1 * (k - llo[2]))] =
                  host[(i + io - glo[0]) + g0 * ((j + jo - glo[1]) + g1 * (k - glo[2]))];
      } else if (field == LES_WIND_PROFILE) {  // (1:jp, 1:kp)
        for (int k = 0; k < T->cfg.kp; ++k)
          for (int j = 0; j < T->jpl; ++j)
            loc[j + (size_t)T->jpl * k] = host[(j + jo) + (size_t)T->cfg.jp * k];
      } else {
        const int ax = tiled_axis(field), off = ax == 0 ? io : (ax == 1 ? jo : 0);
        for (int i = llo[0]; i <= lhi[0]; ++i) loc[i - llo[0]] = host[i + off - glo[0]];
      }
      const int r = les_upload(h, field, loc.data());
      if (r) return r;
    }
  if (field == LES_P) T->p_consistent = false;
  if (field == LES_DX1) T->gdx1.assign(host, host + gn);
  if (field == LES_DY1) T->gdy1.assign(host, host + gn);
  if (field == LES_DZN) T->gdzn.assign(host, host + gn);
  if (field == LES_DX1 || field == LES_DY1 || field == LES_DZN) return tiled_write_area(T);
  return LES_OK;
}

// every tile's interior, and the halos on the domain's faces, -> the whole-domain array
extern "C" int les_tiled_download(les_tiled* T, int field, float* host) {
  if (!T || !host || field < 0 || field >= LES_NFIELDS) return LES_E_INVALID;
  int glo[3], ghi[3], llo[3], lhi[3];
  les_field_bounds(field, T->cfg.ip, T->cfg.jp, T->cfg.kp, glo, ghi);
  les_field_bounds(field, T->ipl, T->jpl, T->cfg.kp, llo, lhi);
  const long g0 = ghi[0] - glo[0] + 1, g1 = ghi[1] - glo[1] + 1;
  int r = les_tiled_sync(T);
  if (r) return r;
  for (int b = 0; b < T->sj; ++b)
    for (int a = 0; a < T->si; ++a) {
      les_handle* h = T->t[a + T->si * b];
      const int io = a * T->ipl, jo = b * T->jpl;
      std::vector<float> loc((size_t)h->cnt[field]);
      if ((r = les_download(h, field, loc.data()))) return r;
      if (is_3d(field)) {
        const long l0 = lhi[0] - llo[0] + 1, l1 = lhi[1] - llo[1] + 1;
        for (int k = llo[2]; k <= lhi[2]; ++k)
          for (int j = llo[1]; j <= lhi[1]; ++j) {
            if ((j < 1 && b > 0) || (j > T->jpl && b < T->sj - 1)) continue;
            for (int i = llo[0]; i <= lhi[0]; ++i) {
              if ((i < 1 && a > 0) || (i > T->ipl && a < T->si - 1)) continue;
              host[(i + io - glo[0]) + g0 * ((j + jo - glo[1]) + g1 * (k - glo[2]))] =
                  loc[(i - llo[0]) + l0 * ((j - llo[1]) + l1 * (k - llo[2]))];
            }
          }
      } else if (field == LES_WIND_PROFILE) {
        for (int k = 0; k < T->cfg.kp; ++k)
          for (int j = 0; j < T->jpl; ++j)
            host[(j + jo) + (size_t)T->cfg.jp * k] = loc[j + (size_t)T->jpl * k];
      } else {
        const int ax = tiled_axis(field), off = ax == 0 ? io : (ax == 1 ? jo : 0);
        for (int i = llo[0]; i <= lhi[0]; ++i) host[i + off - glo[0]] = loc[i - llo[0]];
      }
    }
  return LES_OK;
}

extern "C" int les_tiled_ifdata(les_tiled* T) {
  if (!T) return LES_E_INVALID;
  hipStream_t st = T->t[0]->st;
  int r = les_tiled_sync(T);
  if (r) return r;
  // ifdata.f95:258-273 as les_ifdata runs it
  tiled_bondv1(T, st, 0, 1);
  tiled_boundp(T, st, 0);
  tiled_velfg_les(T, st, 0);
  return tiled_check(T);
}

extern "C" int les_tiled_step(les_tiled* T, int n) {
  if (!T) return LES_E_INVALID;
  hipStream_t st = T->t[0]->st;
  tiled_bondv1(T, st, 1, n < 1 ? 2 : n);
  tiled_velfg_les(T, st, 1);
  tiled_press(T, st);
  return tiled_check(T);
}

// init.f95:82-107 and les.f95:36 on every tile.  The coefficients are pointwise functions of
// the grid, which the tiles hold sliced with their halos, except delx1, which uses dx1(0) and
// dy1(0) of the whole domain.
extern "C" int les_tiled_init_coefficients(les_tiled* T) {
  if (!T) return LES_E_INVALID;
  if (T->gdx1.empty() || T->gdy1.empty() || T->gdzn.empty())
    return tfail(T, LES_E_STATE, "tiled_init_coefficients: upload dx1, dy1 and dzn first");
  int r;
  std::vector<float> delx1(T->cfg.kp);
  for (int k = 1; k <= T->cfg.kp; ++k)
    delx1[k - 1] = powf(T->gdx1[1] * T->gdy1[0] * T->gdzn[k + 1], 1.0f / 3.0f);
  for (les_handle* h : T->t) {
    if ((r = les_init_coefficients(h))) return r;
    if ((r = les_upload(h, LES_DELX1, delx1.data()))) return r;
  }
  return tiled_write_area(T);
}

// les_tiled_step for n = n0 .. n0+nsteps-1: steps with n <= 1 eagerly, the others as
// replays of one captured step (every launch of a step n >= 2 is the same).  A later upload
// does not change the captured launches (they take the tiles' arrays by address).
extern "C" int les_tiled_run(les_tiled* T, int n0, int nsteps) {
  if (!T || nsteps < 0) return LES_E_INVALID;
  int r;
  for (; nsteps > 0 && n0 <= 1; ++n0, --nsteps)
    if ((r = les_tiled_step(T, n0))) return r;
  if (nsteps == 0) return LES_OK;
  hipStream_t st = T->t[0]->st;
  if (!T->gexec) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return tfail(T, LES_E_HIP, "tiled_run: begin capture");
    r = les_tiled_step(T, 2);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &gr);
    if (r || e != hipSuccess) {
      if (gr) hipGraphDestroy(gr);
      return r ? r : tfail(T, LES_E_HIP, std::string("tiled_run: capture: ") + hipGetErrorString(e));
    }
    T->graph = gr;
    if (hipGraphInstantiate(&T->gexec, T->graph, nullptr, nullptr, 0) != hipSuccess)
      return tfail(T, LES_E_HIP, "tiled_run: graph instantiate");
  }
  for (int q = 0; q < nsteps; ++q)
    if (hipGraphLaunch(T->gexec, st) != hipSuccess) return tfail(T, LES_E_HIP, "tiled_run: launch");
  return LES_OK;
}
