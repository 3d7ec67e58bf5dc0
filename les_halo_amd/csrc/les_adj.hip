// The adjoint form of the LES step: forward and adjoint statement kernels (generated from the
// statement list in les_halo_amd/adjgen.py into les_adj_gen.inc) in FP32 or FP64, the
// reductions the step makes (uout, rhsav, pav) and their adjoints, and a window driver with
// timestep checkpointing: the gradient of an objective (the mean of u over probe points and
// steps) with respect to the inflow wind profile and a per-cell drag coefficient field.
//
// State carried from step to step: u, v, w, p, f, g, h, fold, gold, hold, usum, vsum, wsum.
// Every other array is written by the step before it is read (or never written: a constant).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <initializer_list>
#include <string>
#include <utility>
#include <vector>

#include "les_hip.h"
#include "les_kernels.h"

namespace les_adj_impl {

template <typename T>
struct AdjScal {
  T uout, rhsav, pav;
};
template <typename T>
struct AdjPar {
  T dt, ro, vn, alpha, beta, cs0, omega, area, pco;
};

// Scalar adjoints: the AdjScal<T> (uout, rhsav, pav) followed by ADJ_SLOTS partial sums per
// scalar.  A statement transpose adds its terms of scalar q as one atomic add per workgroup
// into slot (workgroup mod ADJ_SLOTS), so the adds of a domain-wide statement spread over
// many addresses; k_scal_fold adds the slots into the scalar before it is read.
constexpr int ADJ_SLOTS = 64;
constexpr int SB_COUNT = 3 + 3 * ADJ_SLOTS;  // scalars and slots (elements of T)
template <typename T>
__device__ __forceinline__ T* scal_slots(AdjScal<T>* SB, int q) {
  return reinterpret_cast<T*>(SB + 1) + q * ADJ_SLOTS;
}
// every thread of the workgroup calls it (uniform control flow)
template <typename T>
__device__ __forceinline__ void scal_add(AdjScal<T>* SB, int q, T v) {
  __shared__ T ws[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = T(0);
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += ws[w];
    if (s != T(0)) atomicAdd(scal_slots(SB, q) + blockIdx.x % ADJ_SLOTS, s);
  }
  __syncthreads();  // ws is reused by the next call
}
// one wavefront: SB[q] += the sum of its slots, slots zeroed
template <typename T>
__global__ __launch_bounds__(ADJ_SLOTS) void k_scal_fold(AdjScal<T>* SB, int q) {
  T* slot = scal_slots(SB, q);
  T v = slot[threadIdx.x];
  slot[threadIdx.x] = T(0);
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (threadIdx.x == 0) reinterpret_cast<T*>(SB)[q] += v;
}

#include "les_adj_gen.inc"
#include "les_adj_hand.h"

// ---- reductions of the step and their adjoints -----------------------------------------------
constexpr int RED_THREADS = 1024;

// bondv1_ext_wind_profile.f95:67-75 on the pre-outflow u: uout = (max(0, max u(im,:,:)) +
// min(0, min u(im,:,:))) / 2 (one workgroup)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_uout(AdjArr<T> A, AdjScal<T>* S, int im, int jm,
                                                       int km) {
  __shared__ T smax[RED_THREADS], smin[RED_THREADS];
  T mx = T(0), mn = T(0);
  for (int t = threadIdx.x; t < jm * km; t += RED_THREADS) {
    const int j = 1 + t % jm, k = 1 + t / jm;
    const T x = A.u_pre[(long)im + (long)(im + 2) * ((long)(j + 1) + (long)(jm + 3) * k)];
    mx = max(mx, x);
    mn = min(mn, x);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + o]);
      smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) S->uout = (smax[0] + smin[0]) / T(2);
}

// adjoint: the max (if positive) and the min (if negative) take half of uout's adjoint each; the
// element is the first (in k, then j) holding the extreme value
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_uout_bw(AdjArr<T> A, AdjArr<T> B, AdjScal<T>* SB,
                                                          int im, int jm, int km) {
  __shared__ T smax[RED_THREADS], smin[RED_THREADS];
  __shared__ int amax[RED_THREADS], amin[RED_THREADS];
  T mx = T(0), mn = T(0);
  int ax = -1, an = -1;
  for (int t = threadIdx.x; t < jm * km; t += RED_THREADS) {
    const int j = 1 + t % jm, k = 1 + t / jm;
    const T x = A.u_pre[(long)im + (long)(im + 2) * ((long)(j + 1) + (long)(jm + 3) * k)];
    if (x > mx) {
      mx = x;
      ax = t;
    }
    if (x < mn) {
      mn = x;
      an = t;
    }
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  amax[threadIdx.x] = ax;
  amin[threadIdx.x] = an;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const int q = threadIdx.x + o;
      if (smax[q] > smax[threadIdx.x] ||
          (smax[q] == smax[threadIdx.x] && amax[q] >= 0 && (amax[threadIdx.x] < 0 || amax[q] < amax[threadIdx.x]))) {
        smax[threadIdx.x] = smax[q];
        amax[threadIdx.x] = amax[q];
      }
      if (smin[q] < smin[threadIdx.x] ||
          (smin[q] == smin[threadIdx.x] && amin[q] >= 0 && (amin[threadIdx.x] < 0 || amin[q] < amin[threadIdx.x]))) {
        smin[threadIdx.x] = smin[q];
        amin[threadIdx.x] = amin[q];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const T half = SB->uout / T(2);
    auto at = [&](int t) {
      const int j = 1 + t % jm, k = 1 + t / jm;
      return (long)im + (long)(im + 2) * ((long)(j + 1) + (long)(jm + 3) * k);
    };
    if (amax[0] >= 0 && smax[0] > T(0)) B.u[at(amax[0])] += half;
    if (amin[0] >= 0 && smin[0] < T(0)) B.u[at(amin[0])] += half;
    SB->uout = T(0);
  }
}

// which: 0 rhsav = sum dx1*dy1*dzn*rhs / area (press.f95:67-75); 1 pav = sum p*dx1*dy1*dzn / pco
// (press.f95:125-134).  Per-block FP64 partials in a fixed order, then one workgroup adds them
// in a fixed order: deterministic.
template <typename T>
__device__ __forceinline__ T wsum_term(const AdjArr<T>& A, int which, int i, int j, int k, int im,
                                       int jm) {
  const T dx = A.dx1[i + 1], dy = A.dy1[j], dz = A.dzn[k + 1];
  if (which == 0) return dx * dy * dz * A.rhs[(long)i + (long)(im + 2) * ((long)j + (long)(jm + 2) * k)];
  return A.p[(long)i + (long)(im + 3) * ((long)j + (long)(jm + 3) * k)] * dx * dy * dz;
}

template <typename T>
__global__ __launch_bounds__(256) void k_wsum_part(AdjArr<T> A, int which, double* part, int im,
                                                   int jm, int km) {
  __shared__ double s[256];
  const long n = (long)im * jm * km;
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int i = 1 + (int)(t % im), j = 1 + (int)((t / im) % jm), k = 1 + (int)(t / ((long)im * jm));
    acc += (double)wsum_term(A, which, i, j, k, im, jm);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

template <typename T>
__global__ __launch_bounds__(256) void k_wsum_final(const double* part, int nb, int which,
                                                    AdjScal<T>* S, AdjPar<T> P) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int q = threadIdx.x; q < nb; q += 256) acc += part[q];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (which == 0)
      S->rhsav = (T)s[0] / P.area;
    else
      S->pav = (T)s[0] / P.pco;
  }
}

// adjoint: d(sum w*x / c)/dx = w / c
template <typename T>
__global__ __launch_bounds__(256) void k_wsum_bw(AdjArr<T> A, AdjArr<T> B, AdjScal<T>* SB, int which,
                                                 AdjPar<T> P, int im, int jm, int km) {
  const long n = (long)im * jm * km;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int i = 1 + (int)(t % im), j = 1 + (int)((t / im) % jm), k = 1 + (int)(t / ((long)im * jm));
  const T dx = A.dx1[i + 1], dy = A.dy1[j], dz = A.dzn[k + 1];
  if (which == 0)
    B.rhs[(long)i + (long)(im + 2) * ((long)j + (long)(jm + 2) * k)] += SB->rhsav * (dx * dy * dz) / P.area;
  else
    B.p[(long)i + (long)(im + 3) * ((long)j + (long)(jm + 3) * k)] += SB->pav * (dx * dy * dz) / P.pco;
}

template <typename T>
__global__ void k_zero_scal(AdjScal<T>* S) {
  S->uout = S->rhsav = S->pav = T(0);
}

// objective: acc += sum over probes of u at the probe (i, j, k triples)
template <typename T>
__global__ void k_probe_sum(AdjArr<T> A, const int* probes, int np, double* acc, int im, int jm) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int q = 0; q < np; ++q) {
    const int i = probes[3 * q], j = probes[3 * q + 1], k = probes[3 * q + 2];
    s += (double)A.u[(long)i + (long)(im + 2) * ((long)(j + 1) + (long)(jm + 3) * k)];
  }
  *acc += s;
}

template <typename T>
__global__ void k_probe_seed(AdjArr<T> B, const int* probes, int np, T wgt, int im, int jm) {
  if (threadIdx.x != 0) return;
  for (int q = 0; q < np; ++q) {
    const int i = probes[3 * q], j = probes[3 * q + 1], k = probes[3 * q + 2];
    B.u[(long)i + (long)(im + 2) * ((long)(j + 1) + (long)(jm + 3) * k)] += wgt;
  }
}

// Adjoint of one colour sweep of press.f95:91-98 (statement sor<nrd>) in gather form:
//   p(y) <- p(y) + omega*(cn1(y)*(cn2l p(y+ei) + cn2s p(y-ei) + cn3l p(y+ej) + cn3s p(y-ej)
//                                 + cn4l p(y+ek) + cn4s p(y-ek) - rhs(y)) - p(y))
// for the interior cells y of the colour.  Its transpose adds omega*cn1(y)*c*pb(y) to the
// adjoint of every neighbour x of such a y (c the coefficient through which y reads x), then
// scales pb(y) by 1 - omega and takes omega*cn1(y)*pb(y) off rhsb(y).  The first kernel gathers
// into every point x of p's box that is not a cell of the colour (each x reads its up to six
// neighbours of the colour), the second updates the colour's cells: no atomics, and the same
// arithmetic for every neighbour as the scatter form.
template <typename T>
__global__ __launch_bounds__(256) void k_sor_bw_gather(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P,
                                                       int nrd, int im, int jm, int km) {
  const int ni = im + 3, nj = jm + 3, nk = km + 2;  // p(0:im+2, 0:jm+2, 0:km+1)
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)ni * nj * nk) return;
  const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
  auto inside = [&](int a, int b, int c) { return a >= 1 && a <= im && b >= 1 && b <= jm && c >= 1 && c <= km; };
  auto mine = [&](int a, int b, int c) { return inside(a, b, c) && ((a + b + c + nrd) & 1) == 1; };
  if (mine(i, j, k)) return;
  auto P_ = [&](int a, int b, int c) { return (long)a + (long)ni * ((long)b + (long)nj * c); };
  auto C1 = [&](int a, int b, int c) {
    return A.cn1[(long)(a - 1) + (long)im * ((long)(b - 1) + (long)jm * (c - 1))];
  };
  T acc = T(0);
  // y = x - e_i reads x through cn2l(i_y); y = x + e_i through cn2s(i_y); likewise j, k
  if (mine(i - 1, j, k)) acc += P.omega * C1(i - 1, j, k) * A.cn2l[i - 2] * B.p[P_(i - 1, j, k)];
  if (mine(i + 1, j, k)) acc += P.omega * C1(i + 1, j, k) * A.cn2s[i] * B.p[P_(i + 1, j, k)];
  if (mine(i, j - 1, k)) acc += P.omega * C1(i, j - 1, k) * A.cn3l[j - 2] * B.p[P_(i, j - 1, k)];
  if (mine(i, j + 1, k)) acc += P.omega * C1(i, j + 1, k) * A.cn3s[j] * B.p[P_(i, j + 1, k)];
  if (mine(i, j, k - 1)) acc += P.omega * C1(i, j, k - 1) * A.cn4l[k - 2] * B.p[P_(i, j, k - 1)];
  if (mine(i, j, k + 1)) acc += P.omega * C1(i, j, k + 1) * A.cn4s[k] * B.p[P_(i, j, k + 1)];
  if (acc != T(0)) B.p[P_(i, j, k)] += acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sor_bw_scale(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P,
                                                      int nrd, int im, int jm, int km) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)im * jm * km) return;
  const int i = 1 + (int)(t % im), j = 1 + (int)((t / im) % jm), k = 1 + (int)(t / ((long)im * jm));
  if (((i + j + k + nrd) & 1) != 1) return;
  const long ip = (long)i + (long)(im + 3) * ((long)j + (long)(jm + 3) * k);
  const long ir = (long)i + (long)(im + 2) * ((long)j + (long)(jm + 2) * k);
  const T pb = B.p[ip];
  B.rhs[ir] -= P.omega * A.cn1[t] * pb;
  B.p[ip] = (T(1) - P.omega) * pb;
}

// boundp1 (boundp.f95:23-44) as one copy: its i-face then j-face loops leave every halo cell
// (a, b, c) of the ring a in {0, im+1} or b in {0, jm+1} (0 <= a <= im+1, 0 <= b <= jm+1, every
// level c of p) holding p(mi(a), pj(b), c) with mi(0) = 1, mi(im+1) = im, pj(0) = jm,
// pj(jm+1) = 1 (identity otherwise), sources that are never ring cells.  The transpose adds
// each ring cell's adjoint into its source's and zeroes it.  boundp2 (boundp.f95:5-21): the
// levels 0 and km+1 copy levels 1 and km.  These replace the six statement kernels of the two
// loops in the SOR iterations and after pav (same results).
__device__ __forceinline__ bool bp1_cell(long t, int im, int jm, int km, long& dst, long& src) {
  const int ring = 2 * (jm + 2) + 2 * im;
  if (t >= (long)ring * (km + 2)) return false;
  const int r = (int)(t % ring), c = (int)(t / ring);
  int a, b;
  if (r < 2 * (jm + 2)) {
    a = (r < jm + 2) ? 0 : im + 1;
    b = r % (jm + 2);
  } else {
    const int q = r - 2 * (jm + 2);
    b = (q < im) ? 0 : jm + 1;
    a = 1 + q % im;
  }
  const int sa = a == 0 ? 1 : (a == im + 1 ? im : a);
  const int sb = b == 0 ? jm : (b == jm + 1 ? 1 : b);
  const long ni = im + 3, nj = jm + 3;
  dst = a + ni * (b + nj * c);
  src = sa + ni * (sb + nj * c);
  return true;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bp1_fw(AdjArr<T> A, int im, int jm, int km) {
  long dst, src;
  if (bp1_cell((long)blockIdx.x * 256 + threadIdx.x, im, jm, km, dst, src)) A.p[dst] = A.p[src];
}

template <typename T>
__global__ __launch_bounds__(256) void k_bp1_bw(AdjArr<T> B, int im, int jm, int km) {
  long dst, src;
  if (!bp1_cell((long)blockIdx.x * 256 + threadIdx.x, im, jm, km, dst, src)) return;
  const T v = B.p[dst];
  B.p[dst] = T(0);
  if (v != T(0)) atomicAdd(&B.p[src], v);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bp2(AdjArr<T> A, AdjArr<T> B, int adjoint, int im, int jm,
                                             int km) {
  const long ni = im + 3, nj = jm + 3, plane = ni * nj;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * (long)(im + 2) * (jm + 2)) return;
  const long cell = t % ((long)(im + 2) * (jm + 2));
  const int a = (int)(cell % (im + 2)), b = (int)(cell / (im + 2));
  const bool top = t >= (long)(im + 2) * (jm + 2);
  const long dst = a + ni * b + plane * (top ? km + 1 : 0);
  const long src = a + ni * b + plane * (top ? km : 1);
  if (!adjoint) {
    A.p[dst] = A.p[src];
  } else {
    B.p[src] += B.p[dst];  // one ring cell per source: no other thread touches src
    B.p[dst] = T(0);
  }
}

// ops[q0 ..] holds the statements first, first + 1, .., last in this order
bool is_block(const int* ops, int q0, int first, int last) {
  for (int s = first; s <= last; ++s)
    if (ops[q0 + (s - first)] != s) return false;
  return true;
}

// the first statement of the hand-transposed point-wise block that ends at ops[q], or -1
int point_block(const int* ops, int q) {
  static const int blocks[4][2] = {{ADJ_velnw_u, ADJ_velnw_w},
                                   {ADJ_fb_usum, ADJ_fb_h},
                                   {ADJ_adam_save_f, ADJ_adam_old_h},
                                   {ADJ_rhs_uvw, ADJ_rhs_fgh}};
  for (const auto& b : blocks)
    if (ops[q] == b[1] && q >= b[1] - b[0] && is_block(ops, q - (b[1] - b[0]), b[0], b[1]))
      return b[0];
  return -1;
}

// the first statement of the hand-transposed boundary / scalar block that ends at ops[q], or -1
int bc_block(const int* ops, int q) {
  static const int blocks[5][2] = {{ADJ_inflow_u, ADJ_inflow_w},
                                   {ADJ_perj0_u, ADJ_top_w},
                                   {ADJ_bfg_f, ADJ_bfg_h1},
                                   {ADJ_rhs_adj, ADJ_rhs_adj},
                                   {ADJ_pav_adj, ADJ_pav_adj}};
  for (const auto& b : blocks)
    if (ops[q] == b[1] && q >= b[1] - b[0] && is_block(ops, q - (b[1] - b[0]), b[0], b[1]))
      return b[0];
  return -1;
}

bool is_seq(const int* ops, int n, int q, std::initializer_list<int> want) {
  if (q + (int)want.size() > n) return false;
  int z = q;
  for (int w : want)
    if (ops[z++] != w) return false;
  return true;
}

const int CARRIED[] = {ADJ_A_u, ADJ_A_v, ADJ_A_w, ADJ_A_p, ADJ_A_f, ADJ_A_g, ADJ_A_h, ADJ_A_fold,
                       ADJ_A_gold, ADJ_A_hold, ADJ_A_usum, ADJ_A_vsum, ADJ_A_wsum};
constexpr int NCARRIED = sizeof(CARRIED) / sizeof(CARRIED[0]);

}  // namespace les_adj_impl

using namespace les_adj_impl;

struct les_adj {
  les_config cfg;
  int fp64 = 0;
  int im = 0, jm = 0, km = 0;
  hipStream_t st = nullptr;
  std::string err;
  void* A[ADJ_NARRAYS] = {};  // forward arrays (T)
  void* B[ADJ_NARRAYS] = {};  // adjoints (T), active arrays only
  long n[ADJ_NARRAYS] = {};
  void* S = nullptr;   // AdjScal<T>
  void* SB = nullptr;  // adjoint scalars
  double* part = nullptr;
  int npart = 0;
  double* jacc = nullptr;
  int* probes = nullptr;
  int nprobes_cap = 0;
  double area = 0.0, pco = 0.0;
  // FP32: a plain handle on the same grid whose colour-split solve runs the forward's SOR
  // iterations (the same values as the statement sweeps when p's halos are mirror images on
  // entry; the transposed sweeps do not read the iterates, the solve being linear)
  les_handle* ph = nullptr;
  bool ph_tuned = false;
  bool sor_split = false;  // the last forward SOR ran through ph (its transpose does too)
  bool p_mirror = true;  // every halo of p equals its mirror interior cell
  bool use_slots = true;  // les_adj_window keeps per-step copies (les_adj_set_slots)
  // The recomputation of a step skips the SOR solve (its p comes from the forward window's
  // per-step copy instead): the transposes never read the SOR's iterates, the solve being
  // linear, and a recomputed step is bitwise the window's step.
  bool skip_sor = false;
  // the hand-written fused transposes replace the generated ones of the blocks they cover
  // (les_adj_set_transposes; bit 0 the vel2 + velFG block, bit 1 the les block)
  int hand = 3;
  // scratch of the hand-written les transpose: 4 arrays of im*jm*km and the weight tables (T)
  void* hs = nullptr;
};

namespace {

int afail(les_adj* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define ACHK(h, x)                                                                  \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return afail(h, LES_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

size_t tsize(const les_adj* h) { return h->fp64 ? sizeof(double) : sizeof(float); }

template <typename T>
AdjArr<T> arrs(les_adj* h, bool adj) {
  AdjArr<T> a{};
  for (int q = 0; q < ADJ_NARRAYS; ++q) adj_slot(a, q) = static_cast<T*>(adj ? h->B[q] : h->A[q]);
  return a;
}

template <typename T>
AdjPar<T> pars(const les_adj* h) {
  const les_config& c = h->cfg;
  return AdjPar<T>{(T)c.dt, (T)c.ro, (T)c.vn, (T)c.alpha, (T)c.beta, (T)c.cs0, (T)c.omega,
                   (T)h->area, (T)h->pco};
}

// field id of les_hip.h (or LES_ADJ_CD) -> adjoint array index
int array_of_field(int f) {
  static const char* names[LES_NFIELDS] = {
      "p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold", "rhs", "sm",
      "diu1", "diu2", "diu3", "diu4", "diu5", "diu6", "diu7", "diu8", "diu9", nullptr, "bmask1",
      "cmask1", "dmask1", "cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "dx1", "dxs", "dy1",
      "dys", "dzn", "dzs", nullptr, "delx1", "wind"};
  const char* nm = nullptr;
  if (f == LES_ADJ_CD) nm = "cd";
  else if (f >= 0 && f < LES_NFIELDS) nm = names[f];
  if (!nm) return -1;
  for (int q = 0; q < ADJ_NARRAYS; ++q)
    if (!strcmp(ADJ_ARRAY_NAMES[q], nm)) return q;
  return -1;
}

template <typename T>
void run_sum(les_adj* h, int which) {
  AdjArr<T> A = arrs<T>(h, false);
  hipLaunchKernelGGL(k_wsum_part<T>, dim3(h->npart), dim3(256), 0, h->st, A, which, h->part, h->im,
                     h->jm, h->km);
  hipLaunchKernelGGL(k_wsum_final<T>, dim3(1), dim3(256), 0, h->st, h->part, h->npart, which,
                     static_cast<AdjScal<T>*>(h->S), pars<T>(h));
}

// The forward's SOR iterations by the plain handle's split solve (FP32 with mirror-consistent
// p halos); false when it does not apply and the statement sweeps run instead.
template <typename T>
bool split_sor_fw(les_adj* h) {
  h->sor_split = false;
  if (sizeof(T) != sizeof(float) || !h->ph || !h->p_mirror || h->cfg.nmaxp < 1) return false;
  if (!h->ph_tuned) {
    les_tuning t{};
    if (les_autotune(h->ph, 5, &t) != LES_OK) return false;
    h->ph_tuned = true;
  }
  h->sor_split = les::sor_solve_external(h->ph, h->st, static_cast<float*>(h->A[ADJ_A_p]),
                                         static_cast<float*>(h->A[ADJ_A_rhs])) == LES_OK;
  return h->sor_split;
}

// Its transpose after a split forward (the statement transposes otherwise: on p with mirror
// halos both forwards are the same function, so either transpose is its adjoint).
template <typename T>
bool split_sor_bw(les_adj* h) {
  if (sizeof(T) != sizeof(float) || !h->sor_split || !h->B[ADJ_A_p] || !h->B[ADJ_A_rhs])
    return false;
  return les::sor_adjoint_external(h->ph, h->st, static_cast<float*>(h->B[ADJ_A_p]),
                                   static_cast<float*>(h->B[ADJ_A_rhs])) == LES_OK;
}

template <typename T>
void run_ops_fw(les_adj* h, const int* ops, int nops) {
  const AdjArr<T> A = arrs<T>(h, false);
  const AdjPar<T> P = pars<T>(h);
  AdjScal<T>* S = static_cast<AdjScal<T>*>(h->S);
  const long nbp1 = (long)(2 * (h->jm + 2) + 2 * h->im) * (h->km + 2);
  const long nbp2 = 2L * (h->im + 2) * (h->jm + 2);
  for (int q = 0; q < nops; ++q) {
    const int op = ops[q];
    if (is_seq(ops, nops, q, {ADJ_bp1_i0, ADJ_bp1_i1, ADJ_bp1_j0, ADJ_bp1_j1})) {
      hipLaunchKernelGGL(k_bp1_fw<T>, dim3((unsigned)((nbp1 + 255) / 256)), dim3(256), 0, h->st, A,
                         h->im, h->jm, h->km);
      q += 3;
    } else if (is_seq(ops, nops, q, {ADJ_bp2_k0, ADJ_bp2_k1})) {
      hipLaunchKernelGGL(k_bp2<T>, dim3((unsigned)((nbp2 + 255) / 256)), dim3(256), 0, h->st, A, A,
                         0, h->im, h->jm, h->km);
      q += 1;
    } else if (op >= 0) {
      adj_launch_fw<T>(op, h->st, A, S, P, h->im, h->jm, h->km);
    } else if (op == ADJ_OP_SNAP) {
      for (const char* nm : {"u", "v", "w"}) {
        int a = -1, b = -1;
        for (int z = 0; z < ADJ_NARRAYS; ++z) {
          if (!strcmp(ADJ_ARRAY_NAMES[z], nm)) a = z;
          if (!strcmp(ADJ_ARRAY_NAMES[z], (std::string(nm) + "_pre").c_str())) b = z;
        }
        hipMemcpyAsync(h->A[b], h->A[a], h->n[a] * tsize(h), hipMemcpyDeviceToDevice, h->st);
      }
    } else if (op == ADJ_OP_UOUT) {
      hipLaunchKernelGGL(k_uout<T>, dim3(1), dim3(RED_THREADS), 0, h->st, A, S, h->im, h->jm, h->km);
    } else if (op == ADJ_OP_RHSAV) {
      run_sum<T>(h, 0);
    } else if (op == ADJ_OP_PAV) {
      run_sum<T>(h, 1);
    } else if (op == ADJ_OP_SOR) {
      if (h->skip_sor)  // the transpose tries the split form (statement sweeps if it fails)
        h->sor_split = sizeof(T) == sizeof(float) && h->ph && h->p_mirror && h->cfg.nmaxp >= 1;
      else if (!split_sor_fw<T>(h))
        for (int l = 0; l < h->cfg.nmaxp; ++l) run_ops_fw<T>(h, ADJ_SOR_OPS, ADJ_SOR_NOPS);
    }
  }
}

template <typename T>
void run_ops_bw(les_adj* h, const int* ops, int nops) {
  const AdjArr<T> A = arrs<T>(h, false), B = arrs<T>(h, true);
  const AdjPar<T> P = pars<T>(h);
  AdjScal<T>* S = static_cast<AdjScal<T>*>(h->S);
  AdjScal<T>* SB = static_cast<AdjScal<T>*>(h->SB);
  const long ncell = (long)h->im * h->jm * h->km;
  const long nbp1 = (long)(2 * (h->jm + 2) + 2 * h->im) * (h->km + 2);
  const long nbp2 = 2L * (h->im + 2) * (h->jm + 2);
  for (int q = nops - 1; q >= 0; --q) {
    const int op = ops[q];
    if ((h->hand & 1) && op == VEL_BLOCK_LAST && q >= VEL_BLOCK_LAST - VEL_BLOCK_FIRST &&
        is_block(ops, q - (VEL_BLOCK_LAST - VEL_BLOCK_FIRST), VEL_BLOCK_FIRST, VEL_BLOCK_LAST)) {
      launch_vel_block_bw<T>(h->st, A, B, P, h->im, h->jm, h->km);
      q -= VEL_BLOCK_LAST - VEL_BLOCK_FIRST;
    } else if ((h->hand & 2) && bc_block(ops, q) >= 0) {
      const int first = bc_block(ops, q);
      launch_bc_block_bw<T>(first, h->st, B, SB, h->im, h->jm, h->km);
      q -= op - first;
    } else if ((h->hand & 2) && point_block(ops, q) >= 0) {
      const int first = point_block(ops, q);
      launch_point_block_bw<T>(first, h->st, A, B, P, h->im, h->jm, h->km);
      q -= op - first;
    } else if ((h->hand & 2) && op == LES_BLOCK_LAST && q >= LES_BLOCK_LAST - LES_BLOCK_FIRST &&
               is_block(ops, q - (LES_BLOCK_LAST - LES_BLOCK_FIRST), LES_BLOCK_FIRST,
                        LES_BLOCK_LAST)) {
      launch_les_block_bw<T>(h->st, A, B, P, static_cast<T*>(h->hs), h->im, h->jm, h->km);
      q -= LES_BLOCK_LAST - LES_BLOCK_FIRST;
    } else if (q >= 3 && is_seq(ops, nops, q - 3, {ADJ_bp1_i0, ADJ_bp1_i1, ADJ_bp1_j0, ADJ_bp1_j1})) {
      hipLaunchKernelGGL(k_bp1_bw<T>, dim3((unsigned)((nbp1 + 255) / 256)), dim3(256), 0, h->st, B,
                         h->im, h->jm, h->km);
      q -= 3;
    } else if (q >= 1 && is_seq(ops, nops, q - 1, {ADJ_bp2_k0, ADJ_bp2_k1})) {
      hipLaunchKernelGGL(k_bp2<T>, dim3((unsigned)((nbp2 + 255) / 256)), dim3(256), 0, h->st, A, B,
                         1, h->im, h->jm, h->km);
      q -= 1;
    } else if (op == ADJ_sor0 || op == ADJ_sor1) {
      const int nrd = op == ADJ_sor0 ? 0 : 1;
      const long nbox = (long)(h->im + 3) * (h->jm + 3) * (h->km + 2);
      hipLaunchKernelGGL(k_sor_bw_gather<T>, dim3((unsigned)((nbox + 255) / 256)), dim3(256), 0,
                         h->st, A, B, P, nrd, h->im, h->jm, h->km);
      hipLaunchKernelGGL(k_sor_bw_scale<T>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                         h->st, A, B, P, nrd, h->im, h->jm, h->km);
    } else if (op >= 0) {
      adj_launch_bw<T>(op, h->st, A, B, S, SB, P, h->im, h->jm, h->km);
    } else if (op == ADJ_OP_UOUT) {
      hipLaunchKernelGGL(k_scal_fold<T>, dim3(1), dim3(ADJ_SLOTS), 0, h->st, SB, 0);
      hipLaunchKernelGGL(k_uout_bw<T>, dim3(1), dim3(RED_THREADS), 0, h->st, A, B, SB, h->im, h->jm,
                         h->km);
    } else if (op == ADJ_OP_RHSAV || op == ADJ_OP_PAV) {
      const int which = op == ADJ_OP_RHSAV ? 0 : 1;
      hipLaunchKernelGGL(k_scal_fold<T>, dim3(1), dim3(ADJ_SLOTS), 0, h->st, SB, 1 + which);
      hipLaunchKernelGGL(k_wsum_bw<T>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, h->st,
                         A, B, SB, which, P, h->im, h->jm, h->km);
      hipLaunchKernelGGL(k_zero_scal<T>, dim3(1), dim3(1), 0, h->st, SB);
    } else if (op == ADJ_OP_SOR) {
      if (!split_sor_bw<T>(h))
        for (int l = 0; l < h->cfg.nmaxp; ++l) run_ops_bw<T>(h, ADJ_SOR_OPS, ADJ_SOR_NOPS);
    }
  }
}

void step_fw(les_adj* h) {
  if (h->fp64)
    run_ops_fw<double>(h, ADJ_STEP_OPS, ADJ_STEP_NOPS);
  else
    run_ops_fw<float>(h, ADJ_STEP_OPS, ADJ_STEP_NOPS);
}

void step_bw(les_adj* h) {
  if (h->fp64)
    run_ops_bw<double>(h, ADJ_STEP_OPS, ADJ_STEP_NOPS);
  else
    run_ops_bw<float>(h, ADJ_STEP_OPS, ADJ_STEP_NOPS);
}

bool is_carried(int a) {
  for (int q = 0; q < NCARRIED; ++q)
    if (CARRIED[q] == a) return true;
  return false;
}

// Up to ZERO_MAX device arrays cleared in one launch (4-byte words; blockIdx.y: the array).
constexpr int ZERO_MAX = 64;
struct ZeroList {
  uint32_t* dst[ZERO_MAX];
  long words[ZERO_MAX];
};
__global__ __launch_bounds__(256) void k_zero_arrays(ZeroList L) {
  const int a = blockIdx.y;
  uint32_t* __restrict__ d = L.dst[a];
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < L.words[a]; t += (long)gridDim.x * 256)
    d[t] = 0u;
}

// the adjoints of everything but the carried state and the controls (wind, cd) start each
// reversed step at zero (a fill kernel over the list of arrays, not one memset per array)
int zero_temporaries(les_adj* h) {
  ZeroList L{};
  int m = 0;
  auto flush = [&]() {
    if (m) hipLaunchKernelGGL(k_zero_arrays, dim3(256, m), dim3(256), 0, h->st, L);
    m = 0;
  };
  auto add = [&](void* p, size_t bytes) {
    L.dst[m] = static_cast<uint32_t*>(p);
    L.words[m++] = (long)(bytes / 4);
    if (m == ZERO_MAX) flush();
  };
  for (int q = 0; q < ADJ_NARRAYS; ++q)
    if (h->B[q] && !is_carried(q) && q != ADJ_A_wind && q != ADJ_A_cd) add(h->B[q], h->n[q] * tsize(h));
  add(h->SB, SB_COUNT * tsize(h));
  flush();
  ACHK(h, hipGetLastError());
  return LES_OK;
}

// Up to COPY_MAX device arrays copied in one launch (4-byte words; blockIdx.y: the array).
constexpr int COPY_MAX = 16;
struct CopyList {
  const uint32_t* src[COPY_MAX];
  uint32_t* dst[COPY_MAX];
  long words[COPY_MAX];
};
__global__ __launch_bounds__(256) void k_copy_arrays(CopyList L) {
  const int a = blockIdx.y;
  const uint32_t* __restrict__ s = L.src[a];
  uint32_t* __restrict__ d = L.dst[a];
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < L.words[a]; t += (long)gridDim.x * 256)
    d[t] = s[t];
}

// Copies n arrays (bytes[q] each, multiples of 4) on the handle's stream, COPY_MAX per launch.
int copy_arrays(les_adj* h, int n, void* const* dst, const void* const* src, const size_t* bytes) {
  for (int q0 = 0; q0 < n; q0 += COPY_MAX) {
    CopyList L{};
    const int m = std::min(COPY_MAX, n - q0);
    for (int q = 0; q < m; ++q) {
      L.src[q] = static_cast<const uint32_t*>(src[q0 + q]);
      L.dst[q] = static_cast<uint32_t*>(dst[q0 + q]);
      L.words[q] = (long)(bytes[q0 + q] / 4);
    }
    hipLaunchKernelGGL(k_copy_arrays, dim3(1024, m), dim3(256), 0, h->st, L);
  }
  ACHK(h, hipGetLastError());
  return LES_OK;
}

// the carried state to / from a packed buffer
int move_state(les_adj* h, char* buf, bool save) {
  void* dst[NCARRIED];
  const void* src[NCARRIED];
  size_t bytes[NCARRIED];
  for (int q = 0; q < NCARRIED; ++q) {
    const int a = CARRIED[q];
    bytes[q] = h->n[a] * tsize(h);
    dst[q] = save ? (void*)buf : h->A[a];
    src[q] = save ? (const void*)h->A[a] : (const void*)buf;
    buf += bytes[q];
  }
  return copy_arrays(h, NCARRIED, dst, src, bytes);
}
int save_state(les_adj* h, char* dst) { return move_state(h, dst, true); }
int load_state(les_adj* h, const char* src) { return move_state(h, const_cast<char*>(src), false); }

size_t state_bytes(const les_adj* h) {
  size_t b = 0;
  for (int q = 0; q < NCARRIED; ++q) b += h->n[CARRIED[q]] * tsize(h);
  return b;
}

int check(les_adj* h) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return afail(h, LES_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return LES_OK;
}

int set_probes(les_adj* h, const int* probes, int np) {
  if (np > h->nprobes_cap) {
    if (h->probes) hipFree(h->probes);
    h->probes = nullptr;
    ACHK(h, hipMalloc(&h->probes, 3 * sizeof(int) * np));
    h->nprobes_cap = np;
  }
  for (int q = 0; q < np; ++q) {
    const int i = probes[3 * q], j = probes[3 * q + 1], k = probes[3 * q + 2];
    if (i < 1 || i > h->im || j < 1 || j > h->jm || k < 1 || k > h->km)
      return afail(h, LES_E_INVALID, "probe outside the interior");
  }
  ACHK(h, hipMemcpyAsync(h->probes, probes, 3 * sizeof(int) * np, hipMemcpyHostToDevice, h->st));
  return LES_OK;
}

void seed_probes(les_adj* h, int np, double wgt) {
  if (h->fp64)
    hipLaunchKernelGGL(k_probe_seed<double>, dim3(1), dim3(64), 0, h->st, arrs<double>(h, true),
                       h->probes, np, wgt, h->im, h->jm);
  else
    hipLaunchKernelGGL(k_probe_seed<float>, dim3(1), dim3(64), 0, h->st, arrs<float>(h, true),
                       h->probes, np, (float)wgt, h->im, h->jm);
}

void sum_probes(les_adj* h, int np) {
  if (h->fp64)
    hipLaunchKernelGGL(k_probe_sum<double>, dim3(1), dim3(64), 0, h->st, arrs<double>(h, false),
                       h->probes, np, h->jacc, h->im, h->jm);
  else
    hipLaunchKernelGGL(k_probe_sum<float>, dim3(1), dim3(64), 0, h->st, arrs<float>(h, false),
                       h->probes, np, h->jacc, h->im, h->jm);
}

}  // namespace

extern "C" int les_adj_create(const les_config* cfg, int fp64, les_adj** out) {
  if (!cfg || !out || cfg->ip < 2 || cfg->jp < 2 || cfg->kp < 2 || cfg->nmaxp < 0) return LES_E_INVALID;
  if (cfg->sum_order != LES_SUM_EXACT) return LES_E_INVALID;  // the transposes assume exact sums
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LES_E_NODEV;
  les_adj* h = new les_adj();
  h->cfg = *cfg;
  h->fp64 = fp64 ? 1 : 0;
  h->im = cfg->ip;
  h->jm = cfg->jp;
  h->km = cfg->kp;
  auto bail = [&](int code) {
    les_adj_destroy(h);
    return code;
  };
  if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreate(&h->st) != hipSuccess)
    return bail(LES_E_HIP);
  for (int q = 0; q < ADJ_NARRAYS; ++q) {
    h->n[q] = adj_array_size(q, h->im, h->jm, h->km);
    const size_t bytes = h->n[q] * tsize(h);
    if (hipMalloc(&h->A[q], bytes) != hipSuccess) return bail(LES_E_NOMEM);
    if (hipMemset(h->A[q], 0, bytes) != hipSuccess) return bail(LES_E_HIP);
    if (!ADJ_PASSIVE[q]) {
      if (hipMalloc(&h->B[q], bytes) != hipSuccess) return bail(LES_E_NOMEM);
      if (hipMemset(h->B[q], 0, bytes) != hipSuccess) return bail(LES_E_HIP);
    }
  }
  // the drag coefficient of the reference's feedback (beta * u * mask) is 1
  {
    const int a = ADJ_A_cd;
    std::vector<double> one((size_t)h->n[a], 1.0);
    les_adj* hh = h;
    const int r = les_adj_upload(hh, LES_ADJ_CD, one.data());
    if (r) return bail(r);
  }
  h->npart = 512;
  if (hipMalloc(&h->hs, (4 * (size_t)h->im * h->jm * h->km + 5 * (size_t)(h->im + h->jm + h->km + 6)) *
                           tsize(h)) != hipSuccess)
    return bail(LES_E_NOMEM);
  if (hipMalloc(&h->S, 3 * sizeof(double)) != hipSuccess || hipMalloc(&h->SB, SB_COUNT * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->part, h->npart * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->jacc, sizeof(double)) != hipSuccess)
    return bail(LES_E_NOMEM);
  if (hipMemset(h->S, 0, 3 * sizeof(double)) != hipSuccess || hipMemset(h->SB, 0, SB_COUNT * sizeof(double)) != hipSuccess)
    return bail(LES_E_HIP);
  if (!h->fp64) {  // without it (les_create failed) the statement kernels run everything
    les_config pc = *cfg;
    pc.pjuge = 0.0f;  // the statement SOR runs nmaxp iterations, no early exit
    les_handle* ph = nullptr;
    if (les_create(&pc, &ph) == LES_OK) h->ph = ph;
  }
  *out = h;
  return LES_OK;
}

extern "C" int les_adj_destroy(les_adj* h) {
  if (!h) return LES_E_INVALID;
  if (h->st) hipStreamSynchronize(h->st);
  if (h->ph) les_destroy(h->ph);
  for (int q = 0; q < ADJ_NARRAYS; ++q) {
    if (h->A[q]) hipFree(h->A[q]);
    if (h->B[q]) hipFree(h->B[q]);
  }
  for (void* p : {h->S, h->SB, (void*)h->part, (void*)h->jacc, (void*)h->probes, h->hs})
    if (p) hipFree(p);
  if (h->st) hipStreamDestroy(h->st);
  delete h;
  return LES_OK;
}

extern "C" const char* les_adj_last_error(const les_adj* h) { return h ? h->err.c_str() : ""; }

// every halo cell of p(0:ip+2, 0:jp+2, 0:kp+1) in the box (0:ip+1, 0:jp+1, 0:kp+1) equals its
// mirror interior cell (i: 0 -> 1, ip+1 -> ip; j periodic: 0 -> jp, jp+1 -> 1; k: 0 -> 1,
// kp+1 -> kp), bit for bit
static bool p_halos_mirror(const les_adj* h, const std::vector<float>& p) {
  const int ip = h->im, jp = h->jm, kp = h->km;
  auto at = [&](int i, int j, int k) { return p[i + (size_t)(ip + 3) * (j + (size_t)(jp + 3) * k)]; };
  for (int k = 0; k <= kp + 1; ++k)
    for (int j = 0; j <= jp + 1; ++j)
      for (int i = 0; i <= ip + 1; ++i) {
        const bool halo = i == 0 || i == ip + 1 || j == 0 || j == jp + 1 || k == 0 || k == kp + 1;
        if (!halo) continue;
        const int mi = i == 0 ? 1 : (i == ip + 1 ? ip : i);
        const int mj = j == 0 ? jp : (j == jp + 1 ? 1 : j);
        const int mk = k == 0 ? 1 : (k == kp + 1 ? kp : k);
        const float a = at(i, j, k), b = at(mi, mj, mk);
        if (std::memcmp(&a, &b, sizeof(float)) != 0) return false;
      }
  return true;
}

// the fields a plain handle holds (every input of the step but the drag field)
static bool plain_field(int field) { return field >= 0 && field < LES_NFIELDS; }

static int xfer(les_adj* h, int field, double* host, bool up, bool adj) {
  if (!h || !host) return LES_E_INVALID;
  const int a = array_of_field(field);
  if (a < 0) return afail(h, LES_E_INVALID, "field not held by the adjoint form");
  void* dev = adj ? h->B[a] : h->A[a];
  if (!dev) return afail(h, LES_E_INVALID, "no adjoint is kept for this field");
  const long n = h->n[a];
  ACHK(h, hipStreamSynchronize(h->st));
  if (h->fp64) {
    if (up)
      ACHK(h, hipMemcpy(dev, host, n * sizeof(double), hipMemcpyHostToDevice));
    else
      ACHK(h, hipMemcpy(host, dev, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp((size_t)n);
    if (up) {
      for (long q = 0; q < n; ++q) tmp[q] = (float)host[q];
      ACHK(h, hipMemcpy(dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
      if (!adj && field == LES_P) h->p_mirror = p_halos_mirror(h, tmp);
      if (!adj && h->ph && plain_field(field)) {
        h->ph_tuned = false;
        if (les_upload(h->ph, field, tmp.data()) != LES_OK) {  // fall back to the statements
          les_destroy(h->ph);
          h->ph = nullptr;
        }
      }
    } else {
      ACHK(h, hipMemcpy(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost));
      for (long q = 0; q < n; ++q) host[q] = tmp[q];
    }
  }
  return LES_OK;
}

extern "C" int les_adj_upload(les_adj* h, int field, const double* host) {
  return xfer(h, field, const_cast<double*>(host), true, false);
}
extern "C" int les_adj_download(les_adj* h, int field, double* host) {
  return xfer(h, field, host, false, false);
}
extern "C" int les_adj_upload_adjoint(les_adj* h, int field, const double* host) {
  return xfer(h, field, const_cast<double*>(host), true, true);
}
extern "C" int les_adj_download_adjoint(les_adj* h, int field, double* host) {
  return xfer(h, field, host, false, true);
}

extern "C" int les_adj_set_slots(les_adj* h, int enable) {
  if (!h) return LES_E_INVALID;
  h->use_slots = enable != 0;
  return LES_OK;
}

extern "C" int les_adj_set_sums(les_adj* h, double area, double pco) {
  if (!h) return LES_E_INVALID;
  h->area = area;
  h->pco = pco;
  return LES_OK;
}

extern "C" int les_adj_step(les_adj* h, int nsteps) {
  if (!h || nsteps < 0) return LES_E_INVALID;
  if (h->area <= 0.0 || h->pco <= 0.0) return afail(h, LES_E_STATE, "set the volume sums first");
  for (int s = 0; s < nsteps; ++s) step_fw(h);
  return check(h);
}

extern "C" int les_adj_reverse_step(les_adj* h) {
  if (!h) return LES_E_INVALID;
  if (h->area <= 0.0 || h->pco <= 0.0) return afail(h, LES_E_STATE, "set the volume sums first");
  int r;
  step_fw(h);
  if ((r = zero_temporaries(h))) return r;
  step_bw(h);
  return check(h);
}

extern "C" int les_adj_set_transposes(les_adj* h, int hand) {
  if (!h) return LES_E_INVALID;
  h->hand = hand ? 3 : 0;
  return LES_OK;
}

extern "C" int les_adj_zero_adjoints(les_adj* h) {
  if (!h) return LES_E_INVALID;
  for (int q = 0; q < ADJ_NARRAYS; ++q)
    if (h->B[q]) ACHK(h, hipMemsetAsync(h->B[q], 0, h->n[q] * tsize(h), h->st));
  ACHK(h, hipMemsetAsync(h->SB, 0, SB_COUNT * tsize(h), h->st));
  return check(h);
}

extern "C" int les_adj_window(les_adj* h, int nsteps, int ckpt, const int* probes, int nprobes,
                              double* objective) {
  if (!h || nsteps < 1 || ckpt < 1 || !probes || nprobes < 1 || !objective) return LES_E_INVALID;
  if (h->area <= 0.0 || h->pco <= 0.0) return afail(h, LES_E_STATE, "set the volume sums first");
  int r;
  if ((r = set_probes(h, probes, nprobes))) return r;
  const size_t sb = state_bytes(h);
  const int nck = (nsteps + ckpt - 1) / ckpt;
  char *ck = nullptr, *seg = nullptr, *fin = nullptr, *pst = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(h->st);
    if (ck) hipFree(ck);
    if (seg) hipFree(seg);
    if (fin) hipFree(fin);
    if (pst) hipFree(pst);
  };
  if (hipMalloc(&ck, sb * nck) != hipSuccess || hipMalloc(&fin, sb) != hipSuccess) {
    cleanup();
    return afail(h, LES_E_NOMEM, "window: checkpoint memory");
  }
  auto hipfail = [&](hipError_t e, const char* what) {
    cleanup();
    return afail(h, LES_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMemsetAsync(h->jacc, 0, sizeof(double), h->st)) != hipSuccess)
    return hipfail(e, "window: clear the objective");
  // p after every step of the forward window (8 MB per step at 150x150x90): the recomputed
  // steps take it instead of solving again.  Without the memory they solve.
  const size_t pbytes = h->n[ADJ_A_p] * tsize(h);
  if (hipMalloc(&pst, pbytes * nsteps) != hipSuccess) {
    pst = nullptr;
    (void)hipGetLastError();
  }
  auto keep_p = [&](int s) -> int {
    if (!pst) return LES_OK;
    void* dst = pst + pbytes * s;
    const void* src = h->A[ADJ_A_p];
    return copy_arrays(h, 1, &dst, &src, &pbytes);
  };
  // step s again from its state, p from the forward window's copy when there is one
  auto recompute = [&](int s) -> int {
    if (!pst) {
      step_fw(h);
      return LES_OK;
    }
    h->skip_sor = true;
    step_fw(h);
    h->skip_sor = false;
    void* dst = h->A[ADJ_A_p];
    const void* src = pst + pbytes * s;
    return copy_arrays(h, 1, &dst, &src, &pbytes);
  };
  // forward window: a checkpoint every ckpt steps, the objective after every step.  In FP32 the
  // plain handle runs it (its fused kernels, feedbf's beta term times the drag field: the same
  // values as the statement kernels) when it holds the same volume sums; the carried arrays are
  // its fields meanwhile.
  bool plain = false;
  if (h->ph) {
    les_scalars ps{};
    plain = les_get_scalars(h->ph, &ps) == LES_OK && ps.area == (float)h->area &&
            ps.pco == (float)h->pco;
  }
  void* Ac[NCARRIED];
  if (plain) {
    for (int q = 0; q < NCARRIED && plain; ++q) {
      const int a = CARRIED[q];
      int fid = -1;
      for (int f = 0; f < LES_NFIELDS; ++f)
        if (array_of_field(f) == a) fid = f;
      Ac[q] = fid >= 0 ? les::field_ptr(h->ph, fid) : nullptr;
      if (!Ac[q]) plain = false;
    }
  }
  if (plain) {
    {
      void* dst[NCARRIED];
      const void* src[NCARRIED];
      size_t bytes[NCARRIED];
      for (int q = 0; q < NCARRIED; ++q) {
        dst[q] = Ac[q];
        src[q] = h->A[CARRIED[q]];
        bytes[q] = h->n[CARRIED[q]] * tsize(h);
      }
      if ((r = copy_arrays(h, NCARRIED, dst, src, bytes))) return cleanup(), r;
    }
    if (!h->ph_tuned) {  // the fused SOR iteration's tile for this grid
      les_tuning t{};
      if (les_autotune(h->ph, 5, &t) == LES_OK) h->ph_tuned = true;
    }
    les::set_drag(h->ph, static_cast<const float*>(h->A[ADJ_A_cd]));
    les::set_p_consistent(h->ph, h->p_mirror);
    void* Aown[NCARRIED];
    for (int q = 0; q < NCARRIED; ++q) {
      Aown[q] = h->A[CARRIED[q]];
      h->A[CARRIED[q]] = Ac[q];
    }
    auto own_arrays = [&]() {
      for (int q = 0; q < NCARRIED; ++q) h->A[CARRIED[q]] = Aown[q];
      les::set_drag(h->ph, nullptr);
    };
    for (int s = 0; s < nsteps; ++s) {
      if (s % ckpt == 0 && (r = save_state(h, ck + sb * (s / ckpt)))) return own_arrays(), cleanup(), r;
      if ((r = les::step_external(h->ph, h->st, 2)))
        return own_arrays(), cleanup(), afail(h, r, "window: plain step");
      if ((r = keep_p(s))) return own_arrays(), cleanup(), r;
      sum_probes(h, nprobes);
    }
    r = save_state(h, fin);
    own_arrays();
    if (r) return cleanup(), r;
  } else {
    for (int s = 0; s < nsteps; ++s) {
      if (s % ckpt == 0 && (r = save_state(h, ck + sb * (s / ckpt)))) return cleanup(), r;
      step_fw(h);
      if ((r = keep_p(s))) return cleanup(), r;
      sum_probes(h, nprobes);
    }
    if ((r = save_state(h, fin))) return cleanup(), r;
  }
  double jsum = 0.0;
  if ((e = hipMemcpyAsync(&jsum, h->jacc, sizeof(double), hipMemcpyDeviceToHost, h->st)) != hipSuccess ||
      (e = hipStreamSynchronize(h->st)) != hipSuccess)
    return hipfail(e, "window: read the objective");
  // reverse: each segment recomputed from its checkpoint.  With per-step slots (every array the
  // step writes and its scalars uout, rhsav, pav, once per step of a segment) the recomputed
  // steps keep their values and each transposed step reads its own; without them (no memory for the slots) each step is
  // recomputed once more from its saved state before its transpose.
  if ((r = les_adj_zero_adjoints(h))) return cleanup(), r;
  const double wgt = 1.0 / ((double)nprobes * nsteps);
  const int nseg = std::min(ckpt, nsteps);
  size_t off[ADJ_NARRAYS] = {}, slot_bytes = 0;
  for (int q = 0; q < ADJ_NARRAYS; ++q)
    if (ADJ_WRITTEN[q]) {
      off[q] = slot_bytes;
      slot_bytes += (h->n[q] * tsize(h) + 255) / 256 * 256;
    }
  const size_t off_s = slot_bytes;  // the step's scalars (AdjScal)
  slot_bytes += 256;
  char* slots = nullptr;
  if (h->use_slots && hipMalloc(&slots, slot_bytes * nseg) != hipSuccess) {
    slots = nullptr;
    (void)hipGetLastError();  // the failed allocation only selects the recomputing reverse
  }
  if (!slots && hipMalloc(&seg, sb * nseg) != hipSuccess) {  // the states of one segment
    cleanup();
    return afail(h, LES_E_NOMEM, "window: segment memory");
  }
  void* A0[ADJ_NARRAYS];
  std::memcpy(A0, h->A, sizeof(A0));
  void* const S0 = h->S;
  auto slot = [&](int t, int q) { return (void*)(slots + slot_bytes * t + off[q]); };
  auto point = [&](int t) {
    for (int q = 0; q < ADJ_NARRAYS; ++q)
      if (ADJ_WRITTEN[q]) h->A[q] = slot(t, q);
    h->S = slots + slot_bytes * t + off_s;
  };
  auto unslot = [&]() {  // the handle's own arrays back, the slots freed
    std::memcpy(h->A, A0, sizeof(A0));
    h->S = S0;
    if (slots) {
      hipStreamSynchronize(h->st);
      hipFree(slots);
      slots = nullptr;
    }
  };
  auto fail_slots = [&](int code) {
    unslot();
    cleanup();
    return code;
  };
  // FP32 with the hand-written transposes: the segment is recomputed by the plain handle's
  // fused kernels (the stages before the pressure solve; p from the forward window's copy),
  // and each step's slot takes the arrays those transposes read: u, v, w, the snapshot
  // before the outflow, uout, diu1..9 and sm.  Its values are the statement kernels' (the
  // window's forward runs the same kernels).
  const bool plain_rc = plain && h->hand && pst && slots;
  const int kept[] = {ADJ_A_u, ADJ_A_v, ADJ_A_w, ADJ_A_sm, ADJ_A_diu1, ADJ_A_diu2, ADJ_A_diu3,
                      ADJ_A_diu4, ADJ_A_diu5, ADJ_A_diu6, ADJ_A_diu7, ADJ_A_diu8, ADJ_A_diu9};
  constexpr int NKEPT = sizeof(kept) / sizeof(kept[0]);
  void* kept_plain[NKEPT] = {};
  if (plain_rc) {
    const int fid[NKEPT] = {LES_U, LES_V, LES_W, LES_SM, LES_DIU1, LES_DIU2, LES_DIU3,
                            LES_DIU4, LES_DIU5, LES_DIU6, LES_DIU7, LES_DIU8, LES_DIU9};
    for (int q = 0; q < NKEPT; ++q) kept_plain[q] = les::field_ptr(h->ph, fid[q]);
  }
  if (slots) {
    // elements a step reads but never writes keep the values of the handle's arrays (the
    // plain recomputation overwrites every array the hand-written transposes read)
    for (int t = 0; t < nseg && !plain_rc; ++t) {
      void* dst[ADJ_NARRAYS];
      const void* src[ADJ_NARRAYS];
      size_t bytes[ADJ_NARRAYS];
      int m = 0;
      for (int q = 0; q < ADJ_NARRAYS; ++q)
        if (ADJ_WRITTEN[q]) {
          dst[m] = slot(t, q);
          src[m] = A0[q];
          bytes[m++] = h->n[q] * tsize(h);
        }
      if ((r = copy_arrays(h, m, dst, src, bytes))) return fail_slots(r);
    }
    for (int c = nck - 1; c >= 0 && plain_rc; --c) {
      const int s0 = c * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      {  // the checkpoint into the plain handle's fields
        void* dst[NCARRIED];
        const void* src[NCARRIED];
        size_t bytes[NCARRIED];
        const char* b = ck + sb * c;
        for (int q = 0; q < NCARRIED; ++q) {
          dst[q] = Ac[q];
          src[q] = b;
          bytes[q] = h->n[CARRIED[q]] * tsize(h);
          b += bytes[q];
        }
        if ((r = copy_arrays(h, NCARRIED, dst, src, bytes))) return fail_slots(r);
      }
      les::set_drag(h->ph, static_cast<const float*>(h->A[ADJ_A_cd]));
      for (int s = s0; s < s1; ++s) {
        point(s - s0);
        if ((r = les::stages_external(h->ph, h->st, 2, static_cast<float*>(h->A[ADJ_A_u_pre]),
                                      static_cast<float*>(h->A[ADJ_A_v_pre]),
                                      static_cast<float*>(h->A[ADJ_A_w_pre]),
                                      &static_cast<AdjScal<float>*>(h->S)->uout))) {
          les::set_drag(h->ph, nullptr);
          return fail_slots(afail(h, r, "window: plain stages"));
        }
        void* dst[NKEPT + 1];
        const void* src[NKEPT + 1];
        size_t bytes[NKEPT + 1];
        for (int q = 0; q < NKEPT; ++q) {
          dst[q] = h->A[kept[q]];
          src[q] = kept_plain[q];
          bytes[q] = h->n[kept[q]] * tsize(h);
        }
        dst[NKEPT] = Ac[3];  // p after the step, from the forward window
        src[NKEPT] = pst + pbytes * s;
        bytes[NKEPT] = pbytes;
        if ((r = copy_arrays(h, NKEPT + 1, dst, src, bytes))) {
          les::set_drag(h->ph, nullptr);
          return fail_slots(r);
        }
      }
      les::set_drag(h->ph, nullptr);
      // the SOR's transpose through the split form, as after a skipped solve (the statement
      // sweeps if it does not apply)
      h->sor_split = h->p_mirror && h->cfg.nmaxp >= 1;
      for (int s = s1 - 1; s >= s0; --s) {
        point(s - s0);
        if ((r = zero_temporaries(h))) return fail_slots(r);
        seed_probes(h, nprobes, wgt);
        step_bw(h);
      }
    }
    for (int c = nck - 1; c >= 0 && !plain_rc; --c) {
      const int s0 = c * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      for (int s = s0; s < s1; ++s) {
        point(s - s0);
        if (s == s0) {
          if ((r = load_state(h, ck + sb * c))) return fail_slots(r);
        } else {
          void* dst[NCARRIED];
          const void* src[NCARRIED];
          size_t bytes[NCARRIED];
          for (int q = 0; q < NCARRIED; ++q) {
            const int a = CARRIED[q];
            dst[q] = slot(s - s0, a);
            src[q] = slot(s - s0 - 1, a);
            bytes[q] = h->n[a] * tsize(h);
          }
          if ((r = copy_arrays(h, NCARRIED, dst, src, bytes))) return fail_slots(r);
        }
        if ((r = recompute(s))) return fail_slots(r);
      }
      for (int s = s1 - 1; s >= s0; --s) {
        point(s - s0);
        if ((r = zero_temporaries(h))) return fail_slots(r);
        seed_probes(h, nprobes, wgt);
        step_bw(h);
      }
    }
    unslot();
  } else {
    for (int c = nck - 1; c >= 0; --c) {
      const int s0 = c * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      if ((r = load_state(h, ck + sb * c))) return cleanup(), r;
      for (int s = s0; s < s1; ++s) {
        if ((r = save_state(h, seg + sb * (s - s0)))) return cleanup(), r;
        if (s + 1 < s1 && (r = recompute(s))) return cleanup(), r;
      }
      for (int s = s1 - 1; s >= s0; --s) {
        if ((r = load_state(h, seg + sb * (s - s0)))) return cleanup(), r;
        if ((r = recompute(s))) return cleanup(), r;
        if ((r = zero_temporaries(h))) return cleanup(), r;
        seed_probes(h, nprobes, wgt);
        step_bw(h);
      }
    }
  }
  if ((r = load_state(h, fin))) return cleanup(), r;
  if ((e = hipStreamSynchronize(h->st)) != hipSuccess) return hipfail(e, "window");
  *objective = jsum * wgt;
  cleanup();
  return check(h);
}

