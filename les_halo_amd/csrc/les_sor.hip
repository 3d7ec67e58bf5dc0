// Red-black SOR of press.f95:87-115 on a colour-split copy of p and rhs.
//
// The reference sweeps the Fortran array in place with stride 2 along i and refreshes the
// boundary images of p after every colour (boundp1) and every iteration (boundp2).  Here, for
// the duration of the solve, interior cell (i,j,k) is stored in array c = (i+j+k)&1 at
// position ((i-1)>>1, j-1, k-1) of a padded row-major box (rows of nr floats, nr a multiple of
// 4): every row of one colour starts at element 0 and is contiguous, so a sweep reads and
// writes whole cache lines, 4 cells per thread as float4 - 8 bytes per grid cell per sweep.
//
// Boundary images.  When the halos of p are mirror images of the interior on entry (the state
// boundp1 + boundp2 leave, i.e. after every press), the image a cell reads across an i face or
// a k face is, at the time of the read, the cell's own current value (it was refreshed after
// the cell's last update and the cell has not changed since); the image across the periodic j
// face is the current value of the cell at the other end of the row, which for even jp has the
// other colour and was refreshed after the other colour's last sweep.  The sweep therefore
// reads the cell itself for i/k faces and the wrapped row for j faces, keeps no halos, and the
// final gather writes every image from the interior - the reference's state after the last
// boundp2.  Odd jp (same colour across the seam) and inconsistent halos use the in-place path.
//
// cn1 is recomputed from its 1-D factors (init.f95:98-104) in the reference's operation order,
// bit-identical to the stored array; on a uniform horizontal grid (every ax, by, cn2*, cn3*
// equal) the factors reduce to scalars and a per-k table of cn1.
#include "les_kernels.h"

namespace les {

namespace {

// colour of cell (i,j,k) and first interior i of colour c in row (j,k)
__device__ __forceinline__ int first_i(int c, int j, int k) { return ((1 + j + k) & 1) == c ? 1 : 2; }

__device__ __forceinline__ int split_index(const SplitSOR& sp, int i, int j, int k) {
  return ((i - 1) >> 1) + sp.nr * ((j - 1) + sp.nj * (k - 1));
}

}  // namespace

// ---- p, rhs (Fortran layouts) -> split copy; optionally rhs -= rhsav (press.f95:76-82) ----
// A thread takes split entry ii of row (j,k) in both colours, i.e. the two neighbouring cells
// i = 1 + 2ii and 2 + 2ii: its split stores are unit-stride across the wavefront, and its two
// Fortran loads (and rhs stores) fill each other's cache lines.
__global__ __launch_bounds__(256) void k_split_scatter(Dims d, Fields F, SplitSOR sp,
                                                       const DevScalars* sc, int adjust_rhs,
                                                       int with_rhs) {
  const int ii = blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (ii >= sp.nr || j > d.jp) return;
  const int base = ii + sp.nr * ((j - 1) + sp.nj * (k - 1));
  // both colours' loads before any store (a store could alias them for the compiler)
  float pv[2], rv[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int i = min(first_i(c, j, k) + 2 * ii, d.ip);  // clamped: loaded, not used
    pv[c] = F.p[d.P(i, j, k)];
    rv[c] = with_rhs ? F.rhs[d.R(i, j, k)] : 0.0f;
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int i = first_i(c, j, k) + 2 * ii;
    if (i > d.ip) continue;  // padding entries of the split rows stay untouched
    sp.p[c][base] = pv[c];
    if (!with_rhs) continue;
    float r = rv[c];
    if (adjust_rhs) {
      r = r - sc->rhsav;
      F.rhs[d.R(i, j, k)] = r;
    }
    sp.r[c][base] = r;
  }
}

void launch_split_scatter(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                          int adjust_rhs, int with_rhs) {
  dim3 grid((sp.nr + 63) / 64, (d.jp + 3) / 4, d.kp);
  hipLaunchKernelGGL(k_split_scatter, grid, dim3(64, 4), 0, st, d, F, sp, sc, adjust_rhs,
                     with_rhs);
}

// ---- one colour sweep: 4 consecutive cells of one colour row per thread ----
// A thread takes group grp (entries 4 grp .. 4 grp + 3) of colour row j over KC consecutive
// levels, marching up in k: the other colour's rows at k-1 and k stay in registers from the
// levels before, so a level loads the other colour at k+1 and at j +- 1 only.  A sweep that
// sums the residual takes SPLIT_KC_RESID levels a thread (its exact-sum flush then serves four
// times the cells: at 300x300x90 with the residual every iteration, 2.95 ms per step against
// 3.57 with one level, 3.18 with two, 3.41 with eight); a sweep without it one level (2x2 tiles
// of 150x150x90: 2.43 ms per step against 2.54, 2.60 and 2.88 with two, four and eight).
constexpr int SPLIT_KC_RESID = 4;
// L: the workgroup's index in the sweep (after the XCD remap); lds: its limb scratch
template <bool UNIFORM, bool RESID, int KC = RESID ? SPLIT_KC_RESID : 1>
__device__ __forceinline__ void sor_split_body(const Dims& d, const Grid1D& g, const SplitSOR& sp,
                                               const Physics& ph, int nrd, int64_t* partials,
                                               int ng, int nthreads, int L,
                                               unsigned long long* lds,
                                               const SeqTerms& tm = SeqTerms{},
                                               const SplitPush* push = nullptr) {
  if (RESID) block_limbs_init(lds);
  ThreadAcc acc;
  acc.init();
  const int t = L * 256 + threadIdx.x;
  const int tc = min(t, nthreads - 1);  // tail lanes index a valid group
  // tc = grp + ng * row, row = (j-1) + jp * chunk; quotients by multiply-high (exact for the
  // sizes sor_split_divisors admits)
  const int row = (int)__umulhi((unsigned)tc, sp.m_ng);
  const int grp = tc - row * ng;
  const int ch = (int)__umulhi((unsigned)row, sp.m_jp);
  const int j = 1 + row - ch * d.jp;
  const int k0 = 1 + ch * KC, k1 = min(d.kp, k0 + KC - 1);
  const int c = (1 + nrd) & 1;  // nrd = 0 updates i = 1 + mod(k+j,2): odd i+j+k
  const float* po = sp.p[c ^ 1];
  const int pl = sp.nr * sp.nj;
  const int rowoff = sp.nr * (j - 1) + 4 * grp;  // the group in level 1
  // neighbour rows: periodic in j (a tile: its halo rows, j = jp+1 at row index jp and j = 0 at
  // jp+1), the cell itself across k faces
  const int rn = (j == d.jp && !sp.jhalo) ? -sp.nr * (d.jp - 1) : sp.nr;     // j+1
  const int rsj = (j == 1) ? sp.nr * (sp.jhalo ? d.jp + 1 : d.jp - 1) : -sp.nr;  // j-1
  float c2l = 0.0f, c2s = 0.0f, c3l, c3s;
  if (UNIFORM) {
    c2l = g.cn2l[0];
    c2s = g.cn2s[0];
    c3l = g.cn3l[0];
    c3s = g.cn3s[0];
  } else {
    c3l = g.cn3l[j - 1];
    c3s = g.cn3s[j - 1];
  }
  auto ld4 = [](const float* p) { return *reinterpret_cast<const float4*>(p); };
  if (t < nthreads) {
    // the other colour's group at levels k-1 (od) and k (oc); od is not read at k = 1
    float4 oc = ld4(po + rowoff + pl * (k0 - 1));
    float4 od = ld4(po + rowoff + pl * max(k0 - 2, 0));
    for (int k = k0; k <= k1; ++k) {
      const int base = rowoff + pl * (k - 1);
      const float4 ou = ld4(po + base + (k < d.kp ? pl : 0));  // k+1 (k itself at the top)
      const int i0 = first_i(c, j, k) + 8 * grp;  // i of the group's first cell
      if (i0 <= d.ip) {
        // i faces shared with a neighbour tile: its cells p(0,j,k) and p(ip+1,j,k)
        const int ih = (j - 1) + d.jp * (k - 1);
        const float hwv = (sp.hw && i0 == 1) ? sp.hw[ih] : 0.0f;
        const float hev = (sp.he && i0 + 6 >= d.ip) ? sp.he[ih] : 0.0f;
        const float4 own = ld4(sp.p[c] + base);
        const float4 rhs = ld4(sp.r[c] + base);
        const float4 on = ld4(po + base + rn);
        const float4 os = ld4(po + base + rsj);
        const float4 uu = (k < d.kp) ? ou : own;
        const float4 dd = (k > 1) ? od : own;
        // i neighbours: odd i -> east o[ii], west o[ii-1]; even i -> east o[ii+1], west o[ii]
        float xe[4], xw[4];
        const float ov[4] = {oc.x, oc.y, oc.z, oc.w};
        if (i0 & 1) {
          const float prev = (grp > 0) ? po[base - 1] : 0.0f;
          xe[0] = ov[0]; xe[1] = ov[1]; xe[2] = ov[2]; xe[3] = ov[3];
          xw[0] = prev; xw[1] = ov[0]; xw[2] = ov[1]; xw[3] = ov[2];
        } else {
          const float next = (i0 + 7 <= d.ip) ? po[base + 4] : 0.0f;
          xw[0] = ov[0]; xw[1] = ov[1]; xw[2] = ov[2]; xw[3] = ov[3];
          xe[0] = ov[1]; xe[1] = ov[2]; xe[2] = ov[3]; xe[3] = next;
        }
        const float pv[4] = {own.x, own.y, own.z, own.w};
        const float rv[4] = {rhs.x, rhs.y, rhs.z, rhs.w};
        const float nv[4] = {on.x, on.y, on.z, on.w};
        const float sv[4] = {os.x, os.y, os.z, os.w};
        const float uv[4] = {uu.x, uu.y, uu.z, uu.w};
        const float dv[4] = {dd.x, dd.y, dd.z, dd.w};
        float out[4];
        const float c4l = g.cn4l[k - 1], c4s = g.cn4s[k - 1];
        const float cn1k = UNIFORM ? sp.cn1k[k] : 0.0f;
        auto relax = [&](int q, int i, float pe, float pw) {
          const float pc = pv[q];
          float cn1, a2l = c2l, a2s = c2s;
          if (UNIFORM) {
            cn1 = cn1k;
          } else {
            cn1 = 1.0f / (sp.ax[i] + sp.by[j] + sp.cz[k]);
            a2l = g.cn2l[i - 1];
            a2s = g.cn2s[i - 1];
          }
          const float reltmp =
              ph.omega * (cn1 * (a2l * pe + a2s * pw + c3l * nv[q] + c3s * sv[q] + c4l * uv[q] +
                                 c4s * dv[q] - rv[q]) -
                          pc);
          if (RESID) {
            if (tm.t[0])  // the reference-order sum adds the terms afterwards
              tm.t[c][((i - 1) >> 1) + tm.nrt * ((j - 1) + d.jp * (k - 1))] = reltmp * reltmp;
            else
              acc.add(reltmp * reltmp, lds);
          }
          return pc + reltmp;
        };
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = i0 + 2 * q;
          const float pc = pv[q];
          const float pe = (i == d.ip) ? (sp.he ? hev : pc) : xe[q];
          const float pw = (i == 1) ? (sp.hw ? hwv : pc) : xw[q];
          out[q] = (i <= d.ip) ? relax(q, i, pe, pw) : pc;
        }
        const float4 o4 = make_float4(out[0], out[1], out[2], out[3]);
        *reinterpret_cast<float4*>(sp.p[c] + base) = o4;
        if (push) {  // this tile's face cells are its neighbours' halos (same split position)
          const int lev = pl * (k - 1) + 4 * grp;
          if (j == 1 && push->s[c])
            *reinterpret_cast<float4*>(push->s[c] + lev + sp.nr * d.jp) = o4;
          if (j == d.jp && push->n[c])
            *reinterpret_cast<float4*>(push->n[c] + lev + sp.nr * (d.jp + 1)) = o4;
          if (i0 == 1 && push->w) push->w[ih] = out[0];
          const int qe = (d.ip - i0) >> 1;  // the group's cell i = ip, if it has one
          if (push->e && i0 + 6 >= d.ip && i0 + 2 * qe == d.ip)
            push->e[ih] = qe == 0 ? out[0] : (qe == 1 ? out[1] : (qe == 2 ? out[2] : out[3]));
        }
      }
      od = oc;
      oc = ou;
    }
  }
  if (RESID) {  // into the accumulation slots (partials: RedBuf::slots)
    acc.flush(lds);
    block_limbs_atomic(lds, partials, (int)blockIdx.x);
  }
}

template <bool UNIFORM, bool RESID>
__global__ __launch_bounds__(256) void k_sor_split(Dims d, Grid1D g, SplitSOR sp, Physics ph,
                                                   int nrd, int64_t* partials,
                                                   const DevScalars* sc, int ng, int nthreads,
                                                   SeqTerms tm) {
  __shared__ unsigned long long lds[NLIMB];
  if (sc->sor_done) return;
  sor_split_body<UNIFORM, RESID>(d, g, sp, ph, nrd, partials, ng, nthreads,
                                 xcd_tile(blockIdx.x, gridDim.x), lds, tm);
}

// one colour sweep of every tile of a cut domain (blockIdx.y = tile), no residual
template <bool UNIFORM>
__global__ __launch_bounds__(256) void k_sor_split_tiles(Dims d, SorTiles T, Physics ph, int nrd,
                                                         int ng, int nthreads) {
  const int q = blockIdx.y;
  if (T.sc[q]->sor_done) return;
  sor_split_body<UNIFORM, false>(d, T.g[q], T.sp[q], ph, nrd, nullptr, ng, nthreads,
                                 xcd_tile(blockIdx.x, gridDim.x), nullptr, SeqTerms{}, &T.push[q]);
}

// ---- one whole SOR iteration (colour 1 then colour 0, press.f95:90-100 with nrd = 0, 1) per
// launch, on a uniform horizontal grid without residual.
//
// The iteration reads the previous iterate (pin) and writes the next (pout), so workgroups
// never see each other's results.  A workgroup owns TJ rows of j (all of i) and a chunk of k
// levels and marches up in k with two rings of three levels in LDS:
//   sB: the old colour-0 values, rows j0-2 .. j0+TJ+1 (TJ+4 rows);
//   sA: the new colour-1 values, rows j0-1 .. j0+TJ   (TJ+2 rows), recomputed on the two
//       rows next to the tile so that colour 0 at the tile's edge rows sees final colour-1
//       neighbours.
// At level k it stores old colour 0 of level k+2 into the ring, updates colour 1 at k+1, then
// colour 0 at k from the colour-1 ring, and writes both colours of level k.  The global loads
// of the next level (colour 0 at k+3, colour 1 and its rhs at k+2, the colour-0 rhs at k+1) are
// issued into registers while the current level computes, so their latency overlaps the
// barriers.  A thread handles groups of four consecutive cells of a colour row (float4 loads,
// stores and LDS accesses); the groups of all rows of a ring level are numbered together and
// dealt to the workgroup's threads, so few lanes idle when a row is shorter than a wavefront.
// Every cell gets the same FP32 expression as in k_sor_split with the same operand values: the
// result is bit-identical to the two colour sweeps, with 12 B per cell of traffic per
// iteration instead of 16 B.
// Workgroup barrier for LDS hand-offs only: waits for this wavefront's LDS operations, not
// for its outstanding global loads and stores (which __syncthreads() would drain), so the
// prefetched loads of the next level stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The body of one workgroup with logical index L (k_sor_iter, and the ensemble's batched
// launch, which deals the tiles of every member to one grid)
template <int TJ, int NT, bool FIRST>
__device__ __forceinline__ void sor_iter_body(const Dims& d, const Grid1D& g, const SplitSOR& sp,
                                              const Physics& ph, const SorPair& io,
                                              const DevScalars* sc, int kc, int njt, int L) {
  constexpr int NW = SOR_ITER_NRMAX;         // LDS row stride (floats)
  constexpr int NGMAX = SOR_ITER_NRMAX / 4;  // groups of four cells in the longest row
  constexpr int PB = ((TJ + 4) * NGMAX + NT - 1) / NT;  // colour-0 ring groups per thread
  constexpr int PA = ((TJ + 2) * NGMAX + NT - 1) / NT;  // colour-1 ring groups per thread
  constexpr int PU = (TJ * NGMAX + NT - 1) / NT;        // tile groups per thread
  __shared__ float4 sB4[3][TJ + 4][NW / 4];
  __shared__ float4 sA4[3][TJ + 2][NW / 4];
  __shared__ float sC[3][SOR_ITER_KCMAX + 3];  // cn1, cn4l, cn4s of levels k0-1 .. k1+1
  if (sc->sor_done) return;
  // FIRST: the shifts applied to every rhs and p value read (rhs -= rhsav; p -= the pav the
  // gather took off), each the same FP32 subtraction the scatter and the gather make
  float rsh = 0.0f, psh = 0.0f;
  if constexpr (FIRST) {
    rsh = sc->rhsav;
    psh = io.p_shift ? sc->pav_carry : 0.0f;
  }
  auto sub4 = [](float4 a, float s) { return make_float4(a.x - s, a.y - s, a.z - s, a.w - s); };
  const int jt = L % njt, kt = L / njt;
  const int j0 = 1 + jt * TJ;
  const int k0 = 1 + kt * kc;
  if (k0 > d.kp) return;  // padding workgroups of the XCD remap (uniform per workgroup)
  const int k1 = min(d.kp, k0 + kc - 1);
  const int nr = sp.nr, ng = nr >> 2, jp = d.jp, kp = d.kp, ip = d.ip;
  const float c2l = g.cn2l[0], c2s = g.cn2s[0], c3l = g.cn3l[0], c3s = g.cn3s[0];
  const float om = ph.omega;
  const int lstride = nr * jp;
  auto wrapj = [&](int j) { return ((j - 1) % jp + jp) % jp + 1; };  // periodic j
  auto lev = [&](int k) { return lstride * (k - 1); };
  // The level coefficients come from LDS: a global load of them between the prefetches would
  // make the wait for it (the vector memory counter drains in order) a wait for the prefetches.
  for (int t = threadIdx.x; t < k1 - k0 + 3; t += NT) {
    const int k = k0 - 1 + t;
    if (k >= 1 && k <= kp) {
      sC[0][t] = sp.cn1k[k];
      sC[1][t] = g.cn4l[k - 1];
      sC[2][t] = g.cn4s[k - 1];
    }
  }

  // Work items: item t of a ring level is group t % ng of ring row t / ng.  Items past the end
  // are clamped onto the last one for their loads (every lane issues every load, so the number
  // of loads in flight is the same on every path and the compiler's waits count exactly) and
  // skip LDS writes, updates and stores.
  int bRow[PB], bG[PB], bOff[PB];
  bool bOk[PB];
#pragma unroll
  for (int q = 0; q < PB; ++q) {
    const int t0 = threadIdx.x + NT * q;
    bOk[q] = t0 < (TJ + 4) * ng;
    const int t = min(t0, (TJ + 4) * ng - 1);
    bRow[q] = t / ng;
    bG[q] = t - bRow[q] * ng;
    bOff[q] = nr * (wrapj(j0 - 2 + bRow[q]) - 1) + 4 * bG[q];
  }
  int aRow[PA], aG[PA], aJ[PA], aOff[PA];
  bool aOk[PA];
#pragma unroll
  for (int q = 0; q < PA; ++q) {
    const int t0 = threadIdx.x + NT * q;
    aOk[q] = t0 < (TJ + 2) * ng;
    const int t = min(t0, (TJ + 2) * ng - 1);
    aRow[q] = t / ng;
    aG[q] = t - aRow[q] * ng;
    aJ[q] = wrapj(j0 - 1 + aRow[q]);
    aOff[q] = nr * (aJ[q] - 1) + 4 * aG[q];
  }
  const int nUrow = min(TJ, jp - j0 + 1);  // tile rows inside the grid
  int uRow[PU], uG[PU], uOff[PU];
  bool uOk[PU];
#pragma unroll
  for (int q = 0; q < PU; ++q) {
    const int t0 = threadIdx.x + NT * q;
    uOk[q] = t0 < nUrow * ng;
    const int t = min(t0, nUrow * ng - 1);
    uRow[q] = t / ng;
    uG[q] = t - uRow[q] * ng;
    uOff[q] = nr * (j0 + uRow[q] - 1) + 4 * uG[q];
  }

  float4 bq[PB];          // old colour 0 of one level (ring rows)
  float4 ao[PA], ar[PA];  // old colour 1 and its rhs (ring rows)
  float4 br[PU];          // colour-0 rhs (tile rows)
  auto ld4 = [](const float* p) { return *reinterpret_cast<const float4*>(p); };
  auto fetch_b = [&](int k) {
#pragma unroll
    for (int q = 0; q < PB; ++q) bq[q] = ld4(io.pin[0] + bOff[q] + lev(k));
  };
  auto put_b = [&](int k, auto sl) {  // into ring slot sl of level k
#pragma unroll
    for (int q = 0; q < PB; ++q)
      if (bOk[q]) sB4[sl][bRow[q]][bG[q]] = FIRST ? sub4(bq[q], psh) : bq[q];
  };
  auto fetch_a = [&](int k) {
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      ao[q] = ld4(io.pin[1] + aOff[q] + lev(k));
      ar[q] = ld4(sp.r[1] + aOff[q] + lev(k));
    }
  };
  auto fetch_br = [&](int k) {
#pragma unroll
    for (int q = 0; q < PU; ++q) br[q] = ld4(sp.r[0] + uOff[q] + lev(k));
  };
  // Four cells of one colour row: own values pc, rhs rr; the other colour of the same level
  // in LDS (om_: row j, on_/os_: rows j+1/j-1, ou_/od_: levels k+1/k-1).  The first own cell
  // is i0 = f + 8 g; its east/west neighbours are other-colour entries 4g+q+f-1 and 4g+q+f-2.
  // INT (a compile-time bool) says 1 < k < kp, where the k-boundary selections drop out.
  auto relax4 = [&](float4 pc4, float4 rr4, const float4* om_, const float4* on_,
                    const float4* os_, const float4* ou_, const float4* od_, int g4, int f, int k,
                    float cn1, float c4l, float c4s, auto INT) {
    const float4 xm = om_[g4];
    // entry 4g-1 (f == 1) or 4g+4 (f == 2), clamped into the row: only read for cells that
    // take the value (never for i == 1 or i == ip)
    const float xo =
        reinterpret_cast<const float*>(om_)[f == 1 ? max(4 * g4 - 1, 0) : min(4 * g4 + 4, NW - 1)];
    const float4 xn = on_[g4], xs = os_[g4], xu = ou_[g4], xd = od_[g4];
    // xmv[1 + m] is entry 4g+m; xmv[0] stands for 4g-1 (f == 1) and xmv[5] for 4g+4 (f == 2)
    const float xmv[6] = {xo, xm.x, xm.y, xm.z, xm.w, xo};
    const float pcv[4] = {pc4.x, pc4.y, pc4.z, pc4.w};
    const float rrv[4] = {rr4.x, rr4.y, rr4.z, rr4.w};
    const float xnv[4] = {xn.x, xn.y, xn.z, xn.w}, xsv[4] = {xs.x, xs.y, xs.z, xs.w};
    const float xuv[4] = {xu.x, xu.y, xu.z, xu.w}, xdv[4] = {xd.x, xd.y, xd.z, xd.w};
    float out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // computed for every cell and kept where i <= ip (padding cells keep their value), so
      // the LDS reads stay whole float4s outside any branch
      const int i = f + 2 * (4 * g4 + q);
      const float pc = pcv[q];
      const float pe = (i == ip) ? pc : (f == 1 ? xmv[1 + q] : xmv[2 + q]);
      const float pw = (i == 1) ? pc : (f == 1 ? xmv[q] : xmv[1 + q]);
      const float pu = (INT || k < kp) ? xuv[q] : pc;
      const float pd = (INT || k > 1) ? xdv[q] : pc;
      const float reltmp =
          om * (cn1 * (c2l * pe + c2s * pw + c3l * xnv[q] + c3s * xsv[q] + c4l * pu + c4s * pd -
                       rrv[q]) -
                pc);
      out[q] = (i <= ip) ? pc + reltmp : pc;
    }
    return make_float4(out[0], out[1], out[2], out[3]);
  };
  // Ring slot of level k: (k - k0 + 3) % 3.  The main loop passes slots as compile-time
  // constants (std::integral_constant), so LDS addresses are fixed offsets.
  auto slot = [&](int k) { return (k - k0 + 3) % 3; };
  // new colour 1 of level k (slots m, u = k+1, dn = k-1) on ring rows 0 .. TJ+1 -> sA ring
  auto update_a = [&](int k, auto m, auto u, auto dn, auto INT) {
    const int t = k - k0 + 1;
    const float cn1 = sC[0][t], c4l = sC[1][t], c4s = sC[2][t];
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      if (!aOk[q]) continue;
      const int r = aRow[q];
      const int f = ((1 + aJ[q] + k) & 1) == 1 ? 1 : 2;  // first i of colour 1 in this row
      const float4 pc4 = FIRST ? sub4(ao[q], psh) : ao[q];
      const float4 rr4 = FIRST ? sub4(ar[q], rsh) : ar[q];
      const float4 v = relax4(pc4, rr4, sB4[m][r + 1], sB4[m][r + 2], sB4[m][r],
                              sB4[u][r + 1], sB4[dn][r + 1], aG[q], f, k, cn1, c4l, c4s, INT);
      sA4[m][r][aG[q]] = v;
      if constexpr (FIRST) {  // the adjusted rhs of the colour-1 cells this workgroup owns
        if (k >= k0 && k <= k1 && r >= 1 && r <= nUrow)
          *reinterpret_cast<float4*>(io.rout[1] + aOff[q] + lev(k)) = rr4;
      }
    }
  };
  // new colour 0 of level k on the tile rows; writes both colours of level k
  auto update_b = [&](int k, auto m, auto u, auto dn, auto INT) {
    const int t = k - k0 + 1;
    const float cn1 = sC[0][t], c4l = sC[1][t], c4s = sC[2][t];
#pragma unroll
    for (int q = 0; q < PU; ++q) {
      if (!uOk[q]) continue;
      const int r = uRow[q];
      const int f = ((1 + j0 + r + k) & 1) == 0 ? 1 : 2;  // first i of colour 0 in this row
      const float4 pc4 = sB4[m][r + 2][uG[q]];
      const float4 rr4 = FIRST ? sub4(br[q], rsh) : br[q];
      const float4 v = relax4(pc4, rr4, sA4[m][r + 1], sA4[m][r + 2], sA4[m][r],
                              sA4[u][r + 1], sA4[dn][r + 1], uG[q], f, k, cn1, c4l, c4s, INT);
      const int o = uOff[q] + lev(k);
      if constexpr (FIRST) *reinterpret_cast<float4*>(io.rout[0] + o) = rr4;
      *reinterpret_cast<float4*>(io.pout[0] + o) = v;
      *reinterpret_cast<float4*>(io.pout[1] + o) = sA4[m][r + 1][uG[q]];
    }
  };
  const std::false_type EDGE{};
  const std::true_type INNER{};

  // prologue: colour 0 at k0-2 .. k0+1 in the ring, colour 1 at k0-1 and k0.  Fetches of a
  // level outside 1..kp (or past the tile) load a clamped level and are never used.
  fetch_b(max(k0 - 2, 1));
  if (k0 - 2 >= 1) put_b(k0 - 2, slot(k0 - 2));
  fetch_b(max(k0 - 1, 1));
  if (k0 - 1 >= 1) put_b(k0 - 1, slot(k0 - 1));
  fetch_b(k0);
  put_b(k0, slot(k0));
  fetch_b(min(k0 + 1, kp));
  fetch_a(max(k0 - 1, 1));
  lds_barrier();
  if (k0 - 1 >= 1) update_a(k0 - 1, slot(k0 - 1), slot(k0), slot(k0 - 2), EDGE);
  lds_barrier();  // level k0-2's slot is read by update_a(k0-1) and refilled next
  if (k0 + 1 <= kp) put_b(k0 + 1, slot(k0 + 1));
  fetch_a(k0);
  lds_barrier();
  update_a(k0, slot(k0), slot(k0 + 1), slot(k0 - 1), EDGE);
  // registers for the first step
  fetch_b(min(k0 + 2, kp));
  fetch_a(min(k0 + 1, kp));
  fetch_br(k0);
  // One level k with slots s0 = slot(k), s1 = slot(k+1), s2 = slot(k+2) = slot(k-1).  Every
  // step ends with the same three fetches, so the loads in flight at the top of a step are
  // the same whichever step ran before it.
  auto step = [&](int k, auto s0, auto s1, auto s2) {
    lds_barrier();  // sA(k) complete; the ring slots of level k-1 are free
    if (k + 2 <= kp) put_b(k + 2, s2);
    fetch_b(min(k + 3, kp));  // for the next step, in flight
    lds_barrier();
    if (k + 1 < kp)
      update_a(k + 1, s1, s2, s0, INNER);  // 1 < k+1 always
    else if (k + 1 == kp)
      update_a(k + 1, s1, s2, s0, EDGE);
    fetch_a(min(k + 2, kp));
    lds_barrier();
    if (k > 1 && k < kp)
      update_b(k, s0, s1, s2, INNER);
    else
      update_b(k, s0, s1, s2, EDGE);
    fetch_br(min(k + 1, k1));
  };
  const std::integral_constant<int, 0> C0{};
  const std::integral_constant<int, 1> C1{};
  const std::integral_constant<int, 2> C2{};
  for (int k = k0; k <= k1; k += 3) {
    step(k, C0, C1, C2);
    if (k + 1 <= k1) step(k + 1, C1, C2, C0);
    if (k + 2 <= k1) step(k + 2, C2, C0, C1);
  }
}

template <int TJ, int NT, bool FIRST>
__global__ __launch_bounds__(NT) void k_sor_iter(Dims d, Grid1D g, SplitSOR sp, Physics ph,
                                                  SorPair io, const DevScalars* sc, int kc,
                                                  int njt) {
  sor_iter_body<TJ, NT, FIRST>(d, g, sp, ph, io, sc, kc, njt,
                               xcd_tile((int)blockIdx.x, (int)gridDim.x));
}

int sor_iter_supported(Dims d, const SplitSOR& sp) { return sp.uniform && sp.nr <= SOR_ITER_NRMAX; }

// the batched fused iteration: workgroup w (after the XCD remap) is tile w - base of member wgm[w]
template <bool FIRST>
__global__ __launch_bounds__(256) void k_sor_iter_batch(const BatchSor* __restrict__ tab,
                                                        const uint8_t* __restrict__ wgm,
                                                        int total, int which) {
  const int w = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (w >= total) return;
  // the member's operands into registers once (the body's stores cannot alias them there)
  const BatchSor& t = tab[wgm[w]];
  const Dims d = t.d;
  const Grid1D g = t.g;
  const SplitSOR sp = which == 0 ? t.sp : t.sra;
  const Physics ph = t.ph;
  const SorPair io = which == 0 ? t.first : (which == 1 ? t.fwd : t.back);
  const DevScalars* sc = t.sc;
  sor_iter_body<4, 256, FIRST>(d, g, sp, ph, io, sc, t.kc, t.njt, w - t.base);
}

void launch_sor_iter_batch(hipStream_t st, const BatchSor* tab, const uint8_t* wgm, int total,
                           int which) {
  const dim3 grid((unsigned)((total + 7) / 8 * 8));
  if (which == 0)
    hipLaunchKernelGGL(k_sor_iter_batch<true>, grid, dim3(256), 0, st, tab, wgm, total, which);
  else
    hipLaunchKernelGGL(k_sor_iter_batch<false>, grid, dim3(256), 0, st, tab, wgm, total, which);
}

namespace {
template <int TJ, int NT>
void launch_sor_iter_tj(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, SorPair io,
                        DevScalars* sc, int kc, int first) {
  const int nkt = (d.kp + kc - 1) / kc;
  const int njt = (d.jp + TJ - 1) / TJ;
  const int nwg = (njt * nkt + 7) / 8 * 8;
  if (first)
    hipLaunchKernelGGL((k_sor_iter<TJ, NT, true>), dim3(nwg), dim3(NT), 0, st, d, g, sp, ph, io,
                       sc, kc, njt);
  else
    hipLaunchKernelGGL((k_sor_iter<TJ, NT, false>), dim3(nwg), dim3(NT), 0, st, d, g, sp, ph, io,
                       sc, kc, njt);
}
}  // namespace

void launch_sor_iter(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, SorPair io,
                     DevScalars* sc, SorIterShape shape, int first) {
  const int kc = shape.kc > 0 ? min(shape.kc, SOR_ITER_KCMAX) : 10;
  if (shape.tj == 4)
    launch_sor_iter_tj<4, 256>(st, d, g, sp, ph, io, sc, kc, first);
  else if (shape.tj == 16)
    launch_sor_iter_tj<16, 512>(st, d, g, sp, ph, io, sc, kc, first);
  else
    launch_sor_iter_tj<8, 512>(st, d, g, sp, ph, io, sc, kc, first);
}

namespace {
// kc levels per thread (the forward sweep's; one in the transposed sweep)
void split_threads(Dims d, const SplitSOR& sp, int& ng, int& nthreads, int& nwg, int kc = 1) {
  ng = sp.nr / 4;
  nthreads = ng * d.jp * ((d.kp + kc - 1) / kc);
  const int nb = (nthreads + 255) / 256;
  nwg = (nb + 7) / 8 * 8;
}
}  // namespace


// floor(2^32 / n) + 1: umulhi(t, m) == t / n for every t with t * n < 2^32
uint32_t magic_div(uint32_t n) { return (uint32_t)(0xffffffffu / n) + 1u; }

bool sor_split_divisors(Dims d, SplitSOR& sp) {
  const uint64_t ng = (uint64_t)sp.nr / 4, rows = (uint64_t)d.jp * d.kp;
  if (ng * rows * ng >= (1ull << 32) || rows * (uint64_t)d.jp >= (1ull << 32)) return false;
  sp.m_ng = magic_div((uint32_t)ng);
  sp.m_jp = magic_div((uint32_t)d.jp);
  return true;
}

void launch_sor_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, Physics ph, int nrd,
                      int want_resid, RedBuf rb, DevScalars* sc, SeqTerms tm) {
  int ng, nt, nwg;
  split_threads(d, sp, ng, nt, nwg, want_resid ? SPLIT_KC_RESID : 1);
  int64_t* part = rb.slots;
  const dim3 grid(nwg), blk(256);
  if (sp.uniform) {
    if (want_resid)
      hipLaunchKernelGGL((k_sor_split<true, true>), grid, blk, 0, st, d, g, sp, ph, nrd, part, sc, ng, nt, tm);
    else
      hipLaunchKernelGGL((k_sor_split<true, false>), grid, blk, 0, st, d, g, sp, ph, nrd, part, sc, ng, nt, tm);
  } else {
    if (want_resid)
      hipLaunchKernelGGL((k_sor_split<false, true>), grid, blk, 0, st, d, g, sp, ph, nrd, part, sc, ng, nt, tm);
    else
      hipLaunchKernelGGL((k_sor_split<false, false>), grid, blk, 0, st, d, g, sp, ph, nrd, part, sc, ng, nt, tm);
  }
}

void launch_sor_split_tiles(hipStream_t st, Dims d, const SorTiles& T, Physics ph, int nrd) {
  int ng, nt, nwg;
  split_threads(d, T.sp[0], ng, nt, nwg);
  const dim3 grid(nwg, T.ntiles), blk(256);
  bool uniform = true;
  for (int q = 0; q < T.ntiles; ++q) uniform = uniform && T.sp[q].uniform;
  if (uniform)
    hipLaunchKernelGGL((k_sor_split_tiles<true>), grid, blk, 0, st, d, T, ph, nrd, ng, nt);
  else
    hipLaunchKernelGGL((k_sor_split_tiles<false>), grid, blk, 0, st, d, T, ph, nrd, ng, nt);
}

// ---- pav partial sums (press.f95:125-132) from the split copy ----
// Exact partial sums of dx1*dy1*dzn*p over the colour-split copy (press.f95:123-134).  A
// thread takes one group of four split entries of row j in both colours over a chunk of k
// levels (float4 loads along the split rows); padding entries (i > ip) are skipped.
__global__ __launch_bounds__(256) void k_pav_split(Dims d, Grid1D g, SplitSOR sp,
                                                   int64_t* partials, int ng, int kchunk,
                                                   int64_t* slots) {
  __shared__ unsigned long long lds[NLIMB];
  block_limbs_init(lds);
  const int t = blockIdx.x * 256 + threadIdx.x;
  ThreadAcc acc;
  acc.init();
  if (t < ng * d.jp) {
    const int row = t / ng;
    const int grp = t - row * ng;
    const int j = 1 + row;
    const float dy1j = g.DY1(j);
    const int k0 = 1 + blockIdx.y * kchunk, k1 = min(d.kp, k0 + kchunk - 1);
    for (int k = k0; k <= k1; ++k) {
      const float dznk = g.DZN(k);
      const int base = sp.nr * (row + sp.nj * (k - 1)) + 4 * grp;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float4 v = *reinterpret_cast<const float4*>(sp.p[c] + base);
        const float vv[4] = {v.x, v.y, v.z, v.w};
        const int i0 = first_i(c, j, k) + 8 * grp;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = i0 + 2 * q;
          if (i <= d.ip) acc.add(vv[q] * g.DX1(i) * dy1j * dznk, lds);
        }
      }
    }
  }
  acc.flush(lds);
  const int blk = blockIdx.x + gridDim.x * blockIdx.y;
  if (slots)
    block_limbs_atomic(lds, slots, blk);
  else
    block_limbs_store(lds, partials + (long)blk * NLIMB);
}

__global__ __launch_bounds__(FIN_THREADS) void k_pav_split_finalize(const int64_t* partials,
                                                                    int nb, DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_partials(partials, nb, limbs);
  if (threadIdx.x == 0) sc->pav = limbs_to_float(limbs) / sc->pco;
}

__global__ __launch_bounds__(RED_SLOTS * NLIMB) void k_pav_split_finalize_slots(int64_t* slots,
                                                                               DevScalars* sc) {
  __shared__ int64_t limbs[NLIMB];
  sum_slots(slots, limbs);
  if (threadIdx.x == 0) sc->pav = limbs_to_float(limbs) / sc->pco;
}

void launch_pav_split(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, RedBuf rb, DevScalars* sc) {
  const int ng = sp.nr / 4;
  const int nbx = (ng * d.jp + 255) / 256;
  // about 1300 workgroups: enough waves to stream the copy, few partials to add up
  int nkc = std::min(std::max(1300 / nbx, 1), d.kp);
  const int kchunk = (d.kp + nkc - 1) / nkc;
  nkc = (d.kp + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(k_pav_split, dim3(nbx, nkc), dim3(256), 0, st, d, g, sp, rb.partials, ng,
                     kchunk, rb.slots);
  if (rb.slots)
    hipLaunchKernelGGL(k_pav_split_finalize_slots, dim3(1), dim3(RED_SLOTS * NLIMB), 0, st,
                       rb.slots, sc);
  else
    hipLaunchKernelGGL(k_pav_split_finalize, dim3(1), dim3(FIN_THREADS), 0, st, rb.partials,
                       nbx * nkc, sc);
}

// ---- face exchange of the split copies of the tiles of one cut domain (after a colour
// sweep).  Job 0: my row j = jp+1 (row index jp) <- the north neighbour's row j = 1; job 1: my
// row j = 0 (row index jp+1) <- the south neighbour's row j = jp, both colours, every entry; job
// 2: hw(j,k) <- the west neighbour's p(ip,j,k); job 3: he(j,k) <- the east neighbour's
// p(1,j,k).  Tiles have the same even sizes, so a cell's colour and split position along a
// shared face are the same in both tiles; p(ip,j,k) has colour (j+k)&1 (first i 2, entry
// ip/2 - 1) and p(1,j,k) colour (1+j+k)&1 (entry 0).  Sources are interior entries, targets
// halo entries: no job reads what another writes. ----
__global__ __launch_bounds__(256) void k_split_exchange(Dims d, SplitTiles T) {
  const int q = blockIdx.y, job = blockIdx.z;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const int nr = T.nr, nj = T.nj;
  if (job < 2) {  // four entries per thread (rows are float4-aligned: nr % 4 == 0)
    const int ng = nr / 4;
    if (t >= 2L * d.kp * ng) return;
    const int ii = 4 * (int)(t % ng);
    const int rest = (int)(t / ng);
    const int c = rest & 1, k = 1 + (rest >> 1);
    const int src = job == 0 ? T.n[q] : T.s[q];
    const int srow = job == 0 ? 0 : d.jp - 1;
    const int drow = job == 0 ? d.jp : d.jp + 1;
    const long lev = (long)nr * nj * (k - 1);
    *reinterpret_cast<float4*>(T.p[q][c] + ii + nr * drow + lev) =
        *reinterpret_cast<const float4*>(T.p[src][c] + ii + nr * srow + lev);
  } else {
    if (t >= (long)d.jp * d.kp) return;
    const int j = 1 + (int)(t % d.jp), k = 1 + (int)(t / d.jp);
    const long row = (long)nr * ((j - 1) + (long)nj * (k - 1));
    if (job == 2) {
      if (T.hw[q]) T.hw[q][t] = T.p[T.w[q]][(j + k) & 1][(d.ip - 2) / 2 + row];
    } else if (T.he[q]) {
      T.he[q][t] = T.p[T.e[q]][(1 + j + k) & 1][row];
    }
  }
}

void launch_split_exchange(hipStream_t st, Dims d, const SplitTiles& T) {
  const long n = std::max(2L * d.kp * (T.nr / 4), (long)d.jp * d.kp);
  hipLaunchKernelGGL(k_split_exchange, dim3((unsigned)((n + 255) / 256), T.ntiles, 4), dim3(256),
                     0, st, d, T);
}

// ---- split copy -> p over the box [0,ip+1]x[0,jp+1]x[0,kp+1], optionally minus pav
// (press.f95:135-146): every halo cell takes its mirror interior cell, M(0)=1, M(ip+1)=ip in i,
// M(0)=jp, M(jp+1)=1 in j, M(0)=1, M(kp+1)=kp in k, as boundp1 then boundp2 leave them ----
__global__ __launch_bounds__(256) void k_split_gather_adj(Dims d, Fields F, SplitSOR sp,
                                                          DevScalars* sc, int subtract_pav,
                                                          int set_carry) {
  const int a = blockIdx.x * 64 + threadIdx.x;
  const int b = blockIdx.y * 4 + threadIdx.y;
  const int c0 = 2 * blockIdx.z;  // levels c0 and c0 + 1: both loads before the stores
  if (set_carry && a == 0 && b == 0 && c0 == 0) sc->pav_carry = sc->pav;
  if (a > d.ip + 1 || b > d.jp + 1) return;
  const int i = (a == 0) ? 1 : ((a == d.ip + 1) ? d.ip : a);
  const int j = (b == 0) ? d.jp : ((b == d.jp + 1) ? 1 : b);
  float v[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = min(c0 + q, d.kp + 1);
    const int k = (c == 0) ? 1 : ((c == d.kp + 1) ? d.kp : c);
    v[q] = sp.p[(i + j + k) & 1][split_index(sp, i, j, k)];
  }
  const float pav = subtract_pav ? sc->pav : 0.0f;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = c0 + q;
    if (c > d.kp + 1) break;
    F.p[d.P(a, b, c)] = subtract_pav ? v[q] - pav : v[q];
  }
}

void launch_split_gather_adj(hipStream_t st, Dims d, Fields F, SplitSOR sp, DevScalars* sc,
                             int subtract_pav, int set_carry) {
  dim3 grid((d.ip + 2 + 63) / 64, (d.jp + 2 + 3) / 4, (d.kp + 3) / 2);
  hipLaunchKernelGGL(k_split_gather_adj, grid, dim3(64, 4), 0, st, d, F, sp, sc, subtract_pav,
                     set_carry);
}

// F.rhs over the interior from the split rhs arrays sp.r (the caller passes the resident
// solve's adjusted rhs there)
__global__ __launch_bounds__(256) void k_split_gather_rhs(Dims d, Fields F, SplitSOR sp) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  F.rhs[d.R(i, j, k)] = sp.r[(i + j + k) & 1][split_index(sp, i, j, k)];
}

void launch_split_gather_rhs(hipStream_t st, Dims d, Fields F, SplitSOR sp) {
  dim3 grid((d.ip + 63) / 64, (d.jp + 3) / 4, d.kp);
  hipLaunchKernelGGL(k_split_gather_rhs, grid, dim3(64, 4), 0, st, d, F, sp);
}

// ---- transposes of the colour-split solve (the adjoint form's FP32 reverse) ----
// One thread per interior cell (i,j,k) of the Fortran box, i fastest.

// gather: every box cell b of [0,ip+1]x[0,jp+1]x[0,kp+1] took the split value of its mirror
// interior cell, so the split adjoint of cell x is the sum of pb over the cells mirroring x
// (x itself, and across the faces it touches: i = 0 / ip+1, j = jp+1 / 0, k = 0 / kp+1), and
// those pb are consumed.
__global__ __launch_bounds__(256) void k_split_gather_t(Dims d, float* pb, SplitSOR sp) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int is[2] = {i, i == 1 ? 0 : (i == d.ip ? d.ip + 1 : -1)};
  const int js[2] = {j, j == 1 ? d.jp + 1 : (j == d.jp ? 0 : -1)};
  const int ks[2] = {k, k == 1 ? 0 : (k == d.kp ? d.kp + 1 : -1)};
  float v = 0.0f;
  for (int c = 0; c < 2; ++c) {
    if (ks[c] < 0) continue;
    for (int b = 0; b < 2; ++b) {
      if (js[b] < 0) continue;
      for (int a = 0; a < 2; ++a) {
        if (is[a] < 0) continue;
        const long o = d.P(is[a], js[b], ks[c]);
        v = v + pb[o];
        pb[o] = 0.0f;
      }
    }
  }
  const int cc = (i + j + k) & 1;
  const int e = split_index(sp, i, j, k);
  sp.p[cc][e] = v;
  sp.r[cc][e] = 0.0f;
}

// One colour sweep p_c <- p_c + omega (cn1 (sum_n a_n q_n - r) - p_c) reads the other colour
// (q_n) and, across i and k faces, its own value (the mirror image); its transpose, applied in
// reverse sweep order, gives each colour-c cell y, from its adjoint P(y):
//   rhs adjoint:       R(y) += -omega cn1(y) P(y)
//   own:               P(y) <- s(y) P(y), s = 1 - omega + omega cn1(y) (sum of a_n over images)
//   other colour x:    P(x) += omega cn1(y) a_n(y) P(y) for each neighbour x that y reads.
// The kernel for the colour-c cells y combines the own part of one sweep's transpose with the
// other-colour part of the preceding one (in reverse order): it gathers, for y, the terms of
// its other-colour neighbours x (x reads y with x's coefficient toward y: c2l from the west
// neighbour, c2s from the east, c3l from the south, c3s from the north, cn4l(k-1) from below,
// cn4s(k+1) from above, each times omega cn1 of x's level), so every cell is written by one
// thread and no atomics are needed.
// A thread takes four consecutive entries 4g..4g+3 of a colour-c row (j, k), cells
// i = f + 2 (4g + q) with f the row's first i of the colour, through float4 loads of its own
// row and of the other colour's rows at j, j-1, j+1, k-1, k+1 (the sweep's access pattern);
// each cell's terms in the same order as a one-cell form would take them.
__global__ __launch_bounds__(256) void k_split_sweep_t(Dims d, Grid1D g, SplitSOR sp, float om,
                                                       int c, int mode, int nthreads) {
  // items: group grp of row (j, k), grp fastest, dealt densely over the workgroups in the
  // XCD-aware order of the forward sweep; quotients by multiply-high (sor_split_divisors)
  const int ng = sp.nr >> 2;
  const int t = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (t >= nthreads) return;
  const int row = (int)__umulhi((unsigned)t, sp.m_ng);
  const int grp = t - row * ng;
  const int kk = (int)__umulhi((unsigned)row, sp.m_jp);
  const int j = 1 + row - kk * d.jp, k = 1 + kk;
  const int f = first_i(c, j, k);
  const int i0 = f + 8 * grp;  // the group's first cell
  if (i0 > d.ip) return;
  auto ld4 = [](const float* p) { return *reinterpret_cast<const float4*>(p); };
  const float* po = sp.p[c ^ 1];
  const int lev = sp.nr * sp.nj;
  const int base = sp.nr * ((j - 1) + sp.nj * (k - 1)) + 4 * grp;
  const float c2l = g.cn2l[0], c2s = g.cn2s[0], c3l = g.cn3l[0], c3s = g.cn3s[0];
  const float cn1 = sp.cn1k[k];
  const float4 own4 = ld4(sp.p[c] + base);
  const float own[4] = {own4.x, own4.y, own4.z, own4.w};
  float out[4] = {own[0], own[1], own[2], own[3]};
  if (mode != 0) {  // own coefficient and the rhs adjoint
    const float4 r4 = ld4(sp.r[c] + base);
    float rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 2 * q;
      if (i > d.ip) continue;
      float img = 0.0f;
      if (i == 1) img = img + c2s;
      if (i == d.ip) img = img + c2l;
      if (k == 1) img = img + g.cn4s[k - 1];
      if (k == d.kp) img = img + g.cn4l[k - 1];
      const float sc = (1.0f - om) + om * (cn1 * img);
      rv[q] = rv[q] - om * (cn1 * own[q]);
      out[q] = sc * own[q];
    }
    *reinterpret_cast<float4*>(sp.r[c] + base) = make_float4(rv[0], rv[1], rv[2], rv[3]);
  }
  if (mode != 2) {  // the other colour's cells that read these
    const int jn = (j == d.jp) ? 1 : j + 1, js = (j == 1) ? d.jp : j - 1;
    const float4 xm = ld4(po + base);
    const float4 xs = ld4(po + base + sp.nr * (js - j));
    const float4 xn = ld4(po + base + sp.nr * (jn - j));
    const float4 xd = (k > 1) ? ld4(po + base - lev) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 xu = (k < d.kp) ? ld4(po + base + lev) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // other-colour entry 4g-1 (west of the first cell, f == 1) or 4g+4 (east of the last, f == 2)
    const float xo = (f == 1) ? (grp > 0 ? po[base - 1] : 0.0f)
                              : (4 * grp + 4 < sp.nr ? po[base + 4] : 0.0f);
    const float m[6] = {xo, xm.x, xm.y, xm.z, xm.w, xo};  // m[1 + q]: entry 4g+q
    const float sv[4] = {xs.x, xs.y, xs.z, xs.w}, nv[4] = {xn.x, xn.y, xn.z, xn.w};
    const float dv[4] = {xd.x, xd.y, xd.z, xd.w}, uv[4] = {xu.x, xu.y, xu.z, xu.w};
    const float cdn = (k > 1) ? sp.cn1k[k - 1] : 0.0f, c4d = (k > 1) ? g.cn4l[k - 2] : 0.0f;
    const float cup = (k < d.kp) ? sp.cn1k[k + 1] : 0.0f, c4u = (k < d.kp) ? g.cn4s[k] : 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 2 * q;
      if (i > d.ip) continue;
      const float pw = (f == 1) ? m[q] : m[1 + q];      // other colour at i - 1
      const float pe = (f == 1) ? m[1 + q] : m[2 + q];  // other colour at i + 1
      float gsum = 0.0f;
      if (i > 1) gsum = gsum + c2l * pw;
      if (i < d.ip) gsum = gsum + c2s * pe;
      gsum = gsum + c3l * sv[q];
      gsum = gsum + c3s * nv[q];
      float v = om * (cn1 * gsum);
      if (k > 1) v = v + om * (cdn * (c4d * dv[q]));
      if (k < d.kp) v = v + om * (cup * (c4u * uv[q]));
      out[q] = out[q] + v;
    }
  }
  *reinterpret_cast<float4*>(sp.p[c] + base) = make_float4(out[0], out[1], out[2], out[3]);
}

__global__ __launch_bounds__(256) void k_split_scatter_t(Dims d, float* pb, float* rb, SplitSOR sp) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.ip || j > d.jp) return;
  const int cc = (i + j + k) & 1;
  const int e = split_index(sp, i, j, k);
  pb[d.P(i, j, k)] = pb[d.P(i, j, k)] + sp.p[cc][e];
  rb[d.R(i, j, k)] = rb[d.R(i, j, k)] + sp.r[cc][e];
}

void launch_split_gather_t(hipStream_t st, Dims d, float* pb, SplitSOR sp) {
  dim3 grid((d.ip + 63) / 64, (d.jp + 3) / 4, d.kp);
  hipLaunchKernelGGL(k_split_gather_t, grid, dim3(64, 4), 0, st, d, pb, sp);
}

void launch_split_sweep_t(hipStream_t st, Dims d, Grid1D g, SplitSOR sp, float omega, int c,
                          int mode) {
  int ng, nt, nwg;
  split_threads(d, sp, ng, nt, nwg);
  hipLaunchKernelGGL(k_split_sweep_t, dim3(nwg), dim3(256), 0, st, d, g, sp, omega, c, mode, nt);
}

void launch_split_scatter_t(hipStream_t st, Dims d, float* pb, float* rb, SplitSOR sp) {
  dim3 grid((d.ip + 63) / 64, (d.jp + 3) / 4, d.kp);
  hipLaunchKernelGGL(k_split_scatter_t, grid, dim3(64, 4), 0, st, d, pb, rb, sp);
}

}  // namespace les
