// Hand-written transposes of the adjoint form (included by les_adj.hip after the generated
// statement kernels).  Each replaces the generated transposes of a block of consecutive
// statements with a few fused gather-form kernels: every thread owns one element of each
// adjoint it writes and adds up what the block's statements read from that element, so there are
// no atomic adds, and the block's intermediate adjoints stay in registers where the statements
// allow it.  In exact arithmetic they equal the statement transposes (the FP64 test against
// them is the check); in floating point only the order of the additions differs.
//
// The vel2 + velFG block (vel2.f95:48-188, velFG.f95:56-105): nou1..9, diu1..9, cov1..9 on their
// boxes, the halo copies of vel2.f95:125-188, then f, g, h.  Its transpose:
//   k_velfg_bw      f, g, h adjoints -> cov adjoints (stored), diu adjoints (added)
//   k_vel_halo_bw   the halo copies' transposes: each halo element's adjoint into its source
//   k_vel2_bw       cov adjoints -> nou, diu adjoints (pointwise, in registers) -> u, v, w
//                   adjoints; f, g, h adjoints zeroed on the points velFG writes
// The nou adjoints other than through cov are zero (no statement reads nou but cov), so
// k_vel2_bw forms them from the cov adjoints alone, and it recomputes nou from u, v, w: the
// transposes read no nou array.

// Fortran layout of a 3-D array: element (i, j, k) at (i - i0) + ni * ((j - j0) + nj * (k - k0))
struct Lay {
  int i0, j0, k0, ni, nj;
  __device__ __forceinline__ long at(int i, int j, int k) const {
    return (long)(i - i0) + (long)ni * ((long)(j - j0) + (long)nj * (long)(k - k0));
  }
};

// nou, diu, cov q: (lo:im+2, 0:jm+2, 0:km+2), lo = -1 for q = 1, 5 and 0 otherwise
__host__ __device__ inline Lay lay_vel(int q, int im, int jm) {
  const int lo = (q == 1 || q == 5) ? -1 : 0;
  return Lay{lo, 0, 0, im + 3 - lo, jm + 3};
}
// u, v: (0:im+1, -1:jm+1, 0:km+1); w: (0:im+1, -1:jm+1, -1:km+1); f, g, h: (0:im, 0:jm, 0:km)
__host__ __device__ inline Lay lay_uv(int im, int jm) { return Lay{0, -1, 0, im + 2, jm + 3}; }
__host__ __device__ inline Lay lay_w(int im, int jm) { return Lay{0, -1, -1, im + 2, jm + 3}; }
__host__ __device__ inline Lay lay_fgh(int im, int jm) { return Lay{0, 0, 0, im + 1, jm + 1}; }

// the boxes of the block's statements: IN (1:im, 1:jm, 1:km), K1 (k to km+1), KL (k to km-1)
__device__ __forceinline__ bool in_box(int i, int j, int k, int im, int jm, int klo, int khi) {
  return i >= 1 && i <= im && j >= 1 && j <= jm && k >= klo && k <= khi;
}
// box of nou q, diu q, cov q (vel2.f95:48-123): K1 for q = 3, 6; KL for q = 7, 8; else IN
__device__ __forceinline__ int vel_khi(int q, int km) {
  return (q == 3 || q == 6) ? km + 1 : ((q == 7 || q == 8) ? km - 1 : km);
}

template <typename T>
struct VelPtrs {  // the nine arrays of one kind, indexable by q (slot 0 unused)
  T* a[10];
};
template <typename T>
__device__ __forceinline__ VelPtrs<T> diu_of(const AdjArr<T>& X) {
  return VelPtrs<T>{{nullptr, X.diu1, X.diu2, X.diu3, X.diu4, X.diu5, X.diu6, X.diu7, X.diu8, X.diu9}};
}
template <typename T>
__device__ __forceinline__ VelPtrs<T> nou_of(const AdjArr<T>& X) {
  return VelPtrs<T>{{nullptr, X.nou1, X.nou2, X.nou3, X.nou4, X.nou5, X.nou6, X.nou7, X.nou8, X.nou9}};
}
template <typename T>
__device__ __forceinline__ VelPtrs<T> cov_of(const AdjArr<T>& X) {
  return VelPtrs<T>{{nullptr, X.cov1, X.cov2, X.cov3, X.cov4, X.cov5, X.cov6, X.cov7, X.cov8, X.cov9}};
}

// ---- velFG.f95:59-105 transposed.  Thread (i, j, k), 1 <= i <= im+1, 1 <= j <= jm+1,
// 1 <= k <= km+1: the f, g, h points that read element (i, j, k) of a cov or diu array are
// (i, j, k) itself and its -1 neighbour along the array's direction of differencing. ----
template <typename T>
__global__ __launch_bounds__(256) void k_velfg_bw(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P, int im,
                                                  int jm, int km) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > im + 1 || j > jm + 1 || k > km + 1) return;
  const Lay F = lay_fgh(im, jm);
  auto dx1 = [&](int a) { return A.dx1[a + 1]; };
  auto dy1 = [&](int b) { return A.dy1[b]; };
  auto dzn = [&](int c) { return A.dzn[c + 1]; };
  auto dzs = [&](int c) { return A.dzs[c + 1]; };
  // the f, g (box IN) and h (box KL) adjoints at (i, j, k) and at its -1 neighbours, 0 outside
  // the box (loaded at the nearest point of the box either way, so no load waits on a branch)
  auto ld = [&](const T* y, int a, int b, int c, int khi) {
    const bool ok = in_box(a, b, c, im, jm, 1, khi);
    const T v = y[F.at(min(max(a, 1), im), min(max(b, 1), jm), min(max(c, 1), khi))];
    return ok ? v : T(0);
  };
  const T fc = ld(B.f, i, j, k, km), fw = ld(B.f, i - 1, j, k, km), fs = ld(B.f, i, j - 1, k, km),
          fd = ld(B.f, i, j, k - 1, km);
  const T gc = ld(B.g, i, j, k, km), gw = ld(B.g, i - 1, j, k, km), gs = ld(B.g, i, j - 1, k, km),
          gd = ld(B.g, i, j, k - 1, km);
  const T hc = ld(B.h, i, j, k, km - 1), hw = ld(B.h, i - 1, j, k, km - 1),
          hs = ld(B.h, i, j - 1, k, km - 1), hd = ld(B.h, i, j, k - 1, km - 1);
  // a term whose point lies outside the box has a zero adjoint; its grid values are read at
  // indices inside the arrays' bounds (i+1, j+1 clamped)
  const int ic = min(i, im), jc = min(j, jm);
  const T two(2), half(0.5), vn = P.vn;
  T cb[10], db[10];
  // f: covx1 = (dx1(i+1) cov1(i) + dx1(i) cov1(i+1)) / (dx1(i) + dx1(i+1)), covy1, covz1 the
  // means of cov2 along j and cov3 along k; dfu1 = 2 (diu1(i+1) - diu1(i)) / (dx1(i) + dx1(i+1))
  // + (diu2(j+1) - diu2(j)) / dy1(j) + (diu3(k+1) - diu3(k)) / dzn(k)
  cb[1] = -(fc * dx1(ic + 1) / (dx1(ic) + dx1(ic + 1)) + fw * dx1(i - 1) / (dx1(i - 1) + dx1(i)));
  cb[2] = -half * (fc + fs);
  cb[3] = -half * (fc + fd);
  db[1] = vn * (two * fw / (dx1(i - 1) + dx1(i)) - two * fc / (dx1(ic) + dx1(ic + 1)));
  db[2] = vn * (fs / dy1(j - 1) - fc / dy1(j));
  db[3] = vn * (fd / dzn(k - 1) - fc / dzn(k));
  // g: covx1 the mean of cov4 along i, covy1 = (dy1(j+1) cov5(j) + dy1(j) cov5(j+1)) /
  // (dy1(j) + dy1(j+1)), covz1 the mean of cov6 along k; dfv1 = (diu4(i+1) - diu4(i)) / dx1(i)
  // + 2 (diu5(j+1) - diu5(j)) / (dy1(j) + dy1(j+1)) + (diu6(k+1) - diu6(k)) / dzn(k)
  cb[4] = -half * (gc + gw);
  cb[5] = -(gc * dy1(jc + 1) / (dy1(jc) + dy1(jc + 1)) + gs * dy1(j - 1) / (dy1(j - 1) + dy1(j)));
  cb[6] = -half * (gc + gd);
  db[4] = vn * (gw / dx1(i - 1) - gc / dx1(i));
  db[5] = vn * (two * gs / (dy1(j - 1) + dy1(j)) - two * gc / (dy1(jc) + dy1(jc + 1)));
  db[6] = vn * (gd / dzn(k - 1) - gc / dzn(k));
  // h (k <= km-1): covx1, covy1 the means of cov7 along i and cov8 along j, covz1 =
  // (dzn(k+1) cov9(k) + dzn(k) cov9(k+1)) / (dzn(k) + dzn(k+1)); dfw1 = (diu7(i+1) - diu7(i)) /
  // dx1(i) + (diu8(j+1) - diu8(j)) / dy1(j) + (diu9(k+1) - diu9(k)) / dzs(k)
  cb[7] = -half * (hc + hw);
  cb[8] = -half * (hc + hs);
  cb[9] = -(hc * dzn(k + 1) / (dzn(k) + dzn(k + 1)) + hd * dzn(k - 1) / (dzn(k - 1) + dzn(k)));
  db[7] = vn * (hw / dx1(i - 1) - hc / dx1(i));
  db[8] = vn * (hs / dy1(j - 1) - hc / dy1(j));
  db[9] = vn * (hd / dzs(k - 1) - hc / dzs(k));
  const VelPtrs<T> C = cov_of(B), D = diu_of(B);
#pragma unroll
  for (int q = 1; q <= 9; ++q) {
    const long e = lay_vel(q, im, jm).at(i, j, k);
    C.a[q][e] = cb[q];  // the cov adjoints are zero before (nothing later in the step reads cov)
    D.a[q][e] += db[q];
  }
}

// ---- the halo copies' transposes (vel2.f95:125-188), in two launches: first the four copies
// made last (diu2, diu3 at i = im+1 for j, k to jm+1, km+1; diu4, diu6 at j = 0 for i, k to
// im+1, km+1), whose sources include halo elements the other copies write; then the 27 others
// (nou, diu, cov 1, 4, 7 at i = im+1 from i = im; 2, 5, 8 at j = 0 from jm and at jm+1 from 1,
// over k to km, km-1 for 7, 8), independent of one another.  src += dst, dst = 0. ----
template <typename T>
__global__ __launch_bounds__(256) void k_vel_halo_bw(AdjArr<T> B, int im, int jm, int km,
                                                     int late) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;  // the copy
  const VelPtrs<T> D = diu_of(B), N = nou_of(B), C = cov_of(B);
  int q, ext_a, ext_b, kind;  // kind 0: i = im+1 <- im; 1: j = 0 <- jm; 2: j = jm+1 <- 1
  T* y;
  if (late) {
    const int qs[4] = {2, 3, 4, 6};
    q = qs[c];
    kind = c < 2 ? 0 : 1;
    ext_a = kind == 0 ? jm + 1 : im + 1;  // j (kind 0) or i (kind 1) runs from 1
    ext_b = km + 1;
    y = D.a[q];
  } else {
    const int base = c / 3, kindof = c % 3;  // c = 3 * (array q's copy group) + which of nou/diu/cov
    // groups: 0..2 copies at i = im+1 of q = 1, 4, 7; 3..5 at j = 0 of q = 2, 5, 8;
    // 6..8 at j = jm+1 of q = 2, 5, 8
    const int grp = base % 3, sel = base / 3;
    q = sel == 0 ? (grp == 0 ? 1 : grp == 1 ? 4 : 7) : (grp == 0 ? 2 : grp == 1 ? 5 : 8);
    kind = sel;
    ext_a = sel == 0 ? jm : im;
    ext_b = (q == 7 || q == 8) ? km - 1 : km;
    y = kindof == 0 ? N.a[q] : (kindof == 1 ? D.a[q] : C.a[q]);
  }
  if (t >= ext_a * ext_b) return;
  const int a = 1 + t % ext_a, k = 1 + t / ext_a;
  const Lay L = lay_vel(q, im, jm);
  long dst, src;
  if (kind == 0) {
    dst = L.at(im + 1, a, k);
    src = L.at(im, a, k);
  } else if (kind == 1) {
    dst = L.at(a, 0, k);
    src = L.at(a, jm, k);
  } else {
    dst = L.at(a, jm + 1, k);
    src = L.at(a, 1, k);
  }
  y[src] += y[dst];
  y[dst] = T(0);
}

// ---- vel2.f95:48-123 transposed: cov q = nou q * diu q gives nou q's adjoint cov_b * diu q
// and adds cov_b * nou q to diu q's; the nou and diu statements (means and differences of u, v,
// w) are gathered into the u, v, w adjoints.  Thread (i, j, k) over 0..im+1, 0..jm+1, 0..km+1
// owns u, v, w (i, j, k).  Each term names the statement point pt that reads the element and
// the coefficient the statement gives it there. ----
template <typename T>
__global__ __launch_bounds__(256) void k_vel2_bw(AdjArr<T> A, AdjArr<T> B, int im, int jm, int km) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i > im + 1 || j > jm + 1 || k > km + 1) return;
  auto dx1 = [&](int a) { return A.dx1[a + 1]; };
  auto dy1 = [&](int b) { return A.dy1[b]; };
  auto dzn = [&](int c) { return A.dzn[c + 1]; };
  auto dzs = [&](int c) { return A.dzs[c + 1]; };
  const T two(2), half(0.5);
  const VelPtrs<T> Cb = cov_of(B), Db = diu_of(B), Df = diu_of(A);
  const Lay U = lay_uv(im, jm), W = lay_w(im, jm), F = lay_fgh(im, jm);
  // nou q at (a, b, c) from u, v, w (vel2.f95:48-123; the statements' expressions), so the
  // recomputation need not keep the nou arrays
  auto nou = [&](int q, int a, int b, int c) {
    switch (q) {
      case 1: return (A.u[U.at(a - 1, b, c)] + A.u[U.at(a, b, c)]) / two;
      case 5: return (A.v[U.at(a, b - 1, c)] + A.v[U.at(a, b, c)]) / two;
      case 9: return (A.w[W.at(a, b, c - 1)] + A.w[W.at(a, b, c)]) / two;
      case 2: return (dx1(a + 1) * A.v[U.at(a, b - 1, c)] + dx1(a) * A.v[U.at(a + 1, b - 1, c)]) /
                     (dx1(a) + dx1(a + 1));
      case 3: return (dx1(a + 1) * A.w[W.at(a, b, c - 1)] + dx1(a) * A.w[W.at(a + 1, b, c - 1)]) /
                     (dx1(a) + dx1(a + 1));
      case 4: return (dy1(b + 1) * A.u[U.at(a - 1, b, c)] + dy1(b) * A.u[U.at(a - 1, b + 1, c)]) /
                     (dy1(b) + dy1(b + 1));
      case 6: return (dy1(b + 1) * A.w[W.at(a, b, c - 1)] + dy1(b) * A.w[W.at(a, b + 1, c - 1)]) /
                     (dy1(b) + dy1(b + 1));
      case 7: return (dzn(c + 1) * A.u[U.at(a - 1, b, c)] + dzn(c) * A.u[U.at(a - 1, b, c + 1)]) /
                     (dzn(c) + dzn(c + 1));
      default: return (dzn(c + 1) * A.v[U.at(a, b - 1, c)] + dzn(c) * A.v[U.at(a, b - 1, c + 1)]) /
                      (dzn(c) + dzn(c + 1));
    }
  };
  // term: the adjoint of nou q (isn = 1) or diu q at the statement point (a, b, c) times the
  // coefficient the statement gives the element there (kind, at that point), 0 if the point is
  // outside the statement's box.  The loads are made at the nearest point of the box whatever
  // the test says, so the terms' loads issue without waiting on branches.
  auto term = [&](int q, bool isn, int a, int b, int c, int kind) {
    const bool ok = in_box(a, b, c, im, jm, 1, vel_khi(q, km));
    a = min(max(a, 1), im);
    b = min(max(b, 1), jm);
    c = min(max(c, 1), vel_khi(q, km));
    const long e = lay_vel(q, im, jm).at(a, b, c);
    const T cb = Cb.a[q][e];
    const T adj = isn ? cb * Df.a[q][e] : Db.a[q][e] + cb * nou(q, a, b, c);
    T co;
    switch (kind) {
      case 0: co = half; break;
      case 1: co = T(1) / dx1(a); break;
      case 2: co = T(1) / dy1(b); break;
      case 3: co = T(1) / dzn(c); break;
      case 4: co = T(1) / dzs(c - 1); break;
      case 5: co = two / (dx1(a - 1) + dx1(a)); break;
      case 6: co = two / (dy1(b - 1) + dy1(b)); break;
      case 7: co = dx1(a + 1) / (dx1(a) + dx1(a + 1)); break;
      case 8: co = dx1(a) / (dx1(a) + dx1(a + 1)); break;
      case 9: co = dy1(b + 1) / (dy1(b) + dy1(b + 1)); break;
      case 10: co = dy1(b) / (dy1(b) + dy1(b + 1)); break;
      case 11: co = dzn(c + 1) / (dzn(c) + dzn(c + 1)); break;
      default: co = dzn(c) / (dzn(c) + dzn(c + 1)); break;
    }
    return ok ? adj * co : T(0);
  };
  // u: nou1 = (u(i-1) + u(i))/2, diu1 = (u(i) - u(i-1))/dx1(i), diu2 = 2 (u(j) - u(j-1)) /
  // (dy1(j-1) + dy1(j)), diu3 = (u(k) - u(k-1))/dzs(k-1), nou4 = (dy1(j+1) u(i-1,j) + dy1(j)
  // u(i-1,j+1)) / (dy1(j) + dy1(j+1)), nou7 = (dzn(k+1) u(i-1,k) + dzn(k) u(i-1,k+1)) /
  // (dzn(k) + dzn(k+1)); each coefficient at the statement point
  const T ub = term(1, true, i, j, k, 0) + term(1, true, i + 1, j, k, 0) +
               term(1, false, i, j, k, 1) - term(1, false, i + 1, j, k, 1) +
               term(2, false, i, j, k, 6) - term(2, false, i, j + 1, k, 6) +
               term(3, false, i, j, k, 4) - term(3, false, i, j, k + 1, 4) +
               term(4, true, i + 1, j, k, 9) + term(4, true, i + 1, j - 1, k, 10) +
               term(7, true, i + 1, j, k, 11) + term(7, true, i + 1, j, k - 1, 12);
  // v: nou5 = (v(j-1) + v(j))/2, diu5 = (v(j) - v(j-1))/dy1(j), nou2 = (dx1(i+1) v(i,j-1) +
  // dx1(i) v(i+1,j-1)) / (dx1(i) + dx1(i+1)), diu4 = 2 (v(i) - v(i-1)) / (dx1(i-1) + dx1(i)),
  // diu6 = (v(k) - v(k-1))/dzs(k-1), nou8 = (dzn(k+1) v(j-1,k) + dzn(k) v(j-1,k+1)) /
  // (dzn(k) + dzn(k+1))
  const T vb = term(5, true, i, j, k, 0) + term(5, true, i, j + 1, k, 0) +
               term(5, false, i, j, k, 2) - term(5, false, i, j + 1, k, 2) +
               term(2, true, i, j + 1, k, 7) + term(2, true, i - 1, j + 1, k, 8) +
               term(4, false, i, j, k, 5) - term(4, false, i + 1, j, k, 5) +
               term(6, false, i, j, k, 4) - term(6, false, i, j, k + 1, 4) +
               term(8, true, i, j + 1, k, 11) + term(8, true, i, j + 1, k - 1, 12);
  // w: nou9 = (w(k-1) + w(k))/2, diu9 = (w(k) - w(k-1))/dzn(k), nou3 = (dx1(i+1) w(i,k-1) +
  // dx1(i) w(i+1,k-1)) / (dx1(i) + dx1(i+1)), nou6 = (dy1(j+1) w(j,k-1) + dy1(j) w(j+1,k-1)) /
  // (dy1(j) + dy1(j+1)), diu7 = 2 (w(i) - w(i-1)) / (dx1(i-1) + dx1(i)), diu8 = 2 (w(j) -
  // w(j-1)) / (dy1(j-1) + dy1(j))
  const T wb = term(9, true, i, j, k, 0) + term(9, true, i, j, k + 1, 0) +
               term(9, false, i, j, k, 3) - term(9, false, i, j, k + 1, 3) +
               term(3, true, i, j, k + 1, 7) + term(3, true, i - 1, j, k + 1, 8) +
               term(6, true, i, j, k + 1, 9) + term(6, true, i, j - 1, k + 1, 10) +
               term(7, false, i, j, k, 5) - term(7, false, i + 1, j, k, 5) +
               term(8, false, i, j, k, 6) - term(8, false, i, j + 1, k, 6);
  B.u[U.at(i, j, k)] += ub;
  B.v[U.at(i, j, k)] += vb;
  B.w[W.at(i, j, k)] += wb;
  // velFG writes f, g on IN and h on KL: their adjoints before it are zero there
  if (in_box(i, j, k, im, jm, 1, km)) {
    const long e = F.at(i, j, k);
    B.f[e] = T(0);
    B.g[e] = T(0);
    if (k <= km - 1) B.h[e] = T(0);
  }
}

// The vel2 + velFG block of the statement list (ADJ_nou1 .. ADJ_velfg_h, in this order)
constexpr int VEL_BLOCK_FIRST = ADJ_nou1, VEL_BLOCK_LAST = ADJ_velfg_h;
static_assert(VEL_BLOCK_LAST - VEL_BLOCK_FIRST == 60, "the vel2 + velFG block has 61 statements");

template <typename T>
void launch_vel_block_bw(hipStream_t st, const AdjArr<T>& A, const AdjArr<T>& B,
                         const AdjPar<T>& P, int im, int jm, int km) {
  hipLaunchKernelGGL(k_velfg_bw<T>, dim3((im + 1 + 63) / 64, (jm + 1 + 3) / 4, km + 1),
                     dim3(64, 4), 0, st, A, B, P, im, jm, km);
  const int late_n = std::max(jm + 1, im + 1) * (km + 1);
  hipLaunchKernelGGL(k_vel_halo_bw<T>, dim3((late_n + 255) / 256, 4), dim3(256), 0, st, B, im,
                     jm, km, 1);
  const int early_n = std::max(jm, im) * km;
  hipLaunchKernelGGL(k_vel_halo_bw<T>, dim3((early_n + 255) / 256, 27), dim3(256), 0, st, B, im,
                     jm, km, 0);
  hipLaunchKernelGGL(k_vel2_bw<T>, dim3((im + 2 + 63) / 64, (jm + 2 + 3) / 4, km + 2),
                     dim3(64, 4), 0, st, A, B, im, jm, km);
}

// ---- les.f95:33-182 (calc_sm, boundsm, visc) transposed ----------------------------------------
// The block: calc_sm (sm from the diu), boundsm's six copies, then f, g, h += the viscous
// terms.  Each viscous term is (flux2 - flux1) / h along x, y, z with flux = evs * dsum: evs an
// interpolation of sm (one value, or four with the products of two linear weights), dsum one or
// two diu values (2 diu for the own direction).  Its transpose:
//   k_visc_sm_bw    sm's adjoint gathered from every flux that reads sm (alpha * dsum * weight)
//   k_bsm_bw        boundsm's copies transposed (k faces, then j, then i: the reverse order)
//   k_calc_sm_bw    per cell G = smb cd^2 / (2 sqrt(s)) and the three cross terms of s, kept in
//                   a scratch array; sm's adjoint zeroed where calc_sm writes it
//   k_les_diu_bw    the diu adjoints gathered from the fluxes (alpha * evs * coefficient) and
//                   from calc_sm (through the scratch array)
// with alpha = +-(f, g or h adjoint) / h.

__host__ __device__ inline Lay lay_sm(int im, int jm) { return Lay{-1, -1, 0, im + 3, jm + 3}; }

// one viscous flux of les.f95:78-180: statement s (0 f, 1 g, 2 h), face t (0 x2, 1 x1, 2 y2,
// 3 y1, 4 z2, 5 z1); term n = 6 s + t
struct ViscTerm {
  int quad;          // 0: evs = sm at (oi, oj, ok); 1: sum over e, e' of W_outer(e) W_inner(e') sm
  int oi, oj, ok;    // offset of the single sm value
  int wo, wi;        // weight kinds: 0 xi, 1 xm (along i), 2 yi, 3 ym (j), 4 zi (k), 5 zm (k)
  int oo[2], io[2];  // offsets of the outer / inner weights' sm values along their axes
  int nd;            // diu values in dsum
  int dq[2], di[2], dj[2], dk[2];
  int dc;            // coefficient of dsum's diu (2 for the own direction, else 1)
};

__host__ __device__ constexpr ViscTerm visc_term(int n) {
  switch (n) {
    // visc_f: visux2 = sm(i+1) 2 diu1(i+1), visux1 = sm 2 diu1, visuy2/1 = evsy (diu2 + diu4),
    // visuz2/1 = evsz (diu3 + diu7)
    case 0: return {0, 1, 0, 0, 0, 0, {0, 0}, {0, 0}, 1, {1, 0}, {1, 0}, {0, 0}, {0, 0}, 2};
    case 1: return {0, 0, 0, 0, 0, 0, {0, 0}, {0, 0}, 1, {1, 0}, {0, 0}, {0, 0}, {0, 0}, 2};
    case 2: return {1, 0, 0, 0, 2, 0, {0, 1}, {0, 1}, 2, {2, 4}, {0, 1}, {1, 0}, {0, 0}, 1};
    case 3: return {1, 0, 0, 0, 2, 0, {-1, 0}, {0, 1}, 2, {2, 4}, {0, 1}, {0, -1}, {0, 0}, 1};
    case 4: return {1, 0, 0, 0, 4, 0, {0, 1}, {0, 1}, 2, {3, 7}, {0, 1}, {0, 0}, {1, 0}, 1};
    case 5: return {1, 0, 0, 0, 5, 0, {-1, 0}, {0, 1}, 2, {3, 7}, {0, 1}, {0, 0}, {0, -1}, 1};
    // visc_g: evsx (diu2 + diu4) with xi / xm, sm(j+1) 2 diu5(j+1), sm 2 diu5, evsz (diu6 + diu8)
    case 6: return {1, 0, 0, 0, 2, 0, {0, 1}, {0, 1}, 2, {2, 4}, {0, 1}, {1, 0}, {0, 0}, 1};
    case 7: return {1, 0, 0, 0, 2, 1, {0, 1}, {-1, 0}, 2, {2, 4}, {-1, 0}, {1, 0}, {0, 0}, 1};
    case 8: return {0, 0, 1, 0, 0, 0, {0, 0}, {0, 0}, 1, {5, 0}, {0, 0}, {1, 0}, {0, 0}, 2};
    case 9: return {0, 0, 0, 0, 0, 0, {0, 0}, {0, 0}, 1, {5, 0}, {0, 0}, {0, 0}, {0, 0}, 2};
    case 10: return {1, 0, 0, 0, 4, 0, {0, 1}, {0, 1}, 2, {6, 8}, {0, 0}, {0, 1}, {1, 0}, 1};
    case 11: return {1, 0, 0, 0, 5, 0, {-1, 0}, {0, 1}, 2, {6, 8}, {0, 0}, {0, 1}, {0, -1}, 1};
    // visc_h: evsx (diu3 + diu7) with zi x xi / xm, evsy (diu6 + diu8) with zi x yi / ym,
    // sm(k+1) 2 diu9(k+1), sm 2 diu9
    case 12: return {1, 0, 0, 0, 4, 0, {0, 1}, {0, 1}, 2, {3, 7}, {0, 1}, {0, 0}, {1, 0}, 1};
    case 13: return {1, 0, 0, 0, 4, 1, {0, 1}, {-1, 0}, 2, {3, 7}, {-1, 0}, {0, 0}, {1, 0}, 1};
    case 14: return {1, 0, 0, 0, 4, 2, {0, 1}, {0, 1}, 2, {6, 8}, {0, 0}, {0, 1}, {1, 0}, 1};
    case 15: return {1, 0, 0, 0, 4, 3, {0, 1}, {-1, 0}, 2, {6, 8}, {0, 0}, {-1, 0}, {1, 0}, 1};
    case 16: return {0, 0, 0, 1, 0, 0, {0, 0}, {0, 0}, 1, {9, 0}, {0, 0}, {0, 0}, {1, 0}, 2};
    default: return {0, 0, 0, 0, 0, 0, {0, 0}, {0, 0}, 1, {9, 0}, {0, 0}, {0, 0}, {0, 0}, 2};
  }
}

// The interpolation weights and reciprocal spacings the fluxes use, per index of their axis
// (k_les_tab): for i in 1..im, xi0 = dx1(i+1)/(dx1(i)+dx1(i+1)), xi1 = dx1(i)/(dx1(i)+dx1(i+1))
// (the weights of xi's first and second value), xm0 = dx1(i)/(dx1(i-1)+dx1(i)), xm1 =
// dx1(i-1)/(dx1(i-1)+dx1(i)), 1/dx1(i); the same along j with dy1 and along k with dzn.
// Table (axis a, kind c) holds n_a = (im, jm or km) + 2 entries at ((a-base) + c n_a).
template <typename T>
struct LesTab {
  const T* w;
  int im, jm, km;
  __device__ __forceinline__ T at(int axis, int c, int idx) const {
    const int nx = im + 2, ny = jm + 2, nz = km + 2;
    const int base = axis == 0 ? 0 : (axis == 1 ? 5 * nx : 5 * (nx + ny));
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    return w[base + c * n + idx];
  }
  // weight e of interpolation kind w at point (i, j, k)
  __device__ __forceinline__ T weight(int w, int e, int i, int j, int k) const {
    return at(w / 2, 2 * (w % 2) + e, w < 2 ? i : (w < 4 ? j : k));
  }
  __device__ __forceinline__ T rcp(int axis, int idx) const { return at(axis, 4, idx); }
  __device__ __forceinline__ bool in(int i, int j, int k) const { return in_box(i, j, k, im, jm, 1, km); }
  // the nearest cell of IN: the gathers load at it unconditionally and keep the term only for a
  // point inside IN, so the loads of all terms issue without waiting on branches
  __device__ __forceinline__ void clamp(int& i, int& j, int& k) const {
    i = min(max(i, 1), im);
    j = min(max(j, 1), jm);
    k = min(max(k, 1), km);
  }
};

template <typename T>
__global__ __launch_bounds__(256) void k_les_tab(T* w, const T* dx1, const T* dy1, const T* dzn,
                                                 int im, int jm, int km) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int nx = im + 2, ny = jm + 2, nz = km + 2;
  int axis, idx, n, m;
  const T* d;
  int off;  // array index of element 0 of d
  if (t < nx) {
    axis = 0, idx = t, n = nx, m = im, d = dx1, off = 1;
  } else if (t < nx + ny) {
    axis = 1, idx = t - nx, n = ny, m = jm, d = dy1, off = 0;
  } else if (t < nx + ny + nz) {
    axis = 2, idx = t - nx - ny, n = nz, m = km, d = dzn, off = 1;
  } else {
    return;
  }
  const int base = axis == 0 ? 0 : (axis == 1 ? 5 * nx : 5 * (nx + ny));
  T v[5] = {T(0), T(0), T(0), T(0), T(0)};
  if (idx >= 1 && idx <= m) {
    const T dm = d[idx - 1 + off], dc = d[idx + off], dp = d[idx + 1 + off];
    v[0] = dp / (dc + dp);
    v[1] = dc / (dc + dp);
    v[2] = dc / (dm + dc);
    v[3] = dm / (dm + dc);
    v[4] = T(1) / dc;
  }
  for (int c = 0; c < 5; ++c) w[base + c * n + idx] = v[c];
}

template <int Q, typename T>
__device__ __forceinline__ const T* diu_ptr(const AdjArr<T>& X) {
  if constexpr (Q == 1) return X.diu1;
  else if constexpr (Q == 2) return X.diu2;
  else if constexpr (Q == 3) return X.diu3;
  else if constexpr (Q == 4) return X.diu4;
  else if constexpr (Q == 5) return X.diu5;
  else if constexpr (Q == 6) return X.diu6;
  else if constexpr (Q == 7) return X.diu7;
  else if constexpr (Q == 8) return X.diu8;
  else return X.diu9;
}

// alpha of term N at point (a, b, c): +-(f, g or h adjoint) / (dx1(a), dy1(b) or dzn(c))
template <int N, typename T>
__device__ __forceinline__ T visc_alpha(const AdjArr<T>& B, const LesTab<T>& W, const Lay& F,
                                        int a, int b, int c) {
  constexpr int s = N / 6, t = N % 6;
  const long e = F.at(a, b, c);
  const T y = s == 0 ? B.f[e] : (s == 1 ? B.g[e] : B.h[e]);
  const T r = t < 2 ? W.rcp(0, a) : (t < 4 ? W.rcp(1, b) : W.rcp(2, c));
  return (t % 2 == 0 ? y : -y) * r;
}

// the offset of a weight kind's sm value o along its axis, as (di, dj, dk)
__host__ __device__ constexpr int axis_di(int w, int o) { return w <= 1 ? o : 0; }
__host__ __device__ constexpr int axis_dj(int w, int o) { return (w == 2 || w == 3) ? o : 0; }
__host__ __device__ constexpr int axis_dk(int w, int o) { return w >= 4 ? o : 0; }

// term N's share of sm(i, j, k)'s adjoint: alpha * dc * dsum * weight at every point reading it
template <int N, typename T>
__device__ __forceinline__ void visc_sm_term(T& acc, const AdjArr<T>& A, const AdjArr<T>& B,
                                             const LesTab<T>& W, const Lay& F, int i, int j,
                                             int k) {
  constexpr ViscTerm v = visc_term(N);
  const int im = W.im, jm = W.jm;
  auto dsum = [&](int a, int b, int c) {
    T ds = diu_ptr<v.dq[0], T>(A)[lay_vel(v.dq[0], im, jm).at(a + v.di[0], b + v.dj[0], c + v.dk[0])];
    if constexpr (v.nd == 2)
      ds += diu_ptr<v.dq[1], T>(A)[lay_vel(v.dq[1], im, jm).at(a + v.di[1], b + v.dj[1], c + v.dk[1])];
    return ds;
  };
  if constexpr (!v.quad) {
    int a = i - v.oi, b = j - v.oj, c = k - v.ok;
    const bool ok = W.in(a, b, c);
    W.clamp(a, b, c);
    const T t = visc_alpha<N>(B, W, F, a, b, c) * T(v.dc) * dsum(a, b, c);
    acc += ok ? t : T(0);
  } else {
#pragma unroll
    for (int eo = 0; eo < 2; ++eo)
#pragma unroll
      for (int ei = 0; ei < 2; ++ei) {
        int a = i - axis_di(v.wo, v.oo[eo]) - axis_di(v.wi, v.io[ei]);
        int b = j - axis_dj(v.wo, v.oo[eo]) - axis_dj(v.wi, v.io[ei]);
        int c = k - axis_dk(v.wo, v.oo[eo]) - axis_dk(v.wi, v.io[ei]);
        const bool ok = W.in(a, b, c);
        W.clamp(a, b, c);
        const T t = visc_alpha<N>(B, W, F, a, b, c) * T(v.dc) * dsum(a, b, c) *
                    W.weight(v.wo, eo, a, b, c) * W.weight(v.wi, ei, a, b, c);
        acc += ok ? t : T(0);
      }
  }
}

template <typename T, int... N>
__device__ __forceinline__ void visc_sm_terms(std::integer_sequence<int, N...>, T& acc,
                                              const AdjArr<T>& A, const AdjArr<T>& B,
                                              const LesTab<T>& W, const Lay& F, int i, int j,
                                              int k) {
  (visc_sm_term<N, T>(acc, A, B, W, F, i, j, k), ...);
}

template <typename T>
__global__ __launch_bounds__(256) void k_visc_sm_bw(AdjArr<T> A, AdjArr<T> B, LesTab<T> W) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  const int im = W.im, jm = W.jm, km = W.km;
  if (i > im + 1 || j > jm + 1 || k > km + 1) return;
  T acc = T(0);
  visc_sm_terms(std::make_integer_sequence<int, 18>{}, acc, A, B, W, lay_fgh(im, jm), i, j, k);
  B.sm[lay_sm(im, jm).at(i, j, k)] += acc;
}

// boundsm.f95:13-32 transposed, one phase per launch (0: k faces, 1: j faces, 2: i faces), both
// faces of the phase together (blockIdx.y)
template <typename T>
__global__ __launch_bounds__(256) void k_bsm_bw(T* smb, int im, int jm, int km, int phase) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int side = blockIdx.y;
  const Lay L = lay_sm(im, jm);
  long dst, src;
  T sgn = T(1);
  if (phase == 0) {  // sm(i, j, 0) = -sm(i, j, 1), sm(i, j, km+1) = sm(i, j, km): i 0..im+1, j -1..jm+1
    const int ni = im + 2;
    if (t >= ni * (jm + 3)) return;
    const int a = t % ni, b = -1 + t / ni;
    dst = L.at(a, b, side == 0 ? 0 : km + 1);
    src = L.at(a, b, side == 0 ? 1 : km);
    if (side == 0) sgn = T(-1);
  } else if (phase == 1) {  // sm(i, 0, k) = sm(i, 1, k), sm(i, jm+1, k) = sm(i, jm, k): i 0..im+1, k 0..km+1
    const int ni = im + 2;
    if (t >= ni * (km + 2)) return;
    const int a = t % ni, c = t / ni;
    dst = L.at(a, side == 0 ? 0 : jm + 1, c);
    src = L.at(a, side == 0 ? 1 : jm, c);
  } else {  // sm(0, j, k) = sm(1, j, k), sm(im+1, j, k) = sm(im, j, k): j -1..jm+1, k 0..km+1
    const int nj = jm + 3;
    if (t >= nj * (km + 2)) return;
    const int b = -1 + t % nj, c = t / nj;
    dst = L.at(side == 0 ? 0 : im + 1, b, c);
    src = L.at(side == 0 ? 1 : im, b, c);
  }
  smb[src] += sgn * smb[dst];
  smb[dst] = T(0);
}

// calc_sm (les.f95:40-56) at each interior cell: G = smb (cs0 delx1(k))^2 / (2 sqrt(s)) (0 if
// s = 0) and T12, T23, T13 = G (S12, S23, S13) / 2, the sums of the averaged cross derivatives
// whose squares s holds; stored per cell (scratch: G, T12, T23, T13 arrays of im*jm*km)
template <typename T>
__global__ __launch_bounds__(256) void k_calc_sm_bw(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P,
                                                    T* scratch, int im, int jm, int km) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > im || j > jm) return;
  const VelPtrs<T> D = diu_of(A);
  auto d = [&](int q, int a, int b, int c) { return D.a[q][lay_vel(q, im, jm).at(a, b, c)]; };
  const T q4(0.25);
  const T d1 = d(1, i, j, k), d5 = d(5, i, j, k), d9 = d(9, i, j, k);
  const T s12 = (d(2, i - 1, j, k) + d(2, i - 1, j + 1, k) + d(2, i, j, k) + d(2, i, j + 1, k)) * q4 +
                (d(4, i, j, k) + d(4, i, j - 1, k) + d(4, i + 1, j, k) + d(4, i + 1, j - 1, k)) * q4;
  const T s23 = (d(8, i, j, k) + d(8, i, j, k - 1) + d(8, i, j + 1, k) + d(8, i, j + 1, k - 1)) * q4 +
                (d(6, i, j - 1, k) + d(6, i, j - 1, k + 1) + d(6, i, j, k) + d(6, i, j, k + 1)) * q4;
  const T s13 = (d(3, i - 1, j, k) + d(3, i - 1, j, k + 1) + d(3, i, j, k) + d(3, i, j, k + 1)) * q4 +
                (d(7, i, j, k) + d(7, i, j, k - 1) + d(7, i + 1, j, k) + d(7, i + 1, j, k - 1)) * q4;
  const T s = T(2) * (d1 * d1 + d5 * d5 + d9 * d9) + s12 * s12 + s23 * s23 + s13 * s13;
  const long e = lay_sm(im, jm).at(i, j, k);
  const T cd = P.cs0 * A.delx1[k - 1];
  const T g = s > T(0) ? B.sm[e] * (cd * cd) / (T(2) * sqrt(s)) : T(0);
  B.sm[e] = T(0);
  const long n = (long)im * jm * km, c = (i - 1) + (long)im * ((j - 1) + (long)jm * (k - 1));
  scratch[c] = g;
  scratch[n + c] = T(0.5) * g * s12;
  scratch[2 * n + c] = T(0.5) * g * s23;
  scratch[3 * n + c] = T(0.5) * g * s13;
}

// term N's share of the diu adjoints at (i, j, k): alpha * dc * evs at every point whose dsum
// reads diu dq[m] there
template <int N, typename T>
__device__ __forceinline__ void visc_diu_term(T* acc, const AdjArr<T>& A, const AdjArr<T>& B,
                                              const LesTab<T>& W, const Lay& F, const Lay& S,
                                              int i, int j, int k) {
  constexpr ViscTerm v = visc_term(N);
#pragma unroll
  for (int m = 0; m < v.nd; ++m) {
    int a = i - v.di[m], b = j - v.dj[m], c = k - v.dk[m];
    const bool ok = W.in(a, b, c);
    W.clamp(a, b, c);
    T evs;
    if constexpr (!v.quad) {
      evs = A.sm[S.at(a + v.oi, b + v.oj, c + v.ok)];
    } else {
      evs = T(0);
#pragma unroll
      for (int eo = 0; eo < 2; ++eo)
#pragma unroll
        for (int ei = 0; ei < 2; ++ei)
          evs += W.weight(v.wo, eo, a, b, c) * W.weight(v.wi, ei, a, b, c) *
                 A.sm[S.at(a + axis_di(v.wo, v.oo[eo]) + axis_di(v.wi, v.io[ei]),
                           b + axis_dj(v.wo, v.oo[eo]) + axis_dj(v.wi, v.io[ei]),
                           c + axis_dk(v.wo, v.oo[eo]) + axis_dk(v.wi, v.io[ei]))];
    }
    const T t = visc_alpha<N>(B, W, F, a, b, c) * T(v.dc) * evs;
    acc[v.dq[m]] += ok ? t : T(0);
  }
}

template <typename T, int... N>
__device__ __forceinline__ void visc_diu_terms(std::integer_sequence<int, N...>, T* acc,
                                               const AdjArr<T>& A, const AdjArr<T>& B,
                                               const LesTab<T>& W, const Lay& F, const Lay& S,
                                               int i, int j, int k) {
  (visc_diu_term<N, T>(acc, A, B, W, F, S, i, j, k), ...);
}

template <typename T>
__global__ __launch_bounds__(256) void k_les_diu_bw(AdjArr<T> A, AdjArr<T> B, LesTab<T> W,
                                                    const T* scratch) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  const int im = W.im, jm = W.jm, km = W.km;
  if (i > im + 1 || j > jm + 1 || k > km + 1) return;
  T acc[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) acc[q] = T(0);
  visc_diu_terms(std::make_integer_sequence<int, 18>{}, acc, A, B, W, lay_fgh(im, jm),
                 lay_sm(im, jm), i, j, k);
  // calc_sm: diu1, 5, 9 at their own cell (4 G diu), the cross derivatives through T12, T23, T13
  const long n3 = (long)im * jm * km;
  auto sc = [&](int w, int a, int b, int c) {
    const bool ok = W.in(a, b, c);
    W.clamp(a, b, c);
    const T t = scratch[w * n3 + (a - 1) + (long)im * ((b - 1) + (long)jm * (c - 1))];
    return ok ? t : T(0);
  };
  if (W.in(i, j, k)) {
    const T g4 = T(4) * sc(0, i, j, k);
    acc[1] += g4 * A.diu1[lay_vel(1, im, jm).at(i, j, k)];
    acc[5] += g4 * A.diu5[lay_vel(5, im, jm).at(i, j, k)];
    acc[9] += g4 * A.diu9[lay_vel(9, im, jm).at(i, j, k)];
  }
  acc[2] += sc(1, i + 1, j, k) + sc(1, i + 1, j - 1, k) + sc(1, i, j, k) + sc(1, i, j - 1, k);
  acc[4] += sc(1, i, j, k) + sc(1, i, j + 1, k) + sc(1, i - 1, j, k) + sc(1, i - 1, j + 1, k);
  acc[8] += sc(2, i, j, k) + sc(2, i, j, k + 1) + sc(2, i, j - 1, k) + sc(2, i, j - 1, k + 1);
  acc[6] += sc(2, i, j + 1, k) + sc(2, i, j + 1, k - 1) + sc(2, i, j, k) + sc(2, i, j, k - 1);
  acc[3] += sc(3, i + 1, j, k) + sc(3, i + 1, j, k - 1) + sc(3, i, j, k) + sc(3, i, j, k - 1);
  acc[7] += sc(3, i, j, k) + sc(3, i, j, k + 1) + sc(3, i - 1, j, k) + sc(3, i - 1, j, k + 1);
  B.diu1[lay_vel(1, im, jm).at(i, j, k)] += acc[1];
  B.diu2[lay_vel(2, im, jm).at(i, j, k)] += acc[2];
  B.diu3[lay_vel(3, im, jm).at(i, j, k)] += acc[3];
  B.diu4[lay_vel(4, im, jm).at(i, j, k)] += acc[4];
  B.diu5[lay_vel(5, im, jm).at(i, j, k)] += acc[5];
  B.diu6[lay_vel(6, im, jm).at(i, j, k)] += acc[6];
  B.diu7[lay_vel(7, im, jm).at(i, j, k)] += acc[7];
  B.diu8[lay_vel(8, im, jm).at(i, j, k)] += acc[8];
  B.diu9[lay_vel(9, im, jm).at(i, j, k)] += acc[9];
}

// The les block of the statement list (ADJ_calc_sm .. ADJ_visc_h, in this order)
constexpr int LES_BLOCK_FIRST = ADJ_calc_sm, LES_BLOCK_LAST = ADJ_visc_h;
static_assert(LES_BLOCK_LAST - LES_BLOCK_FIRST == 9, "the les block has 10 statements");

template <typename T>
void launch_les_block_bw(hipStream_t st, const AdjArr<T>& A, const AdjArr<T>& B,
                         const AdjPar<T>& P, T* scratch, int im, int jm, int km) {
  // scratch: G, T12, T23, T13 (4 im*jm*km), then the weight tables
  T* tab = scratch + 4 * (long)im * jm * km;
  const int ntab = (im + 2) + (jm + 2) + (km + 2);
  hipLaunchKernelGGL(k_les_tab<T>, dim3((ntab + 255) / 256), dim3(256), 0, st, tab, A.dx1, A.dy1,
                     A.dzn, im, jm, km);
  const LesTab<T> W{tab, im, jm, km};
  const dim3 gx((im + 2 + 63) / 64, (jm + 2 + 3) / 4, km + 2), blk(64, 4);
  hipLaunchKernelGGL(k_visc_sm_bw<T>, gx, blk, 0, st, A, B, W);
  const int nk = (im + 2) * (jm + 3), nj = (im + 2) * (km + 2), ni = (jm + 3) * (km + 2);
  hipLaunchKernelGGL(k_bsm_bw<T>, dim3((nk + 255) / 256, 2), dim3(256), 0, st, B.sm, im, jm, km, 0);
  hipLaunchKernelGGL(k_bsm_bw<T>, dim3((nj + 255) / 256, 2), dim3(256), 0, st, B.sm, im, jm, km, 1);
  hipLaunchKernelGGL(k_bsm_bw<T>, dim3((ni + 255) / 256, 2), dim3(256), 0, st, B.sm, im, jm, km, 2);
  hipLaunchKernelGGL(k_calc_sm_bw<T>, dim3((im + 63) / 64, (jm + 3) / 4, km), blk, 0, st, A, B, P,
                     scratch, im, jm, km);
  hipLaunchKernelGGL(k_les_diu_bw<T>, gx, blk, 0, st, A, B, W, scratch);
}

// ---- the point-wise and two-point statements around the stencils, transposed ------------------
__host__ __device__ inline Lay lay_p(int im, int jm) { return Lay{0, 0, 0, im + 3, jm + 3}; }
__host__ __device__ inline Lay lay_in(int im, int jm) { return Lay{1, 1, 1, im, jm}; }  // fold, fd, cd
__host__ __device__ inline Lay lay_rhs(int im, int jm) { return Lay{0, 0, 0, im + 2, jm + 2}; }

// velnw.f95:27-59: u += dt (f - (p(i+1) - p(i)) / ro / dxs(i)) on IN, v alike along j, w along k
// for k <= km-1.  Transposed: f, g, h adjoints += dt times u's, v's, w's; p's adjoint gathers
// the pressure-gradient terms of the points that read it.  Thread: p element (i, j, k), 1..im+1,
// 1..jm+1, 1..km.
template <typename T>
__global__ __launch_bounds__(256) void k_velnw_bw(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P, int im,
                                                  int jm, int km) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > im + 1 || j > jm + 1 || k > km) return;
  const Lay U = lay_uv(im, jm), W = lay_w(im, jm), F = lay_fgh(im, jm);
  auto dxs = [&](int a) { return A.dxs[a]; };
  auto dys = [&](int b) { return A.dys[b]; };
  auto dzs = [&](int c) { return A.dzs[c + 1]; };
  const T c = P.dt / P.ro;
  // u, v (box IN) and w (box KL) adjoints at a point, 0 outside the box; loaded at the nearest
  // point of the box either way, so no load waits on a branch
  auto ld = [&](const T* y, const Lay& L, int a, int b, int cc, int khi) {
    const bool ok = in_box(a, b, cc, im, jm, 1, khi);
    const T v = y[L.at(min(max(a, 1), im), min(max(b, 1), jm), min(max(cc, 1), khi))];
    return ok ? v : T(0);
  };
  const T ub = ld(B.u, U, i, j, k, km), vb = ld(B.v, U, i, j, k, km), wb = ld(B.w, W, i, j, k, km - 1);
  const T uw = ld(B.u, U, i - 1, j, k, km), vs = ld(B.v, U, i, j - 1, k, km),
          wd = ld(B.w, W, i, j, k - 1, km - 1);
  const int ic = min(i, im), jc = min(j, jm), kc = min(k, km - 1);
  // the gradient terms: p(i) in u(i) (+) and in u(i-1) (-), alike along j and k
  T pb = c * ub / dxs(ic) + c * vb / dys(jc) + c * wb / dzs(kc);
  pb -= c * uw / dxs(i - 1) + c * vs / dys(j - 1) + c * wd / dzs(k - 1);
  if (in_box(i, j, k, im, jm, 1, km)) {
    const long e = F.at(i, j, k);
    B.f[e] += P.dt * ub;
    B.g[e] += P.dt * vb;
    if (k <= km - 1) B.h[e] += P.dt * wb;
  }
  B.p[lay_p(im, jm).at(i, j, k)] += pb;
}

// feedbf.f95:41-65 with the drag field: usum = (usum + u) bmask1, then f += alpha usum dt +
// beta cd u bmask1 (v, vsum, cmask1, g and w, wsum, dmask1, h alike), on IN.  Transposed per
// cell: the f adjoint stays; usum's gains alpha dt f_b, cd's beta u bmask1 f_b, u's beta cd
// bmask1 f_b; then usum's adjoint passes through the mask to the old usum and to u.
template <typename T>
__global__ __launch_bounds__(256) void k_feedbf_bw(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P, int im,
                                                   int jm, int km) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > im || j > jm) return;
  const Lay U = lay_uv(im, jm), W = lay_w(im, jm), F = lay_fgh(im, jm), C = lay_in(im, jm);
  const long e = F.at(i, j, k), ec = C.at(i, j, k);
  const T cd = A.cd[ec];
  const T m[3] = {A.bmask1[Lay{-1, 0, 0, im + 3, jm + 2}.at(i, j, k)],
                  A.cmask1[Lay{0, -1, 0, im + 2, jm + 3}.at(i, j, k)],
                  A.dmask1[Lay{0, 0, 0, im + 2, jm + 2}.at(i, j, k)]};
  T* tb[3] = {B.f, B.g, B.h};
  T* sb[3] = {B.usum, B.vsum, B.wsum};
  const T* a[3] = {A.u, A.v, A.w};
  T* ab[3] = {B.u, B.v, B.w};
  T cdb = T(0);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long ea = q < 2 ? U.at(i, j, k) : W.at(i, j, k);
    const T y = tb[q][e];
    const T s_b = sb[q][e] + P.alpha * P.dt * y;
    cdb += P.beta * a[q][ea] * m[q] * y;
    ab[q][ea] += P.beta * cd * m[q] * y + m[q] * s_b;
    sb[q][e] = m[q] * s_b;
  }
  B.cd[ec] += cdb;
}

// adam.f95:35-43: fd = f; f = 1.5 f - 0.5 fold; fold = fd (g, h alike), on IN.  Transposed per
// cell: f_b' = 1.5 f_b + fold_b + fd_b, fold_b' = -0.5 f_b, fd_b' = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_adam_bw(AdjArr<T> B, int im, int jm, int km) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > im || j > jm) return;
  const long e = lay_fgh(im, jm).at(i, j, k), eo = lay_in(im, jm).at(i, j, k);
  T* fb[3] = {B.f, B.g, B.h};
  T* ob[3] = {B.fold, B.gold, B.hold};
  T* db[3] = {B.fd, B.gd, B.hd};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T y = fb[q][e];
    fb[q][e] = T(1.5) * y + ob[q][eo] + db[q][eo];
    ob[q][eo] = T(-0.5) * y;
    db[q][eo] = T(0);
  }
}

// press.f95:59-64: rhs = div(u, v, w) on IN, then rhs = div(f, g, h) + rhs / dt.  Transposed:
// with r the rhs adjoint, f's gathers r(i)/dx1(i) - r(i+1)/dx1(i+1) (g, h alike along j, k) and
// u's the same of r / dt.  Thread: element (i, j, k), 0..im, 0..jm, 0..km.  The rhs adjoint is
// left as it is: no statement before these reads rhs.
template <typename T>
__global__ __launch_bounds__(256) void k_rhs_bw(AdjArr<T> A, AdjArr<T> B, AdjPar<T> P, int im,
                                                int jm, int km) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i > im || j > jm || k > km) return;
  const Lay R = lay_rhs(im, jm), U = lay_uv(im, jm), W = lay_w(im, jm), F = lay_fgh(im, jm);
  auto dx1 = [&](int a) { return A.dx1[a + 1]; };
  auto dy1 = [&](int b) { return A.dy1[b]; };
  auto dzn = [&](int c) { return A.dzn[c + 1]; };
  // the rhs adjoint at a point of IN, 0 outside (loaded at the nearest point of IN either way)
  auto r = [&](int a, int b, int c) {
    const bool ok = in_box(a, b, c, im, jm, 1, km);
    const T v = B.rhs[R.at(min(max(a, 1), im), min(max(b, 1), jm), min(max(c, 1), km))];
    return ok ? v : T(0);
  };
  const T rc = r(i, j, k), re = r(i + 1, j, k), rn = r(i, j + 1, k), ru = r(i, j, k + 1);
  // grid values at indices inside the arrays (a zero adjoint multiplies the clamped ones)
  const int ia = max(i, 1), ja = max(j, 1), ka = max(k, 1);
  const int ib = min(i + 1, im), jb = min(j + 1, jm), kb = min(k + 1, km);
  const T xf = rc / dx1(ia) - re / dx1(ib), yf = rc / dy1(ja) - rn / dy1(jb),
          zf = rc / dzn(ka) - ru / dzn(kb);
  const T xu = rc / P.dt / dx1(ia) - re / P.dt / dx1(ib), yu = rc / P.dt / dy1(ja) - rn / P.dt / dy1(jb),
          zu = rc / P.dt / dzn(ka) - ru / P.dt / dzn(kb);
  const long e = F.at(i, j, k);
  B.f[e] += xf;
  B.g[e] += yf;
  B.h[e] += zf;
  B.u[U.at(i, j, k)] += xu;
  B.v[U.at(i, j, k)] += yu;
  B.w[W.at(i, j, k)] += zu;
}

// blocks of consecutive statements with these transposes
static_assert(ADJ_velnw_w - ADJ_velnw_u == 2, "velnw block");
static_assert(ADJ_fb_h - ADJ_fb_usum == 5, "feedbf block");
static_assert(ADJ_adam_old_h - ADJ_adam_save_f == 8, "adam block");
static_assert(ADJ_rhs_fgh - ADJ_rhs_uvw == 1, "rhs block");

template <typename T>
void launch_point_block_bw(int first, hipStream_t st, const AdjArr<T>& A, const AdjArr<T>& B,
                           const AdjPar<T>& P, int im, int jm, int km) {
  const dim3 blk(64, 4);
  if (first == ADJ_velnw_u)
    hipLaunchKernelGGL(k_velnw_bw<T>, dim3((im + 1 + 63) / 64, (jm + 1 + 3) / 4, km), blk, 0, st, A,
                       B, P, im, jm, km);
  else if (first == ADJ_fb_usum)
    hipLaunchKernelGGL(k_feedbf_bw<T>, dim3((im + 63) / 64, (jm + 3) / 4, km), blk, 0, st, A, B, P,
                       im, jm, km);
  else if (first == ADJ_adam_save_f)
    hipLaunchKernelGGL(k_adam_bw<T>, dim3((im + 63) / 64, (jm + 3) / 4, km), blk, 0, st, B, im, jm,
                       km);
  else if (first == ADJ_rhs_uvw)
    hipLaunchKernelGGL(k_rhs_bw<T>, dim3((im + 1 + 63) / 64, (jm + 1 + 3) / 4, km + 1), blk, 0, st,
                       A, B, P, im, jm, km);
}

// ---- boundary copies and the scalar subtractions, transposed -----------------------------------
// bondv1_ext_wind_profile.f95's inflow (u = wind at i = 0, 1; v = w = 0 there): the wind
// adjoint gathers u's two inflow adjoints, which (and v's, w's) are consumed.  Thread (j, k).
template <typename T>
__global__ __launch_bounds__(256) void k_inflow_bw(AdjArr<T> B, int im, int jm, int km) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= jm * km) return;
  const int j = 1 + t % jm, k = 1 + t / jm;
  const Lay U = lay_uv(im, jm), W = lay_w(im, jm);
  T s = T(0);
  for (int i = 0; i <= 1; ++i) {
    s += B.u[U.at(i, j, k)];
    B.u[U.at(i, j, k)] = T(0);
    B.v[U.at(i, j, k)] = T(0);
    B.w[W.at(i, j, k)] = T(0);
  }
  B.wind[(j - 1) + (long)jm * (k - 1)] += s;
}

// bondv1.f95:81-137's faces: phase 0 the k faces (u, v at k = 0 the negated, at km+1 the copied
// level; w = 0 at k = 0 and km), phase 1 the periodic j faces (u, v for k 0..km+1, w for 0..km),
// each transposed (the k faces first: they read j faces the periodic copies wrote).
// blockIdx.y: the array and face.  w's k-face statements cover i = -1..im+1 and j = -1..jm+1,
// i = -1 meaning the element before w(0, j, k) in memory, as in the reference's loops.
template <typename T>
__global__ __launch_bounds__(256) void k_bc_uvw_bw(AdjArr<T> B, int im, int jm, int km, int phase) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int a = blockIdx.y >> 1, side = blockIdx.y & 1;  // a: 0 u, 1 v, 2 w
  T* y = a == 0 ? B.u : (a == 1 ? B.v : B.w);
  const Lay L = a == 2 ? lay_w(im, jm) : lay_uv(im, jm);
  if (phase == 0) {
    if (a < 2) {  // i 0..im+1, j 0..jm+1
      const int ni = im + 2;
      if (t >= ni * (jm + 2)) return;
      const int i = t % ni, j = t / ni;
      const long dst = L.at(i, j, side == 0 ? 0 : km + 1), src = L.at(i, j, side == 0 ? 1 : km);
      y[src] += side == 0 ? -y[dst] : y[dst];
      y[dst] = T(0);
    } else {  // i -1..im+1, j -1..jm+1
      const int ni = im + 3;
      if (t >= ni * (jm + 3)) return;
      const int i = -1 + t % ni, j = -1 + t / ni;
      y[L.at(i, j, side == 0 ? 0 : km)] = T(0);
    }
  } else {  // i 0..im+1, k 0..km+1 (w: 0..km); side 0: j = 0 from jm, 1: j = jm+1 from 1
    const int ni = im + 2, nk = a == 2 ? km + 1 : km + 2;
    if (t >= ni * nk) return;
    const int i = t % ni, k = t / ni;
    const long dst = L.at(i, side == 0 ? 0 : jm + 1, k), src = L.at(i, side == 0 ? jm : 1, k);
    y[src] += y[dst];
    y[dst] = T(0);
  }
}

// bondFG.f95:15-32: f(0) = f(1), g(j = 0) = g(jm), h(k = 0) = h(km) = 0, transposed
// (blockIdx.y: the statement)
template <typename T>
__global__ __launch_bounds__(256) void k_bfg_bw(AdjArr<T> B, int im, int jm, int km) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const Lay F = lay_fgh(im, jm);
  switch (blockIdx.y) {
    case 0: {  // j 1..jm, k 1..km
      if (t >= jm * km) return;
      const int j = 1 + t % jm, k = 1 + t / jm;
      B.f[F.at(1, j, k)] += B.f[F.at(0, j, k)];
      B.f[F.at(0, j, k)] = T(0);
      return;
    }
    case 1: {  // i 1..im, k 1..km
      if (t >= im * km) return;
      const int i = 1 + t % im, k = 1 + t / im;
      B.g[F.at(i, jm, k)] += B.g[F.at(i, 0, k)];
      B.g[F.at(i, 0, k)] = T(0);
      return;
    }
    default: {  // i 1..im, j 1..jm
      if (t >= im * jm) return;
      const int i = 1 + t % im, j = 1 + t / im;
      B.h[F.at(i, j, blockIdx.y == 2 ? 0 : km)] = T(0);
      return;
    }
  }
}

// press.f95:82 (rhs -= rhsav) and :139 (p -= pav) transposed: the adjoint stays, and the
// scalar's adjoint (slot q of the scalar adjoints) takes minus its sum over the interior.
// Thread: (i, j) with a chunk of 8 levels (one-dimensional workgroups, as scal_add needs).
template <typename T>
__global__ __launch_bounds__(256) void k_sub_scalar_bw(const T* y, Lay L, AdjScal<T>* SB, int q,
                                                       int im, int jm, int km) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int k0 = 1 + blockIdx.y * 8;
  T s = T(0);
  if (t < im * jm) {
    const int i = 1 + t % im, j = 1 + t / im;
    for (int k = k0; k < k0 + 8 && k <= km; ++k) s += y[L.at(i, j, k)];
  }
  scal_add(SB, q, -s);
}

static_assert(ADJ_inflow_w - ADJ_inflow_u == 2, "inflow block");
static_assert(ADJ_top_w - ADJ_perj0_u == 11, "bondv1 face block");
static_assert(ADJ_bfg_h1 - ADJ_bfg_f == 3, "bondfg block");

template <typename T>
void launch_bc_block_bw(int first, hipStream_t st, const AdjArr<T>& B, AdjScal<T>* SB, int im,
                        int jm, int km) {
  if (first == ADJ_inflow_u) {
    hipLaunchKernelGGL(k_inflow_bw<T>, dim3((jm * km + 255) / 256), dim3(256), 0, st, B, im, jm, km);
  } else if (first == ADJ_perj0_u) {
    const int n0 = (im + 3) * (jm + 3), n1 = (im + 2) * (km + 2);
    hipLaunchKernelGGL(k_bc_uvw_bw<T>, dim3((n0 + 255) / 256, 6), dim3(256), 0, st, B, im, jm, km, 0);
    hipLaunchKernelGGL(k_bc_uvw_bw<T>, dim3((n1 + 255) / 256, 6), dim3(256), 0, st, B, im, jm, km, 1);
  } else if (first == ADJ_bfg_f) {
    const int n = std::max(std::max(jm * km, im * km), im * jm);
    hipLaunchKernelGGL(k_bfg_bw<T>, dim3((n + 255) / 256, 4), dim3(256), 0, st, B, im, jm, km);
  } else if (first == ADJ_rhs_adj || first == ADJ_pav_adj) {
    const bool r = first == ADJ_rhs_adj;
    hipLaunchKernelGGL(k_sub_scalar_bw<T>, dim3((im * jm + 255) / 256, (km + 7) / 8), dim3(256), 0,
                       st, r ? B.rhs : B.p, r ? lay_rhs(im, jm) : lay_p(im, jm), SB, r ? 1 : 2, im,
                       jm, km);
  }
}
