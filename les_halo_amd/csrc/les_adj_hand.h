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
// k_vel2_bw forms them from the cov adjoints alone.

// Fortran layout of a 3-D array: element (i, j, k) at (i - i0) + ni * ((j - j0) + nj * (k - k0))
struct Lay {
  int i0, j0, k0, ni, nj;
  __device__ __forceinline__ long at(int i, int j, int k) const {
    return (long)(i - i0) + (long)ni * ((long)(j - j0) + (long)nj * (long)(k - k0));
  }
};

// nou, diu, cov q: (lo:im+2, 0:jm+2, 0:km+2), lo = -1 for q = 1, 5 and 0 otherwise
__device__ __forceinline__ Lay lay_vel(int q, int im, int jm) {
  const int lo = (q == 1 || q == 5) ? -1 : 0;
  return Lay{lo, 0, 0, im + 3 - lo, jm + 3};
}
// u, v: (0:im+1, -1:jm+1, 0:km+1); w: (0:im+1, -1:jm+1, -1:km+1); f, g, h: (0:im, 0:jm, 0:km)
__device__ __forceinline__ Lay lay_uv(int im, int jm) { return Lay{0, -1, 0, im + 2, jm + 3}; }
__device__ __forceinline__ Lay lay_w(int im, int jm) { return Lay{0, -1, -1, im + 2, jm + 3}; }
__device__ __forceinline__ Lay lay_fgh(int im, int jm) { return Lay{0, 0, 0, im + 1, jm + 1}; }

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
  // the f, g (box IN) and h (box KL) adjoints at (i, j, k) and at its -1 neighbours
  auto ld = [&](const T* y, int a, int b, int c, int khi) {
    return in_box(a, b, c, im, jm, 1, khi) ? y[F.at(a, b, c)] : T(0);
  };
  const T fc = ld(B.f, i, j, k, km), fw = ld(B.f, i - 1, j, k, km), fs = ld(B.f, i, j - 1, k, km),
          fd = ld(B.f, i, j, k - 1, km);
  const T gc = ld(B.g, i, j, k, km), gw = ld(B.g, i - 1, j, k, km), gs = ld(B.g, i, j - 1, k, km),
          gd = ld(B.g, i, j, k - 1, km);
  const T hc = ld(B.h, i, j, k, km - 1), hw = ld(B.h, i - 1, j, k, km - 1),
          hs = ld(B.h, i, j - 1, k, km - 1), hd = ld(B.h, i, j, k - 1, km - 1);
  // a term whose point lies outside the box reads no grid value beyond the arrays' bounds
  const bool ic = i <= im, jc = j <= jm, kc = k <= km, kh = k <= km - 1;
  const T two(2), half(0.5), vn = P.vn;
  T cb[10], db[10];
  // f: covx1 = (dx1(i+1) cov1(i) + dx1(i) cov1(i+1)) / (dx1(i) + dx1(i+1)), covy1, covz1 the
  // means of cov2 along j and cov3 along k; dfu1 = 2 (diu1(i+1) - diu1(i)) / (dx1(i) + dx1(i+1))
  // + (diu2(j+1) - diu2(j)) / dy1(j) + (diu3(k+1) - diu3(k)) / dzn(k)
  cb[1] = -((ic ? fc * dx1(i + 1) / (dx1(i) + dx1(i + 1)) : T(0)) +
            (i > 1 ? fw * dx1(i - 1) / (dx1(i - 1) + dx1(i)) : T(0)));
  cb[2] = -half * (fc + fs);
  cb[3] = -half * (fc + fd);
  db[1] = vn * ((i > 1 ? two * fw / (dx1(i - 1) + dx1(i)) : T(0)) -
                (ic ? two * fc / (dx1(i) + dx1(i + 1)) : T(0)));
  db[2] = vn * ((j > 1 ? fs / dy1(j - 1) : T(0)) - (jc ? fc / dy1(j) : T(0)));
  db[3] = vn * ((k > 1 ? fd / dzn(k - 1) : T(0)) - (kc ? fc / dzn(k) : T(0)));
  // g: covx1 the mean of cov4 along i, covy1 = (dy1(j+1) cov5(j) + dy1(j) cov5(j+1)) /
  // (dy1(j) + dy1(j+1)), covz1 the mean of cov6 along k; dfv1 = (diu4(i+1) - diu4(i)) / dx1(i)
  // + 2 (diu5(j+1) - diu5(j)) / (dy1(j) + dy1(j+1)) + (diu6(k+1) - diu6(k)) / dzn(k)
  cb[4] = -half * (gc + gw);
  cb[5] = -((jc ? gc * dy1(j + 1) / (dy1(j) + dy1(j + 1)) : T(0)) +
            (j > 1 ? gs * dy1(j - 1) / (dy1(j - 1) + dy1(j)) : T(0)));
  cb[6] = -half * (gc + gd);
  db[4] = vn * ((i > 1 ? gw / dx1(i - 1) : T(0)) - (ic ? gc / dx1(i) : T(0)));
  db[5] = vn * ((j > 1 ? two * gs / (dy1(j - 1) + dy1(j)) : T(0)) -
                (jc ? two * gc / (dy1(j) + dy1(j + 1)) : T(0)));
  db[6] = vn * ((k > 1 ? gd / dzn(k - 1) : T(0)) - (kc ? gc / dzn(k) : T(0)));
  // h (k <= km-1): covx1, covy1 the means of cov7 along i and cov8 along j, covz1 =
  // (dzn(k+1) cov9(k) + dzn(k) cov9(k+1)) / (dzn(k) + dzn(k+1)); dfw1 = (diu7(i+1) - diu7(i)) /
  // dx1(i) + (diu8(j+1) - diu8(j)) / dy1(j) + (diu9(k+1) - diu9(k)) / dzs(k)
  cb[7] = -half * (hc + hw);
  cb[8] = -half * (hc + hs);
  cb[9] = -((kh ? hc * dzn(k + 1) / (dzn(k) + dzn(k + 1)) : T(0)) +
            (k > 1 ? hd * dzn(k - 1) / (dzn(k - 1) + dzn(k)) : T(0)));
  db[7] = vn * ((i > 1 ? hw / dx1(i - 1) : T(0)) - (ic ? hc / dx1(i) : T(0)));
  db[8] = vn * ((j > 1 ? hs / dy1(j - 1) : T(0)) - (jc ? hc / dy1(j) : T(0)));
  db[9] = vn * ((k > 1 ? hd / dzs(k - 1) : T(0)) - (kh ? hc / dzs(k) : T(0)));
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
  const VelPtrs<T> Cb = cov_of(B), Db = diu_of(B), Nf = nou_of(A), Df = diu_of(A);
  // the adjoints of nou q and diu q at point (a, b, c) of their box (0 outside it)
  auto noub = [&](int q, int a, int b, int c) {
    if (!in_box(a, b, c, im, jm, 1, vel_khi(q, km))) return T(0);
    const long e = lay_vel(q, im, jm).at(a, b, c);
    return Cb.a[q][e] * Df.a[q][e];
  };
  auto diub = [&](int q, int a, int b, int c) {
    if (!in_box(a, b, c, im, jm, 1, vel_khi(q, km))) return T(0);
    const long e = lay_vel(q, im, jm).at(a, b, c);
    return Db.a[q][e] + Cb.a[q][e] * Nf.a[q][e];
  };
  auto inq = [&](int q, int a, int b, int c) { return in_box(a, b, c, im, jm, 1, vel_khi(q, km)); };
  const T two(2), half(0.5);
  // u: nou1 = (u(i-1) + u(i))/2, diu1 = (u(i) - u(i-1))/dx1(i), diu2 = 2 (u(j) - u(j-1)) /
  // (dy1(j-1) + dy1(j)), diu3 = (u(k) - u(k-1))/dzs(k-1), nou4 = (dy1(j+1) u(i-1,j) + dy1(j)
  // u(i-1,j+1)) / (dy1(j) + dy1(j+1)), nou7 = (dzn(k+1) u(i-1,k) + dzn(k) u(i-1,k+1)) /
  // (dzn(k) + dzn(k+1))
  T ub = half * (noub(1, i, j, k) + noub(1, i + 1, j, k));
  if (inq(1, i, j, k)) ub += diub(1, i, j, k) / dx1(i);
  if (inq(1, i + 1, j, k)) ub -= diub(1, i + 1, j, k) / dx1(i + 1);
  if (inq(2, i, j, k)) ub += two * diub(2, i, j, k) / (dy1(j - 1) + dy1(j));
  if (inq(2, i, j + 1, k)) ub -= two * diub(2, i, j + 1, k) / (dy1(j) + dy1(j + 1));
  if (inq(3, i, j, k)) ub += diub(3, i, j, k) / dzs(k - 1);
  if (inq(3, i, j, k + 1)) ub -= diub(3, i, j, k + 1) / dzs(k);
  if (inq(4, i + 1, j, k)) ub += noub(4, i + 1, j, k) * dy1(j + 1) / (dy1(j) + dy1(j + 1));
  if (inq(4, i + 1, j - 1, k)) ub += noub(4, i + 1, j - 1, k) * dy1(j - 1) / (dy1(j - 1) + dy1(j));
  if (inq(7, i + 1, j, k)) ub += noub(7, i + 1, j, k) * dzn(k + 1) / (dzn(k) + dzn(k + 1));
  if (inq(7, i + 1, j, k - 1)) ub += noub(7, i + 1, j, k - 1) * dzn(k - 1) / (dzn(k - 1) + dzn(k));
  // v: nou5 = (v(j-1) + v(j))/2, diu5 = (v(j) - v(j-1))/dy1(j), nou2 = (dx1(i+1) v(i,j-1) +
  // dx1(i) v(i+1,j-1)) / (dx1(i) + dx1(i+1)), diu4 = 2 (v(i) - v(i-1)) / (dx1(i-1) + dx1(i)),
  // diu6 = (v(k) - v(k-1))/dzs(k-1), nou8 = (dzn(k+1) v(j-1,k) + dzn(k) v(j-1,k+1)) /
  // (dzn(k) + dzn(k+1))
  T vb = half * (noub(5, i, j, k) + noub(5, i, j + 1, k));
  if (inq(5, i, j, k)) vb += diub(5, i, j, k) / dy1(j);
  if (inq(5, i, j + 1, k)) vb -= diub(5, i, j + 1, k) / dy1(j + 1);
  if (inq(2, i, j + 1, k)) vb += noub(2, i, j + 1, k) * dx1(i + 1) / (dx1(i) + dx1(i + 1));
  if (inq(2, i - 1, j + 1, k)) vb += noub(2, i - 1, j + 1, k) * dx1(i - 1) / (dx1(i - 1) + dx1(i));
  if (inq(4, i, j, k)) vb += two * diub(4, i, j, k) / (dx1(i - 1) + dx1(i));
  if (inq(4, i + 1, j, k)) vb -= two * diub(4, i + 1, j, k) / (dx1(i) + dx1(i + 1));
  if (inq(6, i, j, k)) vb += diub(6, i, j, k) / dzs(k - 1);
  if (inq(6, i, j, k + 1)) vb -= diub(6, i, j, k + 1) / dzs(k);
  if (inq(8, i, j + 1, k)) vb += noub(8, i, j + 1, k) * dzn(k + 1) / (dzn(k) + dzn(k + 1));
  if (inq(8, i, j + 1, k - 1)) vb += noub(8, i, j + 1, k - 1) * dzn(k - 1) / (dzn(k - 1) + dzn(k));
  // w: nou9 = (w(k-1) + w(k))/2, diu9 = (w(k) - w(k-1))/dzn(k), nou3 = (dx1(i+1) w(i,k-1) +
  // dx1(i) w(i+1,k-1)) / (dx1(i) + dx1(i+1)), nou6 = (dy1(j+1) w(j,k-1) + dy1(j) w(j+1,k-1)) /
  // (dy1(j) + dy1(j+1)), diu7 = 2 (w(i) - w(i-1)) / (dx1(i-1) + dx1(i)), diu8 = 2 (w(j) -
  // w(j-1)) / (dy1(j-1) + dy1(j))
  T wb = half * (noub(9, i, j, k) + noub(9, i, j, k + 1));
  if (inq(9, i, j, k)) wb += diub(9, i, j, k) / dzn(k);
  if (inq(9, i, j, k + 1)) wb -= diub(9, i, j, k + 1) / dzn(k + 1);
  if (inq(3, i, j, k + 1)) wb += noub(3, i, j, k + 1) * dx1(i + 1) / (dx1(i) + dx1(i + 1));
  if (inq(3, i - 1, j, k + 1)) wb += noub(3, i - 1, j, k + 1) * dx1(i - 1) / (dx1(i - 1) + dx1(i));
  if (inq(6, i, j, k + 1)) wb += noub(6, i, j, k + 1) * dy1(j + 1) / (dy1(j) + dy1(j + 1));
  if (inq(6, i, j - 1, k + 1)) wb += noub(6, i, j - 1, k + 1) * dy1(j - 1) / (dy1(j - 1) + dy1(j));
  if (inq(7, i, j, k)) wb += two * diub(7, i, j, k) / (dx1(i - 1) + dx1(i));
  if (inq(7, i + 1, j, k)) wb -= two * diub(7, i + 1, j, k) / (dx1(i) + dx1(i + 1));
  if (inq(8, i, j, k)) wb += two * diub(8, i, j, k) / (dy1(j - 1) + dy1(j));
  if (inq(8, i, j + 1, k)) wb -= two * diub(8, i, j + 1, k) / (dy1(j) + dy1(j + 1));
  const Lay U = lay_uv(im, jm), W = lay_w(im, jm), F = lay_fgh(im, jm);
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
