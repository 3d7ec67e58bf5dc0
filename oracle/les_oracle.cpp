// CPU oracle for the LES_Halo timestep: a sequential restatement of the reference Fortran
// physics (src/Fortran/*.f95), used ONLY as the checker by tests/, by the smoke check and by
// the CPU-baseline leg of the benchmark.  Nothing in the product library links or calls it.
//
// Every routine keeps the Fortran array bounds, column-major layout (i fastest) and the
// left-to-right evaluation order of the reference expressions, so that an FP32 instantiation
// compiled without FP contraction reproduces the reference bit-for-bit.  An FP64 instantiation
// of the same code is the high-precision oracle.  Flat indexing reproduces the reference's
// out-of-bounds-in-i writes exactly (bondv1.f95:132-137 writes w(-1,j,k)).
//
// Reductions (rhsav, pav, the SOR residual) have two modes:
//   ORC_SUM_SEQ   : sequential accumulation in the working precision (press.f95:65-75,123-134,97)
//   ORC_SUM_EXACT : exact order-independent sum of the same FP32 terms, rounded once to FP32
//                   (the semantics the HIP path implements); FP64 instantiations treat it as SEQ.
//
// Parity pins: see tests/test_oracle_*.py (reference-build fixtures, closed forms, invariants).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

template <typename T>
struct A3 {
  T* d;
  int i0, j0, k0;
  long ni, nj;
  A3(void* p, int i0_, int i1, int j0_, int j1, int k0_)
      : d(static_cast<T*>(p)), i0(i0_), j0(j0_), k0(k0_), ni(i1 - i0_ + 1), nj(j1 - j0_ + 1) {}
  T& operator()(int i, int j, int k) const {
    return d[(long)(i - i0) + ni * ((long)(j - j0) + nj * (long)(k - k0))];
  }
};

template <typename T>
struct A1 {
  T* d;
  int lo;
  A1(void* p, int lo_) : d(static_cast<T*>(p)), lo(lo_) {}
  T& operator()(int i) const { return d[i - lo]; }
};

template <typename T>
struct A2 {  // wind_profile(1:jp,1:kp)
  T* d;
  long nj;
  A2(void* p, int jp) : d(static_cast<T*>(p)), nj(jp) {}
  T& operator()(int j, int k) const { return d[(long)(j - 1) + nj * (long)(k - 1)]; }
};

}  // namespace

extern "C" {

// Pointers to host arrays in the reference's Fortran bounds (main_ocl.f95:40-167).
typedef struct {
  int ip, jp, kp;
  void *p, *u, *v, *w, *usum, *vsum, *wsum, *f, *g, *h, *fold, *gold, *hold, *rhs, *sm;
  void* diu[9];
  void* nou[9];
  void* cov[9];
  void *amask1, *bmask1, *cmask1, *dmask1, *cn1;
  void *dx1, *dxs, *dy1, *dys, *dzn, *dzs, *z2;
  void *cn2l, *cn2s, *cn3l, *cn3s, *cn4l, *cn4s, *delx1, *wind;
  void *fx, *fy, *fz, *dfu1, *dfv1, *dfw1;
} orc_fields;

typedef struct {
  double dt, ro, vn, alpha, beta, cs0, omega, pjuge;
  int nmaxp;
  int sum_mode;  // 0 = sequential, 1 = exact
  double refine_tol;  // > 0: after the SOR, more blocks of nmaxp iterations until the relative
  int refine_max;     //      residual |rhs - L p|_2/|rhs|_2 <= refine_tol (at most refine_max)
  double refine_omega;  // > 0: relaxation of those further blocks (else omega)
} orc_params;

typedef struct {
  double uout, rhsav, area, pav, pco, sor;
  int sor_iters;
  double resid_l2;
  int refine_passes;
} orc_scalars;

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Exact summation of FP32 terms (oracle implementation: a 384-bit two's-complement integer in
// 32-bit limbs with full carry propagation; value = integer * 2^-149).
// ---------------------------------------------------------------------------------------------
namespace {

struct BigAcc {
  static const int NL = 12;
  uint32_t l[NL];
  BigAcc() { std::memset(l, 0, sizeof(l)); }
  void add_shifted(uint64_t mant, int shift, bool neg) {
    // build the 384-bit magnitude mant << shift, then add or subtract it
    uint32_t t[NL];
    std::memset(t, 0, sizeof(t));
    int limb = shift / 32, off = shift % 32;
    unsigned __int128 v = (unsigned __int128)mant << off;
    for (int q = 0; q < 3 && limb + q < NL; ++q) t[limb + q] = (uint32_t)(v >> (32 * q));
    if (neg) {  // two's complement of t
      uint64_t c = 1;
      for (int q = 0; q < NL; ++q) {
        uint64_t s = (uint64_t)(uint32_t)~t[q] + c;
        t[q] = (uint32_t)s;
        c = s >> 32;
      }
    }
    uint64_t c = 0;
    for (int q = 0; q < NL; ++q) {
      uint64_t s = (uint64_t)l[q] + t[q] + c;
      l[q] = (uint32_t)s;
      c = s >> 32;
    }
  }
  void add(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    uint32_t e = (b >> 23) & 0xff, m = b & 0x7fffff;
    if (e == 0xff) return;  // inf/nan are not summed exactly (never produced by finite inputs)
    if (e) m |= 0x800000u;
    if (!m) return;
    int shift = (e ? (int)e : 1) - 1;
    add_shifted(m, shift, (b >> 31) != 0);
  }
  // round-to-nearest-even conversion of the exact value to float
  float to_float() const {
    uint32_t mag[NL];
    bool neg = (l[NL - 1] >> 31) != 0;
    std::memcpy(mag, l, sizeof(mag));
    if (neg) {
      uint64_t c = 1;
      for (int q = 0; q < NL; ++q) {
        uint64_t s = (uint64_t)(uint32_t)~mag[q] + c;
        mag[q] = (uint32_t)s;
        c = s >> 32;
      }
    }
    int top = -1;
    for (int q = NL - 1; q >= 0 && top < 0; --q)
      if (mag[q]) top = q * 32 + 31 - __builtin_clz(mag[q]);
    if (top < 0) return 0.0f;
    auto bit = [&](int pos) -> uint32_t { return pos < 0 ? 0u : (mag[pos / 32] >> (pos % 32)) & 1u; };
    uint32_t out;
    if (top < 24) {  // exactly representable: integer * 2^-149 with < 24 bits
      uint32_t mm = mag[0];
      if (top >= 23) {  // normal with exponent field 1
        out = (1u << 23) | (mm & 0x7fffff);
      } else {
        out = mm;  // subnormal
      }
    } else {
      int drop = top - 23;  // bits below the kept 24
      uint64_t keep = 0;
      for (int q = 0; q < 24; ++q) keep |= (uint64_t)bit(drop + q) << q;
      uint32_t rnd = bit(drop - 1);
      bool sticky = false;
      for (int q = 0; q < drop - 1 && !sticky; ++q) sticky = bit(q) != 0;
      if (rnd && (sticky || (keep & 1))) keep += 1;
      int exp_field = drop + 1;  // value = keep * 2^(drop-149); keep in [2^23, 2^24]
      if (keep == (1u << 24)) {
        keep >>= 1;
        exp_field += 1;
      }
      if (exp_field >= 255) {
        out = 0x7f800000u;
      } else {
        out = ((uint32_t)exp_field << 23) | ((uint32_t)keep & 0x7fffff);
      }
    }
    if (neg) out |= 0x80000000u;
    float r;
    std::memcpy(&r, &out, 4);
    return r;
  }
};

template <typename T>
struct Summer {
  int mode;
  T seq;
  BigAcc big;
  explicit Summer(int m) : mode(m), seq(T(0)) {}
  void add(T x) {
    if (mode == 1 && sizeof(T) == 4)
      big.add((float)x);
    else
      seq = seq + x;
  }
  T value() const { return (mode == 1 && sizeof(T) == 4) ? (T)big.to_float() : seq; }
};

// ---------------------------------------------------------------------------------------------
// Field views
// ---------------------------------------------------------------------------------------------
template <typename T>
struct F {
  int ip, jp, kp, im, jm, km;
  A3<T> p, u, v, w, usum, vsum, wsum, f, g, h, fold, gold, hold, rhs, sm;
  A3<T> diu1, diu2, diu3, diu4, diu5, diu6, diu7, diu8, diu9;
  A3<T> nou1, nou2, nou3, nou4, nou5, nou6, nou7, nou8, nou9;
  A3<T> cov1, cov2, cov3, cov4, cov5, cov6, cov7, cov8, cov9;
  A3<T> amask1, bmask1, cmask1, dmask1, cn1;
  A3<T> fx, fy, fz, dfu1, dfv1, dfw1;
  A1<T> dx1, dxs, dy1, dys, dzn, dzs, z2, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s, delx1;
  A2<T> wind;
  explicit F(const orc_fields& s)
      : ip(s.ip), jp(s.jp), kp(s.kp), im(s.ip), jm(s.jp), km(s.kp),
        p(s.p, 0, ip + 2, 0, jp + 2, 0),
        u(s.u, 0, ip + 1, -1, jp + 1, 0),
        v(s.v, 0, ip + 1, -1, jp + 1, 0),
        w(s.w, 0, ip + 1, -1, jp + 1, -1),
        usum(s.usum, 0, ip, 0, jp, 0),
        vsum(s.vsum, 0, ip, 0, jp, 0),
        wsum(s.wsum, 0, ip, 0, jp, 0),
        f(s.f, 0, ip, 0, jp, 0),
        g(s.g, 0, ip, 0, jp, 0),
        h(s.h, 0, ip, 0, jp, 0),
        fold(s.fold, 1, ip, 1, jp, 1),
        gold(s.gold, 1, ip, 1, jp, 1),
        hold(s.hold, 1, ip, 1, jp, 1),
        rhs(s.rhs, 0, ip + 1, 0, jp + 1, 0),
        sm(s.sm, -1, ip + 1, -1, jp + 1, 0),
        diu1(s.diu[0], -1, ip + 2, 0, jp + 2, 0),
        diu2(s.diu[1], 0, ip + 2, 0, jp + 2, 0),
        diu3(s.diu[2], 0, ip + 2, 0, jp + 2, 0),
        diu4(s.diu[3], 0, ip + 2, 0, jp + 2, 0),
        diu5(s.diu[4], -1, ip + 2, 0, jp + 2, 0),
        diu6(s.diu[5], 0, ip + 2, 0, jp + 2, 0),
        diu7(s.diu[6], 0, ip + 2, 0, jp + 2, 0),
        diu8(s.diu[7], 0, ip + 2, 0, jp + 2, 0),
        diu9(s.diu[8], 0, ip + 2, 0, jp + 2, 0),
        nou1(s.nou[0], -1, ip + 2, 0, jp + 2, 0),
        nou2(s.nou[1], 0, ip + 2, 0, jp + 2, 0),
        nou3(s.nou[2], 0, ip + 2, 0, jp + 2, 0),
        nou4(s.nou[3], 0, ip + 2, 0, jp + 2, 0),
        nou5(s.nou[4], -1, ip + 2, 0, jp + 2, 0),
        nou6(s.nou[5], 0, ip + 2, 0, jp + 2, 0),
        nou7(s.nou[6], 0, ip + 2, 0, jp + 2, 0),
        nou8(s.nou[7], 0, ip + 2, 0, jp + 2, 0),
        nou9(s.nou[8], 0, ip + 2, 0, jp + 2, 0),
        cov1(s.cov[0], -1, ip + 2, 0, jp + 2, 0),
        cov2(s.cov[1], 0, ip + 2, 0, jp + 2, 0),
        cov3(s.cov[2], 0, ip + 2, 0, jp + 2, 0),
        cov4(s.cov[3], 0, ip + 2, 0, jp + 2, 0),
        cov5(s.cov[4], -1, ip + 2, 0, jp + 2, 0),
        cov6(s.cov[5], 0, ip + 2, 0, jp + 2, 0),
        cov7(s.cov[6], 0, ip + 2, 0, jp + 2, 0),
        cov8(s.cov[7], 0, ip + 2, 0, jp + 2, 0),
        cov9(s.cov[8], 0, ip + 2, 0, jp + 2, 0),
        amask1(s.amask1, 0, ip + 1, 0, jp + 1, 0),
        bmask1(s.bmask1, -1, ip + 1, 0, jp + 1, 0),
        cmask1(s.cmask1, 0, ip + 1, -1, jp + 1, 0),
        dmask1(s.dmask1, 0, ip + 1, 0, jp + 1, 0),
        cn1(s.cn1, 1, ip, 1, jp, 1),
        fx(s.fx, 0, ip, 0, jp, 0),
        fy(s.fy, 0, ip, 0, jp, 0),
        fz(s.fz, 0, ip, 0, jp, 0),
        dfu1(s.dfu1, 0, ip, 1, jp, 1),
        dfv1(s.dfv1, 1, ip, 0, jp, 1),
        dfw1(s.dfw1, 1, ip, 1, jp, 1),
        dx1(s.dx1, -1), dxs(s.dxs, 0), dy1(s.dy1, 0), dys(s.dys, 0), dzn(s.dzn, -1),
        dzs(s.dzs, -1), z2(s.z2, 1), cn2l(s.cn2l, 1), cn2s(s.cn2s, 1), cn3l(s.cn3l, 1),
        cn3s(s.cn3s, 1), cn4l(s.cn4l, 1), cn4s(s.cn4s, 1), delx1(s.delx1, 1),
        wind(s.wind, jp) {}
};

// velnw.f95:27-59
template <typename T>
void velnw(F<T>& a, T ro, T dt) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T pz = (-a.p(i, j, k) + a.p(i + 1, j, k)) / ro / a.dxs(i);
        a.u(i, j, k) = a.u(i, j, k) + dt * (a.f(i, j, k) - pz);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T pz = (-a.p(i, j, k) + a.p(i, j + 1, k)) / ro / a.dys(j);
        a.v(i, j, k) = a.v(i, j, k) + dt * (a.g(i, j, k) - pz);
      }
  for (int k = 1; k <= km - 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T pz = (-a.p(i, j, k) + a.p(i, j, k + 1)) / ro / a.dzs(k);
        a.w(i, j, k) = a.w(i, j, k) + dt * (a.h(i, j, k) - pz);
      }
}

// bondv1_ext_wind_profile.f95:27-127 (inflow from wind_profile, n==1 fill, outflow, periodic j,
// top/bottom).  Returns uout.
template <typename T>
T bondv1_ext(F<T>& a, int n, T dt) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int i = 0; i <= 1; ++i)
    for (int k = 1; k <= km; ++k)
      for (int j = 1; j <= jm; ++j) {
        a.u(i, j, k) = a.wind(j, k);
        a.v(i, j, k) = T(0);
        a.w(i, j, k) = T(0);
      }
  if (n == 1) {  // ICAL == 0
    for (int k = 1; k <= km; ++k)
      for (int j = 1; j <= jm; ++j)
        for (int i = 2; i <= im; ++i) {
          a.u(i, j, k) = a.u(1, j, k);
          a.v(i, j, k) = a.v(1, j, k);
          a.w(i, j, k) = a.w(1, j, k);
        }
  }
  T aaa = T(0), bbb = T(0);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j) {
      T x = a.u(im, j, k);
      aaa = std::fmax(aaa, x);  // amax1
      bbb = std::fmin(bbb, x);  // amin1
    }
  T uout = (aaa + bbb) / T(2);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      a.u(im, j, k) = a.u(im, j, k) - dt * uout * (a.u(im, j, k) - a.u(im - 1, j, k)) / a.dxs(im);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      a.v(im + 1, j, k) =
          a.v(im + 1, j, k) - dt * uout * (a.v(im + 1, j, k) - a.v(im, j, k)) / a.dxs(im);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      a.w(im + 1, j, k) =
          a.w(im + 1, j, k) - dt * uout * (a.w(im + 1, j, k) - a.w(im, j, k)) / a.dxs(im);
  for (int k = 0; k <= km + 1; ++k)
    for (int i = 0; i <= im + 1; ++i) {
      a.u(i, 0, k) = a.u(i, jm, k);
      a.u(i, jm + 1, k) = a.u(i, 1, k);
    }
  for (int k = 0; k <= km + 1; ++k)
    for (int i = 0; i <= im + 1; ++i) {
      a.v(i, 0, k) = a.v(i, jm, k);
      a.v(i, jm + 1, k) = a.v(i, 1, k);
    }
  for (int k = 0; k <= km; ++k)
    for (int i = 0; i <= im + 1; ++i) {
      a.w(i, 0, k) = a.w(i, jm, k);
      a.w(i, jm + 1, k) = a.w(i, 1, k);
    }
  for (int j = 0; j <= jm + 1; ++j)
    for (int i = 0; i <= im + 1; ++i) {
      a.u(i, j, 0) = -a.u(i, j, 1);
      a.u(i, j, km + 1) = a.u(i, j, km);
    }
  for (int j = 0; j <= jm + 1; ++j)
    for (int i = 0; i <= im + 1; ++i) {
      a.v(i, j, 0) = -a.v(i, j, 1);
      a.v(i, j, km + 1) = a.v(i, j, km);
    }
  for (int j = -1; j <= jm + 1; ++j)
    for (int i = -1; i <= im + 1; ++i) {  // i = -1 lies outside w's bounds: flat wrap as in Fortran
      a.w(i, j, 0) = T(0);
      a.w(i, j, km) = T(0);
    }
  return uout;
}

// vel2.f95:48-188
template <typename T>
void vel2(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T two = T(2);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou1(i, j, k) = (a.u(i - 1, j, k) + a.u(i, j, k)) / two;
        a.diu1(i, j, k) = (-a.u(i - 1, j, k) + a.u(i, j, k)) / a.dx1(i);
        a.nou5(i, j, k) = (a.v(i, j - 1, k) + a.v(i, j, k)) / two;
        a.diu5(i, j, k) = (-a.v(i, j - 1, k) + a.v(i, j, k)) / a.dy1(j);
        a.nou9(i, j, k) = (a.w(i, j, k - 1) + a.w(i, j, k)) / two;
        a.diu9(i, j, k) = (-a.w(i, j, k - 1) + a.w(i, j, k)) / a.dzn(k);
        a.cov1(i, j, k) = a.nou1(i, j, k) * a.diu1(i, j, k);
        a.cov5(i, j, k) = a.nou5(i, j, k) * a.diu5(i, j, k);
        a.cov9(i, j, k) = a.nou9(i, j, k) * a.diu9(i, j, k);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou2(i, j, k) = (a.dx1(i + 1) * a.v(i, j - 1, k) + a.dx1(i) * a.v(i + 1, j - 1, k)) /
                          (a.dx1(i) + a.dx1(i + 1));
        a.diu2(i, j, k) = two * (-a.u(i, j - 1, k) + a.u(i, j, k)) / (a.dy1(j - 1) + a.dy1(j));
        a.cov2(i, j, k) = a.nou2(i, j, k) * a.diu2(i, j, k);
      }
  for (int k = 1; k <= km + 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou3(i, j, k) = (a.dx1(i + 1) * a.w(i, j, k - 1) + a.dx1(i) * a.w(i + 1, j, k - 1)) /
                          (a.dx1(i) + a.dx1(i + 1));
        a.diu3(i, j, k) = (-a.u(i, j, k - 1) + a.u(i, j, k)) / a.dzs(k - 1);
        a.cov3(i, j, k) = a.nou3(i, j, k) * a.diu3(i, j, k);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou4(i, j, k) = (a.dy1(j + 1) * a.u(i - 1, j, k) + a.dy1(j) * a.u(i - 1, j + 1, k)) /
                          (a.dy1(j) + a.dy1(j + 1));
        a.diu4(i, j, k) = two * (-a.v(i - 1, j, k) + a.v(i, j, k)) / (a.dx1(i - 1) + a.dx1(i));
        a.cov4(i, j, k) = (a.nou4(i, j, k) - T(0)) * a.diu4(i, j, k);
      }
  for (int k = 1; k <= km + 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou6(i, j, k) = (a.dy1(j + 1) * a.w(i, j, k - 1) + a.dy1(j) * a.w(i, j + 1, k - 1)) /
                          (a.dy1(j) + a.dy1(j + 1));
        a.diu6(i, j, k) = (-a.v(i, j, k - 1) + a.v(i, j, k)) / a.dzs(k - 1);
        a.cov6(i, j, k) = a.nou6(i, j, k) * a.diu6(i, j, k);
      }
  for (int k = 1; k <= km - 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou7(i, j, k) = (a.dzn(k + 1) * a.u(i - 1, j, k) + a.dzn(k) * a.u(i - 1, j, k + 1)) /
                          (a.dzn(k) + a.dzn(k + 1));
        a.diu7(i, j, k) = two * (-a.w(i - 1, j, k) + a.w(i, j, k)) / (a.dx1(i - 1) + a.dx1(i));
        a.cov7(i, j, k) = (a.nou7(i, j, k) - T(0)) * a.diu7(i, j, k);
      }
  for (int k = 1; k <= km - 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.nou8(i, j, k) = (a.dzn(k + 1) * a.v(i, j - 1, k) + a.dzn(k) * a.v(i, j - 1, k + 1)) /
                          (a.dzn(k) + a.dzn(k + 1));
        a.diu8(i, j, k) = two * (-a.w(i, j - 1, k) + a.w(i, j, k)) / (a.dy1(j - 1) + a.dy1(j));
        a.cov8(i, j, k) = a.nou8(i, j, k) * a.diu8(i, j, k);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j) {
      a.nou1(im + 1, j, k) = a.nou1(im, j, k);
      a.diu1(im + 1, j, k) = a.diu1(im, j, k);
      a.cov1(im + 1, j, k) = a.cov1(im, j, k);
    }
  for (int k = 1; k <= km; ++k)
    for (int i = 1; i <= im; ++i) {
      a.nou2(i, 0, k) = a.nou2(i, jm, k);
      a.diu2(i, 0, k) = a.diu2(i, jm, k);
      a.cov2(i, 0, k) = a.cov2(i, jm, k);
      a.nou2(i, jm + 1, k) = a.nou2(i, 1, k);
      a.diu2(i, jm + 1, k) = a.diu2(i, 1, k);
      a.cov2(i, jm + 1, k) = a.cov2(i, 1, k);
    }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j) {
      a.nou4(im + 1, j, k) = a.nou4(im, j, k);
      a.diu4(im + 1, j, k) = a.diu4(im, j, k);
      a.cov4(im + 1, j, k) = a.cov4(im, j, k);
    }
  for (int k = 1; k <= km; ++k)
    for (int i = 1; i <= im; ++i) {
      a.nou5(i, 0, k) = a.nou5(i, jm, k);
      a.diu5(i, 0, k) = a.diu5(i, jm, k);
      a.cov5(i, 0, k) = a.cov5(i, jm, k);
      a.nou5(i, jm + 1, k) = a.nou5(i, 1, k);
      a.diu5(i, jm + 1, k) = a.diu5(i, 1, k);
      a.cov5(i, jm + 1, k) = a.cov5(i, 1, k);
    }
  for (int k = 1; k <= km - 1; ++k)
    for (int j = 1; j <= jm; ++j) {
      a.nou7(im + 1, j, k) = a.nou7(im, j, k);
      a.diu7(im + 1, j, k) = a.diu7(im, j, k);
      a.cov7(im + 1, j, k) = a.cov7(im, j, k);
    }
  for (int k = 1; k <= km - 1; ++k)
    for (int i = 1; i <= im; ++i) {
      a.nou8(i, 0, k) = a.nou8(i, jm, k);
      a.diu8(i, 0, k) = a.diu8(i, jm, k);
      a.cov8(i, 0, k) = a.cov8(i, jm, k);
      a.nou8(i, jm + 1, k) = a.nou8(i, 1, k);
      a.diu8(i, jm + 1, k) = a.diu8(i, 1, k);
      a.cov8(i, jm + 1, k) = a.cov8(i, 1, k);
    }
  for (int k = 1; k <= km + 1; ++k)
    for (int j = 1; j <= jm + 1; ++j) {
      a.diu2(im + 1, j, k) = a.diu2(im, j, k);
      a.diu3(im + 1, j, k) = a.diu3(im, j, k);
    }
  for (int k = 1; k <= km + 1; ++k)
    for (int i = 1; i <= im + 1; ++i) {
      a.diu4(i, 0, k) = a.diu4(i, jm, k);
      a.diu6(i, 0, k) = a.diu6(i, jm, k);
    }
}

// velFG.f95:56-105
template <typename T>
void velfg(F<T>& a, T vn) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T two = T(2);
  vel2(a);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T covx1 = (a.dx1(i + 1) * a.cov1(i, j, k) + a.dx1(i) * a.cov1(i + 1, j, k)) /
                  (a.dx1(i) + a.dx1(i + 1));
        T covy1 = (a.cov2(i, j, k) + a.cov2(i, j + 1, k)) / two;
        T covz1 = (a.cov3(i, j, k) + a.cov3(i, j, k + 1)) / two;
        T covc = covx1 + covy1 + covz1;
        a.dfu1(i, j, k) = two * (-a.diu1(i, j, k) + a.diu1(i + 1, j, k)) / (a.dx1(i) + a.dx1(i + 1)) +
                          (-a.diu2(i, j, k) + a.diu2(i, j + 1, k)) / a.dy1(j) +
                          (-a.diu3(i, j, k) + a.diu3(i, j, k + 1)) / a.dzn(k);
        T df = vn * a.dfu1(i, j, k);
        a.f(i, j, k) = (-covc + df);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T covx1 = (a.cov4(i, j, k) + a.cov4(i + 1, j, k)) / two;
        T covy1 = (a.dy1(j + 1) * a.cov5(i, j, k) + a.dy1(j) * a.cov5(i, j + 1, k)) /
                  (a.dy1(j) + a.dy1(j + 1));
        T covz1 = (a.cov6(i, j, k) + a.cov6(i, j, k + 1)) / two;
        T covc = covx1 + covy1 + covz1;
        a.dfv1(i, j, k) = (-a.diu4(i, j, k) + a.diu4(i + 1, j, k)) / a.dx1(i) +
                          two * (-a.diu5(i, j, k) + a.diu5(i, j + 1, k)) / (a.dy1(j) + a.dy1(j + 1)) +
                          (-a.diu6(i, j, k) + a.diu6(i, j, k + 1)) / a.dzn(k);
        T df = vn * a.dfv1(i, j, k);
        a.g(i, j, k) = (-covc + df);
      }
  for (int k = 1; k <= km - 1; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T covx1 = (a.cov7(i, j, k) + a.cov7(i + 1, j, k)) / two;
        T covy1 = (a.cov8(i, j, k) + a.cov8(i, j + 1, k)) / two;
        T covz1 = (a.dzn(k + 1) * a.cov9(i, j, k) + a.dzn(k) * a.cov9(i, j, k + 1)) /
                  (a.dzn(k) + a.dzn(k + 1));
        T covc = covx1 + covy1 + covz1;
        a.dfw1(i, j, k) = (-a.diu7(i, j, k) + a.diu7(i + 1, j, k)) / a.dx1(i) +
                          (-a.diu8(i, j, k) + a.diu8(i, j + 1, k)) / a.dy1(j) +
                          (-a.diu9(i, j, k) + a.diu9(i, j, k + 1)) / a.dzs(k);
        T df = vn * a.dfw1(i, j, k);
        a.h(i, j, k) = (-covc + df);
      }
}

// feedbf.f95:38-67
template <typename T>
void feedbf(F<T>& a, T alpha, T beta, T dt) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.usum(i, j, k) = (a.usum(i, j, k) + a.u(i, j, k)) * a.bmask1(i, j, k);
        a.vsum(i, j, k) = (a.vsum(i, j, k) + a.v(i, j, k)) * a.cmask1(i, j, k);
        a.wsum(i, j, k) = (a.wsum(i, j, k) + a.w(i, j, k)) * a.dmask1(i, j, k);
        T f1x = alpha * a.usum(i, j, k) * dt;
        T f1y = alpha * a.vsum(i, j, k) * dt;
        T f1z = alpha * a.wsum(i, j, k) * dt;
        T f2x = beta * a.u(i, j, k) * a.bmask1(i, j, k);
        T f2y = beta * a.v(i, j, k) * a.cmask1(i, j, k);
        T f2z = beta * a.w(i, j, k) * a.dmask1(i, j, k);
        a.fx(i, j, k) = f1x + f2x;
        a.fy(i, j, k) = f1y + f2y;
        a.fz(i, j, k) = f1z + f2z;
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.f(i, j, k) = a.f(i, j, k) + a.fx(i, j, k);
        a.g(i, j, k) = a.g(i, j, k) + a.fy(i, j, k);
        a.h(i, j, k) = a.h(i, j, k) + a.fz(i, j, k);
      }
}

// boundsm.f95:13-32
template <typename T>
void boundsm(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int k = 0; k <= km + 1; ++k)
    for (int j = -1; j <= jm + 1; ++j) {
      a.sm(0, j, k) = a.sm(1, j, k);
      a.sm(im + 1, j, k) = a.sm(im, j, k);
    }
  for (int k = 0; k <= km + 1; ++k)
    for (int i = 0; i <= im + 1; ++i) {
      a.sm(i, jm + 1, k) = a.sm(i, jm, k);
      a.sm(i, 0, k) = a.sm(i, 1, k);
    }
  for (int j = -1; j <= jm + 1; ++j)
    for (int i = 0; i <= im + 1; ++i) {
      a.sm(i, j, 0) = -a.sm(i, j, 1);
      a.sm(i, j, km + 1) = a.sm(i, j, km);
    }
}

// les.f95:31-60 (delx1 is supplied by the caller: les.f95:33-37 evaluates it with powf, which the
// caller does with the same libm call)
template <typename T>
void les_calc_sm(F<T>& a, T cs0) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T q = T(0.25), two = T(2);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T dudxx1 = a.diu1(i, j, k);
        T dudyx1 = (a.diu2(i - 1, j, k) + a.diu2(i - 1, j + 1, k) + a.diu2(i, j, k) + a.diu2(i, j + 1, k)) * q;
        T dudzx1 = (a.diu3(i - 1, j, k) + a.diu3(i - 1, j, k + 1) + a.diu3(i, j, k) + a.diu3(i, j, k + 1)) * q;
        T dvdxx1 = (a.diu4(i, j, k) + a.diu4(i, j - 1, k) + a.diu4(i + 1, j, k) + a.diu4(i + 1, j - 1, k)) * q;
        T dvdyx1 = a.diu5(i, j, k);
        T dvdzx1 = (a.diu6(i, j - 1, k) + a.diu6(i, j - 1, k + 1) + a.diu6(i, j, k) + a.diu6(i, j, k + 1)) * q;
        T dwdxx1 = (a.diu7(i, j, k) + a.diu7(i, j, k - 1) + a.diu7(i + 1, j, k) + a.diu7(i + 1, j, k - 1)) * q;
        T dwdyx1 = (a.diu8(i, j, k) + a.diu8(i, j, k - 1) + a.diu8(i, j + 1, k) + a.diu8(i, j + 1, k - 1)) * q;
        T dwdzx1 = a.diu9(i, j, k);
        T csx1 = cs0;
        T cd = csx1 * a.delx1(k);
        T s = two * (dudxx1 * dudxx1 + dvdyx1 * dvdyx1 + dwdzx1 * dwdzx1) +
              (dudyx1 + dvdxx1) * (dudyx1 + dvdxx1) + (dwdyx1 + dvdzx1) * (dwdyx1 + dvdzx1) +
              (dudzx1 + dwdxx1) * (dudzx1 + dwdxx1);
        a.sm(i, j, k) = (cd * cd) * std::sqrt(s);
      }
}

// les.f95:78-182
template <typename T>
void les_calc_visc(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T two = T(2);
  auto& dx1 = a.dx1;
  auto& dy1 = a.dy1;
  auto& dzn = a.dzn;
  auto& sm = a.sm;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T evsx2 = sm(i + 1, j, k);
        T evsx1 = sm(i, j, k);
        T evsy2 = (dy1(j + 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1))) +
                   dy1(j) * ((dx1(i + 1) * sm(i, j + 1, k) + dx1(i) * sm(i + 1, j + 1, k)) / (dx1(i) + dx1(i + 1)))) /
                  (dy1(j) + dy1(j + 1));
        T evsy1 = (dy1(j + 1) * ((dx1(i + 1) * sm(i, j - 1, k) + dx1(i) * sm(i + 1, j - 1, k)) / (dx1(i) + dx1(i + 1))) +
                   dy1(j) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1)))) /
                  (dy1(j) + dy1(j + 1));
        T evsz2 = (dzn(k + 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1))) +
                   dzn(k) * ((dx1(i + 1) * sm(i, j, k + 1) + dx1(i) * sm(i + 1, j, k + 1)) / (dx1(i) + dx1(i + 1)))) /
                  (dzn(k) + dzn(k + 1));
        T evsz1 = (dzn(k) * ((dx1(i + 1) * sm(i, j, k - 1) + dx1(i) * sm(i + 1, j, k - 1)) / (dx1(i) + dx1(i + 1))) +
                   dzn(k - 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1)))) /
                  (dzn(k - 1) + dzn(k));
        T visux2 = (evsx2) * two * a.diu1(i + 1, j, k);
        T visux1 = (evsx1) * two * a.diu1(i, j, k);
        T visuy2 = (evsy2) * (a.diu2(i, j + 1, k) + a.diu4(i + 1, j, k));
        T visuy1 = (evsy1) * (a.diu2(i, j, k) + a.diu4(i + 1, j - 1, k));
        T visuz2 = (evsz2) * (a.diu3(i, j, k + 1) + a.diu7(i + 1, j, k));
        T visuz1 = (evsz1) * (a.diu3(i, j, k) + a.diu7(i + 1, j, k - 1));
        T vfu = (visux2 - visux1) / dx1(i) + (visuy2 - visuy1) / dy1(j) + (visuz2 - visuz1) / dzn(k);
        a.f(i, j, k) = (a.f(i, j, k) + vfu);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T evsy2 = sm(i, j + 1, k);
        T evsy1 = sm(i, j, k);
        T evsx2 = (dy1(j + 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1))) +
                   dy1(j) * ((dx1(i + 1) * sm(i, j + 1, k) + dx1(i) * sm(i + 1, j + 1, k)) / (dx1(i) + dx1(i + 1)))) /
                  (dy1(j) + dy1(j + 1));
        T evsx1 = (dy1(j + 1) * ((dx1(i) * sm(i - 1, j, k) + dx1(i - 1) * sm(i, j, k)) / (dx1(i - 1) + dx1(i))) +
                   dy1(j) * ((dx1(i) * sm(i - 1, j + 1, k) + dx1(i - 1) * sm(i, j + 1, k)) / (dx1(i - 1) + dx1(i)))) /
                  (dy1(j) + dy1(j + 1));
        T evsz2 = (dzn(k + 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1))) +
                   dzn(k) * ((dx1(i + 1) * sm(i, j, k + 1) + dx1(i) * sm(i + 1, j, k + 1)) / (dx1(i) + dx1(i + 1)))) /
                  (dzn(k) + dzn(k + 1));
        T evsz1 = (dzn(k) * ((dx1(i + 1) * sm(i, j, k - 1) + dx1(i) * sm(i + 1, j, k - 1)) / (dx1(i) + dx1(i + 1))) +
                   dzn(k - 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1)))) /
                  (dzn(k - 1) + dzn(k));
        T visvx2 = (evsx2) * (a.diu2(i, j + 1, k) + a.diu4(i + 1, j, k));
        T visvx1 = (evsx1) * (a.diu2(i - 1, j + 1, k) + a.diu4(i, j, k));
        T visvy2 = (evsy2) * two * a.diu5(i, j + 1, k);
        T visvy1 = (evsy1) * two * a.diu5(i, j, k);
        T visvz2 = (evsz2) * (a.diu6(i, j, k + 1) + a.diu8(i, j + 1, k));
        T visvz1 = (evsz1) * (a.diu6(i, j, k) + a.diu8(i, j + 1, k - 1));
        T vfv = (visvx2 - visvx1) / dx1(i) + (visvy2 - visvy1) / dy1(j) + (visvz2 - visvz1) / dzn(k);
        a.g(i, j, k) = (a.g(i, j, k) + vfv);
      }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T evsz2 = sm(i, j, k + 1);
        T evsz1 = sm(i, j, k);
        T evsx2 = (dzn(k + 1) * ((dx1(i + 1) * sm(i, j, k) + dx1(i) * sm(i + 1, j, k)) / (dx1(i) + dx1(i + 1))) +
                   dzn(k) * ((dx1(i + 1) * sm(i, j, k + 1) + dx1(i) * sm(i + 1, j, k + 1)) / (dx1(i) + dx1(i + 1)))) /
                  (dzn(k) + dzn(k + 1));
        T evsx1 = (dzn(k + 1) * ((dx1(i) * sm(i - 1, j, k) + dx1(i - 1) * sm(i, j, k)) / (dx1(i - 1) + dx1(i))) +
                   dzn(k) * ((dx1(i) * sm(i - 1, j, k + 1) + dx1(i - 1) * sm(i, j, k + 1)) / (dx1(i - 1) + dx1(i)))) /
                  (dzn(k) + dzn(k + 1));
        T evsy2 = (dzn(k + 1) * ((dy1(j + 1) * sm(i, j, k) + dy1(j) * sm(i, j + 1, k)) / (dy1(j) + dy1(j + 1))) +
                   dzn(k) * ((dy1(j + 1) * sm(i, j, k + 1) + dy1(j) * sm(i, j + 1, k + 1)) / (dy1(j) + dy1(j + 1)))) /
                  (dzn(k) + dzn(k + 1));
        T evsy1 = (dzn(k + 1) * ((dy1(j) * sm(i, j - 1, k) + dy1(j - 1) * sm(i, j, k)) / (dy1(j - 1) + dy1(j))) +
                   dzn(k) * ((dy1(j) * sm(i, j - 1, k + 1) + dy1(j - 1) * sm(i, j, k + 1)) / (dy1(j - 1) + dy1(j)))) /
                  (dzn(k) + dzn(k + 1));
        T viswx2 = (evsx2) * (a.diu3(i, j, k + 1) + a.diu7(i + 1, j, k));
        T viswx1 = (evsx1) * (a.diu3(i - 1, j, k + 1) + a.diu7(i, j, k));
        T viswy2 = (evsy2) * (a.diu6(i, j, k + 1) + a.diu8(i, j + 1, k));
        T viswy1 = (evsy1) * (a.diu6(i, j - 1, k + 1) + a.diu8(i, j, k));
        T viswz2 = (evsz2) * two * a.diu9(i, j, k + 1);
        T viswz1 = (evsz1) * two * a.diu9(i, j, k);
        T vfw = (viswx2 - viswx1) / dx1(i) + (viswy2 - viswy1) / dy1(j) + (viswz2 - viswz1) / dzn(k);
        a.h(i, j, k) = (a.h(i, j, k) + vfw);
      }
}

// adam.f95:32-46
template <typename T>
void adam(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T c15 = T(1.5), c05 = T(0.5);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        T fd = a.f(i, j, k), gd = a.g(i, j, k), hd = a.h(i, j, k);
        a.f(i, j, k) = c15 * a.f(i, j, k) - c05 * a.fold(i, j, k);
        a.g(i, j, k) = c15 * a.g(i, j, k) - c05 * a.gold(i, j, k);
        a.h(i, j, k) = c15 * a.h(i, j, k) - c05 * a.hold(i, j, k);
        a.fold(i, j, k) = fd;
        a.gold(i, j, k) = gd;
        a.hold(i, j, k) = hd;
      }
}

// bondFG.f95:15-32
template <typename T>
void bondfg(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j) a.f(0, j, k) = a.f(1, j, k);
  for (int k = 1; k <= km; ++k)
    for (int i = 1; i <= im; ++i) a.g(i, 0, k) = a.g(i, jm, k);
  for (int j = 1; j <= jm; ++j)
    for (int i = 1; i <= im; ++i) {
      a.h(i, j, 0) = T(0);
      a.h(i, j, km) = T(0);
    }
}

// boundp.f95:23-44
template <typename T>
void boundp1(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int k = 0; k <= km + 1; ++k)
    for (int j = 0; j <= jm + 1; ++j) {
      a.p(0, j, k) = a.p(1, j, k);
      a.p(im + 1, j, k) = a.p(im, j, k);
    }
  for (int k = 0; k <= km + 1; ++k)
    for (int i = 0; i <= im + 1; ++i) {
      a.p(i, 0, k) = a.p(i, jm, k);
      a.p(i, jm + 1, k) = a.p(i, 1, k);
    }
}

// boundp.f95:5-21
template <typename T>
void boundp2(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  for (int j = 0; j <= jm + 1; ++j)
    for (int i = 0; i <= im + 1; ++i) {
      a.p(i, j, 0) = a.p(i, j, 1);
      a.p(i, j, km + 1) = a.p(i, j, km);
    }
}

// press.f95:49-82: bondfg, rhs, rhsav (returns rhsav; area through *area)
template <typename T>
T press_rhsav(F<T>& a, T dt, int sum_mode, T* area_out) {
  const int im = a.im, jm = a.jm, km = a.km;
  bondfg(a);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.rhs(i, j, k) = (-a.u(i - 1, j, k) + a.u(i, j, k)) / a.dx1(i) +
                         (-a.v(i, j - 1, k) + a.v(i, j, k)) / a.dy1(j) +
                         (-a.w(i, j, k - 1) + a.w(i, j, k)) / a.dzn(k);
        a.rhs(i, j, k) = (a.f(i, j, k) - a.f(i - 1, j, k)) / a.dx1(i) +
                         (a.g(i, j, k) - a.g(i, j - 1, k)) / a.dy1(j) +
                         (a.h(i, j, k) - a.h(i, j, k - 1)) / a.dzn(k) + a.rhs(i, j, k) / dt;
      }
  Summer<T> num(sum_mode);
  T area = T(0);  // the volume sum: a grid constant, always in the reference's sequential order
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        num.add(a.dx1(i) * a.dy1(j) * a.dzn(k) * a.rhs(i, j, k));
        area = area + a.dx1(i) * a.dy1(j) * a.dzn(k);
      }
  T rhsav = num.value() / area;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) a.rhs(i, j, k) = a.rhs(i, j, k) - rhsav;
  if (area_out) *area_out = area;
  return rhsav;
}

// press.f95:91-100: one red-black colour sweep; returns the sweep's residual terms via summer.
// The FP64 instantiation (the high-precision reference, not the bitwise restatement) runs the
// levels of a colour in parallel: a colour's cells read only the other colour, so the updated p
// is the same for any order; the residual is summed per level and the levels added in k order,
// so it does not depend on the thread count either.
template <typename T>
void sor_colour(F<T>& a, int nrd, T omega, Summer<T>& sor) {
  const int im = a.im, jm = a.jm, km = a.km;
  if constexpr (std::is_same<T, double>::value) {
    std::vector<double> part((size_t)km + 1, 0.0);
#pragma omp parallel for schedule(static) if ((long)im * jm * km >= (1L << 18))
    for (int k = 1; k <= km; ++k) {
      double s = 0.0;
      for (int j = 1; j <= jm; ++j)
        for (int i = 1 + (k + j + nrd) % 2; i <= im; i += 2) {
          T reltmp = omega * (a.cn1(i, j, k) *
                                  (a.cn2l(i) * a.p(i + 1, j, k) + a.cn2s(i) * a.p(i - 1, j, k) +
                                   a.cn3l(j) * a.p(i, j + 1, k) + a.cn3s(j) * a.p(i, j - 1, k) +
                                   a.cn4l(k) * a.p(i, j, k + 1) + a.cn4s(k) * a.p(i, j, k - 1) -
                                   a.rhs(i, j, k)) -
                              a.p(i, j, k));
          a.p(i, j, k) = a.p(i, j, k) + reltmp;
          s += reltmp * reltmp;
        }
      part[k] = s;
    }
    for (int k = 1; k <= km; ++k) sor.add(part[k]);
    return;
  }
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1 + (k + j + nrd) % 2; i <= im; i += 2) {
        T reltmp = omega * (a.cn1(i, j, k) *
                                (a.cn2l(i) * a.p(i + 1, j, k) + a.cn2s(i) * a.p(i - 1, j, k) +
                                 a.cn3l(j) * a.p(i, j + 1, k) + a.cn3s(j) * a.p(i, j - 1, k) +
                                 a.cn4l(k) * a.p(i, j, k + 1) + a.cn4s(k) * a.p(i, j, k - 1) -
                                 a.rhs(i, j, k)) -
                            a.p(i, j, k));
        a.p(i, j, k) = a.p(i, j, k) + reltmp;
        sor.add(reltmp * reltmp);
      }
}

// press.f95:87-115; returns the number of iterations run, the last residual via *sor_out
template <typename T>
int press_sor(F<T>& a, T omega, T pjuge, int nmaxp, int sum_mode, T* sor_out) {
  int l;
  T sorv = T(0);
  for (l = 1; l <= nmaxp; ++l) {
    Summer<T> sor(sum_mode);
    for (int nrd = 0; nrd <= 1; ++nrd) {
      sor_colour(a, nrd, omega, sor);
      boundp1(a);
    }
    boundp2(a);
    sorv = sor.value();
    if (sorv < pjuge) break;
  }
  if (sor_out) *sor_out = sorv;
  return l > nmaxp ? nmaxp : l;
}

// press.f95:123-146
template <typename T>
T press_pav_adj(F<T>& a, int sum_mode, T* pco_out) {
  const int im = a.im, jm = a.jm, km = a.km;
  Summer<T> num(sum_mode);
  T pco = T(0);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        num.add(a.p(i, j, k) * a.dx1(i) * a.dy1(j) * a.dzn(k));
        pco = pco + a.dx1(i) * a.dy1(j) * a.dzn(k);
      }
  T pav = num.value() / pco;
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) a.p(i, j, k) = a.p(i, j, k) - pav;
  boundp1(a);
  boundp2(a);
  if (pco_out) *pco_out = pco;
  return pav;
}

// Relative L2 residual |rhs - L p|_2 / |rhs|_2 of the Poisson problem whose fixed point the SOR
// of press.f95:91-100 relaxes: L p = sum_nb c_nb p_nb - p/cn1 with the current boundary images
// of p (mirror in i and k, periodic in j after boundp), accumulated in FP64.
template <typename T>
double rel_residual(F<T>& a) {
  double rr = 0.0, bb = 0.0;
  if constexpr (std::is_same<T, double>::value) {  // per-level sums added in k order
    std::vector<double> prr((size_t)a.km + 1, 0.0), pbb((size_t)a.km + 1, 0.0);
#pragma omp parallel for schedule(static) if ((long)a.im * a.jm * a.km >= (1L << 18))
    for (int k = 1; k <= a.km; ++k) {
      double r2 = 0.0, b2 = 0.0;
      for (int j = 1; j <= a.jm; ++j)
        for (int i = 1; i <= a.im; ++i) {
          const double lp = a.cn2l(i) * a.p(i + 1, j, k) + a.cn2s(i) * a.p(i - 1, j, k) +
                            a.cn3l(j) * a.p(i, j + 1, k) + a.cn3s(j) * a.p(i, j - 1, k) +
                            a.cn4l(k) * a.p(i, j, k + 1) + a.cn4s(k) * a.p(i, j, k - 1) -
                            a.p(i, j, k) / a.cn1(i, j, k);
          const double b = a.rhs(i, j, k), r = b - lp;
          r2 += r * r;
          b2 += b * b;
        }
      prr[k] = r2;
      pbb[k] = b2;
    }
    for (int k = 1; k <= a.km; ++k) {
      rr += prr[k];
      bb += pbb[k];
    }
    return bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
  }
  for (int k = 1; k <= a.km; ++k)
    for (int j = 1; j <= a.jm; ++j)
      for (int i = 1; i <= a.im; ++i) {
        const double lp = (double)a.cn2l(i) * a.p(i + 1, j, k) + (double)a.cn2s(i) * a.p(i - 1, j, k) +
                          (double)a.cn3l(j) * a.p(i, j + 1, k) + (double)a.cn3s(j) * a.p(i, j - 1, k) +
                          (double)a.cn4l(k) * a.p(i, j, k + 1) + (double)a.cn4s(k) * a.p(i, j, k - 1) -
                          (double)a.p(i, j, k) / (double)a.cn1(i, j, k);
        const double b = a.rhs(i, j, k), r = b - lp;
        rr += r * r;
        bb += b * b;
      }
  return bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
}

// Weight of cell (i,j,k) in which the Poisson operator of rel_residual (boundary images folded:
// mirror in i and k, periodic in j) is symmetric: w = (dxs(i-1)+dxs(i)) (dys(j-1)+dys(j))
// (dzs(k-1)+dzs(k)), since w(i) cn2l(i) = w(i+1) cn2s(i+1) = 2/dxs(i) with init.f95:82-107's
// coefficients (alike along j and k).  w is the operator's left null vector (its right null
// vector is the constants): L p = rhs has a solution only if sum w rhs = 0.
template <typename T>
double null_weight(F<T>& a, int i, int j, int k) {
  return ((double)a.dxs(i - 1) + (double)a.dxs(i)) * ((double)a.dys(j - 1) + (double)a.dys(j)) *
         ((double)a.dzs(k - 1) + (double)a.dzs(k));
}

// rhs -= sum(w rhs) / sum(w): the rhs made solvable.  press.f95:65-82 removes the volume mean
// (dx1 dy1 dzn); on a stretched vertical grid that leaves sum w rhs != 0, a part of rhs no p
// can produce, so a solve to a tight residual first removes it.  Returns the shift.
template <typename T>
double project_rhs(F<T>& a) {
  double sw = 0.0, swr = 0.0;
  for (int k = 1; k <= a.km; ++k)
    for (int j = 1; j <= a.jm; ++j)
      for (int i = 1; i <= a.im; ++i) {
        const double w = null_weight(a, i, j, k);
        sw += w;
        swr += w * (double)a.rhs(i, j, k);
      }
  const double c = swr / sw;
  for (int k = 1; k <= a.km; ++k)
    for (int j = 1; j <= a.jm; ++j)
      for (int i = 1; i <= a.im; ++i) a.rhs(i, j, k) = (T)((double)a.rhs(i, j, k) - c);
  return c;
}

template <typename T>
void press(F<T>& a, const orc_params& q, orc_scalars* out) {
  T area, pco, sorv;
  T rhsav = press_rhsav(a, (T)q.dt, q.sum_mode, &area);
  // a solve to a tolerance needs a solvable problem
  if (q.refine_tol > 0.0) project_rhs(a);
  int it = press_sor(a, (T)q.omega, (T)q.pjuge, q.nmaxp, q.sum_mode, &sorv);
  int passes = 0;
  double res = 0.0;
  if (q.refine_tol > 0.0) {  // converged solve: further blocks of nmaxp iterations
    res = rel_residual(a);
    while (res > q.refine_tol && passes < q.refine_max) {
      T s2;
      press_sor(a, (T)(q.refine_omega > 0.0 ? q.refine_omega : q.omega), T(0), q.nmaxp,
                q.sum_mode, &s2);
      ++passes;
      res = rel_residual(a);
    }
  }
  T pav = press_pav_adj(a, q.sum_mode, &pco);
  if (out) {
    out->rhsav = rhsav;
    out->area = area;
    out->sor = sorv;
    out->sor_iters = it;
    out->pav = pav;
    out->pco = pco;
    out->resid_l2 = res;
    out->refine_passes = passes;
  }
}

// les.f95:33-109+ with boundsm in between (les.f95:69)
template <typename T>
void les(F<T>& a, T cs0) {
  les_calc_sm(a, cs0);
  boundsm(a);
  les_calc_visc(a);
}

// main_ocl.f95:264-299 (EXTERNAL_WIND_PROFILE, IFBF=1)
template <typename T>
void step(F<T>& a, int n, const orc_params& q, orc_scalars* out) {
  velnw(a, (T)q.ro, (T)q.dt);
  T uout = bondv1_ext(a, n, (T)q.dt);
  velfg(a, (T)q.vn);
  feedbf(a, (T)q.alpha, (T)q.beta, (T)q.dt);
  les(a, (T)q.cs0);
  adam(a);
  press(a, q, out);
  if (out) out->uout = uout;
}

// ifdata.f95:250-273 with the inflow taken from wind_profile (bondv1_ext) and the masks supplied
template <typename T>
void ifdata(F<T>& a, int n, const orc_params& q, orc_scalars* out) {
  T uout = bondv1_ext(a, n, (T)q.dt);
  boundp1(a);
  boundp2(a);
  velfg(a, (T)q.vn);
  feedbf(a, (T)q.alpha, (T)q.beta, (T)q.dt);
  les(a, (T)q.cs0);
  if (out) out->uout = uout;
}

// init.f95:82-107
template <typename T>
void init_coeffs(F<T>& a) {
  const int im = a.im, jm = a.jm, km = a.km;
  const T two = T(2), one = T(1);
  for (int i = 1; i <= im; ++i) {
    a.cn2s(i) = two / (a.dxs(i - 1) * (a.dxs(i - 1) + a.dxs(i)));
    a.cn2l(i) = two / (a.dxs(i) * (a.dxs(i - 1) + a.dxs(i)));
  }
  for (int j = 1; j <= jm; ++j) {
    a.cn3s(j) = two / (a.dys(j - 1) * (a.dys(j - 1) + a.dys(j)));
    a.cn3l(j) = two / (a.dys(j) * (a.dys(j - 1) + a.dys(j)));
  }
  for (int k = 1; k <= km; ++k) {
    T dz1 = a.dzs(k - 1), dz2 = a.dzs(k);
    a.cn4s(k) = two / (dz1 * (dz1 + dz2));
    a.cn4l(k) = two / (dz2 * (dz1 + dz2));
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.cn1(i, j, k) = two / (a.dxs(i - 1) * a.dxs(i)) + two / (a.dys(j - 1) * a.dys(j)) + two / (dz1 * dz2);
        a.cn1(i, j, k) = one / a.cn1(i, j, k);
      }
  }
}

// feedbfm.f95:21-63 without the file read (zbm supplied, bounds (-1:ip+1,-1:jp+1))
template <typename T>
void feedbfm(F<T>& a, const T* zbm_data) {
  const int im = a.im, jm = a.jm, km = a.km;
  A3<T> zbm((void*)zbm_data, -1, a.ip + 1, -1, a.jp + 1, 0);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        a.amask1(i, j, k) = T(1);
        a.bmask1(i, j, k) = T(0);
        a.cmask1(i, j, k) = T(0);
        a.dmask1(i, j, k) = T(0);
      }
  for (int j = 1; j <= jm; ++j)
    for (int i = 1; i <= im; ++i)
      for (int k = 1; k <= km; ++k)
        if (zbm(i, j, 0) > a.z2(k) + T(0.5) * a.dzn(k)) a.amask1(i, j, k) = T(0);
  for (int k = 1; k <= km; ++k)
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i)
        if (a.amask1(i, j, k) == T(0)) {
          a.bmask1(i, j, k) = T(1);
          a.cmask1(i, j, k) = T(1);
          a.dmask1(i, j, k) = T(1);
        }
}

template <typename T>
void delx1_calc(F<T>& a) {
  // les.f95:33-37: (dx1(0)*dy1(0)*dzn(k))**(1./3.)
  for (int k = 1; k <= a.km; ++k) {
    if (sizeof(T) == 4)
      a.delx1(k) = (T)powf((float)(a.dx1(0) * a.dy1(0) * a.dzn(k)), 1.0f / 3.0f);
    else
      a.delx1(k) = (T)pow((double)(a.dx1(0) * a.dy1(0) * a.dzn(k)), 1.0 / 3.0);
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C entry points: orc_<name>_f32 / orc_<name>_f64
// ---------------------------------------------------------------------------------------------
#define ORC_BOTH(NAME, BODY)                                                            \
  extern "C" int NAME##_f32(const orc_fields* s, const orc_params* q, int n, orc_scalars* o) { \
    typedef float T;                                                                    \
    F<T> a(*s);                                                                         \
    (void)q; (void)n; (void)o;                                                          \
    BODY;                                                                               \
    return 0;                                                                           \
  }                                                                                     \
  extern "C" int NAME##_f64(const orc_fields* s, const orc_params* q, int n, orc_scalars* o) { \
    typedef double T;                                                                   \
    F<T> a(*s);                                                                         \
    (void)q; (void)n; (void)o;                                                          \
    BODY;                                                                               \
    return 0;                                                                           \
  }

ORC_BOTH(orc_velnw, velnw<T>(a, (T)q->ro, (T)q->dt))
ORC_BOTH(orc_bondv1, { T uo = bondv1_ext<T>(a, n, (T)q->dt); if (o) o->uout = uo; })
ORC_BOTH(orc_vel2, vel2<T>(a))
ORC_BOTH(orc_velfg, velfg<T>(a, (T)q->vn))
ORC_BOTH(orc_feedbf, feedbf<T>(a, (T)q->alpha, (T)q->beta, (T)q->dt))
ORC_BOTH(orc_les_calc_sm, les_calc_sm<T>(a, (T)q->cs0))
ORC_BOTH(orc_boundsm, boundsm<T>(a))
ORC_BOTH(orc_les_calc_visc, les_calc_visc<T>(a))
ORC_BOTH(orc_les, les<T>(a, (T)q->cs0))
ORC_BOTH(orc_adam, adam<T>(a))
ORC_BOTH(orc_bondfg, bondfg<T>(a))
ORC_BOTH(orc_boundp1, boundp1<T>(a))
ORC_BOTH(orc_boundp2, boundp2<T>(a))
ORC_BOTH(orc_press_rhsav, {
  T area;
  T r = press_rhsav<T>(a, (T)q->dt, q->sum_mode, &area);
  if (o) { o->rhsav = r; o->area = area; }
})
ORC_BOTH(orc_sor_colour, {
  Summer<T> s(q->sum_mode);
  sor_colour<T>(a, n, (T)q->omega, s);
  if (o) o->sor = s.value();
})
ORC_BOTH(orc_press_sor, {
  T sv;
  int it = press_sor<T>(a, (T)q->omega, (T)q->pjuge, q->nmaxp, q->sum_mode, &sv);
  if (o) { o->sor = sv; o->sor_iters = it; }
})
ORC_BOTH(orc_press_pav_adj, {
  T pco;
  T pv = press_pav_adj<T>(a, q->sum_mode, &pco);
  if (o) { o->pav = pv; o->pco = pco; }
})
ORC_BOTH(orc_press, press<T>(a, *q, o))
ORC_BOTH(orc_step, step<T>(a, n, *q, o))
ORC_BOTH(orc_ifdata, ifdata<T>(a, n, *q, o))
ORC_BOTH(orc_init_coeffs, init_coeffs<T>(a))
ORC_BOTH(orc_delx1, delx1_calc<T>(a))

extern "C" int orc_feedbfm_f32(const orc_fields* s, const float* zbm) {
  F<float> a(*s);
  feedbfm<float>(a, zbm);
  return 0;
}
extern "C" int orc_feedbfm_f64(const orc_fields* s, const double* zbm) {
  F<double> a(*s);
  feedbfm<double>(a, zbm);
  return 0;
}

// Exact sum of n FP32 values rounded once to FP32 (exposed for its own tests).
extern "C" float orc_exact_sum_f32(const float* x, long n) {
  BigAcc b;
  for (long i = 0; i < n; ++i) b.add(x[i]);
  return b.to_float();
}

// ---------------------------------------------------------------------------------------------
// aveflow.f95:42-157: time and spanwise statistics of u, v, w, p, sm after each step.
// The eleven (ip,jp,kp) arrays and seven (ip,kp) arrays are flat, i fastest.  The final uwfx
// loop reads aveu and avew at level k-1 = 0 for k = 1, before the arrays' first element; here the
// ip*jp values there read as 0 (flat indexing, so avew(ip+1,jp,0) is avew(1,1,1)), which is
// what the reference computes when the arrays are preceded by zeros.
// ---------------------------------------------------------------------------------------------
extern "C" {
typedef struct {
  void *aveu, *avev, *avew, *avep, *avel, *aveuu, *avevv, *aveww, *avesm, *avesmsm, *uwfx;
  void *avesu, *avesv, *avesw, *avesuu, *avesvv, *avesww, *uwfxs;
} orc_ave;
}

namespace {
template <typename T>
void aveflow(F<T>& a, const orc_ave& s, int n, int n1, int nmax) {
  const int im = a.im, jm = a.jm, km = a.km;
  const long n3 = (long)im * jm * km, n2 = (long)im * km;
  T* aveu = (T*)s.aveu; T* avev = (T*)s.avev; T* avew = (T*)s.avew; T* avep = (T*)s.avep;
  T* avel = (T*)s.avel; T* aveuu = (T*)s.aveuu; T* avevv = (T*)s.avevv; T* aveww = (T*)s.aveww;
  T* avesm = (T*)s.avesm; T* avesmsm = (T*)s.avesmsm; T* uwfx = (T*)s.uwfx;
  T* avesu = (T*)s.avesu; T* avesv = (T*)s.avesv; T* avesw = (T*)s.avesw;
  T* avesuu = (T*)s.avesuu; T* avesvv = (T*)s.avesvv; T* avesww = (T*)s.avesww;
  T* uwfxs = (T*)s.uwfxs;
  auto o3 = [&](int i, int j, int k) { return (long)(i - 1) + (long)im * ((j - 1) + (long)jm * (k - 1)); };
  auto o2 = [&](int i, int k) { return (long)(i - 1) + (long)im * (k - 1); };
  auto rd = [&](const T* x, long off) { return off < 0 ? T(0) : x[off]; };
  if (n == n1) {
    T* z3[] = {aveu, avev, avew, avep, avel, aveuu, avevv, aveww, avesm, avesmsm, uwfx};
    for (T* x : z3) std::fill(x, x + n3, T(0));
    T* z2[] = {avesu, avesv, avesw, avesuu, avesvv, avesww};
    for (T* x : z2) std::fill(x, x + n2, T(0));
  }
  if (n >= n1) {  // aveflow.f95:70-86
    for (int k = 1; k <= km; ++k)
      for (int j = 1; j <= jm; ++j)
        for (int i = 1; i <= im; ++i) {
          const long q = o3(i, j, k);
          const T uu = a.u(i, j, k), vv = a.v(i, j, k), ww = a.w(i, j, k);
          aveu[q] = aveu[q] + uu;
          avev[q] = avev[q] + vv;
          avew[q] = avew[q] + ww;
          avep[q] = avep[q] + a.p(i, j, k);
          aveuu[q] = aveuu[q] + uu * uu;
          avevv[q] = avevv[q] + vv * vv;
          aveww[q] = aveww[q] + ww * ww;
          const T smv = a.sm(i, j, k);
          avesm[q] = avesm[q] + smv;
          avesmsm[q] = avesmsm[q] + smv * smv;
          uwfx[q] = uwfx[q] + T(0.5) * (a.u(i, j, k - 1) + a.u(i, j, k)) * T(0.5) *
                                  (a.w(i, j, k - 1) + a.w(i + 1, j, k - 1));
        }
  }
  if (n != nmax) return;
  const T cnt = (T)(float)(nmax - n1 + 1);  // float(nmax-n1+1)
  T* d3[] = {aveu, avev, avew, avep, avel, aveuu, avevv, aveww, avesm, avesmsm};
  for (T* x : d3)
    for (long q = 0; q < n3; ++q) x[q] = x[q] / cnt;
  for (int k = 1; k <= km; ++k)  // aveflow.f95:104-110
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        const long q = o3(i, j, k);
        uwfx[q] = uwfx[q] / cnt - T(0.5) * (rd(aveu, o3(i, j, k - 1)) + aveu[q]) * T(0.5) *
                                      (rd(avew, o3(i, j, k - 1)) + rd(avew, o3(i + 1, j, k - 1)));
      }
  for (int k = 1; k <= km; ++k)  // aveflow.f95:111-123
    for (int j = 1; j <= jm; ++j)
      for (int i = 1; i <= im; ++i) {
        const long q = o3(i, j, k), r = o2(i, k);
        avesu[r] = avesu[r] + aveu[q];
        avesv[r] = avesv[r] + avev[q];
        avesw[r] = avesw[r] + avew[q];
        avesuu[r] = avesuu[r] + aveuu[q];
        avesvv[r] = avesvv[r] + avevv[q];
        avesww[r] = avesww[r] + aveww[q];
        uwfxs[r] = uwfxs[r] + uwfx[q];
      }
  const T fjm = (T)(float)jm;
  T* s2[] = {avesu, avesv, avesw, avesuu, avesvv, avesww, uwfxs};
  for (T* x : s2)
    for (long r = 0; r < n2; ++r) x[r] = x[r] / fjm;
  for (long q = 0; q < n3; ++q) {  // aveflow.f95:136-149
    aveuu[q] = std::sqrt(std::fabs(aveuu[q] - aveu[q] * aveu[q]));
    avevv[q] = std::sqrt(std::fabs(avevv[q] - avev[q] * avev[q]));
    aveww[q] = std::sqrt(std::fabs(aveww[q] - avew[q] * avew[q]));
  }
  for (long r = 0; r < n2; ++r) {
    avesuu[r] = std::sqrt(std::fabs(avesuu[r] - avesu[r] * avesu[r]));
    avesvv[r] = std::sqrt(std::fabs(avesvv[r] - avesv[r] * avesv[r]));
    avesww[r] = std::sqrt(std::fabs(avesww[r] - avesw[r] * avesw[r]));
  }
}
}  // namespace

extern "C" int orc_aveflow_f32(const orc_fields* s, const orc_ave* st, int n, int n1, int nmax) {
  F<float> a(*s);
  aveflow<float>(a, *st, n, n1, nmax);
  return 0;
}
extern "C" int orc_aveflow_f64(const orc_fields* s, const orc_ave* st, int n, int n1, int nmax) {
  F<double> a(*s);
  aveflow<double>(a, *st, n, n1, nmax);
  return 0;
}
