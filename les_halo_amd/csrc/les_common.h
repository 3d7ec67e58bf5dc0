// Device-side layout helpers and the exact reduction used by every LES_Halo kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace les {

// Interior sizes; every array keeps the reference's Fortran bounds (main_ocl.f95:40-167),
// column-major with i fastest.  Index helpers take Fortran indices.
// Linear indices of the Fortran arrays (column-major, each with its own bounds).  The index
// arithmetic is 32-bit (les_create admits only grids whose largest array, diu1, has fewer than
// 2^30 elements), so a neighbour's index is the centre's plus a constant or a stride and the
// compiler folds it into the load's offset instead of redoing 64-bit products.
struct Dims {
  int ip, jp, kp;
  // p(0:ip+2,0:jp+2,0:kp+1)
  __host__ __device__ long P(int i, int j, int k) const { return i + (ip + 3) * (j + (jp + 3) * k); }
  // u,v(0:ip+1,-1:jp+1,0:kp+1)
  __host__ __device__ long U(int i, int j, int k) const {
    return i + (ip + 2) * ((j + 1) + (jp + 3) * k);
  }
  // w(0:ip+1,-1:jp+1,-1:kp+1)
  __host__ __device__ long W(int i, int j, int k) const {
    return i + (ip + 2) * ((j + 1) + (jp + 3) * (k + 1));
  }
  // f,g,h,usum,vsum,wsum(0:ip,0:jp,0:kp)
  __host__ __device__ long F(int i, int j, int k) const { return i + (ip + 1) * (j + (jp + 1) * k); }
  // fold,gold,hold,cn1(1:ip,1:jp,1:kp)
  __host__ __device__ long O(int i, int j, int k) const {
    return (i - 1) + ip * ((j - 1) + jp * (k - 1));
  }
  // rhs(0:ip+1,0:jp+1,0:kp+1)
  __host__ __device__ long R(int i, int j, int k) const { return i + (ip + 2) * (j + (jp + 2) * k); }
  // sm(-1:ip+1,-1:jp+1,0:kp+1)
  __host__ __device__ long S(int i, int j, int k) const {
    return (i + 1) + (ip + 3) * ((j + 1) + (jp + 3) * k);
  }
  // diu1,diu5(-1:ip+2,0:jp+2,0:kp+2)
  __host__ __device__ long D15(int i, int j, int k) const {
    return (i + 1) + (ip + 4) * (j + (jp + 3) * k);
  }
  // diu2..4,6..9(0:ip+2,0:jp+2,0:kp+2)
  __host__ __device__ long D(int i, int j, int k) const { return i + (ip + 3) * (j + (jp + 3) * k); }
  // amask1,dmask1 / interior mask byte array (1:ip,1:jp,1:kp)
  __host__ __device__ long M(int i, int j, int k) const { return O(i, j, k); }
  // element counts of the arrays above
  __host__ __device__ int nP() const { return (ip + 3) * (jp + 3) * (kp + 2); }
  __host__ __device__ int nU() const { return (ip + 2) * (jp + 3) * (kp + 2); }
  __host__ __device__ int nW() const { return (ip + 2) * (jp + 3) * (kp + 3); }
  __host__ __device__ int nF() const { return (ip + 1) * (jp + 1) * (kp + 1); }
  __host__ __device__ int nO() const { return ip * jp * kp; }
  __host__ __device__ int nR() const { return (ip + 2) * (jp + 2) * (kp + 2); }
  __host__ __device__ int nS() const { return (ip + 3) * (jp + 3) * (kp + 2); }
  __host__ __device__ int nD15() const { return (ip + 4) * (jp + 3) * (kp + 3); }
  __host__ __device__ int nD() const { return (ip + 3) * (jp + 3) * (kp + 3); }
};

// Raw-buffer view of an FP32 device array of n elements: loads and stores address it with a
// 32-bit byte offset from a descriptor held in scalar registers, so an access costs one VALU
// add for its row instead of a 64-bit address computation.  The hardware range check makes
// an access outside [0, n) read 0 or be dropped.  Byte counts and offsets are formed in
// unsigned 32-bit arithmetic: les_create admits arrays below 2^30 elements, so n * 4 and every
// in-range offset fit, and an index below 0 wraps to an offset beyond the range.
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const float* p, int n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0,
                                            (int)((unsigned)n * 4u), 0x00020000)) {}
  __device__ __forceinline__ static int off(long i) { return (int)((unsigned)i * 4u); }
  __device__ __forceinline__ float operator[](long i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off(i), 0, 0));
  }
  __device__ __forceinline__ void st(long i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off(i), 0, 0);
  }
};

// 1-D grid arrays, offsets of their lower bounds: dx1(-1:), dxs(0:), dy1(0:), dys(0:),
// dzn(-1:), dzs(-1:), z2(1:), cn2*(1:), cn3*(1:), cn4*(1:), delx1(1:)
struct Grid1D {
  const float *dx1, *dxs, *dy1, *dys, *dzn, *dzs;
  const float *cn2l, *cn2s, *cn3l, *cn3s, *cn4l, *cn4s, *delx1;
  // per-level divisors along k as (b, RN64(1/b)) in FP64, for zdiv: dzn(k), dzs(k) and
  // dzn(k) + dzn(k+1) (the FP32 sum), indexed like dzn (set on the host from the uploaded grid)
  const double2 *rzn, *rzs, *rzp;
  int p2xy;  // every dx1, dy1, dxs, dys and neighbour sum dx1(i)+dx1(i+1), dy1(j)+dy1(j+1) is a
             // power of two (set on the host; selects xdiv<true>)
  __device__ float DX1(int i) const { return dx1[i + 1]; }
  __device__ float DXS(int i) const { return dxs[i]; }
  __device__ float DY1(int j) const { return dy1[j]; }
  __device__ float DYS(int j) const { return dys[j]; }
  __device__ float DZN(int k) const { return dzn[k + 1]; }
  __device__ float DZS(int k) const { return dzs[k + 1]; }
  __device__ double2 RZN(int k) const { return rzn[k + 1]; }
  __device__ double2 RZS(int k) const { return rzs[k + 1]; }
  __device__ double2 RZP(int k) const { return rzp[k + 1]; }  // dzn(k) + dzn(k+1)
};

// a / b with IEEE rounding, faster when b is a power of two.  For a positive normal b = 2^e
// whose reciprocal is normal (biased exponent 1..253), 1/b is exact and a * (1/b) is the same
// real number as a / b, so both round to the same float (subnormal results included): one
// multiply instead of the division sequence (scale, reciprocal, four FMAs, fixup).
__device__ __forceinline__ bool pow2_normal(float b) {
  const uint32_t bb = __float_as_uint(b);
  return (bb & 0x807fffffu) == 0u && (bb & 0x7f800000u) - 0x00800000u <= 0x7e000000u;
}
__device__ __forceinline__ float pow2_recip(float b) {  // requires pow2_normal(b)
  return __uint_as_float(0x7f000000u - __float_as_uint(b));
}
// per-lane denominators (grid spacings along i and j): the launch knows on the host whether
// every one of them is a power of two (Grid1D::p2xy) and picks P2 accordingly
template <bool P2>
__device__ __forceinline__ float xdiv(float a, float b) {
  if constexpr (P2) return a * pow2_recip(b);
  return a / b;
}

// a / b, bit for bit the FP32 division, for a divisor that comes with z = (b, y = RN64(1/b))
// in FP64 (the vertical spacings, one per level).  For finite nonzero a and b:
//  * q0 = RN64(a * y) is within 2^-52 (relative) of a/b.  A quotient of two floats is either
//    a midpoint between floats or at least 2^-49 (relative) away from every midpoint (write
//    a = A 2^s, b = B 2^t, m = M 2^u with 24-bit A, B and 25-bit odd M: a - m b is a nonzero
//    multiple of 2^min(s, u+t)), so RN32(q0) = RN32(a/b) unless a/b is a midpoint, which it
//    can only be when it is subnormal.
//  * q1 = RN64(q0 + (a - b q0) y) (Markstein's correction; the FMA remainder is exact): the
//    sum before its rounding is within about 2^-104 (relative) of a/b, so where a/b is a
//    midpoint (a double) q1 is that midpoint exactly, and elsewhere q1 is as close to a/b as
//    q0 is; the final rounding then ties to even like the division's.
// v_div_fixup_f32, the last instruction of the division sequence, then gives the division's
// results for zero, infinite and NaN operands (and the sign); for other operands it passes
// its first operand with the quotient's sign, which is RN32(a/b) here as in the division.
// Six VALU instructions (two conversions, three FP64 FMA-class, the fixup) against eleven,
// one of them the quarter-rate reciprocal.
__device__ __forceinline__ float zdiv(float a, float b, double2 z) {
  const double x = (double)a;
  const double q0 = x * z.y;
  const double r = __builtin_fma(-z.x, q0, x);
  const double q1 = __builtin_fma(r, z.y, q0);
  return __builtin_amdgcn_div_fixupf((float)q1, b, a);
}

// ------------------------------------------------------------------------------------------
// Exact order-independent summation of FP32 terms.
//
// A finite float is m * 2^(s-149) with a 24-bit integer m and s = max(E,1)-1 in [0,253].  The
// sum is kept as a fixed-point integer with LSB 2^-149 in NLIMB signed 64-bit limbs holding
// 32-bit digits in carry-save form (value = sum_q limb[q] * 2^(32q)).  Each term adds its low
// and high 32-bit parts to two adjacent limbs, so no carry ever has to be propagated until
// the final normalisation; with |digit| < 2^32 a limb can absorb 2^31 terms.  Any summation
// order gives the same integer, hence the same correctly rounded result.
// ------------------------------------------------------------------------------------------
constexpr int NLIMB = 10;

struct TermParts {
  int limb;      // target limb of the low part (high part goes to limb + 1)
  int64_t lo, hi;
};

__device__ __forceinline__ bool split_term(float x, TermParts& t) {
  uint32_t b = __float_as_uint(x);
  uint32_t e = (b >> 23) & 0xffu;
  uint32_t m = b & 0x7fffffu;
  if (e == 0xffu) return false;  // inf/nan: not representable, dropped
  if (e) m |= 0x800000u;
  if (!m) return false;
  int s = (e ? (int)e : 1) - 1;
  int off = s & 31;
  uint64_t v = (uint64_t)m << off;  // < 2^56
  int64_t lo = (int64_t)(v & 0xffffffffull), hi = (int64_t)(v >> 32);
  if (b >> 31) {
    lo = -lo;
    hi = -hi;
  }
  t.limb = s >> 5;
  t.lo = lo;
  t.hi = hi;
  return true;
}

// Per-thread accumulator: a 4-limb register window [base, base+4) plus a fall-back into the
// block's LDS limbs for terms outside the window.
struct ThreadAcc {
  int64_t r0, r1, r2, r3;
  int base;  // -1 until the first term
  __device__ void init() {
    r0 = r1 = r2 = r3 = 0;
    base = -1;
  }
  __device__ __forceinline__ void add(float x, unsigned long long* lds) {
    TermParts t;
    if (!split_term(x, t)) return;
    if (base < 0) base = min(max(t.limb - 1, 0), NLIMB - 4);
    int d = t.limb - base;
    if (d >= 0 && d <= 2) {
      r0 += (d == 0) ? t.lo : 0;
      r1 += (d == 0) ? t.hi : ((d == 1) ? t.lo : 0);
      r2 += (d == 1) ? t.hi : ((d == 2) ? t.lo : 0);
      r3 += (d == 2) ? t.hi : 0;
    } else {
      atomicAdd(&lds[t.limb], (unsigned long long)t.lo);
      atomicAdd(&lds[t.limb + 1], (unsigned long long)t.hi);
    }
  }
  // Adds the window into the block's LDS limbs.  Called by every lane of the wavefront
  // together (the lanes exchange their windows): the lanes whose windows start at the same
  // limb are summed with cross-lane adds and one lane adds the result, so the block's few
  // LDS words see one atomic per wavefront and window position instead of one per lane.
  __device__ void flush(unsigned long long* lds) {
    const int lane = (int)__lane_id();
    unsigned long long pend = __ballot(base >= 0);
    while (pend) {
      const int leader = __ffsll((long long)pend) - 1;
      const int b = __shfl(base, leader);
      const bool mine = (base == b);
      int64_t s0 = mine ? r0 : 0, s1 = mine ? r1 : 0, s2 = mine ? r2 : 0, s3 = mine ? r3 : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
        s3 += __shfl_xor(s3, o);
      }
      if (lane == leader) {
        if (s0) atomicAdd(&lds[b + 0], (unsigned long long)s0);
        if (s1) atomicAdd(&lds[b + 1], (unsigned long long)s1);
        if (s2) atomicAdd(&lds[b + 2], (unsigned long long)s2);
        if (s3) atomicAdd(&lds[b + 3], (unsigned long long)s3);
      }
      pend &= ~__ballot(mine);
    }
  }
};

// Normalises NLIMB carry-save limbs and rounds the exact value (times 2^-149) to the
// nearest float, ties to even.  An exactly zero sum gives +0.
__device__ inline float limbs_to_float(const int64_t* in) {
  // carry-propagate into 32-bit digits; the final carry is the sign extension
  uint32_t dig[NLIMB + 2];
  int64_t carry = 0;
  for (int q = 0; q < NLIMB; ++q) {
    int64_t v = in[q] + carry;
    dig[q] = (uint32_t)(v & 0xffffffffll);
    carry = v >> 32;  // arithmetic shift
  }
  dig[NLIMB] = (uint32_t)(carry & 0xffffffffll);
  dig[NLIMB + 1] = (uint32_t)((carry >> 32) & 0xffffffffll);
  const int ND = NLIMB + 2;
  bool neg = (dig[ND - 1] >> 31) != 0;
  if (neg) {  // magnitude = two's complement negation
    uint64_t c = 1;
    for (int q = 0; q < ND; ++q) {
      uint64_t s = (uint64_t)(uint32_t)(~dig[q]) + c;
      dig[q] = (uint32_t)s;
      c = s >> 32;
    }
  }
  int top = -1;
  for (int q = ND - 1; q >= 0; --q) {
    if (dig[q]) {
      top = q * 32 + 31 - __clz(dig[q]);
      break;
    }
  }
  if (top < 0) return 0.0f;
  uint32_t out;
  if (top < 24) {  // exact: integer below 2^24 times 2^-149 (subnormal, or exponent field 1)
    out = dig[0];
  } else {
    int drop = top - 23;
    // 24 kept bits starting at bit 'drop'
    uint64_t window = 0;
    int q0 = drop >> 5, o0 = drop & 31;
    uint64_t lo = dig[q0], mid = (q0 + 1 < ND) ? dig[q0 + 1] : 0u;
    window = ((mid << 32) | lo) >> o0;
    uint32_t keep = (uint32_t)(window & 0xffffffu);
    // round bit and sticky bits below it
    int rpos = drop - 1;
    uint32_t rnd = (dig[rpos >> 5] >> (rpos & 31)) & 1u;
    bool sticky = false;
    for (int q = 0; q < (rpos >> 5); ++q) sticky |= dig[q] != 0;
    sticky |= (dig[rpos >> 5] & ((1u << (rpos & 31)) - 1u)) != 0;
    if (rnd && (sticky || (keep & 1u))) keep += 1u;
    uint32_t ef = (uint32_t)drop + 1u;
    if (keep == (1u << 24)) {
      keep >>= 1;
      ef += 1u;
    }
    out = (ef >= 255u) ? 0x7f800000u : ((ef << 23) | (keep & 0x7fffffu));
  }
  if (neg) out |= 0x80000000u;
  return __uint_as_float(out);
}

// Same exact value rounded to double (used by the mixed-precision mode).
__device__ inline double limbs_to_double(const int64_t* in) {
  uint32_t dig[NLIMB + 2];
  int64_t carry = 0;
  for (int q = 0; q < NLIMB; ++q) {
    int64_t v = in[q] + carry;
    dig[q] = (uint32_t)(v & 0xffffffffll);
    carry = v >> 32;
  }
  dig[NLIMB] = (uint32_t)(carry & 0xffffffffll);
  dig[NLIMB + 1] = (uint32_t)((carry >> 32) & 0xffffffffll);
  const int ND = NLIMB + 2;
  bool neg = (dig[ND - 1] >> 31) != 0;
  if (neg) {
    uint64_t c = 1;
    for (int q = 0; q < ND; ++q) {
      uint64_t s = (uint64_t)(uint32_t)(~dig[q]) + c;
      dig[q] = (uint32_t)s;
      c = s >> 32;
    }
  }
  int top = -1;
  for (int q = ND - 1; q >= 0; --q) {
    if (dig[q]) {
      top = q * 32 + 31 - __clz(dig[q]);
      break;
    }
  }
  if (top < 0) return 0.0;
  // hi63 = bits [scale, top] aligned so that its leading bit is bit 62; sticky = bits below
  int scale = top - 62;
  uint64_t hi63 = 0;
  bool sticky = false;
  int lo = scale < 0 ? 0 : scale;
  for (int bpos = top; bpos >= lo; --bpos)
    hi63 = (hi63 << 1) | ((dig[bpos >> 5] >> (bpos & 31)) & 1u);
  if (scale < 0) {
    hi63 <<= -scale;
  } else {
    for (int q = 0; q < (scale >> 5); ++q) sticky |= dig[q] != 0;
    sticky |= (dig[scale >> 5] & ((1u << (scale & 31)) - 1u)) != 0;
  }
  uint64_t mant = hi63 >> 10;  // 53 significant bits
  uint64_t rem = hi63 & 0x3ffull;
  bool up = (rem > 0x200ull) || (rem == 0x200ull && (sticky || (mant & 1ull)));
  if (up) mant += 1;
  double r = ldexp((double)mant, scale + 10 - 149);
  return neg ? -r : r;
}


// ------------------------------------------------------------------------------------------
// Block-level pieces of the exact reductions: every block accumulates its limbs in LDS and
// writes them to partials[block]; a one-workgroup finaliser adds all partials (integers,
// exact) and rounds once.
// ------------------------------------------------------------------------------------------
constexpr int RED_KC = 8;  // k-chunks of the reduction grids

__device__ inline void block_limbs_init(unsigned long long* lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid < NLIMB) lds[tid] = 0ull;
  __syncthreads();
}

__device__ inline void block_limbs_store(unsigned long long* lds, int64_t* out) {
  __syncthreads();
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid < NLIMB) out[tid] = (int64_t)lds[tid];
}

// Sum partials[0..nb) into limbs_out (one workgroup, any size >= NLIMB).  Thread t adds limb
// t % NLIMB of blocks t / NLIMB, + G, + 2G, ... (G groups of NLIMB threads: consecutive
// threads read consecutive words), eight loads in flight at a time.
__device__ inline void sum_partials(const int64_t* partials, int nb, int64_t* limbs_out) {
  __shared__ unsigned long long acc[NLIMB];
  if (threadIdx.x < NLIMB) acc[threadIdx.x] = 0ull;
  __syncthreads();
  const int G = (int)blockDim.x / NLIMB;
  const int q = (int)threadIdx.x % NLIMB, b0 = (int)threadIdx.x / NLIMB;
  if (b0 < G) {
    int64_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int b = b0;
    for (; b + 7 * G < nb; b += 8 * G) {
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] += partials[(long)(b + u * G) * NLIMB + q];
    }
    for (; b < nb; b += G) r[0] += partials[(long)b * NLIMB + q];
    const int64_t s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    if (s) atomicAdd(&acc[q], (unsigned long long)s);
  }
  __syncthreads();
  if (threadIdx.x < NLIMB) limbs_out[threadIdx.x] = (int64_t)acc[threadIdx.x];
  __syncthreads();
}
constexpr int FIN_THREADS = 1024;  // workgroup size of the one-workgroup finalisers

// The same exact sums through a few accumulation slots: block L adds its limbs into slot
// L % RED_SLOTS with 64-bit integer atomics (integer addition: the order does not matter), and
// the finaliser reads the RED_SLOTS x NLIMB words instead of one partial per block, clearing
// each slot word as it reads it (the slots start at zero and are zero after every finalise).
// A slot limb adds up at most (blocks / RED_SLOTS) block limbs, far inside 64 bits.
constexpr int RED_SLOTS = 64;

__device__ inline void block_limbs_atomic(unsigned long long* lds, int64_t* slots, int L) {
  __syncthreads();
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid < NLIMB && lds[tid])
    atomicAdd(reinterpret_cast<unsigned long long*>(slots) + (L % RED_SLOTS) * NLIMB + tid,
              lds[tid]);
}

// Sum the slots into limbs_out and clear them (one workgroup of at least RED_SLOTS * NLIMB
// threads).
__device__ inline void sum_slots(int64_t* slots, int64_t* limbs_out) {
  __shared__ unsigned long long acc[NLIMB];
  if (threadIdx.x < NLIMB) acc[threadIdx.x] = 0ull;
  __syncthreads();
  if (threadIdx.x < RED_SLOTS * NLIMB) {
    const int64_t v = slots[threadIdx.x];
    slots[threadIdx.x] = 0;
    if (v) atomicAdd(&acc[threadIdx.x % NLIMB], (unsigned long long)v);
  }
  __syncthreads();
  if (threadIdx.x < NLIMB) limbs_out[threadIdx.x] = (int64_t)acc[threadIdx.x];
  __syncthreads();
}

// Workgroups are handed to the 8 XCDs (each with its own L2) round-robin.  With a grid padded
// to a multiple of 8, this maps dispatch id w to a logical tile so that every XCD works on one
// contiguous range of tiles: spatial neighbours then share an L2.
__device__ __forceinline__ int xcd_tile(int w, int nwg_padded) {
  const int per = nwg_padded >> 3;
  return (w & 7) * per + (w >> 3);
}

// A 3-D grid of nbx x nby x nbz blocks (i, j, k) launched as a 1-D grid padded to a multiple
// of 8, visited in a cache-friendly logical order: the j blocks are cut into strips of sj,
// the order runs strip by strip, within a strip level by level in k, within a level j then i,
// and every XCD takes one contiguous share of the order (xcd_tile).  A stencil's i, j and k
// neighbour cells are then read by blocks resident on the same XCD (same L2) at nearly the
// same time, and adjacent i blocks share the cache lines at their seams.
struct StripGrid {
  int nbx, nby, nbz, sj;
};
constexpr int STRIP_SJ = 8;
inline StripGrid strip_grid(int nbx, int nby, int nbz) { return StripGrid{nbx, nby, nbz, STRIP_SJ}; }
inline dim3 strip_dim(const StripGrid& s) {
  const long n = (long)s.nbx * s.nby * s.nbz;
  return dim3((unsigned)((n + 7) / 8 * 8));
}
// Block (bx, by, bz) of logical index L in the strip order; false past the grid's blocks.
__device__ __forceinline__ bool strip_map(const StripGrid& s, int L, int& bx, int& by, int& bz) {
  if (L >= s.nbx * s.nby * s.nbz) return false;
  bx = L % s.nbx;
  const int R = L / s.nbx;  // row tile, ordered (strip, bz, j within the strip)
  const int per = s.sj * s.nbz;
  const int nstrip = (s.nby + s.sj - 1) / s.sj;
  const int st = min(R / per, nstrip - 1);
  const int w = min(s.sj, s.nby - st * s.sj);
  const int r = R - st * per;
  bz = r / w;
  by = st * s.sj + (r - bz * w);
  return true;
}

// Logical block (bx, by, bz) of this workgroup and its logical index L; false for padding.
__device__ __forceinline__ bool strip_block(const StripGrid& s, int& bx, int& by, int& bz,
                                            int& L) {
  L = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (L >= s.nbx * s.nby * s.nbz) return false;
  bx = L % s.nbx;
  const int R = L / s.nbx;  // row tile, ordered (strip, bz, j within the strip)
  const int per = s.sj * s.nbz;
  const int nstrip = (s.nby + s.sj - 1) / s.sj;
  const int st = min(R / per, nstrip - 1);
  const int w = min(s.sj, s.nby - st * s.sj);
  const int r = R - st * per;
  bz = r / w;
  by = st * s.sj + (r - bz * w);
  return true;
}

}  // namespace les
