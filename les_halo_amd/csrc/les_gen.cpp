// Synthetic inputs of the benchmark configurations (host only, deterministic).
//
// RNG: SplitMix64 (Steele, Lea & Flood 2014) with uniforms u = (x >> 11) * 2^-53 in [0,1) and
// normals from Box-Muller on (1-u1, u2) using the cosine branch only.  All float arithmetic
// that mirrors the reference (grid.f95, feedbfm.f95, initial_wind_profile.f95) is FP32 in the
// reference's order; the log-law inflow is evaluated in double and rounded once.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/les_hip.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

extern "C" int les_gen_grid(int ip, int jp, int kp, float dx, float dy, float dz0, float z_flat,
                            float ratio, float* dx1, float* dxs, float* dy1, float* dys,
                            float* dzn, float* dzs, float* z2) {
  if (ip < 2 || jp < 2 || kp < 2 || !dx1 || !dxs || !dy1 || !dys || !dzn || !dzs || !z2)
    return LES_E_INVALID;
  // dx1(-1:ip+1), dy1(0:jp+1), dzn/dzs(-1:kp+2), z2(1:kp+2), dxs(0:ip), dys(0:jp)
  for (int i = -1; i <= ip + 1; ++i) dx1[i + 1] = dx;
  for (int j = 0; j <= jp + 1; ++j) dy1[j] = dy;
  auto DZN = [&](int k) -> float& { return dzn[k + 1]; };
  auto DZS = [&](int k) -> float& { return dzs[k + 1]; };
  DZN(1) = dz0;
  float top = dz0;
  for (int k = 2; k <= kp + 1; ++k) {
    if (top + dz0 <= z_flat)
      DZN(k) = dz0;
    else
      DZN(k) = DZN(k - 1) * ratio;
    top = top + DZN(k);
  }
  // grid.f95:60-65
  DZN(kp + 1) = DZN(kp);
  DZN(kp + 2) = DZN(kp + 1);
  DZN(0) = DZN(1);
  DZN(-1) = DZN(0);
  z2[0] = DZN(1);
  for (int k = 2; k <= kp + 2; ++k) z2[k - 1] = z2[k - 2] + DZN(k);
  // grid.f95:67-82
  for (int k = 0; k <= kp; ++k) DZS(k) = DZN(k + 1) / 2.0f + DZN(k) / 2.0f;
  for (int i = 0; i <= ip; ++i) dxs[i] = dx1[i + 1] / 2.0f + dx1[i + 2] / 2.0f;
  for (int j = 0; j <= jp; ++j) dys[j] = dy1[j] / 2.0f + dy1[j + 1] / 2.0f;
  DZS(kp + 1) = DZS(kp);
  DZS(kp + 2) = DZS(kp + 1);
  DZS(-1) = DZS(0);
  return LES_OK;
}

extern "C" int les_gen_terrain(int ip, int jp, float dx, float dy, uint64_t seed, float coverage,
                               float wmin, float wmax, float hmax, float* zbm) {
  if (ip < 1 || jp < 1 || !zbm || coverage < 0.0f || coverage > 1.0f || wmax < wmin || dx <= 0.0f ||
      dy <= 0.0f)
    return LES_E_INVALID;
  const long ni = ip + 3, nj = jp + 3;  // zbm(-1:ip+1,-1:jp+1)
  for (long q = 0; q < ni * nj; ++q) zbm[q] = 0.0f;
  auto Z = [&](int i, int j) -> float& { return zbm[(i + 1) + ni * (j + 1)]; };
  SplitMix64 rng(seed);
  const long target = (long)std::ceil((double)coverage * ip * jp);
  long covered = 0;
  for (long it = 0; covered < target && it < 100000000L; ++it) {
    const int wx = std::max(1, (int)std::lround((wmin + (wmax - wmin) * rng.uniform()) / dx));
    const int wy = std::max(1, (int)std::lround((wmin + (wmax - wmin) * rng.uniform()) / dy));
    const int i0 = 1 + (int)(rng.uniform() * ip);
    const int j0 = 1 + (int)(rng.uniform() * jp);
    const float hgt = (float)(hmax * rng.uniform());
    if (!(hgt > 0.0f)) continue;
    for (int j = j0; j <= std::min(j0 + wy - 1, jp); ++j)
      for (int i = i0; i <= std::min(i0 + wx - 1, ip); ++i) {
        if (Z(i, j) == 0.0f) ++covered;
        if (hgt > Z(i, j)) Z(i, j) = hgt;
      }
  }
  return LES_OK;
}

extern "C" int les_gen_masks(int ip, int jp, int kp, const float* zbm, const float* z2,
                             const float* dzn, float* amask1, float* bmask1, float* cmask1,
                             float* dmask1) {
  if (!zbm || !z2 || !dzn || !amask1 || !bmask1 || !cmask1 || !dmask1) return LES_E_INVALID;
  const long na = (long)(ip + 2) * (jp + 2) * (kp + 2);
  const long nb = (long)(ip + 3) * (jp + 2) * (kp + 2);
  const long nc = (long)(ip + 2) * (jp + 3) * (kp + 2);
  for (long q = 0; q < na; ++q) amask1[q] = dmask1[q] = 0.0f;
  for (long q = 0; q < nb; ++q) bmask1[q] = 0.0f;
  for (long q = 0; q < nc; ++q) cmask1[q] = 0.0f;
  auto A = [&](float* a, int i, int j, int k) -> float& {
    return a[i + (long)(ip + 2) * (j + (long)(jp + 2) * k)];
  };
  auto B = [&](int i, int j, int k) -> float& {
    return bmask1[(i + 1) + (long)(ip + 3) * (j + (long)(jp + 2) * k)];
  };
  auto C = [&](int i, int j, int k) -> float& {
    return cmask1[i + (long)(ip + 2) * ((j + 1) + (long)(jp + 3) * k)];
  };
  const long nz = ip + 3;
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i) A(amask1, i, j, k) = 1.0f;
  // feedbfm.f95:42-50: amask1 = 0 where the building is taller than z2(k)+0.5*dzn(k)
  for (int j = 1; j <= jp; ++j)
    for (int i = 1; i <= ip; ++i) {
      const float hb = zbm[(i + 1) + nz * (j + 1)];
      for (int k = 1; k <= kp; ++k)
        if (hb > z2[k - 1] + 0.5f * dzn[k + 1]) A(amask1, i, j, k) = 0.0f;
    }
  // feedbfm.f95:53-63
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i)
        if (A(amask1, i, j, k) == 0.0f) {
          B(i, j, k) = 1.0f;
          C(i, j, k) = 1.0f;
          A(dmask1, i, j, k) = 1.0f;
        }
  return LES_OK;
}

// feedbfm.f95:31-37: the GIS height file (GIS/Tokyo_20mgrid.txt) holds 100 x 100 building
// heights, one list-directed record per line, rows j = 100 down to 1, i = 1..100 within a row,
// stored at zbm(i+25, j+25) of zbm(-1:ipmax+1,-1:jpmax+1).  A record contributes its first value.
extern "C" int les_gis_read_zbm(const char* path, int ipmax, int jpmax, float* zbm) {
  if (!path || !zbm || ipmax < 125 || jpmax < 125) return LES_E_INVALID;
  FILE* f = std::fopen(path, "r");
  if (!f) return LES_E_INVALID;
  const long ni = ipmax + 3;
  char line[256];
  int status = LES_OK;
  for (int j = 100; j >= 1 && status == LES_OK; --j)
    for (int i = 1; i <= 100; ++i) {
      float v;
      if (!std::fgets(line, sizeof line, f) || std::sscanf(line, "%f", &v) != 1) {
        status = LES_E_INVALID;
        break;
      }
      zbm[(i + 25 + 1) + ni * (long)(j + 25 + 1)] = v;
    }
  std::fclose(f);
  return status;
}

extern "C" int les_gen_inflow(int jp, int kp, const float* z2, const float* dzn, float ustar,
                              float z0, float noise, uint64_t seed, float* wind) {
  if (jp < 1 || kp < 1 || !z2 || !dzn || !wind || z0 <= 0.0f) return LES_E_INVALID;
  SplitMix64 rng(seed);
  for (int k = 1; k <= kp; ++k) {
    const float z = z2[k - 1] + 0.5f * dzn[k + 1];  // bondv1.f95:29 inflow height
    const double base = ((double)ustar / 0.4) * std::log((double)z / (double)z0);
    for (int j = 1; j <= jp; ++j) {
      const double xi = rng.normal();
      wind[(j - 1) + (long)jp * (k - 1)] = (float)(base * (1.0 + (double)noise * xi));
    }
  }
  return LES_OK;
}

extern "C" int les_initial_wind_profile(int jp, int kp, const float* z2, const float* dzn,
                                        float* wind) {
  if (jp < 1 || kp < 1 || !z2 || !dzn || !wind) return LES_E_INVALID;
  const int kfix = std::min(78, kp);
  for (int k = 1; k <= kfix; ++k) {
    const float u_val = 5.0f * powf((z2[k - 1] + 0.5f * dzn[k + 1]) / 600.0f, 0.2f);
    for (int j = 1; j <= jp; ++j) wind[(j - 1) + (long)jp * (k - 1)] = u_val;
  }
  for (int k = 79; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      wind[(j - 1) + (long)jp * (k - 1)] = wind[(j - 1) + (long)jp * (77 - 1)];
  return LES_OK;
}
