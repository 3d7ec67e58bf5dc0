// Restart and output files in the reference's Fortran unformatted sequential format.
//
// data20 (press.f95:192-199): u, v, w, p, usum, vsum, wsum over the interior i = 1..im,
// j = 1..jm, k = 1..km (i fastest), one record each; data21 (adam.f95:24-29): fold, gold, hold,
// fghold.  The ICAL = 1 start (ifdata.f95:227-247) reads data30 = a leading record (n, time)
// followed by the seven data20 records, and data31 = the four data21 records.  A record is a
// 4-byte little-endian byte count, the payload, and the byte count again (the sequential
// unformatted layout of gfortran and flang for records below 2 GiB).
// fghold is never assigned by the reference time loop (adam.f95 only writes it out), so it is
// written as zeros and skipped on reading.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/les_hip.h"

namespace {

const int kData20[7] = {LES_U, LES_V, LES_W, LES_P, LES_USUM, LES_VSUM, LES_WSUM};
const int kData21[3] = {LES_FOLD, LES_GOLD, LES_HOLD};

bool put_record(FILE* f, const void* data, uint32_t nbytes) {
  return std::fwrite(&nbytes, 4, 1, f) == 1 &&
         (nbytes == 0 || std::fwrite(data, 1, nbytes, f) == nbytes) &&
         std::fwrite(&nbytes, 4, 1, f) == 1;
}

bool get_record(FILE* f, void* data, uint32_t nbytes) {
  uint32_t a = 0, b = 0;
  if (std::fread(&a, 4, 1, f) != 1 || a != nbytes) return false;
  if (nbytes && std::fread(data, 1, nbytes, f) != nbytes) return false;
  return std::fread(&b, 4, 1, f) == 1 && b == nbytes;
}

// interior (1:ip,1:jp,1:kp) of a full Fortran-bounded array <-> packed i-fastest block
void interior_copy(const int lo[3], const int hi[3], int ip, int jp, int kp, float* full,
                   float* packed, bool to_packed) {
  const long ni = hi[0] - lo[0] + 1, nj = hi[1] - lo[1] + 1;
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j) {
      float* row = full + (1 - lo[0]) + ni * ((long)(j - lo[1]) + nj * (long)(k - lo[2]));
      float* pk = packed + (long)ip * ((long)(j - 1) + (long)jp * (k - 1));
      if (to_packed)
        std::memcpy(pk, row, sizeof(float) * ip);
      else
        std::memcpy(row, pk, sizeof(float) * ip);
    }
}

struct Dims3 {
  int ip, jp, kp;
};

int dims_of(les_handle* h, Dims3& d) { return les_get_dims(h, &d.ip, &d.jp, &d.kp); }

}  // namespace

extern "C" int les_write_restart(les_handle* h, const char* data20, const char* data21, int n,
                                 float time, int with_header) {
  if (!h || !data20 || !data21) return LES_E_INVALID;
  Dims3 d;
  int r = dims_of(h, d);
  if (r) return r;
  const size_t nin = (size_t)d.ip * d.jp * d.kp;
  if (nin * sizeof(float) >= (size_t)1 << 31) return LES_E_INVALID;  // one-record limit
  const uint32_t nb = (uint32_t)(nin * sizeof(float));
  std::vector<float> full, packed(nin);
  FILE* f = std::fopen(data20, "wb");
  if (!f) return LES_E_INVALID;
  bool ok = true;
  if (with_header) {  // the data30 form: (n, time) first
    unsigned char hdr[8];
    std::memcpy(hdr, &n, 4);
    std::memcpy(hdr + 4, &time, 4);
    ok = put_record(f, hdr, 8);
  }
  for (int q = 0; q < 7 && ok; ++q) {
    int lo[3], hi[3];
    const int64_t cnt = les_field_bounds(kData20[q], d.ip, d.jp, d.kp, lo, hi);
    full.resize((size_t)cnt);
    if ((r = les_download(h, kData20[q], full.data()))) {
      std::fclose(f);
      return r;
    }
    interior_copy(lo, hi, d.ip, d.jp, d.kp, full.data(), packed.data(), true);
    ok = put_record(f, packed.data(), nb);
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return LES_E_INVALID;
  f = std::fopen(data21, "wb");
  if (!f) return LES_E_INVALID;
  for (int q = 0; q < 3 && ok; ++q) {
    if ((r = les_download(h, kData21[q], packed.data()))) {  // fold..hold are (1:ip,1:jp,1:kp)
      std::fclose(f);
      return r;
    }
    ok = put_record(f, packed.data(), nb);
  }
  if (ok) {
    std::fill(packed.begin(), packed.end(), 0.0f);  // fghold
    ok = put_record(f, packed.data(), nb);
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LES_OK : LES_E_INVALID;
}

extern "C" int les_read_restart(les_handle* h, const char* data30, const char* data31, int* n,
                                float* time) {
  if (!h || !data30 || !data31) return LES_E_INVALID;
  Dims3 d;
  int r = dims_of(h, d);
  if (r) return r;
  const size_t nin = (size_t)d.ip * d.jp * d.kp;
  if (nin * sizeof(float) >= (size_t)1 << 31) return LES_E_INVALID;
  const uint32_t nb = (uint32_t)(nin * sizeof(float));
  std::vector<float> full, packed(nin);
  FILE* f = std::fopen(data30, "rb");
  if (!f) return LES_E_INVALID;
  unsigned char hdr[8];
  bool ok = get_record(f, hdr, 8);
  if (ok) {
    if (n) std::memcpy(n, hdr, 4);
    if (time) std::memcpy(time, hdr + 4, 4);
  }
  for (int q = 0; q < 7 && ok; ++q) {
    int lo[3], hi[3];
    const int64_t cnt = les_field_bounds(kData20[q], d.ip, d.jp, d.kp, lo, hi);
    full.resize((size_t)cnt);
    ok = get_record(f, packed.data(), nb);
    if (!ok) break;
    // only the interior is replaced: the halos keep the handle's values, as in the reference,
    // whose init sequence (ifdata.f95:258-273) recomputes them after the read
    if ((r = les_download(h, kData20[q], full.data()))) {
      std::fclose(f);
      return r;
    }
    interior_copy(lo, hi, d.ip, d.jp, d.kp, full.data(), packed.data(), false);
    if ((r = les_upload(h, kData20[q], full.data()))) {
      std::fclose(f);
      return r;
    }
  }
  std::fclose(f);
  if (!ok) return LES_E_INVALID;
  f = std::fopen(data31, "rb");
  if (!f) return LES_E_INVALID;
  for (int q = 0; q < 3 && ok; ++q) {
    ok = get_record(f, packed.data(), nb);
    if (ok && (r = les_upload(h, kData21[q], packed.data()))) {
      std::fclose(f);
      return r;
    }
  }
  if (ok) ok = get_record(f, packed.data(), nb);  // fghold: read and dropped
  std::fclose(f);
  return ok ? LES_OK : LES_E_INVALID;
}
