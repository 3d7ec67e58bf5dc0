// Snapshot output of the reference time loop, written from the host copies of the fields.
//
// NetCDF (module_LES_write_netcdf.f95): init_netcdf_file creates six files in the netCDF
// classic format -- LES_output.nc and LES_output_p.nc (pressure), LES_output_u/v/w.nc
// (velocity_x/y/z) and LES_output_uvwsum.nc (velocity_sum_x/y/z) -- each with the dimensions
// level, latitude, longitude and a fixed-length time dimension of NTIMESTEPS = 20
// (module_LES_write_netcdf.f95:24-32,124-151), the coordinate variables latitude, longitude
// (float, values 1..n) and time (double, values 1..20) with their units (:160-256), and the
// data variables over (time, level, latitude, longitude) spanning each array's full Fortran
// bounds (:230-250,303-314).  write_to_netcdf_file stores one time slab of every variable
// (:318-345).  The files are written here directly in the published classic format (netCDF
// "CDF-1"; "CDF-2", the 64-bit-offset variant, when a variable starts beyond 2 GiB): a header
// of dimensions, global attributes (none) and variables, then each variable's data, big-endian,
// laid out as the netCDF library lays them out with its default alignment (nc_enddef: every
// variable starts at the next multiple of 4 after the header or the previous variable, fill
// mode on, so unwritten time slabs hold the default float fill value 9.96921e36).
//
// anime (anime.f95:26-49): at n == n0, n == nmax or every 1000 steps, data22 gets the records
// (im, jm, km) and the cell coordinates x = dxl(i-1)+dx1(i), z = z2(k), y = dyl(j-1)+dy1(j)
// (in the order x, z, y, one record), data23 the records (im, jm, km, 4) and the cell-centred
// 0.5*(u(i-1)+u(i)), 0.5*(w(k-1)+w(k)), 0.5*(v(j-1)+v(j)) and amask1 over the interior (one
// record), as Fortran unformatted sequential files.  dxl and dyl are the running sums of dx1
// and dy1 from 0 (grid.f95:26-38), in FP32 in that order.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/les_hip.h"

namespace {

constexpr int NC_DOUBLE = 6, NC_FLOAT = 5, NC_CHAR = 2;
constexpr uint32_t NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C;
constexpr uint32_t FILL_FLOAT_BITS = 0x7cf00000u;  // 9.9692099683868690e+36f (NC_FILL_FLOAT)

// big-endian byte stream
struct Out {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void pad4() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad4();
  }
};

struct NcVar {
  std::string name, units;
  int type;               // NC_FLOAT or NC_DOUBLE
  std::vector<int> dims;  // dimension ids, slowest first
  uint64_t nelems = 0, begin = 0;
};

struct NcFile {
  std::string path;
  std::vector<std::pair<std::string, uint32_t>> dims;
  std::vector<NcVar> vars;
  int version = 1;
  uint64_t size = 0;

  int add_dim(const std::string& n, uint32_t len) {
    dims.emplace_back(n, len);
    return (int)dims.size() - 1;
  }
  int add_var(const std::string& n, int type, std::vector<int> d, const std::string& units) {
    NcVar v;
    v.name = n;
    v.type = type;
    v.dims = std::move(d);
    v.units = units;
    v.nelems = 1;
    for (int id : v.dims) v.nelems *= dims[id].second;
    vars.push_back(v);
    return (int)vars.size() - 1;
  }
  static uint64_t vsize_bytes(const NcVar& v) {
    const uint64_t s = v.nelems * (v.type == NC_DOUBLE ? 8 : 4);
    return (s + 3) / 4 * 4;
  }
  std::vector<unsigned char> header() const {
    Out o;
    o.b.push_back('C');
    o.b.push_back('D');
    o.b.push_back('F');
    o.b.push_back((unsigned char)version);
    o.u32(0);  // numrecs: no record dimension
    o.u32(NC_DIMENSION);
    o.u32((uint32_t)dims.size());
    for (const auto& d : dims) {
      o.name(d.first);
      o.u32(d.second);
    }
    o.u32(0);  // no global attributes (ABSENT)
    o.u32(0);
    o.u32(NC_VARIABLE);
    o.u32((uint32_t)vars.size());
    for (const auto& v : vars) {
      o.name(v.name);
      o.u32((uint32_t)v.dims.size());
      for (int id : v.dims) o.u32((uint32_t)id);
      o.u32(NC_ATTRIBUTE);
      o.u32(1);
      o.name("units");
      o.u32(NC_CHAR);
      o.name(v.units);  // nelems = string length, chars padded to 4
      o.u32((uint32_t)v.type);
      const uint64_t vs = vsize_bytes(v);
      o.u32(vs >= 0xfffffffcull ? 0xffffffffu : (uint32_t)vs);
      if (version == 1)
        o.u32((uint32_t)v.begin);
      else
        o.u64(v.begin);
    }
    return o.b;
  }
  // data offsets as nc_enddef assigns them (header, then the variables in order, 4-aligned)
  void layout() {
    for (int pass = 0; pass < 2; ++pass) {  // the header length depends on the offset width
      uint64_t off = (header().size() + 3) / 4 * 4;
      for (auto& v : vars) {
        v.begin = off;
        off += vsize_bytes(v);
      }
      size = off;
      const bool wide = !vars.empty() && vars.back().begin > 0x7fffffffull;
      version = wide ? 2 : 1;
    }
  }
};

bool write_at(FILE* f, uint64_t off, const void* p, size_t n) {
  return fseeko(f, (off_t)off, SEEK_SET) == 0 && std::fwrite(p, 1, n, f) == n;
}

void to_be32(const float* src, size_t n, std::vector<unsigned char>& dst) {
  dst.resize(n * 4);
  for (size_t q = 0; q < n; ++q) {
    uint32_t v;
    std::memcpy(&v, src + q, 4);
    dst[4 * q] = (unsigned char)(v >> 24);
    dst[4 * q + 1] = (unsigned char)(v >> 16);
    dst[4 * q + 2] = (unsigned char)(v >> 8);
    dst[4 * q + 3] = (unsigned char)v;
  }
}

// Creates the file: header, coordinate values, every data variable filled with the fill value.
bool create_file(NcFile& nf, const std::vector<int>& coord_vars) {
  nf.layout();
  FILE* f = std::fopen(nf.path.c_str(), "wb");
  if (!f) return false;
  const std::vector<unsigned char> h = nf.header();
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
  std::vector<unsigned char> buf;
  for (size_t q = 0; q < nf.vars.size() && ok; ++q) {
    const NcVar& v = nf.vars[q];
    const bool coord = std::find(coord_vars.begin(), coord_vars.end(), (int)q) != coord_vars.end();
    if (coord) {  // 1, 2, ..., n
      if (v.type == NC_DOUBLE) {
        buf.resize(v.nelems * 8);
        for (uint64_t e = 0; e < v.nelems; ++e) {
          const double x = (double)(e + 1);
          uint64_t bits;
          std::memcpy(&bits, &x, 8);
          for (int s = 0; s < 8; ++s) buf[8 * e + s] = (unsigned char)(bits >> (56 - 8 * s));
        }
      } else {
        std::vector<float> x(v.nelems);
        for (uint64_t e = 0; e < v.nelems; ++e) x[e] = (float)(e + 1);
        to_be32(x.data(), x.size(), buf);
      }
      ok = write_at(f, v.begin, buf.data(), buf.size());
    } else {  // fill mode: the whole variable holds the fill value until written
      const size_t chunk = 1 << 20;
      buf.resize(chunk * 4);
      for (size_t e = 0; e < chunk; ++e) {
        buf[4 * e] = (unsigned char)(FILL_FLOAT_BITS >> 24);
        buf[4 * e + 1] = (unsigned char)(FILL_FLOAT_BITS >> 16);
        buf[4 * e + 2] = (unsigned char)(FILL_FLOAT_BITS >> 8);
        buf[4 * e + 3] = (unsigned char)FILL_FLOAT_BITS;
      }
      ok = fseeko(f, (off_t)v.begin, SEEK_SET) == 0;
      for (uint64_t e = 0; e < v.nelems && ok; e += chunk) {
        const size_t n = (size_t)std::min<uint64_t>(chunk, v.nelems - e);
        ok = std::fwrite(buf.data(), 4, n, f) == n;
      }
    }
    // padding to the next 4-byte boundary is zero (only the double time variable can need none)
  }
  // the file ends at the last variable's padded end
  if (ok && nf.size > 0) {
    const unsigned char z[4] = {0, 0, 0, 0};
    const uint64_t last = nf.vars.back().begin + nf.vars.back().nelems * 4;
    if (nf.size > last) ok = write_at(f, last, z, (size_t)(nf.size - last));
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok;
}

}  // namespace

struct les_netcdf {
  int ip, jp, kp, ntimes;
  NcFile files[6];  // output, p, u, v, w, uvwsum
};

extern "C" int les_netcdf_create(const char* dir, int ip, int jp, int kp, int ntimes,
                                 les_netcdf** out) {
  if (!dir || !out || ip < 1 || jp < 1 || kp < 1 || ntimes < 1) return LES_E_INVALID;
  *out = nullptr;
  les_netcdf* nc = new les_netcdf();
  nc->ip = ip;
  nc->jp = jp;
  nc->kp = kp;
  nc->ntimes = ntimes;
  const std::string base = std::string(dir) + "/LES_output";
  const char* suffix[6] = {"", "_p", "_u", "_v", "_w", "_uvwsum"};
  // (lon, lat, lvl) extents of each file (module_LES_write_netcdf.f95:26-32,124-151)
  const uint32_t ext[6][3] = {{(uint32_t)ip + 3, (uint32_t)jp + 3, (uint32_t)kp + 2},
                              {(uint32_t)ip + 3, (uint32_t)jp + 3, (uint32_t)kp + 2},
                              {(uint32_t)ip + 2, (uint32_t)jp + 3, (uint32_t)kp + 2},
                              {(uint32_t)ip + 2, (uint32_t)jp + 3, (uint32_t)kp + 2},
                              {(uint32_t)ip + 2, (uint32_t)jp + 3, (uint32_t)kp + 3},
                              {(uint32_t)ip + 1, (uint32_t)jp + 1, (uint32_t)kp + 1}};
  const char* names[6][3] = {{"pressure", nullptr, nullptr},    {"pressure", nullptr, nullptr},
                             {"velocity_x", nullptr, nullptr},  {"velocity_y", nullptr, nullptr},
                             {"velocity_z", nullptr, nullptr},
                             {"velocity_sum_x", "velocity_sum_y", "velocity_sum_z"}};
  for (int q = 0; q < 6; ++q) {
    NcFile& nf = nc->files[q];
    nf.path = base + suffix[q] + ".nc";
    const int lvl = nf.add_dim("level", ext[q][2]);
    const int lat = nf.add_dim("latitude", ext[q][1]);
    const int lon = nf.add_dim("longitude", ext[q][0]);
    const int tim = nf.add_dim("time", (uint32_t)ntimes);
    nf.add_var("latitude", NC_FLOAT, {lat}, "degrees_north");
    nf.add_var("longitude", NC_FLOAT, {lon}, "degrees_east");
    nf.add_var("time", NC_DOUBLE, {tim}, "hours since 2014-08-05 16:00:00");
    for (int v = 0; v < 3 && names[q][v]; ++v)
      nf.add_var(names[q][v], NC_FLOAT, {tim, lvl, lat, lon}, q <= 1 ? "hPa" : "m/s");
    if (!create_file(nf, {0, 1, 2})) {
      delete nc;
      return LES_E_INVALID;
    }
  }
  *out = nc;
  return LES_OK;
}

extern "C" int les_netcdf_write_arrays(les_netcdf* nc, int t, const float* p, const float* u,
                                       const float* v, const float* w, const float* usum,
                                       const float* vsum, const float* wsum) {
  if (!nc || t < 1 || t > nc->ntimes || !p || !u || !v || !w || !usum || !vsum || !wsum)
    return LES_E_INVALID;
  const float* data[6][3] = {{p, nullptr, nullptr}, {p, nullptr, nullptr},
                             {u, nullptr, nullptr}, {v, nullptr, nullptr},
                             {w, nullptr, nullptr}, {usum, vsum, wsum}};
  std::vector<unsigned char> buf;
  for (int q = 0; q < 6; ++q) {
    NcFile& nf = nc->files[q];
    FILE* f = std::fopen(nf.path.c_str(), "r+b");
    if (!f) return LES_E_INVALID;
    bool ok = true;
    for (int s = 0; s < 3 && data[q][s] && ok; ++s) {
      const NcVar& var = nf.vars[3 + s];
      const uint64_t slab = var.nelems / (uint64_t)nc->ntimes;  // one time level
      to_be32(data[q][s], (size_t)slab, buf);
      ok = write_at(f, var.begin + (uint64_t)(t - 1) * slab * 4, buf.data(), buf.size());
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return LES_E_INVALID;
  }
  return LES_OK;
}

extern "C" int les_netcdf_write(les_netcdf* nc, les_handle* h, int t) {
  if (!nc || !h) return LES_E_INVALID;
  int ip, jp, kp, r;
  if ((r = les_get_dims(h, &ip, &jp, &kp))) return r;
  if (ip != nc->ip || jp != nc->jp || kp != nc->kp) return LES_E_INVALID;
  const int fields[7] = {LES_P, LES_U, LES_V, LES_W, LES_USUM, LES_VSUM, LES_WSUM};
  std::vector<float> a[7];
  for (int q = 0; q < 7; ++q) {
    int lo[3], hi[3];
    a[q].resize((size_t)les_field_bounds(fields[q], ip, jp, kp, lo, hi));
    if ((r = les_download(h, fields[q], a[q].data()))) return r;
  }
  return les_netcdf_write_arrays(nc, t, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                 a[4].data(), a[5].data(), a[6].data());
}

extern "C" int les_netcdf_close(les_netcdf* nc) {
  delete nc;  // every write closes its files; nothing is buffered
  return LES_OK;
}

// ---- anime.f95:26-49 ----
namespace {

bool put_record(FILE* f, const void* data, uint64_t nbytes) {
  if (nbytes >= (uint64_t)1 << 31) return false;  // single-record form only
  const uint32_t nb = (uint32_t)nbytes;
  return std::fwrite(&nb, 4, 1, f) == 1 && (nb == 0 || std::fwrite(data, 1, nb, f) == nb) &&
         std::fwrite(&nb, 4, 1, f) == 1;
}

}  // namespace

extern "C" int les_anime_arrays(int n, int n0, int nmax, int ip, int jp, int kp, const float* dx1,
                                const float* dy1, const float* z2, const float* u,
                                const float* v, const float* w, const float* amask1,
                                const char* data22, const char* data23) {
  if (!dx1 || !dy1 || !z2 || !u || !v || !w || !amask1 || !data22 || !data23 || ip < 1 ||
      jp < 1 || kp < 1)
    return LES_E_INVALID;
  if (!(n == n0 || n == nmax || n % 1000 == 0)) return LES_OK;  // anime.f95:26
  const size_t N = (size_t)ip * jp * kp;
  // dxl(0:ip), dyl(0:jp) as grid.f95:26-38 computes them; dx1 is (-1:ip+1), dy1 (0:jp+1)
  std::vector<float> dxl(ip + 1), dyl(jp + 1);
  dxl[0] = 0.0f;
  for (int i = 1; i <= ip; ++i) dxl[i] = dxl[i - 1] + dx1[i + 1];
  dyl[0] = 0.0f;
  for (int j = 1; j <= jp; ++j) dyl[j] = dyl[j - 1] + dy1[j];
  std::vector<float> rec(4 * N);
  float* a1 = rec.data();
  float* a3 = a1 + N;
  float* a2 = a3 + N;
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i) {
        const size_t e = (size_t)(i - 1) + (size_t)ip * ((j - 1) + (size_t)jp * (k - 1));
        a1[e] = dxl[i - 1] + dx1[i + 1];
        a2[e] = dyl[j - 1] + dy1[j];
        a3[e] = z2[k - 1];  // z2(1:kp+2)
      }
  FILE* f = std::fopen(data22, "wb");
  if (!f) return LES_E_INVALID;
  const int32_t h3[3] = {ip, jp, kp};
  bool ok = put_record(f, h3, 12) && put_record(f, rec.data(), 3 * N * 4);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return LES_E_INVALID;
  // u, v (0:ip+1,-1:jp+1,0:kp+1), w (0:ip+1,-1:jp+1,-1:kp+1), amask1 (0:ip+1,0:jp+1,0:kp+1)
  const size_t ui = ip + 2, uj = jp + 3;
  auto U = [&](const float* x, int i, int j, int k) { return x[i + ui * ((j + 1) + uj * k)]; };
  auto W = [&](int i, int j, int k) { return w[i + ui * ((j + 1) + uj * (k + 1))]; };
  auto A = [&](int i, int j, int k) {
    return amask1[i + (size_t)(ip + 2) * (j + (size_t)(jp + 2) * k)];
  };
  float* cu = rec.data();
  float* cw = cu + N;
  float* cv = cw + N;
  float* cm = cv + N;
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= jp; ++j)
      for (int i = 1; i <= ip; ++i) {
        const size_t e = (size_t)(i - 1) + (size_t)ip * ((j - 1) + (size_t)jp * (k - 1));
        cu[e] = 0.5f * (U(u, i - 1, j, k) + U(u, i, j, k));
        cw[e] = 0.5f * (W(i, j, k - 1) + W(i, j, k));
        cv[e] = 0.5f * (U(v, i, j - 1, k) + U(v, i, j, k));
        cm[e] = A(i, j, k);
      }
  f = std::fopen(data23, "wb");
  if (!f) return LES_E_INVALID;
  const int32_t h4[4] = {ip, jp, kp, 4};
  ok = put_record(f, h4, 16) && put_record(f, rec.data(), 4 * N * 4);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LES_OK : LES_E_INVALID;
}

extern "C" int les_anime(les_handle* h, int n, int n0, int nmax, const char* data22,
                         const char* data23) {
  if (!h) return LES_E_INVALID;
  int ip, jp, kp, r;
  if ((r = les_get_dims(h, &ip, &jp, &kp))) return r;
  if (!(n == n0 || n == nmax || n % 1000 == 0)) return LES_OK;
  const int fields[7] = {LES_DX1, LES_DY1, LES_Z2, LES_U, LES_V, LES_W, LES_AMASK1};
  std::vector<float> a[7];
  for (int q = 0; q < 7; ++q) {
    int lo[3], hi[3];
    a[q].resize((size_t)les_field_bounds(fields[q], ip, jp, kp, lo, hi));
    if ((r = les_download(h, fields[q], a[q].data()))) return r;
  }
  return les_anime_arrays(n, n0, nmax, ip, jp, kp, a[0].data(), a[1].data(), a[2].data(),
                          a[3].data(), a[4].data(), a[5].data(), a[6].data(), data22, data23);
}
