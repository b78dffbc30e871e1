// dwf_io.cpp -- gauge configuration files (host side).
//
// "Binary" (GaugeConfig("Binary"), src/lib/IO/gaugeConfig.cpp:61-107, with
// FieldIO_Binary, src/lib/IO/fieldIO_Binary.cpp:29-176, and the ILDG format,
// src/lib/IO/io_format_gauge.h:43-70): for every lexical site x + Nx (y + Ny
// (z + Nz t)), the 4 link matrices U_0..U_3 of that site, 18 doubles each
// (element ab = ri + 2 (b + 3 a)), big-endian.  On one process the
// "Binary_Parallel" files of the reference have the same layout.
// In memory the field is the corelib Field_G: U[mu][site][18].
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dwf_hip.h"

namespace {

inline uint64_t bswap64(uint64_t v) {
  v = ((v & 0x00000000ffffffffull) << 32) | (v >> 32);
  v = ((v & 0x0000ffff0000ffffull) << 16) | ((v >> 16) & 0x0000ffff0000ffffull);
  v = ((v & 0x00ff00ff00ff00ffull) << 8) | ((v >> 8) & 0x00ff00ff00ff00ffull);
  return v;
}

inline bool host_is_little_endian() {
  const uint16_t one = 1;
  uint8_t b;
  std::memcpy(&b, &one, 1);
  return b == 1;
}

// big-endian file value <-> host double
inline double from_file(uint64_t w) {
  if (host_is_little_endian()) w = bswap64(w);
  double d;
  std::memcpy(&d, &w, 8);
  return d;
}
inline uint64_t to_file(double d) {
  uint64_t w;
  std::memcpy(&w, &d, 8);
  return host_is_little_endian() ? bswap64(w) : w;
}

size_t volume(const int* dims) {
  return (size_t)dims[0] * dims[1] * dims[2] * dims[3];
}

}  // namespace

extern "C" {

int dwf_gauge_read_binary(const char* path, const int lattice_size[4], double* U_host) {
  if (!path || !lattice_size || !U_host) return DWF_ERR_INVALID_ARG;
  const size_t V = volume(lattice_size);
  FILE* f = std::fopen(path, "rb");
  if (!f) return DWF_ERR_INVALID_ARG;
  std::vector<uint64_t> site(4 * 18);
  int rc = DWF_OK;
  for (size_t s = 0; s < V && rc == DWF_OK; ++s) {
    if (std::fread(site.data(), 8, site.size(), f) != site.size()) {
      rc = DWF_ERR_INVALID_ARG;   // short file
      break;
    }
    for (int mu = 0; mu < 4; ++mu)
      for (int e = 0; e < 18; ++e) U_host[((size_t)mu * V + s) * 18 + e] = from_file(site[mu * 18 + e]);
  }
  std::fclose(f);
  return rc;
}

int dwf_gauge_write_binary(const char* path, const int lattice_size[4], const double* U_host) {
  if (!path || !lattice_size || !U_host) return DWF_ERR_INVALID_ARG;
  const size_t V = volume(lattice_size);
  FILE* f = std::fopen(path, "wb");
  if (!f) return DWF_ERR_INVALID_ARG;
  std::vector<uint64_t> site(4 * 18);
  int rc = DWF_OK;
  for (size_t s = 0; s < V; ++s) {
    for (int mu = 0; mu < 4; ++mu)
      for (int e = 0; e < 18; ++e) site[mu * 18 + e] = to_file(U_host[((size_t)mu * V + s) * 18 + e]);
    if (std::fwrite(site.data(), 8, site.size(), f) != site.size()) {
      rc = DWF_ERR_INVALID_ARG;
      break;
    }
  }
  if (std::fclose(f) != 0) rc = DWF_ERR_INVALID_ARG;
  return rc;
}

}  // extern "C"
