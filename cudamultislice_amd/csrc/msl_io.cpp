// msl_io.cpp -- on-disk output: 32-bit float TIFF files as the reference's GUI
// writes them through LibTiff.NET ("CUDA Simulation GUI/MainWindow.xaml.cs":
// 684-890): one sample per pixel, IEEE float (SampleFormat 3), 32 bits, one
// strip holding every row, top-left origin, min-is-black, no compression.
// Written here directly (baseline TIFF 6.0, little-endian), with no library.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "msl_host.hpp"

using namespace mslh;

namespace {

void put16(std::vector<uint8_t>& b, uint16_t v) {
  b.push_back((uint8_t)(v & 0xff));
  b.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t>& b, uint32_t v) {
  for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}

// one IFD entry: SHORT (type 3) or LONG (type 4), count 1, value inline
void entry(std::vector<uint8_t>& b, uint16_t tag, uint16_t type, uint32_t value) {
  put16(b, tag);
  put16(b, type);
  put32(b, 1);
  if (type == 3) {
    put16(b, (uint16_t)value);
    put16(b, 0);
  } else {
    put32(b, value);
  }
}

int write_tiff(const char* path, int w, int h, const float* data) {
  if (!path || !data || w <= 0 || h <= 0) return MSL_ERR_ARG;
  const uint64_t bytes = (uint64_t)w * h * 4;
  if (bytes > 0xFFFFFFF0ULL) return MSL_ERR_UNSUPPORTED;
  constexpr uint16_t kEntries = 13;
  const uint32_t ifd = 8, data_off = ifd + 2 + 12 * kEntries + 4;
  std::vector<uint8_t> b;
  b.reserve(data_off);
  b.push_back('I');
  b.push_back('I');
  put16(b, 42);
  put32(b, ifd);
  put16(b, kEntries);  // tags in ascending order
  entry(b, 256, 4, (uint32_t)w);               // ImageWidth
  entry(b, 257, 4, (uint32_t)h);               // ImageLength
  entry(b, 258, 3, 32);                        // BitsPerSample
  entry(b, 259, 3, 1);                         // Compression: none
  entry(b, 262, 3, 1);                         // Photometric: min-is-black
  entry(b, 266, 3, 1);                         // FillOrder: MSB2LSB
  entry(b, 273, 4, data_off);                  // StripOffsets
  entry(b, 274, 3, 1);                         // Orientation: top-left
  entry(b, 277, 3, 1);                         // SamplesPerPixel
  entry(b, 278, 4, (uint32_t)h);               // RowsPerStrip: one strip
  entry(b, 279, 4, (uint32_t)bytes);           // StripByteCounts
  entry(b, 284, 3, 1);                         // PlanarConfiguration: contiguous
  entry(b, 339, 3, 3);                         // SampleFormat: IEEE float
  put32(b, 0);                                 // no further IFD
  FILE* f = std::fopen(path, "wb");
  if (!f) return MSL_ERR_IO;
  bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  // float32 little-endian (the host byte order on x86-64 and AArch64 Linux)
  for (int y = 0; y < h && ok; ++y) ok = std::fwrite(data + (size_t)y * w, 4, (size_t)w, f) == (size_t)w;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MSL_OK : MSL_ERR_IO;
}

}  // namespace

extern "C" {

msl_status msl_write_tiff(const char* path, int width, int height, const float* data) {
  return write_tiff(path, width, height, data);
}

msl_status msl_save_tiff(msl_sim* s, int what, const char* path) {
  API_BEGIN(s)
  require(path != nullptr, MSL_ERR_ARG, "null path");
  const int w = s->resX, h = s->resY;
  const size_t n = s->npix();
  std::vector<float> img(n);
  if (what == MSL_TIFF_EW_RE || what == MSL_TIFF_EW_IM) {
    require(s->have_waves, MSL_ERR_STATE, "no exit wave");
    std::vector<float2> ew(n);
    d2h(s, ew.data(), s->Psi.p, n * sizeof(float2));
    for (size_t i = 0; i < n; ++i) img[i] = what == MSL_TIFF_EW_RE ? ew[i].x : ew[i].y;
  } else if (what == MSL_TIFF_IMAGE) {
    require(s->have_image, MSL_ERR_STATE, "SimulateImage first");
    if (s->have_detimg) {
      // GetImValue per pixel: readout noise (draw = pixel index) + |image|
      std::vector<float2> z(n);
      d2h(s, z.data(), s->detimg.p, n * sizeof(float2));
      for (size_t i = 0; i < n; ++i) {
        float r = 4.f;
        msl_readout_noise(s, (long long)i, &r);
        img[i] = r + std::sqrt(z[i].x * z[i].x + z[i].y * z[i].y);
      }
    } else {
      d2h(s, img.data(), s->image.p, n * sizeof(float));
    }
  } else if (what == MSL_TIFF_TDS) {
    require(s->have_tds, MSL_ERR_STATE, "AllocTDS / AddTDSWaves first");
    d2h(s, img.data(), s->tds.p, n * sizeof(float));
  } else {
    require(false, MSL_ERR_ARG, "unknown export");
  }
  const int st = write_tiff(path, w, h, img.data());
  require(st == MSL_OK, st, std::string("cannot write ") + path);
  API_END(s)
}

}  // extern "C"
