// Host helpers on exported images: the noisiness measurement
// (noisiness.rs:16-74) and the PNG writer (main.rs:193-221).
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "bundling.h"

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  if (!err || errlen == 0) return;
  const size_t n = std::min(errlen - 1, msg.size());
  std::memcpy(err, msg.data(), n);
  err[n] = '\0';
}

// validate_buffer (noisiness.rs:16-26)
bool validate(const uint8_t* rgba, uint32_t w, uint32_t h, char* err, size_t errlen) {
  if (w == 0 || h == 0) {
    set_err(err, errlen, "image dimensions must be nonzero");
    return false;
  }
  if (!rgba) {
    set_err(err, errlen,
            "invalid RGBA buffer length for " + std::to_string(w) + "x" + std::to_string(h));
    return false;
  }
  return true;
}

void put_be32(std::vector<uint8_t>& out, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) out.push_back((uint8_t)(v >> s));
}

void put_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n) {
  put_be32(out, (uint32_t)n);
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  if (n) out.insert(out.end(), data, data + n);
  const uLong crc = crc32(0L, out.data() + at, (uInt)(n + 4));
  put_be32(out, (uint32_t)crc);
}

}  // namespace

extern "C" {

int bt_noisiness_background(const uint8_t* rgba, uint32_t width, uint32_t height, uint8_t* rgb,
                            char* err, size_t errlen) {
  if (!rgb) return BT_E_INVALID;
  if (!validate(rgba, width, height, err, errlen)) return BT_E_INVALID;
  const uint8_t* p0 = rgba;
  if (p0[3] != 255) {
    set_err(err, errlen, "blank frame must be opaque");
    return BT_E_INVALID;
  }
  const size_t n = (size_t)width * height;
  for (size_t i = 1; i < n; ++i) {
    const uint8_t* p = rgba + 4 * i;
    if (p[0] != p0[0] || p[1] != p0[1] || p[2] != p0[2] || p[3] != 255) {
      set_err(err, errlen, "blank frame must have uniform RGB and opaque alpha");
      return BT_E_INVALID;
    }
  }
  rgb[0] = p0[0];
  rgb[1] = p0[1];
  rgb[2] = p0[2];
  return BT_OK;
}

int bt_noisiness_measure(const uint8_t* rgba, uint32_t width, uint32_t height,
                         const uint8_t* rgb, bt_noisiness* out, char* err, size_t errlen) {
  if (!rgb || !out) return BT_E_INVALID;
  if (!validate(rgba, width, height, err, errlen)) return BT_E_INVALID;
  const uint64_t total = (uint64_t)width * height;
  uint64_t colored = 0;
  for (uint64_t i = 0; i < total; ++i) {
    const uint8_t* p = rgba + 4 * i;
    colored += (p[0] != rgb[0] || p[1] != rgb[1] || p[2] != rgb[2]) ? 1u : 0u;
  }
  out->colored_pixels = colored;
  out->background_pixels = total - colored;
  out->total_pixels = total;
  out->noisiness = (double)colored / (double)total;
  out->colored_to_background_ratio = out->background_pixels > 0
                                         ? (double)colored / (double)out->background_pixels
                                         : std::numeric_limits<double>::quiet_NaN();
  return BT_OK;
}

int bt_write_png(const char* path, const uint8_t* rgba, uint32_t width, uint32_t height,
                 char* err, size_t errlen) {
  if (!path) return BT_E_INVALID;
  if (!validate(rgba, width, height, err, errlen)) return BT_E_INVALID;
  const size_t row = (size_t)width * 4;
  std::vector<uint8_t> raw((row + 1) * height);
  for (uint32_t y = 0; y < height; ++y) {
    raw[y * (row + 1)] = 0;  // filter type None
    std::memcpy(&raw[y * (row + 1) + 1], rgba + y * row, row);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) {
    set_err(err, errlen, "deflate failed");
    return BT_E_INVALID;
  }
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, width);
  put_be32(ihdr, height);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, filter 0, no interlace
  put_chunk(png, "IHDR", ihdr.data(), ihdr.size());
  put_chunk(png, "IDAT", z.data(), zlen);
  put_chunk(png, "IEND", nullptr, 0);
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_err(err, errlen, std::string("create ") + path);
    return BT_E_INVALID;
  }
  const bool ok = std::fwrite(png.data(), 1, png.size(), fp) == png.size();
  if (std::fclose(fp) != 0 || !ok) {
    set_err(err, errlen, std::string("write ") + path);
    return BT_E_INVALID;
  }
  return BT_OK;
}

}  // extern "C"
