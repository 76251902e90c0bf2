// The output stage: Output.encode / Image.toRgb8 (image.zig:6-72), writePpm
// (image.zig:75-80) and the stored-block PNG writer (png.zig:10-131).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "zt_host.h"

namespace {

// Tonemap.apply (image.zig:15-22).
double tonemap(uint32_t op, double c) {
  const double x = std::fmax(c, 0.0);
  switch (op) {
    case ZT_TONEMAP_REINHARD:
      return x / (1.0 + x);
    case ZT_TONEMAP_ACES:
      return (x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14);
    default:
      return x;
  }
}

inline double clamp01(double v) { return std::fmin(std::fmax(v, 0.0), 1.0); }

// Output.encode (image.zig:30-37).
uint8_t encode(const zt_output& o, double linear) {
  const double v = std::isnan(linear) ? 0.0 : linear;
  const double mapped = clamp01(tonemap(o.tonemap, v * o.exposure));
  const double corrected = std::pow(mapped, 1.0 / o.gamma);
  return (uint8_t)std::round(clamp01(corrected) * 255.0);
}

// CRC-32 (ISO-HDLC, the PNG chunk checksum) and Adler-32 (the zlib trailer),
// written out from their definitions (RFC 1952 / RFC 1950).
struct Crc32 {
  uint32_t table[256];
  Crc32() {
    for (uint32_t n = 0; n < 256; ++n) {
      uint32_t c = n;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[n] = c;
    }
  }
  uint32_t update(uint32_t crc, const uint8_t* p, size_t n) const {
    uint32_t c = ~crc;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
  }
};
const Crc32& crcTable() {
  static const Crc32 t;
  return t;
}

struct Adler32 {
  uint32_t a = 1, b = 0;
  void update(const uint8_t* p, size_t n) {
    while (n > 0) {
      const size_t k = n < 5552 ? n : 5552;  // largest run with no 32-bit overflow
      for (size_t i = 0; i < k; ++i) {
        a += p[i];
        b += a;
      }
      a %= 65521;
      b %= 65521;
      p += k;
      n -= k;
    }
  }
  uint32_t final() const { return (b << 16) | a; }
};

void be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* kind, const uint8_t* data, size_t n) {  // png.zig:59-70
  be32(out, (uint32_t)n);
  const size_t at = out.size();
  out.insert(out.end(), kind, kind + 4);
  if (n) out.insert(out.end(), data, data + n);
  be32(out, crcTable().update(0, out.data() + at, 4 + n));
}

int deliver(const std::vector<uint8_t>& bytes, uint8_t* dst, uint64_t cap, uint64_t* len) {
  *len = bytes.size();
  if (!dst) return ZT_OK;
  if (cap < bytes.size()) return ZT_ERR_BUFFER_TOO_SMALL;
  std::memcpy(dst, bytes.data(), bytes.size());
  return ZT_OK;
}

constexpr size_t kMaxStoredBlock = 65535;  // png.zig:8

}  // namespace

extern "C" {

int zt_encode_rgb8(const double* linear, uint64_t count, const zt_output* out, uint8_t* rgb) {
  if ((count && (!linear || !rgb)) || !out) return ZT_ERR_INVALID_ARGUMENT;
  if (out->tonemap > ZT_TONEMAP_ACES) return ZT_ERR_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < count; ++i) rgb[i] = encode(*out, linear[i]);
  return ZT_OK;
}

int zt_write_ppm(uint64_t width, uint64_t height, const uint8_t* rgb, uint8_t* dst, uint64_t cap, uint64_t* len) {
  if (!len || ((width * height) != 0 && !rgb)) return ZT_ERR_INVALID_ARGUMENT;
  char head[64];
  const int h = std::snprintf(head, sizeof head, "P6\n%llu %llu\n255\n", (unsigned long long)width,
                              (unsigned long long)height);
  std::vector<uint8_t> bytes(head, head + h);
  bytes.insert(bytes.end(), rgb, rgb + width * height * 3);
  return deliver(bytes, dst, cap, len);
}

int zt_write_png(uint64_t width, uint64_t height, const uint8_t* rgb, uint8_t* dst, uint64_t cap, uint64_t* len) {
  if (!len) return ZT_ERR_INVALID_ARGUMENT;
  if (width == 0 || height == 0 || width > 0x7FFFFFFFull || height > 0x7FFFFFFFull) return ZT_ERR_INVALID_DIMENSIONS;
  if (!rgb) return ZT_ERR_INVALID_ARGUMENT;
  const uint64_t raw_len = (width * 3 + 1) * height;
  const uint64_t blocks = (raw_len + kMaxStoredBlock - 1) / kMaxStoredBlock;
  const uint64_t zlib_len = 2 + raw_len + 5 * blocks + 4;
  if (zlib_len > 0x7FFFFFFFull) return ZT_ERR_IMAGE_TOO_LARGE;
  std::vector<uint8_t> out;
  out.reserve(8 + 25 + 12 + zlib_len + 12);
  const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  out.insert(out.end(), sig, sig + 8);
  std::vector<uint8_t> ihdr;
  be32(ihdr, (uint32_t)width);
  be32(ihdr, (uint32_t)height);
  const uint8_t rest[5] = {8, 2, 0, 0, 0};  // bit depth, truecolour, deflate, filter 0, no interlace
  ihdr.insert(ihdr.end(), rest, rest + 5);
  chunk(out, "IHDR", ihdr.data(), ihdr.size());
  // IDAT: zlib header, stored blocks over the filtered scanlines, Adler-32.
  std::vector<uint8_t> z;
  z.reserve(zlib_len);
  z.push_back(0x78);
  z.push_back(0x01);
  Adler32 adler;
  uint64_t remaining = raw_len, block_left = 0;
  auto put = [&](const uint8_t* p, uint64_t n) {
    while (n > 0) {
      if (block_left == 0) {
        const uint16_t bl = (uint16_t)std::min<uint64_t>(remaining, kMaxStoredBlock);
        z.push_back(bl == remaining ? 1 : 0);  // BFINAL, BTYPE = stored
        z.push_back((uint8_t)(bl & 0xFF));
        z.push_back((uint8_t)(bl >> 8));
        z.push_back((uint8_t)(~bl & 0xFF));
        z.push_back((uint8_t)((uint16_t)~bl >> 8));
        block_left = bl;
      }
      const uint64_t k = std::min(n, block_left);
      z.insert(z.end(), p, p + k);
      adler.update(p, k);
      block_left -= k;
      remaining -= k;
      p += k;
      n -= k;
    }
  };
  const uint8_t filter_none = 0;
  for (uint64_t y = 0; y < height; ++y) {
    put(&filter_none, 1);
    put(rgb + y * width * 3, width * 3);
  }
  be32(z, adler.final());
  chunk(out, "IDAT", z.data(), z.size());
  chunk(out, "IEND", nullptr, 0);
  return deliver(out, dst, cap, len);
}

}  // extern "C"
