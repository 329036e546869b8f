// Film sinks on the host side (include/pyre_gen.h): the conversions of the
// reference's Film::save_png / save_aov_pngs / save_exr
// (crates/pyre/src/film.rs:117-264) over PixelSample arrays, and PNG / EXR
// writers for the packed buffers (these and the device packing of a
// session's film, pyre_session_pack, produce the same bytes).
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pyre_gen.h"
#include "pmath.hpp"

using namespace pyre;

namespace {

// Rust's `f32 as u8`: truncation, saturating, NaN -> 0.
uint8_t u8_sat(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

// film.rs:188-194: clamp to [0, 1] (glam's NaN-ignoring max/min), gamma
// 1/2.2 by powf, times 255.
uint8_t gamma8(float x) {
  const float c = rmin(rmax(x, 0.0f), 1.0f);
  return u8_sat(powf(c, 1.0f / 2.2f) * 255.0f);
}

// film.rs:138-143: 0.5 * (n + 1), clamped, times 255.
uint8_t normal8(float n) { return u8_sat(rmin(rmax(n * 0.5f + 0.5f, 0.0f), 1.0f) * 255.0f); }

// ----------------------------------------------------------------- PNG --

uint32_t crc32_of(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t r = i;
      for (int k = 0; k < 8; ++k) r = (r & 1) ? 0xEDB88320u ^ (r >> 1) : r >> 1;
      table[i] = r;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
  put_be32(out, (uint32_t)data.size());
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  put_be32(out, crc32_of(out.data() + start, out.size() - start) ^ 0xFFFFFFFFu);
}

// 8-bit grey (channels 1) or RGB (channels 3); filter 0 on every row, the
// zlib stream in stored blocks (lossless, no compression).
std::vector<uint8_t> encode_png(uint32_t w, uint32_t h, uint32_t channels, const uint8_t* px) {
  std::vector<uint8_t> raw;
  const size_t row = (size_t)w * channels;
  raw.reserve((row + 1) * h);
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), px + y * row, px + (y + 1) * row);
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  size_t off = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + n == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((uint8_t)n);
    z.push_back((uint8_t)(n >> 8));
    z.push_back((uint8_t)~n);
    z.push_back((uint8_t)(~n >> 8));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
    off += n;
  } while (off < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  put_be32(z, (b << 16) | a);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, w);
  put_be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, (uint8_t)(channels == 3 ? 2 : 0), 0, 0, 0});
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  return out;
}

// ----------------------------------------------------------------- EXR --

void put_le32(std::vector<uint8_t>& v, uint32_t x) {
  for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k)));
}

void put_str(std::vector<uint8_t>& v, const char* s) { v.insert(v.end(), s, s + std::strlen(s) + 1); }

void attr(std::vector<uint8_t>& v, const char* name, const char* type, const std::vector<uint8_t>& value) {
  put_str(v, name);
  put_str(v, type);
  put_le32(v, (uint32_t)value.size());
  v.insert(v.end(), value.begin(), value.end());
}

const char* const kExrNames[10] = {"B", "G", "N.X", "N.Y", "N.Z", "R", "Z", "albedo.B", "albedo.G", "albedo.R"};

// Channel k of the sorted set belongs to the AOV group: 0 beauty, 1 albedo,
// 2 normal, 4 depth (the mask bits of pyre_write_exr).
uint32_t exr_group(int k) { return k == 2 || k == 3 || k == 4 ? PYRE_AOV_NORMAL : k == 6 ? PYRE_AOV_DEPTH : k >= 7 ? PYRE_AOV_ALBEDO : 0u; }

// Single-part scanline OpenEXR, FLOAT channels in name order, one scanline
// per chunk, no compression.
std::vector<uint8_t> encode_exr(uint32_t w, uint32_t h, uint32_t mask, const float* planes) {
  std::vector<int> ch;
  for (int k = 0; k < 10; ++k)
    if (exr_group(k) == 0 || (mask & exr_group(k))) ch.push_back(k);
  std::vector<uint8_t> out = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
  std::vector<uint8_t> chl;
  for (int k : ch) {
    put_str(chl, kExrNames[k]);
    put_le32(chl, 2);  // FLOAT
    chl.insert(chl.end(), {0, 0, 0, 0});
    put_le32(chl, 1);
    put_le32(chl, 1);
  }
  chl.push_back(0);
  attr(out, "channels", "chlist", chl);
  attr(out, "compression", "compression", {0});
  std::vector<uint8_t> box;
  put_le32(box, 0);
  put_le32(box, 0);
  put_le32(box, w - 1);
  put_le32(box, h - 1);
  attr(out, "dataWindow", "box2i", box);
  attr(out, "displayWindow", "box2i", box);
  attr(out, "lineOrder", "lineOrder", {0});
  std::vector<uint8_t> one;
  const float f1 = 1.0f, f0 = 0.0f;
  uint32_t b1, b0;
  std::memcpy(&b1, &f1, 4);
  std::memcpy(&b0, &f0, 4);
  put_le32(one, b1);
  attr(out, "pixelAspectRatio", "float", one);
  std::vector<uint8_t> c2;
  put_le32(c2, b0);
  put_le32(c2, b0);
  attr(out, "screenWindowCenter", "v2f", c2);
  attr(out, "screenWindowWidth", "float", one);
  out.push_back(0);
  const size_t table = out.size();
  out.resize(table + 8 * (size_t)h);
  const size_t npix = (size_t)w * h;
  for (uint32_t y = 0; y < h; ++y) {
    const uint64_t pos = out.size();
    std::memcpy(out.data() + table + 8 * (size_t)y, &pos, 8);
    put_le32(out, y);
    put_le32(out, (uint32_t)(ch.size() * w * 4));
    for (int k : ch) {
      const float* src = planes + (size_t)k * npix + (size_t)y * w;
      const size_t at = out.size();
      out.resize(at + 4 * (size_t)w);
      std::memcpy(out.data() + at, src, 4 * (size_t)w);
    }
  }
  return out;
}

int write_file(const char* path, const std::vector<uint8_t>& bytes, char* err, size_t err_len) {
  FILE* f = std::fopen(path, "wb");
  bool ok = f && std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
  if (f) ok = std::fclose(f) == 0 && ok;
  if (ok) return PYRE_OK;
  if (err && err_len) std::snprintf(err, err_len, "io: %s (%s)", std::strerror(errno), path);
  return PYRE_EIO;
}

}  // namespace

extern "C" {

int pyre_film_pack(const pyre_pixel_sample* px, uint32_t width, uint32_t height, uint32_t what, void* out) {
  if (!px || !out || width == 0 || height == 0) return PYRE_EINVAL;
  const size_t n = (size_t)width * height;
  uint8_t* o8 = (uint8_t*)out;
  switch (what) {
    case PYRE_PACK_BEAUTY_PNG:
    case PYRE_PACK_ALBEDO_PNG:
      for (size_t i = 0; i < n; ++i) {
        const float* v = what == PYRE_PACK_BEAUTY_PNG ? px[i].radiance : px[i].albedo;
        for (int c = 0; c < 3; ++c) o8[3 * i + c] = gamma8(v[c]);
      }
      return PYRE_OK;
    case PYRE_PACK_NORMAL_PNG:
      for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) o8[3 * i + c] = normal8(px[i].normal[c]);
      return PYRE_OK;
    case PYRE_PACK_DEPTH_PNG: {
      // film.rs:149-170: min/max over depth > 0 (0 marks a miss), near = bright
      float lo = kInf, hi = 0.0f;
      for (size_t i = 0; i < n; ++i)
        if (px[i].depth > 0.0f) {
          lo = rmin(lo, px[i].depth);
          hi = rmax(hi, px[i].depth);
        }
      const float range = rmax(hi - lo, 1e-6f);
      for (size_t i = 0; i < n; ++i) {
        const float d = px[i].depth;
        const float v = d > 0.0f ? rmin(rmax(1.0f - (d - lo) / range, 0.0f), 1.0f) : 0.0f;
        o8[i] = u8_sat(v * 255.0f);
      }
      return PYRE_OK;
    }
    case PYRE_PACK_EXR: {
      float* f = (float*)out;
      for (size_t i = 0; i < n; ++i) {
        const pyre_pixel_sample& s = px[i];
        const float v[10] = {s.radiance[2], s.radiance[1], s.normal[0], s.normal[1], s.normal[2],
                             s.radiance[0], s.depth,       s.albedo[2], s.albedo[1], s.albedo[0]};
        for (int k = 0; k < 10; ++k) f[(size_t)k * n + i] = v[k];
      }
      return PYRE_OK;
    }
    default:
      return PYRE_EINVAL;
  }
}

int pyre_write_png(const char* path, uint32_t width, uint32_t height, uint32_t channels, const uint8_t* px,
                   char* err, size_t err_len) {
  if (!path || !px || width == 0 || height == 0 || (channels != 1 && channels != 3)) return PYRE_EINVAL;
  return write_file(path, encode_png(width, height, channels, px), err, err_len);
}

int pyre_write_exr(const char* path, uint32_t width, uint32_t height, uint32_t aov_mask, const float* planes,
                   char* err, size_t err_len) {
  if (!path || !planes || width == 0 || height == 0 || (aov_mask & ~7u)) return PYRE_EINVAL;
  return write_file(path, encode_exr(width, height, aov_mask, planes), err, err_len);
}

}  // extern "C"
