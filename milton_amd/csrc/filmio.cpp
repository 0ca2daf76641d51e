// Image files: the writers behind mpt_film_save and mpt_image_write_* (the
// file side of FileRenderOutput::_save, renderers/outputs/FileRenderOutput.cpp:
// 57-76 -> Image::save, common/image/Image.cpp:103-129, which picks the
// format from the file name's extension) and the reader behind
// mpt_image_read (Image::load, Image.cpp:67-101) for environment images.
// Host code only.
//
// Formats (all rows top to bottom unless the format says otherwise):
//   PNG   8-bit RGB, one IDAT of zlib-compressed scanlines with filter 0
//   PPM   binary P6, maxval 255
//   RGBE  Radiance .hdr: "#?RGBE" header, FORMAT=32-bit_rle_rgbe,
//         "-Y H +X W", scanlines run-length encoded per channel when
//         8 <= W < 32768, flat otherwise (HDRUtils::saveHDR,
//         common/image/HDRUtils.cpp:114-148, over the published RGBE codec)
//   PFM   "PF", W H, scale -1 (little-endian), RGB float rows bottom to top
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <new>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"

namespace {

// Images larger than this many pixels are refused when read.
constexpr uint64_t kMaxReadPixels = 1ull << 28;

enum Format { F_NONE, F_PNG, F_PPM, F_RGBE, F_PFM };

Format format_of(const char *path) {
  const std::string p(path);
  const size_t dot = p.rfind('.');
  if (dot == std::string::npos) return F_NONE;
  std::string ext = p.substr(dot + 1);
  for (char &c : ext) c = (char)std::tolower((unsigned char)c);
  if (ext == "png") return F_PNG;
  if (ext == "ppm") return F_PPM;
  if (ext == "hdr" || ext == "rgbe") return F_RGBE;
  if (ext == "pfm") return F_PFM;
  return F_NONE;
}

struct File {
  FILE *f;
  explicit File(const char *path) : f(std::fopen(path, "wb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
  bool put(const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
  bool close() {
    const bool ok = std::fclose(f) == 0;
    f = nullptr;
    return ok;
  }
};

void be32(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

void png_chunk(std::vector<uint8_t> &out, const char *type, const uint8_t *data, size_t n) {
  be32(out, (uint32_t)n);
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  if (n) out.insert(out.end(), data, data + n);
  be32(out, (uint32_t)crc32(0, out.data() + at, (uInt)(n + 4)));
}

bool write_png(const char *path, uint32_t w, uint32_t h, const uint8_t *rgba) {
  const size_t row = 1 + 3 * (size_t)w;
  std::vector<uint8_t> raw(row * h);
  for (uint32_t y = 0; y < h; ++y) {
    uint8_t *r = raw.data() + row * y;
    r[0] = 0;   // filter: none
    for (uint32_t x = 0; x < w; ++x)
      std::memcpy(r + 1 + 3 * (size_t)x, rgba + 4 * ((size_t)y * w + x), 3);
  }
  uLongf zn = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zn);
  if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 3) != Z_OK) return false;
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  be32(ihdr, w);
  be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});   // 8-bit RGB, deflate, filter 0, no interlace
  png_chunk(out, "IHDR", ihdr.data(), ihdr.size());
  png_chunk(out, "IDAT", z.data(), zn);
  png_chunk(out, "IEND", nullptr, 0);
  File f(path);
  return f.f && f.put(out.data(), out.size()) && f.close();
}

bool write_ppm(const char *path, uint32_t w, uint32_t h, const uint8_t *rgba) {
  File f(path);
  if (!f.f) return false;
  char head[64];
  const int n = std::snprintf(head, sizeof head, "P6\n%u %u\n255\n", w, h);
  if (!f.put(head, (size_t)n)) return false;
  std::vector<uint8_t> rgb(3 * (size_t)w);
  for (uint32_t y = 0; y < h; ++y) {
    for (uint32_t x = 0; x < w; ++x) std::memcpy(&rgb[3 * (size_t)x], rgba + 4 * ((size_t)y * w + x), 3);
    if (!f.put(rgb.data(), rgb.size())) return false;
  }
  return f.close();
}

// One pixel in shared-exponent form: the largest component's binary exponent
// e (frexp: m = f 2^e, f in [0.5, 1)) is stored as e + 128 and each
// component as floor(c 2^(8-e)) (exact: a power-of-two scaling), so the
// largest lands in [128, 255]; a largest component below 1e-32, or a NaN,
// is stored as zero, and negative components as 0.
void to_rgbe(float r, float g, float b, uint8_t *o) {
  const float m = std::max(r, std::max(g, b));
  if (!(m >= 1e-32f)) {
    o[0] = o[1] = o[2] = o[3] = 0;
    return;
  }
  int e;
  (void)std::frexp(m, &e);
  const float c[3] = {r, g, b};
  for (int i = 0; i < 3; ++i) o[i] = c[i] > 0 ? (uint8_t)std::ldexp((double)c[i], 8 - e) : 0;
  o[3] = (uint8_t)(e + 128);
}

// Run-length encodes one channel of a scanline: a run of 4..127 equal bytes
// becomes (128 + length, byte); anything else goes in literal blocks of at
// most 128 bytes (length, bytes).
void rle_channel(const uint8_t *v, uint32_t n, std::vector<uint8_t> &out) {
  uint32_t i = 0;
  while (i < n) {
    // the next run of at least 4 at or after i
    uint32_t run_at = i, run_len = 0;
    while (run_at < n) {
      run_len = 1;
      while (run_at + run_len < n && run_len < 127 && v[run_at + run_len] == v[run_at]) ++run_len;
      if (run_len >= 4) break;
      run_at += run_len;
      run_len = 0;
    }
    while (i < run_at) {
      const uint32_t lit = std::min<uint32_t>(128, run_at - i);
      out.push_back((uint8_t)lit);
      out.insert(out.end(), v + i, v + i + lit);
      i += lit;
    }
    if (run_len >= 4) {
      out.push_back((uint8_t)(128 + run_len));
      out.push_back(v[run_at]);
      i = run_at + run_len;
    }
  }
}

bool write_rgbe(const char *path, uint32_t w, uint32_t h, const float *rgb) {
  File f(path);
  if (!f.f) return false;
  char head[128];
  const int n = std::snprintf(head, sizeof head, "#?RGBE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
  if (!f.put(head, (size_t)n)) return false;
  std::vector<uint8_t> px(4 * (size_t)w), ch(w), out;
  for (uint32_t y = 0; y < h; ++y) {
    const float *s = rgb + 3 * (size_t)y * w;
    for (uint32_t x = 0; x < w; ++x) to_rgbe(s[3 * x], s[3 * x + 1], s[3 * x + 2], &px[4 * (size_t)x]);
    out.clear();
    if (w < 8 || w > 0x7fff) {
      out.insert(out.end(), px.begin(), px.end());
    } else {
      out.insert(out.end(), {2, 2, (uint8_t)(w >> 8), (uint8_t)(w & 0xff)});
      for (int c = 0; c < 4; ++c) {
        for (uint32_t x = 0; x < w; ++x) ch[x] = px[4 * (size_t)x + c];
        rle_channel(ch.data(), w, out);
      }
    }
    if (!f.put(out.data(), out.size())) return false;
  }
  return f.close();
}

bool write_pfm(const char *path, uint32_t w, uint32_t h, const float *rgb) {
  File f(path);
  if (!f.f) return false;
  char head[64];
  const int n = std::snprintf(head, sizeof head, "PF\n%u %u\n-1.0\n", w, h);
  if (!f.put(head, (size_t)n)) return false;
  static_assert(sizeof(float) == 4, "PFM stores 4-byte floats");
  for (uint32_t y = h; y--;)   // bottom row first
    if (!f.put(rgb + 3 * (size_t)y * w, 12 * (size_t)w)) return false;
  return f.close();
}

// ---------------------------------------------------------------------------
// Readers.
// ---------------------------------------------------------------------------

struct In {
  FILE *f;
  explicit In(const char *path) : f(std::fopen(path, "rb")) {}
  ~In() {
    if (f) std::fclose(f);
  }
  bool line(std::string &s) {   // one line without its '\n'; false at EOF
    s.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF && c != '\n') s.push_back((char)c);
    return c != EOF || !s.empty();
  }
  bool get(void *p, size_t n) { return std::fread(p, 1, n, f) == n; }
};

// RGBE_ReadHeader as the reference's copy of the codec reads it
// (common/image/rgbe.cpp:159-216): lines up to the first blank one, one of
// them "FORMAT=32-bit_rle_rgbe", then "-Y height +X width".
bool read_rgbe(const char *path, uint32_t &w, uint32_t &h, std::vector<float> &rgb) {
  In in(path);
  if (!in.f) return false;
  std::string s;
  bool format = false;
  for (;;) {
    if (!in.line(s)) return false;
    if (s.empty()) break;
    if (s == "FORMAT=32-bit_rle_rgbe") format = true;
  }
  if (!format || !in.line(s)) return false;
  int hh = 0, ww = 0;
  if (std::sscanf(s.c_str(), "-Y %d +X %d", &hh, &ww) != 2 || hh <= 0 || ww <= 0) return false;
  if ((uint64_t)hh * (uint64_t)ww > kMaxReadPixels) return false;
  w = (uint32_t)ww;
  h = (uint32_t)hh;
  rgb.assign(3 * (size_t)w * h, 0.0f);
  std::vector<uint8_t> px(4 * (size_t)w);
  for (uint32_t y = 0; y < h; ++y) {
    bool flat = w < 8 || w > 0x7fff;
    size_t have = 0;   // pixels of this scanline already in px
    if (!flat) {
      uint8_t b[4];
      if (!in.get(b, 4)) return false;
      if (b[0] != 2 || b[1] != 2 || (b[2] & 0x80)) {
        // not run-length encoded: these four bytes are the first pixel
        std::memcpy(px.data(), b, 4);
        have = 1;
        flat = true;
      } else {
        if ((((uint32_t)b[2] << 8) | b[3]) != w) return false;
        for (int c = 0; c < 4; ++c) {
          uint32_t x = 0;
          while (x < w) {
            uint8_t hd[2];
            if (!in.get(hd, 2)) return false;
            if (hd[0] > 128) {
              const uint32_t n = hd[0] - 128u;
              if (n > w - x) return false;
              for (uint32_t i = 0; i < n; ++i) px[4 * (size_t)(x + i) + c] = hd[1];
              x += n;
            } else {
              const uint32_t n = hd[0];
              if (n == 0 || n > w - x) return false;
              px[4 * (size_t)x + c] = hd[1];
              for (uint32_t i = 1; i < n; ++i) {
                uint8_t v;
                if (!in.get(&v, 1)) return false;
                px[4 * (size_t)(x + i) + c] = v;
              }
              x += n;
            }
          }
        }
      }
    }
    if (flat && !in.get(px.data() + 4 * have, 4 * (w - have))) return false;
    float *out = rgb.data() + 3 * (size_t)y * w;
    for (uint32_t x = 0; x < w; ++x) {
      const uint8_t *q = &px[4 * (size_t)x];
      // rgbe2float: byte * 2^(e - 136), zero when e = 0
      const float f = q[3] ? (float)std::ldexp(1.0, (int)q[3] - 136) : 0.0f;
      for (int c = 0; c < 3; ++c) out[3 * x + c] = q[c] * f;
    }
  }
  return true;
}

// Header tokens of PFM / PPM: whitespace-separated, '#' comments (PPM).
bool token(In &in, std::string &t) {
  t.clear();
  int c;
  while ((c = std::fgetc(in.f)) != EOF) {
    if (c == '#' && t.empty()) {
      while ((c = std::fgetc(in.f)) != EOF && c != '\n') {
      }
      continue;
    }
    if (std::isspace(c)) {
      if (!t.empty()) return true;   // consumes exactly one whitespace after the token
      continue;
    }
    t.push_back((char)c);
  }
  return !t.empty();
}

bool read_pfm(const char *path, uint32_t &w, uint32_t &h, std::vector<float> &rgb) {
  In in(path);
  if (!in.f) return false;
  std::string magic, sw, sh, ss;
  if (!token(in, magic) || (magic != "PF" && magic != "Pf") || !token(in, sw) || !token(in, sh) || !token(in, ss))
    return false;
  const long ww = std::atol(sw.c_str()), hh = std::atol(sh.c_str());
  const double scale = std::atof(ss.c_str());
  if (ww <= 0 || hh <= 0 || scale == 0 || (uint64_t)ww * (uint64_t)hh > kMaxReadPixels) return false;
  w = (uint32_t)ww;
  h = (uint32_t)hh;
  const int ch = magic == "PF" ? 3 : 1;
  const bool big = scale > 0;
  std::vector<uint8_t> row(4 * (size_t)ch * w);
  rgb.assign(3 * (size_t)w * h, 0.0f);
  for (uint32_t y = h; y--;) {   // bottom row first
    if (!in.get(row.data(), row.size())) return false;
    for (size_t i = 0; i < (size_t)ch * w; ++i) {
      uint8_t b[4];
      std::memcpy(b, &row[4 * i], 4);
      if (big) {
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
      }
      float v;
      std::memcpy(&v, b, 4);
      const size_t x = i / ch;
      if (ch == 3) rgb[3 * ((size_t)y * w + x) + i % 3] = v;
      else
        for (int c = 0; c < 3; ++c) rgb[3 * ((size_t)y * w + x) + c] = v;
    }
  }
  return true;
}

bool read_ppm(const char *path, uint32_t &w, uint32_t &h, std::vector<float> &rgb) {
  In in(path);
  if (!in.f) return false;
  std::string magic, sw, sh, sm;
  if (!token(in, magic) || magic != "P6" || !token(in, sw) || !token(in, sh) || !token(in, sm)) return false;
  const long ww = std::atol(sw.c_str()), hh = std::atol(sh.c_str()), mx = std::atol(sm.c_str());
  if (ww <= 0 || hh <= 0 || mx != 255 || (uint64_t)ww * (uint64_t)hh > kMaxReadPixels) return false;
  w = (uint32_t)ww;
  h = (uint32_t)hh;
  std::vector<uint8_t> px(3 * (size_t)w * h);
  if (!in.get(px.data(), px.size())) return false;
  rgb.resize(px.size());
  // Rgba32 -> Vector3: byte / 255 (common/image/Rgba.inl:32-36)
  for (size_t i = 0; i < px.size(); ++i) rgb[i] = (float)(px[i] / 255.0);
  return true;
}

}  // namespace

extern "C" {

int mpt_image_read(const char *path, uint32_t *width, uint32_t *height, float *rgb, size_t capacity) {
  if (!path || !width || !height) return MPT_E_INVALID;
  uint32_t w = 0, h = 0;
  std::vector<float> img;
  bool ok;
  try {
    switch (format_of(path)) {
      case F_RGBE:
        ok = read_rgbe(path, w, h, img);
        break;
      case F_PFM:
        ok = read_pfm(path, w, h, img);
        break;
      case F_PPM:
        ok = read_ppm(path, w, h, img);
        break;
      default:
        ok = false;
    }
  } catch (const std::bad_alloc &) {
    return MPT_E_NOMEM;
  }
  if (!ok) return MPT_E_IO;
  *width = w;
  *height = h;
  if (!rgb || capacity < img.size()) return MPT_E_INVALID;
  std::memcpy(rgb, img.data(), img.size() * sizeof(float));
  return MPT_OK;
}

uint32_t mpt_tonemap_from_name(const char *name) {
  if (!name) return MPT_TONEMAP_REINHARD;
  const std::string s(name);
  return (s == "reinhard" || s == "default") ? MPT_TONEMAP_REINHARD : MPT_TONEMAP_GAMMA;
}

int mpt_image_write_ldr(const char *path, uint32_t w, uint32_t h, const uint8_t *rgba) {
  if (!path || !rgba || !w || !h) return MPT_E_INVALID;
  try {
    switch (format_of(path)) {
      case F_PNG:
        return write_png(path, w, h, rgba) ? MPT_OK : MPT_E_IO;
      case F_PPM:
        return write_ppm(path, w, h, rgba) ? MPT_OK : MPT_E_IO;
      case F_RGBE:
      case F_PFM: {
        // Rgba32 -> RgbaHDR: byte / 255 (common/image/Rgba.inl:32-36)
        std::vector<float> rgb(3 * (size_t)w * h);
        for (size_t i = 0; i < (size_t)w * h; ++i)
          for (int c = 0; c < 3; ++c) rgb[3 * i + c] = (float)(rgba[4 * i + c] / 255.0);
        return mpt_image_write_hdr(path, w, h, rgb.data());
      }
      default:
        return MPT_E_IO;
    }
  } catch (const std::bad_alloc &) {
    return MPT_E_NOMEM;
  }
}

int mpt_image_write_hdr(const char *path, uint32_t w, uint32_t h, const float *rgb) {
  if (!path || !rgb || !w || !h) return MPT_E_INVALID;
  try {
    switch (format_of(path)) {
      case F_RGBE:
        return write_rgbe(path, w, h, rgb) ? MPT_OK : MPT_E_IO;
      case F_PFM:
        return write_pfm(path, w, h, rgb) ? MPT_OK : MPT_E_IO;
      default:
        return MPT_E_IO;
    }
  } catch (const std::bad_alloc &) {
    return MPT_E_NOMEM;
  }
}

}  // extern "C"
