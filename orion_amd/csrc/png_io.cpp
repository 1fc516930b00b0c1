// 8-bit PNG reading and writing for environment maps.  orion loads envmap
// images through stb_image's stbi_load(path, &w, &h, &channels, 0) and keeps
// 3- or 4-channel results (texture/imagetexture.cpp:22-48); this reader
// returns the same bytes and channel count for 8-bit, non-interlaced PNGs
// (grey 1, grey+alpha 2, RGB 3, RGBA 4, palette 3, or 4 with tRNS), the
// forms envmap images come in.  Inflate is zlib's.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_scene.h"

namespace orion_amd {
namespace {

uint32_t be32(const unsigned char *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
void put32(std::vector<unsigned char> *o, uint32_t v) {
  o->push_back((unsigned char)(v >> 24)); o->push_back((unsigned char)(v >> 16));
  o->push_back((unsigned char)(v >> 8)); o->push_back((unsigned char)v);
}
int paeth(int a, int b, int c) {
  int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

}  // namespace

bool png_read(const std::string &path, int *width, int *height, int *channels, std::vector<unsigned char> *px,
              std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) { *err = "image not found - " + path; return false; }
  std::vector<unsigned char> file;
  unsigned char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) file.insert(file.end(), buf, buf + n);
  std::fclose(f);
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0) { *err = path + ": not a PNG file"; return false; }
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<unsigned char> idat, plte, trns;
  size_t pos = 8;
  bool end = false;
  while (!end) {
    if (pos + 12 > file.size()) { *err = path + ": truncated PNG"; return false; }
    uint32_t len = be32(&file[pos]);
    if (len > file.size() - pos - 12) { *err = path + ": truncated PNG chunk"; return false; }
    const unsigned char *type = &file[pos + 4], *data = &file[pos + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13) { *err = path + ": bad IHDR"; return false; }
      w = be32(data); h = be32(data + 4);
      depth = data[8]; ctype = data[9]; interlace = data[12];
      if (data[10] != 0 || data[11] != 0) { *err = path + ": unknown PNG compression/filter method"; return false; }
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(data, data + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(data, data + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      end = true;
    }
    pos += 12 + (size_t)len;
  }
  if (w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15)) { *err = path + ": unsupported PNG size"; return false; }
  if (depth != 8) { *err = path + ": only 8-bit PNG images are supported"; return false; }
  if (interlace != 0) { *err = path + ": interlaced PNG images are not supported"; return false; }
  int spp;  // samples per pixel in the file
  switch (ctype) {
    case 0: spp = 1; break;
    case 2: spp = 3; break;
    case 3: spp = 1; break;
    case 4: spp = 2; break;
    case 6: spp = 4; break;
    default: *err = path + ": unknown PNG colour type"; return false;
  }
  if (ctype == 3 && (plte.empty() || plte.size() % 3)) { *err = path + ": palette image without PLTE"; return false; }
  const size_t stride = (size_t)w * spp;
  std::vector<unsigned char> raw((stride + 1) * h);
  uLongf rawlen = (uLongf)raw.size();
  if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size()) {
    *err = path + ": corrupt PNG image data";
    return false;
  }
  std::vector<unsigned char> img(stride * h);
  for (uint32_t y = 0; y < h; ++y) {
    const unsigned char *in = &raw[y * (stride + 1)];
    unsigned char *out = &img[y * stride];
    const unsigned char *prev = y ? &img[(y - 1) * stride] : nullptr;
    int ft = in[0];
    ++in;
    for (size_t x = 0; x < stride; ++x) {
      int a = x >= (size_t)spp ? out[x - spp] : 0;
      int b = prev ? prev[x] : 0;
      int c = (prev && x >= (size_t)spp) ? prev[x - spp] : 0;
      int v;
      switch (ft) {
        case 0: v = in[x]; break;
        case 1: v = in[x] + a; break;
        case 2: v = in[x] + b; break;
        case 3: v = in[x] + ((a + b) >> 1); break;
        case 4: v = in[x] + paeth(a, b, c); break;
        default: *err = path + ": bad PNG filter type"; return false;
      }
      out[x] = (unsigned char)v;
    }
  }
  if (ctype == 3) {
    int ch = trns.empty() ? 3 : 4;
    px->resize((size_t)w * h * ch);
    size_t np = plte.size() / 3;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
      size_t k = img[i];
      if (k >= np) { *err = path + ": palette index out of range"; return false; }
      for (int c = 0; c < 3; ++c) (*px)[i * ch + c] = plte[3 * k + c];
      if (ch == 4) (*px)[i * ch + 3] = k < trns.size() ? trns[k] : 255;
    }
    *channels = ch;
  } else {
    *px = std::move(img);
    *channels = spp;
  }
  *width = (int)w;
  *height = (int)h;
  return true;
}

bool png_write_rgb8(const std::string &path, int width, int height, const unsigned char *rgb, std::string *err) {
  std::vector<unsigned char> raw;
  raw.reserve((size_t)height * (3 * (size_t)width + 1));
  for (int y = 0; y < height; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb + (size_t)y * 3 * width, rgb + (size_t)(y + 1) * 3 * width);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) { *err = "PNG compression failed"; return false; }
  z.resize(zlen);
  std::vector<unsigned char> o = {137, 80, 78, 71, 13, 10, 26, 10};
  auto chunk = [&](const char *type, const std::vector<unsigned char> &d) {
    put32(&o, (uint32_t)d.size());
    size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), d.begin(), d.end());
    put32(&o, (uint32_t)crc32(0L, &o[start], (uInt)(o.size() - start)));
  };
  std::vector<unsigned char> ihdr;
  put32(&ihdr, (uint32_t)width);
  put32(&ihdr, (uint32_t)height);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  chunk("IHDR", ihdr);
  chunk("IDAT", z);
  chunk("IEND", {});
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { *err = "cannot write " + path; return false; }
  bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) *err = "cannot write " + path;
  return ok;
}

}  // namespace orion_amd
