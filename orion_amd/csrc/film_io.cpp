// Film output: RenderTarget::getAveraged on host buffers and the image files
// an averaged film is written to.
//
//   BMP  saveAsBmp (texture/texture.cpp:77-137): 'BM', a 12-byte file header
//        and a 40-byte info header (the layout the reference's structs have
//        where long is 32 bits, its MSVC build; under LP64 the same structs
//        grow to 24 + 80 bytes and the file is not a BMP), then 32-bit BGRA
//        rows bottom-up, each channel (char)(clamp(v, 0, 1) * 255), alpha 255.
//   PFM  FP32 RGB, "PF" header, little-endian scale -1, bottom row first.
//   EXR  OpenEXR single-part scanline image, no compression, FLOAT channels
//        B, G, R (alphabetical, as the format requires), one line per block.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"
#include "host_scene.h"

namespace orion_amd {
namespace {

void le16(std::vector<unsigned char> *o, uint32_t v) {
  o->push_back((unsigned char)v); o->push_back((unsigned char)(v >> 8));
}
void le32(std::vector<unsigned char> *o, uint32_t v) {
  for (int k = 0; k < 4; ++k) o->push_back((unsigned char)(v >> (8 * k)));
}
void le64(std::vector<unsigned char> *o, uint64_t v) {
  for (int k = 0; k < 8; ++k) o->push_back((unsigned char)(v >> (8 * k)));
}
void lef(std::vector<unsigned char> *o, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  le32(o, u);
}
void str0(std::vector<unsigned char> *o, const char *s) { o->insert(o->end(), s, s + std::strlen(s) + 1); }

// color_FloatToChar (texture.cpp:70-73): clamp to [0, 1] then truncate f * 255
unsigned char to_byte(float f) {
  f = f < 0 ? 0 : (f > 1.0f ? 1.0f : f);
  if (f != f) return 0;  // NaN: x86's truncating conversion yields 0x80000000, low byte 0
  return (unsigned char)(int)(f * 255);
}

bool write_file(const std::string &path, const std::vector<unsigned char> &b, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { *err = "cannot write " + path; return false; }
  bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) *err = "cannot write " + path;
  return ok;
}

std::vector<unsigned char> bmp_bytes(int W, int H, const float *rgb) {
  const uint32_t bytes = 4u * (uint32_t)W * (uint32_t)H;
  std::vector<unsigned char> o;
  o.reserve(54 + bytes);
  o.push_back('B'); o.push_back('M');
  le32(&o, 54 + bytes);  // bfSize = bfOffBits + bytes
  le16(&o, 0); le16(&o, 0);
  le32(&o, 54);          // bfOffBits = sizeof(file header) + sizeof(info header) + 2
  le32(&o, 40);          // biSize
  le32(&o, (uint32_t)W); le32(&o, (uint32_t)H);
  le16(&o, 1); le16(&o, 32);
  for (int k = 0; k < 6; ++k) le32(&o, 0);  // compression, size image, ppm x/y, colours used/important
  size_t base = o.size();
  o.resize(base + bytes);
  for (int j = 0; j < H; ++j) {
    size_t off = base + (size_t)(H - j - 1) * W * 4;
    for (int i = 0; i < W; ++i) {
      const float *s = rgb + 3 * ((size_t)j * W + i);
      o[off++] = to_byte(s[2]);
      o[off++] = to_byte(s[1]);
      o[off++] = to_byte(s[0]);
      o[off++] = 255;
    }
  }
  return o;
}

std::vector<unsigned char> pfm_bytes(int W, int H, const float *rgb) {
  char head[64];
  int n = std::snprintf(head, sizeof(head), "PF\n%d %d\n-1.0\n", W, H);
  std::vector<unsigned char> o(head, head + n);
  for (int y = H - 1; y >= 0; --y)
    for (size_t k = 0; k < 3 * (size_t)W; ++k) lef(&o, rgb[3 * (size_t)y * W + k]);
  return o;
}

std::vector<unsigned char> exr_bytes(int W, int H, const float *rgb) {
  std::vector<unsigned char> o;
  le32(&o, 20000630);  // magic 76 2f 31 01
  le32(&o, 2);         // version 2, single-part scanline
  auto attr = [&](const char *name, const char *type, const std::vector<unsigned char> &v) {
    str0(&o, name);
    str0(&o, type);
    le32(&o, (uint32_t)v.size());
    o.insert(o.end(), v.begin(), v.end());
  };
  std::vector<unsigned char> ch;
  for (const char *c : {"B", "G", "R"}) {
    str0(&ch, c);
    le32(&ch, 2);  // FLOAT
    ch.push_back(0); ch.push_back(0); ch.push_back(0); ch.push_back(0);  // pLinear + reserved
    le32(&ch, 1); le32(&ch, 1);  // x / y sampling
  }
  ch.push_back(0);
  attr("channels", "chlist", ch);
  attr("compression", "compression", {0});
  std::vector<unsigned char> box;
  le32(&box, 0); le32(&box, 0); le32(&box, (uint32_t)(W - 1)); le32(&box, (uint32_t)(H - 1));
  attr("dataWindow", "box2i", box);
  attr("displayWindow", "box2i", box);
  attr("lineOrder", "lineOrder", {0});
  std::vector<unsigned char> one;
  lef(&one, 1.0f);
  attr("pixelAspectRatio", "float", one);
  std::vector<unsigned char> ctr;
  lef(&ctr, 0.f); lef(&ctr, 0.f);
  attr("screenWindowCenter", "v2f", ctr);
  attr("screenWindowWidth", "float", one);
  o.push_back(0);  // end of header
  const uint64_t line = 8 + 12 * (uint64_t)W;
  uint64_t first = o.size() + 8 * (uint64_t)H;
  for (int y = 0; y < H; ++y) le64(&o, first + line * y);
  for (int y = 0; y < H; ++y) {
    le32(&o, (uint32_t)y);
    le32(&o, 12u * (uint32_t)W);
    for (int c : {2, 1, 0})
      for (int x = 0; x < W; ++x) lef(&o, rgb[3 * ((size_t)y * W + x) + c]);
  }
  return o;
}

}  // namespace
}  // namespace orion_amd

using namespace orion_amd;

extern "C" {

int orion_film_average(int width, int height, const float *rgb_sum, const int *count, float *rgb) {
  if (width <= 0 || height <= 0 || !rgb_sum || !count || !rgb) { set_error("bad argument"); return ORION_ERR_ARG; }
  for (size_t i = 0; i < (size_t)width * height; ++i)
    for (int c = 0; c < 3; ++c) rgb[3 * i + c] = rgb_sum[3 * i + c] / count[i];  // Spectrum / int
  return ORION_OK;
}

int orion_film_write(const char *path, int format, int width, int height, const float *rgb) {
  if (!path || !rgb || width <= 0 || height <= 0 || (int64_t)width * height > (1 << 28)) {
    set_error("bad argument");
    return ORION_ERR_ARG;
  }
  std::vector<unsigned char> b;
  if (format == ORION_IMAGE_BMP) b = bmp_bytes(width, height, rgb);
  else if (format == ORION_IMAGE_PFM) b = pfm_bytes(width, height, rgb);
  else if (format == ORION_IMAGE_EXR) b = exr_bytes(width, height, rgb);
  else { set_error("unknown image format"); return ORION_ERR_ARG; }
  std::string err;
  if (!write_file(path, b, &err)) { set_error(err); return ORION_ERR_IO; }
  return ORION_OK;
}

}  // extern "C"
