// File formats of the reference's example pipeline, host side (no GPU):
//
//   load_unary  - the compressed ".c_unary" unaries of ProbImage
//                 (external/ProbImage/probimage.cpp:140-165 decompress,
//                 :64-78 boostToProb) turned into -log probabilities as
//                 load_unary does (src/file_storage.cpp:262-291);
//   load_image  - an uncompressed 24 / 32 bpp BMP as interleaved RGB, pixel
//                 (x, y) at (x + y W) * 3 (src/file_storage.cpp:197-219, which
//                 reads it with OpenCV; here a BMP reader of our own);
//   save_map    - the argmax labelling coloured with the MSRC palette
//                 (src/file_storage.cpp:332-380), written as PNG (stored
//                 deflate blocks, no compression library needed) or as a
//                 binary PPM when the path ends in ".ppm".
//
// Errors throw std::runtime_error.  Header-only; needs densecrf_ho_amd.hpp
// for MatrixXf.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "densecrf_ho_amd.hpp"

namespace dcrf_ho {

struct img_size {
  int width;
  int height;
};

namespace io_detail {

inline std::vector<unsigned char> read_all(const std::string& path) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("cannot open " + path);
  std::vector<unsigned char> buf;
  unsigned char tmp[1 << 16];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  std::fclose(fp);
  return buf;
}
inline uint32_t be32(const unsigned char* p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}
inline uint32_t le32(const unsigned char* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
inline uint16_t le16(const unsigned char* p) { return (uint16_t)(p[0] | p[1] << 8); }

// CRC-32 (PNG chunks) and Adler-32 (zlib stream), from their definitions
inline uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xffffffffu) {
  for (size_t i = 0; i < n; ++i) {
    c ^= p[i];
    for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
  }
  return c;
}
inline void put_be32(std::vector<unsigned char>& o, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}

}  // namespace io_detail

// The MSRC colour of each label (the dataset's colour code; label 21 and up:
// black), as the reference's save_map uses it (include/file_storage.hpp:15-38).
inline const unsigned char* msrc_palette() {
  static const unsigned char p[22 * 3] = {
      128, 0,   0,    //
      0,   128, 0,    //
      128, 128, 0,    //
      0,   0,   128,  //
      0,   128, 128,  //
      128, 128, 128,  //
      192, 0,   0,    //
      64,  128, 0,    //
      192, 128, 0,    //
      64,  0,   128,  //
      192, 0,   128,  //
      64,  128, 128,  //
      192, 128, 128,  //
      0,   64,  0,    //
      128, 64,  0,    //
      0,   192, 0,    //
      128, 64,  128,  //
      0,   192, 128,  //
      128, 192, 128,  //
      64,  64,  0,    //
      192, 64,  0,    //
      0,   0,   0};
  return p;
}

// ProbImage::decompress (probimage.cpp:140-165): big-endian header {width,
// height, depth, M, eps bits}, M x depth int32 keys, width x height uint32
// key ids; value = key * eps.  Then boostToProb (:64-78) and -log.
inline MatrixXf load_unary(const std::string& path, img_size& size, int max_label = -1) {
  using namespace io_detail;
  const std::vector<unsigned char> b = read_all(path);
  if (b.size() < 20) throw std::runtime_error("load_unary: truncated header in " + path);
  const uint32_t W = be32(&b[0]), H = be32(&b[4]), D = be32(&b[8]), M = be32(&b[12]);
  const uint32_t eb = be32(&b[16]);
  float eps;
  std::memcpy(&eps, &eb, 4);
  const size_t need = 20 + (size_t)M * D * 4 + (size_t)W * H * 4;
  if (W == 0 || H == 0 || D == 0 || b.size() < need) throw std::runtime_error("load_unary: truncated file " + path);
  const unsigned char* keys = &b[20];
  const unsigned char* ids = keys + (size_t)M * D * 4;
  std::vector<float> p((size_t)W * H * D);
  for (size_t i = 0; i < (size_t)W * H; ++i) {
    const uint32_t id = be32(ids + 4 * i);
    if (id >= M) throw std::runtime_error("load_unary: key id out of range in " + path);
    for (uint32_t j = 0; j < D; ++j) p[i * D + j] = (float)(int32_t)be32(keys + 4 * ((size_t)id * D + j)) * eps;
  }
  for (size_t i = 0; i < (size_t)W * H; ++i) {  // boostToProb: exp and 1/sum as the C library's double routines
    float* dp = &p[i * D];
    float mx = dp[0];
    for (uint32_t j = 1; j < D; ++j)
      if (mx < dp[j]) mx = dp[j];
    float nm = 0;
    for (uint32_t j = 0; j < D; ++j) nm += (dp[j] = (float)std::exp((double)(dp[j] - mx)));
    nm = (float)(1.0 / nm);
    for (uint32_t j = 0; j < D; ++j) dp[j] *= nm;
  }
  if (size.width <= 0 && size.height <= 0) size = img_size{(int)W, (int)H};
  if (size.width != (int)W || size.height != (int)H) throw std::runtime_error("load_unary: size differs from the file");
  if (max_label <= 0) max_label = (int)D;
  if (max_label > (int)D) throw std::runtime_error("load_unary: more labels than the file holds");
  MatrixXf u(max_label, (long)W * H);
  for (size_t i = 0; i < (size_t)W * H; ++i)
    for (int k = 0; k < max_label; ++k) u(k, (long)i) = (float)-std::log((double)p[i * D + k]);
  return u;
}

// Uncompressed BMP (BITMAPINFOHEADER or later, 24 or 32 bits per pixel,
// bottom-up or top-down rows) as RGB, pixel (x, y) at (x + y W) * 3.  The
// caller frees the result with delete[], as with the reference's load_image.
inline unsigned char* load_image(const std::string& path, img_size& size) {
  using namespace io_detail;
  const std::vector<unsigned char> b = read_all(path);
  if (b.size() < 54 || b[0] != 'B' || b[1] != 'M') throw std::runtime_error("load_image: not a BMP file: " + path);
  const uint32_t off = le32(&b[10]);
  const int32_t w = (int32_t)le32(&b[18]), h0 = (int32_t)le32(&b[22]);
  const uint16_t bpp = le16(&b[28]);
  const uint32_t comp = le32(&b[30]);
  if ((bpp != 24 && bpp != 32) || !(comp == 0 || (comp == 3 && bpp == 32)))
    throw std::runtime_error("load_image: only uncompressed 24 / 32 bpp BMP is supported: " + path);
  const bool bottom_up = h0 > 0;
  const int h = bottom_up ? h0 : -h0;
  const size_t stride = ((size_t)w * bpp / 8 + 3) / 4 * 4;
  if (w <= 0 || h <= 0 || b.size() < off + stride * h) throw std::runtime_error("load_image: truncated file " + path);
  if (size.width == -1) size.width = w;
  if (size.height == -1) size.height = h;
  if (size.width != w || size.height != h) throw std::runtime_error("load_image: size differs from the unaries");
  unsigned char* rgb = new unsigned char[(size_t)w * h * 3];
  for (int y = 0; y < h; ++y) {
    const unsigned char* row = &b[off + stride * (bottom_up ? h - 1 - y : y)];
    for (int x = 0; x < w; ++x) {
      const unsigned char* px = row + (size_t)x * (bpp / 8);  // B, G, R (, A)
      unsigned char* o = rgb + ((size_t)x + (size_t)y * w) * 3;
      o[0] = px[2];
      o[1] = px[1];
      o[2] = px[0];
    }
  }
  return rgb;
}

// argmax labelling (first maximum) coloured with the MSRC palette
inline void save_map(const MatrixXf& Q, const img_size& size, const std::string& path,
                     const std::string& dataset = "MSRC") {
  using namespace io_detail;
  if (dataset != "MSRC") throw std::runtime_error("save_map: only the MSRC palette is provided");
  const int W = size.width, H = size.height;
  if ((long)W * H != Q.cols()) throw std::runtime_error("save_map: size differs from the field");
  const unsigned char* pal = msrc_palette();
  std::vector<unsigned char> rgb((size_t)W * H * 3);
  for (long i = 0; i < Q.cols(); ++i) {
    int lbl = 0;
    for (long k = 1; k < Q.rows(); ++k)
      if (Q(k, i) > Q(lbl, i)) lbl = (int)k;
    const int c = lbl < 22 ? lbl : 21;
    for (int ch = 0; ch < 3; ++ch) rgb[(size_t)i * 3 + ch] = pal[3 * c + ch];
  }
  std::vector<unsigned char> out;
  const bool ppm = path.size() >= 4 && path.compare(path.size() - 4, 4, ".ppm") == 0;
  if (ppm) {
    const std::string hdr = "P6\n" + std::to_string(W) + " " + std::to_string(H) + "\n255\n";
    out.assign(hdr.begin(), hdr.end());
    out.insert(out.end(), rgb.begin(), rgb.end());
  } else {
    // PNG: IHDR (8-bit RGB), one IDAT holding a zlib stream of stored
    // (uncompressed) deflate blocks over the filtered rows (filter 0), IEND
    std::vector<unsigned char> raw;
    raw.reserve((size_t)H * (W * 3 + 1));
    for (int y = 0; y < H; ++y) {
      raw.push_back(0);
      raw.insert(raw.end(), rgb.begin() + (size_t)y * W * 3, rgb.begin() + (size_t)(y + 1) * W * 3);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    uint32_t a = 1, bsum = 0;
    for (unsigned char v : raw) {
      a = (a + v) % 65521u;
      bsum = (bsum + a) % 65521u;
    }
    for (size_t pos = 0; pos < raw.size(); pos += 65535) {
      const size_t n = std::min<size_t>(65535, raw.size() - pos);
      z.push_back(pos + n >= raw.size() ? 1 : 0);
      z.push_back((unsigned char)(n & 0xff));
      z.push_back((unsigned char)(n >> 8));
      z.push_back((unsigned char)(~n & 0xff));
      z.push_back((unsigned char)((~n >> 8) & 0xff));
      z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
    }
    put_be32(z, (bsum << 16) | a);
    auto chunk = [&](const char* type, const std::vector<unsigned char>& data) {
      put_be32(out, (uint32_t)data.size());
      std::vector<unsigned char> td(type, type + 4);
      td.insert(td.end(), data.begin(), data.end());
      out.insert(out.end(), td.begin(), td.end());
      put_be32(out, crc32(td.data(), td.size()) ^ 0xffffffffu);
    };
    const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out.assign(sig, sig + 8);
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
  }
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw std::runtime_error("save_map: cannot write " + path);
  const size_t wr = std::fwrite(out.data(), 1, out.size(), fp);
  std::fclose(fp);
  if (wr != out.size()) throw std::runtime_error("save_map: short write to " + path);
}

}  // namespace dcrf_ho
