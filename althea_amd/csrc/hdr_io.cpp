// RGBE (.hdr) image files: the format of the reference's environment maps
// (Content/HDRI_Skybox/*.hdr) and of its prefiltered-map cache
// (Content/PrecomputedMaps/<env>/*.hdr), which it reads and writes through
// stb_image (Src/Utilities.cpp:216-255).  The Radiance format as published:
// a text header ("#?RADIANCE" or "#?RGBE", FORMAT=32-bit_rle_rgbe, a blank
// line, "-Y height +X width"), then scanlines of 4-byte RGBE pixels, either
// flat or run-length coded per channel (scanline tag 2, 2, width >> 8,
// width & 255; a count byte > 128 repeats the next byte count - 128 times,
// otherwise count literal bytes follow).
//
// Conversion as stb does it: value = byte * 2^(E - 136) for E != 0, else 0;
// writing takes m, e = frexp(max(r, g, b)) and stores byte = floor(c * m * 256
// / max) and E = e + 128 (0 for a maximum below 1e-32).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace althea {
namespace {

struct File {
  std::FILE* f = nullptr;
  explicit File(std::FILE* h) : f(h) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool read_line(std::FILE* f, std::string& line) {
  line.clear();
  for (int c; (c = std::fgetc(f)) != EOF;) {
    if (c == '\n') return true;
    line.push_back(static_cast<char>(c));
    if (line.size() > 4096) return false;
  }
  return !line.empty();
}

// Parses the header; leaves `f` at the first pixel byte.
althea_status read_header(std::FILE* f, const char* path, int& w, int& h) {
  std::string line;
  if (!read_line(f, line) || (line.rfind("#?RADIANCE", 0) != 0 && line.rfind("#?RGBE", 0) != 0)) {
    set_error("hdr: %s is not a Radiance RGBE file", path);
    return ALTHEA_ERROR_IO;
  }
  for (;;) {
    if (!read_line(f, line)) {
      set_error("hdr: %s: truncated header", path);
      return ALTHEA_ERROR_IO;
    }
    if (line.empty()) break;
    if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe") {
      set_error("hdr: %s: unsupported %s", path, line.c_str());
      return ALTHEA_ERROR_UNSUPPORTED;
    }
  }
  if (!read_line(f, line) || std::sscanf(line.c_str(), "-Y %d +X %d", &h, &w) != 2 || w < 1 || h < 1 ||
      w > 65536 || h > 65536) {
    set_error("hdr: %s: unsupported resolution line (only -Y h +X w)", path);
    return ALTHEA_ERROR_UNSUPPORTED;
  }
  return ALTHEA_OK;
}

void rgbe_to_float(const unsigned char* p, float* out) {
  if (p[3] != 0) {
    const float f = std::ldexp(1.0f, static_cast<int>(p[3]) - 136);
    out[0] = p[0] * f;
    out[1] = p[1] * f;
    out[2] = p[2] * f;
  } else {
    out[0] = out[1] = out[2] = 0.0f;
  }
  out[3] = 1.0f;
}

void float_to_rgbe(const float* c, unsigned char* p) {
  const float mx = std::fmax(c[0], std::fmax(c[1], c[2]));
  if (!(mx >= 1e-32f)) {
    p[0] = p[1] = p[2] = p[3] = 0;
    return;
  }
  int e;
  const float scale = std::frexp(mx, &e) * 256.0f / mx;
  p[0] = static_cast<unsigned char>(std::fmax(c[0], 0.0f) * scale);
  p[1] = static_cast<unsigned char>(std::fmax(c[1], 0.0f) * scale);
  p[2] = static_cast<unsigned char>(std::fmax(c[2], 0.0f) * scale);
  p[3] = static_cast<unsigned char>(e + 128);
}

// One run-length coded channel of a scanline: runs of 3+ equal bytes become
// (128 + n, byte), n <= 127; the rest go out as literal blocks of <= 128.
void write_channel(std::vector<unsigned char>& out, const unsigned char* v, int n) {
  int i = 0;
  while (i < n) {
    int run = 1;
    while (i + run < n && run < 127 && v[i + run] == v[i]) ++run;
    if (run >= 3) {
      out.push_back(static_cast<unsigned char>(128 + run));
      out.push_back(v[i]);
      i += run;
      continue;
    }
    int j = i;  // literal block up to the next run of 3 or 128 bytes
    while (j < n && j - i < 128) {
      if (j + 2 < n && v[j] == v[j + 1] && v[j] == v[j + 2]) break;
      ++j;
    }
    out.push_back(static_cast<unsigned char>(j - i));
    out.insert(out.end(), v + i, v + j);
    i = j;
  }
}

}  // namespace
}  // namespace althea

using namespace althea;

extern "C" {

althea_status althea_hdr_info(const char* path, int* width, int* height) {
  ALTHEA_REQUIRE(path && width && height, "hdr_info: null argument");
  File f(std::fopen(path, "rb"));
  if (!f.f) {
    set_error("hdr: %s does not exist", path);
    return ALTHEA_ERROR_NOT_FOUND;
  }
  return read_header(f.f, path, *width, *height);
}

althea_status althea_hdr_read(const char* path, float* rgba, size_t capacity) {
  ALTHEA_REQUIRE(path && rgba, "hdr_read: null argument");
  File f(std::fopen(path, "rb"));
  if (!f.f) {
    set_error("hdr: %s does not exist", path);
    return ALTHEA_ERROR_NOT_FOUND;
  }
  int w = 0, h = 0;
  althea_status st = read_header(f.f, path, w, h);
  if (st != ALTHEA_OK) return st;
  ALTHEA_REQUIRE(capacity >= static_cast<size_t>(w) * h * 4, "hdr_read: %s needs %zu floats, capacity %zu", path,
                 static_cast<size_t>(w) * h * 4, capacity);
  std::vector<unsigned char> line(static_cast<size_t>(w) * 4);
  auto truncated = [&]() {
    set_error("hdr: %s: truncated or corrupt pixel data", path);
    return ALTHEA_ERROR_IO;
  };
  for (int y = 0; y < h; ++y) {
    float* row = rgba + static_cast<size_t>(y) * w * 4;
    unsigned char tag[4];
    if (std::fread(tag, 1, 4, f.f) != 4) return truncated();
    const bool rle = w >= 8 && w < 32768 && tag[0] == 2 && tag[1] == 2 && (tag[2] & 0x80) == 0;
    if (!rle) {
      // Flat pixels; the four bytes read are the first pixel.
      std::memcpy(line.data(), tag, 4);
      if (w > 1 && std::fread(line.data() + 4, 1, static_cast<size_t>(w - 1) * 4, f.f) != static_cast<size_t>(w - 1) * 4)
        return truncated();
    } else {
      if (((tag[2] << 8) | tag[3]) != w) return truncated();
      for (int c = 0; c < 4; ++c) {
        int x = 0;
        while (x < w) {
          int count = std::fgetc(f.f);
          if (count == EOF || count == 0) return truncated();
          if (count > 128) {
            count -= 128;
            const int v = std::fgetc(f.f);
            if (v == EOF || x + count > w) return truncated();
            for (int k = 0; k < count; ++k) line[(x + k) * 4 + c] = static_cast<unsigned char>(v);
          } else {
            if (x + count > w) return truncated();
            for (int k = 0; k < count; ++k) {
              const int v = std::fgetc(f.f);
              if (v == EOF) return truncated();
              line[(x + k) * 4 + c] = static_cast<unsigned char>(v);
            }
          }
          x += count;
        }
      }
    }
    for (int x = 0; x < w; ++x) rgbe_to_float(line.data() + 4 * x, row + 4 * x);
  }
  return ALTHEA_OK;
}

althea_status althea_hdr_write(const char* path, const float* rgba, int width, int height) {
  ALTHEA_REQUIRE(path && rgba && width >= 1 && height >= 1 && width <= 65536 && height <= 65536,
                 "hdr_write: invalid argument");
  File f(std::fopen(path, "wb"));
  if (!f.f) {
    set_error("hdr: cannot create %s", path);
    return ALTHEA_ERROR_IO;
  }
  std::string head = "#?RADIANCE\n# Written by althea\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(height) +
                     " +X " + std::to_string(width) + "\n";
  std::vector<unsigned char> out(head.begin(), head.end());
  std::vector<unsigned char> px(static_cast<size_t>(width) * 4), chan(width);
  const bool rle = width >= 8 && width < 32768;
  for (int y = 0; y < height; ++y) {
    const float* row = rgba + static_cast<size_t>(y) * width * 4;
    for (int x = 0; x < width; ++x) float_to_rgbe(row + 4 * x, px.data() + 4 * x);
    if (!rle) {
      out.insert(out.end(), px.begin(), px.end());
      continue;
    }
    out.push_back(2);
    out.push_back(2);
    out.push_back(static_cast<unsigned char>(width >> 8));
    out.push_back(static_cast<unsigned char>(width & 255));
    for (int c = 0; c < 4; ++c) {
      for (int x = 0; x < width; ++x) chan[x] = px[4 * x + c];
      write_channel(out, chan.data(), width);
    }
  }
  if (std::fwrite(out.data(), 1, out.size(), f.f) != out.size()) {
    set_error("hdr: short write to %s", path);
    return ALTHEA_ERROR_IO;
  }
  return ALTHEA_OK;
}

}  // extern "C"
