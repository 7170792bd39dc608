// Oracle (test infrastructure only): the reference's film output stage on
// the CPU, the checker for the library's ar_film_to_rgb8 / ar_film_save_exr.
//   orc_film_rgb8   Film::to_rgb_image (agent-ray src/film.rs:283-303) with
//                   tonemap_pixel (film.rs:153-180); powf is the host libm's,
//                   as Rust's f32::powf is.
//   orc_f32_to_f16  half::f16::from_f32, the conversion Film::save_exr uses
//                   (film.rs:323-325; `half` is a dependency of the `exr`
//                   crate 1.74, not vendored): IEEE binary16, round to
//                   nearest even, NaN -> quiet NaN keeping sign and the top
//                   9 payload bits.
//   orc_exr_blocks  the bytes OpenEXR's ZIP compression deflates for each
//                   block of 16 scan lines of a B, G, R half image (the
//                   `exr` crate's ZIP16 encoding, film.rs:330-341): per line
//                   the B, G, R halves little endian; bytes split even-first;
//                   each byte after the first replaced by (b[i] - b[i-1] +
//                   128) mod 256.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// glam Vec3A::clamp(ZERO, ONE): max, then min (a NaN lane becomes 0).
float clamp01(float v) { return std::fmin(std::fmax(v, 0.0f), 1.0f); }

void tonemap_pixel(float& r, float& g, float& b, uint32_t tm, float white_point) {
  if (tm == 0) {  // ToneMapper::Clamp (film.rs:156)
    r = clamp01(r), g = clamp01(g), b = clamp01(b);
    return;
  }
  float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
  float scale;
  if (tm == 1) {  // ToneMapper::Reinhard (film.rs:157-165)
    if (lum > 1e-6f)
      scale = (lum / (1.0f + lum)) / lum;
    else
      scale = 1.0f / (1.0f + lum);
  } else {  // ToneMapper::ReinhardExtended (film.rs:166-178)
    float lw2 = white_point * white_point;
    float numerator = lum * (1.0f + lum / lw2);
    float mapped_lum = numerator / (1.0f + lum);
    scale = lum > 1e-6f ? mapped_lum / lum : 1.0f;
  }
  r = clamp01(r * scale), g = clamp01(g * scale), b = clamp01(b * scale);
}

// (x * 255.0 + 0.5).min(255.0) as u8 (film.rs:297-299): f32::min returns
// the non-NaN operand; `as u8` truncates and saturates, NaN -> 0.
uint8_t quantize(float x) {
  float v = std::fmin(x * 255.0f + 0.5f, 255.0f);
  if (!(v >= 0.0f)) return 0;
  return (uint8_t)v;
}

uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, exp = (x >> 23) & 0xffu, man = x & 0x7fffffu;
  if (exp == 0xffu) {
    if (man == 0) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | 0x7e00u | (man >> 13));
  }
  int e = (int)exp - 112;  // rebias 127 -> 15
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    // |f| = (man | 2^23) * 2^(exp - 150); as a multiple of 2^-24 that is
    // (man | 2^23) * 2^(e - 14): drop 14 - e bits, rounding to even.
    if (14 - e > 24) return (uint16_t)sign;
    uint64_t m = man | 0x800000u;
    int drop = 14 - e;
    uint64_t q = m >> drop, r = m - (q << drop), halfway = 1ull << (drop - 1);
    if (r > halfway || (r == halfway && (q & 1))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t q = ((uint32_t)e << 10) | (man >> 13), r = man & 0x1fffu;
  if (r > 0x1000u || (r == 0x1000u && (q & 1))) ++q;
  return (uint16_t)(sign | q);
}

}  // namespace

extern "C" {

void orc_film_rgb8(const float* film, uint64_t npix, uint32_t tm, float white_point, float gamma, float scale,
                   uint8_t* out) {
  float inv_gamma = 1.0f / gamma;
  for (uint64_t i = 0; i < npix; ++i) {
    float r = film[3 * i] * scale, g = film[3 * i + 1] * scale, b = film[3 * i + 2] * scale;
    tonemap_pixel(r, g, b, tm, white_point);
    out[3 * i] = quantize(std::pow(r, inv_gamma));
    out[3 * i + 1] = quantize(std::pow(g, inv_gamma));
    out[3 * i + 2] = quantize(std::pow(b, inv_gamma));
  }
}

void orc_f32_to_f16(const float* in, uint64_t n, uint16_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = f32_to_f16(in[i]);
}

// out: width*height*6 bytes, block b (16 lines) at b * 16 * 6 * width.
void orc_exr_blocks(const float* film, uint32_t width, uint32_t height, float scale, uint8_t* out) {
  const uint64_t line_bytes = 6ull * width;
  for (uint32_t y0 = 0; y0 < height; y0 += 16) {
    uint32_t lines = height - y0 < 16 ? height - y0 : 16;
    uint64_t n = lines * line_bytes;
    std::vector<uint8_t> raw(n);
    uint64_t k = 0;
    for (uint32_t y = y0; y < y0 + lines; ++y)
      for (int c = 2; c >= 0; --c)  // channels sorted by name: B, G, R
        for (uint32_t x = 0; x < width; ++x) {
          uint16_t h = f32_to_f16(film[3 * ((uint64_t)y * width + x) + c] * scale);
          raw[k++] = (uint8_t)(h & 0xff);
          raw[k++] = (uint8_t)(h >> 8);
        }
    uint8_t* dst = out + (uint64_t)y0 * line_bytes;
    uint64_t even = (n + 1) / 2;
    std::vector<uint8_t> t(n);
    for (uint64_t i = 0; i < n; ++i) t[(i & 1) ? even + i / 2 : i / 2] = raw[i];
    for (uint64_t i = 0; i < n; ++i) dst[i] = i == 0 ? t[0] : (uint8_t)((int)t[i] - (int)t[i - 1] + 128 + 256);
  }
}

}  // extern "C"
