// Film output stage: tone mapping, gamma and 8-bit encoding
// (Film::to_rgb_image, agent-ray src/film.rs:283-303, with tonemap_pixel,
// film.rs:153-180) and the half-float OpenEXR file (Film::save_exr,
// film.rs:310-346), computed on the device so that only 3 (8-bit) or 6
// (half) bytes per pixel cross PCIe instead of 12.
//
// The reference writes its EXR with the `exr` crate (1.74, not vendored):
// one scan-line part, channels B, G, R (sorted by name) as HALF, ZIP
// compression over blocks of 16 scan lines, increasing Y.  Per block the
// uncompressed bytes are, for each line, the line's B halves, then G, then R
// (little endian); OpenEXR's ZIP step splits them into even and odd bytes
// (even first), replaces each byte after the first by (b[i] - b[i-1] + 128)
// mod 256, and deflates the result with zlib; a block whose compressed size
// is not smaller than its raw size is stored raw.  The device produces the
// split-and-predicted bytes of every block; the host deflates the blocks
// on its threads and writes the header, the offset table and the chunks.
//
// f32 -> f16 is IEEE round to nearest even (the `half` crate's from_f32):
// overflow to infinity, gradual underflow to subnormals, NaN kept as a
// quiet NaN with its sign and top payload bits.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ar_internal.h"
#include "ar_kernels.h"

namespace ar {

constexpr uint32_t kExrLines = 16;  // scan lines per ZIP block

__host__ __device__ __forceinline__ uint16_t f32_to_f16(float f) {
  const uint32_t x = __builtin_bit_cast(uint32_t, f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t exp = (x >> 23) & 0xffu;
  uint32_t man = x & 0x7fffffu;
  if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? (0x0200u | (man >> 13)) : 0u));
  const int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    const uint32_t shift = (uint32_t)(14 - e);
    uint32_t h = man >> shift;
    const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
  const uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // a carry rounds up into the exponent
  return (uint16_t)h;
}

// glam's Vec3A::clamp(ZERO, ONE) = max then min; a NaN lane becomes 0.
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// tonemap_pixel (film.rs:153-180), in the reference's operation order.
__device__ __forceinline__ void tonemap(float& r, float& g, float& b, uint32_t tm, float white_point) {
  if (tm != AR_TONEMAP_CLAMP) {
    const float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
    float scale;
    if (tm == AR_TONEMAP_REINHARD) {
      scale = lum > 1e-6f ? (lum / (1.0f + lum)) / lum : 1.0f / (1.0f + lum);
    } else {
      const float lw2 = white_point * white_point;
      const float mapped = lum * (1.0f + lum / lw2) / (1.0f + lum);
      scale = lum > 1e-6f ? mapped / lum : 1.0f;
    }
    r *= scale;
    g *= scale;
    b *= scale;
  }
  r = clamp01(r);
  g = clamp01(g);
  b = clamp01(b);
}

// (v * 255 + 0.5).min(255) as u8: Rust's min keeps the non-NaN operand and
// its float -> u8 cast saturates (negative and NaN give 0).
__device__ __forceinline__ uint8_t to_u8(float encoded) {
  const float v = fminf(encoded * 255.0f + 0.5f, 255.0f);
  return v >= 0.0f ? (uint8_t)(uint32_t)v : (uint8_t)0;
}

__device__ __forceinline__ uint32_t encode_channel(float v, float inv_gamma) { return to_u8(powf(v, inv_gamma)); }

// One pixel per thread (any alignment).
__global__ void k_film_rgb8(const float* __restrict__ film, uint8_t* __restrict__ out, uint32_t npix, uint32_t tm,
                            float white_point, float inv_gamma, float scale) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
    float r = film[3 * (size_t)i] * scale, g = film[3 * (size_t)i + 1] * scale, b = film[3 * (size_t)i + 2] * scale;
    tonemap(r, g, b, tm, white_point);
    out[3 * (size_t)i] = (uint8_t)encode_channel(r, inv_gamma);
    out[3 * (size_t)i + 1] = (uint8_t)encode_channel(g, inv_gamma);
    out[3 * (size_t)i + 2] = (uint8_t)encode_channel(b, inv_gamma);
  }
}

// Four pixels per thread: three 16-byte loads (12 floats) and three 4-byte
// stores (12 bytes); film 16-byte and out 4-byte aligned.  Pixels past the
// last whole group of four go through k_film_rgb8.
__global__ void k_film_rgb8_x4(const float4* __restrict__ film, uint32_t* __restrict__ out, uint32_t groups,
                               uint32_t tm, float white_point, float inv_gamma, float scale) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= groups) return;
  const float4 a = film[3 * (size_t)q], b = film[3 * (size_t)q + 1], c = film[3 * (size_t)q + 2];
  const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
  uint32_t w[3] = {0u, 0u, 0u};
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    float r = v[3 * p] * scale, g = v[3 * p + 1] * scale, bl = v[3 * p + 2] * scale;
    tonemap(r, g, bl, tm, white_point);
    const uint32_t e[3] = {encode_channel(r, inv_gamma), encode_channel(g, inv_gamma), encode_channel(bl, inv_gamma)};
#pragma unroll
    for (int k = 0; k < 3; ++k) w[(3 * p + k) >> 2] |= e[k] << (8 * ((3 * p + k) & 3));
  }
  out[3 * (size_t)q] = w[0];
  out[3 * (size_t)q + 1] = w[1];
  out[3 * (size_t)q + 2] = w[2];
}

// Byte k of a block's raw data (lines y0.., each line B, G, R halves).
__device__ __forceinline__ uint32_t exr_raw_byte(const float* __restrict__ film, uint32_t w, uint32_t y0, float scale,
                                                 uint32_t k) {
  const uint32_t line_bytes = 6u * w;
  const uint32_t line = k / line_bytes, rem = k - line * line_bytes;
  const uint32_t ch = rem / (2u * w), x = (rem - ch * 2u * w) >> 1;
  const float v = film[3 * ((size_t)(y0 + line) * w + x) + (2u - ch)] * scale;  // B, G, R planes
  const uint16_t h = f32_to_f16(v);
  return (k & 1u) ? (uint32_t)(h >> 8) : (uint32_t)(h & 0xffu);
}

// Split-and-predicted bytes of every block, blocks back to back (block b
// starts at byte b * 16 * 6w; the last block may hold fewer lines).
__global__ void k_film_exr_blocks(const float* __restrict__ film, uint8_t* __restrict__ out, uint32_t w, uint32_t h,
                                  float scale, uint64_t total) {
  const uint64_t block_bytes = (uint64_t)kExrLines * 6u * w;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t blk = (uint32_t)(i / block_bytes);
    const uint32_t j = (uint32_t)(i - (uint64_t)blk * block_bytes);
    const uint32_t y0 = blk * kExrLines;
    const uint32_t lines = h - y0 < kExrLines ? h - y0 : kExrLines;
    const uint32_t n = lines * 6u * w;  // raw bytes of this block
    const uint32_t even = (n + 1u) / 2u;
    // split: position j holds raw byte 2j (first half) or 2(j - even) + 1
    auto split = [&](uint32_t p) { return p < even ? 2u * p : 2u * (p - even) + 1u; };
    uint32_t v = exr_raw_byte(film, w, y0, scale, split(j));
    if (j > 0) v = (v - exr_raw_byte(film, w, y0, scale, split(j - 1)) + 128u) & 0xffu;
    out[i] = (uint8_t)v;
  }
}

static uint32_t film_grid(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(g, 1), (uint64_t)device_cus() * 16);
}

// ---- EXR container (OpenEXR 2.0 file layout, single scan-line part) -------

struct ByteWriter {
  std::vector<uint8_t> b;
  void u8(uint8_t v) { b.push_back(v); }
  void i32(int32_t v) { for (int k = 0; k < 4; ++k) b.push_back((uint8_t)((uint32_t)v >> (8 * k))); }
  void u64(uint64_t v) { for (int k = 0; k < 8; ++k) b.push_back((uint8_t)(v >> (8 * k))); }
  void f32(float v) { i32((int32_t)__builtin_bit_cast(uint32_t, v)); }
  void str(const char* s) { b.insert(b.end(), s, s + std::strlen(s) + 1); }
  void bytes(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
  void attr(const char* name, const char* type, const ByteWriter& v) {
    str(name);
    str(type);
    i32((int32_t)v.b.size());
    bytes(v.b.data(), v.b.size());
  }
};

static std::vector<uint8_t> exr_header(uint32_t w, uint32_t h) {
  ByteWriter f, v;
  f.i32(20000630);  // magic
  f.i32(2);         // version 2, single-part scan lines, short names
  for (const char* c : {"B", "G", "R"}) {
    v.str(c);
    v.i32(1);  // HALF
    v.i32(0);  // pLinear 0 + 3 reserved bytes
    v.i32(1);  // x sampling
    v.i32(1);  // y sampling
  }
  v.u8(0);
  f.attr("channels", "chlist", v);
  v.b.clear();
  v.u8(3);  // ZIP_COMPRESSION (16 scan lines per block)
  f.attr("compression", "compression", v);
  v.b.clear();
  for (int32_t k : {0, 0, (int32_t)w - 1, (int32_t)h - 1}) v.i32(k);
  f.attr("dataWindow", "box2i", v);
  f.attr("displayWindow", "box2i", v);
  v.b.clear();
  v.u8(0);  // INCREASING_Y
  f.attr("lineOrder", "lineOrder", v);
  v.b.clear();
  v.f32(1.0f);
  f.attr("pixelAspectRatio", "float", v);
  v.b.clear();
  v.f32(0.0f);
  v.f32(0.0f);
  f.attr("screenWindowCenter", "v2f", v);
  v.b.clear();
  v.f32(1.0f);
  f.attr("screenWindowWidth", "float", v);
  v.b.clear();
  v.bytes("beauty", 6);  // LayerAttributes::named("beauty") (film.rs:333)
  f.attr("name", "string", v);
  f.u8(0);  // end of header
  return f.b;
}

}  // namespace ar

using namespace ar;

extern "C" ar_status ar_film_to_rgb8(const float* film, uint32_t width, uint32_t height, uint32_t tonemapper,
                                     float white_point, float gamma, float scale, uint8_t* out, uint32_t flags,
                                     void* stream) {
  if (!film || !out) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: NULL argument");
  if (width == 0 || height == 0) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: empty film");
  if ((uint64_t)width * height >= (1ull << 31)) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: film too large");
  if (tonemapper > AR_TONEMAP_REINHARD_EXTENDED) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: unknown tone mapper");
  if (!(gamma > 0.0f) || !std::isfinite(gamma)) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: gamma must be > 0");
  if (!std::isfinite(scale)) return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: scale must be finite");
  if (tonemapper == AR_TONEMAP_REINHARD_EXTENDED && !(white_point > 0.0f))
    return fail(AR_ERR_INVALID_ARG, "ar_film_to_rgb8: white_point must be > 0");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t npix = width * height;
  const size_t in_bytes = (size_t)npix * 3 * sizeof(float), out_bytes = (size_t)npix * 3;
  const bool host = flags & AR_HOST_PTRS;
  DevBuf buf;
  const float* din = film;
  uint8_t* dout = out;
  if (host) {
    AR_HIP(buf.alloc(in_bytes + out_bytes));
    AR_HIP(hipMemcpyAsync(buf.p, film, in_bytes, hipMemcpyHostToDevice, st));
    din = buf.as<float>();
    dout = buf.as<uint8_t>() + in_bytes;
  }
  const float inv_gamma = 1.0f / gamma;
  const bool aligned = ((uintptr_t)din % 16 == 0) && ((uintptr_t)dout % 4 == 0);
  const uint32_t groups = aligned ? npix / 4 : 0, done = 4 * groups;
  if (groups)
    hipLaunchKernelGGL(k_film_rgb8_x4, dim3((groups + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<const float4*>(din), reinterpret_cast<uint32_t*>(dout), groups, tonemapper,
                       white_point, inv_gamma, scale);
  if (done < npix)
    hipLaunchKernelGGL(k_film_rgb8, dim3(film_grid(npix - done)), dim3(256), 0, st, din + 3 * (size_t)done,
                       dout + 3 * (size_t)done, npix - done, tonemapper, white_point, inv_gamma, scale);
  AR_HIP(hipGetLastError());
  if (host) AR_HIP(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, st));
  if (host || !(flags & AR_ASYNC)) AR_HIP(hipStreamSynchronize(st));
  return AR_OK;
}

extern "C" ar_status ar_film_save_exr(const float* film, uint32_t width, uint32_t height, float scale,
                                      const char* path, uint32_t flags, void* stream) try {
  if (!film || !path) return fail(AR_ERR_INVALID_ARG, "ar_film_save_exr: NULL argument");
  if (width == 0 || height == 0) return fail(AR_ERR_INVALID_ARG, "ar_film_save_exr: empty film");
  if (width > (1u << 24) || height > (1u << 24) || (uint64_t)width * height >= (1ull << 31))
    return fail(AR_ERR_INVALID_ARG, "ar_film_save_exr: film too large");
  if (!std::isfinite(scale)) return fail(AR_ERR_INVALID_ARG, "ar_film_save_exr: scale must be finite");
  hipStream_t st = (hipStream_t)stream;
  const uint64_t npix = (uint64_t)width * height;
  const size_t in_bytes = npix * 3 * sizeof(float);
  const uint64_t raw_total = npix * 6;  // split-and-predicted bytes of all blocks
  const uint32_t nblk = (height + kExrLines - 1) / kExrLines;
  const uint64_t block_bytes = (uint64_t)kExrLines * 6u * width;
  DevBuf in, dev;
  const float* din = film;
  if (flags & AR_HOST_PTRS) {
    AR_HIP(in.alloc(in_bytes));
    AR_HIP(hipMemcpyAsync(in.p, film, in_bytes, hipMemcpyHostToDevice, st));
    din = in.as<float>();
  }
  AR_HIP(dev.alloc(raw_total));
  void* pinned = nullptr;
  AR_HIP(hipHostMalloc(&pinned, raw_total, hipHostMallocDefault));
  struct PinnedFree {
    void* p;
    ~PinnedFree() { (void)hipHostFree(p); }
  } pf{pinned};
  hipLaunchKernelGGL(k_film_exr_blocks, dim3(film_grid(raw_total)), dim3(256), 0, st, din, dev.as<uint8_t>(), width,
                     height, scale, raw_total);
  AR_HIP(hipGetLastError());
  AR_HIP(hipMemcpyAsync(pinned, dev.p, raw_total, hipMemcpyDeviceToHost, st));
  AR_HIP(hipStreamSynchronize(st));
  // deflate the blocks on the host threads
  std::vector<std::vector<uint8_t>> packed(nblk);
  std::atomic<uint32_t> next{0};
  std::atomic<bool> zfail{false};
  auto work = [&]() {
    for (uint32_t b; (b = next.fetch_add(1)) < nblk;) {
      const uint32_t lines = std::min(kExrLines, height - b * kExrLines);
      const uLong n = (uLong)lines * 6u * width;
      const uint8_t* src = (const uint8_t*)pinned + (uint64_t)b * block_bytes;
      uLongf cap = compressBound(n);
      std::vector<uint8_t> z(cap);
      if (compress2(z.data(), &cap, src, n, 4) != Z_OK) {  // level 4: speed over size
        zfail = true;
        return;
      }
      if (cap < n) {
        z.resize(cap);
        packed[b] = std::move(z);
      } else {
        // not smaller: the block is stored raw (un-split, un-predicted)
        std::vector<uint8_t> raw(n);
        const uint32_t even = (uint32_t)((n + 1) / 2);
        uint8_t prev = 0;
        std::vector<uint8_t> t(n);
        for (uLong i = 0; i < n; ++i) {
          prev = i == 0 ? src[0] : (uint8_t)(src[i] + prev - 128);
          t[i] = prev;
        }
        for (uLong i = 0; i < n; ++i) raw[i] = (i & 1) ? t[even + i / 2] : t[i / 2];
        packed[b] = std::move(raw);
      }
    }
  };
  const unsigned nth = std::max(1u, std::min({std::thread::hardware_concurrency(), 16u, nblk}));
  std::vector<std::thread> pool;
  for (unsigned k = 1; k < nth; ++k) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (zfail) return fail(AR_ERR_DEVICE, "ar_film_save_exr: zlib compression failed");
  std::vector<uint8_t> hdr = exr_header(width, height);
  std::vector<uint64_t> offsets(nblk);
  uint64_t pos = hdr.size() + 8ull * nblk;
  for (uint32_t b = 0; b < nblk; ++b) {
    offsets[b] = pos;
    pos += 8 + packed[b].size();
  }
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return fail(AR_ERR_INVALID_ARG, std::string("ar_film_save_exr: cannot open ") + path);
  bool ok = std::fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
  ByteWriter tbl;
  for (uint64_t o : offsets) tbl.u64(o);
  ok = ok && std::fwrite(tbl.b.data(), 1, tbl.b.size(), fp) == tbl.b.size();
  for (uint32_t b = 0; b < nblk && ok; ++b) {
    ByteWriter ch;
    ch.i32((int32_t)(b * kExrLines));
    ch.i32((int32_t)packed[b].size());
    ok = std::fwrite(ch.b.data(), 1, 8, fp) == 8 &&
         std::fwrite(packed[b].data(), 1, packed[b].size(), fp) == packed[b].size();
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) return fail(AR_ERR_INVALID_ARG, std::string("ar_film_save_exr: write failed: ") + path);
  return AR_OK;
} catch (...) {
  return ar::fail_exception("ar_film_save_exr");
}
