// Oracle (test infrastructure only): gradient-domain path tracer restated
// from the reference's src/integrators/gradient_domain_pt.rs.
#include <chrono>

#include "orc_render.h"

namespace orc {

// ReplaySampler (gradient_domain_pt.rs:91-161): records every value the
// base path draws; replays the tape for the offset paths and, beyond the
// tape, continues the inner per-pixel stream.
struct ReplaySampler : Sampler {
  Sampler* inner;
  std::vector<float> tape;
  size_t cursor = 0;
  bool recording = true;
  explicit ReplaySampler(Sampler* in) : inner(in) { tape.reserve(128); }
  void reset_for_new_sample() {
    tape.clear();
    cursor = 0;
    recording = true;
  }
  void start_replay() {
    recording = false;
    cursor = 0;
  }
  float next_1d() override {
    if (recording) {
      float v = inner->next_1d();
      tape.push_back(v);
      return v;
    }
    float v = cursor < tape.size() ? tape[cursor] : inner->next_1d();
    cursor++;
    return v;
  }
};

// trace_with_gradients (gradient_domain_pt.rs:266-342) for one sample.
// Counter mode: offsets use the base pixel's (pixel, sample) key from
// dimension 2 on, which is the tape replay without a tape.
static void trace_with_gradients(const Scene& sc, const Camera& cam, Sampler& base_smp, ReplaySampler* rep,
                                 uint64_t seed, uint32_t px, uint32_t py, uint32_t si, uint32_t w, uint32_t h,
                                 const PtConfig& cfg, V3* base, V3* dx, V3* dy, TraceCounters* c) {
  float jx, jy;
  V3 b;
  if (rep) {
    rep->reset_for_new_sample();
    jx = rep->next_1d();
    jy = rep->next_1d();
    float u = ((float)px + jx) / (float)w, v = 1.0f - ((float)py + jy) / (float)h;
    b = pt_li(sc, cam.generate_ray(u, v), *rep, cfg, c);
  } else {
    jx = base_smp.next_1d();
    jy = base_smp.next_1d();
    float u = ((float)px + jx) / (float)w, v = 1.0f - ((float)py + jy) / (float)h;
    b = pt_li(sc, cam.generate_ray(u, v), base_smp, cfg, c);
  }
  auto offset = [&](uint32_t ox, uint32_t oy) -> V3 {
    float u = ((float)ox + jx) / (float)w, v = 1.0f - ((float)oy + jy) / (float)h;
    if (rep) {
      rep->start_replay();
      rep->next_1d();
      rep->next_1d();
      return pt_li(sc, cam.generate_ray(u, v), *rep, cfg, c);
    }
    CounterSampler cs(seed, py * w + px, si, 0);
    cs.dim = 2;
    return pt_li(sc, cam.generate_ray(u, v), cs, cfg, c);
  };
  V3 right = px + 1 < w ? offset(px + 1, py) : b;
  V3 down = py + 1 < h ? offset(px, py + 1) : b;
  *base = b;
  *dx = right - b;
  *dy = down - b;
}

// screened_poisson_reconstruct (gradient_domain_pt.rs:449-630), red-black
// SOR in fp32; sign = +1 is the reference RHS, -1 the textbook one.
void poisson_sor(const float* P, const float* gx, const float* gy, uint32_t w, uint32_t h, float alpha,
                 uint32_t iterations, float omega, int sign, float* out) {
  size_t n = (size_t)w * h;
  std::vector<float> img(3 * n), rhs(3 * n);
  for (int ch = 0; ch < 3; ++ch) {
    for (uint32_t y = 0; y < h; ++y)
      for (uint32_t x = 0; x < w; ++x) {
        size_t i = (size_t)y * w + x;
        img[ch * n + i] = P[3 * i + ch];
        float div = gx[3 * i + ch];
        if (x > 0) div -= gx[3 * (i - 1) + ch];
        div += gy[3 * i + ch];
        if (y > 0) div -= gy[3 * (i - w) + ch];
        rhs[ch * n + i] = alpha * P[3 * i + ch] + (sign >= 0 ? div : -div);
      }
  }
  for (uint32_t it = 0; it < iterations; ++it) {
    for (uint32_t parity = 0; parity < 2; ++parity) {
      for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
          if (((x + y) & 1u) != parity) continue;
          size_t i = (size_t)y * w + x;
          for (int ch = 0; ch < 3; ++ch) {
            const float* im = &img[ch * n];
            uint32_t nn = 0;
            float sum = 0.0f;
            if (x > 0) { sum += im[i - 1]; ++nn; }
            if (x + 1 < w) { sum += im[i + 1]; ++nn; }
            if (y > 0) { sum += im[i - w]; ++nn; }
            if (y + 1 < h) { sum += im[i + w]; ++nn; }
            float diag = alpha + (float)nn;
            float inv = 1.0f / diag;
            float gs = (rhs[ch * n + i] + sum) * inv;
            img[ch * n + i] = (1.0f - omega) * img[ch * n + i] + omega * gs;
          }
        }
    }
  }
  for (size_t i = 0; i < n; ++i)
    for (int ch = 0; ch < 3; ++ch) out[3 * i + ch] = std::fmax(img[ch * n + i], 0.0f);
}

}  // namespace orc

using namespace orc;

extern "C" {

// GDPT pass 1 (gradient_domain_pt.rs:346-393): primal, dx, dy films
// (accum * (1/spp)).  Returns wall seconds.
// Pixels [x0, x1) x [y0, y1) only (x1, y1 = 0: to the edge); the others are
// left untouched.
double orc_render_gdpt_ex(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                          uint32_t rr_depth, int rng_mode, uint64_t seed, int nthreads, uint32_t x0, uint32_t x1,
                          uint32_t y0, uint32_t y1, float* primal, float* dxo, float* dyo) {
  Scene* sc = (Scene*)s;
  Camera cam;
  const float* c = cam16;
  cam.origin = V3(c[0], c[1], c[2]);
  cam.llc = V3(c[3], c[4], c[5]);
  cam.horiz = V3(c[6], c[7], c[8]);
  cam.vert = V3(c[9], c[10], c[11]);
  cam.forward = V3(c[12], c[13], c[14]);
  cam.area = c[15];
  PtConfig cfg;
  cfg.spp = spp;
  cfg.max_depth = max_depth;
  cfg.rr_depth = rr_depth;
  auto t0 = std::chrono::steady_clock::now();
  if (x1 == 0 || x1 > w) x1 = w;
  if (y1 == 0 || y1 > h) y1 = h;
  parallel_rows(y1 - y0, nthreads, [&](uint32_t yy) {
    const uint32_t y = y0 + yy;
    for (uint32_t x = x0; x < x1; ++x) {
      RefSampler inner(RefSampler::pixel_seed(x, y));
      ReplaySampler rep(&inner);
      V3 pa, xa, ya;
      for (uint32_t si = 0; si < spp; ++si) {
        V3 b, gx, gy;
        CounterSampler cs(seed, y * w + x, si, 0);
        trace_with_gradients(*sc, cam, cs, rng_mode == RNG_REF ? &rep : nullptr, seed, x, y, si, w, h, cfg, &b, &gx,
                             &gy, nullptr);
        pa = pa + b;
        xa = xa + gx;
        ya = ya + gy;
      }
      float inv = 1.0f / (float)spp;
      size_t o = 3 * ((size_t)y * w + x);
      V3 vals[3] = {pa * inv, xa * inv, ya * inv};
      float* outs[3] = {primal, dxo, dyo};
      for (int k = 0; k < 3; ++k) {
        outs[k][o] = vals[k].x;
        outs[k][o + 1] = vals[k].y;
        outs[k][o + 2] = vals[k].z;
      }
    }
  });
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_render_gdpt(void* s, const float* cam16, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                       uint32_t rr_depth, int rng_mode, uint64_t seed, int nthreads, float* primal, float* dxo,
                       float* dyo) {
  return orc_render_gdpt_ex(s, cam16, w, h, spp, max_depth, rr_depth, rng_mode, seed, nthreads, 0, 0, 0, 0, primal,
                            dxo, dyo);
}

double orc_poisson_sor(const float* P, const float* gx, const float* gy, uint32_t w, uint32_t h, float alpha,
                       uint32_t iterations, float omega, int sign, float* out) {
  auto t0 = std::chrono::steady_clock::now();
  poisson_sor(P, gx, gy, w, h, alpha, iterations, omega, sign, out);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
