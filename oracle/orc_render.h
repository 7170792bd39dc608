// Oracle (test infrastructure only): scene model, camera, BSDFs, lights,
// samplers and integrators restated from the reference renderer.
#pragma once
#include <atomic>
#include <functional>
#include <thread>

#include "orc.h"

namespace orc {

// ---------------------------------------------------------------------------
// Materials: closed set mirroring the reference's Bsdf impls
// (surfaces/lambertian.rs, surfaces/microfacet.rs) plus an emission term
// (extension: the reference has no emitters; see QuadLight below).
// ---------------------------------------------------------------------------
enum MatType : uint32_t { MAT_LAMBERT = 0, MAT_CONDUCTOR = 1, MAT_DIELECTRIC = 2 };

struct Material {
  uint32_t type = MAT_LAMBERT;
  V3 albedo;       // Lambertian albedo, or conductor F0
  int tex = -1;    // Lambertian albedo texture (Lambertian::with_texture), -1: constant albedo
  float roughness = 0.5f;
  float eta = 1.5f;
  V3 emission;     // radiance emitted on the front side (extension)
  int light_id = -1;
  bool emissive() const { return !is_zero(emission); }
};

// IEC 61966-2-1 sRGB -> linear (texture.rs:144-151).
float srgb_to_linear(float x);

// ImageTexture (texture.rs:65-135): linear texels, row-major, top row
// first, converted from 8-bit sRGB at load; bilinear, repeat wrap.
struct Texture {
  uint32_t w = 0, h = 0;
  std::vector<V3> data;
  void load_rgb8(const uint8_t* rgb, uint32_t width, uint32_t height);  // texture.rs:74-93
  V3 fetch(int x, int y) const;                                            // texture.rs:126-134
  V3 sample(float u, float v) const;                                       // texture.rs:97-122
};

struct BsdfSample {
  V3 wi;
  float pdf;
  V3 f;
};

V3 bsdf_eval(const Material& m, V3 wi, V3 wo);
bool bsdf_sample(const Material& m, V3 wo, float u_sel, float ux, float uy, BsdfSample* out);
float bsdf_pdf(const Material& m, V3 wi, V3 wo);

// ---------------------------------------------------------------------------
// Lights (lights.rs, lights/point.rs) + quad area light (extension).
// ---------------------------------------------------------------------------
enum LightType : uint32_t { LIGHT_POINT = 0, LIGHT_QUAD = 1 };

struct LightSample {
  V3 wi;
  float dist;
  V3 li;
  float pdf;
  float cos_l = 0.0f;  // quad: cosine at the light (0 for a point light)
};

struct EmissionSample {
  Ray ray;
  V3 le;
  float pdf_pos, pdf_dir;
  V3 n_light;
};

struct Light {
  uint32_t type = LIGHT_POINT;
  V3 pos;        // point position, or quad corner
  V3 intensity;  // point intensity (W/sr), or quad radiance
  V3 eu, ev;     // quad edges
  V3 n;          // quad normal = normalize(cross(eu, ev))
  float area = 0;
  bool sample(V3 p, float ux, float uy, LightSample* out) const;
  V3 power() const;
  bool is_delta() const { return type == LIGHT_POINT; }
  bool sample_emission(float upx, float upy, float udx, float udy, EmissionSample* out) const;
  float pdf_emission_dir(V3 dir) const;
  // Solid-angle pdf of sampling point x (with normal n) from p; 0 if behind.
  float pdf_li(V3 p, V3 x, float dist) const;
};

struct LightDist {  // lights.rs:115-235
  bool power = true;
  std::vector<float> pmf, cdf;
  void build(const std::vector<Light>& lights, bool use_power);
  void sample_index(float u, uint32_t* idx, float* pmf_out) const;
};

// ---------------------------------------------------------------------------
// Scene (scene.rs)
// ---------------------------------------------------------------------------
struct Object {
  std::shared_ptr<Mesh> mesh;
  std::unique_ptr<Blas> blas;
  M4 xf;
  M3 nmat;
  std::vector<uint32_t> mats;  // material ids by slot
};

struct ShadingPoint {
  V3 p, n;       // n: shading normal flipped toward wo
  V3 ng_world;   // unflipped world normal (front-face test for emitters)
  Onb onb;
  float tu, tv;  // tex_uv (scene.rs:55-64)
  Material mtl;  // the material with a textured albedo resolved at tex_uv
  const Material* mat;  // &mtl
  uint32_t inst, prim;
};

struct Scene {
  std::vector<Material> materials;
  std::vector<Texture> textures;
  std::vector<std::unique_ptr<Object>> objects;
  std::vector<Light> lights;
  LightDist dist;
  bool has_dist = false;
  Tlas tlas;
  void commit(int dist_mode);  // dist_mode: 0 none, 1 uniform, 2 power
  bool intersect(const Ray& r, float* t, ShadingPoint* sp, TraceCounters* c = nullptr) const;
  bool occluded(const Ray& r, TraceCounters* c = nullptr) const {
    return tlas.intersect(r, nullptr, c);
  }
};

// ---------------------------------------------------------------------------
// Camera (cameras.rs:78-249)
// ---------------------------------------------------------------------------
struct Camera {
  V3 origin, llc, horiz, vert, forward;
  float area;
  static Camera lookat(V3 from, V3 at, V3 up, float vfov, float aspect);
  Ray generate_ray(float u, float v) const {
    V3 d = ((llc + horiz * u) + vert * v) - origin;  // uv.x*H == H*uv.x bitwise
    return Ray::make(origin, d, 0.001f, kF32Max);
  }
  struct WeSample {
    float px, py;
    V3 wi;
    float dist, we, pdf;
  };
  bool sample_we(V3 p, uint32_t w, uint32_t h, WeSample* out) const;
  float pdf_we(const Ray& r) const;
};

// ---------------------------------------------------------------------------
// Samplers (sampler.rs).
// ---------------------------------------------------------------------------
struct Sampler {
  virtual ~Sampler() {}
  virtual float next_1d() = 0;
};

// rand 0.10 SmallRng as restated: Xoshiro256++ seeded by SplitMix64 from the
// per-pixel hash (sampler.rs:79-83), f32 = (next_u64 >> 40) * 2^-24.
// Parity unpinned (the rand crate is not available here).
struct RefSampler : Sampler {
  uint64_t s[4];
  static uint64_t pixel_seed(uint32_t x, uint32_t y) {
    return ((uint64_t)y * 2654435761ull) ^ ((uint64_t)x * 805459861ull);
  }
  explicit RefSampler(uint64_t seed);
  uint64_t next_u64();
  float next_1d() override { return (float)(next_u64() >> 40) * (1.0f / 16777216.0f); }
};

// Counter-based sampler keyed on (pixel, sample, dimension): the generator
// the HIP library uses.  Philox4x32-10, pinned by the Random123 known-answer
// vectors (tests/test_oracle.py).
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
struct CounterSampler : Sampler {
  uint32_t key[2];
  uint32_t pixel, sample, stream, dim;
  CounterSampler(uint64_t seed, uint32_t pixel_, uint32_t sample_, uint32_t stream_)
      : pixel(pixel_), sample(sample_), stream(stream_), dim(0) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
  }
  float next_1d() override;
};

enum RngMode : int { RNG_REF = 0, RNG_COUNTER = 1 };

// ---------------------------------------------------------------------------
// Integrators
// ---------------------------------------------------------------------------
struct PtConfig {
  uint32_t spp = 64, max_depth = 8, rr_depth = 3;
  // Test-only switches (pins, not reference behaviour): emitter estimator
  // 0 = MIS (light sampling + BSDF hits, power heuristic), 1 = light sampling
  // only, 2 = BSDF hits only; only_len > 0 keeps the contributions of paths
  // of exactly only_len vertices (camera and light included).
  int pt_mode = 0;
  uint32_t only_len = 0;
};

float power_heuristic(float a, float b);
V3 pt_li(const Scene& sc, Ray ray, Sampler& s, const PtConfig& cfg, TraceCounters* c);

void parallel_rows(uint32_t h, int nthreads, const std::function<void(uint32_t)>& fn);

}  // namespace orc
