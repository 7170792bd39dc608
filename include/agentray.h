/*
 * agentray.h — C ABI of the MI355X-native ray-traversal and path-sampling
 * library (libagentray.so).  Plain pointers and sizes only; every function
 * returns an ar_status (0 = AR_OK) and leaves a thread-local message for
 * ar_last_error() on failure.
 *
 * The entry points replace, one for one, the interfaces of the reference
 * renderer (shiinamiyuki/agent-ray, Rust) that sit on the traversal and
 * path-sampling hot path.  Each declaration cites the reference interface it
 * stands in for.  A Rust `extern "C"` block (or ctypes, or cgo) binds this
 * header directly; see INTEGRATION.md.
 *
 * Ownership: the caller owns every input/output buffer; the library owns the
 * opaque handles.  Pointers are device pointers unless AR_HOST_PTRS is set
 * in `flags` (then the library stages the copies itself).  A committed scene
 * is immutable and may be used from several threads; each call runs on the
 * HIP stream passed in (NULL = the library's own stream) and is synchronous
 * unless AR_ASYNC is set.
 */
#ifndef AGENTRAY_H
#define AGENTRAY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ar_status;
#define AR_OK 0
#define AR_ERR_INVALID_ARG (-1)
#define AR_ERR_UNSUPPORTED_TYPE (-2)
#define AR_ERR_OUT_OF_BUDGET (-3)
#define AR_ERR_DEVICE (-4)
#define AR_ERR_QUEUE_OVERFLOW (-5)
#define AR_ERR_STATE (-6)
#define AR_ERR_CANCELLED (-7)
#define AR_ERR_NOT_FOUND (-8)

/* flags */
#define AR_HOST_PTRS 0x1u    /* buffers are host memory */
#define AR_ASYNC 0x2u        /* do not synchronize the stream before returning */
#define AR_REF_ORDER 0x4u    /* traverse left child first (reference DFS order) */
#define AR_TUNE_CLOSEST 0x10u /* ar_scene_autotune: tune the closest-hit launch only */
#define AR_TUNE_ANYHIT 0x20u  /* ar_scene_autotune: tune the any-hit launch only */
#define AR_COUNT_ANYHIT 0x40u /* ar_trace_count: count the any-hit (ar_occluded) launch */

/* ---- value types ------------------------------------------------------ */

/* Ray, 32 bytes: origin, t_min, direction (unit length, as produced by the
 * reference's Ray::new, geometry.rs:12-20), t_max. */
typedef struct ar_ray {
  float origin[3];
  float t_min;
  float dir[3];
  float t_max;
} ar_ray;

/* Closest hit, 8 bytes: world-space t and primitive id (prim_id = UINT32_MAX
 * and t = +inf on a miss).  Mirrors RayHit{t, prim_id} (accel/bvh.rs:10-20);
 * instance ids go to an optional separate array. */
typedef struct ar_hit {
  float t;
  uint32_t prim_id;
} ar_hit;

/* Material types: the closed set of Bsdf impls (surfaces.rs:25-29). */
#define AR_MAT_LAMBERTIAN 0u /* surfaces/lambertian.rs:17-65 (albedo)          */
#define AR_MAT_CONDUCTOR 1u  /* surfaces/microfacet.rs:86-153 (f0, roughness)  */
#define AR_MAT_DIELECTRIC 2u /* surfaces/microfacet.rs:163-307 (eta, roughness) */
typedef struct ar_material {
  uint32_t type;
  float albedo[3]; /* Lambertian albedo, or conductor F0 */
  float roughness; /* GGX alpha, clamped to [1e-3, 1] as the reference does */
  float eta;       /* dielectric interior IOR */
  uint32_t albedo_texture; /* Lambertian::with_texture (lambertian.rs:28-33): a texture id from
                              ar_scene_add_texture, 0 for the constant albedo */
} ar_material;

/* Light types: PointLight (lights/point.rs:22-100) and a quad area light
 * (extension; the reference has no emitters).  A quad light also inserts its
 * emissive geometry (two triangles, normal = cross(edge_u, edge_v)). */
#define AR_LIGHT_POINT 0u
#define AR_LIGHT_QUAD 1u
typedef struct ar_light {
  uint32_t type;
  float position[3];  /* point position, or quad corner */
  float intensity[3]; /* point intensity I (W/sr), or quad radiance L */
  float edge_u[3];
  float edge_v[3];
} ar_light;

/* Light selection (lights.rs:115-235). */
#define AR_LIGHTDIST_NONE 0u
#define AR_LIGHTDIST_UNIFORM 1u
#define AR_LIGHTDIST_POWER 2u

/* Pinhole camera state (PinholeCamera, cameras.rs:78-87), 16 floats. */
typedef struct ar_camera {
  float origin[3];
  float lower_left[3];
  float horizontal[3];
  float vertical[3];
  float forward[3];
  float area;
} ar_camera;

/* Integrators (integrators/mod.rs:19-27 implementors). */
#define AR_INTEGRATOR_PT 0u   /* PathTracer, path_tracer.rs:159-191            */
#define AR_INTEGRATOR_BDPT 1u /* BidirectionalPathTracer, bdpt.rs:1203-1357   */
#define AR_INTEGRATOR_GDPT 2u /* GradientDomainPathTracer, gdpt.rs:345-415    */

/* PT execution schedule. */
#define AR_SCHED_WAVEFRONT 0u  /* persistent wavefront stages over device queues */
#define AR_SCHED_MEGAKERNEL 1u /* one thread per path (the plain GPU kernel)     */

typedef struct ar_render_config {
  uint32_t integrator;
  uint32_t spp;
  uint32_t max_depth;
  uint32_t rr_depth;
  uint64_t seed;          /* counter-RNG key (Philox4x32-10) */
  uint32_t schedule;      /* AR_SCHED_* */
  uint32_t sample_chunk;  /* samples per pixel per wave (0 = auto) */
  /* BDPT (BdptConfig, bidirectional_path_tracer.rs:102-133) */
  uint32_t mis_mode;      /* 0 = power heuristic, 1 = uniform (MisWeightMode, bidirectional_path_tracer.rs:91-99);
                             2 = textbook power heuristic (no reference counterpart): every reverse pdf taken once,
                             the camera cosine and the light pick's pmf included, so the weights of a path sum to
                             one and the image equals the path tracer's in expectation */
  float mis_beta;
  /* GDPT (GdptConfig, gradient_domain_pt.rs:47-76) */
  float alpha;
  uint32_t poisson_iterations; /* CG iterations */
  int32_t poisson_sign;        /* +1: reference RHS sign; -1: textbook */
  uint32_t gdpt_outputs;       /* 1: also return primal, dx, dy films */
} ar_render_config;

typedef struct ar_render_stats {
  double seconds;        /* wall time of the render call */
  uint64_t paths;        /* camera paths traced (W*H*spp, x3 for GDPT) */
  uint64_t rays;         /* every traversal query: extension + shadow */
  uint64_t shadow_rays;
  uint32_t queue_overflow;
  uint32_t reserved;
} ar_render_stats;

typedef struct ar_scene_info {
  uint32_t n_objects;
  uint32_t n_triangles;
  uint32_t n_nodes;      /* binary nodes (two child boxes each) over all BLAS + TLAS */
  uint32_t n_lights;
  uint32_t max_depth;    /* deepest BLAS */
  uint64_t device_bytes; /* scene bytes resident in HBM */
} ar_scene_info;

typedef struct ar_scene ar_scene;

/* ---- library ----------------------------------------------------------- */
const char* ar_last_error(void);
ar_status ar_init(int device);
const char* ar_version(void);

/* ---- scene (scene.rs:113-168, accel/bvh.rs:317-448) --------------------- */
/* Scene::new's inputs are collected, then committed once. */
ar_status ar_scene_create(ar_scene** out);
ar_status ar_scene_destroy(ar_scene* s);
/* Material table entry; *out_id is the material id used by meshes. */
ar_status ar_scene_add_material(ar_scene* s, const ar_material* m, uint32_t* out_id);
/* SceneObject::new (scene.rs:29-35) + TriangleMesh (primitives/mesh.rs:7-16):
 * positions (nv x 3 floats), indices (nt x 3), optional vertex normals
 * (nv x 3 or NULL), material_slots (len 1 or nt), material ids by slot,
 * and a column-major 4x4 local-to-world transform (NULL = identity). */
ar_status ar_scene_add_mesh(ar_scene* s, const float* positions, uint32_t nv, const uint32_t* indices,
                            uint32_t nt, const float* normals, const uint32_t* material_slots,
                            uint32_t n_slots, const uint32_t* material_ids, uint32_t n_materials,
                            const float* transform, uint32_t* out_object_id);
ar_status ar_scene_add_light(ar_scene* s, const ar_light* l, uint32_t* out_light_id);
ar_status ar_scene_set_light_distribution(ar_scene* s, uint32_t mode);
/* BLASAccel::build per object + TLAS::build (bvh.rs:326-410): host SAH build,
 * flatten to the device layout, upload.  Immutable afterwards. */
ar_status ar_scene_commit(ar_scene* s);
ar_status ar_scene_get_info(const ar_scene* s, ar_scene_info* info);
/* Diagnostic, host only (no device needed): the library's SAH build over n
 * boxes (6 floats: lo, hi) in the reference's binary form (bvh.rs:32-51):
 * out_boxes (6 per node), out_meta (left_or_start, right_or_end, count per
 * node; capacity 2n-1 nodes), out_prims (n), node count and tree depth. */
ar_status ar_bvh_build_host(const float* boxes, uint32_t n, float* out_boxes, uint32_t* out_meta,
                            uint32_t* out_prims, uint32_t* out_nodes, uint32_t* out_depth);
/* Bottom-level tree builders for ar_scene_commit.  AR_BUILDER_SAH is the
 * reference's full-sweep SAH on the host (bvh.rs:71-236), for both tree
 * forms.  AR_BUILDER_SAH_SPLIT (the default) keeps that tree for
 * AR_REF_ORDER and builds the near-first 4-wide trees from an SAH tree with
 * spatial splits (a triangle crossing a split plane is referenced from both
 * sides, at most 2 references per triangle on average), for each mesh whose
 * SAH cost this lowers by at least 5% (the others keep the reference's
 * tree): about 3x the build time, 23% faster closest-hit traversal on a
 * 2^20-triangle soup, no change on a room of boxes.
 * AR_BUILDER_LBVH builds each mesh's tree on the device (Morton order, radix
 * sort, Karras hierarchy; leaves of up to four triangles): much faster to
 * build, slower to trace.  All three give the same query answers.  The TLAS
 * is always SAH.  No reference counterpart. */
#define AR_BUILDER_SAH 0u
#define AR_BUILDER_LBVH 1u
#define AR_BUILDER_SAH_SPLIT 2u
ar_status ar_scene_set_builder(ar_scene* s, uint32_t builder);
/* The device builder over n boxes, in ar_bvh_build_host's output form. */
ar_status ar_bvh_build_device(const float* boxes, uint32_t n, float* out_boxes, uint32_t* out_meta,
                              uint32_t* out_prims, uint32_t* out_nodes, uint32_t* out_depth);

/* ---- launch configuration and autotuning (no reference counterpart: the
 * reference runs on rayon's CPU pool) ---------------------------------------
 * Traversal kernels come in (workgroup size, LDS stack depth) variants
 * (64/8, 64/16, 128/8, 128/16, 256/8, 256/12) and a persistent grid of blocks_per_cu
 * workgroups per CU.  The autotuner times the candidates on a ray sample
 * (device pointers) and keeps the fastest for this scene; ar_trace,
 * ar_occluded and the wavefront stages then use it. */
typedef struct ar_trace_config {
  uint32_t workgroup_size;
  uint32_t lds_stack;
  uint32_t blocks_per_cu;
  uint32_t refill_min; /* 1..64: idle lanes a wave gathers before it takes new rays */
  uint32_t search_min; /* 0..63: a wave's node phase goes on while more lanes than this have no leaf */
  uint32_t park_slots; /* 3 or 4: triangle leaves a lane parks in one node phase */
  uint32_t balance;    /* 1: a wave that has walked its static half of the rays takes further chunks of 64
                          from a counter shared by the launch, so the waves end together (launches of at
                          least 4 chunks per wave); 0: static shares only, no atomics.  Default at commit:
                          1 for scenes of at least 4096 triangles, 0 below (cheap rays make the counter's
                          atomics cost more than the balance gains). */
  float ms; /* measured time of the winning candidate (autotune only) */
} ar_trace_config;
ar_status ar_scene_set_trace_config(ar_scene* s, const ar_trace_config* c);  /* closest and any hit */
ar_status ar_scene_get_trace_config(const ar_scene* s, ar_trace_config* c);  /* closest hit */
/* Any-hit launches (ar_occluded, the path tracers' shadow stage) keep their
 * own configuration: ar_scene_autotune tunes it separately. */
ar_status ar_scene_set_anyhit_config(ar_scene* s, const ar_trace_config* c);
ar_status ar_scene_get_anyhit_config(const ar_scene* s, ar_trace_config* c);
ar_status ar_scene_autotune(ar_scene* s, const ar_ray* sample_rays, uint64_t n, uint32_t flags, void* stream,
                            ar_trace_config* best);

/* Queue chunk of ar_render (no reference counterpart): how many paths (BDPT:
 * samples) one wavefront pass holds when cfg->sample_chunk is 0.  Times one
 * chunk of the given job at each of five sizes (2^22..2^26 paths; BDPT
 * 2^19..2^23 samples; at most spp samples per pixel), keeps the smallest
 * within 3 % of the best rate for this scene and integrator, and returns it.
 * The scene's ar_render calls of that integrator use it afterwards. */
ar_status ar_render_autotune_chunk(ar_scene* s, const ar_camera* cam, uint32_t width, uint32_t height,
                                   const ar_render_config* cfg, void* stream, uint64_t* best_target);

/* ---- traversal (TLAS::intersect, bvh.rs:413-447; Scene::occluded,
 * scene.rs:166-168) ------------------------------------------------------- */
/* Closest hit for n rays.  inst_ids may be NULL. */
ar_status ar_trace(const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst_ids,
                   uint32_t flags, void* stream);
/* No counterpart in the reference.  ar_trace (same answers, bit for bit)
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc.  It allocates
 * nothing: the walker's chunk counters (its first 256 bytes, zero between
 * calls) and the traversal stack spill live in a fixed 64 MiB array in each
 * GPU's memory, held until the process ends, and the grid is reduced to fit
 * it.  A scene whose spill does not fit even the smallest grid is rejected
 * with AR_ERR_OUT_OF_BUDGET; host pointers (AR_HOST_PTRS) with
 * AR_ERR_INVALID_ARG.  Device pointers only.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU race and must be serialised by
 * the caller. */
ar_status ar_trace_noalloc(const ar_scene* s, const ar_ray* rays, uint64_t n, ar_hit* hits, uint32_t* inst_ids,
                           uint32_t flags, void* stream);
/* Any hit (boolean equal to the reference's full closest-hit is_some()). */
ar_status ar_occluded(const ar_scene* s, const ar_ray* rays, uint64_t n, uint32_t* occluded,
                      uint32_t flags, void* stream);

/* Diagnostic, no reference counterpart: the traversal work of one ar_trace
 * (or, with AR_COUNT_ANYHIT, ar_occluded) launch over the same rays, counted
 * by an instrumented build of the same kernel run with the same grid and
 * launch configuration.  Device pointers only; the rays are traced once more
 * and the results discarded.  Record sizes: node 64 B, triangle 48 B,
 * instance 96 B, spilled stack entry 8 B. */
typedef struct ar_trace_counts {
  uint64_t rays;
  uint64_t node_visits;  /* 4-wide or binary node records read */
  uint64_t tri_tests;    /* triangle records read and tested */
  uint64_t inst_tests;   /* instance records read at TLAS leaves */
  uint64_t spill_writes; /* stack entries pushed beyond the LDS stack */
  uint64_t spill_reads;  /* spilled stack entries popped */
  uint32_t grid;         /* workgroups of the launch */
  uint32_t workgroup_size;
} ar_trace_counts;
ar_status ar_trace_count(const ar_scene* s, const ar_ray* rays, uint64_t n, uint32_t flags, void* stream,
                         ar_trace_counts* out);

/* ---- camera (cameras.rs:91-149) ----------------------------------------- */
ar_status ar_camera_lookat(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov_deg,
                           float aspect, ar_camera* out);
ar_status ar_camera_eye_angle(const float eye[3], float yaw_deg, float pitch_deg, float vfov_deg,
                              float aspect, ar_camera* out);

/* ---- rendering (Integrator::render, integrators/mod.rs:19-27) ----------- */
/* out_rgb: width*height*3 floats, row-major, top row first (the Film the
 * reference returns).  For GDPT with gdpt_outputs=1, out_aux receives the
 * primal, dx and dy films (3 x width*height*3 floats). */
ar_status ar_render(const ar_scene* s, const ar_camera* cam, uint32_t width, uint32_t height,
                    const ar_render_config* cfg, float* out_rgb, float* out_aux, ar_render_stats* stats,
                    uint32_t flags, void* stream);

/* ---- screened Poisson (gradient_domain_pt.rs:449-630) ------------------- */
/* Conjugate gradient on (alpha + n_neigh) I - sum(neigh) = alpha P + sign*div G,
 * fp32 vectors, fp64 dot products and convergence test, warm start P, output
 * clamped at 0.  Films are width*height*3 floats. */
ar_status ar_poisson_solve(const float* primal, const float* dx, const float* dy, uint32_t width,
                           uint32_t height, float alpha, uint32_t iterations, int32_t sign, float* out,
                           double* final_residual, uint32_t flags, void* stream);

/* ---- textures and OBJ import (texture.rs, importer.rs) --------------------- */
/* ImageTexture::load (texture.rs:74-93) from decoded 8-bit RGB (host memory,
 * row-major, top row first): texels are converted from sRGB to linear
 * (texture.rs:144-151) and sampled bilinearly with repeat wrap
 * (texture.rs:97-135).  Returns an id >= 1 for ar_material.albedo_texture. */
ar_status ar_scene_add_texture(ar_scene* s, const uint8_t* rgb, uint32_t width, uint32_t height, uint32_t* out_id);
/* TriangleMesh.tex_coords (primitives/mesh.rs:10) of an object added by
 * ar_scene_add_mesh: nv (u, v) pairs, one per vertex.  Without them a
 * texture is addressed by the hit's barycentrics (scene.rs:55-64). */
ar_status ar_scene_set_mesh_texcoords(ar_scene* s, uint32_t object_id, const float* uv, uint32_t nv);
/* image::open(path).into_rgb8() for PNG files (8- and 16-bit; gray, RGB,
 * palette, gray + alpha, RGBA; non-interlaced and Adam7): *rgb receives
 * width*height*3 bytes, freed with ar_image_free. */
ar_status ar_image_load_png(const char* path, uint8_t** rgb, uint32_t* width, uint32_t* height);
void ar_image_free(uint8_t* rgb);
/* load_obj_scene (importer.rs:167-241): the OBJ's models become objects of
 * the scene (one per model, split by object/group and material as tobj's
 * GPU_LOAD_OPTIONS do: triangulated, one index per vertex), MTL materials
 * are converted by convert_material (importer.rs:87-135; map_Kd textures
 * loaded relative to the OBJ's directory), `transform` (16 floats,
 * column-major, NULL = identity) applies to every object.  *n_objects
 * receives the number of objects added. */
ar_status ar_scene_load_obj(ar_scene* s, const char* path, const float* transform, uint32_t* n_objects);

/* ---- film output (film.rs:133-346) -------------------------------------- */
/* Films are width*height*3 floats (row-major, top row first), each pixel
 * multiplied by `scale` first (1/spp normalisation in the reference). */
#define AR_TONEMAP_CLAMP 0u             /* ToneMapper::Clamp (film.rs:141-142)            */
#define AR_TONEMAP_REINHARD 1u          /* ToneMapper::Reinhard (film.rs:143-144)         */
#define AR_TONEMAP_REINHARD_EXTENDED 2u /* ToneMapper::ReinhardExtended (film.rs:145-150) */
/* Film::to_rgb_image (film.rs:283-303): tone map (tonemap_pixel,
 * film.rs:153-180), encode with 1/gamma, quantise to 8 bits; out receives
 * width*height*3 bytes (RGB).  white_point is read by
 * AR_TONEMAP_REINHARD_EXTENDED only. */
ar_status ar_film_to_rgb8(const float* film, uint32_t width, uint32_t height, uint32_t tonemapper,
                          float white_point, float gamma, float scale, uint8_t* out, uint32_t flags, void* stream);
/* Film::save_exr (film.rs:310-346): linear half-float OpenEXR file, channels
 * B, G, R, ZIP compression in blocks of 16 scan lines.  Synchronises the
 * stream (the file is written on return). */
ar_status ar_film_save_exr(const float* film, uint32_t width, uint32_t height, float scale, const char* path,
                           uint32_t flags, void* stream);

/* ---- serving (no reference counterpart: the request loop above
 * Integrator::render) ------------------------------------------------------ */
typedef struct ar_server ar_server;
typedef struct ar_server_config {
  uint64_t hbm_budget_bytes; /* admission budget for per-job device state */
  uint32_t max_active_jobs;  /* job slots rendering concurrently */
  uint32_t max_concurrent_jobs; /* jobs whose chunks are on the device at once, each on its own
                                  stream (1..4; 0 = min(3, max_active_jobs)) */
  uint32_t pool_paths;       /* paths in flight in the shared path-tracer queues (0 = 2^22;
                                AR_POOL_OFF: every job renders in chunks of its own) */
  uint32_t pool_min_paths;   /* PT / GDPT wavefront jobs of at least this many paths share the
                                pool; smaller ones run alone as one chunk (0 = 2^22) */
  uint32_t pool_fill_paths;  /* while a pooled job drains its last paths, the other pooled jobs
                                keep at most this many alive (0 = pool_paths / 8) */
  uint32_t chunk_paths;      /* PT / GDPT paths per chunk of a job outside the pool (0 = 2^23) */
  uint32_t bdpt_chunk_samples; /* BDPT samples per chunk (0 = 2^22); connection passes of a quarter */
} ar_server_config;
#define AR_POOL_OFF 0xFFFFFFFFu
typedef struct ar_job_desc {
  uint32_t width, height;
  ar_camera camera;
  ar_render_config config;
} ar_job_desc;
typedef struct ar_job_stats {
  double submit_time, start_time, finish_time; /* seconds, server clock */
  uint64_t paths, rays;
  int32_t state;
} ar_job_stats;
#define AR_JOB_QUEUED 0
#define AR_JOB_RUNNING 1
#define AR_JOB_DONE 2
#define AR_JOB_CANCELLED 3
#define AR_JOB_FAILED 4

ar_status ar_server_create(const ar_scene* s, const ar_server_config* cfg, ar_server** out);
ar_status ar_server_destroy(ar_server* srv);
ar_status ar_job_submit(ar_server* srv, const ar_job_desc* job, uint64_t* out_job_id);
ar_status ar_job_poll(ar_server* srv, uint64_t job_id, int32_t* out_state);
/* Blocks until the job ends; copies its film (width*height*3 floats, host). */
ar_status ar_job_wait(ar_server* srv, uint64_t job_id, float* out_rgb, ar_job_stats* stats);
ar_status ar_job_cancel(ar_server* srv, uint64_t job_id);
double ar_server_clock(const ar_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* AGENTRAY_H */
