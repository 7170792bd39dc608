// Library-level entry points: errors, device init, camera construction.
#include <new>
#include <cmath>

#include "ar_internal.h"

namespace ar {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

ar_status fail(ar_status code, const std::string& msg) {
  g_error = msg;
  return code;
}

ar_status fail_exception(const char* fn) {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(AR_ERR_OUT_OF_BUDGET, std::string(fn) + ": out of host memory");
  } catch (const std::exception& e) {
    return fail(AR_ERR_INVALID_ARG, std::string(fn) + ": " + e.what());
  } catch (...) {
    return fail(AR_ERR_INVALID_ARG, std::string(fn) + ": unknown exception");
  }
}

hipStream_t pick_stream(const ar_scene* s, void* stream) {
  if (stream) return (hipStream_t)stream;
  return s ? s->stream : nullptr;
}

}  // namespace ar

using namespace ar;

extern "C" {

const char* ar_last_error(void) { return g_error.c_str(); }

const char* ar_version(void) { return "agentray-mi355x 0.1 (gfx950)"; }

ar_status ar_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(AR_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(AR_ERR_INVALID_ARG, "device index out of range");
  AR_HIP(hipSetDevice(device));
  return noalloc_setup_current_device();
}

// PinholeCamera::from_lookat (cameras.rs:91-123).
ar_status ar_camera_lookat(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov,
                           float aspect, ar_camera* out) {
  if (!lookfrom || !lookat || !vup || !out) return fail(AR_ERR_INVALID_ARG, "NULL argument");
  const float rads_per_deg = kPi / 180.0f;
  float theta = vfov * rads_per_deg;
  float h = std::tan(theta / 2.0f);
  float vh = 2.0f * h;
  float vw = aspect * vh;
  f3 from = mk(lookfrom[0], lookfrom[1], lookfrom[2]);
  f3 at = mk(lookat[0], lookat[1], lookat[2]);
  f3 up = mk(vup[0], vup[1], vup[2]);
  f3 w = normalize(from - at);
  f3 u = normalize(cross(up, w));
  f3 v = cross(w, u);
  f3 horiz = u * vw, vert = v * vh;
  f3 llc = ((from - horiz / 2.0f) - vert / 2.0f) - w;
  f3 fwd = -w;
  f3 vals[5] = {from, llc, horiz, vert, fwd};
  float* o = out->origin;
  for (int i = 0; i < 5; ++i) {
    o[3 * i] = vals[i].x;
    o[3 * i + 1] = vals[i].y;
    o[3 * i + 2] = vals[i].z;
  }
  out->area = vw * vh;
  return AR_OK;
}

// PinholeCamera::from_eye_angle (cameras.rs:131-142).
ar_status ar_camera_eye_angle(const float eye[3], float yaw, float pitch, float vfov, float aspect,
                              ar_camera* out) {
  if (!eye || !out) return fail(AR_ERR_INVALID_ARG, "NULL argument");
  const float rads_per_deg = kPi / 180.0f;
  float ry = yaw * rads_per_deg, rp = pitch * rads_per_deg;
  float sy = std::sin(ry), cy = std::cos(ry), sp = std::sin(rp), cp = std::cos(rp);
  f3 fwd = mk(sy * cp, sp, -cy * cp);
  f3 e = mk(eye[0], eye[1], eye[2]);
  f3 at = e + fwd;
  float a[3] = {at.x, at.y, at.z};
  float up[3] = {0.0f, 1.0f, 0.0f};
  return ar_camera_lookat(eye, a, up, vfov, aspect, out);
}

}  // extern "C"
