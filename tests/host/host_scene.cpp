// Test helper (built by tests/test_importer.py with g++, never shipped):
// reads back what ar_scene_load_obj / ar_scene_add_texture put into an
// uncommitted scene (meshes, materials, textures), so the CPU suite can
// check the importer without a GPU.
#include <cstring>

#include "../../agent-ray_amd/csrc/ar_internal.h"

extern "C" {

// counts[0..2] = meshes, materials, textures
int host_scene_counts(ar_scene* s, uint32_t* counts) {
  counts[0] = (uint32_t)s->meshes.size();
  counts[1] = (uint32_t)s->materials.size();
  counts[2] = (uint32_t)s->tex_info.size();
  return 0;
}

// sizes[0..5] = vertices, triangles, has normals, has uvs, slots, materials
int host_scene_mesh_sizes(ar_scene* s, uint32_t i, uint32_t* sizes) {
  const ar::HostMesh& m = s->meshes[i];
  sizes[0] = (uint32_t)(m.pos.size() / 3);
  sizes[1] = (uint32_t)(m.idx.size() / 3);
  sizes[2] = m.nrm.empty() ? 0 : 1;
  sizes[3] = m.uv.empty() ? 0 : 1;
  sizes[4] = (uint32_t)m.slots.size();
  sizes[5] = (uint32_t)m.mats.size();
  return 0;
}

int host_scene_mesh(ar_scene* s, uint32_t i, float* pos, float* nrm, float* uv, uint32_t* idx, uint32_t* slots,
                    uint32_t* mats, float* xf) {
  const ar::HostMesh& m = s->meshes[i];
  std::memcpy(pos, m.pos.data(), m.pos.size() * 4);
  if (!m.nrm.empty()) std::memcpy(nrm, m.nrm.data(), m.nrm.size() * 4);
  if (!m.uv.empty()) std::memcpy(uv, m.uv.data(), m.uv.size() * 4);
  std::memcpy(idx, m.idx.data(), m.idx.size() * 4);
  std::memcpy(slots, m.slots.data(), m.slots.size() * 4);
  std::memcpy(mats, m.mats.data(), m.mats.size() * 4);
  std::memcpy(xf, m.xf, 64);
  return 0;
}

// out[0..6] = type, albedo r g b, roughness, eta, texture (-1 none)
int host_scene_material(ar_scene* s, uint32_t i, float* out) {
  const ar::DMaterial& d = s->materials[i];
  out[0] = (float)d.type;
  for (int k = 0; k < 3; ++k) out[1 + k] = d.albedo[k];
  out[4] = d.roughness;
  out[5] = d.eta;
  out[6] = (float)d.tex;
  return 0;
}

// wh[0..1] = width, height; texels: width*height*3 linear floats (if non-null)
int host_scene_texture(ar_scene* s, uint32_t i, uint32_t* wh, float* texels) {
  const uint4 t = s->tex_info[i];
  wh[0] = t.y;
  wh[1] = t.z;
  if (texels)
    for (uint32_t k = 0; k < t.y * t.z; ++k) {
      const float4 c = s->texels[t.x + k];
      texels[3 * k] = c.x, texels[3 * k + 1] = c.y, texels[3 * k + 2] = c.z;
    }
  return 0;
}

}  // extern "C"
