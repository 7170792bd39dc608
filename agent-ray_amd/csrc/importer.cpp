// Scene ingestion on the host: PNG decoding (the `image` crate's
// image::open(..).into_rgb8(), which ImageTexture::load uses,
// agent-ray src/texture.rs:74-93) and the OBJ/MTL importer
// (load_obj_scene, src/importer.rs:167-241, reading the file the way the
// `tobj` crate does with GPU_LOAD_OPTIONS: polygons triangulated as fans,
// one index per distinct position/texcoord/normal triple, a new model at
// every `o` / `g` and `usemtl` once the current one has faces).
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ar_internal.h"

using namespace ar;

namespace {

// ---- PNG ---------------------------------------------------------------------

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// Undo the per-scanline filters of one (sub)image in place; rows of `rb`
// bytes, each preceded by its filter type byte; bpp = bytes per complete
// pixel (at least 1).
bool unfilter(uint8_t* data, uint32_t rows, size_t rb, uint32_t bpp, std::vector<uint8_t>& out) {
  out.assign(rows * rb, 0);
  for (uint32_t y = 0; y < rows; ++y) {
    const uint8_t ft = data[y * (rb + 1)];
    const uint8_t* in = data + y * (rb + 1) + 1;
    uint8_t* cur = out.data() + y * rb;
    const uint8_t* prev = y ? cur - rb : nullptr;
    for (size_t i = 0; i < rb; ++i) {
      const int a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
      int v = in[i];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: return false;
      }
      cur[i] = (uint8_t)v;
    }
  }
  return true;
}

struct Png {
  uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> plte;
};

// Sample k of a row of `depth`-bit samples, scaled to 8 bits (low depths
// replicate their bits: 1-bit 0/255, 2-bit x85, 4-bit x17; 16-bit rounds
// v / 257).
uint8_t sample8(const uint8_t* row, size_t k, uint32_t depth, bool scale_low) {
  switch (depth) {
    case 16: {
      const uint32_t v = (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
      return (uint8_t)((v + 128) / 257);
    }
    case 8: return row[k];
    default: {
      const uint32_t per = 8 / depth, sh = 8 - depth * (uint32_t)(k % per + 1);
      const uint32_t v = (row[k / per] >> sh) & ((1u << depth) - 1u);
      return scale_low ? (uint8_t)(v * (255u / ((1u << depth) - 1u))) : (uint8_t)v;
    }
  }
}

// One decoded row of the image's samples -> RGB8 pixels at x positions
// x0, x0 + dx, ... of output row y.
void row_to_rgb(const Png& p, const uint8_t* row, uint32_t n, uint8_t* rgb, uint32_t y, uint32_t x0, uint32_t dx) {
  const uint32_t ch = p.ctype == 0 ? 1 : p.ctype == 2 ? 3 : p.ctype == 3 ? 1 : p.ctype == 4 ? 2 : 4;
  for (uint32_t i = 0; i < n; ++i) {
    uint8_t* o = rgb + 3 * ((size_t)y * p.w + x0 + (size_t)i * dx);
    if (p.ctype == 3) {
      const uint32_t idx = sample8(row, i, p.depth, false);
      for (int k = 0; k < 3; ++k) o[k] = 3 * idx + k < p.plte.size() ? p.plte[3 * idx + k] : 0;
    } else if (p.ctype == 0 || p.ctype == 4) {
      o[0] = o[1] = o[2] = sample8(row, (size_t)i * ch, p.depth, true);
    } else {
      for (int k = 0; k < 3; ++k) o[k] = sample8(row, (size_t)i * ch + k, p.depth, true);
    }
  }
}

ar_status decode_png(const std::vector<uint8_t>& f, std::vector<uint8_t>& rgb, uint32_t* W, uint32_t* H) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail(AR_ERR_UNSUPPORTED_TYPE, "not a PNG file");
  Png p;
  std::vector<uint8_t> z;
  size_t pos = 8;
  bool ihdr = false, iend = false;
  while (pos + 8 <= f.size() && !iend) {
    const uint32_t len = be32(&f[pos]);
    const std::string type((const char*)&f[pos + 4], 4);
    if (pos + 12 + (size_t)len > f.size()) return fail(AR_ERR_INVALID_ARG, "truncated PNG chunk");
    const uint8_t* d = &f[pos + 8];
    if (type == "IHDR") {
      if (len < 13) return fail(AR_ERR_INVALID_ARG, "bad IHDR");
      p.w = be32(d);
      p.h = be32(d + 4);
      p.depth = d[8];
      p.ctype = d[9];
      p.interlace = d[12];
      ihdr = true;
    } else if (type == "PLTE") {
      p.plte.assign(d, d + len);
    } else if (type == "IDAT") {
      z.insert(z.end(), d, d + len);
    } else if (type == "IEND") {
      iend = true;
    }
    pos += 12 + (size_t)len;
  }
  if (!ihdr || p.w == 0 || p.h == 0) return fail(AR_ERR_INVALID_ARG, "PNG without a valid IHDR");
  const bool ok_depth = (p.ctype == 0 && (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8 || p.depth == 16)) ||
                        (p.ctype == 3 && (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8)) ||
                        ((p.ctype == 2 || p.ctype == 4 || p.ctype == 6) && (p.depth == 8 || p.depth == 16));
  if (!ok_depth || p.interlace > 1) return fail(AR_ERR_UNSUPPORTED_TYPE, "unsupported PNG colour type / depth");
  if ((uint64_t)p.w * p.h > (1ull << 28)) return fail(AR_ERR_INVALID_ARG, "PNG too large");
  const uint32_t ch = p.ctype == 0 ? 1 : p.ctype == 2 ? 3 : p.ctype == 3 ? 1 : p.ctype == 4 ? 2 : 4;
  const uint32_t bits = ch * p.depth, bpp = std::max(1u, bits / 8);
  auto rowbytes = [&](uint32_t n) { return ((size_t)n * bits + 7) / 8; };
  // Adam7 passes (x0, y0, dx, dy); one pass covering everything otherwise
  const uint32_t A[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  const int npass = p.interlace ? 7 : 1;
  size_t raw_size = 0;
  for (int k = 0; k < npass; ++k) {
    const uint32_t x0 = p.interlace ? A[k][0] : 0, y0 = p.interlace ? A[k][1] : 0;
    const uint32_t dx = p.interlace ? A[k][2] : 1, dy = p.interlace ? A[k][3] : 1;
    const uint32_t pw = p.w > x0 ? (p.w - x0 + dx - 1) / dx : 0, ph = p.h > y0 ? (p.h - y0 + dy - 1) / dy : 0;
    if (pw && ph) raw_size += (size_t)ph * (rowbytes(pw) + 1);
  }
  std::vector<uint8_t> raw(raw_size);
  uLongf got = (uLongf)raw_size;
  if (uncompress(raw.data(), &got, z.data(), (uLong)z.size()) != Z_OK || got != raw_size)
    return fail(AR_ERR_INVALID_ARG, "PNG image data does not inflate to the expected size");
  rgb.assign((size_t)p.w * p.h * 3, 0);
  size_t off = 0;
  std::vector<uint8_t> rows;
  for (int k = 0; k < npass; ++k) {
    const uint32_t x0 = p.interlace ? A[k][0] : 0, y0 = p.interlace ? A[k][1] : 0;
    const uint32_t dx = p.interlace ? A[k][2] : 1, dy = p.interlace ? A[k][3] : 1;
    const uint32_t pw = p.w > x0 ? (p.w - x0 + dx - 1) / dx : 0, ph = p.h > y0 ? (p.h - y0 + dy - 1) / dy : 0;
    if (!pw || !ph) continue;
    const size_t rb = rowbytes(pw);
    if (!unfilter(raw.data() + off, ph, rb, bpp, rows)) return fail(AR_ERR_INVALID_ARG, "bad PNG filter type");
    for (uint32_t r = 0; r < ph; ++r) row_to_rgb(p, rows.data() + r * rb, pw, rgb.data(), y0 + r * dy, x0, dx);
    off += (size_t)ph * (rb + 1);
  }
  *W = p.w;
  *H = p.h;
  return AR_OK;
}

// Regular files only: a directory opens on Linux and fails on the first read.
ar_status read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(AR_ERR_NOT_FOUND, "cannot open " + path);
  struct stat sb;
  if (fstat(fileno(f), &sb) != 0 || !S_ISREG(sb.st_mode)) {
    std::fclose(f);
    return fail(AR_ERR_NOT_FOUND, "not a regular file: " + path);
  }
  out.resize((size_t)sb.st_size);
  const size_t got = out.empty() ? 0 : std::fread(out.data(), 1, out.size(), f);
  std::fclose(f);
  if (got != out.size()) return fail(AR_ERR_NOT_FOUND, "short read: " + path);
  return AR_OK;
}

// ---- OBJ / MTL ------------------------------------------------------------------

struct Mtl {  // the tobj::Material fields convert_material reads
  std::string name;
  bool has_kd = false, has_ks = false, has_ns = false, has_d = false, has_ni = false;
  float kd[3] = {0.5f, 0.5f, 0.5f}, ks[3] = {0, 0, 0}, ns = 0, d = 1, ni = 1.5f;
  std::string map_kd;
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

void parse_mtl(const std::string& text, std::vector<Mtl>& mats) {
  std::istringstream in(text);
  std::string line;
  Mtl* cur = nullptr;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') continue;
    std::string rest;
    std::getline(ls, rest);
    rest = trim(rest);
    std::istringstream rs(rest);
    if (key == "newmtl") {
      mats.emplace_back();
      mats.back().name = rest;
      cur = &mats.back();
    } else if (!cur) {
      continue;
    } else if (key == "Kd") {
      rs >> cur->kd[0] >> cur->kd[1] >> cur->kd[2];
      cur->has_kd = true;
    } else if (key == "Ks") {
      rs >> cur->ks[0] >> cur->ks[1] >> cur->ks[2];
      cur->has_ks = true;
    } else if (key == "Ns") {
      rs >> cur->ns;
      cur->has_ns = true;
    } else if (key == "Ni") {
      rs >> cur->ni;
      cur->has_ni = true;
    } else if (key == "d") {
      rs >> cur->d;
      cur->has_d = true;
    } else if (key == "Tr") {  // transparency = 1 - dissolve
      float tr = 0;
      rs >> tr;
      cur->d = 1.0f - tr;
      cur->has_d = true;
    } else if (key == "map_Kd") {
      cur->map_kd = rest;  // the rest of the line names the file
    }
  }
}

// convert_material (importer.rs:87-135) -> an ar_material; the texture is
// loaded from `base`-relative `map_Kd` (backslashes as separators).
ar_status convert_material(ar_scene* s, const Mtl& m, const std::string& base, uint32_t* out_id) {
  auto ns_to_roughness = [](float ns) {  // importer.rs:75-77
    return std::min(std::max(std::sqrt(2.0f / (ns + 2.0f)), 1e-3f), 1.0f);
  };
  const float ns = std::max(m.ns, 0.0f);
  const float ni = std::max(m.ni, 1.0f);
  const float rough = ns_to_roughness(ns);
  ar_material am{};
  am.roughness = rough;
  am.eta = ni;
  if (m.d < 0.5f) {  // transmissive: dielectric (importer.rs:99-102)
    am.type = AR_MAT_DIELECTRIC;
    am.albedo[0] = am.albedo[1] = am.albedo[2] = 1.0f;
    return ar_scene_add_material(s, &am, out_id);
  }
  const float lum = 0.2126f * m.ks[0] + 0.7152f * m.ks[1] + 0.0722f * m.ks[2];  // importer.rs:80-82
  if (lum > 0.04f) {  // glossy: conductor with F0 = Ks (importer.rs:104-108)
    am.type = AR_MAT_CONDUCTOR;
    for (int k = 0; k < 3; ++k) am.albedo[k] = m.ks[k];
    return ar_scene_add_material(s, &am, out_id);
  }
  am.type = AR_MAT_LAMBERTIAN;  // diffuse (importer.rs:110-134)
  for (int k = 0; k < 3; ++k) am.albedo[k] = m.kd[k];
  am.roughness = 1.0f;
  am.eta = 1.0f;
  if (!m.map_kd.empty()) {
    std::string rel = m.map_kd;
    for (char& c : rel)
      if (c == '\\') c = '/';
    std::string path = base.empty() ? rel : base + "/" + rel;
    uint8_t* rgb = nullptr;
    uint32_t w = 0, h = 0;
    if (ar_image_load_png(path.c_str(), &rgb, &w, &h) == AR_OK) {
      uint32_t tid = 0;
      ar_status e = ar_scene_add_texture(s, rgb, w, h, &tid);
      ar_image_free(rgb);
      if (e != AR_OK) return e;
      am.albedo_texture = tid;
    } else {
      std::fprintf(stderr, "[importer] warning: could not load texture '%s'\n", path.c_str());
    }
  }
  return ar_scene_add_material(s, &am, out_id);
}

struct Model {
  std::vector<float> pos, uv, nrm;
  std::vector<uint32_t> idx;
  bool has_uv = true, has_nrm = true;
  int mat = -1;
  std::map<std::tuple<long, long, long>, uint32_t> remap;
};

}  // namespace

extern "C" {

static ar_status image_load_png_body(const char* path, uint8_t** rgb, uint32_t* width, uint32_t* height) {
  if (!path || !rgb || !width || !height) return fail(AR_ERR_INVALID_ARG, "ar_image_load_png: NULL argument");
  std::vector<uint8_t> f, out;
  ar_status e = read_file(path, f);
  if (e != AR_OK) return e;
  e = decode_png(f, out, width, height);
  if (e != AR_OK) return e;
  *rgb = (uint8_t*)std::malloc(out.size());
  if (!*rgb) return fail(AR_ERR_OUT_OF_BUDGET, "ar_image_load_png: out of host memory");
  std::memcpy(*rgb, out.data(), out.size());
  return AR_OK;
}

void ar_image_free(uint8_t* rgb) { std::free(rgb); }

static ar_status scene_load_obj_body(ar_scene* s, const char* path, const float* transform,
                                     uint32_t* n_objects) {
  if (!s || !path) return fail(AR_ERR_INVALID_ARG, "ar_scene_load_obj: NULL argument");
  if (s->committed) return fail(AR_ERR_STATE, "scene already committed");
  std::vector<uint8_t> bytes;
  ar_status e = read_file(path, bytes);
  if (e != AR_OK) return e;
  const std::string spath(path);
  const size_t slash = spath.find_last_of('/');
  const std::string base = slash == std::string::npos ? std::string() : spath.substr(0, slash);
  std::vector<float> V, T, N;  // file-wide positions, texcoords, normals
  std::vector<Mtl> mtls;
  std::vector<Model> models;
  Model cur;
  int cur_mat = -1;
  auto flush = [&]() {
    if (!cur.idx.empty()) models.push_back(std::move(cur));
    cur = Model();
    cur.mat = cur_mat;
  };
  auto resolve = [](long i, size_t n) -> long { return i > 0 ? i - 1 : (i < 0 ? (long)n + i : -1); };
  std::istringstream in(std::string(bytes.begin(), bytes.end()));
  std::string line;
  uint32_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key) || key[0] == '#') continue;
    if (key == "v" || key == "vn") {
      float x = 0, y = 0, z = 0;
      ls >> x >> y >> z;
      std::vector<float>& dst = key == "v" ? V : N;
      dst.insert(dst.end(), {x, y, z});
    } else if (key == "vt") {
      float u = 0, v = 0;
      ls >> u >> v;
      T.insert(T.end(), {u, v});
    } else if (key == "f") {
      std::vector<uint32_t> poly;
      std::string tok;
      while (ls >> tok) {
        long vi = 0, ti = 0, ni = 0;
        const size_t a = tok.find('/'), b = a == std::string::npos ? a : tok.find('/', a + 1);
        vi = std::strtol(tok.c_str(), nullptr, 10);
        if (a != std::string::npos && a + 1 < tok.size() && tok[a + 1] != '/') ti = std::strtol(tok.c_str() + a + 1, nullptr, 10);
        if (b != std::string::npos && b + 1 < tok.size()) ni = std::strtol(tok.c_str() + b + 1, nullptr, 10);
        const long pv = resolve(vi, V.size() / 3), pt = resolve(ti, T.size() / 2), pn = resolve(ni, N.size() / 3);
        if (pv < 0 || (size_t)pv >= V.size() / 3 || (ti && (pt < 0 || (size_t)pt >= T.size() / 2)) ||
            (ni && (pn < 0 || (size_t)pn >= N.size() / 3)))
          return fail(AR_ERR_INVALID_ARG, "ar_scene_load_obj: face index out of range on line " + std::to_string(lineno));
        const auto key3 = std::make_tuple(pv, ti ? pt : -1L, ni ? pn : -1L);
        auto it = cur.remap.find(key3);
        uint32_t id;
        if (it == cur.remap.end()) {
          id = (uint32_t)(cur.pos.size() / 3);
          cur.remap.emplace(key3, id);
          cur.pos.insert(cur.pos.end(), {V[3 * pv], V[3 * pv + 1], V[3 * pv + 2]});
          if (ti) cur.uv.insert(cur.uv.end(), {T[2 * pt], T[2 * pt + 1]});
          else cur.has_uv = false;
          if (ni) cur.nrm.insert(cur.nrm.end(), {N[3 * pn], N[3 * pn + 1], N[3 * pn + 2]});
          else cur.has_nrm = false;
        } else {
          id = it->second;
        }
        poly.push_back(id);
      }
      for (size_t k = 2; k < poly.size(); ++k) cur.idx.insert(cur.idx.end(), {poly[0], poly[k - 1], poly[k]});
    } else if (key == "o" || key == "g") {
      flush();
    } else if (key == "usemtl") {
      std::string name;
      std::getline(ls, name);
      name = trim(name);
      int id = -1;
      for (size_t k = 0; k < mtls.size(); ++k)
        if (mtls[k].name == name) id = (int)k;
      if (!cur.idx.empty()) flush();
      cur_mat = id;
      cur.mat = id;
    } else if (key == "mtllib") {
      std::string name;
      std::getline(ls, name);
      name = trim(name);
      std::vector<uint8_t> mb;
      if (read_file(base.empty() ? name : base + "/" + name, mb) == AR_OK) parse_mtl(std::string(mb.begin(), mb.end()), mtls);
    }
  }
  flush();
  // all MTL materials converted up front; no MTL: one mid-grey Lambertian
  std::vector<uint32_t> mat_ids;
  for (const Mtl& m : mtls) {
    uint32_t id = 0;
    e = convert_material(s, m, base, &id);
    if (e != AR_OK) return e;
    mat_ids.push_back(id);
  }
  if (mat_ids.empty()) {
    ar_material d{};
    d.type = AR_MAT_LAMBERTIAN;
    d.albedo[0] = d.albedo[1] = d.albedo[2] = 0.5f;
    d.roughness = 1.0f;
    d.eta = 1.0f;
    uint32_t id = 0;
    e = ar_scene_add_material(s, &d, &id);
    if (e != AR_OK) return e;
    mat_ids.push_back(id);
  }
  uint32_t added = 0;
  for (const Model& m : models) {
    const uint32_t slot = (uint32_t)std::min<long>(m.mat < 0 ? 0 : m.mat, (long)mat_ids.size() - 1);
    uint32_t obj = 0;
    e = ar_scene_add_mesh(s, m.pos.data(), (uint32_t)(m.pos.size() / 3), m.idx.data(), (uint32_t)(m.idx.size() / 3),
                          (m.has_nrm && !m.nrm.empty()) ? m.nrm.data() : nullptr, &slot, 1, mat_ids.data(),
                          (uint32_t)mat_ids.size(), transform, &obj);
    if (e != AR_OK) return e;
    if (m.has_uv && !m.uv.empty()) {
      e = ar_scene_set_mesh_texcoords(s, obj, m.uv.data(), (uint32_t)(m.uv.size() / 2));
      if (e != AR_OK) return e;
    }
    ++added;
  }
  if (n_objects) *n_objects = added;
  return AR_OK;
}

// The readers' containers and streams can throw (host memory, stream
// errors); nothing may unwind through the C ABI.
ar_status ar_image_load_png(const char* path, uint8_t** rgb, uint32_t* width, uint32_t* height) {
  try {
    return image_load_png_body(path, rgb, width, height);
  } catch (const std::bad_alloc&) {
    return fail(AR_ERR_OUT_OF_BUDGET, "ar_image_load_png: out of host memory");
  } catch (const std::exception& e) {
    return fail(AR_ERR_INVALID_ARG, std::string("ar_image_load_png: ") + e.what());
  }
}

ar_status ar_scene_load_obj(ar_scene* s, const char* path, const float* transform, uint32_t* n_objects) {
  try {
    return scene_load_obj_body(s, path, transform, n_objects);
  } catch (const std::bad_alloc&) {
    return fail(AR_ERR_OUT_OF_BUDGET, "ar_scene_load_obj: out of host memory");
  } catch (const std::exception& e) {
    return fail(AR_ERR_INVALID_ARG, std::string("ar_scene_load_obj: ") + e.what());
  }
}

}  // extern "C"
