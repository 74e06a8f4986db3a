// Scene builder: transforms, Falcor's TriangleMesh generators, normal generation,
// flattening into the pass's triangle soup, and the C ABI of include/rs_import.h.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>

#include "ingest.h"

namespace ingest {

// ---------------------------------------------------------------- transforms

m4 mul(const m4& a, const m4& b) {
  m4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float s = 0.f;
      for (int k = 0; k < 4; k++) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

m4 transpose(const m4& a) {
  m4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = a.m[j][i];
  return r;
}

// Gauss-Jordan with partial pivoting in double precision.
m4 inverse(const m4& a) {
  double t[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      t[i][j] = a.m[i][j];
      t[i][j + 4] = i == j ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(t[r][c]) > std::fabs(t[piv][c])) piv = r;
    if (std::fabs(t[piv][c]) < 1e-30) throw std::runtime_error("singular transform");
    if (piv != c)
      for (int j = 0; j < 8; j++) std::swap(t[c][j], t[piv][j]);
    double inv = 1.0 / t[c][c];
    for (int j = 0; j < 8; j++) t[c][j] *= inv;
    for (int r = 0; r < 4; r++) {
      if (r == c) continue;
      double f = t[r][c];
      if (f != 0.0)
        for (int j = 0; j < 8; j++) t[r][j] -= f * t[c][j];
    }
  }
  m4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = (float)t[i][j + 4];
  return r;
}

f3 xform_point(const m4& a, f3 p) {
  float v[4] = {p.x, p.y, p.z, 1.f}, o[4];
  for (int i = 0; i < 4; i++) o[i] = a.m[i][0] * v[0] + a.m[i][1] * v[1] + a.m[i][2] * v[2] + a.m[i][3] * v[3];
  if (o[3] != 1.f && o[3] != 0.f) return {o[0] / o[3], o[1] / o[3], o[2] / o[3]};
  return {o[0], o[1], o[2]};
}

f3 xform_vector(const m4& a, f3 v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}

float det3(const m4& a) {
  const auto& m = a.m;
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

m4 translation(f3 t) {
  m4 r;
  r.m[0][3] = t.x;
  r.m[1][3] = t.y;
  r.m[2][3] = t.z;
  return r;
}

m4 scaling(f3 s) {
  m4 r;
  r.m[0][0] = s.x;
  r.m[1][1] = s.y;
  r.m[2][2] = s.z;
  return r;
}

m4 rotation(float radians, f3 axis) {
  f3 a = normalize(axis);
  float c = std::cos(radians), s = std::sin(radians), t = 1.f - c;
  m4 r;
  r.m[0][0] = t * a.x * a.x + c;
  r.m[0][1] = t * a.x * a.y - s * a.z;
  r.m[0][2] = t * a.x * a.z + s * a.y;
  r.m[1][0] = t * a.x * a.y + s * a.z;
  r.m[1][1] = t * a.y * a.y + c;
  r.m[1][2] = t * a.y * a.z - s * a.x;
  r.m[2][0] = t * a.x * a.z - s * a.y;
  r.m[2][1] = t * a.y * a.z + s * a.x;
  r.m[2][2] = t * a.z * a.z + c;
  return r;
}

m4 look_at_lh(f3 eye, f3 target, f3 up) {
  f3 z = normalize(target - eye);
  f3 x = normalize(cross(up, z));
  f3 y = cross(z, x);
  m4 r;
  float rows[3][3] = {{x.x, x.y, x.z}, {y.x, y.y, y.z}, {z.x, z.y, z.z}};
  f3 axes[3] = {x, y, z};
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) r.m[i][j] = rows[i][j];
    r.m[i][3] = -dot(axes[i], eye);
  }
  return r;
}

f3 fresnel_conductor(f3 eta, f3 k, float cosTheta) {
  auto one = [&](float e, float kk) {
    float c2 = cosTheta * cosTheta, s2 = 1.f - c2;
    float e2 = e * e, k2 = kk * kk;
    float t0 = e2 - k2 - s2;
    float a2b2 = std::sqrt(t0 * t0 + 4.f * e2 * k2);
    float t1 = a2b2 + c2;
    float a = std::sqrt(0.5f * (a2b2 + t0));
    float t2 = 2.f * a * cosTheta;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = c2 * a2b2 + s2 * s2;
    float t4 = t2 * s2;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
  };
  return {one(eta.x, k.x), one(eta.y, k.y), one(eta.z, k.z)};
}

// ---------------------------------------------------------------- TriangleMesh generators

Mesh mesh_quad(float sx, float sy) {
  Mesh m;
  m.name = "quad";
  float hx = 0.5f * sx, hy = 0.5f * sy;
  m.positions = {{-hx, 0.f, -hy}, {hx, 0.f, -hy}, {-hx, 0.f, hy}, {hx, 0.f, hy}};
  m.normals.assign(4, f3{0.f, 1.f, 0.f});
  m.indices = {2, 1, 0, 1, 2, 3};
  m.frontFaceCW = sx * sy < 0.f;
  return m;
}

Mesh mesh_disk(float radius, uint32_t segments) {
  Mesh m;
  m.name = "disk";
  const float kPi = 3.14159265358979323846f;
  m.positions.push_back({0.f, 0.f, 0.f});
  for (uint32_t i = 0; i < segments; i++) {
    float phi = float(i) / float(segments) * 2.f * kPi;
    float c = std::cos(phi), s = -std::sin(phi);
    m.positions.push_back({c * radius, 0.f, s * radius});
    m.indices.insert(m.indices.end(), {0u, i + 1, ((i + 1) % segments) + 1});
  }
  m.normals.assign(m.positions.size(), f3{0.f, 1.f, 0.f});
  return m;
}

Mesh mesh_cube(f3 size) {
  static const float P[6][4][3] = {
      {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}},
      {{-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, 0.5f}},
      {{-0.5f, 0.5f, -0.5f}, {-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}},
      {{0.5f, 0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}},
      {{-0.5f, 0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, -0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}},
      {{0.5f, 0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}},
  };
  static const float N[6][3] = {{0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}};
  Mesh m;
  m.name = "cube";
  f3 sign{size.x < 0.f ? -1.f : 1.f, size.y < 0.f ? -1.f : 1.f, size.z < 0.f ? -1.f : 1.f};
  m.frontFaceCW = size.x * size.y * size.z < 0.f;
  for (uint32_t f = 0; f < 6; f++) {
    uint32_t base = (uint32_t)m.positions.size();
    m.indices.insert(m.indices.end(), {base, base + 2, base + 1, base, base + 3, base + 2});
    for (int j = 0; j < 4; j++) {
      m.positions.push_back({P[f][j][0] * size.x, P[f][j][1] * size.y, P[f][j][2] * size.z});
      m.normals.push_back({N[f][0] * sign.x, N[f][1] * sign.y, N[f][2] * sign.z});
    }
  }
  return m;
}

Mesh mesh_sphere(float radius, uint32_t segU, uint32_t segV) {
  Mesh m;
  m.name = "sphere";
  const float kPi = 3.14159265358979323846f;
  for (uint32_t v = 0; v <= segV; v++)
    for (uint32_t u = 0; u <= segU; u++) {
      float theta = (u / float(segU)) * 2.f * kPi;
      float phi = (v / float(segV)) * kPi;
      f3 dir{std::cos(theta) * std::sin(phi), std::cos(phi), std::sin(theta) * std::sin(phi)};
      m.positions.push_back(dir * radius);
      m.normals.push_back(dir);
    }
  for (uint32_t v = 0; v < segV; v++)
    for (uint32_t u = 0; u < segU; u++) {
      uint32_t i0 = v * (segU + 1) + u;
      uint32_t i1 = v * (segU + 1) + (u + 1) % (segU + 1);
      uint32_t i2 = (v + 1) * (segU + 1) + u;
      uint32_t i3 = (v + 1) * (segU + 1) + (u + 1) % (segU + 1);
      m.indices.insert(m.indices.end(), {i0, i1, i2, i2, i1, i3});
    }
  return m;
}

void flat_normals(Mesh& mesh) {
  Mesh out;
  out.name = mesh.name;
  out.frontFaceCW = mesh.frontFaceCW;
  size_t nt = mesh.indices.size() / 3;
  out.positions.reserve(3 * nt);
  out.normals.reserve(3 * nt);
  out.indices.reserve(3 * nt);
  for (size_t t = 0; t < nt; t++) {
    f3 p[3];
    for (int k = 0; k < 3; k++) p[k] = mesh.positions[mesh.indices[3 * t + k]];
    f3 n = normalize(cross(p[1] - p[0], p[2] - p[0]));
    if (mesh.frontFaceCW) n = n * -1.f;
    for (int k = 0; k < 3; k++) {
      out.indices.push_back((uint32_t)out.positions.size());
      out.positions.push_back(p[k]);
      out.normals.push_back(n);
    }
  }
  mesh = std::move(out);
}

void smooth_normals(Mesh& mesh) {
  std::vector<f3> acc(mesh.positions.size());
  for (size_t t = 0; t + 2 < mesh.indices.size(); t += 3) {
    uint32_t i[3] = {mesh.indices[t], mesh.indices[t + 1], mesh.indices[t + 2]};
    f3 n = cross(mesh.positions[i[1]] - mesh.positions[i[0]], mesh.positions[i[2]] - mesh.positions[i[0]]);
    if (mesh.frontFaceCW) n = n * -1.f;
    for (int k = 0; k < 3; k++) acc[i[k]] = acc[i[k]] + n;  // |cross| = 2 x area: area weighting
  }
  for (auto& n : acc) n = normalize(n);
  mesh.normals = std::move(acc);
}

// ---------------------------------------------------------------- builder

void Builder::warn(const std::string& msg) {
  log += msg;
  log += '\n';
}

uint32_t Builder::add_material(const Material& m) {
  materials.push_back(m);
  dirty = true;
  return (uint32_t)materials.size() - 1;
}

uint32_t Builder::add_mesh(Mesh mesh) {
  for (uint32_t i : mesh.indices)
    if (i >= mesh.positions.size()) throw std::runtime_error("mesh '" + mesh.name + "': vertex index out of range");
  if (mesh.indices.size() % 3) throw std::runtime_error("mesh '" + mesh.name + "': index count not a multiple of 3");
  if (mesh.normals.size() != mesh.positions.size()) flat_normals(mesh);
  meshes.push_back(std::move(mesh));
  dirty = true;
  return (uint32_t)meshes.size() - 1;
}

void Builder::add_instance(uint32_t mesh, const m4& transform, uint32_t material) {
  if (mesh >= meshes.size()) throw std::runtime_error("instance of an unknown mesh");
  if (material >= materials.size()) throw std::runtime_error("instance with an unknown material");
  instances.push_back({mesh, transform, material});
  dirty = true;
}

void Builder::set_env_rgb(const Image& img, float intensity, const float tint[3], const float rot[3]) {
  envW = img.width;
  envH = img.height;
  envTexels.assign((size_t)envW * envH * 4, 0.f);
  for (size_t i = 0; i < (size_t)envW * envH; i++) {
    for (int c = 0; c < 3; c++) envTexels[4 * i + c] = img.rgb[3 * i + c];
    envTexels[4 * i + 3] = 1.f;
  }
  envIntensity = intensity;
  for (int c = 0; c < 3; c++) {
    envTint[c] = tint ? tint[c] : 1.f;
    envRotation[c] = rot ? rot[c] : 0.f;
  }
  hasEnv = true;
}

// SceneBuilder's per-instance vertex transform (SceneBuilder.cpp:1700-1710): positions by
// the instance matrix, normals by its inverse transpose, renormalised; a mirroring
// transform (negative determinant) flips the winding, as does a clockwise mesh
// (Scene.slang:395-405 isWorldFrontFaceCW), which the soup bakes into the index order.
void Builder::flatten() {
  if (!dirty) return;
  outPositions.clear();
  outNormals.clear();
  outIndices.clear();
  outTriMaterial.clear();
  outInstances.clear();
  for (const Instance& inst : instances) {
    const Mesh& mesh = meshes[inst.mesh];
    outInstances.push_back(rs_instance_desc{(uint32_t)(outPositions.size() / 3), (uint32_t)mesh.positions.size(),
                                            (uint32_t)outTriMaterial.size(), (uint32_t)(mesh.indices.size() / 3)});
    m4 nrm = transpose(inverse(inst.transform));
    bool flip = mesh.frontFaceCW != (det3(inst.transform) < 0.f);
    uint32_t base = (uint32_t)(outPositions.size() / 3);
    for (size_t v = 0; v < mesh.positions.size(); v++) {
      f3 p = xform_point(inst.transform, mesh.positions[v]);
      f3 n = normalize(xform_vector(nrm, mesh.normals[v]));
      outPositions.insert(outPositions.end(), {p.x, p.y, p.z});
      outNormals.insert(outNormals.end(), {n.x, n.y, n.z});
    }
    for (size_t t = 0; t < mesh.indices.size(); t += 3) {
      uint32_t a = mesh.indices[t], b = mesh.indices[t + 1], c = mesh.indices[t + 2];
      if (flip) std::swap(b, c);
      outIndices.insert(outIndices.end(), {base + a, base + b, base + c});
      outTriMaterial.push_back(inst.material);
    }
  }
  outMaterials.clear();
  for (const Material& m : materials) outMaterials.push_back(m.m);
  dirty = false;
}

std::string dirname_of(const std::string& path) {
  size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

std::string join_path(const std::string& dir, const std::string& rel) {
  if (!rel.empty() && rel[0] == '/') return rel;
  return dir + "/" + rel;
}

}  // namespace ingest

// ---------------------------------------------------------------- C ABI

using namespace ingest;

struct rs_builder {
  Builder b;
};

namespace {

m4 read_transform(const float* t) { return t ? m4::from_rows(t) : m4::identity(); }

template <class F>
int32_t guarded(rs_builder* h, F&& f) {
  if (!h) return RS_E_INVALID_ARG;
  try {
    return f();
  } catch (const std::exception& e) {
    h->b.warn(std::string("error: ") + e.what());
    return RS_E_IMPORT;
  }
}

}  // namespace

extern "C" {

rs_builder* rs_builder_create(void) { return new rs_builder(); }

void rs_builder_destroy(rs_builder* h) { delete h; }

const char* rs_builder_log(const rs_builder* h) { return h ? h->b.log.c_str() : ""; }

int32_t rs_builder_import(rs_builder* h, const char* path) {
  if (!path) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    std::string p(path);
    std::string ext = p.size() >= 5 ? p.substr(p.find_last_of('.') + 1) : "";
    std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
    if (ext == "pbrt") import_pbrt(h->b, p);
    else if (ext == "xml") import_mitsuba(h->b, p);
    else {
      h->b.warn("error: unsupported scene format '" + ext + "' (" + p + ")");
      return RS_E_UNSUPPORTED;
    }
    return RS_OK;
  });
}

int32_t rs_builder_add_material(rs_builder* h, const rs_material* m) {
  if (!m) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Material mat;
    mat.m = *m;
    return (int32_t)h->b.add_material(mat);
  });
}

int32_t rs_builder_add_mesh(rs_builder* h, const float* positions, const float* normals, uint32_t nv,
                            const uint32_t* indices, uint32_t nt, const float* transform, uint32_t material,
                            uint32_t frontFaceCW) {
  if (!positions || !indices || nv == 0 || nt == 0) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Mesh m;
    m.positions.resize(nv);
    for (size_t v = 0; v < nv; v++) m.positions[v] = {positions[3 * v], positions[3 * v + 1], positions[3 * v + 2]};
    if (normals) {
      m.normals.resize(nv);
      for (size_t v = 0; v < nv; v++) m.normals[v] = {normals[3 * v], normals[3 * v + 1], normals[3 * v + 2]};
    }
    m.indices.assign(indices, indices + 3 * (size_t)nt);
    m.frontFaceCW = frontFaceCW != 0;
    uint32_t id = h->b.add_mesh(std::move(m));
    h->b.add_instance(id, read_transform(transform), material);
    return (int32_t)id;
  });
}

int32_t rs_builder_add_primitive(rs_builder* h, uint32_t kind, const float* params, const float* transform,
                                 uint32_t material) {
  if (!params) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Mesh m;
    switch (kind) {
      case RS_PRIM_QUAD: m = mesh_quad(params[0], params[1]); break;
      case RS_PRIM_DISK: m = mesh_disk(params[0], (uint32_t)params[1]); break;
      case RS_PRIM_CUBE: m = mesh_cube({params[0], params[1], params[2]}); break;
      case RS_PRIM_SPHERE: m = mesh_sphere(params[0], (uint32_t)params[1], (uint32_t)params[2]); break;
      default: h->b.warn("error: unknown primitive kind"); return RS_E_INVALID_ARG;
    }
    uint32_t id = h->b.add_mesh(std::move(m));
    h->b.add_instance(id, read_transform(transform), material);
    return (int32_t)id;
  });
}

int32_t rs_builder_add_mesh_file(rs_builder* h, const char* path, const float* transform, uint32_t material,
                                 uint32_t smoothNormals) {
  if (!path) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Mesh m = read_mesh_file(path);
    if (m.normals.size() != m.positions.size()) {
      if (smoothNormals) smooth_normals(m);
      else flat_normals(m);
    }
    uint32_t id = h->b.add_mesh(std::move(m));
    h->b.add_instance(id, read_transform(transform), material);
    return (int32_t)id;
  });
}

int32_t rs_builder_add_instance(rs_builder* h, uint32_t mesh, const float* transform, uint32_t material) {
  return guarded(h, [&]() -> int32_t {
    h->b.add_instance(mesh, read_transform(transform), material);
    return RS_OK;
  });
}

int32_t rs_builder_set_camera(rs_builder* h, const rs_import_camera* c) {
  if (!h || !c) return RS_E_INVALID_ARG;
  Camera& cam = h->b.camera;
  cam.present = true;
  cam.position = {c->position[0], c->position[1], c->position[2]};
  cam.target = {c->target[0], c->target[1], c->target[2]};
  cam.up = {c->up[0], c->up[1], c->up[2]};
  cam.focalLength = c->focalLength;
  cam.frameHeight = c->frameHeight;
  cam.apertureRadius = c->apertureRadius;
  cam.focalDistance = c->focalDistance;
  return RS_OK;
}

int32_t rs_builder_set_env_map(rs_builder* h, const rs_env_map_desc* d) {
  if (!h || !d || !d->texels || d->width == 0 || d->height == 0) return RS_E_INVALID_ARG;
  Builder& b = h->b;
  b.envW = d->width;
  b.envH = d->height;
  b.envTexels.assign(d->texels, d->texels + (size_t)4 * d->width * d->height);
  b.envIntensity = d->intensity;
  for (int c = 0; c < 3; c++) {
    b.envTint[c] = d->tint[c];
    b.envRotation[c] = d->rotation[c];
  }
  b.hasEnv = true;
  return RS_OK;
}

int32_t rs_builder_set_env_map_file(rs_builder* h, const char* path, float intensity, const float tint[3],
                                    const float rotation[3], uint32_t octahedral) {
  if (!path) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Image img = read_pfm(path);
    if (octahedral) img = octahedral_to_latlong(img, 2 * img.width);
    h->b.set_env_rgb(img, intensity, tint, rotation);
    return RS_OK;
  });
}

int32_t rs_builder_set_film(rs_builder* h, uint32_t width, uint32_t height) {
  if (!h) return RS_E_INVALID_ARG;
  h->b.filmW = width;
  h->b.filmH = height;
  return RS_OK;
}

int32_t rs_builder_set_diffuse_model(rs_builder* h, uint32_t model) {
  if (!h || model > RS_DIFFUSE_LAMBERT) return RS_E_INVALID_ARG;
  h->b.diffuseModel = model;
  h->b.dirty = true;
  return RS_OK;
}

int32_t rs_builder_get_scene(rs_builder* h, rs_scene_desc* out) {
  if (!out) return RS_E_INVALID_ARG;
  return guarded(h, [&]() -> int32_t {
    Builder& b = h->b;
    if (b.instances.empty()) {
      b.warn("error: the scene has no geometry");
      return RS_E_NO_SCENE;
    }
    b.flatten();
    out->positions = b.outPositions.data();
    out->normals = b.outNormals.data();
    out->numVertices = (uint32_t)(b.outPositions.size() / 3);
    out->indices = b.outIndices.data();
    out->triMaterial = b.outTriMaterial.data();
    out->numTriangles = (uint32_t)b.outTriMaterial.size();
    out->materials = b.outMaterials.data();
    out->numMaterials = (uint32_t)b.outMaterials.size();
    out->diffuseModel = b.diffuseModel;
    out->instances = b.outInstances.data();
    out->numInstances = (uint32_t)b.outInstances.size();
    return RS_OK;
  });
}

int32_t rs_builder_get_camera(const rs_builder* h, rs_import_camera* out) {
  if (!h || !out) return RS_E_INVALID_ARG;
  const Camera& c = h->b.camera;
  std::memset(out, 0, sizeof(*out));
  out->present = c.present ? 1u : 0u;
  if (!c.present) return RS_E_INVALID_ARG;
  f3 v[3] = {c.position, c.target, c.up};
  float* dst[3] = {out->position, out->target, out->up};
  for (int i = 0; i < 3; i++) {
    dst[i][0] = v[i].x;
    dst[i][1] = v[i].y;
    dst[i][2] = v[i].z;
  }
  out->focalLength = c.focalLength;
  out->frameHeight = c.frameHeight;
  out->apertureRadius = c.apertureRadius;
  out->focalDistance = c.focalDistance;
  return RS_OK;
}

int32_t rs_builder_get_env_map(const rs_builder* h, rs_env_map_desc* out) {
  if (!h || !out) return RS_E_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  const Builder& b = h->b;
  if (!b.hasEnv) return RS_E_INVALID_ARG;
  out->texels = b.envTexels.data();
  out->width = b.envW;
  out->height = b.envH;
  out->intensity = b.envIntensity;
  for (int c = 0; c < 3; c++) {
    out->tint[c] = b.envTint[c];
    out->rotation[c] = b.envRotation[c];
  }
  return RS_OK;
}

int32_t rs_builder_get_film(const rs_builder* h, uint32_t* width, uint32_t* height) {
  if (!h || !width || !height) return RS_E_INVALID_ARG;
  *width = h->b.filmW;
  *height = h->b.filmH;
  return RS_OK;
}

}  // extern "C"
