// Mesh and image files: PLY and Wavefront OBJ (what Falcor's TriangleMesh::createFromFile
// reads through Assimp for the importers, Source/Falcor/Scene/TriangleMesh.cpp:196-279),
// PFM images, and the equal-area octahedral to lat-long conversion of environment maps.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "../rs/envmap.h"
#include "ingest.h"

namespace ingest {

namespace {

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  return s;
}

std::vector<char> read_all(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error(path + ": cannot open");
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// ---- PLY

enum class PlyType { I8, U8, I16, U16, I32, U32, F32, F64 };

PlyType ply_type(const std::string& t, const std::string& path) {
  static const std::map<std::string, PlyType> kTypes = {
      {"char", PlyType::I8},    {"int8", PlyType::I8},     {"uchar", PlyType::U8},  {"uint8", PlyType::U8},
      {"short", PlyType::I16},  {"int16", PlyType::I16},   {"ushort", PlyType::U16}, {"uint16", PlyType::U16},
      {"int", PlyType::I32},    {"int32", PlyType::I32},   {"uint", PlyType::U32},  {"uint32", PlyType::U32},
      {"float", PlyType::F32},  {"float32", PlyType::F32}, {"double", PlyType::F64}, {"float64", PlyType::F64}};
  auto it = kTypes.find(t);
  if (it == kTypes.end()) throw std::runtime_error(path + ": unknown PLY type '" + t + "'");
  return it->second;
}

size_t ply_size(PlyType t) {
  switch (t) {
    case PlyType::I8:
    case PlyType::U8: return 1;
    case PlyType::I16:
    case PlyType::U16: return 2;
    case PlyType::I32:
    case PlyType::U32:
    case PlyType::F32: return 4;
    case PlyType::F64: return 8;
  }
  return 0;
}

struct PlyProperty {
  std::string name;
  PlyType type;
  bool list = false;
  PlyType countType = PlyType::U8;
};

struct PlyElement {
  std::string name;
  size_t count = 0;
  std::vector<PlyProperty> props;
};

// Reads values of a PLY body in one of the three encodings.
struct PlyReader {
  const char* p;
  const char* end;
  int format;  // 0 ascii, 1 little endian, 2 big endian
  std::string path;

  double value(PlyType t) {
    if (format == 0) {
      while (p < end && std::isspace((unsigned char)*p)) p++;
      if (p >= end) throw std::runtime_error(path + ": PLY body ends early");
      char* e = nullptr;
      double v = std::strtod(p, &e);
      if (e == p) throw std::runtime_error(path + ": malformed PLY value");
      p = e;
      return v;
    }
    size_t n = ply_size(t);
    if ((size_t)(end - p) < n) throw std::runtime_error(path + ": PLY body ends early");
    unsigned char b[8];
    std::memcpy(b, p, n);
    p += n;
    if (format == 2) std::reverse(b, b + n);  // this host is little endian
    switch (t) {
      case PlyType::I8: return (double)(int8_t)b[0];
      case PlyType::U8: return (double)b[0];
      case PlyType::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
      case PlyType::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
      case PlyType::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
      case PlyType::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
      case PlyType::F32: { float v; std::memcpy(&v, b, 4); return v; }
      case PlyType::F64: { double v; std::memcpy(&v, b, 8); return v; }
    }
    return 0.0;
  }
};

}  // namespace

Mesh read_ply(const std::string& path) {
  std::vector<char> data = read_all(path);
  const char* begin = data.data();
  const char* end = begin + data.size();
  const char* hdrEnd = nullptr;
  static const char kEnd[] = "end_header";
  for (const char* q = begin; q + sizeof(kEnd) - 1 <= end; q++)
    if (std::memcmp(q, kEnd, sizeof(kEnd) - 1) == 0) {
      hdrEnd = q + sizeof(kEnd) - 1;
      break;
    }
  if (data.size() < 3 || std::memcmp(begin, "ply", 3) != 0 || !hdrEnd)
    throw std::runtime_error(path + ": not a PLY file");
  while (hdrEnd < end && *hdrEnd != '\n') hdrEnd++;
  if (hdrEnd < end) hdrEnd++;
  std::istringstream hdr(std::string(begin, hdrEnd));
  std::string line;
  int format = -1;
  std::vector<PlyElement> elements;
  while (std::getline(hdr, line)) {
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "format") {
      std::string f;
      ls >> f;
      format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
      if (format < 0) throw std::runtime_error(path + ": unknown PLY format '" + f + "'");
    } else if (kw == "element") {
      PlyElement e;
      ls >> e.name >> e.count;
      elements.push_back(e);
    } else if (kw == "property") {
      if (elements.empty()) throw std::runtime_error(path + ": PLY property before any element");
      std::string t;
      ls >> t;
      PlyProperty pr;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> pr.name;
        pr.list = true;
        pr.countType = ply_type(ct, path);
        pr.type = ply_type(it, path);
      } else {
        pr.type = ply_type(t, path);
        ls >> pr.name;
      }
      elements.back().props.push_back(pr);
    }
  }
  if (format < 0) throw std::runtime_error(path + ": PLY header has no format");

  Mesh mesh;
  mesh.name = path;
  PlyReader rd{hdrEnd, end, format, path};
  bool hasNormals = false;
  for (const PlyElement& e : elements) {
    if (e.name == "vertex") {
      int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1;
      for (size_t k = 0; k < e.props.size(); k++) {
        const std::string& n = e.props[k].name;
        if (n == "x") ix = (int)k;
        else if (n == "y") iy = (int)k;
        else if (n == "z") iz = (int)k;
        else if (n == "nx") inx = (int)k;
        else if (n == "ny") iny = (int)k;
        else if (n == "nz") inz = (int)k;
      }
      if (ix < 0 || iy < 0 || iz < 0) throw std::runtime_error(path + ": PLY vertices lack x, y or z");
      hasNormals = inx >= 0 && iny >= 0 && inz >= 0;
      mesh.positions.resize(e.count);
      if (hasNormals) mesh.normals.resize(e.count);
      std::vector<double> vals(e.props.size());
      for (size_t v = 0; v < e.count; v++) {
        for (size_t k = 0; k < e.props.size(); k++) {
          const PlyProperty& pr = e.props[k];
          if (pr.list) {
            size_t n = (size_t)rd.value(pr.countType);
            for (size_t j = 0; j < n; j++) rd.value(pr.type);
            vals[k] = 0.0;
          } else {
            vals[k] = rd.value(pr.type);
          }
        }
        mesh.positions[v] = {(float)vals[ix], (float)vals[iy], (float)vals[iz]};
        if (hasNormals) mesh.normals[v] = {(float)vals[inx], (float)vals[iny], (float)vals[inz]};
      }
    } else if (e.name == "face") {
      for (size_t f = 0; f < e.count; f++) {
        for (const PlyProperty& pr : e.props) {
          if (!pr.list) {
            rd.value(pr.type);
            continue;
          }
          size_t n = (size_t)rd.value(pr.countType);
          std::vector<uint32_t> poly(n);
          for (size_t j = 0; j < n; j++) poly[j] = (uint32_t)rd.value(pr.type);
          if (pr.name != "vertex_indices" && pr.name != "vertex_index") continue;
          for (size_t j = 1; j + 1 < n; j++) mesh.indices.insert(mesh.indices.end(), {poly[0], poly[j], poly[j + 1]});
        }
      }
    } else {
      for (size_t i = 0; i < e.count; i++)
        for (const PlyProperty& pr : e.props) {
          size_t n = pr.list ? (size_t)rd.value(pr.countType) : 1;
          for (size_t j = 0; j < n; j++) rd.value(pr.type);
        }
    }
  }
  for (uint32_t i : mesh.indices)
    if (i >= mesh.positions.size()) throw std::runtime_error(path + ": PLY face index out of range");
  if (!hasNormals) mesh.normals.clear();
  return mesh;
}

// Wavefront OBJ: v, vn and f records (1-based or negative relative indices; v/vt/vn,
// v//vn and v forms); other records are ignored.  A vertex is the (position, normal)
// pair a face corner names; faces are triangulated as fans.
Mesh read_obj(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error(path + ": cannot open");
  std::vector<f3> P, N;
  Mesh mesh;
  mesh.name = path;
  std::unordered_map<uint64_t, uint32_t> corner;
  bool anyMissingNormal = false;
  std::string line;
  size_t lineNo = 0;
  auto resolve = [&](long i, size_t n) -> long {
    long r = i > 0 ? i - 1 : (long)n + i;
    if (i == 0 || r < 0 || r >= (long)n)
      throw std::runtime_error(path + ":" + std::to_string(lineNo) + ": OBJ index out of range");
    return r;
  };
  while (std::getline(f, line)) {
    lineNo++;
    std::istringstream ls(line);
    std::string kw;
    if (!(ls >> kw) || kw[0] == '#') continue;
    if (kw == "v") {
      f3 p;
      ls >> p.x >> p.y >> p.z;
      P.push_back(p);
    } else if (kw == "vn") {
      f3 n;
      ls >> n.x >> n.y >> n.z;
      N.push_back(n);
    } else if (kw == "f") {
      std::vector<uint32_t> poly;
      std::string tok;
      while (ls >> tok) {
        long vi = 0, ni = 0;
        size_t s1 = tok.find('/');
        vi = std::stol(tok.substr(0, s1));
        if (s1 != std::string::npos) {
          size_t s2 = tok.find('/', s1 + 1);
          if (s2 != std::string::npos && s2 + 1 < tok.size()) ni = std::stol(tok.substr(s2 + 1));
        }
        long pv = resolve(vi, P.size());
        long pn = ni ? resolve(ni, N.size()) : -1;
        if (pn < 0) anyMissingNormal = true;
        uint64_t key = ((uint64_t)pv << 32) | (uint32_t)(pn + 1);
        auto it = corner.find(key);
        uint32_t idx;
        if (it == corner.end()) {
          idx = (uint32_t)mesh.positions.size();
          corner.emplace(key, idx);
          mesh.positions.push_back(P[pv]);
          mesh.normals.push_back(pn >= 0 ? N[pn] : f3{});
        } else {
          idx = it->second;
        }
        poly.push_back(idx);
      }
      for (size_t j = 1; j + 1 < poly.size(); j++) mesh.indices.insert(mesh.indices.end(), {poly[0], poly[j], poly[j + 1]});
    }
  }
  if (anyMissingNormal) mesh.normals.clear();
  return mesh;
}

Mesh read_mesh_file(const std::string& path) {
  std::string ext = lower(path.substr(path.find_last_of('.') + 1));
  if (ext == "ply") return read_ply(path);
  if (ext == "obj") return read_obj(path);
  throw std::runtime_error(path + ": unsupported mesh format '" + ext + "' (PLY and OBJ are read)");
}

// PFM: "PF" (RGB) or "Pf" (grey), width height, scale (negative: little endian), then
// float rows from the bottom of the image up.
Image read_pfm(const std::string& path) {
  std::vector<char> data = read_all(path);
  std::string hdr;
  size_t pos = 0, fields = 0;
  std::string tok[4];
  while (pos < data.size() && fields < 4) {
    while (pos < data.size() && std::isspace((unsigned char)data[pos])) pos++;
    size_t s = pos;
    while (pos < data.size() && !std::isspace((unsigned char)data[pos])) pos++;
    tok[fields++] = std::string(data.data() + s, pos - s);
  }
  pos++;  // the single whitespace after the scale
  if (fields < 4 || (tok[0] != "PF" && tok[0] != "Pf")) throw std::runtime_error(path + ": not a PFM file");
  Image img;
  img.width = (uint32_t)std::stoul(tok[1]);
  img.height = (uint32_t)std::stoul(tok[2]);
  float scale = std::stof(tok[3]);
  int channels = tok[0] == "PF" ? 3 : 1;
  size_t n = (size_t)img.width * img.height * channels;
  if (data.size() < pos + 4 * n) throw std::runtime_error(path + ": PFM data ends early");
  std::vector<float> raw(n);
  std::memcpy(raw.data(), data.data() + pos, 4 * n);
  if (scale > 0.f)
    for (float& v : raw) {
      unsigned char b[4];
      std::memcpy(b, &v, 4);
      std::reverse(b, b + 4);
      std::memcpy(&v, b, 4);
    }
  img.rgb.resize((size_t)img.width * img.height * 3);
  for (uint32_t y = 0; y < img.height; y++)
    for (uint32_t x = 0; x < img.width; x++) {
      size_t src = ((size_t)(img.height - 1 - y) * img.width + x) * channels;
      size_t dst = ((size_t)y * img.width + x) * 3;
      for (int c = 0; c < 3; c++) img.rgb[dst + c] = raw[src + (channels == 3 ? c : 0)];
    }
  return img;
}

// EnvMapConverter.cs.slang: each lat-long texel centre's direction, mapped to the
// equal-area octahedral square, sampled bilinearly (clamped at the border).
Image octahedral_to_latlong(const Image& oct, uint32_t width) {
  Image out;
  out.width = width;
  out.height = std::max(1u, width / 2);
  out.rgb.resize((size_t)out.width * out.height * 3);
  auto texel = [&](int x, int y, int c) {
    x = std::min(std::max(x, 0), (int)oct.width - 1);
    y = std::min(std::max(y, 0), (int)oct.height - 1);
    return oct.rgb[((size_t)y * oct.width + x) * 3 + c];
  };
  for (uint32_t y = 0; y < out.height; y++)
    for (uint32_t x = 0; x < out.width; x++) {
      rs::vec2 ll{(x + 0.5f) / out.width, (y + 0.5f) / out.height};
      rs::vec3 dir = rs::latlong_to_world(ll);
      rs::vec2 uv = rs::ndir_to_oct_ea(dir);
      float fx = uv.x * oct.width - 0.5f, fy = uv.y * oct.height - 0.5f;
      int x0 = (int)std::floor(fx), y0 = (int)std::floor(fy);
      float ax = fx - x0, ay = fy - y0;
      for (int c = 0; c < 3; c++) {
        float v = (1.f - ax) * (1.f - ay) * texel(x0, y0, c) + ax * (1.f - ay) * texel(x0 + 1, y0, c) +
                  (1.f - ax) * ay * texel(x0, y0 + 1, c) + ax * ay * texel(x0 + 1, y0 + 1, c);
        out.rgb[((size_t)y * out.width + x) * 3 + c] = v;
      }
    }
  return out;
}

}  // namespace ingest
