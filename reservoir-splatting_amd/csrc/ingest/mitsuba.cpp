// Mitsuba 3 XML scene files: a small XML reader, the object/property model with
// <default> substitution and <ref> resolution, and the conversion of
// Source/plugins/importers/MitsubaImporter/MitsubaImporter.cpp:262-790 into builder
// materials, meshes, camera and environment map.
//
// Conversion (StandardMaterial targets; the modelled pass has no PBRT material family):
//   * diffuse -> base colour = reflectance, roughness 1, metallic 0 (the reference uses its
//     Lambertian PBRTDiffuseMaterial; an import sets the scene's diffuse model to Lambert);
//   * plastic / roughplastic -> base colour = diffuse_reflectance, IoR = int/ext, roughness
//     0 or sqrt(alpha) (alpha default 0.1), as the reference;
//   * conductor / roughconductor -> metallic, base colour = normal-incidence Fresnel of
//     eta + ik (the reference keeps eta/k in a PBRTConductorMaterial), roughness 0 or
//     sqrt(alpha); without eta/k a perfect mirror;
//   * dielectric / thindielectric / roughdielectric -> IoR int/ext; transmission is not
//     modelled, so the surface keeps its specular reflection only;
//   * twosided -> the nested BSDF made double-sided;
//   * area emitter nested in a shape with a diffuse BSDF -> emissive = radiance, base
//     colour kept, metallic 0, roughness 1 (:559-590; other BSDFs are an error, as there);
//   * perspective / thinlens sensor -> camera node to_world * flipZ, focal length from
//     fov on a film width of 24 / height * width mm (:598-660), film 768 x 576 default;
//   * constant emitter -> a one-texel environment map; envmap (PFM here) -> lat-long map
//     with intensity = scale and rotation = EulerXYZ(to_world * flipZ) (:662-724);
//   * shapes obj / ply (smooth normals unless face_normals), sphere (radius, center),
//     disk and rectangle (to_world * YtoZ), cube (size 2) (:452-521).
// Transform operations pre-multiply in document order (Parser.h:903-1017).  <matrix> is
// read row-major as Mitsuba defines it (the reference's parser stores it transposed).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>

#include "ingest.h"

namespace ingest {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------- XML

struct XmlNode {
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XmlNode>> children;
  int line = 0;
  const std::string* attr(const std::string& n) const {
    for (const auto& a : attrs)
      if (a.first == n) return &a.second;
    return nullptr;
  }
};

class XmlReader {
 public:
  XmlReader(const std::string& path) : path_(path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(path + ": cannot open");
    s_.assign((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  }

  std::unique_ptr<XmlNode> document() {
    skip_misc();
    auto root = element();
    skip_misc();
    return root;
  }

 private:
  std::string path_, s_;
  size_t p_ = 0;
  int line_ = 1;

  [[noreturn]] void error(const std::string& msg) {
    throw std::runtime_error(path_ + ":" + std::to_string(line_) + ": " + msg);
  }
  void adv(size_t n = 1) {
    for (size_t i = 0; i < n && p_ < s_.size(); i++)
      if (s_[p_++] == '\n') line_++;
  }
  bool starts(const char* t) const { return s_.compare(p_, std::strlen(t), t) == 0; }
  void ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) adv();
  }
  void skip_until(const char* t) {
    while (p_ < s_.size() && !starts(t)) adv();
    if (p_ >= s_.size()) error(std::string("missing '") + t + "'");
    adv(std::strlen(t));
  }
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!")) skip_until(">");
      else break;
    }
  }
  std::string name() {
    size_t b = p_;
    while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_' || s_[p_] == '-' ||
                              s_[p_] == ':' || s_[p_] == '.'))
      adv();
    if (b == p_) error("expected a name");
    return s_.substr(b, p_ - b);
  }
  static std::string unescape(const std::string& v) {
    std::string o;
    for (size_t i = 0; i < v.size(); i++) {
      if (v[i] == '&') {
        size_t e = v.find(';', i);
        if (e != std::string::npos) {
          std::string ent = v.substr(i + 1, e - i - 1);
          const char* rep = ent == "lt" ? "<" : ent == "gt" ? ">" : ent == "amp" ? "&" : ent == "quot" ? "\"" :
                            ent == "apos" ? "'" : nullptr;
          if (rep) {
            o += rep;
            i = e;
            continue;
          }
        }
      }
      o += v[i];
    }
    return o;
  }
  std::unique_ptr<XmlNode> element() {
    if (!starts("<")) error("expected an element");
    adv();
    auto n = std::make_unique<XmlNode>();
    n->line = line_;
    n->tag = name();
    for (;;) {
      ws();
      if (starts("/>")) {
        adv(2);
        return n;
      }
      if (starts(">")) {
        adv();
        break;
      }
      std::string an = name();
      ws();
      if (!starts("=")) error("expected '=' after attribute '" + an + "'");
      adv();
      ws();
      char q = p_ < s_.size() ? s_[p_] : 0;
      if (q != '"' && q != '\'') error("expected a quoted attribute value");
      adv();
      size_t b = p_;
      while (p_ < s_.size() && s_[p_] != q) adv();
      if (p_ >= s_.size()) error("unterminated attribute value");
      n->attrs.push_back({an, unescape(s_.substr(b, p_ - b))});
      adv();
    }
    for (;;) {
      while (p_ < s_.size() && s_[p_] != '<') adv();  // text content is ignored
      if (p_ >= s_.size()) error("unterminated element '" + n->tag + "'");
      if (starts("<!--")) {
        skip_until("-->");
        continue;
      }
      if (starts("</")) {
        adv(2);
        std::string end = name();
        if (end != n->tag) error("mismatched closing tag '" + end + "' for '" + n->tag + "'");
        ws();
        if (!starts(">")) error("expected '>'");
        adv();
        return n;
      }
      n->children.push_back(element());
    }
  }
};

// ---------------------------------------------------------------- object model

enum class Cls { Scene, BSDF, Shape, Emitter, Sensor, Film, Texture, Medium, Other };

Cls class_of(const std::string& tag) {
  if (tag == "scene") return Cls::Scene;
  if (tag == "bsdf") return Cls::BSDF;
  if (tag == "shape") return Cls::Shape;
  if (tag == "emitter") return Cls::Emitter;
  if (tag == "sensor") return Cls::Sensor;
  if (tag == "film") return Cls::Film;
  if (tag == "texture") return Cls::Texture;
  if (tag == "medium") return Cls::Medium;
  return Cls::Other;  // integrator, sampler, rfilter, phase, volume, ...
}

struct Object;

struct Props {
  std::map<std::string, float> floats;
  std::map<std::string, int64_t> ints;
  std::map<std::string, bool> bools;
  std::map<std::string, std::string> strings;
  std::map<std::string, f3> vectors, colors;
  std::map<std::string, m4> transforms;
  std::vector<std::pair<std::string, const Object*>> refs;  // named or nested children
};

struct Object {
  Cls cls = Cls::Other;
  std::string type, id;
  Props props;
  int line = 0;
};

class MitsubaScene {
 public:
  MitsubaScene(Builder& b, const std::string& path) : b_(b), path_(path), dir_(dirname_of(path)) {}

  void run() {
    XmlReader rd(path_);
    auto doc = rd.document();
    if (doc->tag != "scene") error(doc->line, "root element must be <scene>");
    const Object* scene = parse_object(*doc);
    build(*scene);
  }

 private:
  Builder& b_;
  std::string path_, dir_;
  std::vector<std::unique_ptr<Object>> objects_;
  std::map<std::string, const Object*> ids_;
  std::map<std::string, std::string> defaults_;
  std::set<std::string> warned_;

  [[noreturn]] void error(int line, const std::string& msg) {
    throw std::runtime_error(path_ + ":" + std::to_string(line) + ": " + msg);
  }
  void warn_once(const std::string& msg) {
    if (warned_.insert(msg).second) b_.warn(path_ + ": " + msg);
  }

  std::string subst(const std::string& v) {
    std::string o;
    for (size_t i = 0; i < v.size(); i++) {
      if (v[i] == '$') {
        size_t e = i + 1;
        while (e < v.size() && (std::isalnum((unsigned char)v[e]) || v[e] == '_')) e++;
        auto it = defaults_.find(v.substr(i + 1, e - i - 1));
        if (it != defaults_.end()) {
          o += it->second;
          i = e - 1;
          continue;
        }
      }
      o += v[i];
    }
    return o;
  }
  std::string attr(const XmlNode& n, const std::string& a, bool required = true) {
    const std::string* v = n.attr(a);
    if (!v) {
      if (required) error(n.line, "<" + n.tag + "> needs attribute '" + a + "'");
      return "";
    }
    return subst(*v);
  }
  float to_float(const XmlNode& n, const std::string& s) {
    char* e = nullptr;
    float v = std::strtof(s.c_str(), &e);
    while (e && *e && std::isspace((unsigned char)*e)) e++;
    if (!e || *e || s.empty()) error(n.line, "could not parse floating point value '" + s + "'");
    return v;
  }
  std::vector<float> floats(const XmlNode& n, const std::string& s) {
    std::vector<float> v;
    std::string t;
    std::istringstream ss(s);
    std::string tok;
    std::string cleaned = s;
    std::replace(cleaned.begin(), cleaned.end(), ',', ' ');
    std::istringstream cs(cleaned);
    while (cs >> tok) v.push_back(to_float(n, tok));
    return v;
  }
  f3 xyz(const XmlNode& n, float def) {
    std::string v = attr(n, "value", false);
    if (!v.empty()) {
      if (n.attr("x") || n.attr("y") || n.attr("z")) error(n.line, "can't mix and match 'value' and 'x'/'y'/'z'");
      std::vector<float> f = floats(n, v);
      if (f.size() == 1) return {f[0], f[0], f[0]};
      if (f.size() == 3) return {f[0], f[1], f[2]};
      error(n.line, "'value' needs one or three components");
    }
    auto comp = [&](const char* c) {
      std::string s = attr(n, c, false);
      return s.empty() ? def : to_float(n, s);
    };
    return {comp("x"), comp("y"), comp("z")};
  }
  f3 named_vec(const XmlNode& n, const std::string& a) {
    std::vector<float> f = floats(n, attr(n, a));
    if (f.size() != 3) error(n.line, "'" + a + "' needs three components");
    return {f[0], f[1], f[2]};
  }

  m4 parse_transform(const XmlNode& tn) {
    m4 t;
    for (const auto& c : tn.children) {
      const XmlNode& op = *c;
      if (op.tag == "translate") {
        t = mul(translation(xyz(op, 0.f)), t);
      } else if (op.tag == "scale") {
        t = mul(scaling(xyz(op, 1.f)), t);
      } else if (op.tag == "rotate") {
        f3 axis = xyz(op, 0.f);
        float angle = to_float(op, attr(op, "angle"));
        t = mul(rotation(angle / 180.f * kPi, axis), t);
      } else if (op.tag == "lookat") {
        f3 origin = named_vec(op, "origin"), target = named_vec(op, "target");
        f3 up = op.attr("up") ? named_vec(op, "up") : f3{};
        f3 dir = normalize(target - origin);
        if (length(up) == 0.f) {
          // buildFrame: any vector orthogonal to dir
          up = std::fabs(dir.x) > std::fabs(dir.z) ? f3{-dir.y, dir.x, 0.f} : f3{0.f, -dir.z, dir.y};
        }
        f3 right = normalize(cross(normalize(up), dir));
        f3 newUp = cross(dir, right);
        m4 la;
        f3 cols[4] = {right, newUp, dir, origin};
        for (int j = 0; j < 4; j++) {
          la.m[0][j] = cols[j].x;
          la.m[1][j] = cols[j].y;
          la.m[2][j] = cols[j].z;
        }
        t = mul(la, t);
      } else if (op.tag == "matrix") {
        std::vector<float> f = floats(op, attr(op, "value"));
        m4 m;
        if (f.size() == 16) {
          for (int i = 0; i < 16; i++) m.m[i / 4][i % 4] = f[i];
        } else if (f.size() == 9) {
          for (int i = 0; i < 9; i++) m.m[i / 3][i % 3] = f[i];
        } else {
          error(op.line, "matrix needs 9 or 16 values");
        }
        t = mul(m, t);
      } else {
        error(op.line, "unexpected <" + op.tag + "> in a transform");
      }
    }
    return t;
  }

  const Object* parse_object(const XmlNode& n) {
    auto obj = std::make_unique<Object>();
    obj->cls = class_of(n.tag);
    obj->type = attr(n, "type", false);
    obj->id = attr(n, "id", false);
    obj->line = n.line;
    for (const auto& cp : n.children) {
      const XmlNode& c = *cp;
      const std::string& tag = c.tag;
      if (tag == "default") {
        std::string nm = attr(c, "name");
        if (!defaults_.count(nm)) defaults_[nm] = attr(c, "value");
        continue;
      }
      if (tag == "include") {
        warn_once("<include> is not supported and ignored");
        continue;
      }
      if (tag == "ref") {
        std::string id = attr(c, "id");
        auto it = ids_.find(id);
        if (it == ids_.end()) error(c.line, "reference to unknown object '" + id + "'");
        obj->props.refs.push_back({attr(c, "name", false), it->second});
        continue;
      }
      if (class_of(tag) != Cls::Other || tag == "integrator" || tag == "sampler" || tag == "rfilter" ||
          tag == "phase" || tag == "volume") {
        const Object* child = parse_object(c);
        obj->props.refs.push_back({attr(c, "name", false), child});
        continue;
      }
      std::string nm = attr(c, "name");
      if (tag == "float") obj->props.floats[nm] = to_float(c, attr(c, "value"));
      else if (tag == "integer") obj->props.ints[nm] = (int64_t)std::stoll(attr(c, "value"));
      else if (tag == "boolean") obj->props.bools[nm] = attr(c, "value") == "true";
      else if (tag == "string") obj->props.strings[nm] = attr(c, "value");
      else if (tag == "point" || tag == "vector") obj->props.vectors[nm] = xyz(c, 0.f);
      else if (tag == "rgb" || tag == "color") {
        std::vector<float> f = floats(c, attr(c, "value"));
        if (f.size() == 1) f = {f[0], f[0], f[0]};
        if (f.size() != 3) error(c.line, "RGB value requires one or three components");
        obj->props.colors[nm] = {f[0], f[1], f[2]};
      } else if (tag == "spectrum") {
        std::vector<float> f = floats(c, attr(c, "value"));
        if (f.size() != 1) {
          warn_once("sampled spectra are not converted; using 1");
          f = {1.f};
        }
        obj->props.colors[nm] = {f[0], f[0], f[0]};
      } else if (tag == "transform") {
        obj->props.transforms[nm] = parse_transform(c);
      } else {
        error(c.line, "unexpected <" + tag + ">");
      }
    }
    const Object* raw = obj.get();
    if (!obj->id.empty()) ids_[obj->id] = raw;
    objects_.push_back(std::move(obj));
    return raw;
  }

  // ------------------------------------------------------------ conversion

  float ior(const Props& p, const std::string& n, const std::string& def) {
    auto it = p.floats.find(n);
    if (it != p.floats.end()) return it->second;
    static const std::map<std::string, float> kIOR = {
        {"vacuum", 1.0f},        {"helium", 1.000036f},       {"hydrogen", 1.000132f}, {"air", 1.000277f},
        {"carbon dioxide", 1.00045f}, {"water", 1.3330f},     {"acetone", 1.36f},      {"ethanol", 1.361f},
        {"carbon tetrachloride", 1.461f}, {"glycerol", 1.4729f}, {"benzene", 1.501f}, {"silicone oil", 1.52045f},
        {"bromine", 1.661f},     {"water ice", 1.31f},        {"fused quartz", 1.458f}, {"pyrex", 1.470f},
        {"acrylic glass", 1.49f}, {"polypropylene", 1.49f},   {"bk7", 1.5046f},        {"sodium chloride", 1.544f},
        {"amber", 1.55f},        {"pet", 1.5750f},            {"diamond", 2.419f}};
    auto s = p.strings.find(n);
    std::string name = s != p.strings.end() ? s->second : def;
    std::transform(name.begin(), name.end(), name.begin(), ::tolower);
    auto t = kIOR.find(name);
    if (t == kIOR.end()) {
      warn_once("'" + name + "' is not a valid IOR name");
      return 0.f;
    }
    return t->second;
  }

  // A colour or float property; textures fall back to the default with a warning.
  f3 color(const Props& p, const std::string& n, f3 def) {
    auto c = p.colors.find(n);
    if (c != p.colors.end()) return c->second;
    auto f = p.floats.find(n);
    if (f != p.floats.end()) return {f->second, f->second, f->second};
    for (const auto& r : p.refs)
      if (r.first == n && r.second->cls == Cls::Texture) {
        warn_once("textures are not supported; using a constant value for '" + n + "'");
        return def;
      }
    return def;
  }

  float alpha(const Props& p) {
    if (p.strings.count("distribution")) warn_once("parameter 'distribution' is not supported");
    if (p.bools.count("sample_visible")) warn_once("parameter 'sample_visible' is not supported");
    f3 a = color(p, "alpha", {0.1f, 0.1f, 0.1f});
    return a.x;
  }

  struct Bsdf {
    Material mat;
    bool diffuse = false;  // the reference's PBRTDiffuse (needed under an area emitter)
  };

  bool build_bsdf(const Object& o, Bsdf& out) {
    const Props& p = o.props;
    Material& mat = out.mat;
    mat = Material();
    mat.name = o.id;
    rs_material& m = mat.m;
    auto base = [&](f3 c) {
      m.baseColor[0] = c.x;
      m.baseColor[1] = c.y;
      m.baseColor[2] = c.z;
    };
    const std::string& t = o.type;
    if (t == "diffuse") {
      base(color(p, "reflectance", {0.5f, 0.5f, 0.5f}));
      m.roughness = 1.f;
      m.metallic = 0.f;
      out.diffuse = true;
      b_.diffuseModel = RS_DIFFUSE_LAMBERT;
    } else if (t == "dielectric" || t == "roughdielectric" || t == "thindielectric") {
      float ii = ior(p, "int_ior", "bk7"), ei = ior(p, "ext_ior", "air");
      if (p.floats.count("specular_reflectance")) warn_once("parameter 'specular_reflectance' is not supported");
      if (p.floats.count("specular_transmittance")) warn_once("parameter 'specular_transmittance' is not supported");
      m.doubleSided = 1;
      m.roughness = t == "roughdielectric" ? std::sqrt(alpha(p)) : 0.f;
      m.ior = ii / ei;
      base({0.f, 0.f, 0.f});
      warn_once("'" + t + "': specular transmission is not modelled; imported as its specular reflection");
    } else if (t == "conductor" || t == "roughconductor") {
      if (p.floats.count("specular_reflectance")) warn_once("parameter 'specular_reflectance' is not supported");
      if (p.colors.count("eta") && p.colors.count("k")) base(fresnel_conductor(p.colors.at("eta"), p.colors.at("k"), 1.f));
      else base({1.f, 1.f, 1.f});
      m.metallic = 1.f;
      m.doubleSided = 1;
      float a = 0.f;
      if (t == "roughconductor") a = alpha(p);
      m.roughness = std::sqrt(a);
    } else if (t == "plastic" || t == "roughplastic") {
      base(color(p, "diffuse_reflectance", {0.5f, 0.5f, 0.5f}));
      float ii = ior(p, "int_ior", "polypropylene"), ei = ior(p, "ext_ior", "air");
      if (p.bools.count("nonlinear")) warn_once("parameter 'nonlinear' is not supported");
      m.roughness = t == "roughplastic" ? std::sqrt(alpha(p)) : 0.f;
      m.ior = ii / ei;
    } else if (t == "twosided") {
      bool found = false;
      for (const auto& r : p.refs)
        if (r.second->cls == Cls::BSDF) {
          if (found) error(o.line, "'twosided' BSDF can only have one nested BSDF");
          if (!build_bsdf(*r.second, out)) return false;
          out.mat.m.doubleSided = 1;
          found = true;
        }
      return found;
    } else {
      warn_once("BSDF type '" + t + "' is not supported");
      return false;
    }
    return true;
  }

  void build_shape(const Object& o) {
    const Props& p = o.props;
    m4 toWorld = p.transforms.count("to_world") ? p.transforms.at("to_world") : m4::identity();
    if (p.bools.count("flip_normals")) warn_once("parameter 'flip_normals' is not supported");
    m4 yToZ;
    yToZ.m[1][1] = 0.f;
    yToZ.m[1][2] = 1.f;
    yToZ.m[2][1] = 1.f;
    yToZ.m[2][2] = 0.f;
    Mesh mesh;
    m4 transform = toWorld;
    const std::string& t = o.type;
    if (t == "obj" || t == "ply") {
      auto fn = p.strings.find("filename");
      if (fn == p.strings.end()) error(o.line, "shape '" + t + "' needs a filename");
      mesh = read_mesh_file(join_path(dir_, fn->second));
      bool faceNormals = p.bools.count("face_normals") && p.bools.at("face_normals");
      if (mesh.normals.size() != mesh.positions.size()) {
        if (faceNormals) flat_normals(mesh);
        else smooth_normals(mesh);
      }
      if (p.bools.count("flip_tex_coords")) warn_once("parameter 'flip_tex_coords' is not supported");
    } else if (t == "sphere") {
      f3 center = p.vectors.count("center") ? p.vectors.at("center") : f3{};
      float radius = p.floats.count("radius") ? p.floats.at("radius") : 1.f;
      mesh = mesh_sphere(radius, 32, 16);
      transform = mul(toWorld, translation(center));
    } else if (t == "disk") {
      mesh = mesh_disk(1.f, 32);
      transform = mul(toWorld, yToZ);
    } else if (t == "rectangle") {
      mesh = mesh_quad(2.f, 2.f);
      transform = mul(toWorld, yToZ);
    } else if (t == "cube") {
      mesh = mesh_cube({2.f, 2.f, 2.f});
    } else {
      warn_once("shape type '" + t + "' is not supported");
      return;
    }
    Bsdf bsdf;
    bool hasBsdf = false;
    for (const auto& r : p.refs)
      if (r.second->cls == Cls::BSDF) {
        if (hasBsdf) error(o.line, "shape can only have one BSDF");
        hasBsdf = build_bsdf(*r.second, bsdf);
      }
    if (!hasBsdf) bsdf = Bsdf();  // a default StandardMaterial
    for (const auto& r : p.refs)
      if (r.second->cls == Cls::Medium) warn_once("participating media are not supported and ignored");
    for (const auto& r : p.refs)
      if (r.second->cls == Cls::Emitter && r.second->type == "area") {
        if (!bsdf.diffuse) error(o.line, "shape with area emitter must have a diffuse material");
        f3 rad = color(r.second->props, "radiance", {0.f, 0.f, 0.f});
        Material em;
        em.name = o.id;
        em.m.emissive[0] = rad.x;
        em.m.emissive[1] = rad.y;
        em.m.emissive[2] = rad.z;
        em.m.metallic = 0.f;
        em.m.roughness = 1.f;
        for (int c = 0; c < 3; c++) em.m.baseColor[c] = bsdf.mat.m.baseColor[c];
        bsdf.mat = em;
      }
    uint32_t mat = b_.add_material(bsdf.mat);
    uint32_t id = b_.add_mesh(std::move(mesh));
    b_.add_instance(id, transform, mat);
  }

  static m4 flip_z() { return scaling({1.f, 1.f, -1.f}); }

  void build_sensor(const Object& o) {
    const Props& p = o.props;
    m4 toWorld = p.transforms.count("to_world") ? p.transforms.at("to_world") : m4::identity();
    uint32_t W = 768, H = 576;
    for (const auto& r : p.refs)
      if (r.second->cls == Cls::Film) {
        const Props& fp = r.second->props;
        if (fp.ints.count("width")) W = (uint32_t)fp.ints.at("width");
        if (fp.ints.count("height")) H = (uint32_t)fp.ints.at("height");
      }
    b_.filmW = W;
    b_.filmH = H;
    if (o.type != "perspective" && o.type != "thinlens") {
      warn_once("sensor type '" + o.type + "' is not supported");
      return;
    }
    if (p.floats.count("focal_length") && p.floats.count("fov"))
      error(o.line, "cannot specify both 'focal_length' and 'fov'");
    Camera& c = b_.camera;
    c.present = true;
    c.focalLength = p.floats.count("focal_length") ? p.floats.at("focal_length") : 50.f;
    if (p.floats.count("fov")) {
      float filmWidth = (24.f / (float)H) * (float)W;
      c.focalLength = fov_y_to_focal_length(p.floats.at("fov") * kPi / 180.f, filmWidth);
    }
    if (p.strings.count("fov_axis")) warn_once("parameter 'fov_axis' is not supported");
    c.frameHeight = 24.f;
    c.focalDistance = p.floats.count("focus_distance") ? p.floats.at("focus_distance") : 1.f;
    c.apertureRadius = p.floats.count("aperture_radius") ? p.floats.at("aperture_radius") : 0.f;
    m4 node = mul(toWorld, flip_z());
    f3 pos{node.m[0][3], node.m[1][3], node.m[2][3]};
    f3 fwd{-node.m[0][2], -node.m[1][2], -node.m[2][2]};
    c.position = pos;
    c.target = pos + normalize(fwd);
    c.up = normalize(f3{node.m[0][1], node.m[1][1], node.m[2][1]});
  }

  static void euler_xyz_degrees(const m4& t, float out[3]) {
    float sy = std::max(-1.f, std::min(1.f, t.m[0][2]));
    float y = std::asin(sy), x, z;
    if (std::fabs(sy) < 0.9999f) {
      x = std::atan2(-t.m[1][2], t.m[2][2]);
      z = std::atan2(-t.m[0][1], t.m[0][0]);
    } else {
      x = std::atan2(t.m[2][1], t.m[1][1]);
      z = 0.f;
    }
    out[0] = x * 180.f / kPi;
    out[1] = y * 180.f / kPi;
    out[2] = z * 180.f / kPi;
  }

  void build_emitter(const Object& o) {
    const Props& p = o.props;
    m4 toWorld = p.transforms.count("to_world") ? p.transforms.at("to_world") : m4::identity();
    const float one[3] = {1.f, 1.f, 1.f}, zero[3] = {0.f, 0.f, 0.f};
    if (o.type == "area") error(o.line, "'area' emitter needs to be nested in a shape");
    if (o.type == "constant") {
      if (b_.hasEnv) error(o.line, "only one envmap can be added");
      f3 r = color(p, "radiance", {1.f, 1.f, 1.f});
      Image img;
      img.width = img.height = 1;
      img.rgb = {r.x, r.y, r.z};
      b_.set_env_rgb(img, 1.f, one, zero);
    } else if (o.type == "envmap") {
      if (b_.hasEnv) error(o.line, "only one envmap can be added");
      auto fn = p.strings.find("filename");
      if (fn == p.strings.end()) error(o.line, "envmap needs a filename");
      Image img = read_pfm(join_path(dir_, fn->second));
      float scale = p.floats.count("scale") ? p.floats.at("scale") : 1.f;
      float rot[3];
      euler_xyz_degrees(mul(toWorld, flip_z()), rot);
      b_.set_env_rgb(img, scale, one, rot);
    } else {
      warn_once("emitter type '" + o.type + "' is not supported");
    }
  }

  void build(const Object& scene) {
    for (const auto& r : scene.props.refs) {
      const Object& c = *r.second;
      switch (c.cls) {
        case Cls::Sensor: build_sensor(c); break;
        case Cls::Emitter: build_emitter(c); break;
        case Cls::Shape: build_shape(c); break;
        default: break;
      }
    }
  }
};

}  // namespace

void import_mitsuba(Builder& b, const std::string& path) {
  MitsubaScene scene(b, path);
  scene.run();
}

}  // namespace ingest
