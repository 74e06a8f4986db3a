// pbrt-v4 scene files: tokenizer, directive parser with the graphics-state stack, and the
// conversion into builder materials, meshes, camera and environment map.
//
// Parsing follows the pbrt-v4 file format as the reference's PBRTImporter reads it
// (Source/plugins/importers/PBRTImporter/Parser.cpp, Builder.cpp:217-530: the current
// transform post-multiplied by Translate / Scale / Rotate / LookAt / ConcatTransform,
// Transform and ConcatTransform transposed, the camera transform the inverse of the
// transform at Camera, WorldBegin resetting it, AttributeBegin/End saving the graphics
// state, ObjectBegin/End/ObjectInstance instancing).  The conversion follows
// PBRTImporter.cpp:589-1672 and 1794-1905 with StandardMaterial targets (the importer's
// default, usePBRTMaterials = false).  Three of its behaviours are kept as written:
//   * getRoughness (:424-463) broadcasts uroughness + vroughness into both components,
//     so a "roughness" r gives alpha r + r (sqrt(r + r) when remapped);
//   * a disk (:1449-1459) is placed by translate(height) * YtoZ only; the shape's own
//     transform is not applied;
//   * the diffuse area light's "scale" (:1388-1416) is read but not applied.
// Unsupported input is reported in the builder's log, as the reference's logWarning.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <set>
#include <sstream>
#include <variant>

#include "ingest.h"

namespace ingest {

namespace {

constexpr float kPi = 3.14159265358979323846f;

struct Loc {
  std::string file;
  int line = 0;
  std::string str() const { return file + ":" + std::to_string(line); }
};

[[noreturn]] void fail(const Loc& loc, const std::string& msg) { throw std::runtime_error(loc.str() + ": " + msg); }

// ---------------------------------------------------------------- tokenizer

struct Token {
  std::string text;
  bool quoted = false;
  Loc loc;
};

class Tokenizer {
 public:
  Tokenizer(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(path + ": cannot open");
    src_.assign((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    loc_.file = path;
    loc_.line = 1;
  }

  bool next(Token& t) {
    for (;;) {
      if (pos_ >= src_.size()) return false;
      char c = src_[pos_];
      if (c == '\n') {
        loc_.line++;
        pos_++;
      } else if (std::isspace((unsigned char)c)) {
        pos_++;
      } else if (c == '#') {
        while (pos_ < src_.size() && src_[pos_] != '\n') pos_++;
      } else {
        break;
      }
    }
    t = Token{};
    t.loc = loc_;
    char c = src_[pos_];
    if (c == '[' || c == ']') {
      t.text = std::string(1, c);
      pos_++;
      return true;
    }
    if (c == '"') {
      pos_++;
      t.quoted = true;
      while (pos_ < src_.size() && src_[pos_] != '"') {
        char d = src_[pos_++];
        if (d == '\n') fail(t.loc, "newline in a string");
        if (d == '\\' && pos_ < src_.size()) {
          char e = src_[pos_++];
          switch (e) {
            case 'n': d = '\n'; break;
            case 't': d = '\t'; break;
            default: d = e;
          }
        }
        t.text += d;
      }
      if (pos_ >= src_.size()) fail(t.loc, "unterminated string");
      pos_++;
      return true;
    }
    while (pos_ < src_.size()) {
      char d = src_[pos_];
      if (std::isspace((unsigned char)d) || d == '"' || d == '[' || d == ']' || d == '#') break;
      t.text += d;
      pos_++;
    }
    return true;
  }

 private:
  std::string src_;
  size_t pos_ = 0;
  Loc loc_;
};

// ---------------------------------------------------------------- parameters

struct Param {
  std::string type, name;
  std::vector<double> nums;
  std::vector<std::string> strs;
  Loc loc;
};

struct Params {
  std::vector<Param> list;

  const Param* find(const std::string& name) const {
    for (const Param& p : list)
      if (p.name == name) return &p;
    return nullptr;
  }
  bool has(const std::string& name) const { return find(name) != nullptr; }
  bool has_type(const std::string& name, const std::string& type) const {
    const Param* p = find(name);
    return p && p->type == type;
  }
  float get_float(const std::string& name, float def) const {
    const Param* p = find(name);
    if (!p || (p->type != "float" && p->type != "integer") || p->nums.empty()) return def;
    return (float)p->nums[0];
  }
  int get_int(const std::string& name, int def) const {
    const Param* p = find(name);
    if (!p || p->type != "integer" || p->nums.empty()) return def;
    return (int)p->nums[0];
  }
  bool get_bool(const std::string& name, bool def) const {
    const Param* p = find(name);
    if (!p || p->type != "bool" || p->strs.empty()) return def;
    return p->strs[0] == "true";
  }
  std::string get_string(const std::string& name, const std::string& def) const {
    const Param* p = find(name);
    if (!p || (p->type != "string" && p->type != "texture") || p->strs.empty()) return def;
    return p->strs[0];
  }
  std::vector<double> get_nums(const std::string& name) const {
    const Param* p = find(name);
    return p ? p->nums : std::vector<double>{};
  }
  f3 get_point3(const std::string& name, f3 def) const {
    const Param* p = find(name);
    if (!p || p->nums.size() < 3) return def;
    return {(float)p->nums[0], (float)p->nums[1], (float)p->nums[2]};
  }
  std::vector<std::string> names() const {
    std::vector<std::string> n;
    for (const Param& p : list) n.push_back(p.name);
    return n;
  }
};

// pbrt-v4 parameter types and the kind of value each holds.
bool numeric_type(const std::string& t) {
  static const std::set<std::string> k = {"integer", "float",   "point2", "vector2", "point3", "vector3", "normal",
                                          "normal3", "point",   "vector", "rgb",     "color",  "blackbody"};
  return k.count(t) != 0;
}

// ---------------------------------------------------------------- scene entities

// A material reference: an index into the unnamed materials, a name, or none (default).
struct MatRef {
  int index = -1;
  std::string name;
  bool none() const { return index < 0 && name.empty(); }
};

struct MaterialEntity {
  std::string name, type;
  Params params;
  Loc loc;
};

struct ShapeEntity {
  std::string type;
  Params params;
  m4 transform;
  bool reverse = false;
  MatRef material;
  int areaLight = -1;
  Loc loc;
};

struct AreaLightEntity {
  std::string type;
  Params params;
  Loc loc;
};

struct LightEntity {
  std::string type;
  Params params;
  m4 transform;
  Loc loc;
};

struct TextureEntity {
  std::string name, cls;
  Params params;
  Loc loc;
};

struct GraphicsState {
  m4 ctm;
  bool reverse = false;
  MatRef material;
  std::string areaLightType;
  Params areaLightParams;
  Loc areaLightLoc;
};

class PbrtScene {
 public:
  explicit PbrtScene(Builder& b) : b_(b) {}

  void parse_file(const std::string& path) {
    if (depth_ > 32) throw std::runtime_error(path + ": Include nesting too deep");
    depth_++;
    Tokenizer tk(path);
    dirs_.push_back(dirname_of(path));
    Token t;
    have_ = false;
    while (read(tk, t)) directive(tk, t);
    dirs_.pop_back();
    depth_--;
  }

  void finish() {
    if (!inWorld_) throw std::runtime_error("end of files before 'WorldBegin'");
    if (!stack_.empty()) throw std::runtime_error("missing end to AttributeBegin");
    build();
  }

 private:
  Builder& b_;
  int depth_ = 0;
  std::vector<std::string> dirs_;
  bool inWorld_ = false;
  GraphicsState gs_;
  struct StackEntry {
    bool object;
    GraphicsState state;
    Loc loc;
  };
  std::vector<StackEntry> stack_;
  std::map<std::string, m4> namedCoordSys_;
  // The token after the current directive's arguments, read ahead.
  Token ahead_;
  bool have_ = false;

  // Entities (BasicScene).
  bool hasCamera_ = false;
  std::string cameraType_;
  Params cameraParams_;
  m4 cameraTransform_;  // worldFromCamera
  Loc cameraLoc_;
  std::vector<MaterialEntity> materials_;
  std::vector<std::pair<std::string, MaterialEntity>> namedMaterials_;
  std::map<std::string, TextureEntity> floatTextures_, spectrumTextures_;
  std::vector<LightEntity> lights_;
  std::vector<AreaLightEntity> areaLights_;
  std::vector<ShapeEntity> shapes_;
  std::map<std::string, std::vector<ShapeEntity>> objects_;
  std::vector<std::pair<std::string, std::pair<m4, Loc>>> instances_;
  std::vector<ShapeEntity>* activeObject_ = nullptr;
  std::set<std::string> warned_;

  void warn(const Loc& loc, const std::string& msg) { b_.warn(loc.str() + ": " + msg); }

  bool read(Tokenizer& tk, Token& t) {
    if (have_) {
      t = ahead_;
      have_ = false;
      return true;
    }
    return tk.next(t);
  }
  void unread(const Token& t) {
    ahead_ = t;
    have_ = true;
  }

  Token expect_string(Tokenizer& tk, const Token& at) {
    Token t;
    if (!read(tk, t) || !t.quoted) fail(at.loc, "'" + at.text + "' expects a quoted string");
    return t;
  }

  double expect_number(Tokenizer& tk, const Token& at) {
    Token t;
    if (!read(tk, t) || t.quoted || t.text == "[" || t.text == "]") fail(at.loc, "'" + at.text + "' expects a number");
    char* e = nullptr;
    double v = std::strtod(t.text.c_str(), &e);
    if (!e || *e) fail(t.loc, "malformed number '" + t.text + "'");
    return v;
  }

  std::vector<double> numbers(Tokenizer& tk, const Token& at, size_t n) {
    Token t;
    bool bracket = false;
    if (read(tk, t)) {
      if (t.text == "[" && !t.quoted) bracket = true;
      else unread(t);
    }
    std::vector<double> v;
    for (size_t i = 0; i < n; i++) v.push_back(expect_number(tk, at));
    if (bracket && (!read(tk, t) || t.text != "]")) fail(at.loc, "missing ']'");
    return v;
  }

  // A parameter list: "type name" value | [ values ].
  Params params(Tokenizer& tk) {
    Params ps;
    Token t;
    while (read(tk, t)) {
      if (!t.quoted) {
        unread(t);
        break;
      }
      std::istringstream decl(t.text);
      Param p;
      p.loc = t.loc;
      decl >> p.type >> p.name;
      if (p.name.empty()) fail(t.loc, "malformed parameter declaration '" + t.text + "'");
      if (p.type == "point") p.type = "point3";
      if (p.type == "vector") p.type = "vector3";
      if (p.type == "normal") p.type = "normal3";
      if (p.type == "color") p.type = "rgb";
      Token v;
      if (!read(tk, v)) fail(t.loc, "parameter '" + p.name + "' has no value");
      std::vector<Token> vals;
      if (v.text == "[" && !v.quoted) {
        for (;;) {
          if (!read(tk, v)) fail(t.loc, "missing ']'");
          if (v.text == "]" && !v.quoted) break;
          vals.push_back(v);
        }
      } else {
        vals.push_back(v);
      }
      for (const Token& x : vals) {
        if (x.quoted) {
          p.strs.push_back(x.text);
        } else if (x.text == "true" || x.text == "false") {
          p.strs.push_back(x.text);
        } else {
          char* e = nullptr;
          double d = std::strtod(x.text.c_str(), &e);
          if (!e || *e) fail(x.loc, "malformed value '" + x.text + "'");
          p.nums.push_back(d);
        }
      }
      if (numeric_type(p.type) && !p.strs.empty()) fail(t.loc, "parameter '" + p.name + "' expects numbers");
      ps.list.push_back(std::move(p));
    }
    return ps;
  }

  void apply(const m4& m) { gs_.ctm = mul(gs_.ctm, m); }

  void directive(Tokenizer& tk, const Token& t) {
    if (t.quoted) fail(t.loc, "unexpected string \"" + t.text + "\"");
    const std::string& d = t.text;
    if (d == "Translate") {
      auto v = numbers(tk, t, 3);
      apply(translation({(float)v[0], (float)v[1], (float)v[2]}));
    } else if (d == "Scale") {
      auto v = numbers(tk, t, 3);
      apply(scaling({(float)v[0], (float)v[1], (float)v[2]}));
    } else if (d == "Rotate") {
      auto v = numbers(tk, t, 4);
      apply(rotation((float)v[0] * kPi / 180.f, {(float)v[1], (float)v[2], (float)v[3]}));
    } else if (d == "LookAt") {
      auto v = numbers(tk, t, 9);
      apply(look_at_lh({(float)v[0], (float)v[1], (float)v[2]}, {(float)v[3], (float)v[4], (float)v[5]},
                       {(float)v[6], (float)v[7], (float)v[8]}));
    } else if (d == "Transform" || d == "ConcatTransform") {
      auto v = numbers(tk, t, 16);
      float r[16];
      for (int i = 0; i < 16; i++) r[i] = (float)v[i];
      m4 m = transpose(m4::from_rows(r));  // pbrt writes matrices for row vectors
      if (d == "Transform") gs_.ctm = m;
      else apply(m);
    } else if (d == "Identity") {
      gs_.ctm = m4::identity();
    } else if (d == "CoordinateSystem") {
      namedCoordSys_[expect_string(tk, t).text] = gs_.ctm;
    } else if (d == "CoordSysTransform") {
      Token n = expect_string(tk, t);
      auto it = namedCoordSys_.find(n.text);
      if (it != namedCoordSys_.end()) gs_.ctm = it->second;
      else warn(t.loc, "couldn't find named coordinate system '" + n.text + "'");
    } else if (d == "Camera") {
      Token n = expect_string(tk, t);
      Params ps = params(tk);
      if (inWorld_) fail(t.loc, "Camera inside the world block");
      m4 worldFromCamera = inverse(gs_.ctm);
      namedCoordSys_["camera"] = worldFromCamera;
      hasCamera_ = true;
      cameraType_ = n.text;
      cameraParams_ = ps;
      cameraTransform_ = worldFromCamera;
      cameraLoc_ = t.loc;
    } else if (d == "Film") {
      expect_string(tk, t);
      Params ps = params(tk);
      b_.filmW = (uint32_t)ps.get_int("xresolution", 1280);
      b_.filmH = (uint32_t)ps.get_int("yresolution", 720);
    } else if (d == "Sampler" || d == "Integrator" || d == "PixelFilter" || d == "Accelerator" ||
               d == "ColorSpace" || d == "MakeNamedMedium") {
      expect_string(tk, t);
      if (d == "MakeNamedMedium") warn(t.loc, "participating media are not supported and ignored");
      params(tk);
    } else if (d == "Option") {
      Token n = expect_string(tk, t);
      Token v;
      read(tk, v);
      warn(t.loc, "Option '" + n.text + "' is not supported and ignored");
    } else if (d == "MediumInterface") {
      expect_string(tk, t);
      Token v;
      if (read(tk, v) && !v.quoted) unread(v);
    } else if (d == "TransformTimes") {
      numbers(tk, t, 2);
      warn(t.loc, "animated transforms are not supported; the start transform is used");
    } else if (d == "ActiveTransform") {
      Token v;
      read(tk, v);
      warn(t.loc, "ActiveTransform is not supported; all transforms are updated together");
    } else if (d == "WorldBegin") {
      inWorld_ = true;
      gs_.ctm = m4::identity();
      namedCoordSys_["world"] = gs_.ctm;
    } else if (d == "WorldEnd") {
      // pbrt-v3 form; nothing to do.
    } else if (d == "AttributeBegin" || d == "TransformBegin") {
      stack_.push_back({false, gs_, t.loc});
    } else if (d == "AttributeEnd" || d == "TransformEnd") {
      if (stack_.empty()) {
        warn(t.loc, "unmatched AttributeEnd encountered; ignoring it");
        return;
      }
      if (stack_.back().object) fail(t.loc, "mismatched nesting: open ObjectBegin at AttributeEnd");
      if (d == "TransformEnd") gs_.ctm = stack_.back().state.ctm;
      else gs_ = stack_.back().state;
      stack_.pop_back();
    } else if (d == "Attribute") {
      expect_string(tk, t);
      params(tk);
      warn(t.loc, "Attribute defaults are not supported and ignored");
    } else if (d == "ReverseOrientation") {
      gs_.reverse = !gs_.reverse;
    } else if (d == "Material") {
      Token n = expect_string(tk, t);
      Params ps = params(tk);
      materials_.push_back({"", n.text, ps, t.loc});
      gs_.material = MatRef{(int)materials_.size() - 1, ""};
    } else if (d == "MakeNamedMaterial") {
      Token n = expect_string(tk, t);
      Params ps = params(tk);
      std::string type = ps.get_string("type", "");
      for (const auto& nm : namedMaterials_)
        if (nm.first == n.text) fail(t.loc, "named material '" + n.text + "' redefined");
      namedMaterials_.push_back({n.text, MaterialEntity{n.text, type, ps, t.loc}});
    } else if (d == "NamedMaterial") {
      Token n = expect_string(tk, t);
      gs_.material = MatRef{-1, n.text};
    } else if (d == "Texture") {
      Token n = expect_string(tk, t);
      Token ty = expect_string(tk, t);
      Token cls = expect_string(tk, t);
      Params ps = params(tk);
      TextureEntity te{n.text, cls.text, ps, t.loc};
      if (ty.text == "float") floatTextures_[n.text] = te;
      else if (ty.text == "spectrum" || ty.text == "color") spectrumTextures_[n.text] = te;
      else fail(t.loc, "texture type must be 'float' or 'spectrum'");
    } else if (d == "LightSource") {
      Token n = expect_string(tk, t);
      Params ps = params(tk);
      lights_.push_back({n.text, ps, gs_.ctm, t.loc});
    } else if (d == "AreaLightSource") {
      Token n = expect_string(tk, t);
      gs_.areaLightType = n.text;
      gs_.areaLightParams = params(tk);
      gs_.areaLightLoc = t.loc;
    } else if (d == "Shape") {
      Token n = expect_string(tk, t);
      Params ps = params(tk);
      if (!inWorld_) fail(t.loc, "Shape outside the world block");
      ShapeEntity s{n.text, ps, gs_.ctm, gs_.reverse, gs_.material, -1, t.loc};
      if (!gs_.areaLightType.empty()) {
        if (activeObject_) warn(t.loc, "area lights not supported with object instancing");
        areaLights_.push_back({gs_.areaLightType, gs_.areaLightParams, gs_.areaLightLoc});
        s.areaLight = (int)areaLights_.size() - 1;
      }
      if (activeObject_) activeObject_->push_back(std::move(s));
      else shapes_.push_back(std::move(s));
    } else if (d == "ObjectBegin") {
      Token n = expect_string(tk, t);
      if (activeObject_) fail(t.loc, "ObjectBegin called inside of instance definition");
      stack_.push_back({true, gs_, t.loc});
      if (objects_.count(n.text)) fail(t.loc, "object '" + n.text + "' redefined");
      activeObject_ = &objects_[n.text];
    } else if (d == "ObjectEnd") {
      if (!activeObject_) fail(t.loc, "ObjectEnd called outside of instance definition");
      if (stack_.empty() || !stack_.back().object) fail(t.loc, "mismatched nesting at ObjectEnd");
      gs_ = stack_.back().state;
      stack_.pop_back();
      activeObject_ = nullptr;
    } else if (d == "ObjectInstance") {
      Token n = expect_string(tk, t);
      if (activeObject_) fail(t.loc, "ObjectInstance can't be called inside instance definition");
      instances_.push_back({n.text, {gs_.ctm, t.loc}});
    } else if (d == "Include" || d == "Import") {
      Token n = expect_string(tk, t);
      Token saved;
      bool hadAhead = have_;
      if (hadAhead) saved = ahead_;
      have_ = false;
      parse_file(join_path(dirs_.back(), n.text));
      if (hadAhead) unread(saved);
    } else {
      fail(t.loc, "unknown directive '" + d + "'");
    }
  }

  // ------------------------------------------------------------ conversion

  void warn_unsupported_params(const Params& ps, const std::vector<std::string>& names) {
    for (const std::string& n : names)
      if (const Param* p = ps.find(n)) warn(p->loc, "parameter '" + n + "' is not supported and ignored");
  }

  // An RGB value of a spectrum parameter (rgb; blackbody and sampled or named spectra are
  // not converted here: reported, and the default is used).
  f3 spectrum_rgb(const Params& ps, const std::string& name, f3 def) {
    const Param* p = ps.find(name);
    if (!p) return def;
    if (p->type == "rgb" && p->nums.size() >= 3) return {(float)p->nums[0], (float)p->nums[1], (float)p->nums[2]};
    if (p->type == "float" && !p->nums.empty()) {
      float v = (float)p->nums[0];
      return {v, v, v};
    }
    warn(p->loc, "spectrum '" + name + "' of type '" + p->type + "' is not converted to RGB here; using " +
                     std::to_string(def.x) + "," + std::to_string(def.y) + "," + std::to_string(def.z));
    return def;
  }

  // A spectrum texture or constant: the constant value, or the default with a warning for
  // a texture that is not constant (no texture mapping in the modelled materials).
  f3 spectrum_constant(const Params& ps, const std::string& name, f3 def) {
    const Param* p = ps.find(name);
    if (p && p->type == "texture") {
      std::string tex = p->strs.empty() ? "" : p->strs[0];
      auto it = spectrumTextures_.find(tex);
      if (it == spectrumTextures_.end()) fail(p->loc, "cannot find spectrum texture named '" + tex + "'");
      if (it->second.cls == "constant") return spectrum_rgb(it->second.params, "value", {1.f, 1.f, 1.f});
      warn(p->loc, "texture '" + tex + "' (" + it->second.cls + ") is not supported; using a constant value");
      return def;
    }
    return spectrum_rgb(ps, name, def);
  }

  bool float_param(const Params& ps, const std::string& name, float& out) {
    const Param* p = ps.find(name);
    if (!p) return false;
    if (p->type == "texture") {
      std::string tex = p->strs.empty() ? "" : p->strs[0];
      auto it = floatTextures_.find(tex);
      if (it == floatTextures_.end()) fail(p->loc, "cannot find float texture named '" + tex + "'");
      if (it->second.cls == "constant") {
        out = it->second.params.get_float("value", 1.f);
        return true;
      }
      warn(p->loc, "non-constant roughness is not supported; using 0.5");
      out = -1.f;  // marks the non-constant fallback
      return true;
    }
    out = ps.get_float(name, 0.f);
    return true;
  }

  // getRoughness / getScalarRoughness (PBRTImporter.cpp:424-479): NDF alpha.
  float scalar_roughness(const Params& ps, const std::string& r = "roughness", const std::string& ur = "uroughness",
                         const std::string& vr = "vroughness") {
    float u = 0.f, v = 0.f;
    if (!float_param(ps, ur, u)) float_param(ps, r, u);
    if (!float_param(ps, vr, v)) float_param(ps, r, v);
    if (u < 0.f || v < 0.f) return 0.5f;
    float a = u + v;  // the importer's float2{u + v} broadcast
    if (ps.get_bool("remaproughness", true)) a = std::sqrt(a);
    return a;
  }

  float scalar_eta(const Params& ps, const std::string& name = "eta") {
    const Param* p = ps.find(name);
    if (!p) return 1.5f;
    if (p->type == "float" && !p->nums.empty()) return (float)p->nums[0];
    f3 rgb = spectrum_rgb(ps, name, {1.5f, 1.5f, 1.5f});
    return (rgb.x + rgb.y + rgb.z) / 3.f;
  }

  // getConductorEtaK + getConductorSpecularAlbedo (:517-587): normal-incidence Fresnel.
  f3 conductor_albedo(const Params& ps, const Loc& loc, const std::string& refl = "reflectance",
                      const std::string& etaN = "eta", const std::string& kN = "k") {
    bool hasR = ps.has(refl), hasEta = ps.has(etaN), hasK = ps.has(kN);
    if (hasR && (hasEta || hasK)) fail(loc, "both '" + refl + "' and '" + etaN + "' and '" + kN + "' can't be provided");
    f3 eta, k;
    if (hasR) {
      f3 r = spectrum_constant(ps, refl, {0.5f, 0.5f, 0.5f});
      auto clampr = [](float x) { return std::min(std::max(x, 0.f), 0.9999f); };
      r = {clampr(r.x), clampr(r.y), clampr(r.z)};
      eta = {1.f, 1.f, 1.f};
      k = {2.f * std::sqrt(r.x) / std::sqrt(1.f - r.x), 2.f * std::sqrt(r.y) / std::sqrt(1.f - r.y),
           2.f * std::sqrt(r.z) / std::sqrt(1.f - r.z)};
    } else {
      // The reference's defaults are the named spectra metal-Cu-eta / metal-Cu-k converted
      // to RGB; the commonly used RGB values of copper stand in for that conversion.
      const f3 kCuEta{0.200438f, 0.924033f, 1.10221f}, kCuK{3.91295f, 2.45285f, 2.14219f};
      if (!hasEta || !hasK) warn(loc, "conductor without '" + etaN + "'/'" + kN + "': using RGB copper constants");
      eta = hasEta ? spectrum_constant(ps, etaN, kCuEta) : kCuEta;
      k = hasK ? spectrum_constant(ps, kN, kCuK) : kCuK;
    }
    return fresnel_conductor(eta, k, 1.f);
  }

  // createMaterial (:975-1352), StandardMaterial targets.  Returns false for the
  // 'interface' material (no surface).
  bool convert_material(const MaterialEntity& e, Material& mat, bool isAreaLight) {
    const Params& ps = e.params;
    const std::string& type = e.type;
    mat = Material();
    mat.name = e.name;
    rs_material& m = mat.m;
    auto setBase = [&](f3 c) {
      m.baseColor[0] = c.x;
      m.baseColor[1] = c.y;
      m.baseColor[2] = c.z;
    };
    (void)isAreaLight;
    if (type == "" || type == "none" || type == "interface") {
      if (type != "interface") warn(e.loc, "material type '" + type + "' is deprecated, use 'interface' instead");
      return false;
    } else if (type == "diffuse") {
      warn_unsupported_params(ps, {"displacement"});
      setBase(spectrum_constant(ps, "reflectance", {0.5f, 0.5f, 0.5f}));
      m.metallic = 0.f;
      m.roughness = 1.f;
      m.doubleSided = 1;
    } else if (type == "coateddiffuse") {
      warn_unsupported_params(ps, {"thickness", "eta", "maxdepth", "nsamples", "g", "albedo", "displacement"});
      setBase(spectrum_constant(ps, "reflectance", {0.5f, 0.5f, 0.5f}));
      m.metallic = 0.f;
      m.roughness = std::sqrt(scalar_roughness(ps));
      m.doubleSided = 1;
    } else if (type == "conductor") {
      warn_unsupported_params(ps, {"displacement"});
      setBase(conductor_albedo(ps, e.loc));
      m.metallic = 1.f;
      m.roughness = std::sqrt(scalar_roughness(ps));
      m.doubleSided = 1;
    } else if (type == "coatedconductor") {
      warn_unsupported_params(ps, {"interface.roughness", "interface.uroughness", "interface.vroughness",
                                   "interface.eta", "maxdepth", "nsamples", "g", "albedo", "displacement"});
      setBase(conductor_albedo(ps, e.loc, "conductor.reflectance", "conductor.eta", "conductor.k"));
      m.metallic = 1.f;
      m.roughness =
          std::sqrt(scalar_roughness(ps, "conductor.roughness", "conductor.uroughness", "conductor.vroughness"));
      m.doubleSided = 1;
    } else if (type == "dielectric" || type == "thindielectric") {
      warn_unsupported_params(ps, {"displacement"});
      m.metallic = 0.f;
      m.roughness = type == "dielectric" ? std::sqrt(scalar_roughness(ps)) : 0.f;
      m.ior = scalar_eta(ps);
      // StandardMaterial specularTransmission = 1 in the reference; the modelled BSDF is
      // opaque, so the surface keeps the dielectric's specular reflection only.
      setBase({0.f, 0.f, 0.f});
      warn(e.loc, "'" + type + "': specular transmission is not modelled; imported as its specular reflection");
    } else if (type == "diffusetransmission") {
      warn_unsupported_params(ps, {"displacement", "scale"});
      setBase(spectrum_constant(ps, "reflectance", {0.25f, 0.25f, 0.25f}));
      m.metallic = 0.f;
      m.roughness = 1.f;
      m.doubleSided = 1;
      warn(e.loc, "'diffusetransmission': diffuse transmission is not modelled");
    } else if (type == "hair" || type == "measured" || type == "subsurface" || type == "mix") {
      warn(e.loc, "material type '" + type + "' is not supported; using the default material");
      m.doubleSided = 1;
    } else {
      fail(e.loc, "unknown material type '" + type + "'");
    }
    if (ps.has("normalmap")) warn(e.loc, "normal maps are not supported and ignored");
    return true;
  }

  // Material ids in the builder, created in the reference's order: named, then unnamed.
  std::map<std::string, int> namedIds_;
  std::vector<int> unnamedIds_;
  int defaultId_ = -1;

  int material_id(const MatRef& ref) {
    int id = -2;
    if (ref.index >= 0) id = unnamedIds_[(size_t)ref.index];
    else if (!ref.name.empty()) {
      auto it = namedIds_.find(ref.name);
      if (it == namedIds_.end()) throw std::runtime_error("named material '" + ref.name + "' not defined");
      id = it->second;
    }
    if (id >= 0) return id;
    // No material, or 'interface' (no surface): getMaterial (:233-260) falls back to a
    // double-sided default StandardMaterial.
    if (defaultId_ < 0) {
      Material d;
      d.name = "Default";
      d.m.doubleSided = 1;
      defaultId_ = (int)b_.add_material(d);
    }
    return defaultId_;
  }

  const MaterialEntity* material_entity(const MatRef& ref) {
    if (ref.index >= 0) return &materials_[(size_t)ref.index];
    for (const auto& nm : namedMaterials_)
      if (nm.first == ref.name) return &nm.second;
    return nullptr;
  }

  // createAreaLight (:1388-1416) on a fresh material of the shape (createShape :1645-1664).
  int area_light_material(const ShapeEntity& s) {
    const AreaLightEntity& al = areaLights_[(size_t)s.areaLight];
    if (al.type != "diffuse") fail(al.loc, "unknown area light type '" + al.type + "'");
    warn_unsupported_params(al.params, {"twosided", "power", "filename"});
    Material mat;
    const MaterialEntity* me = s.material.none() ? nullptr : material_entity(s.material);
    if (me) {
      if (!convert_material(*me, mat, true)) {
        warn(s.loc, "area light on an 'interface' material: emissive default material used");
        mat = Material();
        mat.m.doubleSided = 1;
      }
    } else {
      mat.name = "Emissive" + std::to_string(s.areaLight);
      mat.m.baseColor[0] = mat.m.baseColor[1] = mat.m.baseColor[2] = 0.f;
      mat.m.roughness = 0.f;
    }
    f3 L = spectrum_rgb(al.params, "L", {1.f, 1.f, 1.f});
    mat.m.emissive[0] = L.x;
    mat.m.emissive[1] = L.y;
    mat.m.emissive[2] = L.z;
    return (int)b_.add_material(mat);
  }

  static void apply_transform(Mesh& mesh, const m4& t) {
    m4 nrm = transpose(inverse(t));
    for (auto& p : mesh.positions) p = xform_point(t, p);
    for (auto& n : mesh.normals) n = normalize(xform_vector(nrm, n));
    if (det3(t) < 0.f) mesh.frontFaceCW = !mesh.frontFaceCW;
  }

  // createShape (:1418-1672): an object-space mesh and its transform, or false.
  bool convert_shape(const ShapeEntity& s, Mesh& mesh, m4& transform) {
    const Params& ps = s.params;
    warn_unsupported_params(ps, {"alpha"});
    transform = s.transform;
    if (s.type == "sphere") {
      warn_unsupported_params(ps, {"zmin", "zmax", "phimax"});
      mesh = mesh_sphere(ps.get_float("radius", 1.f), 32, 16);
    } else if (s.type == "disk") {
      warn_unsupported_params(ps, {"innerradius", "phimax"});
      mesh = mesh_disk(ps.get_float("radius", 1.f), 32);
      m4 yToZ;
      yToZ.m[1][1] = 0.f;
      yToZ.m[1][2] = 1.f;
      yToZ.m[2][1] = 1.f;
      yToZ.m[2][2] = 0.f;
      apply_transform(mesh, mul(translation({0.f, 0.f, ps.get_float("height", 0.f)}), yToZ));
      transform = m4::identity();  // the importer does not set the shape transform here
    } else if (s.type == "trianglemesh") {
      warn_unsupported_params(ps, {"S", "faceIndices"});
      std::vector<double> idx = ps.get_nums("indices"), P = ps.get_nums("P"), N = ps.get_nums("N");
      if (P.empty()) {
        warn(s.loc, "vertex positions 'P' missing; skipping");
        return false;
      }
      size_t nv = P.size() / 3;
      if (idx.empty()) {
        if (nv == 3) idx = {0, 1, 2};
        else {
          warn(s.loc, "vertex indices 'indices' missing; skipping");
          return false;
        }
      }
      if (idx.size() % 3) {
        warn(s.loc, "number of vertex indices is not a multiple of 3; discarding the remainder");
        idx.resize(idx.size() - idx.size() % 3);
      }
      if (!N.empty() && N.size() != P.size()) {
        warn(s.loc, "number of 'N' elements must match number of 'P' elements; discarding 'N'");
        N.clear();
      }
      for (double i : idx)
        if (i < 0 || i >= (double)nv) {
          warn(s.loc, "vertex index out of bounds; skipping");
          return false;
        }
      mesh.name = "trianglemesh";
      for (size_t v = 0; v < nv; v++) mesh.positions.push_back({(float)P[3 * v], (float)P[3 * v + 1], (float)P[3 * v + 2]});
      for (size_t v = 0; v < N.size() / 3; v++) mesh.normals.push_back({(float)N[3 * v], (float)N[3 * v + 1], (float)N[3 * v + 2]});
      for (double i : idx) mesh.indices.push_back((uint32_t)i);
      if (mesh.normals.empty()) {
        // The reference stores zero normals here (NaN shading); flat normals instead.
        warn(s.loc, "trianglemesh without 'N': using flat normals");
        flat_normals(mesh);
      }
    } else if (s.type == "plymesh") {
      warn_unsupported_params(ps, {"displacement", "displacement.edgelength"});
      std::string file = ps.get_string("filename", "");
      mesh = read_mesh_file(join_path(sceneDir_, file));
      if (mesh.normals.size() != mesh.positions.size()) flat_normals(mesh);
      mesh.name = file;
    } else if (s.type == "loopsubdiv") {
      warn_unsupported_params(ps, {"scheme"});
      std::vector<double> idx = ps.get_nums("indices"), P = ps.get_nums("P");
      if (idx.empty()) fail(s.loc, "missing vertex indices in 'indices'");
      if (P.empty()) fail(s.loc, "missing vertex positions in 'P'");
      std::vector<f3> pos;
      for (size_t v = 0; v + 2 < P.size(); v += 3) pos.push_back({(float)P[v], (float)P[v + 1], (float)P[v + 2]});
      std::vector<uint32_t> ind;
      for (double i : idx) ind.push_back((uint32_t)i);
      mesh = loop_subdivide(ps.get_int("levels", 3), pos, ind);
    } else if (s.type == "cylinder" || s.type == "bilinearmesh" || s.type == "curve") {
      warn(s.loc, "shape type '" + s.type + "' is not supported");
      return false;
    } else {
      fail(s.loc, "unknown shape type '" + s.type + "'");
    }
    if (s.reverse) mesh.frontFaceCW = !mesh.frontFaceCW;
    return true;
  }

  std::string sceneDir_;

 public:
  void set_scene_dir(const std::string& d) { sceneDir_ = d; }
  static Mesh loop_subdivide(int levels, const std::vector<f3>& P, const std::vector<uint32_t>& indices);

 private:
  void build() {
    // Materials: named first, then unnamed (buildScene :1807-1813).
    for (const auto& nm : namedMaterials_) {
      Material mat;
      namedIds_[nm.first] = convert_material(nm.second, mat, false) ? (int)b_.add_material(mat) : -1;
    }
    for (const auto& me : materials_) {
      Material mat;
      unnamedIds_.push_back(convert_material(me, mat, false) ? (int)b_.add_material(mat) : -1);
    }
    // Camera (createCamera :589-641, node transform worldFromCamera * InvertZ :1816-1822).
    if (hasCamera_) {
      if (cameraType_ == "perspective") {
        warn_unsupported_params(cameraParams_, {"frameaspectratio", "screenwindow"});
        Camera& c = b_.camera;
        c.present = true;
        c.apertureRadius = cameraParams_.get_float("lensradius", 0.f);
        c.focalDistance = cameraParams_.get_float("focaldistance", 1e6f);
        c.frameHeight = 24.f;
        c.focalLength = fov_y_to_focal_length(cameraParams_.get_float("fov", 90.f) * kPi / 180.f, 24.f);
        m4 node = mul(cameraTransform_, scaling({1.f, 1.f, -1.f}));
        f3 pos{node.m[0][3], node.m[1][3], node.m[2][3]};
        f3 fwd{-node.m[0][2], -node.m[1][2], -node.m[2][2]};
        c.position = pos;
        c.target = pos + normalize(fwd);
        c.up = normalize(f3{node.m[0][1], node.m[1][1], node.m[2][1]});
      } else if (cameraType_ == "orthographic" || cameraType_ == "realistic" || cameraType_ == "spherical") {
        warn(cameraLoc_, "camera type '" + cameraType_ + "' is not supported");
      } else {
        fail(cameraLoc_, "unknown camera type '" + cameraType_ + "'");
      }
    }
    // Lights (createLight :643-750).
    for (const LightEntity& l : lights_) {
      if (l.type == "infinite") {
        warn_unsupported_params(l.params, {"portal", "illuminance"});
        const Param* L = l.params.find("L");
        std::string file = l.params.get_string("filename", "");
        float scale = l.params.get_float("scale", 1.f);
        if (L && !file.empty()) fail(l.loc, "can't specify both emission 'L' and 'filename' for infinite light");
        if (b_.hasEnv) {
          warn(l.loc, "no support for multiple infinite lights; discarding this light");
          continue;
        }
        const float one[3] = {1.f, 1.f, 1.f}, zero[3] = {0.f, 0.f, 0.f};
        if (L) {
          // A constant infinite light: a one-texel map (its 'scale' is not applied).
          f3 c = spectrum_rgb(l.params, "L", {1.f, 1.f, 1.f});
          Image img;
          img.width = img.height = 1;
          img.rgb = {c.x, c.y, c.z};
          b_.set_env_rgb(img, 1.f, one, zero);
        } else if (!file.empty()) {
          Image img = read_pfm(join_path(sceneDir_, file));
          warn(l.loc, "environment map is converted from equal-area octahedral to lat-long parametrization");
          Image ll = octahedral_to_latlong(img, 2 * img.width);
          float rot[3];
          euler_xyz_degrees(l.transform, rot);
          b_.set_env_rgb(ll, scale, one, rot);
        }
      } else if (l.type == "distant" || l.type == "point" || l.type == "spot" || l.type == "goniometric" ||
                 l.type == "projection") {
        warn(l.loc, "light type '" + l.type + "' (an analytic light) is not modelled by the pass; ignored");
      } else {
        fail(l.loc, "unknown light type '" + l.type + "'");
      }
    }
    // Shapes (:1846-1856) and instances (:1883-1903).
    for (const ShapeEntity& s : shapes_) add_shape(s, m4::identity());
    std::map<std::string, std::vector<std::pair<uint32_t, std::pair<m4, uint32_t>>>> defs;
    for (const auto& inst : instances_) {
      auto it = objects_.find(inst.first);
      if (it == objects_.end()) fail(inst.second.second, "object instance '" + inst.first + "' not defined");
      auto d = defs.find(inst.first);
      if (d == defs.end()) {
        std::vector<std::pair<uint32_t, std::pair<m4, uint32_t>>> meshes;
        for (const ShapeEntity& s : it->second) {
          Mesh mesh;
          m4 tr;
          if (!convert_shape(s, mesh, tr)) continue;
          int mat = s.areaLight >= 0 ? area_light_material(s) : material_id(s.material);
          meshes.push_back({b_.add_mesh(std::move(mesh)), {tr, (uint32_t)mat}});
        }
        d = defs.emplace(inst.first, std::move(meshes)).first;
      }
      for (const auto& mm : d->second) b_.add_instance(mm.first, mul(inst.second.first, mm.second.first), mm.second.second);
    }
  }

  void add_shape(const ShapeEntity& s, const m4& parent) {
    Mesh mesh;
    m4 tr;
    if (!convert_shape(s, mesh, tr)) return;
    int mat = s.areaLight >= 0 ? area_light_material(s) : material_id(s.material);
    uint32_t id = b_.add_mesh(std::move(mesh));
    b_.add_instance(id, mul(parent, tr), (uint32_t)mat);
  }

  // extractEulerAngleXYZ of the light transform, in degrees (EnvMap rotation order XYZ).
  static void euler_xyz_degrees(const m4& t, float out[3]) {
    // R = Rx * Ry * Rz (column vectors): r02 = sin(y).
    float sy = std::max(-1.f, std::min(1.f, t.m[0][2]));
    float y = std::asin(sy);
    float x, z;
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
};

}  // namespace

Mesh loop_subdivide_mesh(int levels, const std::vector<f3>& P, const std::vector<uint32_t>& indices);

Mesh PbrtScene::loop_subdivide(int levels, const std::vector<f3>& P, const std::vector<uint32_t>& indices) {
  return loop_subdivide_mesh(levels, P, indices);
}

void import_pbrt(Builder& b, const std::string& path) {
  PbrtScene scene(b);
  scene.set_scene_dir(dirname_of(path));
  scene.parse_file(path);
  scene.finish();
}

}  // namespace ingest
