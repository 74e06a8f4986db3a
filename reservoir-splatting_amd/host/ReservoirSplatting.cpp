// C++ host mirror of the reference's ReservoirSplatting render pass (see ReservoirSplatting.h).
#include "ReservoirSplatting.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace rs_amd {

namespace {

// Property keys of the pass (RS/ReservoirSplatting.cpp:99-131).
const std::string kSamplesPerPixel = "samplesPerPixel";
const std::string kMaxSurfaceBounces = "maxSurfaceBounces";
const std::string kMaxDiffuseBounces = "maxDiffuseBounces";
const std::string kMaxSpecularBounces = "maxSpecularBounces";
const std::string kMaxTransmissionBounces = "maxTransmissionBounces";
const std::string kFixedSeed = "fixedSeed";
const std::string kUseRussianRoulette = "useRussianRoulette";
const std::string kMISHeuristic = "misHeuristic";
const std::string kMISPowerExponent = "misPowerExponent";
const std::string kAdjustShadingNormals = "adjustShadingNormals";
const std::string kSpecularRoughnessThreshold = "specularRoughnessThreshold";
const std::string kPrimaryLodMode = "primaryLodMode";
const std::string kEnableTemporalResampling = "enableTemporalResampling";
const std::string kEnableSpatialResampling = "enableSpatialResampling";
const std::string kTemporalReuse = "temporalReuse";
const std::string kScatterBackupMISOption = "scatterBackupMISOption";
const std::string kGatherOption = "gatherOption";
const std::string kNumTimePartitions = "numTimePartitions";
// Options the reference sets from its UI (ReSTIRParams, CameraParams, RS/ReservoirSplatting.h:
// 164-184) and this build's own (quirk bits, stage capture), accepted as properties too.
const std::string kNumSpatialIterations = "numSpatialIterations";
const std::string kNeighborCount = "neighborCount";
const std::string kSpatialGatherRadius = "spatialGatherRadius";
const std::string kArtificialFrameTime = "artificialFrameTime";
const std::string kUseConfidenceTemporally = "useConfidenceWeightsTemporally";
const std::string kUseConfidenceSpatially = "useConfidenceWeightsSpatially";
const std::string kGBufferAdjustShadingNormals = "gbufferAdjustShadingNormals";
const std::string kCaptureStages = "captureStages";
const std::string kQuirkMask = "quirkMask";
// Keys the reference reads that have no effect on this path (sample generator and emissive
// sampler choice, alpha test, nested dielectrics, texture LOD, output sizing, colour format).
const char* const kInertKeys[] = {"sampleGenerator",  "emissiveSampler",    "lightBVHOptions",
                                  "useAlphaTest",     "maxNestedMaterials", "useLightsInDielectricVolumes",
                                  "disableCaustics",  "lodBias",            "outputSize",
                                  "fixedOutputSize",  "colorFormat"};

constexpr uint32_t kMaxSamplesPerPixel = 16;  // RS/Params.slang:68
constexpr uint32_t kMaxBounces = 254;         // RS/Params.slang:70

// FALCOR_ENUM_INFO names (RS/Params.slang:55-59, RS/ShiftOptions.slang:42-83).
struct EnumName {
  const char* name;
  uint32_t value;
};
const EnumName kMISNames[] = {{"Balance", 0}, {"PowerTwo", 1}, {"PowerExp", 2}};
const EnumName kTemporalNames[] = {{"GatherOnly", RS_TEMPORAL_GATHER_ONLY},
                                   {"ScatterOnly", RS_TEMPORAL_SCATTER_ONLY},
                                   {"ScatterBackup", RS_TEMPORAL_SCATTER_BACKUP},
                                   {"MultiScatter", RS_TEMPORAL_MULTI_SCATTER}};
const EnumName kBackupNames[] = {{"Balance", 0}, {"Pairwise", 1}};
const EnumName kGatherNames[] = {{"Fast", RS_GATHER_FAST}, {"Clamped", RS_GATHER_CLAMPED}, {"Robust", RS_GATHER_ROBUST}};

std::string kind_of(const Properties::Value& v) {
  switch (v.index()) {
    case 0: return "bool";
    case 1: return "integer";
    case 2: return "number";
    default: return "string";
  }
}

uint32_t to_uint(const std::string& key, const Properties::Value& v) {
  if (const bool* b = std::get_if<bool>(&v)) return *b ? 1u : 0u;
  if (const int64_t* i = std::get_if<int64_t>(&v)) {
    if (*i < 0 || *i > (int64_t)std::numeric_limits<uint32_t>::max())
      throw Error(RS_E_INVALID_ARG, "Property '" + key + "' is out of the range of an unsigned integer.");
    return (uint32_t)*i;
  }
  if (const double* d = std::get_if<double>(&v)) {
    if (!(*d >= 0.0) || *d > (double)std::numeric_limits<uint32_t>::max() || std::floor(*d) != *d)
      throw Error(RS_E_INVALID_ARG, "Property '" + key + "' is not an unsigned integer.");
    return (uint32_t)*d;
  }
  throw Error(RS_E_INVALID_ARG, "Property '" + key + "' expects a number, got a " + kind_of(v) + ".");
}

float to_float(const std::string& key, const Properties::Value& v) {
  if (const double* d = std::get_if<double>(&v)) return (float)*d;
  if (const int64_t* i = std::get_if<int64_t>(&v)) return (float)*i;
  throw Error(RS_E_INVALID_ARG, "Property '" + key + "' expects a number, got a " + kind_of(v) + ".");
}

bool to_bool(const std::string& key, const Properties::Value& v) {
  if (const bool* b = std::get_if<bool>(&v)) return *b;
  if (const int64_t* i = std::get_if<int64_t>(&v)) return *i != 0;
  throw Error(RS_E_INVALID_ARG, "Property '" + key + "' expects a bool, got a " + kind_of(v) + ".");
}

template <size_t N>
uint32_t to_enum(const std::string& key, const Properties::Value& v, const EnumName (&names)[N]) {
  if (const std::string* s = std::get_if<std::string>(&v)) {
    for (const EnumName& e : names)
      if (*s == e.name) return e.value;
    throw Error(RS_E_INVALID_ARG, "Invalid enum value '" + *s + "' for property '" + key + "'.");
  }
  const uint32_t x = to_uint(key, v);
  for (const EnumName& e : names)
    if (x == e.value) return x;
  throw Error(RS_E_INVALID_ARG, "Invalid enum value " + std::to_string(x) + " for property '" + key + "'.");
}

template <size_t N>
std::string enum_name(uint32_t value, const EnumName (&names)[N]) {
  for (const EnumName& e : names)
    if (e.value == value) return e.name;
  return std::to_string(value);
}

}  // namespace

// ------------------------------------------------------------------ Properties

Properties::Properties(std::initializer_list<std::pair<std::string, Value>> items) {
  for (const auto& kv : items) (*this)[kv.first] = kv.second;
}

bool Properties::has(const std::string& key) const {
  for (const auto& kv : mItems)
    if (kv.first == key) return true;
  return false;
}

Properties::Value& Properties::operator[](const std::string& key) {
  for (auto& kv : mItems)
    if (kv.first == key) return kv.second;
  mItems.emplace_back(key, Value(false));
  return mItems.back().second;
}

const Properties::Value& Properties::at(const std::string& key) const {
  for (const auto& kv : mItems)
    if (kv.first == key) return kv.second;
  throw Error(RS_E_INVALID_ARG, "Property '" + key + "' does not exist.");
}

// ------------------------------------------------------------------ ReservoirSplatting

ReservoirSplatting::ReservoirSplatting(const Properties& props, Logger log) : mLog(std::move(log)) {
  rs_default_options(&mOptions);
  const int32_t code = rs_create(&mOptions, &mCtx);
  if (code != RS_OK && code != RS_E_BAD_OPTION) {
    if (mCtx) rs_destroy(mCtx);
    mCtx = nullptr;
    throw Error(code, "ReservoirSplatting: rs_create failed");
  }
  rs_get_options(mCtx, &mOptions);
  if (props.size()) setProperties(props);
}

ReservoirSplatting::~ReservoirSplatting() {
  if (mCtx) rs_destroy(mCtx);
}

void ReservoirSplatting::warn(const std::string& msg) {
  mWarnings.push_back(msg);
  if (mLog) mLog(msg);
}

void ReservoirSplatting::check(int32_t code) const {
  if (code != RS_OK) throw Error(code, rs_last_error(mCtx));
}

// rs_set_options, whose own clamps (partition count, quirk bits) come back as warnings.
void ReservoirSplatting::apply(const rs_options& o) {
  const int32_t code = rs_set_options(mCtx, &o);
  if (code == RS_E_BAD_OPTION) warn(rs_last_error(mCtx));
  else check(code);
  rs_get_options(mCtx, &mOptions);
}

void ReservoirSplatting::setProperties(const Properties& props) {
  // parseProperties (RS/ReservoirSplatting.cpp:192-258).
  rs_options o = mOptions;
  bool rayCones = false;
  for (const auto& [key, value] : props) {
    if (key == kSamplesPerPixel) o.samplesPerPixel = to_uint(key, value);
    else if (key == kMaxSurfaceBounces) o.maxSurfaceBounces = to_uint(key, value);
    else if (key == kMaxDiffuseBounces) o.maxDiffuseBounces = to_uint(key, value);
    else if (key == kMaxSpecularBounces) o.maxSpecularBounces = to_uint(key, value);
    else if (key == kMaxTransmissionBounces) o.maxTransmissionBounces = to_uint(key, value);
    else if (key == kFixedSeed) {
      o.fixedSeed = to_uint(key, value);
      o.useFixedSeed = 1;
    } else if (key == kUseRussianRoulette) o.useRussianRoulette = to_bool(key, value);
    else if (key == kMISHeuristic) o.misHeuristic = to_enum(key, value, kMISNames);
    else if (key == kMISPowerExponent) o.misPowerExponent = to_float(key, value);
    else if (key == kAdjustShadingNormals) o.adjustShadingNormals = to_bool(key, value);
    else if (key == kSpecularRoughnessThreshold) o.specularRoughnessThreshold = to_float(key, value);
    else if (key == kPrimaryLodMode) {
      const std::string* s = std::get_if<std::string>(&value);
      rayCones = s && *s == "RayCones";
    } else if (key == kEnableTemporalResampling) o.enableTemporalResampling = to_bool(key, value);
    else if (key == kEnableSpatialResampling) o.enableSpatialResampling = to_bool(key, value);
    else if (key == kTemporalReuse) o.temporalReuse = to_enum(key, value, kTemporalNames);
    else if (key == kScatterBackupMISOption) o.scatterBackupMIS = to_enum(key, value, kBackupNames);
    else if (key == kGatherOption) o.gatherMechanism = to_enum(key, value, kGatherNames);
    else if (key == kNumTimePartitions) o.numTimePartitions = to_uint(key, value);
    else if (key == kNumSpatialIterations) o.numSpatialIterations = to_uint(key, value);
    else if (key == kNeighborCount) o.neighborCount = to_uint(key, value);
    else if (key == kSpatialGatherRadius) o.spatialGatherRadius = to_float(key, value);
    else if (key == kArtificialFrameTime) o.artificialFrameTime = to_float(key, value);
    else if (key == kUseConfidenceTemporally) o.useConfidenceWeightsTemporally = to_bool(key, value);
    else if (key == kUseConfidenceSpatially) o.useConfidenceWeightsSpatially = to_bool(key, value);
    else if (key == kGBufferAdjustShadingNormals) o.gbufferAdjustShadingNormals = to_bool(key, value);
    else if (key == kCaptureStages) o.captureStages = to_bool(key, value);
    else if (key == kQuirkMask) o.quirkMask = to_uint(key, value);
    else if (std::find(std::begin(kInertKeys), std::end(kInertKeys), key) != std::end(kInertKeys)) continue;
    else warn("Unknown property '" + key + "' in ReservoirSplatting properties.");
  }
  if (props.has(kMaxSurfaceBounces)) {
    // Bounce counts not given explicitly follow maxSurfaceBounces.
    if (!props.has(kMaxDiffuseBounces)) o.maxDiffuseBounces = o.maxSurfaceBounces;
    if (!props.has(kMaxSpecularBounces)) o.maxSpecularBounces = o.maxSurfaceBounces;
    if (!props.has(kMaxTransmissionBounces)) o.maxTransmissionBounces = o.maxSurfaceBounces;
  } else {
    o.maxSurfaceBounces = std::max(o.maxDiffuseBounces, std::max(o.maxSpecularBounces, o.maxTransmissionBounces));
  }
  const bool raise = o.maxSurfaceBounces < o.maxDiffuseBounces || o.maxSurfaceBounces < o.maxSpecularBounces ||
                     o.maxSurfaceBounces < o.maxTransmissionBounces;
  if (props.has(kMaxSurfaceBounces) && raise)
    warn("'" + kMaxSurfaceBounces + "' is set lower than '" + kMaxDiffuseBounces + "', '" + kMaxSpecularBounces +
         "' or '" + kMaxTransmissionBounces + "' and will be increased.");

  // validateOptions (:260-301), each clamp with the reference's warning.
  if (o.specularRoughnessThreshold < 0.f || o.specularRoughnessThreshold > 1.f) {
    warn("'specularRoughnessThreshold' has invalid value. Clamping to range [0,1].");
    o.specularRoughnessThreshold = std::clamp(o.specularRoughnessThreshold, 0.f, 1.f);
  }
  if (o.samplesPerPixel < 1 || o.samplesPerPixel > kMaxSamplesPerPixel) {
    warn("'samplesPerPixel' must be in the range [1, " + std::to_string(kMaxSamplesPerPixel) +
         "]. Clamping to this range.");
    o.samplesPerPixel = std::clamp(o.samplesPerPixel, 1u, kMaxSamplesPerPixel);
  }
  auto clampBounces = [&](uint32_t& bounces, const std::string& name) {
    if (bounces > kMaxBounces) {
      warn("'" + name + "' exceeds the maximum supported bounces. Clamping to " + std::to_string(kMaxBounces) + ".");
      bounces = kMaxBounces;
    }
  };
  clampBounces(o.maxSurfaceBounces, kMaxSurfaceBounces);
  clampBounces(o.maxDiffuseBounces, kMaxDiffuseBounces);
  clampBounces(o.maxSpecularBounces, kMaxSpecularBounces);
  clampBounces(o.maxTransmissionBounces, kMaxTransmissionBounces);
  o.maxSurfaceBounces =
      std::max(o.maxSurfaceBounces, std::max(o.maxDiffuseBounces, std::max(o.maxSpecularBounces, o.maxTransmissionBounces)));
  if (rayCones) warn("Unsupported tex lod mode. Defaulting to Mip0.");
  apply(o);
}

Properties ReservoirSplatting::getProperties() const {
  const rs_options& o = mOptions;
  Properties p;
  p[kSamplesPerPixel] = (int64_t)o.samplesPerPixel;
  p[kMaxSurfaceBounces] = (int64_t)o.maxSurfaceBounces;
  p[kMaxDiffuseBounces] = (int64_t)o.maxDiffuseBounces;
  p[kMaxSpecularBounces] = (int64_t)o.maxSpecularBounces;
  p[kMaxTransmissionBounces] = (int64_t)o.maxTransmissionBounces;
  if (o.useFixedSeed) p[kFixedSeed] = (int64_t)o.fixedSeed;
  p[kUseRussianRoulette] = o.useRussianRoulette != 0;
  p[kMISHeuristic] = enum_name(o.misHeuristic, kMISNames);
  p[kMISPowerExponent] = (double)o.misPowerExponent;
  p[kAdjustShadingNormals] = o.adjustShadingNormals != 0;
  p[kSpecularRoughnessThreshold] = (double)o.specularRoughnessThreshold;
  p[kEnableTemporalResampling] = o.enableTemporalResampling != 0;
  p[kEnableSpatialResampling] = o.enableSpatialResampling != 0;
  p[kTemporalReuse] = enum_name(o.temporalReuse, kTemporalNames);
  p[kScatterBackupMISOption] = enum_name(o.scatterBackupMIS, kBackupNames);
  p[kGatherOption] = enum_name(o.gatherMechanism, kGatherNames);
  p[kNumTimePartitions] = (int64_t)o.numTimePartitions;
  p[kNumSpatialIterations] = (int64_t)o.numSpatialIterations;
  p[kNeighborCount] = (int64_t)o.neighborCount;
  p[kSpatialGatherRadius] = (double)o.spatialGatherRadius;
  p[kArtificialFrameTime] = (double)o.artificialFrameTime;
  if (o.quirkMask != RS_QUIRKS_REFERENCE) p[kQuirkMask] = (int64_t)o.quirkMask;
  return p;
}

void ReservoirSplatting::setScene(const rs_scene_desc& scene) { check(rs_set_scene(mCtx, &scene)); }
void ReservoirSplatting::setEnvMap(const rs_env_map_desc& envMap) { check(rs_set_env_map(mCtx, &envMap)); }
void ReservoirSplatting::clearEnvMap() { check(rs_clear_env_map(mCtx)); }

void ReservoirSplatting::execute(const rs_camera& camera, const rs_frame_io& io, void* stream) {
  check(rs_render_frame(mCtx, &camera, &io, stream));
}

void ReservoirSplatting::reset() { check(rs_reset(mCtx)); }

PixelStats ReservoirSplatting::getPixelStats() const {
  rs_stats s{};
  check(rs_get_stats(mCtx, &s));
  return PixelStats{s.raysClosestHit, s.raysVisibility, s.raysClosestHit + s.raysVisibility};
}

bool ReservoirSplatting::getUseFixedSeed() const { return mOptions.useFixedSeed != 0; }

void ReservoirSplatting::setUseFixedSeed(bool value) {
  rs_options o = mOptions;
  o.useFixedSeed = value ? 1u : 0u;
  apply(o);
}

uint32_t ReservoirSplatting::getFixedSeed() const { return mOptions.fixedSeed; }

void ReservoirSplatting::setFixedSeed(uint32_t value) {
  rs_options o = mOptions;
  o.fixedSeed = value;
  apply(o);
}

rs_frame_flags ReservoirSplatting::getDictionary() const {
  rs_frame_flags f{};
  check(rs_get_frame_flags(mCtx, &f));
  return f;
}

}  // namespace rs_amd
