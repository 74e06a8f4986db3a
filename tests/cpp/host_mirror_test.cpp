// Tests of the C++ host mirror of the ReservoirSplatting pass (rs_amd::ReservoirSplatting).
//
// CPU cases check the pass's property surface against the reference's parseProperties /
// validateOptions / getProperties (RS/ReservoirSplatting.cpp:182-344): defaults, bounce
// defaulting, every clamp and its warning, unknown and inert keys, enum names, the fixed
// seed, round trips and the errors execute raises.  GPU cases render the golden Cornell box
// through the mirror and compare every reservoir, reconnection record, colour and ray count
// with the oracle, bit for bit.
//
// Usage: host_mirror_test cpu | gpu | <case name>   (exit status 0 when every case passes)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/rs.h"
#include "../../include/rs_import.h"
#include "../../reservoir-splatting_amd/host/ReservoirSplatting.h"

// The oracle's C API (oracle/rs_oracle.cpp), linked as the checker.
extern "C" {
typedef struct rs_cpu_ctx rs_cpu_ctx;
typedef struct rs_cpu_frame_extra {
  const uint32_t* vbuffer;
  uint32_t vbufferPrimitiveIndexBits;
  uint32_t sceneFlags;
  uint32_t* rayCount;
  uint32_t* pathLength;
} rs_cpu_frame_extra;
int32_t rs_cpu_create(const rs_options* opts, rs_cpu_ctx** out);
void rs_cpu_destroy(rs_cpu_ctx* h);
int32_t rs_cpu_set_scene(rs_cpu_ctx* h, const rs_scene_desc* sd);
int32_t rs_cpu_render_frame(rs_cpu_ctx* h, const rs_camera* cam, uint32_t W, uint32_t H, const float* mvecIn,
                            float* colorOut, const rs_cpu_frame_extra* ex);
int32_t rs_cpu_read_buffer(rs_cpu_ctx* h, int which, void* host, size_t* bytes);
int32_t rs_cpu_get_stats(rs_cpu_ctx* h, rs_stats* out);
}

using rs_amd::Error;
using rs_amd::Properties;
using rs_amd::ReservoirSplatting;

namespace {

int g_failures = 0;
std::string g_case;

#define CHECK(cond)                                                                        \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      std::printf("  %s:%d: %s: CHECK(%s) failed\n", __FILE__, __LINE__, g_case.c_str(), #cond); \
      g_failures++;                                                                        \
    }                                                                                      \
  } while (0)

template <class F>
int32_t thrown_code(F&& f) {
  try {
    f();
  } catch (const Error& e) {
    return e.code();
  }
  return RS_OK;
}

bool has_warning(const ReservoirSplatting& rs, const std::string& text) {
  for (const std::string& w : rs.warnings())
    if (w == text) return true;
  return false;
}

int64_t as_int(const Properties::Value& v) { return std::get<int64_t>(v); }
const std::string& as_str(const Properties::Value& v) { return std::get<std::string>(v); }

// ------------------------------------------------------------------ CPU cases

void cpu_defaults() {
  // The header defaults (RS/ReservoirSplatting.h:124-184).
  ReservoirSplatting rs;
  const Properties p = rs.getProperties();
  CHECK(as_int(p.at("samplesPerPixel")) == 1);
  CHECK(as_int(p.at("maxSurfaceBounces")) == 9);
  CHECK(as_int(p.at("maxDiffuseBounces")) == 9);
  CHECK(std::get<bool>(p.at("useRussianRoulette")));
  CHECK(as_str(p.at("misHeuristic")) == "Balance");
  CHECK(as_str(p.at("temporalReuse")) == "GatherOnly");
  CHECK(as_str(p.at("gatherOption")) == "Fast");
  CHECK(as_str(p.at("scatterBackupMISOption")) == "Balance");
  CHECK(as_int(p.at("numTimePartitions")) == 2);
  CHECK(std::fabs(std::get<double>(p.at("specularRoughnessThreshold")) - 0.25) < 1e-7);
  CHECK(!p.has("fixedSeed"));
  CHECK(!rs.getUseFixedSeed());
  CHECK(rs.warnings().empty());
}

void cpu_script_properties() {
  // The graph script's pass (scripts/ReservoirSplatting.py:17-20) with its ReSTIR keys.
  ReservoirSplatting rs({{"samplesPerPixel", int64_t(1)},
                         {"maxSurfaceBounces", int64_t(6)},
                         {"temporalReuse", std::string("ScatterBackup")},
                         {"gatherOption", std::string("Clamped")},
                         {"scatterBackupMISOption", std::string("Pairwise")},
                         {"misHeuristic", std::string("PowerExp")},
                         {"misPowerExponent", 3.0},
                         {"numTimePartitions", int64_t(8)}});
  const rs_options& o = rs.options();
  CHECK(o.temporalReuse == RS_TEMPORAL_SCATTER_BACKUP);
  CHECK(o.gatherMechanism == RS_GATHER_CLAMPED);
  CHECK(o.scatterBackupMIS == RS_BACKUP_MIS_PAIRWISE);
  CHECK(o.misHeuristic == RS_MIS_POWER_EXP && o.misPowerExponent == 3.f);
  CHECK(o.numTimePartitions == 8);
  // maxSurfaceBounces given alone sets the three bounce kinds (:234-240)
  CHECK(o.maxDiffuseBounces == 6 && o.maxSpecularBounces == 6 && o.maxTransmissionBounces == 6);
  CHECK(rs.warnings().empty());
}

void cpu_bounce_defaulting() {
  // Without maxSurfaceBounces it becomes the largest of the three (:241-245).
  ReservoirSplatting a({{"maxDiffuseBounces", int64_t(3)}, {"maxSpecularBounces", int64_t(12)},
                        {"maxTransmissionBounces", int64_t(5)}});
  CHECK(a.options().maxSurfaceBounces == 12);
  CHECK(a.warnings().empty());
  // Given lower than another kind: warned, then raised in validateOptions (:247-256, :284-286).
  ReservoirSplatting b({{"maxSurfaceBounces", int64_t(2)}, {"maxDiffuseBounces", int64_t(5)}});
  CHECK(b.options().maxSurfaceBounces == 5 && b.options().maxDiffuseBounces == 5);
  CHECK(b.options().maxSpecularBounces == 2 && b.options().maxTransmissionBounces == 2);
  CHECK(has_warning(b, "'maxSurfaceBounces' is set lower than 'maxDiffuseBounces', 'maxSpecularBounces' or "
                       "'maxTransmissionBounces' and will be increased."));
}

void cpu_clamps() {
  // validateOptions (:260-301): each clamp applied and warned.
  std::vector<std::string> logged;
  ReservoirSplatting rs({{"samplesPerPixel", int64_t(40)},
                         {"specularRoughnessThreshold", 1.5},
                         {"maxSurfaceBounces", int64_t(300)},
                         {"maxDiffuseBounces", int64_t(300)}},
                        [&](const std::string& m) { logged.push_back(m); });
  const rs_options& o = rs.options();
  CHECK(o.samplesPerPixel == 16);
  CHECK(o.specularRoughnessThreshold == 1.f);
  CHECK(o.maxSurfaceBounces == 254 && o.maxDiffuseBounces == 254);
  CHECK(has_warning(rs, "'samplesPerPixel' must be in the range [1, 16]. Clamping to this range."));
  CHECK(has_warning(rs, "'specularRoughnessThreshold' has invalid value. Clamping to range [0,1]."));
  CHECK(has_warning(rs, "'maxSurfaceBounces' exceeds the maximum supported bounces. Clamping to 254."));
  CHECK(has_warning(rs, "'maxDiffuseBounces' exceeds the maximum supported bounces. Clamping to 254."));
  CHECK(logged == rs.warnings());
  ReservoirSplatting z({{"samplesPerPixel", int64_t(0)}, {"specularRoughnessThreshold", -0.5}});
  CHECK(z.options().samplesPerPixel == 1 && z.options().specularRoughnessThreshold == 0.f);
  // the library's own clamp (partition count) arrives as a warning too
  ReservoirSplatting k({{"numTimePartitions", int64_t(500)}});
  CHECK(k.options().numTimePartitions == 64);
  CHECK(k.warnings().size() == 1);
}

void cpu_unknown_and_inert_keys() {
  ReservoirSplatting rs({{"hybridMISOption", std::string("Pairwise")},  // the script's misspelt key (quirk Q4)
                         {"sampleGenerator", int64_t(1)},
                         {"useAlphaTest", true},
                         {"colorFormat", std::string("RGBA32F")},
                         {"primaryLodMode", std::string("RayCones")}});
  CHECK(rs.warnings().size() == 2);
  CHECK(has_warning(rs, "Unknown property 'hybridMISOption' in ReservoirSplatting properties."));
  CHECK(has_warning(rs, "Unsupported tex lod mode. Defaulting to Mip0."));
  CHECK(rs.options().scatterBackupMIS == RS_BACKUP_MIS_BALANCE);  // the misspelt key changed nothing
}

void cpu_enum_errors() {
  CHECK(thrown_code([] { ReservoirSplatting rs({{"temporalReuse", std::string("MultiSplat")}}); }) ==
        RS_E_INVALID_ARG);
  CHECK(thrown_code([] { ReservoirSplatting rs({{"gatherOption", int64_t(7)}}); }) == RS_E_INVALID_ARG);
  CHECK(thrown_code([] { ReservoirSplatting rs({{"samplesPerPixel", std::string("four")}}); }) == RS_E_INVALID_ARG);
  CHECK(thrown_code([] { ReservoirSplatting rs({{"samplesPerPixel", int64_t(-1)}}); }) == RS_E_INVALID_ARG);
  // enums by value as well as by name
  ReservoirSplatting rs({{"temporalReuse", int64_t(RS_TEMPORAL_MULTI_SCATTER)}});
  CHECK(rs.options().temporalReuse == RS_TEMPORAL_MULTI_SCATTER);
  // a failed setProperties leaves the options as they were
  CHECK(thrown_code([&] { rs.setProperties({{"samplesPerPixel", int64_t(2)}, {"misHeuristic", std::string("x")}}); }) ==
        RS_E_INVALID_ARG);
  CHECK(rs.options().samplesPerPixel == 1);
}

void cpu_fixed_seed() {
  // fixedSeed as a property turns useFixedSeed on (:207); the pybind properties (:142-156).
  ReservoirSplatting rs({{"fixedSeed", int64_t(7)}});
  CHECK(rs.getUseFixedSeed() && rs.getFixedSeed() == 7);
  CHECK(as_int(rs.getProperties().at("fixedSeed")) == 7);
  rs.setUseFixedSeed(false);
  CHECK(!rs.getUseFixedSeed() && !rs.getProperties().has("fixedSeed"));
  rs.setFixedSeed(11);
  rs.setUseFixedSeed(true);
  CHECK(rs.options().useFixedSeed == 1 && rs.options().fixedSeed == 11);
}

void cpu_round_trip() {
  ReservoirSplatting a({{"temporalReuse", std::string("MultiScatter")},
                        {"numTimePartitions", int64_t(4)},
                        {"neighborCount", int64_t(3)},
                        {"numSpatialIterations", int64_t(2)},
                        {"spatialGatherRadius", 12.5},
                        {"artificialFrameTime", 1.0 / 48.0},
                        {"fixedSeed", int64_t(5)},
                        {"quirkMask", int64_t(0)}});
  ReservoirSplatting b(a.getProperties());
  CHECK(b.warnings().empty());
  CHECK(std::memcmp(&a.options(), &b.options(), sizeof(rs_options)) == 0);
}

void cpu_execute_errors() {
  ReservoirSplatting rs;
  rs_camera cam{};
  cam.focalLength = 21.f;
  cam.frameHeight = 24.f;
  cam.aspectRatio = 1.f;
  cam.focalDistance = 10.f;
  cam.target[2] = -1.f;
  rs_frame_io io{};
  io.width = 5000;
  io.height = 16;
  // setFrameDim (:364): frames above 4096 pixels per side are refused
  try {
    rs.execute(cam, io);
    CHECK(false);
  } catch (const Error& e) {
    CHECK(e.code() == RS_E_FRAME_TOO_LARGE);
    CHECK(std::string(e.what()) == "Frame dimensions up to 4096 pixels width/height are supported.");
  }
  io.width = 16;
  CHECK(thrown_code([&] { rs.execute(cam, io); }) == RS_E_NO_SCENE);
}

// ------------------------------------------------------------------ GPU cases

struct Cornell {
  rs_builder* b = nullptr;
  rs_scene_desc desc{};
  rs_import_camera cam{};
  Cornell(const std::string& root) {
    b = rs_builder_create();
    if (rs_builder_import(b, (root + "/tests/golden/scenes/cornell.pbrt").c_str()) != RS_OK ||
        rs_builder_get_scene(b, &desc) != RS_OK || rs_builder_get_camera(b, &cam) != RS_OK)
      throw Error(RS_E_INVALID_ARG, std::string("cornell.pbrt: ") + rs_builder_log(b));
  }
  ~Cornell() { rs_builder_destroy(b); }
  rs_camera frame_camera(uint32_t frame, uint32_t W, uint32_t H, float shutter) const {
    rs_camera c{};
    for (int i = 0; i < 3; i++) {
      c.position[i] = cam.position[i];
      c.target[i] = cam.target[i];
    }
    c.position[0] += 0.01f * (float)frame;  // the camera slides, so temporal reuse reprojects
    c.focalLength = cam.focalLength;
    c.frameHeight = cam.frameHeight;
    c.aspectRatio = (float)W / (float)H;
    c.focalDistance = cam.focalDistance;
    c.apertureRadius = cam.apertureRadius;
    c.shutterSpeed = shutter;
    return c;
  }
};

std::vector<uint8_t> gpu_buffer(const ReservoirSplatting& rs, int which, size_t bytes) {
  std::vector<uint8_t> v(bytes);
  size_t n = bytes;
  if (rs_read_buffer(rs.context(), which, v.data(), &n) != RS_OK || n != bytes) v.clear();
  return v;
}

std::vector<uint8_t> cpu_buffer(rs_cpu_ctx* c, int which, size_t bytes) {
  std::vector<uint8_t> v(bytes);
  size_t n = bytes;
  if (rs_cpu_read_buffer(c, which, v.data(), &n) != RS_OK || n != bytes) v.clear();
  return v;
}

std::string g_root = ".";

// Frames of the Cornell box through the mirror against the oracle, bit for bit.
void gpu_frames(const Properties& props, float shutter) {
  const uint32_t W = 64, H = 48;
  Cornell scene(g_root);
  ReservoirSplatting rs(props);
  rs.setScene(scene.desc);
  rs_cpu_ctx* cpu = nullptr;
  CHECK(rs_cpu_create(&rs.options(), &cpu) == RS_OK);
  CHECK(rs_cpu_set_scene(cpu, &scene.desc) == RS_OK);
  for (uint32_t f = 0; f < 4; f++) {
    const rs_camera cam = scene.frame_camera(f, W, H, shutter);
    rs_frame_io io{};
    io.width = W;
    io.height = H;
    rs.execute(cam, io);
    CHECK(rs_cpu_render_frame(cpu, &cam, W, H, nullptr, nullptr, nullptr) == RS_OK);
    const struct {
      int which;
      size_t bpp;
    } bufs[] = {{RS_BUF_HISTORY_RESERVOIRS, 32}, {RS_BUF_HISTORY_RECON, 112}, {RS_BUF_COLOR, 16},
                {RS_BUF_MVEC, 8},                {RS_BUF_VBUFFER, 12}};
    for (const auto& b : bufs) {
      const std::vector<uint8_t> g = gpu_buffer(rs, b.which, b.bpp * W * H), c = cpu_buffer(cpu, b.which, b.bpp * W * H);
      CHECK(!g.empty() && g == c);
      if (g != c) std::printf("    frame %u, buffer %d differs\n", f, b.which);
    }
    rs_stats cs{};
    rs_cpu_get_stats(cpu, &cs);
    const rs_amd::PixelStats ps = rs.getPixelStats();
    CHECK(ps.closestHitRays == cs.raysClosestHit && ps.visibilityRays == cs.raysVisibility);
    CHECK(ps.totalRays == ps.closestHitRays + ps.visibilityRays && ps.totalRays > 0);
    const rs_frame_flags d = rs.getDictionary();
    CHECK(d.suppressAutoReset == 0 && d.shouldAccumulate == 1);
  }
  rs_cpu_destroy(cpu);
}

void gpu_cornell_scatter_backup() {
  gpu_frames({{"temporalReuse", std::string("ScatterBackup")}, {"gatherOption", std::string("Clamped")},
              {"maxSurfaceBounces", int64_t(4)}},
             0.f);
}

void gpu_cornell_gather_robust() {
  gpu_frames({{"temporalReuse", std::string("GatherOnly")}, {"gatherOption", std::string("Robust")},
              {"maxSurfaceBounces", int64_t(4)}},
             0.f);
}

void gpu_cornell_multi_scatter() {
  gpu_frames({{"temporalReuse", std::string("MultiScatter")}, {"numTimePartitions", int64_t(3)},
              {"maxSurfaceBounces", int64_t(4)}, {"fixedSeed", int64_t(3)}},
             1.f / 24.f);
}

void gpu_reset_restarts() {
  // reset() (:813-820): the next frame starts over; two passes reset at the same frame agree.
  const uint32_t W = 32, H = 32;
  Cornell scene(g_root);
  Properties props{{"temporalReuse", std::string("ScatterOnly")}, {"maxSurfaceBounces", int64_t(3)}};
  ReservoirSplatting a(props), b(props);
  a.setScene(scene.desc);
  b.setScene(scene.desc);
  rs_frame_io io{};
  io.width = W;
  io.height = H;
  for (uint32_t f = 0; f < 3; f++) a.execute(scene.frame_camera(f, W, H, 0.f), io);
  a.reset();
  a.setFixedSeed(9);
  a.setUseFixedSeed(true);
  b.setFixedSeed(9);
  b.setUseFixedSeed(true);
  a.execute(scene.frame_camera(5, W, H, 0.f), io);
  b.execute(scene.frame_camera(5, W, H, 0.f), io);
  CHECK(gpu_buffer(a, RS_BUF_HISTORY_RESERVOIRS, 32 * W * H) == gpu_buffer(b, RS_BUF_HISTORY_RESERVOIRS, 32 * W * H));
  CHECK(a.getPixelStats().totalRays == b.getPixelStats().totalRays);
}

struct Case {
  const char* name;
  const char* kind;
  void (*fn)();
};

const Case kCases[] = {
    {"defaults", "cpu", cpu_defaults},
    {"script_properties", "cpu", cpu_script_properties},
    {"bounce_defaulting", "cpu", cpu_bounce_defaulting},
    {"clamps", "cpu", cpu_clamps},
    {"unknown_and_inert_keys", "cpu", cpu_unknown_and_inert_keys},
    {"enum_errors", "cpu", cpu_enum_errors},
    {"fixed_seed", "cpu", cpu_fixed_seed},
    {"round_trip", "cpu", cpu_round_trip},
    {"execute_errors", "cpu", cpu_execute_errors},
    {"cornell_scatter_backup", "gpu", gpu_cornell_scatter_backup},
    {"cornell_gather_robust", "gpu", gpu_cornell_gather_robust},
    {"cornell_multi_scatter", "gpu", gpu_cornell_multi_scatter},
    {"reset_restarts", "gpu", gpu_reset_restarts},
};

}  // namespace

int main(int argc, char** argv) {
  const std::string which = argc > 1 ? argv[1] : "cpu";
  if (argc > 2) g_root = argv[2];
  int run = 0;
  for (const Case& c : kCases) {
    if (which != c.kind && which != c.name) continue;
    g_case = c.name;
    const int before = g_failures;
    std::printf("case %s (%s)\n", c.name, c.kind);
    std::fflush(stdout);
    try {
      c.fn();
    } catch (const std::exception& e) {
      std::printf("  %s: unexpected exception: %s\n", c.name, e.what());
      g_failures++;
    }
    std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", c.name);
    std::fflush(stdout);
    run++;
  }
  if (run == 0) {
    std::printf("no case matches '%s'\n", which.c_str());
    return 2;
  }
  return g_failures == 0 ? 0 : 1;
}
