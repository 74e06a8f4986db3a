// Driver of the oracle for its sanitizer build (make -C oracle sanitize): every temporal mode,
// gather mechanism and option branch, the frozen scene, instance motion, a caller VBuffer,
// the environment map, per-pixel stats, the FP64 replay and the accumulator, on a small
// procedural room, so that AddressSanitizer and UndefinedBehaviorSanitizer watch every code
// path of the oracle.  Test infrastructure only; exits non-zero on any failed call.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/rs.h"

extern "C" {
typedef struct rs_cpu_ctx rs_cpu_ctx;
typedef struct rs_cpu_frame_extra {
  const uint32_t* vbuffer;
  uint32_t vbufferPrimitiveIndexBits;
  uint32_t sceneFlags;
  uint32_t* rayCount;
  uint32_t* pathLength;
} rs_cpu_frame_extra;
int32_t rs_cpu_create(const rs_options*, rs_cpu_ctx**);
void rs_cpu_destroy(rs_cpu_ctx*);
int32_t rs_cpu_set_threads(int);
int32_t rs_cpu_set_scene(rs_cpu_ctx*, const rs_scene_desc*);
int32_t rs_cpu_update_instances(rs_cpu_ctx*, const float*, uint32_t);
int32_t rs_cpu_set_frozen(rs_cpu_ctx*, int);
int32_t rs_cpu_set_env_map(rs_cpu_ctx*, const rs_env_map_desc*);
int32_t rs_cpu_reset(rs_cpu_ctx*);
int32_t rs_cpu_render_frame(rs_cpu_ctx*, const rs_camera*, uint32_t, uint32_t, const float*, float*,
                            const rs_cpu_frame_extra*);
int32_t rs_cpu_read_buffer(rs_cpu_ctx*, int, void*, size_t*);
int32_t rs_cpu_scatter_counts(rs_cpu_ctx*, float*, size_t*);
int32_t rs_cpu_replay(rs_cpu_ctx*, const double*, const float*, double*, double*, double*);
int32_t rs_cpu_accum_create(const rs_accum_options*, void**);
void rs_cpu_accum_destroy(void*);
int32_t rs_cpu_accum_execute(void*, const float*, float*, uint32_t, uint32_t, int, int);
}

namespace {

int g_failures = 0;

void check(int32_t code, const char* what) {
  if (code != RS_OK) {
    std::printf("FAIL %s: %d\n", what, code);
    g_failures++;
  }
}

struct Soup {
  std::vector<float> P, N;
  std::vector<uint32_t> I, M;
  std::vector<rs_instance_desc> inst;
  void quad(const float o[3], const float u[3], const float v[3], uint32_t mat, uint32_t n) {
    rs_instance_desc d{(uint32_t)(P.size() / 3), (n + 1) * (n + 1), (uint32_t)(I.size() / 3), 2 * n * n};
    float nr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    float l = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    for (uint32_t j = 0; j <= n; j++)
      for (uint32_t i = 0; i <= n; i++) {
        float a = (float)i / n, b = (float)j / n;
        for (int k = 0; k < 3; k++) {
          P.push_back(o[k] + a * u[k] + b * v[k]);
          N.push_back(nr[k] / l);
        }
      }
    for (uint32_t j = 0; j < n; j++)
      for (uint32_t i = 0; i < n; i++) {
        uint32_t a = d.firstVertex + j * (n + 1) + i, b = a + 1, c = a + n + 1, e = c + 1;
        uint32_t t[6] = {a, b, e, a, e, c};
        for (uint32_t x : t) I.push_back(x);
        M.push_back(mat);
        M.push_back(mat);
      }
    inst.push_back(d);
  }
};

rs_material material(float r, float g, float b, float rough, float emit, float metal) {
  rs_material m{};
  m.baseColor[0] = r;
  m.baseColor[1] = g;
  m.baseColor[2] = b;
  m.roughness = rough;
  m.emissive[0] = m.emissive[1] = m.emissive[2] = emit;
  m.metallic = metal;
  m.ior = 1.5f;
  m.activeLobes = 0xff;
  m.doubleSided = 1;
  return m;
}

rs_camera camera(uint32_t f, uint32_t W, uint32_t H, float aperture, float shutter) {
  rs_camera c{};
  c.position[0] = 0.02f * f;
  c.position[2] = 3.5f;
  c.target[0] = 0.02f * f;
  c.focalLength = 0.5f * 24.f / std::tan(20.f * 3.14159265f / 180.f);
  c.frameHeight = 24.f;
  c.aspectRatio = (float)W / (float)H;
  c.focalDistance = 3.5f;
  c.apertureRadius = aperture;
  c.shutterSpeed = shutter;
  return c;
}

}  // namespace

int main() {
  rs_cpu_set_threads(2);
  Soup s;
  const float floorO[3] = {-1, -1, 1}, floorU[3] = {0, 0, -2}, floorV[3] = {2, 0, 0};
  const float backO[3] = {-1, -1, -1}, backU[3] = {0, 2, 0}, backV[3] = {2, 0, 0};
  const float leftO[3] = {-1, -1, 1}, leftU[3] = {0, 2, 0}, leftV[3] = {0, 0, -2};
  const float boxO[3] = {-0.3f, -1, -0.3f}, boxU[3] = {0, 0.8f, 0}, boxV[3] = {0.6f, 0, 0};
  const float lightO[3] = {-0.25f, 0.999f, -0.25f}, lightU[3] = {0.5f, 0, 0}, lightV[3] = {0, 0, 0.5f};
  s.quad(floorO, floorU, floorV, 0, 3);
  s.quad(backO, backU, backV, 1, 3);
  s.quad(leftO, leftU, leftV, 0, 2);
  s.quad(boxO, boxU, boxV, 2, 2);
  s.quad(lightO, lightU, lightV, 3, 1);
  std::vector<rs_material> mats = {material(0.6f, 0.5f, 0.4f, 1.f, 0.f, 0.f), material(0.4f, 0.6f, 0.5f, 0.2f, 0.f, 0.f),
                                   material(0.8f, 0.8f, 0.8f, 0.35f, 0.f, 0.6f), material(0.5f, 0.5f, 0.5f, 1.f, 15.f, 0.f)};
  rs_scene_desc sd{};
  sd.positions = s.P.data();
  sd.normals = s.N.data();
  sd.numVertices = (uint32_t)(s.P.size() / 3);
  sd.indices = s.I.data();
  sd.triMaterial = s.M.data();
  sd.numTriangles = (uint32_t)(s.I.size() / 3);
  sd.materials = mats.data();
  sd.numMaterials = (uint32_t)mats.size();
  sd.diffuseModel = RS_DIFFUSE_FROSTBITE;
  sd.instances = s.inst.data();
  sd.numInstances = (uint32_t)s.inst.size();

  const uint32_t W = 20, H = 14, P = W * H;
  std::vector<float> color(4 * P), accOut(4 * P);
  std::vector<uint32_t> rays(P), lengths(P), vb(4 * P);
  std::vector<float> env(16 * 8 * 4, 0.5f);
  struct Variant {
    uint32_t mode, gather, backupMIS, mis, spp, neighbours, iterations, fixedSeed, confidence, rr, quirks;
    float aperture, shutter;
  };
  const Variant variants[] = {
      {0, 0, 0, 0, 1, 1, 1, 0, 1, 1, RS_QUIRKS_REFERENCE, 0.f, 0.f},
      {0, 1, 0, 1, 2, 3, 2, 0, 1, 1, RS_QUIRKS_REFERENCE, 0.05f, 0.f},
      {0, 2, 0, 0, 1, 1, 1, 0, 0, 1, 0u, 0.05f, 0.f},
      {1, 1, 0, 2, 1, 0, 1, 1, 1, 0, RS_QUIRKS_REFERENCE, 0.f, 0.f},
      {2, 1, 1, 0, 3, 2, 2, 0, 1, 1, RS_QUIRKS_REFERENCE, 0.05f, 0.f},
      {2, 2, 0, 0, 1, 1, 1, 0, 1, 1, 0u, 0.f, 0.f},
      {2, 0, 1, 0, 1, 1, 1, 0, 1, 1, RS_QUIRKS_REFERENCE, 0.f, 0.f},
      {3, 1, 0, 0, 1, 1, 1, 0, 1, 1, RS_QUIRKS_REFERENCE, 0.04f, 1.f / 24.f},
      {3, 1, 0, 0, 2, 2, 1, 0, 1, 1, 0u, 0.f, 1.f / 30.f},
  };
  int v = 0;
  for (const Variant& x : variants) {
    rs_options o{};
    o.samplesPerPixel = x.spp;
    o.maxSurfaceBounces = o.maxDiffuseBounces = o.maxSpecularBounces = o.maxTransmissionBounces = 4;
    o.useRussianRoulette = x.rr;
    o.misHeuristic = x.mis;
    o.misPowerExponent = 2.f;
    o.specularRoughnessThreshold = 0.25f;
    o.gbufferAdjustShadingNormals = 1;
    o.enableTemporalResampling = 1;
    o.temporalReuse = x.mode;
    o.gatherMechanism = x.gather;
    o.scatterBackupMIS = x.backupMIS;
    o.numTimePartitions = 3;
    o.useConfidenceWeightsTemporally = x.confidence;
    o.enableSpatialResampling = 1;
    o.numSpatialIterations = x.iterations;
    o.neighborCount = x.neighbours;
    o.useConfidenceWeightsSpatially = x.confidence;
    o.spatialGatherRadius = 30.f;
    o.artificialFrameTime = 1.f / 60.f;
    o.useFixedSeed = x.fixedSeed;
    o.fixedSeed = 3;
    o.captureStages = 1;
    o.quirkMask = x.quirks;
    rs_cpu_ctx* c = nullptr;
    check(rs_cpu_create(&o, &c), "create");
    check(rs_cpu_set_scene(c, &sd), "set_scene");
    if (v % 2) {
      rs_env_map_desc e{env.data(), 16, 8, 1.f, {1.f, 1.f, 1.f}, {0.f, 30.f, 0.f}};
      check(rs_cpu_set_env_map(c, &e), "set_env_map");
    }
    for (uint32_t f = 0; f < 6; f++) {
      std::vector<float> X(12 * s.inst.size(), 0.f);
      for (size_t k = 0; k < s.inst.size(); k++) {
        float a = 0.05f * f * (k == 3), t = 0.01f * f * (k == 3);
        X[12 * k + 0] = std::cos(a);
        X[12 * k + 2] = std::sin(a);
        X[12 * k + 3] = t;
        X[12 * k + 5] = 1.f;
        X[12 * k + 8] = -std::sin(a);
        X[12 * k + 10] = std::cos(a);
      }
      check(rs_cpu_update_instances(c, X.data(), (uint32_t)s.inst.size()), "update_instances");
      rs_cpu_set_frozen(c, f == 3 || f == 4);
      if (f == 4) rs_cpu_reset(c);
      rs_camera cam = camera(f, W, H, x.aperture, x.shutter);
      rs_cpu_frame_extra ex{nullptr, 0, f == 5 ? (uint32_t)RS_SCENE_SHOULD_RESET : 0u, rays.data(), lengths.data()};
      if (f == 2) {  // the pass's own primary hits as a caller VBuffer
        std::vector<uint32_t> hits(3 * P);
        size_t bytes = hits.size() * 4;
        check(rs_cpu_read_buffer(c, RS_BUF_VBUFFER, hits.data(), &bytes), "read vbuffer");
        for (uint32_t i = 0; i < P; i++) {
          uint32_t prim = hits[3 * i], inst = 0;
          while (inst + 1 < s.inst.size() && prim != 0xffffffffu && prim >= s.inst[inst + 1].firstTriangle) inst++;
          vb[4 * i] = prim == 0xffffffffu ? 0u : ((1u << 29) | inst);
          vb[4 * i + 1] = prim == 0xffffffffu ? 0u : prim - s.inst[inst].firstTriangle;
          vb[4 * i + 2] = hits[3 * i + 1];
          vb[4 * i + 3] = hits[3 * i + 2];
        }
        ex.vbuffer = vb.data();
      }
      check(rs_cpu_render_frame(c, &cam, W, H, nullptr, color.data(), &ex), "render_frame");
      for (int b = 0; b <= RS_BUF_FINAL_RECON; b++) {
        std::vector<uint8_t> buf(112 * 3 * (size_t)P);
        size_t bytes = buf.size();
        (void)rs_cpu_read_buffer(c, b, buf.data(), &bytes);
      }
      std::vector<float> counts(P);
      size_t n = P;
      check(rs_cpu_scatter_counts(c, counts.data(), &n), "scatter_counts");
    }
    std::vector<double> params(4 * mats.size()), img(3 * P), grad(4 * mats.size());
    for (size_t m = 0; m < mats.size(); m++) {
      for (int k = 0; k < 3; k++) params[4 * m + k] = mats[m].baseColor[k];
      params[4 * m + 3] = mats[m].roughness;
    }
    double loss = 0.0;
    check(rs_cpu_replay(c, params.data(), color.data(), img.data(), grad.data(), &loss), "replay");
    rs_cpu_destroy(c);
    v++;
  }
  rs_accum_options ao{1, RS_ACCUM_SINGLE_COMPENSATED, 3, RS_ACCUM_OVERFLOW_EMA};
  void* acc = nullptr;
  check(rs_cpu_accum_create(&ao, &acc), "accum_create");
  for (int i = 0; i < 5; i++) check(rs_cpu_accum_execute(acc, color.data(), accOut.data(), W, H, 1, i == 4), "accum");
  rs_cpu_accum_destroy(acc);
  std::printf("sanitize driver: %d variants, %d failures\n", v, g_failures);
  return g_failures == 0 ? 0 : 1;
}
