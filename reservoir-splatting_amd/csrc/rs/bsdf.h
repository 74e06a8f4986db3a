// StandardMaterial (MetalRough) and StandardBSDF with the ReSTIR lobe mask.
//
// Follows Source/Falcor/Rendering/Materials/StandardMaterial.slang:50-133 (instance
// setup), StandardMaterialInstance.slang:63-160 (eval / sample / evalPdf in the local
// frame with hemisphere checks), BSDFs/StandardBSDF.slang:128-347 (lobe probabilities,
// lobe mask semantics: mask 0 means all lobes; sample() adds the other lobes' pdf only
// when the mask is 0), BSDFs/FrostbiteDiffuseBRDF.slang:39-104,
// BSDFs/LambertDiffuseBRDF.slang:37-97, BSDFs/SpecularMicrofacet.slang:49-184 (GGX with
// VNDF sampling, height-correlated Smith masking), IsotropicGGX.slang:46-206,
// Fresnel.slang:42-51 and Source/Falcor/Scene/Material/ShadingUtils.slang:111-245
// (normal flip, shading-normal adjustment, back-face-black hemisphere test).
// Only opaque materials are modelled (no transmission lobes are ever active).
#pragma once
#include "scene.h"
#include "rng.h"

namespace rs {

static constexpr float kMinCosTheta = 1e-6f;
static constexpr float kMinGGXAlpha = 0.0064f;

RS_HD float pow5(float x) {
  float x2 = x * x;
  return x2 * x2 * x;
}
RS_HD float fresnel_schlick(float f0, float f90, float c) { return f0 + (f90 - f0) * pow5(fmax_(1.f - c, 0.f)); }
RS_HD vec3 fresnel_schlick(vec3 f0, float f90, float c) {
  float p = pow5(fmax_(1.f - c, 0.f));
  return vec3{f0.x + (f90 - f0.x) * p, f0.y + (f90 - f0.y) * p, f0.z + (f90 - f0.z) * p};
}

RS_HD float ggx_ndf(float alpha, float c) {
  float a2 = alpha * alpha;
  float d = ((c * a2 - c) * c + 1.f);
  return a2 / (d * d * kPi);
}
RS_HD float ggx_g1(float alphaSqr, float c) {
  if (c <= 0.f) return 0.f;
  float c2 = c * c;
  float tan2 = fmax_(1.f - c2, 0.f) / c2;
  return 2.f / (1.f + fsqrt(1.f + alphaSqr * tan2));
}
RS_HD float ggx_lambda(float alphaSqr, float c) {
  if (c <= 0.f) return 0.f;
  float c2 = c * c;
  float tan2 = fmax_(1.f - c2, 0.f) / c2;
  return 0.5f * (-1.f + fsqrt(1.f + alphaSqr * tan2));
}
RS_HD float ggx_masking_correlated(float alpha, float ci, float co) {
  float a2 = alpha * alpha;
  float li = ggx_lambda(a2, ci), lo = ggx_lambda(a2, co);
  return 1.f / (1.f + li + lo);
}
RS_HD float ggx_vndf_pdf(float alpha, vec3 wi, vec3 h) {
  float G1 = ggx_g1(alpha * alpha, wi.z);
  float D = ggx_ndf(alpha, h.z);
  return G1 * D * fmax_(0.f, dot(wi, h)) / wi.z;
}
// Dupuy & Benyoub 2023, "Sampling Visible GGX Normals with Spherical Caps".
RS_HD vec3 ggx_vndf_sample(float alpha, vec3 wi, vec2 u, float& pdf) {
  vec3 Vh = normalize(vec3{alpha * wi.x, alpha * wi.y, wi.z});
  float phi = k2Pi * u.x;
  float z = __builtin_fmaf(1.f - u.y, 1.f + Vh.z, -Vh.z);
  float sinTheta = fsqrt(fsaturate(1.f - z * z));
  float sp, cp;
  fsincos(phi, sp, cp);
  vec3 c{sinTheta * cp, sinTheta * sp, z};
  vec3 Nh = c + Vh;
  vec3 h = normalize(vec3{alpha * Nh.x, alpha * Nh.y, Nh.z});
  pdf = ggx_vndf_pdf(alpha, wi, h);
  return h;
}

// StandardBSDFData restricted to the opaque case.
struct BSDFData {
  vec3 diffuse;
  vec3 specular;
  float roughness;
  float metallic;
  float eta;
};

struct MaterialInstance {
  vec3 T, B, N;  // shading frame after flip / adjustment
  BSDFData data;
  vec3 emission;
  uint32_t diffuseModel;

  RS_HD vec3 toLocal(vec3 v) const { return vec3{dot(v, T), dot(v, B), dot(v, N)}; }
  RS_HD vec3 fromLocal(vec3 v) const { return T * v.x + B * v.y + N * v.z; }
};

// setupMaterialInstance + flipShadingNormal (+ adjustShadingNormal when `adjust`), with the
// base colour and roughness given explicitly (the frozen replay evaluates the material at
// other parameter values than the scene's).
RS_HD MaterialInstance make_material_instance_with(const SceneView& s, const ShadingData& sd, bool adjust, vec3 base,
                                                   float roughness) {
  const Material& m = s.materials[sd.materialID];
  MaterialInstance mi;
  float ior = m.ior;
  mi.data.eta = sd.frontFacing ? (sd.IoR / ior) : (ior / sd.IoR);
  float f = (ior - 1.f) / (ior + 1.f);
  float F0 = f * f;
  float metal = m.metallic;
  mi.data.diffuse = vec3{base.x + (0.f - base.x) * metal, base.y + (0.f - base.y) * metal, base.z + (0.f - base.z) * metal};
  mi.data.specular = vec3{F0 + (base.x - F0) * metal, F0 + (base.y - F0) * metal, F0 + (base.z - F0) * metal};
  mi.data.roughness = roughness;
  mi.data.metallic = metal;
  mi.T = sd.T;
  mi.B = sd.B;
  mi.N = sd.N;
  if (!sd.frontFacing && sd.doubleSided) mi.N = -mi.N;
  if (adjust) {
    vec3 Ng = sd.orientedFaceNormal();
    float signN = dot(mi.N, Ng) >= 0.f ? 1.f : -1.f;
    vec3 Ns = mi.N * signN;
    const float kCosThetaThreshold = 0.1f;
    float cosTheta = dot(sd.V, Ns);
    if (cosTheta <= kCosThetaThreshold) {
      float t = fsaturate(cosTheta * (1.f / kCosThetaThreshold));
      vec3 n = lerp(Ng, Ns, t) * signN;
      // ShadingFrame(n, tangentW) with tangentW = (invented T, +1).
      vec3 N = normalize(n);
      vec3 tw = sd.T;
      vec3 Tn = normalize(tw - N * dot(tw, N));
      mi.N = N;
      mi.T = Tn;
      mi.B = cross(N, Tn);
    }
  }
  mi.emission = sd.frontFacing ? m.emissive : vec3{0.f, 0.f, 0.f};
  mi.diffuseModel = s.diffuseModel;
  return mi;
}

RS_HD MaterialInstance make_material_instance(const SceneView& s, const ShadingData& sd, bool adjust) {
  const Material& m = s.materials[sd.materialID];
  return make_material_instance_with(s, sd, adjust, m.baseColor, m.roughness);
}

// StandardBSDF, constructed per query from the local incident direction.
struct StandardBSDF {
  vec3 diffAlbedo;
  float diffRoughness;
  vec3 specAlbedo;
  float alpha;
  uint32_t activeLobes;
  uint32_t lobeMask;
  uint32_t diffuseModel;
  float pDiffuseReflection, pSpecularReflection;

  RS_HD void init(vec3 wi, uint32_t mask, uint32_t active, const BSDFData& d, uint32_t model) {
    lobeMask = mask;
    diffuseModel = model;
    diffAlbedo = d.diffuse;
    diffRoughness = d.roughness;
    float a = d.roughness * d.roughness;
    a = fmax_(a, kMinGGXAlpha);
    alpha = lobeMask == 5 ? 0.f : a;
    activeLobes = active;
    specAlbedo = d.specular;
    const float specTrans = 0.f, diffTrans = 0.f;
    float metallicBRDF = d.metallic * (1.f - specTrans);
    float dielectricBSDF = (1.f - d.metallic) * (1.f - specTrans);
    float diffuseWeight = luminance(d.diffuse);
    float specularWeight = luminance(fresnel_schlick(d.specular, 1.f, wi.z));
    pDiffuseReflection = (active & kLobeDiffuseReflection) ? diffuseWeight * dielectricBSDF * (1.f - diffTrans) : 0.f;
    pSpecularReflection =
        (active & (kLobeSpecularReflection | kLobeDeltaReflection)) ? specularWeight * (metallicBRDF + dielectricBSDF) : 0.f;
    float norm = pDiffuseReflection + 0.f + pSpecularReflection + 0.f;
    if (norm > 0.f) {
      norm = 1.f / norm;
      pDiffuseReflection *= norm;
      pSpecularReflection *= norm;
    }
  }

  // ---- diffuse lobe
  RS_HD vec3 frostbite_weight(vec3 wi, vec3 wo) const {
    vec3 h = normalize(wi + wo);
    float woDotH = dot(wo, h);
    float energyBias = 0.f + (0.5f - 0.f) * diffRoughness;
    float energyFactor = 1.f + (1.f / 1.51f - 1.f) * diffRoughness;
    float fd90 = energyBias + 2.f * woDotH * woDotH * diffRoughness;
    float wiScatter = fresnel_schlick(1.f, fd90, wi.z);
    float woScatter = fresnel_schlick(1.f, fd90, wo.z);
    return diffAlbedo * wiScatter * woScatter * energyFactor;
  }
  RS_HD vec3 diffuse_eval(vec3 wi, vec3 wo) const {
    if (fmin_(wi.z, wo.z) < kMinCosTheta) return v3s(0.f);
    if (diffuseModel == kDiffuseLambert) return kInvPi * diffAlbedo * wo.z;
    return frostbite_weight(wi, wo) * kInvPi * wo.z;
  }
  RS_HD float diffuse_pdf(vec3 wi, vec3 wo) const {
    if (fmin_(wi.z, wo.z) < kMinCosTheta) return 0.f;
    return kInvPi * wo.z;
  }
  RS_HD bool diffuse_sample(vec3 wi, vec3& wo, float& pdf, vec3& weight, uint32_t& lobe, Rng& sg) const {
    wo = sample_cosine_hemisphere_concentric(next2(sg), pdf);
    lobe = kLobeDiffuseReflection;
    if (fmin_(wi.z, wo.z) < kMinCosTheta) {
      weight = v3s(0.f);
      return false;
    }
    weight = diffuseModel == kDiffuseLambert ? diffAlbedo : frostbite_weight(wi, wo);
    return true;
  }

  // ---- specular (GGX) lobe
  RS_HD bool has_spec_lobe() const { return (activeLobes & kLobeSpecularReflection) != 0; }
  RS_HD vec3 spec_eval(vec3 wi, vec3 wo) const {
    if (fmin_(wi.z, wo.z) < kMinCosTheta) return v3s(0.f);
    if (alpha == 0.f) return v3s(0.f);
    if (!has_spec_lobe()) return v3s(0.f);
    vec3 h = normalize(wi + wo);
    float wiDotH = dot(wi, h);
    float D = ggx_ndf(alpha, h.z);
    float G = ggx_masking_correlated(alpha, wi.z, wo.z);
    vec3 F = fresnel_schlick(specAlbedo, 1.f, wiDotH);
    return F * D * G * 0.25f / wi.z;
  }
  RS_HD float spec_pdf(vec3 wi, vec3 wo) const {
    if (fmin_(wi.z, wo.z) < kMinCosTheta) return 0.f;
    if (alpha == 0.f) return 0.f;
    if (!has_spec_lobe()) return 0.f;
    vec3 h = normalize(wi + wo);
    float wiDotH = dot(wi, h);
    float pdf = ggx_vndf_pdf(alpha, wi, h);
    return pdf / (4.f * wiDotH);
  }
  RS_HD bool spec_sample(vec3 wi, vec3& wo, float& pdf, vec3& weight, uint32_t& lobe, Rng& sg) const {
    wo = v3s(0.f);
    weight = v3s(0.f);
    pdf = 0.f;
    lobe = kLobeSpecularReflection;
    if (wi.z < kMinCosTheta) return false;
    if (alpha == 0.f) {
      if (!(activeLobes & kLobeDeltaReflection)) return false;
      wo = vec3{-wi.x, -wi.y, wi.z};
      pdf = 0.f;
      weight = fresnel_schlick(specAlbedo, 1.f, wi.z);
      lobe = kLobeDeltaReflection;
      return true;
    }
    if (!has_spec_lobe()) return false;
    vec3 h = ggx_vndf_sample(alpha, wi, next2(sg), pdf);
    float wiDotH = dot(wi, h);
    wo = h * (2.f * wiDotH) - wi;
    if (wo.z < kMinCosTheta) return false;
    float G = ggx_masking_correlated(alpha, wi.z, wo.z);
    float GOverG1wo = G * (1.f + ggx_lambda(alpha * alpha, wi.z));
    vec3 F = fresnel_schlick(specAlbedo, 1.f, wiDotH);
    pdf /= (4.f * wiDotH);
    weight = F * GOverG1wo;
    lobe = kLobeSpecularReflection;
    return true;
  }

  // ---- mixture
  RS_HD vec3 eval(vec3 wi, vec3 wo) const {
    uint32_t mask = lobeMask == 0 ? 0xfu : lobeMask;
    vec3 r = v3s(0.f);
    if (pDiffuseReflection > 0.f && (mask & 1u)) r += (1.f - 0.f) * (1.f - 0.f) * diffuse_eval(wi, wo);
    if (pSpecularReflection > 0.f && (mask & 4u)) r += (1.f - 0.f) * spec_eval(wi, wo);
    return r;
  }
  RS_HD float eval_pdf(vec3 wi, vec3 wo) const {
    uint32_t mask = lobeMask == 0 ? 0xfu : lobeMask;
    float pdf = 0.f;
    if (pDiffuseReflection > 0.f && (mask & 1u)) pdf += pDiffuseReflection * diffuse_pdf(wi, wo);
    if (pSpecularReflection > 0.f && (mask & 4u)) pdf += pSpecularReflection * spec_pdf(wi, wo);
    return pdf;
  }
  RS_HD bool sample(vec3 wi, vec3& wo, float& pdf, vec3& weight, uint32_t& lobe, Rng& sg) const {
    wo = v3s(0.f);
    weight = v3s(0.f);
    pdf = 0.f;
    lobe = kLobeDiffuseReflection;
    bool valid = false;
    float uSelect = next1(sg);
    if (uSelect < pDiffuseReflection) {
      valid = diffuse_sample(wi, wo, pdf, weight, lobe, sg);
      weight /= pDiffuseReflection;
      weight *= (1.f - 0.f) * (1.f - 0.f);
      pdf *= pDiffuseReflection;
      if (lobeMask == 0) {
        if (pSpecularReflection > 0.f) pdf += pSpecularReflection * spec_pdf(wi, wo);
      }
    } else if (uSelect < pDiffuseReflection + 0.f) {
      // diffuse transmission: never active for opaque materials
    } else if (uSelect < pDiffuseReflection + 0.f + pSpecularReflection) {
      valid = spec_sample(wi, wo, pdf, weight, lobe, sg);
      weight /= pSpecularReflection;
      weight *= (1.f - 0.f);
      pdf *= pSpecularReflection;
      if (lobeMask == 0) {
        if (pDiffuseReflection > 0.f) pdf += pDiffuseReflection * diffuse_pdf(wi, wo);
      }
    }
    return valid;
  }
};

// StandardBSDF::getLobeTypes for the opaque case.
RS_HD uint32_t bsdf_lobe_types(const BSDFData& d) {
  uint32_t lt = kLobeSpecularReflection;
  if (d.diffuse.x > 0.f || d.diffuse.y > 0.f || d.diffuse.z > 0.f) lt |= kLobeDiffuseReflection;
  return lt;
}

RS_HD bool valid_hemisphere(const ShadingData& sd, const MaterialInstance& mi, vec3 wiL, vec3 woL, vec3 wo) {
  bool isTransmission = wiL.z * woL.z < 0.f;
  if (fmin_(wiL.z, fabs_(woL.z)) < kMinCosTheta) return false;
  bool wiTop = sd.frontFacing;
  bool woTop = dot(wo, sd.faceN) >= 0.f;
  if (isTransmission) {
    if (wiTop == woTop) return false;
  } else {
    if (wiTop != woTop) return false;
  }
  bool shadingTop = dot(mi.N, sd.faceN) >= 0.f;
  if (wiTop != shadingTop) return false;
  return true;
}

struct BSDFSample {
  vec3 wo;
  float pdf;
  vec3 weight;
  uint32_t lobeType;
};

// IMaterialInstance wrappers; `lobeMask` is the current ReSTIR lobe mask in the
// material header (ShadingData::lobeMask).
RS_HD vec3 mi_eval(const ShadingData& sd, const MaterialInstance& mi, vec3 wo) {
  vec3 wiL = mi.toLocal(sd.V), woL = mi.toLocal(wo);
  if (!valid_hemisphere(sd, mi, wiL, woL, wo)) return v3s(0.f);
  StandardBSDF b;
  b.init(wiL, sd.lobeMask, sd.activeLobes, mi.data, mi.diffuseModel);
  return b.eval(wiL, woL);
}
RS_HD float mi_eval_pdf(const ShadingData& sd, const MaterialInstance& mi, vec3 wo) {
  vec3 wiL = mi.toLocal(sd.V), woL = mi.toLocal(wo);
  if (!valid_hemisphere(sd, mi, wiL, woL, wo)) return 0.f;
  StandardBSDF b;
  b.init(wiL, sd.lobeMask, sd.activeLobes, mi.data, mi.diffuseModel);
  return b.eval_pdf(wiL, woL);
}
RS_HD bool mi_sample(const ShadingData& sd, const MaterialInstance& mi, Rng& sg, BSDFSample& r) {
  vec3 wiL = mi.toLocal(sd.V);
  vec3 woL;
  r.wo = v3s(0.f);
  StandardBSDF b;
  b.init(wiL, sd.lobeMask, sd.activeLobes, mi.data, mi.diffuseModel);
  if (!b.sample(wiL, woL, r.pdf, r.weight, r.lobeType, sg)) return false;
  r.wo = mi.fromLocal(woL);
  if (!valid_hemisphere(sd, mi, wiL, woL, r.wo)) return false;
  return true;
}

// The BSDF value of mi_eval and its derivatives with respect to the material's base colour
// and roughness, every other input held fixed.  Channel c of the value depends on base
// colour channel c only (diffuse albedo base*(1-metallic), specular albedo
// F0 + (base-F0)*metallic), so dBase holds d f_c / d base_c.  The roughness derivative
// goes through the Frostbite diffuse terms and the GGX alpha = max(r^2, 0.0064) of the
// normal distribution and the height-correlated masking term.
struct BSDFGrad {
  vec3 f;
  vec3 dBase;
  vec3 dRough;
};

RS_HD BSDFGrad mi_eval_grad(const ShadingData& sd, const MaterialInstance& mi, vec3 wo, float metallic) {
  BSDFGrad g{v3s(0.f), v3s(0.f), v3s(0.f)};
  vec3 wi = mi.toLocal(sd.V), woL = mi.toLocal(wo);
  if (!valid_hemisphere(sd, mi, wi, woL, wo)) return g;
  StandardBSDF b;
  b.init(wi, sd.lobeMask, sd.activeLobes, mi.data, mi.diffuseModel);
  g.f = b.eval(wi, woL);
  if (fmin_(wi.z, woL.z) < kMinCosTheta) return g;
  const uint32_t mask = b.lobeMask == 0 ? 0xfu : b.lobeMask;
  const float r = b.diffRoughness;
  if (b.pDiffuseReflection > 0.f && (mask & 1u)) {
    const float kd = (1.f - metallic) * kInvPi * woL.z;
    if (b.diffuseModel == kDiffuseLambert) {
      g.dBase += v3s(kd);
    } else {
      vec3 h = normalize(wi + woL);
      float woDotH = dot(woL, h);
      float ef = 1.f + (1.f / 1.51f - 1.f) * r, dEf = 1.f / 1.51f - 1.f;
      float fd90 = 0.5f * r + 2.f * woDotH * woDotH * r, dFd90 = 0.5f + 2.f * woDotH * woDotH;
      float pi5 = pow5(fmax_(1.f - wi.z, 0.f)), po5 = pow5(fmax_(1.f - woL.z, 0.f));
      float si = 1.f + (fd90 - 1.f) * pi5, so = 1.f + (fd90 - 1.f) * po5;
      float S = si * so * ef;
      float dS = dFd90 * pi5 * so * ef + si * dFd90 * po5 * ef + si * so * dEf;
      g.dBase += v3s(kd * S);
      g.dRough += b.diffAlbedo * (dS * kInvPi * woL.z);
    }
  }
  if (b.pSpecularReflection > 0.f && (mask & 4u) && b.alpha != 0.f && b.has_spec_lobe()) {
    vec3 h = normalize(wi + woL);
    float wiDotH = dot(wi, h);
    float a = b.alpha, a2 = a * a;
    float c = h.z, c2 = c * c;
    float d = c2 * (a2 - 1.f) + 1.f;
    float D = a2 / (kPi * d * d);
    float dDda2 = (d - 2.f * a2 * c2) / (kPi * d * d * d);
    auto tan2 = [](float cz) { return fmax_(1.f - cz * cz, 0.f) / (cz * cz); };
    float ti = tan2(wi.z), to = tan2(woL.z);
    float qi = fsqrt(1.f + a2 * ti), qo = fsqrt(1.f + a2 * to);
    float G = 1.f / (1.f + 0.5f * (qi - 1.f) + 0.5f * (qo - 1.f));
    float dGda2 = -G * G * (0.25f * ti / qi + 0.25f * to / qo);
    float p5 = pow5(fmax_(1.f - wiDotH, 0.f));
    vec3 F = fresnel_schlick(b.specAlbedo, 1.f, wiDotH);
    float k = 0.25f / wi.z;
    // d alpha / d r = 2r while r^2 is above the alpha floor
    float dA2dr = (r * r > kMinGGXAlpha) ? 2.f * a * 2.f * r : 0.f;
    g.dBase += v3s(metallic * (1.f - p5) * D * G * k);
    g.dRough += F * ((dDda2 * G + D * dGda2) * dA2dr * k);
  }
  return g;
}

// BSDFProperties::roughness (clamped like StandardMaterialInstance::getProperties).
RS_HD float mi_roughness(const MaterialInstance& mi) {
  float a = mi.data.roughness * mi.data.roughness;
  return a < kMinGGXAlpha ? 0.f : mi.data.roughness;
}

}  // namespace rs
