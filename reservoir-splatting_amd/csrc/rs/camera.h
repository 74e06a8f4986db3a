// Thin-lens camera and the keyframe camera manager of the pass.
//
// CameraData mirrors the fields of Source/Falcor/Scene/Camera/CameraData.slang:36-72
// that the pass reads.  Rays follow computeRayThinlensSubPixelRandom
// (Source/Falcor/Scene/Camera/Camera.slang:146-168).  The keyframe ring and the time
// interpolation (position lerp, forward-vector slerp by quaternion) follow
// Source/RenderPasses/ReservoirSplatting/CameraManager.slang:33-114; the host fills the
// ring as in RS/ReservoirSplatting.cpp:1360-1388 (all slots on frame 0 or reset, one
// slot per frame otherwise) with K = 2 + ceil(shutter / frameTime) slots (:1633-1641).
#pragma once
#include "sampling.h"

namespace rs {

static constexpr int kMaxCameraKeyframes = 16;

struct CameraData {
  vec3 posW;
  vec3 up;
  vec3 cameraU, cameraV, cameraW;
  float focalLength;    // mm
  float frameHeight;    // mm
  float aspectRatio;    // width / height
  float focalDistance;  // scene units
  float apertureRadius;
  float shutterSpeed;
  float _pad;
};

struct CameraManager {
  float artificialFrameTime;
  uint32_t numKeyframes;
  uint32_t _pad0, _pad1;
  vec3 positions[kMaxCameraKeyframes];
  vec3 targets[kMaxCameraKeyframes];
};

struct Ray {
  vec3 origin;
  float tMin;
  vec3 dir;
  float tMax;
};

static constexpr float kRayTMax = 1e30f;

RS_HD vec3 angular_lerp(vec3 prevForward, vec3 currForward, float t) {
  vec3 axis = normalize(cross(currForward, prevForward));
  float angle = facos(dot(currForward, prevForward));
  if (angle < 0.001f) return currForward;
  quat q = rotate_angle_axis(angle * t, axis);
  return rotate_vector(currForward, q);
}

RS_HD void recalc_camera_vectors(CameraData& cam, vec3 forward) {
  cam.cameraW = forward * cam.focalDistance;
  cam.cameraU = normalize(cross(cam.cameraW, cam.up));
  cam.cameraV = normalize(cross(cam.cameraU, cam.cameraW));
  float fovY = cam.focalLength == 0.f ? 0.f : 2.f * fatan(0.5f * cam.frameHeight / cam.focalLength);
  float th = ftan(fovY * 0.5f);
  float ulen = cam.focalDistance * th * cam.aspectRatio;
  float vlen = cam.focalDistance * th;
  cam.cameraU = cam.cameraU * ulen;
  cam.cameraV = cam.cameraV * vlen;
}

// Orient `cam` at `time` (seconds before the current frame) using the keyframe ring.
RS_HD void orient_camera(const CameraManager& cm, CameraData& cam, uint32_t frameCount, float time) {
  int K = (int)cm.numKeyframes;
  int currIndex = (int)(frameCount % (uint32_t)K);
  int frameIndex = (int)ffloor(time / cm.artificialFrameTime);
  int first = ((currIndex - frameIndex - 1 + K) % K + K) % K;
  int second = (first + 1 + K) % K;
  float t = (time - (float)frameIndex * cm.artificialFrameTime) / cm.artificialFrameTime;
  vec3 p0 = cm.positions[first], t0 = cm.targets[first];
  vec3 p1 = cm.positions[second], t1 = cm.targets[second];
  cam.up = vec3{0.f, 1.f, 0.f};
  cam.posW = lerp(p1, p0, t);
  vec3 f0 = normalize(t0 - p0);
  vec3 f1 = normalize(t1 - p1);
  recalc_camera_vectors(cam, angular_lerp(f0, f1, t));
}

RS_HD Ray camera_ray_subpixel(const CameraData& cam, uint32_t W, uint32_t H, vec2 fractionalPixel, vec2 lensSample) {
  const float eps = 0.001f;
  vec2 fp{fclamp(fractionalPixel.x, eps, (float)W - eps), fclamp(fractionalPixel.y, eps, (float)H - eps)};
  vec2 p{fp.x / (float)W, fp.y / (float)H};
  vec2 ndc{2.f * p.x + -1.f, -2.f * p.y + 1.f};
  Ray ray;
  ray.origin = cam.posW;
  ray.dir = cam.cameraU * ndc.x + cam.cameraV * ndc.y + cam.cameraW;
  vec3 target = ray.origin + ray.dir;
  ray.origin = ray.origin + (normalize(cam.cameraU) * lensSample.x + normalize(cam.cameraV) * lensSample.y) * cam.apertureRadius;
  ray.dir = normalize(target - ray.origin);
  ray.tMin = 0.f;
  ray.tMax = kRayTMax;
  return ray;
}

}  // namespace rs
