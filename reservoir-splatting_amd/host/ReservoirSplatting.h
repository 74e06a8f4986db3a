// C++ host mirror of the reference's ReservoirSplatting render pass over the MI355X library
// (include/rs.h).  It keeps the pass's C++ surface: construction from Properties, the
// property keys and their parsing, validation and warnings (RS/ReservoirSplatting.cpp:99-131,
// 182-344), setScene (:378), execute (:399-509), reset (:813-820) and the pybind properties
// pixelStats, useFixedSeed and fixedSeed (:142-156).  Errors the reference raises with
// FALCOR_THROW are thrown as rs_amd::Error; its logWarning messages go to the warning log.
#pragma once

#include <cstdint>
#include <functional>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "../../include/rs.h"

namespace rs_amd {

// Falcor::Properties as the pass reads it: string keys to bool, integer, floating-point or
// string values, in the order they were set (a render-graph script's dictionary).  Enum
// options are given by name ("ScatterBackup") or by value.
class Properties {
 public:
  using Value = std::variant<bool, int64_t, double, std::string>;

  Properties() = default;
  Properties(std::initializer_list<std::pair<std::string, Value>> items);

  bool has(const std::string& key) const;
  // The value of `key`, inserted (as false) when absent.
  Value& operator[](const std::string& key);
  const Value& at(const std::string& key) const;  // throws Error when absent
  size_t size() const { return mItems.size(); }

  std::vector<std::pair<std::string, Value>>::const_iterator begin() const { return mItems.begin(); }
  std::vector<std::pair<std::string, Value>>::const_iterator end() const { return mItems.end(); }

 private:
  std::vector<std::pair<std::string, Value>> mItems;
};

// FALCOR_THROW's counterpart: the library's RS_E_* code and message.
class Error : public std::runtime_error {
 public:
  Error(int32_t code, const std::string& msg) : std::runtime_error(msg), mCode(code) {}
  int32_t code() const { return mCode; }

 private:
  int32_t mCode;
};

// pixelStats (PixelStats): the last frame's ray counts, with logTraceRay semantics.
struct PixelStats {
  uint64_t closestHitRays = 0;
  uint64_t visibilityRays = 0;
  uint64_t totalRays = 0;
};

class ReservoirSplatting {
 public:
  using Logger = std::function<void(const std::string&)>;

  // createPass("ReservoirSplatting", props).  Warnings go to `log` when given and are kept
  // in warnings() either way.
  explicit ReservoirSplatting(const Properties& props = Properties(), Logger log = nullptr);
  ~ReservoirSplatting();
  ReservoirSplatting(const ReservoirSplatting&) = delete;
  ReservoirSplatting& operator=(const ReservoirSplatting&) = delete;

  // parseProperties + validateOptions.
  void setProperties(const Properties& props);
  // The options as properties: the reference's rendering, sampling and material keys, and
  // (unlike the reference, which drops them) the ReSTIR options, so a round trip keeps the mode.
  Properties getProperties() const;

  // setScene: the scene as the library's triangle soup; an environment map beside it.
  void setScene(const rs_scene_desc& scene);
  void setEnvMap(const rs_env_map_desc& envMap);
  void clearEnvMap();

  // execute: one frame of `camera` over the frame I/O (device pointers, sizes), on `stream`
  // (a hipStream_t; null for the default stream).
  void execute(const rs_camera& camera, const rs_frame_io& io, void* stream = nullptr);

  void reset();
  PixelStats getPixelStats() const;
  bool getUseFixedSeed() const;
  void setUseFixedSeed(bool value);
  uint32_t getFixedSeed() const;
  void setFixedSeed(uint32_t value);

  // The render-graph dictionary entries the last frame wrote (:503-506).
  rs_frame_flags getDictionary() const;

  const std::vector<std::string>& warnings() const { return mWarnings; }
  const rs_options& options() const { return mOptions; }
  rs_ctx* context() const { return mCtx; }

 private:
  void warn(const std::string& msg);
  void apply(const rs_options& o);
  void check(int32_t code) const;

  rs_ctx* mCtx = nullptr;
  rs_options mOptions{};
  Logger mLog;
  std::vector<std::string> mWarnings;
};

}  // namespace rs_amd
