// autotune.hpp -- launch-variant tuning of a device-resident model's step,
// with a JSON cache, for the C++ drop-in path.
//
// The counterpart of the reference's GPU autotuner
// (include/openpfc/runtime/common/gpu_autotune.hpp:249-333) and of its
// environment:
//   OPENPFC_GPU_AUTOTUNE_DISABLE=1   keep the library's default variants;
//   OPENPFC_GPU_AUTOTUNE_MODE        auto (default): take the cached choice,
//                                    or tune (pfc_hip_pfc_autotune) and
//                                    store it; cache_only: a missing entry
//                                    throws std::runtime_error;
//                                    fallback_only: defaults, no cache;
//   OPENPFC_GPU_AUTOTUNE_CACHE_DIR   directory of the cache file (default
//                                    "."), named <device>_autotune_cache.json.
// An entry is keyed by model, precision, extents and the mean-split mode and
// holds the variant name chosen for every step role; an entry naming a
// variant the library no longer offers counts as missing.  What is tuned
// here is the library's launch shape per pass (radix plan, lines per block,
// layout), where the reference tunes block sizes of its elementwise kernels.
#pragma once

#include "core.hpp"
#include "json.hpp"
#include "world.hpp"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>

namespace pfc::gpu {

inline std::string autotune_env(const char *name, const char *fallback) {
  const char *v = std::getenv(name);
  return v ? std::string(v) : std::string(fallback);
}

// "<arch>_<CUs>cu" of the current device, e.g. "gfx950_256cu".
inline std::string device_id() {
  int dev = 0;
  pfc::detail::check_hip(hipGetDevice(&dev), "autotune: hipGetDevice");
  hipDeviceProp_t p{};
  pfc::detail::check_hip(hipGetDeviceProperties(&p, dev), "autotune: hipGetDeviceProperties");
  std::string arch = p.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  return arch + "_" + std::to_string(p.multiProcessorCount) + "cu";
}

inline std::string cache_file() {
  return autotune_env("OPENPFC_GPU_AUTOTUNE_CACHE_DIR", ".") + "/" + device_id() +
         "_autotune_cache.json";
}

inline std::string step_key(const std::string &model, int precision, const Int3 &n,
                            bool mean_split) {
  return model + "/" + (precision == PFC_HIP_F64 ? "f64" : "f32") + "/" +
         std::to_string(n[0]) + "x" + std::to_string(n[1]) + "x" + std::to_string(n[2]) +
         (mean_split ? "/mean_split" : "");
}

// Select the variants of a cached entry; false (nothing changed) when any
// role's name is not offered by the plan.
inline bool apply_entry(pfc_hip_plan_t plan, const json::Value &entry) {
  if (!entry.is_object()) return false;
  std::vector<std::pair<int, int>> picks;
  for (int role = 0; role < 64; ++role) {
    int count = 0;
    if (pfc_hip_variant_count(plan, role, &count) != PFC_HIP_OK) break;
    if (count <= 0) continue;
    const json::Value *v = entry.find(std::to_string(role));
    if (!v) return false;
    int found = -1;
    for (int i = 0; i < count && found < 0; ++i) {
      const char *nm = pfc_hip_variant_name(plan, role, i);
      if (nm && v->is_string() && v->as_string() == nm) found = i;
    }
    if (found < 0) return false;
    picks.emplace_back(role, found);
  }
  for (auto [role, idx] : picks)
    pfc::detail::check_abi(pfc_hip_variant_set(plan, role, idx), "autotune: variant_set");
  return true;
}

inline json::Value current_entry(pfc_hip_plan_t plan) {
  json::Value e = json::Value::object();
  for (int role = 0; role < 64; ++role) {
    int count = 0, idx = 0;
    if (pfc_hip_variant_count(plan, role, &count) != PFC_HIP_OK) break;
    if (count <= 0) continue;
    pfc::detail::check_abi(pfc_hip_variant_get(plan, role, &idx), "autotune: variant_get");
    const char *nm = pfc_hip_variant_name(plan, role, idx);
    e[std::to_string(role)] = json::Value::string(nm ? nm : "");
  }
  return e;
}

inline json::Value load_cache(const std::string &path) {
  std::ifstream f(path);
  if (!f) return json::Value::object();
  std::stringstream ss;
  ss << f.rdbuf();
  try {
    json::Value v = json::parse(ss.str());
    if (v.is_object()) return v;
  } catch (const std::exception &) {
  }
  return json::Value::object(); // unreadable cache: as if empty
}

// Written to a temporary name and renamed, so a reader never sees half a file.
inline void store_cache(const std::string &path, const json::Value &v) {
  const std::string tmp = path + ".tmp" + std::to_string((long)std::rand());
  {
    std::ofstream f(tmp);
    if (!f) return; // a read-only cache directory only loses the cache
    f << v.dump(1) << "\n";
  }
  std::error_code ec;
  std::filesystem::rename(tmp, path, ec);
  if (ec) std::filesystem::remove(tmp, ec);
}

enum class TuneOutcome { defaults, cached, tuned };

// Tune (or restore from the cache) the step variants of `plan`, whose
// operators are set, timing full steps on `psi` (restored afterwards).
inline TuneOutcome tune_step(pfc_hip_plan_t plan, void *psi, const std::string &key,
                             hipStream_t stream) {
  if (autotune_env("OPENPFC_GPU_AUTOTUNE_DISABLE", "0") == "1") return TuneOutcome::defaults;
  const std::string mode = autotune_env("OPENPFC_GPU_AUTOTUNE_MODE", "auto");
  if (mode == "fallback_only") return TuneOutcome::defaults;
  if (mode != "auto" && mode != "cache_only")
    throw std::invalid_argument("OPENPFC_GPU_AUTOTUNE_MODE must be auto, cache_only or "
                                "fallback_only, got '" + mode + "'");
  const std::string path = cache_file();
  json::Value cache = load_cache(path);
  if (const json::Value *entries = cache.find("entries"))
    if (const json::Value *e = entries->find(key))
      if (apply_entry(plan, *e)) return TuneOutcome::cached;
  if (mode == "cache_only")
    throw std::runtime_error("No cached configuration for kernel: " + key + " (" + path + ")");
  pfc::detail::check_abi(pfc_hip_pfc_autotune(plan, psi, 3, stream, nullptr, 0),
                         "autotune: pfc_hip_pfc_autotune");
  cache = load_cache(path); // another process may have stored entries meanwhile
  cache["device_id"] = json::Value::string(device_id());
  json::Value &entries = cache["entries"];
  if (!entries.is_object()) entries = json::Value::object();
  entries[key] = current_entry(plan);
  store_cache(path, cache);
  return TuneOutcome::tuned;
}

} // namespace pfc::gpu
