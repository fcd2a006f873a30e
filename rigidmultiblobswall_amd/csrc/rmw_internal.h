// Shared host-side helpers of librmw_hip: error reporting, alignment,
// device properties.  Every extern "C" entry point catches everything and
// converts it to a status code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/rmw_hip.h"

namespace rmw {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define RMW_CHECK_HIP(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      throw ::rmw::Error(RMW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define RMW_REQUIRE(cond, msg)                                    \
  do {                                                            \
    if (!(cond)) throw ::rmw::Error(RMW_E_ARG, std::string(msg)); \
  } while (0)

// Run `body` and map exceptions to status codes.
template <typename F>
int guarded(F&& body)
{
  try {
    body();
    return RMW_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RMW_E_ARG;
  } catch (...) {
    set_last_error("unknown error");
    return RMW_E_ARG;
  }
}

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

struct DeviceProps {
  int device = -1;
  int cu_count = 0;
  int clock_khz = 0;
  int warp = 64;
  std::string arch;
};

// Cached per current device.
const DeviceProps& device_props();

// Pair-kernel timing (rmw_profile_enable / rmw_profile_collect): when
// enabled, prof_begin/prof_end bracket one pair-kernel launch with HIP
// events recorded on its stream.
// kind: 0 wall RPY, 1 different radii, 2 blob-blob forces.
struct ProfRec { hipEvent_t a, b; int kind; };
constexpr int kProfKinds = 3;
void prof_begin(ProfRec& rec, bool& on, hipStream_t st, int kind = 0);
void prof_end(const ProfRec& rec, bool on, hipStream_t st);

}  // namespace rmw
