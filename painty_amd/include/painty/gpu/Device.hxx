// The device context shared by the drop-in classes, and the mapping of C ABI
// status codes onto the exceptions the reference throws.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "painty_hip.h"

namespace painty {
namespace gpu {

// Throws the reference's exception type for a failed call:
// std::invalid_argument where the reference throws it, std::runtime_error
// for device failures (which the CPU reference cannot have).
inline void check(int status) {
  if (status == PNT_OK) return;
  const std::string msg = pnt_last_error();
  if (status == PNT_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error("painty_hip: " + msg);
}

// The context (one device, one HIP stream) every drop-in object uses: device
// 0 unless Device::select(d) is called before the first object is made.
class Device {
 public:
  static pnt_ctx get() { return slot(-1); }
  static void select(int device) { slot(device); }

 private:
  static pnt_ctx slot(int device) {
    struct Holder {
      pnt_ctx ctx = nullptr;
      explicit Holder(int d) { check(pnt_ctx_create(d, &ctx)); }
      ~Holder() { pnt_ctx_destroy(ctx); }
    };
    static Holder h(device < 0 ? 0 : device);
    return h.ctx;
  }
};

// Images cross the boundary as FP32; the device stores FP32.
template <class T>
inline std::vector<float> to_f32(const T* p, std::size_t n) {
  std::vector<float> out(n);
  for (std::size_t i = 0; i < n; ++i) out[i] = static_cast<float>(p[i]);
  return out;
}
template <class T>
inline void from_f32(const std::vector<float>& in, T* p) {
  for (std::size_t i = 0; i < in.size(); ++i) p[i] = static_cast<T>(in[i]);
}

// State of a host mirror of device data:
//   Stale     -- the device is newer; the next read downloads
//   Clean     -- mirror == device
//   HostDirty -- handed out by a mutable getter; uploaded before the next
//                device operation that reads it
enum class Mirror { Stale, Clean, HostDirty };

}  // namespace gpu
}  // namespace painty
