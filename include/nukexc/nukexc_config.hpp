// nukexc/nukexc_config.hpp -- C++ shim over libnukexc_hip.so.
//
// Replaces the reference's nukexc_config.hpp (Kokkos ExecSpace, views,
// Point; nukexc_config.hpp:50-112) for the XC path.  Device arrays are owned
// through the C ABI (nkx_malloc / nkx_free), so code that includes these
// headers needs neither HIP nor Kokkos headers; link with -lnukexc_hip.
// Errors from the ABI are rethrown as std::runtime_error, as the reference
// throws (e.g. xc_integrals.hpp:81-84).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nukexc_hip.h"

namespace Nukexc {

// ---- status handling -------------------------------------------------------
inline void check(int status, const nkx_ctx *ctx = nullptr) {
  if (status != NKX_OK)
    throw std::runtime_error(std::string("nukexc: ") + nkx_last_error(ctx) +
                             " (status " + std::to_string(status) + ")");
}

// ---- execution context -----------------------------------------------------
// The reference runs every call on a default Kokkos execution-space instance
// (xc_integrals.hpp:78) after Kokkos::initialize (standalone.cxx:252).  Here
// each host thread gets one nkx_ctx on the current device on first use.
class Context {
 public:
  explicit Context(int device = 0) { check(nkx_create(device, nullptr, &h_)); }
  ~Context() { nkx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  nkx_ctx *get() const { return h_; }
  void fence() const { check(nkx_sync(h_), h_); }

 private:
  nkx_ctx *h_ = nullptr;
};

inline int &default_device() {
  static int dev = 0;
  return dev;
}

inline Context &context() {
  thread_local std::unique_ptr<Context> ctx(new Context(default_device()));
  return *ctx;
}

inline void initialize(int device = 0) {
  default_device() = device;
  (void)context();
}

// Execution-space instance (Kokkos::DefaultExecutionSpace in the reference,
// nukexc_config.hpp:50; `ExecSpace space;` at density.hpp:58 and
// standalone.cxx).  It names the calling thread's ctx, whose one HIP stream
// orders every call; work given to a space is asynchronous until fence(), as
// a Kokkos kernel launched on an instance is.
class ExecSpace {
 public:
  ExecSpace() : ctx_(context().get()) {}
  nkx_ctx *ctx() const { return ctx_; }
  void fence() const { check(nkx_sync(ctx_), ctx_); }

 private:
  nkx_ctx *ctx_;
};

// ---- points and device arrays ----------------------------------------------
struct Point {  // ArborX::Point<3, double> in the reference
  double x[3] = {0, 0, 0};
  double &operator[](int i) { return x[i]; }
  const double &operator[](int i) const { return x[i]; }
};
static_assert(sizeof(Point) == 3 * sizeof(double), "Point must be three packed doubles");

// Owning device array with the extents of a Kokkos view.  rank-2 arrays use
// LayoutLeft like the reference's DeviceView2DLeft (nukexc_config.hpp:97-98):
// element (i, j) at i + j * extent(0).
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n0, size_t n1 = 1) { allocate(n0, n1); }
  DeviceArray(DeviceArray &&o) noexcept { swap(o); }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    swap(o);
    return *this;
  }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() {
    if (p_) nkx_free(context().get(), p_);
  }

  void allocate(size_t n0, size_t n1 = 1) {
    if (p_) nkx_free(context().get(), p_);
    p_ = nullptr;
    n0_ = n0;
    n1_ = n1;
    void *v = nullptr;
    check(nkx_malloc(context().get(), size() * sizeof(T), &v), context().get());
    p_ = static_cast<T *>(v);
  }
  size_t extent(int r) const { return r == 0 ? n0_ : n1_; }
  size_t size() const { return n0_ * n1_; }
  T *data() const { return p_; }

  void from_host(const T *src) {
    check(nkx_memcpy_h2d(context().get(), p_, src, size() * sizeof(T)), context().get());
  }
  void from_host(const std::vector<T> &src) {
    if (src.size() != size()) throw std::runtime_error("DeviceArray::from_host: size mismatch");
    from_host(src.data());
  }
  std::vector<T> to_host() const {
    std::vector<T> out(size());
    if (size())
      check(nkx_memcpy_d2h(context().get(), out.data(), p_, size() * sizeof(T)),
            context().get());
    return out;
  }

 private:
  void swap(DeviceArray &o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n0_, o.n0_);
    std::swap(n1_, o.n1_);
  }
  T *p_ = nullptr;
  size_t n0_ = 0, n1_ = 0;
};

using DeviceView1D = DeviceArray<double>;
using DeviceView2DLeft = DeviceArray<double>;

}  // namespace Nukexc
