// Internal helpers shared by the libwarpmi translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

void wm_set_error(const std::string &msg);

struct wm_error : std::runtime_error {
  int code;
  wm_error(const std::string &m, int c = -1) : std::runtime_error(m), code(c) {}
};

#define WM_HIP(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      throw wm_error(std::string(#x) + ": " + hipGetErrorString(e_), -2); \
  } while (0)

#define WM_REQUIRE(cond, msg)                      \
  do {                                             \
    if (!(cond)) throw wm_error(std::string(msg)); \
  } while (0)

// Runs body, turning exceptions into the ABI's error codes.
#define WM_ENTRY(...)                      \
  try {                                    \
    __VA_ARGS__;                           \
    return 0;                              \
  } catch (const wm_error &e) {            \
    wm_set_error(e.what());                \
    return e.code;                         \
  } catch (const std::exception &e) {      \
    wm_set_error(e.what());                \
    return -3;                             \
  }

static inline hipStream_t wm_stream(void *s) { return (hipStream_t)s; }

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Owning device array (hipMalloc'ed, zero-filled on allocation).  With a
// lead of k elements, p points k elements past the start of the allocation
// (the red-black arrays use it to put element 1 of every row on a 128-byte
// boundary).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  T *raw = nullptr;
  size_t n = 0, lead = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (raw) (void)hipFree(raw);
    p = raw = nullptr;
    n = lead = 0;
  }
  void alloc(size_t count, size_t lead_ = 0) {
    if (count == n && p && lead_ == lead) return;
    release();
    if (count == 0) return;
    WM_HIP(hipMalloc((void **)&raw, sizeof(T) * (count + lead_)));
    WM_HIP(hipMemset(raw, 0, sizeof(T) * (count + lead_)));
    p = raw + lead_;
    n = count;
    lead = lead_;
  }
  // a red-black array: element 1 of each row 128-byte aligned when the row
  // length is a multiple of 128 bytes
  void alloc_rb(size_t count) { alloc(count, 128 / sizeof(T) - 1); }
  // exchanges the two arrays (the ping-pong buffers of the fused SOR pass)
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(raw, o.raw);
    std::swap(n, o.n);
    std::swap(lead, o.lead);
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (!v.empty()) WM_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
};

// Launch check: reports the launch error of the kernel just enqueued.
#define WM_LAUNCHED() WM_HIP(hipGetLastError())

// Physical constant used by the solver: eps0 = 1/(mu0*clight*clight),
// mu0 = 4*pi/10000000 (scripts/warp.py:198-202; pywarp90/top.v:172-175).
static inline double wm_eps0() {
  double pi = 3.14159265358979323;
  double clight = 2.99792458e+8;
  double mu0 = 4 * pi / 10000000;
  return 1 / (mu0 * clight * clight);
}
