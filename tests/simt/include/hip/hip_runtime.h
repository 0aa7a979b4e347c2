// Host (CPU) stand-in for the HIP runtime header, used only to compile the
// perturbation kernels' device headers for the CPU emulation harness
// (tests/simt/simt_host.cpp): one wavefront is 64 host threads, and every
// cross-lane operation (DPP moves, ds_bpermute, readlane, ballot, barriers)
// exchanges values through a per-wave buffer between two barrier phases.
// This lets the device code run under AddressSanitizer and
// UndefinedBehaviorSanitizer.  A cross-lane operation executed by only part
// of a wave (divergent control flow) never completes its barrier: the
// harness's watchdog reports that as an error.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __constant__ static const
// LDS arrays live in kernel bodies only, which the harness never runs (it calls
// the per-item device functions with arrays it shares between a wave's threads)
#define __shared__
// LDS pointers are plain host pointers here
#define DEB_LDS

// kernel bodies refer to these; the harness never runs a kernel body
struct simt_dim3 {
  int x, y, z;
};
inline constexpr simt_dim3 blockIdx{0, 0, 0}, threadIdx{0, 0, 0}, blockDim{64, 1, 1}, gridDim{1, 1, 1};

using hipError_t = int;
using hipStream_t = void *;
constexpr hipError_t hipSuccess = 0;
inline const char *hipGetErrorString(hipError_t) { return "host emulation"; }

namespace simt {

// Two exchange buffers used in turn: after the barrier of exchange k every
// lane has finished reading the buffer of exchange k - 1, so exchange k + 1
// may overwrite it -- one barrier per exchange.
struct Wave {
  std::barrier<> bar{64};
  uint64_t slot[2][64];
};
inline thread_local int lane = 0;
inline thread_local int parity = 0;
inline thread_local Wave *wave = nullptr;

// every lane publishes v; returns the value lane src published
inline uint64_t exchange(uint64_t v, int src) {
  uint64_t *buf = wave->slot[parity];
  parity ^= 1;
  buf[lane] = v;
  wave->bar.arrive_and_wait();
  return buf[src];
}

// source lane of a DPP move (-1: out of the wave or the row)
inline int dpp_source(int ctrl, int i) {
  if (ctrl <= 0xff) return (i & ~3) | ((ctrl >> (2 * (i & 3))) & 3);  // quad_perm
  switch (ctrl) {
    case 0x130: return i < 63 ? i + 1 : -1;                            // wave_shl:1
    case 0x138: return i > 0 ? i - 1 : -1;                             // wave_shr:1
    case 0x140: return (i & ~15) | (15 - (i & 15));                    // row_mirror
    case 0x141: return (i & ~7) | (7 - (i & 7));                       // row_half_mirror
    default: std::abort();                                             // not used by the kernels
  }
}

}  // namespace simt

inline int __lane_id() { return simt::lane; }

inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int, int, bool bound_zero) {
  const int s = simt::dpp_source(ctrl, simt::lane);
  const uint64_t v = simt::exchange((uint32_t)src, s < 0 ? 0 : s);
  if (s < 0) return bound_zero ? 0 : old;
  return (int)(uint32_t)v;
}
inline int __builtin_amdgcn_ds_bpermute(int addr, int src) {
  return (int)(uint32_t)simt::exchange((uint32_t)src, (addr >> 2) & 63);
}
inline int __builtin_amdgcn_readlane(int x, int l) { return (int)(uint32_t)simt::exchange((uint32_t)x, l); }
inline int __builtin_amdgcn_readfirstlane(int x) { return (int)(uint32_t)simt::exchange((uint32_t)x, 0); }
inline uint64_t __ballot(int pred) {
  uint64_t *buf = simt::wave->slot[simt::parity];
  simt::parity ^= 1;
  buf[simt::lane] = pred ? 1 : 0;
  simt::wave->bar.arrive_and_wait();
  uint64_t m = 0;
  for (int i = 0; i < 64; ++i) m |= buf[i] << i;
  return m;
}
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
inline void __builtin_amdgcn_s_barrier() { simt::wave->bar.arrive_and_wait(); }
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

inline int __double2loint(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  return (int)(uint32_t)b;
}
inline int __double2hiint(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  return (int)(uint32_t)(b >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const uint64_t b = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double x;
  std::memcpy(&x, &b, 8);
  return x;
}

template <class T>
inline T atomicAdd(T *p, T v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T atomicCAS(T *p, T expected, T desired) {
  __atomic_compare_exchange_n(p, &expected, desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return expected;
}
template <class T>
inline T atomicExch(T *p, T v) {
  return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST);
}

using std::exp;
using std::fabs;
using std::fma;
using std::fmax;
using std::max;
using std::min;
using std::fmin;
using std::isfinite;
using std::isinf;
using std::log;
using std::pow;
using std::sqrt;
