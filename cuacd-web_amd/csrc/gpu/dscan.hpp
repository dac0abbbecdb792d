// dscan.hpp — device-wide exclusive scan in one pass (HIP, gfx950).
//
// A persistent grid takes chunks of SCAN_CH consecutive elements in ticket
// order (an atomic counter), so every chunk below a block's current one is
// owned by a block that is already running. Each chunk scans its elements in
// the workgroup (wave prefix sums through DPP/shuffles, wave totals through
// LDS), publishes its aggregate, then looks back over its predecessors'
// published values (decoupled look-back: an inclusive prefix ends the walk,
// an aggregate adds to it and the walk goes on) and publishes its own
// inclusive prefix. Values are NF fields of T, summed with T's arithmetic
// (unsigned fields wrap modulo 2^bits, so differences of prefixes stay exact).
//
// The walk waits on a predecessor at most SCAN_SPIN times; past that it sets
// the control block's error word and stops waiting, so the kernel always ends
// and the host reports the error instead of using the result.
#pragma once
#include "dnum.hpp"

namespace dk {

constexpr uint32_t SCAN_T = 256;   // threads per block
constexpr uint32_t SCAN_IPT = 4;   // elements per thread
constexpr uint32_t SCAN_CH = SCAN_T * SCAN_IPT;
constexpr uint32_t SCAN_SPIN = 1u << 22;

struct ScanCtl {  // ticket and error are cleared before every scan (k_scan_reset)
  uint32_t ticket;
  uint32_t error;
  uint32_t sticky;  // every error since the owner last cleared it
  uint32_t pad;
};

// chunk status words (0 none, 1 aggregate, 2 inclusive prefix) and the control block
__global__ void k_scan_reset(uint32_t* __restrict__ status, uint32_t nStatus, ScanCtl* __restrict__ ctl) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nStatus; i += gridDim.x * blockDim.x) status[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->sticky |= ctl->error;
    ctl->ticket = 0;
    ctl->error = 0;
  }
}

template <typename T>
__device__ __forceinline__ void st_relaxed(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d) {
  return __shfl_up(v, d, 64);
}

// Scans S::count() elements. S provides
//   typedef T; static constexpr int NF;
//   uint32_t count() const;                        number of elements (may read device memory)
//   void load(uint32_t i, T* v) const;             the NF values of element i (i < count())
//   void store(uint32_t i, const T* excl) const;   consumer of element i's exclusive prefix
//   void total(const T* tot) const;                called once, with the sum of all elements
// agg / inc hold NF values per chunk.
template <typename S>
__device__ void scan_chunks(const S& s, uint32_t* __restrict__ status, typename S::T* __restrict__ agg,
                            typename S::T* __restrict__ inc, ScanCtl* __restrict__ ctl) {
  typedef typename S::T T;
  constexpr int NF = S::NF;
  constexpr uint32_t NW = SCAN_T / 64;
  __shared__ T s_w[NW][NF];
  __shared__ T s_pre[NF];
  __shared__ uint32_t s_chunk;
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t n = s.count();
  const uint32_t nChunks = (n + SCAN_CH - 1) / SCAN_CH;
  for (;;) {
    if (tid == 0) s_chunk = atomicAdd(&ctl->ticket, 1u);
    __syncthreads();
    const uint32_t c = s_chunk;
    __syncthreads();
    if (c >= nChunks) break;
    // thread-local values and their sum
    T v[SCAN_IPT][NF], run[NF];
    for (int f = 0; f < NF; f++) run[f] = 0;
    const uint32_t base = c * SCAN_CH + tid * SCAN_IPT;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_IPT; k++) {
      if (base + k < n) s.load(base + k, v[k]);
      else
        for (int f = 0; f < NF; f++) v[k][f] = 0;
      for (int f = 0; f < NF; f++) run[f] += v[k][f];
    }
    // wave inclusive scan of the thread sums, then the wave totals through LDS
    T incl[NF];
    for (int f = 0; f < NF; f++) {
      T x = run[f];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const T y = shfl_up_t(x, d);
        if (lane >= (uint32_t)d) x += y;
      }
      incl[f] = x;
      if (lane == 63) s_w[w][f] = x;
    }
    __syncthreads();
    T before[NF], chunkTot[NF];
    for (int f = 0; f < NF; f++) {
      T b = 0, t = 0;
      for (uint32_t q = 0; q < NW; q++) {
        const T y = s_w[q][f];
        b += q < w ? y : T(0);
        t += y;
      }
      before[f] = b + incl[f] - run[f];  // exclusive prefix of this thread within the chunk
      chunkTot[f] = t;
    }
    // publish, look back, publish the inclusive prefix
    if (tid == 0) {
      T pre[NF];
      for (int f = 0; f < NF; f++) pre[f] = 0;
      if (c == 0) {
        for (int f = 0; f < NF; f++) st_relaxed(&inc[f], chunkTot[f]);
        __hip_atomic_store(&status[0], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        for (int f = 0; f < NF; f++) st_relaxed(&agg[(size_t)c * NF + f], chunkTot[f]);
        __hip_atomic_store(&status[c], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t p = c - 1;
        for (;;) {
          uint32_t st = 0, spins = 0;
          while ((st = __hip_atomic_load(&status[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
            if (++spins >= SCAN_SPIN) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (st == 0) {
            atomicOr(&ctl->error, 1u);
            break;
          }
          const T* src = st == 2 ? inc : agg;
          for (int f = 0; f < NF; f++) pre[f] += ld_relaxed(&src[(size_t)p * NF + f]);
          if (st == 2 || p == 0) break;
          p--;
        }
        for (int f = 0; f < NF; f++) st_relaxed(&inc[(size_t)c * NF + f], pre[f] + chunkTot[f]);
        __hip_atomic_store(&status[c], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      for (int f = 0; f < NF; f++) s_pre[f] = pre[f];
      if (c == nChunks - 1) {
        T tot[NF];
        for (int f = 0; f < NF; f++) tot[f] = pre[f] + chunkTot[f];
        s.total(tot);
      }
    }
    __syncthreads();
    T ex[NF];
    for (int f = 0; f < NF; f++) ex[f] = s_pre[f] + before[f];
#pragma unroll
    for (uint32_t k = 0; k < SCAN_IPT; k++) {
      if (base + k < n) s.store(base + k, ex);
      for (int f = 0; f < NF; f++) ex[f] += v[k][f];
    }
    __syncthreads();
  }
  // an empty input still reports its (zero) total
  if (n == 0 && blockIdx.x == 0 && tid == 0) {
    T tot[NF];
    for (int f = 0; f < NF; f++) tot[f] = 0;
    s.total(tot);
  }
}

}  // namespace dk
