// isp_device.hpp - device helpers shared by the kernel files: wave sums and
// the fence-free grid completion of the persistent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "isp_internal.hpp"

namespace isp_amd {

// Wave64 sum of a 64-bit value.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
  return v;
}

// Grid-wide completion without cache-wide fences. Blocks publish their
// partial sums with device-scope atomics whose results each thread consumes
// (folded into `token`), so every atomic has been performed at the point of
// coherence before the thread reaches the barrier; only then does one thread
// count the block's contribution. The block that completes the count is the
// last one: it reads the totals back and clears them for the next frame. (An
// agent-scope fence here would write back and invalidate the whole L2 of the
// block's XCD.) `flag` is a block-shared int.
//
// complete_items: `count` of `total` contributions counted on one counter.
__device__ __forceinline__ bool complete_items(unsigned int* done, unsigned int count, unsigned int total,
                                               unsigned int token, int* flag) {
  // Consuming the token makes the atomics return their values and the wave wait
  // for them here (the asm reads a register and touches no memory).
  __asm__ volatile("" : : "v"(token));
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int before = atomicAdd(done, count);
    *flag = before + count == total ? 1 : 0;
  }
  __syncthreads();
  return *flag != 0;
}

// complete_grid: one contribution per block of the grid, counted in two levels
// so that no counter takes more than a kSlots-th of the blocks: block b counts
// on done[b % kSlots]; the block that completes its slot's group (whose size
// the grid fixes) counts on `top`; the one that completes `top` is the last.
__device__ __forceinline__ bool complete_grid(LineCounter* done, LineCounter* top, unsigned int token, int* flag) {
  __asm__ volatile("" : : "v"(token));
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int g = gridDim.x;
    const unsigned int x = blockIdx.x % kSlots;
    const unsigned int groups = g < kSlots ? g : kSlots;
    const unsigned int members = (g - x + kSlots - 1) / kSlots;
    int last = 0;
    if (atomicAdd(&done[x].n, 1U) + 1U == members) last = atomicAdd(&top->n, 1U) + 1U == groups ? 1 : 0;
    *flag = last;
  }
  __syncthreads();
  return *flag != 0;
}

// The last block's reset of complete_grid's counters (every block has counted).
__device__ __forceinline__ void reset_grid(LineCounter* done, LineCounter* top) {
  if (threadIdx.x < kSlots) atomicExch(&done[threadIdx.x].n, 0U);
  if (threadIdx.x == 0) atomicExch(&top->n, 0U);
}

// Reads and clears the kSlots copies of an accumulator, returning their sum.
template <typename T>
__device__ __forceinline__ T take_slots(T* first, std::size_t stride) {
  T total = 0;
#pragma unroll
  for (int x = 0; x < kSlots; ++x) total += atomicExch(first + x * stride, T(0));
  return total;
}

__device__ __forceinline__ unsigned int fold(unsigned long long v) {
  return static_cast<unsigned int>(v) ^ static_cast<unsigned int>(v >> 32);
}

}  // namespace isp_amd
