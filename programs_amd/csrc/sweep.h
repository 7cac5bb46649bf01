// Device helpers of the one-launch triangular sweeps (krylov.hip: the
// spray-simulation factor on the slot tables; flowlab.hip: flowlab's factor
// on CSR rows).  Rows are taken in level order in chunks of BS positions; a
// resident grid takes the chunks from a ticket counter, so every row a chunk
// waits for belongs to a chunk taken earlier, by a block that is running or
// has finished.  An output starts as kSweepMark (a signalling NaN: no
// arithmetic result has this pattern) and a row waits until each value it
// needs has been replaced.  Waiting lanes never block their wave: the wave
// loops until all its lanes are finished.  Every wait is bounded
// (kSweepSpin passes per chunk); past the bound the kernel records an error
// and its blocks stop taking chunks.
#pragma once
#include "dev.h"

namespace fvk {

constexpr unsigned long long kSweepMark = 0x7FF4DEADBEEF0001ull;
constexpr unsigned kSweepSpin = 1u << 20;

__device__ __forceinline__ unsigned long long ld_agent(const double *p) {
  return __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v) {
  __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// The chunk loop: row.start(pos) loads the row at level-order position pos,
// row.advance() takes the values that have arrived (in the row's order) and
// returns true once the row's result is written.
template <class Row>
__device__ __forceinline__ void sweep_chunks(int n, int *ticket, int *err, Row &row) {
  __shared__ int s_chunk, s_abort;
  if (threadIdx.x == 0) s_abort = 0;
  for (;;) {
    if (threadIdx.x == 0) s_chunk = atomicAdd(ticket, 1);
    __syncthreads();
    const int chunk = s_chunk;
    const int abort = s_abort;
    __syncthreads();
    if (abort || (long)chunk * BS >= (long)n) return;
    const int pos = chunk * BS + threadIdx.x;
    bool fin = pos >= n;
    if (!fin) row.start(pos);
    unsigned spin = 0, idle = 0;
    while (!__all(fin)) {
      bool moved = false;
      if (!fin) {
        fin = row.advance();
        moved = fin;
      }
      if (++spin > kSweepSpin) {
        if (!fin) { *err = 1; s_abort = 1; }
        break;
      }
      idle = __any(moved) ? 0 : idle + 1;
      if (!__all(fin)) {
        if (idle < 8) __builtin_amdgcn_s_sleep(1);
        else __builtin_amdgcn_s_sleep(8);
      }
    }
    __syncthreads();
  }
}

// at most this many lower (upper) entries per row are kept in registers
constexpr int kSweepSlots = 8;

}  // namespace fvk
