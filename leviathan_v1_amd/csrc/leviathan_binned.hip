// leviathan_binned.hip -- phase 1 of the two-phase ("binned") step, for
// graphs without index locality (Barabasi-Albert and other expander-like
// topologies).
//
// The direct step gathers theta_j(t - d) with one random 2- or 4-byte load
// per edge; on a graph whose edges join random node pairs every such load
// misses the caches and moves a whole 128-byte line (measured: ~124 bytes of
// fetch per edge at N = 10^7).  The binned step splits the gather:
//
//   phase 1, scatter_kernel (this file): walk the graph by SOURCE (CSC order).
//     Lanes on consecutive out-edges of a run of consecutive sources read the
//     sources' history once and write h_e = history[j][t - d_e] into a
//     destination-bucketed array.  The block first places its values in LDS
//     in ascending slot order (a host-computed rank per edge): its values for
//     one bucket are one contiguous run of slots, so they leave as a few
//     whole-run writes rather than one L2 request per edge.  A bucket holds the incoming edges of a run
//     of consecutive rows (about 1 MB of values), ordered by source, so each
//     bucket has one write front per XCD that advances with the source sweep;
//     with a few hundred buckets the fronts stay in L2 until their lines are
//     full and leave as whole lines.
//   phase 2, step_kernel with the bucketed gather (leviathan_kernels.hip): the
//     usual row stream, reading each edge's value through a 32-bit position
//     into its bucket.  Rows are walked in order and each XCD owns a
//     contiguous run of row blocks, so the bucket being read (1 MB) stays in
//     that XCD's 4 MB L2 and is fetched from HBM about once.
//
// Per edge the HBM traffic is delay 1 + rank 2 + position 4 + h write 2|4 (phase 1) and
// position 4 + h read 2|4 + weight read/write 8 (phase 2), plus one pass over
// the history ring, instead of a cache line per edge.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "leviathan_internal.h"
#include "step_kernel.h"

namespace lev {

template <class H, int LAYOUT>
__global__ __launch_bounds__(kBlock) void scatter_kernel(ScatterArgs a) {
  constexpr int EPT = kScatterEpt;
  constexpr int CAP = kBlock * EPT;
  __shared__ uint8_t s_esrc[CAP];
  __shared__ H s_val[CAP];  // the block's values in slot order
  const int tid = threadIdx.x;
  const int32_t b = xcd_block(blockIdx.x, a.num_blocks, a.xcd_remap);
  const int4 blk = a.blocks[b];
  const int32_t s0 = blk.x, s1 = blk.y;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const int32_t c0 = a.csc_ptr[s0];
  const int32_t c1 = a.csc_ptr[s1];
  const int32_t nnz = c1 - c0;
  const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
  H* __restrict__ hbuf = reinterpret_cast<H*>(a.hbuf);
  const int64_t pitch = a.ring_pitch;

  if (!(blk.w & kBlockLongRow)) {
    const int32_t nsrc = s1 - s0;
    if (tid < nsrc) {
      const int32_t rs = a.csc_ptr[s0 + tid] - c0;
      const int32_t re = a.csc_ptr[s0 + tid + 1] - c0;
      for (int32_t q = rs; q < re; ++q) s_esrc[q] = (uint8_t)tid;
    }
    __syncthreads();
    int32_t dd[EPT];
    int32_t rk[EPT];
    uint32_t pp[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) {
        dd[k] = __builtin_nontemporal_load(a.delay + c0 + le);
        rk[k] = __builtin_nontemporal_load(a.rank + c0 + le);
        pp[k] = __builtin_nontemporal_load(a.pos + c0 + le);
      }
    }
    // read by source (lanes of one source share its history line), place in slot order
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) {
        int32_t slot = cur - dd[k];
        slot += (slot < 0) ? S : 0;
        // node-major: a source's history line is only reused by this block's lanes, so it is
        // read with the streaming hint and leaves the L2 to the buckets' write fronts
        const H* src = ring + ring_index<LAYOUT>(pitch, s0 + s_esrc[le], slot);
        s_val[rk[k]] = LAYOUT == 0 ? __builtin_nontemporal_load(src) : *src;
      }
    }
    __syncthreads();
    // write in slot order: the values bound for one bucket leave as one contiguous run
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) hbuf[pp[k]] = s_val[le];
    }
  } else {
    // one source with more than CAP out-edges: the whole workgroup strides over them
    for (int32_t le = tid; le < nnz; le += kBlock) {
      const int32_t d = __builtin_nontemporal_load(a.delay + c0 + le);
      const uint32_t p = __builtin_nontemporal_load(a.pos + c0 + le);
      int32_t slot = cur - d;
      slot += (slot < 0) ? S : 0;
      hbuf[p] = ring[ring_index<LAYOUT>(pitch, s0, slot)];  // a hub's line: reused, cached
    }
  }
}

hipError_t launch_scatter(int history, int layout, const ScatterArgs& a, hipStream_t s) {
  if (a.num_blocks <= 0) return hipSuccess;
  const dim3 grid(a.num_blocks), block(kBlock);
  if (history == 0) {
    if (layout == 0)
      scatter_kernel<float, 0><<<grid, block, 0, s>>>(a);
    else
      scatter_kernel<float, 1><<<grid, block, 0, s>>>(a);
  } else {
    if (layout == 0)
      scatter_kernel<uint16_t, 0><<<grid, block, 0, s>>>(a);
    else
      scatter_kernel<uint16_t, 1><<<grid, block, 0, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace lev
