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
  __shared__ H s_val[CAP];                  // the block's values in slot order
  __shared__ int32_t s_base[kScatterRuns];  // slot bases of the block's first runs
  const int tid = threadIdx.x;
  const int32_t b = xcd_block(blockIdx.x, a.num_blocks, a.xcd_remap);
  const int4 blk = a.blocks[b];  // carries its out-edge range: no csc_ptr round trip first
  const int32_t s0 = blk.x, c0 = blk.y, nnz = blk.z - blk.y;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
  H* __restrict__ hbuf = reinterpret_cast<H*>(a.hbuf);
  const int64_t pitch = a.ring_pitch;

  if (!(blk.w & kBlockLongRow)) {
    if constexpr (LAYOUT == 0) {
      // node-major: the block's out-edges are in source order; map each edge to its source
      const int32_t nsrc = (blk.w >> 1) & 511;
      int32_t rs = 0, re = 0;
      if (tid < nsrc) {
        rs = a.csc_ptr[s0 + tid] - c0;
        re = a.csc_ptr[s0 + tid + 1] - c0;
      }
      // a source with more than kSerialRow out-edges is mapped by its whole wavefront
      const bool heavy = tid < nsrc && re - rs > kSerialRow;
      if (tid < nsrc && !heavy)
        for (int32_t q = rs; q < re; ++q) s_esrc[q] = (uint8_t)tid;
      const int lane = tid & 63;
      for (uint64_t m = __ballot(heavy); m; m &= m - 1) {
        const int src = __ffsll((unsigned long long)m) - 1;
        const int32_t hs = __builtin_amdgcn_readlane(rs, src), he = __builtin_amdgcn_readlane(re, src);
        for (int32_t q = hs + lane; q < he; q += 64) s_esrc[q] = (uint8_t)((tid & ~63) + src);
      }
      __syncthreads();
    }
    // Every load is unconditional (a lane past the block's edges reads out-edge
    // 0 and the block's first source, and discards the values), so all of a
    // lane's loads are in flight at once; with a branch around each one the
    // compiler waited for every history load before issuing the next.
    // the block's runs of consecutive slots: slot-order index k of run r goes
    // to slot run_base[r] + k; the first kScatterRuns bases are staged in LDS
    const int32_t rb = a.run_ptr[b], nr = min(a.run_ptr[b + 1] - rb, kScatterRuns);
    int32_t tb[kScatterRuns / kBlock];
#pragma unroll
    for (int j = 0; j < kScatterRuns / kBlock; ++j) {
      const int32_t r = j * kBlock + tid;
      tb[j] = a.run_base[r < nr ? rb + r : 0];
    }
    // node-major: dd holds the delay; slot-major: the source index in the block << 8 | delay
    int32_t dd[EPT];
    int32_t rk[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      const int32_t c = le < nnz ? c0 + le : 0;
      if constexpr (LAYOUT == 0) {
        dd[k] = __builtin_nontemporal_load(a.delay + c);
      } else {
        dd[k] = le < nnz ? (int32_t)__builtin_nontemporal_load(a.dsrc + c) : 0;
      }
      rk[k] = __builtin_nontemporal_load(a.rank + c);
    }
    // node-major: edges in source order, so the lanes of one source share its
    // history line (read with the streaming hint: only this block reuses it,
    // and it leaves the L2 to the buckets' write fronts).  Slot-major: edges in
    // (delay, source) order, so the lanes of one load read one slot row for
    // consecutive sources.
    H v[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      int32_t slot = cur - (dd[k] & 255);
      slot += (slot < 0) ? S : 0;
      int32_t node;
      if constexpr (LAYOUT == 0) {
        node = s0 + (le < nnz ? (int32_t)s_esrc[le] : 0);
      } else {
        node = s0 + (dd[k] >> 8);
      }
      const H* src = ring + ring_index<LAYOUT>(pitch, node, slot);
      v[k] = LAYOUT == 0 ? __builtin_nontemporal_load(src) : *src;
    }
    // place in slot order
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) s_val[rk[k]] = v[k];
    }
#pragma unroll
    for (int j = 0; j < kScatterRuns / kBlock; ++j) {
      const int32_t r = j * kBlock + tid;
      if (r < nr) s_base[r] = tb[j];
    }
    // each slot-order index's run, loaded once the delays and ranks are no
    // longer live (fewer registers)
    int32_t ri[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      ri[k] = __builtin_nontemporal_load(a.run_idx + (le < nnz ? c0 + le : 0));
    }
    __syncthreads();
    // write in slot order: the values bound for one bucket leave as one contiguous run
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) {
        const int32_t base = ri[k] < kScatterRuns ? s_base[ri[k]] : a.run_base[rb + ri[k]];
        hbuf[base + le] = s_val[le];
      }
    }
  } else {
    // one source with more than CAP out-edges: the whole workgroup strides over
    // them, four edges per lane at a time with unconditional loads
    for (int32_t base = 0; base < nnz; base += 4 * kBlock) {
      int32_t d[4];
      uint32_t p[4];
      H v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        const int32_t c = c0 + (le < nnz ? le : 0);
        d[k] = __builtin_nontemporal_load(a.delay + c);
        p[k] = __builtin_nontemporal_load(a.pos + c);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        int32_t slot = cur - d[k];
        slot += (slot < 0) ? S : 0;
        v[k] = ring[ring_index<LAYOUT>(pitch, s0, slot)];  // a hub's line: reused, cached
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        if (le < nnz) hbuf[p[k]] = v[k];
      }
    }
  }
}

// Regroup pass of the staged schedule: kRegroupChunksPerWg consecutive chunks
// of the buckets per workgroup.  The loads of all its chunks are issued
// first (one memory round trip for the group); then, chunk by chunk, the
// values are placed in LDS in staged-slot order (host rank), each position
// learns its run from a block-wide scan of the run heads, and the values
// leave run by run: consecutive lanes write consecutive staged slots.
template <class H>
__global__ __launch_bounds__(kBlock) void regroup_kernel(RegroupArgs a) {
  constexpr int PER = kRegroupChunk / kBlock;
  constexpr int G = kRegroupChunksPerWg;
  __shared__ H s_val[kRegroupChunk];
  __shared__ uint16_t s_run[kRegroupChunk];  // run heads, then each position's run
  __shared__ int2 s_runs[kRegroupRuns];
  __shared__ int32_t s_wsum[kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const H* __restrict__ src = reinterpret_cast<const H*>(a.src);
  H* __restrict__ dst = reinterpret_cast<H*>(a.dst);
  const int32_t k0 = blockIdx.x * G;
  const int32_t ng = min(G, a.num_chunks - k0);

  int4 ch[G];
  H v[G][PER];
  int32_t rk[G][PER];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    ch[g] = g < ng ? a.chunks[k0 + g] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < ch[g].y) {
        v[g][k] = src[ch[g].x + l];
        rk[g][k] = a.rank[ch[g].x + l];
      }
    }
  }

#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t n = ch[g].y, r0 = ch[g].z, nr = ch[g].w;
    if (g > 0) __syncthreads();  // the previous chunk's values have left the LDS
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < n) s_run[l] = 0;
    }
    for (int32_t r = tid; r < nr; r += kBlock) s_runs[r] = a.runs[r0 + r];
    __syncthreads();
    for (int32_t r = tid; r < nr; r += kBlock) s_run[s_runs[r].x] = 1;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < n) s_val[rk[g][k]] = v[g][k];
    }
    __syncthreads();

    // exclusive scan of the heads: lane tid owns positions [PER * tid, PER * tid + PER)
    int32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t l = PER * tid + j;
      cnt += (l < n) ? (int32_t)s_run[l] : 0;
    }
    int32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    int32_t run = incl - cnt;
    for (int w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t l = PER * tid + j;
      if (l < n) {
        run += s_run[l];
        s_run[l] = (uint16_t)(run - 1);
      }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < n) {
        const int2 rr = s_runs[s_run[l]];
        dst[rr.y + (l - rr.x)] = s_val[l];
      }
    }
  }
}

hipError_t launch_regroup(int history, const RegroupArgs& a, hipStream_t s) {
  if (a.num_chunks <= 0) return hipSuccess;
  const dim3 grid((a.num_chunks + kRegroupChunksPerWg - 1) / kRegroupChunksPerWg), block(kBlock);
  if (history == 0)
    regroup_kernel<float><<<grid, block, 0, s>>>(a);
  else
    regroup_kernel<uint16_t><<<grid, block, 0, s>>>(a);
  return hipGetLastError();
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
