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
// Per edge of a short source block the HBM traffic is source-and-delay code 2
// (delay 1 on the node-major ring) + rank and run index 4 + h write 2|4
// (phase 1) and position 4 + h read 2|4 + weight read/write 8 (phase 2), plus
// one pass over the history ring and the block's run bases, instead of a cache
// line per edge.  Edges of sources beyond one block read delay 1 + position 4.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "leviathan_internal.h"
#include "step_kernel.h"

namespace lev {

// slot of theta(t + k - d) on a ring of S slots, cur = t % S, 0 <= d <= S - 2
// and 0 <= k < S (the host allows at most S steps per pass), so that
// -S < cur + k - d < 2 * S
__device__ __forceinline__ int32_t ahead_slot(int32_t cur, int32_t k, int32_t d, int32_t S) {
  int32_t slot = cur + k - d;
  slot += (slot < 0) ? S : 0;
  slot -= (slot >= S) ? S : 0;
  return slot;
}

// How a short source block reads the history ring (MODE of scatter_kernel):
// 64-bit element indices, 32-bit byte offsets from the ring's base, or (slot-
// major rings of at most scatter_window_slots slots) a window of the ring in LDS.
constexpr int kScatterGather64 = 0, kScatterOff32 = 1, kScatterWindow = 2;

// K = number of consecutive steps whose values one pass gathers (1, 2 or 4):
// slot p of hbuf holds the K values of its edge side by side (HVec), step
// t + k's in component k.  Values of edges with d < k are read from slots
// this pass must not trust (not yet written for step t + k); early_kernel
// overwrites them before that step's row pass.
// kScatterOff32 (node-major rings below 2^32 bytes): each history read is a
// 32-bit byte offset from the ring's base (the scalar-base form of the load),
// one multiply per edge and then one slot step per further step with one wrap
// test, in place of a 64-bit multiply per read.  At BA N = 10^7, uint16, four
// steps per pass: node-major 2.40 -> 2.20 ms per pass.
// kScatterWindow (slot-major rings): the block's sources' window of the whole
// ring is loaded into LDS with coalesced 16-byte loads that depend only on the
// block descriptor, and the edges read their values there.  At BA N = 10^7,
// uint16, four steps per pass: 0.401 -> 0.332 ms per step for the scatter and
// its early gathers (the gathers issued 64 dependent 2-byte loads per lane).
template <class H, int LAYOUT, int K, int MODE>
__global__ __launch_bounds__(kBlock) void scatter_kernel(ScatterArgs a) {
  // a source block holds up to CAP out-edges, EPT per lane.  (With K = 4 a
  // 512-lane workgroup, 8 edges per lane at 92 instead of 164 VGPRs, took
  // 1.92 ms per pass at BA N = 10^7, u16, against 1.82 ms with 256 lanes.)
  using V = HVec<H, K>;
  constexpr int T = kBlock;
  constexpr int CAP = kBlock * kScatterEpt;
  constexpr int EPT = CAP / T;
  static_assert(MODE != kScatterWindow || LAYOUT == 1, "the window form reads a slot-major ring");
  static_assert(MODE != kScatterOff32 || LAYOUT == 0, "32-bit offsets are used for node-major rings");
  constexpr int WSLOTS = scatter_window_slots(sizeof(H));
  // the block's value vectors in slot order (gathers); the window form's window of the ring is
  // dynamic shared memory of S * kScatterWinCols elements
  constexpr size_t kRawBytes = MODE == kScatterWindow ? 16 : sizeof(V) * (CAP + 1);
  __shared__ uint8_t s_esrc[LAYOUT == 0 ? CAP : 1];  // node-major: each edge's source in the block
  __shared__ __attribute__((aligned(16))) char s_raw[kRawBytes];
  extern __shared__ __attribute__((aligned(16))) char s_dyn[];
  V* const s_val = reinterpret_cast<V*>(s_raw);
  __shared__ int32_t s_base[kScatterRuns];  // slot bases of the block's first runs
  const int tid = threadIdx.x;
  const int32_t b = xcd_block(blockIdx.x, a.num_blocks, a.xcd_remap);
  const int4 blk = a.blocks[b];  // carries its out-edge range: no csc_ptr round trip first
  const int32_t s0 = blk.x, c0 = blk.y, nnz = blk.z - blk.y;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
  V* __restrict__ hbuf = reinterpret_cast<V*>(a.hbuf);
  const int64_t pitch = a.ring_pitch;

  if (!(blk.w & kBlockLongRow)) {
    if constexpr (MODE == kScatterWindow) {
      // Window form (slot-major ring).  The block's out-edges are stored in the
      // order of their value slots, so consecutive lanes write consecutive slots
      // of a run straight from registers: no rank, no placement in LDS.  Each
      // edge's code holds its source in the block << 8 | delay (low half) and
      // its run of consecutive slots (high half).  The block's sources
      // [s0, s0 + nsrc) in every slot row -- an 8-aligned window of
      // kScatterWinCols columns per slot -- come in with coalesced 16-byte
      // loads that wait for nothing but the block's descriptor, and each edge
      // reads its K values from LDS.  The ring holds kScatterWinCols elements of
      // padding past its end for the last slot row's window.
      H* const s_win = reinterpret_cast<H*>(s_dyn);
      constexpr int EPC = 16 / (int)sizeof(H);          // elements per 16-byte chunk
      constexpr int CPR = kScatterWinCols / EPC;        // chunks per window row
      constexpr int NCH = (WSLOTS * CPR + T - 1) / T;   // chunk loads per lane at most
      const int32_t rb = a.run_ptr[b], nr_all = a.run_ptr[b + 1] - rb, nr = min(nr_all, kScatterRuns);
      int32_t tb[kScatterRuns / T];
#pragma unroll
      for (int j = 0; j < kScatterRuns / T; ++j) {
        const int32_t r = j * T + tid;
        tb[j] = a.run_base[r < nr ? rb + r : 0];
      }
      uint32_t code[EPT];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * T + tid;
        code[k] = __builtin_nontemporal_load(a.rank_run + (uint32_t)(le < nnz ? c0 + le : 0));
      }
      // (Loading only the chunks that hold the block's sources, and 16-byte
      // loads typed as a clang vector rather than uint4, each made the pass
      // slower: 0.338 -> 0.361 and -> 0.406 ms per step with the early gathers.)
      const int32_t w0 = s0 & ~7;
      const int32_t nch = S * CPR;
      uint4 ch[NCH];
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        const int32_t id = j * T + tid < nch ? j * T + tid : 0;
        const int32_t r = id / CPR, c = id % CPR;
        ch[j] = *reinterpret_cast<const uint4*>(ring + (int64_t)r * pitch + w0 + c * EPC);
      }
      // Ring slot r goes to window row (r - cur - 1) mod S.  Step t + q's value
      // of an edge with delay d, ring slot cur + q - d, is then row S - 1 - d + q
      // for every q <= d: an edge's K values lie in consecutive rows from
      // S - 1 - d, with no wrap to compute.  For q > d the value does not exist
      // yet (an early edge, gathered later into the array the row pass reads):
      // its row is clamped to S - 1, inside the window.  (Per value this is an
      // add and a minimum where the wrapped slot took about ten instructions.)
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        const int32_t id = j * T + tid;
        if (id < nch) {
          int32_t row = id / CPR - cur - 1;
          row += (row < 0) ? S : 0;
          *reinterpret_cast<uint4*>(s_win + row * kScatterWinCols + (id % CPR) * EPC) = ch[j];
        }
      }
#pragma unroll
      for (int j = 0; j < kScatterRuns / T; ++j) {
        const int32_t r = j * T + tid;
        if (r < nr) s_base[r] = tb[j];
      }
      __syncthreads();
      // the last window row at the block's first source
      const int32_t top = (S - 1) * kScatterWinCols + (s0 - w0);
      V pv[EPT];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t last = top + (int32_t)((code[k] >> 8) & 255u);         // row S - 1, the edge's column
        const int32_t first = last - (int32_t)(code[k] & 255u) * kScatterWinCols;  // row S - 1 - d
#pragma unroll
        for (int q = 0; q < K; ++q) pv[k].x[q] = s_win[min(first + q * kScatterWinCols, last)];
      }
      // the block's runs of consecutive slots: index le of run r goes to slot run_base[r] + le
      // (a lane past the block's edges loaded edge 0's code and stores nothing)
      if (nr_all <= kScatterRuns) {
        int32_t at[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) at[k] = s_base[(code[k] >> 16) & (kScatterRuns - 1)] + k * T + tid;
#pragma unroll
        for (int k = 0; k < EPT; ++k)
          if (k * T + tid < nnz) hbuf[at[k]] = pv[k];
      } else {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * T + tid;
          if (le < nnz) hbuf[a.run_base[rb + (int32_t)(code[k] >> 16)] + le] = pv[k];
        }
      }
    } else {
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
      const int32_t rb = a.run_ptr[b], nr_all = a.run_ptr[b + 1] - rb, nr = min(nr_all, kScatterRuns);
      int32_t tb[kScatterRuns / T];
#pragma unroll
      for (int j = 0; j < kScatterRuns / T; ++j) {
        const int32_t r = j * T + tid;
        tb[j] = a.run_base[r < nr ? rb + r : 0];
      }
      // node-major: dd holds the delay; slot-major: the source index in the block << 8 | delay
      int32_t dd[EPT];
      uint32_t rk[EPT];  // rank | run index << 16
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * T + tid;
        // unsigned 32-bit element indices: the loads take the scalar-base form
        const uint32_t c = (uint32_t)(le < nnz ? c0 + le : 0);
        if constexpr (LAYOUT == 0) {
          dd[k] = __builtin_nontemporal_load(a.delay + c);
        } else {
          const int32_t code = __builtin_nontemporal_load(a.dsrc + c);
          dd[k] = le < nnz ? code : 0;
        }
        rk[k] = __builtin_nontemporal_load(a.rank_run + c);
      }
      // node-major: edges in source order, so the lanes of one source share its
      // history line (read with the streaming hint: only this block reuses it,
      // and it leaves the L2 to the buckets' write fronts).  Slot-major: edges
      // in (delay, source) order, so the lanes of one load read one slot row for
      // consecutive sources; for q > 0 those rows were read for q - 1 by the
      // neighbouring delay and are in cache.
      // the K steps' values of every edge, all loads in flight at once, packed per edge
      V pv[EPT];
      if constexpr (MODE == kScatterOff32) {
        // slot of step t + q is slot0 + q, less S once it passes the ring's end
        const char* base = reinterpret_cast<const char*>(a.ring);
        const uint32_t wrap_bytes = (uint32_t)S * (uint32_t)sizeof(H);  // a whole turn of the ring
        uint32_t off0[EPT];
        int32_t slot0[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * T + tid;
          const int32_t d = dd[k] & 255;
          slot0[k] = cur - d + (cur < d ? S : 0);  // slot of step t: -S < cur - d < S
          const int32_t node = s0 + (le < nnz ? (int32_t)s_esrc[le] : 0);
          off0[k] = ((uint32_t)node * (uint32_t)pitch + (uint32_t)slot0[k]) * (uint32_t)sizeof(H);
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
          H v[EPT];
#pragma unroll
          for (int k = 0; k < EPT; ++k) {
            const uint32_t off = off0[k] + (uint32_t)(q * sizeof(H)) - (slot0[k] >= S - q ? wrap_bytes : 0u);
            v[k] = __builtin_nontemporal_load(reinterpret_cast<const H*>(base + off));
          }
#pragma unroll
          for (int k = 0; k < EPT; ++k) pv[k].x[q] = v[k];
        }
      } else {
        // One step's loads go into registers of their own and are packed
        // afterwards.  (Small changes of this code's shape moved the pass a lot:
        // loading straight into the vectors made the compiler issue the uint16
        // history loads one at a time, 2.2 instead of 1.6 ms per four-step
        // pass, and a lambda that returned the source node took the FP32
        // two-step pass from 1.33 to 1.51 ms, both at BA N = 10^7.)
        auto load_values = [&](H (&v)[EPT], int q) {
#pragma unroll
          for (int k = 0; k < EPT; ++k) {
            const int32_t le = k * T + tid;
            const int32_t slot = ahead_slot(cur, q, dd[k] & 255, S);
            int32_t node;
            if constexpr (LAYOUT == 0) {
              node = s0 + (le < nnz ? (int32_t)s_esrc[le] : 0);
            } else {
              node = s0 + (dd[k] >> 8);
            }
            const H* src = ring + ring_index<LAYOUT>(pitch, node, slot);
            v[k] = LAYOUT == 0 ? __builtin_nontemporal_load(src) : *src;
          }
        };
#pragma unroll
        for (int q = 0; q < K; ++q) {
          H v[EPT];
          load_values(v, q);
#pragma unroll
          for (int k = 0; k < EPT; ++k) pv[k].x[q] = v[k];
        }
      }
      // place in slot order; a lane past the block's edges writes the spare
      // slot CAP, so that no load is sunk into a branch and waited for alone
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * T + tid;
        s_val[le < nnz ? (rk[k] & 0xffff) : CAP] = pv[k];
      }
#pragma unroll
      for (int j = 0; j < kScatterRuns / T; ++j) {
        const int32_t r = j * T + tid;
        if (r < nr) s_base[r] = tb[j];
      }
      __syncthreads();
      // write in slot order: the values bound for one bucket leave as one contiguous run.
      // The branch on the run count is uniform over the block: with a per-lane
      // choice between s_base and run_base the compiler read both through one
      // flat load of a selected pointer, and every store then waited for all
      // earlier loads and stores (vmcnt(0)).
      // All slots and values are read from LDS first, into registers of their
      // own, and the stores follow: with one register set reused per store,
      // each store's source registers were guarded by a wait for every
      // earlier store (vmcnt(0)).
      if (nr_all <= kScatterRuns) {
        int32_t at[EPT];
        V val[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * T + tid;  // < CAP: every LDS read is in range
          at[k] = s_base[(rk[k] >> 16) & (kScatterRuns - 1)] + le;
          val[k] = s_val[le];
        }
#pragma unroll
        for (int k = 0; k < EPT; ++k)
          if (k * T + tid < nnz) hbuf[at[k]] = val[k];
      } else {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * T + tid;
          if (le < nnz) hbuf[a.run_base[rb + (int32_t)(rk[k] >> 16)] + le] = s_val[le];
        }
      }
    }
  } else {
    // one source with more than CAP out-edges: the whole workgroup strides over
    // them, four edges per lane at a time with unconditional loads
    for (int32_t base = 0; base < nnz; base += 4 * T) {
      int32_t d[4];
      uint32_t p[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * T + tid;
        const int32_t c = c0 + (le < nnz ? le : 0);
        d[k] = __builtin_nontemporal_load(a.delay + c);
        p[k] = __builtin_nontemporal_load(a.pos + c);
      }
      V v[4];
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          v[k].x[q] = ring[ring_index<LAYOUT>(pitch, s0, ahead_slot(cur, q, d[k], S))];  // a hub's line: reused, cached
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * T + tid;
        if (le < nnz) hbuf[p[k]] = v[k];
      }
    }
  }
}

// The values of the edges with delay < k for step t + k of a multi-step
// scatter: edge (j, d) reads theta_j(t + k - d), which step t + k - d - 1
// writes.  The gather runs once step t + k - 1 is done, or (step_off = 1, a
// graph without delay-0 edges) once step t + k - 2 is.  One lane per edge, in
// (delay, source) order so that the lanes of a load read one slot row for
// ascending sources.
template <class H, int LAYOUT>
__global__ __launch_bounds__(kBlock) void early_kernel(EarlyArgs a) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.count) return;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)((*a.step + a.step_off) % S);
  int32_t slot = cur - (int32_t)a.delay[i];
  slot += (slot < 0) ? S : 0;
  const H* ring = reinterpret_cast<const H*>(a.ring);
  reinterpret_cast<H*>(a.hbuf)[(int64_t)a.pos[i] * a.hstride] = ring[ring_index<LAYOUT>(a.ring_pitch, a.src[i], slot)];
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

// One vector of a scatter pass's uint16 values (4 or 8 bytes) loaded with the
// streaming (nontemporal) hint.
template <class V>
__device__ __forceinline__ V stream_load(const V* p) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "a 4- or 8-byte vector");
  V v;
  if constexpr (sizeof(V) == 4) {
    const uint32_t raw = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
    __builtin_memcpy(&v, &raw, 4);
  } else {
    const uint64_t raw = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
    __builtin_memcpy(&v, &raw, 8);
  }
  return v;
}

// Regroup of the values of K consecutive steps at once (several steps per
// scatter pass): one chunk per workgroup.  Each bucket slot holds its edge's
// K values side by side, so one load moves all K; the rank, the run heads'
// scan and each position's staged slot are computed once, and the K values
// leave to the K steps' staged arrays.
// PACK (uint16 values, paired row passes): the values of steps 2p and 2p + 1
// leave together as one 4-byte pair to pair array p ([K / 2][stride] pairs),
// which the paired row pass reads with one load per slot for both steps: K / 2
// rounds of placement and stores instead of K.
template <class H, int K, bool PACK>
__global__ __launch_bounds__(kBlock) void regroup_multi_kernel(RegroupArgs a) {
  using V = HVec<H, K>;
  static_assert(!PACK || (sizeof(H) == 2 && K % 2 == 0), "pairs of uint16 values");
  constexpr int PER = kRegroupChunk / kBlock;
  // The values are placed in LDS one step (PACK: one pair of steps) at a time,
  // two barriers per round.  (Placing whole vectors, one LDS write and read per
  // position, needs 32 KB per workgroup at two uint16 steps: 4 instead of 6
  // workgroups per CU, and the pass took 0.71 instead of 0.61 ms at BA N = 10^7.)
  // PACK's 16 KB of pairs fill the 32 KB that five workgroups per CU leave each
  // with the run heads and records, so the scan's wavefront sums borrow the
  // pair buffer before the first round is placed.
  using P = typename std::conditional<PACK, uint32_t, H>::type;
  __shared__ P s_val[kRegroupChunk];          // one round's values (pairs) in staged-slot order
  __shared__ uint16_t s_run[kRegroupChunk];  // run heads, then each position's run
  __shared__ int2 s_runs[kRegroupRuns];
  __shared__ int32_t s_wsum_own[PACK ? 1 : kBlock / 64];
  int32_t* const s_wsum = PACK ? reinterpret_cast<int32_t*>(s_val) : s_wsum_own;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const V* __restrict__ src = reinterpret_cast<const V*>(a.src);
  // each XCD takes a contiguous run of chunks: the few chunks that fill one
  // staged line (consecutive source ranges of a bucket) meet in one L2.
  // Measured at BA N = 10^7, u16, two steps: 0.83 -> 0.72 ms, writes 1.56 ->
  // 1.17 GB per pass.  (The one-step regroup, two chunks per workgroup, got
  // no faster from it: 0.45 ms, writes 0.78 -> 0.86 GB.)
  const int4 ch = a.chunks[xcd_block(blockIdx.x, a.num_chunks, 1)];
  const int32_t n = ch.y, r0 = ch.z, nr = ch.w;

  // unconditional loads: a lane past the chunk reads element 0 and discards it.
  // uint16 values: the chunk's values and ranks are read once, with the
  // streaming hint, and then leave the L2 to the staged lines that neighbouring
  // chunks fill in turn (BA N = 10^7, four steps per pass: 0.224 -> 0.203 ms
  // per step).  FP32 values read without it (the hint: 0.492 -> 0.519), and the
  // same hint on the scatter's window loads or the row pass's weight loads made
  // those passes slower (uint16: 0.243 -> 0.353 and 0.395 -> 0.450).
  V v[PER];
  int32_t rk[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int32_t l = k * kBlock + tid;
    const int64_t c = l < n ? (int64_t)ch.x + l : 0;
    if constexpr (sizeof(H) == 2) {
      rk[k] = __builtin_nontemporal_load(a.rank + c);
      v[k] = stream_load(src + c);
    } else {
      rk[k] = a.rank[c];
      v[k] = src[c];
    }
  }
  // round q's value (PACK: pair) of position k
  auto value = [&](int k, int q) -> P {
    if constexpr (PACK)
      return (uint32_t)v[k].x[2 * q] | ((uint32_t)v[k].x[2 * q + 1] << 16);
    else
      return v[k].x[q];
  };
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int32_t l = k * kBlock + tid;
    if (l < n) s_run[l] = 0;
  }
  for (int32_t r = tid; r < nr; r += kBlock) s_runs[r] = a.runs[r0 + r];
  __syncthreads();
  for (int32_t r = tid; r < nr; r += kBlock) s_run[s_runs[r].x] = 1;
  if constexpr (!PACK) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < n) s_val[rk[k]] = value(k, 0);
    }
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
  __syncthreads();  // (PACK: every lane has read the wavefront sums from the pair buffer)

  // each position's staged slot, then the rounds' stores
  int32_t off[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int32_t l = k * kBlock + tid;
    off[k] = 0;
    if (l < n) {
      const int2 rr = s_runs[s_run[l]];
      off[k] = rr.y + (l - rr.x);
    }
  }
#pragma unroll
  for (int q = 0; q < (PACK ? K / 2 : K); ++q) {
    if (q > 0 || PACK) {
      if (q > 0) __syncthreads();  // the previous round's values have left the LDS
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int32_t l = k * kBlock + tid;
        if (l < n) s_val[rk[k]] = value(k, q);
      }
      __syncthreads();
    }
    P* __restrict__ out = reinterpret_cast<P*>(a.dst) + q * a.stride;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t l = k * kBlock + tid;
      if (l < n) out[off[k]] = s_val[l];
    }
  }
}

template <int K>
static void regroup_multi_launch(int history, const RegroupArgs& a, hipStream_t s) {
  const dim3 grid(a.num_chunks), block(kBlock);
  if (history == 0)
    regroup_multi_kernel<float, K, false><<<grid, block, 0, s>>>(a);
  else if (a.pack_pairs)
    regroup_multi_kernel<uint16_t, K, true><<<grid, block, 0, s>>>(a);
  else
    regroup_multi_kernel<uint16_t, K, false><<<grid, block, 0, s>>>(a);
}

hipError_t launch_regroup(int history, const RegroupArgs& a, hipStream_t s) {
  if (a.num_chunks <= 0) return hipSuccess;
  if (a.pack_pairs && (history == 0 || a.steps % 2 != 0)) return hipErrorInvalidValue;
  switch (a.steps) {
    case 1: {
      const dim3 grid((a.num_chunks + kRegroupChunksPerWg - 1) / kRegroupChunksPerWg), block(kBlock);
      if (history == 0)
        regroup_kernel<float><<<grid, block, 0, s>>>(a);
      else
        regroup_kernel<uint16_t><<<grid, block, 0, s>>>(a);
      break;
    }
    case 2: regroup_multi_launch<2>(history, a, s); break;
    case kMaxScatterSteps: regroup_multi_launch<kMaxScatterSteps>(history, a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class H, int K>
static void scatter_launch(int layout, const ScatterArgs& a, hipStream_t s) {
  const dim3 grid(a.num_blocks), block(kBlock);
  const bool off32 = a.ring_bytes < (int64_t(1) << 32);
  // (the 32-bit offset gathers were slower than the 64-bit ones on a slot-major ring: BA N = 10^7, FP32,
  // two steps per pass, 0.77 against 0.66 ms per step)
  if (scatter_uses_window(layout, a.n_slots, sizeof(H)))
    scatter_kernel<H, 1, K, kScatterWindow>
        <<<grid, block, (size_t)a.n_slots * kScatterWinCols * sizeof(H), s>>>(a);
  else if (layout != 0)
    scatter_kernel<H, 1, K, kScatterGather64><<<grid, block, 0, s>>>(a);
  else if (off32)
    scatter_kernel<H, 0, K, kScatterOff32><<<grid, block, 0, s>>>(a);
  else
    scatter_kernel<H, 0, K, kScatterGather64><<<grid, block, 0, s>>>(a);
}

template <int K>
static void scatter_launch(int history, int layout, const ScatterArgs& a, hipStream_t s) {
  if (history == 0)
    scatter_launch<float, K>(layout, a, s);
  else
    scatter_launch<uint16_t, K>(layout, a, s);
}

hipError_t launch_scatter(int history, int layout, const ScatterArgs& a, hipStream_t s) {
  if (a.num_blocks <= 0) return hipSuccess;
  switch (a.steps) {
    case 1: scatter_launch<1>(history, layout, a, s); break;
    case 2: scatter_launch<2>(history, layout, a, s); break;
    case kMaxScatterSteps: scatter_launch<kMaxScatterSteps>(history, layout, a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_early(int history, int layout, const EarlyArgs& a, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const dim3 grid((a.count + kBlock - 1) / kBlock), block(kBlock);
  if (history == 0) {
    if (layout == 0)
      early_kernel<float, 0><<<grid, block, 0, s>>>(a);
    else
      early_kernel<float, 1><<<grid, block, 0, s>>>(a);
  } else {
    if (layout == 0)
      early_kernel<uint16_t, 0><<<grid, block, 0, s>>>(a);
    else
      early_kernel<uint16_t, 1><<<grid, block, 0, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace lev
