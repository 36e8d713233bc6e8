// leviathan_kernels.hip -- gfx950 kernels of the fused LEVIATHAN step.
//
// One step of the reference (src/cuda/leviathan_kernel.cu:391-438) is four
// launches: prep (:93-111), csr_physics (:114-167), integrate_and_reduce
// (:170-234) and final_reduce_r (:237-264), plus a host-side gain update
// (:430-435) behind a stream synchronisation.  Here one step is two launches:
//
//   step_kernel      prep + delayed gather + sin coupling + Hebbian update +
//                    integrate + history write + FP64 partial sums of
//                    cos/sin, for one row block per 256-lane workgroup;
//   finalize_kernel  FP64 sum of the partials of each network, r, the FP64
//                    gain controller, the step counter and, when a hub
//                    statistics handle is attached, the hub state of the
//                    new phases.
//
// The fusion is legal because a row only writes its own state (theta,
// theta_hat, omega_hat, its incoming edges' weights and its own history
// slot), and the history ring has S = max_delay + 2 slots: the slot written
// in step t, (t+1) % S, is never one a delay 0..max_delay reads in step t.
// The reference needs the kernel boundary because its ring has max_delay+1
// slots and the write slot aliases the max_delay read slot.
//
// Every stream is read and written with plain loads and stores: on the
// 10^7-node graph the streaming (nontemporal) hint made the row pass no
// faster (0-7 % slower across runs).
//
// Row blocks (built on the host) are either a run of up to 256 consecutive
// rows holding at most 256*EPT incoming edges in total (stream mode: the
// edges of the whole run are streamed with consecutive lanes on consecutive
// edges, so column indices, delays and weights move in full cache lines
// whatever the degrees), or a single row with more edges (long-row mode:
// the whole workgroup strides over that row).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "leviathan_internal.h"
#include "step_kernel.h"

namespace lev {

// GATHER = 0: the delayed phase of edge e is read from the history ring at
// (col[e], t - delay[e]).  GATHER = 1: it was written by the scatter kernel
// to hbuf[hpos[e]] (binned schedule); col and delay are not read.
// GATHER = 2 (staged schedule): the scatter kernel wrote the values of a row
// block's edges into the block's own range [e0, e1) of hbuf, ordered by
// source.  The block reads that range with coalesced loads into LDS and picks
// edge e's value at e0 + lpos[e] there; a long row's range is in CSR order and
// is read straight.
struct Item {
  int32_t r0, e0, nnz, nrows, net;
  bool is_long;
};

__device__ __forceinline__ Item unpack_item(int4 v) {
  Item it;
  it.r0 = v.x;
  it.e0 = v.y;
  it.nnz = v.z - v.y;
  it.nrows = (v.w >> 1) & 511;
  it.is_long = (v.w & kBlockLongRow) != 0;
  it.net = v.w >> 10;
  return it;
}

template <class H, int LAYOUT, int EPT, int GATHER>
__global__ __launch_bounds__(kBlock) void step_kernel(StepArgs a) {
  constexpr int CAP = kBlock * EPT;
  __shared__ float s_theta[kBlock];
  __shared__ float s_gate[kBlock];
  __shared__ uint8_t s_erow[CAP];
  __shared__ float s_contrib[CAP];
  __shared__ double s_red[2][kBlock / 64];

  const int tid = threadIdx.x;
  // Row blocks to workgroups.  The binned row pass reads its rows' bucket, so
  // each XCD takes one contiguous run of blocks and the bucket stays in its L2.
  // The other row passes take the blocks in order, dealt round-robin, so that
  // the eight XCDs stream through neighbouring addresses: BA N = 10^7, uint16,
  // staged/8 0.603 -> 0.565 ms; the FP32 direct/8 drop-in 4.73 -> 4.66 ms; the
  // 256-network Watts-Strogatz ensemble, direct/16, 1.26 -> 1.24 ms (its
  // gathers' locality between neighbouring rows did not need the runs).
  int32_t b;
  if constexpr (GATHER != 1) {
    b = (int32_t)blockIdx.x;
    if (b >= a.num_items) return;
  } else {
    const int32_t x = blockIdx.x % kNumXcd;
    b = a.xcd_split[x] + (int32_t)(blockIdx.x / kNumXcd);
    if (b >= a.xcd_split[x + 1]) return;  // this XCD's run is shorter than the longest
  }
  const Item it = unpack_item(a.items[b]);  // carries its edge range: no row_ptr round trip first
  if (it.nrows == 0) {  // padding item
    if (threadIdx.x == 0) a.partials[b] = make_double2(0.0, 0.0);
    return;
  }
  const int32_t r0 = it.r0, r1 = it.r0 + it.nrows, net = it.net;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const int32_t nxt = (cur + 1 == S) ? 0 : cur + 1;
  const float g = (float)a.ctrl[net].g;
  const float dt = a.dt;
  const int32_t e0 = it.e0;
  const int32_t nnz = it.nnz;
  const int32_t nrows = r1 - r0;
  const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
  const H* __restrict__ hbuf = reinterpret_cast<const H*>(a.hbuf);
  const int64_t pitch = a.ring_pitch;
  float u_t = 0.0f;
  if (a.drive_b) u_t = a.drive_u[step % kDriveSteps];

  double my_c = 0.0, my_s = 0.0;

  if (!it.is_long) {
    // ---------------- stream mode ----------------
    RowState rs_;
    int32_t rs = 0, re = 0;
    float om = 0.0f;  // natural frequency (plus drive), loaded with the row's state: a load after
                      // the last barrier would add a memory round trip to every row block
    if (tid < nrows) {
      const int32_t i = r0 + tid;
      om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[tid] = sizeof(H) == 2 ? rs_.th * kInvTwoPi : rs_.th;  // uint16: revolutions (edge_update_rev)
      s_gate[tid] = rs_.gate;
      rs = a.row_ptr[i] - e0;
      re = a.row_ptr[i + 1] - e0;
    }
    // Rows of up to kSerialRow edges are handled by their own lane, longer
    // ones by their whole wavefront, one after another (a lane looping over a
    // long row would hold its wavefront for the row's whole length).
    const int lane = tid & 63;
    const bool heavy = tid < nrows && re - rs > kSerialRow;
    if (tid < nrows && !heavy)
      for (int32_t q = rs; q < re; ++q) s_erow[q] = (uint8_t)tid;
    for (uint64_t m = __ballot(heavy); m; m &= m - 1) {
      const int src = __ffsll((unsigned long long)m) - 1;
      const int32_t hs = __builtin_amdgcn_readlane(rs, src), he = __builtin_amdgcn_readlane(re, src);
      for (int32_t q = hs + lane; q < he; q += 64) s_erow[q] = (uint8_t)((tid & ~63) + src);
    }

    // The loads are unconditional: a lane past the block's edges reads edge 0
    // (every edge array holds at least one element) and discards the value.
    // Straight-line code keeps all of a lane's loads in flight at once; with a
    // branch around each load the compiler waited for every load before
    // issuing the next.
    int32_t jj[EPT];
    int32_t dd[EPT];
    float ww[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      const int32_t e = le < nnz ? e0 + le : 0;
      if constexpr (GATHER == 0) {
        jj[k] = a.col[e];
        dd[k] = a.delay[e];
      } else if constexpr (GATHER == 1) {
        jj[k] = (int32_t)a.hpos[e];
      } else {
        jj[k] = (int32_t)a.lpos[e];
      }
      ww[k] = a.w[e];
    }
    if constexpr (GATHER == 2) {
      // s_contrib holds the block's staged values until the edge pass below
      H sv[EPT];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * kBlock + tid;
        sv[k] = hbuf[le < nnz ? e0 + le : 0];
      }
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * kBlock + tid;
        if (le < nnz) s_contrib[le] = sizeof(H) == 2 ? (float)sv[k] : HistCodec<H>::decode(sv[k]);  // uint16: the code
      }
    }
    __syncthreads();

    float hh[EPT];
    if constexpr (GATHER == 2) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int32_t le = k * kBlock + tid;
        hh[k] = le < nnz ? s_contrib[jj[k]] : 0.0f;
      }
    } else {
      // every lane's column, delay or slot is a valid one (edge 0 for a lane past the edges)
      H hr[EPT];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        if constexpr (GATHER == 0) {
          int32_t slot = cur - dd[k];
          slot += (slot < 0) ? S : 0;
          hr[k] = ring[ring_index<LAYOUT>(pitch, jj[k], slot)];
        } else {
          // the bucketed array keeps a scatter pass's steps side by side: hstride values per slot
          hr[k] = hbuf[(int64_t)(uint32_t)jj[k] * a.hstride];
        }
      }
#pragma unroll
      for (int k = 0; k < EPT; ++k) hh[k] = sizeof(H) == 2 ? (float)hr[k] : HistCodec<H>::decode(hr[k]);
    }
    if constexpr (GATHER == 2) __syncthreads();  // every staged value is read before it is overwritten
    // the lane's weight stores address the weights by a 32-bit element index
    // from the kernel argument's base (E < 2^31), so no store needs 64-bit
    // address arithmetic of its own
    const uint32_t wlane = (uint32_t)(e0 + tid);
    // The edge arithmetic runs on every lane (a lane past the block's edges
    // works on row 0 and a defined value and stores nothing), so that only the
    // two stores are predicated and the constants stay in registers across the
    // unrolled edges.
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;  // < CAP: the LDS read is in range
      const bool live = le < nnz;
      const int lr = live ? (int)s_erow[le] : 0;
      float c;
      float wn;
      if constexpr (sizeof(H) == 2)
        wn = edge_update_rev(hh[k], s_theta[lr], s_gate[lr], ww[k], dt, a.eta, a.lam, a.w_min, a.w_max, &c);
      else
        wn = edge_update(hh[k], s_theta[lr], s_gate[lr], ww[k], dt, a.eta, a.lam, a.w_min, a.w_max, &c);
      if (live) {
        s_contrib[le] = c;  // coupling uses the pre-update weight (:154)
        a.w[wlane + k * kBlock] = wn;
      }
    }
    __syncthreads();

    // Row sums in CSR order: an FP64 sum by the row's lane, rounded once (two
    // VALU instructions per term where Neumaier's FP32 sum took about eight, in
    // a loop the wave runs for its longest row), or for a long row lane-strided
    // Neumaier partials combined by a fixed FP64 tree.
    float coupling = 0.0f;
    if (tid < nrows && !heavy) {
      double acc = 0.0;
      for (int32_t q = rs; q < re; ++q) acc += (double)s_contrib[q];
      coupling = (float)acc;
    }
    for (uint64_t m = __ballot(heavy); m; m &= m - 1) {
      const int src = __ffsll((unsigned long long)m) - 1;
      const int32_t hs = __builtin_amdgcn_readlane(rs, src), he = __builtin_amdgcn_readlane(re, src);
      CompSum acc;
      for (int32_t q = hs + lane; q < he; q += 64) acc.add(s_contrib[q]);
      const double tot = wave_sum((double)acc.s + (double)acc.c);
      if (lane == src) coupling = (float)tot;
    }

    if (tid < nrows) {
      const int32_t i = r0 + tid;
      const float x = row_integrate(rs_, om, g, coupling, dt, a.alpha);
      a.theta[i] = x;
      reinterpret_cast<H*>(a.ring)[ring_index<LAYOUT>(pitch, i, nxt)] = HistCodec<H>::encode(x);
      a.theta_hat[i] = rs_.thh_new;
      a.omega_hat[i] = rs_.omh_new;
      float sx, cx;
      sincos_fast(x, &sx, &cx);
      my_c = (double)cx;
      my_s = (double)sx;
    }
  } else {
    // ---------------- long-row mode: one row, whole workgroup ----------------
    const int32_t i = r0;
    RowState rs_;
    float om = 0.0f;
    if (tid == 0) {
      om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[0] = rs_.th;
      s_gate[0] = rs_.gate;
    }
    __syncthreads();
    const float th_i = s_theta[0];
    const float th_rev = th_i * kInvTwoPi;
    const float gate = s_gate[0];
    CompSum acc;
    for (int32_t base = 0; base < nnz; base += kBlock * 4) {
      // unconditional loads (edge e0 for a lane past the row's end), as in stream mode
      int32_t jj[4], dd[4];
      float ww[4];
      H hr[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        const int32_t e = e0 + (le < nnz ? le : 0);
        if constexpr (GATHER == 0) {
          jj[k] = a.col[e];
          dd[k] = a.delay[e];
        } else if constexpr (GATHER == 1) {
          jj[k] = (int32_t)a.hpos[e];
        } else {
          jj[k] = e;
        }
        ww[k] = a.w[e];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if constexpr (GATHER == 0) {
          int32_t slot = cur - dd[k];
          slot += (slot < 0) ? S : 0;
          hr[k] = ring[ring_index<LAYOUT>(pitch, jj[k], slot)];
        } else if constexpr (GATHER == 1) {
          hr[k] = hbuf[(int64_t)(uint32_t)jj[k] * a.hstride];
        } else {
          hr[k] = hbuf[(uint32_t)jj[k]];
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        if (le < nnz) {
          float c;
          float wn;
          if constexpr (sizeof(H) == 2)
            wn = edge_update_rev((float)hr[k], th_rev, gate, ww[k], dt, a.eta, a.lam, a.w_min, a.w_max, &c);
          else
            wn = edge_update(HistCodec<H>::decode(hr[k]), th_i, gate, ww[k], dt, a.eta, a.lam, a.w_min, a.w_max, &c);
          acc.add(c);
          a.w[e0 + le] = wn;
        }
      }
    }
    // FP64 tree over the lanes' compensated partials (fixed order).
    double part = wave_sum((double)acc.s + (double)acc.c);
    if ((tid & 63) == 0) s_red[0][tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
#pragma unroll
      for (int q = 0; q < kBlock / 64; ++q) tot += s_red[0][q];
      const float x = row_integrate(rs_, om, g, (float)tot, dt, a.alpha);
      a.theta[i] = x;
      reinterpret_cast<H*>(a.ring)[ring_index<LAYOUT>(pitch, i, nxt)] = HistCodec<H>::encode(x);
      a.theta_hat[i] = rs_.thh_new;
      a.omega_hat[i] = rs_.omh_new;
      float sx, cx;
      sincos_fast(x, &sx, &cx);
      my_c = (double)cx;
      my_s = (double)sx;
    }
    __syncthreads();  // s_red is reused below
  }

  // Block partial of sum cos / sum sin of the new phases, in FP64.
  const double wc = wave_sum(my_c);
  const double ws = wave_sum(my_s);
  if ((tid & 63) == 0) {
    s_red[0][tid >> 6] = wc;
    s_red[1][tid >> 6] = ws;
  }
  __syncthreads();
  if (tid == 0) {
    double c = 0.0, s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
      c += s_red[0][q];
      s += s_red[1][q];
    }
    a.partials[b] = make_double2(c, s);
  }
}

// Sum of kChunkPartials consecutive partials per workgroup, in a fixed order.
__global__ __launch_bounds__(kChunkThreads) void chunk_sum_kernel(const double2* partials, int32_t n,
                                                                  double2* chunks) {
  __shared__ double s_red[2][kChunkThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t base = blockIdx.x * kChunkPartials;
  double2 p[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int32_t q = base + k * kChunkThreads + tid;
    p[k] = q < n ? partials[q] : make_double2(0.0, 0.0);
  }
  double c = 0.0, s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    c += p[k].x;
    s += p[k].y;
  }
  c = wave_sum(c);
  s = wave_sum(s);
  if (lane == 0) {
    s_red[0][wave] = c;
    s_red[1][wave] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double tc = 0.0, ts = 0.0;
    for (int w = 0; w < kChunkThreads / 64; ++w) {
      tc += s_red[0][w];
      ts += s_red[1][w];
    }
    chunks[blockIdx.x] = make_double2(tc, ts);
  }
}

// r and the homeostatic gain controller (leviathan_kernel.cu:237-264 and
// :430-435), in FP64, for every network; then the step counter advances.
__global__ __launch_bounds__(kFinalizeThreads) void finalize_kernel(FinalizeArgs a) {
  __shared__ double s_red[2][kFinalizeThreads / 64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t step = *a.step;
  const int64_t hslot = step % kRHistSteps;
  // Hub transition statistics of an attached PhiHandle, from the phases the
  // row pass just wrote: the last wavefront (of workgroup 0) issues its loads
  // now and records the state after its share of the reduction.
  const bool hub_wave = a.phi.st != nullptr && blockIdx.x == 0 && wave == kFinalizeThreads / 64 - 1;
  HubLoad hub{};
  if (hub_wave) hub = hub_load(a.phi, a.theta, lane);

  auto finish = [&](int net, double c, double s) {
    NetCtrl& nc = a.ctrl[net];
    const double r = sqrt(c * c + s * s) / (double)nc.n_nodes;
    nc.r = r;
    a.r_hist[hslot * a.n_nets + net] = (float)r;
    if (nc.controller) {
      double gn = nc.g + (double)a.dt * (nc.beta * (nc.r0 - r) - nc.mu * (nc.g - nc.g_rest));
      gn = fmin(nc.g_max, gn);
      gn = fmax(nc.g_min, gn);
      nc.g = gn;
    }
  };

  if (a.n_nets == 1) {
    // eight loads in flight per thread, added in index order
    double c = 0.0, s = 0.0;
    int32_t q = tid;
    for (; q + 7 * kFinalizeThreads < a.num_blocks; q += 8 * kFinalizeThreads) {
      double2 p[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) p[k] = a.partials[q + k * kFinalizeThreads];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        c += p[k].x;
        s += p[k].y;
      }
    }
    for (; q < a.num_blocks; q += kFinalizeThreads) {
      const double2 p = a.partials[q];
      c += p.x;
      s += p.y;
    }
    c = wave_sum(c);
    s = wave_sum(s);
    if (lane == 0) {
      s_red[0][wave] = c;
      s_red[1][wave] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double tc = 0.0, ts = 0.0;
      for (int q = 0; q < kFinalizeThreads / 64; ++q) {
        tc += s_red[0][q];
        ts += s_red[1][q];
      }
      finish(0, tc, ts);
    }
  } else {
    // one wavefront per network, spread over a grid of workgroups; each lane
    // adds the partials q = lane, lane + 64, ... of the network's long and
    // stream ranges in index order, eight loads in flight
    const int net = blockIdx.x * (kFinalizeThreads / 64) + wave;
    if (net < a.n_nets) {
      const int4 rg = a.ranges[net];
      double c = 0.0, s = 0.0;
      auto add_range = [&](int32_t lo, int32_t hi) {
        int32_t q = lo + lane;
        for (; q + 7 * 64 < hi; q += 8 * 64) {
          double2 p[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) p[k] = a.partials[q + k * 64];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            c += p[k].x;
            s += p[k].y;
          }
        }
        for (; q < hi; q += 64) {
          const double2 p = a.partials[q];
          c += p.x;
          s += p.y;
        }
      };
      add_range(rg.x, rg.y);
      add_range(rg.z, rg.w);
      c = wave_sum(c);
      s = wave_sum(s);
      if (lane == 0) finish(net, c, s);
    }
  }
  if (hub_wave) hub_record(a.phi, hub, lane);
  __syncthreads();
  if (tid == 0) {
    if (gridDim.x == 1) {
      *a.step = step + 1;
    } else {
      // every workgroup has read the step counter (at its start, before the
      // barrier above); the last one to arrive advances it and re-arms the count
      const unsigned int arrived = atomicAdd(a.fin_done, 1u);
      if (arrived == gridDim.x - 1) {
        *a.fin_done = 0u;
        *a.step = step + 1;
      }
    }
  }
}

template <class H, int LAYOUT>
__global__ void fill_ring_kernel(H* ring, int64_t pitch, int32_t n_slots, const float* theta, int32_t n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n * n_slots;
  if (idx >= total) return;
  int32_t node, slot;
  if constexpr (LAYOUT == 0) {
    node = (int32_t)(idx / n_slots);
    slot = (int32_t)(idx % n_slots);
  } else {
    slot = (int32_t)(idx / n);
    node = (int32_t)(idx % n);
  }
  ring[ring_index<LAYOUT>(pitch, node, slot)] = HistCodec<H>::encode(theta[node]);
}

template <class H, int LAYOUT>
__global__ void write_slot_kernel(H* ring, int64_t pitch, int32_t slot, const float* in, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ring[ring_index<LAYOUT>(pitch, i, slot)] = HistCodec<H>::encode(in[i]);
}

template <class H, int LAYOUT>
__global__ void read_slot_kernel(const H* ring, int64_t pitch, int32_t slot, float* out, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = HistCodec<H>::decode(ring[ring_index<LAYOUT>(pitch, i, slot)]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <class H, int LAYOUT, int GATHER>
static hipError_t launch_step_t(int ept, const StepArgs& a, hipStream_t s) {
  int32_t longest = 0;
  for (int x = 0; x < kNumXcd; ++x) longest = std::max(longest, a.xcd_split[x + 1] - a.xcd_split[x]);
  const dim3 grid((unsigned)(GATHER != 1 ? a.num_items : kNumXcd * longest)), block(kBlock);
  switch (ept) {
    case 4:
      step_kernel<H, LAYOUT, 4, GATHER><<<grid, block, 0, s>>>(a);
      break;
    case 8:
      step_kernel<H, LAYOUT, 8, GATHER><<<grid, block, 0, s>>>(a);
      break;
    case 16:
      step_kernel<H, LAYOUT, 16, GATHER><<<grid, block, 0, s>>>(a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class H, int LAYOUT>
static hipError_t launch_step_g(int ept, int gather, const StepArgs& a, hipStream_t s) {
  switch (gather) {
    case 0:
      return launch_step_t<H, LAYOUT, 0>(ept, a, s);
    case 1:
      return launch_step_t<H, LAYOUT, 1>(ept, a, s);
    case 2:
      return launch_step_t<H, LAYOUT, 2>(ept, a, s);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_step(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s) {
  if (a.num_items <= 0) return hipSuccess;
  if (history == 0) {
    return layout == 0 ? launch_step_g<float, 0>(ept, gather, a, s) : launch_step_g<float, 1>(ept, gather, a, s);
  }
  return layout == 0 ? launch_step_g<uint16_t, 0>(ept, gather, a, s) : launch_step_g<uint16_t, 1>(ept, gather, a, s);
}

hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s) {
  if (a.n_nets == 1 && a.chunks && a.num_blocks >= kChunkMinBlocks) {
    const int32_t n_chunks = (a.num_blocks + kChunkPartials - 1) / kChunkPartials;
    chunk_sum_kernel<<<n_chunks, kChunkThreads, 0, s>>>(a.partials, a.num_blocks, a.chunks);
    FinalizeArgs b = a;
    b.partials = a.chunks;
    b.num_blocks = n_chunks;
    finalize_kernel<<<1, kFinalizeThreads, 0, s>>>(b);
  } else if (a.n_nets > 1) {
    constexpr int kNetsPerWg = kFinalizeThreads / 64;
    finalize_kernel<<<(a.n_nets + kNetsPerWg - 1) / kNetsPerWg, kFinalizeThreads, 0, s>>>(a);
  } else {
    finalize_kernel<<<1, kFinalizeThreads, 0, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_fill_ring(int history, int layout, void* ring, int64_t pitch, int32_t n_slots,
                            const float* theta, int32_t n, hipStream_t s) {
  const int64_t total = (int64_t)n * n_slots;
  if (total == 0) return hipSuccess;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  if (history == 0) {
    if (layout == 0)
      fill_ring_kernel<float, 0><<<grid, block, 0, s>>>((float*)ring, pitch, n_slots, theta, n);
    else
      fill_ring_kernel<float, 1><<<grid, block, 0, s>>>((float*)ring, pitch, n_slots, theta, n);
  } else {
    if (layout == 0)
      fill_ring_kernel<uint16_t, 0><<<grid, block, 0, s>>>((uint16_t*)ring, pitch, n_slots, theta, n);
    else
      fill_ring_kernel<uint16_t, 1><<<grid, block, 0, s>>>((uint16_t*)ring, pitch, n_slots, theta, n);
  }
  return hipGetLastError();
}

hipError_t launch_write_slot(int history, int layout, void* ring, int64_t pitch, int32_t slot, const float* in,
                             int32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (history == 0) {
    if (layout == 0)
      write_slot_kernel<float, 0><<<grid, block, 0, s>>>((float*)ring, pitch, slot, in, n);
    else
      write_slot_kernel<float, 1><<<grid, block, 0, s>>>((float*)ring, pitch, slot, in, n);
  } else {
    if (layout == 0)
      write_slot_kernel<uint16_t, 0><<<grid, block, 0, s>>>((uint16_t*)ring, pitch, slot, in, n);
    else
      write_slot_kernel<uint16_t, 1><<<grid, block, 0, s>>>((uint16_t*)ring, pitch, slot, in, n);
  }
  return hipGetLastError();
}

hipError_t launch_read_slot(int history, int layout, const void* ring, int64_t pitch, int32_t slot,
                            float* out, int32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (history == 0) {
    if (layout == 0)
      read_slot_kernel<float, 0><<<grid, block, 0, s>>>((const float*)ring, pitch, slot, out, n);
    else
      read_slot_kernel<float, 1><<<grid, block, 0, s>>>((const float*)ring, pitch, slot, out, n);
  } else {
    if (layout == 0)
      read_slot_kernel<uint16_t, 0><<<grid, block, 0, s>>>((const uint16_t*)ring, pitch, slot, out, n);
    else
      read_slot_kernel<uint16_t, 1><<<grid, block, 0, s>>>((const uint16_t*)ring, pitch, slot, out, n);
  }
  return hipGetLastError();
}

}  // namespace lev
