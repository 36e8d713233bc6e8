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
// Paired row passes (graphs without delay-0 edges, leviathan_step_n): two
// consecutive steps take pair_kernel (step t in full and step t+1's prep,
// Hebbian update and coupling sum), finalize_kernel (step t), and
// pair_finish_kernel (step t+1's integration with the gain the finalize just
// set, and both steps' history slots), then finalize_kernel (step t+1).
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
// GATHER = 3 (staged schedule, paired row pass, uint16 history): as 2, with the
// pair's two steps' values side by side (one 4-byte pair per staged slot,
// written so by the regroup), so each slot is one load for both steps.
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

// Stream mode: the LDS byte map from each of the block's edges to its row.
// Rows of up to kSerialRow edges are filled by their own lane, longer ones by
// their whole wavefront, one after another (a lane looping over a long row
// would hold its wavefront for the row's whole length).  `mine`: the lane owns
// a row, whose edges are [rs, re) of the block.
__device__ __forceinline__ void fill_edge_rows(uint8_t* s_erow, int tid, bool mine, bool heavy, int32_t rs,
                                               int32_t re) {
  const int lane = tid & 63;
  if (mine && !heavy) {
    int32_t q = rs;
    for (; q + 4 <= re; q += 4) {  // four stores per trip of the loop
      s_erow[q] = (uint8_t)tid;
      s_erow[q + 1] = (uint8_t)tid;
      s_erow[q + 2] = (uint8_t)tid;
      s_erow[q + 3] = (uint8_t)tid;
    }
    for (; q < re; ++q) s_erow[q] = (uint8_t)tid;
  }
  for (uint64_t m = __ballot(heavy); m; m &= m - 1) {
    const int src = __ffsll((unsigned long long)m) - 1;
    const int32_t hs = __builtin_amdgcn_readlane(rs, src), he = __builtin_amdgcn_readlane(re, src);
    for (int32_t q = hs + lane; q < he; q += 64) s_erow[q] = (uint8_t)((tid & ~63) + src);
  }
}

// Stream mode: each row's coupling sum of its edges' contributions in CSR
// order: an FP64 sum by the row's lane, rounded once (two VALU instructions
// per term where Neumaier's FP32 sum took about eight, in a loop the wave runs
// for its longest row), or for a row of more than kSerialRow edges
// lane-strided Neumaier partials combined by a fixed FP64 tree.  The loops
// read eight (a lane's row) or four (a wavefront's row) terms before adding
// them in order: one LDS round trip per group of terms instead of per term,
// and the same sums (BA N = 10^7, uint16 staged/8 with paired passes: 0.907
// -> 0.877 ms per step with four terms and the row-map fill unrolled alike,
// 0.869 with eight).
__device__ __forceinline__ float row_coupling(const float* s_contrib, int tid, bool mine, bool heavy, int32_t rs,
                                              int32_t re) {
  const int lane = tid & 63;
  float coupling = 0.0f;
  if (mine && !heavy) {
    double acc = 0.0;
    int32_t q = rs;
    for (; q + 8 <= re; q += 8) {
      float c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) c[u] = s_contrib[q + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += (double)c[u];
    }
    for (; q < re; ++q) acc += (double)s_contrib[q];
    coupling = (float)acc;
  }
  for (uint64_t m = __ballot(heavy); m; m &= m - 1) {
    const int src = __ffsll((unsigned long long)m) - 1;
    const int32_t hs = __builtin_amdgcn_readlane(rs, src), he = __builtin_amdgcn_readlane(re, src);
    CompSum acc;
    int32_t q = hs + lane;
    for (; q + 192 < he; q += 256) {
      const float c0 = s_contrib[q], c1 = s_contrib[q + 64], c2 = s_contrib[q + 128], c3 = s_contrib[q + 192];
      acc.add(c0);
      acc.add(c1);
      acc.add(c2);
      acc.add(c3);
    }
    for (; q < he; q += 64) acc.add(s_contrib[q]);
    const double tot = wave_sum((double)acc.s + (double)acc.c);
    if (lane == src) coupling = (float)tot;
  }
  return coupling;
}

// FP64 sums over a wavefront of the lanes' cos and sin of their rows' new
// phases (lanes without a row hold +0.0).  A row's cosine is never zero (pi/2
// is not a float), so a wavefront whose cosines are all +0.0 holds no row and
// its sums are exactly +0.0: it skips the reductions.
__device__ __forceinline__ void wave_sums_cs(double my_c, double my_s, double* wc, double* ws) {
  if (__ballot(my_c != 0.0) == 0) {
    *wc = 0.0;
    *ws = 0.0;
  } else {
    *wc = wave_sum(my_c);
    *ws = wave_sum(my_s);
  }
}

// The block's FP64 partial of sum cos / sum sin of its new phases (lanes
// without a row add zeros), summed in a fixed order: DPP wavefront sums, then
// the four wavefronts in turn.
__device__ __forceinline__ void store_block_partial(double my_c, double my_s, double (*s_red)[kBlock / 64],
                                                    double2* out) {
  const int tid = threadIdx.x;
  double wc, ws;
  wave_sums_cs(my_c, my_s, &wc, &ws);
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
    *out = make_double2(c, s);
  }
}

// Long-row mode: the whole workgroup strides over one row's nnz edges from e0,
// updating each weight and returning the lane's compensated partial of the
// coupling.  The delayed values come from `hb`: the ring read at slot
// cur - delay (GATHER 0), the bucketed array (GATHER 1, component already
// added to hb) or the row's range of the staged array in CSR order (GATHER 2).
// A lane visits the same edges whatever the step, so a second sweep over the
// row reads back only the weights the lane itself stored.  (hb is not
// __restrict__: the paired pass writes a long row's early values into the
// staged arrays before it sweeps them.)
template <class H, int LAYOUT, int GATHER>
__device__ __forceinline__ CompSum long_row_sweep(const StepArgs& a, const H* hb, int32_t cur,
                                                  int32_t e0, int32_t nnz, float th_i, float gate, float dt) {
  const int tid = threadIdx.x;
  const int32_t S = a.n_slots;
  const int64_t pitch = a.ring_pitch;
  const float th_rev = th_i * kInvTwoPi;
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
        hr[k] = hb[ring_index<LAYOUT>(pitch, jj[k], slot)];
      } else if constexpr (GATHER == 1) {
        hr[k] = hb[(int64_t)(uint32_t)jj[k] * a.hstride];
      } else {
        hr[k] = hb[(uint32_t)jj[k] * (GATHER == 3 ? 2u : 1u)];  // packed pairs: every other value
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
  return acc;
}

// Long-row mode: the coupling of one row from its lanes' partials, by a fixed
// FP64 tree (wavefront sums, then the four wavefronts in turn), returned to
// lane 0.
__device__ __forceinline__ float long_row_coupling(const CompSum& acc, double (*s_red)[kBlock / 64]) {
  const int tid = threadIdx.x;
  const double part = wave_sum((double)acc.s + (double)acc.c);
  if ((tid & 63) == 0) s_red[0][tid >> 6] = part;
  __syncthreads();
  double tot = 0.0;
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) tot += s_red[0][q];
  }
  __syncthreads();  // s_red is reused
  return (float)tot;
}

// Row block of workgroup blockIdx.x.  The binned row pass reads its rows'
// bucket, so each XCD takes one contiguous run of blocks and the bucket stays
// in its L2.  The other row passes take the blocks in order, dealt round-robin,
// so that the eight XCDs stream through neighbouring addresses: BA N = 10^7,
// uint16, staged/8 0.603 -> 0.565 ms; the FP32 direct/8 drop-in 4.73 -> 4.66
// ms; the 256-network Watts-Strogatz ensemble, direct/16, 1.26 -> 1.24 ms (its
// gathers' locality between neighbouring rows did not need the runs).
// Returns -1 for a workgroup without a block.
template <int GATHER>
__device__ __forceinline__ int32_t row_block_of(const StepArgs& a) {
  if constexpr (GATHER != 1) {
    const int32_t b = (int32_t)blockIdx.x;
    return b < a.num_items ? b : -1;
  } else {
    const int32_t x = blockIdx.x % kNumXcd;
    const int32_t b = a.xcd_split[x] + (int32_t)(blockIdx.x / kNumXcd);
    return b < a.xcd_split[x + 1] ? b : -1;  // this XCD's run is shorter than the longest
  }
}

// Stream mode: each lane's EPT edges of the block (edge le = k * kBlock + tid)
// and the value each edge reads.  The loads are unconditional: a lane past the
// block's edges reads edge 0 (every edge array holds at least one element) and
// discards the value.  Straight-line code keeps all of a lane's loads in
// flight at once; with a branch around each load the compiler waited for every
// load before issuing the next.
//
// NV values per edge are read: those of the step (cur: the step's slot) and,
// with NV = 2, those of the next step (the ring one slot later, the bucketed
// slot's next component, or the next staged array a.hbuf_next).  The staged
// schedule stages the block's values in LDS (s_v[v], as codes for uint16) and
// reads edge e's at lpos[e]; the caller's barrier then separates the staging
// from the reads, so this is done in two calls (STAGE then PICK).
template <class H, int LAYOUT, int EPT, int GATHER, int NV>
struct StreamEdges {
  int32_t jj[EPT];  // column (GATHER 0), value slot (GATHER 1) or staged position (GATHER 2)
  int32_t dd[EPT];
  float ww[EPT];
  float hh[NV][EPT];  // the delayed values, decoded (uint16: the code as a float)

  __device__ __forceinline__ void load(const StepArgs& a, int32_t e0, int32_t nnz, int32_t cur, float* const* s_v) {
    const int tid = threadIdx.x;
    const int32_t S = a.n_slots;
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
    if constexpr (GATHER >= 2) {
      // the block's staged values go to LDS (s_v[v]) until the edge pass
      H sv[NV][EPT];
      if constexpr (GATHER == 3) {
        static_assert(NV == 2 && sizeof(H) == 2, "packed pairs hold two uint16 values per slot");
        const uint32_t* __restrict__ src = reinterpret_cast<const uint32_t*>(a.hbuf);
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * kBlock + tid;
          const uint32_t p = src[le < nnz ? e0 + le : 0];
          sv[0][k] = (H)(p & 0xFFFFu);
          sv[1][k] = (H)(p >> 16);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const H* __restrict__ src = reinterpret_cast<const H*>(v == 0 ? a.hbuf : a.hbuf_next);
#pragma unroll
          for (int k = 0; k < EPT; ++k) {
            const int32_t le = k * kBlock + tid;
            sv[v][k] = src[le < nnz ? e0 + le : 0];
          }
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          const int32_t le = k * kBlock + tid;
          if (le < nnz) s_v[v][le] = sizeof(H) == 2 ? (float)sv[v][k] : HistCodec<H>::decode(sv[v][k]);
        }
    } else {
      // every lane's column, delay or slot is a valid one (edge 0 for a lane past the edges)
      const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
      const H* __restrict__ hbuf = reinterpret_cast<const H*>(a.hbuf);
      const int64_t pitch = a.ring_pitch;
      H hr[NV][EPT];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        if constexpr (GATHER == 0) {
          int32_t slot = cur - dd[k];
          slot += (slot < 0) ? S : 0;
          hr[0][k] = ring[ring_index<LAYOUT>(pitch, jj[k], slot)];
          if constexpr (NV == 2) {
            const int32_t slot2 = slot + 1 == S ? 0 : slot + 1;
            hr[1][k] = ring[ring_index<LAYOUT>(pitch, jj[k], slot2)];
          }
        } else {
          // the bucketed array keeps a scatter pass's steps side by side: hstride values per slot
          const int64_t at = (int64_t)(uint32_t)jj[k] * a.hstride;
          hr[0][k] = hbuf[at];
          if constexpr (NV == 2) hr[1][k] = hbuf[at + 1];
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int k = 0; k < EPT; ++k) hh[v][k] = sizeof(H) == 2 ? (float)hr[v][k] : HistCodec<H>::decode(hr[v][k]);
    }
  }

  // staged schedule, after the barrier that follows load(): each edge's values
  // from LDS.  A lane past the block's edges holds edge 0's slot, which lies in
  // the block capacity, so every read is in range and they issue together.
  __device__ __forceinline__ void pick(float* const* s_v) {
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int k = 0; k < EPT; ++k) hh[v][k] = s_v[v][jj[k]];
  }

  // The Hebbian update of the lane's edges with values hh[v]: each edge's
  // contribution w_old * sin(delta) goes to s_contrib, and its new weight to
  // ww (and, with store, to memory).  The edge arithmetic runs on every lane
  // (a lane past the block's edges works on whatever row its byte-map entry
  // names and stores nothing), so that only the stores are predicated and the
  // constants stay in registers across the unrolled edges.  s_theta holds each
  // row's phase (revolutions for uint16 history) and s_gate its gate.  The
  // byte map is read for every lane (for a lane past the block's edges a stale
  // entry, which still addresses s_theta in range), so the edges' row reads,
  // then their phase and gate reads, issue together: one LDS round trip each
  // instead of two per edge (BA N = 10^7, uint16 staged/8, paired passes:
  // 0.870 -> 0.865 ms per step).
  template <int V>
  __device__ __forceinline__ void update(const StepArgs& a, int32_t e0, int32_t nnz, const uint8_t* s_erow,
                                         const float* s_theta, const float* s_gate, float* s_contrib, bool store) {
    const int tid = threadIdx.x;
    // the lane's weight stores address the weights by a 32-bit element index
    // from the kernel argument's base (E < 2^31), so no store needs 64-bit
    // address arithmetic of its own
    const uint32_t wlane = (uint32_t)(e0 + tid);
    int lr[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) lr[k] = s_erow[k * kBlock + tid];  // < CAP: in range
    float th[EPT], gt[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      th[k] = s_theta[lr[k]];
      gt[k] = s_gate[lr[k]];
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      const bool live = le < nnz;
      float c;
      float wn;
      if constexpr (sizeof(H) == 2)
        wn = edge_update_rev(hh[V][k], th[k], gt[k], ww[k], a.dt, a.eta, a.lam, a.w_min, a.w_max, &c);
      else
        wn = edge_update(hh[V][k], th[k], gt[k], ww[k], a.dt, a.eta, a.lam, a.w_min, a.w_max, &c);
      ww[k] = wn;
      if (live) {
        s_contrib[le] = c;  // coupling uses the pre-update weight (:154)
        if (store) a.w[wlane + k * kBlock] = wn;
      }
    }
  }
};

// One step of the row blocks (prep, gather, coupling, Hebbian update,
// integration, history write, block partial of r).
template <class H, int LAYOUT, int EPT, int GATHER>
__global__ __launch_bounds__(kBlock) void step_kernel(StepArgs a) {
  constexpr int CAP = kBlock * EPT;
  __shared__ float s_theta[kBlock];
  __shared__ float s_gate[kBlock];
  __shared__ uint8_t s_erow[CAP];
  __shared__ float s_contrib[CAP];
  __shared__ double s_red[2][kBlock / 64];

  const int tid = threadIdx.x;
  const int32_t b = row_block_of<GATHER>(a);
  if (b < 0) return;
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
    const bool mine = tid < nrows;
    if (mine) {
      const int32_t i = r0 + tid;
      om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[tid] = sizeof(H) == 2 ? rs_.th * kInvTwoPi : rs_.th;  // uint16: revolutions (edge_update_rev)
      s_gate[tid] = rs_.gate;
      rs = a.row_ptr[i] - e0;
      re = a.row_ptr[i + 1] - e0;
    }
    const bool heavy = mine && re - rs > kSerialRow;
    fill_edge_rows(s_erow, tid, mine, heavy, rs, re);

    StreamEdges<H, LAYOUT, EPT, GATHER, 1> ed;
    float* const s_v[1] = {s_contrib};  // staged schedule: s_contrib holds the values until the edge pass
    ed.load(a, e0, nnz, cur, s_v);
    __syncthreads();
    if constexpr (GATHER == 2) {
      ed.pick(s_v);
      __syncthreads();  // every staged value is read before it is overwritten
    }
    ed.template update<0>(a, e0, nnz, s_erow, s_theta, s_gate, s_contrib, true);
    __syncthreads();

    const float coupling = row_coupling(s_contrib, tid, mine, heavy, rs, re);
    if (mine) {
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
    const H* hb = reinterpret_cast<const H*>(GATHER == 0 ? a.ring : a.hbuf);
    const CompSum acc = long_row_sweep<H, LAYOUT, GATHER>(a, hb, cur, e0, nnz, s_theta[0], s_gate[0], dt);
    const float coupling = long_row_coupling(acc, s_red);
    if (tid == 0) {
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
  }
  store_block_partial(my_c, my_s, s_red, &a.partials[b]);
}

// Staged schedule, second pair (steps k, k + 1, k > 0) of a multi-step block:
// an early edge's values of both steps straight from the ring, at slots
// cur - delay and cur + 1 - delay (the step counter is at step k).  Step k's
// value replaces the scatter's only for delay < k; for the others the first
// load reads step k + 1's slot too (same line, no memory traffic of its own,
// and no branch around a load) and its value is not used.
template <class H, int LAYOUT>
__device__ __forceinline__ void early_pair_values(const StepArgs& a, int2 rec, int32_t cur, H* v0, H* v1) {
  const H* __restrict__ ring = reinterpret_cast<const H*>(a.ring);
  const int32_t S = a.n_slots;
  const int32_t d = rec.y & 3;
  int32_t slot = cur - d;
  slot += (slot < 0) ? S : 0;
  const int32_t slot2 = slot + 1 == S ? 0 : slot + 1;
  *v0 = ring[ring_index<LAYOUT>(a.ring_pitch, rec.x, d < a.early_k ? slot : slot2)];
  *v1 = ring[ring_index<LAYOUT>(a.ring_pitch, rec.x, slot2)];
}

// The paired row pass: steps t and t + 1 of the row blocks in one pass, for
// graphs without a delay-0 edge.  Step t + 1 reads theta_j(t + 1 - d) with
// d >= 1, which exists before step t's row pass, and each row's own state
// after step t is the workgroup's, so the workgroup runs step t in full and
// then step t + 1's prep, Hebbian update and coupling sum, reading and writing
// each weight once for the two steps.  Only step t + 1's integration waits:
// it needs the gain g(t + 1), which the finalize of step t computes from the
// order parameter of all of step t's new phases.  The pass leaves each row's
// step t + 1 coupling sum and prediction error in a.pair, and
// pair_finish_kernel integrates after that finalize.  The history slot of
// step t's new phase is written by pair_finish_kernel too, with step t + 1's:
// nothing reads it before the next pair (step t + 1 reads delays >= 1), and on
// a node-major ring the two slots share a line, which then leaves the L2 once.
//
// In the second pair of a multi-step block of the staged schedule, the pass
// also gathers its block's early edges (a.early_k > 0): the block patches
// their values in its LDS copy of the staged values (stream mode) or in its
// range of the staged arrays before sweeping it (long-row mode), where
// early_kernel would have written them to the staged arrays beforehand.
//
// Every value is computed as the single-step kernel computes it, from the
// same operands in the same order (the weights after step t stay in
// registers; the coupling sum and the error are stored as the FP32 values the
// integration reads), so phases, weights, history, r and gain are
// bit-identical to two single steps.
template <class H, int LAYOUT, int EPT, int GATHER>
__global__ __launch_bounds__(kBlock) void pair_kernel(StepArgs a) {
  constexpr int CAP = kBlock * EPT;
  __shared__ float s_theta[kBlock];
  __shared__ float s_gate[kBlock];
  __shared__ uint8_t s_erow[CAP];
  __shared__ float s_contrib[CAP];
  __shared__ float s_next[GATHER >= 2 ? CAP : 1];  // staged schedule: step t + 1's values
  __shared__ double s_red[2][kBlock / 64];

  const int tid = threadIdx.x;
  const int32_t b = row_block_of<GATHER>(a);
  if (b < 0) return;
  const Item it = unpack_item(a.items[b]);
  if (it.nrows == 0) {  // padding item
    if (threadIdx.x == 0) a.partials[b] = make_double2(0.0, 0.0);
    return;
  }
  const int32_t r0 = it.r0, net = it.net;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const int32_t nxt = (cur + 1 == S) ? 0 : cur + 1;
  const float g = (float)a.ctrl[net].g;
  const float dt = a.dt;
  const int32_t e0 = it.e0;
  const int32_t nnz = it.nnz;
  const int32_t nrows = it.nrows;
  float u_t = 0.0f;
  if (a.drive_b) u_t = a.drive_u[step % kDriveSteps];

  double my_c = 0.0, my_s = 0.0;

  if (!it.is_long) {
    // ---------------- stream mode ----------------
    RowState rs_;
    int32_t rs = 0, re = 0;
    float om = 0.0f;
    const bool mine = tid < nrows;
    if (mine) {
      const int32_t i = r0 + tid;
      om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[tid] = sizeof(H) == 2 ? rs_.th * kInvTwoPi : rs_.th;
      s_gate[tid] = rs_.gate;
      rs = a.row_ptr[i] - e0;
      re = a.row_ptr[i + 1] - e0;
    }
    const bool heavy = mine && re - rs > kSerialRow;
    fill_edge_rows(s_erow, tid, mine, heavy, rs, re);

    // staged schedule, second pair of a block: this lane's first early edge, fetched with the other loads
    int32_t eq = 0, eq1 = 0;
    int2 erec = make_int2(0, -1);
    H ev0{}, ev1{};
    if constexpr (GATHER >= 2) {
      if (a.early_k > 0) {
        eq = a.early_ptr[b] + tid;
        eq1 = a.early_ptr[b + 1];
        if (eq < eq1) {
          erec = a.early[eq];
          early_pair_values<H, LAYOUT>(a, erec, cur, &ev0, &ev1);
        }
      }
    }
    StreamEdges<H, LAYOUT, EPT, GATHER, 2> ed;
    float* const s_v[2] = {s_contrib, s_next};
    ed.load(a, e0, nnz, cur, s_v);
    __syncthreads();
    if constexpr (GATHER >= 2) {
      if (a.early_k > 0) {  // uniform: patch the staged copies once every staging store is done
        auto as_lds = [](H v) { return sizeof(H) == 2 ? (float)v : HistCodec<H>::decode(v); };
        while (eq < eq1) {  // the fetched record, then the lane's next one (a block with > kBlock)
          const int32_t lp = erec.y >> 2;
          if ((erec.y & 3) < a.early_k) s_contrib[lp] = as_lds(ev0);
          s_next[lp] = as_lds(ev1);
          eq += kBlock;
          if (eq < eq1) {
            erec = a.early[eq];
            early_pair_values<H, LAYOUT>(a, erec, cur, &ev0, &ev1);
          }
        }
        __syncthreads();
      }
      ed.pick(s_v);
      __syncthreads();  // every staged value is read before s_contrib is overwritten
    }

    // step t: the new weights stay in the lanes' registers
    ed.template update<0>(a, e0, nnz, s_erow, s_theta, s_gate, s_contrib, false);
    __syncthreads();
    const float coupling = row_coupling(s_contrib, tid, mine, heavy, rs, re);
    RowState rs2;
    if (mine) {
      const int32_t i = r0 + tid;
      const float x = row_integrate(rs_, om, g, coupling, dt, a.alpha);
      a.theta[i] = x;  // its history slot is written by the node pass (below)
      float sx, cx;
      sincos_fast(x, &sx, &cx);
      my_c = (double)cx;
      my_s = (double)sx;
      // step t + 1's prep from the state step t leaves (every edge of step t
      // read s_theta and s_gate before the barrier ahead of the row sums)
      rs2 = row_prep(x, rs_.thh_new, rs_.omh_new, dt, a.kappa_w, a.gamma);
      a.theta_hat[i] = rs2.thh_new;
      a.omega_hat[i] = rs2.omh_new;
      s_theta[tid] = sizeof(H) == 2 ? rs2.th * kInvTwoPi : rs2.th;
      s_gate[tid] = rs2.gate;
    }
    __syncthreads();  // every row sum has read s_contrib

    // step t + 1 from the weights step t left
    ed.template update<1>(a, e0, nnz, s_erow, s_theta, s_gate, s_contrib, true);
    __syncthreads();
    const float coupling2 = row_coupling(s_contrib, tid, mine, heavy, rs, re);
    if (mine) a.pair[r0 + tid] = make_float2(coupling2, rs2.eps);
  } else {
    // ---------------- long-row mode: one row, whole workgroup, two sweeps ----------------
    const int32_t i = r0;
    RowState rs_, rs2;
    float om = 0.0f;
    if constexpr (GATHER >= 2) {
      // second pair of a block: the row's early values go to its range of the staged arrays (CSR
      // order; packed pairs: every other value from each step's start), which the sweeps read
      // after the barrier below
      if (a.early_k > 0) {
        H* const h0 = reinterpret_cast<H*>(a.hbuf);
        H* const h1 = reinterpret_cast<H*>(a.hbuf_next);
        for (int32_t q = a.early_ptr[b] + tid; q < a.early_ptr[b + 1]; q += kBlock) {
          const int2 rec = a.early[q];
          H v0, v1;
          early_pair_values<H, LAYOUT>(a, rec, cur, &v0, &v1);
          const int64_t slot = (int64_t)(e0 + (rec.y >> 2)) * (GATHER == 3 ? 2 : 1);
          if ((rec.y & 3) < a.early_k) h0[slot] = v0;
          h1[slot] = v1;
        }
      }
    }
    if (tid == 0) {
      om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[0] = rs_.th;
      s_gate[0] = rs_.gate;
    }
    __syncthreads();
    const H* hb = reinterpret_cast<const H*>(GATHER == 0 ? a.ring : a.hbuf);
    const CompSum acc = long_row_sweep<H, LAYOUT, GATHER>(a, hb, cur, e0, nnz, s_theta[0], s_gate[0], dt);
    const float coupling = long_row_coupling(acc, s_red);
    if (tid == 0) {
      const float x = row_integrate(rs_, om, g, coupling, dt, a.alpha);
      a.theta[i] = x;  // its history slot is written by the node pass (below)
      float sx, cx;
      sincos_fast(x, &sx, &cx);
      my_c = (double)cx;
      my_s = (double)sx;
      rs2 = row_prep(x, rs_.thh_new, rs_.omh_new, dt, a.kappa_w, a.gamma);
      a.theta_hat[i] = rs2.thh_new;
      a.omega_hat[i] = rs2.omh_new;
      s_theta[0] = rs2.th;
      s_gate[0] = rs2.gate;
    }
    __syncthreads();
    // step t + 1: the ring one slot later, the bucketed slot's next component, or the next staged
    // array (packed pairs: the value after step t's)
    const H* hb2 = GATHER == 0   ? reinterpret_cast<const H*>(a.ring)
                   : GATHER == 1 ? reinterpret_cast<const H*>(a.hbuf) + 1
                                 : reinterpret_cast<const H*>(a.hbuf_next);
    const CompSum acc2 = long_row_sweep<H, LAYOUT, GATHER>(a, hb2, nxt, e0, nnz, s_theta[0], s_gate[0], dt);
    const float coupling2 = long_row_coupling(acc2, s_red);
    if (tid == 0) a.pair[i] = make_float2(coupling2, rs2.eps);
  }
  store_block_partial(my_c, my_s, s_red, &a.partials[b]);
}

// Second half of a paired row pass: step t + 1's integration of every row,
// after the finalize of step t has set g(t + 1) and advanced the step counter,
// from the coupling sum and prediction error pair_kernel left.  It writes the
// history slots of both steps' new phases (theta(t + 1), which pair_kernel
// left in theta, and theta(t + 2)).  A workgroup
// takes kFinishBlocks row blocks, with each block's rows on the lanes the
// single-step kernel puts them on (a long row on lane 0), so each block
// partial of r is the single-step kernel's to the bit.  All its loads are
// issued first.  One block per workgroup (10^5 workgroups of ~100 rows at BA
// N = 10^7) was bound by workgroup dispatch: 0.104 ms for 0.2 GB.
template <class H, int LAYOUT>
__global__ __launch_bounds__(kBlock) void pair_finish_kernel(StepArgs a) {
  constexpr int NB = kFinishBlocks;
  __shared__ double s_red[NB][2][kBlock / 64];
  const int tid = threadIdx.x;
  const int32_t b0 = (int32_t)blockIdx.x * NB;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);  // the slot of theta(t + 1), the pass's starting phases
  const int32_t nxt = (cur + 1 == S) ? 0 : cur + 1;
  const float u_t = a.drive_b ? a.drive_u[step % kDriveSteps] : 0.0f;
  int32_t row[NB];
  float th[NB], om[NB], g[NB];
  float2 ce[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    row[j] = -1;
    if (b0 + j < a.num_items) {
      const Item it = unpack_item(a.items[b0 + j]);
      const int32_t rows = it.nrows == 0 ? 0 : it.is_long ? 1 : it.nrows;
      if (tid < rows) {
        const int32_t i = it.r0 + tid;
        row[j] = i;
        th[j] = a.theta[i];
        om[j] = a.omega[i];
        if (a.drive_b) om[j] = fmaf(a.drive_b[i], u_t, om[j]);
        ce[j] = a.pair[i];
        g[j] = (float)a.ctrl[it.net].g;
      }
    }
  }
  double my_c[NB], my_s[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    my_c[j] = 0.0;
    my_s[j] = 0.0;
    if (row[j] >= 0) {
      RowState rs;
      rs.th = th[j];
      rs.eps = ce[j].y;
      const float x = row_integrate(rs, om[j], g[j], ce[j].x, a.dt, a.alpha);
      a.theta[row[j]] = x;
      H* const ring = reinterpret_cast<H*>(a.ring);
      ring[ring_index<LAYOUT>(a.ring_pitch, row[j], cur)] = HistCodec<H>::encode(th[j]);
      ring[ring_index<LAYOUT>(a.ring_pitch, row[j], nxt)] = HistCodec<H>::encode(x);
      float sx, cx;
      sincos_fast(x, &sx, &cx);
      my_c[j] = (double)cx;
      my_s[j] = (double)sx;
    }
  }
  // each block's partial as store_block_partial forms it: wavefront sums, then the four in turn
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double wc, ws;
    wave_sums_cs(my_c[j], my_s[j], &wc, &ws);
    if ((tid & 63) == 0) {
      s_red[j][0][tid >> 6] = wc;
      s_red[j][1][tid >> 6] = ws;
    }
  }
  __syncthreads();
  if (tid < NB && b0 + tid < a.num_items) {
    double c = 0.0, s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
      c += s_red[tid][0][q];
      s += s_red[tid][1][q];
    }
    a.partials[b0 + tid] = make_double2(c, s);
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

template <class H, int LAYOUT, int GATHER, bool PAIR>
static hipError_t launch_step_t(int ept, const StepArgs& a, hipStream_t s) {
  int32_t longest = 0;
  for (int x = 0; x < kNumXcd; ++x) longest = std::max(longest, a.xcd_split[x + 1] - a.xcd_split[x]);
  const dim3 grid((unsigned)(GATHER != 1 ? a.num_items : kNumXcd * longest)), block(kBlock);
  switch (ept) {
    case 4:
      if constexpr (PAIR) pair_kernel<H, LAYOUT, 4, GATHER><<<grid, block, 0, s>>>(a);
      else step_kernel<H, LAYOUT, 4, GATHER><<<grid, block, 0, s>>>(a);
      break;
    case 8:
      if constexpr (PAIR) pair_kernel<H, LAYOUT, 8, GATHER><<<grid, block, 0, s>>>(a);
      else step_kernel<H, LAYOUT, 8, GATHER><<<grid, block, 0, s>>>(a);
      break;
    case 16:
      if constexpr (PAIR) pair_kernel<H, LAYOUT, 16, GATHER><<<grid, block, 0, s>>>(a);
      else step_kernel<H, LAYOUT, 16, GATHER><<<grid, block, 0, s>>>(a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class H, int LAYOUT, bool PAIR>
static hipError_t launch_step_g(int ept, int gather, const StepArgs& a, hipStream_t s) {
  switch (gather) {
    case 0:
      return launch_step_t<H, LAYOUT, 0, PAIR>(ept, a, s);
    case 1:
      return launch_step_t<H, LAYOUT, 1, PAIR>(ept, a, s);
    case 2:
      return launch_step_t<H, LAYOUT, 2, PAIR>(ept, a, s);
    case 3:  // packed pairs: the paired pass over uint16 history only
      if constexpr (PAIR && sizeof(H) == 2) return launch_step_t<H, LAYOUT, 3, PAIR>(ept, a, s);
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
}

template <bool PAIR>
static hipError_t launch_step_p(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s) {
  if (history == 0)
    return layout == 0 ? launch_step_g<float, 0, PAIR>(ept, gather, a, s)
                       : launch_step_g<float, 1, PAIR>(ept, gather, a, s);
  return layout == 0 ? launch_step_g<uint16_t, 0, PAIR>(ept, gather, a, s)
                     : launch_step_g<uint16_t, 1, PAIR>(ept, gather, a, s);
}

hipError_t launch_step(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s) {
  if (a.num_items <= 0) return hipSuccess;
  return launch_step_p<false>(history, layout, ept, gather, a, s);
}

hipError_t launch_pair(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s) {
  if (a.num_items <= 0) return hipSuccess;
  if (!a.pair || (gather >= 2 && !a.hbuf_next)) return hipErrorInvalidValue;
  return launch_step_p<true>(history, layout, ept, gather, a, s);
}

hipError_t launch_pair_finish(int history, int layout, const StepArgs& a, hipStream_t s) {
  if (a.num_items <= 0) return hipSuccess;
  if (!a.pair) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((a.num_items + kFinishBlocks - 1) / kFinishBlocks)), block(kBlock);
  if (history == 0) {
    if (layout == 0)
      pair_finish_kernel<float, 0><<<grid, block, 0, s>>>(a);
    else
      pair_finish_kernel<float, 1><<<grid, block, 0, s>>>(a);
  } else {
    if (layout == 0)
      pair_finish_kernel<uint16_t, 0><<<grid, block, 0, s>>>(a);
    else
      pair_finish_kernel<uint16_t, 1><<<grid, block, 0, s>>>(a);
  }
  return hipGetLastError();
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
