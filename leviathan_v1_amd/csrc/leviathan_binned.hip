// leviathan_binned.hip -- the two-phase ("binned") step for graphs without
// index locality (Barabasi-Albert and other expander-like topologies).
//
// The direct step (leviathan_kernels.hip) gathers theta_j(t - d) with one
// random 2- or 4-byte load per edge; on a graph whose edges join random node
// pairs every such load misses the caches and moves a whole cache line.  The
// binned step moves each delayed phase through HBM as part of a stream
// instead (propagation blocking):
//
//   phase 1, scatter_kernel: walk the graph by SOURCE (CSC order).  Lanes on
//     consecutive out-edges of a run of consecutive sources read the sources'
//     history lines once and write h_e = history[j][t - d_e] into a
//     destination-bucketed array: bucket b holds the incoming edges of a run
//     of consecutive rows, ordered by source, so the writes of one source run
//     land in a narrow, advancing window of every bucket and leave L2 as
//     whole lines.
//   phase 2, bucket_kernel: one workgroup per bucket stages the bucket's
//     h values in LDS (one coalesced read), then walks its rows in CSR order
//     (sub-blocks of <= 256 rows and <= 2048 edges, consecutive lanes on
//     consecutive edges) reading h through a per-edge 16-bit permutation, and
//     does the coupling, Hebbian update, integration and history write with
//     exactly the arithmetic of the direct step.  Rows too heavy for a
//     sub-block form single-row buckets whose h values are already in CSR
//     order (a row's columns are sorted, so source order is CSR order).
//
// Per edge the HBM traffic is delay 1 + position 4 + h write 2|4 (phase 1) and
// h read 2|4 + permutation 2 + weight read/write 8 (phase 2), plus one pass
// over the history ring, instead of a cache line per edge.
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
    uint32_t pp[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) {
        dd[k] = __builtin_nontemporal_load(a.delay + c0 + le);
        pp[k] = __builtin_nontemporal_load(a.pos + c0 + le);
      }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int32_t le = k * kBlock + tid;
      if (le < nnz) {
        int32_t slot = cur - dd[k];
        slot += (slot < 0) ? S : 0;
        hbuf[pp[k]] = ring[ring_index<LAYOUT>(pitch, s0 + s_esrc[le], slot)];
      }
    }
  } else {
    // one source with more than CAP out-edges: the whole workgroup strides over them
    for (int32_t le = tid; le < nnz; le += kBlock) {
      const int32_t d = __builtin_nontemporal_load(a.delay + c0 + le);
      const uint32_t p = __builtin_nontemporal_load(a.pos + c0 + le);
      int32_t slot = cur - d;
      slot += (slot < 0) ? S : 0;
      hbuf[p] = ring[ring_index<LAYOUT>(pitch, s0, slot)];
    }
  }
}

template <class H, int LAYOUT>
__global__ __launch_bounds__(kBlock) void bucket_kernel(BucketArgs ba) {
  constexpr int CAP = kBlock * kBucketSubEpt;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  H* s_h = reinterpret_cast<H*>(s_dyn);
  __shared__ float s_theta[kBlock];
  __shared__ float s_gate[kBlock];
  __shared__ uint8_t s_erow[CAP];
  __shared__ float s_contrib[CAP];
  __shared__ double s_red[2][kBlock / 64];

  const StepArgs& a = ba.s;
  const int tid = threadIdx.x;
  const int32_t b = xcd_block(blockIdx.x, a.num_blocks, a.xcd_remap);
  const int4 blk = a.blocks[b];
  const int32_t r0 = blk.x, r1 = blk.y, net = blk.z;
  const int64_t step = *a.step;
  const int32_t S = a.n_slots;
  const int32_t cur = (int32_t)(step % S);
  const int32_t nxt = (cur + 1 == S) ? 0 : cur + 1;
  const float g = (float)a.ctrl[net].g;
  const float dt = a.dt;
  const int64_t pitch = a.ring_pitch;
  const H* __restrict__ hbuf = reinterpret_cast<const H*>(ba.hbuf);
  H* __restrict__ ring_w = reinterpret_cast<H*>(a.ring);
  float u_t = 0.0f;
  if (a.drive_b) u_t = a.drive_u[step % kDriveSteps];
  const int32_t e0 = a.row_ptr[r0];
  const int32_t e1 = a.row_ptr[r1];
  const int32_t nnz = e1 - e0;

  double my_c = 0.0, my_s = 0.0;

  if (!(blk.w & kBlockLongRow)) {
    // stage the bucket's delayed phases (bucket order = source order)
    for (int32_t q = tid; q < nnz; q += kBlock) s_h[q] = __builtin_nontemporal_load(hbuf + e0 + q);
    const int2 sbi = ba.sb_index[b];
    for (int32_t sb = 0; sb < sbi.y; ++sb) {
      const int32_t ra = ba.sb_bounds[sbi.x + sb];
      const int32_t rb = ba.sb_bounds[sbi.x + sb + 1];
      const int32_t ea = a.row_ptr[ra];
      const int32_t sn = a.row_ptr[rb] - ea;
      const int32_t nrows = rb - ra;
      RowState rs_;
      int32_t rs = 0, re = 0;
      if (tid < nrows) {
        const int32_t i = ra + tid;
        rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
        s_theta[tid] = rs_.th;
        s_gate[tid] = rs_.gate;
        rs = a.row_ptr[i] - ea;
        re = a.row_ptr[i + 1] - ea;
        for (int32_t q = rs; q < re; ++q) s_erow[q] = (uint8_t)tid;
      }
      __syncthreads();  // also orders the staging loads above before the LDS reads below
      uint16_t pq[kBucketSubEpt];
      float ww[kBucketSubEpt];
#pragma unroll
      for (int k = 0; k < kBucketSubEpt; ++k) {
        const int32_t le = k * kBlock + tid;
        if (le < sn) {
          pq[k] = __builtin_nontemporal_load(ba.perm + ea + le);
          ww[k] = __builtin_nontemporal_load(a.w + ea + le);
        }
      }
#pragma unroll
      for (int k = 0; k < kBucketSubEpt; ++k) {
        const int32_t le = k * kBlock + tid;
        if (le < sn) {
          const int lr = s_erow[le];
          float c;
          const float wn = edge_update(HistCodec<H>::decode(s_h[pq[k]]), s_theta[lr], s_gate[lr], ww[k], dt, a.eta,
                                       a.lam, a.w_min, a.w_max, &c);
          s_contrib[le] = c;
          __builtin_nontemporal_store(wn, a.w + ea + le);
        }
      }
      __syncthreads();
      if (tid < nrows) {
        const int32_t i = ra + tid;
        CompSum acc;
        for (int32_t q = rs; q < re; ++q) acc.add(s_contrib[q]);
        float om = a.omega[i];
        if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
        const float x = row_integrate(rs_, om, g, acc.value(), dt, a.alpha);
        a.theta[i] = x;
        ring_w[ring_index<LAYOUT>(pitch, i, nxt)] = HistCodec<H>::encode(x);
        a.theta_hat[i] = rs_.thh_new;
        a.omega_hat[i] = rs_.omh_new;
        float sx, cx;
        sincosf(x, &sx, &cx);
        my_c += (double)cx;
        my_s += (double)sx;
      }
      __syncthreads();  // s_theta / s_erow / s_contrib are reused by the next sub-block
    }
  } else {
    // single heavy row: its h values are in CSR order already
    const int32_t i = r0;
    RowState rs_;
    if (tid == 0) {
      rs_ = row_prep(a.theta[i], a.theta_hat[i], a.omega_hat[i], dt, a.kappa_w, a.gamma);
      s_theta[0] = rs_.th;
      s_gate[0] = rs_.gate;
    }
    __syncthreads();
    const float th_i = s_theta[0];
    const float gate = s_gate[0];
    CompSum acc;
    for (int32_t base = 0; base < nnz; base += kBlock * 4) {
      float ww[4], hh[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        if (le < nnz) {
          ww[k] = __builtin_nontemporal_load(a.w + e0 + le);
          hh[k] = HistCodec<H>::decode(__builtin_nontemporal_load(hbuf + e0 + le));
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int32_t le = base + k * kBlock + tid;
        if (le < nnz) {
          float c;
          const float wn = edge_update(hh[k], th_i, gate, ww[k], dt, a.eta, a.lam, a.w_min, a.w_max, &c);
          acc.add(c);
          __builtin_nontemporal_store(wn, a.w + e0 + le);
        }
      }
    }
    double part = wave_sum((double)acc.s + (double)acc.c);
    if ((tid & 63) == 0) s_red[0][tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
#pragma unroll
      for (int q = 0; q < kBlock / 64; ++q) tot += s_red[0][q];
      float om = a.omega[i];
      if (a.drive_b) om = fmaf(a.drive_b[i], u_t, om);
      const float x = row_integrate(rs_, om, g, (float)tot, dt, a.alpha);
      a.theta[i] = x;
      ring_w[ring_index<LAYOUT>(pitch, i, nxt)] = HistCodec<H>::encode(x);
      a.theta_hat[i] = rs_.thh_new;
      a.omega_hat[i] = rs_.omh_new;
      float sx, cx;
      sincosf(x, &sx, &cx);
      my_c = (double)cx;
      my_s = (double)sx;
    }
    __syncthreads();
  }

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

hipError_t launch_bucket(int history, int layout, const BucketArgs& a, size_t lds_bytes, hipStream_t s) {
  if (a.s.num_blocks <= 0) return hipSuccess;
  const dim3 grid(a.s.num_blocks), block(kBlock);
  if (history == 0) {
    if (layout == 0)
      bucket_kernel<float, 0><<<grid, block, lds_bytes, s>>>(a);
    else
      bucket_kernel<float, 1><<<grid, block, lds_bytes, s>>>(a);
  } else {
    if (layout == 0)
      bucket_kernel<uint16_t, 0><<<grid, block, lds_bytes, s>>>(a);
    else
      bucket_kernel<uint16_t, 1><<<grid, block, lds_bytes, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace lev
