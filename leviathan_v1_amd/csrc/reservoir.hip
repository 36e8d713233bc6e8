// reservoir.hip -- feature capture and readout kernels (gfx950).
//
// The reference's reservoir (python/reservoir.py:45-86) copies all N phases to
// the host after every step, builds cos/sin features there and fits
// sklearn's Ridge in FP64 on the CPU.  Here the features of the chosen nodes
// are written on the device after every step (capture_kernel), and the ridge
// normal equations are formed on the device:
//
//   colsum_kernel        sum_t w_t X[t][f] in FP64 (means, X^T y, X^T v)
//   matvec_kernel        sum_f (X[t][f] - mu_f) w_f in FP64 (residuals)
//   center_bf16_kernel   Xc^T = (X - mu)^T rounded to BF16, feature-major,
//                        zero-padded to whole 256-feature / 64-sample tiles
//   gram_kernel          G = Xc^T Xc on the matrix cores: v_mfma_f32_32x32x16_bf16
//                        (BF16 inputs, FP32 accumulation); 128x256 tiles on or
//                        above the diagonal, LDS double buffering, mirrored.
//
// The readout factorises G + alpha I once in FP64 on the device (rocSOLVER,
// host code in leviathan_api.hip) and refines the solution with residuals
// b - (Xc^T Xc + alpha I) w computed by the FP64 kernels above, so the BF16
// Gram only preconditions and the answer is the FP64 ridge solution.  For
// features too ill-conditioned for that, the Gram is formed in FP64
// (center_fp64_kernel, then rocBLAS DSYRK) instead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "leviathan_internal.h"
#include "step_kernel.h"

namespace lev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// After the finalize of step t (the step counter then reads t + 1) the new
// phases of the feature nodes become row t - base of X.
__global__ void capture_kernel(const float* __restrict__ theta, const int32_t* __restrict__ nodes, int32_t n_nodes,
                               float* __restrict__ X, int64_t max_rows, const int64_t* __restrict__ step,
                               const int64_t* __restrict__ base) {
  const int64_t row = *step - *base - 1;
  if (row < 0 || row >= max_rows) return;
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_nodes) return;
  float s, c;
  sincosf(theta[nodes[f]], &s, &c);
  float* out = X + row * (int64_t)(2 * n_nodes);
  out[f] = c;
  out[n_nodes + f] = s;
}

// partial[c][f] = sum over rows of chunk c of w[t] * X[row0 + t][f]
__global__ void colsum_kernel(const float* __restrict__ X, int64_t ld, int64_t row0, int64_t nrows,
                              const double* __restrict__ w, int32_t F, int64_t rows_per_chunk,
                              double* __restrict__ partial) {
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int64_t t0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t t1 = t0 + rows_per_chunk < nrows ? t0 + rows_per_chunk : nrows;
  double acc = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const double x = (double)X[(row0 + t) * ld + f];
    acc += w ? w[t] * x : x;
  }
  partial[(int64_t)blockIdx.y * F + f] = acc;
}

__global__ void colsum_final_kernel(const double* __restrict__ partial, int32_t chunks, int32_t F,
                                    double* __restrict__ out) {
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  double acc = 0.0;
  for (int32_t c = 0; c < chunks; ++c) acc += partial[(int64_t)c * F + f];
  out[f] = acc;
}

// out[t] = sum_f (X[row0 + t][f] - mean[f]) * w[f]; one wave per row.
__global__ __launch_bounds__(256) void matvec_kernel(const float* __restrict__ X, int64_t ld, int64_t row0,
                                                     int64_t nrows, const double* __restrict__ mean,
                                                     const double* __restrict__ w, int32_t F,
                                                     double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= nrows) return;
  const float* row = X + (row0 + t) * ld;
  double acc = 0.0;
  for (int32_t f = lane; f < F; f += 64) acc += ((double)row[f] - mean[f]) * w[f];
  acc = wave_sum(acc);
  if (lane == 0) out[t] = acc;
}

__device__ __forceinline__ uint16_t to_bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// XcT[f][t] = bf16(X[row0 + t][f] - mean[f]) for f < F, t < nrows; zero in
// the padding (F <= f < Fp, nrows <= t < Tp).  64x64 tiles through LDS.
__global__ __launch_bounds__(256) void center_bf16_kernel(const float* __restrict__ X, int64_t ld, int64_t row0,
                                                          int64_t nrows, const double* __restrict__ mean,
                                                          int32_t F, uint16_t* __restrict__ XcT, int64_t Tp,
                                                          uint16_t* __restrict__ XcT_lo) {
  __shared__ float tile[64][65];
  const int32_t f0 = blockIdx.x * 64;
  const int64_t t0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int64_t t = t0 + r;
    const int32_t f = f0 + tx;
    float v = 0.0f;
    if (t < nrows && f < F) v = (float)((double)X[(row0 + t) * ld + f] - mean[f]);
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int32_t f = f0 + r;
    const int64_t t = t0 + tx;
    const float x = tile[tx][r];
    const uint16_t hi = to_bf16_rne(x);
    XcT[(int64_t)f * Tp + t] = hi;
    // split form: lo = bf16(x - hi), so hi + lo carries ~16 significant bits
    if (XcT_lo) XcT_lo[(int64_t)f * Tp + t] = to_bf16_rne(x - __uint_as_float((uint32_t)hi << 16));
  }
}

// Split-BF16 Gram: G = Hi Hi^T + Hi Lo^T + Lo Hi^T with X = Hi + Lo, three
// BF16 products per block on the matrix cores, FP32 accumulation; the
// relative error of the inputs drops from 2^-9 to ~2^-17.  128x128 tiles on
// or above the diagonal, four waves of 64x64, fragments read straight from
// the L2 (the layouts of gram_kernel's comment apply).
__global__ __launch_bounds__(256) void gram_split_kernel(const uint16_t* __restrict__ Hi,
                                                         const uint16_t* __restrict__ Lo, int64_t Tp, int32_t nt,
                                                         int32_t F, float* __restrict__ G) {
  int32_t b = blockIdx.x, ti = 0;
  while (b >= nt - ti) {
    b -= nt - ti;
    ++ti;
  }
  const int32_t tj = ti + b;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int32_t ib = ti * 128 + (wave >> 1) * 64;
  const int32_t jb = tj * 128 + (wave & 1) * 64;
  const int lm = lane & 31;
  const int64_t lk = (lane >> 5) * 8;
  const int64_t oa0 = (int64_t)(ib + lm) * Tp + lk, oa1 = (int64_t)(ib + 32 + lm) * Tp + lk;
  const int64_t ob0 = (int64_t)(jb + lm) * Tp + lk, ob1 = (int64_t)(jb + 32 + lm) * Tp + lk;
  f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};
  const int64_t steps = Tp / 16;
  for (int64_t s = 0; s < steps; ++s) {
    const int64_t k = 16 * s;
    const bf16x8 ha0 = *reinterpret_cast<const bf16x8*>(Hi + oa0 + k);
    const bf16x8 ha1 = *reinterpret_cast<const bf16x8*>(Hi + oa1 + k);
    const bf16x8 hb0 = *reinterpret_cast<const bf16x8*>(Hi + ob0 + k);
    const bf16x8 hb1 = *reinterpret_cast<const bf16x8*>(Hi + ob1 + k);
    const bf16x8 la0 = *reinterpret_cast<const bf16x8*>(Lo + oa0 + k);
    const bf16x8 la1 = *reinterpret_cast<const bf16x8*>(Lo + oa1 + k);
    const bf16x8 lb0 = *reinterpret_cast<const bf16x8*>(Lo + ob0 + k);
    const bf16x8 lb1 = *reinterpret_cast<const bf16x8*>(Lo + ob1 + k);
    c00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la0, hb0, c00, 0, 0, 0);
    c00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha0, lb0, c00, 0, 0, 0);
    c00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha0, hb0, c00, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la0, hb1, c01, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha0, lb1, c01, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha0, hb1, c01, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la1, hb0, c10, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha1, lb0, c10, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha1, hb0, c10, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la1, hb1, c11, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha1, lb1, c11, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha1, hb1, c11, 0, 0, 0);
  }
  const f32x16 cc[2][2] = {{c00, c01}, {c10, c11}};
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int32_t i = ib + 32 * p + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
        const int32_t j = jb + 32 * q + lm;
        // (i, j) and (j, i) sum the three products in different orders: keep
        // i <= j and mirror it, so the result is exactly symmetric
        if (i < F && j < F && (ti != tj || i <= j)) {
          G[(int64_t)i * F + j] = cc[p][q][v];
          G[(int64_t)j * F + i] = cc[p][q][v];
        }
      }
}

// G = XcT XcT^T on the matrix cores.  Workgroup tile: 128 rows (i) x 256
// columns (j); tiles entirely below the diagonal are skipped, and inside a
// tile only i <= j is stored, together with its mirror (j, i).  Four waves,
// each a 64 x 128 sub-tile = 2 x 4 blocks of v_mfma_f32_32x32x16_bf16 (eight
// 32x32 FP32 accumulators).  The K dimension (samples) advances 32 at a time
// through double-buffered LDS: while the matrix cores work on one slice the
// next slice is already in registers.  LDS rows are padded from 64 to 80
// bytes so the 16-byte fragment reads of 32 consecutive rows spread over the
// banks.
//
// MFMA operand fragments: lane l holds row (l % 32) of its 32-row block and
// samples 8 * (l / 32) .. +7 of the 16-sample step; accumulator register v of
// lane l is element (8 * (v / 4) + 4 * (l / 32) + v % 4, l % 32) of the block.
constexpr int kGramBM = 128, kGramBN = 256, kGramBK = 32;
constexpr int kGramPitch = 40;  // bf16 per LDS row: 32 samples + 8 padding

__global__ __launch_bounds__(256) void gram_kernel(const uint16_t* __restrict__ XcT, int64_t Tp, int32_t nI,
                                                   int32_t nJ, int32_t F, float* __restrict__ G) {
  __shared__ __align__(16) uint16_t sA[2][kGramBM * kGramPitch];
  __shared__ __align__(16) uint16_t sB[2][kGramBN * kGramPitch];
  // tile enumeration: row block I needs column blocks J >= I / 2
  int32_t b = blockIdx.x, I = 0;
  while (b >= nJ - I / 2) {
    b -= nJ - I / 2;
    ++I;
  }
  const int32_t J = I / 2 + b;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int32_t i0 = I * kGramBM, j0 = J * kGramBN;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 128;  // wave's sub-tile inside the workgroup tile

  // global -> register staging: A slice is 128 rows x 64 B (512 chunks of 16 B),
  // B slice 256 rows x 64 B (1024 chunks); thread tid takes chunks tid + 256 q.
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  u32x4 ra[2], rb[4];
  auto load_slice = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      ra[q] = *reinterpret_cast<const u32x4*>(XcT + (int64_t)(i0 + row) * Tp + k0 + part * 8);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      rb[q] = *reinterpret_cast<const u32x4*>(XcT + (int64_t)(j0 + row) * Tp + k0 + part * 8);
    }
  };
  auto store_slice = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      *reinterpret_cast<u32x4*>(&sA[buf][row * kGramPitch + part * 8]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = tid + 256 * q, row = c >> 2, part = c & 3;
      *reinterpret_cast<u32x4*>(&sB[buf][row * kGramPitch + part * 8]) = rb[q];
    }
  };

  f32x16 acc[2][4];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = f32x16{};

  const int64_t nk = Tp / kGramBK;
  load_slice(0);
  store_slice(0);
  __syncthreads();
  const int lrow = lane & 31, lk = (lane >> 5) * 8;
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) load_slice((kt + 1) * kGramBK);
#pragma unroll
    for (int kk = 0; kk < kGramBK; kk += 16) {
      bf16x8 fa[2], fb[4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
        fa[p] = *reinterpret_cast<const bf16x8*>(&sA[buf][(wr + 32 * p + lrow) * kGramPitch + kk + lk]);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        fb[q] = *reinterpret_cast<const bf16x8*>(&sB[buf][(wc + 32 * q + lrow) * kGramPitch + kk + lk]);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc[p][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[p], fb[q], acc[p][q], 0, 0, 0);
    }
    if (kt + 1 < nk) store_slice(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int32_t i = i0 + wr + 32 * p + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
        const int32_t j = j0 + wc + 32 * q + lrow;
        if (i <= j && j < F) {
          G[(int64_t)i * F + j] = acc[p][q][v];
          if (i != j) G[(int64_t)j * F + i] = acc[p][q][v];
        }
      }
}

// ---------------------------------------------------------------------------

hipError_t launch_capture(const float* theta, const int32_t* nodes, int32_t n_nodes, float* X, int64_t max_rows,
                          const int64_t* step, const int64_t* base, hipStream_t s) {
  const dim3 grid((n_nodes + 255) / 256), block(256);
  capture_kernel<<<grid, block, 0, s>>>(theta, nodes, n_nodes, X, max_rows, step, base);
  return hipGetLastError();
}

hipError_t launch_colsum(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* w, int32_t F,
                         double* partial, int32_t chunks, double* out, hipStream_t s) {
  const int64_t rpc = (nrows + chunks - 1) / chunks;
  const dim3 grid((F + 255) / 256, chunks), block(256);
  colsum_kernel<<<grid, block, 0, s>>>(X, ld, row0, nrows, w, F, rpc, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  colsum_final_kernel<<<dim3((F + 255) / 256), block, 0, s>>>(partial, chunks, F, out);
  return hipGetLastError();
}

hipError_t launch_matvec(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                         const double* w, int32_t F, double* out, hipStream_t s) {
  const dim3 grid((unsigned)((nrows + 3) / 4)), block(256);
  matvec_kernel<<<grid, block, 0, s>>>(X, ld, row0, nrows, mean, w, F, out);
  return hipGetLastError();
}

hipError_t launch_center_bf16(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                              int32_t F, int32_t Fp, uint16_t* XcT, int64_t Tp, uint16_t* XcT_lo, hipStream_t s) {
  const dim3 grid(Fp / 64, (unsigned)(Tp / 64)), block(256);
  center_bf16_kernel<<<grid, block, 0, s>>>(X, ld, row0, nrows, mean, F, XcT, Tp, XcT_lo);
  return hipGetLastError();
}

hipError_t launch_gram(const uint16_t* XcT, int64_t Tp, int32_t Fp, int32_t F, float* G, hipStream_t s) {
  // Fp is a multiple of 256 and Tp of 64 (both zero-padded by center_bf16_kernel)
  const int32_t nI = Fp / kGramBM, nJ = Fp / kGramBN;
  int32_t tiles = 0;
  for (int32_t I = 0; I < nI; ++I) tiles += nJ - I / 2;
  gram_kernel<<<dim3(tiles), dim3(256), 0, s>>>(XcT, Tp, nI, nJ, F, G);
  return hipGetLastError();
}

// C = G widened to FP64.
__global__ __launch_bounds__(256) void widen_kernel(const float* G, int64_t n, double* C) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    C[q] = (double)G[q];
}

// A = C + diag_add I: the matrix the readout factorises.
__global__ __launch_bounds__(256) void shifted_kernel(const double* C, int32_t F, double diag_add, double* A) {
  const int64_t n = (int64_t)F * F;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = (int32_t)(q / F), j = (int32_t)(q % F);
    A[q] = C[q] + (i == j ? diag_add : 0.0);
  }
}

// The mean of C's diagonal (one workgroup, a fixed order).
__global__ __launch_bounds__(256) void diag_mean_kernel(const double* C, int32_t F, double* diag_mean) {
  __shared__ double s_part[256 / 64];
  double acc = 0.0;
  for (int32_t i = threadIdx.x; i < F; i += 256) acc += C[(int64_t)i * F + i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *diag_mean = (s_part[0] + s_part[1] + s_part[2] + s_part[3]) / (double)F;
}

// Xc = X[row0 : row0 + nrows] - mean in FP64, [nrows][F] row-major (for the
// FP64 Gram: column-major F x nrows with leading dimension F).
__global__ __launch_bounds__(256) void center_fp64_kernel(const float* __restrict__ X, int64_t ld, int64_t row0,
                                                          int64_t nrows, const double* __restrict__ mean, int32_t F,
                                                          double* __restrict__ Xc) {
  const int64_t n = nrows * F;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = q / F;
    const int32_t f = (int32_t)(q % F);
    Xc[q] = (double)X[(row0 + t) * ld + f] - mean[f];
  }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

hipError_t launch_widen(const float* G, int64_t n, double* C, hipStream_t s) {
  widen_kernel<<<grid_for(n), 256, 0, s>>>(G, n, C);
  return hipGetLastError();
}

hipError_t launch_shifted(const double* C, int32_t F, double diag_add, double* A, hipStream_t s) {
  shifted_kernel<<<grid_for((int64_t)F * F), 256, 0, s>>>(C, F, diag_add, A);
  return hipGetLastError();
}

hipError_t launch_diag_mean(const double* C, int32_t F, double* diag_mean, hipStream_t s) {
  diag_mean_kernel<<<1, 256, 0, s>>>(C, F, diag_mean);
  return hipGetLastError();
}

hipError_t launch_center_fp64(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                              int32_t F, double* Xc, hipStream_t s) {
  center_fp64_kernel<<<grid_for(nrows * F), 256, 0, s>>>(X, ld, row0, nrows, mean, F, Xc);
  return hipGetLastError();
}

hipError_t launch_gram_split(const uint16_t* Hi, const uint16_t* Lo, int64_t Tp, int32_t Fp, int32_t F, float* G,
                             hipStream_t s) {
  const int32_t nt = Fp / 128;
  gram_split_kernel<<<dim3(nt * (nt + 1) / 2), dim3(256), 0, s>>>(Hi, Lo, Tp, nt, F, G);
  return hipGetLastError();
}

}  // namespace lev
