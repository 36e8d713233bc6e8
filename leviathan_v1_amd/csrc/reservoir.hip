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
//                        zero-padded to whole 128-feature / 16-sample tiles
//   gram_kernel          G = Xc^T Xc on the matrix cores: v_mfma_f32_32x32x16_bf16
//                        (BF16 inputs, FP32 accumulation); one 128x128 tile per
//                        workgroup, upper-triangle tiles only, mirrored.
//
// The host factorises G + alpha I once in FP64 and refines the solution with
// residuals b - (Xc^T Xc + alpha I) w computed by the FP64 kernels above, so
// the BF16 Gram only preconditions and the answer is the FP64 ridge solution.
#include <hip/hip_runtime.h>
#include <stdint.h>

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
                                                          int32_t F, uint16_t* __restrict__ XcT, int64_t Tp) {
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
    XcT[(int64_t)f * Tp + t] = to_bf16_rne(tile[tx][r]);
  }
}

// G[i][j] = sum_t XcT[i][t] XcT[j][t] for the 128x128 tile (ti, tj), ti <= tj,
// written to both (i, j) and (j, i).  Four waves, each a 64x64 quarter as
// 2x2 blocks of v_mfma_f32_32x32x16_bf16.  Operand fragments: lane l holds
// row (l % 32) of its 32-row block and the 8 consecutive samples
// k0 + 8 * (l / 32) .. +7; accumulator register v of lane l is element
// (8 * (v / 4) + 4 * (l / 32) + v % 4, l % 32) of the 32x32 block.
__global__ __launch_bounds__(256) void gram_kernel(const uint16_t* __restrict__ XcT, int64_t Tp, int32_t nt,
                                                   int32_t F, float* __restrict__ G) {
  // upper-triangle tile enumeration, row by row
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
  const bf16x8* a0 = reinterpret_cast<const bf16x8*>(XcT + (int64_t)(ib + lm) * Tp + lk);
  const bf16x8* a1 = reinterpret_cast<const bf16x8*>(XcT + (int64_t)(ib + 32 + lm) * Tp + lk);
  const bf16x8* b0 = reinterpret_cast<const bf16x8*>(XcT + (int64_t)(jb + lm) * Tp + lk);
  const bf16x8* b1 = reinterpret_cast<const bf16x8*>(XcT + (int64_t)(jb + 32 + lm) * Tp + lk);
  f32x16 c00 = {}, c01 = {}, c10 = {}, c11 = {};
  const int64_t steps = Tp / 16;  // Tp is a multiple of 64
#pragma unroll 4
  for (int64_t s = 0; s < steps; ++s) {
    const bf16x8 va0 = a0[2 * s], va1 = a1[2 * s], vb0 = b0[2 * s], vb1 = b1[2 * s];
    c00 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va0, vb0, c00, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va0, vb1, c01, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va1, vb0, c10, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va1, vb1, c11, 0, 0, 0);
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
        if (i < F && j < F) {
          G[(int64_t)i * F + j] = cc[p][q][v];
          if (ti != tj) G[(int64_t)j * F + i] = cc[p][q][v];
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
                              int32_t F, int32_t Fp, uint16_t* XcT, int64_t Tp, hipStream_t s) {
  const dim3 grid(Fp / 64, (unsigned)(Tp / 64)), block(256);
  center_bf16_kernel<<<grid, block, 0, s>>>(X, ld, row0, nrows, mean, F, XcT, Tp);
  return hipGetLastError();
}

hipError_t launch_gram(const uint16_t* XcT, int64_t Tp, int32_t Fp, int32_t F, float* G, hipStream_t s) {
  const int32_t nt = Fp / 128;
  const dim3 grid(nt * (nt + 1) / 2), block(256);
  gram_kernel<<<grid, block, 0, s>>>(XcT, Tp, nt, F, G);
  return hipGetLastError();
}

}  // namespace lev
