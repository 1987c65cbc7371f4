// Measurement of the MFMA alternative for the sum-factorization contractions:
// the 3-point 1D basis contraction y = S x (S the 3 x 3 table of the Q2
// basis at the Gauss points, x the 3 nodal values of one line of a cell) of
// many lines, FP32, done
//   variant 0: as VALU FMAs, one line per thread (the kernels' form);
//   variant 1: on the matrix cores, packed as dense as a 3 x 3 operator allows:
//              A = blockdiag(S x 5) padded to 16 x 16, B = 16 lines of each of
//              5 line groups (rows: group x component, row 15 zero), one
//              16 x 16 x 16 product = 4 v_mfma_f32_16x16x4f32 for 80 lines.
// Each line is contracted reps times (VALU: y <- S y + x, a dependent chain;
// MFMA: C <- A B + C with four independent accumulators), so that the time is
// the contraction rate, not memory.  Useful work is 9 FMA per line and
// repetition in both variants.  No counterpart in the reference.

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aspect_mfs.h"
#include "mfs_internal.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// the Q2 interpolation table at the 3 Gauss points of [0, 1]
__host__ __device__ inline float s_tab(int q, int i) {
  const float g = 0.5f * 0.7745966692414834f;  // Gauss points 0.5 -+ g, 0.5
  const float x = q == 0 ? 0.5f - g : (q == 1 ? 0.5f : 0.5f + g);
  const float n[3] = {0.0f, 0.5f, 1.0f};
  float v = 1.0f;
  for (int j = 0; j < 3; ++j)
    if (j != i) v *= (x - n[j]) / (n[i] - n[j]);
  return v;
}

// lines are stored component-planar: x[comp * n + line]
__global__ void k_probe_fma(const long long n, const int reps, const float *__restrict__ x, float *__restrict__ y) {
  const long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (l >= n) return;
  float S[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int i = 0; i < 3; ++i) S[q][i] = s_tab(q, i);
  const float x0 = x[l], x1 = x[n + l], x2 = x[2 * n + l];
  float a0 = x0, a1 = x1, a2 = x2;
  for (int r = 0; r < reps; ++r) {
    const float b0 = S[0][0] * a0 + S[0][1] * a1 + S[0][2] * a2 + x0;
    const float b1 = S[1][0] * a0 + S[1][1] * a1 + S[1][2] * a2 + x1;
    const float b2 = S[2][0] * a0 + S[2][1] * a1 + S[2][2] * a2 + x2;
    a0 = b0;
    a1 = b1;
    a2 = b2;
  }
  y[l] = a0;
  y[n + l] = a1;
  y[2 * n + l] = a2;
}

// one wave per tile of 80 lines (5 groups of 16); lane layout of
// v_mfma_f32_16x16x4f32 on a 64-wide wave: A[row = lane % 16][k = lane / 16],
// B[k = lane / 16][col = lane % 16], result rows 4 (lane / 16) + j, column
// lane % 16 in element j.  n is a multiple of 80 (checked by the host).
__global__ void __launch_bounds__(256) k_probe_mfma(const long long n, const int reps, const float *__restrict__ x,
                                                    float *__restrict__ y) {
  const long long tile = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / 64;
  if (tile * 80 >= n) return;
  const int lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
  float a[4], b[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = kq + 4 * s, row = col;  // A: this lane's row is lane % 16
    a[s] = (row < 15 && k < 15 && row / 3 == k / 3) ? s_tab(row % 3, k % 3) : 0.0f;
    // B: this lane's k is kq + 4 s, its column the line col of group k / 3
    b[s] = k < 15 ? x[(long long)(k % 3) * n + tile * 80 + (k / 3) * 16 + col] : 0.0f;
  }
  f4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int r = 0; r < reps; ++r) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b[s], acc[u], 0, 0, 0);
  }
  // every accumulator holds reps * S x; write the sum of the four (4 reps * S x)
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = 4 * kq + j;
    if (row >= 15) continue;
    const float v = acc[0][j] + acc[1][j] + acc[2][j] + acc[3][j];
    y[(long long)(row % 3) * n + tile * 80 + (row / 3) * 16 + col] = v;
  }
}

}  // namespace

extern "C" int mfs_probe_contraction(int variant, long long n_lines, int reps, double *seconds, double *out) {
  try {
    if (variant != 0 && variant != 1) throw std::invalid_argument("mfs_probe_contraction: variant 0 (FMA) or 1 (MFMA)");
    if (n_lines <= 0 || n_lines % 80 != 0 || n_lines > (1LL << 28))
      throw std::invalid_argument("mfs_probe_contraction: n_lines must be a positive multiple of 80, at most 2^28");
    if (reps < 1 || !seconds) throw std::invalid_argument("mfs_probe_contraction: reps >= 1 and an output required");
    std::vector<float> hx(3 * n_lines);
    for (long long i = 0; i < 3 * n_lines; ++i) hx[i] = (float)((i * 2654435761LL) % 1000) * 1e-3f - 0.5f;
    float *dx = nullptr, *dy = nullptr;
    MFS_HIP(hipMalloc(reinterpret_cast<void **>(&dx), 3 * n_lines * sizeof(float)));
    MFS_HIP(hipMalloc(reinterpret_cast<void **>(&dy), 3 * n_lines * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
      MFS_HIP(hipMemcpy(dx, hx.data(), 3 * n_lines * sizeof(float), hipMemcpyHostToDevice));
      MFS_HIP(hipEventCreate(&e0));
      MFS_HIP(hipEventCreate(&e1));
      const unsigned blocks = variant == 0 ? (unsigned)((n_lines + 255) / 256) : (unsigned)((n_lines / 80 * 64 + 255) / 256);
      auto launch = [&](int rr) {
        if (variant == 0)
          hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(256), 0, 0, n_lines, rr, dx, dy);
        else
          hipLaunchKernelGGL(k_probe_mfma, dim3(blocks), dim3(256), 0, 0, n_lines, rr, dx, dy);
        MFS_HIP(hipGetLastError());
      };
      launch(reps);  // warm-up
      MFS_HIP(hipEventRecord(e0, 0));
      launch(reps);
      MFS_HIP(hipEventRecord(e1, 0));
      MFS_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
      *seconds = ms * 1e-3;
      if (out) {
        // one repetition, for the cross-check of the two variants: FMA gives
        // S x + x, MFMA 4 S x
        launch(1);
        std::vector<float> hy(3 * n_lines);
        MFS_HIP(hipMemcpy(hy.data(), dy, 3 * n_lines * sizeof(float), hipMemcpyDeviceToHost));
        for (long long i = 0; i < 3 * n_lines; ++i) out[i] = hy[i];
      }
    } catch (...) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipFree(dx);
      (void)hipFree(dy);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return MFS_OK;
  } catch (const std::exception &e) {
    mfs::set_last_error(e.what());
    return MFS_ERROR;
  }
}
