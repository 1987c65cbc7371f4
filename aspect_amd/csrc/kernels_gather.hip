// gfx950 kernels of the split A-block apply on 3D degree-2 box levels (the
// A block of every V-cycle level) and the dispatch of the 3D Q2 x Q1 Stokes
// and A-block applies to the plane kernel (kernels_planes.hip).
//
// Split form: the plane kernel writes each tile's in-tile node sums to a work
// array [tile][comp][lz][ly][lx] in one launch over all tiles; a node-wise pass
// sums the at most 8 tiles sharing each DoF in a fixed order (deterministic,
// no tile colours) and applies the operator mode, or, fused, one
// Jacobi-Chebyshev step of the smoother.
//
// Restates ABlockOperator::local_apply / apply_add
// (source/simulator/solver/matrix_free_operators.cc:1029-1064) for the sums
// over cells, and deal.II's PreconditionChebyshev step for the fused form.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mfs_internal.h"
#include "tile_common.h"

namespace mfs {
namespace {

// Sum over the at most 8 tiles sharing velocity node g of component comp,
// from the work array [tile][comp][lz][ly][lx] of in-tile sums, in a fixed
// order (deterministic).  Per direction the candidates of node g are
// (tile g/8 - 1, local node 8) for g mod 8 == 0, g > 0, and (tile g/8, local
// node g mod 8) while g/8 < ntiles.  32-bit offsets (checked by the launcher).
template <typename T>
__device__ __forceinline__ T tile_gather_sum(const LevelDesc &L, const T *__restrict__ work, int comp,
                                             const int (&g)[3]) {
  constexpr int TN = 3 * 729;
  int tt[3][2], ll[3][2];
  bool ok[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    tt[d][0] = (g[d] >> 3) - 1;
    ll[d][0] = 8;
    ok[d][0] = (g[d] & 7) == 0 && g[d] > 0;
    tt[d][1] = g[d] >> 3;
    ll[d][1] = g[d] & 7;
    ok[d][1] = (g[d] >> 3) < L.ntiles[d];
  }
  const int sty = L.ntiles[0], stz = L.ntiles[0] * L.ntiles[1];
  T val = 0;
#pragma unroll
  for (int sz = 0; sz < 2; ++sz)
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
      for (int sx = 0; sx < 2; ++sx) {
        const bool v = ok[2][sz] && ok[1][sy] && ok[0][sx];
        const int tile = tt[0][sx] + sty * tt[1][sy] + stz * tt[2][sz];
        const int wi = v ? tile * TN + ((comp * 9 + ll[2][sz]) * 9 + ll[1][sy]) * 9 + ll[0][sx] : 0;
        const T c = work[wi];
        val += v ? c : (T)0;
      }
  return val;
}

// Second pass of the split A-block apply: one thread per velocity DoF adds
// the tiles sharing the node and applies the mode; constrained rows take src
// (identity).  Grid: x over the nodes of an xy plane, y over (z, comp).
template <int MODE, typename T>
__global__ void k_gather_tiles(const LevelDesc L, const T *__restrict__ work, const T *__restrict__ src_u,
                               const T *__restrict__ rhs_u, T *__restrict__ dst_u) {
  const unsigned plane = (unsigned)(L.nv[0] * L.nv[1]);
  // XCD-aware order of the (x, y) blocks: node rows that share lines of the work array meet in one L2
  const unsigned m = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned by = m / gridDim.x, bx = m - by * gridDim.x;
  const unsigned pidx = bx * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = (int)by;  // gz + nv2 * comp
  const int comp = row / L.nv[2], gz = row - comp * L.nv[2];
  const int gy = (int)(pidx / (unsigned)L.nv[0]), gx = (int)pidx - gy * L.nv[0];
  const int g[3] = {gx, gy, gz};
  const T val = tile_gather_sum<T>(L, work, comp, g);
  const long long gi = comp * L.vnodes + pidx + (long long)plane * gz;
  T vv = val;
  if (con_bf(L, comp, gx, gy, gz)) vv = src_u[gi];
  if constexpr (MODE == MODE_SET)
    dst_u[gi] = vv;
  else if constexpr (MODE == MODE_RESID)
    dst_u[gi] = rhs_u[gi] - vv;
  else
    dst_u[gi] += vv;
}

// The same second pass fused with a Jacobi-Chebyshev update: r = b - A x at
// each DoF (not stored), then d = dinv r inv_theta (first step of a sweep
// from a nonzero x) or d = f1 d + f2 dinv r, and x += d -- the arithmetic of
// v_cheb_first / v_cheb_step, so the result is that of the unfused sequence.
template <typename T>
__global__ void k_gather_cheb(const LevelDesc L, const T *__restrict__ work, T *x, const T *__restrict__ b,
                              const T *__restrict__ dinv, T *__restrict__ d, const T f1, const T f2,
                              const int first) {
  const unsigned plane = (unsigned)(L.nv[0] * L.nv[1]);
  // XCD-aware order of the (x, y) blocks: node rows that share lines of the work array meet in one L2
  const unsigned m = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned by = m / gridDim.x, bx = m - by * gridDim.x;
  const unsigned pidx = bx * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = (int)by;
  const int comp = row / L.nv[2], gz = row - comp * L.nv[2];
  const int gy = (int)(pidx / (unsigned)L.nv[0]), gx = (int)pidx - gy * L.nv[0];
  const int g[3] = {gx, gy, gz};
  const T val = tile_gather_sum<T>(L, work, comp, g);
  const long long gi = comp * L.vnodes + pidx + (long long)plane * gz;
  const T xi = x[gi];
  const T r = b[gi] - (con_bf(L, comp, gx, gy, gz) ? xi : val);
  const T dd = first ? dinv[gi] * r * f2 : f1 * d[gi] + f2 * dinv[gi] * r;
  d[gi] = dd;
  x[gi] = xi + dd;
}

// Second pass of the split Stokes apply for the pressure: one thread per
// pressure node sums the at most 8 tiles sharing it from work_p
// [tile][plz][ply][plx] (candidates per direction: tile g/4 - 1 with local
// node 4 for g mod 4 == 0, g > 0, and tile g/4 with local node g mod 4) in a
// fixed order, then applies the mode.
template <int MODE, typename T>
__global__ void k_gather_p(const LevelDesc L, const T *__restrict__ work_p, const T *__restrict__ rhs_p,
                           T *__restrict__ dst_p) {
  const unsigned plane = (unsigned)(L.np[0] * L.np[1]);
  const unsigned m = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned by = m / gridDim.x, bx = m - by * gridDim.x;
  const unsigned pidx = bx * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int gz = (int)by;
  const int gy = (int)(pidx / (unsigned)L.np[0]), gx = (int)pidx - gy * L.np[0];
  const int g[3] = {gx, gy, gz};
  int tt[3][2], ll[3][2];
  bool ok[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    tt[d][0] = (g[d] >> 2) - 1;
    ll[d][0] = 4;
    ok[d][0] = (g[d] & 3) == 0 && g[d] > 0;
    tt[d][1] = g[d] >> 2;
    ll[d][1] = g[d] & 3;
    ok[d][1] = (g[d] >> 2) < L.ntiles[d];
  }
  const int sty = L.ntiles[0], stz = L.ntiles[0] * L.ntiles[1];
  T val = 0;
#pragma unroll
  for (int sz = 0; sz < 2; ++sz)
#pragma unroll
    for (int sy = 0; sy < 2; ++sy)
#pragma unroll
      for (int sx = 0; sx < 2; ++sx) {
        const bool v = ok[2][sz] && ok[1][sy] && ok[0][sx];
        const int tile = tt[0][sx] + sty * tt[1][sy] + stz * tt[2][sz];
        const T c = work_p[v ? tile * 125 + (ll[2][sz] * 5 + ll[1][sy]) * 5 + ll[0][sx] : 0];
        val += v ? c : (T)0;
      }
  const long long gi = pidx + (long long)plane * gz;
  if constexpr (MODE == MODE_SET)
    dst_p[gi] = val;
  else if constexpr (MODE == MODE_RESID)
    dst_p[gi] = rhs_p[gi] - val;
  else
    dst_p[gi] += val;
}

// split Stokes apply: one launch of the plane kernel over all tiles, then the
// node-wise sums of the velocity and the pressure
template <typename T>
void launch_stokes_split(Mode mode, const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru,
                         const T *rp, T *du, T *dp, T *work, hipStream_t st) {
  if (split_work_size(L) >= (1LL << 31) || (long long)L.nv[0] * L.nv[1] >= (1LL << 31))
    throw std::runtime_error("level too large for the split Stokes apply (32-bit work offsets)");
  T *const work_p = work + split_work_size(L);
  launch_planes_split_stokes<T>(L, su, sp, eta, pq, work, work_p, st);
  const dim3 grid((unsigned)(((long long)L.nv[0] * L.nv[1] + 511) / 512), (unsigned)(3 * L.nv[2]));
  const dim3 gridp((unsigned)(((long long)L.np[0] * L.np[1] + 255) / 256), (unsigned)L.np[2]);
  if (mode == MODE_SET) {
    hipLaunchKernelGGL((k_gather_tiles<MODE_SET, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
    hipLaunchKernelGGL((k_gather_p<MODE_SET, T>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  } else if (mode == MODE_RESID) {
    hipLaunchKernelGGL((k_gather_tiles<MODE_RESID, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
    hipLaunchKernelGGL((k_gather_p<MODE_RESID, T>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  } else {
    hipLaunchKernelGGL((k_gather_tiles<MODE_ADD, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
    hipLaunchKernelGGL((k_gather_p<MODE_ADD, T>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  }
  MFS_HIP(hipGetLastError());
}

// split A-block apply: per-cell contributions of all tiles in one launch, then the node-wise sum
template <typename T>
void launch_ablock_split(Mode mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                         T *work, hipStream_t st) {
  launch_planes_split<T>(L, su, eta, pq, work, st);
  if (split_work_size(L) >= (1LL << 31) || (long long)L.nv[0] * L.nv[1] >= (1LL << 31))
    throw std::runtime_error("level too large for the split A-block apply (32-bit work offsets)");
  // 512 nodes (two rows of x) per workgroup: a 128-byte line of the work
  // array holds parts of several node rows, and one CU then consumes more of it
  const dim3 grid((unsigned)(((long long)L.nv[0] * L.nv[1] + 511) / 512), (unsigned)(3 * L.nv[2]));
  if (mode == MODE_SET)
    hipLaunchKernelGGL((k_gather_tiles<MODE_SET, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
  else if (mode == MODE_RESID)
    hipLaunchKernelGGL((k_gather_tiles<MODE_RESID, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
  else
    hipLaunchKernelGGL((k_gather_tiles<MODE_ADD, T>), grid, dim3(512), 0, st, L, work, su, ru, du);
  MFS_HIP(hipGetLastError());
}

template <typename T>
void launch_ablock_cheb_split(const LevelDesc &L, T *x, const T *b, const T *eta, int pq, const T *dinv, T *d, T f1,
                              T f2, int first, T *work, hipStream_t st) {
  launch_planes_split<T>(L, (const T *)x, eta, pq, work, st);
  if (split_work_size(L) >= (1LL << 31) || (long long)L.nv[0] * L.nv[1] >= (1LL << 31))
    throw std::runtime_error("level too large for the split A-block apply (32-bit work offsets)");
  const dim3 grid((unsigned)(((long long)L.nv[0] * L.nv[1] + 511) / 512), (unsigned)(3 * L.nv[2]));
  hipLaunchKernelGGL((k_gather_cheb<T>), grid, dim3(512), 0, st, L, (const T *)work, x, b, dinv, d, f1, f2, first);
  MFS_HIP(hipGetLastError());
}

}  // namespace

long long split_work_size(const LevelDesc &L) { return (long long)L.ntiles_total * 3 * 729; }
long long stokes_split_work_size(const LevelDesc &L) { return (long long)L.ntiles_total * (3 * 729 + 125); }

template <typename T>
bool launch_cell_q2(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                       int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                       T *split_work) {
  if (L.dim != 3 || L.k != 2 || L.geo) return false;  // the plane kernel assumes the uniform box
  if (op == OP_ABLOCK && split_work) {
    launch_ablock_split<T>(mode, L, src_u, eta, eta_per_qp, rhs_u, dst_u, split_work, st);
    return true;
  }
  if (op == OP_STOKES && split_work) {
    launch_stokes_split<T>(mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, split_work, st);
    return true;
  }
  return launch_cell_planes<T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
}
template <typename T>
bool launch_ablock_cheb(const LevelDesc &L, T *x, const T *b, const T *eta, int eta_per_qp, const T *dinv, T *d,
                        T f1, T f2, int first, T *split_work, hipStream_t st) {
  if (L.dim != 3 || L.k != 2 || !split_work || L.geo) return false;
  launch_ablock_cheb_split<T>(L, x, b, eta, eta_per_qp, dinv, d, f1, f2, first, split_work, st);
  return true;
}
template bool launch_ablock_cheb<double>(const LevelDesc &, double *, const double *, const double *, int,
                                         const double *, double *, double, double, int, double *, hipStream_t);
template bool launch_ablock_cheb<float>(const LevelDesc &, float *, const float *, const float *, int, const float *,
                                        float *, float, float, int, float *, hipStream_t);
template bool launch_cell_q2<double>(Op, Mode, const LevelDesc &, const double *, const double *,
                                        const double *, int, const double *, const double *, double *, double *,
                                        hipStream_t, double *);
template bool launch_cell_q2<float>(Op, Mode, const LevelDesc &, const float *, const float *, const float *,
                                       int, const float *, const float *, float *, float *, hipStream_t, float *);

}  // namespace mfs
