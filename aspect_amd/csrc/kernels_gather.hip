// gfx950 node-wise passes of the split applies on 3D degree-2 box levels
// (the FP64 Stokes operator of the active level, the A block of every V-cycle
// level) and the dispatch of the 3D Q2 x Q1 Stokes and A-block applies to the
// plane kernel (kernels_planes.hip).
//
// Split form: one launch of the plane kernel over all tiles (no tile
// colours).  Applies (SET / RESID / ADD): the tile-interior nodes are
// finished by that launch and the in-tile sums of the tile-boundary nodes go
// to a compact work array; a pass over the tile-boundary nodes sums the at
// most 8 tiles sharing each in a fixed order (deterministic) and applies the
// mode.  Chebyshev-fused A-block step: every in-tile sum goes to the work
// array and a pass over all DoFs sums the tiles and applies one
// Jacobi-Chebyshev step of the smoother (that pass streams every DoF of four
// vectors anyway, so it also takes the tile-interior sums).
//
// Restates ABlockOperator::local_apply / apply_add
// (source/simulator/solver/matrix_free_operators.cc:1029-1064) and
// StokesOperator::local_apply (:376-547) for the sums over cells, and
// deal.II's PreconditionChebyshev step for the fused form.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mfs_internal.h"
#include "tile_common.h"

namespace mfs {
namespace {

// Sum over the at most 8 tiles sharing tile-boundary velocity node g of
// component comp, from the compact work array [tile][comp][386] of in-tile
// sums (tile_bnd_index), in a fixed order (deterministic).  Per direction the
// candidates of node g are (tile g/8 - 1, local node 8) for g mod 8 == 0,
// g > 0, and (tile g/8, local node g mod 8) while g/8 < ntiles.  32-bit
// offsets (checked by the launcher).
template <typename T>
__device__ __forceinline__ T tile_gather_sum(const LevelDesc &L, const T *__restrict__ work, int comp,
                                             const int (&g)[3]) {
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
        const int wi = v ? (tile * 3 + comp) * kTileBnd + tile_bnd_index(ll[0][sx], ll[1][sy], ll[2][sz]) : 0;
        const T c = work[wi];
        val += v ? c : (T)0;
      }
  return val;
}

// Tile-boundary nodes (some coordinate a multiple of 8) of a level, in two
// launches.  Full planes (sparse = 0): gz = 8 k, every (gx, gy) of the plane.
// Other planes (sparse = 1): gz = the k-th value not a multiple of 8
// (8 (k / 7) + k mod 7 + 1); first the full rows gy = 8 r (nyr nv0 nodes,
// x fastest), then the nodes gx = 8 c of the other rows (gy fastest).
// Grid: x over the nodes of a plane, y over (comp, plane); the blocks run in
// XCD-aware order so that node rows sharing lines of the work array meet in
// one L2.  Returns false for threads past the plane's nodes.
__device__ __forceinline__ bool bnd_node(const LevelDesc &L, const int sparse, int &comp, int (&g)[3]) {
  const unsigned m = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned by = m / gridDim.x, bx = m - by * gridDim.x;
  const unsigned i = bx * blockDim.x + threadIdx.x;
  const int nzp = (L.nv[2] - 1) / 8 + 1;
  const int planes = sparse ? L.nv[2] - nzp : nzp;
  comp = (int)by / planes;
  const int k = (int)by - comp * planes;
  g[2] = sparse ? 8 * (k / 7) + k % 7 + 1 : 8 * k;
  if (!sparse) {
    if (i >= (unsigned)(L.nv[0] * L.nv[1])) return false;
    g[1] = (int)(i / (unsigned)L.nv[0]);
    g[0] = (int)i - g[1] * L.nv[0];
    return true;
  }
  const int nyr = (L.nv[1] - 1) / 8 + 1, nxc = (L.nv[0] - 1) / 8 + 1;
  const unsigned nrow = (unsigned)(nyr * L.nv[0]), ncol = (unsigned)(nxc * (L.nv[1] - nyr));
  if (i >= nrow + ncol) return false;
  if (i < nrow) {
    const int r = (int)(i / (unsigned)L.nv[0]);
    g[1] = 8 * r;
    g[0] = (int)i - r * L.nv[0];
  } else {  // along y first: the x-face entries of a tile are contiguous in ly
    const int ny = L.nv[1] - nyr, j = (int)(i - nrow), c = j / ny, mm = j - c * ny;
    g[0] = 8 * c;
    g[1] = 8 * (mm / 7) + mm % 7 + 1;
  }
  return true;
}
// launch shape of bnd_node: (blocks along x, blocks along y) for 256 threads
inline dim3 bnd_grid(const LevelDesc &L, int sparse) {
  const int nzp = (L.nv[2] - 1) / 8 + 1;
  if (!sparse) return dim3((unsigned)(((long long)L.nv[0] * L.nv[1] + 255) / 256), (unsigned)(3 * nzp));
  const int nyr = (L.nv[1] - 1) / 8 + 1, nxc = (L.nv[0] - 1) / 8 + 1;
  const long long n = (long long)nyr * L.nv[0] + (long long)nxc * (L.nv[1] - nyr);
  return dim3((unsigned)((n + 255) / 256), (unsigned)(3 * (L.nv[2] - nzp)));
}

// Second pass of the split A-block (and Stokes velocity) apply: one thread
// per tile-boundary velocity DoF adds the tiles sharing the node and applies
// the mode; constrained rows take src (identity).
template <int MODE, typename T>
__global__ void k_gather_tiles(const LevelDesc L, const int sparse, const T *__restrict__ work,
                               const T *__restrict__ src_u, const T *__restrict__ rhs_u, T *__restrict__ dst_u) {
  int comp, g[3];
  if (!bnd_node(L, sparse, comp, g)) return;
  const T val = tile_gather_sum<T>(L, work, comp, g);
  const long long gi = comp * L.vnodes + g[0] + (long long)L.nv[0] * (g[1] + (long long)L.nv[1] * g[2]);
  const T vv = con_bf(L, comp, g[0], g[1], g[2]) ? src_u[gi] : val;
  if constexpr (MODE == MODE_SET)
    dst_u[gi] = vv;
  else if constexpr (MODE == MODE_RESID)
    dst_u[gi] = rhs_u[gi] - vv;
  else
    dst_u[gi] += vv;
}

template <typename T>
void launch_gather(int mode, const LevelDesc &L, const T *work, const T *su, const T *ru, T *du, hipStream_t st) {
  for (int sparse = 0; sparse < 2; ++sparse) {
    const dim3 grid = bnd_grid(L, sparse);
    if (grid.x == 0 || grid.y == 0) continue;
    if (mode == MODE_SET)
      hipLaunchKernelGGL((k_gather_tiles<MODE_SET, T>), grid, dim3(256), 0, st, L, sparse, work, su, ru, du);
    else if (mode == MODE_RESID)
      hipLaunchKernelGGL((k_gather_tiles<MODE_RESID, T>), grid, dim3(256), 0, st, L, sparse, work, su, ru, du);
    else
      hipLaunchKernelGGL((k_gather_tiles<MODE_ADD, T>), grid, dim3(256), 0, st, L, sparse, work, su, ru, du);
  }
  MFS_HIP(hipGetLastError());
}

// Sum over the at most 8 tiles sharing velocity node g of component comp from
// the full work array [tile][comp][lz][ly][lx] (the Chebyshev-fused step),
// candidates and order as in tile_gather_sum
template <typename T>
__device__ __forceinline__ T tile_gather_sum_full(const LevelDesc &L, const T *__restrict__ work, int comp,
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

// Second pass of the Chebyshev-fused split A-block step: one thread per
// velocity DoF sums the tiles sharing the node, r = b - A x (not stored;
// constrained rows r = b - x), then d = dinv r f2 (first step of a sweep from
// a nonzero x) or d = f1 d + f2 dinv r, and x += d -- the arithmetic of
// v_cheb_first / v_cheb_step, so the result is that of the unfused sequence.
// Grid: x over the nodes of an xy plane, y over (comp, z), XCD-aware order.
template <typename T>
__global__ void k_gather_cheb(const LevelDesc L, const T *__restrict__ work, const ChebArgs<T> ch) {
  const unsigned plane = (unsigned)(L.nv[0] * L.nv[1]);
  const unsigned m = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned by = m / gridDim.x, bx = m - by * gridDim.x;
  const unsigned pidx = bx * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = (int)by;
  const int comp = row / L.nv[2], gz = row - comp * L.nv[2];
  const int gy = (int)(pidx / (unsigned)L.nv[0]), gx = (int)pidx - gy * L.nv[0];
  const int g[3] = {gx, gy, gz};
  const T val = tile_gather_sum_full<T>(L, work, comp, g);
  const long long gi = comp * L.vnodes + pidx + (long long)plane * gz;
  const T xi = ch.x[gi];
  const T r = ch.b[gi] - (con_bf(L, comp, gx, gy, gz) ? xi : val);
  const T dd = ch.first ? ch.dinv[gi] * r * ch.f2 : ch.f1 * ch.d[gi] + ch.f2 * ch.dinv[gi] * r;
  ch.d[gi] = dd;
  ch.x[gi] = xi + dd;
}

// Work offsets and node indices of the split applies are 32-bit
inline void check_split_offsets(const LevelDesc &L) {
  if ((long long)L.ntiles_total * (3 * 729) >= (1LL << 31) || (long long)L.nv[0] * L.nv[1] >= (1LL << 31) ||
      (long long)L.np[0] * L.np[1] >= (1LL << 31))
    throw std::runtime_error("level too large for the split apply (32-bit work offsets)");
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

// split Stokes apply (FP64): one launch of the plane kernel over all tiles,
// then the node-wise sums over the tiles of the tile-boundary velocity nodes
// and of every pressure node
void launch_stokes_split(Mode mode, const LevelDesc &L, const double *su, const double *sp, const double *eta,
                         int pq, const double *ru, const double *rp, double *du, double *dp, double *work,
                         hipStream_t st) {
  check_split_offsets(L);
  double *const work_p = work + split_apply_work_size(L);
  launch_planes_split_stokes(mode, L, su, sp, eta, pq, ru, rp, du, work, work_p, st);
  launch_gather<double>(mode, L, work, su, ru, du, st);
  const dim3 gridp((unsigned)(((long long)L.np[0] * L.np[1] + 255) / 256), (unsigned)L.np[2]);
  if (mode == MODE_SET)
    hipLaunchKernelGGL((k_gather_p<MODE_SET, double>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  else if (mode == MODE_RESID)
    hipLaunchKernelGGL((k_gather_p<MODE_RESID, double>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  else
    hipLaunchKernelGGL((k_gather_p<MODE_ADD, double>), gridp, dim3(256), 0, st, L, work_p, rp, dp);
  MFS_HIP(hipGetLastError());
}

// split A-block apply: all tiles in one launch (tile-interior nodes finished
// there), then the node-wise sum over the tile-boundary nodes
template <typename T>
void launch_ablock_split(Mode mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                         T *work, hipStream_t st) {
  check_split_offsets(L);
  launch_planes_split<T>(mode, L, su, eta, pq, ru, du, work, st);
  launch_gather<T>(mode, L, work, su, ru, du, st);
}

template <typename T>
void launch_ablock_cheb_split(const LevelDesc &L, T *x, const T *b, const T *eta, int pq, const T *dinv, T *d, T f1,
                              T f2, int first, T *work, hipStream_t st) {
  check_split_offsets(L);
  launch_planes_split<T>(SPLIT_FULL_SUMS, L, (const T *)x, eta, pq, nullptr, nullptr, work, st);
  const ChebArgs<T> ch{x, b, dinv, d, f1, f2, first};
  const dim3 grid((unsigned)(((long long)L.nv[0] * L.nv[1] + 511) / 512), (unsigned)(3 * L.nv[2]));
  hipLaunchKernelGGL((k_gather_cheb<T>), grid, dim3(512), 0, st, L, (const T *)work, ch);
  MFS_HIP(hipGetLastError());
}

}  // namespace

// work arrays of the split A-block applies (0: the level has none): 3D
// degree 2 full (Chebyshev-fused step) and compact (applies), 2D degree 2
// full in-tile sums (2 x 17^2 per tile, both uses); none on a mapped level
long long split_work_size(const LevelDesc &L) {
  if (L.k != 2 || L.geo) return 0;
  return L.dim == 3 ? (long long)L.ntiles_total * 3 * 729 : (long long)L.ntiles_total * 2 * 289;
}
long long split_apply_work_size(const LevelDesc &L) {
  if (L.k != 2 || L.geo) return 0;
  return L.dim == 3 ? (long long)L.ntiles_total * 3 * kTileBnd : (long long)L.ntiles_total * 2 * 289;
}
long long stokes_split_work_size(const LevelDesc &L) { return (long long)L.ntiles_total * (3 * kTileBnd + 125); }

template <typename T>
bool launch_cell_q2(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                       int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                       T *split_work) {
  if (L.dim != 3 || L.k != 2 || L.geo) return false;  // the plane kernel assumes the uniform box
  if (op == OP_ABLOCK && split_work) {
    launch_ablock_split<T>(mode, L, src_u, eta, eta_per_qp, rhs_u, dst_u, split_work, st);
    return true;
  }
  if constexpr (std::is_same<T, double>::value) {
    if (op == OP_STOKES && split_work) {
      launch_stokes_split(mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, split_work, st);
      return true;
    }
  }
  return launch_cell_planes<T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
}
template <typename T>
bool launch_ablock_cheb(const LevelDesc &L, T *x, const T *b, const T *eta, int eta_per_qp, const T *dinv, T *d,
                        T f1, T f2, int first, T *split_work, hipStream_t st) {
  if (L.k != 2 || !split_work || L.geo) return false;
  if (L.dim == 2) {
    const ChebArgs<T> ch{x, b, dinv, d, f1, f2, first};
    launch_ablock_split2<T>(SPLIT_CHEB, L, (const T *)x, eta, eta_per_qp, b, x, split_work, ch, st);
    return true;
  }
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
