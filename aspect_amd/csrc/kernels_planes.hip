// gfx950 plane kernel for the 3D Q2 x Q1 Stokes operator and A block on
// uniform box levels.
//
// Restates, per cell:
//   StokesOperator::local_apply   source/simulator/solver/matrix_free_operators.cc:376-547
//   ABlockOperator::inner_cell_operation / local_apply   :974-1052
//
// Work decomposition: one workgroup per 4x4x4 tile of cells, one thread per
// (cell, plane): plane b is the cell's node plane kz = b on input and output
// and its quadrature plane qz = b in between.  The 64 cells of a tile form one
// wave, so b is uniform over a wave (3 waves per workgroup).  The x and y
// directions of the sum factorization stay in the thread's registers (a 3 x 3
// plane per component); only the z direction goes through LDS:
//   1. the thread reads its 3 x 9 nodes from global memory, interpolates them
//      in x and y and stores the plane [comp][qy qx] to LDS;
//   2. it reads the three planes of its cell and forms the values at its
//      quadrature plane (S) and their z derivatives (G = D S) at once;
//   3. x and y derivatives (D) in registers, the quadrature-point operation,
//      and the x and y tests (D^T) into registers; the z fluxes go to LDS;
//   4. the z test of the fluxes (G^T) and the z integration (S^T) of the x/y
//      test sums, both from the three planes of the cell;
//   5. S^T in y and x in registers: the cell's node plane kz = b, stored to
//      LDS as per-cell contributions [comp][kz][jy][i][cell];
//   6. tile write-out: every tile node sums its at most 8 cells in a fixed
//      order (deterministic), in tile colours (lowest colour stores, later
//      colours add), or in the split form into a work array of in-tile sums.
// Per tile this is about 350 LDS reads and 1.5k FMA-class instructions per
// thread for 192 threads; the exchange array is [comp][k][plane][cell]
// (41.5 KB in FP64), two workgroups per CU in FP64.
//
// Collocation form: S (Q x N) interpolates nodal values to Gauss points, D
// (Q x Q) differentiates on the Gauss points (D S = G); tests apply D^T, G^T
// and S^T.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mfs_internal.h"
#include "tile_common.h"

namespace mfs {
namespace {

constexpr int PL_THREADS = 192;  // 64 cells x 3 planes
// internal split modes (every tile in one launch).  PL_SPLIT + m (m =
// MODE_SET, MODE_RESID, MODE_ADD): the tile-interior nodes (no other tile
// touches them) are finished with mode m, the in-tile sums of the
// tile-boundary nodes go to the compact work array [tile][comp][386]
// (tile_bnd_index).  PL_SPLIT_FULL: every in-tile node sum goes to the work
// array [tile][comp][lz][ly][lx] (the Chebyshev-fused A-block step, whose
// node-wise pass streams every DoF anyway).
constexpr int PL_SPLIT = 3;
constexpr int PL_SPLIT_FULL = 6;

template <typename T>
struct PlanesBlocks {  // workgroups per CU the register budget is tuned for
  static constexpr int value = std::is_same<T, double>::value ? 2 : 4;
};

// LDS barrier that also fences the instruction scheduler: without the
// sched_barriers the compiler sinks the arithmetic on values read before the
// barrier to after it and keeps all of the reads live across it (81 FP64
// values in the z steps, which spilled to scratch)
__device__ __forceinline__ void plane_barrier() {
  __builtin_amdgcn_sched_barrier(0);
  lds_barrier();
  __builtin_amdgcn_sched_barrier(0);
}

// Pins a value to a register at this point of the program: the compiler may
// not defer its computation past a later barrier (an empty asm statement
// with the value as an in/out operand; no instruction is emitted)
template <typename T>
__device__ __forceinline__ void pin(T &x) {
  asm volatile("" : "+v"(x));
}

// c0 x0 + c1 x1 + c2 x2 with compile-time coefficients (zero terms vanish)
template <typename T>
__device__ __forceinline__ T dot3c(T c0, T c1, T c2, T x0, T x1, T x2) {
  T r = c0 == T(0) ? T(0) : c0 * x0;
  r = c1 == T(0) ? r : (c1 == T(1) ? r + x1 : r + c1 * x1);
  r = c2 == T(0) ? r : (c2 == T(1) ? r + x2 : r + c2 * x2);
  return r;
}

// Write-out row r of a tile (243 rows (comp, lz, ly) of 9 nodes along x):
// the 147 rows with 1 <= ly, lz <= 7 first (comp, lz, ly lexicographic), so
// that they are the first rows of threads 0..146, then the 96 rows on a z or
// y face (per comp: lz = 0 and lz = 8 with every ly, then ly = 0 and ly = 8
// with 1 <= lz <= 7)
__device__ __forceinline__ void wrow(int r, int &comp, int &lz, int &ly) {
  if (r < 147) {
    comp = r / 49;
    const int q = r - 49 * comp;
    lz = 1 + q / 7;
    ly = 1 + q % 7;
  } else {
    const int f = r - 147;
    comp = f / 32;
    const int q = f - 32 * comp;
    if (q < 18) {
      lz = q < 9 ? 0 : 8;
      ly = q < 9 ? q : q - 9;
    } else {
      lz = 1 + ((q - 18) >> 1);
      ly = (q & 1) ? 8 : 0;
    }
  }
}

// Compact split modes in FP32: the operand of the 7 tile-interior nodes lx =
// 1..7 of this thread's first write-out row (rows 0..146 are the interior
// rows, wrow), rhs (RESID) or dst (ADD), is loaded at the start of the kernel
// so that its latency overlaps the tile's arithmetic.  In FP64 the registers
// would spill: the write-out loads it.
template <int MODE, typename T>
struct PlanePrefetch {
  static constexpr int N =
      std::is_same<T, float>::value && (MODE == PL_SPLIT + MODE_RESID || MODE == PL_SPLIT + MODE_ADD) ? 1 : 0;
};

// Tile write-out by node rows.  Row r = tid + 192 pass (r < 243, wrow): comp,
// lz, ly; the 9 nodes lx of the row sum their contributing cells from the
// per-cell contributions xb [comp][kz][jy][i][cell] (zero slot at 81 * 64):
// per (z, y) candidate pair (node_cands) the cell row's 4 x 3 values go to
// node lx = 2 cx + i.  Colour modes: the lowest colour touching a node
// stores, later colours add; constrained rows take src (SET), rhs - src
// (RESID), dst + src (ADD) from their first writer.  Compact split modes:
// rows on a z or y face, and the two x-face nodes of the other rows, go to the
// compact work array; the 7 tile-interior nodes of those rows are finished
// here with the mode (constrained DoFs, only at the upper end of a partial
// tile, as above).  Full split: every row to the full work array.
template <int MODE, typename T, bool INTERIOR>
__device__ __forceinline__ void planes_write_rows(const LevelDesc &L, const T *xb, const int tid, const int tx,
                                                  const int ty, const int tz, const int px, const int py,
                                                  const int pz, const long long tile_id, const T *__restrict__ src_u,
                                                  const T *__restrict__ rhs_u, T *__restrict__ dst_u,
                                                  T *__restrict__ work_u, const T (&pf)[1][7]) {
  constexpr int NPF = PlanePrefetch<MODE, T>::N;
  constexpr int CT = 64, ZS = 81 * CT;
  constexpr bool SPLIT = (MODE >= PL_SPLIT && MODE < PL_SPLIT_FULL);
  constexpr bool FULL = (MODE == PL_SPLIT_FULL);
  constexpr int IM = MODE - PL_SPLIT;  // compact split: mode of the tile-interior nodes
  const int ncx = INTERIOR ? 4 : min(4, L.n[0] - 4 * tx), ncy = INTERIOR ? 4 : min(4, L.n[1] - 4 * ty),
            ncz = INTERIOR ? 4 : min(4, L.n[2] - 4 * tz);
  const int sy = L.nv[0], sz = L.nv[0] * L.nv[1];
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int row = tid + PL_THREADS * pass;
    if (row >= 243) break;
    int comp, lz, ly;
    wrow(row, comp, lz, ly);
    int oy[2], oz[2];
    bool wy[2], wz[2];
    node_cands(ly, 4, 3 * CT, ncy, oy, wy);
    node_cands(lz, 16, 9 * CT, ncz, oz, wz);
    T val[9];
#pragma unroll
    for (int lx = 0; lx < 9; ++lx) val[lx] = 0;
#pragma unroll
    for (int za = 0; za < 2; ++za)
#pragma unroll
      for (int yb = 0; yb < 2; ++yb) {
        const bool v = wz[za] && wy[yb];
        if (INTERIOR) {
          // a missing (z, y) pair reads valid slots with weight 0 (no address selects)
          const T f = v ? (T)1 : (T)0;
          const int base = comp * 27 * CT + (v ? oz[za] + oy[yb] : 0);
#pragma unroll
          for (int ccx = 0; ccx < 4; ++ccx)
#pragma unroll
            for (int i = 0; i < 3; ++i) val[2 * ccx + i] += f * xb[base + i * CT + ccx];
        } else {
          const int base = comp * 27 * CT + oz[za] + oy[yb];
#pragma unroll
          for (int ccx = 0; ccx < 4; ++ccx)
#pragma unroll
            for (int i = 0; i < 3; ++i) val[2 * ccx + i] += xb[v && ccx < ncx ? base + i * CT + ccx : ZS];
        }
      }
    if constexpr (FULL) {
      T *const wt = work_u + tile_id * (3 * 729) + ((comp * 9 + lz) * 9 + ly) * 9;
#pragma unroll
      for (int lx = 0; lx < 9; ++lx) wt[lx] = val[lx];
    } else if constexpr (SPLIT) {
      T *const wt = work_u + tile_id * (3 * kTileBnd) + comp * kTileBnd;
      if (lz == 0 || lz == 8 || ly == 0 || ly == 8) {
        const int i0 = tile_bnd_index(0, ly, lz);
#pragma unroll
        for (int lx = 0; lx < 9; ++lx) wt[i0 + lx] = val[lx];
      } else {  // an interior row (only in pass 0: row = tid < 147, wrow)
        const int i0 = tile_bnd_index(0, ly, lz);
        wt[i0] = val[0];
        wt[i0 + 49] = val[8];
        // tile-interior nodes lx = 1..7: no other tile touches them
        const int gy = 8 * ty + ly, gz = 8 * tz + lz;
        const bool rowin = INTERIOR || (gy < L.nv[1] && gz < L.nv[2]);
        const long long rb =
            (comp == 0 ? 0 : (comp == 1 ? L.vnodes : 2 * L.vnodes)) + 8 * tx + ((long long)sy * gy + (long long)sz * gz);
#pragma unroll
        for (int lx = 1; lx < 8; ++lx) {
          const int gx = 8 * tx + lx;
          if (!INTERIOR && !(rowin && gx < L.nv[0])) continue;
          const long long gi = rb + lx;
          const bool con = !INTERIOR && con_bf(L, comp, gx, gy, gz);
          const T add = con ? src_u[gi] : val[lx];
          if constexpr (IM == MODE_SET)
            dst_u[gi] = add;
          else if constexpr (IM == MODE_RESID)
            dst_u[gi] = (NPF ? pf[0][lx - 1] : rhs_u[gi]) - add;
          else
            dst_u[gi] = (NPF ? pf[0][lx - 1] : dst_u[gi]) + add;
        }
      }
    } else {
      const int gy = 8 * ty + ly, gz = 8 * tz + lz;
      const bool shy = INTERIOR ? (ly == 0 || ly == 8) : ((ly == 0 && ty > 0) || (ly == 8 && ty + 1 < L.ntiles[1]));
      const bool shz = INTERIOR ? (lz == 0 || lz == 8) : ((lz == 0 && tz > 0) || (lz == 8 && tz + 1 < L.ntiles[2]));
      const bool rowin = INTERIOR || (gy < L.nv[1] && gz < L.nv[2]);
      const bool firstyz = !((shy && py) || (shz && pz));
      T *const drow = dst_u + (comp == 0 ? 0 : (comp == 1 ? L.vnodes : 2 * L.vnodes)) + 8 * tx +
                      ((long long)sy * gy + (long long)sz * gz);
      const T *const rrow = rhs_u + (drow - dst_u);
      const T *const srow = src_u + (drow - dst_u);
#pragma unroll
      for (int lx = 0; lx < 9; ++lx) {
        const int gx = 8 * tx + lx;
        const bool in = rowin && (INTERIOR || gx < L.nv[0]);
        const bool shx = INTERIOR ? (lx == 0 || lx == 8) : ((lx == 0 && tx > 0) || (lx == 8 && tx + 1 < L.ntiles[0]));
        const bool first = firstyz && !(shx && px);
        // operands through selected pointers (one load each; a node with
        // nothing to read loads the row's first node, a cached line)
        const int lxs = in ? lx : 0;
        T base;  // SET: 0 or dst; RESID: rhs or dst; ADD: dst
        if constexpr (MODE == MODE_SET) {
          const T old = drow[first ? 0 : lxs];
          base = first ? (T)0 : old;
        } else if constexpr (MODE == MODE_RESID) {
          base = (first ? rrow : (const T *)drow)[lxs];
        } else {
          base = drow[lxs];
        }
        if (INTERIOR) {
          drow[lx] = MODE == MODE_RESID ? base - val[lx] : base + val[lx];
        } else {
          if (!in) continue;
          const bool con = con_bf(L, comp, gx, gy, gz);
          if (con && !first) continue;
          // constrained rows (first writer): SET src, RESID rhs - src, ADD dst + src
          const T add = con ? srow[lx] : val[lx];
          drow[lx] = MODE == MODE_RESID ? base - add : base + add;
        }
      }
    }
  }
}

template <int OP, int MODE, typename T>
__global__ void __launch_bounds__(PL_THREADS, PlanesBlocks<T>::value)
    k_planes(const LevelDesc L, const int color, const T *__restrict__ src_u, const T *__restrict__ src_p,
             const T *__restrict__ eta, const int eta_per_qp, const T *__restrict__ rhs_u,
             const T *__restrict__ rhs_p, T *__restrict__ dst_u, T *__restrict__ dst_p, T *__restrict__ work_u,
             T *__restrict__ work_p) {
  using B = tab::Basis<T, 2>;
  constexpr int CT = 64;
  constexpr bool STOKES = (OP == OP_STOKES);
  constexpr bool SPLIT = (MODE >= PL_SPLIT);  // compact or full split: every tile in one launch
  constexpr int IM = MODE - PL_SPLIT;
  constexpr int XN = 3 * 9 * 3 * CT;  // [comp][k][plane][cell]; [XN]: zero slot of the write-out gathers
  constexpr int WN = 9 * 3 * CT;      // one component's planes [k][plane][cell]
  __shared__ T sX[XN + 1];            // z fluxes, then x/y test sums, then per-cell node contributions
  __shared__ T sW[2][WN];             // x/y-interpolated node planes, one component at a time
  T *const sQ = sW[0];                // pressure test shares [node][plane][cell] (after the last component)

  const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
  int tx, ty, tz;
  if constexpr (SPLIT) {
    // XCD-aware tile order: each XCD runs a contiguous range of tiles
    const int bid = (int)xcd_order(blockIdx.x, gridDim.x);
    tx = bid % L.ntiles[0];
    ty = (bid / L.ntiles[0]) % L.ntiles[1];
    tz = bid / (L.ntiles[0] * L.ntiles[1]);
  } else {
    const int hx = (L.ntiles[0] - px + 1) >> 1, hy = (L.ntiles[1] - py + 1) >> 1;
    int bid = blockIdx.x;
    tx = 2 * (bid % hx) + px;
    bid /= hx;
    ty = 2 * (bid % hy) + py;
    bid /= hy;
    tz = 2 * bid + pz;
  }
  const long long tile_id = tx + (long long)L.ntiles[0] * (ty + (long long)L.ntiles[1] * tz);
  const int tid = threadIdx.x;
  const int c = tid & (CT - 1);
  const int b = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cx = c & 3, cy = (c >> 2) & 3, cz = c >> 4;
  const int vx0 = 8 * tx, vy0 = 8 * ty, vz0 = 8 * tz;
  const bool interior =
      vx0 > 0 && vy0 > 0 && vz0 > 0 && vx0 + 9 < L.nv[0] && vy0 + 9 < L.nv[1] && vz0 + 9 < L.nv[2];
  auto XI = [&](int comp, int k, int pl) { return ((comp * 9 + k) * 3 + pl) * CT + c; };
  auto WI = [&](int k, int pl) { return (k * 3 + pl) * CT + c; };
  const int sy = L.nv[0], sz = L.nv[0] * L.nv[1];  // node strides (a tile spans < 2^31 entries)
  const long long tb = vx0 + (long long)sy * vy0 + (long long)sz * vz0;

  // ---- 1. this thread's nodes (2cx + i, 2cy + j, 2cz + b), constrained
  // velocity DoFs read as zero (plain_read: as stored)
  T u[3][9];
  {
    const T *const stile = src_u + tb;
    const int lz = 2 * cz + b;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const long long cofs = comp == 0 ? 0 : (comp == 1 ? L.vnodes : 2 * L.vnodes);
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int lx = 2 * cx + i, ly = 2 * cy + j;
          bool take = true, in = true;
          if (!interior) {
            const int gx = vx0 + lx, gy = vy0 + ly, gz = vz0 + lz;
            in = gx < L.nv[0] && gy < L.nv[1] && gz < L.nv[2];
            take = in && (L.plain_read || !con_bf(L, comp, gx, gy, gz));
          }
          const T val = stile[in ? cofs + (lx + sy * ly + sz * lz) : 0];
          u[comp][3 * j + i] = take ? val : (T)0;
        }
    }
  }
  T pn[8];  // pressure nodes of the cell (i, j, k) -> i + 2 j + 4 k
  if constexpr (STOKES) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int gx = 4 * tx + cx + i, gy = 4 * ty + cy + j, gz = 4 * tz + cz + k;
          const bool in = gx < L.np[0] && gy < L.np[1] && gz < L.np[2];
          pn[i + 2 * j + 4 * k] = src_p[in ? gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz) : 0];
        }
  }
  // viscosity at this thread's quadrature points (qx, qy, qz = b)
  T et[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) et[k] = eta[eta_per_qp ? (tile_id * 27 + k + 9 * b) * CT + c : tile_id * CT + c];
  // compact split modes (FP32): the operand of the tile-interior nodes of
  // this thread's first write-out row (PlanePrefetch)
  constexpr int NPF = PlanePrefetch<MODE, T>::N;
  T pf[1][7];
  if constexpr (NPF > 0) {
    int pcomp = 0, plz = 1, ply = 1;
    wrow(tid, pcomp, plz, ply);
    const int gy = vy0 + ply, gz = vz0 + plz;
    const bool rowok = tid < 147 && (interior || (gy < L.nv[1] && gz < L.nv[2]));
    const long long rb = (pcomp == 0 ? 0 : (pcomp == 1 ? L.vnodes : 2 * L.vnodes)) + vx0 +
                         ((long long)sy * gy + (long long)sz * gz);
#pragma unroll
    for (int lx = 1; lx < 8; ++lx) {
      const bool ok = rowok && (interior || vx0 + lx < L.nv[0]);
      const long long gi = ok ? rb + lx : 0;
      pf[0][lx - 1] = IM == MODE_RESID ? rhs_u[gi] : dst_u[gi];
    }
  }

  // wave-uniform coefficients: rows and columns of the 1D tables for plane b,
  // and the differentiation matrix with the inverse mesh sizes folded in
  // (A0 = D / h_x, A1 = D / h_y; z: G / h_z), so that gradients and tests
  // need no separate scaling
  using LC = LevelConst<T>;
  T Sb[3], Gz[3], STb[3], GzT[3], SPb[2], A0[3][3], A1[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Sb[k] = B::S(b, k);
    Gz[k] = B::G(b, k) * LC::ih(L, 2);
    STb[k] = B::S(k, b);
    GzT[k] = B::G(k, b) * LC::ih(L, 2);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      A0[k][r] = B::D(k, r) * LC::ih(L, 0);
      A1[k][r] = B::D(k, r) * LC::ih(L, 1);
    }
  }
  SPb[0] = B::SP(b, 0);
  SPb[1] = B::SP(b, 1);

  // x and y interpolation of one component's node plane, to sW[buf]
  auto interp_xy = [&](const T (&uc)[9], T *w) {
    T ax[3][3];  // [j][qx]
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        ax[j][qx] = dot3c(B::S(qx, 0), B::S(qx, 1), B::S(qx, 2), uc[3 * j], uc[3 * j + 1], uc[3 * j + 2]);
#pragma unroll
    for (int qy = 0; qy < 3; ++qy)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        w[WI(3 * qy + qx, b)] = dot3c(B::S(qy, 0), B::S(qy, 1), B::S(qy, 2), ax[0][qx], ax[1][qx], ax[2][qx]);
  };
  // pressure nodes interpolated in z to the plane b: [j][i]
  T pzv[4] = {0, 0, 0, 0};
  if constexpr (STOKES) {
#pragma unroll
    for (int n = 0; n < 4; ++n) pzv[n] = SPb[0] * pn[n] + SPb[1] * pn[n + 4];
  }

  // ---- 2./3. per velocity component cc (the operator is linear in u): values
  // and z derivatives at the quadrature plane b from the three node planes,
  // x and y derivatives in registers, and the contributions of grad u_cc to
  // the fluxes sigma_ik JxW (sigma = eta (g + g^T), g_ik = d u_i / d x_k),
  // tested at once: x and y tests (D^T / h) into the registers t, z fluxes
  // into sX (own slots, scaled by 1/h_z in step 4).  div u accumulates for
  // the pressure and compressibility terms.
  const T s = LC::s(L);
  const T wb = LC::vol(L) * (b == 0 ? B::W(0) : (b == 1 ? B::W(1) : B::W(2)));
  T t[3][9];  // x and y test sums at the quadrature plane
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k) t[comp][k] = 0;
  T dv[9];  // div u at the quadrature points
#pragma unroll
  for (int k = 0; k < 9; ++k) dv[k] = 0;
  // test of flux value f of output component i in direction k at (qx, qy);
  // z fluxes: the first contribution to a slot stores
  auto test = [&](int i, int k, int qx, int qy, T f, bool first_z) {
    if (k == 0) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
        if (B::D(qx, r) != T(0)) t[i][3 * qy + r] += A0[qx][r] * f;
    } else if (k == 1) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
        if (B::D(qy, r) != T(0)) t[i][3 * r + qx] += A1[qy][r] * f;
    } else {
      T &slot = sX[XI(i, 3 * qy + qx, b)];
      slot = first_z ? f : slot + f;
    }
  };
  auto dsum = [&](const T (&a)[3][3], int q, T x0, T x1, T x2) {
    T r = a[q][0] * x0;
    if (B::D(q, 1) != T(0)) r += a[q][1] * x1;
    return r + a[q][2] * x2;
  };
  interp_xy(u[0], sW[0]);
  plane_barrier();
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    T v[9], dz[9];
    const T *const w = sW[cc & 1];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const T w0 = w[WI(k, 0)], w1 = w[WI(k, 1)], w2 = w[WI(k, 2)];
      v[k] = Sb[0] * w0 + Sb[1] * w1 + Sb[2] * w2;
      dz[k] = Gz[0] * w0 + Gz[1] * w1 + Gz[2] * w2;
    }
    // next component's planes into the other buffer (last read before the previous barrier)
    if (cc < 2) interp_xy(u[cc + 1], sW[(cc + 1) & 1]);
#pragma unroll
    for (int qy = 0; qy < 3; ++qy)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx) {
        const int q = 3 * qy + qx;
        const T cq = et[q] * (wb * (B::W(qy) * B::W(qx)));
        T g[3];
        g[0] = dsum(A0, qx, v[3 * qy], v[3 * qy + 1], v[3 * qy + 2]);
        g[1] = dsum(A1, qy, v[qx], v[3 + qx], v[6 + qx]);
        g[2] = dz[q];
        dv[q] += g[cc];
        // sigma_{cc,k} += eta g_k (twice for k = cc), sigma_{k,cc} += eta g_k (k != cc)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const T ck = cq * g[k];
          test(cc, k, qx, qy, k == cc ? ck + ck : ck, k == 2);
          if (k != cc) test(k, cc, qx, qy, ck, false);
        }
      }
    if (cc < 2) plane_barrier();
  }
  // pressure and compressibility: sigma_ii -= s p + 2/3 eta div u; pressure test
  T hq[4] = {0, 0, 0, 0};  // pressure test shares (i, j) of the plane, before the z weights
  if (STOKES || L.compressible) {
#pragma unroll
    for (int qy = 0; qy < 3; ++qy)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx) {
        const int q = 3 * qy + qx;
        const T JxW = wb * (B::W(qy) * B::W(qx));
        T pd = L.compressible ? (T)(2.0 / 3.0) * et[q] * JxW * dv[q] : (T)0;
        if constexpr (STOKES) {
          const T y0 = B::SP(qy, 0) * pzv[0] + B::SP(qy, 1) * pzv[2];
          const T y1 = B::SP(qy, 0) * pzv[1] + B::SP(qy, 1) * pzv[3];
          const T pq = B::SP(qx, 0) * y0 + B::SP(qx, 1) * y1;
          pd += s * JxW * pq;
          const T fp = -s * JxW * dv[q];
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) hq[i + 2 * j] += (B::SP(qx, i) * B::SP(qy, j)) * fp;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) test(i, i, qx, qy, -pd, false);
      }
  }
  plane_barrier();  // every thread is done with sW
  if constexpr (STOKES) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int n = 0; n < 4; ++n) sQ[((n + 4 * k) * 3 + b) * CT + c] = SPb[k] * hq[n];
  }

  // ---- 4. z test of the fluxes (G^T / h_z), then z integration of the x/y test sums (S^T)
  T o[3][9];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k)
      o[comp][k] = GzT[0] * sX[XI(comp, k, 0)] + GzT[1] * sX[XI(comp, k, 1)] + GzT[2] * sX[XI(comp, k, 2)];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k) pin(o[comp][k]);
  plane_barrier();
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k) sX[XI(comp, k, b)] = t[comp][k];
  plane_barrier();
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k)
      o[comp][k] += STb[0] * sX[XI(comp, k, 0)] + STb[1] * sX[XI(comp, k, 1)] + STb[2] * sX[XI(comp, k, 2)];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int k = 0; k < 9; ++k) pin(o[comp][k]);
  plane_barrier();

  // ---- 5. y and x integration (S^T): node plane kz = b of the cell; per-cell
  // contributions [comp][kz][jy][i][cell]
  if (tid == 0) sX[XN] = 0;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    T hy[3][3];  // [jy][qx]
#pragma unroll
    for (int jy = 0; jy < 3; ++jy)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        hy[jy][qx] = dot3c(B::S(0, jy), B::S(1, jy), B::S(2, jy), o[comp][qx], o[comp][3 + qx], o[comp][6 + qx]);
#pragma unroll
    for (int jy = 0; jy < 3; ++jy)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        sX[(((comp * 3 + b) * 3 + jy) * 3 + i) * CT + c] =
            dot3c(B::S(0, i), B::S(1, i), B::S(2, i), hy[jy][0], hy[jy][1], hy[jy][2]);
  }
  plane_barrier();

  // ---- 6. tile write-out: one thread per tile node row (comp, lz, ly), 243
  // rows in two passes; the row's 9 nodes sum their cells in a fixed order:
  // per (z, y) candidate cell pair the 12 contributions of the cell row (4
  // cells x 3 local x nodes) are read once, each into its node.  Pressure:
  // the threads without a second row (64..188) take one tile pressure node.
  if (interior)
    planes_write_rows<MODE, T, true>(L, sX, tid, tx, ty, tz, px, py, pz, tile_id, src_u, rhs_u, dst_u, work_u, pf);
  else
    planes_write_rows<MODE, T, false>(L, sX, tid, tx, ty, tz, px, py, pz, tile_id, src_u, rhs_u, dst_u, work_u, pf);
  if constexpr (STOKES) {
    if (tid >= 64 && tid < 64 + 125) {
      const int t = tid - 64;
      const int plx = t % 5, ply = (t / 5) % 5, plz = t / 25;
      const int gx = 4 * tx + plx, gy = 4 * ty + ply, gz = 4 * tz + plz;
      const bool in = gx < L.np[0] && gy < L.np[1] && gz < L.np[2];
      const bool shx = (plx == 0 && tx > 0) || (plx == 4 && tx + 1 < L.ntiles[0]);
      const bool shy = (ply == 0 && ty > 0) || (ply == 4 && ty + 1 < L.ntiles[1]);
      const bool shz = (plz == 0 && tz > 0) || (plz == 4 && tz + 1 < L.ntiles[2]);
      const bool first = !((shx && px) || (shy && py) || (shz && pz));
      const long long gi = in ? gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz) : 0;
      T pold = 0;
      if constexpr (SPLIT)
        (void)IM;
      else if constexpr (MODE == MODE_SET)
        pold = first ? (T)0 : dst_p[gi];
      else if constexpr (MODE == MODE_RESID)
        pold = (first ? rhs_p : dst_p)[gi];
      else
        pold = dst_p[gi];
      const int ncx = min(4, L.n[0] - 4 * tx), ncy = min(4, L.n[1] - 4 * ty), ncz = min(4, L.n[2] - 4 * tz);
      T val = 0;
#pragma unroll
      for (int e3 = 0; e3 < 2; ++e3) {
        const int ccz = plz - 1 + e3;
        const bool vz = ccz >= 0 && ccz < ncz;
#pragma unroll
        for (int e2 = 0; e2 < 2; ++e2) {
          const int ccy = ply - 1 + e2;
          const bool vy = ccy >= 0 && ccy < ncy;
#pragma unroll
          for (int e1 = 0; e1 < 2; ++e1) {
            const int ccx = plx - 1 + e1;
            const bool vv = vz && vy && ccx >= 0 && ccx < ncx;
            const int cell = vv ? ccx + 4 * (ccy + 4 * ccz) : 0;
            const int node = (1 - e1) + 2 * (1 - e2) + 4 * (1 - e3);
            const T pv = sQ[(node * 3 + 0) * CT + cell] + sQ[(node * 3 + 1) * CT + cell] + sQ[(node * 3 + 2) * CT + cell];
            val += vv ? pv : (T)0;
          }
        }
      }
      if constexpr (SPLIT)
        work_p[tile_id * 125 + t] = val;  // in-tile sums [tile][plz][ply][plx]
      else if (in)
        dst_p[gi] = MODE == MODE_RESID ? pold - val : pold + val;
    }
  }
}

template <int OP, int MODE, typename T>
void launch_planes_m(const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru, const T *rp,
                     T *du, T *dp, hipStream_t st) {
  for (int color = 0; color < 8; ++color) {
    const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
    const long long hx = (L.ntiles[0] - px + 1) / 2, hy = (L.ntiles[1] - py + 1) / 2, hz = (L.ntiles[2] - pz + 1) / 2;
    const long long nb = hx * hy * hz;
    if (nb <= 0) continue;
    hipLaunchKernelGGL((k_planes<OP, MODE, T>), dim3((unsigned)nb), dim3(PL_THREADS), 0, st, L, color, su, sp, eta,
                       pq, ru, rp, du, dp, (T *)nullptr, (T *)nullptr);
  }
  MFS_HIP(hipGetLastError());
}

}  // namespace

// First pass of the split A-block apply with the plane kernel, every tile in
// one launch: mode MODE_SET / RESID / ADD (compact split: tile-interior nodes
// finished, tile-boundary in-tile sums to work[tile][comp][386]) or
// SPLIT_FULL_SUMS (all in-tile sums to work[tile][comp][lz][ly][lx])
template <typename T>
void launch_planes_split(int mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                         T *work, hipStream_t st) {
  const dim3 grid((unsigned)L.ntiles_total), block(PL_THREADS);
  const T *np_ = nullptr;
  T *nw = nullptr;
  switch (mode) {
    case MODE_SET:
      hipLaunchKernelGGL((k_planes<OP_ABLOCK, PL_SPLIT + MODE_SET, T>), grid, block, 0, st, L, 0, su, np_, eta, pq,
                         ru, np_, du, nw, work, nw);
      break;
    case MODE_RESID:
      hipLaunchKernelGGL((k_planes<OP_ABLOCK, PL_SPLIT + MODE_RESID, T>), grid, block, 0, st, L, 0, su, np_, eta, pq,
                         ru, np_, du, nw, work, nw);
      break;
    case MODE_ADD:
      hipLaunchKernelGGL((k_planes<OP_ABLOCK, PL_SPLIT + MODE_ADD, T>), grid, block, 0, st, L, 0, su, np_, eta, pq,
                         ru, np_, du, nw, work, nw);
      break;
    case SPLIT_FULL_SUMS:
      hipLaunchKernelGGL((k_planes<OP_ABLOCK, PL_SPLIT_FULL, T>), grid, block, 0, st, L, 0, su, np_, eta, pq, np_,
                         np_, nw, nw, work, nw);
      break;
    default:
      throw std::runtime_error("launch_planes_split: unknown mode");
  }
  MFS_HIP(hipGetLastError());
}
template void launch_planes_split<double>(int, const LevelDesc &, const double *, const double *, int,
                                          const double *, double *, double *, hipStream_t);
template void launch_planes_split<float>(int, const LevelDesc &, const float *, const float *, int, const float *,
                                         float *, float *, hipStream_t);

// The same for the FP64 Stokes operator (compact split): velocity as above,
// pressure in-tile sums to work_p[tile][plz][ply][plx] (5^3 per tile, all nodes)
void launch_planes_split_stokes(int mode, const LevelDesc &L, const double *su, const double *sp,
                                const double *eta, int pq, const double *ru, const double *rp, double *du,
                                double *work_u, double *work_p, hipStream_t st) {
  const dim3 grid((unsigned)L.ntiles_total), block(PL_THREADS);
  double *nd = nullptr;
  if (mode == MODE_SET)
    hipLaunchKernelGGL((k_planes<OP_STOKES, PL_SPLIT + MODE_SET, double>), grid, block, 0, st, L, 0, su, sp, eta, pq,
                       ru, rp, du, nd, work_u, work_p);
  else if (mode == MODE_RESID)
    hipLaunchKernelGGL((k_planes<OP_STOKES, PL_SPLIT + MODE_RESID, double>), grid, block, 0, st, L, 0, su, sp, eta,
                       pq, ru, rp, du, nd, work_u, work_p);
  else
    hipLaunchKernelGGL((k_planes<OP_STOKES, PL_SPLIT + MODE_ADD, double>), grid, block, 0, st, L, 0, su, sp, eta, pq,
                       ru, rp, du, nd, work_u, work_p);
  MFS_HIP(hipGetLastError());
}

template <typename T>
bool launch_cell_planes(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                        int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st) {
  if (L.dim != 3 || L.k != 2 || (op != OP_STOKES && op != OP_ABLOCK)) return false;
  if (9LL * L.nv[0] * L.nv[1] >= (1LL << 31)) return false;  // 32-bit offsets within a tile
  auto run = [&](auto opc) {
    constexpr int OPC = decltype(opc)::value;
    if (mode == MODE_SET)
      launch_planes_m<OPC, MODE_SET, T>(L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
    else if (mode == MODE_RESID)
      launch_planes_m<OPC, MODE_RESID, T>(L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
    else
      launch_planes_m<OPC, MODE_ADD, T>(L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  };
  if (op == OP_STOKES)
    run(std::integral_constant<int, OP_STOKES>());
  else
    run(std::integral_constant<int, OP_ABLOCK>());
  return true;
}
template bool launch_cell_planes<double>(Op, Mode, const LevelDesc &, const double *, const double *,
                                         const double *, int, const double *, const double *, double *, double *,
                                         hipStream_t);
template bool launch_cell_planes<float>(Op, Mode, const LevelDesc &, const float *, const float *, const float *,
                                        int, const float *, const float *, float *, float *, hipStream_t);

}  // namespace mfs
