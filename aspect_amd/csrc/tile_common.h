// Tile configuration and helpers shared by the cell kernels (Q_k x Q_{k-1}
// on uniform box levels).
#pragma once
#include <hip/hip_runtime.h>

#include "basis_tables.h"
#include "mfs_internal.h"

namespace mfs {

// ------------------------------------------------------------------ configuration
template <int DIM, int K>
struct Tile;
template <>
struct Tile<3, 2> {
  static constexpr int TX = 4, TY = 4, TZ = 4;
};
template <>
struct Tile<3, 3> {
  static constexpr int TX = 4, TY = 4, TZ = 2;
};
template <>
struct Tile<2, 2> {
  static constexpr int TX = 8, TY = 8, TZ = 1;
};
template <>
struct Tile<2, 3> {
  static constexpr int TX = 8, TY = 8, TZ = 1;
};

template <int DIM, int K>
struct Cfg {
  static constexpr int N = K + 1, Q = K + 1, NP = K;
  static constexpr int ZN = DIM == 3 ? N : 1, ZQ = DIM == 3 ? Q : 1, ZNP = DIM == 3 ? NP : 1;
  static constexpr int TX = Tile<DIM, K>::TX, TY = Tile<DIM, K>::TY, TZ = Tile<DIM, K>::TZ;
  static constexpr int CT = TX * TY * TZ;
  static constexpr int THREADS = CT * Q;
  static constexpr int VX = K * TX + 1, VY = K * TY + 1, VZ = DIM == 3 ? K * TZ + 1 : 1;
  static constexpr int PX = (K - 1) * TX + 1, PY = (K - 1) * TY + 1, PZ = DIM == 3 ? (K - 1) * TZ + 1 : 1;
  static constexpr int VN = VX * VY * VZ, PN = PX * PY * PZ;
  static constexpr int NQ = ZQ * Q * Q, NPL = ZNP * NP * NP;
  static constexpr int QS = ZQ * Q;  // quadrature points of one y-slice
  static constexpr int XBUF = DIM * ZQ * Q * Q * CT;
  static constexpr int PBUF = DIM * ZN * N * N * CT + NPL * Q * CT;
  static constexpr int NBUF = DIM * VN + PN;
  static constexpr int SMEM = (XBUF > NBUF ? (XBUF > PBUF ? XBUF : PBUF) : (NBUF > PBUF ? NBUF : PBUF));
};

__device__ __forceinline__ bool is_con(const LevelDesc &L, int comp, int gx, int gy, int gz) {
  const int g[3] = {gx, gy, gz};
  for (int d = 0; d < L.dim; ++d) {
    int b;
    if (g[d] == 0)
      b = L.bc[2 * d];
    else if (g[d] == L.nv[d] - 1)
      b = L.bc[2 * d + 1];
    else
      continue;
    if (b == 2 || (b == 1 && comp == d)) return true;
  }
  return false;
}


// Velocity part of the tile write-out.  xb holds the per-cell nodal
// contributions, layout [comp][kz][jy][i][cell].  Every tile node sums its
// contributing cells in a fixed order (deterministic) and is written out:
// the lowest tile colour touching a node stores, later colours add.
// Constrained rows follow the identity semantics of the operator.
// MODE_TSPLIT: the in-tile sums go to the work array dst_u[tile][comp][node]
// instead (constraints applied by the node-wise pass).
template <int DIM, int K, int OP, int MODE, typename T, int NTH>
__device__ __forceinline__ void tile_write_u(const LevelDesc &L, const T *xb, const int tid, const int tx,
                                             const int ty, const int tz, const int px, const int py, const int pz,
                                             const T *__restrict__ src_u, const T *__restrict__ rhs_u,
                                             T *__restrict__ dst_u) {
  using C = Cfg<DIM, K>;
  constexpr int N = C::N, ZN = C::ZN, CT = C::CT;
  const int vx0 = K * C::TX * tx, vy0 = K * C::TY * ty, vz0 = DIM == 3 ? K * C::TZ * tz : 0;
  for (int t = tid; t < DIM * C::VN; t += NTH) {
    const int comp = t / C::VN, r = t - comp * C::VN;
    const int lx = r % C::VX, ly = (r / C::VX) % C::VY, lz = r / (C::VX * C::VY);
    const int gx = vx0 + lx, gy = vy0 + ly, gz = vz0 + lz;
    if (gx >= L.nv[0] || gy >= L.nv[1] || gz >= L.nv[2]) continue;
    // contributing (cell, local index) per direction: at most two
    int ccx[2], iix[2], ncx = 0, ccy[2], iiy[2], ncy = 0, ccz[2], iiz[2], ncz = 0;
    if (lx % K == 0 && lx > 0) { ccx[ncx] = lx / K - 1; iix[ncx++] = K; }
    if (lx / K < C::TX) { ccx[ncx] = lx / K; iix[ncx++] = lx % K; }
    if (ly % K == 0 && ly > 0) { ccy[ncy] = ly / K - 1; iiy[ncy++] = K; }
    if (ly / K < C::TY) { ccy[ncy] = ly / K; iiy[ncy++] = ly % K; }
    if (DIM == 3) {
      if (lz % K == 0 && lz > 0) { ccz[ncz] = lz / K - 1; iiz[ncz++] = K; }
      if (lz / K < C::TZ) { ccz[ncz] = lz / K; iiz[ncz++] = lz % K; }
    } else {
      ccz[0] = 0;
      iiz[0] = 0;
      ncz = 1;
    }
    T val = 0;
    for (int a = 0; a < ncz; ++a) {
      if (tz * C::TZ + ccz[a] >= (DIM == 3 ? L.n[2] : 1)) continue;
      for (int bb = 0; bb < ncy; ++bb) {
        if (ty * C::TY + ccy[bb] >= L.n[1]) continue;
        for (int e = 0; e < ncx; ++e) {
          if (tx * C::TX + ccx[e] >= L.n[0]) continue;
          const int cell = ccx[e] + C::TX * (ccy[bb] + C::TY * ccz[a]);
          val += xb[(((comp * ZN + iiz[a]) * N + iiy[bb]) * N + iix[e]) * CT + cell];
        }
      }
    }
    if constexpr (MODE == MODE_TSPLIT) {  // in-tile sum to the work array (dst_u)
      dst_u[(tx + (long long)L.ntiles[0] * (ty + (long long)L.ntiles[1] * tz)) * (DIM * C::VN) + t] = val;
      continue;
    }
    const bool shx = (lx == 0 && tx > 0) || (lx == C::VX - 1 && tx + 1 < L.ntiles[0]);
    const bool shy = (ly == 0 && ty > 0) || (ly == C::VY - 1 && ty + 1 < L.ntiles[1]);
    const bool shz = DIM == 3 && ((lz == 0 && tz > 0) || (lz == C::VZ - 1 && tz + 1 < L.ntiles[2]));
    const bool first = !((shx && px) || (shy && py) || (shz && pz));
    const long long gi = comp * L.vnodes + gx + (long long)L.nv[0] * (gy + (long long)L.nv[1] * gz);
    T vv;
    if (is_con(L, comp, gx, gy, gz)) {
      if (!first) continue;
      if constexpr (OP == OP_ADIAG || OP == OP_LDIAG)
        vv = 1;
      else if constexpr (OP == OP_BT || OP == OP_RHS)
        vv = 0;
      else
        vv = src_u[gi];
    } else {
      vv = val;
    }
    if constexpr (MODE == MODE_SET)
      dst_u[gi] = first ? vv : dst_u[gi] + vv;
    else if constexpr (MODE == MODE_RESID)
      dst_u[gi] = first ? rhs_u[gi] - vv : dst_u[gi] - vv;
    else
      dst_u[gi] += vv;
  }
}

// ------------------------------------------------------------------ 3D degree-2 tiles (4x4x4 cells, 9^3 nodes)
// constrained velocity DoF test without branches (faces: 0 open, 1 no normal
// flux, 2 zero velocity)
__device__ __forceinline__ bool con_bf(const LevelDesc &L, int comp, int gx, int gy, int gz) {
  const int g[3] = {gx, gy, gz};
  bool con = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int b = g[d] == 0 ? L.bc[2 * d] : (g[d] == L.nv[d] - 1 ? L.bc[2 * d + 1] : 0);
    con = con || b == 2 || (b == 1 && comp == d);
  }
  return con;
}

// Workgroup barrier for LDS exchanges only: the fences order LDS accesses
// (wait for this wave's LDS operations), not global ones, so global loads
// issued ahead of the barrier (prefetches for later steps) stay in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// XCD-aware block order.  The hardware deals the workgroups of a launch
// round-robin to the 8 XCDs, each with its own L2; block b of nb is renumbered
// so that XCD x runs the contiguous range [x q + min(x, r), (x + 1) q +
// min(x + 1, r)) of nb = 8 q + r (a bijection of [0, nb)).  Blocks that touch
// neighbouring data (tiles sharing the cache lines of their common nodes,
// node rows sharing lines of a work array) then mostly meet in one L2.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned nb) {
  const unsigned q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}

// Compact index of a tile-boundary node (at least one local index 0 or 8) of
// a 4x4x4 degree-2 tile, per component: the two z faces [lz][ly][lx] (162),
// then the two y faces without their z-face rows [ly][lz - 1][lx] (126), then
// the two x faces without the other faces' nodes [lx][lz - 1][ly - 1] (98):
// 386 values.  Rows along x stay contiguous on the z and y faces.
constexpr int kTileBnd = 386;
__device__ __forceinline__ int tile_bnd_index(int lx, int ly, int lz) {
  if (lz == 0 || lz == 8) return (lz == 8 ? 81 : 0) + 9 * ly + lx;
  if (ly == 0 || ly == 8) return 162 + (ly == 8 ? 63 : 0) + 9 * (lz - 1) + lx;
  return 288 + (lx == 8 ? 49 : 0) + 7 * (lz - 1) + (ly - 1);
}

// Node of a 4x4x4 degree-2 tile: component and local indices (9 per direction).
struct TileNode {
  int comp, lx, ly, lz;
  bool valid;
};

// Contributing cells of tile node l along one direction (local cell stride
// sc, local-node stride si in the per-cell contribution layout): candidate 0
// is (l/2 - 1, local node 2), present for even l > 0; candidate 1 is
// (l/2, local node l mod 2), present while l/2 < nc.
__device__ __forceinline__ void node_cands(int l, int sc, int si, int nc, int off[2], bool ok[2]) {
  off[0] = 2 * si + ((l >> 1) - 1) * sc;
  ok[0] = (l & 1) == 0 && l > 0;
  off[1] = (l & 1) * si + (l >> 1) * sc;
  ok[1] = (l >> 1) < nc;
}

// Output node of the write-out sweep: offset from the tile origin, whether
// it lies in the level, whether this tile is the first (lowest colour) to
// write it, whether it is a constrained row.  INTERIOR: the tile holds no
// node of the domain boundary (no constrained node, no partial cell range).
struct OutNode {
  long long gi;
  bool in, first, con;
};
template <bool INTERIOR>
__device__ __forceinline__ OutNode out_node(const LevelDesc &L, const TileNode &n, int tx, int ty, int tz, int px,
                                            int py, int pz) {
  constexpr int VX = 9;
  OutNode o;
  const int gx = 8 * tx + n.lx, gy = 8 * ty + n.ly, gz = 8 * tz + n.lz;
  o.in = INTERIOR || (gx < L.nv[0] && gy < L.nv[1] && gz < L.nv[2]);
  const bool shx = INTERIOR ? (n.lx == 0 || n.lx == VX - 1)
                            : ((n.lx == 0 && tx > 0) || (n.lx == VX - 1 && tx + 1 < L.ntiles[0]));
  const bool shy = INTERIOR ? (n.ly == 0 || n.ly == VX - 1)
                            : ((n.ly == 0 && ty > 0) || (n.ly == VX - 1 && ty + 1 < L.ntiles[1]));
  const bool shz = INTERIOR ? (n.lz == 0 || n.lz == VX - 1)
                            : ((n.lz == 0 && tz > 0) || (n.lz == VX - 1 && tz + 1 < L.ntiles[2]));
  o.first = !((shx && px) || (shy && py) || (shz && pz));
  const int sy = L.nv[0], sz = L.nv[0] * L.nv[1];  // node strides (a tile spans < 2^31 entries)
  o.gi = (n.comp == 0 ? 0 : (n.comp == 1 ? L.vnodes : 2 * L.vnodes)) + (n.lx + sy * n.ly + sz * n.lz);
  o.con = INTERIOR ? false : (o.in && con_bf(L, n.comp, gx, gy, gz));
  return o;
}

// Sum over the at most 8 cells of a 4x4x4 degree-2 tile that contribute to
// tile node n, in a fixed order.  The 8 candidate reads are unconditional: a
// missing candidate (or a thread past the last node) reads xb[ZS], a slot
// that holds zero, so all lanes run the same instruction stream.  xb:
// per-cell contributions [comp][kz][jy][i][cell] followed by the zero slot.
template <typename T, bool INTERIOR>
__device__ __forceinline__ T tile_node_sum(const LevelDesc &L, const T *xb, const TileNode &n, const int tx,
                                           const int ty, const int tz) {
  constexpr int CT = 64, ZS = 81 * CT;
  const int ncx = INTERIOR ? 4 : min(4, L.n[0] - 4 * tx), ncy = INTERIOR ? 4 : min(4, L.n[1] - 4 * ty),
            ncz = INTERIOR ? 4 : min(4, L.n[2] - 4 * tz);
  int ox[2], oy[2], oz[2];
  bool wx[2], wy[2], wz[2];
  node_cands(n.lx, 1, CT, ncx, ox, wx);
  node_cands(n.ly, 4, 3 * CT, ncy, oy, wy);
  node_cands(n.lz, 16, 9 * CT, ncz, oz, wz);
  const int cb = n.comp * 27 * CT;
  T val = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 2; ++e) val += xb[n.valid && wz[a] && wy[b] && wx[e] ? cb + oz[a] + oy[b] + ox[e] : ZS];
  return val;
}


}  // namespace mfs
