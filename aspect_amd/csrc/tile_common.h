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
    const bool shx = (lx == 0 && tx > 0) || (lx == C::VX - 1 && tx + 1 < L.ntiles[0]);
    const bool shy = (ly == 0 && ty > 0) || (ly == C::VY - 1 && ty + 1 < L.ntiles[1]);
    const bool shz = DIM == 3 && ((lz == 0 && tz > 0) || (lz == C::VZ - 1 && tz + 1 < L.ntiles[2]));
    const bool first = !((shx && px) || (shy && py) || (shz && pz));
    const long long gi = comp * L.vnodes + gx + (long long)L.nv[0] * (gy + (long long)L.nv[1] * gz);
    T vv;
    if (is_con(L, comp, gx, gy, gz)) {
      if (!first) continue;
      if constexpr (OP == OP_ADIAG)
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

}  // namespace mfs
