// gfx950 line kernel for the 3D Q2 x Q1 Stokes operator and A block on
// uniform box levels (the hot operators of the solve).
//
// Restates, per cell:
//   StokesOperator::local_apply   source/simulator/solver/matrix_free_operators.cc:376-547
//   ABlockOperator::inner_cell_operation / local_apply   :974-1052
//
// Work decomposition: one workgroup per 4x4x4 tile of cells, one thread per
// (cell, line): a line is one of the 3 x 3 x-lines of a cell's nodes or
// quadrature points, indexed (a, b) = (y index, z index).  The 64 cells of a
// tile form one wave, so (a, b) is uniform over a wave.  The x direction of
// the sum factorization stays in registers; the y and z directions go through
// LDS, one exchange array [comp][q][line][cell] per step.  FP64: three arrays
// in rotation, one barrier per step (a workgroup needs 134 KB of LDS, one per
// CU).  FP32: two arrays and one more barrier (41.5 KB, two workgroups per
// CU).  The first step interpolates in x and y at once, straight from the
// tile's nodes.  Per-thread state is one line (3 values per component), which
// keeps register use low enough for several waves per SIMD.  Collocation form:
// S (Q x N) interpolates nodal values to Gauss points, D (Q x Q)
// differentiates on the Gauss points (D S = G); tests apply D^T and S^T.
//
// Tile write-out: every tile node sums its cells from LDS in a fixed order
// (deterministic); tile colours (lowest colour stores, later colours add), or
// in the split form a work array of in-tile sums that a node-wise pass adds
// over the tiles sharing a node.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mfs_internal.h"
#include "tile_common.h"

namespace mfs {
namespace {

constexpr int LN_THREADS = 576;  // 64 cells x 9 lines
// internal mode of k_lines: every tile in one launch, in-tile node sums
// written to a work array [tile][comp][lz][ly][lx], summed by k_gather_tiles
constexpr int LN_MODE_SPLIT = 3;

template <typename T>
struct LinesBlocks {  // workgroups per CU the register budget is tuned for
  static constexpr int value = std::is_same<T, double>::value ? 1 : 2;
};

// acc + c x for a coefficient c that is a compile-time constant after
// unrolling: zero and unit entries of the 1D tables cost nothing
template <typename T>
__device__ __forceinline__ T madd(T acc, T c, T x) {
  return c == T(0) ? acc : (c == T(1) ? acc + x : acc + c * x);
}
template <typename T>
__device__ __forceinline__ T dot3(T c0, T c1, T c2, T x0, T x1, T x2) {
  return madd(madd(c0 == T(0) ? T(0) : c0 * x0, c1, x1), c2, x2);
}

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

// r / 9 for 0 <= r < 81 as a multiply and shift
__device__ __forceinline__ int div9(int r) { return (r * 57) >> 9; }
constexpr bool div9_exact() {
  for (int r = 0; r < 81; ++r)
    if (((r * 57) >> 9) != r / 9) return false;
  return true;
}
static_assert(div9_exact(), "div9 is exact on 0..80");

// Node of a 4x4x4 degree-2 tile (9^3 nodes per component) handled by thread
// tid in pass k of a sweep of the 576 threads over the 3 x 729 tile nodes:
// t = tid + 576 k = 9 row + lx with row = comp * 81 + lz * 9 + ly, so
// consecutive threads take consecutive x nodes.
struct TileNode {
  int comp, lx, ly, lz;
  bool valid;
};
__device__ __forceinline__ TileNode tile_node(int tid, int k) {
  TileNode n;
  const int row = tid / 9 + 64 * k;
  n.valid = row < 243;
  n.comp = row >= 162 ? 2 : (row >= 81 ? 1 : 0);
  const int rr = row - 81 * n.comp;
  n.lz = div9(rr);
  n.ly = rr - 9 * n.lz;
  n.lx = tid - 9 * (tid / 9);
  return n;
}

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

constexpr int LN_WPASSES = (3 * 729 + LN_THREADS - 1) / LN_THREADS;  // write-out passes per thread

// Pressure node t (of the 5^3 of a tile) in the write-out: global index,
// whether it lies in the level, whether this tile is its first writer.
struct PNode {
  int lx, ly, lz;
  long long gi;
  bool in, first;
};
__device__ __forceinline__ PNode p_node(const LevelDesc &L, int t, int tx, int ty, int tz, int px, int py, int pz) {
  constexpr int PX = 5;
  PNode o;
  o.lx = t % PX;
  o.ly = (t / PX) % PX;
  o.lz = t / (PX * PX);
  const int gx = 4 * tx + o.lx, gy = 4 * ty + o.ly, gz = 4 * tz + o.lz;
  o.in = gx < L.np[0] && gy < L.np[1] && gz < L.np[2];
  const bool shx = (o.lx == 0 && tx > 0) || (o.lx == PX - 1 && tx + 1 < L.ntiles[0]);
  const bool shy = (o.ly == 0 && ty > 0) || (o.ly == PX - 1 && ty + 1 < L.ntiles[1]);
  const bool shz = (o.lz == 0 && tz > 0) || (o.lz == PX - 1 && tz + 1 < L.ntiles[2]);
  o.first = !((shx && px) || (shy && py) || (shz && pz));
  o.gi = gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz);
  return o;
}

// Addresses of the global operands of the velocity write-out, which are
// loaded before the tile's computation so that their latency overlaps it
// (earlier colours are earlier launches, so dst is final for them).  Per
// node: op -> dst (SET, not the first writer), rhs (RESID, first writer) or
// dst (otherwise); sp -> src on constrained rows (boundary tiles).  A lane
// with nothing to load points at the tile origin (one cached line).
template <int MODE, typename T, bool INTERIOR>
__device__ __forceinline__ void lines_prefetch_addr(const LevelDesc &L, const int tid, const int tx, const int ty,
                                                    const int tz, const int px, const int py, const int pz,
                                                    const T *src_u, const T *rhs_u, const T *dst_u,
                                                    const T *(&op)[LN_WPASSES], const T *(&sp)[LN_WPASSES]) {
  const long long tb = 8 * tx + (long long)L.nv[0] * (8 * ty + (long long)L.nv[1] * (8 * tz));
#pragma unroll
  for (int k = 0; k < LN_WPASSES; ++k) {
    const TileNode n = tile_node(tid, k);
    const OutNode o = out_node<INTERIOR>(L, n, tx, ty, tz, px, py, pz);
    const bool act = n.valid && o.in;
    const long long g = tb + (act ? o.gi : 0);
    if constexpr (MODE == MODE_SET)
      op[k] = dst_u + (act && !o.first ? g : tb);
    else if constexpr (MODE == MODE_RESID)
      op[k] = (o.first ? rhs_u : dst_u) + g;
    else
      op[k] = dst_u + g;
    sp[k] = src_u + (!INTERIOR && act && o.con && o.first ? g : tb);
  }
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

// Velocity write-out of a 4x4x4 tile, degree 2: every tile node sums its
// contributing cells (tile_node_sum) and is written to dst; old/sold: the
// prefetched operands.  Colours, constrained rows and modes as in
// tile_write_u.
template <int MODE, typename T, bool INTERIOR>
__device__ __forceinline__ void lines_write_u(const LevelDesc &L, const T *xb, const int tid, const int tx,
                                              const int ty, const int tz, const int px, const int py, const int pz,
                                              const T (&old)[LN_WPASSES], const T (&sold)[LN_WPASSES],
                                              T *__restrict__ dst_u) {
  const long long tb = 8 * tx + (long long)L.nv[0] * (8 * ty + (long long)L.nv[1] * (8 * tz));
  T *const dtile = dst_u + tb;
#pragma unroll
  for (int k = 0; k < LN_WPASSES; ++k) {
    const TileNode n = tile_node(tid, k);
    const T val = tile_node_sum<T, INTERIOR>(L, xb, n, tx, ty, tz);
    const OutNode o = out_node<INTERIOR>(L, n, tx, ty, tz, px, py, pz);
    // constrained rows (first writer): SET src, RESID rhs - src, ADD dst + src
    const T base = (MODE == MODE_SET && o.first) ? (T)0 : old[k];
    T out = MODE == MODE_RESID ? base - val : base + val;
    if constexpr (!INTERIOR) {
      const T cv = MODE == MODE_SET ? sold[k] : (MODE == MODE_RESID ? old[k] - sold[k] : old[k] + sold[k]);
      out = o.con ? cv : out;
    }
    if (n.valid && o.in && !(o.con && !o.first)) dtile[o.gi] = out;
  }
}

template <int OP, int MODE, typename T>
__global__ void __launch_bounds__(LN_THREADS, LinesBlocks<T>::value)
    k_lines(const LevelDesc L, const int color, const T *__restrict__ src_u, const T *__restrict__ src_p,
            const T *__restrict__ eta, const int eta_per_qp, const T *__restrict__ rhs_u,
            const T *__restrict__ rhs_p, T *__restrict__ dst_u, T *__restrict__ dst_p) {
  using C = Cfg<3, 2>;
  using B = tab::Basis<T, 2>;
  using BD = tab::Basis<double, 2>;
  constexpr int CT = C::CT, NL = 9;
  constexpr int XA = 9 * NL * CT;  // [comp][q][line][cell]
  constexpr int XP = 2 * NL * CT;  // pressure tile, then pressure shares [i][line][cell]
  constexpr bool STOKES = (OP == OP_STOKES);
  static_assert(CT == 64 && CT * NL == LN_THREADS, "one wave per line");
  static_assert(3 * C::VN <= XA && C::PN <= XP, "tile load fits the exchange arrays");
  // FP32: two exchange arrays (one more barrier in the test steps) keep the
  // workgroup at <= 53 KB of LDS, the most two 576-thread workgroups per CU
  // get; FP64 (one workgroup per CU in any case): three arrays, one barrier
  // per step.  [XA]: zero slot of the write-out gathers.
  constexpr bool TWO = std::is_same<T, float>::value;
  __shared__ T sA[XA + 1], sB[XA + 1];
  __shared__ T sC3[TWO ? 1 : XA];
  T *const sC = TWO ? sB : sC3;  // tile nodes (FP32: in sB), FP64: also the z fluxes
  __shared__ T sP[STOKES ? XP : 1];

  constexpr bool SPLIT = (MODE == LN_MODE_SPLIT);  // all tiles in one launch, contributions to dst_u
  const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
  int tx, ty, tz;
  if constexpr (SPLIT) {
    const int bid = blockIdx.x;
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
  const int ln = __builtin_amdgcn_readfirstlane(tid / CT);
  const int a = ln % 3, b = ln / 3;
  const int cx = c % C::TX, cy = (c / C::TX) % C::TY, cz = c / (C::TX * C::TY);
  const int vx0 = 2 * C::TX * tx, vy0 = 2 * C::TY * ty, vz0 = 2 * C::TZ * tz;
  const int ox0 = C::TX * tx, oy0 = C::TY * ty, oz0 = C::TZ * tz;
  auto XI = [&](int comp, int q, int l) { return ((comp * 3 + q) * NL + l) * CT + c; };

  // ---- 0. cooperative tile load (constrained velocity DoFs read as zero);
  // interior tiles (no node on the domain boundary) skip the bounds and
  // constraint tests
  if (tid == 0) {
    sA[XA] = 0;
    sB[XA] = 0;
  }
  const bool interior = vx0 > 0 && vy0 > 0 && vz0 > 0 && vx0 + C::VX < L.nv[0] && vy0 + C::VY < L.nv[1] &&
                        vz0 + C::VZ < L.nv[2];
  // tile values, stored to LDS after the later loads have been issued
  T nd[LN_WPASSES];
  bool take[LN_WPASSES];
  T pnd = 0;
  {
    // the loads of all passes are issued before the LDS stores (a lane past
    // the last node, or at a node outside the level, reads the tile origin);
    // rows t in [3 * 729, 4 * 576) of sC receive unused values
    static_assert(C::VX == 9 && C::VY == 9 && C::VZ == 9, "tile_node assumes 9^3 nodes");
    static_assert(LN_WPASSES * LN_THREADS <= XA, "load passes fit sC");
    const long long tb = vx0 + (long long)L.nv[0] * (vy0 + (long long)L.nv[1] * vz0);
    const T *const stile = src_u + tb;
    const int sy = L.nv[0], sz = L.nv[0] * L.nv[1];
    long long gofs[LN_WPASSES];
    if (interior) {
#pragma unroll
      for (int k = 0; k < LN_WPASSES; ++k) {
        const TileNode n = tile_node(tid, k);
        gofs[k] =
            n.valid ? (n.comp == 0 ? 0 : (n.comp == 1 ? L.vnodes : 2 * L.vnodes)) + (n.lx + sy * n.ly + sz * n.lz)
                    : 0;
        take[k] = true;
      }
    } else {
#pragma unroll
      for (int k = 0; k < LN_WPASSES; ++k) {
        const TileNode n = tile_node(tid, k);
        const int gx = vx0 + n.lx, gy = vy0 + n.ly, gz = vz0 + n.lz;
        const bool in = n.valid && gx < L.nv[0] && gy < L.nv[1] && gz < L.nv[2];
        take[k] = in && (L.plain_read || !con_bf(L, n.comp, gx, gy, gz));
        gofs[k] =
            in ? (n.comp == 0 ? 0 : (n.comp == 1 ? L.vnodes : 2 * L.vnodes)) + (n.lx + sy * n.ly + sz * n.lz) : 0;
      }
    }
#pragma unroll
    for (int k = 0; k < LN_WPASSES; ++k) nd[k] = stile[gofs[k]];
  }
  bool pin = false;
  if constexpr (STOKES) {
    const int t = min(tid, C::PN - 1);
    const int lx = t % C::PX, ly = (t / C::PX) % C::PY, lz = t / (C::PX * C::PY);
    const int gx = ox0 + lx, gy = oy0 + ly, gz = oz0 + lz;
    pin = gx < L.np[0] && gy < L.np[1] && gz < L.np[2];
    pnd = src_p[pin ? gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz) : 0];
  }
  // keep the tile loads ahead of the operand loads below: memory returns in
  // issue order, so the LDS stores of the tile then wait for the tile only
  __builtin_amdgcn_sched_barrier(0);
  T wold[LN_WPASSES], wsold[LN_WPASSES];
  T pold[2] = {0, 0};  // pressure write-out operands of wave 8 (nodes tid - 512, tid - 448)
  if constexpr (!SPLIT) {
    const T *op[LN_WPASSES], *sp[LN_WPASSES];
    if (interior)
      lines_prefetch_addr<MODE, T, true>(L, tid, tx, ty, tz, px, py, pz, src_u, rhs_u, dst_u, op, sp);
    else
      lines_prefetch_addr<MODE, T, false>(L, tid, tx, ty, tz, px, py, pz, src_u, rhs_u, dst_u, op, sp);
#pragma unroll
    for (int k = 0; k < LN_WPASSES; ++k) wold[k] = *op[k];
    if (!interior) {
#pragma unroll
      for (int k = 0; k < LN_WPASSES; ++k) wsold[k] = *sp[k];
    }
    if constexpr (STOKES) {
      if (ln == 8) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int t = tid - 8 * CT + j * CT;
          const PNode o = p_node(L, min(t, C::PN - 1), tx, ty, tz, px, py, pz);
          const bool act = t < C::PN && o.in;
          if constexpr (MODE == MODE_SET) {
            pold[j] = dst_p[act && !o.first ? o.gi : 0];  // used only when not the first writer
          } else if constexpr (MODE == MODE_RESID) {
            pold[j] = (o.first ? rhs_p : dst_p)[act ? o.gi : 0];
          } else {
            pold[j] = dst_p[act ? o.gi : 0];
          }
        }
      }
    }
  }
  // viscosity at this thread's quadrature points (used in the gradient step)
  T et_q[3];
#pragma unroll
  for (int qx = 0; qx < 3; ++qx)
    et_q[qx] = eta[eta_per_qp ? (tile_id * 27 + qx + 3 * ln) * CT + c : tile_id * CT + c];
#pragma unroll
  for (int k = 0; k < LN_WPASSES; ++k) sC[tid + k * LN_THREADS] = take[k] ? nd[k] : (T)0;
  if constexpr (STOKES) {
    if (tid < C::PN) sP[tid] = pin ? pnd : (T)0;
  }
  lds_barrier();

  // wave-uniform rows and columns of the 1D tables
  T Sa[3], Sb[3], Da[3], Db[3], DTa[3], DTb[3], STa[3], STb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Sa[k] = B::S(a, k);
    Sb[k] = B::S(b, k);
    Da[k] = B::D(a, k);
    Db[k] = B::D(b, k);
    DTa[k] = B::D(k, a);
    DTb[k] = B::D(k, b);
    STa[k] = B::S(k, a);
    STb[k] = B::S(k, b);
  }

  // ---- 1. x and y interpolation: line (qy, kz) = (a, b), from the three node
  // lines (jy, kz), jy = 0..2, of the tile in LDS; pressure at the quadrature
  // line (qy, qz) = (a, b)
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    T wy[3] = {0, 0, 0};  // y-interpolated node line
#pragma unroll
    for (int jy = 0; jy < 3; ++jy) {
      const T *row = sC + comp * C::VN + ((2 * cz + b) * C::VY + (2 * cy + jy)) * C::VX + 2 * cx;
#pragma unroll
      for (int i = 0; i < 3; ++i) wy[i] += Sa[jy] * row[i];
    }
#pragma unroll
    for (int qx = 0; qx < 3; ++qx) sA[XI(comp, qx, ln)] = dot3(B::S(qx, 0), B::S(qx, 1), B::S(qx, 2), wy[0], wy[1], wy[2]);
  }
  T pq[3] = {0, 0, 0};
  if constexpr (STOKES) {
    const T SPa[2] = {a == 0 ? B::SP(0, 0) : (a == 1 ? B::SP(1, 0) : B::SP(2, 0)),
                      a == 0 ? B::SP(0, 1) : (a == 1 ? B::SP(1, 1) : B::SP(2, 1))};
    const T SPb[2] = {b == 0 ? B::SP(0, 0) : (b == 1 ? B::SP(1, 0) : B::SP(2, 0)),
                      b == 0 ? B::SP(0, 1) : (b == 1 ? B::SP(1, 1) : B::SP(2, 1))};
    T pyz[2];  // pressure nodes interpolated in y and z to (a, b), per x node
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      T acc = 0;
#pragma unroll
      for (int kz = 0; kz < 2; ++kz)
#pragma unroll
        for (int jy = 0; jy < 2; ++jy)
          acc += SPb[kz] * SPa[jy] * sP[((cz + kz) * C::PY + (cy + jy)) * C::PX + cx + i];
      pyz[i] = acc;
    }
#pragma unroll
    for (int qx = 0; qx < 3; ++qx) pq[qx] = B::SP(qx, 0) * pyz[0] + B::SP(qx, 1) * pyz[1];
  }
  lds_barrier();

  // ---- 2. z interpolation: quadrature line (qy, qz) = (a, b)
  T v[3][3];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int qx = 0; qx < 3; ++qx) {
      v[comp][qx] = Sb[0] * sA[XI(comp, qx, a)] + Sb[1] * sA[XI(comp, qx, a + 3)] + Sb[2] * sA[XI(comp, qx, a + 6)];
      sB[XI(comp, qx, ln)] = v[comp][qx];
    }
  lds_barrier();

  // ---- 3. gradients, quadrature-point operation, x test; y and z fluxes to LDS
  // Physical gradient g_ik = d_ik ih_k (d: derivative on the unit cell), stress
  // sigma = eta (g + g^T) - (s p [+ 2/3 eta div u]) I, and the fluxes that the
  // tests take, f_ik = sigma_ik JxW ih_k, folded into
  //   f_ik = c (w_kk d_ik + w_ik d_ki) (i != k),  f_ii = 2 c w_ii d_ii - P JxW ih_i
  // with c = eta JxW and w_ik = ih_i ih_k.
  using LC = LevelConst<T>;
  const T s = LC::s(L);
  const T wab = LC::vol(L) * (a == 0 ? B::W(0) : (a == 1 ? B::W(1) : B::W(2))) *
                (b == 0 ? B::W(0) : (b == 1 ? B::W(1) : B::W(2)));
  T tacc[3][3];
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int r = 0; r < 3; ++r) tacc[comp][r] = 0;
  T hp0 = 0, hp1 = 0;
  T fz[3][3];  // FP32: z fluxes, to LDS after the next barrier
#pragma unroll
  for (int qx = 0; qx < 3; ++qx) {
    const T JxW = wab * B::W(qx);
    const T cq = et_q[qx] * JxW;
    T d[3][3];  // d[comp][direction]
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      d[comp][0] = dot3(B::D(qx, 0), B::D(qx, 1), B::D(qx, 2), v[comp][0], v[comp][1], v[comp][2]);
      d[comp][1] = Da[0] * sB[XI(comp, qx, 3 * b)] + Da[1] * sB[XI(comp, qx, 3 * b + 1)] +
                   Da[2] * sB[XI(comp, qx, 3 * b + 2)];
      d[comp][2] = Db[0] * sB[XI(comp, qx, a)] + Db[1] * sB[XI(comp, qx, a + 3)] + Db[2] * sB[XI(comp, qx, a + 6)];
    }
    const T div = d[0][0] * LC::ih(L, 0) + d[1][1] * LC::ih(L, 1) + d[2][2] * LC::ih(L, 2);
    T pdiag = 0;  // P JxW
    if (L.compressible) pdiag = (T)(2.0 / 3.0) * cq * div;
    if constexpr (STOKES) pdiag += s * pq[qx] * JxW;
    T flux[3][3];
    if constexpr (std::is_same<T, float>::value) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (i == k)
            flux[i][i] = (T)2 * (cq * LC::ihw(L, i, i)) * d[i][i] - pdiag * LC::ih(L, i);
          else
            flux[i][k] = (cq * LC::ihw(L, k, k)) * d[i][k] + (cq * LC::ihw(L, i, k)) * d[k][i];
        }
    } else {
      // same fluxes with fewer live values (the FP64 kernel is register-bound)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) d[i][k] *= LC::ih(L, k);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          flux[i][k] = ((i == k ? (T)2 * cq * d[i][i] - pdiag : cq * (d[i][k] + d[k][i]))) * LC::ih(L, k);
    }
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
#pragma unroll
      for (int r = 0; r < 3; ++r) tacc[comp][r] = madd(tacc[comp][r], B::D(qx, r), flux[comp][0]);
      sA[XI(comp, qx, ln)] = flux[comp][1];
      if constexpr (TWO)
        fz[comp][qx] = flux[comp][2];
      else
        sC[XI(comp, qx, ln)] = flux[comp][2];
    }
    if constexpr (STOKES) {
      const T fp = -s * div * JxW;
      hp0 += B::SP(qx, 0) * fp;
      hp1 += B::SP(qx, 1) * fp;
    }
  }
  if constexpr (STOKES) {
    sP[ln * CT + c] = hp0;
    sP[(NL + ln) * CT + c] = hp1;
  }
  lds_barrier();

  // ---- 4. y and z tests at the quadrature line (qy, qz) = (a, b); pressure:
  // waves 0..7 integrate the x-integrated shares over y and z to the cell's
  // pressure node (i, j, k) = (ln & 1, (ln >> 1) & 1, ln >> 2)
  T pnode = 0;
  if constexpr (STOKES) {
    if (ln < 8) {
      const int pi = ln & 1, pj = (ln >> 1) & 1, pk = ln >> 2;
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        const T wy = pj ? B::SP(l % 3, 1) : B::SP(l % 3, 0), wz = pk ? B::SP(l / 3, 1) : B::SP(l / 3, 0);
        pnode += wy * wz * sP[(pi * NL + l) * CT + c];
      }
    }
  }
  if constexpr (TWO) {
    // 4a: y test from sA; the z fluxes go to sB (read last in step 3)
    T tq[3][3];
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx) {
        sB[XI(comp, qx, ln)] = fz[comp][qx];
        tq[comp][qx] = tacc[comp][qx] + DTa[0] * sA[XI(comp, qx, 3 * b)] + DTa[1] * sA[XI(comp, qx, 3 * b + 1)] +
                       DTa[2] * sA[XI(comp, qx, 3 * b + 2)];
      }
    lds_barrier();
    // 4b: z test from sB; the sums go to sA (read last in 4a)
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        sA[XI(comp, qx, ln)] = tq[comp][qx] + DTb[0] * sB[XI(comp, qx, a)] + DTb[1] * sB[XI(comp, qx, a + 3)] +
                               DTb[2] * sB[XI(comp, qx, a + 6)];
  } else {
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        sB[XI(comp, qx, ln)] = tacc[comp][qx] + DTa[0] * sA[XI(comp, qx, 3 * b)] +
                               DTa[1] * sA[XI(comp, qx, 3 * b + 1)] + DTa[2] * sA[XI(comp, qx, 3 * b + 2)] +
                               DTb[0] * sC[XI(comp, qx, a)] + DTb[1] * sC[XI(comp, qx, a + 3)] +
                               DTb[2] * sC[XI(comp, qx, a + 6)];
  }
  lds_barrier();

  // ---- 5. z integration: line (qy, kz) = (a, b); pressure node values to
  // sP[node][cell] (no thread reads sP in this step)
  if constexpr (STOKES) {
    if (ln < 8) sP[ln * CT + c] = pnode;
  }
  {
    // FP32: test sums in sA, result to sB; FP64: the other way round
    const T *const in5 = TWO ? sA : sB;
    T *const out5 = TWO ? sB : sA;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
#pragma unroll
      for (int qx = 0; qx < 3; ++qx)
        out5[XI(comp, qx, ln)] =
            STb[0] * in5[XI(comp, qx, a)] + STb[1] * in5[XI(comp, qx, a + 3)] + STb[2] * in5[XI(comp, qx, a + 6)];
  }
  lds_barrier();

  // ---- 6. y integration: node line (jy, kz) = (a, b), then x; per-cell
  // contributions [comp][kz][jy][i][cell] to xb (FP32: sA, FP64: sB)
  const T *const in6 = TWO ? sB : sA;
  T *const xb = TWO ? sA : sB;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    T hyv[3];
#pragma unroll
    for (int qx = 0; qx < 3; ++qx)
      hyv[qx] = STa[0] * in6[XI(comp, qx, 3 * b)] + STa[1] * in6[XI(comp, qx, 3 * b + 1)] +
                STa[2] * in6[XI(comp, qx, 3 * b + 2)];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T o =
          madd(madd(B::S(0, i) == T(0) ? T(0) : B::S(0, i) * hyv[0], B::S(1, i), hyv[1]), B::S(2, i), hyv[2]);
      xb[(((comp * 3 + b) * 3 + a) * 3 + i) * CT + c] = o;
    }
  }
  lds_barrier();

  if constexpr (SPLIT) {
    // in-tile node sums of all 3 x 9^3 tile nodes to the work array
    // [tile][comp][lz][ly][lx]; k_gather_tiles adds the tiles sharing a node.
    // One thread per node row (comp, lz, ly): the (y, z) cell pairs are set
    // up once per row and the x pairs are compile-time offsets, so the 12
    // contributions of each (y, z) cell row are read once for all 9 nodes.
    constexpr int ZS = 81 * CT;
    if (tid < 3 * 81) {
      const int comp = tid >= 162 ? 2 : (tid >= 81 ? 1 : 0);
      const int rr = tid - 81 * comp, lz = div9(rr), ly = rr - 9 * lz;
      const int ncx = interior ? 4 : min(4, L.n[0] - 4 * tx), ncy = interior ? 4 : min(4, L.n[1] - 4 * ty),
                ncz = interior ? 4 : min(4, L.n[2] - 4 * tz);
      int oy[2], oz[2], ox[2];
      bool wy[2], wz[2], wx[2];
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
          const int base = comp * 27 * CT + oz[za] + oy[yb];
#pragma unroll
          for (int lx = 0; lx < 9; ++lx) {
            node_cands(lx, 1, CT, ncx, ox, wx);  // compile-time offsets; ncx only masks
#pragma unroll
            for (int e = 0; e < 2; ++e) val[lx] += xb[v && wx[e] ? base + ox[e] : ZS];
          }
        }
      T *const wt = dst_u + tile_id * (3 * C::VN) + tid * 9;
#pragma unroll
      for (int lx = 0; lx < 9; ++lx) wt[lx] = val[lx];
    }
    return;
  }

  // ---- 7. tile write-out
  if (interior)
    lines_write_u<MODE, T, true>(L, xb, tid, tx, ty, tz, px, py, pz, wold, wsold, dst_u);
  else
    lines_write_u<MODE, T, false>(L, xb, tid, tx, ty, tz, px, py, pz, wold, wsold, dst_u);
  if constexpr (STOKES) {
    const int ncx = min(C::TX, L.n[0] - C::TX * tx), ncy = min(C::TY, L.n[1] - C::TY * ty),
              ncz = min(C::TZ, L.n[2] - C::TZ * tz);
    // wave 8 (its velocity write-out ends a pass early): two pressure nodes per thread
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (ln != 8) break;
      const int t = tid - 8 * CT + j * CT;
      const PNode o = p_node(L, min(t, C::PN - 1), tx, ty, tz, px, py, pz);
      // contributing cells per direction: (l - 1, local node 1) and (l, local node 0);
      // sP[node][cell] with node = i + 2 j + 4 k
      T val = 0;
#pragma unroll
      for (int e3 = 0; e3 < 2; ++e3) {
        const int ccz = o.lz - 1 + e3;
        const bool vz = ccz >= 0 && ccz < ncz;
#pragma unroll
        for (int e2 = 0; e2 < 2; ++e2) {
          const int ccy = o.ly - 1 + e2;
          const bool vy = ccy >= 0 && ccy < ncy;
#pragma unroll
          for (int e1 = 0; e1 < 2; ++e1) {
            const int ccx = o.lx - 1 + e1;
            const bool v = vz && vy && ccx >= 0 && ccx < ncx;
            const int cell = v ? ccx + C::TX * (ccy + C::TY * ccz) : 0;
            const T pv = sP[((1 - e1) + 2 * (1 - e2) + 4 * (1 - e3)) * CT + cell];
            val += v ? pv : (T)0;
          }
        }
      }
      const T base = (MODE == MODE_SET && o.first) ? (T)0 : pold[j];
      if (t < C::PN && o.in) dst_p[o.gi] = MODE == MODE_RESID ? base - val : base + val;
    }
  }
}

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
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = blockIdx.y;  // gz + nv2 * comp
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
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = blockIdx.y;
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

template <int OP, int MODE, typename T>
void launch_lines_m(const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru, const T *rp,
                    T *du, T *dp, hipStream_t st) {
  for (int color = 0; color < 8; ++color) {
    const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
    const long long hx = (L.ntiles[0] - px + 1) / 2, hy = (L.ntiles[1] - py + 1) / 2, hz = (L.ntiles[2] - pz + 1) / 2;
    const long long nb = hx * hy * hz;
    if (nb <= 0) continue;
    hipLaunchKernelGGL((k_lines<OP, MODE, T>), dim3((unsigned)nb), dim3(LN_THREADS), 0, st, L, color, su, sp, eta,
                       pq, ru, rp, du, dp);
  }
  MFS_HIP(hipGetLastError());
}

// split A-block apply: per-cell contributions of all tiles in one launch, then the node-wise sum
template <typename T>
void launch_ablock_split(Mode mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                         T *work, hipStream_t st) {
  hipLaunchKernelGGL((k_lines<OP_ABLOCK, LN_MODE_SPLIT, T>), dim3((unsigned)L.ntiles_total), dim3(LN_THREADS), 0, st,
                     L, 0, su, (const T *)nullptr, eta, pq, (const T *)nullptr, (const T *)nullptr, work, (T *)nullptr);
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
  hipLaunchKernelGGL((k_lines<OP_ABLOCK, LN_MODE_SPLIT, T>), dim3((unsigned)L.ntiles_total), dim3(LN_THREADS), 0, st,
                     L, 0, (const T *)x, (const T *)nullptr, eta, pq, (const T *)nullptr, (const T *)nullptr, work,
                     (T *)nullptr);
  if (split_work_size(L) >= (1LL << 31) || (long long)L.nv[0] * L.nv[1] >= (1LL << 31))
    throw std::runtime_error("level too large for the split A-block apply (32-bit work offsets)");
  const dim3 grid((unsigned)(((long long)L.nv[0] * L.nv[1] + 511) / 512), (unsigned)(3 * L.nv[2]));
  hipLaunchKernelGGL((k_gather_cheb<T>), grid, dim3(512), 0, st, L, (const T *)work, x, b, dinv, d, f1, f2, first);
  MFS_HIP(hipGetLastError());
}

template <int OP, typename T>
void launch_lines_o(Mode mode, const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru,
                    const T *rp, T *du, T *dp, hipStream_t st) {
  if (mode == MODE_SET)
    launch_lines_m<OP, MODE_SET, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);
  else if (mode == MODE_RESID)
    launch_lines_m<OP, MODE_RESID, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);
  else
    launch_lines_m<OP, MODE_ADD, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);
}

}  // namespace

long long split_work_size(const LevelDesc &L) { return (long long)L.ntiles_total * 3 * 729; }

template <typename T>
bool launch_cell_lines(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                       int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                       T *split_work) {
  if (L.dim != 3 || L.k != 2) return false;
  if (op == OP_ABLOCK && split_work) {
    launch_ablock_split<T>(mode, L, src_u, eta, eta_per_qp, rhs_u, dst_u, split_work, st);
    return true;
  }
  if (op == OP_STOKES)
    launch_lines_o<OP_STOKES, T>(mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else if (op == OP_ABLOCK)
    launch_lines_o<OP_ABLOCK, T>(mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else
    return false;
  return true;
}
template <typename T>
bool launch_ablock_cheb(const LevelDesc &L, T *x, const T *b, const T *eta, int eta_per_qp, const T *dinv, T *d,
                        T f1, T f2, int first, T *split_work, hipStream_t st) {
  if (L.dim != 3 || L.k != 2 || !split_work) return false;
  launch_ablock_cheb_split<T>(L, x, b, eta, eta_per_qp, dinv, d, f1, f2, first, split_work, st);
  return true;
}
template bool launch_ablock_cheb<double>(const LevelDesc &, double *, const double *, const double *, int,
                                         const double *, double *, double, double, int, double *, hipStream_t);
template bool launch_ablock_cheb<float>(const LevelDesc &, float *, const float *, const float *, int, const float *,
                                        float *, float, float, int, float *, hipStream_t);
template bool launch_cell_lines<double>(Op, Mode, const LevelDesc &, const double *, const double *,
                                        const double *, int, const double *, const double *, double *, double *,
                                        hipStream_t, double *);
template bool launch_cell_lines<float>(Op, Mode, const LevelDesc &, const float *, const float *, const float *,
                                       int, const float *, const float *, float *, float *, hipStream_t, float *);

}  // namespace mfs
