// gfx950 kernels for the matrix-free Q_k x Q_{k-1} Stokes operators on uniform
// box levels, the level transfers, viscosity tables and the vector kernels of
// the Chebyshev / Krylov iterations.
//
// Operator kernels restate, per cell:
//   StokesOperator::local_apply   source/simulator/solver/matrix_free_operators.cc:376-547
//   BTBlockOperator::local_apply  :668-701
//   MassMatrixOperator::inner_cell_operation  :808-845
//   ABlockOperator::inner_cell_operation      :974-1009
//   compute_diagonal of A and of the mass     :1068-1098, :861-908
//
// Work decomposition: one workgroup per tile of cells (Tile<dim,k>), one thread
// per (cell, y-slice).  A thread keeps the x and z directions of its slice in
// registers; the y contraction of the sum factorization goes through LDS.
// Contributions of the cells of a tile are assembled in LDS in 2^dim cell-parity
// phases (cells of one parity share no node), so the tile sum is
// deterministic.  Tiles are launched in 2^dim colours (tile parity); tiles of
// one colour share no node, so a tile writes every node of its box without
// atomics: the lowest colour touching a node stores it, later colours add.

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "mfs_internal.h"
#include "tile_common.h"

namespace mfs {

// ------------------------------------------------------------------ 1D tables
struct BasisConst {
  double S[4][4];   // velocity basis value  [q][node]
  double G[4][4];   // velocity basis derivative on [0,1]
  double SP[4][4];  // pressure basis value
  double S2[4][4], G2[4][4], SP2[4][4];  // squares (diagonal kernels)
  double D[4][4];   // collocation derivative on the Gauss points: D[q][r] = l_r'(x_q)
  double w[4];      // Gauss weights on [0,1]
  double x[4];      // Gauss points on [0,1]
  double W[7][4];   // prolongation weights, velocity degree k:  [r][m], r < 2k+1
  double WP[5][3];  // prolongation weights, pressure degree k-1
};
__constant__ BasisConst c_basis[2];  // index k-2

// the cell kernel's tables in its own arithmetic type (no conversions at use)
template <typename T>
struct BasisT {
  T S[4][4], G[4][4], D[4][4], SP[4][4], S2[4][4], G2[4][4], SP2[4][4], w[4];
};
__constant__ BasisT<double> c_tabd[2];
__constant__ BasisT<float> c_tabf[2];
template <typename T, int K>
__device__ __forceinline__ const BasisT<T> &tb() {
  if constexpr (std::is_same<T, double>::value)
    return c_tabd[K - 2];
  else
    return c_tabf[K - 2];
}

namespace {
double lag(const std::vector<double> &pts, int i, double x) {
  double v = 1.0;
  for (size_t j = 0; j < pts.size(); ++j)
    if ((int)j != i) v *= (x - pts[j]) / (pts[i] - pts[j]);
  return v;
}
double lagd(const std::vector<double> &pts, int i, double x) {
  double s = 0.0;
  for (size_t m = 0; m < pts.size(); ++m) {
    if ((int)m == i) continue;
    double t = 1.0 / (pts[i] - pts[m]);
    for (size_t j = 0; j < pts.size(); ++j)
      if ((int)j != i && j != m) t *= (x - pts[j]) / (pts[i] - pts[j]);
    s += t;
  }
  return s;
}
// Gauss-Legendre on [0,1] by Newton iteration on P_n
void gauss(int n, std::vector<double> &x, std::vector<double> &w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}
// FE_Q support points: Gauss-Lobatto (endpoints + roots of P'_k)
std::vector<double> lobatto(int k) {
  std::vector<double> x(k + 1);
  x[0] = 0;
  x[k] = 1;
  for (int i = 1; i < k; ++i) {
    double z = -std::cos(M_PI * i / k);
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int m = 2; m <= k; ++m) {
        double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      double d1 = k * (z * p1 - p0) / (z * z - 1.0);
      double d2 = (2.0 * z * d1 - k * (k + 1.0) * p1) / (1.0 - z * z);
      double dz = d1 / d2;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
  return x;
}
// prolongation weights of a degree kk Lagrange basis: fine node r (0..2kk) of a
// coarse cell sits at xi = (r/kk + pts[r%kk]) / 2; weights are the coarse basis there
void prolong_table(int kk, double *tab, int ld) {
  std::vector<double> pts = lobatto(kk);
  for (int r = 0; r <= 2 * kk; ++r) {
    int fc = r < 2 * kk ? r / kk : 1;
    int i = r < 2 * kk ? r - kk * fc : kk;
    double xi = 0.5 * (fc + pts[i]);
    for (int m = 0; m <= kk; ++m) {
      double v = lag(pts, m, xi);
      tab[r * ld + m] = std::fabs(v) < 1e-15 ? 0.0 : v;
    }
  }
}
}  // namespace

void level_derive(LevelDesc &L) {
  L.vol = 1.0;
  for (int d = 0; d < 3; ++d) {
    L.ih[d] = 1.0 / L.h[d];
    L.vol *= L.h[d];
  }
  for (int i = 0; i < 3; ++i) {
    L.ih_f[i] = (float)L.ih[i];
    for (int k = 0; k < 3; ++k) {
      L.ihw[i][k] = L.ih[i] * L.ih[k];
      L.ihw_f[i][k] = (float)L.ihw[i][k];
    }
  }
  L.vol_f = (float)L.vol;
  L.s_f = (float)L.s;
}

void upload_basis_tables() {
  static bool done = false;
  if (done) return;
  BasisConst b[2] = {};
  for (int k = 2; k <= 3; ++k) {
    BasisConst &B = b[k - 2];
    std::vector<double> qx, qw;
    gauss(k + 1, qx, qw);
    std::vector<double> vp = lobatto(k), pp = lobatto(k - 1);
    for (int q = 0; q <= k; ++q) {
      B.w[q] = qw[q];
      B.x[q] = qx[q];
      for (int i = 0; i <= k; ++i) {
        B.S[q][i] = lag(vp, i, qx[q]);
        B.G[q][i] = lagd(vp, i, qx[q]);
        B.S2[q][i] = B.S[q][i] * B.S[q][i];
        B.G2[q][i] = B.G[q][i] * B.G[q][i];
      }
      for (int i = 0; i < k; ++i) {
        B.SP[q][i] = lag(pp, i, qx[q]);
        B.SP2[q][i] = B.SP[q][i] * B.SP[q][i];
      }
    }
    for (int q = 0; q <= k; ++q)
      for (int r = 0; r <= k; ++r) B.D[q][r] = lagd(qx, r, qx[q]);
    prolong_table(k, &B.W[0][0], 4);
    prolong_table(k - 1, &B.WP[0][0], 3);
  }
  MFS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_basis), b, sizeof(b)));
  BasisT<double> td[2];
  BasisT<float> tf[2];
  for (int k = 0; k < 2; ++k) {
    for (int q = 0; q < 4; ++q) {
      td[k].w[q] = b[k].w[q];
      tf[k].w[q] = (float)b[k].w[q];
      for (int i = 0; i < 4; ++i) {
        td[k].S[q][i] = b[k].S[q][i], td[k].G[q][i] = b[k].G[q][i], td[k].D[q][i] = b[k].D[q][i];
        td[k].SP[q][i] = b[k].SP[q][i], td[k].S2[q][i] = b[k].S2[q][i], td[k].G2[q][i] = b[k].G2[q][i];
        td[k].SP2[q][i] = b[k].SP2[q][i];
        tf[k].S[q][i] = (float)b[k].S[q][i], tf[k].G[q][i] = (float)b[k].G[q][i], tf[k].D[q][i] = (float)b[k].D[q][i];
        tf[k].SP[q][i] = (float)b[k].SP[q][i], tf[k].S2[q][i] = (float)b[k].S2[q][i];
        tf[k].G2[q][i] = (float)b[k].G2[q][i], tf[k].SP2[q][i] = (float)b[k].SP2[q][i];
      }
    }
  }
  MFS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tabd), td, sizeof(td)));
  MFS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tf, sizeof(tf)));
  done = true;
}

int tile_cells(int dim, int k, int out[3]) {
  if (dim == 3 && k == 2) {
    out[0] = Tile<3, 2>::TX, out[1] = Tile<3, 2>::TY, out[2] = Tile<3, 2>::TZ;
  } else if (dim == 3 && k == 3) {
    out[0] = Tile<3, 3>::TX, out[1] = Tile<3, 3>::TY, out[2] = Tile<3, 3>::TZ;
  } else if (dim == 2 && k == 2) {
    out[0] = Tile<2, 2>::TX, out[1] = Tile<2, 2>::TY, out[2] = 1;
  } else if (dim == 2 && k == 3) {
    out[0] = Tile<2, 3>::TX, out[1] = Tile<2, 3>::TY, out[2] = 1;
  } else {
    throw std::runtime_error("unsupported (dim, degree)");
  }
  return out[0] * out[1] * out[2];
}
int tile_threads(int dim, int k) {
  int t[3];
  return tile_cells(dim, k, t) * (k + 1);
}

#define CB tb<T, K>()

// ------------------------------------------------------------------ the cell kernel
// Collocation form of the sum factorization: nodal values are interpolated to
// quadrature-point values with S (Q x N) along every direction, and
// derivatives are taken on the Gauss points with D (Q x Q), D S = G.  Only the y
// direction crosses threads (LDS); x and z stay in registers.
#ifndef MFS_MINW_F32
#define MFS_MINW_F32 1
#endif
#ifndef MFS_MINW_F64
#define MFS_MINW_F64 1
#endif
template <typename T>
struct MinWaves {
  static constexpr int value = std::is_same<T, double>::value ? MFS_MINW_F64 : MFS_MINW_F32;
};
template <int DIM, int K, int OP, int MODE, typename T>
__global__ void __launch_bounds__((Cfg<DIM, K>::THREADS), (MinWaves<T>::value))
    k_cell(const LevelDesc L, const int color, const T *__restrict__ src_u, const T *__restrict__ src_p,
           const T *__restrict__ eta, const int eta_per_qp, const T *__restrict__ rhs_u,
           const T *__restrict__ rhs_p, T *__restrict__ dst_u, T *__restrict__ dst_p) {
  using C = Cfg<DIM, K>;
  constexpr int N = C::N, Q = C::Q, NP = C::NP, ZN = C::ZN, ZQ = C::ZQ, ZNP = C::ZNP, CT = C::CT;
  constexpr bool NEED_U = (OP == OP_STOKES || OP == OP_ABLOCK || OP == OP_LAPLACE);
  constexpr bool NEED_P = (OP == OP_STOKES || OP == OP_BT || OP == OP_MASS);
  constexpr bool OUT_U = (OP == OP_STOKES || OP == OP_ABLOCK || OP == OP_BT || OP == OP_ADIAG || OP == OP_RHS ||
                          OP == OP_LAPLACE || OP == OP_LDIAG);
  constexpr bool OUT_P = (OP == OP_STOKES || OP == OP_MASS || OP == OP_MDIAG);
  constexpr bool GRAD_TEST = (OP == OP_STOKES || OP == OP_ABLOCK || OP == OP_BT || OP == OP_LAPLACE);
  constexpr bool SQDIAG = (OP == OP_ADIAG || OP == OP_LDIAG);  // closed-form velocity diagonals
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  T *nb = sm;                // [DIM][VZ][VY][VX] tile velocity nodes (load phase)
  T *pb = sm + DIM * C::VN;  // [PZ][PY][PX] tile pressure nodes (load phase)
  T *xb = sm;                // exchange buffer, then per-cell contributions

  const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
  int tx, ty, tz;
  if constexpr (MODE == MODE_CELLS) {  // a batch of cells of a general mesh
    tx = ty = tz = 0;
  } else if constexpr (MODE == MODE_TSPLIT) {  // every tile in one launch
    int b = blockIdx.x;
    tx = b % L.ntiles[0];
    b /= L.ntiles[0];
    ty = b % L.ntiles[1];
    tz = b / L.ntiles[1];
  } else {
    const int hx = (L.ntiles[0] - px + 1) >> 1, hy = (L.ntiles[1] - py + 1) >> 1;
    int b = blockIdx.x;
    tx = 2 * (b % hx) + px;
    b /= hx;
    ty = 2 * (b % hy) + py;
    b /= hy;
    tz = 2 * b + pz;
  }
  const long long tile_id =
      MODE == MODE_CELLS ? (long long)xcd_order(blockIdx.x, gridDim.x)  // batches sharing nodes meet in one L2
                         : tx + (long long)L.ntiles[0] * (ty + (long long)L.ntiles[1] * tz);
  const int tid = threadIdx.x;
  const int c = tid % CT;
  // the y-slice index is uniform over a wave when a slice spans whole waves
  const int j = (CT % 64 == 0) ? __builtin_amdgcn_readfirstlane(tid / CT) : tid / CT;
  const int cx = c % C::TX, cy = (c / C::TX) % C::TY, cz = c / (C::TX * C::TY);
  const int vx0 = K * C::TX * tx, vy0 = K * C::TY * ty, vz0 = DIM == 3 ? K * C::TZ * tz : 0;
  const int ox0 = (K - 1) * C::TX * tx, oy0 = (K - 1) * C::TY * ty, oz0 = DIM == 3 ? (K - 1) * C::TZ * tz : 0;

  // ---- 1. cooperative tile load (constrained velocity DoFs read as zero);
  // general mesh: the cells' nodes through the cell lists, [comp][a][CT] and [a][CT]
  constexpr int NV = ZN * N * N;
  if constexpr (MODE == MODE_CELLS) {
    if constexpr (NEED_U) {
      for (int t = tid; t < DIM * NV * CT; t += C::THREADS) {
        const int comp = t / (NV * CT), r = t - comp * NV * CT;
        nb[t] = src_u[comp * L.vnodes + L.cell_u[tile_id * NV * CT + r]];
      }
    }
    if constexpr (NEED_P) {
      T *pc = sm + DIM * NV * CT;
      for (int t = tid; t < C::NPL * CT; t += C::THREADS) pc[t] = src_p[L.cell_p[tile_id * C::NPL * CT + t]];
    }
  } else if constexpr (NEED_U) {
    for (int t = tid; t < DIM * C::VN; t += C::THREADS) {
      const int comp = t / C::VN, r = t - comp * C::VN;
      const int lx = r % C::VX, ly = (r / C::VX) % C::VY, lz = r / (C::VX * C::VY);
      const int gx = vx0 + lx, gy = vy0 + ly, gz = vz0 + lz;
      T v = 0;
      if (gx < L.nv[0] && gy < L.nv[1] && gz < L.nv[2] && (L.plain_read || !is_con(L, comp, gx, gy, gz)))
        v = src_u[comp * L.vnodes + gx + (long long)L.nv[0] * (gy + (long long)L.nv[1] * gz)];
      nb[t] = v;
    }
  }
  if constexpr (NEED_P && MODE != MODE_CELLS) {
    for (int t = tid; t < C::PN; t += C::THREADS) {
      const int lx = t % C::PX, ly = (t / C::PX) % C::PY, lz = t / (C::PX * C::PY);
      const int gx = ox0 + lx, gy = oy0 + ly, gz = oz0 + lz;
      T v = 0;
      if (gx < L.np[0] && gy < L.np[1] && gz < L.np[2])
        v = src_p[gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz)];
      pb[t] = v;
    }
  }
  if constexpr (NEED_U || NEED_P) __syncthreads();

  T u[DIM][ZN][N];
  T p[ZNP][NP][NP];
  if constexpr (NEED_U) {
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int kz = 0; kz < ZN; ++kz)
#pragma unroll
        for (int i = 0; i < N; ++i)
          u[comp][kz][i] = MODE == MODE_CELLS
                               ? nb[(comp * NV + i + N * (j + N * kz)) * CT + c]
                               : nb[comp * C::VN + ((K * cz + kz) * C::VY + (K * cy + j)) * C::VX + K * cx + i];
  }
  if constexpr (NEED_P) {
#pragma unroll
    for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
      for (int jy = 0; jy < NP; ++jy)
#pragma unroll
        for (int i = 0; i < NP; ++i)
          p[kz][jy][i] = MODE == MODE_CELLS
                             ? sm[DIM * NV * CT + (i + NP * (jy + NP * kz)) * CT + c]
                             : pb[(((K - 1) * cz + kz) * C::PY + ((K - 1) * cy + jy)) * C::PX + (K - 1) * cx + i];
  }
  __syncthreads();  // nb/pb now reused as exchange buffer

  const T ih[3] = {(T)(1.0 / L.h[0]), (T)(1.0 / L.h[1]), (T)(DIM == 3 ? 1.0 / L.h[2] : 1.0)};
  const double jac = L.h[0] * L.h[1] * (DIM == 3 ? L.h[2] : 1.0);
  const T s = (T)L.s;
  // exchange slot of (component, qz, qx) for y-slice jj of this cell
  auto X = [&](int comp, int qz, int qx, int jj) -> T & {
    return xb[(((comp * ZQ + qz) * Q + qx) * Q + jj) * CT + c];
  };

  // ---- 2. quadrature-point values of u on slice qy = j
  T v[DIM][ZQ][Q];
  if constexpr (NEED_U) {
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp) {
      T vx[ZN][Q];
#pragma unroll
      for (int kz = 0; kz < ZN; ++kz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          T a = 0;
#pragma unroll
          for (int i = 0; i < N; ++i) a += CB.S[qx][i] * u[comp][kz][i];
          vx[kz][qx] = a;
        }
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          T a = 0;
          if constexpr (DIM == 3) {
#pragma unroll
            for (int kz = 0; kz < N; ++kz) a += CB.S[qz][kz] * vx[kz][qx];
          } else {
            a = vx[0][qx];
          }
          X(comp, qz, qx, j) = a;
        }
    }
    __syncthreads();
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          T a = 0;
#pragma unroll
          for (int jj = 0; jj < N; ++jj) a += CB.S[j][jj] * X(comp, qz, qx, jj);
          v[comp][qz][qx] = a;
        }
    __syncthreads();
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) X(comp, qz, qx, j) = v[comp][qz][qx];
    __syncthreads();
  }

  // ---- 3. plane by plane: gradients, quadrature-point operation, test-function sums
  T g[DIM][ZQ][Q];     // velocity test-function sums in quadrature-point space (gradient tests)
  T pacc[ZNP][NP][NP]; // pressure test-function sums of this slice
#pragma unroll
  for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
    for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) g[comp][qz][qx] = 0;
#pragma unroll
  for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
    for (int jy = 0; jy < NP; ++jy)
#pragma unroll
      for (int i = 0; i < NP; ++i) pacc[kz][jy][i] = 0;

#pragma unroll
  for (int qz = 0; qz < ZQ; ++qz) {
    T fy[DIM][Q];  // y-flux of this plane, stashed into LDS after the plane's u reads
    T fz_or_val[Q];
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) {
      const int qn = qx + Q * (j + Q * qz);
      const double wq = c_basis[K - 2].w[qx] * c_basis[K - 2].w[j] * (DIM == 3 ? c_basis[K - 2].w[qz] : 1.0);
      // general geometry: J^{-1} (row-major, e * DIM + d) and JxW per quadrature point
      T Jv[DIM * DIM];
      T JxW = (T)(jac * wq);
      if (L.geo) {
        constexpr int G = DIM * DIM + 1;
        const double *gq = L.geo + (tile_id * C::NQ + qn) * (long long)G * CT + c;
#pragma unroll
        for (int e = 0; e < DIM * DIM; ++e) Jv[e] = (T)gq[e * CT];
        JxW = (T)gq[DIM * DIM * CT];
      }
      T et = 1;
      if constexpr (OP != OP_BT && OP != OP_LAPLACE && OP != OP_LDIAG)
        et = eta_per_qp ? eta[(tile_id * C::NQ + qn) * CT + c] : eta[tile_id * CT + c];
      T pqv = 0;
      if constexpr (NEED_P) {
#pragma unroll
        for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
          for (int jy = 0; jy < NP; ++jy)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
              const T wz = DIM == 3 ? CB.SP[qz][kz] : (T)1;
              pqv += wz * CB.SP[j][jy] * CB.SP[qx][i] * p[kz][jy][i];
            }
      }
      T gr[DIM][DIM];
      if constexpr (NEED_U) {
#pragma unroll
        for (int comp = 0; comp < DIM; ++comp) {
          T dx = 0, dy = 0, dz = 0;
#pragma unroll
          for (int r = 0; r < Q; ++r) {
            dx += CB.D[qx][r] * v[comp][qz][r];
            dy += CB.D[j][r] * X(comp, qz, qx, r);
            if constexpr (DIM == 3) dz += CB.D[qz][r] * v[comp][r][qx];
          }
          if (L.geo) {
            const T rd[3] = {dx, dy, dz};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              T a = 0;
#pragma unroll
              for (int e = 0; e < DIM; ++e) a += Jv[e * DIM + d] * rd[e];
              gr[comp][d] = a;
            }
          } else {
            gr[comp][0] = dx * ih[0];
            gr[comp][1] = dy * ih[1];
            if constexpr (DIM == 3) gr[comp][2] = dz * ih[2];
          }
        }
      }
      T fp = 0;
      if constexpr (GRAD_TEST) {
        T flux[DIM][DIM];
        if constexpr (OP == OP_STOKES || OP == OP_ABLOCK) {
          T div = 0;
#pragma unroll
          for (int a = 0; a < DIM; ++a) div += gr[a][a];
#pragma unroll
          for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int bb = 0; bb < DIM; ++bb) {
              T sig = et * (gr[a][bb] + gr[bb][a]);
              if (a == bb) {
                if (L.compressible) sig -= 2 * et / 3 * div;
                if constexpr (OP == OP_STOKES) sig -= s * pqv;
              }
              flux[a][bb] = sig * JxW;
            }
          if constexpr (OP == OP_STOKES) fp = -s * div * JxW;
        } else if constexpr (OP == OP_LAPLACE) {  // grad u : grad v, component by component
#pragma unroll
          for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int bb = 0; bb < DIM; ++bb) flux[a][bb] = gr[a][bb] * JxW;
        } else {  // B^T: -s p I
#pragma unroll
          for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int bb = 0; bb < DIM; ++bb) flux[a][bb] = a == bb ? -s * pqv * JxW : (T)0;
        }
        // fluxes against the reference derivatives of the test functions:
        // f_e = sum_d J^{-1}[e][d] f_d (box: f_e / h_e)
#pragma unroll
        for (int comp = 0; comp < DIM; ++comp) {
          T fr[DIM];
          if (L.geo) {
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
              T a = 0;
#pragma unroll
              for (int d = 0; d < DIM; ++d) a += Jv[e * DIM + d] * flux[comp][d];
              fr[e] = a;
            }
          } else {
#pragma unroll
            for (int e = 0; e < DIM; ++e) fr[e] = flux[comp][e] * ih[e];
          }
          // x and z tests accumulate in registers, y test is stashed
#pragma unroll
          for (int r = 0; r < Q; ++r) {
            g[comp][qz][r] += CB.D[qx][r] * fr[0];
            if constexpr (DIM == 3) g[comp][r][qx] += CB.D[qz][r] * fr[2];
          }
          fy[comp][qx] = fr[1];
        }
      } else if constexpr (OP == OP_MASS) {
        fp = s * s / et * pqv * JxW;
      } else if constexpr (OP == OP_MDIAG) {
        fp = s * s / et * JxW;
      } else if constexpr (OP == OP_RHS || SQDIAG) {
        fz_or_val[qx] = et * JxW;  // rho JxW for the right-hand side, eta JxW (or JxW) for the diagonals
      }
      if constexpr (OUT_P) {
        constexpr bool SQ = (OP == OP_MDIAG);
#pragma unroll
        for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
          for (int jy = 0; jy < NP; ++jy)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
              const T wz = DIM == 3 ? (SQ ? CB.SP2[qz][kz] : CB.SP[qz][kz]) : (T)1;
              const T wy = SQ ? CB.SP2[j][jy] : CB.SP[j][jy];
              const T wx = SQ ? CB.SP2[qx][i] : CB.SP[qx][i];
              pacc[kz][jy][i] += wz * wy * wx * fp;
            }
      }
    }
    if constexpr (GRAD_TEST) {
      if constexpr (NEED_U) __syncthreads();  // every slice has read this plane's u values
#pragma unroll
      for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) X(comp, qz, qx, j) = fy[comp][qx];
    } else if constexpr (OP == OP_RHS || SQDIAG) {
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) X(0, qz, qx, j) = fz_or_val[qx];
    }
  }

  // ---- 4. y tests across slices, back to nodal values
  T out[DIM][ZN][N];
  if constexpr (GRAD_TEST) {
    __syncthreads();
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          T a = 0;
#pragma unroll
          for (int qy = 0; qy < Q; ++qy) a += CB.D[qy][j] * X(comp, qz, qx, qy);
          g[comp][qz][qx] += a;
        }
    __syncthreads();
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) X(comp, qz, qx, j) = g[comp][qz][qx];
    __syncthreads();
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp) {
      T h[ZQ][Q];
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          T a = 0;
#pragma unroll
          for (int qy = 0; qy < Q; ++qy) a += CB.S[qy][j] * X(comp, qz, qx, qy);
          h[qz][qx] = a;
        }
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T hx[ZQ];
#pragma unroll
        for (int qz = 0; qz < ZQ; ++qz) {
          T a = 0;
#pragma unroll
          for (int qx = 0; qx < Q; ++qx) a += CB.S[qx][i] * h[qz][qx];
          hx[qz] = a;
        }
#pragma unroll
        for (int kz = 0; kz < ZN; ++kz) {
          T a = 0;
          if constexpr (DIM == 3) {
#pragma unroll
            for (int qz = 0; qz < ZQ; ++qz) a += CB.S[qz][kz] * hx[qz];
          } else {
            a = hx[0];
          }
          out[comp][kz][i] = a;
        }
      }
    }
  } else if constexpr (OP == OP_RHS || SQDIAG) {
    __syncthreads();
    constexpr bool SQ = SQDIAG;
    T bs[ZQ][Q], bg[ZQ][Q];
#pragma unroll
    for (int qz = 0; qz < ZQ; ++qz)
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T s0 = 0, s1 = 0;
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) {
          const T f = X(0, qz, qx, qy);
          s0 += (SQ ? CB.S2[qy][j] : CB.S[qy][j]) * f;
          if constexpr (SQ) s1 += CB.G2[qy][j] * f;
        }
        bs[qz][qx] = s0;
        bg[qz][qx] = s1;
      }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T rs[ZQ], rg[ZQ], rsg[ZQ];
#pragma unroll
      for (int qz = 0; qz < ZQ; ++qz) {
        T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
          if constexpr (SQ) {
            a0 += CB.G2[qx][i] * bs[qz][qx];  // D_x: G2 x, S2 y
            a1 += CB.S2[qx][i] * bg[qz][qx];  // D_y: S2 x, G2 y
            a2 += CB.S2[qx][i] * bs[qz][qx];  // D_z: S2 x, S2 y
          } else {
            a0 += CB.S[qx][i] * bs[qz][qx];
          }
        }
        rs[qz] = a0;
        rg[qz] = a1;
        rsg[qz] = a2;
      }
#pragma unroll
      for (int kz = 0; kz < ZN; ++kz) {
        T a0 = 0, a1 = 0, a2 = 0;
        if constexpr (DIM == 3) {
#pragma unroll
          for (int qz = 0; qz < ZQ; ++qz) {
            if constexpr (SQ) {
              a0 += CB.S2[qz][kz] * rs[qz];
              a1 += CB.S2[qz][kz] * rg[qz];
              a2 += CB.G2[qz][kz] * rsg[qz];
            } else {
              a0 += CB.S[qz][kz] * rs[qz];
            }
          }
        } else {
          a0 = rs[0];
          a1 = rg[0];
        }
        if constexpr (SQ) {
          const T Dx = a0 * ih[0] * ih[0], Dy = a1 * ih[1] * ih[1], Dz = a2 * ih[2] * ih[2];
          const T sum = Dx + Dy + (DIM == 3 ? Dz : (T)0);
          // 2 eta eps(phi):eps(phi) = eta (|grad N|^2 + (dN/dx_c)^2)  [- 2/3 eta (dN/dx_c)^2 compressible]
          // the Laplace diagonal: |grad N|^2 for every component
          const T fc = OP == OP_LDIAG ? (T)0 : (L.compressible ? (T)(1.0 / 3.0) : (T)1);
          out[0][kz][i] = sum + fc * Dx;
          out[1][kz][i] = sum + fc * Dy;
          if constexpr (DIM == 3) out[2][kz][i] = sum + fc * Dz;
        } else {
#pragma unroll
          for (int comp = 0; comp < DIM; ++comp) out[comp][kz][i] = (T)L.gravity[comp] * a0;
        }
      }
    }
  }

  // ---- 5. per-cell contributions to LDS, then every tile node sums its
  // contributors in a fixed order (deterministic) and is written out
  constexpr int CU_OFF = 0;                         // [comp][kz][j][i][CT]
  constexpr int CP_OFF = DIM * ZN * N * N * CT;     // [kz][jy][i][Q][CT] pressure shares per slice
  if constexpr (MODE == MODE_CELLS) {
    if constexpr (OUT_U) {
#pragma unroll
      for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
        for (int kz = 0; kz < ZN; ++kz)
#pragma unroll
          for (int i = 0; i < N; ++i)
            dst_u[((tile_id * DIM + comp) * NV + i + N * (j + N * kz)) * CT + c] = out[comp][kz][i];
    }
    if constexpr (OUT_P) {
      __syncthreads();
#pragma unroll
      for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
        for (int jy = 0; jy < NP; ++jy)
#pragma unroll
          for (int i = 0; i < NP; ++i) xb[CP_OFF + ((((kz * NP + jy) * NP + i) * Q + j) * CT) + c] = pacc[kz][jy][i];
      __syncthreads();
      for (int t = tid; t < C::NPL * CT; t += C::THREADS) {
        const int a = t / CT, cc = t - a * CT;
        T val = 0;
#pragma unroll
        for (int jj = 0; jj < Q; ++jj) val += xb[CP_OFF + (a * Q + jj) * CT + cc];
        dst_p[tile_id * C::NPL * CT + t] = val;
      }
    }
    return;
  }
  __syncthreads();
  if constexpr (OUT_U) {
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp)
#pragma unroll
      for (int kz = 0; kz < ZN; ++kz)
#pragma unroll
        for (int i = 0; i < N; ++i) xb[CU_OFF + (((comp * ZN + kz) * N + j) * N + i) * CT + c] = out[comp][kz][i];
  }
  if constexpr (OUT_P) {
#pragma unroll
    for (int kz = 0; kz < ZNP; ++kz)
#pragma unroll
      for (int jy = 0; jy < NP; ++jy)
#pragma unroll
        for (int i = 0; i < NP; ++i) xb[CP_OFF + ((((kz * NP + jy) * NP + i) * Q + j) * CT) + c] = pacc[kz][jy][i];
  }
  __syncthreads();

  if constexpr (OUT_U)
    tile_write_u<DIM, K, OP, MODE, T, C::THREADS>(L, xb + CU_OFF, tid, tx, ty, tz, px, py, pz, src_u, rhs_u,
                                                  dst_u);
  if constexpr (OUT_P) {
    constexpr int KP = K - 1;
    for (int t = tid; t < C::PN; t += C::THREADS) {
      const int lx = t % C::PX, ly = (t / C::PX) % C::PY, lz = t / (C::PX * C::PY);
      const int gx = ox0 + lx, gy = oy0 + ly, gz = oz0 + lz;
      if (gx >= L.np[0] || gy >= L.np[1] || gz >= L.np[2]) continue;
      int ccx[2], iix[2], ncx = 0, ccy[2], iiy[2], ncy = 0, ccz[2], iiz[2], ncz = 0;
      if (lx % KP == 0 && lx > 0) { ccx[ncx] = lx / KP - 1; iix[ncx++] = KP; }
      if (lx / KP < C::TX) { ccx[ncx] = lx / KP; iix[ncx++] = lx % KP; }
      if (ly % KP == 0 && ly > 0) { ccy[ncy] = ly / KP - 1; iiy[ncy++] = KP; }
      if (ly / KP < C::TY) { ccy[ncy] = ly / KP; iiy[ncy++] = ly % KP; }
      if (DIM == 3) {
        if (lz % KP == 0 && lz > 0) { ccz[ncz] = lz / KP - 1; iiz[ncz++] = KP; }
        if (lz / KP < C::TZ) { ccz[ncz] = lz / KP; iiz[ncz++] = lz % KP; }
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
#pragma unroll
            for (int jj = 0; jj < Q; ++jj)
              val += xb[CP_OFF + ((((iiz[a] * NP + iiy[bb]) * NP + iix[e]) * Q + jj) * CT) + cell];
          }
        }
      }
      const bool shx = (lx == 0 && tx > 0) || (lx == C::PX - 1 && tx + 1 < L.ntiles[0]);
      const bool shy = (ly == 0 && ty > 0) || (ly == C::PY - 1 && ty + 1 < L.ntiles[1]);
      const bool shz = DIM == 3 && ((lz == 0 && tz > 0) || (lz == C::PZ - 1 && tz + 1 < L.ntiles[2]));
      const bool first = !((shx && px) || (shy && py) || (shz && pz));
      const long long gi = gx + (long long)L.np[0] * (gy + (long long)L.np[1] * gz);
      if constexpr (MODE == MODE_SET)
        dst_p[gi] = first ? val : dst_p[gi] + val;
      else if constexpr (MODE == MODE_RESID)
        dst_p[gi] = first ? rhs_p[gi] - val : dst_p[gi] - val;
      else
        dst_p[gi] += val;
    }
  }
}

template <int DIM, int K, int OP, int MODE, typename T>
void launch_dkm(const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru, const T *rp,
                T *du, T *dp, hipStream_t st) {
  using C = Cfg<DIM, K>;
  const size_t smem = (size_t)C::SMEM * sizeof(T);
  for (int color = 0; color < (1 << DIM); ++color) {
    const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
    const long long hx = (L.ntiles[0] - px + 1) / 2, hy = (L.ntiles[1] - py + 1) / 2,
                    hz = DIM == 3 ? (L.ntiles[2] - pz + 1) / 2 : 1;
    const long long nb = hx * hy * hz;
    if (nb <= 0) continue;
    hipLaunchKernelGGL((k_cell<DIM, K, OP, MODE, T>), dim3((unsigned)nb), dim3(C::THREADS), smem, st, L, color,
                       su, sp, eta, pq, ru, rp, du, dp);
  }
  MFS_HIP(hipGetLastError());
}

// Split A-block apply with the general cell kernel (2D degree 2): every tile
// in one launch writing its in-tile node sums to work[tile][comp][ly][lx]
// (2 x 17^2 per 8x8-cell tile), then one thread per velocity DoF sums the at
// most 4 tiles sharing it in a fixed order (per direction: tile g/16 - 1,
// local node 16 for g mod 16 == 0, g > 0; tile g/16, local node g mod 16 while
// g/16 < ntiles) and applies the mode (constrained rows take src), or a
// Jacobi-Chebyshev step (ChebArgs; constrained rows r = b - x), the
// arithmetic of v_cheb_first / v_cheb_step.
constexpr int GM2_CHEB = SPLIT_CHEB;
// General meshes: one launch of the cell kernel over all batches of 64 cells
// (MODE_CELLS), degree 2; the per-cell contributions go to cu / cp.
template <int DIM, int OP, typename T>
static void launch_cells_d(const LevelDesc &L, long long nbatch, const T *su, const T *sp, const T *eta, T *cu,
                           T *cp, hipStream_t st) {
  using C = Cfg<DIM, 2>;
  const size_t smem = (size_t)C::SMEM * sizeof(T);
  hipLaunchKernelGGL((k_cell<DIM, 2, OP, MODE_CELLS, T>), dim3((unsigned)nbatch), dim3(C::THREADS), smem, st, L, 0,
                     su, sp, eta, 1, (const T *)nullptr, (const T *)nullptr, cu, cp);
  MFS_HIP(hipGetLastError());
}
template <int DIM, typename T>
static void launch_cells_dim(Op op, const LevelDesc &L, long long nbatch, const T *su, const T *sp, const T *eta,
                             T *cu, T *cp, hipStream_t st) {
  switch (op) {
    case OP_STOKES: launch_cells_d<DIM, OP_STOKES, T>(L, nbatch, su, sp, eta, cu, cp, st); break;
    case OP_ABLOCK: launch_cells_d<DIM, OP_ABLOCK, T>(L, nbatch, su, sp, eta, cu, cp, st); break;
    case OP_BT: launch_cells_d<DIM, OP_BT, T>(L, nbatch, su, sp, eta, cu, cp, st); break;
    case OP_MASS: launch_cells_d<DIM, OP_MASS, T>(L, nbatch, su, sp, eta, cu, cp, st); break;
    default: throw std::invalid_argument("general mesh: unsupported operator");
  }
}
template <typename T>
void launch_cells(Op op, const LevelDesc &L, long long nbatch, const T *su, const T *sp, const T *eta, T *cu, T *cp,
                  hipStream_t st) {
  if (nbatch <= 0) return;
  if (L.k != 2 || !L.geo || !L.cell_u || !L.cell_p) throw std::invalid_argument("general mesh: degree 2 cell lists and geometry required");
  if (L.dim == 3)
    launch_cells_dim<3, T>(op, L, nbatch, su, sp, eta, cu, cp, st);
  else
    launch_cells_dim<2, T>(op, L, nbatch, su, sp, eta, cu, cp, st);
}
template void launch_cells<double>(Op, const LevelDesc &, long long, const double *, const double *, const double *,
                                   double *, double *, hipStream_t);

template <int MODE, typename T>
__global__ void k_gather_tiles2(const LevelDesc L, const T *__restrict__ work, const T *__restrict__ src_u,
                                const T *__restrict__ rhs_u, T *__restrict__ dst_u, const ChebArgs<T> ch) {
  using C = Cfg<2, 2>;
  constexpr int SX = 2 * C::TX, SY = 2 * C::TY;  // tile node spans
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 2 * L.vnodes) return;
  const int comp = (int)(i / L.vnodes);
  const long long node = i - comp * L.vnodes;
  const int gy = (int)(node / L.nv[0]), gx = (int)(node - (long long)gy * L.nv[0]);
  int tt[2][2], ll[2][2];
  bool ok[2][2];
  const int g[2] = {gx, gy}, sp[2] = {SX, SY};
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    tt[d][0] = g[d] / sp[d] - 1;
    ll[d][0] = sp[d];
    ok[d][0] = g[d] % sp[d] == 0 && g[d] > 0;
    tt[d][1] = g[d] / sp[d];
    ll[d][1] = g[d] % sp[d];
    ok[d][1] = g[d] / sp[d] < L.ntiles[d];
  }
  T val = 0;
#pragma unroll
  for (int sy = 0; sy < 2; ++sy)
#pragma unroll
    for (int sx = 0; sx < 2; ++sx) {
      const bool v = ok[1][sy] && ok[0][sx];
      const long long tile = tt[0][sx] + (long long)L.ntiles[0] * tt[1][sy];
      const long long wi = v ? (tile * 2 + comp) * C::VN + ll[0][sx] + C::VX * ll[1][sy] : 0;
      const T c = work[wi];
      val += v ? c : (T)0;
    }
  const bool con = is_con(L, comp, gx, gy, 0);
  if constexpr (MODE == GM2_CHEB) {
    const T xi = ch.x[i];
    const T r = ch.b[i] - (con ? xi : val);
    const T dd = ch.first ? ch.dinv[i] * r * ch.f2 : ch.f1 * ch.d[i] + ch.f2 * ch.dinv[i] * r;
    ch.d[i] = dd;
    ch.x[i] = xi + dd;
  } else {
    const T vv = con ? src_u[i] : val;
    if constexpr (MODE == MODE_SET)
      dst_u[i] = vv;
    else if constexpr (MODE == MODE_RESID)
      dst_u[i] = rhs_u[i] - vv;
    else
      dst_u[i] += vv;
  }
}

template <typename T>
void launch_ablock_split2(int mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                          T *work, const ChebArgs<T> &ch, hipStream_t st) {
  using C = Cfg<2, 2>;
  if (L.ntiles_total * 2 * C::VN >= (1LL << 40)) throw std::runtime_error("level too large for the split apply");
  const size_t smem = (size_t)C::SMEM * sizeof(T);
  hipLaunchKernelGGL((k_cell<2, 2, OP_ABLOCK, MODE_TSPLIT, T>), dim3((unsigned)L.ntiles_total), dim3(C::THREADS),
                     smem, st, L, 0, su, (const T *)nullptr, eta, pq, (const T *)nullptr, (const T *)nullptr, work,
                     (T *)nullptr);
  const dim3 grid((unsigned)((2 * L.vnodes + 255) / 256));
  if (mode == MODE_SET)
    hipLaunchKernelGGL((k_gather_tiles2<MODE_SET, T>), grid, dim3(256), 0, st, L, work, su, ru, du, ch);
  else if (mode == MODE_RESID)
    hipLaunchKernelGGL((k_gather_tiles2<MODE_RESID, T>), grid, dim3(256), 0, st, L, work, su, ru, du, ch);
  else if (mode == MODE_ADD)
    hipLaunchKernelGGL((k_gather_tiles2<MODE_ADD, T>), grid, dim3(256), 0, st, L, work, su, ru, du, ch);
  else
    hipLaunchKernelGGL((k_gather_tiles2<GM2_CHEB, T>), grid, dim3(256), 0, st, L, work, su, ru, du, ch);
  MFS_HIP(hipGetLastError());
}
template void launch_ablock_split2<double>(int, const LevelDesc &, const double *, const double *, int,
                                           const double *, double *, double *, const ChebArgs<double> &, hipStream_t);
template void launch_ablock_split2<float>(int, const LevelDesc &, const float *, const float *, int, const float *,
                                          float *, float *, const ChebArgs<float> &, hipStream_t);

// Pressure mass operator (degree 1, one viscosity value per cell), evaluated
// per node: out(n) = s^2 |K| sum over the <= 2^dim cells c around n of
// (1/eta_c) sum_m M1(n, m) p_m, M1 the Q1 reference mass matrix (tensor
// product of [[1/3, 1/6], [1/6, 1/3]], which the Gauss rule of
// MassMatrixOperator integrates exactly).  One thread per node and one launch
// per apply: no tile colours, deterministic by construction.  The cell weight
// w_c is 1/eta_c: eta from the tile-blocked cellwise table (MASS_TILES) or
// 1/eta from a natural-order table (MASS_INV).
constexpr int MASS_TILES = 0, MASS_INV = 1;
template <int DIM, int SRC, typename T>
__device__ __forceinline__ T mass_node(const LevelDesc &L, const T *__restrict__ src, const T *__restrict__ ceta,
                                       const int gx, const int gy, const int gz) {
  constexpr int NZ = DIM == 3 ? 3 : 1, OZ = DIM == 3 ? 2 : 1;
  const int sy = L.np[0], sz = L.np[0] * L.np[1];  // pressure node strides
  const int cy = L.n[0], cz = L.n[0] * L.n[1];     // cell strides
  const bool inner = gx > 0 && gy > 0 && gx + 1 < L.np[0] && gy + 1 < L.np[1] &&
                     (DIM == 2 || (gz > 0 && gz + 1 < L.np[2]));
  const int n0 = gx + sy * gy + sz * gz;  // this node
  // the 3^dim pressure values around the node (zero outside the level)
  T pv[NZ][3][3];
  if (inner) {  // away from the level boundary: constant offsets from the node
    const T *const p0 = src + n0;
#pragma unroll
    for (int dz = 0; dz < NZ; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) pv[dz][dy][dx] = p0[(dx - 1) + sy * (dy - 1) + sz * (DIM == 3 ? dz - 1 : 0)];
  } else {
#pragma unroll
    for (int dz = 0; dz < NZ; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const int oz = DIM == 3 ? dz - 1 : 0;
          const int x = gx - 1 + dx, y = gy - 1 + dy, z = gz + oz;
          const bool in = x >= 0 && y >= 0 && x < L.np[0] && y < L.np[1] && (DIM == 2 || (z >= 0 && z < L.np[2]));
          const T v = src[in ? n0 + (dx - 1) + sy * (dy - 1) + sz * oz : 0];
          pv[dz][dy][dx] = in ? v : (T)0;
        }
  }
  constexpr T m1[2] = {T(1) / T(3), T(1) / T(6)};  // Q1 mass on [0,1]: 1/3 on the diagonal, 1/6 off it
  T acc = 0;
#pragma unroll
  for (int oz = 0; oz < OZ; ++oz)
#pragma unroll
    for (int oy = 0; oy < 2; ++oy)
#pragma unroll
      for (int ox = 0; ox < 2; ++ox) {
        // cell g - 1 + o; the node is its local node 1 - o, its nodes are pv[o + j]
        const int c[3] = {gx - 1 + ox, gy - 1 + oy, DIM == 3 ? gz - 1 + oz : 0};
        const bool valid = inner || (c[0] >= 0 && c[1] >= 0 && c[0] < L.n[0] && c[1] < L.n[1] &&
                                     (DIM == 2 || (c[2] >= 0 && c[2] < L.n[2])));
        T cs = 0;
#pragma unroll
        for (int jz = 0; jz < OZ; ++jz)
#pragma unroll
          for (int jy = 0; jy < 2; ++jy)
#pragma unroll
            for (int jx = 0; jx < 2; ++jx) {
              const T wz = DIM == 3 ? m1[(1 - oz) != jz] : (T)1;
              cs += m1[(1 - ox) != jx] * m1[(1 - oy) != jy] * wz * pv[DIM == 3 ? oz + jz : 0][oy + jy][ox + jx];
            }
        if constexpr (SRC == MASS_INV) {
          const int c0 = (gx - 1) + cy * (gy - 1) + (DIM == 3 ? cz * (gz - 1) : 0);  // cell (g - 1)
          const T w = ceta[valid ? c0 + ox + cy * oy + cz * oz : 0];
          acc += valid ? cs * w : (T)0;
        } else {
          // tile-blocked cellwise viscosity: 4x4x4 tiles in 3D, 8x8 in 2D (64 cells)
          constexpr int TS = DIM == 3 ? 4 : 8, TSH = DIM == 3 ? 2 : 3;
          const int tx = c[0] >> TSH, ty = c[1] >> TSH, tz = DIM == 3 ? c[2] >> TSH : 0;
          const int ti = (tx + L.ntiles[0] * (ty + L.ntiles[1] * tz)) * 64 + (c[0] & (TS - 1)) +
                         TS * ((c[1] & (TS - 1)) + (DIM == 3 ? TS * (c[2] & (TS - 1)) : 0));
          const T e = ceta[valid ? ti : 0];
          acc += valid ? cs / e : (T)0;
        }
      }
  using LC = LevelConst<T>;
  return LC::vol(L) * LC::s(L) * LC::s(L) * acc;
}

// grid: x over the nodes of an xy plane, y over z (32-bit offsets, checked by the launcher)
template <int DIM, int SRC, int MODE, typename T>
__global__ void k_mass_nodes(const LevelDesc L, const T *__restrict__ src, const T *__restrict__ ceta,
                             const T *__restrict__ rhs, T *__restrict__ dst) {
  const unsigned plane = (unsigned)(L.np[0] * L.np[1]);
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int gz = blockIdx.y;
  const int gy = (int)(pidx / (unsigned)L.np[0]), gx = (int)pidx - gy * L.np[0];
  const T v = mass_node<DIM, SRC, T>(L, src, ceta, gx, gy, gz);
  const long long i = pidx + (long long)plane * gz;
  if constexpr (MODE == MODE_SET)
    dst[i] = v;
  else if constexpr (MODE == MODE_RESID)
    dst[i] = rhs[i] - v;
  else
    dst[i] += v;
}
// One Chebyshev step of the Schur smoother fused with its residual:
//   r = b - M x_in,  d = (first ? 0 : f1 d) + f2 D^-1 r,  x_out = x_in + d.
// x_in is read around every node, so x_out is a different buffer (the caller
// alternates the two).
template <int DIM, typename T>
__global__ void k_mass_cheb(const LevelDesc L, const T *__restrict__ xin, const T *__restrict__ ieta,
                            const T *__restrict__ b, const T *__restrict__ dinv, T *__restrict__ d,
                            T *__restrict__ xout, T f1, T f2, int first) {
  const unsigned plane = (unsigned)(L.np[0] * L.np[1]);
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int gz = blockIdx.y;
  const int gy = (int)(pidx / (unsigned)L.np[0]), gx = (int)pidx - gy * L.np[0];
  const long long i = pidx + (long long)plane * gz;
  const T bi = b[i], di = dinv[i], dold = first ? T(0) : d[i];
  const T v = mass_node<DIM, MASS_INV, T>(L, xin, ieta, gx, gy, gz);
  const T r = bi - v;
  const T dd = first ? di * r * f2 : f1 * dold + f2 * di * r;
  d[i] = dd;
  xout[i] = xin[i] + dd;
}
// The whole Schur smoothing of a small level in one workgroup: two x buffers
// and the inverse cell viscosities in LDS, b, D^-1 and d in registers, the steps of
// the Chebyshev sequence (seq: per step f1, f2, first) separated by workgroup
// barriers, x written back at the end.  Each step is the arithmetic of
// k_mass_cheb on the LDS copies; a sweep from x = 0 is a step with x_in = 0
// (r = b).  One launch instead of one per step on levels that are
// launch-bound.
template <int DIM, typename T>
__global__ void __launch_bounds__(1024) k_mass_smooth_small(const LevelDesc L, T *__restrict__ x,
                                                            const T *__restrict__ b, const T *__restrict__ dinv,
                                                            const T *__restrict__ ieta, const MassChebSeq seq,
                                                            const int x_zero) {
  // each thread owns the nodes tid + 1024 k (k < NPT): b, D^-1 and d of its
  // nodes stay in registers, their coordinates are computed once; x (two
  // buffers) and 1/eta live in LDS
  constexpr int NPT = DIM == 2 ? 4 : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *const sm = reinterpret_cast<T *>(smem_raw);
  const int n = (int)L.pnodes, nc = (int)L.cells;
  T *xa = sm, *xb = sm + n, *sie = sm + 2 * n;
  int gxyz[NPT][3];
  T bi[NPT], di[NPT], dd[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int i = threadIdx.x + 1024 * k;
    const int ii = i < n ? i : 0;
    const int np01 = L.np[0] * L.np[1];
    gxyz[k][2] = ii / np01;
    const int rem = ii - gxyz[k][2] * np01;
    gxyz[k][1] = rem / L.np[0];
    gxyz[k][0] = rem - gxyz[k][1] * L.np[0];
    bi[k] = b[ii];
    di[k] = dinv[ii];
    dd[k] = 0;
    if (i < n) xa[i] = x_zero ? (T)0 : x[i];
  }
  for (int c = threadIdx.x; c < nc; c += 1024) sie[c] = ieta[c];
  __syncthreads();
  for (int s = 0; s < seq.n; ++s) {
    const T f1 = (T)seq.f1[s], f2 = (T)seq.f2[s];
    const int first = seq.first[s];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = threadIdx.x + 1024 * k;
      if (i < n) {
        const T v = mass_node<DIM, MASS_INV, T>(L, xa, sie, gxyz[k][0], gxyz[k][1], gxyz[k][2]);
        const T r = bi[k] - v;
        dd[k] = first ? di[k] * r * f2 : f1 * dd[k] + f2 * di[k] * r;
        xb[i] = xa[i] + dd[k];
      }
    }
    __syncthreads();
    T *t = xa;
    xa = xb;
    xb = t;
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int i = threadIdx.x + 1024 * k;
    if (i < n) x[i] = xa[i];
  }
}

template <typename T>
bool launch_mass_smooth_small(const LevelDesc &L, T *x, const T *b, const T *dinv, const T *ieta,
                              const MassChebSeq &seq, bool x_zero, hipStream_t st) {
  if (!ieta || L.k != 2 || L.geo || seq.n < 1 || seq.n > MassChebSeq::MAX) return false;
  // where one workgroup beats a launch per step: 2D levels up to 64^2
  // pressure nodes, 3D levels up to 9^3 (a 3D step costs 27 reads per node)
  if (L.pnodes > (L.dim == 2 ? 4096 : 729)) return false;
  const size_t bytes = (size_t)(2 * L.pnodes + L.cells) * sizeof(T);
  if (bytes > kSmallSmoothLds) return false;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(1), dim3(1024), bytes, st, L, x, b, dinv, ieta, seq, x_zero ? 1 : 0);
  };
  if (L.dim == 3)
    go(k_mass_smooth_small<3, T>);
  else
    go(k_mass_smooth_small<2, T>);
  MFS_HIP(hipGetLastError());
  return true;
}
// dynamic LDS above 64 KB for the single-workgroup smoother (set once per
// device, before any launch or stream capture: mfs_create)
void init_small_smoother() {
  MFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mass_smooth_small<2, float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallSmoothLds));
  MFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mass_smooth_small<3, float>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallSmoothLds));
  MFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mass_smooth_small<2, double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallSmoothLds));
  MFS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mass_smooth_small<3, double>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallSmoothLds));
}
template bool launch_mass_smooth_small<double>(const LevelDesc &, double *, const double *, const double *,
                                               const double *, const MassChebSeq &, bool, hipStream_t);
template bool launch_mass_smooth_small<float>(const LevelDesc &, float *, const float *, const float *, const float *,
                                              const MassChebSeq &, bool, hipStream_t);

// natural-order inverse cell viscosity from a tile-blocked cellwise table
// (degree 2: 4x4x4 tiles in 3D, 8x8 in 2D)
template <typename T>
__global__ void k_inv_eta_nat(const LevelDesc L, const T *__restrict__ eta_tiles, T *__restrict__ out) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= L.cells) return;
  const int cx = (int)(c % L.n[0]), cy = (int)((c / L.n[0]) % L.n[1]), cz = (int)(c / ((long long)L.n[0] * L.n[1]));
  const int TS = L.dim == 3 ? 4 : 8;
  const long long ti = (cx / TS + (long long)L.ntiles[0] * (cy / TS + (long long)L.ntiles[1] * (cz / TS))) * 64 +
                       (cx % TS) + TS * ((cy % TS) + TS * (cz % TS));
  out[c] = (T)1 / eta_tiles[ti];
}

// node-wise mass kernels: degree 2 (Q1 pressure), box geometry, 32-bit offsets
inline bool mass_nodes_ok(const LevelDesc &L) {
  if (L.k != 2 || L.geo) return false;
  if (L.pnodes >= (1LL << 31) || L.ntiles_total * 64 >= (1LL << 31) || L.cells >= (1LL << 31))
    throw std::runtime_error("level too large for the node-wise mass apply (32-bit offsets)");
  return true;
}
inline dim3 mass_grid(const LevelDesc &L) {
  return dim3((unsigned)(((long long)L.np[0] * L.np[1] + 255) / 256), (unsigned)L.np[2]);
}
template <int SRC, typename T>
void launch_mass_nodes_src(Mode mode, const LevelDesc &L, const T *src, const T *ceta, const T *rhs, T *dst,
                           hipStream_t st) {
  const dim3 grid = mass_grid(L);
  auto go = [&](auto dimc) {
    constexpr int D = decltype(dimc)::value;
    if (mode == MODE_SET)
      hipLaunchKernelGGL((k_mass_nodes<D, SRC, MODE_SET, T>), grid, dim3(256), 0, st, L, src, ceta, rhs, dst);
    else if (mode == MODE_RESID)
      hipLaunchKernelGGL((k_mass_nodes<D, SRC, MODE_RESID, T>), grid, dim3(256), 0, st, L, src, ceta, rhs, dst);
    else
      hipLaunchKernelGGL((k_mass_nodes<D, SRC, MODE_ADD, T>), grid, dim3(256), 0, st, L, src, ceta, rhs, dst);
  };
  if (L.dim == 3)
    go(std::integral_constant<int, 3>());
  else
    go(std::integral_constant<int, 2>());
  MFS_HIP(hipGetLastError());
}

template <typename T>
bool launch_mass_inv(Mode mode, const LevelDesc &L, const T *src, const T *ieta, const T *rhs, T *dst,
                     hipStream_t st) {
  if (!ieta || !mass_nodes_ok(L)) return false;
  launch_mass_nodes_src<MASS_INV, T>(mode, L, src, ieta, rhs, dst, st);
  return true;
}
template <typename T>
bool launch_mass_cheb(const LevelDesc &L, const T *xin, const T *ieta, const T *b, const T *dinv, T *d, T *xout,
                      T f1, T f2, int first, hipStream_t st) {
  if (!ieta || !mass_nodes_ok(L)) return false;
  if (xin == xout) throw std::runtime_error("launch_mass_cheb: x_in and x_out must differ");
  const dim3 grid = mass_grid(L);
  if (L.dim == 3)
    hipLaunchKernelGGL((k_mass_cheb<3, T>), grid, dim3(256), 0, st, L, xin, ieta, b, dinv, d, xout, f1, f2, first);
  else
    hipLaunchKernelGGL((k_mass_cheb<2, T>), grid, dim3(256), 0, st, L, xin, ieta, b, dinv, d, xout, f1, f2, first);
  MFS_HIP(hipGetLastError());
  return true;
}
template bool launch_mass_cheb<double>(const LevelDesc &, const double *, const double *, const double *,
                                       const double *, double *, double *, double, double, int, hipStream_t);
template bool launch_mass_cheb<float>(const LevelDesc &, const float *, const float *, const float *, const float *,
                                      float *, float *, float, float, int, hipStream_t);
template bool launch_mass_inv<double>(Mode, const LevelDesc &, const double *, const double *, const double *,
                                      double *, hipStream_t);
template bool launch_mass_inv<float>(Mode, const LevelDesc &, const float *, const float *, const float *, float *,
                                     hipStream_t);
template <typename T>
void launch_inv_eta_nat(const LevelDesc &L, const T *eta_tiles, T *out, hipStream_t st) {
  const unsigned nb = (unsigned)((L.cells + 255) / 256);
  hipLaunchKernelGGL((k_inv_eta_nat<T>), dim3(nb), dim3(256), 0, st, L, eta_tiles, out);
  MFS_HIP(hipGetLastError());
}
template void launch_inv_eta_nat<double>(const LevelDesc &, const double *, double *, hipStream_t);
template void launch_inv_eta_nat<float>(const LevelDesc &, const float *, float *, hipStream_t);

template <typename T>
bool launch_mass_nodes(Mode mode, const LevelDesc &L, const T *src, const T *eta, int eta_per_qp, const T *rhs,
                       T *dst, hipStream_t st) {
  if (eta_per_qp || !mass_nodes_ok(L)) return false;
  static_assert(Tile<3, 2>::TX == 4 && Tile<3, 2>::TY == 4 && Tile<3, 2>::TZ == 4, "tile-blocked eta layout");
  static_assert(Tile<2, 2>::TX == 8 && Tile<2, 2>::TY == 8, "tile-blocked eta layout");
  launch_mass_nodes_src<MASS_TILES, T>(mode, L, src, eta, rhs, dst, st);
  return true;
}

template <int DIM, int K, typename T>
void launch_dk(Op op, Mode mode, const LevelDesc &L, const T *su, const T *sp, const T *eta, int pq, const T *ru,
               const T *rp, T *du, T *dp, hipStream_t st) {
#define MFS_CASE(O)                                                                          \
  case O:                                                                                    \
    if (mode == MODE_SET)                                                                    \
      launch_dkm<DIM, K, O, MODE_SET, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);           \
    else if (mode == MODE_RESID)                                                             \
      launch_dkm<DIM, K, O, MODE_RESID, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);         \
    else                                                                                     \
      launch_dkm<DIM, K, O, MODE_ADD, T>(L, su, sp, eta, pq, ru, rp, du, dp, st);           \
    break;
  switch (op) {
    MFS_CASE(OP_STOKES)
    MFS_CASE(OP_ABLOCK)
    MFS_CASE(OP_BT)
    MFS_CASE(OP_MASS)
    MFS_CASE(OP_ADIAG)
    MFS_CASE(OP_MDIAG)
    MFS_CASE(OP_RHS)
    MFS_CASE(OP_LAPLACE)
    MFS_CASE(OP_LDIAG)
  }
#undef MFS_CASE
}

// B^T block for Q2 x Q1 (the cell operator with test gradient and flux
// -s p I), evaluated per velocity node.  On a uniform box level every cell has
// the same Jacobian, and the 3-point Gauss rule integrates the products of a
// Q2 basis function (or its derivative) with a Q1 one exactly, so the operator
// is a tensor-product stencil over the pressure nodes:
//   (B^T p)_{a}(g) = -s |K| / h_a  sum_m p_m  prod_d T_d(g_d, m_d),
// T_d the 1D "derivative" table for d = a and the 1D "mass" table otherwise,
// summed over the (at most two) cells around g_d.  On [0,1] with Q2 nodes
// 0, 1/2, 1 and Q1 nodes 0, 1:
//   mass  int phi_i psi_m : [[1/6, 0], [1/3, 1/3], [0, 1/6]],
//   deriv int phi_i' psi_m: [[-5/6, -1/6], [2/3, -2/3], [1/6, 5/6]].
// A cell-midpoint node (g odd, cell c = (g-1)/2) takes pressure nodes c, c+1;
// a vertex node (g even, P = g/2) takes P-1 .. P+1 from the cells P-1 (local
// node 2) and P (local node 0) that exist.  Constrained velocity rows are zero.
// One thread per pair of velocity nodes along x (all components), one launch,
// fixed order.
template <typename T>
__device__ __forceinline__ void bt_taps(int g, int n, int &q0, T cm[3], T cd[3]) {
  // branch-free: neighbouring lanes alternate between midpoint and vertex nodes
  const bool odd = g & 1;
  const int P = g >> 1;
  const T l = P >= 1 ? T(1) : T(0), r = P < n ? T(1) : T(0);
  q0 = odd ? P : P - 1;
  cm[0] = odd ? T(1) / T(3) : T(0);
  cm[1] = odd ? T(1) / T(3) : (l + r) * (T(1) / T(6));
  cm[2] = 0;
  cd[0] = odd ? T(2) / T(3) : l * (T(1) / T(6));
  cd[1] = odd ? -T(2) / T(3) : (l - r) * (T(5) / T(6));
  cd[2] = odd ? T(0) : -r * (T(1) / T(6));
}
template <int DIM, int MODE, typename T, typename TO>
__global__ void __launch_bounds__(256) k_bt_nodes(const LevelDesc L, const T *__restrict__ p,
                                                  const T *__restrict__ rhs, TO *__restrict__ dst) {
  // arithmetic in T; the result is rounded to TO once, at the store.  One
  // thread per pair of nodes along x: the vertex node 2P and the midpoint node
  // 2P+1 (none for P = n0) share the pressure columns P-1 .. P+1.
  // grid: x over the pairs of an xy plane, y over z (32-bit offsets, checked
  // by the launcher)
  const int npair = L.n[0] + 1;
  const unsigned plane = (unsigned)(npair * L.nv[1]);
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int gz = blockIdx.y;
  const int gy = (int)(pidx / (unsigned)npair), P = (int)pidx - gy * npair;
  const bool has_odd = P < L.n[0];
  const long long nodeE = 2 * P + (long long)L.nv[0] * (gy + (long long)L.nv[1] * gz);
  // the streamed operand first, so its latency overlaps the pressure gathers
  T prevE[DIM], prevO[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const long long i = a * L.vnodes + nodeE;
    if constexpr (MODE == MODE_RESID) {
      prevE[a] = rhs[i];
      prevO[a] = has_odd ? rhs[i + 1] : T(0);
    } else if constexpr (MODE == MODE_ADD) {
      prevE[a] = (T)dst[i];
      prevO[a] = has_odd ? (T)dst[i + 1] : T(0);
    } else {
      prevE[a] = prevO[a] = 0;
    }
  }
  // x taps over the columns P-1, P, P+1
  const T l = P >= 1 ? T(1) : T(0), r = P < L.n[0] ? T(1) : T(0);
  const T cmE[3] = {T(0), (l + r) * (T(1) / T(6)), T(0)};
  const T cdE[3] = {l * (T(1) / T(6)), (l - r) * (T(5) / T(6)), -r * (T(1) / T(6))};
  const T cmO[3] = {T(0), T(1) / T(3), T(1) / T(3)};
  const T cdO[3] = {T(0), T(2) / T(3), -T(2) / T(3)};
  int q0[3];
  T cm[3][3], cd[3][3];
  bt_taps<T>(gy, L.n[1], q0[1], cm[1], cd[1]);
  if constexpr (DIM == 3) bt_taps<T>(gz, L.n[2], q0[2], cm[2], cd[2]);
  // clamped pressure offsets (a clamped tap has a zero coefficient)
  unsigned xo[3], yo[3], zo[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    xo[t] = (unsigned)min(max(P - 1 + t, 0), L.np[0] - 1);
    yo[t] = (unsigned)(min(max(q0[1] + t, 0), L.np[1] - 1) * L.np[0]);
    zo[t] = DIM == 3 ? (unsigned)(min(max(q0[2] + t, 0), L.np[2] - 1)) * (unsigned)(L.np[0] * L.np[1]) : 0u;
  }
  T pv[DIM == 3 ? 3 : 1][3][3];
#pragma unroll
  for (int tz = 0; tz < (DIM == 3 ? 3 : 1); ++tz)
#pragma unroll
    for (int ty = 0; ty < 3; ++ty)
#pragma unroll
      for (int tx = 0; tx < 3; ++tx) pv[tz][ty][tx] = p[zo[tz] + yo[ty] + xo[tx]];
  T valE[3] = {0, 0, 0}, valO[3] = {0, 0, 0};
#pragma unroll
  for (int tz = 0; tz < (DIM == 3 ? 3 : 1); ++tz) {
    const T mz = DIM == 3 ? cm[2][tz] : T(1), dz = DIM == 3 ? cd[2][tz] : T(0);
#pragma unroll
    for (int ty = 0; ty < 3; ++ty) {
      T rmE = 0, rdE = 0, rmO = 0, rdO = 0;  // x contractions with the mass and derivative taps
#pragma unroll
      for (int tx = 0; tx < 3; ++tx) {
        rmE += cmE[tx] * pv[tz][ty][tx];
        rdE += cdE[tx] * pv[tz][ty][tx];
        rmO += cmO[tx] * pv[tz][ty][tx];
        rdO += cdO[tx] * pv[tz][ty][tx];
      }
      valE[0] += rdE * cm[1][ty] * mz;
      valE[1] += rmE * cd[1][ty] * mz;
      valO[0] += rdO * cm[1][ty] * mz;
      valO[1] += rmO * cd[1][ty] * mz;
      if constexpr (DIM == 3) {
        valE[2] += rmE * cm[1][ty] * dz;
        valO[2] += rmO * cm[1][ty] * dz;
      }
    }
  }
  // boundary faces a node lies on: bit 2d (low face), bit 2d+1 (high face);
  // the midpoint node is never on an x face
  int onyz = 0;
  onyz |= (gy == 0 ? 1 << 2 : 0) | (gy == L.nv[1] - 1 ? 2 << 2 : 0);
  if constexpr (DIM == 3) onyz |= (gz == 0 ? 1 << 4 : 0) | (gz == L.nv[2] - 1 ? 2 << 4 : 0);
  const int onE = onyz | (P == 0 ? 1 : 0) | (2 * P == L.nv[0] - 1 ? 2 : 0);
  using LC = LevelConst<T>;
  const T sv = -LC::s(L) * LC::vol(L);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    bool conE = false, conO = false;
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      const bool hit = L.bc[f] == 2 || (L.bc[f] == 1 && (f >> 1) == a);
      conE = conE || (((onE >> f) & 1) && hit);
      conO = conO || (((onyz >> f) & 1) && hit);
    }
    const T ih = LC::ih(L, a);
    const T vE = conE ? T(0) : sv * ih * valE[a];
    const T vO = conO ? T(0) : sv * ih * valO[a];
    const long long i = a * L.vnodes + nodeE;
    if constexpr (MODE == MODE_SET) {
      dst[i] = (TO)vE;
      if (has_odd) dst[i + 1] = (TO)vO;
    } else if constexpr (MODE == MODE_RESID) {
      dst[i] = (TO)(prevE[a] - vE);
      if (has_odd) dst[i + 1] = (TO)(prevO[a] - vO);
    } else {
      dst[i] = (TO)(prevE[a] + vE);
      if (has_odd) dst[i + 1] = (TO)(prevO[a] + vO);
    }
  }
}
template <typename T, typename TO>
bool launch_bt_nodes(Mode mode, const LevelDesc &L, const T *src_p, const T *rhs_u, TO *dst_u, hipStream_t st) {
  if (L.k != 2 || L.geo) return false;
  if ((long long)L.nv[0] * L.nv[1] >= (1LL << 31) || L.pnodes >= (1LL << 31))
    throw std::runtime_error("level too large for the node-wise B^T apply (32-bit offsets)");
  const dim3 grid((unsigned)(((long long)(L.n[0] + 1) * L.nv[1] + 255) / 256), (unsigned)(L.dim == 3 ? L.nv[2] : 1));
#define MFS_BT(D, M) hipLaunchKernelGGL((k_bt_nodes<D, M, T, TO>), grid, dim3(256), 0, st, L, src_p, rhs_u, dst_u)
  if (L.dim == 3) {
    if (mode == MODE_SET) MFS_BT(3, MODE_SET);
    else if (mode == MODE_RESID) MFS_BT(3, MODE_RESID);
    else MFS_BT(3, MODE_ADD);
  } else {
    if (mode == MODE_SET) MFS_BT(2, MODE_SET);
    else if (mode == MODE_RESID) MFS_BT(2, MODE_RESID);
    else MFS_BT(2, MODE_ADD);
  }
#undef MFS_BT
  MFS_HIP(hipGetLastError());
  return true;
}

template bool launch_bt_nodes<double, float>(Mode, const LevelDesc &, const double *, const double *, float *,
                                             hipStream_t);
template bool launch_bt_nodes<double, double>(Mode, const LevelDesc &, const double *, const double *, double *,
                                              hipStream_t);

template <typename T>
void launch_cell_op(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                    int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                    T *split_work) {
  check_op_args(op, mode, src_u, src_p, eta, rhs_u, rhs_p, dst_u, dst_p);
  if (launch_cell_q2<T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st, split_work))
    return;
  if (op == OP_ABLOCK && split_work && L.dim == 2 && L.k == 2 && !L.geo) {
    launch_ablock_split2<T>(mode, L, src_u, eta, eta_per_qp, rhs_u, dst_u, split_work, ChebArgs<T>{}, st);
    return;
  }
  if ((op == OP_ADIAG || op == OP_LDIAG) && L.geo) {
    launch_vdiag_geo<T>(op, L, eta, eta_per_qp, dst_u, st);
    return;
  }
  if (op == OP_MASS && launch_mass_nodes<T>(mode, L, src_p, eta, eta_per_qp, rhs_p, dst_p, st)) return;
  if (op == OP_BT && launch_bt_nodes<T, T>(mode, L, src_p, rhs_u, dst_u, st)) return;
  if (L.dim == 3 && L.k == 2)
    launch_dk<3, 2, T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else if (L.dim == 3 && L.k == 3)
    launch_dk<3, 3, T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else if (L.dim == 2 && L.k == 2)
    launch_dk<2, 2, T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else if (L.dim == 2 && L.k == 3)
    launch_dk<2, 3, T>(op, mode, L, src_u, src_p, eta, eta_per_qp, rhs_u, rhs_p, dst_u, dst_p, st);
  else
    throw std::runtime_error("unsupported (dim, degree)");
}
template void launch_cell_op<double>(Op, Mode, const LevelDesc &, const double *, const double *, const double *,
                                     int, const double *, const double *, double *, double *, hipStream_t, double *);
template void launch_cell_op<float>(Op, Mode, const LevelDesc &, const float *, const float *, const float *, int,
                                    const float *, const float *, float *, float *, hipStream_t, float *);

// ------------------------------------------------------------------ free-surface face term
// StokesOperator::local_apply_boundary_face (matrix_free_operators.cc:563-602):
// on a free-surface face (normal direction d, side), -(t . u)(phi . n) over the
// face.  Pass 1, one thread per (face cell, face quadrature point): the value
// s = JxW (t . u)(x_q) with u read from the face's nodes (constrained DoFs as
// zero unless plain_read).  Pass 2, one thread per face node: the rows of the
// normal component add sign * (-n_d) sum over the <= 2^(dim-1) face cells at
// the node of sum_q phi_node(x_q) s_q, in a fixed order; constrained rows are
// skipped.  Face cells are lexicographic over the tangential directions t1 < t2
// (t1 fastest), face quadrature points likewise.
template <int DIM, int K>
__global__ void k_fs_values(const LevelDesc L, const int face, const double *__restrict__ stab,
                            const double *__restrict__ src_u, double *__restrict__ sv) {
  using B = tab::Basis<double, K>;
  constexpr int N = K + 1, Q = K + 1, NFQ = DIM == 3 ? Q * Q : Q, NFN = DIM == 3 ? N * N : N;
  const int d = face / 2, side = face % 2;
  const int t1 = d == 0 ? 1 : 0, t2 = DIM == 3 ? (d == 2 ? 1 : 2) : 0;
  const long long n1 = L.n[t1], n2 = DIM == 3 ? L.n[t2] : 1;
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= n1 * n2 * NFQ) return;
  const int q = (int)(idx % NFQ);
  const long long cell = idx / NFQ;
  const int c1 = (int)(cell % n1), c2 = (int)(cell / n1);
  const int q1 = q % Q, q2 = q / Q;
  const double area = L.h[t1] * (DIM == 3 ? L.h[t2] : 1.0);
  const double w = area * B::W(q1) * (DIM == 3 ? B::W(q2) : 1.0);
  int g[3] = {0, 0, 0};
  g[d] = side == 0 ? 0 : L.nv[d] - 1;
  double tu = 0;
#pragma unroll
  for (int a = 0; a < NFN; ++a) {
    const int a1 = a % N, a2 = a / N;
    g[t1] = K * c1 + a1;
    if (DIM == 3) g[t2] = K * c2 + a2;
    const double phi = B::S(q1, a1) * (DIM == 3 ? B::S(q2, a2) : 1.0);
    const long long node = g[0] + (long long)L.nv[0] * (g[1] + (long long)L.nv[1] * g[2]);
#pragma unroll
    for (int comp = 0; comp < DIM; ++comp) {
      const bool con = !L.plain_read && is_con(L, comp, g[0], g[1], g[2]);
      const double uv = con ? 0.0 : src_u[comp * L.vnodes + node];
      tu += stab[idx * DIM + comp] * phi * uv;
    }
  }
  sv[idx] = w * tu;
}

template <int DIM, int K>
__global__ void k_fs_apply(const LevelDesc L, const int face, const double *__restrict__ sv, const double sign,
                           double *__restrict__ dst_u) {
  using B = tab::Basis<double, K>;
  constexpr int Q = K + 1, NFQ = DIM == 3 ? Q * Q : Q;
  const int d = face / 2, side = face % 2;
  const int t1 = d == 0 ? 1 : 0, t2 = DIM == 3 ? (d == 2 ? 1 : 2) : 0;
  const long long m1 = L.nv[t1], m2 = DIM == 3 ? L.nv[t2] : 1;
  const long long n1 = L.n[t1], n2 = DIM == 3 ? L.n[t2] : 1;
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= m1 * m2) return;
  const int g1 = (int)(idx % m1), g2 = (int)(idx / m1);
  int g[3] = {0, 0, 0};
  g[d] = side == 0 ? 0 : L.nv[d] - 1;
  g[t1] = g1;
  if (DIM == 3) g[t2] = g2;
  if (is_con(L, d, g[0], g[1], g[2])) return;
  double val = 0;
  // contributing face cells per tangential direction: (g/K - 1, local K) when g % K == 0 and g > 0, (g/K, g % K)
#pragma unroll
  for (int e2 = 0; e2 < (DIM == 3 ? 2 : 1); ++e2) {
    int c2 = 0, a2 = 0;
    if (DIM == 3) {
      c2 = e2 == 0 ? g2 / K - 1 : g2 / K;
      a2 = e2 == 0 ? K : g2 % K;
      if (e2 == 0 && !(g2 % K == 0 && g2 > 0)) continue;
      if (c2 >= n2) continue;
    }
#pragma unroll
    for (int e1 = 0; e1 < 2; ++e1) {
      const int c1 = e1 == 0 ? g1 / K - 1 : g1 / K, a1 = e1 == 0 ? K : g1 % K;
      if (e1 == 0 && !(g1 % K == 0 && g1 > 0)) continue;
      if (c1 >= n1) continue;
      const double *s = sv + ((long long)c1 + n1 * c2) * NFQ;
#pragma unroll
      for (int q = 0; q < NFQ; ++q) {
        const int q1 = q % Q, q2 = q / Q;
        val += B::S(q1, a1) * (DIM == 3 ? B::S(q2, a2) : 1.0) * s[q];
      }
    }
  }
  const double nsign = side == 0 ? -1.0 : 1.0;
  dst_u[d * L.vnodes + g[0] + (long long)L.nv[0] * (g[1] + (long long)L.nv[1] * g[2])] += sign * (-nsign) * val;
}

long long free_surface_cells(const LevelDesc &L, int face) {
  const int d = face / 2;
  long long c = 1;
  for (int e = 0; e < L.dim; ++e)
    if (e != d) c *= L.n[e];
  return c;
}

void launch_free_surface(const LevelDesc &L, const int faces[6], const double *stab, const double *src_u,
                         double *sv, double sign, double *dst_u, hipStream_t st) {
  const int nfq = L.dim == 3 ? (L.k + 1) * (L.k + 1) : L.k + 1;
  long long off = 0;
  for (int f = 0; f < 2 * L.dim; ++f) {
    if (!faces[f]) continue;
    const int d = f / 2;
    const long long cells = free_surface_cells(L, f);
    long long nodes = 1;
    for (int e = 0; e < L.dim; ++e)
      if (e != d) nodes *= L.nv[e];
    const long long nv = cells * nfq;
    const unsigned g1 = (unsigned)((nv + 255) / 256), g2 = (unsigned)((nodes + 255) / 256);
    const double *tb = stab + off * L.dim;
    if (L.dim == 3 && L.k == 2) {
      hipLaunchKernelGGL((k_fs_values<3, 2>), dim3(g1), dim3(256), 0, st, L, f, tb, src_u, sv);
      hipLaunchKernelGGL((k_fs_apply<3, 2>), dim3(g2), dim3(256), 0, st, L, f, sv, sign, dst_u);
    } else if (L.dim == 3 && L.k == 3) {
      hipLaunchKernelGGL((k_fs_values<3, 3>), dim3(g1), dim3(256), 0, st, L, f, tb, src_u, sv);
      hipLaunchKernelGGL((k_fs_apply<3, 3>), dim3(g2), dim3(256), 0, st, L, f, sv, sign, dst_u);
    } else if (L.dim == 2 && L.k == 2) {
      hipLaunchKernelGGL((k_fs_values<2, 2>), dim3(g1), dim3(256), 0, st, L, f, tb, src_u, sv);
      hipLaunchKernelGGL((k_fs_apply<2, 2>), dim3(g2), dim3(256), 0, st, L, f, sv, sign, dst_u);
    } else {
      hipLaunchKernelGGL((k_fs_values<2, 3>), dim3(g1), dim3(256), 0, st, L, f, tb, src_u, sv);
      hipLaunchKernelGGL((k_fs_apply<2, 3>), dim3(g2), dim3(256), 0, st, L, f, sv, sign, dst_u);
    }
    off += nv;
  }
  MFS_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ transfers
// prolongation: one thread per fine node (all components)
template <int DIM, typename T>
__global__ void k_prolongate(const LevelDesc Lc, const LevelDesc Lf, const int kind, const T *__restrict__ src,
                             T *__restrict__ dst, const int add) {
  const int kidx = Lf.k - 2;
  const bool pres = kind == 1;
  const int kk = pres ? Lf.k - 1 : Lf.k;
  const int *nf = pres ? Lf.np : Lf.nv;
  const int *ncn = pres ? Lc.np : Lc.nv;
  const long long nfine = (long long)nf[0] * nf[1] * nf[2];
  const long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (f >= nfine) return;
  const int g[3] = {(int)(f % nf[0]), (int)((f / nf[0]) % nf[1]), (int)(f / ((long long)nf[0] * nf[1]))};
  int e[3], r[3];
  for (int d = 0; d < 3; ++d) {
    if (d >= DIM) {
      e[d] = 0;
      r[d] = 0;
      continue;
    }
    e[d] = g[d] / (2 * kk);
    r[d] = g[d] - 2 * kk * e[d];
    if (e[d] == Lc.n[d]) {
      e[d] = Lc.n[d] - 1;
      r[d] = 2 * kk;
    }
  }
  const int ncomp = pres ? 1 : DIM;
  const long long ncoarse = (long long)ncn[0] * ncn[1] * ncn[2];
  for (int comp = 0; comp < ncomp; ++comp) {
    double acc = 0;
    for (int mz = 0; mz <= (DIM == 3 ? kk : 0); ++mz) {
      const double wz = DIM == 3 ? (pres ? c_basis[kidx].WP[r[2]][mz] : c_basis[kidx].W[r[2]][mz]) : 1.0;
      if (wz == 0.0) continue;
      const int cz = DIM == 3 ? kk * e[2] + mz : 0;
      for (int my = 0; my <= kk; ++my) {
        const double wy = pres ? c_basis[kidx].WP[r[1]][my] : c_basis[kidx].W[r[1]][my];
        if (wy == 0.0) continue;
        const int cy = kk * e[1] + my;
        for (int mx = 0; mx <= kk; ++mx) {
          const double wx = pres ? c_basis[kidx].WP[r[0]][mx] : c_basis[kidx].W[r[0]][mx];
          if (wx == 0.0) continue;
          const int cx = kk * e[0] + mx;
          if (!pres && is_con(Lc, comp, cx, cy, cz)) continue;
          acc += wz * wy * wx * (double)src[comp * ncoarse + cx + (long long)ncn[0] * (cy + (long long)ncn[1] * cz)];
        }
      }
    }
    if (!pres && is_con(Lf, comp, g[0], g[1], g[2])) acc = 0.0;
    const long long gi = comp * nfine + f;
    dst[gi] = add ? (T)(dst[gi] + (T)acc) : (T)acc;
  }
}

// restriction (transpose of prolongation): one thread per coarse node (all components)
template <int DIM, typename T>
__global__ void k_restrict(const LevelDesc Lc, const LevelDesc Lf, const int kind, const T *__restrict__ src,
                           T *__restrict__ dst) {
  const int kidx = Lf.k - 2;
  const bool pres = kind == 1;
  const int kk = pres ? Lf.k - 1 : Lf.k;
  const int *nf = pres ? Lf.np : Lf.nv;
  const int *ncn = pres ? Lc.np : Lc.nv;
  const long long ncoarse = (long long)ncn[0] * ncn[1] * ncn[2];
  const long long cidx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cidx >= ncoarse) return;
  const int m[3] = {(int)(cidx % ncn[0]), (int)((cidx / ncn[0]) % ncn[1]), (int)(cidx / ((long long)ncn[0] * ncn[1]))};
  // per direction: fine nodes g in [2m - 2kk, 2m + 2kk] with nonzero weight
  int gl[3][13];
  double wl[3][13];
  int cnt[3];
  for (int d = 0; d < 3; ++d) {
    cnt[d] = 0;
    if (d >= DIM) {
      gl[d][0] = 0;
      wl[d][0] = 1.0;
      cnt[d] = 1;
      continue;
    }
    for (int g = 2 * m[d] - 2 * kk; g <= 2 * m[d] + 2 * kk; ++g) {
      if (g < 0 || g >= nf[d]) continue;
      int e = g / (2 * kk), r = g - 2 * kk * e;
      if (e == Lc.n[d]) {
        e = Lc.n[d] - 1;
        r = 2 * kk;
      }
      const int off = m[d] - kk * e;
      if (off < 0 || off > kk) continue;
      const double w = pres ? c_basis[kidx].WP[r][off] : c_basis[kidx].W[r][off];
      if (w == 0.0) continue;
      gl[d][cnt[d]] = g;
      wl[d][cnt[d]] = w;
      ++cnt[d];
    }
  }
  const long long nfine = (long long)nf[0] * nf[1] * nf[2];
  const int ncomp = pres ? 1 : DIM;
  for (int comp = 0; comp < ncomp; ++comp) {
    double acc = 0;
    if (pres || !is_con(Lc, comp, m[0], m[1], m[2])) {
      for (int a = 0; a < cnt[2]; ++a)
        for (int bb = 0; bb < cnt[1]; ++bb)
          for (int cc = 0; cc < cnt[0]; ++cc) {
            if (!pres && is_con(Lf, comp, gl[0][cc], gl[1][bb], gl[2][a])) continue;
            acc += wl[2][a] * wl[1][bb] * wl[0][cc] *
                   (double)src[comp * nfine + gl[0][cc] + (long long)nf[0] * (gl[1][bb] + (long long)nf[1] * gl[2][a])];
          }
    }
    dst[comp * ncoarse + cidx] = (T)acc;
  }
}

// 3D transfers with fixed tap patterns for Q2 velocity (KK = 2, nodes 0, 1/2, 1)
// and Q1 pressure (KK = 1).  Prolongation row of fine local node r in its
// coarse cell: the Lagrange basis at r / (2 KK); the zero taps are skipped at
// compile time.
template <int KK>
struct ProlongTaps;
template <>
struct ProlongTaps<2> {
  static constexpr double W[5][3] = {
      {1.0, 0.0, 0.0}, {0.375, 0.75, -0.125}, {0.0, 1.0, 0.0}, {-0.125, 0.75, 0.375}, {0.0, 0.0, 1.0}};
};
template <>
struct ProlongTaps<1> {
  static constexpr double W[3][2] = {{1.0, 0.0}, {0.5, 0.5}, {0.0, 1.0}};
};

__device__ __forceinline__ bool con_node(const LevelDesc &L, int comp, int gx, int gy, int gz) {
  const int g[3] = {gx, gy, gz};
  bool con = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int b = g[d] == 0 ? L.bc[2 * d] : (g[d] == L.nv[d] - 1 ? L.bc[2 * d + 1] : 0);
    con = con || b == 2 || (b == 1 && comp == d);
  }
  return con;
}

// Separable transfers (degree 2): the tensor-product prolongation P = Px Py Pz
// (and R = P^T) applied one direction per pass through two work arrays (2D:
// Px Py, one work array).  Each pass
// is a 1D stencil along one axis of an (n0, n1, n2) array per component:
// restriction, coarse node m takes the fine nodes 2m + o (a coarse vertex: o = 0
// (1), +-1 (3/8), +-3 (-1/8) for KK = 2, o = 0 (1), +-1 (1/2) for KK = 1; a cell
// midpoint node: o = 0 (1), +-1 (3/4)), prolongation, fine node g takes its coarse cell's nodes with
// the row of ProlongTaps.  Constrained DoFs: the first restriction pass reads
// constrained fine nodes as zero and the last one zeroes constrained coarse
// nodes; the first prolongation pass reads constrained coarse nodes as zero
// and the last one zeroes constrained fine nodes.
struct Dims3 {
  int n[3];
};
// Grid of the passes: x over the nodes of an xy plane of the output, y over
// (z, comp); 32-bit offsets (the launcher checks the sizes).  Arithmetic in T.
template <int KK, typename T>
__global__ void k_restrict_pass(const LevelDesc Lc, const LevelDesc Lf, const int ax, const Dims3 din, const Dims3 dout,
                                const int ncomp, const int first, const int last, const T *__restrict__ in,
                                T *__restrict__ out) {
  const unsigned plane = (unsigned)(dout.n[0] * dout.n[1]);
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = blockIdx.y;  // z + nz * comp
  const int comp = row / dout.n[2];
  const int o1 = (int)(pidx / (unsigned)dout.n[0]);
  const int o[3] = {(int)pidx - o1 * dout.n[0], o1, row - comp * dout.n[2]};
  const int stride = ax == 0 ? 1 : (ax == 1 ? din.n[0] : din.n[0] * din.n[1]);
  const int m = o[ax];
  int base[3] = {o[0], o[1], o[2]};
  base[ax] = 0;
  const int b0 = comp * din.n[0] * din.n[1] * din.n[2] + base[0] + din.n[0] * (base[1] + din.n[1] * base[2]);
  constexpr bool VEL = KK == 2;
  constexpr int NO = VEL ? 5 : 3;
  const bool vertex = !VEL || (m & 1) == 0;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    int off;
    T wv;
    if constexpr (VEL) {
      off = k == 0 ? -3 : (k == 1 ? -1 : (k == 2 ? 0 : (k == 3 ? 1 : 3)));
      wv = vertex ? (k == 2 ? T(1) : (k == 1 || k == 3 ? T(0.375) : T(-0.125)))
                  : (k == 2 ? T(1) : (k == 1 || k == 3 ? T(0.75) : T(0)));
    } else {
      off = k - 1;
      wv = k == 1 ? T(1) : T(0.5);
    }
    const int g = 2 * m + off;
    bool ok = wv != T(0) && g >= 0 && g < din.n[ax];
    if (VEL && first) {
      int f[3] = {o[0], o[1], o[2]};
      f[ax] = g;
      ok = ok && !con_node(Lf, comp, f[0], f[1], f[2]);
    }
    const T v = in[ok ? b0 + g * stride : b0];
    acc += ok ? wv * v : T(0);
  }
  if (last && VEL && con_node(Lc, comp, o[0], o[1], o[2])) acc = 0;
  out[(long long)row * plane + pidx] = acc;
}
template <int KK, typename T>
__global__ void k_prolong_pass(const LevelDesc Lc, const LevelDesc Lf, const int ax, const Dims3 din, const Dims3 dout,
                               const int ncomp, const int first, const int last, const int add,
                               const T *__restrict__ in, T *__restrict__ out) {
  const unsigned plane = (unsigned)(dout.n[0] * dout.n[1]);
  const unsigned pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= plane) return;
  const int row = blockIdx.y;  // z + nz * comp
  const int comp = row / dout.n[2];
  const int o1 = (int)(pidx / (unsigned)dout.n[0]);
  const int o[3] = {(int)pidx - o1 * dout.n[0], o1, row - comp * dout.n[2]};
  const int stride = ax == 0 ? 1 : (ax == 1 ? din.n[0] : din.n[0] * din.n[1]);
  const int g = o[ax];
  const int ncell = Lc.n[ax];
  int e = g / (2 * KK), rr = g - 2 * KK * e;
  if (e == ncell) {
    e = ncell - 1;
    rr = 2 * KK;
  }
  int base[3] = {o[0], o[1], o[2]};
  base[ax] = 0;
  const int b0 = comp * din.n[0] * din.n[1] * din.n[2] + base[0] + din.n[0] * (base[1] + din.n[1] * base[2]);
  constexpr bool VEL = KK == 2;
  T acc = 0;
#pragma unroll
  for (int mm = 0; mm <= KK; ++mm) {
    T wv = (T)ProlongTaps<KK>::W[0][mm];
#pragma unroll
    for (int q = 1; q <= 2 * KK; ++q) wv = rr == q ? (T)ProlongTaps<KK>::W[q][mm] : wv;
    const int cidx = KK * e + mm;
    bool ok = wv != T(0);
    if (VEL && first) {
      int c[3] = {o[0], o[1], o[2]};
      c[ax] = cidx;
      ok = ok && !con_node(Lc, comp, c[0], c[1], c[2]);
    }
    const T v = in[b0 + cidx * stride];
    acc += ok ? wv * v : T(0);
  }
  if (last && VEL && con_node(Lf, comp, o[0], o[1], o[2])) acc = 0;
  const long long oi = (long long)row * plane + pidx;
  out[oi] = add ? (T)(out[oi] + acc) : acc;
}

template <typename T>
static void run_sep_transfer(bool prolong, int kind, const LevelDesc &Lc, const LevelDesc &Lf, const T *src, T *dst,
                             bool add, T *work, hipStream_t st) {
  const bool vel = kind == 0;
  const int ncomp = vel ? Lf.dim : 1;
  if (Lf.dim == 2) {  // two passes through one work array (c, f)
    Dims3 dc, df, d1;
    for (int d = 0; d < 3; ++d) {
      dc.n[d] = vel ? Lc.nv[d] : Lc.np[d];
      df.n[d] = vel ? Lf.nv[d] : Lf.np[d];
    }
    if ((long long)df.n[0] * df.n[1] * ncomp >= (1LL << 31))
      throw std::runtime_error("level too large for the separable transfers (32-bit offsets)");
    d1 = prolong ? dc : df;
    if (prolong)
      d1.n[1] = df.n[1];  // coarse -> (c, f) -> fine
    else
      d1.n[0] = dc.n[0];  // fine -> (c, f) -> coarse
    auto blocks = [&](const Dims3 &d) {
      return dim3((unsigned)(((long long)d.n[0] * d.n[1] + 255) / 256), (unsigned)ncomp);
    };
    if (!prolong) {
      if (vel) {
        hipLaunchKernelGGL((k_restrict_pass<2, T>), blocks(d1), dim3(256), 0, st, Lc, Lf, 0, df, d1, ncomp, 1, 0, src,
                           work);
        hipLaunchKernelGGL((k_restrict_pass<2, T>), blocks(dc), dim3(256), 0, st, Lc, Lf, 1, d1, dc, ncomp, 0, 1,
                           (const T *)work, dst);
      } else {
        hipLaunchKernelGGL((k_restrict_pass<1, T>), blocks(d1), dim3(256), 0, st, Lc, Lf, 0, df, d1, ncomp, 0, 0, src,
                           work);
        hipLaunchKernelGGL((k_restrict_pass<1, T>), blocks(dc), dim3(256), 0, st, Lc, Lf, 1, d1, dc, ncomp, 0, 0,
                           (const T *)work, dst);
      }
    } else {
      if (vel) {
        hipLaunchKernelGGL((k_prolong_pass<2, T>), blocks(d1), dim3(256), 0, st, Lc, Lf, 1, dc, d1, ncomp, 1, 0, 0, src,
                           work);
        hipLaunchKernelGGL((k_prolong_pass<2, T>), blocks(df), dim3(256), 0, st, Lc, Lf, 0, d1, df, ncomp, 0, 1,
                           add ? 1 : 0, (const T *)work, dst);
      } else {
        hipLaunchKernelGGL((k_prolong_pass<1, T>), blocks(d1), dim3(256), 0, st, Lc, Lf, 1, dc, d1, ncomp, 0, 0, 0, src,
                           work);
        hipLaunchKernelGGL((k_prolong_pass<1, T>), blocks(df), dim3(256), 0, st, Lc, Lf, 0, d1, df, ncomp, 0, 0,
                           add ? 1 : 0, (const T *)work, dst);
      }
    }
    MFS_HIP(hipGetLastError());
    return;
  }
  Dims3 dc, df;
  for (int d = 0; d < 3; ++d) {
    dc.n[d] = vel ? Lc.nv[d] : Lc.np[d];
    df.n[d] = vel ? Lf.nv[d] : Lf.np[d];
  }
  if ((long long)df.n[0] * df.n[1] * df.n[2] * ncomp >= (1LL << 31))
    throw std::runtime_error("level too large for the separable transfers (32-bit offsets)");
  auto blocks = [&](const Dims3 &d) {
    return dim3((unsigned)(((long long)d.n[0] * d.n[1] + 255) / 256), (unsigned)(d.n[2] * ncomp));
  };
  if (!prolong) {
    // x, y, z: fine -> (c, f, f) -> (c, c, f) -> coarse
    Dims3 d1 = df, d2 = df;
    d1.n[0] = dc.n[0];
    d2.n[0] = dc.n[0];
    d2.n[1] = dc.n[1];
    T *t1 = work, *t2 = work + (long long)d1.n[0] * d1.n[1] * d1.n[2] * ncomp;
    const T *ins[3] = {src, t1, t2};
    T *outs[3] = {t1, t2, dst};
    const Dims3 di[3] = {df, d1, d2}, dd[3] = {d1, d2, dc};
    for (int ax = 0; ax < 3; ++ax) {
      if (vel)
        hipLaunchKernelGGL((k_restrict_pass<2, T>), blocks(dd[ax]), dim3(256), 0, st, Lc, Lf, ax, di[ax], dd[ax],
                           ncomp, ax == 0 ? 1 : 0, ax == 2 ? 1 : 0, ins[ax], outs[ax]);
      else
        hipLaunchKernelGGL((k_restrict_pass<1, T>), blocks(dd[ax]), dim3(256), 0, st, Lc, Lf, ax, di[ax], dd[ax],
                           ncomp, 0, 0, ins[ax], outs[ax]);
    }
  } else {
    // z, y, x: coarse -> (c, c, f) -> (c, f, f) -> fine
    Dims3 d1 = dc, d2 = dc;
    d1.n[2] = df.n[2];
    d2.n[2] = df.n[2];
    d2.n[1] = df.n[1];
    T *t1 = work, *t2 = work + (long long)d1.n[0] * d1.n[1] * d1.n[2] * ncomp;
    const T *ins[3] = {src, t1, t2};
    T *outs[3] = {t1, t2, dst};
    const Dims3 di[3] = {dc, d1, d2}, dd[3] = {d1, d2, df};
    const int axes[3] = {2, 1, 0};
    for (int p = 0; p < 3; ++p) {
      if (vel)
        hipLaunchKernelGGL((k_prolong_pass<2, T>), blocks(dd[p]), dim3(256), 0, st, Lc, Lf, axes[p], di[p], dd[p],
                           ncomp, p == 0 ? 1 : 0, p == 2 ? 1 : 0, p == 2 && add ? 1 : 0, ins[p], outs[p]);
      else
        hipLaunchKernelGGL((k_prolong_pass<1, T>), blocks(dd[p]), dim3(256), 0, st, Lc, Lf, axes[p], di[p], dd[p],
                           ncomp, 0, 0, p == 2 && add ? 1 : 0, ins[p], outs[p]);
    }
  }
  MFS_HIP(hipGetLastError());
}
long long transfer_work_size(const LevelDesc &Lc, const LevelDesc &Lf) {
  // two work arrays: velocity (c, f, f) and (c, c, f) per component (restriction) or
  // (c, c, f) and (c, f, f) (prolongation); sized for the larger of the two orders
  const long long a = (long long)Lc.nv[0] * Lf.nv[1] * Lf.nv[2], b = (long long)Lc.nv[0] * Lc.nv[1] * Lf.nv[2];
  return 3 * (a + b);
}

template <typename T>
void launch_prolongate(int kind, const LevelDesc &Lc, const LevelDesc &Lf, const T *src_c, T *dst_f, bool add,
                       hipStream_t st, T *work) {
  if (work && Lf.k == 2) {
    run_sep_transfer<T>(true, kind, Lc, Lf, src_c, dst_f, add, work, st);
    return;
  }
  const long long n = kind == 0 ? Lf.vnodes : Lf.pnodes;
  const int bs = 256;
  const unsigned nb = (unsigned)((n + bs - 1) / bs);
  if (Lf.dim == 3)
    hipLaunchKernelGGL((k_prolongate<3, T>), dim3(nb), dim3(bs), 0, st, Lc, Lf, kind, src_c, dst_f, add ? 1 : 0);
  else
    hipLaunchKernelGGL((k_prolongate<2, T>), dim3(nb), dim3(bs), 0, st, Lc, Lf, kind, src_c, dst_f, add ? 1 : 0);
  MFS_HIP(hipGetLastError());
}
template <typename T>
void launch_restrict(int kind, const LevelDesc &Lc, const LevelDesc &Lf, const T *src_f, T *dst_c,
                     hipStream_t st, T *work) {
  if (work && Lf.k == 2) {
    run_sep_transfer<T>(false, kind, Lc, Lf, src_f, dst_c, false, work, st);
    return;
  }
  const long long n = kind == 0 ? Lc.vnodes : Lc.pnodes;
  const int bs = 256;
  const unsigned nb = (unsigned)((n + bs - 1) / bs);
  if (Lf.dim == 3)
    hipLaunchKernelGGL((k_restrict<3, T>), dim3(nb), dim3(bs), 0, st, Lc, Lf, kind, src_f, dst_c);
  else
    hipLaunchKernelGGL((k_restrict<2, T>), dim3(nb), dim3(bs), 0, st, Lc, Lf, kind, src_f, dst_c);
  MFS_HIP(hipGetLastError());
}
template void launch_prolongate<double>(int, const LevelDesc &, const LevelDesc &, const double *, double *, bool,
                                        hipStream_t, double *);
template void launch_prolongate<float>(int, const LevelDesc &, const LevelDesc &, const float *, float *, bool,
                                       hipStream_t, float *);
template void launch_restrict<double>(int, const LevelDesc &, const LevelDesc &, const double *, double *,
                                      hipStream_t, double *);
template void launch_restrict<float>(int, const LevelDesc &, const LevelDesc &, const float *, float *,
                                     hipStream_t, float *);

// ------------------------------------------------------------------ viscosity tables
__device__ __forceinline__ long long tile_index(const LevelDesc &L, int gcx, int gcy, int gcz, int TX, int TY,
                                                int TZ) {
  const int tx = gcx / TX, ty = gcy / TY, tz = gcz / TZ;
  const int cx = gcx - tx * TX, cy = gcy - ty * TY, cz = gcz - tz * TZ;
  const long long tile = tx + (long long)L.ntiles[0] * (ty + (long long)L.ntiles[1] * tz);
  return tile * (TX * TY * TZ) + cx + TX * (cy + TY * cz);
}

// ------------------------------------------------------------------ general geometry
long long geo_table_size(const LevelDesc &L) {
  int t[3];
  const int ct = tile_cells(L.dim, L.k, t);
  const int nq = L.dim == 3 ? (L.k + 1) * (L.k + 1) * (L.k + 1) : (L.k + 1) * (L.k + 1);
  return L.ntiles_total * nq * (L.dim * L.dim + 1) * ct;
}

// natural [cell][q][g] -> [((tile * NQ + q) * G + g) * CT + c]; one thread per (cell, q)
__global__ void k_geo_to_tiles(const LevelDesc L, const double *__restrict__ nat, const int nq, const int TX,
                               const int TY, const int TZ, double *__restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= L.cells * nq) return;
  const long long cell = idx / nq;
  const int q = (int)(idx - cell * nq);
  const int gcx = (int)(cell % L.n[0]), gcy = (int)((cell / L.n[0]) % L.n[1]),
            gcz = (int)(cell / ((long long)L.n[0] * L.n[1]));
  const long long CT = TX * TY * TZ;
  const long long ti = tile_index(L, gcx, gcy, gcz, TX, TY, TZ);
  const long long tile = ti / CT, c = ti - tile * CT;
  const int G = L.dim * L.dim + 1;
  for (int g = 0; g < G; ++g) out[((tile * nq + q) * G + g) * CT + c] = nat[idx * G + g];
}

void launch_geo_to_tiles(const LevelDesc &L, const double *nat, double *tiled, hipStream_t st) {
  int t[3];
  tile_cells(L.dim, L.k, t);
  const int nq = L.dim == 3 ? (L.k + 1) * (L.k + 1) * (L.k + 1) : (L.k + 1) * (L.k + 1);
  v_fill<double>(tiled, 0.0, geo_table_size(L), st);
  const long long n = L.cells * nq;
  hipLaunchKernelGGL(k_geo_to_tiles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, nat, nq, t[0], t[1],
                     t[2], tiled);
  MFS_HIP(hipGetLastError());
}

// Velocity diagonals under a general geometry (A block: eta (|grad N|^2 +
// (d N / d x_c)^2) [- 2/3 eta (d N / d x_c)^2 compressible]; Laplace:
// |grad N|^2), grad N = J^{-T} grad_ref N at every quadrature point.  One
// thread per (cell of one parity colour, local node): cells of a colour share
// no node, so the launch adds to dst without conflicts.
template <int DIM, int K, int OP, typename T>
__global__ void k_vdiag_geo(const LevelDesc L, const int color, const T *__restrict__ eta, const int eta_per_qp,
                            T *__restrict__ dst) {
  using B = tab::Basis<double, K>;
  using C = Cfg<DIM, K>;
  constexpr int N = K + 1, Q = K + 1, NV = DIM == 3 ? N * N * N : N * N, NQ = C::NQ, G = DIM * DIM + 1;
  const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
  const long long hx = (L.n[0] - px + 1) / 2, hy = (L.n[1] - py + 1) / 2, hz = DIM == 3 ? (L.n[2] - pz + 1) / 2 : 1;
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= hx * hy * hz * NV) return;
  const int a = (int)(idx % NV);
  const long long ci = idx / NV;
  const int cx = (int)(2 * (ci % hx) + px), cy = (int)(2 * ((ci / hx) % hy) + py),
            cz = DIM == 3 ? (int)(2 * (ci / (hx * hy)) + pz) : 0;
  const long long ti = tile_index(L, cx, cy, cz, C::TX, C::TY, C::TZ);
  const long long tile = ti / C::CT, c = ti - tile * C::CT;
  const int ai[3] = {a % N, (a / N) % N, a / (N * N)};
  double acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) acc[d] = 0;
  for (int q = 0; q < NQ; ++q) {
    const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
    double r[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      double v = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) v *= d == e ? B::G(qi[d], ai[d]) : B::S(qi[d], ai[d]);
      r[e] = v;
    }
    const double *gq = L.geo + (tile * NQ + q) * (long long)G * C::CT + c;
    double gr[DIM], g2 = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double v = 0;
#pragma unroll
      for (int e = 0; e < DIM; ++e) v += gq[(e * DIM + d) * C::CT] * r[e];
      gr[d] = v;
      g2 += v * v;
    }
    const double JxW = gq[(G - 1) * C::CT];
    if constexpr (OP == OP_LDIAG) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) acc[d] += g2 * JxW;
    } else {
      const double et = (double)(eta_per_qp ? eta[(tile * NQ + q) * C::CT + c] : eta[tile * C::CT + c]);
      const double fc = L.compressible ? 1.0 / 3.0 : 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) acc[d] += et * (g2 + fc * gr[d] * gr[d]) * JxW;
    }
  }
  const int g[3] = {K * cx + ai[0], K * cy + ai[1], DIM == 3 ? K * cz + ai[2] : 0};
  const long long node = g[0] + (long long)L.nv[0] * (g[1] + (long long)L.nv[1] * g[2]);
#pragma unroll
  for (int d = 0; d < DIM; ++d)
    if (!is_con(L, d, g[0], g[1], g[2])) dst[d * L.vnodes + node] += (T)acc[d];
}

// constrained velocity DoFs of a diagonal: 1
template <typename T>
__global__ void k_con_one(const LevelDesc L, T *__restrict__ dst) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= L.dim * L.vnodes) return;
  const int comp = (int)(idx / L.vnodes);
  const long long node = idx - comp * L.vnodes;
  const int gx = (int)(node % L.nv[0]), gy = (int)((node / L.nv[0]) % L.nv[1]),
            gz = (int)(node / ((long long)L.nv[0] * L.nv[1]));
  if (is_con(L, comp, gx, gy, gz)) dst[idx] = (T)1;
}

template <int DIM, int K, typename T>
void launch_vdiag_geo_dk(Op op, const LevelDesc &L, const T *eta, int pq, T *dst, hipStream_t st) {
  constexpr int N = K + 1, NV = DIM == 3 ? N * N * N : N * N;
  v_fill<T>(dst, (T)0, L.dim * L.vnodes, st);
  for (int color = 0; color < (1 << DIM); ++color) {
    const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
    const long long n = (long long)((L.n[0] - px + 1) / 2) * ((L.n[1] - py + 1) / 2) *
                        (DIM == 3 ? (L.n[2] - pz + 1) / 2 : 1) * NV;
    if (n <= 0) continue;
    const dim3 grid((unsigned)((n + 127) / 128));
    if (op == OP_LDIAG)
      hipLaunchKernelGGL((k_vdiag_geo<DIM, K, OP_LDIAG, T>), grid, dim3(128), 0, st, L, color, eta, pq, dst);
    else
      hipLaunchKernelGGL((k_vdiag_geo<DIM, K, OP_ADIAG, T>), grid, dim3(128), 0, st, L, color, eta, pq, dst);
  }
  const long long nu = L.dim * L.vnodes;
  hipLaunchKernelGGL(k_con_one<T>, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, st, L, dst);
  MFS_HIP(hipGetLastError());
}
template <typename T>
void launch_vdiag_geo(Op op, const LevelDesc &L, const T *eta, int pq, T *dst, hipStream_t st) {
  if (L.dim == 3 && L.k == 2)
    launch_vdiag_geo_dk<3, 2, T>(op, L, eta, pq, dst, st);
  else if (L.dim == 3 && L.k == 3)
    launch_vdiag_geo_dk<3, 3, T>(op, L, eta, pq, dst, st);
  else if (L.dim == 2 && L.k == 2)
    launch_vdiag_geo_dk<2, 2, T>(op, L, eta, pq, dst, st);
  else
    launch_vdiag_geo_dk<2, 3, T>(op, L, eta, pq, dst, st);
}


// natural [cell][q] (or [cell]) -> tile-blocked; one thread per (cell, q)
template <typename T>
__global__ void k_eta_to_tiles(const LevelDesc L, const double *__restrict__ eta_nat, const int per_qp, const int nq,
                               const int TX, const int TY, const int TZ, T *__restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int nqq = per_qp ? nq : 1;
  if (idx >= L.cells * nqq) return;
  const long long cell = idx / nqq;
  const int q = (int)(idx - cell * nqq);
  const int gcx = (int)(cell % L.n[0]), gcy = (int)((cell / L.n[0]) % L.n[1]),
            gcz = (int)(cell / ((long long)L.n[0] * L.n[1]));
  const long long ti = tile_index(L, gcx, gcy, gcz, TX, TY, TZ);
  const long long CT = TX * TY * TZ;
  const long long tile = ti / CT, c = ti - tile * CT;
  out[(tile * nqq + q) * CT + c] = (T)eta_nat[idx];
}

// cellwise averaging of natural per-qp values into a tile-blocked per-cell table
template <typename T>
__global__ void k_eta_average(const LevelDesc L, const double *__restrict__ eta_nat, const int nq, const int Q,
                              const int mode, const int TX, const int TY, const int TZ, T *__restrict__ out) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cell >= L.cells) return;
  const int kidx = L.k - 2;
  double acc = 0, wsum = 0;
  bool zero = false;
  for (int q = 0; q < nq; ++q) {
    const double v = eta_nat[cell * nq + q];
    if (mode == 1)
      acc += v;
    else if (mode == 2) {
      if (v == 0.0) zero = true;
      else acc += 1.0 / v;
    } else if (mode == 3)
      acc += log(v);
    else {
      const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
      const double w = c_basis[kidx].w[qx] * c_basis[kidx].w[qy] * (L.dim == 3 ? c_basis[kidx].w[qz] : 1.0);
      acc += w * v;
      wsum += w;
    }
  }
  double r;
  if (mode == 1)
    r = acc / nq;
  else if (mode == 2)
    r = zero ? 0.0 : 1.0 / (acc / nq);
  else if (mode == 3)
    r = exp(acc / nq);
  else
    r = acc / wsum;
  const int gcx = (int)(cell % L.n[0]), gcy = (int)((cell / L.n[0]) % L.n[1]),
            gcz = (int)(cell / ((long long)L.n[0] * L.n[1]));
  out[tile_index(L, gcx, gcy, gcz, TX, TY, TZ)] = (T)r;
}

// quadrature-weighted cell mean of a tile-blocked per-qp table
template <typename T, typename S>
__global__ void k_eta_qp_to_cell(const LevelDesc L, const S *__restrict__ in, const int Q, const int nq,
                                 const int TX, const int TY, const int TZ, T *__restrict__ out) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cell >= L.cells) return;
  const int kidx = L.k - 2;
  const int gcx = (int)(cell % L.n[0]), gcy = (int)((cell / L.n[0]) % L.n[1]),
            gcz = (int)(cell / ((long long)L.n[0] * L.n[1]));
  const long long CT = TX * TY * TZ;
  const long long ti = tile_index(L, gcx, gcy, gcz, TX, TY, TZ);
  const long long tile = ti / CT, c = ti - tile * CT;
  double acc = 0, wsum = 0;
  // cellwise L2 projection onto constants: JxW-weighted mean (box: weights w_q)
  const int G = L.dim * L.dim + 1;
  for (int q = 0; q < nq; ++q) {
    const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
    const double w = L.geo ? L.geo[((tile * nq + q) * G + G - 1) * CT + c]
                           : c_basis[kidx].w[qx] * c_basis[kidx].w[qy] * (L.dim == 3 ? c_basis[kidx].w[qz] : 1.0);
    acc += w * (double)in[(tile * nq + q) * CT + c];
    wsum += w;
  }
  out[ti] = (T)(acc / wsum);
}

// coarse cell = arithmetic mean of its 2^dim children (FE_DGQ(0) restriction)
template <typename T>
__global__ void k_eta_restrict(const LevelDesc Lc, const LevelDesc Lf, const int TX, const int TY, const int TZ,
                               const T *__restrict__ fine, T *__restrict__ coarse) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cell >= Lc.cells) return;
  const int gcx = (int)(cell % Lc.n[0]), gcy = (int)((cell / Lc.n[0]) % Lc.n[1]),
            gcz = (int)(cell / ((long long)Lc.n[0] * Lc.n[1]));
  double acc = 0;
  int cnt = 0;
  for (int dz = 0; dz < (Lc.dim == 3 ? 2 : 1); ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        acc += (double)fine[tile_index(Lf, 2 * gcx + dx, 2 * gcy + dy, Lc.dim == 3 ? 2 * gcz + dz : 0, TX, TY, TZ)];
        ++cnt;
      }
  coarse[tile_index(Lc, gcx, gcy, gcz, TX, TY, TZ)] = (T)(acc / cnt);
}

static void tiles_of(const LevelDesc &L, int t[3]) { tile_cells(L.dim, L.k, t); }

// DGQ1 ("project to Q1 only viscosity") tables.  Cellwise L2 projection of the
// per-qp values onto Q1: with the dual basis psi_0 = 4 - 6x, psi_1 = 6x - 2 of
// Q1 on [0,1], vertex coefficient v = sum_q w_q eta_q prod_d psi_{v_d}(x_q,d)
// (Utilities::project_cellwise, utilities.cc:2738-2800).  Coefficients are in
// natural cell order, [cell][2^dim] with x fastest.
__global__ void k_eta_q1_project(const LevelDesc L, const double *__restrict__ eta_nat, double *__restrict__ coef) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cell >= L.cells) return;
  const BasisConst &B = c_basis[L.k - 2];
  const int Q = L.k + 1, nq = L.dim == 3 ? Q * Q * Q : Q * Q, nv = 1 << L.dim;
  double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < nq; ++q) {
    const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
    double wq = eta_nat[cell * nq + q];
    for (int d = 0; d < L.dim; ++d) wq *= B.w[qi[d]];
    for (int v = 0; v < nv; ++v) {
      double t = wq;
      for (int d = 0; d < L.dim; ++d) {
        const double x = B.x[qi[d]];
        t *= ((v >> d) & 1) ? 6.0 * x - 2.0 : 4.0 - 6.0 * x;
      }
      c[v] += t;
    }
  }
  for (int v = 0; v < nv; ++v) coef[cell * nv + v] = c[v];
}
// parent coefficients = L2 projection of the children's piecewise-Q1 function:
// per direction child k contributes R_k[a][j], R_0 = [[3/4, 1/2], [-1/4, 0]],
// R_1 = [[0, -1/4], [1/2, 3/4]] (parent vertex a, child vertex j)
__global__ void k_eta_q1_restrict(const LevelDesc Lc, const LevelDesc Lf, const double *__restrict__ fine,
                                  double *__restrict__ coarse) {
  const long long cell = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (cell >= Lc.cells) return;
  const double R[2][2][2] = {{{0.75, 0.5}, {-0.25, 0.0}}, {{0.0, -0.25}, {0.5, 0.75}}};
  const int dim = Lc.dim, nv = 1 << dim;
  const int gcx = (int)(cell % Lc.n[0]), gcy = (int)((cell / Lc.n[0]) % Lc.n[1]),
            gcz = (int)(cell / ((long long)Lc.n[0] * Lc.n[1]));
  double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int ch = 0; ch < nv; ++ch) {
    const int k[3] = {ch & 1, (ch >> 1) & 1, (ch >> 2) & 1};
    const long long fc = (2LL * gcx + k[0]) +
                         Lf.n[0] * ((2LL * gcy + k[1]) + (long long)Lf.n[1] * (dim == 3 ? 2LL * gcz + k[2] : 0));
    for (int a = 0; a < nv; ++a)
      for (int j = 0; j < nv; ++j) {
        double w = 1.0;
        for (int d = 0; d < dim; ++d) w *= R[k[d]][(a >> d) & 1][(j >> d) & 1];
        c[a] += w * fine[fc * nv + j];
      }
  }
  for (int a = 0; a < nv; ++a) coarse[cell * nv + a] = c[a];
}
// Q1 values at the quadrature points, clamped to [lo, hi], tile-blocked per-qp table
template <typename T>
__global__ void k_eta_q1_eval(const LevelDesc L, const double *__restrict__ coef, const double lo, const double hi,
                              const int TX, const int TY, const int TZ, T *__restrict__ out) {
  const int Q = L.k + 1, nq = L.dim == 3 ? Q * Q * Q : Q * Q, nv = 1 << L.dim;
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= L.cells * nq) return;
  const long long cell = idx / nq;
  const int q = (int)(idx - cell * nq);
  const BasisConst &B = c_basis[L.k - 2];
  const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
  double val = 0;
  for (int v = 0; v < nv; ++v) {
    double t = coef[cell * nv + v];
    for (int d = 0; d < L.dim; ++d) {
      const double x = B.x[qi[d]];
      t *= ((v >> d) & 1) ? x : 1.0 - x;
    }
    val += t;
  }
  val = fmin(fmax(val, lo), hi);
  const int gcx = (int)(cell % L.n[0]), gcy = (int)((cell / L.n[0]) % L.n[1]),
            gcz = (int)(cell / ((long long)L.n[0] * L.n[1]));
  const long long ti = tile_index(L, gcx, gcy, gcz, TX, TY, TZ);
  const long long CT = TX * TY * TZ;
  const long long tile = ti / CT, c = ti - tile * CT;
  out[(tile * nq + q) * CT + c] = (T)val;
}



template <typename T>
void launch_eta_to_tiles(const LevelDesc &L, const double *eta_nat, int per_qp, T *eta_tiles, hipStream_t st) {
  int t[3];
  tiles_of(L, t);
  const int nq = L.dim == 3 ? (L.k + 1) * (L.k + 1) * (L.k + 1) : (L.k + 1) * (L.k + 1);
  const long long n = L.cells * (per_qp ? nq : 1);
  hipLaunchKernelGGL((k_eta_to_tiles<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, eta_nat, per_qp,
                     nq, t[0], t[1], t[2], eta_tiles);
  MFS_HIP(hipGetLastError());
}
template <typename T>
void launch_eta_average(const LevelDesc &L, const double *eta_nat_qp, int mode, T *out, hipStream_t st) {
  int t[3];
  tiles_of(L, t);
  const int Q = L.k + 1;
  const int nq = L.dim == 3 ? Q * Q * Q : Q * Q;
  hipLaunchKernelGGL((k_eta_average<T>), dim3((unsigned)((L.cells + 255) / 256)), dim3(256), 0, st, L, eta_nat_qp,
                     nq, Q, mode, t[0], t[1], t[2], out);
  MFS_HIP(hipGetLastError());
}
template <typename T, typename S>
void launch_eta_qp_to_cell(const LevelDesc &L, const S *in, T *out, hipStream_t st) {
  int t[3];
  tiles_of(L, t);
  const int Q = L.k + 1;
  const int nq = L.dim == 3 ? Q * Q * Q : Q * Q;
  hipLaunchKernelGGL((k_eta_qp_to_cell<T, S>), dim3((unsigned)((L.cells + 255) / 256)), dim3(256), 0, st, L, in, Q,
                     nq, t[0], t[1], t[2], out);
  MFS_HIP(hipGetLastError());
}
template <typename T>
void launch_eta_restrict(const LevelDesc &Lc, const LevelDesc &Lf, const T *fine, T *coarse, hipStream_t st) {
  int t[3];
  tiles_of(Lc, t);
  hipLaunchKernelGGL((k_eta_restrict<T>), dim3((unsigned)((Lc.cells + 255) / 256)), dim3(256), 0, st, Lc, Lf, t[0],
                     t[1], t[2], fine, coarse);
  MFS_HIP(hipGetLastError());
}
void launch_eta_q1_project(const LevelDesc &L, const double *eta_nat_qp, double *coef, hipStream_t st) {
  hipLaunchKernelGGL(k_eta_q1_project, dim3((unsigned)((L.cells + 255) / 256)), dim3(256), 0, st, L, eta_nat_qp, coef);
  MFS_HIP(hipGetLastError());
}
void launch_eta_q1_restrict(const LevelDesc &Lc, const LevelDesc &Lf, const double *fine, double *coarse,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_eta_q1_restrict, dim3((unsigned)((Lc.cells + 255) / 256)), dim3(256), 0, st, Lc, Lf, fine,
                     coarse);
  MFS_HIP(hipGetLastError());
}
template <typename T>
void launch_eta_q1_eval(const LevelDesc &L, const double *coef, double lo, double hi, T *out, hipStream_t st) {
  int t[3];
  tiles_of(L, t);
  const int Q = L.k + 1, nq = L.dim == 3 ? Q * Q * Q : Q * Q;
  const long long n = L.cells * nq;
  hipLaunchKernelGGL((k_eta_q1_eval<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, coef, lo, hi, t[0],
                     t[1], t[2], out);
  MFS_HIP(hipGetLastError());
}
template void launch_eta_q1_eval<double>(const LevelDesc &, const double *, double, double, double *, hipStream_t);
template void launch_eta_q1_eval<float>(const LevelDesc &, const double *, double, double, float *, hipStream_t);
template void launch_eta_to_tiles<double>(const LevelDesc &, const double *, int, double *, hipStream_t);
template void launch_eta_to_tiles<float>(const LevelDesc &, const double *, int, float *, hipStream_t);
template void launch_eta_average<double>(const LevelDesc &, const double *, int, double *, hipStream_t);
template void launch_eta_average<float>(const LevelDesc &, const double *, int, float *, hipStream_t);
template void launch_eta_qp_to_cell<double, double>(const LevelDesc &, const double *, double *, hipStream_t);
template void launch_eta_qp_to_cell<float, double>(const LevelDesc &, const double *, float *, hipStream_t);
template void launch_eta_restrict<double>(const LevelDesc &, const LevelDesc &, const double *, double *,
                                          hipStream_t);
template void launch_eta_restrict<float>(const LevelDesc &, const LevelDesc &, const float *, float *, hipStream_t);

// ------------------------------------------------------------------ vector kernels
static inline unsigned nblocks(long long n, int bs = 256) {
  long long b = (n + bs - 1) / bs;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (unsigned)(b < 1 ? 1 : b);
}
#define GRID_STRIDE(i, n) \
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

template <typename T>
__global__ void k_fill(T *x, T v, long long n) {
  GRID_STRIDE(i, n) x[i] = v;
}
template <typename A, typename B>
__global__ void k_cast(const A *__restrict__ x, B *__restrict__ y, long long n) {
  GRID_STRIDE(i, n) y[i] = (B)x[i];
}
template <typename T>
__global__ void k_axpby(T a, const T *__restrict__ x, T b, T *__restrict__ y, long long n) {
  GRID_STRIDE(i, n) y[i] = a * x[i] + b * y[i];
}
template <typename T>
__global__ void k_invert(T *x, long long n) {
  GRID_STRIDE(i, n) x[i] = (T)1 / x[i];
}
template <typename T>
__global__ void k_cheb_first(const T *__restrict__ dinv, const T *__restrict__ r, T *__restrict__ d,
                             T *__restrict__ x, T inv_theta, int x_zero, long long n) {
  GRID_STRIDE(i, n) {
    const T dd = dinv[i] * r[i] * inv_theta;
    d[i] = dd;
    x[i] = x_zero ? dd : x[i] + dd;
  }
}
template <typename T>
__global__ void k_cheb_step(const T *__restrict__ dinv, const T *__restrict__ r, T *__restrict__ d,
                            T *__restrict__ x, T f1, T f2, long long n) {
  GRID_STRIDE(i, n) {
    const T dd = f1 * d[i] + f2 * dinv[i] * r[i];
    d[i] = dd;
    x[i] += dd;
  }
}
template <typename T>
__global__ void k_neg_copy(const T *__restrict__ x, T *__restrict__ y, long long n) {
  GRID_STRIDE(i, n) y[i] = -x[i];
}

template <typename T>
void v_fill(T *x, T v, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_fill<T>), dim3(nblocks(n)), dim3(256), 0, st, x, v, n);
}
template <typename A, typename B>
void v_cast(const A *x, B *y, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<A, B>), dim3(nblocks(n)), dim3(256), 0, st, x, y, n);
}
template <typename T>
void v_axpby(T a, const T *x, T b, T *y, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_axpby<T>), dim3(nblocks(n)), dim3(256), 0, st, a, x, b, y, n);
}
template <typename T>
void v_invert(T *x, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_invert<T>), dim3(nblocks(n)), dim3(256), 0, st, x, n);
}
template <typename T>
void v_cheb_first(const T *dinv, const T *r, T *d, T *x, T inv_theta, bool x_zero, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cheb_first<T>), dim3(nblocks(n)), dim3(256), 0, st, dinv, r, d, x, inv_theta,
                     x_zero ? 1 : 0, n);
}
template <typename T>
void v_cheb_step(const T *dinv, const T *r, T *d, T *x, T f1, T f2, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cheb_step<T>), dim3(nblocks(n)), dim3(256), 0, st, dinv, r, d, x, f1, f2, n);
}
template <typename T>
void v_scale_neg_copy(const T *x, T *y, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_neg_copy<T>), dim3(nblocks(n)), dim3(256), 0, st, x, y, n);
}

#define INST_T(T)                                                                  \
  template void v_fill<T>(T *, T, long long, hipStream_t);                         \
  template void v_axpby<T>(T, const T *, T, T *, long long, hipStream_t);          \
  template void v_invert<T>(T *, long long, hipStream_t);                          \
  template void v_cheb_first<T>(const T *, const T *, T *, T *, T, bool, long long, hipStream_t); \
  template void v_cheb_step<T>(const T *, const T *, T *, T *, T, T, long long, hipStream_t);     \
  template void v_scale_neg_copy<T>(const T *, T *, long long, hipStream_t);
INST_T(double)
INST_T(float)
template void v_cast<double, float>(const double *, float *, long long, hipStream_t);
template void v_cast<float, double>(const float *, double *, long long, hipStream_t);
template void v_cast<double, double>(const double *, double *, long long, hipStream_t);
template void v_cast<float, float>(const float *, float *, long long, hipStream_t);

// IDR(s) shadow-space vector k: uniform in [-1, 1) from a counter hash
// (splitmix64 finalizer) of (k, i); the oracle draws the same bits.
__global__ void k_idr_shadow(double *__restrict__ q, const int k, const long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long z = (unsigned long long)(k + 1) * 0x9E3779B97F4A7C15ull + (unsigned long long)i * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  q[i] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
}
void v_idr_shadow(double *q, int k, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_idr_shadow, dim3(nblocks(n)), dim3(256), 0, st, q, k, n);
  MFS_HIP(hipGetLastError());
}

// ---- deterministic reductions: fixed grid of DOT_BLOCKS blocks, fixed-order tree
constexpr int DOT_BLOCKS = 2048;  // 8 workgroups per CU: enough loads in flight to stream HBM
constexpr int DOT_THREADS = 256;
int dot_work_size() { return DOT_BLOCKS; }

__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = DOT_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

template <typename T>
__global__ void __launch_bounds__(DOT_THREADS) k_dot_partial(const T *__restrict__ x, const T *__restrict__ y,
                                                             long long n, double *__restrict__ part) {
  __shared__ double sh[DOT_THREADS];
  double acc = 0;
  GRID_STRIDE(i, n) acc += (double)x[i] * (double)y[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(DOT_THREADS) k_sum_partials(const double *__restrict__ part, int m, int nblk,
                                                              double *__restrict__ out) {
  __shared__ double sh[DOT_THREADS];
  // one block per output value
  const int v = blockIdx.x;
  double acc = 0;
  for (int i = threadIdx.x; i < nblk; i += DOT_THREADS) acc += part[(long long)v * nblk + i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[v] = s;
  (void)m;
}

template <typename T>
void v_dot(const T *x, const T *y, long long n, double *work, double *out, hipStream_t st) {
  hipLaunchKernelGGL((k_dot_partial<T>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, st, x, y, n, work);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(DOT_THREADS), 0, st, work, 1, DOT_BLOCKS, out);
}
template void v_dot<double>(const double *, const double *, long long, double *, double *, hipStream_t);
template void v_dot<float>(const float *, const float *, long long, double *, double *, hipStream_t);

void v_sum_partials(const double *work, int m, double *h, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_partials, dim3(m), dim3(DOT_THREADS), 0, st, work, m, DOT_BLOCKS, h);
}
// Fused Gram-Schmidt kernels over up to MV_MAX Krylov vectors held in chunks of
// 8 (vector i = p[i / 8] + (i % 8) * ld).  k_mdot_all: partial dots of every
// vector with w, 16 vectors per sweep over w, wave-shuffle reductions in a
// fixed order; k_maxpy_all: w += sign * sum_i h_i V_i, one read of each
// vector and one read-modify-write of w.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__global__ void __launch_bounds__(DOT_THREADS) k_mdot_all(VecSet V, long long ld, int m, const double *__restrict__ w,
                                                          long long n, double *__restrict__ part) {
  constexpr int G = 16;
  __shared__ double sh[G][DOT_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int g0 = 0; g0 < m; g0 += G) {
    double acc[G];
#pragma unroll
    for (int k = 0; k < G; ++k) acc[k] = 0;
    GRID_STRIDE(i, n) {
      const double wi = w[i];
#pragma unroll
      for (int k = 0; k < G; ++k)
        if (g0 + k < m) acc[k] += V.p[(g0 + k) >> 3][((g0 + k) & 7) * ld + i] * wi;
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const double r = wave_sum(acc[k]);
      if (lane == 0) sh[k][wv] = r;
    }
    __syncthreads();
    if (threadIdx.x < G && g0 + (int)threadIdx.x < m) {
      double t = 0;
#pragma unroll
      for (int q = 0; q < DOT_THREADS / 64; ++q) t += sh[threadIdx.x][q];
      part[(long long)(g0 + threadIdx.x) * DOT_BLOCKS + blockIdx.x] = t;
    }
    __syncthreads();
  }
}
__global__ void k_maxpy_all(VecSet V, long long ld, int m, const double *__restrict__ h, double *__restrict__ w,
                            long long n, double sign) {
  __shared__ double hs[MV_MAX];
  for (int k = threadIdx.x; k < m; k += blockDim.x) hs[k] = h[k];
  __syncthreads();
  GRID_STRIDE(i, n) {
    double acc = 0;
    for (int k = 0; k < m; ++k) acc += hs[k] * V.p[k >> 3][(k & 7) * ld + i];
    w[i] += sign * acc;
  }
}
// Delayed classical Gram-Schmidt (one reduction and one update sweep per Arnoldi
// step).  Both kernels give every wave chunks of DCGS_E * 64 consecutive entries
// (lane l owns entries base + e * 64 + l, e < DCGS_E), chunk c = wave, wave +
// waves, ...: the chunk's entries of v and w stay in registers while the wave
// streams every basis vector over the chunk, so each vector is read once per
// sweep with eight 512-byte loads in flight per wave.  The basis vectors are
// streamed with nontemporal loads (and the update's results written with
// nontemporal stores): nothing re-reads them before the next sweep, and
// keeping them out of the caches makes both kernels 4-10% faster.
constexpr int DCGS_E = 8;
constexpr int DCGS_BLOCKS = 1024;  // 4 workgroups of 4 waves per CU (86 VGPRs)
__device__ __forceinline__ double wave_allsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Reduction rows s_k = V_k . v and t_k = V_k . w for k < j, then v . v, v . w,
// w . w.  Per chunk and k the wave sums its lanes' partial products (butterfly,
// fixed order) and lane k % 64 accumulates the total, so a wave keeps all 2 j
// rows in four registers per lane; the four waves of a workgroup are added in
// a fixed order into part[row * DCGS_BLOCKS + block].
__global__ void __launch_bounds__(256) k_dcgs_reduce(VecSet V, long long ld, int j, const double *__restrict__ v,
                                                     const double *__restrict__ w, long long n,
                                                     double *__restrict__ part) {
  constexpr int E = DCGS_E;
  __shared__ double sh[4][2 * MV_MAX + 3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long CH = 64LL * E;
  const long long nch = (n + CH - 1) / CH;
  const long long waves = (long long)gridDim.x * 4;
  double aS0 = 0, aS1 = 0, aT0 = 0, aT1 = 0, vv = 0, vw = 0, ww = 0;
  for (long long c = blockIdx.x * 4LL + wv; c < nch; c += waves) {
    const long long base = c * CH + lane;
    const bool full = (c + 1) * CH <= n;
    double vr[E], wr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const long long idx = base + e * 64;
      const bool ok = full || idx < n;
      vr[e] = ok ? v[idx] : 0.0;
      wr[e] = ok ? w[idx] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      vv += vr[e] * vr[e];
      vw += vr[e] * wr[e];
      ww += wr[e] * wr[e];
    }
    for (int k = 0; k < j; ++k) {
      const double *x = V.p[k >> 3] + (long long)(k & 7) * ld + base;
      double s = 0, t = 0;
      if (full) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double xe = __builtin_nontemporal_load(x + e * 64);
          s += xe * vr[e];
          t += xe * wr[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double xe = base + e * 64 < n ? x[e * 64] : 0.0;
          s += xe * vr[e];
          t += xe * wr[e];
        }
      }
      s = wave_allsum(s);
      t = wave_allsum(t);
      if (lane == (k & 63)) {
        if (k < 64) {
          aS0 += s;
          aT0 += t;
        } else {
          aS1 += s;
          aT1 += t;
        }
      }
    }
  }
  if (lane < j) {
    sh[wv][lane] = aS0;
    sh[wv][j + lane] = aT0;
  }
  if (lane + 64 < j) {
    sh[wv][lane + 64] = aS1;
    sh[wv][j + lane + 64] = aT1;
  }
  vv = wave_allsum(vv);
  vw = wave_allsum(vw);
  ww = wave_allsum(ww);
  if (lane == 0) {
    sh[wv][2 * j] = vv;
    sh[wv][2 * j + 1] = vw;
    sh[wv][2 * j + 2] = ww;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < 2 * j + 3; r += 256)
    part[(long long)r * DCGS_BLOCKS + blockIdx.x] = ((sh[0][r] + sh[1][r]) + sh[2][r]) + sh[3][r];
}
// v <- c1 v + sum_k a1_k V_k;  w <- c2 w + c3 v_old + sum_k a2_k V_k; coef = [a1 (j), a2 (j), c1, c2, c3]
__global__ void __launch_bounds__(256) k_dcgs_update(VecSet V, long long ld, int j, const double *__restrict__ coef,
                                                     double *__restrict__ v, double *__restrict__ w, long long n) {
  constexpr int E = DCGS_E;
  __shared__ double cf[2 * MV_MAX + 3];
  for (int k = threadIdx.x; k < 2 * j + 3; k += blockDim.x) cf[k] = coef[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double c1 = cf[2 * j], c2 = cf[2 * j + 1], c3 = cf[2 * j + 2];
  const long long CH = 64LL * E;
  const long long nch = (n + CH - 1) / CH;
  const long long waves = (long long)gridDim.x * 4;
  for (long long c = blockIdx.x * 4LL + wv; c < nch; c += waves) {
    const long long base = c * CH + lane;
    const bool full = (c + 1) * CH <= n;
    double vr[E], wr[E], s1[E], s2[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const long long idx = base + e * 64;
      const bool ok = full || idx < n;
      vr[e] = ok ? v[idx] : 0.0;
      wr[e] = ok ? w[idx] : 0.0;
      s1[e] = s2[e] = 0;
    }
    if (full) {
#pragma unroll 2
      for (int k = 0; k < j; ++k) {
        const double *x = V.p[k >> 3] + (long long)(k & 7) * ld + base;
        const double a1 = cf[k], a2 = cf[j + k];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double xe = __builtin_nontemporal_load(x + e * 64);
          s1[e] += a1 * xe;
          s2[e] += a2 * xe;
        }
      }
    } else {
      for (int k = 0; k < j; ++k) {
        const double *x = V.p[k >> 3] + (long long)(k & 7) * ld + base;
        const double a1 = cf[k], a2 = cf[j + k];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double xe = base + e * 64 < n ? x[e * 64] : 0.0;
          s1[e] += a1 * xe;
          s2[e] += a2 * xe;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const long long idx = base + e * 64;
      if (full || idx < n) {
        __builtin_nontemporal_store(c1 * vr[e] + s1[e], v + idx);
        __builtin_nontemporal_store(c2 * wr[e] + c3 * vr[e] + s2[e], w + idx);
      }
    }
  }
}
void v_dcgs_reduce(const VecSet &V, long long ld, int j, const double *v, const double *w, long long n, double *work,
                   double *rows, hipStream_t st) {
  if (j < 0 || j > MV_MAX) throw std::runtime_error("v_dcgs_reduce: too many vectors");
  hipLaunchKernelGGL(k_dcgs_reduce, dim3(DCGS_BLOCKS), dim3(256), 0, st, V, ld, j, v, w, n, work);
  hipLaunchKernelGGL(k_sum_partials, dim3(2 * j + 3), dim3(DOT_THREADS), 0, st, work, 2 * j + 3, DCGS_BLOCKS, rows);
  MFS_HIP(hipGetLastError());
}
void v_dcgs_update(const VecSet &V, long long ld, int j, const double *coef, double *v, double *w, long long n,
                   hipStream_t st) {
  if (j < 0 || j > MV_MAX) throw std::runtime_error("v_dcgs_update: too many vectors");
  hipLaunchKernelGGL(k_dcgs_update, dim3(DCGS_BLOCKS), dim3(256), 0, st, V, ld, j, coef, v, w, n);
  MFS_HIP(hipGetLastError());
}
void v_mdot_all(const VecSet &V, long long ld, int m, const double *w, long long n, double *work, hipStream_t st) {
  if (m > MV_MAX) throw std::runtime_error("v_mdot_all: too many vectors");
  hipLaunchKernelGGL(k_mdot_all, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, st, V, ld, m, w, n, work);
  MFS_HIP(hipGetLastError());
}
void v_maxpy_all(const VecSet &V, long long ld, int m, const double *h, double *w, long long n, double sign,
                 hipStream_t st) {
  if (m > MV_MAX) throw std::runtime_error("v_maxpy_all: too many vectors");
  hipLaunchKernelGGL(k_maxpy_all, dim3(std::min<long long>(nblocks(n), 8192)), dim3(256), 0, st, V, ld, m, h, w, n,
                     sign);
  MFS_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_mul(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ c, long long n) {
  GRID_STRIDE(i, n) c[i] = a[i] * b[i];
}
template <typename T>
void v_mul(const T *a, const T *b, T *c, long long n, hipStream_t st) {
  hipLaunchKernelGGL((k_mul<T>), dim3(nblocks(n)), dim3(256), 0, st, a, b, c, n);
}
template void v_mul<double>(const double *, const double *, double *, long long, hipStream_t);
template void v_mul<float>(const float *, const float *, float *, long long, hipStream_t);
__global__ void k_scale(double *x, double a, long long n) {
  GRID_STRIDE(i, n) x[i] *= a;
}
void v_scale(double *x, double a, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(256), 0, st, x, a, n);
}
__global__ void k_scaled_copy_dev(const double *__restrict__ x, const double *__restrict__ a, double *__restrict__ y,
                                  long long n) {
  const double inv = 1.0 / a[0];
  GRID_STRIDE(i, n) y[i] = x[i] * inv;
}
void v_scaled_copy_dev(const double *x, const double *a, double *y, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_copy_dev, dim3(nblocks(n)), dim3(256), 0, st, x, a, y, n);
}

template <typename T>
__global__ void k_zero_con(const LevelDesc L, T *u) {
  const long long n = L.vnodes;
  GRID_STRIDE(i, n) {
    const int g[3] = {(int)(i % L.nv[0]), (int)((i / L.nv[0]) % L.nv[1]), (int)(i / ((long long)L.nv[0] * L.nv[1]))};
    for (int comp = 0; comp < L.dim; ++comp)
      if (is_con(L, comp, g[0], g[1], g[2])) u[comp * n + i] = 0;
  }
}
template <typename T>
void v_zero_constrained(const LevelDesc &L, T *u, hipStream_t st) {
  hipLaunchKernelGGL((k_zero_con<T>), dim3(nblocks(L.vnodes)), dim3(256), 0, st, L, u);
}
template void v_zero_constrained<double>(const LevelDesc &, double *, hipStream_t);
template void v_zero_constrained<float>(const LevelDesc &, float *, hipStream_t);

}  // namespace mfs
