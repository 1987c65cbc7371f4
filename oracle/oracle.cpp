// ORACLE — test infrastructure only.
//
// CPU FP64 restatement of ASPECT's matrix-free Stokes solve on box meshes.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library; the product (aspect_amd/) never links or calls it.
//
// What it restates (reference citations are relative to the ASPECT tree):
//   * StokesOperator::local_apply       source/simulator/solver/matrix_free_operators.cc:376-547
//   * BTBlockOperator::local_apply      matrix_free_operators.cc:668-701
//   * MassMatrixOperator::inner_cell_operation / compute_diagonal   :808-845, :861-908
//   * ABlockOperator::inner_cell_operation / compute_diagonal       :974-1009, :1068-1098
//   * MatrixFreeOperators::Base::vmult semantics (deal.II, not vendored): constrained
//     source entries read as zero, constrained rows are identity rows.
//   * level viscosity (DGQ0 projection, children averaged)  stokes_matrix_free_local_smoothing.cc:211-294
//   * smoother set-up, V-cycle, block preconditioner, GMRES -> FGMRES sequence
//                                       stokes_matrix_free_local_smoothing.cc:758-1446,
//                                       include/aspect/simulator/solver/block_stokes_preconditioner.h:43-352
//   * buoyancy right-hand side          source/simulator/assemblers/stokes.cc:393-396
//   * weak form (matrix-based restatement used by the brute-force path)  assemblers/stokes.cc:431-441
//
// Two independent evaluation paths live here: a brute-force dense element
// matrix path (orc_cell_matrices) and a sum-factorized path (orc_*_vmult).
// The tests pin one against the other and both against closed forms.
//
// Vector layout (shared with the product's C-ABI): velocity block is
// component-major, nodes lexicographic with x fastest; pressure block follows.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
typedef struct {
  int dim;              // 2 or 3
  int degree;           // velocity degree k (2 or 3); pressure degree k-1
  int coarse[3];        // cells per direction on level 0
  int n_levels;         // finest level = n_levels-1, cells = coarse << (n_levels-1)
  double extent[3];     // box extents (origin at 0)
  int bc[6];            // per face x-,x+,y-,y+,z-,z+: 0 open, 1 no normal flux, 2 zero velocity
  int compressible;     // 1: subtract (2/3) eta div u I
  double pressure_scaling;
} orc_problem;

typedef struct {
  double tolerance;          // "Linear solver tolerance"
  int n_cheap_steps;         // "Number of cheap Stokes solver steps"
  int n_expensive_steps;     // "Maximum number of expensive Stokes solver steps"
  int restart;               // "GMRES solver restart length"
  double a_block_tolerance;  // "Linear solver A block tolerance"
  double s_block_tolerance;  // "Linear solver S block tolerance"
  int level_eta_mode;        // 0: DGQ0 (cell means); 1: DGQ1 ("project to Q1 only viscosity")
  int cheap_krylov;          // "Krylov method for cheap solver steps": 0 GMRES, 1 IDR(s)
  int idr_s;                 // "IDR(s) parameter"
  int solve_newton_system;   // zero initial guess, tolerance from ||rhs||, no pressure normalization
                             // (stokes_matrix_free_local_smoothing.cc:950-969, 1028-1041, 1443-1444)
} orc_solve_params;

typedef struct {
  int cheap_iterations;
  int expensive_iterations;
  int a_preconditioner_applications;
  int s_preconditioner_applications;
  int converged;
  double initial_nonlinear_residual;
  double final_linear_residual;
  double solver_tolerance;
  double r0_norm;
  double pressure_normalization_adjustment;
} orc_solve_result;
}

namespace {

// ---------------------------------------------------------------- 1D rules
// Gauss-Legendre points/weights on [0,1] (Newton on P_n).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      double pn = (n == 1) ? z : p1, pnm1 = (n == 1) ? 1.0 : p0;
      double dp = n * (z * pn - pnm1) / (z * z - 1.0);
      double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    double pn = (n == 1) ? z : p1, pnm1 = (n == 1) ? 1.0 : p0;
    double dp = n * (z * pn - pnm1) / (z * z - 1.0);
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);  // 2/((1-z^2)P'^2) scaled by 1/2
  }
}

// Gauss-Lobatto support points of FE_Q(k) on [0,1]: endpoints plus roots of P'_k.
std::vector<double> lobatto_points(int k) {
  std::vector<double> x(k + 1);
  x[0] = 0.0;
  x[k] = 1.0;
  for (int i = 1; i < k; ++i) {
    double z = -std::cos(M_PI * i / k);
    for (int it = 0; it < 100; ++it) {
      // P'_k(z) and P''_k(z) via Legendre recursion
      double p0 = 1.0, p1 = z;
      for (int m = 2; m <= k; ++m) {
        double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      double pk = p1, pkm1 = p0;
      double d1 = k * (z * pk - pkm1) / (z * z - 1.0);
      double d2 = (2.0 * z * d1 - k * (k + 1.0) * pk) / (1.0 - z * z);
      double dz = d1 / d2;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
  return x;
}

struct Lagrange1D {
  std::vector<double> pts;
  double value(int i, double x) const {
    double v = 1.0;
    for (size_t j = 0; j < pts.size(); ++j)
      if ((int)j != i) v *= (x - pts[j]) / (pts[i] - pts[j]);
    return v;
  }
  double deriv(int i, double x) const {
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
};

// ---------------------------------------------------------------- mesh levels
struct Level {
  int dim = 3, k = 2;
  int n[3] = {1, 1, 1};    // cells per direction
  double h[3] = {1, 1, 1};
  int nv[3] = {1, 1, 1};   // velocity nodes per direction (k*n+1)
  int np[3] = {1, 1, 1};   // pressure nodes per direction ((k-1)*n+1)
  long long cells = 0, vnodes = 0, pnodes = 0;
  std::vector<uint8_t> vcon;  // dim*vnodes flags: constrained velocity DoF
  // general geometry (empty: the uniform box): per cell, per quadrature point,
  // J^{-1} row-major (d xi_e / d x_d at [e * dim + d]) and JxW (with the weight)
  std::vector<double> geo;
  int G() const { return dim * dim + 1; }
  long long nu() const { return (long long)dim * vnodes; }
  long long np_() const { return pnodes; }
};

// per-level geometry tables registered by orc_set_geometry (test infrastructure)
std::vector<std::vector<double>> g_geometry;

Level make_level(const orc_problem &p, int l) {
  Level L;
  L.dim = p.dim;
  L.k = p.degree;
  for (int d = 0; d < 3; ++d) {
    if (d < p.dim) {
      L.n[d] = p.coarse[d] << l;
      L.h[d] = p.extent[d] / L.n[d];
      L.nv[d] = L.k * L.n[d] + 1;
      L.np[d] = (L.k - 1) * L.n[d] + 1;
    } else {
      L.n[d] = 1;
      L.h[d] = 1.0;
      L.nv[d] = 1;
      L.np[d] = 1;
    }
  }
  L.cells = (long long)L.n[0] * L.n[1] * L.n[2];
  if (l < (int)g_geometry.size() && !g_geometry[l].empty()) {
    const int nq1 = p.degree + 1;
    const long long need = L.cells * (p.dim == 3 ? nq1 * nq1 * nq1 : nq1 * nq1) * (p.dim * p.dim + 1);
    if ((long long)g_geometry[l].size() != need) throw std::runtime_error("geometry table size mismatch");
    L.geo = g_geometry[l];
  }
  L.vnodes = (long long)L.nv[0] * L.nv[1] * L.nv[2];
  L.pnodes = (long long)L.np[0] * L.np[1] * L.np[2];
  L.vcon.assign((size_t)(L.dim * L.vnodes), 0);
  for (long long nidx = 0; nidx < L.vnodes; ++nidx) {
    int ix[3] = {(int)(nidx % L.nv[0]), (int)((nidx / L.nv[0]) % L.nv[1]),
                 (int)(nidx / ((long long)L.nv[0] * L.nv[1]))};
    for (int d = 0; d < L.dim; ++d) {
      for (int side = 0; side < 2; ++side) {
        bool on = side == 0 ? ix[d] == 0 : ix[d] == L.nv[d] - 1;
        if (!on) continue;
        int b = p.bc[2 * d + side];
        if (b == 2)
          for (int c = 0; c < L.dim; ++c) L.vcon[(size_t)c * L.vnodes + nidx] = 1;
        else if (b == 1)
          L.vcon[(size_t)d * L.vnodes + nidx] = 1;
      }
    }
  }
  return L;
}

// ---------------------------------------------------------------- reference element
struct RefElement {
  int dim, k, N, Q, NP;  // nodes 1D (velocity), qps 1D, pressure nodes 1D
  std::vector<double> qx, qw;
  Lagrange1D vb, pb;
  std::vector<double> S, G, SP, GP;  // Q x N, Q x N, Q x NP, Q x NP
  RefElement(int dim_, int k_) : dim(dim_), k(k_), N(k_ + 1), Q(k_ + 1), NP(k_) {
    gauss_legendre(Q, qx, qw);
    vb.pts = lobatto_points(k);
    pb.pts = lobatto_points(k - 1);
    S.resize(Q * N);
    G.resize(Q * N);
    SP.resize(Q * NP);
    GP.resize(Q * NP);
    for (int q = 0; q < Q; ++q) {
      for (int i = 0; i < N; ++i) {
        S[q * N + i] = vb.value(i, qx[q]);
        G[q * N + i] = vb.deriv(i, qx[q]);
      }
      for (int i = 0; i < NP; ++i) {
        SP[q * NP + i] = pb.value(i, qx[q]);
        GP[q * NP + i] = pb.deriv(i, qx[q]);
      }
    }
  }
  int nq() const { return dim == 2 ? Q * Q : Q * Q * Q; }
  int nvloc() const { return dim == 2 ? N * N : N * N * N; }
  int nploc() const { return dim == 2 ? NP * NP : NP * NP * NP; }
  double qweight(int q) const {
    double w = 1.0;
    for (int d = 0; d < dim; ++d) {
      w *= qw[q % Q];
      q /= Q;
    }
    return w;
  }
};

// global velocity node of local node a in cell c; pressure likewise
inline void cell_coords(const Level &L, long long c, int cc[3]) {
  cc[0] = (int)(c % L.n[0]);
  cc[1] = (int)((c / L.n[0]) % L.n[1]);
  cc[2] = (int)(c / ((long long)L.n[0] * L.n[1]));
}
// J^{-1} (row-major, e * dim + d) and JxW of cell c at quadrature point q:
// the general table, or the box values diag(1/h) and |K| w_q
inline double geometry(const Level &L, const RefElement &re, long long c, int q, double Jinv[9]) {
  const int dim = L.dim;
  if (!L.geo.empty()) {
    const double *g = &L.geo[((size_t)c * re.nq() + q) * L.G()];
    for (int i = 0; i < dim * dim; ++i) Jinv[i] = g[i];
    return g[dim * dim];
  }
  double vol = 1.0;
  for (int d = 0; d < dim; ++d) vol *= L.h[d];
  for (int e = 0; e < dim; ++e)
    for (int d = 0; d < dim; ++d) Jinv[e * dim + d] = e == d ? 1.0 / L.h[d] : 0.0;
  return vol * re.qweight(q);
}
// physical gradient from reference derivatives: g_d = sum_e Jinv[e][d] r_e
inline void to_physical(int dim, const double Jinv[9], const double *r, double *g) {
  for (int d = 0; d < dim; ++d) {
    double a = 0;
    for (int e = 0; e < dim; ++e) a += Jinv[e * dim + d] * r[e];
    g[d] = a;
  }
}

inline long long vnode(const Level &L, const int cc[3], int a, int N) {
  int i = a % N, j = (a / N) % N, kk = a / (N * N);
  if (L.dim == 2) kk = 0;
  long long gx = (long long)L.k * cc[0] + i, gy = (long long)L.k * cc[1] + j,
            gz = L.dim == 3 ? (long long)L.k * cc[2] + kk : 0;
  return gx + L.nv[0] * (gy + (long long)L.nv[1] * gz);
}
inline long long pnode(const Level &L, const int cc[3], int a, int NP) {
  int i = a % NP, j = (a / NP) % NP, kk = a / (NP * NP);
  if (L.dim == 2) kk = 0;
  long long gx = (long long)(L.k - 1) * cc[0] + i, gy = (long long)(L.k - 1) * cc[1] + j,
            gz = L.dim == 3 ? (long long)(L.k - 1) * cc[2] + kk : 0;
  return gx + L.np[0] * (gy + (long long)L.np[1] * gz);
}

// ---------------------------------------------------------------- sum factorization
// out[b] for tensor index b (m per dim) = sum_a prod_d M_d[b_d][a_d] in[a] (n per dim).
// M_d is m x n row-major.  Generic in dim (2 or 3), runtime sizes (small).
void tensor_apply(int dim, int m, int n, const double *const *M, const double *in, double *out,
                  double *work) {
  if (dim == 2) {
    // x then y
    double *t = work;  // [n(y)][m(x)]
    for (int j = 0; j < n; ++j)
      for (int bx = 0; bx < m; ++bx) {
        double s = 0;
        for (int a = 0; a < n; ++a) s += M[0][bx * n + a] * in[j * n + a];
        t[j * m + bx] = s;
      }
    for (int by = 0; by < m; ++by)
      for (int bx = 0; bx < m; ++bx) {
        double s = 0;
        for (int a = 0; a < n; ++a) s += M[1][by * n + a] * t[a * m + bx];
        out[by * m + bx] = s;
      }
  } else {
    double *t1 = work;             // [n z][n y][m x]
    double *t2 = work + n * n * m;  // [n z][m y][m x]
    for (int kz = 0; kz < n; ++kz)
      for (int j = 0; j < n; ++j)
        for (int bx = 0; bx < m; ++bx) {
          double s = 0;
          for (int a = 0; a < n; ++a) s += M[0][bx * n + a] * in[(kz * n + j) * n + a];
          t1[(kz * n + j) * m + bx] = s;
        }
    for (int kz = 0; kz < n; ++kz)
      for (int by = 0; by < m; ++by)
        for (int bx = 0; bx < m; ++bx) {
          double s = 0;
          for (int a = 0; a < n; ++a) s += M[1][by * n + a] * t1[(kz * n + a) * m + bx];
          t2[(kz * m + by) * m + bx] = s;
        }
    for (int bz = 0; bz < m; ++bz)
      for (int by = 0; by < m; ++by)
        for (int bx = 0; bx < m; ++bx) {
          double s = 0;
          for (int a = 0; a < n; ++a) s += M[2][bz * n + a] * t2[(a * m + by) * m + bx];
          out[(bz * m + by) * m + bx] = s;
        }
  }
}

struct CellKernels {
  RefElement re;
  std::vector<double> St, Gt, SPt;  // transposes (N x Q, NP x Q)
  explicit CellKernels(int dim, int k) : re(dim, k) {
    St.resize(re.N * re.Q);
    Gt.resize(re.N * re.Q);
    SPt.resize(re.NP * re.Q);
    for (int q = 0; q < re.Q; ++q) {
      for (int i = 0; i < re.N; ++i) {
        St[i * re.Q + q] = re.S[q * re.N + i];
        Gt[i * re.Q + q] = re.G[q * re.N + i];
      }
      for (int i = 0; i < re.NP; ++i) SPt[i * re.Q + q] = re.SP[q * re.NP + i];
    }
  }

  // gradient of a scalar nodal field at all qps: grad[d*nq + q]
  void grad(const double *u, const double h[3], double *grad, double *work) const {
    const int dim = re.dim, nq = re.nq();
    for (int d = 0; d < dim; ++d) {
      const double *M[3];
      for (int e = 0; e < dim; ++e) M[e] = (e == d) ? re.G.data() : re.S.data();
      tensor_apply(dim, re.Q, re.N, M, u, grad + d * nq, work);
      for (int q = 0; q < nq; ++q) grad[d * nq + q] /= h[d];
    }
  }
  // out[a] += sum_q sum_d f[d][q] dphi_a/dx_d(q)  (f already multiplied by JxW)
  void grad_T_add(const double *f, const double h[3], double *out, double *work) const {
    const int dim = re.dim, nq = re.nq(), nv = re.nvloc();
    double *tmp = work;            // nq
    double *res = work + nq;       // nv
    double *w2 = work + nq + nv;   // tensor_apply scratch
    for (int d = 0; d < dim; ++d) {
      const double *M[3];
      for (int e = 0; e < dim; ++e) M[e] = (e == d) ? Gt.data() : St.data();
      for (int q = 0; q < nq; ++q) tmp[q] = f[d * nq + q] / h[d];
      tensor_apply(dim, re.N, re.Q, M, tmp, res, w2);
      for (int a = 0; a < nv; ++a) out[a] += res[a];
    }
  }
  void pvalues(const double *p, double *out, double *work) const {
    const double *M[3] = {re.SP.data(), re.SP.data(), re.SP.data()};
    tensor_apply(re.dim, re.Q, re.NP, M, p, out, work);
  }
  void pvalues_T_add(const double *f, double *out, double *work) const {
    const double *M[3] = {SPt.data(), SPt.data(), SPt.data()};
    const int npl = re.nploc();
    double *res = work;
    tensor_apply(re.dim, re.NP, re.Q, M, f, res, work + npl);
    for (int a = 0; a < npl; ++a) out[a] += res[a];
  }
};

struct OpInput {
  const Level *L;
  const double *eta;   // per cell (n_eta=1) or per qp
  int eta_per_qp;
  double s;
  int compressible;
  int plain = 0;       // read constrained DoFs as stored (read_dof_values_plain)
};

// which ops to do in the cell loop
enum { OP_STOKES = 0, OP_ABLOCK = 1, OP_BT = 2, OP_MASS = 3, OP_LAPLACE = 4 };

// Generic cell loop with coloring over 2^dim parities; dst accumulates (constrained rows untouched).
void cell_loop(const OpInput &in, const CellKernels &K, int op, const double *src_u, const double *src_p,
               double *dst_u, double *dst_p) {
  const Level &L = *in.L;
  const RefElement &re = K.re;
  const int dim = L.dim, nq = re.nq(), nv = re.nvloc(), npl = re.nploc();
  double JxW0 = 1.0;
  for (int d = 0; d < dim; ++d) JxW0 *= L.h[d];
  const int ncolors = 1 << dim;
  static const double unit_h[3] = {1.0, 1.0, 1.0};
  for (int color = 0; color < ncolors; ++color) {
#pragma omp parallel
    {
      std::vector<double> u(dim * nv), gu(dim * dim * nq), f(dim * dim * nq), pl(npl), pq(nq), fp(nq),
          out_u(dim * nv), out_p(npl), work(4 * 64 * 64);
      std::vector<long long> vidx(nv), pidx(npl);
#pragma omp for schedule(static)
      for (long long c = 0; c < L.cells; ++c) {
        int cc[3];
        cell_coords(L, c, cc);
        int col = (cc[0] & 1) | ((cc[1] & 1) << 1) | ((dim == 3 ? (cc[2] & 1) : 0) << 2);
        if (col != color) continue;
        for (int a = 0; a < nv; ++a) vidx[a] = vnode(L, cc, a, re.N);
        for (int a = 0; a < npl; ++a) pidx[a] = pnode(L, cc, a, re.NP);
        const bool need_u = (op == OP_STOKES || op == OP_ABLOCK || op == OP_LAPLACE);
        const bool need_p = (op == OP_STOKES || op == OP_BT || op == OP_MASS);
        if (need_u)
          for (int cmp = 0; cmp < dim; ++cmp)
            for (int a = 0; a < nv; ++a) {
              size_t g = (size_t)cmp * L.vnodes + vidx[a];
              u[cmp * nv + a] = (L.vcon[g] && !in.plain) ? 0.0 : src_u[g];
            }
        if (need_p)
          for (int a = 0; a < npl; ++a) pl[a] = src_p[pidx[a]];
        // reference derivatives (unit h) under a general geometry, transformed per point below
        const double *hh = L.geo.empty() ? L.h : unit_h;
        if (need_u)
          for (int cmp = 0; cmp < dim; ++cmp) K.grad(&u[cmp * nv], hh, &gu[cmp * dim * nq], work.data());
        if (need_p) K.pvalues(pl.data(), pq.data(), work.data());
        // quadrature point physics
        for (int q = 0; q < nq; ++q) {
          // B^T and the Laplace operator have no coefficient
          const double eta = (op == OP_BT || op == OP_LAPLACE) ? 1.0 : in.eta[in.eta_per_qp ? c * nq + q : c];
          double Jinv[9];
          const double JxW = L.geo.empty() ? JxW0 * re.qweight(q) : geometry(L, re, c, q, Jinv);
          if (!L.geo.empty() && need_u)
            for (int cmp = 0; cmp < dim; ++cmp) {
              double r[3], g[3];
              for (int e = 0; e < dim; ++e) r[e] = gu[(cmp * dim + e) * nq + q];
              to_physical(dim, Jinv, r, g);
              for (int d = 0; d < dim; ++d) gu[(cmp * dim + d) * nq + q] = g[d];
            }
          // gu[cmp][d][q] = d u_cmp / d x_d
          double eps[3][3] = {{0}};
          double div = 0;
          if (need_u && op != OP_MASS) {
            for (int a = 0; a < dim; ++a)
              for (int b = 0; b < dim; ++b)
                eps[a][b] = 0.5 * (gu[(a * dim + b) * nq + q] + gu[(b * dim + a) * nq + q]);
            for (int a = 0; a < dim; ++a) div += eps[a][a];
          }
          double sig[3][3] = {{0}};
          if (op == OP_STOKES || op == OP_ABLOCK) {
            for (int a = 0; a < dim; ++a)
              for (int b = 0; b < dim; ++b) sig[a][b] = 2.0 * eta * eps[a][b];
            if (in.compressible)
              for (int a = 0; a < dim; ++a) sig[a][a] -= 2.0 * eta / 3.0 * div;
          }
          if (op == OP_LAPLACE)  // grad u : grad v, component by component
            for (int a = 0; a < dim; ++a)
              for (int b = 0; b < dim; ++b) sig[a][b] = gu[(a * dim + b) * nq + q];
          if (op == OP_STOKES || op == OP_BT)
            for (int a = 0; a < dim; ++a) sig[a][a] -= in.s * pq[q];
          if (op == OP_STOKES) fp[q] = -in.s * div * JxW;
          if (op == OP_MASS) fp[q] = in.s * in.s / eta * pq[q] * JxW;
          if (op != OP_MASS) {
            for (int a = 0; a < dim; ++a)
              for (int b = 0; b < dim; ++b) f[(a * dim + b) * nq + q] = sig[a][b] * JxW;
            if (!L.geo.empty())  // test with reference derivatives: f_e = sum_d Jinv[e][d] f_d
              for (int a = 0; a < dim; ++a) {
                double fd[3], fe[3] = {0, 0, 0};
                for (int d = 0; d < dim; ++d) fd[d] = f[(a * dim + d) * nq + q];
                for (int e = 0; e < dim; ++e)
                  for (int d = 0; d < dim; ++d) fe[e] += Jinv[e * dim + d] * fd[d];
                for (int e = 0; e < dim; ++e) f[(a * dim + e) * nq + q] = fe[e];
              }
          }
        }
        if (op != OP_MASS) {
          std::fill(out_u.begin(), out_u.end(), 0.0);
          for (int cmp = 0; cmp < dim; ++cmp)
            K.grad_T_add(&f[cmp * dim * nq], hh, &out_u[cmp * nv], work.data());
          for (int cmp = 0; cmp < dim; ++cmp)
            for (int a = 0; a < nv; ++a) {
              size_t g = (size_t)cmp * L.vnodes + vidx[a];
              if (!L.vcon[g]) dst_u[g] += out_u[cmp * nv + a];
            }
        }
        if (op == OP_STOKES || op == OP_MASS) {
          std::fill(out_p.begin(), out_p.end(), 0.0);
          K.pvalues_T_add(fp.data(), out_p.data(), work.data());
          for (int a = 0; a < npl; ++a) dst_p[pidx[a]] += out_p[a];
        }
      }
    }
  }
}

// ---------------------------------------------------------------- problem context
struct Context {
  orc_problem p;
  std::vector<Level> levels;
  CellKernels K;
  explicit Context(const orc_problem &pp) : p(pp), K(pp.dim, pp.degree) {
    if (p.dim != 2 && p.dim != 3) throw std::runtime_error("dim must be 2 or 3");
    if (p.degree < 2 || p.degree > 3) throw std::runtime_error("degree must be 2 or 3");
    if (p.n_levels < 1) throw std::runtime_error("n_levels must be >= 1");
    for (int l = 0; l < p.n_levels; ++l) levels.push_back(make_level(p, l));
  }
};

// deal.II Base::vmult semantics: dst = A src on free DoFs, dst[K] = src[K].
void stokes_vmult(const Context &C, int l, const double *eta, int per_qp, const double *src, double *dst) {
  const Level &L = C.levels[l];
  const long long nu = L.nu();
  std::fill(dst, dst + nu + L.pnodes, 0.0);
  OpInput in{&L, eta, per_qp, C.p.pressure_scaling, C.p.compressible};
  cell_loop(in, C.K, OP_STOKES, src, src + nu, dst, dst + nu);
  for (long long i = 0; i < nu; ++i)
    if (L.vcon[i]) dst[i] = src[i];
}
// correct_stokes_rhs (stokes_matrix_free_local_smoothing.cc:617-754): the
// Stokes operator applied to u0 with constrained values read as stored, negated;
// constrained velocity rows receive nothing (integrate_scatter drops them).
void stokes_rhs_correction(const Context &C, int l, const double *eta, int per_qp, const double *u0, double *dst) {
  const Level &L = C.levels[l];
  const long long nu = L.nu(), n = nu + L.pnodes;
  std::fill(dst, dst + n, 0.0);
  OpInput in{&L, eta, per_qp, C.p.pressure_scaling, C.p.compressible};
  in.plain = 1;
  cell_loop(in, C.K, OP_STOKES, u0, u0 + nu, dst, dst + nu);
  for (long long i = 0; i < n; ++i) dst[i] = -dst[i];
}
void ablock_vmult(const Context &C, int l, const double *eta, int per_qp, const double *src, double *dst) {
  const Level &L = C.levels[l];
  const long long nu = L.nu();
  std::fill(dst, dst + nu, 0.0);
  OpInput in{&L, eta, per_qp, C.p.pressure_scaling, C.p.compressible};
  cell_loop(in, C.K, OP_ABLOCK, src, nullptr, dst, nullptr);
  for (long long i = 0; i < nu; ++i)
    if (L.vcon[i]) dst[i] = src[i];
}
// B^T block as used by BlockSchurPreconditioner: dst_u = B^T p on free rows,
// constrained rows take the (zero) velocity input of the block vector.
void bt_vmult(const Context &C, int l, const double *src_p, double *dst_u) {
  const Level &L = C.levels[l];
  const long long nu = L.nu();
  std::fill(dst_u, dst_u + nu, 0.0);
  static const double dummy_eta = 1.0;
  OpInput in{&L, &dummy_eta, 0, C.p.pressure_scaling, C.p.compressible};
  cell_loop(in, C.K, OP_BT, nullptr, src_p, dst_u, nullptr);
}
void mass_vmult(const Context &C, int l, const double *eta, int per_qp, const double *src, double *dst) {
  const Level &L = C.levels[l];
  std::fill(dst, dst + L.pnodes, 0.0);
  OpInput in{&L, eta, per_qp, C.p.pressure_scaling, C.p.compressible};
  cell_loop(in, C.K, OP_MASS, nullptr, src, nullptr, dst);
}

// deal.II MatrixFreeOperators::LaplaceOperator<dim, k, k+1, dim> (vector
// Laplace, coefficient 1) of the mesh-deformation solve
// (source/mesh_deformation/interface.cc:1094-1095), Base::vmult semantics
void laplace_vmult(const Context &C, int l, const double *src, double *dst) {
  const Level &L = C.levels[l];
  const long long nu = L.nu();
  std::fill(dst, dst + nu, 0.0);
  OpInput in{&L, nullptr, 0, C.p.pressure_scaling, 0};
  cell_loop(in, C.K, OP_LAPLACE, src, nullptr, dst, nullptr);
  for (long long i = 0; i < nu; ++i)
    if (L.vcon[i]) dst[i] = src[i];
}
// its diagonal: |grad N|^2 integrated over the cells of the node (the same
// for every component), 1 on constrained DoFs
void laplace_diagonal(const Context &C, int l, double *diag) {
  const Level &L = C.levels[l];
  const RefElement &re = C.K.re;
  const int dim = L.dim, nq = re.nq(), nv = re.nvloc(), N = re.N, Q = re.Q;
  std::fill(diag, diag + L.nu(), 0.0);
  double JxW0 = 1.0;
  for (int d = 0; d < dim; ++d) JxW0 *= L.h[d];
  for (long long c = 0; c < L.cells; ++c) {
    int cc[3];
    cell_coords(L, c, cc);
    for (int a = 0; a < nv; ++a) {
      const int ai[3] = {a % N, (a / N) % N, a / (N * N)};
      double acc = 0;
      for (int q = 0; q < nq; ++q) {
        const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
        double r[3] = {0, 0, 0}, gr[3] = {0, 0, 0}, Jinv[9];
        for (int d = 0; d < dim; ++d) {
          double v = 1.0;
          for (int e = 0; e < dim; ++e) v *= (e == d) ? re.G[qi[e] * N + ai[e]] : re.S[qi[e] * N + ai[e]];
          r[d] = v;
        }
        const double JxW = geometry(L, re, c, q, Jinv);
        to_physical(dim, Jinv, r, gr);
        double g2 = 0;
        for (int d = 0; d < dim; ++d) g2 += gr[d] * gr[d];
        acc += g2 * JxW;
      }
      const long long gn = vnode(L, cc, a, N);
      for (int cmp = 0; cmp < dim; ++cmp) diag[(size_t)cmp * L.vnodes + gn] += acc;
    }
  }
  for (long long i = 0; i < L.nu(); ++i)
    if (L.vcon[i]) diag[i] = 1.0;
}

// Diagonal of the A block from the closed form of 2 eta eps(phi):eps(phi) for
// phi = N e_c on an axis-aligned cell:  eta (|grad N|^2 + (d_c N)^2)
// [- (2/3) eta (d_c N)^2 when compressible]; identity entries on constrained
// DoFs (ABlockOperator::compute_diagonal + set_constrained_entries_to_one,
// matrix_free_operators.cc:1068-1098).  Returns the diagonal, not its inverse.
void ablock_diagonal(const Context &C, int l, const double *eta, int per_qp, double *diag) {
  const Level &L = C.levels[l];
  const RefElement &re = C.K.re;
  const int dim = L.dim, nq = re.nq(), nv = re.nvloc(), N = re.N, Q = re.Q;
  std::fill(diag, diag + L.nu(), 0.0);
  double JxW0 = 1.0;
  for (int d = 0; d < dim; ++d) JxW0 *= L.h[d];
  for (long long c = 0; c < L.cells; ++c) {
    int cc[3];
    cell_coords(L, c, cc);
    for (int a = 0; a < nv; ++a) {
      const int ai[3] = {a % N, (a / N) % N, a / (N * N)};
      const long long gn = vnode(L, cc, a, N);
      double acc[3] = {0, 0, 0};
      for (int q = 0; q < nq; ++q) {
        const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
        double r[3] = {0, 0, 0}, gr[3] = {0, 0, 0}, Jinv[9];
        for (int d = 0; d < dim; ++d) {
          double v = 1.0;
          for (int e = 0; e < dim; ++e) v *= (e == d) ? re.G[qi[e] * N + ai[e]] : re.S[qi[e] * N + ai[e]];
          r[d] = v;
        }
        const double JxW = geometry(L, re, c, q, Jinv);
        to_physical(dim, Jinv, r, gr);
        double g2 = 0;
        for (int d = 0; d < dim; ++d) g2 += gr[d] * gr[d];
        const double et = eta[per_qp ? c * nq + q : c];
        for (int cmp = 0; cmp < dim; ++cmp) {
          double v = et * (g2 + gr[cmp] * gr[cmp]);
          if (C.p.compressible) v -= 2.0 / 3.0 * et * gr[cmp] * gr[cmp];
          acc[cmp] += v * JxW;
        }
      }
      for (int cmp = 0; cmp < dim; ++cmp) diag[(size_t)cmp * L.vnodes + gn] += acc[cmp];
    }
  }
  for (long long i = 0; i < L.nu(); ++i)
    if (L.vcon[i]) diag[i] = 1.0;
}

// Diagonal of the s^2/eta weighted pressure mass (MassMatrixOperator::compute_diagonal,
// matrix_free_operators.cc:861-908).  No pressure constraints on box meshes.
void mass_diagonal(const Context &C, int l, const double *eta, int per_qp, double *diag) {
  const Level &L = C.levels[l];
  const RefElement &re = C.K.re;
  const int dim = L.dim, nq = re.nq(), npl = re.nploc(), NP = re.NP, Q = re.Q;
  const double s = C.p.pressure_scaling;
  std::fill(diag, diag + L.pnodes, 0.0);
  double JxW0 = 1.0;
  for (int d = 0; d < dim; ++d) JxW0 *= L.h[d];
  for (long long c = 0; c < L.cells; ++c) {
    int cc[3];
    cell_coords(L, c, cc);
    for (int a = 0; a < npl; ++a) {
      const int ai[3] = {a % NP, (a / NP) % NP, a / (NP * NP)};
      double acc = 0;
      for (int q = 0; q < nq; ++q) {
        const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
        double v = 1.0;
        for (int e = 0; e < dim; ++e) v *= re.SP[qi[e] * NP + ai[e]];
        const double et = eta[per_qp ? c * nq + q : c];
        double Jinv[9];
        acc += s * s / et * v * v * geometry(L, re, c, q, Jinv);
      }
      diag[pnode(L, cc, a, NP)] += acc;
    }
  }
}

// ---------------------------------------------------------------- brute-force element matrices
// Dense local Stokes matrix [A B^T; B 0] and local mass matrix, evaluated by
// looping over every basis function pair at every quadrature point
// (source/simulator/assemblers/stokes.cc:431-441 weak form; mass s^2/eta from
// matrix_free_operators.cc:831,840).  Local ordering: velocity (component, node
// lexicographic), then pressure nodes.
void cell_matrices(const Context &C, int l, const double *eta, int per_qp, double *stokes, double *mass,
                   long long *dofs) {
  const Level &L = C.levels[l];
  const RefElement &re = C.K.re;
  const int dim = L.dim, nq = re.nq(), nv = re.nvloc(), npl = re.nploc(), N = re.N, NP = re.NP, Q = re.Q;
  const int nloc = dim * nv + npl;
  const double s = C.p.pressure_scaling;
  double JxW0 = 1.0;
  for (int d = 0; d < dim; ++d) JxW0 *= L.h[d];
  std::vector<double> gN(nv * 3), vP(npl);
  for (long long c = 0; c < L.cells; ++c) {
    int cc[3];
    cell_coords(L, c, cc);
    double *Ks = stokes + (size_t)c * nloc * nloc;
    double *Km = mass + (size_t)c * npl * npl;
    std::fill(Ks, Ks + (size_t)nloc * nloc, 0.0);
    std::fill(Km, Km + (size_t)npl * npl, 0.0);
    for (int q = 0; q < nq; ++q) {
      const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
      double Jinv[9];
      const double JxW = geometry(L, re, c, q, Jinv);
      const double et = eta[per_qp ? c * nq + q : c];
      for (int a = 0; a < nv; ++a) {
        const int ai[3] = {a % N, (a / N) % N, a / (N * N)};
        double r[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
          double v = 1.0;
          for (int e = 0; e < dim; ++e)
            v *= (e == d) ? re.vb.deriv(ai[e], re.qx[qi[e]]) : re.vb.value(ai[e], re.qx[qi[e]]);
          r[d] = v;
        }
        to_physical(dim, Jinv, r, &gN[a * 3]);
      }
      for (int a = 0; a < npl; ++a) {
        const int ai[3] = {a % NP, (a / NP) % NP, a / (NP * NP)};
        double v = 1.0;
        for (int e = 0; e < dim; ++e) v *= re.pb.value(ai[e], re.qx[qi[e]]);
        vP[a] = v;
      }
      // symmetric gradients eps(phi_{c,a})_{ij} = 1/2 (delta_ic dN_a/dx_j + delta_jc dN_a/dx_i)
      for (int ci = 0; ci < dim; ++ci)
        for (int a = 0; a < nv; ++a) {
          double ea[3][3] = {{0}};
          for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j)
              ea[i][j] = 0.5 * ((i == ci ? gN[a * 3 + j] : 0.0) + (j == ci ? gN[a * 3 + i] : 0.0));
          const double diva = gN[a * 3 + ci];
          const int row = ci * nv + a;
          for (int cj = 0; cj < dim; ++cj)
            for (int b = 0; b < nv; ++b) {
              double eb[3][3] = {{0}};
              for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                  eb[i][j] = 0.5 * ((i == cj ? gN[b * 3 + j] : 0.0) + (j == cj ? gN[b * 3 + i] : 0.0));
              double ee = 0;
              for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) ee += ea[i][j] * eb[i][j];
              double v = 2.0 * et * ee;
              if (C.p.compressible) v -= 2.0 / 3.0 * et * diva * gN[b * 3 + cj];
              Ks[(size_t)row * nloc + cj * nv + b] += v * JxW;
            }
          for (int b = 0; b < npl; ++b) {
            const double bt = -s * vP[b] * diva * JxW;
            Ks[(size_t)row * nloc + dim * nv + b] += bt;    // B^T
            Ks[(size_t)(dim * nv + b) * nloc + row] += bt;  // B
          }
        }
      for (int a = 0; a < npl; ++a)
        for (int b = 0; b < npl; ++b) Km[a * npl + b] += s * s / et * vP[a] * vP[b] * JxW;
    }
    long long *dd = dofs + (size_t)c * nloc;
    for (int ci = 0; ci < dim; ++ci)
      for (int a = 0; a < nv; ++a) dd[ci * nv + a] = (long long)ci * L.vnodes + vnode(L, cc, a, N);
    for (int a = 0; a < npl; ++a) dd[dim * nv + a] = L.nu() + pnode(L, cc, a, NP);
  }
}

// ---------------------------------------------------------------- transfers
// 1D interpolation from a coarse Lagrange grid (nc cells, degree kk, support
// points pts) to the once-refined grid: fine node g sits at (e_f + pts[i]) in
// fine-cell units, i.e. at half that in coarse-cell units; its weights are the
// coarse basis functions evaluated there.
struct Interp1D {
  int nf = 0, nc = 0;
  std::vector<std::vector<std::pair<int, double>>> rows;  // per fine node: (coarse node, weight)
};
Interp1D make_interp(int ncells_coarse, int kk, const Lagrange1D &basis) {
  Interp1D P;
  const int ncf = 2 * ncells_coarse;
  P.nf = kk * ncf + 1;
  P.nc = kk * ncells_coarse + 1;
  P.rows.resize(P.nf);
  for (int g = 0; g < P.nf; ++g) {
    int ef = std::min(g / kk, ncf - 1);
    int i = g - kk * ef;
    double pos = 0.5 * (ef + basis.pts[i]);
    int ec = std::min((int)std::floor(pos), ncells_coarse - 1);
    double xi = pos - ec;
    for (int j = 0; j <= kk; ++j) {
      double w = basis.value(j, xi);
      if (std::fabs(w) > 1e-14) P.rows[g].push_back({kk * ec + j, w});
    }
  }
  return P;
}

// tensor-product prolongation of one scalar field: coarse (ncx,ncy,ncz) -> fine
void prolongate_scalar(int dim, const Interp1D *P, const double *src, double *dst) {
  int nc[3] = {P[0].nc, dim > 1 ? P[1].nc : 1, dim > 2 ? P[2].nc : 1};
  int nf[3] = {P[0].nf, dim > 1 ? P[1].nf : 1, dim > 2 ? P[2].nf : 1};
  std::vector<double> t1((size_t)nf[0] * nc[1] * nc[2]), t2((size_t)nf[0] * nf[1] * nc[2]);
  for (int z = 0; z < nc[2]; ++z)
    for (int y = 0; y < nc[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        double s = 0;
        for (auto &e : P[0].rows[x]) s += e.second * src[((size_t)z * nc[1] + y) * nc[0] + e.first];
        t1[((size_t)z * nc[1] + y) * nf[0] + x] = s;
      }
  for (int z = 0; z < nc[2]; ++z)
    for (int y = 0; y < nf[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        double s = 0;
        if (dim > 1)
          for (auto &e : P[1].rows[y]) s += e.second * t1[((size_t)z * nc[1] + e.first) * nf[0] + x];
        else
          s = t1[x];
        t2[((size_t)z * nf[1] + y) * nf[0] + x] = s;
      }
  for (int z = 0; z < nf[2]; ++z)
    for (int y = 0; y < nf[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        double s = 0;
        if (dim > 2)
          for (auto &e : P[2].rows[z]) s += e.second * t2[((size_t)e.first * nf[1] + y) * nf[0] + x];
        else
          s = t2[(size_t)y * nf[0] + x];
        dst[((size_t)z * nf[1] + y) * nf[0] + x] = s;
      }
}
// transpose: fine -> coarse (adds into dst)
void restrict_scalar_add(int dim, const Interp1D *P, const double *src, double *dst) {
  int nc[3] = {P[0].nc, dim > 1 ? P[1].nc : 1, dim > 2 ? P[2].nc : 1};
  int nf[3] = {P[0].nf, dim > 1 ? P[1].nf : 1, dim > 2 ? P[2].nf : 1};
  std::vector<double> t2((size_t)nf[0] * nf[1] * nc[2], 0.0), t1((size_t)nf[0] * nc[1] * nc[2], 0.0);
  for (int z = 0; z < nf[2]; ++z)
    for (int y = 0; y < nf[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        const double v = src[((size_t)z * nf[1] + y) * nf[0] + x];
        if (dim > 2)
          for (auto &e : P[2].rows[z]) t2[((size_t)e.first * nf[1] + y) * nf[0] + x] += e.second * v;
        else
          t2[(size_t)y * nf[0] + x] += v;
      }
  for (int z = 0; z < nc[2]; ++z)
    for (int y = 0; y < nf[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        const double v = t2[((size_t)z * nf[1] + y) * nf[0] + x];
        if (dim > 1)
          for (auto &e : P[1].rows[y]) t1[((size_t)z * nc[1] + e.first) * nf[0] + x] += e.second * v;
        else
          t1[x] += v;
      }
  for (int z = 0; z < nc[2]; ++z)
    for (int y = 0; y < nc[1]; ++y)
      for (int x = 0; x < nf[0]; ++x) {
        const double v = t1[((size_t)z * nc[1] + y) * nf[0] + x];
        for (auto &e : P[0].rows[x]) dst[((size_t)z * nc[1] + y) * nc[0] + e.first] += e.second * v;
      }
}

struct Transfer {
  Interp1D Pv[3], Pp[3];
};
Transfer make_transfer(const Context &C, int fine_level) {
  const Level &Lc = C.levels[fine_level - 1];
  Transfer T;
  for (int d = 0; d < C.p.dim; ++d) {
    T.Pv[d] = make_interp(Lc.n[d], C.p.degree, C.K.re.vb);
    T.Pp[d] = make_interp(Lc.n[d], C.p.degree - 1, C.K.re.pb);
  }
  return T;
}
// velocity (kind 0) or pressure (kind 1) prolongation with constrained DoFs
// excluded on both levels (the level operators carry identity rows there).
void prolongate(const Context &C, int fine_level, int kind, const double *src, double *dst) {
  const Level &Lf = C.levels[fine_level], &Lc = C.levels[fine_level - 1];
  Transfer T = make_transfer(C, fine_level);
  if (kind == 0) {
    std::vector<double> s(src, src + Lc.nu());
    for (long long i = 0; i < Lc.nu(); ++i)
      if (Lc.vcon[i]) s[i] = 0.0;
    for (int c = 0; c < C.p.dim; ++c)
      prolongate_scalar(C.p.dim, T.Pv, s.data() + (size_t)c * Lc.vnodes, dst + (size_t)c * Lf.vnodes);
    for (long long i = 0; i < Lf.nu(); ++i)
      if (Lf.vcon[i]) dst[i] = 0.0;
  } else {
    prolongate_scalar(C.p.dim, T.Pp, src, dst);
  }
}
void restrict_(const Context &C, int fine_level, int kind, const double *src, double *dst) {
  const Level &Lf = C.levels[fine_level], &Lc = C.levels[fine_level - 1];
  Transfer T = make_transfer(C, fine_level);
  if (kind == 0) {
    std::vector<double> s(src, src + Lf.nu());
    for (long long i = 0; i < Lf.nu(); ++i)
      if (Lf.vcon[i]) s[i] = 0.0;
    std::fill(dst, dst + Lc.nu(), 0.0);
    for (int c = 0; c < C.p.dim; ++c)
      restrict_scalar_add(C.p.dim, T.Pv, s.data() + (size_t)c * Lf.vnodes, dst + (size_t)c * Lc.vnodes);
    for (long long i = 0; i < Lc.nu(); ++i)
      if (Lc.vcon[i]) dst[i] = 0.0;
  } else {
    std::fill(dst, dst + Lc.pnodes, 0.0);
    restrict_scalar_add(C.p.dim, T.Pp, src, dst);
  }
}

// ---------------------------------------------------------------- level viscosity
// DGQ0 projection of the active viscosity (JxW-weighted cell mean, which is
// the cell value itself when the table is already cell-wise), then the FE_DGQ(0)
// restriction to coarser levels = arithmetic mean of the 2^dim children
// (stokes_matrix_free_local_smoothing.cc:211-263, utilities.cc:2738-2800).
std::vector<std::vector<double>> level_viscosity(const Context &C, const double *eta, int per_qp) {
  const int nl = C.p.n_levels, dim = C.p.dim;
  const RefElement &re = C.K.re;
  const int nq = re.nq();
  std::vector<std::vector<double>> lv(nl);
  const Level &Lf = C.levels[nl - 1];
  lv[nl - 1].resize(Lf.cells);
  for (long long c = 0; c < Lf.cells; ++c) {
    if (!per_qp) {
      lv[nl - 1][c] = eta[c];
      continue;
    }
    // cellwise L2 projection onto constants: JxW-weighted mean
    double s = 0, w = 0, Jinv[9];
    for (int q = 0; q < nq; ++q) {
      const double wq = Lf.geo.empty() ? re.qweight(q) : geometry(Lf, re, c, q, Jinv);
      s += wq * eta[c * nq + q];
      w += wq;
    }
    lv[nl - 1][c] = s / w;
  }
  for (int l = nl - 2; l >= 0; --l) {
    const Level &L = C.levels[l], &F = C.levels[l + 1];
    lv[l].resize(L.cells);
    for (long long c = 0; c < L.cells; ++c) {
      int cc[3];
      cell_coords(L, c, cc);
      double s = 0;
      int cnt = 0;
      for (int dz = 0; dz < (dim == 3 ? 2 : 1); ++dz)
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            long long fc = (2LL * cc[0] + dx) + F.n[0] * ((2LL * cc[1] + dy) + (long long)F.n[1] * (2LL * cc[2] + dz));
            s += lv[l + 1][fc];
            ++cnt;
          }
      lv[l][c] = s / cnt;
    }
  }
  return lv;
}

// ---------------------------------------------------------------- DGQ1 viscosity ("project to Q1")
// Cellwise L2 projection of the quadrature-point viscosity onto Q1 (FE_DGQ(1),
// Utilities::project_cellwise, utilities.cc:2738-2800, called from
// matrix_free_operators.cc:200-270), values back at the quadrature points
// clamped to the [min, max] of the evaluated viscosities (:308-326).  Level
// tables (local_smoothing.cc:211-294): the children's piecewise-Q1 function
// is projected in L2 onto the parent's Q1 space (the FE_DGQ restriction,
// which for DGQ0 is the children mean), then evaluated at the quadrature
// points and clamped the same way.
// Coefficients are per cell at the 2^dim vertices, x fastest.
struct Q1Tables {
  double Rk[2][2][2];  // 1D child->parent projection [child][parent vertex][child vertex]
};
Q1Tables q1_tables() {
  Q1Tables T{};
  std::vector<double> gx, gw;
  gauss_legendre(4, gx, gw);  // exact for the cubic integrands
  for (int k = 0; k < 2; ++k)
    for (int a = 0; a < 2; ++a)
      for (int j = 0; j < 2; ++j) {
        double sum = 0;
        for (size_t q = 0; q < gx.size(); ++q) {
          const double t = gx[q], x = 0.5 * (t + k);
          const double psi = a == 0 ? 4.0 - 6.0 * x : 6.0 * x - 2.0;  // dual basis of Q1 on [0,1]
          const double phi = j == 0 ? 1.0 - t : t;
          sum += gw[q] * psi * phi;
        }
        T.Rk[k][a][j] = 0.5 * sum;
      }
  return T;
}
// projection of one cell's nq values: c_v = sum_q w_q eta_q prod_d psi_{v_d}(x_q,d)
void q1_project_cell(const RefElement &re, const double *eta_q, double *coef) {
  const int dim = re.dim, nv = 1 << dim, Q = re.Q;
  for (int v = 0; v < nv; ++v) coef[v] = 0.0;
  for (int q = 0; q < re.nq(); ++q) {
    const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
    for (int v = 0; v < nv; ++v) {
      double w = re.qweight(q) * eta_q[q];
      for (int d = 0; d < dim; ++d) {
        const double x = re.qx[qi[d]];
        w *= ((v >> d) & 1) ? 6.0 * x - 2.0 : 4.0 - 6.0 * x;
      }
      coef[v] += w;
    }
  }
}
void q1_eval_cell(const RefElement &re, const double *coef, double lo, double hi, double *eta_q) {
  const int dim = re.dim, nv = 1 << dim, Q = re.Q;
  for (int q = 0; q < re.nq(); ++q) {
    const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
    double val = 0;
    for (int v = 0; v < nv; ++v) {
      double w = coef[v];
      for (int d = 0; d < dim; ++d) {
        const double x = re.qx[qi[d]];
        w *= ((v >> d) & 1) ? x : 1.0 - x;
      }
      val += w;
    }
    eta_q[q] = std::min(std::max(val, lo), hi);
  }
}
// per-qp tables of every level (level 0 first) from the raw active per-qp values
std::vector<std::vector<double>> level_viscosity_q1(const Context &C, const double *eta_qp) {
  const int nl = C.p.n_levels, dim = C.p.dim, nv = 1 << dim;
  const RefElement &re = C.K.re;
  const int nq = re.nq();
  const Level &Lf = C.levels[nl - 1];
  for (const Level &L : C.levels)
    if (!L.geo.empty()) throw std::runtime_error("Q1 viscosity projection: box geometry only");
  double lo = eta_qp[0], hi = eta_qp[0];
  for (long long i = 0; i < Lf.cells * nq; ++i) {
    lo = std::min(lo, eta_qp[i]);
    hi = std::max(hi, eta_qp[i]);
  }
  const Q1Tables T = q1_tables();
  std::vector<std::vector<double>> coef(nl), out(nl);
  coef[nl - 1].resize(Lf.cells * nv);
  for (long long c = 0; c < Lf.cells; ++c) q1_project_cell(re, eta_qp + c * nq, &coef[nl - 1][c * nv]);
  for (int l = nl - 2; l >= 0; --l) {
    const Level &L = C.levels[l], &F = C.levels[l + 1];
    coef[l].assign(L.cells * nv, 0.0);
    for (long long c = 0; c < L.cells; ++c) {
      int cc[3];
      cell_coords(L, c, cc);
      for (int ch = 0; ch < nv; ++ch) {
        const int k[3] = {ch & 1, (ch >> 1) & 1, (ch >> 2) & 1};
        const long long fc = (2LL * cc[0] + k[0]) +
                             F.n[0] * ((2LL * cc[1] + k[1]) + (long long)F.n[1] * (dim == 3 ? 2LL * cc[2] + k[2] : 0));
        for (int a = 0; a < nv; ++a)
          for (int j = 0; j < nv; ++j) {
            double w = 1.0;
            for (int d = 0; d < dim; ++d) w *= T.Rk[k[d]][(a >> d) & 1][(j >> d) & 1];
            coef[l][c * nv + a] += w * coef[l + 1][fc * nv + j];
          }
      }
    }
  }
  for (int l = 0; l < nl; ++l) {
    const Level &L = C.levels[l];
    out[l].resize(L.cells * nq);
    for (long long c = 0; c < L.cells; ++c) q1_eval_cell(re, &coef[l][c * nv], lo, hi, &out[l][c * nq]);
  }
  return out;
}

// ---------------------------------------------------------------- vector helpers
using Vec = std::vector<double>;
using Op = std::function<void(const double *, double *)>;
double dot(const Vec &a, const Vec &b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (long long i = 0; i < (long long)a.size(); ++i) s += a[i] * b[i];
  return s;
}
double norm(const Vec &a) { return std::sqrt(dot(a, a)); }

// min / max eigenvalue of a symmetric tridiagonal matrix by Sturm-sequence bisection
int sturm_count(const Vec &d, const Vec &e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    double off = i ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i ? off / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0) ++cnt;
  }
  return cnt;
}
double tridiag_eig(const Vec &d, const Vec &e, int which /*0 min, 1 max*/) {
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < d.size(); ++i) {
    double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < d.size() ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const int target = which == 0 ? 1 : (int)d.size();
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (sturm_count(d, e, mid) >= target)
      hi = mid;
    else
      lo = mid;
    if (hi - lo <= 1e-15 * std::max(1.0, std::fabs(hi))) break;
  }
  return 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- Chebyshev smoother
// Jacobi-preconditioned Chebyshev iteration of the first kind (Saad, Iterative
// Methods, Alg. 12.1) with the eigenvalue interval from a CG-Lanczos estimate
// on D^{-1}A, as set up at stokes_matrix_free_local_smoothing.cc:771-849
// (degree 4, smoothing range 15, 10 CG iterations on levels; degree 8, range
// 1e-3, 100 iterations on level 0).  Eigenvalue estimate: CG started from the
// mean-free pattern (i mod 11), tolerance 1e-2, lambda_max scaled by 1.2.
struct Chebyshev {
  Op A;
  Vec dinv;
  int degree = 4;
  double lmax = 1, lmin = 1, theta = 1, delta = 0;
  int cg_its = 0;
  void estimate(int max_its, double range) {
    const size_t n = dinv.size();
    Vec b(n), x(n, 0.0), r, z(n), p(n), q(n);
    for (size_t i = 0; i < n; ++i) b[i] = (double)(i % 11);
    double mean = 0;
    for (size_t i = 0; i < n; ++i) mean += b[i];
    mean /= (double)n;
    for (size_t i = 0; i < n; ++i) b[i] -= mean;
    r = b;
    Vec alphas, betas;
    double res = norm(r);
    if (res >= 1e-2) {
      for (size_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
      double rz = dot(r, z);
      p = z;
      for (int it = 1; it <= max_its; ++it) {
        A(p.data(), q.data());
        const double pq = dot(p, q);
        const double alpha = rz / pq;
        for (size_t i = 0; i < n; ++i) r[i] -= alpha * q[i];
        alphas.push_back(alpha);
        res = norm(r);
        if (res < 1e-2 || it == max_its) break;
        for (size_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        const double rz2 = dot(r, z);
        const double beta = rz2 / rz;
        rz = rz2;
        for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        betas.push_back(beta);
      }
    }
    cg_its = (int)alphas.size();
    if (alphas.empty()) {
      lmin = lmax = 1.0;
    } else {
      const size_t m = alphas.size();
      Vec d(m), e(m > 1 ? m - 1 : 0);
      d[0] = 1.0 / alphas[0];
      for (size_t k = 1; k < m; ++k) d[k] = 1.0 / alphas[k] + betas[k - 1] / alphas[k - 1];
      for (size_t k = 0; k + 1 < m; ++k) e[k] = std::sqrt(betas[k]) / alphas[k];
      lmin = tridiag_eig(d, e, 0);
      lmax = 1.2 * tridiag_eig(d, e, 1);
    }
    const double alpha = range > 1.0 ? lmax / range : std::min(0.9 * lmax, lmin);
    delta = 0.5 * (lmax - alpha);
    theta = 0.5 * (lmax + alpha);
  }
  // x <- x + Chebyshev polynomial correction; x_is_zero selects the first
  // step without the residual matvec (deal.II's vmult vs step).
  void run(Vec &x, const Vec &b, bool x_is_zero) const {
    const size_t n = b.size();
    Vec r(n), d(n), t(n);
    if (x_is_zero) std::fill(x.begin(), x.end(), 0.0);
    // d_0 = (1/theta) D^{-1} r_0
    if (x_is_zero)
      r = b;
    else {
      A(x.data(), t.data());
      for (size_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
    }
    for (size_t i = 0; i < n; ++i) {
      d[i] = dinv[i] * r[i] / theta;
      x[i] += d[i];
    }
    if (degree < 2 || std::fabs(delta) < 1e-40) return;
    const double sigma = theta / delta;
    double rho_old = delta / theta;
    for (int k = 1; k < degree; ++k) {
      A(x.data(), t.data());
      const double rho = 1.0 / (2.0 * sigma - rho_old);
      const double f1 = rho * rho_old, f2 = 2.0 * rho / delta;
      for (size_t i = 0; i < n; ++i) {
        d[i] = f1 * d[i] + f2 * dinv[i] * (b[i] - t[i]);
        x[i] += d[i];
      }
      rho_old = rho;
    }
  }
};

// ---------------------------------------------------------------- multigrid
struct MG {
  const Context *C;
  int kind;  // 0 velocity A block, 1 pressure mass
  int steps = 1;  // smoother steps (mg::SmootherRelaxation steps)
  std::vector<Op> ops;
  std::vector<Chebyshev> smoothers;
  std::vector<size_t> sizes;
  size_t size(int l) const { return sizes[l]; }
  void smooth(int l, Vec &x, const Vec &b, bool x_zero) const {
    for (int s = 0; s < steps; ++s) smoothers[l].run(x, b, x_zero && s == 0);
  }
  void vcycle_level(int l, Vec &x, const Vec &b) const {
    if (l == 0) {
      smooth(0, x, b, true);
      return;
    }
    smooth(l, x, b, true);
    Vec t(size(l));
    ops[l](x.data(), t.data());
    for (size_t i = 0; i < t.size(); ++i) t[i] = b[i] - t[i];
    Vec bc(size(l - 1)), xc(size(l - 1), 0.0);
    restrict_(*C, l, kind, t.data(), bc.data());
    vcycle_level(l - 1, xc, bc);
    Vec e(size(l));
    prolongate(*C, l, kind, xc.data(), e.data());
    for (size_t i = 0; i < x.size(); ++i) x[i] += e[i];
    smooth(l, x, b, false);
  }
  void vmult(Vec &x, const Vec &b) const {
    x.assign(size((int)ops.size() - 1), 0.0);
    vcycle_level((int)ops.size() - 1, x, b);
  }
};

// ---------------------------------------------------------------- Krylov solvers
struct SolveStatus {
  bool converged = false;
  int its = 0;
  double res = 0;
  int prec_calls = 0;
};
// Restarted GMRES with right preconditioning (non-flexible: x += P(V y) at each
// restart) or FGMRES (flexible: x += Z y).  Convergence on the unpreconditioned
// residual estimate from the Givens rotations.  Classical Gram-Schmidt with a
// second pass when the first cancelled more than 1 - 1/sqrt(2) of the norm.
SolveStatus gmres(const Op &A, const Op &P, Vec &x, const Vec &b, double tol, int maxit, int restart,
                  bool flexible) {
  SolveStatus st;
  const size_t n = b.size();
  int acc = 0;
  while (true) {
    Vec r(n), t(n);
    A(x.data(), t.data());
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
    double rho = norm(r);
    st.res = rho;
    if (rho <= tol) {
      st.converged = true;
      st.its = acc;
      return st;
    }
    if (acc >= maxit) {
      st.its = acc;
      return st;
    }
    std::vector<Vec> V, Z;
    V.push_back(r);
    for (size_t i = 0; i < n; ++i) V[0][i] /= rho;
    std::vector<Vec> H;  // columns
    Vec cs, sn, g(1, rho);
    bool done = false;
    int j = 0;
    for (; j < restart; ++j) {
      ++acc;
      Vec z(n), w(n);
      P(V[j].data(), z.data());
      ++st.prec_calls;
      A(z.data(), w.data());
      if (flexible) Z.push_back(z);
      Vec h(j + 2, 0.0);
      // classical Gram-Schmidt; a second pass when the first cancelled more
      // than 1 - 1/sqrt(2) of the norm (Daniel-Gragg-Kaufman-Stewart test)
      const double wn0 = norm(w);
      for (int pass = 0; pass < 2; ++pass) {
        Vec hp(j + 1);
        for (int i = 0; i <= j; ++i) hp[i] = dot(V[i], w);
        for (int i = 0; i <= j; ++i) {
          h[i] += hp[i];
          for (size_t k = 0; k < n; ++k) w[k] -= hp[i] * V[i][k];
        }
        if (pass == 0 && norm(w) >= 0.70710678118654752 * wn0) break;
      }
      h[j + 1] = norm(w);
      for (int i = 0; i < j; ++i) {
        double a = h[i], bb = h[i + 1];
        h[i] = cs[i] * a + sn[i] * bb;
        h[i + 1] = -sn[i] * a + cs[i] * bb;
      }
      double den = std::hypot(h[j], h[j + 1]);
      double c = den > 0 ? h[j] / den : 1.0, s = den > 0 ? h[j + 1] / den : 0.0;
      cs.push_back(c);
      sn.push_back(s);
      const double hj1 = h[j + 1];
      h[j] = den;
      h[j + 1] = 0.0;
      g.push_back(-s * g[j]);
      g[j] = c * g[j];
      H.push_back(h);
      rho = std::fabs(g[j + 1]);
      st.res = rho;
      if (hj1 > 0) {
        Vec v(n);
        for (size_t k = 0; k < n; ++k) v[k] = w[k] / hj1;
        V.push_back(v);
      }
      if (rho <= tol || hj1 == 0.0) {
        st.converged = true;
        done = true;
        ++j;
        break;
      }
      if (acc >= maxit) {
        done = true;
        ++j;
        break;
      }
    }
    // back substitution H y = g (upper triangular, size j)
    const int m = j;
    Vec y(m, 0.0);
    for (int i = m - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < m; ++k) s -= H[k][i] * y[k];
      y[i] = s / H[i][i];
    }
    if (flexible) {
      for (int i = 0; i < m; ++i)
        for (size_t k = 0; k < n; ++k) x[k] += y[i] * Z[i][k];
    } else {
      Vec comb(n, 0.0), pz(n);
      for (int i = 0; i < m; ++i)
        for (size_t k = 0; k < n; ++k) comb[k] += y[i] * V[i][k];
      P(comb.data(), pz.data());
      ++st.prec_calls;
      for (size_t k = 0; k < n; ++k) x[k] += pz[k];
    }
    if (done) {
      st.its = acc;
      return st;
    }
  }
}

// Shadow-space entry k of IDR(s): uniform in [-1, 1) from a counter hash
// (splitmix64 finalizer), so every implementation draws the same bits.
inline double idr_shadow_entry(int k, long long i) {
  uint64_t z = (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull + (uint64_t)i * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

// IDR(s) with biorthogonalization and right preconditioning (van Gijzen and
// Sonneveld, ACM TOMS 38(1), 2011, Algorithm 913), the scheme of deal.II's
// SolverIDR used for the cheap steps (stokes_matrix_free_local_smoothing.cc:1291-1301):
// one step = s inner updates plus one dimension-reduction update, each with
// one preconditioner application; convergence on the recursively updated
// residual, checked after every update.  The shadow space is s orthonormalized
// vectors from idr_shadow_entry (deal.II draws normal variates; any random
// space serves the method).
SolveStatus idrs(const Op &A, const Op &P, Vec &x, const Vec &b, double tol, int maxit, int s) {
  SolveStatus st;
  const size_t n = b.size();
  Vec r(n), t(n);
  A(x.data(), t.data());
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
  double res = norm(r);
  st.res = res;
  if (res <= tol) {
    st.converged = true;
    return st;
  }
  std::vector<Vec> Qs(s, Vec(n)), G(s, Vec(n, 0.0)), U(s, Vec(n, 0.0));
  for (int k = 0; k < s; ++k) {
    for (size_t i = 0; i < n; ++i) Qs[k][i] = idr_shadow_entry(k, (long long)i);
    for (int j = 0; j < k; ++j) {  // modified Gram-Schmidt
      const double c = dot(Qs[k], Qs[j]);
      for (size_t i = 0; i < n; ++i) Qs[k][i] -= c * Qs[j][i];
    }
    const double nq = norm(Qs[k]);
    for (size_t i = 0; i < n; ++i) Qs[k][i] /= nq;
  }
  std::vector<double> M(s * s, 0.0);  // M(i,k) = Q_i . G_k, lower triangular after each step
  for (int i = 0; i < s; ++i) M[i * s + i] = 1.0;
  double omega = 1.0;
  Vec v(n), uhat(n), vhat(n);
  int step = 0;
  while (true) {
    ++step;
    std::vector<double> f(s);
    for (int i = 0; i < s; ++i) f[i] = dot(Qs[i], r);
    for (int k = 0; k < s; ++k) {
      // solve the lower-triangular system M(k:s, k:s) c = f(k:s)
      std::vector<double> c(s - k);
      for (int i = k; i < s; ++i) {
        double a = f[i];
        for (int j = k; j < i; ++j) a -= M[i * s + j] * c[j - k];
        c[i - k] = a / M[i * s + i];
      }
      for (size_t q = 0; q < n; ++q) {
        double a = r[q];
        for (int i = k; i < s; ++i) a -= c[i - k] * G[i][q];
        v[q] = a;
      }
      P(v.data(), vhat.data());
      ++st.prec_calls;
      for (size_t q = 0; q < n; ++q) {
        double a = omega * vhat[q];
        for (int i = k; i < s; ++i) a += c[i - k] * U[i][q];
        uhat[q] = a;
      }
      A(uhat.data(), G[k].data());
      for (int i = 0; i < k; ++i) {  // make G_k orthogonal to Q_0..Q_{k-1}
        const double alpha = dot(Qs[i], G[k]) / M[i * s + i];
        for (size_t q = 0; q < n; ++q) {
          G[k][q] -= alpha * G[i][q];
          uhat[q] -= alpha * U[i][q];
        }
      }
      U[k] = uhat;
      for (int i = k; i < s; ++i) M[i * s + k] = dot(Qs[i], G[k]);
      if (M[k * s + k] == 0.0) {  // breakdown
        st.its = step;
        return st;
      }
      const double beta = f[k] / M[k * s + k];
      for (size_t q = 0; q < n; ++q) {
        r[q] -= beta * G[k][q];
        x[q] += beta * U[k][q];
      }
      res = norm(r);
      st.res = res;
      if (res <= tol) {
        st.converged = true;
        st.its = step;
        return st;
      }
      if (step >= maxit) {
        st.its = step;
        return st;
      }
      for (int i = k + 1; i < s; ++i) f[i] -= beta * M[i * s + k];
    }
    // dimension reduction
    P(r.data(), uhat.data());
    ++st.prec_calls;
    A(uhat.data(), v.data());
    const double vv = dot(v, v);
    if (vv == 0.0) {
      st.its = step;
      return st;
    }
    omega = dot(v, r) / vv;
    for (size_t q = 0; q < n; ++q) {
      r[q] -= omega * v[q];
      x[q] += omega * uhat[q];
    }
    res = norm(r);
    st.res = res;
    if (res <= tol) {
      st.converged = true;
      st.its = step;
      return st;
    }
    if (step >= maxit) {
      st.its = step;
      return st;
    }
  }
}

// preconditioned CG (deal.II SolverCG semantics: absolute tolerance, x0 given)
SolveStatus cg(const Op &A, const Op &P, Vec &x, const Vec &b, double tol, int maxit) {
  SolveStatus st;
  const size_t n = b.size();
  Vec r(n), z(n), p(n), q(n);
  A(x.data(), q.data());
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  double res = norm(r);
  st.res = res;
  if (res <= tol) {
    st.converged = true;
    return st;
  }
  P(r.data(), z.data());
  double rz = dot(r, z);
  p = z;
  for (int it = 1; it <= maxit; ++it) {
    A(p.data(), q.data());
    const double alpha = rz / dot(p, q);
    for (size_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    res = norm(r);
    st.its = it;
    st.res = res;
    if (res <= tol) {
      st.converged = true;
      return st;
    }
    P(r.data(), z.data());
    const double rz2 = dot(r, z);
    const double beta = rz2 / rz;
    rz = rz2;
    for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return st;
}

// ---------------------------------------------------------------- the solve
struct Solver {
  const Context &C;
  const double *eta;
  int per_qp;
  int lq = 0;  // level tables per quadrature point (DGQ1)
  std::vector<std::vector<double>> lv;
  MG mgA, mgS;
  // q1: "project to Q1 only viscosity" (eta must be per qp); the active table
  // becomes the top level table
  Solver(const Context &c, const double *eta_, int per_qp_, bool q1 = false) : C(c), eta(eta_), per_qp(per_qp_) {
    if (q1) {
      if (!per_qp) throw std::runtime_error("Q1 projection needs per-quadrature-point viscosity");
      lv = level_viscosity_q1(C, eta);
      lq = 1;
      eta = lv.back().data();
    } else {
      lv = level_viscosity(C, eta, per_qp);
    }
    const int nl = C.p.n_levels;
    mgA.C = mgS.C = &C;
    mgA.kind = 0;
    mgS.kind = 1;
    mgA.steps = 1;
    mgS.steps = 4;  // mg::SmootherRelaxation mg_smoother_Schur(4), local_smoothing.cc:803-804
    for (int l = 0; l < nl; ++l) {
      const double *le = lv[l].data();
      mgA.ops.push_back([this, l, le](const double *s, double *d) { ablock_vmult(C, l, le, lq, s, d); });
      mgS.ops.push_back([this, l, le](const double *s, double *d) { mass_vmult(C, l, le, lq, s, d); });
      mgA.sizes.push_back(C.levels[l].nu());
      mgS.sizes.push_back(C.levels[l].pnodes);
    }
  }
  void build_preconditioner() {
    const int nl = C.p.n_levels;
    mgA.smoothers.resize(nl);
    mgS.smoothers.resize(nl);
    for (int l = 0; l < nl; ++l) {
      Chebyshev &a = mgA.smoothers[l], &s = mgS.smoothers[l];
      a.A = mgA.ops[l];
      s.A = mgS.ops[l];
      a.dinv.resize(mgA.sizes[l]);
      s.dinv.resize(mgS.sizes[l]);
      ablock_diagonal(C, l, lv[l].data(), lq, a.dinv.data());
      mass_diagonal(C, l, lv[l].data(), lq, s.dinv.data());
      for (auto &v : a.dinv) v = 1.0 / v;
      for (auto &v : s.dinv) v = 1.0 / v;
      const bool coarse = (l == 0);
      a.degree = s.degree = coarse ? 8 : 4;
      a.estimate(coarse ? 100 : 10, coarse ? 1e-3 : 15.0);
      s.estimate(coarse ? 100 : 10, coarse ? 1e-3 : 15.0);
    }
  }
};

}  // namespace

// ====================================================================== C API
extern "C" {

static thread_local std::string g_err;
const char *orc_last_error() { return g_err.c_str(); }

#define ORC_TRY(...)                  \
  try {                               \
    __VA_ARGS__;                      \
    return 0;                         \
  } catch (const std::exception &e) { \
    g_err = e.what();                 \
    return 1;                         \
  }

int orc_sizes(const orc_problem *p, int level, long long *out) {
  ORC_TRY({
    Level L = make_level(*p, level);
    out[0] = L.cells;
    out[1] = L.nu();
    out[2] = L.pnodes;
    out[3] = L.vnodes;
  })
}
int orc_constrained(const orc_problem *p, int level, unsigned char *out) {
  ORC_TRY({
    Level L = make_level(*p, level);
    std::memcpy(out, L.vcon.data(), L.vcon.size());
  })
}
int orc_qp_points(const orc_problem *p, int level, double *xyz) {
  ORC_TRY({
    Level L = make_level(*p, level);
    RefElement re(p->dim, p->degree);
    const int nq = re.nq(), Q = re.Q;
    for (long long c = 0; c < L.cells; ++c) {
      int cc[3];
      cell_coords(L, c, cc);
      for (int q = 0; q < nq; ++q) {
        const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
        for (int d = 0; d < p->dim; ++d) xyz[((size_t)c * nq + q) * p->dim + d] = (cc[d] + re.qx[qi[d]]) * L.h[d];
      }
    }
  })
}
int orc_quadrature_weights(const orc_problem *p, int level, double *jxw) {
  ORC_TRY({
    Level L = make_level(*p, level);
    RefElement re(p->dim, p->degree);
    double J = 1;
    for (int d = 0; d < p->dim; ++d) J *= L.h[d];
    for (int q = 0; q < re.nq(); ++q) jxw[q] = J * re.qweight(q);
  })
}
int orc_velocity_support_points(const orc_problem *p, int level, double *xyz) {
  ORC_TRY({
    Level L = make_level(*p, level);
    RefElement re(p->dim, p->degree);
    for (long long nidx = 0; nidx < L.vnodes; ++nidx) {
      long long ix[3] = {nidx % L.nv[0], (nidx / L.nv[0]) % L.nv[1], nidx / ((long long)L.nv[0] * L.nv[1])};
      for (int d = 0; d < p->dim; ++d) {
        long long cell = std::min<long long>(ix[d] / p->degree, L.n[d] - 1);
        int i = (int)(ix[d] - p->degree * cell);
        xyz[nidx * p->dim + d] = (cell + re.vb.pts[i]) * L.h[d];
      }
    }
  })
}
int orc_pressure_support_points(const orc_problem *p, int level, double *xyz) {
  ORC_TRY({
    Level L = make_level(*p, level);
    RefElement re(p->dim, p->degree);
    const int kp = p->degree - 1;
    for (long long nidx = 0; nidx < L.pnodes; ++nidx) {
      long long ix[3] = {nidx % L.np[0], (nidx / L.np[0]) % L.np[1], nidx / ((long long)L.np[0] * L.np[1])};
      for (int d = 0; d < p->dim; ++d) {
        long long cell = std::min<long long>(ix[d] / kp, L.n[d] - 1);
        int i = (int)(ix[d] - kp * cell);
        xyz[nidx * p->dim + d] = (cell + re.pb.pts[i]) * L.h[d];
      }
    }
  })
}

int orc_stokes_vmult(const orc_problem *p, int level, const double *eta, int per_qp, const double *src,
                     double *dst) {
  ORC_TRY({
    Context C(*p);
    stokes_vmult(C, level, eta, per_qp, src, dst);
  })
}
int orc_ablock_vmult(const orc_problem *p, int level, const double *eta, int per_qp, const double *src,
                     double *dst) {
  ORC_TRY({
    Context C(*p);
    ablock_vmult(C, level, eta, per_qp, src, dst);
  })
}
int orc_bt_vmult(const orc_problem *p, int level, const double *src_p, double *dst_u) {
  ORC_TRY({
    Context C(*p);
    bt_vmult(C, level, src_p, dst_u);
  })
}
int orc_mass_vmult(const orc_problem *p, int level, const double *eta, int per_qp, const double *src,
                   double *dst) {
  ORC_TRY({
    Context C(*p);
    mass_vmult(C, level, eta, per_qp, src, dst);
  })
}
int orc_ablock_diagonal(const orc_problem *p, int level, const double *eta, int per_qp, double *diag) {
  ORC_TRY({
    Context C(*p);
    ablock_diagonal(C, level, eta, per_qp, diag);
  })
}
int orc_mass_diagonal(const orc_problem *p, int level, const double *eta, int per_qp, double *diag) {
  ORC_TRY({
    Context C(*p);
    mass_diagonal(C, level, eta, per_qp, diag);
  })
}
int orc_cell_matrices(const orc_problem *p, int level, const double *eta, int per_qp, double *stokes,
                      double *mass, long long *dofs) {
  ORC_TRY({
    Context C(*p);
    cell_matrices(C, level, eta, per_qp, stokes, mass, dofs);
  })
}
int orc_prolongate(const orc_problem *p, int fine_level, int kind, const double *src, double *dst) {
  ORC_TRY({
    Context C(*p);
    prolongate(C, fine_level, kind, src, dst);
  })
}
int orc_restrict(const orc_problem *p, int fine_level, int kind, const double *src, double *dst) {
  ORC_TRY({
    Context C(*p);
    restrict_(C, fine_level, kind, src, dst);
  })
}
// level tables: out holds sum_l cells(l) values, level 0 first
int orc_level_viscosity(const orc_problem *p, const double *eta, int per_qp, double *out) {
  ORC_TRY({
    Context C(*p);
    auto lv = level_viscosity(C, eta, per_qp);
    size_t off = 0;
    for (auto &v : lv) {
      std::memcpy(out + off, v.data(), v.size() * sizeof(double));
      off += v.size();
    }
  })
}
// DGQ1 ("project to Q1") tables: out holds sum_l cells(l) * nq values, level 0
// first; the last level is the active table
int orc_level_viscosity_q1(const orc_problem *p, const double *eta_qp, double *out) {
  ORC_TRY({
    Context C(*p);
    auto lv = level_viscosity_q1(C, eta_qp);
    size_t off = 0;
    for (auto &v : lv) {
      std::memcpy(out + off, v.data(), v.size() * sizeof(double));
      off += v.size();
    }
  })
}
// Chebyshev set-up numbers per level: [lmin, lmax, theta, delta, cg_its] for A then S
int orc_smoother_setup(const orc_problem *p, const double *eta, int per_qp, double *out) {
  ORC_TRY({
    Context C(*p);
    Solver S(C, eta, per_qp);
    S.build_preconditioner();
    for (int l = 0; l < p->n_levels; ++l) {
      const Chebyshev *cs[2] = {&S.mgA.smoothers[l], &S.mgS.smoothers[l]};
      for (int k = 0; k < 2; ++k) {
        double *o = out + (l * 2 + k) * 5;
        o[0] = cs[k]->lmin;
        o[1] = cs[k]->lmax;
        o[2] = cs[k]->theta;
        o[3] = cs[k]->delta;
        o[4] = cs[k]->cg_its;
      }
    }
  })
}
// one V-cycle of the A (kind 0) or S (kind 1) multigrid preconditioner on the finest level
int orc_vcycle(const orc_problem *p, const double *eta, int per_qp, int kind, const double *src, double *dst) {
  ORC_TRY({
    Context C(*p);
    Solver S(C, eta, per_qp);
    S.build_preconditioner();
    const MG &mg = kind == 0 ? S.mgA : S.mgS;
    const int top = p->n_levels - 1;
    Vec b(src, src + mg.size(top)), x;
    mg.vmult(x, b);
    std::memcpy(dst, x.data(), x.size() * sizeof(double));
  })
}
// Mesh-deformation displacement solve (MeshDeformationHandler::
// compute_mesh_displacements_gmg_for_degree, source/mesh_deformation/interface.cc:1076-1360):
// Chebyshev smoothers of degree 5 over [1.2 lambda_max / 15, 1.2 lambda_max]
// from 10 eigenvalue-CG iterations on levels > 0; on level 0 Chebyshev as a
// solver: 50 eigenvalue iterations, range 1e-3 and the degree chosen by the
// Chebyshev error bound (deal.II's degree = invalid_unsigned_int):
//   sigma = (1 - sqrt(1/r)) / (1 + sqrt(1/r)), r = lambda_max / alpha,
//   degree = 1 + floor(log(1/eps + sqrt(1/eps^2 - 1)) / log(1/sigma)), eps = 1e-3;
// the coarse grid applies that smoother (MGCoarseGridApplySmoother).
struct MeshMG {
  MG mg;
  void build(const Context &C) {
    const int nl = C.p.n_levels;
    mg.C = &C;
    mg.kind = 0;
    mg.steps = 1;
    for (int l = 0; l < nl; ++l) {
      mg.ops.push_back([&C, l](const double *s, double *d) { laplace_vmult(C, l, s, d); });
      mg.sizes.push_back(C.levels[l].nu());
    }
    mg.smoothers.resize(nl);
    for (int l = 0; l < nl; ++l) {
      Chebyshev &a = mg.smoothers[l];
      a.A = mg.ops[l];
      a.dinv.resize(mg.sizes[l]);
      laplace_diagonal(C, l, a.dinv.data());
      for (auto &v : a.dinv) v = 1.0 / v;
      if (l > 0) {
        a.degree = 5;
        a.estimate(10, 15.0);
      } else {
        a.estimate(50, 1e-3);
        const double alpha = a.theta - a.delta;
        const double r = a.lmax / alpha;
        const double sigma = (1.0 - std::sqrt(1.0 / r)) / (1.0 + std::sqrt(1.0 / r));
        const double eps = 1e-3;
        a.degree = sigma > 0 ? 1 + (int)(std::log(1.0 / eps + std::sqrt(1.0 / eps / eps - 1.0)) / std::log(1.0 / sigma))
                             : 1;
      }
    }
  }
};
// CG with the V-cycle (SolverCG, interface.cc:1316-1340); rhs = -A u0 with u0
// read as stored and constrained rows zero (:1130-1156); solution = x + u0 on
// the constrained DoFs (constraints.distribute, :1342).  Returns iterations.
int mesh_solve(const Context &C, const double *u0, double tolerance, double *sol, double *res_out) {
  const int top = C.p.n_levels - 1;
  const Level &L = C.levels[top];
  const size_t n = (size_t)L.nu();
  MeshMG M;
  M.build(C);
  Vec rhs(n, 0.0);
  {
    OpInput in{&L, nullptr, 0, 1.0, 0};
    in.plain = 1;
    cell_loop(in, C.K, OP_LAPLACE, u0, nullptr, rhs.data(), nullptr);
    for (size_t i = 0; i < n; ++i) rhs[i] = L.vcon[i] ? 0.0 : -rhs[i];
  }
  const double tol = tolerance * norm(rhs);
  Vec x(n, 0.0), r = rhs, z, p, q(n);
  M.mg.vmult(z, r);
  p = z;
  double rz = dot(r, z), res = norm(r);
  int its = 0;
  const int max_its = (int)std::min<size_t>(5 * n, 1000000);
  while (res > tol && its < max_its) {
    laplace_vmult(C, top, p.data(), q.data());
    const double alpha = rz / dot(p, q);
    for (size_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    ++its;
    res = norm(r);
    if (res <= tol) break;
    M.mg.vmult(z, r);
    const double rz2 = dot(r, z);
    const double beta = rz2 / rz;
    rz = rz2;
    for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  for (size_t i = 0; i < n; ++i) sol[i] = L.vcon[i] ? u0[i] : x[i];
  if (res_out) {
    res_out[0] = res;
    res_out[1] = tol;
  }
  return its;
}

// Free-surface stabilization term of the Stokes operator on the active level
// (StokesOperator::local_apply_boundary_face, matrix_free_operators.cc:563-602):
// on every face marked in faces[6] (x-, x+, y-, y+, z-, z+),
//   dst_u[i] += sign * ( - sum_q JxW_q (t_q . u(x_q)) (phi_i(x_q) . n) ),
// constrained source DoFs read as zero (as stored when plain), constrained rows
// skipped (integrate_scatter).  stab: per marked face in face order, per face
// cell (lexicographic over the two other directions, the lower one fastest),
// per face quadrature point (QGauss(k+1) tensor, lower direction fastest), dim
// values t (ASPECT: rho dt theta |g| g_hat, local_smoothing.cc:555-608).
// Written as a plain quadrature loop with the Lagrange basis evaluated pointwise.
void free_surface_apply(const Context &C, int l, const int faces[6], const double *stab, const double *src_u,
                        double *dst_u, double sign, int plain) {
  const Level &L = C.levels[l];
  if (!L.geo.empty()) throw std::runtime_error("free-surface term: box geometry only");
  const RefElement &re = C.K.re;
  const int dim = L.dim, N = re.N, Q = re.Q;
  const double *tab = stab;
  for (int f = 0; f < 2 * dim; ++f) {
    if (!faces[f]) continue;
    const int d = f / 2, side = f % 2;
    int t1 = -1, t2 = -1;  // tangential directions, lower first
    for (int e = 0; e < dim; ++e)
      if (e != d) (t1 < 0 ? t1 : t2) = e;
    const int n1 = L.n[t1], n2 = dim == 3 ? L.n[t2] : 1;
    const double area = L.h[t1] * (dim == 3 ? L.h[t2] : 1.0);
    const int nfq = dim == 3 ? Q * Q : Q, nfn = dim == 3 ? N * N : N;
    const double nsign = side == 0 ? -1.0 : 1.0;  // outward normal n = nsign e_d
    for (int c2 = 0; c2 < n2; ++c2)
      for (int c1 = 0; c1 < n1; ++c1) {
        // global node of face-local node (a1, a2)
        auto gnode = [&](int a1, int a2) {
          int g[3] = {0, 0, 0};
          g[d] = side == 0 ? 0 : L.nv[d] - 1;
          g[t1] = L.k * c1 + a1;
          if (dim == 3) g[t2] = L.k * c2 + a2;
          return (long long)g[0] + L.nv[0] * ((long long)g[1] + (long long)L.nv[1] * g[2]);
        };
        const double *tq = tab + (size_t)(c1 + (size_t)n1 * c2) * nfq * dim;
        for (int q = 0; q < nfq; ++q) {
          const int q1 = q % Q, q2 = q / Q;
          const double w = area * re.qw[q1] * (dim == 3 ? re.qw[q2] : 1.0);
          double tu = 0;
          for (int a = 0; a < nfn; ++a) {
            const int a1 = a % N, a2 = a / N;
            const double phi = re.vb.value(a1, re.qx[q1]) * (dim == 3 ? re.vb.value(a2, re.qx[q2]) : 1.0);
            const long long gn = gnode(a1, a2);
            for (int comp = 0; comp < dim; ++comp) {
              const size_t gi = (size_t)comp * L.vnodes + gn;
              const double uv = (L.vcon[gi] && !plain) ? 0.0 : src_u[gi];
              tu += tq[q * dim + comp] * phi * uv;
            }
          }
          for (int a = 0; a < nfn; ++a) {
            const int a1 = a % N, a2 = a / N;
            const double phi = re.vb.value(a1, re.qx[q1]) * (dim == 3 ? re.vb.value(a2, re.qx[q2]) : 1.0);
            const size_t gi = (size_t)d * L.vnodes + gnode(a1, a2);
            if (!L.vcon[gi]) dst_u[gi] += sign * (-w * tu * phi * nsign);
          }
        }
      }
    tab += (size_t)n1 * n2 * nfq * dim;
  }
}

// CPU timing of the multigrid preconditioner: set-up (diagonals, eigenvalue
// estimates) once, then reps V-cycles of kind 0 (A block) or 1 (Schur mass);
// seconds[0] = set-up, seconds[1] = mean V-cycle time
int orc_time_vcycle(const orc_problem *p, const double *eta, int per_qp, int kind, int reps, double *seconds) {
  ORC_TRY({
    Context C(*p);
    const auto t0 = std::chrono::steady_clock::now();
    Solver S(C, eta, per_qp);
    S.build_preconditioner();
    const auto t1 = std::chrono::steady_clock::now();
    const MG &mg = kind == 0 ? S.mgA : S.mgS;
    const int top = p->n_levels - 1;
    Vec b(mg.size(top)), x;
    for (size_t i = 0; i < b.size(); ++i) b[i] = std::sin(0.37 * (double)i);
    for (int r = 0; r < std::max(1, reps); ++r) mg.vmult(x, b);
    const auto t2 = std::chrono::steady_clock::now();
    seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    seconds[1] = std::chrono::duration<double>(t2 - t1).count() / std::max(1, reps);
  })
}
// geometry tables of the levels used by every following call (natural order:
// cells lexicographic, quadrature points lexicographic, dim*dim + 1 values)
int orc_set_geometry(int level, long long n, const double *geo) {
  ORC_TRY({
    if (level < 0 || level > 30) throw std::runtime_error("level out of range");
    if ((int)g_geometry.size() <= level) g_geometry.resize(level + 1);
    g_geometry[level].assign(geo, geo + n);
  })
}
int orc_clear_geometry() {
  ORC_TRY({ g_geometry.clear(); })
}
// mesh-deformation pieces: the problem's bc are the mesh constraints
// (0 none, 1 no normal displacement, 2 zero / prescribed displacement)
int orc_laplace_vmult(const orc_problem *p, int level, const double *src, double *dst) {
  ORC_TRY({
    Context C(*p);
    laplace_vmult(C, level, src, dst);
  })
}
int orc_laplace_diagonal(const orc_problem *p, int level, double *diag) {
  ORC_TRY({
    Context C(*p);
    laplace_diagonal(C, level, diag);
  })
}
// out[l*5 + {0..4}] = lambda_min, lambda_max, theta, delta, degree (eigenvalue-CG
// iterations in out[nl*5 + l])
int orc_mesh_smoother_info(const orc_problem *p, double *out) {
  ORC_TRY({
    Context C(*p);
    MeshMG M;
    M.build(C);
    const int nl = p->n_levels;
    for (int l = 0; l < nl; ++l) {
      const Chebyshev &a = M.mg.smoothers[l];
      out[l * 5 + 0] = a.lmin;
      out[l * 5 + 1] = a.lmax;
      out[l * 5 + 2] = a.theta;
      out[l * 5 + 3] = a.delta;
      out[l * 5 + 4] = a.degree;
      out[nl * 5 + l] = a.cg_its;
    }
  })
}
int orc_mesh_vcycle(const orc_problem *p, const double *src, double *dst) {
  ORC_TRY({
    Context C(*p);
    MeshMG M;
    M.build(C);
    const int top = p->n_levels - 1;
    Vec b(src, src + M.mg.size(top)), x;
    M.mg.vmult(x, b);
    std::memcpy(dst, x.data(), x.size() * sizeof(double));
  })
}
int orc_mesh_solve(const orc_problem *p, const double *u0, double tolerance, double *sol, int *iterations,
                   double *res) {
  ORC_TRY({
    Context C(*p);
    *iterations = mesh_solve(C, u0, tolerance, sol, res);
  })
}
// the free-surface face term alone, added to dst (N_u values): sign +1 for the
// operator, -1 with plain reads for the right-hand-side correction of
// correct_stokes_rhs (local_smoothing.cc:709-744)
int orc_free_surface_apply(const orc_problem *p, const int *faces, const double *stab, const double *src_u,
                           double *dst_u, double sign, int plain) {
  ORC_TRY({
    Context C(*p);
    free_surface_apply(C, p->n_levels - 1, faces, stab, src_u, dst_u, sign, plain);
  })
}
// single Chebyshev smoother application (x0 = 0) on a level
int orc_chebyshev(const orc_problem *p, const double *eta, int per_qp, int kind, int level, const double *src,
                  double *dst) {
  ORC_TRY({
    Context C(*p);
    Solver S(C, eta, per_qp);
    S.build_preconditioner();
    const MG &mg = kind == 0 ? S.mgA : S.mgS;
    Vec b(src, src + mg.size(level)), x(mg.size(level), 0.0);
    mg.smoothers[level].run(x, b, true);
    std::memcpy(dst, x.data(), x.size() * sizeof(double));
  })
}

// rhs correction -A(u0) of correct_stokes_rhs on the active level
int orc_stokes_rhs_correction(const orc_problem *p, const double *eta, int per_qp, const double *u0, double *out) {
  ORC_TRY({
    Context C(*p);
    stokes_rhs_correction(C, p->n_levels - 1, eta, per_qp, u0, out);
  })
}

// Buoyancy right-hand side rhs_u = int rho g . phi (constrained rows zero),
// rhs_p = 0 (incompressible), source/simulator/assemblers/stokes.cc:393-396.
int orc_rhs_buoyancy(const orc_problem *p, const double *rho_qp, const double *gvec, double *rhs) {
  ORC_TRY({
    Context C(*p);
    const Level &L = C.levels[p->n_levels - 1];
    const RefElement &re = C.K.re;
    const int dim = p->dim, nq = re.nq(), nv = re.nvloc(), N = re.N, Q = re.Q;
    std::fill(rhs, rhs + L.nu() + L.pnodes, 0.0);
    double J = 1;
    for (int d = 0; d < dim; ++d) J *= L.h[d];
    for (long long c = 0; c < L.cells; ++c) {
      int cc[3];
      cell_coords(L, c, cc);
      for (int a = 0; a < nv; ++a) {
        const int ai[3] = {a % N, (a / N) % N, a / (N * N)};
        double acc = 0;
        for (int q = 0; q < nq; ++q) {
          const int qi[3] = {q % Q, (q / Q) % Q, q / (Q * Q)};
          double v = 1.0;
          for (int e = 0; e < dim; ++e) v *= re.S[qi[e] * N + ai[e]];
          double Jinv[9];
          acc += rho_qp[c * nq + q] * v * (L.geo.empty() ? J * re.qweight(q) : geometry(L, re, c, q, Jinv));
        }
        const long long gn = vnode(L, cc, a, N);
        for (int cmp = 0; cmp < dim; ++cmp) rhs[(size_t)cmp * L.vnodes + gn] += gvec[cmp] * acc;
      }
    }
    for (long long i = 0; i < L.nu(); ++i)
      if (L.vcon[i]) rhs[i] = 0.0;
  })
}

// The handler's solve (stokes_matrix_free_local_smoothing.cc:758-1446) on box
// meshes: initial guess x0 (velocity constrained entries zeroed, pressure
// divided by s), residual-based tolerance (:977-1041), cheap GMRES or IDR(s)
// with one V-cycle per block (:1262-1313), FGMRES with inner CG solves on failure
// (:1314-1397), then pressure *= s and volume pressure normalization.
int orc_solve(const orc_problem *p, const double *eta, int per_qp, const orc_solve_params *prm,
              const double *rhs, const double *x0, double *solution, orc_solve_result *res) {
  ORC_TRY({
    if (prm->cheap_krylov != 0 && prm->cheap_krylov != 1) throw std::runtime_error("unknown cheap Krylov method");
    if (prm->cheap_krylov == 1 && prm->idr_s < 1) throw std::runtime_error("IDR(s) parameter must be >= 1");
    Context C(*p);
    const int top = p->n_levels - 1;
    const Level &L = C.levels[top];
    const long long nu = L.nu(), np = L.pnodes, n = nu + np;
    const double s = p->pressure_scaling;
    if (prm->level_eta_mode != 0 && prm->level_eta_mode != 1) throw std::runtime_error("unknown level viscosity mode");
    Solver S(C, eta, per_qp, prm->level_eta_mode == 1);
    S.build_preconditioner();
    // active viscosity table: the input, or its clamped Q1 projection
    if (S.lq) {
      eta = S.eta;
      per_qp = 1;
    }
    Op Aop = [&](const double *src, double *dst) { stokes_vmult(C, top, eta, per_qp, src, dst); };
    const bool newton = prm->solve_newton_system != 0;
    Vec b(rhs, rhs + n), x(x0, x0 + n);
    if (newton) std::fill(x.begin(), x.end(), 0.0);
    for (long long i = 0; i < nu; ++i)
      if (L.vcon[i]) x[i] = 0.0;
    for (long long i = 0; i < np; ++i) x[nu + i] /= s;
    Vec t(n);
    if (newton) {
      // Newton update: tolerance and initial residual from ||rhs|| (:1028-1041)
      double r2 = 0;
      for (long long i = 0; i < n; ++i) r2 += b[i] * b[i];
      res->r0_norm = res->initial_nonlinear_residual = std::sqrt(r2);
    } else {
      Aop(x.data(), t.data());
      double r2 = 0;
      for (long long i = 0; i < n; ++i) r2 += (b[i] - t[i]) * (b[i] - t[i]);
      res->initial_nonlinear_residual = std::sqrt(r2);
      Vec xu0 = x;
      std::fill(xu0.begin(), xu0.begin() + nu, 0.0);
      Aop(xu0.data(), t.data());
      double ru = 0, rp = 0;
      for (long long i = 0; i < nu; ++i) ru += (b[i] - t[i]) * (b[i] - t[i]);
      for (long long i = 0; i < np; ++i) rp += b[nu + i] * b[nu + i];
      res->r0_norm = std::sqrt(ru + rp);
    }
    const double tol = prm->tolerance * res->r0_norm;
    res->solver_tolerance = tol;
    int a_apps = 0, s_apps = 0;
    // cheap block preconditioner
    Op Pcheap = [&](const double *src, double *dst) {
      Vec sp(src + nu, src + n), xp;
      S.mgS.vmult(xp, sp);
      ++s_apps;
      for (long long i = 0; i < np; ++i) dst[nu + i] = -xp[i];
      Vec bt(nu);
      bt_vmult(C, top, dst + nu, bt.data());
      Vec tu(nu);
      for (long long i = 0; i < nu; ++i) tu[i] = L.vcon[i] ? src[i] : src[i] - bt[i];
      Vec xu;
      S.mgA.vmult(xu, tu);
      ++a_apps;
      std::copy(xu.begin(), xu.end(), dst);
    };
    Op Pexp = [&](const double *src, double *dst) {
      Vec sp(src + nu, src + n), xp(np, 0.0);
      const double nsp = norm(sp);
      if (nsp > 1e-50) {
        Op M = [&](const double *a, double *b2) { mass_vmult(C, top, eta, per_qp, a, b2); };
        Op Pm = [&](const double *a, double *b2) {
          Vec bb(a, a + np), xx;
          S.mgS.vmult(xx, bb);
          std::copy(xx.begin(), xx.end(), b2);
        };
        SolveStatus st = cg(M, Pm, xp, sp, nsp * prm->s_block_tolerance, 100);
        if (!st.converged) throw std::runtime_error("Schur complement CG did not converge");
        s_apps += st.its;
      }
      for (long long i = 0; i < np; ++i) dst[nu + i] = -xp[i];
      Vec bt(nu);
      bt_vmult(C, top, dst + nu, bt.data());
      Vec tu(nu);
      for (long long i = 0; i < nu; ++i) tu[i] = L.vcon[i] ? src[i] : src[i] - bt[i];
      Op Aa = [&](const double *a, double *b2) { ablock_vmult(C, top, eta, per_qp, a, b2); };
      Op Pa = [&](const double *a, double *b2) {
        Vec bb(a, a + nu), xx;
        S.mgA.vmult(xx, bb);
        std::copy(xx.begin(), xx.end(), b2);
      };
      Vec xu(nu, 0.0);
      SolveStatus st = cg(Aa, Pa, xu, tu, norm(tu) * prm->a_block_tolerance, 10000);
      if (!st.converged) throw std::runtime_error("A block CG did not converge");
      a_apps += st.its;
      std::copy(xu.begin(), xu.end(), dst);
    };
    res->cheap_iterations = 0;
    res->expensive_iterations = 0;
    bool ok = false;
    if (prm->n_cheap_steps > 0) {
      SolveStatus st = prm->cheap_krylov == 1
                           ? idrs(Aop, Pcheap, x, b, tol, prm->n_cheap_steps, prm->idr_s)
                           : gmres(Aop, Pcheap, x, b, tol, prm->n_cheap_steps, prm->restart, false);
      res->cheap_iterations = st.its;
      res->final_linear_residual = st.res;
      ok = st.converged;
    }
    res->a_preconditioner_applications = a_apps;
    res->s_preconditioner_applications = s_apps;
    if (!ok && prm->n_expensive_steps > 0) {
      int a0 = a_apps, s0 = s_apps;
      SolveStatus st = gmres(Aop, Pexp, x, b, tol, prm->n_expensive_steps, prm->restart, true);
      res->expensive_iterations = st.its;
      res->final_linear_residual = st.res;
      ok = st.converged;
      res->a_preconditioner_applications = a0 + (a_apps - a0);
      res->s_preconditioner_applications = s0 + (s_apps - s0);
    }
    res->converged = ok ? 1 : 0;
    for (long long i = 0; i < np; ++i) x[nu + i] *= s;
    // volume pressure normalization: subtract (int p)/|Omega| (not for a Newton update)
    res->pressure_normalization_adjustment = 0;
    if (!newton) {
      Vec ones(np, 1.0), w(np);
      std::vector<double> unit(L.cells, 1.0);
      // int p = p . (M_1 1) with the unweighted pressure mass (s=1, eta=1)
      orc_problem q = *p;
      q.pressure_scaling = 1.0;
      Context C1(q);
      mass_vmult(C1, top, unit.data(), 0, ones.data(), w.data());
      double ip = 0, vol = 0;
      for (long long i = 0; i < np; ++i) {
        ip += x[nu + i] * w[i];
        vol += w[i];
      }
      const double adj = ip / vol;
      for (long long i = 0; i < np; ++i) x[nu + i] -= adj;
      res->pressure_normalization_adjustment = -adj;
    }
    std::copy(x.begin(), x.end(), solution);
  })
}

}  // extern "C"
