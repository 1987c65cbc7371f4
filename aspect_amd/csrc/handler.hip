// Host side of the MI355X matrix-free Stokes solver: the StokesMatrixFreeHandler
// shape (setup_dofs / assemble / build_preconditioner / solve) over device data,
// and the C-ABI declared in include/aspect_mfs.h.
//
// Reference structure followed (ASPECT, source/simulator/solver/):
//   setup_dofs              stokes_matrix_free_local_smoothing.cc:1450-1851
//   evaluate_material_model :173-294, fill_active_cell_data matrix_free_operators.cc:137-330
//   build_preconditioner    :1855-1867
//   solve                   :758-1446 (smoothers :771-849, tolerance :977-1041,
//                           cheap GMRES :1262-1313, expensive FGMRES :1314-1397,
//                           post-steps :1406-1445)
//   block preconditioner    include/aspect/simulator/solver/block_stokes_preconditioner.h:43-352
//
// Numerics restated (not copied) from the published algorithms: Jacobi
// preconditioned Chebyshev semi-iteration with a CG-Lanczos eigenvalue
// estimate, V-cycle multigrid, right-preconditioned restarted GMRES / flexible
// GMRES with classical Gram-Schmidt (second pass when the first cancelled), and CG.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/aspect_mfs.h"
#include "mfs_internal.h"

using namespace mfs;

namespace mfs {
// Pointers each operator kernel dereferences, by (op, mode): the velocity
// operators read src_u and write dst_u, the pressure ones src_p / dst_p, the
// Stokes operator both; every operator except B^T reads a cell table (the
// density for OP_RHS); RESID reads the right-hand side of what it writes.
void check_op_args(Op op, Mode mode, const void *src_u, const void *src_p, const void *eta, const void *rhs_u,
                   const void *rhs_p, const void *dst_u, const void *dst_p) {
  if (op < OP_STOKES || op > OP_LDIAG) throw std::invalid_argument("unknown operator code");
  if (mode < MODE_SET || mode > MODE_ADD) throw std::invalid_argument("unknown operator mode");
  const bool rd_u = op == OP_STOKES || op == OP_ABLOCK || op == OP_LAPLACE,
             rd_p = op == OP_STOKES || op == OP_BT || op == OP_MASS;
  const bool wr_u = op == OP_STOKES || op == OP_ABLOCK || op == OP_BT || op == OP_ADIAG || op == OP_RHS ||
                    op == OP_LAPLACE || op == OP_LDIAG;
  const bool wr_p = op == OP_STOKES || op == OP_MASS || op == OP_MDIAG;
  const bool diag = op == OP_ADIAG || op == OP_MDIAG || op == OP_RHS || op == OP_LDIAG;
  const bool resid = mode == MODE_RESID && !diag;
  const bool coef = op != OP_BT && op != OP_LAPLACE && op != OP_LDIAG;  // operators with a cell table
  const char *names[7] = {"src_u", "src_p", "eta", "rhs_u", "rhs_p", "dst_u", "dst_p"};
  const void *ptrs[7] = {src_u, src_p, eta, rhs_u, rhs_p, dst_u, dst_p};
  const bool need[7] = {rd_u, rd_p, coef, resid && wr_u, resid && wr_p, wr_u, wr_p};
  static const char *opn[9] = {"Stokes", "A block", "B^T", "mass", "A diagonal", "mass diagonal", "buoyancy rhs",
                               "Laplace", "Laplace diagonal"};
  for (int i = 0; i < 7; ++i)
    if (need[i] && !ptrs[i])
      throw std::invalid_argument(std::string(opn[op]) + " operator: null " + names[i] + " pointer");
}
}  // namespace mfs

namespace {

thread_local std::string g_err;
}  // namespace

// error message of the calling thread's last failed C-ABI call (shared with
// the general-mesh entry points)
void mfs::set_last_error(const char *msg) { g_err = msg ? msg : ""; }

namespace {

template <typename T>
struct DBuf {
  T *p = nullptr;
  long long n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~DBuf() { release(); }
  void alloc(long long nn) {
    release();
    if (nn > 0) MFS_HIP(hipMalloc(reinterpret_cast<void **>(&p), (size_t)nn * sizeof(T)));
    n = nn;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  long long bytes() const { return n * (long long)sizeof(T); }
};

struct ChebInfo {
  double lmin = 1, lmax = 1, theta = 1, delta = 0;
  int degree = 4, its = 0;
};

LevelDesc make_level_desc(const mfs_box &b, int l) {
  LevelDesc L{};
  L.dim = b.dim;
  L.k = b.degree;
  int t[3];
  tile_cells(b.dim, b.degree, t);
  L.cells = 1;
  L.vnodes = 1;
  L.pnodes = 1;
  L.ntiles_total = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < b.dim) {
      L.n[d] = b.coarse[d] << l;
      L.h[d] = b.extent[d] / L.n[d];
      L.nv[d] = b.degree * L.n[d] + 1;
      L.np[d] = (b.degree - 1) * L.n[d] + 1;
      L.ntiles[d] = (L.n[d] + t[d] - 1) / t[d];
    } else {
      L.n[d] = 1;
      L.h[d] = 1.0;
      L.nv[d] = 1;
      L.np[d] = 1;
      L.ntiles[d] = 1;
    }
    L.cells *= L.n[d];
    L.vnodes *= L.nv[d];
    L.pnodes *= L.np[d];
    L.ntiles_total *= L.ntiles[d];
  }
  L.s = b.pressure_scaling;
  level_derive(L);
  for (int f = 0; f < 6; ++f) L.bc[f] = (f < 2 * b.dim) ? b.bc[f] : 0;
  L.compressible = b.compressible;
  return L;
}

int n_qp(const mfs_box &b) {
  const int q = b.degree + 1;
  return b.dim == 3 ? q * q * q : q * q;
}
int cells_per_tile_of(const LevelDesc &L) {
  int t[3];
  return tile_cells(L.dim, L.k, t);
}
int qp_per_cell_of(const LevelDesc &L) { return L.dim == 3 ? (L.k + 1) * (L.k + 1) * (L.k + 1) : (L.k + 1) * (L.k + 1); }
int cells_per_tile(const mfs_box &b) {
  int t[3];
  return tile_cells(b.dim, b.degree, t);
}

// ------------------------------------------------------------------ multigrid in precision T
template <typename T>
struct GMGLevel {
  LevelDesc L;
  long long nu = 0, np = 0;
  DBuf<T> eta, dinvA, dinvS;
  DBuf<T> ietan;  // 1 / cellwise viscosity in natural cell order (Schur level operator), set by build()
  DBuf<T> ux, ub, ur, ud, px, pb, pr, pd;
  ChebInfo chA, chS;
};

template <typename T>
struct GMG {
  hipStream_t st;
  std::vector<std::unique_ptr<GMGLevel<T>>> lv;
  DBuf<double> *dot_work = nullptr, *dot_out = nullptr;
  int top = 0;
  // V-cycles replayed from a captured hipGraph (one per kind); the capture
  // holds kernel arguments (buffers, Chebyshev coefficients), so it is dropped
  // whenever those change
  bool graphs = true;
  hipGraphExec_t vexec[2] = {nullptr, nullptr};
  DBuf<T> xwork;  // work arrays of the separable transfers (sized for the finest pair)
  // A-block applies in the V-cycle run as two launches over all tiles through
  // this work array (per-cell contributions, then node-wise sums) instead of
  // eight tile colours: on coarse levels a colour launch is bound by the latency
  // of one tile, and on the fine level the work-array traffic costs less than
  // the colour tails and the in-kernel gather
  DBuf<T> swork;
  T *split_for(int) { return swork.p; }
  bool ietan_ok = false;  // GMGLevel::ietan matches the current level viscosity
  // vector Laplace hierarchy of the mesh-deformation solve (kind 0 only): the
  // level operator is the Laplace operator, its diagonal the Laplace diagonal,
  // and the smoothers follow interface.cc:1263-1290
  bool laplace = false;
  // Where the level tables came from (the FP64 copy below is rebuilt from them)
  const double *eta_src = nullptr;
  int eta_src_pq = 0;
  const std::vector<DBuf<double>> *q1_src = nullptr;
  double q1_lo = 0, q1_hi = 0;
  // FP32 hierarchies: an FP64 copy of the level tables runs the Chebyshev
  // set-up (diagonals and eigenvalue estimates) and, kept for level 0 only,
  // the coarse-grid solve of every V-cycle
  std::unique_ptr<GMG<double>> shadow;
  void drop_graphs() {
    for (hipGraphExec_t &e : vexec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }
  GMG() = default;
  GMG(const GMG &) = delete;
  GMG &operator=(const GMG &) = delete;
  ~GMG() { drop_graphs(); }

  void setup(const mfs_box &b, hipStream_t s, DBuf<double> *work, DBuf<double> *out) {
    drop_graphs();
    st = s;
    dot_work = work;
    dot_out = out;
    lv.clear();
    const int ct = cells_per_tile(b);
    for (int l = 0; l < b.n_levels; ++l) {
      auto g = std::make_unique<GMGLevel<T>>();
      g->L = make_level_desc(b, l);
      g->nu = (long long)b.dim * g->L.vnodes;
      g->np = g->L.pnodes;
      g->eta.alloc(g->L.ntiles_total * ct);
      v_fill<T>(g->eta.p, (T)1, g->eta.n, st);
      for (DBuf<T> *v : {&g->dinvA, &g->ux, &g->ub, &g->ur, &g->ud}) {
        v->alloc(g->nu);
        v_fill<T>(v->p, (T)0, v->n, st);
      }
      for (DBuf<T> *v : {&g->dinvS, &g->px, &g->pb, &g->pr, &g->pd}) {
        v->alloc(g->np);
        v_fill<T>(v->p, (T)0, v->n, st);
      }
      lv.push_back(std::move(g));
    }
    top = b.n_levels - 1;
    if (top > 0) xwork.alloc(transfer_work_size(lv[top - 1]->L, lv[top]->L));
    long long sw = 0;
    for (auto &g : lv)
      sw = std::max(sw, split_work_size(g->L));
    if (sw > 0 && !laplace) swork.alloc(sw);
  }

  int eta_pq = 0;  // level tables per quadrature point (DGQ1) or per cell (DGQ0)

  // DGQ0 level viscosity: finest from the active table, coarser by averaging children
  void set_eta(const double *active, int active_per_qp) {
    drop_graphs();
    ietan_ok = false;  // recomputed by build()
    eta_pq = 0;
    eta_src = active;
    eta_src_pq = active_per_qp;
    for (auto &g : lv)
      if (g->eta.n != g->L.ntiles_total * cells_per_tile_of(g->L)) g->eta.alloc(g->L.ntiles_total * cells_per_tile_of(g->L));
    GMGLevel<T> &f = *lv[top];
    if (active_per_qp)
      launch_eta_qp_to_cell<T, double>(f.L, active, f.eta.p, st);
    else
      v_cast<double, T>(active, f.eta.p, f.eta.n, st);
    for (int l = top - 1; l >= 0; --l) launch_eta_restrict<T>(lv[l]->L, lv[l + 1]->L, lv[l + 1]->eta.p, lv[l]->eta.p, st);
  }

  // DGQ1 level viscosity: per-qp tables from the Q1 coefficients of every level
  void set_eta_q1(const std::vector<DBuf<double>> &coef, double lo, double hi) {
    drop_graphs();
    ietan_ok = false;
    eta_pq = 1;
    q1_src = &coef;
    q1_lo = lo;
    q1_hi = hi;
    for (int l = 0; l <= top; ++l) {
      GMGLevel<T> &g = *lv[l];
      const long long na = g.L.ntiles_total * cells_per_tile_of(g.L) * qp_per_cell_of(g.L);
      if (g.eta.n != na) {
        g.eta.alloc(na);
        v_fill<T>(g.eta.p, (T)1, na, st);
      }
      launch_eta_q1_eval<T>(g.L, coef[l].p, lo, hi, g.eta.p, st);
    }
  }

  void applyA(int l, Mode mode, const T *x, const T *rhs, T *dst) { applyA_work(l, mode, x, rhs, dst, split_for(l)); }
  // the same apply with the split path's per-cell work array supplied by the caller
  void applyA_work(int l, Mode mode, const T *x, const T *rhs, T *dst, T *work) {
    GMGLevel<T> &g = *lv[l];
    launch_cell_op<T>(laplace ? OP_LAPLACE : OP_ABLOCK, mode, g.L, x, nullptr, g.eta.p, eta_pq, rhs, nullptr, dst,
                      nullptr, st, laplace ? nullptr : work);
  }
  void applyS(int l, Mode mode, const T *x, const T *rhs, T *dst) {
    GMGLevel<T> &g = *lv[l];
    if (ietan_ok && !eta_pq && g.ietan.p && launch_mass_inv<T>(mode, g.L, x, g.ietan.p, rhs, dst, st)) return;
    launch_cell_op<T>(OP_MASS, mode, g.L, nullptr, x, g.eta.p, eta_pq, nullptr, rhs, nullptr, dst, st);
  }
  void apply(int kind, int l, Mode mode, const T *x, const T *rhs, T *dst) {
    if (kind == 0)
      applyA(l, mode, x, rhs, dst);
    else
      applyS(l, mode, x, rhs, dst);
  }
  double dot(const T *a, const T *b, long long n) {
    v_dot<T>(a, b, n, dot_work->p, dot_out->p, st);
    double h = 0;
    MFS_HIP(hipMemcpyAsync(&h, dot_out->p, sizeof(double), hipMemcpyDeviceToHost, st));
    MFS_HIP(hipStreamSynchronize(st));
    return h;
  }

  // Chebyshev eigenvalue estimate: preconditioned CG on D^{-1}A started from
  // the mean-free pattern (i mod 11); Lanczos tridiagonal from the CG
  // coefficients; lambda_max scaled by 1.2 (set-up at local_smoothing.cc:771-849).
  static double tridiag_eig(const std::vector<double> &d, const std::vector<double> &e, bool want_max) {
    double lo = 1e300, hi = -1e300;
    for (size_t i = 0; i < d.size(); ++i) {
      double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < d.size() ? std::fabs(e[i]) : 0.0);
      lo = std::min(lo, d[i] - r);
      hi = std::max(hi, d[i] + r);
    }
    const int target = want_max ? (int)d.size() : 1;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      int cnt = 0;
      double q = 1.0;
      for (size_t i = 0; i < d.size(); ++i) {
        q = d[i] - mid - (i ? e[i - 1] * e[i - 1] / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) ++cnt;
      }
      if (cnt >= target)
        hi = mid;
      else
        lo = mid;
      if (hi - lo <= 1e-15 * std::max(1.0, std::fabs(hi))) break;
    }
    return 0.5 * (lo + hi);
  }

  void estimate(int kind, int l, ChebInfo &ch, int max_its, double range) {
    GMGLevel<T> &g = *lv[l];
    const long long n = kind == 0 ? g.nu : g.np;
    T *dinv = kind == 0 ? g.dinvA.p : g.dinvS.p;
    T *r = kind == 0 ? g.ur.p : g.pr.p;
    T *z = kind == 0 ? g.ud.p : g.pd.p;
    T *p = kind == 0 ? g.ux.p : g.px.p;
    T *q = kind == 0 ? g.ub.p : g.pb.p;
    std::vector<T> pat(n);
    double mean = 0;
    for (long long i = 0; i < n; ++i) mean += (double)(i % 11);
    mean /= (double)n;
    for (long long i = 0; i < n; ++i) pat[i] = (T)((double)(i % 11) - mean);
    MFS_HIP(hipMemcpyAsync(r, pat.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
    std::vector<double> alphas, betas;
    double res = std::sqrt(dot(r, r, n));
    if (res >= 1e-2) {
      v_cheb_first<T>(dinv, r, z, p, (T)1, true, n, st);  // z = p = D^{-1} r
      double rz = dot(r, z, n);
      for (int it = 1; it <= max_its; ++it) {
        apply(kind, l, MODE_SET, p, nullptr, q);
        const double alpha = rz / dot(p, q, n);
        v_axpby<T>((T)(-alpha), q, (T)1, r, n, st);
        alphas.push_back(alpha);
        res = std::sqrt(dot(r, r, n));
        if (res < 1e-2 || it == max_its) break;
        v_mul<T>(dinv, r, z, n, st);
        const double rz2 = dot(r, z, n);
        const double beta = rz2 / rz;
        rz = rz2;
        v_axpby<T>((T)1, z, (T)beta, p, n, st);
        betas.push_back(beta);
      }
    }
    ch.its = (int)alphas.size();
    if (alphas.empty()) {
      ch.lmin = ch.lmax = 1.0;
    } else {
      const size_t m = alphas.size();
      std::vector<double> d(m), e(m > 1 ? m - 1 : 0);
      d[0] = 1.0 / alphas[0];
      for (size_t k = 1; k < m; ++k) d[k] = 1.0 / alphas[k] + betas[k - 1] / alphas[k - 1];
      for (size_t k = 0; k + 1 < m; ++k) e[k] = std::sqrt(betas[k]) / alphas[k];
      ch.lmin = tridiag_eig(d, e, false);
      ch.lmax = 1.2 * tridiag_eig(d, e, true);
    }
    const double a = range > 1.0 ? ch.lmax / range : std::min(0.9 * ch.lmax, ch.lmin);
    ch.delta = 0.5 * (ch.lmax - a);
    ch.theta = 0.5 * (ch.lmax + a);
  }

  // diagonals (build_preconditioner) and Chebyshev set-up (solve() :771-849)
  void build(const mfs_box &b) {
    drop_graphs();
    if (laplace) {
      build_laplace();
      return;
    }
    for (int l = 0; l <= top; ++l) {
      GMGLevel<T> &g = *lv[l];
      if (!eta_pq && g.L.k == 2 && !g.L.geo) {
        if (g.ietan.n != g.L.cells) g.ietan.alloc(g.L.cells);
        launch_inv_eta_nat<T>(g.L, g.eta.p, g.ietan.p, st);
      }
      ietan_ok = true;
      launch_cell_op<T>(OP_ADIAG, MODE_SET, g.L, nullptr, nullptr, g.eta.p, eta_pq, nullptr, nullptr, g.dinvA.p,
                        nullptr, st);
      launch_cell_op<T>(OP_MDIAG, MODE_SET, g.L, nullptr, nullptr, g.eta.p, eta_pq, nullptr, nullptr, nullptr,
                        g.dinvS.p, st);
      v_invert<T>(g.dinvA.p, g.nu, st);
      v_invert<T>(g.dinvS.p, g.np, st);
    }
    if constexpr (std::is_same<T, float>::value) {
      build_shadow(b);
      return;
    }
    (void)b;
    for (int l = 0; l <= top; ++l) {
      GMGLevel<T> &g = *lv[l];
      const bool coarse = l == 0;
      g.chA.degree = g.chS.degree = coarse ? 8 : 4;
      estimate(0, l, g.chA, coarse ? 100 : 10, coarse ? 1e-3 : 15.0);
      estimate(1, l, g.chS, coarse ? 100 : 10, coarse ? 1e-3 : 15.0);
    }
  }

  // Mesh-deformation smoothers (interface.cc:1263-1290): Laplace diagonals;
  // Chebyshev degree 5 over [1.2 lambda_max / 15, 1.2 lambda_max] from 10
  // eigenvalue iterations on levels > 0; on level 0 Chebyshev as a solver:
  // 50 iterations, range 1e-3, the degree from the Chebyshev error bound
  // (deal.II's degree = invalid_unsigned_int): with r = lambda_max / alpha,
  // sigma = (1 - r^-1/2) / (1 + r^-1/2), degree = 1 + floor(log(1/eps +
  // sqrt(1/eps^2 - 1)) / log(1/sigma)), eps = 1e-3
  void build_laplace() {
    for (int l = 0; l <= top; ++l) {
      GMGLevel<T> &g = *lv[l];
      launch_cell_op<T>(OP_LDIAG, MODE_SET, g.L, nullptr, nullptr, nullptr, 0, nullptr, nullptr, g.dinvA.p, nullptr,
                        st);
      v_invert<T>(g.dinvA.p, g.nu, st);
    }
    for (int l = 0; l <= top; ++l) {
      GMGLevel<T> &g = *lv[l];
      if (l > 0) {
        g.chA.degree = 5;
        estimate(0, l, g.chA, 10, 15.0);
      } else {
        estimate(0, 0, g.chA, 50, 1e-3);
        const double alpha = g.chA.theta - g.chA.delta;
        const double r = g.chA.lmax / alpha;
        const double sigma = (1.0 - std::sqrt(1.0 / r)) / (1.0 + std::sqrt(1.0 / r));
        const double eps = 1e-3;
        g.chA.degree =
            sigma > 0 ? 1 + (int)(std::log(1.0 / eps + std::sqrt(1.0 / eps / eps - 1.0)) / std::log(1.0 / sigma)) : 1;
      }
    }
  }

  // FP64 copy of the hierarchy: full set-up, Chebyshev intervals copied to
  // this hierarchy's levels, then everything but level 0 released
  void build_shadow(const mfs_box &b) {
    shadow.reset();
    auto sh = std::make_unique<GMG<double>>();
    sh->graphs = false;
    sh->setup(b, st, dot_work, dot_out);
    for (int l = 0; l <= top; ++l) sh->lv[l]->L.geo = lv[l]->L.geo;
    if (eta_pq)
      sh->set_eta_q1(*q1_src, q1_lo, q1_hi);
    else
      sh->set_eta(eta_src, eta_src_pq);
    sh->build(b);
    for (int l = 0; l <= top; ++l) {
      lv[l]->chA = sh->lv[l]->chA;
      lv[l]->chS = sh->lv[l]->chS;
    }
    MFS_HIP(hipStreamSynchronize(st));
    sh->lv.resize(1);
    sh->top = 0;
    sh->xwork.release();
    const LevelDesc &L0 = sh->lv[0]->L;
    sh->swork.release();
    if (split_work_size(L0) > 0) sh->swork.alloc(split_work_size(L0));
    shadow = std::move(sh);
  }

  // level-0 solve (the degree-8 Chebyshev coarse smoother, :855-860) in FP64
  void coarse_fp64(int kind, int steps) {
    GMGLevel<T> &g = *lv[0];
    GMGLevel<double> &c = *shadow->lv[0];
    const long long n = kind == 0 ? g.nu : g.np;
    double *cx = kind == 0 ? c.ux.p : c.px.p, *cb = kind == 0 ? c.ub.p : c.pb.p;
    v_cast<T, double>(kind == 0 ? g.ub.p : g.pb.p, cb, n, st);
    shadow->smooth(kind, 0, cx, cb, true, steps);
    v_cast<double, T>(cx, kind == 0 ? g.ux.p : g.px.p, n, st);
  }

  // Chebyshev semi-iteration (Saad, Alg. 12.1) with Jacobi preconditioning;
  // steps applications, the first from zero when x_zero.
  void smooth(int kind, int l, T *x, const T *b, bool x_zero, int steps) {
    GMGLevel<T> &g = *lv[l];
    const ChebInfo &ch = kind == 0 ? g.chA : g.chS;
    const long long n = kind == 0 ? g.nu : g.np;
    T *dinv = kind == 0 ? g.dinvA.p : g.dinvS.p;
    T *r = kind == 0 ? g.ur.p : g.pr.p;
    T *d = kind == 0 ? g.ud.p : g.pd.p;
    if (kind == 1 && ietan_ok && !eta_pq && g.ietan.p && g.L.k == 2 && !g.L.geo) {
      smooth_schur_fused(g, ch, x, b, x_zero, steps);
      return;
    }
    // residual + update in one pass where the A-block apply is split
    auto fused = [&](T f1, T f2, int first) {
      return kind == 0 && !laplace &&
             launch_ablock_cheb<T>(g.L, x, b, g.eta.p, eta_pq, dinv, d, f1, f2, first, split_for(l), st);
    };
    for (int s = 0; s < steps; ++s) {
      const bool zero = x_zero && s == 0;
      const T *res = b;
      if (zero) {
        v_cheb_first<T>(dinv, res, d, x, (T)(1.0 / ch.theta), true, n, st);
      } else if (!fused((T)0, (T)(1.0 / ch.theta), 1)) {
        apply(kind, l, MODE_RESID, x, b, r);
        res = r;
        v_cheb_first<T>(dinv, res, d, x, (T)(1.0 / ch.theta), false, n, st);
      }
      if (ch.degree < 2 || std::fabs(ch.delta) < 1e-40) continue;
      const double sigma = ch.theta / ch.delta;
      double rho_old = ch.delta / ch.theta;
      for (int k = 1; k < ch.degree; ++k) {
        const double rho = 1.0 / (2.0 * sigma - rho_old);
        const double f1 = rho * rho_old, f2 = 2.0 * rho / ch.delta;
        if (!fused((T)f1, (T)f2, 0)) {
          apply(kind, l, MODE_RESID, x, b, r);
          v_cheb_step<T>(dinv, r, d, x, (T)f1, (T)f2, n, st);
        }
        rho_old = rho;
      }
    }
  }

  // Schur smoother with the residual fused into every Chebyshev step
  // (launch_mass_cheb).  The step reads x around each node, so it writes the
  // other of two buffers: x and the level's residual buffer, which is free
  // during smoothing.  The first buffer is chosen so that the last step lands
  // in x.
  void smooth_schur_fused(GMGLevel<T> &g, const ChebInfo &ch, T *x, const T *b, bool x_zero, int steps) {
    const bool inner = ch.degree >= 2 && std::fabs(ch.delta) >= 1e-40;
    {  // a level that fits one workgroup: the whole sequence in one launch
      MassChebSeq seq{};
      const double inv_theta = 1.0 / ch.theta;
      bool fits = true;
      for (int s = 0; s < steps && fits; ++s) {
        auto push = [&](double f1, double f2, int first) {
          if (seq.n >= MassChebSeq::MAX) {
            fits = false;
            return;
          }
          seq.f1[seq.n] = f1;
          seq.f2[seq.n] = f2;
          seq.first[seq.n++] = first;
        };
        push(0.0, inv_theta, 1);
        if (!inner) continue;
        const double sigma = ch.theta / ch.delta;
        double rho_old = ch.delta / ch.theta;
        for (int k = 1; k < ch.degree; ++k) {
          const double rho = 1.0 / (2.0 * sigma - rho_old);
          push(rho * rho_old, 2.0 * rho / ch.delta, 0);
          rho_old = rho;
        }
      }
      if (fits && launch_mass_smooth_small<T>(g.L, x, b, g.dinvS.p, g.ietan.p, seq, x_zero, st)) return;
    }
    int swaps = 0;
    for (int s = 0; s < steps; ++s) swaps += (x_zero && s == 0 ? 0 : 1) + (inner ? ch.degree - 1 : 0);
    T *cur = x, *alt = g.pr.p;
    if (x_zero && (swaps & 1)) std::swap(cur, alt);
    const T inv_theta = (T)(1.0 / ch.theta);
    for (int s = 0; s < steps; ++s) {
      if (x_zero && s == 0) {
        v_cheb_first<T>(g.dinvS.p, b, g.pd.p, cur, inv_theta, true, g.np, st);
      } else {
        launch_mass_cheb<T>(g.L, cur, g.ietan.p, b, g.dinvS.p, g.pd.p, alt, (T)0, inv_theta, 1, st);
        std::swap(cur, alt);
      }
      if (!inner) continue;
      const double sigma = ch.theta / ch.delta;
      double rho_old = ch.delta / ch.theta;
      for (int k = 1; k < ch.degree; ++k) {
        const double rho = 1.0 / (2.0 * sigma - rho_old);
        const double f1 = rho * rho_old, f2 = 2.0 * rho / ch.delta;
        launch_mass_cheb<T>(g.L, cur, g.ietan.p, b, g.dinvS.p, g.pd.p, alt, (T)f1, (T)f2, 0, st);
        std::swap(cur, alt);
        rho_old = rho;
      }
    }
    if (cur != x) MFS_HIP(hipMemcpyAsync(x, cur, g.np * sizeof(T), hipMemcpyDeviceToDevice, st));
  }

  void vcycle_level(int kind, int l) {
    GMGLevel<T> &g = *lv[l];
    T *x = kind == 0 ? g.ux.p : g.px.p;
    T *b = kind == 0 ? g.ub.p : g.pb.p;
    T *r = kind == 0 ? g.ur.p : g.pr.p;
    const int steps = kind == 0 ? 1 : 4;
    if (l == 0) {
      if constexpr (std::is_same<T, float>::value) {
        if (shadow) {
          coarse_fp64(kind, steps);
          return;
        }
      }
      smooth(kind, 0, x, b, true, steps);
      return;
    }
    smooth(kind, l, x, b, true, steps);
    apply(kind, l, MODE_RESID, x, b, r);
    GMGLevel<T> &c = *lv[l - 1];
    launch_restrict<T>(kind, c.L, g.L, r, kind == 0 ? c.ub.p : c.pb.p, st, xwork.p);
    vcycle_level(kind, l - 1);
    launch_prolongate<T>(kind, c.L, g.L, kind == 0 ? c.ux.p : c.px.p, x, true, st, xwork.p);
    smooth(kind, l, x, b, false, steps);
  }

  // PreconditionMG::vmult in FP64 at the boundary, precision T inside
  void vmult(int kind, const double *src, double *dst) {
    GMGLevel<T> &g = *lv[top];
    const long long n = kind == 0 ? g.nu : g.np;
    v_cast<double, T>(src, input(kind), n, st);
    vmult_from_input(kind, dst);
  }
  // the fine-level right-hand side of a V-cycle (a caller may fill it directly)
  T *input(int kind) { return kind == 0 ? lv[top]->ub.p : lv[top]->pb.p; }
  // V-cycle on the right-hand side already in input(kind), result cast to dst
  void vmult_from_input(int kind, double *dst) {
    GMGLevel<T> &g = *lv[top];
    const long long n = kind == 0 ? g.nu : g.np;
    if (graphs) {
      if (!vexec[kind]) {
        hipGraph_t gr = nullptr;
        MFS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          vcycle_level(kind, top);
        } catch (...) {
          (void)hipStreamEndCapture(st, &gr);
          if (gr) (void)hipGraphDestroy(gr);
          throw;
        }
        MFS_HIP(hipStreamEndCapture(st, &gr));
        const hipError_t e = hipGraphInstantiate(&vexec[kind], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        MFS_HIP(e);
      }
      MFS_HIP(hipGraphLaunch(vexec[kind], st));
    } else {
      vcycle_level(kind, top);
    }
    v_cast<T, double>(kind == 0 ? g.ux.p : g.px.p, dst, n, st);
  }
};

}  // namespace

// ------------------------------------------------------------------ the handle
struct mfs_handle {
  mfs_box box{};
  int prec = MFS_GMG_FP64;
  hipStream_t st = nullptr;
  std::vector<LevelDesc> L;
  int top = 0;
  long long nu = 0, np = 0, n = 0;
  int nq = 0, ct = 0;
  DBuf<double> eta_nat, eta_active, ones_cell;
  std::vector<DBuf<double>> q1_coef;  // DGQ1 coefficients per level ("project to Q1")
  int active_per_qp = 0;
  bool have_eta = false, have_prec = false;
  double eta_min = 0, eta_max = 0;
  GMG<float> g32;
  GMG<double> g64;
  DBuf<double> dot_work, dot_out, hbuf;
  DBuf<double> w1, w2, w3, w4, w5, pnorm_w;
  std::vector<double *> vchunks, zchunks;
  int a_apps = 0, s_apps = 0;
  // free-surface stabilization of the active-level Stokes operator
  int fs_faces[6] = {0, 0, 0, 0, 0, 0};
  bool fs_on = false;
  DBuf<double> fs_stab, fs_sv;
  // mesh-deformation hierarchy (built for the mesh constraints of the last call)
  std::unique_ptr<GMG<double>> gmesh;
  int gmesh_bc[6] = {-1, -1, -1, -1, -1, -1};
  // general geometry: tile-blocked tables per level (empty: the uniform box)
  std::vector<DBuf<double>> geo_lv;
  // work array of the split FP64 Stokes apply on the active level (3D degree 2)
  DBuf<double> swork_stokes;

  ~mfs_handle() {
    for (double *p : vchunks) (void)hipFree(p);
    for (double *p : zchunks) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }

  const LevelDesc &Ltop() const { return L[top]; }

  void gmg_vmult(int kind, const double *src, double *dst) {
    if (prec == MFS_GMG_FP32)
      g32.vmult(kind, src, dst);
    else
      g64.vmult(kind, src, dst);
    if (kind == 0)
      ++a_apps;
    else
      ++s_apps;
  }

  double dot(const double *a, const double *b, long long nn) {
    v_dot<double>(a, b, nn, dot_work.p, dot_out.p, st);
    double h = 0;
    MFS_HIP(hipMemcpyAsync(&h, dot_out.p, sizeof(double), hipMemcpyDeviceToHost, st));
    MFS_HIP(hipStreamSynchronize(st));
    return h;
  }

  // ---- active-level FP64 operators
  // 3D degree 2 on the box: the Stokes apply runs split (one plane-kernel
  // launch over all tiles, then node-wise sums) with this work array
  double *stokes_work() {
    const LevelDesc &T = Ltop();
    if (T.dim != 3 || T.k != 2 || T.geo) return nullptr;
    const long long n = stokes_split_work_size(T);
    if (swork_stokes.n != n) swork_stokes.alloc(n);
    return swork_stokes.p;
  }
  void stokes(Mode mode, const double *x, const double *rhs, double *dst) {
    launch_cell_op<double>(OP_STOKES, mode, Ltop(), x, x + nu, eta_active.p, active_per_qp, rhs,
                           rhs ? rhs + nu : nullptr, dst, dst + nu, st, stokes_work());
    if (fs_on) launch_free_surface(Ltop(), fs_faces, fs_stab.p, x, fs_sv.p, mode == MODE_RESID ? -1.0 : 1.0, dst, st);
  }
  void ablock_active(Mode mode, const double *x, const double *rhs, double *dst) {
    launch_cell_op<double>(OP_ABLOCK, mode, Ltop(), x, nullptr, eta_active.p, active_per_qp, rhs, nullptr, dst,
                           nullptr, st);
  }
  void mass_active(Mode mode, const double *x, const double *rhs, double *dst) {
    launch_cell_op<double>(OP_MASS, mode, Ltop(), nullptr, x, eta_active.p, active_per_qp, nullptr, rhs, nullptr,
                           dst, st);
  }

  // BlockSchurPreconditioner::vmult, cheap: p = -S^{-1} r_p; t = r_u - B^T p; u = A^{-1} t
  // (degree 2: t is written straight into the V-cycle's fine-level right-hand
  // side in the V-cycle precision, which is the FP64 t rounded once, as the
  // cast at the V-cycle entry would)
  void block_prec_cheap(const double *src, double *dst) {
    gmg_vmult(1, src + nu, w5.p);
    v_scale_neg_copy<double>(w5.p, dst + nu, np, st);
    const bool fused = prec == MFS_GMG_FP32
                           ? launch_bt_nodes<double, float>(MODE_RESID, Ltop(), dst + nu, src, g32.input(0), st)
                           : launch_bt_nodes<double, double>(MODE_RESID, Ltop(), dst + nu, src, g64.input(0), st);
    if (fused) {
      if (prec == MFS_GMG_FP32)
        g32.vmult_from_input(0, dst);
      else
        g64.vmult_from_input(0, dst);
      ++a_apps;
      return;
    }
    launch_cell_op<double>(OP_BT, MODE_RESID, Ltop(), nullptr, dst + nu, nullptr, 0, src, nullptr, w5.p, nullptr,
                           st);
    gmg_vmult(0, w5.p, dst);
  }

  // preconditioned CG (inner solves of the expensive preconditioner)
  int cg(int kind, const double *b, double *x, long long nn, double tol, int maxit, double *r, double *z, double *p,
         double *q) {
    auto A = [&](const double *in, double *out) {
      if (kind == 0)
        ablock_active(MODE_SET, in, nullptr, out);
      else
        mass_active(MODE_SET, in, nullptr, out);
    };
    auto P = [&](const double *in, double *out) {
      if (prec == MFS_GMG_FP32)
        g32.vmult(kind, in, out);
      else
        g64.vmult(kind, in, out);
    };
    v_fill<double>(x, 0.0, nn, st);
    MFS_HIP(hipMemcpyAsync(r, b, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    double res = std::sqrt(dot(r, r, nn));
    if (res <= tol) return 0;
    P(r, z);
    double rz = dot(r, z, nn);
    MFS_HIP(hipMemcpyAsync(p, z, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int it = 1; it <= maxit; ++it) {
      A(p, q);
      const double alpha = rz / dot(p, q, nn);
      v_axpby<double>(alpha, p, 1.0, x, nn, st);
      v_axpby<double>(-alpha, q, 1.0, r, nn, st);
      res = std::sqrt(dot(r, r, nn));
      if (res <= tol) return it;
      P(r, z);
      const double rz2 = dot(r, z, nn);
      const double beta = rz2 / rz;
      rz = rz2;
      v_axpby<double>(1.0, z, beta, p, nn, st);
    }
    throw std::runtime_error("inner CG solve of the expensive Stokes preconditioner did not converge");
  }

  // BlockSchurPreconditioner::vmult, expensive: inner CG solves to relative tolerances
  double a_tol = 1e-2, s_tol = 1e-6;
  DBuf<double> cu1, cu2, cu3, cu4, cp1, cp2, cp3, cp4;
  void block_prec_expensive(const double *src, double *dst) {
    const double nsp = std::sqrt(dot(src + nu, src + nu, np));
    if (nsp > 1e-50) {
      s_apps += cg(1, src + nu, w5.p, np, nsp * s_tol, 100, cp1.p, cp2.p, cp3.p, cp4.p);
      v_scale_neg_copy<double>(w5.p, dst + nu, np, st);
    } else {
      v_fill<double>(dst + nu, 0.0, np, st);
    }
    launch_cell_op<double>(OP_BT, MODE_RESID, Ltop(), nullptr, dst + nu, nullptr, 0, src, nullptr, w5.p, nullptr,
                           st);
    const double nt = std::sqrt(dot(w5.p, w5.p, nu));
    a_apps += cg(0, w5.p, dst, nu, nt * a_tol, 10000, cu1.p, cu2.p, cu3.p, cu4.p);
  }

  double *vec(std::vector<double *> &chunks, int i) {
    const int c = i / 8;
    while ((int)chunks.size() <= c) {
      double *p = nullptr;
      MFS_HIP(hipMalloc(reinterpret_cast<void **>(&p), (size_t)8 * n * sizeof(double)));
      chunks.push_back(p);
    }
    return chunks[c] + (long long)(i % 8) * n;
  }

  // vectors [i0, i0 + MV_MAX) of a chunked basis
  VecSet vecset(std::vector<double *> &chunks, int i0, int m) {
    VecSet V{};
    for (int c = 0; c * 8 < m; ++c) V.p[c] = vec(chunks, i0 + 8 * c);
    return V;
  }
  // w += sign * sum_{i<m} hdev[i] B_i over a chunked basis
  void basis_axpy(std::vector<double *> &chunks, int m, const double *hdev, double *w, double sign) {
    for (int i0 = 0; i0 < m; i0 += MV_MAX) {
      const int mm = std::min(MV_MAX, m - i0);
      v_maxpy_all(vecset(chunks, i0, mm), n, mm, hdev + i0, w, n, sign, st);
    }
  }

  // h = V[0..m)^T w ; w -= V h   (one classical Gram-Schmidt pass), h to host
  void cgs_pass(int m, double *w, std::vector<double> &h) {
    for (int i0 = 0; i0 < m; i0 += MV_MAX) {
      const int mm = std::min(MV_MAX, m - i0);
      v_mdot_all(vecset(vchunks, i0, mm), n, mm, w, n, dot_work.p + (long long)i0 * dot_work_size(), st);
    }
    v_sum_partials(dot_work.p, m, hbuf.p, st);
    basis_axpy(vchunks, m, hbuf.p, w, -1.0);
    h.resize(m);
    MFS_HIP(hipMemcpyAsync(h.data(), hbuf.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    MFS_HIP(hipStreamSynchronize(st));
  }

  static constexpr double REORTH = 0.70710678118654752;  // reorthogonalize when |w'| < |w| / sqrt(2)

  struct KrylovStatus {
    bool converged = false;
    int its = 0;
    double res = 0;
  };

  // restarted right-preconditioned GMRES (flexible=false) or FGMRES (flexible=true)
  template <typename Prec>
  KrylovStatus gmres(const double *b, double *x, double tol, int maxit, int restart, bool flexible, Prec P) {
    KrylovStatus ks;
    int acc = 0;
    std::vector<double> h, hh;
    while (true) {
      double *v0 = vec(vchunks, 0);
      stokes(MODE_RESID, x, b, v0);
      double rho = std::sqrt(dot(v0, v0, n));
      ks.res = rho;
      if (acc == 0) record(rho);
      if (rho <= tol) {
        ks.converged = true;
        ks.its = acc;
        return ks;
      }
      if (acc >= maxit) {
        ks.its = acc;
        return ks;
      }
      v_scale(v0, 1.0 / rho, n, st);
      std::vector<std::vector<double>> H;
      std::vector<double> cs, sn, g(1, rho);
      bool done = false;
      int j = 0;
      for (; j < restart; ++j) {
        ++acc;
        double *z = flexible ? vec(zchunks, j) : w1.p;
        P(vec(vchunks, j), z);
        double *w = vec(vchunks, j + 1);
        stokes(MODE_SET, z, nullptr, w);
        std::vector<double> col(j + 2, 0.0);
        // classical Gram-Schmidt, repeated once when the pass cancelled more
        // than 1 - 1/sqrt(2) of the norm (Daniel-Gragg-Kaufman-Stewart test)
        const double wn0 = std::sqrt(dot(w, w, n));
        cgs_pass(j + 1, w, h);
        for (int i = 0; i <= j; ++i) col[i] = h[i];
        double hn = std::sqrt(dot(w, w, n));
        if (hn < REORTH * wn0) {
          cgs_pass(j + 1, w, hh);
          for (int i = 0; i <= j; ++i) col[i] += hh[i];
          hn = std::sqrt(dot(w, w, n));
        }
        col[j + 1] = hn;
        for (int i = 0; i < j; ++i) {
          const double a = col[i], bb = col[i + 1];
          col[i] = cs[i] * a + sn[i] * bb;
          col[i + 1] = -sn[i] * a + cs[i] * bb;
        }
        const double den = std::hypot(col[j], col[j + 1]);
        const double c = den > 0 ? col[j] / den : 1.0, s = den > 0 ? col[j + 1] / den : 0.0;
        cs.push_back(c);
        sn.push_back(s);
        col[j] = den;
        col[j + 1] = 0.0;
        g.push_back(-s * g[j]);
        g[j] = c * g[j];
        H.push_back(col);
        rho = std::fabs(g[j + 1]);
        ks.res = rho;
        record(rho);
        if (hn > 0) v_scale(w, 1.0 / hn, n, st);
        if (rho <= tol || hn == 0.0) {
          ks.converged = true;
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
      const int m = j;
      std::vector<double> y(m, 0.0);
      for (int i = m - 1; i >= 0; --i) {
        double s = g[i];
        for (int k = i + 1; k < m; ++k) s -= H[k][i] * y[k];
        y[i] = s / H[i][i];
      }
      MFS_HIP(hipMemcpyAsync(hbuf.p, y.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
      if (flexible) {
        basis_axpy(zchunks, m, hbuf.p, x, 1.0);
      } else {
        v_fill<double>(w2.p, 0.0, n, st);
        basis_axpy(vchunks, m, hbuf.p, w2.p, 1.0);
        P(w2.p, w3.p);
        v_axpby<double>(1.0, w3.p, 1.0, x, n, st);
      }
      MFS_HIP(hipStreamSynchronize(st));
      if (done) {
        ks.its = acc;
        return ks;
      }
    }
  }

  // Restarted flexible GMRES (x += Z y) whose Arnoldi process uses delayed
  // classical Gram-Schmidt (Swirydowicz et al., Numer. Linear Algebra Appl.
  // 28(2), 2021; Bielich et al., Parallel Comput. 112, 2022): step j computes
  // w = A P v_j with v_j still provisional (one projection pass, approximate
  // norm), then ONE reduction gives V^T v_j, V^T w and the three inner
  // products of v_j and w; from them v_j is reorthogonalized (the second
  // Gram-Schmidt pass, one step late), column j-1 of the Hessenberg matrix is
  // corrected and finalized, and w is projected; ONE sweep updates v_j and w.
  // Same iterates as GMRES with classical Gram-Schmidt twice, at half its
  // basis traffic.  Restart lengths up to MV_MAX - 1.
  template <typename Prec>
  KrylovStatus fgmres_dcgs(const double *b, double *x, double tol, int maxit, int restart, Prec P) {
    KrylovStatus ks;
    int acc = 0;
    std::vector<double> red, coef;
    while (true) {
      double *v0 = vec(vchunks, 0);
      stokes(MODE_RESID, x, b, v0);
      double rho = std::sqrt(dot(v0, v0, n));
      ks.res = rho;
      if (acc == 0) record(rho);
      if (rho <= tol) {
        ks.converged = true;
        ks.its = acc;
        return ks;
      }
      if (acc >= maxit) {
        ks.its = acc;
        return ks;
      }
      v_scale(v0, 1.0 / rho, n, st);
      std::vector<std::vector<double>> H;  // H[c]: column c, rows 0..c+1 (finalized columns are rotated)
      std::vector<double> cs, sn, g(1, rho);
      // Givens rotations of a column against the first c rotations, then its own
      auto rotate = [&](std::vector<double> col, int c, std::vector<double> &cs_, std::vector<double> &sn_,
                        std::vector<double> &g_) {
        for (int i = 0; i < c; ++i) {
          const double a = col[i], bb = col[i + 1];
          col[i] = cs_[i] * a + sn_[i] * bb;
          col[i + 1] = -sn_[i] * a + cs_[i] * bb;
        }
        const double den = std::hypot(col[c], col[c + 1]);
        const double cc = den > 0 ? col[c] / den : 1.0, ss = den > 0 ? col[c + 1] / den : 0.0;
        cs_.push_back(cc);
        sn_.push_back(ss);
        col[c] = den;
        col[c + 1] = 0.0;
        g_.push_back(-ss * g_[c]);
        g_[c] = cc * g_[c];
        return col;
      };
      bool done = false, provisional_conv = false;
      int m = 0;  // finalized columns
      for (int j = 0;; ++j) {
        double *vj = vec(vchunks, j);
        const bool at_limit = j == restart || acc >= maxit;
        const bool finalize_only = at_limit || provisional_conv;
        double *w = vj;
        if (!finalize_only) {
          ++acc;
          double *z = vec(zchunks, j);
          P(vj, z);
          w = vec(vchunks, j + 1);
          stokes(MODE_SET, z, nullptr, w);
        }
        const VecSet V = vecset(vchunks, 0, j);
        v_dcgs_reduce(V, n, j, vj, w, n, dot_work.p, hbuf.p, st);
        const int rows = 2 * j + 3;
        red.resize(rows);
        MFS_HIP(hipMemcpyAsync(red.data(), hbuf.p, rows * sizeof(double), hipMemcpyDeviceToHost, st));
        MFS_HIP(hipStreamSynchronize(st));
        const double *sv = red.data(), *tv = red.data() + j;
        const double alpha = red[2 * j], beta = red[2 * j + 1], gamma = red[2 * j + 2];
        double ss = 0, st_ = 0, tt = 0;
        for (int k = 0; k < j; ++k) {
          ss += sv[k] * sv[k];
          st_ += sv[k] * tv[k];
          tt += tv[k] * tv[k];
        }
        // the first basis vector is normalized exactly
        const double nu = j == 0 ? 1.0 : std::sqrt(std::max(alpha - ss, 0.0));
        if (j > 0 && m < j) {
          // finalize column j-1: its component along the provisional v_j
          // splits into V s and nu v_j(final)
          std::vector<double> &col = H[j - 1];
          const double hj = col[j];
          for (int k = 0; k < j; ++k) col[k] += hj * sv[k];
          col[j] = hj * nu;
          col = rotate(col, j - 1, cs, sn, g);
          ks.res = std::fabs(g[j]);
          record(ks.res);
          m = j;
          if (ks.res <= tol || nu == 0.0) {
            ks.converged = true;
            done = true;
          }
        }
        if (done || at_limit) break;
        if (finalize_only) {  // the provisional estimate was too optimistic: do the full step now
          provisional_conv = false;
          --j;
          continue;
        }
        // provisional column j and the next provisional basis vector
        const double hjj = (beta - st_) / nu;
        std::vector<double> col(j + 2);
        for (int k = 0; k < j; ++k) col[k] = tv[k];
        col[j] = hjj;
        const double r2 = gamma - tt - hjj * hjj;
        const double rn = std::sqrt(std::max(r2, 1e-30 * gamma));
        col[j + 1] = rn;
        H.push_back(col);
        coef.assign(rows, 0.0);
        for (int k = 0; k < j; ++k) {
          coef[k] = -sv[k] / nu;
          coef[j + k] = -(tv[k] - sv[k] * hjj / nu) / rn;
        }
        coef[2 * j] = 1.0 / nu;
        coef[2 * j + 1] = 1.0 / rn;
        coef[2 * j + 2] = -hjj / (nu * rn);
        MFS_HIP(hipMemcpyAsync(hbuf.p, coef.data(), rows * sizeof(double), hipMemcpyHostToDevice, st));
        v_dcgs_update(V, n, j, hbuf.p, vj, w, n, st);
        // residual estimate with the provisional column decides whether the
        // next step only finalizes
        std::vector<double> cs2 = cs, sn2 = sn, g2 = g;
        rotate(col, j, cs2, sn2, g2);
        provisional_conv = std::fabs(g2[j + 1]) <= tol;
      }
      // least squares with the m finalized columns, x += Z y
      std::vector<double> y(m, 0.0);
      for (int i = m - 1; i >= 0; --i) {
        double sum = g[i];
        for (int k = i + 1; k < m; ++k) sum -= H[k][i] * y[k];
        y[i] = sum / H[i][i];
      }
      if (m > 0) {
        MFS_HIP(hipMemcpyAsync(hbuf.p, y.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
        basis_axpy(zchunks, m, hbuf.p, x, 1.0);
      }
      MFS_HIP(hipStreamSynchronize(st));
      if (done) {
        ks.its = acc;
        return ks;
      }
    }
  }

  // IDR(s) with biorthogonalization (van Gijzen and Sonneveld, ACM TOMS 38(1),
  // 2011, Algorithm 913), deal.II's SolverIDR scheme for the cheap steps
  // (stokes_matrix_free_local_smoothing.cc:1291-1301).  One step = s inner
  // updates + one dimension-reduction update, one preconditioner application
  // each; convergence on the recursive residual after every update.  x and r
  // are always updated with the same U_k and G_k = A U_k, so the residual stays
  // consistent with x whatever the precision of the preconditioner.
  template <typename Prec>
  KrylovStatus idrs(const double *b, double *x, double tol, int maxit, int s, Prec P) {
    KrylovStatus ks;
    double *r = vec(vchunks, 3 * s), *v = vec(vchunks, 3 * s + 1), *uhat = vec(vchunks, 3 * s + 2),
           *vhat = vec(vchunks, 3 * s + 3);
    auto Q = [&](int i) { return vec(vchunks, i); };
    auto G = [&](int i) { return vec(vchunks, s + i); };
    auto U = [&](int i) { return vec(vchunks, 2 * s + i); };
    const size_t bytes = (size_t)n * sizeof(double);
    stokes(MODE_RESID, x, b, r);
    double res = std::sqrt(dot(r, r, n));
    ks.res = res;
    record(res);
    if (res <= tol) {
      ks.converged = true;
      return ks;
    }
    for (int k = 0; k < s; ++k) {  // shadow space, modified Gram-Schmidt
      v_idr_shadow(Q(k), k, n, st);
      for (int j = 0; j < k; ++j) v_axpby<double>(-dot(Q(k), Q(j), n), Q(j), 1.0, Q(k), n, st);
      v_scale(Q(k), 1.0 / std::sqrt(dot(Q(k), Q(k), n)), n, st);
      v_fill<double>(G(k), 0.0, n, st);
      v_fill<double>(U(k), 0.0, n, st);
    }
    std::vector<double> M(s * s, 0.0), f(s), c(s);
    for (int i = 0; i < s; ++i) M[i * s + i] = 1.0;
    double omega = 1.0;
    int step = 0;
    while (true) {
      ++step;
      for (int i = 0; i < s; ++i) f[i] = dot(Q(i), r, n);
      for (int k = 0; k < s; ++k) {
        for (int i = k; i < s; ++i) {  // M(k:s, k:s) is lower triangular
          double a = f[i];
          for (int j = k; j < i; ++j) a -= M[i * s + j] * c[j];
          c[i] = a / M[i * s + i];
        }
        MFS_HIP(hipMemcpyAsync(v, r, bytes, hipMemcpyDeviceToDevice, st));
        for (int i = k; i < s; ++i) v_axpby<double>(-c[i], G(i), 1.0, v, n, st);
        P(v, vhat);
        MFS_HIP(hipMemcpyAsync(uhat, vhat, bytes, hipMemcpyDeviceToDevice, st));
        v_scale(uhat, omega, n, st);
        for (int i = k; i < s; ++i) v_axpby<double>(c[i], U(i), 1.0, uhat, n, st);
        stokes(MODE_SET, uhat, nullptr, G(k));
        for (int i = 0; i < k; ++i) {  // G_k orthogonal to Q_0..Q_{k-1}
          const double alpha = dot(Q(i), G(k), n) / M[i * s + i];
          v_axpby<double>(-alpha, G(i), 1.0, G(k), n, st);
          v_axpby<double>(-alpha, U(i), 1.0, uhat, n, st);
        }
        MFS_HIP(hipMemcpyAsync(U(k), uhat, bytes, hipMemcpyDeviceToDevice, st));
        for (int i = k; i < s; ++i) M[i * s + k] = dot(Q(i), G(k), n);
        if (M[k * s + k] == 0.0) {  // breakdown: leave it to the expensive phase
          ks.its = step;
          return ks;
        }
        const double beta = f[k] / M[k * s + k];
        v_axpby<double>(-beta, G(k), 1.0, r, n, st);
        v_axpby<double>(beta, U(k), 1.0, x, n, st);
        res = std::sqrt(dot(r, r, n));
        ks.res = res;
        record(res);
        if (res <= tol) {
          ks.converged = true;
          ks.its = step;
          return ks;
        }
        if (step >= maxit) {
          ks.its = step;
          return ks;
        }
        for (int i = k + 1; i < s; ++i) f[i] -= beta * M[i * s + k];
      }
      // dimension reduction
      P(r, uhat);
      stokes(MODE_SET, uhat, nullptr, v);
      const double vv = dot(v, v, n);
      if (vv == 0.0) {
        ks.its = step;
        return ks;
      }
      omega = dot(v, r, n) / vv;
      v_axpby<double>(-omega, v, 1.0, r, n, st);
      v_axpby<double>(omega, uhat, 1.0, x, n, st);
      res = std::sqrt(dot(r, r, n));
      ks.res = res;
      record(res);
      if (res <= tol) {
        ks.converged = true;
        ks.its = step;
        return ks;
      }
      if (step >= maxit) {
        ks.its = step;
        return ks;
      }
    }
  }

  // the handler's solve (device vectors)
  int solve_dev(const double *rhs, const double *x0, double *sol, const mfs_solve_params &prm, mfs_solve_result &res) {
    if (!have_eta) throw std::runtime_error("mfs_solve: viscosity not set (call mfs_set_viscosity)");
    if (!have_prec) throw std::runtime_error("mfs_solve: preconditioner not built (call mfs_build_preconditioner)");
    if (prm.restart < 1 || prm.restart > 1000) throw std::runtime_error("mfs_solve: restart must be in [1, 1000]");
    if (prm.n_cheap_steps < 0 || prm.n_expensive_steps < 0) throw std::runtime_error("mfs_solve: negative step limit");
    if (prm.cheap_krylov != MFS_KRYLOV_GMRES && prm.cheap_krylov != MFS_KRYLOV_IDRS)
      throw std::runtime_error("mfs_solve: unknown Krylov method for the cheap steps");
    if (prm.cheap_krylov == MFS_KRYLOV_IDRS && (prm.idr_s < 1 || prm.idr_s > 8))
      throw std::runtime_error("mfs_solve: IDR(s) parameter must be in [1, 8]");
    std::memset(&res, 0, sizeof(res));
    const double s = box.pressure_scaling;
    a_apps = s_apps = 0;
    a_tol = prm.a_block_tolerance;
    s_tol = prm.s_block_tolerance;
    if (prm.failure_strategy != MFS_FAILURE_ABORT && prm.failure_strategy != MFS_FAILURE_CONTINUE)
      throw std::runtime_error("mfs_solve: unknown linear solver failure strategy");
    const bool newton = prm.solve_newton_system != 0;
    if (newton || !x0) {
      // Newton update: the best guess is zero (:950-969)
      v_fill<double>(sol, 0.0, n, st);
    } else {
      // initial guess: constrained velocity entries zeroed, pressure / s (:944-975)
      MFS_HIP(hipMemcpyAsync(sol, x0, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      v_zero_constrained<double>(Ltop(), sol, st);
      v_scale(sol + nu, 1.0 / s, np, st);
    }
    if (newton) {
      // tolerance and initial nonlinear residual from the Newton right-hand side (:1028-1041)
      res.r0_norm = std::sqrt(dot(rhs, rhs, nu) + dot(rhs + nu, rhs + nu, np));
      res.initial_nonlinear_residual = res.r0_norm;
    } else {
      // residuals for the nonlinear residual and the solver tolerance (:977-1025)
      stokes(MODE_RESID, sol, rhs, w1.p);
      res.initial_nonlinear_residual = std::sqrt(dot(w1.p, w1.p, n));
      MFS_HIP(hipMemcpyAsync(w2.p, sol, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      v_fill<double>(w2.p, 0.0, nu, st);
      stokes(MODE_RESID, w2.p, rhs, w1.p);
      const double ru2 = dot(w1.p, w1.p, nu);
      const double rp2 = dot(rhs + nu, rhs + nu, np);
      res.r0_norm = std::sqrt(ru2 + rp2);
    }
    const double tol = prm.tolerance * res.r0_norm;
    res.solver_tolerance = tol;
    hipEvent_t e0, e1;
    MFS_HIP(hipEventCreate(&e0));
    MFS_HIP(hipEventCreate(&e1));
    MFS_HIP(hipEventRecord(e0, st));
    bool ok = false;
    std::vector<double> hist_cheap, hist_exp;  // residual history per phase (SolverControl history data)
    if (prm.n_cheap_steps > 0) {
      history = &hist_cheap;
      // The cheap phase stores the preconditioned vectors (flexible update
      // x += Z y): with an FP32 V-cycle the preconditioner is linear only up to
      // FP32 rounding, and the flexible form keeps the residual estimate exact.
      const bool flexible = prec == MFS_GMG_FP32;
      auto pc = [&](const double *a, double *b) { block_prec_cheap(a, b); };
      KrylovStatus ks = prm.cheap_krylov == MFS_KRYLOV_IDRS ? idrs(rhs, sol, tol, prm.n_cheap_steps, prm.idr_s, pc)
                        : prm.restart < MV_MAX ? fgmres_dcgs(rhs, sol, tol, prm.n_cheap_steps, prm.restart, pc)
                                               : gmres(rhs, sol, tol, prm.n_cheap_steps, prm.restart, flexible, pc);
      res.cheap_iterations = ks.its;
      res.final_linear_residual = ks.res;
      res.cheap_final_residual = ks.res;
      ok = ks.converged;
      history = nullptr;
    }
    hipEvent_t em;
    MFS_HIP(hipEventCreate(&em));
    MFS_HIP(hipEventRecord(em, st));
    if (!ok && prm.n_expensive_steps > 0) {
      for (DBuf<double> *v : {&cu1, &cu2, &cu3, &cu4})
        if (v->n != nu) v->alloc(nu);
      for (DBuf<double> *v : {&cp1, &cp2, &cp3, &cp4})
        if (v->n != np) v->alloc(np);
      auto pe = [&](const double *a, double *b) { block_prec_expensive(a, b); };
      history = &hist_exp;
      KrylovStatus ks = prm.restart < MV_MAX ? fgmres_dcgs(rhs, sol, tol, prm.n_expensive_steps, prm.restart, pe)
                                             : gmres(rhs, sol, tol, prm.n_expensive_steps, prm.restart, true, pe);
      res.expensive_iterations = ks.its;
      res.final_linear_residual = ks.res;
      ok = ks.converged;
      history = nullptr;
    }
    MFS_HIP(hipEventRecord(e1, st));
    MFS_HIP(hipEventSynchronize(e1));
    float ms = 0, msc = 0;
    MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    MFS_HIP(hipEventElapsedTime(&msc, e0, em));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(em);
    res.solve_seconds = ms * 1e-3;
    res.cheap_seconds = msc * 1e-3;
    res.a_preconditioner_applications = a_apps;
    res.s_preconditioner_applications = s_apps;
    res.converged = ok ? 1 : 0;
    // post-steps (:1406-1445): pressure *= s, volume normalization (not for a Newton update)
    v_scale(sol + nu, s, np, st);
    if (!newton) {
      const double ip = dot(sol + nu, pnorm_w.p, np);
      const double vol = dot(pnorm_w.p, ones_p(), np);
      const double adj = ip / vol;
      v_fill<double>(w3.p, 1.0, np, st);
      v_axpby<double>(-adj, w3.p, 1.0, sol + nu, np, st);
      res.pressure_normalization_adjustment = -adj;
    }
    MFS_HIP(hipStreamSynchronize(st));
    if (ok) return MFS_OK;
    // Linear solver failure strategy (:1353-1395)
    if (prm.failure_strategy == MFS_FAILURE_CONTINUE) return MFS_OK;
    g_err = failure_message(prm, res, hist_cheap, hist_exp);
    return MFS_NO_CONVERGENCE;
  }

  // Utilities::throw_linear_solver_failure_exception (source/utilities.cc:2828-2886):
  // the residual history of every phase that ran goes to history_file, the
  // message states the initial, final and required residuals
  static std::string failure_message(const mfs_solve_params &prm, const mfs_solve_result &res,
                                     const std::vector<double> &hc, const std::vector<double> &he) {
    std::vector<const std::vector<double> *> phases;
    if (prm.n_cheap_steps > 0) phases.push_back(&hc);
    if (prm.n_expensive_steps > 0) phases.push_back(&he);
    char buf[512];
    std::string msg =
        "The iterative Stokes solver in StokesMatrixFreeHandler::solve did not converge.\n\n";
    if (!phases.empty() && !phases.front()->empty()) {
      std::snprintf(buf, sizeof(buf), "The initial residual was: %e\n", phases.front()->front());
      msg += buf;
    }
    if (!phases.empty() && !phases.back()->empty()) {
      std::snprintf(buf, sizeof(buf), "The final residual is: %e\n", phases.back()->back());
      msg += buf;
    }
    std::snprintf(buf, sizeof(buf), "The required residual for convergence is: %e\n", res.solver_tolerance);
    msg += buf;
    if (prm.history_file && prm.history_file[0]) {
      if (FILE *f = std::fopen(prm.history_file, "w")) {
        bool first = true;
        for (const std::vector<double> *ph : phases) {
          if (ph->empty()) continue;
          if (!first) std::fputc('\n', f);
          first = false;
          for (size_t j = 0; j < ph->size(); ++j) std::fprintf(f, "%zu %.17g\n", j, (*ph)[j]);
        }
        std::fclose(f);
        msg += std::string("See ") + prm.history_file + " for the full convergence history.\n";
      } else {
        msg += std::string("The convergence history could not be written to ") + prm.history_file + ".\n";
      }
    }
    std::snprintf(buf, sizeof(buf),
                  "\nThe solver reported the following error:\nIterative method reported convergence failure in "
                  "step %d. The residual in the last step was %e.",
                  res.cheap_iterations + res.expensive_iterations, res.final_linear_residual);
    msg += buf;
    return msg;
  }
  std::vector<double> *history = nullptr;
  void record(double r) {
    if (history) history->push_back(r);
  }
  DBuf<double> ones_p_buf;
  const double *ones_p() {
    if (ones_p_buf.n != np) {
      ones_p_buf.alloc(np);
      v_fill<double>(ones_p_buf.p, 1.0, np, st);
    }
    return ones_p_buf.p;
  }
};

namespace {

// Work array of the allocation-free A-block apply.  The name below is a
// host-side stand-in for the device array: host code uses only the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch()).
constexpr long long kNoAllocScratchBytes = MFS_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of scratch on the current device, resolved once per device
// (at mfs_create, or at the first allocation-free call on a device).
void *noalloc_scratch() {
  int dev = 0;
  MFS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// Support points of FE_Q(k) on [0, 1] (Gauss-Lobatto; k <= 3 in closed form)
std::vector<double> fe_q_support(int k) {
  if (k == 1) return {0.0, 1.0};
  if (k == 2) return {0.0, 0.5, 1.0};
  const double a = 0.5 * (1.0 - 1.0 / std::sqrt(5.0));
  return {0.0, a, 1.0 - a, 1.0};
}

// Lattice index along one direction of coordinate x on nc cells of size h
// with FE_Q(k) nodes pts: k c + i for the node at cell c, local point i
// (points within 1e-6 cell sizes); -1 if x is not a node
long long lattice_node(double x, double h, long long nc, const std::vector<double> &pts) {
  const double t = x / h;
  if (!(t > -1e-6 && t < nc + 1e-6)) return -1;
  long long c = (long long)std::floor(t + 1e-6);
  if (c >= nc) c = nc - 1;
  if (c < 0) c = 0;
  const double xi = t - (double)c;
  const int k = (int)pts.size() - 1;
  for (int i = 0; i <= k; ++i)
    if (std::fabs(xi - pts[i]) < 1e-6) return k * c + i;
  return -1;
}

void create(const mfs_box *b, int prec, mfs_handle **out) {
  if (!b || !out) throw std::runtime_error("mfs_create: null argument");
  if (b->dim != 2 && b->dim != 3) throw std::runtime_error("mfs_create: dim must be 2 or 3");
  if (b->degree != 2 && b->degree != 3) throw std::runtime_error("mfs_create: degree must be 2 or 3");
  if (b->n_levels < 1 || b->n_levels > 16) throw std::runtime_error("mfs_create: n_levels out of range");
  for (int d = 0; d < b->dim; ++d) {
    if (b->coarse[d] < 1) throw std::runtime_error("mfs_create: coarse cells must be >= 1");
    if (!(b->extent[d] > 0)) throw std::runtime_error("mfs_create: extents must be positive");
    if (((long long)b->coarse[d] << (b->n_levels - 1)) > (1LL << 20))
      throw std::runtime_error("mfs_create: level too large");
  }
  for (int f = 0; f < 2 * b->dim; ++f)
    if (b->bc[f] < 0 || b->bc[f] > 2) throw std::runtime_error("mfs_create: invalid boundary type");
  if (!(b->pressure_scaling > 0)) throw std::runtime_error("mfs_create: pressure scaling must be positive");
  if (prec != MFS_GMG_FP64 && prec != MFS_GMG_FP32) throw std::runtime_error("mfs_create: invalid GMG precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    throw std::runtime_error("mfs_create: no HIP device available");
  upload_basis_tables();
  (void)noalloc_scratch();
  static std::once_flag smoother_once;
  std::call_once(smoother_once, init_small_smoother);
  auto h = std::make_unique<mfs_handle>();
  h->box = *b;
  for (int d = b->dim; d < 3; ++d) {
    h->box.coarse[d] = 1;
    h->box.extent[d] = 1.0;
  }
  h->prec = prec;
  MFS_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  for (int l = 0; l < b->n_levels; ++l) h->L.push_back(make_level_desc(h->box, l));
  h->top = b->n_levels - 1;
  h->nu = (long long)b->dim * h->Ltop().vnodes;
  h->np = h->Ltop().pnodes;
  h->n = h->nu + h->np;
  h->nq = n_qp(h->box);
  h->ct = cells_per_tile(h->box);
  h->dot_work.alloc((long long)dot_work_size() * 2048);
  h->dot_out.alloc(16);
  h->hbuf.alloc(2048);
  for (DBuf<double> *v : {&h->w1, &h->w2, &h->w3, &h->w4, &h->w5}) {
    v->alloc(h->n);
    v_fill<double>(v->p, 0.0, v->n, h->st);
  }
  if (prec == MFS_GMG_FP32)
    h->g32.setup(h->box, h->st, &h->dot_work, &h->dot_out);
  else
    h->g64.setup(h->box, h->st, &h->dot_work, &h->dot_out);
  // pressure normalization weights: int psi_i (unit mass applied to ones)
  {
    LevelDesc Lu = h->Ltop();
    Lu.s = 1.0;
    level_derive(Lu);
    h->ones_cell.alloc(Lu.ntiles_total * h->ct);
    v_fill<double>(h->ones_cell.p, 1.0, h->ones_cell.n, h->st);
    h->pnorm_w.alloc(h->np);
    v_fill<double>(h->w3.p, 1.0, h->np, h->st);
    launch_cell_op<double>(OP_MASS, MODE_SET, Lu, nullptr, h->w3.p, h->ones_cell.p, 0, nullptr, nullptr, nullptr,
                           h->pnorm_w.p, h->st);
  }
  MFS_HIP(hipStreamSynchronize(h->st));
  *out = h.release();
}

void set_viscosity(mfs_handle *h, const double *eta_qp, int averaging) {
  if (!h) throw std::invalid_argument("null mfs_handle");
  if (!eta_qp) throw std::runtime_error("mfs_set_viscosity: null viscosity");
  if (averaging < 0 || averaging > MFS_AVERAGE_PROJECT_Q1)
    throw std::runtime_error("mfs_set_viscosity: invalid averaging");
  const LevelDesc &L = h->Ltop();
  const long long nqv = L.cells * h->nq;
  double mn = 1e300, mx = -1e300;
  for (long long i = 0; i < nqv; ++i) {
    if (!(eta_qp[i] > 0) || !std::isfinite(eta_qp[i]))
      throw std::runtime_error("mfs_set_viscosity: viscosities must be positive and finite");
  }
  if (h->eta_nat.n != nqv) h->eta_nat.alloc(nqv);
  MFS_HIP(hipMemcpyAsync(h->eta_nat.p, eta_qp, nqv * sizeof(double), hipMemcpyHostToDevice, h->st));
  const bool q1 = averaging == MFS_AVERAGE_PROJECT_Q1;
  if (q1)
    for (const LevelDesc &Ll : h->L)
      if (Ll.geo) throw std::runtime_error("mfs_set_viscosity: project-to-Q1 averaging needs the box geometry");
  h->active_per_qp = (averaging == MFS_AVERAGE_NONE || q1) ? 1 : 0;
  const long long na = L.ntiles_total * h->ct * (h->active_per_qp ? h->nq : 1);
  if (h->eta_active.n != na) h->eta_active.alloc(na);
  v_fill<double>(h->eta_active.p, 1.0, na, h->st);
  if (q1) {
    // Q1 coefficients on every level, active table = clamped values at the
    // quadrature points (matrix_free_operators.cc:308-326)
    for (long long i = 0; i < nqv; ++i) {
      mn = std::min(mn, eta_qp[i]);
      mx = std::max(mx, eta_qp[i]);
    }
    const int nlev = (int)h->L.size(), nv = 1 << L.dim;
    h->q1_coef.resize(nlev);
    for (int l = 0; l < nlev; ++l)
      if (h->q1_coef[l].n != h->L[l].cells * nv) h->q1_coef[l].alloc(h->L[l].cells * nv);
    launch_eta_q1_project(L, h->eta_nat.p, h->q1_coef[nlev - 1].p, h->st);
    for (int l = nlev - 2; l >= 0; --l)
      launch_eta_q1_restrict(h->L[l], h->L[l + 1], h->q1_coef[l + 1].p, h->q1_coef[l].p, h->st);
    launch_eta_q1_eval<double>(L, h->q1_coef[nlev - 1].p, mn, mx, h->eta_active.p, h->st);
  } else if (h->active_per_qp) {
    launch_eta_to_tiles<double>(L, h->eta_nat.p, 1, h->eta_active.p, h->st);
    for (long long i = 0; i < nqv; ++i) {
      mn = std::min(mn, eta_qp[i]);
      mx = std::max(mx, eta_qp[i]);
    }
  } else {
    launch_eta_average<double>(L, h->eta_nat.p, averaging, h->eta_active.p, h->st);
    // min/max over the averaged values (Viscosity range report)
    std::vector<double> tmp(na);
    MFS_HIP(hipMemcpyAsync(tmp.data(), h->eta_active.p, na * sizeof(double), hipMemcpyDeviceToHost, h->st));
    MFS_HIP(hipStreamSynchronize(h->st));
    // padding entries are 1; scan only real cells via the natural order recomputation
    for (long long c = 0; c < L.cells; ++c) {
      double v;
      const double *e = eta_qp + c * h->nq;
      if (averaging == MFS_AVERAGE_ARITHMETIC) {
        double s = 0;
        for (int q = 0; q < h->nq; ++q) s += e[q];
        v = s / h->nq;
      } else if (averaging == MFS_AVERAGE_HARMONIC) {
        double s = 0;
        for (int q = 0; q < h->nq; ++q) s += 1.0 / e[q];
        v = 1.0 / (s / h->nq);
      } else {
        double s = 0;
        for (int q = 0; q < h->nq; ++q) s += std::log(e[q]);
        v = std::exp(s / h->nq);
      }
      mn = std::min(mn, v);
      mx = std::max(mx, v);
    }
  }
  h->eta_min = mn;
  h->eta_max = mx;
  if (q1) {
    if (h->prec == MFS_GMG_FP32)
      h->g32.set_eta_q1(h->q1_coef, mn, mx);
    else
      h->g64.set_eta_q1(h->q1_coef, mn, mx);
  } else if (h->prec == MFS_GMG_FP32) {
    h->g32.set_eta(h->eta_active.p, h->active_per_qp);
  } else {
    h->g64.set_eta(h->eta_active.p, h->active_per_qp);
  }
  MFS_HIP(hipStreamSynchronize(h->st));
  h->have_eta = true;
  h->have_prec = false;
}

// host <-> device helpers
struct HostIO {
  mfs_handle *h;
  void in(double *d, const double *hsrc, long long n) {
    if (!hsrc) throw std::invalid_argument("null host input vector");
    MFS_HIP(hipMemcpyAsync(d, hsrc, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  }
  void out(double *hdst, const double *d, long long n) {
    if (!hdst) throw std::invalid_argument("null host output vector");
    MFS_HIP(hipMemcpyAsync(hdst, d, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    MFS_HIP(hipStreamSynchronize(h->st));
  }
};

void need_handle(const mfs_handle *h) {
  if (!h) throw std::invalid_argument("null mfs_handle");
}
void check_level(const mfs_handle *h, int level, bool allow_active) {
  need_handle(h);
  if (level == -1 && allow_active) return;
  if (level < 0 || level > h->top) throw std::runtime_error("level out of range");
}
void need_eta(const mfs_handle *h) {
  need_handle(h);
  if (!h->have_eta) throw std::runtime_error("viscosity not set (call mfs_set_viscosity)");
}
void need_prec(const mfs_handle *h) {
  need_handle(h);
  if (!h->have_prec) throw std::runtime_error("preconditioner not built");
}
// host buffers of a host-vector entry point
void need_ptrs(std::initializer_list<const void *> ps, const char *fn) {
  for (const void *p : ps)
    if (!p) throw std::invalid_argument(std::string(fn) + ": null argument");
}

template <typename F>
int guard(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
    return MFS_ERROR;
  }
}

// run a level operation in precision T on FP64 host vectors
template <typename T, typename F>
void level_op_host(mfs_handle *h, GMG<T> &g, long long nin, long long nout, const double *src, double *dst, F f) {
  HostIO io{h};
  io.in(h->w1.p, src, nin);
  DBuf<T> a, b;
  a.alloc(std::max(nin, 1LL));
  b.alloc(std::max(nout, 1LL));
  v_cast<double, T>(h->w1.p, a.p, nin, h->st);
  f(a.p, b.p);
  v_cast<T, double>(b.p, h->w2.p, nout, h->st);
  io.out(dst, h->w2.p, nout);
  (void)g;
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

const char *mfs_last_error(void) { return g_err.c_str(); }

int mfs_validate_operator_arguments(int op, int mode, const void *src_u, const void *src_p, const void *eta,
                                    const void *rhs_u, const void *rhs_p, const void *dst_u, const void *dst_p) {
  return guard([&] {
    check_op_args((Op)op, (Mode)mode, src_u, src_p, eta, rhs_u, rhs_p, dst_u, dst_p);
    return MFS_OK;
  });
}

int mfs_box_dof_index(const mfs_box *box, long long n, const double *points, const int *component,
                      long long *index) {
  return guard([&] {
    if (!box || (n > 0 && (!points || !component || !index)))
      throw std::invalid_argument("mfs_box_dof_index: null argument");
    const mfs_box &b = *box;
    if (b.dim != 2 && b.dim != 3) throw std::invalid_argument("mfs_box_dof_index: dim must be 2 or 3");
    if (b.degree != 2 && b.degree != 3) throw std::invalid_argument("mfs_box_dof_index: degree must be 2 or 3");
    if (b.n_levels < 1 || b.n_levels > 16) throw std::invalid_argument("mfs_box_dof_index: n_levels out of range");
    const int k = b.degree;
    long long nc[3] = {1, 1, 1}, nv[3] = {1, 1, 1}, npn[3] = {1, 1, 1};
    double h[3] = {1, 1, 1};
    for (int d = 0; d < b.dim; ++d) {
      if (b.coarse[d] < 1 || !(b.extent[d] > 0)) throw std::invalid_argument("mfs_box_dof_index: invalid box");
      nc[d] = (long long)b.coarse[d] << (b.n_levels - 1);
      h[d] = b.extent[d] / (double)nc[d];
      nv[d] = k * nc[d] + 1;
      npn[d] = (k - 1) * nc[d] + 1;
    }
    const long long vnodes = nv[0] * nv[1] * nv[2], nu = b.dim * vnodes, np = npn[0] * npn[1] * npn[2];
    const std::vector<double> pv = fe_q_support(k), pp = fe_q_support(k - 1);
    std::vector<uint8_t> seen((size_t)(nu + np), 0);
    for (long long i = 0; i < n; ++i) {
      const int c = component[i];
      if (c < 0 || c > b.dim)
        throw std::invalid_argument("mfs_box_dof_index: component " + std::to_string(c) + " of DoF " +
                                    std::to_string(i) + " out of range");
      const bool vel = c < b.dim;
      long long g[3] = {0, 0, 0};
      for (int d = 0; d < b.dim; ++d) {
        g[d] = lattice_node(points[3 * i + d], h[d], nc[d], vel ? pv : pp);
        if (g[d] < 0)
          throw std::invalid_argument("mfs_box_dof_index: support point of DoF " + std::to_string(i) +
                                      " is not a node of the box");
      }
      const long long *m = vel ? nv : npn;
      const long long lex = g[0] + m[0] * (g[1] + m[1] * g[2]);
      const long long idx = vel ? c * vnodes + lex : nu + lex;
      if (seen[(size_t)idx]) throw std::invalid_argument("mfs_box_dof_index: two DoFs at one node (DoF " +
                                                         std::to_string(i) + ")");
      seen[(size_t)idx] = 1;
      index[i] = idx;
    }
    return MFS_OK;
  });
}

int mfs_create(const mfs_box *box, int gmg_precision, mfs_handle **out) {
  return guard([&] {
    create(box, gmg_precision, out);
    return MFS_OK;
  });
}

int mfs_destroy(mfs_handle *h) {
  return guard([&] {
    delete h;
    return MFS_OK;
  });
}

int mfs_name(const mfs_handle *h, char *buf, int len) {
  return guard([&] {
    if (!h || !buf || len <= 0) throw std::runtime_error("mfs_name: bad argument");
    std::snprintf(buf, (size_t)len, "%s", "GMG");
    return MFS_OK;
  });
}

int mfs_sizes(const mfs_handle *h, int level, long long out[4]) {
  return guard([&] {
    check_level(h, level, false);
    need_ptrs({out}, "mfs_sizes");
    const LevelDesc &L = h->L[level];
    out[0] = L.cells;
    out[1] = (long long)L.dim * L.vnodes;
    out[2] = L.pnodes;
    out[3] = L.vnodes;
    return MFS_OK;
  });
}

int mfs_set_viscosity(mfs_handle *h, const double *eta_qp, int averaging) {
  return guard([&] {
    set_viscosity(h, eta_qp, averaging);
    return MFS_OK;
  });
}

int mfs_set_pressure_scaling(mfs_handle *h, double s) {
  return guard([&] {
    need_handle(h);
    if (!(s > 0)) throw std::runtime_error("pressure scaling must be positive");
    h->box.pressure_scaling = s;
    for (auto &L : h->L) {
      L.s = s;
      level_derive(L);
    }
    auto upd = [&](auto &g) {
      for (auto &lv : g.lv) {
        lv->L.s = s;
        level_derive(lv->L);
      }
    };
    upd(h->g32);
    upd(h->g64);
    h->have_prec = false;
    return MFS_OK;
  });
}

// volume weights int psi_i of the pressure normalization on the active level
void pressure_weights(mfs_handle *h) {
  LevelDesc Lu = h->Ltop();
  Lu.s = 1.0;
  level_derive(Lu);
  Lu.geo = h->Ltop().geo;
  v_fill<double>(h->w3.p, 1.0, h->np, h->st);
  launch_cell_op<double>(OP_MASS, MODE_SET, Lu, nullptr, h->w3.p, h->ones_cell.p, 0, nullptr, nullptr, nullptr,
                         h->pnorm_w.p, h->st);
}

void set_geometry(mfs_handle *h, int level, const double *geo) {
  need_handle(h);
  check_level(h, level, false);
  if (!geo) throw std::invalid_argument("mfs_set_geometry: null table");
  for (int f = 0; f < 2 * h->box.dim; ++f)
    if (h->box.bc[f] == MFS_BC_TANGENTIAL)
      throw std::invalid_argument("mfs_set_geometry: no-normal-flux faces need axis-aligned normals (box geometry)");
  if (h->fs_on) throw std::invalid_argument("mfs_set_geometry: the free-surface term needs the box geometry");
  const LevelDesc &L = h->L[level];
  const int nq = h->nq;
  const long long n = L.cells * nq * (L.dim * L.dim + 1);
  if ((int)h->geo_lv.size() != h->top + 1) h->geo_lv.resize(h->top + 1);
  DBuf<double> nat;
  nat.alloc(n);
  MFS_HIP(hipMemcpyAsync(nat.p, geo, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  DBuf<double> &t = h->geo_lv[level];
  if (t.n != geo_table_size(L)) t.alloc(geo_table_size(L));
  launch_geo_to_tiles(L, nat.p, t.p, h->st);
  MFS_HIP(hipStreamSynchronize(h->st));
  // every descriptor of this level: the handler's, the multigrid's (graphs hold
  // them by value, so they are dropped), the mesh-deformation hierarchy's
  h->L[level].geo = t.p;
  auto upd = [&](auto &g) {
    if ((int)g.lv.size() > level) {
      g.lv[level]->L.geo = t.p;
      g.drop_graphs();
      g.shadow.reset();
    }
  };
  if (h->prec == MFS_GMG_FP32)
    upd(h->g32);
  else
    upd(h->g64);
  h->gmesh.reset();
  h->have_prec = false;
  h->have_eta = false;  // the level tables depend on JxW: mfs_set_viscosity again
  if (level == h->top) {
    pressure_weights(h);
    MFS_HIP(hipStreamSynchronize(h->st));
  }
}

void set_free_surface(mfs_handle *h, const int *faces, const double *stab) {
  need_handle(h);
  for (const auto &L : h->L)
    if (L.geo) throw std::invalid_argument("mfs_set_free_surface: box geometry only");
  if (!faces) throw std::invalid_argument("mfs_set_free_surface: null faces");
  const LevelDesc &L = h->Ltop();
  const int nfq = L.dim == 3 ? (L.k + 1) * (L.k + 1) : L.k + 1;
  long long tot = 0, mx = 0;
  int f6[6] = {0, 0, 0, 0, 0, 0};
  for (int f = 0; f < 6; ++f) {
    if (faces[f] != 0 && faces[f] != 1) throw std::invalid_argument("mfs_set_free_surface: face flags must be 0 or 1");
    if (f >= 2 * L.dim && faces[f]) throw std::invalid_argument("mfs_set_free_surface: face beyond the dimension");
    f6[f] = faces[f];
    if (f6[f]) {
      const long long c = free_surface_cells(L, f) * nfq;
      tot += c;
      mx = std::max(mx, c);
    }
  }
  h->fs_on = false;
  std::memcpy(h->fs_faces, f6, sizeof(f6));
  if (tot == 0) return;
  if (!stab) throw std::invalid_argument("mfs_set_free_surface: null stabilization table");
  h->fs_stab.alloc(tot * L.dim);
  h->fs_sv.alloc(mx);
  MFS_HIP(hipMemcpyAsync(h->fs_stab.p, stab, tot * L.dim * sizeof(double), hipMemcpyHostToDevice, h->st));
  MFS_HIP(hipStreamSynchronize(h->st));
  h->fs_on = true;
}

// MeshDeformationHandler::compute_mesh_displacements_gmg_for_degree
// (source/mesh_deformation/interface.cc:1076-1360) on the handle's levels:
// rhs = -A u0c (u0 read as stored, constrained rows zero) with u0c the
// constrained part of u0; CG with the Laplace V-cycle to tolerance * ||rhs||
// within 5 * N iterations; solution = x + u0c.
void mesh_setup(mfs_handle *h, const int *bc) {
  need_handle(h);
  if (!bc) throw std::invalid_argument("mesh deformation: null boundary types");
  mfs_box mb = h->box;
  for (int f = 0; f < 6; ++f) {
    if (bc[f] < 0 || bc[f] > 3) throw std::invalid_argument("mesh deformation: boundary type must be 0..3");
    mb.bc[f] = f < 2 * mb.dim ? (bc[f] == 3 ? 2 : bc[f]) : 0;
    if (mb.bc[f] == MFS_BC_TANGENTIAL && h->Ltop().geo)
      throw std::invalid_argument("mesh deformation: no-normal-displacement faces need the box geometry");
  }
  mb.compressible = 0;
  if (h->gmesh && std::memcmp(h->gmesh_bc, mb.bc, sizeof(mb.bc)) == 0) return;
  auto g = std::make_unique<GMG<double>>();
  g->laplace = true;
  g->setup(mb, h->st, &h->dot_work, &h->dot_out);
  for (int l = 0; l <= g->top; ++l) g->lv[l]->L.geo = h->L[l].geo;
  g->build(mb);
  MFS_HIP(hipStreamSynchronize(h->st));
  std::memcpy(h->gmesh_bc, mb.bc, sizeof(mb.bc));
  h->gmesh = std::move(g);
}

int mesh_solve_dev(mfs_handle *h, const int *bc, const double *d_u0, double tolerance, double *d_sol, int *its,
                   double *res) {
  mesh_setup(h, bc);
  GMG<double> &g = *h->gmesh;
  const int top = g.top;
  const LevelDesc &L = g.lv[top]->L;
  const long long n = g.lv[top]->nu;
  hipStream_t st = h->st;
  DBuf<double> u0c, rhs, x, r, z, pv, q;
  for (DBuf<double> *v : {&u0c, &rhs, &x, &r, &z, &pv, &q}) v->alloc(n);
  // constrained part of u0: u0 - (u0 with constrained entries zeroed)
  MFS_HIP(hipMemcpyAsync(z.p, d_u0, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  v_zero_constrained<double>(L, z.p, st);
  MFS_HIP(hipMemcpyAsync(u0c.p, d_u0, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  v_axpby<double>(-1.0, z.p, 1.0, u0c.p, n, st);
  LevelDesc Lp = L;
  Lp.plain_read = 1;
  launch_cell_op<double>(OP_LAPLACE, MODE_SET, Lp, u0c.p, nullptr, nullptr, 0, nullptr, nullptr, rhs.p, nullptr, st);
  v_zero_constrained<double>(L, rhs.p, st);
  v_scale(rhs.p, -1.0, n, st);
  const double rn = std::sqrt(g.dot(rhs.p, rhs.p, n));
  const double tol = tolerance * rn;
  v_fill<double>(x.p, 0.0, n, st);
  MFS_HIP(hipMemcpyAsync(r.p, rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  double rr = rn;
  int k = 0;
  const long long max_its = std::min<long long>(5 * n, 1000000);
  if (rr > tol) {
    g.vmult(0, r.p, z.p);
    MFS_HIP(hipMemcpyAsync(pv.p, z.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    double rz = g.dot(r.p, z.p, n);
    while (k < max_its) {
      g.applyA(top, MODE_SET, pv.p, nullptr, q.p);
      const double alpha = rz / g.dot(pv.p, q.p, n);
      v_axpby<double>(alpha, pv.p, 1.0, x.p, n, st);
      v_axpby<double>(-alpha, q.p, 1.0, r.p, n, st);
      ++k;
      rr = std::sqrt(g.dot(r.p, r.p, n));
      if (rr <= tol) break;
      g.vmult(0, r.p, z.p);
      const double rz2 = g.dot(r.p, z.p, n);
      const double beta = rz2 / rz;
      rz = rz2;
      v_axpby<double>(1.0, z.p, beta, pv.p, n, st);
    }
  }
  // solution = x + u0c (x is zero on the constrained DoFs)
  v_axpby<double>(1.0, x.p, 1.0, u0c.p, n, st);
  MFS_HIP(hipMemcpyAsync(d_sol, u0c.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  MFS_HIP(hipStreamSynchronize(st));
  if (its) *its = k;
  if (res) {
    res[0] = rr;
    res[1] = tol;
  }
  return rr <= tol ? MFS_OK : MFS_NO_CONVERGENCE;
}

void build_prec(mfs_handle *h) {
  need_eta(h);
  if (h->prec == MFS_GMG_FP32)
    h->g32.build(h->box);
  else
    h->g64.build(h->box);
  MFS_HIP(hipStreamSynchronize(h->st));
  h->have_prec = true;
}

int mfs_set_geometry(mfs_handle *h, int level, const double *geo) {
  return guard([&] {
    set_geometry(h, level, geo);
    return MFS_OK;
  });
}

int mfs_set_free_surface(mfs_handle *h, const int faces[6], const double *stab) {
  return guard([&] {
    set_free_surface(h, faces, stab);
    return MFS_OK;
  });
}

int mfs_mesh_deformation_solve(mfs_handle *h, const int bc[6], const double *u0, double tolerance, double *sol,
                               int *iterations, double *residuals) {
  return guard([&] {
    need_handle(h);
    need_ptrs({u0, sol}, "mfs_mesh_deformation_solve");
    const long long n = h->nu;
    DBuf<double> du0, dsol;
    du0.alloc(n);
    dsol.alloc(n);
    MFS_HIP(hipMemcpyAsync(du0.p, u0, n * sizeof(double), hipMemcpyHostToDevice, h->st));
    const int rc = mesh_solve_dev(h, bc, du0.p, tolerance, dsol.p, iterations, residuals);
    MFS_HIP(hipMemcpyAsync(sol, dsol.p, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    MFS_HIP(hipStreamSynchronize(h->st));
    if (rc != MFS_OK) g_err = "mesh deformation: CG did not reach the tolerance";
    return rc;
  });
}

int mfs_mesh_laplace_vmult(mfs_handle *h, const int bc[6], int level, const double *src, double *dst) {
  return guard([&] {
    mesh_setup(h, bc);
    check_level(h, level, false);
    need_ptrs({src, dst}, "mfs_mesh_laplace_vmult");
    GMG<double> &g = *h->gmesh;
    const long long n = g.lv[level]->nu;
    DBuf<double> a, b;
    a.alloc(n);
    b.alloc(n);
    MFS_HIP(hipMemcpyAsync(a.p, src, n * sizeof(double), hipMemcpyHostToDevice, h->st));
    g.applyA(level, MODE_SET, a.p, nullptr, b.p);
    MFS_HIP(hipMemcpyAsync(dst, b.p, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    MFS_HIP(hipStreamSynchronize(h->st));
    return MFS_OK;
  });
}

int mfs_mesh_vcycle(mfs_handle *h, const int bc[6], const double *src, double *dst, double *smoother_info) {
  return guard([&] {
    mesh_setup(h, bc);
    need_ptrs({src, dst}, "mfs_mesh_vcycle");
    GMG<double> &g = *h->gmesh;
    const long long n = g.lv[g.top]->nu;
    DBuf<double> a, b;
    a.alloc(n);
    b.alloc(n);
    MFS_HIP(hipMemcpyAsync(a.p, src, n * sizeof(double), hipMemcpyHostToDevice, h->st));
    g.vmult(0, a.p, b.p);
    MFS_HIP(hipMemcpyAsync(dst, b.p, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    MFS_HIP(hipStreamSynchronize(h->st));
    if (smoother_info)
      for (int l = 0; l <= g.top; ++l) {
        const ChebInfo &c = g.lv[l]->chA;
        double *o = smoother_info + 6 * l;
        o[0] = c.lmin;
        o[1] = c.lmax;
        o[2] = c.theta;
        o[3] = c.delta;
        o[4] = c.degree;
        o[5] = c.its;
      }
    return MFS_OK;
  });
}

int mfs_build_preconditioner(mfs_handle *h) {
  return guard([&] {
    build_prec(h);
    return MFS_OK;
  });
}

int mfs_stokes_vmult(mfs_handle *h, const double *src, double *dst) {
  return guard([&] {
    need_eta(h);
    HostIO io{h};
    io.in(h->w1.p, src, h->n);
    h->stokes(MODE_SET, h->w1.p, nullptr, h->w2.p);
    io.out(dst, h->w2.p, h->n);
    return MFS_OK;
  });
}

int mfs_stokes_vmult_device(mfs_handle *h, const double *d_src, double *d_dst) {
  return guard([&] {
    need_eta(h);
    need_ptrs({d_src, d_dst}, "mfs_stokes_vmult_device");
    h->stokes(MODE_SET, d_src, nullptr, d_dst);
    MFS_HIP(hipGetLastError());
    return MFS_OK;
  });
}

int mfs_ablock_vmult(mfs_handle *h, int level, const double *src, double *dst) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, true);
    if (level == -1) {
      HostIO io{h};
      io.in(h->w1.p, src, h->nu);
      h->ablock_active(MODE_SET, h->w1.p, nullptr, h->w2.p);
      io.out(dst, h->w2.p, h->nu);
      return MFS_OK;
    }
    const long long nu = (long long)h->box.dim * h->L[level].vnodes;
    auto run = [&](auto &g) {
      level_op_host(h, g, nu, nu, src, dst, [&](auto *a, auto *b) { g.applyA(level, MODE_SET, a, nullptr, b); });
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_ablock_vmult_device(mfs_handle *h, int level, const void *d_src, void *d_dst) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, false);
    need_ptrs({d_src, d_dst}, "mfs_ablock_vmult_device");
    if (h->prec == MFS_GMG_FP32)
      h->g32.applyA(level, MODE_SET, (const float *)d_src, nullptr, (float *)d_dst);
    else
      h->g64.applyA(level, MODE_SET, (const double *)d_src, nullptr, (double *)d_dst);
    MFS_HIP(hipGetLastError());
    return MFS_OK;
  });
}

int mfs_ablock_vmult_device_noalloc(mfs_handle *h, int level, const void *d_src, void *d_dst) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, false);
    need_ptrs({d_src, d_dst}, "mfs_ablock_vmult_device_noalloc");
    const LevelDesc &L = h->L[level];
    const long long esz = h->prec == MFS_GMG_FP32 ? (long long)sizeof(float) : (long long)sizeof(double);
    const long long need = split_apply_work_size(L) * esz;
    if (need > kNoAllocScratchBytes)
      throw std::invalid_argument("mfs_ablock_vmult_device_noalloc: level needs " + std::to_string(need) +
                                  " bytes of work array, more than the " + std::to_string(kNoAllocScratchBytes) +
                                  "-byte scratch array");
    void *work = noalloc_scratch();
    if (h->prec == MFS_GMG_FP32)
      h->g32.applyA_work(level, MODE_SET, (const float *)d_src, nullptr, (float *)d_dst, (float *)work);
    else
      h->g64.applyA_work(level, MODE_SET, (const double *)d_src, nullptr, (double *)d_dst, (double *)work);
    MFS_HIP(hipGetLastError());
    return MFS_OK;
  });
}

int mfs_mass_vmult(mfs_handle *h, int level, const double *src, double *dst) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, true);
    if (level == -1) {
      HostIO io{h};
      io.in(h->w1.p, src, h->np);
      h->mass_active(MODE_SET, h->w1.p, nullptr, h->w2.p);
      io.out(dst, h->w2.p, h->np);
      return MFS_OK;
    }
    const long long np = h->L[level].pnodes;
    auto run = [&](auto &g) {
      level_op_host(h, g, np, np, src, dst, [&](auto *a, auto *b) { g.applyS(level, MODE_SET, a, nullptr, b); });
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_bt_vmult(mfs_handle *h, const double *src_p, double *dst_u) {
  return guard([&] {
    need_handle(h);
    HostIO io{h};
    io.in(h->w1.p, src_p, h->np);
    launch_cell_op<double>(OP_BT, MODE_SET, h->Ltop(), nullptr, h->w1.p, nullptr, 0, nullptr, nullptr, h->w2.p,
                           nullptr, h->st);
    io.out(dst_u, h->w2.p, h->nu);
    return MFS_OK;
  });
}

int mfs_ablock_diagonal(mfs_handle *h, int level, double *diag) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, false);
    const long long nu = (long long)h->box.dim * h->L[level].vnodes;
    auto run = [&](auto &g) {
      using T = std::remove_pointer_t<decltype(g.lv[0]->eta.p)>;
      DBuf<T> d;
      d.alloc(nu);
      launch_cell_op<T>(OP_ADIAG, MODE_SET, g.lv[level]->L, nullptr, nullptr, g.lv[level]->eta.p, g.eta_pq, nullptr,
                        nullptr, d.p, nullptr, h->st);
      v_cast<T, double>(d.p, h->w2.p, nu, h->st);
      HostIO{h}.out(diag, h->w2.p, nu);
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_mass_diagonal(mfs_handle *h, int level, double *diag) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, false);
    const long long np = h->L[level].pnodes;
    auto run = [&](auto &g) {
      using T = std::remove_pointer_t<decltype(g.lv[0]->eta.p)>;
      DBuf<T> d;
      d.alloc(np);
      launch_cell_op<T>(OP_MDIAG, MODE_SET, g.lv[level]->L, nullptr, nullptr, g.lv[level]->eta.p, g.eta_pq, nullptr,
                        nullptr, nullptr, d.p, h->st);
      v_cast<T, double>(d.p, h->w2.p, np, h->st);
      HostIO{h}.out(diag, h->w2.p, np);
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_prolongate(mfs_handle *h, int fine_level, int kind, const double *src, double *dst) {
  return guard([&] {
    need_handle(h);
    if (fine_level < 1 || fine_level > h->top) throw std::runtime_error("fine_level out of range");
    if (kind != 0 && kind != 1) throw std::runtime_error("kind must be 0 or 1");
    const LevelDesc &Lc = h->L[fine_level - 1], &Lf = h->L[fine_level];
    const long long nc = kind == 0 ? (long long)h->box.dim * Lc.vnodes : Lc.pnodes;
    const long long nf = kind == 0 ? (long long)h->box.dim * Lf.vnodes : Lf.pnodes;
    HostIO io{h};
    io.in(h->w1.p, src, nc);
    DBuf<double> work;
    work.alloc(transfer_work_size(Lc, Lf));
    launch_prolongate<double>(kind, Lc, Lf, h->w1.p, h->w2.p, false, h->st, work.p);
    io.out(dst, h->w2.p, nf);
    return MFS_OK;
  });
}

int mfs_restrict(mfs_handle *h, int fine_level, int kind, const double *src, double *dst) {
  return guard([&] {
    need_handle(h);
    if (fine_level < 1 || fine_level > h->top) throw std::runtime_error("fine_level out of range");
    if (kind != 0 && kind != 1) throw std::runtime_error("kind must be 0 or 1");
    const LevelDesc &Lc = h->L[fine_level - 1], &Lf = h->L[fine_level];
    const long long nc = kind == 0 ? (long long)h->box.dim * Lc.vnodes : Lc.pnodes;
    const long long nf = kind == 0 ? (long long)h->box.dim * Lf.vnodes : Lf.pnodes;
    HostIO io{h};
    io.in(h->w1.p, src, nf);
    DBuf<double> work;
    work.alloc(transfer_work_size(Lc, Lf));
    launch_restrict<double>(kind, Lc, Lf, h->w1.p, h->w2.p, h->st, work.p);
    io.out(dst, h->w2.p, nc);
    return MFS_OK;
  });
}

int mfs_level_viscosity(mfs_handle *h, int level, double *out) {
  return guard([&] {
    need_eta(h);
    check_level(h, level, false);
    const LevelDesc &L = h->L[level];
    int t[3];
    tile_cells(L.dim, L.k, t);
    auto run = [&](auto &g) {
      using T = std::remove_pointer_t<decltype(g.lv[0]->eta.p)>;
      std::vector<T> tiles(g.lv[level]->eta.n);
      MFS_HIP(hipMemcpyAsync(tiles.data(), g.lv[level]->eta.p, tiles.size() * sizeof(T), hipMemcpyDeviceToHost,
                             h->st));
      MFS_HIP(hipStreamSynchronize(h->st));
      const long long CT = (long long)t[0] * t[1] * t[2];
      const int nq = g.eta_pq ? qp_per_cell_of(L) : 1;
      for (long long c = 0; c < L.cells; ++c) {
        const int gx = (int)(c % L.n[0]), gy = (int)((c / L.n[0]) % L.n[1]), gz = (int)(c / ((long long)L.n[0] * L.n[1]));
        const long long tile = gx / t[0] + (long long)L.ntiles[0] * (gy / t[1] + (long long)L.ntiles[1] * (gz / t[2]));
        const long long ci = gx % t[0] + t[0] * (gy % t[1] + t[1] * (gz % t[2]));
        for (int q = 0; q < nq; ++q) out[c * nq + q] = (double)tiles[(tile * nq + q) * CT + ci];
      }
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_smoother_info(mfs_handle *h, double *out) {
  return guard([&] {
    need_prec(h);
    need_ptrs({out}, "mfs_smoother_info");
    auto run = [&](auto &g) {
      for (int l = 0; l <= h->top; ++l) {
        const ChebInfo *cs[2] = {&g.lv[l]->chA, &g.lv[l]->chS};
        for (int k = 0; k < 2; ++k) {
          double *o = out + (l * 2 + k) * 5;
          o[0] = cs[k]->lmin;
          o[1] = cs[k]->lmax;
          o[2] = cs[k]->theta;
          o[3] = cs[k]->delta;
          o[4] = cs[k]->its;
        }
      }
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_chebyshev(mfs_handle *h, int kind, int level, const double *src, double *dst) {
  return guard([&] {
    need_prec(h);
    check_level(h, level, false);
    if (kind != 0 && kind != 1) throw std::invalid_argument("kind must be 0 (A block) or 1 (Schur)");
    const long long nn = kind == 0 ? (long long)h->box.dim * h->L[level].vnodes : h->L[level].pnodes;
    auto run = [&](auto &g) {
      level_op_host(h, g, nn, nn, src, dst, [&](auto *a, auto *b) {
        auto &lv = *g.lv[level];
        (void)lv;
        g.smooth(kind, level, b, a, true, 1);
      });
    };
    if (h->prec == MFS_GMG_FP32)
      run(h->g32);
    else
      run(h->g64);
    return MFS_OK;
  });
}

int mfs_vcycle(mfs_handle *h, int kind, const double *src, double *dst) {
  return guard([&] {
    need_prec(h);
    if (kind != 0 && kind != 1) throw std::invalid_argument("kind must be 0 (A block) or 1 (Schur)");
    const long long nn = kind == 0 ? h->nu : h->np;
    HostIO io{h};
    io.in(h->w1.p, src, nn);
    if (h->prec == MFS_GMG_FP32)
      h->g32.vmult(kind, h->w1.p, h->w2.p);
    else
      h->g64.vmult(kind, h->w1.p, h->w2.p);
    io.out(dst, h->w2.p, nn);
    return MFS_OK;
  });
}

int mfs_block_preconditioner(mfs_handle *h, const double *src, double *dst) {
  return guard([&] {
    need_prec(h);
    HostIO io{h};
    io.in(h->w3.p, src, h->n);
    h->block_prec_cheap(h->w3.p, h->w4.p);
    io.out(dst, h->w4.p, h->n);
    return MFS_OK;
  });
}

int mfs_solve_device(mfs_handle *h, const double *d_rhs, const double *d_x0, double *d_sol,
                     const mfs_solve_params *prm, mfs_solve_result *res) {
  return guard([&] {
    need_handle(h);
    need_ptrs({prm, res, d_rhs, d_sol}, "mfs_solve_device");
    return h->solve_dev(d_rhs, d_x0, d_sol, *prm, *res);
  });
}

int mfs_solve(mfs_handle *h, const double *rhs, const double *x0, double *sol, const mfs_solve_params *prm,
              mfs_solve_result *res) {
  return guard([&] {
    need_handle(h);
    need_ptrs({prm, res, rhs, sol}, "mfs_solve");
    DBuf<double> drhs, dx0, dsol;
    drhs.alloc(h->n);
    dsol.alloc(h->n);
    MFS_HIP(hipMemcpyAsync(drhs.p, rhs, h->n * sizeof(double), hipMemcpyHostToDevice, h->st));
    if (x0) {
      dx0.alloc(h->n);
      MFS_HIP(hipMemcpyAsync(dx0.p, x0, h->n * sizeof(double), hipMemcpyHostToDevice, h->st));
    }
    const int rc = h->solve_dev(drhs.p, x0 ? dx0.p : nullptr, dsol.p, *prm, *res);
    HostIO{h}.out(sol, dsol.p, h->n);
    return rc;
  });
}

static void assemble_rhs_dev(mfs_handle *h, const double *d_rho_nat, const double *gravity, double *d_rhs) {
  if (!gravity) throw std::runtime_error("mfs_assemble_buoyancy_rhs: null gravity");
  LevelDesc L = h->Ltop();
  for (int d = 0; d < 3; ++d) L.gravity[d] = d < h->box.dim ? gravity[d] : 0.0;
  DBuf<double> rho;
  rho.alloc(L.ntiles_total * h->ct * h->nq);
  v_fill<double>(rho.p, 0.0, rho.n, h->st);
  launch_eta_to_tiles<double>(L, d_rho_nat, 1, rho.p, h->st);
  launch_cell_op<double>(OP_RHS, MODE_SET, L, nullptr, nullptr, rho.p, 1, nullptr, nullptr, d_rhs, nullptr, h->st);
  v_fill<double>(d_rhs + h->nu, 0.0, h->np, h->st);
  MFS_HIP(hipStreamSynchronize(h->st));
}

// correct_stokes_rhs (local_smoothing.cc:617-754): rhs += -A(u0), the Stokes
// operator applied to u0 with its constrained values read as stored; the
// constrained velocity rows of rhs are left as they are.
void correct_rhs_dev(mfs_handle *h, const double *d_u0, double *d_rhs) {
  need_eta(h);
  LevelDesc Lp = h->Ltop();
  Lp.plain_read = 1;
  launch_cell_op<double>(OP_STOKES, MODE_SET, Lp, d_u0, d_u0 + h->nu, h->eta_active.p, h->active_per_qp, nullptr,
                         nullptr, h->w1.p, h->w1.p + h->nu, h->st);
  // free-surface term with plain reads of u0 (local_smoothing.cc:709-744)
  if (h->fs_on) launch_free_surface(Lp, h->fs_faces, h->fs_stab.p, d_u0, h->fs_sv.p, 1.0, h->w1.p, h->st);
  v_zero_constrained<double>(h->Ltop(), h->w1.p, h->st);
  v_axpby<double>(-1.0, h->w1.p, 1.0, d_rhs, h->n, h->st);
  MFS_HIP(hipStreamSynchronize(h->st));
}

int mfs_correct_stokes_rhs(mfs_handle *h, const double *u0, double *rhs) {
  return guard([&] {
    need_handle(h);
    need_ptrs({u0, rhs}, "mfs_correct_stokes_rhs");
    HostIO io{h};
    io.in(h->w2.p, u0, h->n);
    io.in(h->w3.p, rhs, h->n);
    correct_rhs_dev(h, h->w2.p, h->w3.p);
    io.out(rhs, h->w3.p, h->n);
    return MFS_OK;
  });
}

int mfs_assemble_buoyancy_rhs(mfs_handle *h, const double *rho_qp, const double *gravity, double *rhs) {
  return guard([&] {
    need_handle(h);
    need_ptrs({rho_qp, rhs}, "mfs_assemble_buoyancy_rhs");
    const long long nqv = h->Ltop().cells * h->nq;
    DBuf<double> rho, out;
    rho.alloc(nqv);
    out.alloc(h->n);
    MFS_HIP(hipMemcpyAsync(rho.p, rho_qp, nqv * sizeof(double), hipMemcpyHostToDevice, h->st));
    assemble_rhs_dev(h, rho.p, gravity, out.p);
    HostIO{h}.out(rhs, out.p, h->n);
    return MFS_OK;
  });
}

int mfs_assemble_buoyancy_rhs_device(mfs_handle *h, const double *d_rho_qp, const double *gravity, double *d_rhs) {
  return guard([&] {
    need_handle(h);
    need_ptrs({d_rho_qp, d_rhs}, "mfs_assemble_buoyancy_rhs_device");
    assemble_rhs_dev(h, d_rho_qp, gravity, d_rhs);
    return MFS_OK;
  });
}

int mfs_time_operator(mfs_handle *h, int which, int level, const void *d_src, void *d_dst, int reps,
                      double *avg_seconds) {
  return guard([&] {
    need_eta(h);
    if (reps < 1) throw std::runtime_error("mfs_time_operator: reps must be >= 1");
    need_ptrs({d_src, d_dst, avg_seconds}, "mfs_time_operator");
    if ((which >= 3) && !h->have_prec) throw std::runtime_error("preconditioner not built");
    if (which == 1 || which == 2) check_level(h, level, false);
    auto body = [&]() {
      switch (which) {
        case 0:
          h->stokes(MODE_SET, (const double *)d_src, nullptr, (double *)d_dst);
          break;
        case 1:
        case 2:
          if (h->prec == MFS_GMG_FP32)
            h->g32.apply(which - 1, level, MODE_SET, (const float *)d_src, nullptr, (float *)d_dst);
          else
            h->g64.apply(which - 1, level, MODE_SET, (const double *)d_src, nullptr, (double *)d_dst);
          break;
        case 3:
        case 4:
          if (h->prec == MFS_GMG_FP32)
            h->g32.vmult(which - 3, (const double *)d_src, (double *)d_dst);
          else
            h->g64.vmult(which - 3, (const double *)d_src, (double *)d_dst);
          break;
        case 5:
          h->block_prec_cheap((const double *)d_src, (double *)d_dst);
          break;
        default:
          throw std::runtime_error("mfs_time_operator: unknown operator");
      }
    };
    hipEvent_t e0, e1;
    MFS_HIP(hipEventCreate(&e0));
    MFS_HIP(hipEventCreate(&e1));
    MFS_HIP(hipEventRecord(e0, h->st));
    for (int r = 0; r < reps; ++r) body();
    MFS_HIP(hipEventRecord(e1, h->st));
    MFS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_seconds = ms * 1e-3 / reps;
    return MFS_OK;
  });
}

int mfs_set_graph_capture(mfs_handle *h, int enable) {
  return guard([&] {
    if (!h) throw std::runtime_error("mfs_set_graph_capture: null handle");
    h->g32.drop_graphs();
    h->g64.drop_graphs();
    h->g32.graphs = h->g64.graphs = enable != 0;
    return MFS_OK;
  });
}

int mfs_synchronize(mfs_handle *h) {
  return guard([&] {
    need_handle(h);
    MFS_HIP(hipStreamSynchronize(h->st));
    return MFS_OK;
  });
}

int mfs_memory_consumption(const mfs_handle *h, long long out[3]) {
  return guard([&] {
    need_handle(h);
    need_ptrs({out}, "mfs_memory_consumption");
    long long vec = 0, cell = h->eta_active.bytes() + h->eta_nat.bytes();
    for (const DBuf<double> *v : {&h->w1, &h->w2, &h->w3, &h->w4, &h->w5, &h->pnorm_w}) vec += v->bytes();
    auto acc = [&](const auto &g) {
      for (const auto &lv : g.lv) {
        cell += lv->eta.bytes();
        vec += lv->dinvA.bytes() + lv->dinvS.bytes() + lv->ux.bytes() + lv->ub.bytes() + lv->ur.bytes() +
               lv->ud.bytes() + lv->px.bytes() + lv->pb.bytes() + lv->pr.bytes() + lv->pd.bytes();
      }
    };
    acc(h->g32);
    acc(h->g64);
    out[0] = vec;
    out[1] = cell;
    out[2] = (long long)(h->vchunks.size() + h->zchunks.size()) * 8 * h->n * (long long)sizeof(double);
    return MFS_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
// Independent problems of uneven sizes solved together: one handle (and HIP
// stream) per problem, driven by host worker threads, so that the kernels of
// many small solves run concurrently on the device.
struct mfs_ensemble {
  std::vector<mfs_handle *> members;
  std::vector<int> last_its;  // iterations of each member in the previous ensemble solve (scheduling)
  ~mfs_ensemble() {
    for (mfs_handle *h : members) delete h;
  }
};

namespace {
// run f(i) for i < n on up to n_threads workers; the first error is rethrown
template <typename F>
void run_parallel(int n, int n_threads, F f) {
  std::atomic<int> next{0};
  std::string first_error;
  std::atomic<bool> failed{false};
  auto worker = [&]() {
    for (int i = next++; i < n; i = next++) {
      try {
        f(i);
      } catch (const std::exception &e) {
        if (!failed.exchange(true)) first_error = "ensemble member " + std::to_string(i) + ": " + e.what();
      }
    }
  };
  const int nt = std::max(1, std::min(n_threads, n));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (std::thread &t : pool) t.join();
  if (failed) throw std::runtime_error(first_error);
}
}  // namespace

int mfs_ensemble_create(int n, const mfs_box *boxes, int gmg_precision, mfs_ensemble **out) {
  return guard([&] {
    if (n < 1 || !boxes || !out) throw std::runtime_error("mfs_ensemble_create: bad arguments");
    auto e = std::make_unique<mfs_ensemble>();
    e->members.assign(n, nullptr);
    for (int i = 0; i < n; ++i) create(&boxes[i], gmg_precision, &e->members[i]);
    *out = e.release();
    return MFS_OK;
  });
}

int mfs_ensemble_destroy(mfs_ensemble *e) {
  return guard([&] {
    delete e;
    return MFS_OK;
  });
}

int mfs_ensemble_member(mfs_ensemble *e, int i, mfs_handle **h) {
  return guard([&] {
    if (!e || !h || i < 0 || i >= (int)e->members.size()) throw std::runtime_error("mfs_ensemble_member: bad index");
    *h = e->members[i];
    return MFS_OK;
  });
}

int mfs_ensemble_setup(mfs_ensemble *e, const double *const *eta_qp, int averaging, int n_threads) {
  return guard([&] {
    if (!e || !eta_qp) throw std::runtime_error("mfs_ensemble_setup: null argument");
    run_parallel((int)e->members.size(), n_threads, [&](int i) {
      set_viscosity(e->members[i], eta_qp[i], averaging);
      build_prec(e->members[i]);
    });
    return MFS_OK;
  });
}

int mfs_ensemble_solve(mfs_ensemble *e, const double *const *rhs, double *const *sol, const mfs_solve_params *prm,
                       mfs_solve_result *res, int n_threads, double *seconds) {
  return guard([&] {
    if (!e || !rhs || !sol || !prm || !res) throw std::runtime_error("mfs_ensemble_solve: null argument");
    const int n = (int)e->members.size();
    std::vector<int> rc(n, MFS_OK);
    for (mfs_handle *h : e->members) MFS_HIP(hipStreamSynchronize(h->st));
    // longest-first order (LPT): the workers take members by decreasing
    // expected cost, DoFs times the member's iteration count of the previous
    // ensemble solve (DoFs alone on the first), so that the member on the
    // critical path does not start last
    if ((int)e->last_its.size() != n) e->last_its.assign(n, 1);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (double)e->members[a]->n * e->last_its[a] > (double)e->members[b]->n * e->last_its[b];
    });
    const auto t0 = std::chrono::steady_clock::now();
    run_parallel(n, n_threads, [&](int k) {
      const int i = order[k];
      mfs_handle *h = e->members[i];
      DBuf<double> drhs, dsol;
      drhs.alloc(h->n);
      dsol.alloc(h->n);
      MFS_HIP(hipMemcpyAsync(drhs.p, rhs[i], h->n * sizeof(double), hipMemcpyHostToDevice, h->st));
      rc[i] = h->solve_dev(drhs.p, nullptr, dsol.p, *prm, res[i]);
      HostIO{h}.out(sol[i], dsol.p, h->n);
    });
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) e->last_its[i] = std::max(1, res[i].cheap_iterations + res[i].expensive_iterations);
    for (int i = 0; i < n; ++i)
      if (rc[i] != MFS_OK) return MFS_NO_CONVERGENCE;
    return MFS_OK;
  });
}
