// ORACLE — test infrastructure only (CPU baseline of bench.py and tests/).
//
// CPU FP64 Stokes operator apply for 3D Q2xQ1 on box meshes, vectorized the
// way deal.II's matrix-free CPU path is: a batch of 8 cells fills the lanes of
// one SIMD vector (VectorizedArray<double> over cell batches), the whole
// sum-factorized cell operation runs on those vectors, and OpenMP threads take
// batches of one colour at a time (batches of one colour share no node, so the
// scatter needs no atomics).  Same operator as oracle.cpp's stokes_vmult:
//   StokesOperator::local_apply     source/simulator/solver/matrix_free_operators.cc:376-547
//   MatrixFreeOperators::Base::vmult: constrained source entries read as zero,
//   constrained rows copy the source.
// The sum factorization is in collocation form: nodal values -> Gauss-point
// values with S in each direction, derivatives with the Gauss-point
// differentiation matrix D (D S = G), integration with the transposes.
// The batch kernel is compiled for AVX-512, AVX2+FMA and baseline x86-64 and the
// loader picks the widest one the host supports.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int W = 8;  // cells per batch (SIMD lanes)
typedef double vd __attribute__((vector_size(W * sizeof(double))));

struct Tables {
  double S[3][3];   // S[q][i]: Q2 Lagrange basis (nodes 0, 1/2, 1) at Gauss point q
  double D[3][3];   // D[q][r]: derivative of the Gauss-point Lagrange basis r at Gauss point q
  double SP[3][2];  // Q1 basis (nodes 0, 1) at Gauss point q
  double w[3];      // Gauss weights on [0, 1]
};

Tables make_tables() {
  Tables t{};
  const double g = std::sqrt(0.6) / 2.0;
  const double x[3] = {0.5 - g, 0.5, 0.5 + g};
  t.w[0] = t.w[2] = 5.0 / 18.0;
  t.w[1] = 8.0 / 18.0;
  const double nodes[3] = {0.0, 0.5, 1.0};
  auto lag = [](const double *pts, int n, int i, double xx) {
    double v = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != i) v *= (xx - pts[j]) / (pts[i] - pts[j]);
    return v;
  };
  auto dlag = [](const double *pts, int n, int i, double xx) {
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == i) continue;
      double v = 1.0 / (pts[i] - pts[m]);
      for (int j = 0; j < n; ++j)
        if (j != i && j != m) v *= (xx - pts[j]) / (pts[i] - pts[j]);
      s += v;
    }
    return s;
  };
  for (int q = 0; q < 3; ++q) {
    for (int i = 0; i < 3; ++i) {
      t.S[q][i] = lag(nodes, 3, i, x[q]);
      t.D[q][i] = dlag(x, 3, i, x[q]);
    }
    t.SP[q][0] = 1.0 - x[q];
    t.SP[q][1] = x[q];
  }
  return t;
}

struct Plan {
  int n[3];
  int nv[3], np[3];
  double h[3];
  long long vnodes, pnodes;
  double s;
  int compressible;
  std::vector<uint8_t> vcon;  // 3 * vnodes, component-major
  Tables T;
};

// out[idx(i)] (+)= sum_j M[i][j] in[idx(j)] along direction DIR of a 3x3x3 array
// (x fastest); TR applies M^T
template <int DIR, bool TR, bool ADD>
inline __attribute__((always_inline)) void line3(const double (&M)[3][3], const vd *in, vd *out) {
  constexpr int st = DIR == 0 ? 1 : DIR == 1 ? 3 : 9;
  constexpr int s1 = DIR == 0 ? 3 : 1, s2 = DIR == 2 ? 3 : 9;
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a) {
      const int base = a * s1 + b * s2;
      const vd v0 = in[base], v1 = in[base + st], v2 = in[base + 2 * st];
      for (int i = 0; i < 3; ++i) {
        const double m0 = TR ? M[0][i] : M[i][0], m1 = TR ? M[1][i] : M[i][1], m2 = TR ? M[2][i] : M[i][2];
        const vd r = m0 * v0 + m1 * v1 + m2 * v2;
        if (ADD)
          out[base + i * st] += r;
        else
          out[base + i * st] = r;
      }
    }
}

// Q1 <-> Gauss points along DIR: 2 nodes -> 3 points, or with TR 3 -> 2.
// Directions before DIR have 3 entries, those after it 2 (x is converted
// first going forward and last going back).
constexpr int ext(int d, int dir, int nd) { return d < dir ? 3 : d == dir ? nd : 2; }
template <int DIR, bool TR>
inline __attribute__((always_inline)) void line32(const double (&M)[3][2], const vd *in, vd *out) {
  constexpr int ni = TR ? 3 : 2, no = TR ? 2 : 3;
  constexpr int si[3] = {1, ext(0, DIR, ni), ext(0, DIR, ni) * ext(1, DIR, ni)};
  constexpr int so[3] = {1, ext(0, DIR, no), ext(0, DIR, no) * ext(1, DIR, no)};
  constexpr int o1 = DIR == 0 ? 1 : 0, o2 = DIR == 2 ? 1 : 2;
  for (int b = 0; b < ext(o2, DIR, ni); ++b)
    for (int a = 0; a < ext(o1, DIR, ni); ++a) {
      const int bi = a * si[o1] + b * si[o2], bo = a * so[o1] + b * so[o2];
      for (int i = 0; i < no; ++i) {
        vd r = {};
        for (int j = 0; j < ni; ++j) r += (TR ? M[j][i] : M[i][j]) * in[bi + j * si[DIR]];
        out[bo + i * so[DIR]] = r;
      }
    }
}

// One batch of up to W cells along x in the cell row (cy, cz), starting at cx0.
__attribute__((target_clones("avx512f", "arch=haswell", "default"))) void stokes_batch(
    const Plan &P, long long cx0, int cy, int cz, int lanes, const double *eta, int per_qp, const double *src,
    double *dst) {
  const Tables &T = P.T;
  const long long nvx = P.nv[0], nvxy = (long long)P.nv[0] * P.nv[1];
  const long long npx = P.np[0], npxy = (long long)P.np[0] * P.np[1];
  const double *su = src, *sp = src + 3 * P.vnodes;
  double *du = dst, *dp = dst + 3 * P.vnodes;
  long long vbase[W], pbase[W], cell[W];
  for (int l = 0; l < W; ++l) {
    const long long cx = cx0 + std::min(l, lanes - 1);
    vbase[l] = 2 * cx + nvx * (2LL * cy) + nvxy * (2LL * cz);
    pbase[l] = cx + npx * cy + npxy * cz;
    cell[l] = cx + (long long)P.n[0] * (cy + (long long)P.n[1] * cz);
  }
  // gather velocity (constrained entries as zero) and pressure
  vd u[3][27], p[8];
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          vd v;
          for (int l = 0; l < W; ++l) {
            const long long g = c * P.vnodes + vbase[l] + i + nvx * j + nvxy * k;
            v[l] = P.vcon[g] ? 0.0 : su[g];
          }
          u[c][i + 3 * j + 9 * k] = v;
        }
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        vd v;
        for (int l = 0; l < W; ++l) v[l] = sp[pbase[l] + i + npx * j + npxy * k];
        p[i + 2 * j + 4 * k] = v;
      }
  // values at the Gauss points, then collocation derivatives
  vd t0[27], t1[27], uq[27], g[3][3][27];
  for (int c = 0; c < 3; ++c) {
    line3<0, false, false>(T.S, u[c], t0);
    line3<1, false, false>(T.S, t0, t1);
    line3<2, false, false>(T.S, t1, uq);
    line3<0, false, false>(T.D, uq, g[c][0]);
    line3<1, false, false>(T.D, uq, g[c][1]);
    line3<2, false, false>(T.D, uq, g[c][2]);
  }
  vd pa[12], pb[18], pq[27];
  line32<0, false>(T.SP, p, pa);
  line32<1, false>(T.SP, pa, pb);
  line32<2, false>(T.SP, pb, pq);
  // quadrature-point operation: f = (2 eta eps(u) [- 2/3 eta div u I] - s p I) JxW, fp = -s div u JxW
  const double ih[3] = {1.0 / P.h[0], 1.0 / P.h[1], 1.0 / P.h[2]};
  const double vol = P.h[0] * P.h[1] * P.h[2];
  vd fp[27];
  for (int q = 0; q < 27; ++q) {
    const double jxw = vol * T.w[q % 3] * T.w[(q / 3) % 3] * T.w[q / 9];
    vd e;
    for (int l = 0; l < W; ++l) e[l] = per_qp ? eta[cell[l] * 27 + q] : eta[cell[l]];
    vd gr[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) gr[c][d] = g[c][d][q] * ih[d];
    const vd div = gr[0][0] + gr[1][1] + gr[2][2];
    const vd e2 = 2.0 * e * jxw;
    vd diag = -P.s * pq[q] * jxw;
    if (P.compressible) diag -= (2.0 / 3.0) * e * div * jxw;
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        vd f = 0.5 * e2 * (gr[c][d] + gr[d][c]);
        if (c == d) f += diag;
        g[c][d][q] = f * ih[d];
      }
    fp[q] = -P.s * div * jxw;
  }
  // integrate: D^T along each direction summed, then S^T in all three
  vd out[3][27];
  for (int c = 0; c < 3; ++c) {
    line3<0, true, false>(T.D, g[c][0], t0);
    line3<1, true, true>(T.D, g[c][1], t0);
    line3<2, true, true>(T.D, g[c][2], t0);
    line3<2, true, false>(T.S, t0, t1);
    line3<1, true, false>(T.S, t1, t0);
    line3<0, true, false>(T.S, t0, out[c]);
  }
  vd qa[18], qb[12], op[8];
  line32<2, true>(T.SP, fp, qa);
  line32<1, true>(T.SP, qa, qb);
  line32<0, true>(T.SP, qb, op);
  // scatter lane by lane (neighbouring lanes share nodes), constrained rows skipped
  for (int l = 0; l < lanes; ++l) {
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i < 3; ++i) {
            const long long gi = c * P.vnodes + vbase[l] + i + nvx * j + nvxy * k;
            if (!P.vcon[gi]) du[gi] += out[c][i + 3 * j + 9 * k][l];
          }
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) dp[pbase[l] + i + npx * j + npxy * k] += op[i + 2 * j + 4 * k][l];
  }
}

thread_local std::string g_err;

}  // namespace

extern "C" {

typedef struct {
  int dim;
  int degree;
  int coarse[3];
  int n_levels;
  double extent[3];
  int bc[6];
  int compressible;
  double pressure_scaling;
} orc_port_problem;

const char *orc_port_last_error() { return g_err.c_str(); }

// Instruction set of the batch kernel the loader picks on this host.
const char *orc_port_isa() {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return "avx512f";
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return "avx2+fma";
  return "x86-64 baseline";
}

// Thread count of the oracle's OpenMP loops (omp_set_num_threads).
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
  return 0;
}
int orc_get_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Plan for the finest level of a 3D degree-2 box problem.
int orc_port_create(const orc_port_problem *pp, void **out) {
  try {
    if (pp->dim != 3 || pp->degree != 2) throw std::runtime_error("CPU port: 3D degree 2 only");
    auto P = std::make_unique<Plan>();
    const int l = pp->n_levels - 1;
    for (int d = 0; d < 3; ++d) {
      P->n[d] = pp->coarse[d] << l;
      P->h[d] = pp->extent[d] / P->n[d];
      P->nv[d] = 2 * P->n[d] + 1;
      P->np[d] = P->n[d] + 1;
    }
    P->vnodes = (long long)P->nv[0] * P->nv[1] * P->nv[2];
    P->pnodes = (long long)P->np[0] * P->np[1] * P->np[2];
    P->s = pp->pressure_scaling;
    P->compressible = pp->compressible;
    P->T = make_tables();
    P->vcon.assign((size_t)(3 * P->vnodes), 0);
    for (long long i = 0; i < P->vnodes; ++i) {
      const int ix[3] = {(int)(i % P->nv[0]), (int)((i / P->nv[0]) % P->nv[1]),
                         (int)(i / ((long long)P->nv[0] * P->nv[1]))};
      for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
          if (!(side == 0 ? ix[d] == 0 : ix[d] == P->nv[d] - 1)) continue;
          const int b = pp->bc[2 * d + side];
          if (b == 2)
            for (int c = 0; c < 3; ++c) P->vcon[(size_t)c * P->vnodes + i] = 1;
          else if (b == 1)
            P->vcon[(size_t)d * P->vnodes + i] = 1;
        }
    }
    *out = P.release();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

int orc_port_destroy(void *plan) {
  delete static_cast<Plan *>(plan);
  return 0;
}

// dst = A src with Base::vmult semantics; threads <= 0: OpenMP default
int orc_port_stokes_vmult(void *plan, const double *eta, int per_qp, const double *src, double *dst, int threads) {
  const Plan &P = *static_cast<Plan *>(plan);
  const long long nu = 3 * P.vnodes, n = nu + P.pnodes;
  const long long nbx = (P.n[0] + W - 1) / W;
#ifdef _OPENMP
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#else
  const int nt = 1;
  (void)threads;
#endif
#pragma omp parallel num_threads(nt)
  {
#pragma omp for schedule(static)
    for (long long i = 0; i < n; ++i) dst[i] = 0.0;
    for (int color = 0; color < 8; ++color) {
      const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
      const long long hx = (nbx - px + 1) / 2, hy = (P.n[1] - py + 1) / 2, hz = (P.n[2] - pz + 1) / 2;
#pragma omp for schedule(static)
      for (long long b = 0; b < hx * hy * hz; ++b) {
        const long long bx = px + 2 * (b % hx), cy = py + 2 * ((b / hx) % hy), cz = pz + 2 * (b / (hx * hy));
        const long long cx0 = bx * W;
        const int lanes = (int)std::min<long long>(W, P.n[0] - cx0);
        stokes_batch(P, cx0, (int)cy, (int)cz, lanes, eta, per_qp, src, dst);
      }
    }
#pragma omp for schedule(static)
    for (long long i = 0; i < nu; ++i)
      if (P.vcon[i]) dst[i] = src[i];
  }
  return 0;
}

}  // extern "C"
