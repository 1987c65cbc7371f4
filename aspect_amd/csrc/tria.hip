// Stokes level operators on general meshes: cells given by lists of node ids
// (any topology: locally refined boxes, the spherical shell), the geometry at
// every quadrature point, and affine constraints (hanging nodes, Dirichlet,
// no normal flux) given as closed constraint lines.
//
// Reference: the level operators of the global-coarsening handler
// (source/simulator/solver/stokes_matrix_free_global_coarsening.cc:1603-1625,
// operators of matrix_free_operators.cc:376-547, :668-701, :808-845,
// :974-1009) on a MatrixFree object built with the hanging-node, boundary
// and no-normal-flux constraints of :1481-1598.
//
// One apply is three steps on the handle's stream:
//   1. resolve: x~ = C x (a constrained DoF takes the weighted sum of its
//      masters; deal.II's read_dof_values with constraints);
//   2. the general cell kernel in MODE_CELLS: a workgroup per batch of 64
//      cells reads x~ through the cell lists and writes every cell's nodal
//      contributions to a per-cell buffer (no atomics, no colouring);
//   3. assemble: every free DoF sums the contributions of its node's cells
//      and, through the transposed constraint lines, of the constrained DoFs
//      it is a master of (distribute_local_to_global); constrained rows are
//      identity rows (MatrixFreeOperators::Base::vmult), zero for B^T.
// Sums run in a fixed order, so an apply is deterministic.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aspect_mfs.h"
#include "mfs_internal.h"
#include "tile_common.h"

using namespace mfs;

namespace {

constexpr int CT = 64;  // cells per batch (the cell kernel's tile size for degree 2)

template <typename T>
struct Buf {
  T *p = nullptr;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept : p(o.p) { o.p = nullptr; }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (n) MFS_HIP(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T)));
  }
  void upload(const std::vector<T> &v, hipStream_t st) {
    alloc(v.size());
    if (!v.empty()) MFS_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  }
};

// x~[i] = x[i] on free DoFs, sum_k w_k x[m_k] on constrained ones
__global__ void k_resolve(const long long n, const double *__restrict__ src, const int *__restrict__ row,
                          const long long *__restrict__ off, const long long *__restrict__ mas,
                          const double *__restrict__ w, double *__restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = row[i];
  if (r < 0) {
    out[i] = src[i];
    return;
  }
  double a = 0;
  for (long long k = off[r]; k < off[r + 1]; ++k) a += w[k] * src[mas[k]];
  out[i] = a;
}

// sum of the cell contributions of one node (component offset coff)
__device__ __forceinline__ double node_sum(const double *__restrict__ cb, const long long *__restrict__ iptr,
                                           const int *__restrict__ ioff, long long node, long long coff) {
  double a = 0;
  for (long long k = iptr[node]; k < iptr[node + 1]; ++k) a += cb[ioff[k] + coff];
  return a;
}

// dst[i] for the DoFs of one block (nn nodes per component, component stride
// cs in the contribution buffer): free DoFs their own sum plus the weighted
// sums of the constrained DoFs they are a master of; constrained DoFs
// src[i] (ident) or 0
// Fused modes of the assemble pass (the smoother's steps on a level):
// AS_RESID dst = b - K src; AS_CHEB one Jacobi-Chebyshev step on x = src (in
// place) with d: r = b - K x, d = first ? dinv r / theta : f1 d + f2 dinv r,
// x += d (the arithmetic of v_cheb_first / v_cheb_step)
enum { AS_SET = 0, AS_RESID = 1, AS_CHEB = 2 };
struct ChebStep {
  const double *b, *dinv;
  double *d;
  double f1, f2;  // first step: f2 = 1 / theta
  int first;
};
template <int MODE>
__global__ void k_assemble(const long long n, const long long nn, const long long cs, const double *__restrict__ cb,
                           const long long *__restrict__ iptr, const int *__restrict__ ioff,
                           const int *__restrict__ row, const long long *__restrict__ tptr,
                           const long long *__restrict__ tdof, const double *__restrict__ tw,
                           const double *__restrict__ src, double *__restrict__ dst, const int ident,
                           const ChebStep ch) {
  // XCD-aware order: neighbouring nodes gather from the same cells' contribution
  // lines, so contiguous node ranges go to one XCD (one L2)
  const long long i = xcd_order(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a;
  if (row[i] >= 0) {
    a = ident ? src[i] : 0.0;
  } else {
    const long long comp = i / nn, node = i - comp * nn;
    a = node_sum(cb, iptr, ioff, node, comp * cs);
    for (long long k = tptr[i]; k < tptr[i + 1]; ++k) {
      const long long c = tdof[k], cc = c / nn;
      a += tw[k] * node_sum(cb, iptr, ioff, c - cc * nn, cc * cs);
    }
  }
  if constexpr (MODE == AS_SET) {
    dst[i] = a;
  } else if constexpr (MODE == AS_RESID) {
    dst[i] = 1.0 * ch.b[i] + -1.0 * a;
  } else {
    const double r = 1.0 * ch.b[i] + -1.0 * a;
    const double dd = ch.first ? ch.dinv[i] * r * ch.f2 : ch.f1 * ch.d[i] + ch.f2 * ch.dinv[i] * r;
    ch.d[i] = dd;
    dst[i] = src[i] + dd;
  }
}

// Cell pressure mass matrices s^2/eta (psi_i, psi_j) of a degree-2 level (Q1
// pressure, QGauss(3)), the 2^dim x 2^dim symmetric matrix of every cell as
// its upper triangle, layout [(batch * NT + t) * 64 + c]; computed once per
// viscosity so that the many mass applies of the Schur smoother read one
// small matrix per cell instead of the geometry at every point
template <int DIM>
__global__ void k_tria_mass_matrix(const long long ncells, const double *__restrict__ geo,
                                   const double *__restrict__ eta, const double s2, double *__restrict__ mat) {
  constexpr int NP = 1 << DIM, NQ = DIM == 3 ? 27 : 9, G = DIM * DIM + 1, NT = NP * (NP + 1) / 2;
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const long long b = c / 64, cc = c - b * 64;
  const double xg[3] = {0.5 - 0.5 * 0.7745966692414834, 0.5, 0.5 + 0.5 * 0.7745966692414834};
  double m[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) m[t] = 0;
  for (int q = 0; q < NQ; ++q) {
    const int qx = q % 3, qy = (q / 3) % 3, qz = q / 9;
    const double w = s2 * geo[((b * NQ + q) * G + G - 1) * 64 + cc] / eta[(b * NQ + q) * 64 + cc];
    double psi[NP];
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      const double vx = (a & 1) ? xg[qx] : 1.0 - xg[qx];
      const double vy = (a & 2) ? xg[qy] : 1.0 - xg[qy];
      const double vz = DIM == 3 ? ((a & 4) ? xg[qz] : 1.0 - xg[qz]) : 1.0;
      psi[a] = vx * vy * vz;
    }
    int t = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
      for (int j = i; j < NP; ++j) m[t++] += w * psi[i] * psi[j];
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) mat[(b * NT + t) * 64 + cc] = m[t];
}

// per-cell pressure mass contributions from the stored cell matrices,
// [(batch * NP + a) * 64 + c] as the cell kernel writes them
// (constrained pressure DoFs read through their lines: row / off / mas / w)
template <int DIM>
__global__ void k_tria_mass_apply(const long long ncells, const int *__restrict__ cell_p,
                                  const double *__restrict__ mat, const double *__restrict__ x,
                                  const int *__restrict__ row, const long long *__restrict__ off,
                                  const long long *__restrict__ mas, const double *__restrict__ w,
                                  double *__restrict__ out) {
  constexpr int NP = 1 << DIM, NT = NP * (NP + 1) / 2;
  // XCD-aware order: neighbouring cells share pressure nodes
  const long long c = xcd_order(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const long long b = c / 64, cc = c - b * 64;
  double p[NP], y[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    const int node = cell_p[(b * NP + a) * 64 + cc];
    const int r = row[node];
    double v;
    if (r < 0) {
      v = x[node];
    } else {
      v = 0;
      for (long long k = off[r]; k < off[r + 1]; ++k) v += w[k] * x[mas[k]];
    }
    p[a] = v;
    y[a] = 0;
  }
  int t = 0;
#pragma unroll
  for (int i = 0; i < NP; ++i)
#pragma unroll
    for (int j = i; j < NP; ++j) {
      const double v = mat[(b * NT + t++) * 64 + cc];
      y[i] += v * p[j];
      if (j != i) y[j] += v * p[i];
    }
#pragma unroll
  for (int a = 0; a < NP; ++a) out[(b * NP + a) * 64 + cc] = y[a];
}

// Constraint lines of one block on the device, with their transpose.
struct DevCon {
  long long n = 0;
  Buf<int> row;
  Buf<long long> off, mas, tptr, tdof;
  Buf<double> w, tw;
  std::vector<char> fixed;  // host copy: 1 on constrained DoFs

  // host-side check of the lines (no device call): DoFs and masters in range,
  // every DoF constrained once, no master constrained, finite weights
  static void validate(long long ndof, const mfs_constraints *c, const char *what) {
    DevCon tmp;
    tmp.build(ndof, c, what);
  }

  void set(long long ndof, const mfs_constraints *c, hipStream_t st, const char *what) {
    build(ndof, c, what);
    row.upload(h_row, st);
    off.upload(h_off, st);
    mas.upload(h_mas, st);
    w.upload(h_w, st);
    tptr.upload(h_tptr, st);
    tdof.upload(h_tdof, st);
    tw.upload(h_tw, st);
    MFS_HIP(hipStreamSynchronize(st));
    h_w.clear();
    h_tptr.clear();
    h_tdof.clear();
    h_tw.clear();
  }

  std::vector<int> h_row;
  std::vector<long long> h_off, h_mas, h_tptr, h_tdof;
  std::vector<double> h_w, h_tw;

  void build(long long ndof, const mfs_constraints *c, const char *what) {
    n = ndof;
    std::vector<int> &r = h_row;
    std::vector<long long> &o = h_off, &m = h_mas;
    std::vector<double> &ww = h_w;
    r.assign(ndof, -1);
    o.assign(1, 0);
    m.clear();
    ww.clear();
    fixed.assign(ndof, 0);
    const long long nl = c ? c->n_lines : 0;
    if (nl < 0) throw std::invalid_argument(std::string(what) + " constraints: negative line count");
    if (nl > 0 && (!c->dofs || !c->offsets || (!c->masters && c->offsets[nl] > 0) ||
                   (!c->weights && c->offsets[nl] > 0)))
      throw std::invalid_argument(std::string(what) + " constraints: null array");
    for (long long l = 0; l < nl; ++l) {
      const long long d = c->dofs[l];
      if (d < 0 || d >= ndof) throw std::invalid_argument(std::string(what) + " constraints: DoF out of range");
      if (r[d] >= 0) throw std::invalid_argument(std::string(what) + " constraints: DoF constrained twice");
      r[d] = (int)l;
      fixed[d] = 1;
      if (c->offsets[l + 1] < c->offsets[l])
        throw std::invalid_argument(std::string(what) + " constraints: offsets not increasing");
    }
    for (long long l = 0; l < nl; ++l) {
      for (long long k = c->offsets[l]; k < c->offsets[l + 1]; ++k) {
        const long long mm = c->masters[k];
        if (mm < 0 || mm >= ndof) throw std::invalid_argument(std::string(what) + " constraints: master out of range");
        if (fixed[mm])
          throw std::invalid_argument(std::string(what) +
                                      " constraints: a master is itself constrained (close the constraints first)");
        if (!std::isfinite(c->weights[k])) throw std::invalid_argument(std::string(what) + " constraints: weight not finite");
        m.push_back(mm);
        ww.push_back(c->weights[k]);
      }
      o.push_back((long long)m.size());
    }
    // transpose: for every master, the (constrained DoF, weight) pairs, in line order
    std::vector<long long> &tp = h_tptr, &td = h_tdof;
    std::vector<double> &tv = h_tw;
    tp.assign(ndof + 1, 0);
    td.assign(m.size(), 0);
    tv.assign(m.size(), 0.0);
    for (long long k = 0; k < (long long)m.size(); ++k) ++tp[m[k] + 1];
    for (long long i = 0; i < ndof; ++i) tp[i + 1] += tp[i];
    std::vector<long long> fill(tp.begin(), tp.end() - 1);
    for (long long l = 0; l < nl; ++l)
      for (long long k = o[l]; k < o[l + 1]; ++k) {
        const long long pos = fill[m[k]]++;
        td[pos] = c->dofs[l];
        tv[pos] = ww[k];
      }
  }
};

}  // namespace

struct mfs_tria {
  int dim = 0;
  long long ncells = 0, nbatch = 0, nn_u = 0, nn_p = 0, ndof_u = 0, ndof_p = 0;
  int nv = 0, npl = 0, nq = 0;
  hipStream_t st = nullptr;
  LevelDesc L{};
  Buf<int> cell_u, cell_p;
  Buf<double> geo, eta, cbu, cbp, xu, xp, io_src, io_dst;
  Buf<long long> iptr_u, iptr_p;
  Buf<int> ioff_u, ioff_p;  // offsets into the per-cell buffers (checked < 2^31 at creation)
  DevCon con_u, con_p;
  std::vector<int> h_cell_u, h_cell_p;  // host copies of the cell lists (diagonal colouring)
  bool have_eta = false;
  Buf<double> mass_mat;  // cell pressure mass matrices for the current viscosity
  bool have_mass = false;

  // the cell mass matrices of the current viscosity table (on stream s)
  void update_mass(hipStream_t s) {
    const int np = dim == 3 ? 8 : 4, nt = np * (np + 1) / 2;
    if (!mass_mat.p) mass_mat.alloc(nbatch * nt * CT);
    const double s2 = L.s * L.s;
    if (dim == 3)
      hipLaunchKernelGGL(k_tria_mass_matrix<3>, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, ncells,
                         geo.p, eta.p, s2, mass_mat.p);
    else
      hipLaunchKernelGGL(k_tria_mass_matrix<2>, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, ncells,
                         geo.p, eta.p, s2, mass_mat.p);
    MFS_HIP(hipGetLastError());
    have_mass = true;
  }

  ~mfs_tria() {
    if (st) (void)hipStreamDestroy(st);
  }

  // node -> (offset of its contribution in the per-cell buffer, component 0)
  static void incidence(const int *cells, long long ncells, int per, long long nn, int comps, std::vector<long long> &ptr,
                        std::vector<int> &off) {
    ptr.assign(nn + 1, 0);
    for (long long c = 0; c < ncells * per; ++c) ++ptr[cells[c] + 1];
    for (long long i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
    off.assign(ptr[nn], 0);
    std::vector<long long> fill(ptr.begin(), ptr.end() - 1);
    for (long long c = 0; c < ncells; ++c) {
      const long long b = c / CT, cc = c % CT;
      for (int a = 0; a < per; ++a) off[fill[cells[c * per + a]]++] = (int)(((b * comps) * per + a) * CT + cc);
    }
  }

  void create(const mfs_tria_desc *d, const mfs_constraints *cu, const mfs_constraints *cp) {
    if (!d) throw std::invalid_argument("mfs_tria_create: null description");
    if (d->dim != 2 && d->dim != 3) throw std::invalid_argument("mfs_tria_create: dim must be 2 or 3");
    if (d->degree != 2) throw std::invalid_argument("mfs_tria_create: general meshes support velocity degree 2");
    if (d->n_cells <= 0 || d->n_nodes_u <= 0 || d->n_nodes_p <= 0)
      throw std::invalid_argument("mfs_tria_create: empty mesh");
    if (!d->cell_nodes_u || !d->cell_nodes_p || !d->geometry)
      throw std::invalid_argument("mfs_tria_create: null cell lists or geometry");
    dim = d->dim;
    ncells = d->n_cells;
    nbatch = (ncells + CT - 1) / CT;
    nn_u = d->n_nodes_u;
    nn_p = d->n_nodes_p;
    ndof_u = dim * nn_u;
    ndof_p = nn_p;
    nv = dim == 3 ? 27 : 9;
    npl = dim == 3 ? 8 : 4;
    nq = nv;
    if (nn_u > 0x7fffffffLL || nn_p > 0x7fffffffLL) throw std::invalid_argument("mfs_tria_create: too many nodes");
    for (long long i = 0; i < ncells * nv; ++i)
      if (d->cell_nodes_u[i] < 0 || d->cell_nodes_u[i] >= nn_u)
        throw std::invalid_argument("mfs_tria_create: velocity node id out of range");
    for (long long i = 0; i < ncells * npl; ++i)
      if (d->cell_nodes_p[i] < 0 || d->cell_nodes_p[i] >= nn_p)
        throw std::invalid_argument("mfs_tria_create: pressure node id out of range");
    const int G = dim * dim + 1;
    for (long long i = 0; i < ncells * nq * G; ++i)
      if (!std::isfinite(d->geometry[i])) throw std::invalid_argument("mfs_tria_create: geometry not finite");
    DevCon::validate(ndof_u, cu, "velocity");
    DevCon::validate(ndof_p, cp, "pressure");
    MFS_HIP(hipStreamCreate(&st));
    upload_basis_tables();

    // batch layouts [(batch * per + a) * CT + c]; padded cells read node 0 with
    // zero geometry (their contributions are zero and never assembled)
    std::vector<int> bu(nbatch * nv * CT, 0), bp(nbatch * npl * CT, 0);
    std::vector<double> bg(nbatch * nq * G * CT, 0.0);
    for (long long c = 0; c < ncells; ++c) {
      const long long b = c / CT, cc = c % CT;
      for (int a = 0; a < nv; ++a) bu[(b * nv + a) * CT + cc] = d->cell_nodes_u[c * nv + a];
      for (int a = 0; a < npl; ++a) bp[(b * npl + a) * CT + cc] = d->cell_nodes_p[c * npl + a];
      for (int q = 0; q < nq; ++q)
        for (int g = 0; g < G; ++g) bg[((b * nq + q) * G + g) * CT + cc] = d->geometry[(c * nq + q) * G + g];
    }
    h_cell_u.assign(d->cell_nodes_u, d->cell_nodes_u + ncells * nv);
    h_cell_p.assign(d->cell_nodes_p, d->cell_nodes_p + ncells * npl);
    cell_u.upload(bu, st);
    cell_p.upload(bp, st);
    geo.upload(bg, st);
    if (nbatch * dim * nv * CT >= 0x7fffffffLL) throw std::invalid_argument("mfs_tria_create: too many cells");
    std::vector<long long> ptr;
    std::vector<int> off;
    incidence(d->cell_nodes_u, ncells, nv, nn_u, dim, ptr, off);
    iptr_u.upload(ptr, st);
    ioff_u.upload(off, st);
    incidence(d->cell_nodes_p, ncells, npl, nn_p, 1, ptr, off);
    iptr_p.upload(ptr, st);
    ioff_p.upload(off, st);
    MFS_HIP(hipStreamSynchronize(st));
    con_u.set(ndof_u, cu, st, "velocity");
    con_p.set(ndof_p, cp, st, "pressure");
    cbu.alloc(nbatch * dim * nv * CT);
    cbp.alloc(nbatch * npl * CT);
    xu.alloc(ndof_u);
    xp.alloc(ndof_p);
    eta.alloc(nbatch * nq * CT);
    io_src.alloc(ndof_u + ndof_p);
    io_dst.alloc(ndof_u + ndof_p);

    L.dim = dim;
    L.k = 2;
    L.vnodes = nn_u;
    L.pnodes = nn_p;
    L.cells = ncells;
    L.s = d->pressure_scaling;
    L.compressible = d->compressible ? 1 : 0;
    for (int i = 0; i < 3; ++i) L.h[i] = 1.0;
    level_derive(L);
    L.geo = geo.p;
    L.cell_u = cell_u.p;
    L.cell_p = cell_p.p;
  }

  void set_viscosity(const double *eta_qp) {
    if (!eta_qp) throw std::invalid_argument("mfs_tria_set_viscosity: null viscosity");
    std::vector<double> be(nbatch * nq * CT, 1.0);
    for (long long c = 0; c < ncells; ++c)
      for (int q = 0; q < nq; ++q) {
        const double v = eta_qp[c * nq + q];
        if (!(v > 0) || !std::isfinite(v)) throw std::invalid_argument("mfs_tria_set_viscosity: viscosity must be positive");
        be[((c / CT) * nq + q) * CT + c % CT] = v;
      }
    MFS_HIP(hipMemcpyAsync(eta.p, be.data(), be.size() * sizeof(double), hipMemcpyHostToDevice, st));
    update_mass(st);
    MFS_HIP(hipStreamSynchronize(st));
    have_eta = true;
  }

  void resolve(const double *src, double *out, DevCon &c, hipStream_t s) {
    const unsigned nb = (unsigned)((c.n + 255) / 256);
    hipLaunchKernelGGL(k_resolve, dim3(nb), dim3(256), 0, s, c.n, src, c.row.p, c.off.p, c.mas.p, c.w.p, out);
    MFS_HIP(hipGetLastError());
  }
  void assemble(const double *cb, long long nn, long long cs, const Buf<long long> &ip, const Buf<int> &io,
                DevCon &c, const double *src, double *dst, int ident, hipStream_t s, int mode = AS_SET,
                const ChebStep &ch = ChebStep{}) {
    const unsigned nb = (unsigned)((c.n + 255) / 256);
    if (mode == AS_SET)
      hipLaunchKernelGGL(k_assemble<AS_SET>, dim3(nb), dim3(256), 0, s, c.n, nn, cs, cb, ip.p, io.p, c.row.p,
                         c.tptr.p, c.tdof.p, c.tw.p, src, dst, ident, ch);
    else if (mode == AS_RESID)
      hipLaunchKernelGGL(k_assemble<AS_RESID>, dim3(nb), dim3(256), 0, s, c.n, nn, cs, cb, ip.p, io.p, c.row.p,
                         c.tptr.p, c.tdof.p, c.tw.p, src, dst, ident, ch);
    else
      hipLaunchKernelGGL(k_assemble<AS_CHEB>, dim3(nb), dim3(256), 0, s, c.n, nn, cs, cb, ip.p, io.p, c.row.p,
                         c.tptr.p, c.tdof.p, c.tw.p, src, dst, ident, ch);
    MFS_HIP(hipGetLastError());
  }

  // level operator of the smoother (kind 0: A block, 1: pressure mass) on x,
  // its result assembled in mode `mode` (AS_SET: dst = K x; AS_RESID: dst =
  // ch.b - K x; AS_CHEB: a Jacobi-Chebyshev step on x in place, dst = x)
  void level_op(int kind, const double *x, double *dst, int mode, const ChebStep &ch, hipStream_t s) {
    if (kind == 0) {
      resolve(x, xu.p, con_u, s);
      launch_cells<double>(OP_ABLOCK, L, nbatch, xu.p, nullptr, eta.p, cbu.p, cbp.p, s);
      assemble(cbu.p, nn_u, (long long)nv * CT, iptr_u, ioff_u, con_u, x, dst, 1, s, mode, ch);
      return;
    }
    if (!have_mass) throw std::runtime_error("pressure mass matrices not set");
    const unsigned nb = (unsigned)((ncells + 255) / 256);
    if (dim == 3)
      hipLaunchKernelGGL(k_tria_mass_apply<3>, dim3(nb), dim3(256), 0, s, ncells, cell_p.p, mass_mat.p, x,
                         con_p.row.p, con_p.off.p, con_p.mas.p, con_p.w.p, cbp.p);
    else
      hipLaunchKernelGGL(k_tria_mass_apply<2>, dim3(nb), dim3(256), 0, s, ncells, cell_p.p, mass_mat.p, x,
                         con_p.row.p, con_p.off.p, con_p.mas.p, con_p.w.p, cbp.p);
    MFS_HIP(hipGetLastError());
    assemble(cbp.p, nn_p, 0, iptr_p, ioff_p, con_p, x, dst, 1, s, mode, ch);
  }

  // device vectors: op MFS_OP_STOKES [u|p] -> [u|p], ABLOCK u -> u, BT p -> u, MASS p -> p
  void vmult(int op, const double *src, double *dst) { vmult(op, src, dst, st); }
  // plain: source values read as stored (no constraint resolution) and
  // constrained rows left zero (correct_stokes_rhs)
  void vmult(int op, const double *src, double *dst, hipStream_t s, bool plain = false) {
    if (op != MFS_OP_STOKES && op != MFS_OP_ABLOCK && op != MFS_OP_BT && op != MFS_OP_MASS)
      throw std::invalid_argument("mfs_tria_vmult: operator must be Stokes, A block, B^T or mass");
    if (!src || !dst) throw std::invalid_argument("mfs_tria_vmult: null vector");
    if (op != MFS_OP_BT && !have_eta) throw std::runtime_error("viscosity not set (call mfs_tria_set_viscosity)");
    const bool rd_u = op == MFS_OP_STOKES || op == MFS_OP_ABLOCK;
    const bool rd_p = op == MFS_OP_STOKES || op == MFS_OP_BT || op == MFS_OP_MASS;
    const bool wr_u = op != MFS_OP_MASS, wr_p = op == MFS_OP_STOKES || op == MFS_OP_MASS;
    const double *su = rd_u ? src : nullptr;
    const double *sp = op == MFS_OP_STOKES ? src + ndof_u : (rd_p ? src : nullptr);
    double *du = wr_u ? dst : nullptr;
    double *dp = op == MFS_OP_STOKES ? dst + ndof_u : (wr_p ? dst : nullptr);
    const bool mass_fused = op == MFS_OP_MASS && have_mass && !plain;  // resolve inside the mass kernel
    if (plain) {
      if (rd_u) MFS_HIP(hipMemcpyAsync(xu.p, su, ndof_u * sizeof(double), hipMemcpyDeviceToDevice, s));
      if (rd_p) MFS_HIP(hipMemcpyAsync(xp.p, sp, ndof_p * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      if (rd_u) resolve(su, xu.p, con_u, s);
      if (rd_p && !mass_fused) resolve(sp, xp.p, con_p, s);
    }
    if (op == MFS_OP_MASS && have_mass && !plain) {
      const unsigned nb = (unsigned)((ncells + 255) / 256);
      if (dim == 3)
        hipLaunchKernelGGL(k_tria_mass_apply<3>, dim3(nb), dim3(256), 0, s, ncells, cell_p.p, mass_mat.p, sp,
                           con_p.row.p, con_p.off.p, con_p.mas.p, con_p.w.p, cbp.p);
      else
        hipLaunchKernelGGL(k_tria_mass_apply<2>, dim3(nb), dim3(256), 0, s, ncells, cell_p.p, mass_mat.p, sp,
                           con_p.row.p, con_p.off.p, con_p.mas.p, con_p.w.p, cbp.p);
      MFS_HIP(hipGetLastError());
    } else {
      launch_cells<double>((Op)op, L, nbatch, rd_u ? xu.p : nullptr, rd_p ? xp.p : nullptr, eta.p, cbu.p, cbp.p, s);
    }
    if (wr_u)
      assemble(cbu.p, nn_u, (long long)nv * CT, iptr_u, ioff_u, con_u, su, du, (op == MFS_OP_BT || plain) ? 0 : 1, s);
    if (wr_p) assemble(cbp.p, nn_p, 0, iptr_p, ioff_p, con_p, sp, dp, plain ? 0 : 1, s);
  }

  // Greedy colouring of the DoFs of one block (kind 0 velocity, 1 pressure)
  // such that no two DoFs of a colour meet in one cell's effective DoF set
  // (its nodes and the masters of their constraint lines): A applied to the
  // indicator of a colour then returns the diagonal on that colour.  DoFs of
  // one node get distinct colours (node colour * components + component);
  // constrained DoFs get -1.  Returns the number of colours.
  int colour_dofs(int kind, std::vector<int> &colour) const {
    const std::vector<int> &cells = kind == 0 ? h_cell_u : h_cell_p;
    const int per = kind == 0 ? nv : npl, comps = kind == 0 ? dim : 1;
    const long long nn = kind == 0 ? nn_u : nn_p;
    const DevCon &c = kind == 0 ? con_u : con_p;
    // master nodes of the lines of every node
    std::vector<std::vector<int>> mnodes(nn);
    for (long long dof = 0; dof < c.n; ++dof) {
      const int r = c.h_row[dof];
      if (r < 0) continue;
      for (long long k = c.h_off[r]; k < c.h_off[r + 1]; ++k) mnodes[dof % nn].push_back((int)(c.h_mas[k] % nn));
    }
    std::vector<std::vector<int>> eff(ncells);
    std::vector<std::vector<int>> inv(nn);
    for (long long e = 0; e < ncells; ++e) {
      std::vector<int> &v = eff[e];
      for (int a = 0; a < per; ++a) {
        const int n = cells[e * per + a];
        v.push_back(n);
        v.insert(v.end(), mnodes[n].begin(), mnodes[n].end());
      }
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      for (int n : v) inv[n].push_back((int)e);
    }
    std::vector<int> ncol(nn, -1), stamp;
    int ncolours = 0;
    for (long long n = 0; n < nn; ++n) {
      for (int e : inv[n])
        for (int m : eff[e])
          if (ncol[m] >= 0) {
            if ((int)stamp.size() <= ncol[m]) stamp.resize(ncol[m] + 1, -1);
            stamp[ncol[m]] = (int)n;
          }
      int k = 0;
      while (k < (int)stamp.size() && stamp[k] == (int)n) ++k;
      ncol[n] = k;
      ncolours = std::max(ncolours, k + 1);
    }
    colour.assign(comps * nn, -1);
    for (int cc = 0; cc < comps; ++cc)
      for (long long n = 0; n < nn; ++n)
        if (c.h_row[cc * nn + n] < 0) colour[cc * nn + n] = ncol[n] * comps + cc;
    return ncolours * comps;
  }

  long long n_in(int op) const {
    return op == MFS_OP_STOKES ? ndof_u + ndof_p : (op == MFS_OP_ABLOCK ? ndof_u : ndof_p);
  }
  long long n_out(int op) const {
    return op == MFS_OP_STOKES ? ndof_u + ndof_p : (op == MFS_OP_MASS ? ndof_p : ndof_u);
  }
};

namespace {
template <typename F>
int tguard(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return MFS_ERROR;
  }
}
void need(const mfs_tria *t) {
  if (!t) throw std::invalid_argument("null mfs_tria");
}
}  // namespace

extern "C" {

int mfs_tria_create(const mfs_tria_desc *desc, const mfs_constraints *velocity, const mfs_constraints *pressure,
                    mfs_tria **out) {
  return tguard([&] {
    if (!out) throw std::invalid_argument("mfs_tria_create: null output");
    *out = nullptr;
    mfs_tria *t = new mfs_tria();
    try {
      t->create(desc, velocity, pressure);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
    return MFS_OK;
  });
}

int mfs_tria_destroy(mfs_tria *t) {
  return tguard([&] {
    if (t) {
      (void)hipStreamSynchronize(t->st);
      delete t;
    }
    return MFS_OK;
  });
}

int mfs_tria_sizes(const mfs_tria *t, long long out[3]) {
  return tguard([&] {
    need(t);
    if (!out) throw std::invalid_argument("mfs_tria_sizes: null output");
    out[0] = t->ncells;
    out[1] = t->ndof_u;
    out[2] = t->ndof_p;
    return MFS_OK;
  });
}

int mfs_tria_set_viscosity(mfs_tria *t, const double *eta_qp) {
  return tguard([&] {
    need(t);
    t->set_viscosity(eta_qp);
    return MFS_OK;
  });
}

int mfs_tria_vmult(mfs_tria *t, int op, const double *src, double *dst) {
  return tguard([&] {
    need(t);
    if (!src || !dst) throw std::invalid_argument("mfs_tria_vmult: null vector");
    const long long ni = t->n_in(op), no = t->n_out(op);
    MFS_HIP(hipMemcpyAsync(t->io_src.p, src, ni * sizeof(double), hipMemcpyHostToDevice, t->st));
    t->vmult(op, t->io_src.p, t->io_dst.p);
    MFS_HIP(hipMemcpyAsync(dst, t->io_dst.p, no * sizeof(double), hipMemcpyDeviceToHost, t->st));
    MFS_HIP(hipStreamSynchronize(t->st));
    return MFS_OK;
  });
}

int mfs_tria_correct_stokes_rhs(mfs_tria *t, const double *u0, double *rhs) {
  return tguard([&] {
    need(t);
    if (!u0 || !rhs) throw std::invalid_argument("mfs_tria_correct_stokes_rhs: null vector");
    if (!t->have_eta) throw std::runtime_error("viscosity not set (call mfs_tria_set_viscosity)");
    const long long n = t->ndof_u + t->ndof_p;
    Buf<double> r;
    r.alloc(n);
    MFS_HIP(hipMemcpyAsync(t->io_src.p, u0, n * sizeof(double), hipMemcpyHostToDevice, t->st));
    MFS_HIP(hipMemcpyAsync(r.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, t->st));
    t->vmult(MFS_OP_STOKES, t->io_src.p, t->io_dst.p, t->st, true);
    v_axpby<double>(-1.0, t->io_dst.p, 1.0, r.p, n, t->st);  // rhs -= K_plain u0
    MFS_HIP(hipMemcpyAsync(rhs, r.p, n * sizeof(double), hipMemcpyDeviceToHost, t->st));
    MFS_HIP(hipStreamSynchronize(t->st));
    return MFS_OK;
  });
}

int mfs_tria_get_viscosity(mfs_tria *t, double *eta_qp) {
  return tguard([&] {
    need(t);
    if (!eta_qp) throw std::invalid_argument("mfs_tria_get_viscosity: null output");
    if (!t->have_eta) throw std::runtime_error("viscosity not set");
    std::vector<double> be(t->nbatch * t->nq * CT);
    MFS_HIP(hipMemcpyAsync(be.data(), t->eta.p, be.size() * sizeof(double), hipMemcpyDeviceToHost, t->st));
    MFS_HIP(hipStreamSynchronize(t->st));
    for (long long c = 0; c < t->ncells; ++c)
      for (int q = 0; q < t->nq; ++q) eta_qp[c * t->nq + q] = be[((c / CT) * t->nq + q) * CT + c % CT];
    return MFS_OK;
  });
}

int mfs_tria_time(mfs_tria *t, int op, const double *d_src, double *d_dst, int reps, double *avg_seconds) {
  return tguard([&] {
    need(t);
    if (reps < 1 || !avg_seconds) throw std::invalid_argument("mfs_tria_time: reps >= 1 and an output required");
    hipEvent_t e0, e1;
    MFS_HIP(hipEventCreate(&e0));
    MFS_HIP(hipEventCreate(&e1));
    MFS_HIP(hipEventRecord(e0, t->st));
    for (int r = 0; r < reps; ++r) t->vmult(op, d_src, d_dst);
    MFS_HIP(hipEventRecord(e1, t->st));
    MFS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_seconds = ms * 1e-3 / reps;
    return MFS_OK;
  });
}

int mfs_tria_vmult_device(mfs_tria *t, int op, const double *d_src, double *d_dst) {
  return tguard([&] {
    need(t);
    t->vmult(op, d_src, d_dst);
    MFS_HIP(hipStreamSynchronize(t->st));
    return MFS_OK;
  });
}

}  // extern "C"

// ====================================================================== multigrid on general levels
namespace {

// y = A x (add: y += A x), CSR
__global__ void k_spmv(const long long rows, const long long *__restrict__ ptr, const long long *__restrict__ col,
                       const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
                       const int add) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double a = 0;
  for (long long k = ptr[i]; k < ptr[i + 1]; ++k) a += val[k] * x[col[k]];
  y[i] = add ? y[i] + a : a;
}
__global__ void k_probe_set(const long long n, const int *__restrict__ colour, const int c, double *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = colour[i] == c ? 1.0 : 0.0;
}
__global__ void k_probe_get(const long long n, const int *__restrict__ colour, const int c,
                            const double *__restrict__ y, double *__restrict__ diag) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n && colour[i] == c) diag[i] = y[i];
}
__global__ void k_zero_rows(const long long n, const int *__restrict__ row, double *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n && row[i] >= 0) x[i] = 0.0;
}
__global__ void k_invert(const long long n, double *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = 1.0 / x[i];
}
inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// Active table of a general level from the viscosity at the quadrature points
// (natural [cell][q]): mode 0 as given, 1/2/3 the arithmetic / harmonic /
// geometric mean of the cell's values at every point (material averaging,
// source/material_model/interface.cc:556-680); and the DGQ0 value of the cell
// (JxW-weighted mean of the table: the cellwise L2 projection onto constants)
__global__ void k_tria_active_eta(const long long ncells, const int nq, const int G, const double *__restrict__ eta,
                                  const int mode, const double *__restrict__ geo, double *__restrict__ table,
                                  double *__restrict__ cellval) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const long long b = c / 64, cc = c - b * 64;
  double acc = 0;
  bool zero = false;
  for (int q = 0; q < nq; ++q) {
    const double v = eta[c * nq + q];
    if (mode == 2) {
      if (v == 0.0) zero = true;
      else acc += 1.0 / v;
    } else if (mode == 3) {
      acc += log(v);
    } else {
      acc += v;
    }
  }
  const double avg = mode == 0 ? 0.0 : mode == 1 ? acc / nq : mode == 2 ? (zero ? 0.0 : nq / acc) : exp(acc / nq);
  double s = 0, ws = 0;
  for (int q = 0; q < nq; ++q) {
    const double v = mode == 0 ? eta[c * nq + q] : avg;
    table[(b * nq + q) * 64 + cc] = v;
    const double w = geo[((b * nq + q) * G + G - 1) * 64 + cc];
    s += w * v;
    ws += w;
  }
  cellval[c] = s / ws;
}
// coarse cell value = mean of its fine cells (the FE_DGQ(0) restriction: the
// children of a refined cell, or the same cell when it was not refined)
__global__ void k_tria_restrict(const long long ncoarse, const long long *__restrict__ ptr,
                                const long long *__restrict__ idx, const double *__restrict__ fine,
                                double *__restrict__ coarse) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= ncoarse) return;
  double s = 0;
  for (long long k = ptr[c]; k < ptr[c + 1]; ++k) s += fine[idx[k]];
  coarse[c] = s / (double)(ptr[c + 1] - ptr[c]);
}
__global__ void k_tria_fill_eta(const long long ncells, const int nq, const double *__restrict__ cellval,
                                double *__restrict__ table) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= ncells * nq) return;
  const long long c = i / nq, q = i - c * nq, b = c / 64, cc = c - b * 64;
  table[(b * nq + q) * 64 + cc] = cellval[c];
}

struct DevCsr {
  long long rows = 0, cols = 0;
  Buf<long long> ptr, col;
  Buf<double> val;
  void set(long long r, long long c, const std::vector<long long> &p, const std::vector<long long> &ci,
           const std::vector<double> &v, hipStream_t st) {
    rows = r;
    cols = c;
    ptr.upload(p, st);
    col.upload(ci, st);
    val.upload(v, st);
  }
  void apply(const double *x, double *y, bool add, hipStream_t st) const {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_spmv, dim3(nblk(rows)), dim3(256), 0, st, rows, ptr.p, col.p, val.p, x, y, add ? 1 : 0);
    MFS_HIP(hipGetLastError());
  }
};

struct Cheb {
  double lmin = 1, lmax = 1, theta = 1, delta = 0;
  int degree = 4, its = 0;
};

struct TLevel {
  mfs_tria *t = nullptr;
  long long n[2] = {0, 0};
  Buf<double> dinv[2], x[2], b[2], r[2], d[2];
  Buf<int> colour[2];
  int ncolours[2] = {0, 0};
  Cheb ch[2];
  DevCsr P[2], R[2];  // prolongation from the next coarser level and its transpose
  Buf<long long> kid_ptr, kid_idx;  // cells of the next finer level inside each cell of this level
  Buf<double> cellval;              // DGQ0 viscosity of every cell
};

}  // namespace

struct mfs_tria_gmg {
  hipStream_t st = nullptr;
  std::vector<TLevel> lv;
  Buf<double> dwork, dout, rtmp;
  bool built = false;

  // each V-cycle kind captured once into a graph after build() (the level
  // right-hand side and solution buffers are fixed, so the graph replays
  // unchanged); dropped when the smoothers change
  hipGraphExec_t vexec[2] = {nullptr, nullptr};
  bool graphs = true;
  void drop_graphs() {
    for (auto &g : vexec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }

  ~mfs_tria_gmg() {
    drop_graphs();
    if (st) (void)hipStreamDestroy(st);
  }

  double dot(const double *a, const double *b, long long n) {
    v_dot<double>(a, b, n, dwork.p, dout.p, st);
    double h = 0;
    MFS_HIP(hipMemcpyAsync(&h, dout.p, sizeof(double), hipMemcpyDeviceToHost, st));
    MFS_HIP(hipStreamSynchronize(st));
    return h;
  }
  void op(int kind, int l, const double *x, double *y) {
    lv[l].t->vmult(kind == 0 ? MFS_OP_ABLOCK : MFS_OP_MASS, x, y, st);
  }

  static void csr_from(const mfs_csr &m, long long rows, long long cols, std::vector<long long> &p,
                       std::vector<long long> &c, std::vector<double> &v, const char *what) {
    if (m.rows != rows || m.cols != cols)
      throw std::invalid_argument(std::string(what) + ": prolongation size does not match the levels");
    if (!m.row_ptr || (m.row_ptr[rows] > 0 && (!m.col_idx || !m.values)))
      throw std::invalid_argument(std::string(what) + ": null CSR array");
    p.assign(m.row_ptr, m.row_ptr + rows + 1);
    if (p[0] != 0) throw std::invalid_argument(std::string(what) + ": row_ptr[0] must be 0");
    for (long long i = 0; i < rows; ++i)
      if (p[i + 1] < p[i]) throw std::invalid_argument(std::string(what) + ": row_ptr not increasing");
    c.assign(m.col_idx, m.col_idx + p[rows]);
    v.assign(m.values, m.values + p[rows]);
    for (long long k = 0; k < p[rows]; ++k)
      if (c[k] < 0 || c[k] >= cols) throw std::invalid_argument(std::string(what) + ": column out of range");
  }
  static void transpose(long long rows, long long cols, const std::vector<long long> &p, const std::vector<long long> &c,
                        const std::vector<double> &v, std::vector<long long> &tp, std::vector<long long> &tc,
                        std::vector<double> &tv) {
    tp.assign(cols + 1, 0);
    for (long long k = 0; k < p[rows]; ++k) ++tp[c[k] + 1];
    for (long long j = 0; j < cols; ++j) tp[j + 1] += tp[j];
    tc.assign(p[rows], 0);
    tv.assign(p[rows], 0.0);
    std::vector<long long> fill(tp.begin(), tp.end() - 1);
    for (long long i = 0; i < rows; ++i)
      for (long long k = p[i]; k < p[i + 1]; ++k) {
        const long long pos = fill[c[k]]++;
        tc[pos] = i;
        tv[pos] = v[k];
      }
  }

  bool have_parents = false;

  void create(int n_levels, mfs_tria *const *levels, const mfs_csr *pu, const mfs_csr *pp,
              const long long *const *parents) {
    if (n_levels < 1 || !levels) throw std::invalid_argument("mfs_tria_gmg_create: no levels");
    if (n_levels > 1 && (!pu || !pp)) throw std::invalid_argument("mfs_tria_gmg_create: null prolongations");
    for (int l = 0; l < n_levels; ++l) {
      if (!levels[l]) throw std::invalid_argument("mfs_tria_gmg_create: null level");
      if (levels[l]->dim != levels[0]->dim) throw std::invalid_argument("mfs_tria_gmg_create: levels of mixed dim");
    }
    std::vector<std::vector<long long>> P[2][3], R[2][3];
    lv.resize(n_levels);
    std::vector<long long> p, c, tp, tc;
    std::vector<double> v, tv;
    MFS_HIP(hipStreamCreate(&st));
    for (int l = 0; l < n_levels; ++l) {
      TLevel &L = lv[l];
      L.t = levels[l];
      L.n[0] = L.t->ndof_u;
      L.n[1] = L.t->ndof_p;
      for (int k = 0; k < 2; ++k) {
        L.dinv[k].alloc(L.n[k]);
        L.x[k].alloc(L.n[k]);
        L.b[k].alloc(L.n[k]);
        L.r[k].alloc(L.n[k]);
        L.d[k].alloc(L.n[k]);
      }
      if (l == 0) continue;
      for (int k = 0; k < 2; ++k) {
        const mfs_csr &m = k == 0 ? pu[l - 1] : pp[l - 1];
        csr_from(m, L.n[k], lv[l - 1].n[k], p, c, v, k == 0 ? "velocity" : "pressure");
        transpose(L.n[k], lv[l - 1].n[k], p, c, v, tp, tc, tv);
        L.P[k].set(L.n[k], lv[l - 1].n[k], p, c, v, st);
        L.R[k].set(lv[l - 1].n[k], L.n[k], tp, tc, tv, st);
      }
    }
    if (parents) {
      // children lists of every level from the fine-to-coarse cell maps
      for (int l = 1; l < n_levels; ++l) {
        const long long nf = lv[l].t->ncells, nc = lv[l - 1].t->ncells;
        const long long *par = parents[l - 1];
        if (!par) throw std::invalid_argument("mfs_tria_gmg_create: null parent map");
        std::vector<long long> ptr(nc + 1, 0), idx(nf);
        for (long long f = 0; f < nf; ++f) {
          if (par[f] < 0 || par[f] >= nc) throw std::invalid_argument("mfs_tria_gmg_create: parent out of range");
          ++ptr[par[f] + 1];
        }
        for (long long c = 0; c < nc; ++c) {
          if (ptr[c + 1] == 0) throw std::invalid_argument("mfs_tria_gmg_create: a coarse cell without fine cells");
          ptr[c + 1] += ptr[c];
        }
        std::vector<long long> fill(ptr.begin(), ptr.end() - 1);
        for (long long f = 0; f < nf; ++f) idx[fill[par[f]]++] = f;
        lv[l - 1].kid_ptr.upload(ptr, st);
        lv[l - 1].kid_idx.upload(idx, st);
      }
      for (auto &L : lv) L.cellval.alloc(L.t->ncells);
      have_parents = true;
    }
    long long nmax = 0;
    for (auto &L : lv) nmax = std::max(nmax, L.n[0] + L.n[1]);
    dwork.alloc(dot_work_size());
    dout.alloc(1);
    rtmp.alloc(nmax);
    MFS_HIP(hipStreamSynchronize(st));
  }

  // viscosity of every level from the finest level's values at the quadrature
  // points (the global-coarsening evaluate_material_model,
  // stokes_matrix_free_global_coarsening.cc:173-330: active table with the
  // material averaging, its DGQ0 projection transferred down the levels,
  // one value per cell on the levels)
  void set_viscosity(const double *eta_qp, int averaging) {
    if (!have_parents) throw std::runtime_error("mfs_tria_gmg_set_viscosity: created without parent maps");
    if (!eta_qp) throw std::invalid_argument("mfs_tria_gmg_set_viscosity: null viscosity");
    if (averaging < 0 || averaging > 3)
      throw std::invalid_argument("mfs_tria_gmg_set_viscosity: averaging must be none, arithmetic, harmonic or geometric");
    TLevel &F = lv.back();
    mfs_tria *t = F.t;
    const long long nc = t->ncells;
    for (long long i = 0; i < nc * t->nq; ++i)
      if (!(eta_qp[i] > 0) || !std::isfinite(eta_qp[i]))
        throw std::invalid_argument("mfs_tria_gmg_set_viscosity: viscosity must be positive");
    Buf<double> nat;
    nat.alloc(nc * t->nq);
    MFS_HIP(hipMemcpyAsync(nat.p, eta_qp, nc * t->nq * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tria_active_eta, dim3(nblk(nc)), dim3(256), 0, st, nc, t->nq, t->dim * t->dim + 1, nat.p,
                       averaging, t->geo.p, t->eta.p, F.cellval.p);
    for (int l = (int)lv.size() - 2; l >= 0; --l) {
      TLevel &C = lv[l];
      const long long n = C.t->ncells;
      hipLaunchKernelGGL(k_tria_restrict, dim3(nblk(n)), dim3(256), 0, st, n, C.kid_ptr.p, C.kid_idx.p,
                         lv[l + 1].cellval.p, C.cellval.p);
      hipLaunchKernelGGL(k_tria_fill_eta, dim3(nblk(n * C.t->nq)), dim3(256), 0, st, n, C.t->nq, C.cellval.p,
                         C.t->eta.p);
    }
    for (auto &L : lv) L.t->update_mass(st);
    MFS_HIP(hipGetLastError());
    MFS_HIP(hipStreamSynchronize(st));
    for (auto &L : lv) L.t->have_eta = true;
    drop_graphs();
    built = false;
  }

  // exact diagonal of C^T A C (identity on constrained DoFs) by probing
  // colour classes with the level operator
  void diagonal(int kind, int l) {
    TLevel &L = lv[l];
    const long long n = L.n[kind];
    if (L.ncolours[kind] == 0) {
      std::vector<int> col;
      L.ncolours[kind] = L.t->colour_dofs(kind, col);
      L.colour[kind].upload(col, st);
    }
    v_fill<double>(L.dinv[kind].p, 1.0, n, st);
    for (int cc = 0; cc < L.ncolours[kind]; ++cc) {
      hipLaunchKernelGGL(k_probe_set, dim3(nblk(n)), dim3(256), 0, st, n, L.colour[kind].p, cc, L.x[kind].p);
      op(kind, l, L.x[kind].p, L.r[kind].p);
      hipLaunchKernelGGL(k_probe_get, dim3(nblk(n)), dim3(256), 0, st, n, L.colour[kind].p, cc, L.r[kind].p,
                         L.dinv[kind].p);
    }
    MFS_HIP(hipGetLastError());
  }

  static double tridiag_eig(const std::vector<double> &d, const std::vector<double> &e, bool want_max) {
    // bisection with Sturm counts on the symmetric tridiagonal (d, e)
    double lo = 1e300, hi = -1e300;
    for (size_t i = 0; i < d.size(); ++i) {
      const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < d.size() ? std::fabs(e[i]) : 0.0);
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

  // Chebyshev set-up: Jacobi-preconditioned CG from the mean-free pattern
  // (i mod 11), stopped at |r| < 1e-2 or max_its; Lanczos tridiagonal from the
  // CG coefficients; lambda_max scaled by 1.2
  void estimate(int kind, int l, int max_its, double range) {
    TLevel &L = lv[l];
    Cheb &ch = L.ch[kind];
    const long long n = L.n[kind];
    double *dinv = L.dinv[kind].p, *r = L.r[kind].p, *z = L.d[kind].p, *p = L.x[kind].p, *q = L.b[kind].p;
    std::vector<double> pat(n);
    double mean = 0;
    for (long long i = 0; i < n; ++i) mean += (double)(i % 11);
    mean /= (double)n;
    for (long long i = 0; i < n; ++i) pat[i] = (double)(i % 11) - mean;
    MFS_HIP(hipMemcpyAsync(r, pat.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    std::vector<double> alphas, betas;
    double res = std::sqrt(dot(r, r, n));
    if (res >= 1e-2) {
      v_cheb_first<double>(dinv, r, z, p, 1.0, true, n, st);
      double rz = dot(r, z, n);
      for (int it = 1; it <= max_its; ++it) {
        op(kind, l, p, q);
        const double alpha = rz / dot(p, q, n);
        v_axpby<double>(-alpha, q, 1.0, r, n, st);
        alphas.push_back(alpha);
        res = std::sqrt(dot(r, r, n));
        if (res < 1e-2 || it == max_its) break;
        v_mul<double>(dinv, r, z, n, st);
        const double rz2 = dot(r, z, n);
        const double beta = rz2 / rz;
        rz = rz2;
        v_axpby<double>(1.0, z, beta, p, n, st);
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

  void build() {
    drop_graphs();
    for (int l = 0; l < (int)lv.size(); ++l) {
      for (int k = 0; k < 2; ++k) {
        if (!lv[l].t->have_eta) throw std::runtime_error("mfs_tria_gmg_build: viscosity not set on every level");
        diagonal(k, l);
        hipLaunchKernelGGL(k_invert, dim3(nblk(lv[l].n[k])), dim3(256), 0, st, lv[l].n[k], lv[l].dinv[k].p);
        const bool coarse = l == 0;
        lv[l].ch[k].degree = coarse ? 8 : 4;
        estimate(k, l, coarse ? 100 : 10, coarse ? 1e-3 : 15.0);
      }
    }
    MFS_HIP(hipStreamSynchronize(st));
    built = true;
  }

  // Chebyshev semi-iteration with Jacobi preconditioning, `steps`
  // applications, the first from zero when x_zero
  void smooth(int kind, int l, double *x, const double *b, bool x_zero, int steps) {
    TLevel &L = lv[l];
    const Cheb &ch = L.ch[kind];
    const long long n = L.n[kind];
    double *dinv = L.dinv[kind].p, *r = L.r[kind].p, *d = L.d[kind].p;
    (void)r;
    // residual and Chebyshev update fused into the assemble pass of the level operator
    auto step = [&](double f1, double f2, int first) {
      ChebStep cs{b, dinv, d, f1, f2, first};
      L.t->level_op(kind, x, x, AS_CHEB, cs, st);
    };
    for (int s = 0; s < steps; ++s) {
      if (x_zero && s == 0)
        v_cheb_first<double>(dinv, b, d, x, 1.0 / ch.theta, true, n, st);
      else
        step(0.0, 1.0 / ch.theta, 1);
      if (ch.degree < 2 || std::fabs(ch.delta) < 1e-40) continue;
      const double sigma = ch.theta / ch.delta;
      double rho_old = ch.delta / ch.theta;
      for (int k = 1; k < ch.degree; ++k) {
        const double rho = 1.0 / (2.0 * sigma - rho_old);
        step(rho * rho_old, 2.0 * rho / ch.delta, 0);
        rho_old = rho;
      }
    }
  }

  void vcycle_level(int kind, int l) {
    TLevel &L = lv[l];
    const int steps = kind == 0 ? 1 : 4;
    double *x = L.x[kind].p, *b = L.b[kind].p, *r = L.r[kind].p;
    smooth(kind, l, x, b, true, steps);
    if (l == 0) return;
    ChebStep cs{b, nullptr, nullptr, 0.0, 0.0, 0};
    L.t->level_op(kind, x, r, AS_RESID, cs, st);  // r = b - A x
    TLevel &C = lv[l - 1];
    L.R[kind].apply(r, C.b[kind].p, false, st);
    vcycle_level(kind, l - 1);
    L.P[kind].apply(C.x[kind].p, x, true, st);
    smooth(kind, l, x, b, false, steps);
  }
  // one V-cycle on device vectors of the finest level
  void vcycle(int kind, const double *src, double *dst) {
    if (!built) throw std::runtime_error("multigrid not built (call mfs_tria_gmg_build)");
    TLevel &L = lv.back();
    MFS_HIP(hipMemcpyAsync(L.b[kind].p, src, L.n[kind] * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (graphs) {
      if (!vexec[kind]) {
        hipGraph_t gr = nullptr;
        MFS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          vcycle_level(kind, (int)lv.size() - 1);
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
      vcycle_level(kind, (int)lv.size() - 1);
    }
    MFS_HIP(hipMemcpyAsync(dst, L.x[kind].p, L.n[kind] * sizeof(double), hipMemcpyDeviceToDevice, st));
  }

  // BlockSchurPreconditioner::vmult, cheap variant (block_stokes_preconditioner.h:220-255):
  // p = -S^{-1} r_p, u = A^{-1} (r_u - B^T p), one V-cycle each
  void block_prec(const double *src, double *dst) {
    TLevel &L = lv.back();
    const long long nu = L.n[0], np = L.n[1];
    vcycle(1, src + nu, rtmp.p);
    v_scale_neg_copy<double>(rtmp.p, dst + nu, np, st);
    L.t->vmult(MFS_OP_BT, dst + nu, rtmp.p, st);
    v_axpby<double>(1.0, src, -1.0, rtmp.p, nu, st);  // r_u - B^T p
    vcycle(0, rtmp.p, dst);
  }

  // preconditioned CG on the finest level's A block (kind 0) or mass (kind 1)
  // with one V-cycle as preconditioner, from zero, to |r| <= tol; returns
  // the iterations (the inner solves of the expensive preconditioner)
  int cg(int kind, const double *b, double *x, double tol, int maxit) {
    TLevel &L = lv.back();
    const long long n = L.n[kind];
    const int l = (int)lv.size() - 1;
    for (Buf<double> *v : {&cg_r[kind], &cg_z[kind], &cg_p[kind], &cg_q[kind]})
      if (!v->p) v->alloc(n);
    double *r = cg_r[kind].p, *z = cg_z[kind].p, *p = cg_p[kind].p, *q = cg_q[kind].p;
    v_fill<double>(x, 0.0, n, st);
    MFS_HIP(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    double res = std::sqrt(dot(r, r, n));
    if (res <= tol) return 0;
    vcycle(kind, r, z);
    double rz = dot(r, z, n);
    MFS_HIP(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int it = 1; it <= maxit; ++it) {
      op(kind, l, p, q);
      const double alpha = rz / dot(p, q, n);
      v_axpby<double>(alpha, p, 1.0, x, n, st);
      v_axpby<double>(-alpha, q, 1.0, r, n, st);
      res = std::sqrt(dot(r, r, n));
      if (res <= tol) return it;
      vcycle(kind, r, z);
      const double rz2 = dot(r, z, n);
      const double beta = rz2 / rz;
      rz = rz2;
      v_axpby<double>(1.0, z, beta, p, n, st);
    }
    return maxit;
  }
  Buf<double> cg_r[2], cg_z[2], cg_p[2], cg_q[2], cg_t;
  double a_tol = 1e-2, s_tol = 1e-6;
  long long a_apps = 0, s_apps = 0;

  // BlockSchurPreconditioner::vmult, expensive variant: S^{-1} by CG to
  // s_tol |r_p| (at most 100 iterations; skipped when |r_p| <= 1e-50), A^{-1}
  // by CG to a_tol |r_u - B^T p| (block_stokes_preconditioner.h:220-255, 308-352)
  void block_prec_expensive(const double *src, double *dst) {
    TLevel &L = lv.back();
    const long long nu = L.n[0], np = L.n[1];
    if (!cg_t.p) cg_t.alloc(std::max(nu, np));
    const double nsp = std::sqrt(dot(src + nu, src + nu, np));
    if (nsp > 1e-50) {
      s_apps += cg(1, src + nu, cg_t.p, nsp * s_tol, 100);
      v_scale_neg_copy<double>(cg_t.p, dst + nu, np, st);
    } else {
      v_fill<double>(dst + nu, 0.0, np, st);
    }
    L.t->vmult(MFS_OP_BT, dst + nu, rtmp.p, st);
    v_axpby<double>(1.0, src, -1.0, rtmp.p, nu, st);
    const double nt = std::sqrt(dot(rtmp.p, rtmp.p, nu));
    a_apps += cg(0, rtmp.p, dst, nt * a_tol, 10000);
  }

  // one classical Gram-Schmidt pass against the first m basis vectors (one
  // fused partial-dot sweep, one fixed-order reduction, one fused update),
  // coefficients to the host
  Buf<double> gs_work, gs_h;
  void cgs_pass(const double *Vbase, long long n, int m, double *w, std::vector<double> &h) {
    VecSet V{};
    for (int c = 0; c * 8 < m; ++c) V.p[c] = Vbase + (long long)8 * c * n;
    v_mdot_all(V, n, m, w, n, gs_work.p, st);
    v_sum_partials(gs_work.p, m, gs_h.p, st);
    v_maxpy_all(V, n, m, gs_h.p, w, n, -1.0, st);
    h.resize(m);
    MFS_HIP(hipMemcpyAsync(h.data(), gs_h.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    MFS_HIP(hipStreamSynchronize(st));
  }

  // right-preconditioned restarted GMRES from the initial guess in x,
  // keeping the preconditioned vectors (flexible: the expensive
  // preconditioner is not a fixed linear map); classical Gram-Schmidt with a
  // second pass when the first cancelled more than 1/sqrt(2) of |w|
  template <typename Prec>
  void gmres(const double *b, double *x, double tol, int maxit, int restart, Prec prec, int &its, double &res) {
    TLevel &L = lv.back();
    const long long n = L.n[0] + L.n[1];
    if (restart + 1 > MV_MAX) throw std::invalid_argument("mfs_tria_solve: restart must be below 127");
    Buf<double> Vb, w;
    Vb.alloc((restart + 1) * n);
    std::vector<Buf<double>> Z(restart);
    for (auto &z : Z) z.alloc(n);
    w.alloc(n);
    if (!gs_work.p) {
      gs_work.alloc((long long)MV_MAX * dot_work_size());
      gs_h.alloc(MV_MAX);
    }
    auto Vp = [&](int i) { return Vb.p + (long long)i * n; };
    std::vector<double> hv, hv2;
    its = 0;
    L.t->vmult(MFS_OP_STOKES, x, Vp(0), st);
    v_axpby<double>(1.0, b, -1.0, Vp(0), n, st);  // r = b - K x
    double beta = std::sqrt(dot(Vp(0), Vp(0), n));
    while (its < maxit && beta > tol) {
      std::vector<double> H((restart + 1) * restart, 0.0), cs(restart), sn(restart), g(restart + 1, 0.0);
      auto h = [&](int i, int k) -> double & { return H[i * restart + k]; };
      v_axpby<double>(0.0, Vp(0), 1.0 / beta, Vp(0), n, st);
      g[0] = beta;
      int k = 0;
      while (k < restart && its < maxit) {
        prec(Vp(k), Z[k].p);
        L.t->vmult(MFS_OP_STOKES, Z[k].p, w.p, st);
        const double w0 = std::sqrt(dot(w.p, w.p, n));
        cgs_pass(Vb.p, n, k + 1, w.p, hv);
        double wn = std::sqrt(dot(w.p, w.p, n));
        if (wn < 0.70710678118654752 * w0) {
          cgs_pass(Vb.p, n, k + 1, w.p, hv2);
          for (int i = 0; i <= k; ++i) hv[i] += hv2[i];
          wn = std::sqrt(dot(w.p, w.p, n));
        }
        for (int i = 0; i <= k; ++i) h(i, k) = hv[i];
        h(k + 1, k) = wn;
        MFS_HIP(hipMemcpyAsync(Vp(k + 1), w.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (h(k + 1, k) > 0) v_axpby<double>(0.0, w.p, 1.0 / h(k + 1, k), Vp(k + 1), n, st);
        for (int i = 0; i < k; ++i) {
          const double t = cs[i] * h(i, k) + sn[i] * h(i + 1, k);
          h(i + 1, k) = -sn[i] * h(i, k) + cs[i] * h(i + 1, k);
          h(i, k) = t;
        }
        const double den = std::hypot(h(k, k), h(k + 1, k));
        cs[k] = h(k, k) / den;
        sn[k] = h(k + 1, k) / den;
        h(k, k) = den;
        h(k + 1, k) = 0.0;
        g[k + 1] = -sn[k] * g[k];
        g[k] = cs[k] * g[k];
        ++k;
        ++its;
        if (std::fabs(g[k]) <= tol) break;
      }
      std::vector<double> y(k, 0.0);
      for (int i = k - 1; i >= 0; --i) {
        double a = g[i];
        for (int j = i + 1; j < k; ++j) a -= h(i, j) * y[j];
        y[i] = a / h(i, i);
      }
      for (int i = 0; i < k; ++i) v_axpby<double>(y[i], Z[i].p, 1.0, x, n, st);
      beta = std::fabs(g[k]);
      if (its < maxit && beta > tol) {
        // restart from the true residual
        L.t->vmult(MFS_OP_STOKES, x, Vp(0), st);
        v_axpby<double>(1.0, b, -1.0, Vp(0), n, st);
        beta = std::sqrt(dot(Vp(0), Vp(0), n));
      }
    }
    res = beta;
    MFS_HIP(hipStreamSynchronize(st));
  }
};

extern "C" {

int mfs_tria_gmg_create(int n_levels, mfs_tria *const *levels, const mfs_csr *prolong_u, const mfs_csr *prolong_p,
                        const long long *const *parents, mfs_tria_gmg **out) {
  return tguard([&] {
    if (!out) throw std::invalid_argument("mfs_tria_gmg_create: null output");
    *out = nullptr;
    mfs_tria_gmg *g = new mfs_tria_gmg();
    try {
      g->create(n_levels, levels, prolong_u, prolong_p, parents);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
    return MFS_OK;
  });
}

int mfs_tria_gmg_destroy(mfs_tria_gmg *g) {
  return tguard([&] {
    if (g) {
      (void)hipStreamSynchronize(g->st);
      delete g;
    }
    return MFS_OK;
  });
}

int mfs_tria_gmg_set_viscosity(mfs_tria_gmg *g, const double *eta_qp, int averaging) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    g->set_viscosity(eta_qp, averaging);
    return MFS_OK;
  });
}

int mfs_tria_gmg_set_graph_capture(mfs_tria_gmg *g, int enable) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    MFS_HIP(hipStreamSynchronize(g->st));
    g->drop_graphs();
    g->graphs = enable != 0;
    return MFS_OK;
  });
}

int mfs_tria_gmg_build(mfs_tria_gmg *g) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    g->build();
    return MFS_OK;
  });
}

int mfs_tria_gmg_info(mfs_tria_gmg *g, int level, int kind, double out[4], double *diagonal) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    if (!g->built) throw std::runtime_error("multigrid not built (call mfs_tria_gmg_build)");
    if (level < 0 || level >= (int)g->lv.size() || (kind != 0 && kind != 1))
      throw std::invalid_argument("mfs_tria_gmg_info: level or kind out of range");
    const TLevel &L = g->lv[level];
    if (out) {
      out[0] = L.ch[kind].lmin;
      out[1] = L.ch[kind].lmax;
      out[2] = L.ch[kind].its;
      out[3] = L.ncolours[kind];
    }
    if (diagonal) {
      // the diagonal itself (the smoother stores its inverse)
      Buf<double> tmp;
      tmp.alloc(L.n[kind]);
      MFS_HIP(hipMemcpyAsync(tmp.p, L.dinv[kind].p, L.n[kind] * sizeof(double), hipMemcpyDeviceToDevice, g->st));
      hipLaunchKernelGGL(k_invert, dim3(nblk(L.n[kind])), dim3(256), 0, g->st, L.n[kind], tmp.p);
      MFS_HIP(hipMemcpyAsync(diagonal, tmp.p, L.n[kind] * sizeof(double), hipMemcpyDeviceToHost, g->st));
      MFS_HIP(hipStreamSynchronize(g->st));
    }
    return MFS_OK;
  });
}

int mfs_tria_gmg_vcycle(mfs_tria_gmg *g, int kind, const double *src, double *dst) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    if (kind != 0 && kind != 1) throw std::invalid_argument("mfs_tria_gmg_vcycle: kind must be 0 (A) or 1 (Schur)");
    if (!src || !dst) throw std::invalid_argument("mfs_tria_gmg_vcycle: null vector");
    const long long n = g->lv.back().n[kind];
    Buf<double> a, b;
    a.alloc(n);
    b.alloc(n);
    MFS_HIP(hipMemcpyAsync(a.p, src, n * sizeof(double), hipMemcpyHostToDevice, g->st));
    g->vcycle(kind, a.p, b.p);
    MFS_HIP(hipMemcpyAsync(dst, b.p, n * sizeof(double), hipMemcpyDeviceToHost, g->st));
    MFS_HIP(hipStreamSynchronize(g->st));
    return MFS_OK;
  });
}

int mfs_tria_solve(mfs_tria_gmg *g, const double *rhs, const double *x0, double *sol, const mfs_solve_params *prm,
                   mfs_solve_result *res) {
  return tguard([&] {
    if (!g) throw std::invalid_argument("null mfs_tria_gmg");
    if (!rhs || !sol || !prm || !res) throw std::invalid_argument("mfs_tria_solve: null argument");
    if (!g->built) throw std::runtime_error("multigrid not built (call mfs_tria_gmg_build)");
    if (prm->restart < 1 || prm->restart > 1000) throw std::invalid_argument("mfs_tria_solve: restart must be in [1, 1000]");
    if (prm->n_cheap_steps < 0 || prm->n_expensive_steps < 0) throw std::invalid_argument("mfs_tria_solve: negative step limit");
    if (prm->failure_strategy != MFS_FAILURE_ABORT && prm->failure_strategy != MFS_FAILURE_CONTINUE)
      throw std::invalid_argument("mfs_tria_solve: unknown linear solver failure strategy");
    std::memset(res, 0, sizeof(*res));
    const long long n = g->lv.back().n[0] + g->lv.back().n[1];
    Buf<double> b, x;
    b.alloc(n);
    x.alloc(n);
    MFS_HIP(hipMemcpyAsync(b.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, g->st));
    TLevel &T = g->lv.back();
    const long long nu = T.n[0], np = T.n[1];
    const double s = T.t->L.s;
    double r0, inr;
    if (x0 && !prm->solve_newton_system) {
      // initial guess: constrained velocity entries zeroed, pressure / s
      // (local_smoothing.cc:944-975); the nonlinear residual |rhs - K x0| and
      // the tolerance from the residual with the velocity zeroed (:977-1025)
      MFS_HIP(hipMemcpyAsync(x.p, x0, n * sizeof(double), hipMemcpyHostToDevice, g->st));
      hipLaunchKernelGGL(k_zero_rows, dim3(nblk(nu)), dim3(256), 0, g->st, nu, T.t->con_u.row.p, x.p);
      v_scale(x.p + nu, 1.0 / s, np, g->st);
      Buf<double> w, z;
      w.alloc(n);
      z.alloc(n);
      T.t->vmult(MFS_OP_STOKES, x.p, w.p, g->st);
      v_axpby<double>(1.0, b.p, -1.0, w.p, n, g->st);
      inr = std::sqrt(g->dot(w.p, w.p, n));
      MFS_HIP(hipMemcpyAsync(z.p, x.p, n * sizeof(double), hipMemcpyDeviceToDevice, g->st));
      v_fill<double>(z.p, 0.0, nu, g->st);
      T.t->vmult(MFS_OP_STOKES, z.p, w.p, g->st);
      v_axpby<double>(1.0, b.p, -1.0, w.p, n, g->st);
      r0 = std::sqrt(g->dot(w.p, w.p, nu) + g->dot(b.p + nu, b.p + nu, np));
    } else {
      // zero initial guess (a Newton update's best guess, :950-969); tolerance from |rhs|
      v_fill<double>(x.p, 0.0, n, g->st);
      r0 = inr = std::sqrt(g->dot(b.p, b.p, n));
    }
    const double tol = prm->tolerance * r0;
    hipEvent_t e0, e1;
    MFS_HIP(hipEventCreate(&e0));
    MFS_HIP(hipEventCreate(&e1));
    MFS_HIP(hipEventRecord(e0, g->st));
    int its = 0, its_exp = 0;
    double fin = inr;
    hipEvent_t em;
    MFS_HIP(hipEventCreate(&em));
    if (prm->n_cheap_steps > 0)
      g->gmres(b.p, x.p, tol, prm->n_cheap_steps, prm->restart,
               [&](const double *a, double *z) { g->block_prec(a, z); }, its, fin);
    MFS_HIP(hipEventRecord(em, g->st));
    const double cheap_res = fin;
    g->a_apps = g->s_apps = 0;
    g->a_tol = prm->a_block_tolerance;
    g->s_tol = prm->s_block_tolerance;
    if (fin > tol && prm->n_expensive_steps > 0)
      g->gmres(b.p, x.p, tol, prm->n_expensive_steps, prm->restart,
               [&](const double *a, double *z) { g->block_prec_expensive(a, z); }, its_exp, fin);
    MFS_HIP(hipEventRecord(e1, g->st));
    MFS_HIP(hipEventSynchronize(e1));
    float ms = 0, ms_cheap = 0;
    MFS_HIP(hipEventElapsedTime(&ms, e0, e1));
    MFS_HIP(hipEventElapsedTime(&ms_cheap, e0, em));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(em);
    v_scale(x.p + nu, s, np, g->st);  // post-step: pressure * s (:1406-1445)
    MFS_HIP(hipMemcpyAsync(sol, x.p, n * sizeof(double), hipMemcpyDeviceToHost, g->st));
    MFS_HIP(hipStreamSynchronize(g->st));
    res->cheap_iterations = its;
    res->expensive_iterations = its_exp;
    res->a_preconditioner_applications = its + (int)g->a_apps;
    res->s_preconditioner_applications = its + (int)g->s_apps;
    res->converged = fin <= tol ? 1 : 0;
    res->initial_nonlinear_residual = inr;
    res->r0_norm = r0;
    res->final_linear_residual = fin;
    res->cheap_final_residual = cheap_res;
    res->solver_tolerance = tol;
    res->solve_seconds = ms * 1e-3;
    res->cheap_seconds = ms_cheap * 1e-3;
    if (res->converged) return MFS_OK;
    // "Linear solver failure strategy" (local_smoothing.cc:1353-1395)
    if (prm->failure_strategy == MFS_FAILURE_CONTINUE) return MFS_OK;
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "the iterative Stokes solver on the general mesh did not converge: initial residual %.6e, "
                  "final residual %.6e, required %.6e",
                  r0, fin, tol);
    set_last_error(msg);
    return MFS_NO_CONVERGENCE;
  });
}

}  // extern "C"
