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
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aspect_mfs.h"
#include "mfs_internal.h"

using namespace mfs;

namespace {

constexpr int CT = 64;  // cells per batch (the cell kernel's tile size for degree 2)

template <typename T>
struct Buf {
  T *p = nullptr;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
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
                                           const long long *__restrict__ ioff, long long node, long long coff) {
  double a = 0;
  for (long long k = iptr[node]; k < iptr[node + 1]; ++k) a += cb[ioff[k] + coff];
  return a;
}

// dst[i] for the DoFs of one block (nn nodes per component, component stride
// cs in the contribution buffer): free DoFs their own sum plus the weighted
// sums of the constrained DoFs they are a master of; constrained DoFs
// src[i] (ident) or 0
__global__ void k_assemble(const long long n, const long long nn, const long long cs, const double *__restrict__ cb,
                           const long long *__restrict__ iptr, const long long *__restrict__ ioff,
                           const int *__restrict__ row, const long long *__restrict__ tptr,
                           const long long *__restrict__ tdof, const double *__restrict__ tw,
                           const double *__restrict__ src, double *__restrict__ dst, const int ident) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (row[i] >= 0) {
    dst[i] = ident ? src[i] : 0.0;
    return;
  }
  const long long comp = i / nn, node = i - comp * nn;
  double a = node_sum(cb, iptr, ioff, node, comp * cs);
  for (long long k = tptr[i]; k < tptr[i + 1]; ++k) {
    const long long c = tdof[k], cc = c / nn;
    a += tw[k] * node_sum(cb, iptr, ioff, c - cc * nn, cc * cs);
  }
  dst[i] = a;
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
    h_row.clear();
    h_off.clear();
    h_mas.clear();
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
  Buf<long long> iptr_u, ioff_u, iptr_p, ioff_p;
  DevCon con_u, con_p;
  bool have_eta = false;

  ~mfs_tria() {
    if (st) (void)hipStreamDestroy(st);
  }

  // node -> (offset of its contribution in the per-cell buffer, component 0)
  static void incidence(const int *cells, long long ncells, int per, long long nn, int comps, std::vector<long long> &ptr,
                        std::vector<long long> &off) {
    ptr.assign(nn + 1, 0);
    for (long long c = 0; c < ncells * per; ++c) ++ptr[cells[c] + 1];
    for (long long i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
    off.assign(ptr[nn], 0);
    std::vector<long long> fill(ptr.begin(), ptr.end() - 1);
    for (long long c = 0; c < ncells; ++c) {
      const long long b = c / CT, cc = c % CT;
      for (int a = 0; a < per; ++a) off[fill[cells[c * per + a]]++] = ((b * comps) * per + a) * CT + cc;
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
    cell_u.upload(bu, st);
    cell_p.upload(bp, st);
    geo.upload(bg, st);
    std::vector<long long> ptr, off;
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
    MFS_HIP(hipStreamSynchronize(st));
    have_eta = true;
  }

  void resolve(const double *src, double *out, DevCon &c) {
    const unsigned nb = (unsigned)((c.n + 255) / 256);
    hipLaunchKernelGGL(k_resolve, dim3(nb), dim3(256), 0, st, c.n, src, c.row.p, c.off.p, c.mas.p, c.w.p, out);
    MFS_HIP(hipGetLastError());
  }
  void assemble(const double *cb, long long nn, long long cs, const Buf<long long> &ip, const Buf<long long> &io,
                DevCon &c, const double *src, double *dst, int ident) {
    const unsigned nb = (unsigned)((c.n + 255) / 256);
    hipLaunchKernelGGL(k_assemble, dim3(nb), dim3(256), 0, st, c.n, nn, cs, cb, ip.p, io.p, c.row.p, c.tptr.p,
                       c.tdof.p, c.tw.p, src, dst, ident);
    MFS_HIP(hipGetLastError());
  }

  // device vectors: op MFS_OP_STOKES [u|p] -> [u|p], ABLOCK u -> u, BT p -> u, MASS p -> p
  void vmult(int op, const double *src, double *dst) {
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
    if (rd_u) resolve(su, xu.p, con_u);
    if (rd_p) resolve(sp, xp.p, con_p);
    launch_cells<double>((Op)op, L, nbatch, rd_u ? xu.p : nullptr, rd_p ? xp.p : nullptr, eta.p, cbu.p, cbp.p, st);
    if (wr_u) assemble(cbu.p, nn_u, (long long)nv * CT, iptr_u, ioff_u, con_u, su, du, op == MFS_OP_BT ? 0 : 1);
    if (wr_p) assemble(cbp.p, nn_p, 0, iptr_p, ioff_p, con_p, sp, dp, 1);
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

int mfs_tria_vmult_device(mfs_tria *t, int op, const double *d_src, double *d_dst) {
  return tguard([&] {
    need(t);
    t->vmult(op, d_src, d_dst);
    MFS_HIP(hipStreamSynchronize(t->st));
    return MFS_OK;
  });
}

}  // extern "C"
