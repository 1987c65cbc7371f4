// Internal interface between the host-side handler and the gfx950 kernels.
// Not part of the C-ABI (see include/aspect_mfs.h for that).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mfs {

#define MFS_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw std::runtime_error(std::string("HIP error '") + hipGetErrorString(e_) + "' at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));                \
  } while (0)

// OP_RHS: buoyancy right-hand side  sum_q rho g . phi JxW  (rho per quadrature point in the eta slot)
// OP_LAPLACE: vector Laplace grad u : grad v (coefficient 1) of the mesh-deformation solve; OP_LDIAG its diagonal
enum Op {
  OP_STOKES = 0,
  OP_ABLOCK = 1,
  OP_BT = 2,
  OP_MASS = 3,
  OP_ADIAG = 4,
  OP_MDIAG = 5,
  OP_RHS = 6,
  OP_LAPLACE = 7,
  OP_LDIAG = 8
};
// SET: dst = A src;  RESID: dst = rhs - A src;  ADD: dst += A src
enum Mode { MODE_SET = 0, MODE_RESID = 1, MODE_ADD = 2 };

// mode value of launch_ablock_split2: a fused Jacobi-Chebyshev step
constexpr int SPLIT_CHEB = 3;
// internal mode of the general cell kernel (A block, 2D degree 2): every tile
// in one launch, in-tile node sums to work[tile][comp][lz][ly][lx]
constexpr int MODE_TSPLIT = 3;
// internal mode of the general cell kernel on a general mesh: a workgroup
// takes a batch of CT cells (cell lists LevelDesc::cell_u / cell_p), reads the
// constraint-resolved source through them and writes every cell's nodal
// contributions to dst[batch][comp][node][CT] (velocity) and
// dst_p[batch][node][CT] (pressure); a node-wise pass assembles them
constexpr int MODE_CELLS = 4;
// mode value of launch_planes_split: every in-tile node sum to the full work
// array [tile][comp][lz][ly][lx] (the Chebyshev-fused A-block step)
constexpr int SPLIT_FULL_SUMS = 3;
// Operands of a Jacobi-Chebyshev step fused into the split A-block apply:
// r = b - A x, d = f2 dinv r (first step) or d = f1 d + f2 dinv r, x += d
// (constrained rows: r = b - x)
template <typename T>
struct ChebArgs {
  T *x;
  const T *b;
  const T *dinv;
  T *d;
  T f1, f2;
  int first;
};

// Geometry of one uniform box level.  Velocity nodes: (k*n+1) per direction,
// pressure nodes ((k-1)*n+1); both lexicographic with x fastest; the velocity
// block is component-major.  Cell data (viscosity) is tile-blocked:
// index = (tile_id * NQ_or_1 + q) * CT + cell_in_tile.
struct LevelDesc {
  int dim, k;
  int n[3];
  int nv[3];
  int np[3];
  int ntiles[3];
  long long vnodes, pnodes, cells, ntiles_total;
  double h[3];
  double ih[3];      // 1 / h
  double ihw[3][3];  // ih_i ih_k
  double vol;        // cell volume h0 h1 h2 (h0 h1 in 2D)
  float ih_f[3], ihw_f[3][3], vol_f, s_f;  // the same in single precision
  double s;          // pressure scaling
  int bc[6];      // 0 open, 1 no normal flux, 2 zero velocity
  int compressible;
  double gravity[3];  // used by OP_RHS only
  int plain_read;     // 1: constrained velocity DoFs are read as stored (read_dof_values_plain)
  // general geometry (nullptr: the uniform box): per quadrature point J^{-1}
  // (row-major, d xi_e / d x_d at e * dim + d) and JxW, tile-blocked
  // [((tile * NQ + q) * (dim * dim + 1) + g) * CT + cell_in_tile]
  const double *geo;
  // general meshes (MODE_CELLS): node ids of the cells of every batch of CT
  // cells, [(batch * 3^dim + a) * CT + c] velocity, [(batch * 2^dim + a) * CT + c]
  // pressure (a lexicographic in the cell, x fastest)
  const int *cell_u;
  const int *cell_p;
};

// Mesh constants of a level in the arithmetic type of a kernel (kernel
// arguments, so they stay scalar operands)
template <typename T>
struct LevelConst;
template <>
struct LevelConst<double> {
  __host__ __device__ static double ih(const LevelDesc &L, int d) { return L.ih[d]; }
  __host__ __device__ static double ihw(const LevelDesc &L, int i, int k) { return L.ihw[i][k]; }
  __host__ __device__ static double vol(const LevelDesc &L) { return L.vol; }
  __host__ __device__ static double s(const LevelDesc &L) { return L.s; }
};
template <>
struct LevelConst<float> {
  __host__ __device__ static float ih(const LevelDesc &L, int d) { return L.ih_f[d]; }
  __host__ __device__ static float ihw(const LevelDesc &L, int i, int k) { return L.ihw_f[i][k]; }
  __host__ __device__ static float vol(const LevelDesc &L) { return L.vol_f; }
  __host__ __device__ static float s(const LevelDesc &L) { return L.s_f; }
};
// fill the derived mesh constants (ih, ihw, vol and the single-precision copies)
void level_derive(LevelDesc &L);

// Tile shapes (cells) per (dim, degree); one workgroup per tile, one thread
// per (cell, y-slice).
int tile_cells(int dim, int k, int out[3]);
int tile_threads(int dim, int k);

void upload_basis_tables();

// message returned by mfs_last_error() on the calling thread
void set_last_error(const char *msg);

// general meshes: the cell kernel in MODE_CELLS over nbatch batches of 64 cells
// (degree 2, FP64; L.cell_u / L.cell_p / L.geo set); ops Stokes, A block, B^T, mass
template <typename T>
void launch_cells(Op op, const LevelDesc &L, long long nbatch, const T *su, const T *sp, const T *eta, T *cu, T *cp,
                  hipStream_t st);

// Host-side argument check of an operator launch: throws std::invalid_argument
// naming the first null pointer the (op, mode) kernel would read or write.
void check_op_args(Op op, Mode mode, const void *src_u, const void *src_p, const void *eta, const void *rhs_u,
                   const void *rhs_p, const void *dst_u, const void *dst_p);

template <typename T>
void launch_cell_op(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                    int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                    T *split_work = nullptr);

// 3D Q2 x Q1 Stokes and A-block applies with the plane kernel; returns false
// (nothing launched) for other (dim, degree, op).
// split_work (A block: split_work_size(L) elements; Stokes:
// stokes_split_work_size(L)): the apply then runs as one launch over all tiles
// instead of eight tile colours: nodes inside a tile (no other tile touches
// them) are finished by that launch, the in-tile sums of the tile-boundary
// nodes go to the work array and a node-wise pass over the tile-boundary
// nodes adds the at most 8 tiles sharing each.  split_work_size covers the
// full in-tile sums of the Chebyshev-fused step (3 x 729 per tile).
template <typename T>
bool launch_cell_q2(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                    int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st,
                    T *split_work = nullptr);
// Split A-block apply on a 2D degree-2 level (general cell kernel, one launch
// over all tiles, then a node-wise pass): mode MODE_SET / RESID / ADD, or
// SPLIT_CHEB with ch (a fused Jacobi-Chebyshev step; su = ch.x)
template <typename T>
void launch_ablock_split2(int mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                          T *work, const ChebArgs<T> &ch, hipStream_t st);
long long split_work_size(const LevelDesc &L);        // A block, also the Chebyshev-fused step
long long split_apply_work_size(const LevelDesc &L);  // A-block applies (SET / RESID / ADD) only
long long stokes_split_work_size(const LevelDesc &L);
// First pass of the split apply: every tile in one launch of the plane kernel
// (mode: MODE_SET / RESID / ADD for the tile-interior nodes, or
// SPLIT_FULL_SUMS; Stokes: FP64, SET / RESID / ADD)
template <typename T>
void launch_planes_split(int mode, const LevelDesc &L, const T *su, const T *eta, int pq, const T *ru, T *du,
                         T *work, hipStream_t st);
void launch_planes_split_stokes(int mode, const LevelDesc &L, const double *su, const double *sp,
                                const double *eta, int pq, const double *ru, const double *rp, double *du,
                                double *work_u, double *work_p, hipStream_t st);
// The plane kernel in tile colours (Stokes or A block, any mode)
template <typename T>
bool launch_cell_planes(Op op, Mode mode, const LevelDesc &L, const T *src_u, const T *src_p, const T *eta,
                        int eta_per_qp, const T *rhs_u, const T *rhs_p, T *dst_u, T *dst_p, hipStream_t st);
// Split A-block apply fused with a Jacobi-Chebyshev update (3D degree 2 with
// split_work only; returns false otherwise): r = b - A x (not stored), then
// d = dinv r f2 (first != 0) or d = f1 d + f2 dinv r, and x += d
template <typename T>
bool launch_ablock_cheb(const LevelDesc &L, T *x, const T *b, const T *eta, int eta_per_qp, const T *dinv, T *d,
                        T f1, T f2, int first, T *split_work, hipStream_t st);

// Transfers between level Lc (coarse) and Lf = refined Lc.  kind 0 velocity, 1 pressure.
// work: nullptr, or transfer_work_size(Lc, Lf) elements for the separable
// (one direction per pass) 3D degree-2 path
template <typename T>
void launch_prolongate(int kind, const LevelDesc &Lc, const LevelDesc &Lf, const T *src_c, T *dst_f, bool add,
                       hipStream_t st, T *work = nullptr);
template <typename T>
void launch_restrict(int kind, const LevelDesc &Lc, const LevelDesc &Lf, const T *src_f, T *dst_c,
                     hipStream_t st, T *work = nullptr);
long long transfer_work_size(const LevelDesc &Lc, const LevelDesc &Lf);

// Pressure mass apply (degree 2, 2D or 3D box, cellwise viscosity) from a
// natural-order table of inverse cell viscosities (launch_inv_eta_nat); false
// for degree 3, a general geometry, or a null ieta
template <typename T>
bool launch_mass_inv(Mode mode, const LevelDesc &L, const T *src, const T *ieta, const T *rhs, T *dst,
                     hipStream_t st);
// one fused Chebyshev step of the Schur smoother on such a level:
// d = (first ? 0 : f1 d) + f2 D^-1 (b - M x_in), x_out = x_in + d (x_out != x_in)
template <typename T>
bool launch_mass_cheb(const LevelDesc &L, const T *xin, const T *ieta, const T *b, const T *dinv, T *d, T *xout,
                      T f1, T f2, int first, hipStream_t st);
template <typename T>
void launch_inv_eta_nat(const LevelDesc &L, const T *eta_tiles, T *out, hipStream_t st);
// The Chebyshev steps of one Schur smoothing: d = (first ? 0 : f1 d) + f2 D^-1 r
struct MassChebSeq {
  static constexpr int MAX = 40;
  int n;
  double f1[MAX], f2[MAX];
  int first[MAX];
};
// LDS budget of the single-workgroup smoother (2 pressure vectors + cells)
constexpr size_t kSmallSmoothLds = 144 * 1024;
// The whole Schur smoothing of a degree-2 level that fits one workgroup's LDS
// in one launch (x from zero when x_zero); false if the level does not fit
template <typename T>
bool launch_mass_smooth_small(const LevelDesc &L, T *x, const T *b, const T *dinv, const T *ieta,
                              const MassChebSeq &seq, bool x_zero, hipStream_t st);
void init_small_smoother();
// B^T for degree 2 as a node-wise stencil (dst_u = op(B^T src_p) by mode, rhs_u
// for RESID), computed in T and stored as TO; false for degree 3
template <typename T, typename TO>
bool launch_bt_nodes(Mode mode, const LevelDesc &L, const T *src_p, const T *rhs_u, TO *dst_u, hipStream_t st);

// viscosity tables
// natural layout eta[cell*nq + q] (or eta[cell]) -> tile-blocked layout of L
template <typename T>
void launch_eta_to_tiles(const LevelDesc &L, const double *eta_nat, int per_qp, T *eta_tiles, hipStream_t st);
// cellwise averaging of natural per-qp values into tile-blocked per-cell values.
// mode: 1 arithmetic, 2 harmonic, 3 geometric (unweighted over quadrature
// points), 4 quadrature-weighted mean (DGQ0 projection)
template <typename T>
void launch_eta_average(const LevelDesc &L, const double *eta_nat_qp, int mode, T *eta_tiles_cell,
                        hipStream_t st);
// per-cell table of fine level -> mean of children on coarse level
template <typename T>
void launch_eta_restrict(const LevelDesc &Lc, const LevelDesc &Lf, const T *fine_cell, T *coarse_cell,
                         hipStream_t st);
// DGQ0 projection (quadrature-weighted mean) of a tile-blocked per-qp table
template <typename T, typename S>
void launch_eta_qp_to_cell(const LevelDesc &L, const S *eta_qp_tiles, T *eta_cell_tiles, hipStream_t st);

// DGQ1 ("project to Q1") tables: per-cell Q1 coefficients [cell][2^dim] in
// natural cell order, their children->parent L2 projection, and the values at
// the quadrature points clamped to [lo, hi] as a tile-blocked per-qp table.
void launch_eta_q1_project(const LevelDesc &L, const double *eta_nat_qp, double *coef, hipStream_t st);
void launch_eta_q1_restrict(const LevelDesc &Lc, const LevelDesc &Lf, const double *fine, double *coarse,
                            hipStream_t st);
template <typename T>
void launch_eta_q1_eval(const LevelDesc &L, const double *coef, double lo, double hi, T *out, hipStream_t st);

// vector kernels
template <typename T>
void v_fill(T *x, T v, long long n, hipStream_t st);
template <typename A, typename B>
void v_cast(const A *x, B *y, long long n, hipStream_t st);
template <typename T>
void v_axpby(T a, const T *x, T b, T *y, long long n, hipStream_t st);  // y = a x + b y
template <typename T>
void v_invert(T *x, long long n, hipStream_t st);
template <typename T>
void v_cheb_first(const T *dinv, const T *r, T *d, T *x, T inv_theta, bool x_zero, long long n,
                  hipStream_t st);  // d = dinv r / theta; x = d (x_zero) or x += d
template <typename T>
void v_cheb_step(const T *dinv, const T *r, T *d, T *x, T f1, T f2, long long n,
                 hipStream_t st);  // d = f1 d + f2 dinv r; x += d
template <typename T>
void v_scale_neg_copy(const T *x, T *y, long long n, hipStream_t st);  // y = -x

// Deterministic dot products: partial sums into work[nblk], then a fixed-order
// final reduction into out[0] (device).  work must hold dot_work_size() doubles.
int dot_work_size();
template <typename T>
void v_dot(const T *x, const T *y, long long n, double *work, double *out, hipStream_t st);
// v_sum_partials reduces m partial-dot rows of work into h (device).
void v_sum_partials(const double *work, int m, double *h, hipStream_t st);
// Krylov vectors in chunks of 8: vector i = p[i / 8] + (i % 8) * ld, i < MV_MAX
constexpr int MV_MAX = 128;
struct VecSet {
  const double *p[MV_MAX / 8];
};
// partial dots (V_i . w) for i < m into work[i * dot_work_size() + block]
void v_mdot_all(const VecSet &V, long long ld, int m, const double *w, long long n, double *work, hipStream_t st);
// delayed classical Gram-Schmidt: reduction rows s (j), t (j), v.v, v.w, w.w
// summed into rows[0 .. 2j+3) on device (work holds the per-workgroup
// partials, at most (2j+3) * dot_work_size() doubles); update v <- c1 v + V a1,
// w <- c2 w + c3 v_old + V a2 with coef = [a1, a2, c1, c2, c3] on device
void v_dcgs_reduce(const VecSet &V, long long ld, int j, const double *v, const double *w, long long n, double *work,
                   double *rows, hipStream_t st);
void v_dcgs_update(const VecSet &V, long long ld, int j, const double *coef, double *v, double *w, long long n,
                   hipStream_t st);
// w += sign * sum_{i<m} h[i] V_i (h on device)
void v_maxpy_all(const VecSet &V, long long ld, int m, const double *h, double *w, long long n, double sign,
                 hipStream_t st);
template <typename T>
void v_mul(const T *a, const T *b, T *c, long long n, hipStream_t st);  // c = a * b
// x *= a
void v_scale(double *x, double a, long long n, hipStream_t st);
// y = x / a (a on device)
void v_scaled_copy_dev(const double *x, const double *a, double *y, long long n, hipStream_t st);
// IDR(s) shadow-space vector k (uniform in [-1, 1), counter hash of (k, i))
void v_idr_shadow(double *q, int k, long long n, hipStream_t st);
// zero constrained velocity entries of a level vector
template <typename T>
void v_zero_constrained(const LevelDesc &L, T *u, hipStream_t st);
// natural [cell][q][dim*dim+1] geometry table -> the tile-blocked layout of LevelDesc::geo
void launch_geo_to_tiles(const LevelDesc &L, const double *nat, double *tiled, hipStream_t st);
long long geo_table_size(const LevelDesc &L);
// velocity diagonals (OP_ADIAG, OP_LDIAG) under a general geometry, constrained entries 1
template <typename T>
void launch_vdiag_geo(Op op, const LevelDesc &L, const T *eta, int pq, T *dst, hipStream_t st);
// Free-surface face term of the Stokes operator (active level): dst_u +=
// sign * F src_u on the faces marked in faces[6]; stab holds dim values per
// face quadrature point of the marked faces (face order, face cells and
// points lexicographic over the tangential directions); sv: work array of
// one value per face quadrature point of the marked faces.  plain_read of L
// selects reads of constrained DoFs as stored.
long long free_surface_cells(const LevelDesc &L, int face);
void launch_free_surface(const LevelDesc &L, const int faces[6], const double *stab, const double *src_u,
                         double *sv, double sign, double *dst_u, hipStream_t st);

}  // namespace mfs
