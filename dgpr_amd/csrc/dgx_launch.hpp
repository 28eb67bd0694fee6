// dgx_launch.hpp -- plain structs and launch entry points shared by the host
// (C ABI) and device translation units of libdgx.
#pragma once

#include <hip/hip_runtime_api.h>

namespace dgx
{

constexpr int MAXN = 4;  // (p+1) <= 4 and (q+1) <= 4

// 1D tables of solution order P on the (P+1)-point Gauss rule of [0,1].
//   g, w       Gauss points / weights
//   B[i][m]    equispaced Lagrange l_m(g_i)          (reference basis)
//   Binv       inverse of B
//   Dg[i][m]   d/dx of the Gauss-node Lagrange l^g_m at g_i
//   L[e][m]    l^g_m(e), e = 0 (x=0) or 1 (x=1)
//   Ld[e][m]   d/dx l^g_m at e
//   Blo, Binvlo  low parts of B and Binv as double-double pairs (hi + lo),
//              used by the state basis conversions
struct PTab
{
  double g[MAXN], w[MAXN];
  double B[MAXN][MAXN], Binv[MAXN][MAXN], Dg[MAXN][MAXN];
  double L[2][MAXN], Ld[2][MAXN];
  double Blo[MAXN][MAXN], Binvlo[MAXN][MAXN];
};

// Geometry-order Q equispaced Lagrange values/derivatives at the Gauss points
// of order P and at the two ends of [0,1].
struct QTab
{
  double Bq[MAXN][MAXN], Bqd[MAXN][MAXN];
  double Be[2][MAXN], Bde[2][MAXN];
};

// Tables of one (p, q) pair for the general-order kernels (dgx_general.hip),
// which cover every order the reference accepts: 0 <= p <= 8, 1 <= q <= 8.
// Same objects as PTab and QTab (Bq[i][c], Bqd[i][c]: geometry-order
// equispaced Lagrange value / derivative of node c at Gauss point i; Be, Bde
// at the two ends of [0,1]).
constexpr int GMAXN = 9;  // (p+1) <= 9 and (q+1) <= 9
struct GenTab
{
  double g[GMAXN], w[GMAXN];
  double B[GMAXN][GMAXN], Binv[GMAXN][GMAXN], Dg[GMAXN][GMAXN];
  double L[2][GMAXN], Ld[2][GMAXN];
  double Blo[GMAXN][GMAXN], Binvlo[GMAXN][GMAXN];
  double Bq[GMAXN][GMAXN], Bqd[GMAXN][GMAXN];
  double Be[2][GMAXN], Bde[2][GMAXN];
};

// Physical constants of one evaluation (parameters + derived boundary data).
struct Phys
{
  double gamma, mu, eta, Rgas, P, Pt, Tt;
  double Ufr[5];
  double source[5];
};

// Structured mesh: nodes [e][(q+1)^3][3], element e = (k*ny + j)*nx + i,
// boundary codes per side -x +x -y +y -z +z.
struct MeshArgs
{
  const double* nodes;
  int nx, ny, nz;
  int nelem;
  int bounds[6];
  int affine;  // every element is a q = 1 parallelepiped (constant Jacobian)
  // affine meshes: the BR2 lifting coefficient of each element face [e][6],
  // constant over a parallelepiped's face (fJ / det J times a table sum);
  // null otherwise
  const double* lamf;
};

// Epilogue of one right-hand-side evaluation:
//   Vout = ca*Vbase + cb*Vin + cf*fterm
// with fterm = f for accmode 0..2 and (acc + f)/6 for accmode 3;
// accmode 1 stores acc = f, accmode 2 updates acc += 2 f.  fout (optional)
// receives f.  All states are Gauss-nodal [e][5][(p+1)^3].  Tin holds the
// face trace records of Vin; Tout (optional, only with Vout) receives those
// of Vout.  A trace record set is [e][6 faces][8][(p+1)^2]: the trace (5)
// and the gradient traction tau(grad u).n with the own outward normal (3).
// With trace32 set the records are floats (the mixed-precision form; the
// buffers keep their double-sized allocation and use its first half).
struct StageArgs
{
  const double* Vin;
  const double* Vbase;
  double* acc;
  double* Vout;
  double* fout;
  const double* Tin;
  double* Tout;
  double ca, cb, cf;
  int accmode;
  int trace32;
  int ctype;  // volume contractions of the tiled p = 3 kernel: 0 FP64, 1 FP32
              // matrix cores, 2 split-BF16 matrix cores (with trace32 set)
  // tiled kernel only (tile 1): when set, normp[block] receives the block's
  // sum of R^2, R = -B3^T W f (the residual norm of this stage's f), for
  // tile_blocks(p, nelem) blocks
  double* normp = nullptr;
};

// One member of a batched stage launch: its mesh, physics, stage arguments
// and lifting coefficients, the first grid block (or work unit) of its
// elements, and its order (read by the ensemble stage kernel, which can hold
// the bodies of every order).
struct BatchItem
{
  MeshArgs mesh;
  Phys ph;
  StageArgs sa;
  const double* lam;
  int block0;
  int p;
};

// waves: occupancy target of the compiled stage kernel (1, 2 or 3 waves per
// SIMD), chosen by default_stage_waves or by timing (dgx_autotune)
// tile 0: the one-wavefront-per-row kernel (faces of all directions mixed in
// each pass); tile 1: the tiled kernel (rows per tile fixed per order, faces
// one direction at a time), compiled for 2 and 3 waves per SIMD
hipError_t rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                        const double* lam, int* status, int waves, hipStream_t st, int tile = 0);
bool tile_valid(int p, int q, int tile, int waves);
// one stage of several meshes of the same (p, q): items_dev (device memory,
// sorted by block0) covers nblocks blocks in total; affine: every member is
// an affine q = 1 mesh (MeshArgs::affine)
hipError_t rhs_batch_dispatch(int p, int q, const BatchItem* items_dev, int nitems, int nblocks,
                              int* status, int waves, bool affine, hipStream_t st);
// grid blocks of one mesh in a stage launch
int stage_blocks(int p, int nelem);
// grid blocks of the tiled kernel (tile 1) on one mesh
int tile_blocks(int p, int nelem);
// fout = f(V; mu = 0), the inviscid right-hand side, from V and its trace
// records Tin (written with any mu: their tractions are not read)
hipError_t inviscid_rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                                 const double* Tin, double* fout, const double* lam, int* status,
                                 hipStream_t st);

// ---- ensemble stage kernel of the one-launch layouts (dgx_ensemble_kernels.hip)
// One launch covers the units (rows or tiles of elements) of every member of
// a q = 1 group: block b runs unit b.  items_dev sorted by first unit
// (block0), which ens_units counts per member.  tile 0: units of one row
// (the one-row body); tile 1: units of the tiled body.  affine: every member
// affine.  psel 0: members of any order (one kernel with every order's
// body); psel 1..3: members of that order only.
int ens_units(int p, int nelem, int tile);
hipError_t ens_stage_dispatch(const BatchItem* items_dev, int nitems, int nunits, int* status,
                              int waves, int tile, bool affine, int psel, hipStream_t st);
hipError_t upload_tables_ensemble(const PTab* pt, const QTab* qt);
// face trace records Tout of the Gauss-nodal state V (and the boundary
// states of its domain-boundary faces, see trace_doubles)
hipError_t trace_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                          double* Tout, int* status, hipStream_t st, bool trace32 = false);
// doubles of one trace record set: nelem*6*8*(p+1)^2 trace records followed
// by 5*(p+1)^2 boundary-state values for each of the mesh's boundary faces
size_t trace_doubles(int p, int nx, int ny, int nz);
// lam[e][f][(p+1)^2]: BR2 lifting coefficient of each element side of each face
hipError_t lift_coef_dispatch(int p, int q, const MeshArgs& m, double* lam, hipStream_t st);
// lamf[e][f] = lam[e][f][0] (affine meshes: the face's constant value)
hipError_t lift_face_dispatch(int p, const double* lam, int nelem, double* lamf, hipStream_t st);
hipError_t resid_out_dispatch(int p, int q, const MeshArgs& m, const double* fg, double* f,
                              double* R, double* partial, hipStream_t st);
hipError_t basis_dispatch(int p, const double* in, double* out, int nelem, int mode,
                          hipStream_t st);
// Sums of n per-workgroup partials (deterministic order).  Above
// SUM_ONE_BLOCK partials the sum runs in two launches (chunks of SUM_CHUNK,
// then the chunk sums): the partial buffer then holds sum_scratch(n) more
// doubles after the n partials.
constexpr int SUM_ONE_BLOCK = 4096, SUM_CHUNK = 2048;
size_t sum_scratch(int n);
hipError_t sum_dispatch(const double* partial, int n, double* out, hipStream_t st);
// *out = sqrt(inv_count * sum of the n partial sums): one step's residual norm
hipError_t norm_dispatch(const double* partial, int n, double inv_count, double* out,
                         hipStream_t st);
// psiV = (d f_g / d V)^T psi on a periodic mesh, V with its trace records Tin;
// pvol, aown, anbr are scratch (a state and two trace-record sets)
hipError_t adjoint_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                            const double* Tin, const double* lam, const double* psi,
                            double* pvol, double* aown, double* anbr, double* psiV,
                            hipStream_t st);
// kinetic energy (per-block partials in partial[resid_out_blocks]) and its
// gradient with respect to the Gauss-nodal state
hipError_t kinetic_energy_dispatch(int p, int q, const MeshArgs& m, const double* V,
                                   double* grad, double* partial, hipStream_t st);
// one pass of the reverse RK stage algebra: sum_out = sum_in + a and
// lk = wa a + wl l (a null: lk = wl l); null outputs skipped
hipError_t adj_tail_dispatch(const double* l, const double* a, const double* sum_in,
                             double* sum_out, double* lk, double wl, double wa, size_t n,
                             hipStream_t st);
// *out = scale * x . (y - z) (z may be NULL: scale * x . y); nblocks
// partial sums in partial (nblocks <= SUM_ONE_BLOCK)
hipError_t dot_dispatch(const double* x, const double* y, const double* z, double scale,
                        size_t n, double* partial, int nblocks, double* out, hipStream_t st);
// derived outputs of a Gauss-nodal state on the (Ns+1)^3 sub-grid of each element
hipError_t sample_dispatch(int p, const double* V, int nelem, int Ns, int var, double gamma,
                           double* out, hipStream_t st);
// running mean update of n doubles
hipError_t mean_dispatch(const double* U, double* mean, size_t n, double old_range,
                         double new_range, hipStream_t st);
hipError_t upload_tables(const PTab* pt, const QTab* qt);
hipError_t copy_dispatch(const void* src, void* dst, size_t bytes, hipStream_t st);
int resid_out_blocks(int p, int q, int nelem);
// per-workgroup sums of R^2 (R the modal residual of f_g) into
// partial[0 .. norm_blocks(p, q, nelem)), for the per-step residual norm
int norm_blocks(int p, int q, int nelem);
hipError_t norm_partials_dispatch(int p, int q, const MeshArgs& m, const double* fg,
                                  double* partial, hipStream_t st);

// ---- general-order kernels (dgx_general.hip) ----
// True when (p, q) is served by the general-order kernels only: p = 0,
// p >= 4 or q >= 4 (the tuned kernels cover 1 <= p, q <= 3).
inline bool general_order(int p, int q)
{
  return p < 1 || p > 3 || q > 3;
}
// tables of every (p, q), 0 <= p <= 8, 1 <= q <= 8: gt[p * 8 + q - 1]
hipError_t upload_gen_tables(const GenTab* gt);
// elements per workgroup of the general-order kernels for (p, q) (several
// at p <= 2, as many as 32 KB of their geometry nodes allow)
int gen_elements_per_block(int p, int q);
// the entry points below mirror the tuned ones for the general orders
// (one Gauss node per lane; gen_elements_per_block elements per workgroup)
hipError_t gen_rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                            const double* lam, int* status, hipStream_t st);
hipError_t gen_trace_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                              double* Tout, int* status, hipStream_t st, bool trace32);
hipError_t gen_lift_coef_dispatch(int p, int q, const MeshArgs& m, double* lam, hipStream_t st);
hipError_t gen_resid_out_dispatch(int p, int q, const MeshArgs& m, const double* fg, double* f,
                                  double* R, double* partial, hipStream_t st);
hipError_t gen_basis_dispatch(int p, const double* in, double* out, int nelem, int mode,
                              hipStream_t st);
hipError_t gen_sample_dispatch(int p, const double* V, int nelem, int Ns, int var, double gamma,
                               double* out, hipStream_t st);

namespace host
{
void build_tables(PTab* pt, QTab* qt);
// GenTab of every (p, q), 0 <= p <= 8, 1 <= q <= 8, at gt[p * 8 + q - 1]
void build_gen_tables(GenTab* gt);
}

}  // namespace dgx
