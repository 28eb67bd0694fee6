// dgx_stage.hpp -- device code of the stage kernels: the 1D tables, element
// and face geometry, the face trace records, the one-row stage body and the
// tiled stage body (dgx_tile.hpp).  Included by the translation units that
// hold stage kernels (dgx_kernels.hip, dgx_ensemble_kernels.hip).
//
// Method (equal in exact arithmetic to DGPR's CPU residual, solve.cpp:150-807):
// the reference integrates with the (p+1)^3-point Gauss rule on the same
// points it uses for the mass matrix (geometry_common.cpp:176-177, 298-299),
// so with B3 = B (x) B (x) B (equispaced-Lagrange values at the Gauss points)
// square and invertible,
//     M_e = B3^T W_e B3,   W_e = diag(w_i w_j w_k detJ),
//     M_e^-1 = B3^-1 W_e^-1 B3^-T.
// The device therefore keeps every state in the Gauss-nodal basis V = B3 U
// and computes f_g = B3 f directly:
//     f_g = s - W_e^-1 [ sum_a Dg_a^T G_a  +  face terms ]
// where G_a is the contravariant volume flux w detJ J^-1 (Q - F) and Dg the
// Gauss-node differentiation matrix.  The BR2 lifting of solve.cpp:402-480
// (DL = M_L^-1 RL, evaluated back at the face points) collapses to a
// pointwise scale at each face point:
//     dL = (1/2) fJ sum_m L_m^2 / (w_m detJ_m) * A(SL, (SL-SR) n)
// (sum over the Gauss nodes on the line normal to the face), and without the
// 1/2 and with A at the boundary state on boundary faces (solve.cpp:642-689).
//
// One 64-lane wavefront processes 64/(p+1)^3 elements (one element at p=3),
// one lane per Gauss node.  Each element evaluates its own six faces, so
// there are no atomics and no face-residual buffers; every shared face is
// evaluated once from each side (as the reference does on its processor-edge
// faces, solve.cpp:400, 530, 607).  The two traces of a face point (state
// and gradient traction) come from trace records that the previous stage
// wrote together with its output state, so no element recomputes a
// neighbour's gradient.
#pragma once

#include <hip/hip_runtime.h>

#include "dgx_physics.hpp"
#include "dgx_vjp.hpp"

namespace dgx
{

// 1D tables per order, one copy per translation unit that includes this
// header (each uploads its own, see upload_tables)
static __constant__ PTab c_ptab[3];     // [P-1]
static __constant__ QTab c_qtab[3][3];  // [P-1][Q-1]

template <int P, int Q>
struct Dims
{
  static constexpr int N   = P + 1;
  static constexpr int NB  = N * N * N;
  static constexpr int NF  = N * N;
  static constexpr int EPW = 64 / NB;
  static constexpr int NQ  = Q + 1;
  static constexpr int NG  = NQ * NQ * NQ;
  static constexpr int NFP = EPW * 6 * NF;
  static constexpr int RGN1 = (15 * NB * EPW > 20 * NFP) ? 15 * NB * EPW : 20 * NFP;
};

// node index of Gauss node (i,j,k) in an element
template <int N>
__device__ __forceinline__ int nid(int i, int j, int k)
{
  return (k * N + j) * N + i;
}

// node on the line normal to face direction nd at position m, tangential
// indices (a,b): x-faces (a=j,b=k), y-faces (a=i,b=k), z-faces (a=i,b=j).
template <int N>
__device__ __forceinline__ int line_node(int nd, int m, int a, int b)
{
  return nd == 0 ? nid<N>(m, a, b) : (nd == 1 ? nid<N>(a, m, b) : nid<N>(a, b, m));
}

// Jacobian J[d][c] = d x_d / d xi_c at a reference point described by 1D
// geometry-basis values (v*) and derivatives (d*) in each direction.
template <int Q>
__device__ __forceinline__ void jacobian(const double* X, const double* vx, const double* dx,
                                         const double* vy, const double* dy,
                                         const double* vz, const double* dz, double J[3][3])
{
  constexpr int NQ = Q + 1;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    J[d][0] = J[d][1] = J[d][2] = 0.0;
#pragma unroll
  for (int cz = 0; cz < NQ; ++cz)
#pragma unroll
    for (int cy = 0; cy < NQ; ++cy)
#pragma unroll
      for (int cx = 0; cx < NQ; ++cx)
      {
        const double* x = X + 3 * ((cz * NQ + cy) * NQ + cx);
        double w0 = dx[cx] * vy[cy] * vz[cz];
        double w1 = vx[cx] * dy[cy] * vz[cz];
        double w2 = vx[cx] * vy[cy] * dz[cz];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
          J[d][0] += x[d] * w0;
          J[d][1] += x[d] * w1;
          J[d][2] += x[d] * w2;
        }
      }
}

__device__ __forceinline__ double det3(const double J[3][3])
{
  return J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] +
         J[0][2] * J[1][0] * J[2][1] - J[0][2] * J[1][1] * J[2][0] -
         J[0][1] * J[1][0] * J[2][2] - J[0][0] * J[1][2] * J[2][1];
}

// Ji[c][d] = d xi_c / d x_d
__device__ __forceinline__ void inv3(const double J[3][3], double det, double Ji[3][3])
{
  double id = rcp(det);
  Ji[0][0] = id * (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
  Ji[0][1] = id * (J[0][2] * J[2][1] - J[0][1] * J[2][2]);
  Ji[0][2] = id * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  Ji[1][0] = id * (J[1][2] * J[2][0] - J[1][0] * J[2][2]);
  Ji[1][1] = id * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
  Ji[1][2] = id * (J[0][2] * J[1][0] - J[0][0] * J[1][2]);
  Ji[2][0] = id * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  Ji[2][1] = id * (J[0][1] * J[2][0] - J[0][0] * J[2][1]);
  Ji[2][2] = id * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
}

// Geometry-basis 1D values for the three directions of a volume node (i,j,k)
// or a face point (end e on direction nd, Gauss indices elsewhere).
__device__ __forceinline__ void geo_vals(const QTab& qt, int nd, int e, int i, int j, int k,
                                         const double** v, const double** d)
{
  int idx[3] = {i, j, k};
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    if (c == nd)
    {
      v[c] = qt.Be[e];
      d[c] = qt.Bde[e];
    }
    else
    {
      v[c] = qt.Bq[idx[c]];
      d[c] = qt.Bqd[idx[c]];
    }
  }
}

template <int Q>
__device__ __forceinline__ void jac_at(const QTab& qt, const double* X, int nd, int e, int i,
                                       int j, int k, double J[3][3])
{
  const double* v[3];
  const double* d[3];
  geo_vals(qt, nd, e, i, j, k, v, d);
  jacobian<Q>(X, v[0], d[0], v[1], d[1], v[2], d[2], J);
}


// k-th of three values, selected by value: a conditional between array
// elements would select between their addresses and keep the array in
// scratch memory.
__device__ __forceinline__ double pick3(int k, double a, double b, double c)
{
  return (k == 0) ? a : ((k == 1) ? b : c);
}

// Outward unit normal of local face (nd, hi) from the Jacobian at a face
// point: +/- (dx/dxi_c1 x dx/dxi_c2) with (c1, c2) cyclic after nd, negated
// on low faces (geometry_common.cpp:416-454).  Returns the surface element.
__device__ __forceinline__ double face_normal(const double J[3][3], int nd, int hi, double* nv)
{
  double u0 = pick3(nd, J[0][1], J[0][2], J[0][0]);
  double u1 = pick3(nd, J[1][1], J[1][2], J[1][0]);
  double u2 = pick3(nd, J[2][1], J[2][2], J[2][0]);
  double v0 = pick3(nd, J[0][2], J[0][0], J[0][1]);
  double v1 = pick3(nd, J[1][2], J[1][0], J[1][1]);
  double v2 = pick3(nd, J[2][2], J[2][0], J[2][1]);
  nv[0] = u1 * v2 - u2 * v1;
  nv[1] = u2 * v0 - u0 * v2;
  nv[2] = u0 * v1 - u1 * v0;
  double fJ = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  double sg = (hi ? 1.0 : -1.0) * rcp(fJ);
  nv[0] *= sg;
  nv[1] *= sg;
  nv[2] *= sg;
  return fJ;
}

// Constant geometry of an affine (parallelepiped) q = 1 element: J = X[e_c]
// - X[0] everywhere.  The host sets mesh.affine when every element of the
// mesh is one (bilinear and trilinear node terms zero to rounding); the
// stage and trace kernels then read J, det J and J^-1 from this per-element
// record (g[0..8] = J, g[9] = det J, g[10..18] = J^-1, and for each
// reference direction d the outward unit normal of the element's high face
// normal to d and its surface element, g[19 + 4d .. 22 + 4d]; the low face
// has the opposite normal) instead of evaluating the geometry basis at every
// node and face point.  Lanes 0..EPW-1 fill one record each; a barrier must
// follow before use.
constexpr int GREC = 31;
__device__ __forceinline__ void affine_record(const double* __restrict__ nodes, int e, int nelem,
                                              double* g)
{
  if (e >= nelem)
    return;
  const double* X = nodes + (size_t)e * 24;
  double J[3][3], Ji[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    J[d][0] = X[3 + d] - X[d];   // node (1,0,0) - node (0,0,0)
    J[d][1] = X[6 + d] - X[d];   // node (0,1,0)
    J[d][2] = X[12 + d] - X[d];  // node (0,0,1)
  }
  const double det = det3(J);
  inv3(J, det, Ji);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
      g[3 * i + j]      = J[i][j];
      g[10 + 3 * i + j] = Ji[i][j];
    }
  g[9] = det;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    g[22 + 4 * d] = face_normal(J, d, 1, g + 19 + 4 * d);
}

// J, J^-1 and det J at a point: from the affine record in the affine
// instantiation (AFF, q = 1 only), else from the geometry basis.  The two
// are separate kernels: the general one is unchanged by the affine path.
template <int Q, bool AFF>
__device__ __forceinline__ void geometry_at(const QTab& qt, const double* X, const double* g,
                                            int nd, int e, int i, int j, int k, double J[3][3],
                                            double Ji[3][3], double& det)
{
  if constexpr (AFF)
  {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
      {
        J[a][b]  = g[3 * a + b];
        Ji[a][b] = g[10 + 3 * a + b];
      }
    det = g[9];
    return;
  }
  jac_at<Q>(qt, X, nd, e, i, j, k, J);
  det = det3(J);
  inv3(J, det, Ji);
}

// J^-1, the outward unit normal nv and the surface element fJ at a point of
// local face (nd, hi): from the affine record in the affine instantiation
// (the face_normal of the element's constant J, bitwise), else from the
// geometry basis at the point.
template <int Q, bool AFF>
__device__ __forceinline__ double face_geometry(const QTab& qt, const double* X, const double* g,
                                                int nd, int hi, int i, int j, int k,
                                                double Ji[3][3], double* nv)
{
  if constexpr (AFF)
  {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        Ji[a][b] = g[10 + 3 * a + b];
    const double* n = g + 19 + 4 * nd;
    const double s  = hi ? 1.0 : -1.0;
    nv[0] = s * n[0];
    nv[1] = s * n[1];
    nv[2] = s * n[2];
    return n[3];
  }
  double J[3][3], det;
  geometry_at<Q, false>(qt, X, g, nd, hi, i, j, k, J, Ji, det);
  return face_normal(J, nd, hi, nv);
}

// Copies the 1D tables of (P, Q) into LDS: the lane-dependent table reads of
// the stage kernels then become LDS reads instead of vector loads.
template <int P, int Q>
__device__ __forceinline__ void stage_tables(PTab* sPT, QTab* sQT)
{
  const double* a = reinterpret_cast<const double*>(&c_ptab[P - 1]);
  double* b       = reinterpret_cast<double*>(sPT);
  for (int x = threadIdx.x; x < (int)(sizeof(PTab) / 8); x += 64)
    b[x] = a[x];
  const double* c = reinterpret_cast<const double*>(&c_qtab[P - 1][Q - 1]);
  double* d       = reinterpret_cast<double*>(sQT);
  for (int x = threadIdx.x; x < (int)(sizeof(QTab) / 8); x += 64)
    d[x] = c[x];
}

// Neighbour of element (i,j,k) across local face f (0:-x 1:+x 2:-y 3:+y
// 4:-z 5:+z) on the structured mesh; returns -1 with *bt set for a
// non-periodic domain boundary (geometry_shared.cpp:199-229).
__device__ __forceinline__ int neighbour(const MeshArgs& m, int i, int j, int k, int f, int* bt)
{
  const int nd = f >> 1, hi = f & 1;
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  const int c  = (nd == 0) ? i : ((nd == 1) ? j : k);
  const int n  = (nd == 0) ? nx : ((nd == 1) ? ny : nz);
  const int stride = (nd == 0) ? 1 : ((nd == 1) ? nx : nx * ny);
  int x = c + (hi ? 1 : -1);
  if (x < 0 || x >= n)
  {
    // boundary code of side f, selected by value (no dynamic indexing)
    const int b01 = hi ? m.bounds[1] : m.bounds[0];
    const int b23 = hi ? m.bounds[3] : m.bounds[2];
    const int b45 = hi ? m.bounds[5] : m.bounds[4];
    const int bc  = (nd == 0) ? b01 : ((nd == 1) ? b23 : b45);
    if (bc != -2)
    {
      *bt = bc;
      return -1;
    }
    x = hi ? 0 : n - 1;  // periodic wrap
  }
  *bt = -1;
  return ((k * ny + j) * nx + i) + (x - c) * stride;
}

// Index of the domain-boundary face on side f of element (i,j,k) among all
// boundary faces: sides in order -x +x -y +y -z +z, each numbered by its two
// tangential element coordinates.
__device__ __forceinline__ int bface_index(const MeshArgs& m, int i, int j, int k, int f)
{
  const int nd = f >> 1, hi = f & 1;
  const int sx = m.ny * m.nz, sy = m.nx * m.nz, sz = m.nx * m.ny;
  const int base = (nd == 0) ? (hi ? sx : 0)
                 : ((nd == 1) ? 2 * sx + (hi ? sy : 0) : 2 * sx + 2 * sy + (hi ? sz : 0));
  const int in = (nd == 0) ? k * m.ny + j : ((nd == 1) ? k * m.nx + i : j * m.nx + i);
  return base + in;
}

// Boundary faces of the whole mesh (all six sides).
__host__ __device__ inline size_t boundary_faces(int nx, int ny, int nz)
{
  return 2 * ((size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny);
}

// Per-element face table built once per block: sNb[l][f] is the neighbour
// across face f, or -1 - (boundary face index) on a domain boundary, whose
// code goes to sBt[l][f].
template <int EPW>
__device__ __forceinline__ void face_table(const MeshArgs& mesh, int e0, int (*sNb)[6],
                                           int (*sBt)[6])
{
  const int tid = threadIdx.x;
  if (tid < EPW * 6)
  {
    const int l = tid / 6, f = tid % 6, ee = e0 + l;
    int nb = 0, bt = -1;
    if (ee < mesh.nelem)
    {
      const int ci = ee % mesh.nx, cj = (ee / mesh.nx) % mesh.ny, ck = ee / (mesh.nx * mesh.ny);
      nb = neighbour(mesh, ci, cj, ck, f, &bt);
      if (nb < 0)
        nb = -1 - bface_index(mesh, ci, cj, ck, f);
    }
    sNb[l][f] = nb;
    sBt[l][f] = bt;
  }
}

// Faces are processed in passes of whole faces, one face point per lane:
// FPP faces of NF points fit in one 64-lane pass.
template <int P>
struct FaceDims
{
  static constexpr int N     = P + 1;
  static constexpr int NF    = N * N;
  static constexpr int NB    = N * N * N;
  static constexpr int EPW   = 64 / NB;
  static constexpr int FPP   = 64 / NF;
  static constexpr int NFACE = EPW * 6;
  static constexpr int NPASS = (NFACE + FPP - 1) / FPP;
  static constexpr int FS    = 17;  // doubles per face point in the face pass
  static constexpr int RG    = (5 * NB * EPW > FS * 64) ? 5 * NB * EPW : FS * 64;
};

// Face trace records of the EPW elements whose Gauss-nodal states are in sV:
// for each local face f and face point, the trace U (5) and the gradient
// traction tau(grad u).n (3) taken with the element's own outward unit
// normal n, stored as Tr[e][f][8][(p+1)^2].  The face point (a, b) indexes
// the Gauss points of the two tangential directions in increasing order, so
// the point of face f of e and of face f^1 of its neighbour coincide
// (basis.cpp:180-192).  On a domain-boundary face the boundary state Ub
// (bound_states.cpp:24-130) of the trace is stored as well, after the trace
// records of all elements: Tb[boundary face][5][(p+1)^2].  TT is the stored
// type: double, or float for the mixed-precision form (the records are
// computed in double and rounded once on the store).  Called by every
// thread of the block; sR needs 4*64 doubles.
template <int P, int Q, typename TT, bool AFF = false>
__device__ __forceinline__ void produce_traces(const MeshArgs& mesh, const Phys& ph,
                                               const PTab& pt, const QTab& qt,
                                               const double (*sV)[5][FaceDims<P>::NB],
                                               const double (*sX)[(Q + 1) * (Q + 1) * (Q + 1) * 3],
                                               const double (*sG)[GREC],
                                               const int (*sNb)[6], const int (*sBt)[6],
                                               double* sR, int e0, TT* __restrict__ Tout,
                                               int* status)
{
  using FD = FaceDims<P>;
  constexpr int N = FD::N, NF = FD::NF;
  TT* Tb = Tout + (size_t)mesh.nelem * 6 * 8 * NF;
  const int tid = threadIdx.x;
  bool ok = true;
  for (int pass = 0; pass < FD::NPASS; ++pass)
  {
    const int slot = tid / NF;
    const int gf   = pass * FD::FPP + slot;
    const int fpt  = tid % NF;
    const int a = fpt % N, b = fpt / N;
    const int l = gf / 6, f = gf % 6;
    const int ee = e0 + l;
    const bool fact = (slot < FD::FPP) && (gf < FD::NFACE) && (ee < mesh.nelem);
    const int nd = f >> 1, hi = f & 1;
    const int fbase = slot * NF;

    double U[5], dn[4];
    if (fact)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double v = 0.0, dv = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          double x = sV[l][r][line_node<N>(nd, m, a, b)];
          v += pt.L[hi][m] * x;
          if (r < 4)
            dv += pt.Ld[hi][m] * x;
        }
        U[r] = v;
        if (r < 4)
        {
          dn[r] = dv;
          sR[tid * 4 + r] = v;
        }
      }
    }
    __syncthreads();
    if (fact)
    {
      const int gi0 = (nd == 0) ? 0 : a;
      const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
      const int gi2 = (nd == 2) ? 0 : b;
      double Ji[3][3], nv[3];
      face_geometry<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, Ji, nv);
      double gr[3][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          d1 += pt.Dg[a][m] * sR[(fbase + m + N * b) * 4 + r];
          d2 += pt.Dg[b][m] * sR[(fbase + a + N * m) * 4 + r];
        }
        const double dr = dn[r];
        gr[0][r] = pick3(nd, dr, d1, d1);
        gr[1][r] = pick3(nd, d1, dr, d2);
        gr[2][r] = pick3(nd, d2, d2, dr);
      }
      double vg[3];
      grad_traction(gr, U, Ji, nv, ph.mu, vg);
      TT* T = Tout + ((size_t)ee * 6 + f) * 8 * NF + fpt;
#pragma unroll
      for (int r = 0; r < 5; ++r)
        T[r * NF] = (TT)U[r];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        T[(5 + i) * NF] = (TT)vg[i];
      const int nb = sNb[l][f];
      if (nb < 0)
      {
        double Ub[5];
        ok = boundary_state(sBt[l][f], U, nv, ph, Ub) && ok;
        TT* B = Tb + (size_t)(-1 - nb) * 5 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          B[r * NF] = (TT)Ub[r];
      }
    }
    __syncthreads();
  }
  if (!ok)
    atomicOr(status, 1);
}

// Trace records of a Gauss-nodal state V (the first stage after the state
// was set from outside, and stand-alone residual evaluations).
template <int P, int Q, typename TT, bool AFF>
__global__ __launch_bounds__(64) void dgx_trace_kernel(MeshArgs mesh, Phys ph,
                                                       const double* __restrict__ V,
                                                       TT* __restrict__ Tout, int* status)
{
  using FD = FaceDims<P>;
  constexpr int NB = FD::NB, EPW = FD::EPW;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);
  __shared__ double sV[EPW][5][NB];
  __shared__ double sX[EPW][NG * 3];
  __shared__ double sG[EPW][GREC];
  __shared__ double sR[4 * 64];
  __shared__ PTab sPT;
  __shared__ QTab sQT;
  __shared__ int sNb[EPW][6], sBt[EPW][6];
  stage_tables<P, Q>(&sPT, &sQT);
  const int tid = threadIdx.x;
  const int le  = tid / NB;
  const int nd0 = tid % NB;
  const int e0  = blockIdx.x * EPW;
  face_table<EPW>(mesh, e0, sNb, sBt);
  if (le < EPW && e0 + le < mesh.nelem)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = V[((size_t)(e0 + le) * 5 + r) * NB + nd0];
  }
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int ee = e0 + x / (NG * 3);
    sX[x / (NG * 3)][x % (NG * 3)] =
        (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  if (AFF && tid < EPW)
    affine_record(mesh.nodes, e0 + tid, mesh.nelem, sG[tid]);
  __syncthreads();
  produce_traces<P, Q, TT, AFF>(mesh, ph, sPT, sQT, sV, sX, sG, sNb, sBt, sR, e0, Tout, status);
}

// LDS of one stage-body block, carved from a byte buffer so that kernels
// holding the bodies of several orders (the ensemble stage kernel) give
// them one shared region: state [EPW][5][NB], own geometry nodes
// [EPW][NG*3], affine-element records [EPW][GREC], the exchange region
// (volume flux of one direction, then one face pass), the 1D tables and the
// face table.
template <int P, int Q>
struct StageLds
{
  using FD = FaceDims<P>;
  static constexpr int EPW = FD::EPW, NB = FD::NB, NG = (Q + 1) * (Q + 1) * (Q + 1);
  static constexpr size_t oV  = 0;
  static constexpr size_t oX  = oV + sizeof(double) * EPW * 5 * NB;
  static constexpr size_t oG  = oX + sizeof(double) * EPW * NG * 3;
  static constexpr size_t oR  = oG + sizeof(double) * EPW * GREC;
  static constexpr size_t oPT = oR + sizeof(double) * FD::RG;
  static constexpr size_t oQT = oPT + sizeof(PTab);
  static constexpr size_t oNb = oQT + sizeof(QTab);
  static constexpr size_t oBt = oNb + sizeof(int) * EPW * 6;
  static constexpr size_t bytes = oBt + sizeof(int) * EPW * 6;
};

// One full right-hand-side evaluation f_g = B3 f of EPW elements fused with
// the RK stage combination of StageArgs, reading the face trace records Tin
// of Vin and (optionally) writing those of Vout to Tout.  lam[e][f][pt]
// holds the BR2 lifting coefficient of element e's side of its face f.
// Body of the stage kernel for block `block` of one mesh, in the LDS region
// lds (StageLds<P, Q>::bytes, 16-byte aligned): shared by the single-mesh
// kernel and the ensemble kernels (batched, and the one-row units of the
// one-launch layouts).  INV: the inviscid right-hand
// side only (Euler volume flux, Roe / boundary Euler face flux; no
// gradients, tractions, BR2 or dual-consistency terms), which equals the
// full one at mu = 0 (whose viscous terms are exact zeros) without reading
// the tractions of the records: the adjoint's d f / d mu.
template <int P, int Q, typename TT, bool AFF = false, bool INV = false>
__device__ __forceinline__ void stage_body_lds(const MeshArgs& mesh, const Phys& ph,
                                               const StageArgs& sa,
                                               const double* __restrict__ lam, int* status,
                                               int block, unsigned char* lds)
{
  using FD = FaceDims<P>;
  using SL = StageLds<P, Q>;
  constexpr int N = FD::N, NB = FD::NB, NF = FD::NF, EPW = FD::EPW, FS = FD::FS;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);

  auto sV  = reinterpret_cast<double(*)[5][NB]>(lds + SL::oV);
  auto sX  = reinterpret_cast<double(*)[NG * 3]>(lds + SL::oX);      // own geometry nodes
  auto sG  = reinterpret_cast<double(*)[GREC]>(lds + SL::oG);        // affine records
  double* sR = reinterpret_cast<double*>(lds + SL::oR);  // volume flux, then one face pass
  PTab& sPT = *reinterpret_cast<PTab*>(lds + SL::oPT);
  QTab& sQT = *reinterpret_cast<QTab*>(lds + SL::oQT);
  auto sNb = reinterpret_cast<int(*)[6]>(lds + SL::oNb);
  auto sBt = reinterpret_cast<int(*)[6]>(lds + SL::oBt);
  stage_tables<P, Q>(&sPT, &sQT);
  const PTab& pt = sPT;
  const QTab& qt = sQT;

  const int tid = threadIdx.x;
  const int le  = tid / NB;
  const int nd0 = tid % NB;
  const int e0  = block * EPW;
  const int e   = e0 + le;
  const TT* __restrict__ Tin = reinterpret_cast<const TT*>(sa.Tin);
  const bool vact = (le < EPW) && (e < mesh.nelem);
  face_table<EPW>(mesh, e0, sNb, sBt);

  // ---- load state and element geometry ----
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = sa.Vin[((size_t)e * 5 + r) * NB + nd0];
  }
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int l  = x / (NG * 3);
    int ee = e0 + l;
    sX[l][x % (NG * 3)] = (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  if (AFF && tid < EPW)
    affine_record(mesh.nodes, e0 + tid, mesh.nelem, sG[tid]);
  __syncthreads();

  const int idx[3] = {nd0 % N, (nd0 / N) % N, nd0 / (N * N)};
  double wJ = 1.0;
  double acc[5] = {0, 0, 0, 0, 0};
  double Ji[3][3];
  double H[3][5];  // physical flux Q - F at this node

  // ---- volume: velocity gradient, fluxes ----
  if (vact)
  {
    double J[3][3], det;
    geometry_at<Q, AFF>(qt, sX[le], sG[le], -1, 0, idx[0], idx[1], idx[2], J, Ji, det);
    wJ = pt.w[idx[0]] * pt.w[idx[1]] * pt.w[idx[2]] * det;
    double Vq[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
      Vq[r] = sV[le][r][nd0];
    if constexpr (INV)
      euler_volume_flux(Vq, ph.gamma, H);
    else
    {
      double gr[3][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          g0 += pt.Dg[idx[0]][m] * sV[le][r][nid<N>(m, idx[1], idx[2])];
          g1 += pt.Dg[idx[1]][m] * sV[le][r][nid<N>(idx[0], m, idx[2])];
          g2 += pt.Dg[idx[2]][m] * sV[le][r][nid<N>(idx[0], idx[1], m)];
        }
        gr[0][r] = g0;
        gr[1][r] = g1;
        gr[2][r] = g2;
      }
      volume_flux(Vq, gr, Ji, ph.gamma, ph.mu, H);
    }
  }

  // ---- volume divergence, one reference direction at a time ----
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    if (vact)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
        sR[(le * 5 + r) * NB + nd0] = wJ * (Ji[c][0] * H[0][r] + Ji[c][1] * H[1][r] + Ji[c][2] * H[2][r]);
    }
    __syncthreads();
    if (vact)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          int id2[3] = {idx[0], idx[1], idx[2]};
          id2[c] = m;
          s += pt.Dg[m][idx[c]] * sR[(le * 5 + r) * NB + nid<N>(id2[0], id2[1], id2[2])];
        }
        acc[r] += s;
      }
    }
    __syncthreads();
  }

  // ---- faces, in passes of FPP whole faces ----
  // Both traces of a face point come from the trace records Tin written with
  // the input state (own face f, and the neighbour's face f^1 at the same
  // face point).  Per face point in sR (stride FS = 17): G (0-4), normal
  // contravariant dcQ (5-8), tangential (9-12, 13-16); the T phase then
  // replaces G by T = G - Dg^T H_t1 - Dg^T H_t2.
  bool ok = true;
  for (int pass = 0; pass < FD::NPASS; ++pass)
  {
    const int slot = tid / NF;
    const int gf   = pass * FD::FPP + slot;
    const int fpt  = tid % NF;
    const int a = fpt % N, b = fpt / N;
    const int l = gf / 6, f = gf % 6;
    const int ee = e0 + l;
    const bool fact = (slot < FD::FPP) && (gf < FD::NFACE) && (ee < mesh.nelem);
    const int nd = f >> 1, hi = f & 1;
    const int fbase = slot * NF;

    FacePoint fpo;
    double hn[4], h1[4], h2[4];
    if (INV && fact)
    {
      // Roe flux (interior) or the Euler flux of the boundary state; no
      // viscous face terms: the contravariant dual-consistency flux is zero
      const int en = sNb[l][f];
      const TT* To = Tin + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      double UL[5], Ub[5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
        UL[r] = To[r * NF];
      const TT* Tn = (en >= 0) ? Tin + ((size_t)en * 6 + (f ^ 1)) * 8 * NF + fpt
                               : Tin + (size_t)mesh.nelem * 6 * 8 * NF +
                                     (size_t)(-1 - en) * 5 * NF + fpt;
#pragma unroll
      for (int r = 0; r < 5; ++r)
        Ub[r] = Tn[r * NF];
      const int gi0 = (nd == 0) ? 0 : a;
      const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
      const int gi2 = (nd == 2) ? 0 : b;
      double JiL[3][3], nv[3];
      const double fJ = face_geometry<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, JiL, nv);
      if (en >= 0)
        roe_flux(UL, Ub, nv, ph.gamma, fpo.G);
      else
        euler_flux_n(Ub, ph.gamma, nv, fpo.G);
      const double sc = fJ * pt.w[a] * pt.w[b];
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        fpo.G[r] *= sc;
        sR[tid * FS + r] = fpo.G[r];
      }
#pragma unroll
      for (int r = 0; r < 12; ++r)
        sR[tid * FS + 5 + r] = 0.0;
    }
    else if (fact)
    {
      const int en = sNb[l][f];
      const TT* To = Tin + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      double UL[5], vgL[3];
#pragma unroll
      for (int r = 0; r < 5; ++r)
        UL[r] = To[r * NF];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        vgL[i] = To[(5 + i) * NF];

      // own geometry at the face point (end hi on direction nd)
      const int gi0 = (nd == 0) ? 0 : a;
      const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
      const int gi2 = (nd == 2) ? 0 : b;
      double JiL[3][3], nv[3];
      const double fJ = face_geometry<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, JiL, nv);
      // affine meshes: the face's constant coefficient
      const double lamL =
          AFF ? mesh.lamf[(size_t)ee * 6 + f] : lam[((size_t)ee * 6 + f) * NF + fpt];

      // neighbour trace (its traction was taken with its own outward normal
      // -n), or the boundary state stored with the traces
      double Ub[5], vgR[3] = {0.0, 0.0, 0.0};
      double lamR = 0.0;
      if (en >= 0)
      {
        const TT* Tn = Tin + ((size_t)en * 6 + (f ^ 1)) * 8 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          Ub[r] = Tn[r * NF];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          vgR[i] = -Tn[(5 + i) * NF];
        lamR = AFF ? mesh.lamf[(size_t)en * 6 + (f ^ 1)]
                   : lam[((size_t)en * 6 + (f ^ 1)) * NF + fpt];
      }
      else
      {
        const TT* B = Tin + (size_t)mesh.nelem * 6 * 8 * NF + (size_t)(-1 - en) * 5 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          Ub[r] = B[r * NF];
      }
      face_point_physics(en >= 0, UL, Ub, vgL, vgR, nv, lamL, lamR, ph, fpo);

      // contravariant dual-consistency flux h_c = Ji[c] . dcQ (rows 1-4)
      const double irL = rcp(UL[0]);
      const double uL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
      const double sc = fJ * pt.w[a] * pt.w[b];
      const double* S = fpo.S;
      const double Sm[3][3] = {{S[0], S[3], S[4]}, {S[3], S[1], S[5]}, {S[4], S[5], S[2]}};
      double hc[3][4];
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
#pragma unroll
        for (int i = 0; i < 3; ++i)
          hc[c][i] = sc * (JiL[c][0] * Sm[0][i] + JiL[c][1] * Sm[1][i] + JiL[c][2] * Sm[2][i]);
        hc[c][3] = uL[0] * hc[c][0] + uL[1] * hc[c][1] + uL[2] * hc[c][2];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        hn[r] = pick3(nd, hc[0][r], hc[1][r], hc[2][r]);
        h1[r] = pick3(nd, hc[1][r], hc[0][r], hc[0][r]);
        h2[r] = pick3(nd, hc[2][r], hc[2][r], hc[1][r]);
      }
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        fpo.G[r] *= sc;
        sR[tid * FS + r] = fpo.G[r];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        sR[tid * FS + 5 + r]  = hn[r];
        sR[tid * FS + 9 + r]  = h1[r];
        sR[tid * FS + 13 + r] = h2[r];
      }
    }
    __syncthreads();
    // T = G - Dg^T(t1) H_t1 - Dg^T(t2) H_t2 (mass row has no viscous part);
    // only slots 9-16 of other points are read, so T overwrites 0-4
    if (fact)
    {
#pragma unroll
      for (int r = 1; r < 5; ++r)
      {
        double t = fpo.G[r];
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          t -= pt.Dg[m][a] * sR[(fbase + m + N * b) * FS + 8 + r];
          t -= pt.Dg[m][b] * sR[(fbase + a + N * m) * FS + 12 + r];
        }
        sR[tid * FS + r] = t;
      }
    }
    __syncthreads();
    // gather this pass's faces at the nodes
    if (vact)
    {
#pragma unroll
      for (int ff = 0; ff < 6; ++ff)
      {
        int g = le * 6 + ff;
        if (g / FD::FPP != pass)
          continue;
        int fnd = ff >> 1, fhi = ff & 1;
        int ft1 = (fnd == 0) ? 1 : 0;
        int ft2 = (fnd == 2) ? 1 : 2;
        int m  = idx[fnd];
        int fp = (g % FD::FPP) * NF + idx[ft2] * N + idx[ft1];
        acc[0] += pt.L[fhi][m] * sR[fp * FS];
#pragma unroll
        for (int r = 1; r < 5; ++r)
          acc[r] += pt.L[fhi][m] * sR[fp * FS + r] - pt.Ld[fhi][m] * sR[fp * FS + 4 + r];
      }
    }
    __syncthreads();
  }

  if (!ok)
    atomicOr(status, 1);

  // ---- form f_g and the RK epilogue ----
  if (vact)
  {
    double iwJ = rcp(wJ);
    size_t base = (size_t)e * 5 * NB + nd0;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      size_t ix = base + (size_t)r * NB;
      double fv = ph.source[r] - acc[r] * iwJ;
      if (sa.fout)
        sa.fout[ix] = fv;
      double ft = fv;
      if (sa.accmode == 1)
        sa.acc[ix] = fv;
      else if (sa.accmode == 2)
        sa.acc[ix] = sa.acc[ix] + 2.0 * fv;
      else if (sa.accmode == 3)
        ft = (sa.acc[ix] + fv) * (1.0 / 6.0);
      if (sa.Vout)
      {
        double o = sa.cf * ft;
        if (sa.ca != 0.0)
          o += sa.ca * sa.Vbase[ix];
        if (sa.cb != 0.0)
          o += sa.cb * sV[le][r][nd0];
        sa.Vout[ix] = o;
        sV[le][r][nd0] = o;  // only this lane reads this node from here on
      }
    }
  }

  // ---- face traces of the new state for the next stage ----
  if (sa.Tout)
  {
    __syncthreads();
    produce_traces<P, Q, TT, AFF>(mesh, ph, pt, qt, sV, sX, sG, sNb, sBt, sR, e0,
                                  reinterpret_cast<TT*>(sa.Tout), status);
  }
}

template <int P, int Q, typename TT, bool AFF = false>
__device__ __forceinline__ void stage_body(const MeshArgs& mesh, const Phys& ph,
                                           const StageArgs& sa, const double* __restrict__ lam,
                                           int* status, int block)
{
  __shared__ __attribute__((aligned(16))) unsigned char lds[StageLds<P, Q>::bytes];
  stage_body_lds<P, Q, TT, AFF>(mesh, ph, sa, lam, status, block, lds);
}

#include "dgx_tile.hpp"

}  // namespace dgx
