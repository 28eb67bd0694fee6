// dgx_kernels.hip -- gfx950 kernels of the FP64 DG residual + RK stage.
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
#include <hip/hip_runtime.h>

#include "dgx_physics.hpp"
#include "dgx_vjp.hpp"

namespace dgx
{

__constant__ PTab c_ptab[3];     // [P-1]
__constant__ QTab c_qtab[3][3];  // [P-1][Q-1]

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


// Constant geometry of an affine (parallelepiped) q = 1 element: J = X[e_c]
// - X[0] everywhere.  The host sets mesh.affine when every element of the
// mesh is one (bilinear and trilinear node terms zero to rounding); the
// stage and trace kernels then read J, det J and J^-1 from this per-element
// record (g[0..8] = J, g[9] = det J, g[10..18] = J^-1) instead of evaluating
// the geometry basis at every node and face point.  Lanes 0..EPW-1 fill one
// record each; a barrier must follow before use.
constexpr int GREC = 19;
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
      double J[3][3], Ji[3][3], det;
      geometry_at<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, J, Ji, det);
      double nv[3];
      face_normal(J, nd, hi, nv);
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

// One full right-hand-side evaluation f_g = B3 f of EPW elements fused with
// the RK stage combination of StageArgs, reading the face trace records Tin
// of Vin and (optionally) writing those of Vout to Tout.  lam[e][f][pt]
// holds the BR2 lifting coefficient of element e's side of its face f.
// Body of the stage kernel for block `block` of one mesh: shared by the
// single-mesh kernel and the batched ensemble kernel.
template <int P, int Q, typename TT, bool AFF = false>
__device__ __forceinline__ void stage_body(const MeshArgs& mesh, const Phys& ph,
                                           const StageArgs& sa, const double* __restrict__ lam,
                                           int* status, int block)
{
  using FD = FaceDims<P>;
  constexpr int N = FD::N, NB = FD::NB, NF = FD::NF, EPW = FD::EPW, FS = FD::FS;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);

  __shared__ double sV[EPW][5][NB];
  __shared__ double sX[EPW][NG * 3];      // own geometry nodes
  __shared__ double sG[EPW][GREC];        // affine-element geometry records
  __shared__ double sR[FD::RG];  // volume flux of one direction, then one face pass
  __shared__ PTab sPT;
  __shared__ QTab sQT;
  __shared__ int sNb[EPW][6], sBt[EPW][6];
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
    double Vq[5], gr[3][4];
#pragma unroll
    for (int r = 0; r < 5; ++r)
      Vq[r] = sV[le][r][nd0];
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
    if (fact)
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
      double J[3][3], JiL[3][3], detL;
      geometry_at<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, J, JiL, detL);
      double nv[3];
      const double fJ = face_normal(J, nd, hi, nv);
      const double lamL = lam[((size_t)ee * 6 + f) * NF + fpt];

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
        lamR = lam[((size_t)en * 6 + (f ^ 1)) * NF + fpt];
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

// W is the occupancy target (waves per SIMD) handed to the register
// allocator; the launch picks one of the compiled targets per (P, Q).  TT is
// the type of the trace records read and written (float: mixed precision).
template <int P, int Q, int W, typename TT, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_rhs_stage_kernel(MeshArgs mesh, Phys ph, StageArgs sa, const double* __restrict__ lam,
                          int* status)
{
  stage_body<P, Q, TT, AFF>(mesh, ph, sa, lam, status, blockIdx.x);
}

// Batched form: the blocks of several independent meshes of the same (P, Q)
// in one launch.  items[] is sorted by first block; each block finds its
// member by a uniform binary search.
template <int P, int Q, int W, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_rhs_batch_kernel(const BatchItem* __restrict__ items, int nitems, int* status)
{
  const int b = blockIdx.x;
  int lo = 0, hi = nitems - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BatchItem& it = items[lo];
  stage_body<P, Q, double, AFF>(it.mesh, it.ph, it.sa, it.lam, status, b - it.block0);
}

#include "dgx_tile.hpp"

// BR2 lifting coefficient of each element's side of each of its faces:
//   lam[e][f][pt] = fJ(pt) * sum_m L_hi[m]^2 / (w_m detJ(line node m))
// (the pointwise form of Phi_f M_e^-1 Phi_f^T W_f; geometry only).
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_lift_coef_kernel(MeshArgs mesh, double* lam)
{
  constexpr int N = P + 1, NF = N * N, NG = (Q + 1) * (Q + 1) * (Q + 1);
  const PTab& pt = c_ptab[P - 1];
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)mesh.nelem * 6 * NF;
  if (t >= total)
    return;
  int fpt = (int)(t % NF);
  int f   = (int)((t / NF) % 6);
  size_t e = t / (6 * NF);
  int a = fpt % N, b = fpt / N;
  int nd = f >> 1, hi = f & 1;
  int t1 = (nd == 0) ? 1 : 0;
  int t2 = (nd == 2) ? 1 : 2;
  const double* X = mesh.nodes + e * NG * 3;
  int gi[3];
  gi[nd] = 0;
  gi[t1] = a;
  gi[t2] = b;
  double J[3][3];
  jac_at<Q>(c_qtab[P - 1][Q - 1], X, nd, hi, gi[0], gi[1], gi[2], J);
  int c1 = (nd + 1) % 3, c2 = (nd + 2) % 3;
  double n0 = J[1][c1] * J[2][c2] - J[2][c1] * J[1][c2];
  double n1 = J[2][c1] * J[0][c2] - J[0][c1] * J[2][c2];
  double n2 = J[0][c1] * J[1][c2] - J[1][c1] * J[0][c2];
  double fJ = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  double s = 0.0;
  for (int m = 0; m < N; ++m)
  {
    int gm[3];
    gm[nd] = m;
    gm[t1] = a;
    gm[t2] = b;
    double Jm[3][3];
    jac_at<Q>(c_qtab[P - 1][Q - 1], X, -1, 0, gm[0], gm[1], gm[2], Jm);
    s += pt.L[hi][m] * pt.L[hi][m] / (pt.w[m] * det3(Jm));
  }
  lam[t] = fJ * s;
}

// ---- basis conversions and residual output ------------------------------

// Applies a 1D n x n matrix along direction dir to every variable of the
// block's elements held in s (layout [EPW][5][NB]); trans selects M^T.
template <int N, int NB, int EPW>
__device__ __forceinline__ void apply_dir(double (*s)[5][NB], const double (*M)[MAXN], int dir,
                                          bool trans, int le, int node, bool act)
{
  double out[5];
  int i = node % N, j = (node / N) % N, k = node / (N * N);
  int idx[3] = {i, j, k};
  if (act)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        int id2[3] = {i, j, k};
        id2[dir] = m;
        double c = trans ? M[m][idx[dir]] : M[idx[dir]][m];
        acc += c * s[le][r][(id2[2] * N + id2[1]) * N + id2[0]];
      }
      out[r] = acc;
    }
  }
  __syncthreads();
  if (act)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      s[le][r][node] = out[r];
  }
  __syncthreads();
}

// Double-double arithmetic (hi + lo) for the state basis conversions: the
// state itself is known to 1 ulp, and the residual's sensitivity to the state
// (the stiff acoustic part of the operator) would amplify a conversion error
// of a few ulps, so V = B3 U is formed with error-free transformations
// (Knuth two-sum, FMA two-product) and rounded once.
struct DD
{
  double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b)
{
  double s  = a + b;
  double bb = s - a;
  double e  = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ __forceinline__ DD dd_fast(double a, double b)
{
  double s = a + b;
  return {s, b - (s - a)};
}

// (x.hi + x.lo) * (c + clo) + acc, all in double-double
__device__ __forceinline__ DD dd_fma(DD x, double c, double clo, DD acc)
{
  double p = x.hi * c;
  double e = fma(x.hi, c, -p);
  e = fma(x.hi, clo, e);
  e = fma(x.lo, c, e);
  DD s = dd_two_sum(acc.hi, p);
  s.lo += acc.lo + e;
  return dd_fast(s.hi, s.lo);
}

// mode 0: V = B3 U      (reference basis -> Gauss-nodal)
// mode 1: U = B3^-1 V   (Gauss-nodal -> reference basis)
// mode 2: G = B3^T g    (a gradient with respect to V -> with respect to U)
// mode 3: g = B3^-T G   (a gradient with respect to U -> with respect to V)
template <int P>
__global__ __launch_bounds__(64) void dgx_basis_kernel(const double* in, double* out, int nelem,
                                                       int mode)
{
  constexpr int N = P + 1, NB = N * N * N, EPW = 64 / NB;
  const PTab& pt = c_ptab[P - 1];
  __shared__ double sh[EPW][5][NB];
  __shared__ double sl[EPW][5][NB];
  int tid = threadIdx.x, le = tid / NB, node = tid % NB;
  int e = blockIdx.x * EPW + le;
  bool act = le < EPW && e < nelem;
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      sh[le][r][node] = in[((size_t)e * 5 + r) * NB + node];
      sl[le][r][node] = 0.0;
    }
  __syncthreads();
  const bool inv = (mode == 1 || mode == 3);
  const double(*M)[MAXN]  = inv ? pt.Binv : pt.B;
  const double(*Ml)[MAXN] = inv ? pt.Binvlo : pt.Blo;
  const bool tr = (mode >= 2);
  int idx[3] = {node % N, (node / N) % N, node / (N * N)};
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    DD o[5];
    if (act)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        DD acc = {0.0, 0.0};
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          int id2[3] = {idx[0], idx[1], idx[2]};
          id2[d] = m;
          int src = (id2[2] * N + id2[1]) * N + id2[0];
          const double c  = tr ? M[m][idx[d]] : M[idx[d]][m];
          const double cl = tr ? Ml[m][idx[d]] : Ml[idx[d]][m];
          acc = dd_fma({sh[le][r][src], sl[le][r][src]}, c, cl, acc);
        }
        o[r] = acc;
      }
    }
    __syncthreads();
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        sh[le][r][node] = o[r].hi;
        sl[le][r][node] = o[r].lo;
      }
    __syncthreads();
  }
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      out[((size_t)e * 5 + r) * NB + node] = sh[le][r][node] + sl[le][r][node];
}

// From f_g (Gauss-nodal): f = B3^-1 f_g, R = -M f = -B3^T W f_g.  Either
// output may be NULL.  If partial is not NULL, partial[block] receives the
// block's sum of R^2 (for the residual norm).
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_resid_out_kernel(MeshArgs mesh, const double* fg,
                                                           double* f, double* R,
                                                           double* partial)
{
  using Dm = Dims<P, Q>;
  constexpr int N = Dm::N, NB = Dm::NB, EPW = Dm::EPW, NG = Dm::NG;
  const PTab& pt = c_ptab[P - 1];
  __shared__ double s[EPW][5][NB];
  __shared__ double sX[EPW][NG * 3];
  __shared__ double red[64];
  int tid = threadIdx.x, le = tid / NB, node = tid % NB;
  int e0 = blockIdx.x * EPW, e = e0 + le;
  bool act = le < EPW && e < mesh.nelem;
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int l = x / (NG * 3);
    int ee = e0 + l;
    sX[l][x % (NG * 3)] = (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  double fgv[5];
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      fgv[r] = fg[((size_t)e * 5 + r) * NB + node];
      s[le][r][node] = fgv[r];
    }
  __syncthreads();
  if (f)
  {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      apply_dir<N, NB, EPW>(s, pt.Binv, d, false, le, node, act);
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
        f[((size_t)e * 5 + r) * NB + node] = s[le][r][node];
    __syncthreads();
  }
  double sq = 0.0;
  if (R || partial)
  {
    int ni = node % N, nj = (node / N) % N, nk = node / (N * N);
    if (act)
    {
      double J[3][3];
      jac_at<Q>(c_qtab[P - 1][Q - 1], sX[le], -1, 0, ni, nj, nk, J);
      double wJ = pt.w[ni] * pt.w[nj] * pt.w[nk] * det3(J);
#pragma unroll
      for (int r = 0; r < 5; ++r)
        s[le][r][node] = -wJ * fgv[r];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d)
      apply_dir<N, NB, EPW>(s, pt.B, d, true, le, node, act);
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double v = s[le][r][node];
        if (R)
          R[((size_t)e * 5 + r) * NB + node] = v;
        sq += v * v;
      }
  }
  if (partial)
  {
    red[tid] = sq;
    __syncthreads();
    if (tid == 0)
    {
      double t = 0.0;
      for (int x = 0; x < 64; ++x)
        t += red[x];
      partial[blockIdx.x] = t;
    }
  }
}

// Deterministic sum of the per-block partial sums into out[0]; with
// scale > 0, out[0] = sqrt(scale * sum) instead (the residual norm of
// residual_norm, time_distributed.cpp:23-44, with scale = 1 / state length).
__global__ __launch_bounds__(256) void dgx_sum_kernel(const double* partial, int n, double* out,
                                                      double scale)
{
  __shared__ double red[256];
  double t = 0.0;
  for (int x = threadIdx.x; x < n; x += 256)
    t += partial[x];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[0] = (scale > 0.0) ? sqrt(scale * red[0]) : red[0];
}

// Plain streaming copy (16 B per lane per iteration).
__global__ __launch_bounds__(256) void dgx_copy_kernel(const double2* __restrict__ src,
                                                      double2* __restrict__ dst, size_t n2)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    dst[i] = src[i];
}

hipError_t copy_dispatch(const void* src, void* dst, size_t bytes, hipStream_t st)
{
  size_t n2 = bytes / 16;
  int nblk = (int)((n2 + 255) / 256);
  if (nblk > 8192)
    nblk = 8192;
  if (nblk < 1)
    nblk = 1;
  hipLaunchKernelGGL(dgx_copy_kernel, dim3(nblk), dim3(256), 0, st, (const double2*)src,
                     (double2*)dst, n2);
  return hipGetLastError();
}

// ---- derived outputs -------------------------------------------------------

// eval_output (render/vrtxgen.cpp:25-59) on the (Ns+1)^3 equispaced sub-grid
// of every element, reference points (i/Ns, j/Ns, k/Ns) with sub-node index
// (Ns+1)^2 k + (Ns+1) j + i (vrtxgen.cpp:294-330).  The Gauss-nodal state is
// interpolated with the Gauss-node Lagrange basis, which spans the same
// polynomials as the reference's equispaced basis.  var: mach (0), rho (1),
// u (2), v (3), w (4); the Mach number keeps the reference's pressure
// (gamma - 1)(rho E - |u|^2 / 2).  One thread per sample point.
template <int P>
__global__ __launch_bounds__(256) void dgx_sample_kernel(const double* __restrict__ V,
                                                         int nelem, int Ns, int var,
                                                         double gamma,
                                                         double* __restrict__ out)
{
  constexpr int N = P + 1, NB = N * N * N;
  const PTab& pt = c_ptab[P - 1];
  const int np1 = Ns + 1, nsub = np1 * np1 * np1;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nelem * nsub)
    return;
  const int node = (int)(t % nsub);
  const size_t e = t / nsub;
  const int ijk[3] = {node % np1, (node / np1) % np1, node / (np1 * np1)};
  double l[3][N];
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    const double x = (double)ijk[c] / (double)Ns;
#pragma unroll
    for (int m = 0; m < N; ++m)
    {
      double v = 1.0;
#pragma unroll
      for (int o = 0; o < N; ++o)
        if (o != m)
          v *= (x - pt.g[o]) / (pt.g[m] - pt.g[o]);
      l[c][m] = v;
    }
  }
  double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double* Ve = V + e * 5 * NB;
#pragma unroll
  for (int kk = 0; kk < N; ++kk)
#pragma unroll
    for (int jj = 0; jj < N; ++jj)
#pragma unroll
      for (int ii = 0; ii < N; ++ii)
      {
        const double w = l[2][kk] * l[1][jj] * l[0][ii];
#pragma unroll
        for (int r = 0; r < 5; ++r)
          st[r] += w * Ve[r * NB + nid<N>(ii, jj, kk)];
      }
  double v = 0.0;
  if (var == 0)
  {
    const double u = st[1] / st[0], vv = st[2] / st[0], ww = st[3] / st[0];
    const double sp = sqrt(u * u + vv * vv + ww * ww);
    const double pr = (gamma - 1.0) * (st[4] - 0.5 * sp * sp);
    v = sp / sqrt(gamma * pr / st[0]);
  }
  else if (var == 1)
    v = st[0];
  else
    v = st[var - 1] / st[0];
  out[t] = v;
}

// update_mean (output.cpp:39-49): mean = (mean * old_range + U) / new_range.
__global__ __launch_bounds__(256) void dgx_mean_kernel(const double* __restrict__ U,
                                                       double* __restrict__ mean, size_t n,
                                                       double old_range, double new_range)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    mean[i] = (mean[i] * old_range + U[i]) / new_range;
}

// ---- discrete adjoint --------------------------------------------------------
#include "dgx_adjoint_kernels.hpp"

// ---- launch helpers (host) -----------------------------------------------

// rows per tile of the tiled stage kernel at each order (0: none compiled):
// p = 3: two one-element rows, so each face subpass is the 64 points of one
// direction of two elements; p = 1: one row of 8 elements (16 faces of 4
// points per direction); p = 2: three rows of 2 elements (12 faces of 9
// points per direction in two subpasses)
constexpr int tile_rows(int p)
{
  return p == 3 ? 2 : (p == 2 ? 3 : (p == 1 ? 1 : 0));
}

template <int P, int Q, typename TT, bool AFF>
hipError_t launch_tile_t(const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                         const double* lam, int* status, int waves, hipStream_t st)
{
  constexpr int R = tile_rows(P);
  constexpr int E = TileDims<P, R>::E;
  int nblk = (m.nelem + E - 1) / E;
  if (waves == 2)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 2, TT, AFF>), dim3(nblk), dim3(64), 0, st,
                       m, ph, sa, lam, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 3, TT, AFF>), dim3(nblk), dim3(64), 0, st,
                       m, ph, sa, lam, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int P, int Q, typename TT, bool AFF>
hipError_t launch_rhs_t(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                        int* status, int waves, int tile, hipStream_t st)
{
  if constexpr (Q == 1)
    if (tile)
      return launch_tile_t<P, Q, TT, AFF>(m, ph, sa, lam, status, waves, st);
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  if (waves == 1)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 1, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else if (waves == 2)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 2, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 3, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// affine q = 1 meshes take the constant-Jacobian instantiation
template <int P, int Q, typename TT>
hipError_t launch_rhs_a(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                        int* status, int waves, int tile, hipStream_t st)
{
  if constexpr (Q == 1)
    if (m.affine)
      return launch_rhs_t<P, Q, TT, true>(m, ph, sa, lam, status, waves, tile, st);
  return launch_rhs_t<P, Q, TT, false>(m, ph, sa, lam, status, waves, tile, st);
}

template <int P, int Q>
hipError_t launch_rhs(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                      int* status, int waves, int tile, hipStream_t st)
{
  if (sa.trace32)
    return launch_rhs_a<P, Q, float>(m, ph, sa, lam, status, waves, tile, st);
  return launch_rhs_a<P, Q, double>(m, ph, sa, lam, status, waves, tile, st);
}

template <int P, int Q, bool AFF>
hipError_t launch_rhs_batch_t(const BatchItem* items, int nitems, int nblk, int* status,
                              int waves, hipStream_t st)
{
  if (waves == 1)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 1, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else if (waves == 2)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 2, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 3, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// groups whose members are all affine q = 1 meshes take the constant-
// Jacobian instantiation
template <int P, int Q>
hipError_t launch_rhs_batch(const BatchItem* items, int nitems, int nblk, int* status, int waves,
                            bool affine, hipStream_t st)
{
  if constexpr (Q == 1)
    if (affine)
      return launch_rhs_batch_t<P, Q, true>(items, nitems, nblk, status, waves, st);
  return launch_rhs_batch_t<P, Q, false>(items, nitems, nblk, status, waves, st);
}

template <int P, int Q>
hipError_t launch_trace(const MeshArgs& m, const Phys& ph, const double* V, double* Tout,
                        int* status, hipStream_t st, bool trace32)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  const bool aff = (Q == 1) && m.affine;
  if (trace32 && aff)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, float, Q == 1>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, reinterpret_cast<float*>(Tout), status);
  else if (trace32)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, float, false>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, reinterpret_cast<float*>(Tout), status);
  else if (aff)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, double, Q == 1>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, Tout, status);
  else
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, double, false>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, Tout, status);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_lift_coef(const MeshArgs& m, double* lam, hipStream_t st)
{
  size_t total = (size_t)m.nelem * 6 * (P + 1) * (P + 1);
  int nblk = (int)((total + 63) / 64);
  hipLaunchKernelGGL((dgx_lift_coef_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, lam);
  return hipGetLastError();
}

template <int P>
hipError_t launch_basis(const double* in, double* out, int nelem, int mode, hipStream_t st)
{
  constexpr int EPW = 64 / ((P + 1) * (P + 1) * (P + 1));
  int nblk = (nelem + EPW - 1) / EPW;
  hipLaunchKernelGGL((dgx_basis_kernel<P>), dim3(nblk), dim3(64), 0, st, in, out, nelem, mode);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_resid_out(const MeshArgs& m, const double* fg, double* f, double* R,
                            double* partial, hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  hipLaunchKernelGGL((dgx_resid_out_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, fg, f, R,
                     partial);
  return hipGetLastError();
}

size_t trace_doubles(int p, int nx, int ny, int nz)
{
  size_t nf = (size_t)(p + 1) * (p + 1);
  return (size_t)nx * ny * nz * 6 * 8 * nf + boundary_faces(nx, ny, nz) * 5 * nf;
}

int resid_out_blocks(int p, int nelem)
{
  int nb = (p + 1) * (p + 1) * (p + 1);
  int epw = 64 / nb;
  return (nelem + epw - 1) / epw;
}

#define DGX_DISPATCH_PQ(p, q, CALL)                                                         \
  switch ((p) * 10 + (q))                                                                   \
  {                                                                                         \
    case 11: return CALL(1, 1);                                                             \
    case 12: return CALL(1, 2);                                                             \
    case 13: return CALL(1, 3);                                                             \
    case 21: return CALL(2, 1);                                                             \
    case 22: return CALL(2, 2);                                                             \
    case 23: return CALL(2, 3);                                                             \
    case 31: return CALL(3, 1);                                                             \
    case 32: return CALL(3, 2);                                                             \
    case 33: return CALL(3, 3);                                                             \
  }                                                                                         \
  return hipErrorInvalidValue;

// the tiled kernels are compiled for q = 1 (with higher-order geometry the
// extra rows of accumulators spill)
bool tile_valid(int p, int q, int tile, int waves)
{
  return tile == 0 || (tile == 1 && q == 1 && tile_rows(p) > 0 && (waves == 2 || waves == 3));
}

hipError_t rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                        const double* lam, int* status, int waves, hipStream_t st, int tile)
{
  if (!tile_valid(p, q, tile, waves))
    return hipErrorInvalidValue;
#define CALL_RHS(P, Q) launch_rhs<P, Q>(m, ph, sa, lam, status, waves, tile, st)
  DGX_DISPATCH_PQ(p, q, CALL_RHS)
#undef CALL_RHS
}

hipError_t rhs_batch_dispatch(int p, int q, const BatchItem* items_dev, int nitems, int nblocks,
                              int* status, int waves, bool affine, hipStream_t st)
{
#define CALL_RB(P, Q) launch_rhs_batch<P, Q>(items_dev, nitems, nblocks, status, waves, affine, st)
  DGX_DISPATCH_PQ(p, q, CALL_RB)
#undef CALL_RB
}

int stage_blocks(int p, int nelem)
{
  int epw = 64 / ((p + 1) * (p + 1) * (p + 1));
  return (nelem + epw - 1) / epw;
}

hipError_t trace_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                          double* Tout, int* status, hipStream_t st, bool trace32)
{
#define CALL_TR(P, Q) launch_trace<P, Q>(m, ph, V, Tout, status, st, trace32)
  DGX_DISPATCH_PQ(p, q, CALL_TR)
#undef CALL_TR
}

hipError_t lift_coef_dispatch(int p, int q, const MeshArgs& m, double* lam, hipStream_t st)
{
#define CALL_LC(P, Q) launch_lift_coef<P, Q>(m, lam, st)
  DGX_DISPATCH_PQ(p, q, CALL_LC)
#undef CALL_LC
}

hipError_t resid_out_dispatch(int p, int q, const MeshArgs& m, const double* fg, double* f,
                              double* R, double* partial, hipStream_t st)
{
#define CALL_RO(P, Q) launch_resid_out<P, Q>(m, fg, f, R, partial, st)
  DGX_DISPATCH_PQ(p, q, CALL_RO)
#undef CALL_RO
}

hipError_t basis_dispatch(int p, const double* in, double* out, int nelem, int mode,
                          hipStream_t st)
{
  switch (p)
  {
    case 1: return launch_basis<1>(in, out, nelem, mode, st);
    case 2: return launch_basis<2>(in, out, nelem, mode, st);
    case 3: return launch_basis<3>(in, out, nelem, mode, st);
  }
  return hipErrorInvalidValue;
}

hipError_t sample_dispatch(int p, const double* V, int nelem, int Ns, int var, double gamma,
                           double* out, hipStream_t st)
{
  size_t total = (size_t)nelem * (Ns + 1) * (Ns + 1) * (Ns + 1);
  int nblk = (int)((total + 255) / 256);
  switch (p)
  {
    case 1:
      hipLaunchKernelGGL(dgx_sample_kernel<1>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    case 2:
      hipLaunchKernelGGL(dgx_sample_kernel<2>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    case 3:
      hipLaunchKernelGGL(dgx_sample_kernel<3>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t mean_dispatch(const double* U, double* mean, size_t n, double old_range,
                         double new_range, hipStream_t st)
{
  int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(dgx_mean_kernel, dim3(nblk), dim3(256), 0, st, U, mean, n, old_range,
                     new_range);
  return hipGetLastError();
}

template <int P, int Q, bool AFF>
hipError_t launch_adjoint_t(const MeshArgs& m, const Phys& ph, const double* V, const double* Tin,
                            const double* lam, const double* psi, double* pvol, double* aown,
                            double* anbr, double* psiV, hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  // the volume part keeps the general geometry: its affine form holds 20
  // more VGPRs (146 vs 126 at p = 2), one wave/SIMD less, and runs slower
  hipLaunchKernelGGL((dgx_adj_main_kernel<P, Q, 0, false>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     Tin, lam, psi, pvol, aown, anbr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL((dgx_adj_main_kernel<P, Q, 1, AFF>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     Tin, lam, psi, pvol, aown, anbr);
  e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL((dgx_adj_trace_kernel<P, Q, AFF>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     aown, anbr, pvol, psiV);
  return hipGetLastError();
}

// affine q = 1 meshes take the constant-Jacobian instantiation
template <int P, int Q>
hipError_t launch_adjoint(const MeshArgs& m, const Phys& ph, const double* V, const double* Tin,
                          const double* lam, const double* psi, double* pvol, double* aown,
                          double* anbr, double* psiV, hipStream_t st)
{
  if constexpr (Q == 1)
    if (m.affine)
      return launch_adjoint_t<P, Q, true>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st);
  return launch_adjoint_t<P, Q, false>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st);
}

hipError_t adjoint_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                            const double* Tin, const double* lam, const double* psi,
                            double* pvol, double* aown, double* anbr, double* psiV,
                            hipStream_t st)
{
#define CALL_ADJ(P, Q) launch_adjoint<P, Q>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st)
  DGX_DISPATCH_PQ(p, q, CALL_ADJ)
#undef CALL_ADJ
}

template <int P, int Q>
hipError_t launch_kinetic_energy(const MeshArgs& m, const double* V, double* grad,
                                 double* partial, hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  hipLaunchKernelGGL((dgx_kinetic_energy_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, V, grad,
                     partial);
  return hipGetLastError();
}

hipError_t kinetic_energy_dispatch(int p, int q, const MeshArgs& m, const double* V,
                                   double* grad, double* partial, hipStream_t st)
{
#define CALL_KE(P, Q) launch_kinetic_energy<P, Q>(m, V, grad, partial, st)
  DGX_DISPATCH_PQ(p, q, CALL_KE)
#undef CALL_KE
}

hipError_t axpby_dispatch(double* y, const double* x, double a, double b, size_t n,
                          hipStream_t st)
{
  int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(dgx_axpby_kernel, dim3(nblk), dim3(256), 0, st, y, x, a, b, n);
  return hipGetLastError();
}

hipError_t adj_tail_dispatch(const double* l, const double* a, const double* sum_in,
                             double* sum_out, double* lk, double wl, double wa, size_t n,
                             hipStream_t st)
{
  int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(dgx_adj_tail_kernel, dim3(nblk), dim3(256), 0, st, l, a, sum_in, sum_out,
                     lk, wl, wa, n);
  return hipGetLastError();
}

hipError_t dot_dispatch(const double* x, const double* y, size_t n, double* partial,
                        int nblocks, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(dgx_dot_kernel, dim3(nblocks), dim3(256), 0, st, x, y, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  return sum_dispatch(partial, nblocks, out, st);
}

hipError_t sum_dispatch(const double* partial, int n, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(dgx_sum_kernel, dim3(1), dim3(256), 0, st, partial, n, out, 0.0);
  return hipGetLastError();
}

hipError_t norm_dispatch(const double* partial, int n, double inv_count, double* out,
                         hipStream_t st)
{
  hipLaunchKernelGGL(dgx_sum_kernel, dim3(1), dim3(256), 0, st, partial, n, out, inv_count);
  return hipGetLastError();
}

hipError_t upload_tables(const PTab* pt, const QTab* qt)
{
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ptab), pt, sizeof(PTab) * 3);
  if (e != hipSuccess)
    return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_qtab), qt, sizeof(QTab) * 9);
}

}  // namespace dgx
