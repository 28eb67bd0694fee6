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
// one lane per Gauss node.  Each element evaluates its own six faces with the
// neighbour's trace recomputed from the neighbour's state, so there are no
// atomics and no face-residual buffers; every shared face is evaluated once
// from each side (as the reference does on its processor-edge faces,
// solve.cpp:400, 530, 607).
#include <hip/hip_runtime.h>

#include "dgx_device.hpp"

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
template <int P, int Q>
__device__ __forceinline__ void geo_vals(int nd, int e, int i, int j, int k, const double** v,
                                         const double** d)
{
  const QTab& qt = c_qtab[P - 1][Q - 1];
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

template <int P, int Q>
__device__ __forceinline__ void jac_at(const double* X, int nd, int e, int i, int j, int k,
                                       double J[3][3])
{
  const double* v[3];
  const double* d[3];
  geo_vals<P, Q>(nd, e, i, j, k, v, d);
  jacobian<Q>(X, v[0], d[0], v[1], d[1], v[2], d[2], J);
}

// Neighbour of element (i,j,k) across local face f (0:-x 1:+x 2:-y 3:+y
// 4:-z 5:+z) on the structured mesh; returns -1 with *bt set for a
// non-periodic domain boundary (geometry_shared.cpp:199-229).
__device__ __forceinline__ int neighbour(const MeshArgs& m, int i, int j, int k, int f, int* bt)
{
  int nd = f >> 1, hi = f & 1;
  int c[3] = {i, j, k};
  int n[3] = {m.nx, m.ny, m.nz};
  int x = c[nd] + (hi ? 1 : -1);
  if (x < 0 || x >= n[nd])
  {
    if (m.bounds[f] != -2)
    {
      *bt = m.bounds[f];
      return -1;
    }
    x = (x + n[nd]) % n[nd];
  }
  c[nd] = x;
  *bt = -1;
  return (c[2] * m.ny + c[1]) * m.nx + c[0];
}

// One full right-hand-side evaluation f_g = B3 f of EPW elements fused with
// the RK stage combination of StageArgs.
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_rhs_stage_kernel(MeshArgs mesh, Phys ph,
                                                           StageArgs sa, int* status)
{
  using Dm = Dims<P, Q>;
  constexpr int N = Dm::N, NB = Dm::NB, NF = Dm::NF, EPW = Dm::EPW, NG = Dm::NG;
  constexpr int NFP = Dm::NFP;
  const PTab& pt = c_ptab[P - 1];

  __shared__ double sV[EPW][5][NB];
  __shared__ double sWJ[EPW][NB];
  __shared__ double sX[EPW][NG * 3];
  __shared__ double sR1[Dm::RGN1];       // volume flux G, then face results
  __shared__ double sUF[NFP][10];        // face traces: own (0-4), neighbour (5-9)

  const int tid = threadIdx.x;
  const int le  = tid / NB;
  const int nd0 = tid % NB;
  const int e0  = blockIdx.x * EPW;
  const int e   = e0 + le;
  const bool vact = (le < EPW) && (e < mesh.nelem);

  // ---- load state and element geometry ----
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = sa.Vin[((size_t)e * 5 + r) * NB + nd0];
  }
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int l = x / (NG * 3);
    int ee = e0 + l;
    sX[l][x % (NG * 3)] = (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  __syncthreads();

  const int ni = nd0 % N, nj = (nd0 / N) % N, nk = nd0 / (N * N);
  double Ji[3][3];
  double wJ = 1.0;
  double acc[5] = {0, 0, 0, 0, 0};

  // ---- volume: gradients, fluxes, contravariant flux G ----
  if (vact)
  {
    double J[3][3];
    jac_at<P, Q>(sX[le], -1, 0, ni, nj, nk, J);
    double det = det3(J);
    inv3(J, det, Ji);
    wJ = pt.w[ni] * pt.w[nj] * pt.w[nk] * det;
    sWJ[le][nd0] = wJ;

    double Vq[5], Gr[3][5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      Vq[r] = sV[le][r][nd0];
      double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        gx += pt.Dg[ni][m] * sV[le][r][nid<N>(m, nj, nk)];
        gy += pt.Dg[nj][m] * sV[le][r][nid<N>(ni, m, nk)];
        gz += pt.Dg[nk][m] * sV[le][r][nid<N>(ni, nj, m)];
      }
      // physical gradient: d/dx_d = sum_c d/dxi_c * Ji[c][d]
#pragma unroll
      for (int d = 0; d < 3; ++d)
        Gr[d][r] = gx * Ji[0][d] + gy * Ji[1][d] + gz * Ji[2][d];
    }
    double F[3][5], Qv[3][5];
    euler_flux(Vq, ph.gamma, F);
    ViscState vs = visc_setup(Vq);
    visc_apply(vs, Gr, ph.mu, Qv);
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double h0 = Qv[0][r] - F[0][r], h1 = Qv[1][r] - F[1][r], h2 = Qv[2][r] - F[2][r];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        sR1[((le * 3 + c) * 5 + r) * NB + nd0] = wJ * (Ji[c][0] * h0 + Ji[c][1] * h1 + Ji[c][2] * h2);
    }
  }
  __syncthreads();

  // ---- volume divergence: acc = sum_c Dg_c^T G_c ----
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        s += pt.Dg[m][ni] * sR1[((le * 3 + 0) * 5 + r) * NB + nid<N>(m, nj, nk)];
        s += pt.Dg[m][nj] * sR1[((le * 3 + 1) * 5 + r) * NB + nid<N>(ni, m, nk)];
        s += pt.Dg[m][nk] * sR1[((le * 3 + 2) * 5 + r) * NB + nid<N>(ni, nj, m)];
      }
      acc[r] = s;
    }
  }
  __syncthreads();

  // ---- faces, pass 1: traces and normal derivatives of both sides ----
  constexpr int FPT = (NFP + 63) / 64;  // face points per thread
  double dnL[FPT][5], dnR[FPT][5];
  int nbrE[FPT], nbrBt[FPT];
#pragma unroll
  for (int s = 0; s < FPT; ++s)
  {
    int fp = tid + 64 * s;
    nbrE[s] = -1;
    nbrBt[s] = 0;
    if (fp >= NFP) continue;
    int l = fp / (6 * NF);
    int f = (fp / NF) % 6;
    int a = fp % N, b = (fp / N) % N;
    int ee = e0 + l;
    if (ee >= mesh.nelem) continue;
    int nd = f >> 1, hi = f & 1;
    int ci = ee % mesh.nx, cj = (ee / mesh.nx) % mesh.ny, ck = ee / (mesh.nx * mesh.ny);
    int bt;
    int en = neighbour(mesh, ci, cj, ck, f, &bt);
    nbrE[s] = en;
    nbrBt[s] = bt;
    double uL[5], dL[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double v = 0.0, dv = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        double x = sV[l][r][line_node<N>(nd, m, a, b)];
        v += pt.L[hi][m] * x;
        dv += pt.Ld[hi][m] * x;
      }
      uL[r] = v;
      dL[r] = dv;
      sUF[fp][r] = v;
      dnL[s][r] = dv;
    }
    if (en >= 0)
    {
      const double* Vn = sa.Vin + (size_t)en * 5 * NB;
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double v = 0.0, dv = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          double x = Vn[r * NB + line_node<N>(nd, m, a, b)];
          v += pt.L[1 - hi][m] * x;
          dv += pt.Ld[1 - hi][m] * x;
        }
        sUF[fp][5 + r] = v;
        dnR[s][r] = dv;
      }
    }
    else
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        sUF[fp][5 + r] = 0.0;
        dnR[s][r] = 0.0;
      }
    }
    (void)uL;
    (void)dL;
  }
  __syncthreads();

  // ---- faces, pass 2: geometry, tangential derivatives, physics ----
  // results per face point (scaled by fJ w_a w_b):
  //   R1[0-4]  G  = F - QH.n          (phi-weighted term)
  //   R1[5-9]  Hn = (Ji dcQ)_normal   (normal-derivative-weighted term)
  //   R1[10-14], R1[15-19] tangential contravariant dcQ components
  int bad = 0;
  double res[FPT][20];
#pragma unroll
  for (int s = 0; s < FPT; ++s)
  {
    int fp = tid + 64 * s;
    if (fp >= NFP) continue;
    int l = fp / (6 * NF);
    int f = (fp / NF) % 6;
    int a = fp % N, b = (fp / N) % N;
    int ee = e0 + l;
    if (ee >= mesh.nelem)
    {
#pragma unroll
      for (int x = 0; x < 20; ++x) res[s][x] = 0.0;
      continue;
    }
    int nd = f >> 1, hi = f & 1;
    int t1 = (nd == 0) ? 1 : 0;
    int t2 = (nd == 2) ? 1 : 2;
    int fbase = fp - a - N * b;  // first point of this face

    // own geometry at the face point
    int gi[3];
    gi[nd] = 0; gi[t1] = a; gi[t2] = b;
    double J[3][3], JiL[3][3];
    jac_at<P, Q>(sX[l], nd, hi, gi[0], gi[1], gi[2], J);
    double detL = det3(J);
    inv3(J, detL, JiL);
    // outward normal: +/- (column c1 x column c2), c1,c2 cyclic after nd
    int c1 = (nd + 1) % 3, c2 = (nd + 2) % 3;
    double nv[3];
    nv[0] = J[1][c1] * J[2][c2] - J[2][c1] * J[1][c2];
    nv[1] = J[2][c1] * J[0][c2] - J[0][c1] * J[2][c2];
    nv[2] = J[0][c1] * J[1][c2] - J[1][c1] * J[0][c2];
    double fJ = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    double sg = (hi ? 1.0 : -1.0) / fJ;
    nv[0] *= sg; nv[1] *= sg; nv[2] *= sg;
    double fw = pt.w[a] * pt.w[b];

    // lifting coefficient of the own side
    double lamL = 0.0;
#pragma unroll
    for (int m = 0; m < N; ++m)
      lamL += pt.L[hi][m] * pt.L[hi][m] * fw / sWJ[l][line_node<N>(nd, m, a, b)];
    lamL *= fJ;

    double UL[5], GL[3][5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      UL[r] = sUF[fp][r];
      double ref[3];
      ref[nd] = dnL[s][r];
      double d1 = 0.0, d2 = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        d1 += pt.Dg[a][m] * sUF[fbase + m + N * b][r];
        d2 += pt.Dg[b][m] * sUF[fbase + a + N * m][r];
      }
      ref[t1] = d1;
      ref[t2] = d2;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        GL[d][r] = ref[0] * JiL[0][d] + ref[1] * JiL[1][d] + ref[2] * JiL[2][d];
    }

    double Gf[5], dcQ[3][5];
    ViscState vL = visc_setup(UL);
    if (nbrE[s] >= 0)
    {
      // neighbour geometry: Jacobian at its face point and along its line
      int en = nbrE[s];
      const double* Xr = mesh.nodes + (size_t)en * NG * 3;
      double JR[3][3], JiR[3][3];
      jac_at<P, Q>(Xr, nd, 1 - hi, gi[0], gi[1], gi[2], JR);
      inv3(JR, det3(JR), JiR);
      double lamR = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        int gm[3];
        gm[nd] = m; gm[t1] = a; gm[t2] = b;
        double Jm[3][3];
        jac_at<P, Q>(Xr, -1, 0, gm[0], gm[1], gm[2], Jm);
        lamR += pt.L[1 - hi][m] * pt.L[1 - hi][m] / (pt.w[m] * det3(Jm));
      }
      lamR *= fJ;

      double UR[5], GR[3][5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        UR[r] = sUF[fp][5 + r];
        double ref[3];
        ref[nd] = dnR[s][r];
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          d1 += pt.Dg[a][m] * sUF[fbase + m + N * b][5 + r];
          d2 += pt.Dg[b][m] * sUF[fbase + a + N * m][5 + r];
        }
        ref[t1] = d1;
        ref[t2] = d2;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          GR[d][r] = ref[0] * JiR[0][d] + ref[1] * JiR[1][d] + ref[2] * JiR[2][d];
      }

      double Fr[5];
      roe_flux(UL, UR, nv, ph.gamma, Fr);
      ViscState vR = visc_setup(UR);
      // D = (UL - UR) n ; dcQ_L = A(UL, D/2) ; QH.n from linearity of A:
      // QH = A(UL, GL/2 - eta lamL D/4) + A(UR, GR/2 - eta lamR D/4)
      double Dj[3][5], argL[3][5], argR[3][5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double jmp = UL[r] - UR[r];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
          Dj[d][r]   = 0.5 * jmp * nv[d];
          argL[d][r] = 0.5 * GL[d][r] - 0.25 * ph.eta * lamL * jmp * nv[d];
          argR[d][r] = 0.5 * GR[d][r] - 0.25 * ph.eta * lamR * jmp * nv[d];
        }
      }
      visc_apply(vL, Dj, ph.mu, dcQ);
      double qL[5], qR[5];
      visc_apply_n(vL, argL, ph.mu, nv, qL);
      visc_apply_n(vR, argR, ph.mu, nv, qR);
#pragma unroll
      for (int r = 0; r < 5; ++r)
        Gf[r] = Fr[r] - (qL[r] + qR[r]);
    }
    else
    {
      double Ub[5];
      if (!boundary_state(nbrBt[s], UL, nv, ph, Ub))
        bad = 1;
      double Fr[5];
      euler_flux_n(Ub, ph.gamma, nv, Fr);
      ViscState vb = visc_setup(Ub);
      double Dj[3][5], argL[3][5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double jmp = UL[r] - Ub[r];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          Dj[d][r] = jmp * nv[d];
      }
      // lifting dL = lamL A(Ub, (UL-Ub) n)  (no 1/2, A at the boundary state)
      double Ql[3][5];
      visc_apply(vb, Dj, ph.mu, Ql);
      visc_apply(vL, Dj, ph.mu, dcQ);
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          argL[d][r] = GL[d][r];
      double qL[5];
      visc_apply_n(vL, argL, ph.mu, nv, qL);
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double dln = Ql[0][r] * nv[0] + Ql[1][r] * nv[1] + Ql[2][r] * nv[2];
        Gf[r] = Fr[r] - (qL[r] - ph.eta * lamL * dln);
      }
    }
    double sc = fJ * fw;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double h[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        h[c] = JiL[c][0] * dcQ[0][r] + JiL[c][1] * dcQ[1][r] + JiL[c][2] * dcQ[2][r];
      res[s][r]      = sc * Gf[r];
      res[s][5 + r]  = sc * h[nd];
      res[s][10 + r] = sc * h[t1];
      res[s][15 + r] = sc * h[t2];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < FPT; ++s)
  {
    int fp = tid + 64 * s;
    if (fp >= NFP) continue;
#pragma unroll
    for (int x = 0; x < 20; ++x)
      sR1[fp * 20 + x] = res[s][x];
  }
  __syncthreads();

  // ---- faces, pass 3: T = G - Dg^T(t1) H_t1 - Dg^T(t2) H_t2 ----
#pragma unroll
  for (int s = 0; s < FPT; ++s)
  {
    int fp = tid + 64 * s;
    if (fp >= NFP) continue;
    int a = fp % N, b = (fp / N) % N;
    int fbase = fp - a - N * b;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double t = res[s][r];
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        t -= pt.Dg[m][a] * sR1[(fbase + m + N * b) * 20 + 10 + r];
        t -= pt.Dg[m][b] * sR1[(fbase + a + N * m) * 20 + 15 + r];
      }
      res[s][r] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < FPT; ++s)
  {
    int fp = tid + 64 * s;
    if (fp >= NFP) continue;
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sR1[fp * 20 + r] = res[s][r];
  }
  __syncthreads();

  if (bad)
    atomicOr(status, 1);

  // ---- gather face terms at the nodes, form f_g, RK epilogue ----
  if (vact)
  {
    int idx[3] = {ni, nj, nk};
    double fg[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double s = acc[r];
#pragma unroll
      for (int f = 0; f < 6; ++f)
      {
        int nd = f >> 1, hi = f & 1;
        int t1 = (nd == 0) ? 1 : 0;
        int t2 = (nd == 2) ? 1 : 2;
        int m = idx[nd], a = idx[t1], b = idx[t2];
        int fp = (le * 6 + f) * NF + b * N + a;
        s += pt.L[hi][m] * sR1[fp * 20 + r] - pt.Ld[hi][m] * sR1[fp * 20 + 5 + r];
      }
      fg[r] = ph.source[r] - s / wJ;
    }
    size_t base = (size_t)e * 5 * NB + nd0;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      size_t ix = base + (size_t)r * NB;
      double f = fg[r];
      if (sa.fout)
        sa.fout[ix] = f;
      double ft = f;
      if (sa.accmode == 1)
        sa.acc[ix] = f;
      else if (sa.accmode == 2)
        sa.acc[ix] = sa.acc[ix] + 2.0 * f;
      else if (sa.accmode == 3)
        ft = (sa.acc[ix] + f) * (1.0 / 6.0);
      if (sa.Vout)
      {
        double o = sa.cf * ft;
        if (sa.ca != 0.0)
          o += sa.ca * sa.Vbase[ix];
        if (sa.cb != 0.0)
          o += sa.cb * sV[le][r][nd0];
        sa.Vout[ix] = o;
      }
    }
  }
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
  const double(*M)[MAXN]  = mode == 0 ? pt.B : pt.Binv;
  const double(*Ml)[MAXN] = mode == 0 ? pt.Blo : pt.Binvlo;
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
          acc = dd_fma({sh[le][r][src], sl[le][r][src]}, M[idx[d]][m], Ml[idx[d]][m], acc);
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
      jac_at<P, Q>(sX[le], -1, 0, ni, nj, nk, J);
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

// Deterministic sum of the per-block partial sums into out[0].
__global__ __launch_bounds__(256) void dgx_sum_kernel(const double* partial, int n, double* out)
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
    out[0] = red[0];
}

// ---- launch helpers (host) -----------------------------------------------

template <int P, int Q>
hipError_t launch_rhs(const MeshArgs& m, const Phys& ph, const StageArgs& sa, int* status,
                      hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, ph, sa,
                     status);
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

hipError_t rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                        int* status, hipStream_t st)
{
#define CALL_RHS(P, Q) launch_rhs<P, Q>(m, ph, sa, status, st)
  DGX_DISPATCH_PQ(p, q, CALL_RHS)
#undef CALL_RHS
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

hipError_t sum_dispatch(const double* partial, int n, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(dgx_sum_kernel, dim3(1), dim3(256), 0, st, partial, n, out);
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
