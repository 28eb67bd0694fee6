// dgx_physics.hpp -- pointwise physics of the stage kernel written once for
// a generic scalar type T: double in the residual kernels, forward-mode dual
// numbers (local Jacobians) in the discrete adjoint.  Arguments that the
// adjoint's dual passes hold fixed (reference-space gradients, gradient
// tractions) have their own type G / V, so a pass over the state carries no
// zero derivatives for them.
#pragma once

#include "dgx_device.hpp"

namespace dgx
{

// ---- viscous flux in velocity-gradient form ---------------------------------
// The Newtonian flux A(s, grad s) of flux.cpp:156-202 depends on the
// conserved-variable gradient only through the velocity gradient
//   Lv[i][d] = d u_i / d x_d = (d m_i/dx_d - u_i d rho/dx_d) / rho
// (there is no heat flux), and every BR2 / dual-consistency argument is a
// jump times the normal, whose velocity gradient is rank one:
//   (U_jump (x) n)  ->  dm (x) n,   dm_i = (jump_m_i - u_i jump_rho) / rho.
// tau(X) = mu (X + X^T) - 2/3 mu tr(X) I;  the energy row is u . tau.

// w = tau(X) n for X with X n = xn, X^T n = xtn, tr X = tr
template <typename T>
__device__ __forceinline__ void tau_n(double mu, const T* xn, const T* xtn, T tr, const double* n,
                                      T* w)
{
  T lt = -(2.0 / 3.0) * mu * tr;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    w[i] = mu * (xn[i] + xtn[i]) + lt * n[i];
}

// velocity gradient from reference-space derivatives gr[c][0..3] of
// (rho, m) and the inverse Jacobian Ji[c][d]
template <typename T, typename G>
__device__ __forceinline__ void vel_grad(const G gr[3][4], const T* u, T ir, const double Ji[3][3],
                                         T Lv[3][3])
{
  T dv[3][3];  // dv[c][i] = d u_i / d xi_c
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      dv[c][i] = (gr[c][1 + i] - gr[c][0] * u[i]) * ir;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      Lv[i][d] = dv[0][i] * Ji[0][d] + dv[1][i] * Ji[1][d] + dv[2][i] * Ji[2][d];
}

// One face point seen from the own side: outputs G = F - QH.n (5) and the
// physical dual-consistency flux dcQ as a symmetric stress S (6: xx yy zz
// xy xz yz) whose energy row is u_L . S.
//
// Interior (solve.cpp:393-626):
//   F   = roe(UL, UR, n)
//   dcQ = A(UL, (UL - (UL+UR)/2) n)            = tau(dmL/2 (x) n)
//   QH  = (A(UL,gL) + A(UR,gR))/2 - eta (dL + dR)/2,
//   dL  = lamL/2 A(UL, (UL-UR) n),  dR = lamR/2 A(UR, (UL-UR) n)
// so QH.n = tau(Lv_L/2 - cL dmL(x)n) n + tau(Lv_R/2 - cR dmR(x)n) n with
// c = eta lam / 4, plus the energy rows u_L.(...) and u_R.(...).
// Boundary (solve.cpp:628-794): Ub = boundary state, F = F(Ub).n,
//   dcQ = tau(dm (x) n) with jump UL - Ub,  QH.n = A(UL, gL).n - eta lamL A(Ub, (UL-Ub) n).n
template <typename T>
struct FacePointT
{
  T G[5];
  T S[6];
};
using FacePoint = FacePointT<double>;

// Gradient traction tau(grad u).n of one side, from the reference-space
// derivatives gr[c][0..3] of (rho, m), the inverse Jacobian Ji and the normal:
// with dv[c][i] = d u_i/d xi_c,  (L n)_i = sum_c dv[c][i] (Ji n)_c,
// (L^T n)_d = sum_c Ji[c][d] (dv[c] . n),  tr L = sum_c dv[c] . Ji[c].
template <typename T, typename G>
__device__ __forceinline__ void grad_traction(const G gr[3][4], const T* U, const double Ji[3][3],
                                              const double* n, double mu, T* vg)
{
  T ir = rcp(U[0]);
  T u[3] = {U[1] * ir, U[2] * ir, U[3] * ir};
  T Ln[3] = {0.0, 0.0, 0.0}, LTn[3] = {0.0, 0.0, 0.0}, tr = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    T dv0 = (gr[c][1] - gr[c][0] * u[0]) * ir;
    T dv1 = (gr[c][2] - gr[c][0] * u[1]) * ir;
    T dv2 = (gr[c][3] - gr[c][0] * u[2]) * ir;
    double mc = Ji[c][0] * n[0] + Ji[c][1] * n[1] + Ji[c][2] * n[2];
    T kc = dv0 * n[0] + dv1 * n[1] + dv2 * n[2];
    Ln[0] += dv0 * mc;
    Ln[1] += dv1 * mc;
    Ln[2] += dv2 * mc;
    LTn[0] += Ji[c][0] * kc;
    LTn[1] += Ji[c][1] * kc;
    LTn[2] += Ji[c][2] * kc;
    tr += dv0 * Ji[c][0] + dv1 * Ji[c][1] + dv2 * Ji[c][2];
  }
  tau_n(mu, Ln, LTn, tr, n, vg);
}

// vgL, vgR: gradient tractions tau(grad u).n of the own and neighbour traces
// (vgR unused on boundary faces).  tau(a (x) n).n = mu a + mu/3 (a.n) n.
// Ub is the neighbour trace (interior) or the boundary state (boundary).
template <typename T, typename V>
__device__ __forceinline__ void face_point_physics(bool interior, const T* UL, const T* Ub,
                                                   const V* vgL, const V* vgR, const double* n,
                                                   double lamL, double lamR, const Phys& ph,
                                                   FacePointT<T>& out)
{
  const double mu = ph.mu;
  T irL = rcp(UL[0]);
  T uL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
  if (interior)
    roe_flux(UL, Ub, n, ph.gamma, out.G);
  else
    euler_flux_n(Ub, ph.gamma, n, out.G);
  T irB = rcp(Ub[0]);
  T uB[3] = {Ub[1] * irB, Ub[2] * irB, Ub[3] * irB};
  T jr = UL[0] - Ub[0];
  T dmL[3], dmB[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    T jm = UL[1 + i] - Ub[1 + i];
    dmL[i] = (jm - jr * uL[i]) * irL;
    dmB[i] = (jm - jr * uB[i]) * irB;
  }
  T dnL = dmL[0] * n[0] + dmL[1] * n[1] + dmL[2] * n[2];
  T dnB = dmB[0] * n[0] + dmB[1] * n[1] + dmB[2] * n[2];

  // own side: hL tau(grad u_L).n - cL tau(dmL (x) n).n (interior: halved,
  // with the penalty; boundary: the full gradient traction)
  double hL = interior ? 0.5 : 1.0;
  double cL = interior ? 0.25 * ph.eta * lamL : 0.0;
  T w[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    w[i] = hL * vgL[i] - cL * mu * (dmL[i] + (1.0 / 3.0) * dnL * n[i]);
  T q[4] = {w[0], w[1], w[2], uL[0] * w[0] + uL[1] * w[1] + uL[2] * w[2]};

  // other side: neighbour (interior) or boundary penalty at the boundary state
  double cB = interior ? 0.25 * ph.eta * lamR : ph.eta * lamL;
  double hB = interior ? 0.5 : 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    w[i] = hB * vgR[i] - cB * mu * (dmB[i] + (1.0 / 3.0) * dnB * n[i]);
  q[0] += w[0];
  q[1] += w[1];
  q[2] += w[2];
  q[3] += uB[0] * w[0] + uB[1] * w[1] + uB[2] * w[2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    out.G[1 + i] -= q[i];
  out.G[4] -= q[3];

  // dual-consistency stress tau(s dmL (x) n), s = 1/2 (interior) or 1
  double s = interior ? 0.5 : 1.0;
  T a0 = s * dmL[0], a1 = s * dmL[1], a2 = s * dmL[2];
  T lt = -(2.0 / 3.0) * mu * s * dnL;
  out.S[0] = 2.0 * mu * a0 * n[0] + lt;
  out.S[1] = 2.0 * mu * a1 * n[1] + lt;
  out.S[2] = 2.0 * mu * a2 * n[2] + lt;
  out.S[3] = mu * (a0 * n[1] + a1 * n[0]);
  out.S[4] = mu * (a0 * n[2] + a2 * n[0]);
  out.S[5] = mu * (a1 * n[2] + a2 * n[1]);
}


// Volume flux H = Q - F (flux.cpp:24-48 Euler, flux.cpp:156-202 Newtonian
// viscous, no heat flux) at a Gauss node, in velocity-gradient form, from the
// node state Vq and the reference-space derivatives gr[c][0..3] of (rho, m).
template <typename T, typename G>
__device__ __forceinline__ void volume_flux(const T* Vq, const G gr[3][4], const double Ji[3][3],
                                            double gamma, double mu, T H[3][5])
{
  T ir = rcp(Vq[0]);
  T u[3] = {Vq[1] * ir, Vq[2] * ir, Vq[3] * ir};
  T Lv[3][3];
  vel_grad(gr, u, ir, Ji, Lv);
  T p  = (gamma - 1.0) * (Vq[4] - 0.5 * Vq[0] * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  T Ht = (Vq[4] + p) * ir;
  T lt = -(2.0 / 3.0) * mu * (Lv[0][0] + Lv[1][1] + Lv[2][2]);
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    T tau[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      tau[i] = mu * (Lv[d][i] + Lv[i][d]);
      if (i == d)
        tau[i] += lt;
    }
    T md    = Vq[1 + d];
    H[d][0] = -md;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      H[d][1 + i] = tau[i] - md * u[i];
      if (i == d)
        H[d][1 + i] -= p;
    }
    H[d][4] = u[0] * tau[0] + u[1] * tau[1] + u[2] * tau[2] - md * Ht;
  }
}

// Inviscid part of volume_flux (H = -F, the Euler flux): the values
// volume_flux gives with mu = 0, whose viscous terms are exact zeros.
template <typename T>
__device__ __forceinline__ void euler_volume_flux(const T* Vq, double gamma, T H[3][5])
{
  T ir = rcp(Vq[0]);
  T u[3] = {Vq[1] * ir, Vq[2] * ir, Vq[3] * ir};
  T p  = (gamma - 1.0) * (Vq[4] - 0.5 * Vq[0] * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  T Ht = (Vq[4] + p) * ir;
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    T md    = Vq[1 + d];
    H[d][0] = -md;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      H[d][1 + i] = -(md * u[i]);
      if (i == d)
        H[d][1 + i] -= p;
    }
    H[d][4] = -(md * Ht);
  }
}

}  // namespace dgx
