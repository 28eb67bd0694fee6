// dgx_vjp.hpp -- reverse sweeps (vector-Jacobian products) of the pointwise
// physics of dgx_physics.hpp, used by the discrete adjoint kernels: the
// volume flux, the interior face-point physics (with roe_flux_vjp of
// dgx_device.hpp) and the gradient traction.  Each is the exact transpose of
// the forward code it names; tests/test_adjoint_local.py checks every one
// against forward-mode dual numbers (dgx_dual.hpp) on random states.
#pragma once

#include "dgx_physics.hpp"

namespace dgx
{

// The dual-consistency stress S of an interior face point, the out.S of
// face_point_physics(true, ...): tau(dmL/2 (x) n) with dmL the velocity-like
// jump (m jump - u_L rho jump)/rho_L.  It needs neither the Roe flux nor the
// tractions, so the adjoint kernel forms it on its own.
__device__ __forceinline__ void interior_dc_stress(const double* UL, const double* UR,
                                                   const double* n, double mu, double* S)
{
  const double irL = rcp(UL[0]);
  const double jr  = UL[0] - UR[0];
  double a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    a[i] = 0.5 * ((UL[1 + i] - UR[1 + i]) - jr * (UL[1 + i] * irL)) * irL;
  const double lt = -(2.0 / 3.0) * mu * (a[0] * n[0] + a[1] * n[1] + a[2] * n[2]);
  S[0] = 2.0 * mu * a[0] * n[0] + lt;
  S[1] = 2.0 * mu * a[1] * n[1] + lt;
  S[2] = 2.0 * mu * a[2] * n[2] + lt;
  S[3] = mu * (a[0] * n[1] + a[1] * n[0]);
  S[4] = mu * (a[0] * n[2] + a[2] * n[0]);
  S[5] = mu * (a[1] * n[2] + a[2] * n[1]);
}

// The dual-consistency stress S of a boundary face point, the out.S of
// face_point_physics(false, ...): tau(dm (x) n) with dm = (m jump - u_L rho
// jump)/rho_L and the jump UL - Ub to the boundary state (no 1/2).
__device__ __forceinline__ void boundary_dc_stress(const double* UL, const double* Ub,
                                                   const double* n, double mu, double* S)
{
  const double irL = rcp(UL[0]);
  const double jr  = UL[0] - Ub[0];
  double a[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    a[i] = ((UL[1 + i] - Ub[1 + i]) - jr * (UL[1 + i] * irL)) * irL;
  const double lt = -(2.0 / 3.0) * mu * (a[0] * n[0] + a[1] * n[1] + a[2] * n[2]);
  S[0] = 2.0 * mu * a[0] * n[0] + lt;
  S[1] = 2.0 * mu * a[1] * n[1] + lt;
  S[2] = 2.0 * mu * a[2] * n[2] + lt;
  S[3] = mu * (a[0] * n[1] + a[1] * n[0]);
  S[4] = mu * (a[0] * n[2] + a[2] * n[0]);
  S[5] = mu * (a[1] * n[2] + a[2] * n[1]);
}

// Volume flux H(Vq, gr) (dgx_physics.hpp): gx[0..4] = d/dVq, gx[5 + 4c + r]
// = d/dgr[c][r] of sum psiH . H, in reverse mode.  Forward: u = m/rho,
// dv[c][i] = (gr[c][1+i] - gr[c][0] u_i)/rho, Lv[i][d] = sum_c dv[c][i] Ji[c][d],
// tau[d][i] = mu (Lv[d][i] + Lv[i][d]) - 2/3 mu tr(Lv) delta_di,
// p = (gamma-1)(E - rho |u|^2/2), Ht = (E + p)/rho,
// H[d][0] = -m_d, H[d][1+i] = tau[d][i] - m_d u_i - p delta_di,
// H[d][4] = u . tau[d] - m_d Ht.
__device__ __forceinline__ void volume_flux_vjp(const double* Vq, const double gr[3][4],
                                                const double Ji[3][3], double gamma, double mu,
                                                const double psiH[3][5], double* gx)
{
  const double gm1 = gamma - 1.0;
  const double ir  = rcp(Vq[0]);
  const double u[3] = {Vq[1] * ir, Vq[2] * ir, Vq[3] * ir};
  double dvr[3][3];  // rho dv
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      dvr[c][i] = gr[c][1 + i] - gr[c][0] * u[i];
  double Lv[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      Lv[i][d] = ir * (dvr[0][i] * Ji[0][d] + dvr[1][i] * Ji[1][d] + dvr[2][i] * Ji[2][d]);
  const double q  = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double p  = gm1 * (Vq[4] - 0.5 * Vq[0] * q);
  const double Ht = (Vq[4] + p) * ir;
  const double lt = -(2.0 / 3.0) * mu * (Lv[0][0] + Lv[1][1] + Lv[2][2]);

  // backward through H: bt = adjoint of tau
  double bm[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0}, bt[3][3];
  double bp = 0.0, bHt = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    const double md = Vq[1 + d];
    bm[d] -= psiH[d][0] + psiH[d][4] * Ht;
    bHt -= psiH[d][4] * md;
    bp -= psiH[d][1 + d];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      const double tdi = mu * (Lv[d][i] + Lv[i][d]) + ((i == d) ? lt : 0.0);
      bt[d][i] = psiH[d][1 + i] + psiH[d][4] * u[i];
      bm[d] -= psiH[d][1 + i] * u[i];
      bu[i] += psiH[d][4] * tdi - psiH[d][1 + i] * md;
    }
  }
  // tau -> Lv -> dv; the gradient outputs and the state part through u, 1/rho
  const double btr = -(2.0 / 3.0) * mu * (bt[0][0] + bt[1][1] + bt[2][2]);
  double bL[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      bL[a][b] = mu * (bt[a][b] + bt[b][a]) + ((a == b) ? btr : 0.0);
  double bir = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    double s0 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      const double pdv = bL[i][0] * Ji[c][0] + bL[i][1] * Ji[c][1] + bL[i][2] * Ji[c][2];
      gx[5 + 4 * c + 1 + i] = pdv * ir;
      s0 += pdv * u[i];
      bu[i] -= pdv * gr[c][0] * ir;
      bir += pdv * dvr[c][i];
    }
    gx[5 + 4 * c] = -s0 * ir;
  }
  // Ht = (E + p)/rho, p = (gamma-1)(E - rho |u|^2 / 2), u = m/rho
  double bE = bHt * ir;
  bp += bHt * ir;
  bir += bHt * (Vq[4] + p);
  bE += gm1 * bp;
  double brho = -0.5 * gm1 * bp * q;
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    bu[i] -= gm1 * bp * Vq[0] * u[i];
    bm[i] += bu[i] * ir;
    bir += bu[i] * Vq[1 + i];
  }
  gx[0] = brho - bir * ir * ir;
  gx[1] = bm[0];
  gx[2] = bm[1];
  gx[3] = bm[2];
  gx[4] = bE;
}

// Interior face-point physics: gx[0..4] d/dUL, [5..9] d/dUR, [10..12] d/dvgL,
// [13..15] d/dvgR of psiG . G + psiS . S, in reverse mode: the transposed
// Roe flux Jacobians (roe_flux_vjp) plus the backward sweep of the viscous,
// BR2 and dual-consistency terms of face_point_physics (interior: h = 1/2 on
// both tractions, c = eta lam / 4 per side, s = 1/2).
__device__ __forceinline__ void face_physics_vjp(const double* UL, const double* UR,
                                                 const double* vgL, const double* vgR,
                                                 const double* n, double lamL, double lamR,
                                                 const Phys& ph, const double* psiG,
                                                 const double* psiS, double* gx)
{
  roe_flux_vjp(UL, UR, n, ph.gamma, psiG, gx, gx + 5);
  const double mu = ph.mu;
  const double cL = 0.25 * ph.eta * lamL, cB = 0.25 * ph.eta * lamR, s = 0.5;
  // forward values
  const double irL = rcp(UL[0]), irB = rcp(UR[0]);
  const double uL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
  const double uB[3] = {UR[1] * irB, UR[2] * irB, UR[3] * irB};
  const double jr = UL[0] - UR[0];
  double jm[3], dmL[3], dmB[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    jm[i]  = UL[1 + i] - UR[1 + i];
    dmL[i] = (jm[i] - jr * uL[i]) * irL;
    dmB[i] = (jm[i] - jr * uB[i]) * irB;
  }
  const double dnL = dmL[0] * n[0] + dmL[1] * n[1] + dmL[2] * n[2];
  const double dnB = dmB[0] * n[0] + dmB[1] * n[1] + dmB[2] * n[2];
  double w[3], w2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    w[i]  = 0.5 * vgL[i] - cL * mu * (dmL[i] + (1.0 / 3.0) * dnL * n[i]);
    w2[i] = 0.5 * vgR[i] - cB * mu * (dmB[i] + (1.0 / 3.0) * dnB * n[i]);
  }
  // G[1+i] -= w + w2,  G[4] -= uL . w + uB . w2
  double buL[3], buB[3], bw[3], bw2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    buL[i] = -psiG[4] * w[i];
    buB[i] = -psiG[4] * w2[i];
    bw[i]  = -psiG[1 + i] - psiG[4] * uL[i];
    bw2[i] = -psiG[1 + i] - psiG[4] * uB[i];
    gx[10 + i] = 0.5 * bw[i];
    gx[13 + i] = 0.5 * bw2[i];
  }
  // w = vg/2 - c mu (dm + dn n / 3)
  const double bwn  = bw[0] * n[0] + bw[1] * n[1] + bw[2] * n[2];
  const double bw2n = bw2[0] * n[0] + bw2[1] * n[1] + bw2[2] * n[2];
  double bdnL = -cL * mu * bwn * (1.0 / 3.0), bdnB = -cB * mu * bw2n * (1.0 / 3.0);
  double bdmL[3], bdmB[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    bdmL[i] = -cL * mu * bw[i];
    bdmB[i] = -cB * mu * bw2[i];
  }
  // S = tau(s dmL (x) n): a = s dmL, lt = -2/3 mu s dnL
  const double* pS = psiS;
  bdmL[0] += s * mu * (2.0 * pS[0] * n[0] + pS[3] * n[1] + pS[4] * n[2]);
  bdmL[1] += s * mu * (2.0 * pS[1] * n[1] + pS[3] * n[0] + pS[5] * n[2]);
  bdmL[2] += s * mu * (2.0 * pS[2] * n[2] + pS[4] * n[0] + pS[5] * n[1]);
  bdnL -= (2.0 / 3.0) * mu * s * (pS[0] + pS[1] + pS[2]);
  // dn = dm . n;  dm = (jm - jr u) ir
  double bjr = 0.0, birL = 0.0, birB = 0.0, bjm[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    bdmL[i] += bdnL * n[i];
    bdmB[i] += bdnB * n[i];
    bjm[i] = bdmL[i] * irL + bdmB[i] * irB;
    bjr -= bdmL[i] * uL[i] * irL + bdmB[i] * uB[i] * irB;
    buL[i] -= bdmL[i] * jr * irL;
    buB[i] -= bdmB[i] * jr * irB;
    birL += bdmL[i] * (jm[i] - jr * uL[i]);
    birB += bdmB[i] * (jm[i] - jr * uB[i]);
  }
  // u = m ir, ir = 1/rho; jm, jr = UL - UR
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    birL += buL[i] * UL[1 + i];
    birB += buB[i] * UR[1 + i];
    gx[1 + i] += bjm[i] + buL[i] * irL;
    gx[6 + i] += -bjm[i] + buB[i] * irB;
  }
  gx[0] += bjr - birL * irL * irL;
  gx[5] += -bjr - birB * irB * irB;
}

// Gradient traction vg = tau(grad u).n: gx[4c + r] = d/dgr[c][r],
// gx[12 + r] = d/dU[r] (r < 4) of psivg . vg.
__device__ __forceinline__ void traction_vjp(const double gr[3][4], const double* U,
                                             const double Ji[3][3], const double* n, double mu,
                                             const double* psivg, double* gx)
{
  // vg = mu (Ln + LTn) - 2/3 mu tr n with Ln[i] = sum_c dv_c[i] m_c,
  // LTn[d] = sum_c Ji[c][d] (dv_c . n), tr = sum_c dv_c . Ji[c], m_c = Ji[c] . n,
  // dv_c[i] = (gr[c][1+i] - gr[c][0] u_i) / rho: reverse sweep, pdv = adjoint
  // of dv_c, then through u = m/rho and 1/rho
  const double ir = rcp(U[0]);
  const double u[3] = {U[1] * ir, U[2] * ir, U[3] * ir};
  const double pn  = psivg[0] * n[0] + psivg[1] * n[1] + psivg[2] * n[2];
  const double ptr = -(2.0 / 3.0) * mu * pn;
  double bu[3] = {0.0, 0.0, 0.0}, bir = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    const double mc = Ji[c][0] * n[0] + Ji[c][1] * n[1] + Ji[c][2] * n[2];
    const double kc = mu * (psivg[0] * Ji[c][0] + psivg[1] * Ji[c][1] + psivg[2] * Ji[c][2]);
    double s0 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      const double pdv = mu * psivg[i] * mc + kc * n[i] + ptr * Ji[c][i];
      gx[4 * c + 1 + i] = pdv * ir;
      s0 += pdv * u[i];
      bu[i] -= pdv * gr[c][0] * ir;
      bir += pdv * (gr[c][1 + i] - gr[c][0] * u[i]);
    }
    gx[4 * c] = -s0 * ir;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
  {
    gx[13 + i] = bu[i] * ir;
    bir += bu[i] * U[1 + i];
  }
  gx[12] = -bir * ir * ir;
}

}  // namespace dgx
