// dgx_device.hpp -- device-side tables and pointwise physics for the gfx950
// DG residual kernels.
//
// Pointwise physics restates DGPR's flux.cpp and bound_states.cpp in FP64:
//   euler_flux      <- analytical_flux (flux.cpp:24-48)
//   roe_flux        <- roe             (flux.cpp:56-154)
//   visc            <- A               (flux.cpp:156-202), split into a
//                      state-dependent setup and a gradient-linear apply
//   boundary_state  <- bound_states.cpp:24-130
#pragma once

#include <hip/hip_runtime.h>

#include "dgx_launch.hpp"

namespace dgx
{

// 1/x for finite x > 0 well inside the normal range (densities, surface and
// volume elements): the hardware estimate v_rcp_f64 (relative error below
// 2^-22) refined by two Newton-Raphson steps r <- r + r (1 - x r), which
// leaves an error within one ulp instead of the IEEE division sequence.
__device__ __forceinline__ double rcp(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Euler flux F[d][r] of conserved state U.
__device__ __forceinline__ void euler_flux(const double* U, double gamma,
                                           double F[3][5])
{
  double rho = U[0];
  double ir  = rcp(rho);
  double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  double p = (gamma - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w));
  double H = (U[4] + p) * ir;
  double m[3] = {U[1], U[2], U[3]};
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    F[d][0] = m[d];
    F[d][1] = m[d] * u;
    F[d][2] = m[d] * v;
    F[d][3] = m[d] * w;
    F[d][1 + d] += p;
    F[d][4] = m[d] * H;
  }
}

// Normal Euler flux F(U).n
template <typename T>
__device__ __forceinline__ void euler_flux_n(const T* U, double gamma, const double* n, T* Fn)
{
  T rho = U[0];
  T ir  = rcp(rho);
  T u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  T p  = (gamma - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w));
  T un = u * n[0] + v * n[1] + w * n[2];
  T mn = rho * un;
  Fn[0] = mn;
  Fn[1] = mn * u + p * n[0];
  Fn[2] = mn * v + p * n[1];
  Fn[3] = mn * w + p * n[2];
  Fn[4] = un * (U[4] + p);
}

// Roe flux with |lambda| entropy fix below 5% of the Roe sound speed.
template <typename T>
__device__ __forceinline__ void roe_flux(const T* Ul, const T* Ur, const double* n, double g,
                                         T* f)
{
  double gm1 = g - 1.0;
  T irl = rcp(Ul[0]), irr = rcp(Ur[0]);
  T ul = Ul[1] * irl, vl = Ul[2] * irl, wl = Ul[3] * irl;
  T ur = Ur[1] * irr, vr = Ur[2] * irr, wr = Ur[3] * irr;
  T pl = gm1 * (Ul[4] - 0.5 * Ul[0] * (ul * ul + vl * vl + wl * wl));
  T pr = gm1 * (Ur[4] - 0.5 * Ur[0] * (ur * ur + vr * vr + wr * wr));
  T Hl = (Ul[4] + pl) * irl;
  T Hr = (Ur[4] + pr) * irr;
  T sl = sqrt(Ul[0]), sr = sqrt(Ur[0]);
  T iden = rcp(sl + sr);
  T u = (sl * ul + sr * ur) * iden;
  T v = (sl * vl + sr * vr) * iden;
  T w = (sl * wl + sr * wr) * iden;
  T H = (sl * Hl + sr * Hr) * iden;
  T dr = Ur[0] - Ul[0], dxm = Ur[1] - Ul[1], dym = Ur[2] - Ul[2];
  T dzm = Ur[3] - Ul[3], drE = Ur[4] - Ul[4];
  T qsq = u * u + v * v + w * w;
  T c   = sqrt(gm1 * (H - 0.5 * qsq));
  T ic  = rcp(c);
  T un  = u * n[0] + v * n[1] + w * n[2];
  T l1 = un + c, l2 = un - c, l3 = un;
  T al1 = fabs(l1), al2 = fabs(l2), al3 = fabs(l3);
  T eps  = 0.05 * c;
  T i2e  = rcp(2.0 * eps);
  if (al1 < eps) al1 = fabs((eps * eps + l1 * l1) * i2e);
  if (al2 < eps) al2 = fabs((eps * eps + l2 * l2) * i2e);
  if (al3 < eps) al3 = fabs((eps * eps + l3 * l3) * i2e);
  T s1 = 0.5 * (al1 + al2);
  T s2 = 0.5 * (al1 - al2);
  T G1 = gm1 * (0.5 * qsq * dr - (u * dxm + v * dym + w * dzm) + drE);
  T G2 = -un * dr + (dxm * n[0] + dym * n[1] + dzm * n[2]);
  T C1 = (G1 * ic * ic) * (s1 - al3) + (G2 * ic) * s2;
  T C2 = (G1 * ic) * s2 + (s1 - al3) * G2;
  T unl = ul * n[0] + vl * n[1] + wl * n[2];
  T unr = ur * n[0] + vr * n[1] + wr * n[2];
  T mnl = Ul[0] * unl, mnr = Ur[0] * unr;
  f[0] = 0.5 * (mnl + mnr) - 0.5 * (al3 * dr + C1);
  f[1] = 0.5 * ((mnl * ul + pl * n[0]) + (mnr * ur + pr * n[0])) -
         0.5 * (al3 * dxm + C1 * u + C2 * n[0]);
  f[2] = 0.5 * ((mnl * vl + pl * n[1]) + (mnr * vr + pr * n[1])) -
         0.5 * (al3 * dym + C1 * v + C2 * n[1]);
  f[3] = 0.5 * ((mnl * wl + pl * n[2]) + (mnr * wr + pr * n[2])) -
         0.5 * (al3 * dzm + C1 * w + C2 * n[2]);
  f[4] = 0.5 * ((mnl * Hl) + (mnr * Hr)) - 0.5 * (al3 * drE + C1 * H + C2 * un);
}

// Reverse mode of roe_flux: gl = d(psi . f)/dUl and gr = d(psi . f)/dUr,
// the transposed Roe flux Jacobians applied to psi.  One forward sweep keeps
// the intermediates, one backward sweep accumulates their adjoints (bX is
// the adjoint of X); the entropy fix is differentiated on the branch the
// forward sweep takes, and |l| with derivative +1 at 0, as the dual numbers
// do.
__device__ __forceinline__ void roe_flux_vjp(const double* Ul, const double* Ur, const double* n,
                                             double g, const double* psi, double* gl,
                                             double* gr)
{
  // ---- forward sweep ----
  const double gm1 = g - 1.0;
  const double irl = rcp(Ul[0]), irr = rcp(Ur[0]);
  const double vl[3] = {Ul[1] * irl, Ul[2] * irl, Ul[3] * irl};
  const double vr[3] = {Ur[1] * irr, Ur[2] * irr, Ur[3] * irr};
  const double ql = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
  const double qr = vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
  const double pl = gm1 * (Ul[4] - 0.5 * Ul[0] * ql);
  const double pr = gm1 * (Ur[4] - 0.5 * Ur[0] * qr);
  const double Hl = (Ul[4] + pl) * irl, Hr = (Ur[4] + pr) * irr;
  const double sl = sqrt(Ul[0]), sr = sqrt(Ur[0]);
  const double iden = rcp(sl + sr);
  double va[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    va[k] = (sl * vl[k] + sr * vr[k]) * iden;
  const double H  = (sl * Hl + sr * Hr) * iden;
  const double dr = Ur[0] - Ul[0], drE = Ur[4] - Ul[4];
  const double dm[3] = {Ur[1] - Ul[1], Ur[2] - Ul[2], Ur[3] - Ul[3]};
  const double qsq = va[0] * va[0] + va[1] * va[1] + va[2] * va[2];
  const double c   = sqrt(gm1 * (H - 0.5 * qsq));
  const double ic  = rcp(c);
  const double un  = va[0] * n[0] + va[1] * n[1] + va[2] * n[2];
  const double lk[3] = {un + c, un - c, un};
  const double eps = 0.05 * c;
  const double i2e = rcp(2.0 * eps);
  double al[3];
  bool fix[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    al[k]  = fabs(lk[k]);
    fix[k] = al[k] < eps;
    if (fix[k])
      al[k] = fabs((eps * eps + lk[k] * lk[k]) * i2e);
  }
  const double s1 = 0.5 * (al[0] + al[1]), s2 = 0.5 * (al[0] - al[1]);
  const double udm = va[0] * dm[0] + va[1] * dm[1] + va[2] * dm[2];
  const double G1 = gm1 * (0.5 * qsq * dr - udm + drE);
  const double G2 = -un * dr + (dm[0] * n[0] + dm[1] * n[1] + dm[2] * n[2]);
  const double s13 = s1 - al[2];
  const double C1 = (G1 * ic * ic) * s13 + (G2 * ic) * s2;
  const double C2 = (G1 * ic) * s2 + s13 * G2;
  const double unl = vl[0] * n[0] + vl[1] * n[1] + vl[2] * n[2];
  const double unr = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
  const double mnl = Ul[0] * unl, mnr = Ur[0] * unr;

  // ---- backward sweep ----
  double bvl[3] = {0.0, 0.0, 0.0}, bvr[3] = {0.0, 0.0, 0.0}, bva[3] = {0.0, 0.0, 0.0};
  double bdm[3] = {0.0, 0.0, 0.0};
  double bmnl = 0.5 * psi[0], bmnr = 0.5 * psi[0];
  double bal3 = -0.5 * psi[0] * dr, bdr = -0.5 * psi[0] * al[2], bC1 = -0.5 * psi[0];
  double bpl = 0.0, bpr = 0.0, bC2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    const double ps = psi[1 + k];
    bmnl += 0.5 * ps * vl[k];
    bvl[k] += 0.5 * ps * mnl;
    bpl += 0.5 * ps * n[k];
    bmnr += 0.5 * ps * vr[k];
    bvr[k] += 0.5 * ps * mnr;
    bpr += 0.5 * ps * n[k];
    bal3 -= 0.5 * ps * dm[k];
    bdm[k] -= 0.5 * ps * al[2];
    bC1 -= 0.5 * ps * va[k];
    bva[k] -= 0.5 * ps * C1;
    bC2 -= 0.5 * ps * n[k];
  }
  const double p4 = psi[4];
  bmnl += 0.5 * p4 * Hl;
  double bHl = 0.5 * p4 * mnl;
  bmnr += 0.5 * p4 * Hr;
  double bHr = 0.5 * p4 * mnr;
  bal3 -= 0.5 * p4 * drE;
  double bdrE = -0.5 * p4 * al[2];
  bC1 -= 0.5 * p4 * H;
  double bH  = -0.5 * p4 * C1;
  bC2 -= 0.5 * p4 * un;
  double bun = -0.5 * p4 * C2;
  // mnl = rhol unl, unl = vl . n (and the right side)
  double brl = bmnl * unl, brr = bmnr * unr;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    bvl[k] += bmnl * Ul[0] * n[k];
    bvr[k] += bmnr * Ur[0] * n[k];
  }
  // C2 = G1 ic s2 + s13 G2,  C1 = G1 ic^2 s13 + G2 ic s2
  double bG1 = bC2 * ic * s2 + bC1 * ic * ic * s13;
  double bG2 = bC2 * s13 + bC1 * ic * s2;
  double bic = bC2 * G1 * s2 + bC1 * (2.0 * G1 * ic * s13 + G2 * s2);
  double bs2 = bC2 * G1 * ic + bC1 * G2 * ic;
  double bs13 = bC2 * G2 + bC1 * G1 * ic * ic;
  double bs1 = bs13;
  bal3 -= bs13;
  // G2 = -un dr + dm . n
  bun -= bG2 * dr;
  bdr -= bG2 * un;
  // G1 = gm1 (qsq dr / 2 - va . dm + drE)
  double bqsq = bG1 * gm1 * 0.5 * dr;
  bdr += bG1 * gm1 * 0.5 * qsq;
  bdrE += bG1 * gm1;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    bdm[k] += bG2 * n[k] - bG1 * gm1 * va[k];
    bva[k] -= bG1 * gm1 * dm[k];
  }
  // s1, s2 from al1, al2; entropy fix; l1 = un + c, l2 = un - c, l3 = un
  const double bal[3] = {0.5 * (bs1 + bs2), 0.5 * (bs1 - bs2), bal3};
  double bc = 0.0, beps = 0.0;
  double bl[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    if (fix[k])
    {
      const double ie = 2.0 * i2e;  // 1 / eps
      bl[k] = bal[k] * lk[k] * ie;
      beps += bal[k] * (0.5 - 0.5 * lk[k] * lk[k] * ie * ie);
    }
    else
      bl[k] = (lk[k] < 0.0) ? -bal[k] : bal[k];
  }
  bun += bl[0] + bl[1] + bl[2];
  bc += bl[0] - bl[1] + 0.05 * beps;
  // ic = 1/c;  c = sqrt(gm1 (H - qsq/2))
  bc -= bic * ic * ic;
  bH += bc * gm1 * 0.5 * ic;
  bqsq -= bc * gm1 * 0.25 * ic;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    bva[k] += bun * n[k] + 2.0 * va[k] * bqsq;
  // Roe averages: va = (sl vl + sr vr) iden, H = (sl Hl + sr Hr) iden
  double bsl = bH * Hl * iden, bsr = bH * Hr * iden;
  double biden = bH * (sl * Hl + sr * Hr);
  bHl += bH * sl * iden;
  bHr += bH * sr * iden;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    bsl += bva[k] * vl[k] * iden;
    bsr += bva[k] * vr[k] * iden;
    biden += bva[k] * (sl * vl[k] + sr * vr[k]);
    bvl[k] += bva[k] * sl * iden;
    bvr[k] += bva[k] * sr * iden;
  }
  bsl -= biden * iden * iden;
  bsr -= biden * iden * iden;
  brl += bsl * 0.5 * rcp(sl);
  brr += bsr * 0.5 * rcp(sr);
  // H_side = (E + p) ir;  p = gm1 (E - rho q / 2)
  double bEl = bHl * irl, bEr = bHr * irr;
  bpl += bHl * irl;
  bpr += bHr * irr;
  double birl = bHl * (Ul[4] + pl), birr = bHr * (Ur[4] + pr);
  bEl += gm1 * bpl;
  bEr += gm1 * bpr;
  brl -= 0.5 * gm1 * bpl * ql;
  brr -= 0.5 * gm1 * bpr * qr;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    bvl[k] -= gm1 * bpl * Ul[0] * vl[k];
    bvr[k] -= gm1 * bpr * Ur[0] * vr[k];
  }
  // v = m ir;  ir = 1/rho
  double bml[3], bmr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    bml[k] = bvl[k] * irl;
    bmr[k] = bvr[k] * irr;
    birl += bvl[k] * Ul[1 + k];
    birr += bvr[k] * Ur[1 + k];
  }
  brl -= birl * irl * irl;
  brr -= birr * irr * irr;
  // differences dr, dm, drE (Ur - Ul)
  gl[0] = brl - bdr;
  gr[0] = brr + bdr;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    gl[1 + k] = bml[k] - bdm[k];
    gr[1 + k] = bmr[k] + bdm[k];
  }
  gl[4] = bEl - bdrE;
  gr[4] = bEr + bdrE;
}

// State-dependent part of the Newtonian viscous flux A(s, grad): the
// velocity and 1/rho.  A is linear in its gradient argument.
struct ViscState
{
  double ir, u, v, w;
};

__device__ __forceinline__ ViscState visc_setup(const double* s)
{
  ViscState vs;
  vs.ir = rcp(s[0]);
  vs.u  = s[1] * vs.ir;
  vs.v  = s[2] * vs.ir;
  vs.w  = s[3] * vs.ir;
  return vs;
}

// Q[d][r] = A(s, G) with conserved-variable gradient G[d][r]; no heat flux,
// lambda = -2/3 mu.
__device__ __forceinline__ void visc_apply(const ViscState& vs, const double G[3][5],
                                           double mu, double Q[3][5])
{
  double lam = -(2.0 / 3.0) * mu;
  double ux = (G[0][1] - G[0][0] * vs.u) * vs.ir;
  double uy = (G[1][1] - G[1][0] * vs.u) * vs.ir;
  double uz = (G[2][1] - G[2][0] * vs.u) * vs.ir;
  double vx = (G[0][2] - G[0][0] * vs.v) * vs.ir;
  double vy = (G[1][2] - G[1][0] * vs.v) * vs.ir;
  double vz = (G[2][2] - G[2][0] * vs.v) * vs.ir;
  double wx = (G[0][3] - G[0][0] * vs.w) * vs.ir;
  double wy = (G[1][3] - G[1][0] * vs.w) * vs.ir;
  double wz = (G[2][3] - G[2][0] * vs.w) * vs.ir;
  double dv  = ux + vy + wz;
  double txx = lam * dv + 2.0 * mu * ux;
  double tyy = lam * dv + 2.0 * mu * vy;
  double tzz = lam * dv + 2.0 * mu * wz;
  double txy = mu * (vx + uy);
  double txz = mu * (uz + wx);
  double tyz = mu * (wy + vz);
  Q[0][0] = 0.0; Q[0][1] = txx; Q[0][2] = txy; Q[0][3] = txz;
  Q[0][4] = vs.u * txx + vs.v * txy + vs.w * txz;
  Q[1][0] = 0.0; Q[1][1] = txy; Q[1][2] = tyy; Q[1][3] = tyz;
  Q[1][4] = vs.u * txy + vs.v * tyy + vs.w * tyz;
  Q[2][0] = 0.0; Q[2][1] = txz; Q[2][2] = tyz; Q[2][3] = tzz;
  Q[2][4] = vs.u * txz + vs.v * tyz + vs.w * tzz;
}

// Normal component A(s, G).n only.
__device__ __forceinline__ void visc_apply_n(const ViscState& vs, const double G[3][5],
                                             double mu, const double* n, double* Qn)
{
  double Q[3][5];
  visc_apply(vs, G, mu, Q);
#pragma unroll
  for (int r = 0; r < 5; ++r)
    Qn[r] = Q[0][r] * n[0] + Q[1][r] * n[1] + Q[2][r] * n[2];
}

// Boundary state Ub from interior trace UL and outward normal n.
// Returns false when the subsonic-inflow Mach quadratic has no
// non-negative root (the reference aborts there on the host).  T is double
// in the residual kernels and a dual number in the adjoint's boundary-face
// kernel (the same expressions, so the double instantiation is unchanged).
template <typename T>
__device__ __forceinline__ bool boundary_state(int bt, const T* UL, const double* n,
                                               const Phys& ph, T* Ub)
{
  double gamma = ph.gamma, gm1 = gamma - 1.0;
  switch (bt)
  {
    case -3:  // freestream
#pragma unroll
      for (int i = 0; i < 5; ++i)
        Ub[i] = ph.Ufr[i];
      return true;
    case -4:  // inviscid wall: remove the normal velocity
    {
      T ir = rcp(UL[0]);
      T up = UL[1] * ir, vp = UL[2] * ir, wp = UL[3] * ir;
      T vdn = up * n[0] + vp * n[1] + wp * n[2];
      Ub[0] = UL[0];
      Ub[1] = UL[0] * (up - vdn * n[0]);
      Ub[2] = UL[0] * (vp - vdn * n[1]);
      Ub[3] = UL[0] * (wp - vdn * n[2]);
      Ub[4] = UL[4];
      return true;
    }
    case -7:  // no-slip wall, energy taken from the interior
      Ub[0] = UL[0]; Ub[1] = 0.0; Ub[2] = 0.0; Ub[3] = 0.0; Ub[4] = UL[4];
      return true;
    case -5:  // subsonic inflow: total T, total P, outgoing Riemann invariant
    {
      double R = ph.Rgas, Tt = ph.Tt, Pt = ph.Pt;
      T ir = rcp(UL[0]);
      T up = UL[1] * ir, vp = UL[2] * ir, wp = UL[3] * ir;
      T Vsq = up * up + vp * vp + wp * wp;
      T pp  = gm1 * (UL[4] - 0.5 * UL[0] * Vsq);
      T unp = up * n[0] + vp * n[1] + wp * n[2];
      T cp  = sqrt(gamma * pp * ir);
      T Jp  = unp + ((2.0 * cp) / gm1);
      T a   = gamma * R * Tt - 0.5 * gm1 * Jp * Jp;
      double b = (4.0 * gamma * R * Tt * -1.0) / gm1;
      T c   = (4.0 * gamma * R * Tt) / (gm1 * gm1) - Jp * Jp;
      T disc = sqrt(b * b - 4.0 * a * c);
      T s0  = (-b + disc) / (2.0 * a);
      T s1  = (-b - disc) / (2.0 * a);
      T Mb  = 0.0;
      bool ok    = true;
      if (s0 >= 0.0 && s1 >= 0.0)
        Mb = (s0 < s1) ? s0 : s1;
      else if (s0 >= 0.0)
        Mb = s0;
      else if (s1 >= 0.0)
        Mb = s1;
      else
        ok = false;
      T Tb   = Tt / (1.0 + 0.5 * gm1 * Mb * Mb);
      T Pb   = Pt * pow(Tb / Tt, gamma / gm1);
      T rhob = Pb / (R * Tb);
      T cb   = sqrt((gamma * Pb) / rhob);
      T ub = Mb * cb * -n[0], vb = Mb * cb * -n[1], wb = Mb * cb * -n[2];
      Ub[0] = rhob; Ub[1] = rhob * ub; Ub[2] = rhob * vb; Ub[3] = rhob * wb;
      Ub[4] = (Pb / gm1) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
      return ok;
    }
    case -6:  // subsonic outflow: fixed static pressure, interior entropy
    {
      double P   = ph.P;
      T rp  = UL[0];
      T ir  = rcp(rp);
      T up = UL[1] * ir, vp = UL[2] * ir, wp = UL[3] * ir;
      T Vsq = up * up + vp * vp + wp * wp;
      T pp  = gm1 * (UL[4] - 0.5 * rp * Vsq);
      T Sp  = pp / pow(rp, gamma);
      T rhob = pow(P / Sp, 1.0 / gamma);
      T cb   = sqrt((gamma * P) / rhob);
      T unp  = up * n[0] + vp * n[1] + wp * n[2];
      T cp   = sqrt(gamma * pp * ir);
      T Jp   = unp + ((2.0 * cp) / gm1);
      T unb  = Jp - ((2.0 * cb) / gm1);
      T ub = up - unp * n[0] + unb * n[0];
      T vb = vp - unp * n[1] + unb * n[1];
      T wb = wp - unp * n[2] + unb * n[2];
      Ub[0] = rhob; Ub[1] = rhob * ub; Ub[2] = rhob * vb; Ub[3] = rhob * wb;
      Ub[4] = (P / gm1) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
      return true;
    }
  }
  return false;
}

}  // namespace dgx
