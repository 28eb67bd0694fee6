// Checks each reverse sweep of libdgx's discrete adjoint against forward-mode
// dual numbers run through the forward physics it transposes, on random
// states (including both entropy-fix branches of the Roe flux):
//   roe_flux_vjp        vs  roe_flux<Dual<10>>
//   face_physics_vjp    vs  face_point_physics<Dual<10>> (+ exact linear
//                           differences for the tractions)
//   volume_flux_vjp     vs  volume_flux<Dual<5>> (+ exact linear differences
//                           for the reference-space gradients)
//   traction_vjp        vs  grad_traction<Dual<4>> (+ the same for gradients)
//   interior_dc_stress  vs  the S output of face_point_physics
// Errors are relative to the largest entry of each gradient; the program
// prints the worst per sweep and exits non-zero above 1e-12.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <random>

#include "dgx_dual.hpp"
#include "dgx_vjp.hpp"

using namespace dgx;

static std::mt19937_64 rng(2024);
static double U1() { return std::uniform_real_distribution<double>(-1.0, 1.0)(rng); }

// conserved state of density ~1, sound speed c0 and velocity of the given
// kind: 0 slow (|u.n| below the entropy-fix threshold), 1 fast, 2 moderate,
// 3 near sonic along x
static void state(double* W, double c0, int kind)
{
  double rho = 1.0 + 0.3 * U1(), vel[3];
  double scale = kind == 0 ? 0.01 : (kind == 1 ? 1.0 : 0.3);
  for (int k = 0; k < 3; ++k)
    vel[k] = scale * U1();
  if (kind == 3)
    vel[0] = c0 * (1.0 + 0.02 * U1());
  double p = rho * c0 * c0 / 1.4;
  W[0] = rho;
  for (int k = 0; k < 3; ++k)
    W[1 + k] = rho * vel[k];
  W[4] = p / 0.4 + 0.5 * rho * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
}

static void unit(double* n)
{
  double s = 0.0;
  for (int k = 0; k < 3; ++k)
  {
    n[k] = U1();
    s += n[k] * n[k];
  }
  for (int k = 0; k < 3; ++k)
    n[k] /= std::sqrt(s);
}

static double rel_worst(const double* got, const double* ref, int m)
{
  double scale = 0.0, w = 0.0;
  for (int j = 0; j < m; ++j)
    scale = std::fmax(scale, std::fabs(ref[j]));
  for (int j = 0; j < m; ++j)
    w = std::fmax(w, std::fabs(got[j] - ref[j]) / scale);
  return w;
}

int main()
{
  const int T = 20000;
  double w_roe = 0.0, w_face = 0.0, w_vol = 0.0, w_tr = 0.0;
  Phys ph{};
  ph.gamma = 1.4;
  ph.mu    = 2e-3;
  ph.eta   = 6.0;
  for (int t = 0; t < T; ++t)
  {
    const int kind = t % 4;
    double UL[5], UR[5], n[3], psiG[5], psiS[6], vgL[3], vgR[3];
    const double c0 = 1.0 + 0.5 * U1();
    state(UL, c0, kind);
    state(UR, c0, kind);
    unit(n);
    if (kind == 3)
      n[0] = 1.0, n[1] = n[2] = 0.0;
    for (int r = 0; r < 5; ++r)
      psiG[r] = U1();
    for (int k = 0; k < 6; ++k)
      psiS[k] = U1();
    for (int k = 0; k < 3; ++k)
      vgL[k] = 0.1 * U1(), vgR[k] = 0.1 * U1();
    const double lamL = 1.0 + 0.5 * U1(), lamR = 1.0 + 0.5 * U1();

    using D10 = Dual<10>;
    D10 ul[5], ur[5];
    for (int r = 0; r < 5; ++r)
    {
      ul[r] = D10(UL[r]);
      ur[r] = D10(UR[r]);
      ul[r].d[r] = 1.0;
      ur[r].d[5 + r] = 1.0;
    }
    // Roe flux
    {
      double gl[10];
      roe_flux_vjp(UL, UR, n, ph.gamma, psiG, gl, gl + 5);
      D10 f[5];
      roe_flux(ul, ur, n, ph.gamma, f);
      double ref[10];
      for (int j = 0; j < 10; ++j)
      {
        ref[j] = 0.0;
        for (int r = 0; r < 5; ++r)
          ref[j] += psiG[r] * f[r].d[j];
      }
      w_roe = std::fmax(w_roe, rel_worst(gl, ref, 10));
    }
    // interior face point
    {
      double gx[16], ref[16];
      face_physics_vjp(UL, UR, vgL, vgR, n, lamL, lamR, ph, psiG, psiS, gx);
      FacePointT<D10> fo;
      face_point_physics(true, ul, ur, vgL, vgR, n, lamL, lamR, ph, fo);
      for (int j = 0; j < 10; ++j)
      {
        ref[j] = 0.0;
        for (int r = 0; r < 5; ++r)
          ref[j] += psiG[r] * fo.G[r].d[j];
        for (int k = 0; k < 6; ++k)
          ref[j] += psiS[k] * fo.S[k].d[j];
      }
      for (int j = 0; j < 6; ++j)  // the tractions enter linearly
      {
        double a[3] = {vgL[0], vgL[1], vgL[2]}, b[3] = {vgR[0], vgR[1], vgR[2]};
        FacePoint f0, f1;
        face_point_physics(true, UL, UR, a, b, n, lamL, lamR, ph, f0);
        if (j < 3)
          a[j] += 1.0;
        else
          b[j - 3] += 1.0;
        face_point_physics(true, UL, UR, a, b, n, lamL, lamR, ph, f1);
        ref[10 + j] = 0.0;
        for (int r = 0; r < 5; ++r)
          ref[10 + j] += psiG[r] * (f1.G[r] - f0.G[r]);
        for (int k = 0; k < 6; ++k)
          ref[10 + j] += psiS[k] * (f1.S[k] - f0.S[k]);
      }
      w_face = std::fmax(w_face, rel_worst(gx, ref, 16));
      // the adjoint's own dual-consistency stress equals face_point_physics's
      FacePoint fp;
      face_point_physics(true, UL, UR, vgL, vgR, n, lamL, lamR, ph, fp);
      double S[6];
      interior_dc_stress(UL, UR, n, ph.mu, S);
      w_face = std::fmax(w_face, rel_worst(S, fp.S, 6));
    }
    // volume flux and gradient traction at a node
    {
      double gr[3][4], Ji[3][3], psiH[3][5], psiv[3];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 4; ++r)
          gr[c][r] = U1();
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d)
          Ji[c][d] = (c == d ? 2.0 : 0.0) + 0.3 * U1();
      for (int d = 0; d < 3; ++d)
        for (int r = 0; r < 5; ++r)
          psiH[d][r] = U1();
      for (int k = 0; k < 3; ++k)
        psiv[k] = U1();
      const double* V = UL;

      double gx[17], ref[17];
      volume_flux_vjp(V, gr, Ji, ph.gamma, ph.mu, psiH, gx);
      using D5 = Dual<5>;
      D5 vq[5], H[3][5];
      for (int r = 0; r < 5; ++r)
      {
        vq[r] = D5(V[r]);
        vq[r].d[r] = 1.0;
      }
      volume_flux(vq, gr, Ji, ph.gamma, ph.mu, H);
      for (int j = 0; j < 5; ++j)
      {
        ref[j] = 0.0;
        for (int d = 0; d < 3; ++d)
          for (int r = 0; r < 5; ++r)
            ref[j] += psiH[d][r] * H[d][r].d[j];
      }
      double gt[16], reft[16];
      traction_vjp(gr, V, Ji, n, ph.mu, psiv, gt);
      using D4 = Dual<4>;
      D4 u4[5], vg[3];
      for (int r = 0; r < 5; ++r)
      {
        u4[r] = D4(V[r]);
        if (r < 4)
          u4[r].d[r] = 1.0;
      }
      grad_traction(gr, u4, Ji, n, ph.mu, vg);
      for (int j = 0; j < 4; ++j)
        reft[12 + j] = psiv[0] * vg[0].d[j] + psiv[1] * vg[1].d[j] + psiv[2] * vg[2].d[j];
      for (int c = 0; c < 3; ++c)  // both are linear in the gradients
        for (int r = 0; r < 4; ++r)
        {
          double z[3][4] = {}, g1[3][4] = {};
          g1[c][r] = 1.0;
          double H0[3][5], H1[3][5], v0[3], v1[3];
          volume_flux(V, z, Ji, ph.gamma, ph.mu, H0);
          volume_flux(V, g1, Ji, ph.gamma, ph.mu, H1);
          grad_traction(z, V, Ji, n, ph.mu, v0);
          grad_traction(g1, V, Ji, n, ph.mu, v1);
          double s = 0.0;
          for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 5; ++k)
              s += psiH[d][k] * (H1[d][k] - H0[d][k]);
          ref[5 + 4 * c + r] = s;
          reft[4 * c + r] = psiv[0] * (v1[0] - v0[0]) + psiv[1] * (v1[1] - v0[1]) +
                            psiv[2] * (v1[2] - v0[2]);
        }
      w_vol = std::fmax(w_vol, rel_worst(gx, ref, 17));
      w_tr  = std::fmax(w_tr, rel_worst(gt, reft, 16));
    }
  }
  std::printf("roe_flux_vjp %.3e\nface_physics_vjp %.3e\nvolume_flux_vjp %.3e\ntraction_vjp %.3e\n",
              w_roe, w_face, w_vol, w_tr);
  const double tol = 1e-12;
  return (w_roe < tol && w_face < tol && w_vol < tol && w_tr < tol) ? 0 : 1;
}
