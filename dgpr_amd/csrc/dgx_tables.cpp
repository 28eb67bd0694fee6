// dgx_tables.cpp -- 1D basis/quadrature tables for the device kernels.
//
// The reference evaluates equispaced Lagrange polynomials on [0,1]
// (basis.cpp:64-170) at a (p+1)-point Gauss-Legendre rule on [0,1]
// (quadrature.cpp:84-295).  This file computes the same objects in long
// double and rounds them once to double:
//   * Gauss-Legendre nodes by Newton iteration on P_n from Chebyshev starting
//     points (the textbook method), weights 2/((1-x^2) P_n'(x)^2) mapped to
//     [0,1];
//   * the Gauss-node differentiation matrix by the barycentric formula of
//     Berrut & Trefethen, "Barycentric Lagrange interpolation", SIAM Review
//     46 (2004), eq. (9.4)-(9.5).
#include <cmath>
#include <cstring>

#include "dgx_launch.hpp"

namespace dgx
{
namespace host
{

typedef long double ld;

static void gauss_legendre01(int n, ld* x, ld* w)
{
  const ld pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; ++j)
  {
    // root j (ascending) of P_n on [-1,1]
    ld t = -cosl(pi * (j + 0.75L) / (n + 0.5L));
    ld dp = 1.0L;
    for (int it = 0; it < 100; ++it)
    {
      ld p0 = 1.0L, p1 = t;
      for (int k = 2; k <= n; ++k)
      {
        ld p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1)
      {
        p1 = t;
        p0 = 1.0L;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0L);
      ld dt = p1 / dp;
      t -= dt;
      if (fabsl(dt) < 1e-19L)
        break;
    }
    // recompute derivative at the converged root
    ld p0 = 1.0L, p1 = t;
    for (int k = 2; k <= n; ++k)
    {
      ld p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (t * p1 - p0) / (t * t - 1.0L);
    x[j] = 0.5L * (t + 1.0L);
    w[j] = 1.0L / ((1.0L - t * t) * dp * dp);  // 2/(...) scaled by 1/2
  }
}

// equispaced Lagrange basis of order p on [0,1]: value and derivative
static ld lag(int i, int p, ld x)
{
  if (p == 0)
    return 1.0L;
  ld xi = (ld)i / p, v = 1.0L;
  for (int j = 0; j <= p; ++j)
    if (j != i)
      v *= (x - (ld)j / p) / (xi - (ld)j / p);
  return v;
}

static ld lagd(int i, int p, ld x)
{
  if (p == 0)
    return 0.0L;
  ld xi = (ld)i / p, s = 0.0L;
  for (int j = 0; j <= p; ++j)
  {
    if (j == i)
      continue;
    ld m = 1.0L / (xi - (ld)j / p);
    for (int k = 0; k <= p; ++k)
      if (k != i && k != j)
        m *= (x - (ld)k / p) / (xi - (ld)k / p);
    s += m;
  }
  return s;
}

// Lagrange polynomial on arbitrary nodes z[0..n-1]
static ld lagz(int i, int n, const ld* z, ld x)
{
  ld v = 1.0L;
  for (int j = 0; j < n; ++j)
    if (j != i)
      v *= (x - z[j]) / (z[i] - z[j]);
  return v;
}

static ld lagzd(int i, int n, const ld* z, ld x)
{
  ld s = 0.0L;
  for (int j = 0; j < n; ++j)
  {
    if (j == i)
      continue;
    ld m = 1.0L / (z[i] - z[j]);
    for (int k = 0; k < n; ++k)
      if (k != i && k != j)
        m *= (x - z[k]) / (z[i] - z[k]);
    s += m;
  }
  return s;
}

static bool invert_small(int n, const ld* A, ld* Ai)
{
  ld M[GMAXN][2 * GMAXN];
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < 2 * n; ++c)
      M[r][c] = (c < n) ? A[r * n + c] : ((c - n == r) ? 1.0L : 0.0L);
  for (int c = 0; c < n; ++c)
  {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabsl(M[r][c]) > fabsl(M[piv][c]))
        piv = r;
    if (M[piv][c] == 0.0L)
      return false;
    if (piv != c)
      for (int k = 0; k < 2 * n; ++k)
      {
        ld t = M[c][k];
        M[c][k] = M[piv][k];
        M[piv][k] = t;
      }
    ld d = M[c][c];
    for (int k = 0; k < 2 * n; ++k)
      M[c][k] /= d;
    for (int r = 0; r < n; ++r)
      if (r != c)
      {
        ld f = M[r][c];
        for (int k = 0; k < 2 * n; ++k)
          M[r][k] -= f * M[c][k];
      }
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c)
      Ai[r * n + c] = M[r][n + c];
  return true;
}

// Solution-order tables of order p (PTab fields; T is PTab or GenTab).
template <typename T>
static void fill_p(int p, T& tab)
{
  const int n = p + 1;
  ld g[GMAXN], w[GMAXN];
  gauss_legendre01(n, g, w);
  ld B[GMAXN * GMAXN], Bi[GMAXN * GMAXN];
  for (int i = 0; i < n; ++i)
  {
    tab.g[i] = (double)g[i];
    tab.w[i] = (double)w[i];
    for (int m = 0; m < n; ++m)
      B[i * n + m] = lag(m, p, g[i]);
  }
  invert_small(n, B, Bi);
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m)
    {
      tab.B[i][m]      = (double)B[i * n + m];
      tab.Binv[i][m]   = (double)Bi[i * n + m];
      tab.Blo[i][m]    = (double)(B[i * n + m] - (ld)tab.B[i][m]);
      tab.Binvlo[i][m] = (double)(Bi[i * n + m] - (ld)tab.Binv[i][m]);
    }
  // barycentric differentiation matrix on the Gauss nodes
  ld lam[GMAXN];
  for (int j = 0; j < n; ++j)
  {
    ld pr = 1.0L;
    for (int k = 0; k < n; ++k)
      if (k != j)
        pr *= (g[j] - g[k]);
    lam[j] = 1.0L / pr;
  }
  for (int i = 0; i < n; ++i)
  {
    ld diag = 0.0L;
    for (int m = 0; m < n; ++m)
    {
      if (m == i)
        continue;
      ld d = (lam[m] / lam[i]) / (g[i] - g[m]);
      tab.Dg[i][m] = (double)d;
      diag -= d;
    }
    tab.Dg[i][i] = (double)diag;
  }
  for (int e = 0; e < 2; ++e)
    for (int m = 0; m < n; ++m)
    {
      tab.L[e][m]  = (double)lagz(m, n, g, (ld)e);
      tab.Ld[e][m] = (double)lagzd(m, n, g, (ld)e);
    }
}

// Geometry-order tables of order q at the Gauss points of order p (QTab
// fields; T is QTab or GenTab).
template <typename T>
static void fill_q(int p, int q, T& tab)
{
  const int n = p + 1;
  ld g[GMAXN], w[GMAXN];
  gauss_legendre01(n, g, w);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c <= q; ++c)
    {
      tab.Bq[i][c]  = (double)lag(c, q, g[i]);
      tab.Bqd[i][c] = (double)lagd(c, q, g[i]);
    }
  for (int e = 0; e < 2; ++e)
    for (int c = 0; c <= q; ++c)
    {
      tab.Be[e][c]  = (double)lag(c, q, (ld)e);
      tab.Bde[e][c] = (double)lagd(c, q, (ld)e);
    }
}

// Fills the three P tables (p = 1..3) and nine (P, Q) tables.
void build_tables(PTab* pt, QTab* qt)
{
  memset(pt, 0, sizeof(PTab) * 3);
  memset(qt, 0, sizeof(QTab) * 9);
  for (int p = 1; p <= 3; ++p)
  {
    fill_p(p, pt[p - 1]);
    for (int q = 1; q <= 3; ++q)
      fill_q(p, q, qt[(p - 1) * 3 + (q - 1)]);
  }
}

// Tables of the general-order kernels: every 0 <= p <= 8, 1 <= q <= 8.
void build_gen_tables(GenTab* gt)
{
  memset(gt, 0, sizeof(GenTab) * 72);
  for (int p = 0; p <= 8; ++p)
    for (int q = 1; q <= 8; ++q)
    {
      GenTab& t = gt[p * 8 + q - 1];
      fill_p(p, t);
      fill_q(p, q, t);
    }
}

}  // namespace host
}  // namespace dgx
