/*
 * dg_oracle.c -- CPU FP64 restatement of DGPR's reference residual and
 * explicit time stepping, used ONLY as a test checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path
 * (libdgx.so) never links, loads or calls it.
 *
 * It restates, in plain C and in the reference's own dense form, the
 * algorithm of these reference files (cited per function below):
 *   solver/quadrature.cpp      Gauss-Legendre rule (Newton iteration)
 *   solver/basis.cpp           equispaced Lagrange basis and its tables
 *   solver/dgmath.cpp          det3x3, inv3x3, matmul, spmm, symmstore_mvmul,
 *                              Gauss-Jordan invert
 *   solver/geometry_common.cpp per-element mass-matrix inverse, vJ, vgrad,
 *                              fgrad, surface normals
 *   solver/geometry_shared.cpp single-domain face enumeration
 *   solver/flux.cpp            Euler flux, Roe flux, viscous flux A
 *   solver/bound_states.cpp    boundary states
 *   solver/solve.cpp           residual() on the shared face lists
 *   solver/time_distributed.cpp RK4 / TVD-RK3 / residual norm
 *
 * The restatement is pinned against the reference's own outputs: its CPU
 * path compiles from its sources with g++ and the image's MPICH
 * (oracle/build.py builds oracle/ref_harness.cpp and solver/dg.cpp into
 * oracle/_ref where the reference is present), and the residuals, states,
 * norms and driver files it writes are committed as data under
 * tests/golden/; the reference tests' known answers and mathematical
 * invariants (tests/) cover the rest.
 *
 * Storage layouts follow the reference exactly:
 *   state U[e][r][bf]  (geometry_common.cpp:467-470)
 *   nodes [e][(q+1)^3][3], ordered z-y-x (geometry_common.cpp:44-48)
 *   Minv  packed row-major upper triangle per element
 *   vJ [e][q], vgrad [e][q][bf][3], fgrad [e][lf][q][bf][3]
 *   n [face][q][3], fJ [face][q] indexed by the global face number
 */
#include <float.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* Arithmetic type.  The oracle proper computes in double, as the reference
 * does.  Built with -DORC_EXTENDED it evaluates the same algorithm in x87
 * extended precision (long double, 64-bit significand; <tgmath.h> makes
 * sqrt, pow, fabs, ... follow the type), which measures how far each double
 * evaluation of the residual lies from the exact value of the discrete
 * operator: the rounding floor of an ill-conditioned case. */
#ifdef ORC_EXTENDED
typedef long double real;
#define ORC_EPS LDBL_EPSILON
#define ORC_MAXV LDBL_MAX
#define ORC_PI 3.141592653589793238462643383279502884L
#else
typedef double real;
#define ORC_EPS DBL_EPSILON
#define ORC_MAXV DBL_MAX
#define ORC_PI 3.14159265358979323846
#endif


typedef int64_t s64;
typedef uint64_t u64;

/* Zero-initialised allocation: the reference's array<T> value-initialises
 * its storage (data_structures.cpp:67, new T[len]()), and its matmul scales
 * the output by beta = 0 before accumulating (dgmath.cpp:103-122), which
 * relies on that. */
static void* zalloc(size_t n)
{
  return calloc(1, n ? n : 1);
}

/* Planted mistakes, for tests that show an invariant or a golden vector
 * catches a plausible error (never set outside those tests):
 *   DC_SIGN       R-side dual-consistency term with the opposite sign
 *   DROP_RFLUX    R side misses the Roe flux term
 *   RK4_HALF      classic RK4 stage 3 with dt/2 instead of dt
 *   BR2_INT_HALF  interior-face lifting without its factor 1/2
 *   BR2_BND_HALF  boundary-face lifting with a factor 1/2
 *   BR2_WALL_SL   boundary-face lifting with A at the interior trace SL
 *                 instead of the boundary state Sb
 *   BR2_ETA_HALF  interior QH penalty eta*(dL + dR) without the 1/2 */
enum
{
  ORC_PLANT_NONE         = 0,
  ORC_PLANT_DC_SIGN      = 1,
  ORC_PLANT_DROP_RFLUX   = 2,
  ORC_PLANT_RK4_HALF     = 3,
  ORC_PLANT_BR2_INT_HALF = 4,
  ORC_PLANT_BR2_BND_HALF = 5,
  ORC_PLANT_BR2_WALL_SL  = 6,
  ORC_PLANT_BR2_ETA_HALF = 7
};
static int g_plant = ORC_PLANT_NONE;

void orc_set_plant(int k)
{
  g_plant = k;
}

/* btype values, boundaries.cpp:26-35 */
enum
{
  ORC_INTERIOR         = -1,
  ORC_PERIODIC         = -2,
  ORC_FREESTREAM       = -3,
  ORC_INVISCID_WALL    = -4,
  ORC_SUBSONIC_INFLOW  = -5,
  ORC_SUBSONIC_OUTFLOW = -6,
  ORC_NO_SLIP_WALL     = -7
};

/* parameters, solver_params.cpp:22-36 */
typedef struct
{
  real gamma, R, mu, c, T, P;
  real Ufr[5];
  real source[5];
  real eta;
} orc_params;

/* ------------------------------------------------------------------------ */
/* quadrature (quadrature.cpp:84-173, 187-295)                               */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre nodes/weights on [is, ie] with n points: Newton iteration on
 * P_n from Chebyshev starting points, Legendre values by the three-term
 * recurrence (quadrature.cpp:84-173). */
int orc_gauss_legendre(real is, real ie, int n, real* x, real* w)
{
  if (n < 2)
  {
    x[0] = is + (ie - is) * (0.0 + 1.0) / 2.0;
    w[0] = 1.0 / (1.0 / (ie - is));
    return 0;
  }
  if (n == 2)
  {
    x[0] = is + (ie - is) * ((-1.0 / sqrt((real)3) + 1.0) / 2.0);
    x[1] = is + (ie - is) * ((+1.0 / sqrt((real)3) + 1.0) / 2.0);
    w[0] = 1.0 / (2.0 / (ie - is));
    w[1] = 1.0 / (2.0 / (ie - is));
    return 0;
  }
  for (int j = 0; j < n; ++j)
    x[j] = cos(((2.0 * (real)(n - j) - 1.0) / (2.0 * (real)n)) * ORC_PI);

  real* pv = (real*)zalloc(sizeof(real) * n * (n + 1));
  real* dp = (real*)zalloc(sizeof(real) * n);
  for (int j = 0; j < n; ++j)
    pv[j] = 1.0;
  real max_diff = HUGE_VAL;
  int iter        = 0;
  while (max_diff > ORC_EPS && iter < 200)
  {
    ++iter;
    for (int j = 0; j < n; ++j)
      pv[j + n] = x[j];
    for (int i = 2; i <= n; ++i)
      for (int j = 0; j < n; ++j)
        pv[j + i * n] = ((2.0 * (real)(i - 1) + 1.0) * x[j] * pv[j + (i - 1) * n] -
                         (real)(i - 1) * pv[j + (i - 2) * n]) /
                        (real)i;
    for (int j = 0; j < n; ++j)
      dp[j] = (-(real)n * x[j] * pv[j + n * n] + (real)n * pv[j + (n - 1) * n]) /
              (1.0 - x[j] * x[j]);
    max_diff = 0.0;
    for (int j = 0; j < n; ++j)
    {
      real xo = x[j];
      x[j] -= pv[j + n * n] / dp[j];
      real d = fabs(x[j] - xo);
      if (d > max_diff)
        max_diff = d;
    }
  }
  for (int j = 0; j < n; ++j)
  {
    w[j] = ((ie - is) / ((1.0 - x[j] * x[j]) * dp[j] * dp[j]));
    x[j] = (is + (ie - is) * ((x[j] + 1.0) / 2.0));
  }
  free(pv);
  free(dp);
  return (max_diff > ORC_EPS) ? 1 : 0;
}

/* point count for a rule exact to polynomial order "order"
 * (quadrature.cpp:189) */
int orc_quad_npoints(int order)
{
  return (int)ceil(((real)order + 1.0) / 2.0);
}

/* ------------------------------------------------------------------------ */
/* Lagrange basis on equispaced nodes of [0,1] (basis.cpp:64-170)            */
/* ------------------------------------------------------------------------ */

real orc_lagrange_node1d(int i, int p)
{
  return (p == 0) ? 0.0 : (real)i / (real)p;
}

real orc_lagrange1d(int i, int p, real x)
{
  real e = 1.0;
  const real xi = (real)i / (real)p;
  for (int j = 0; j < p + 1; ++j)
    if (i != j)
    {
      real xj = (real)j / (real)p;
      e *= (x - xj) / (xi - xj);
    }
  return e;
}

real orc_lagrange1d_deriv(int i, int p, real x)
{
  real a_acc = 0.0;
  const real xi = (real)i / (real)p;
  for (int j = 0; j < p + 1; ++j)
  {
    if (j != i)
    {
      real m_acc = 1.0;
      real xj    = (real)j / (real)p;
      for (int m = 0; m < p + 1; ++m)
        if (m != i && m != j)
        {
          real xm = (real)m / (real)p;
          m_acc *= (x - xm) / (xi - xm);
        }
      a_acc += (1.0 / (xi - xj)) * m_acc;
    }
  }
  return a_acc;
}

static real lagrange3d(int bf, int p, const real* r)
{
  int n = p + 1;
  return orc_lagrange1d(bf % n, p, r[0]) * orc_lagrange1d((bf / n) % n, p, r[1]) *
         orc_lagrange1d(bf / (n * n), p, r[2]);
}

static void lagrange3d_grad(int bf, int p, const real* r, real* g)
{
  int n = p + 1;
  int bx = bf % n, by = (bf / n) % n, bz = bf / (n * n);
  real px = orc_lagrange1d(bx, p, r[0]);
  real py = orc_lagrange1d(by, p, r[1]);
  real pz = orc_lagrange1d(bz, p, r[2]);
  g[0] = orc_lagrange1d_deriv(bx, p, r[0]) * py * pz;
  g[1] = px * orc_lagrange1d_deriv(by, p, r[1]) * pz;
  g[2] = px * py * orc_lagrange1d_deriv(bz, p, r[2]);
}

/* face 2D (s,t) -> volume reference point, basis.cpp:180-192 */
void orc_surf2vol(int face, real s, real t, real* r)
{
  switch (face)
  {
    case 0: r[0] = 0.0; r[1] = s; r[2] = t; break;
    case 1: r[0] = 1.0; r[1] = s; r[2] = t; break;
    case 2: r[0] = t; r[1] = 0.0; r[2] = s; break;
    case 3: r[0] = t; r[1] = 1.0; r[2] = s; break;
    case 4: r[0] = s; r[1] = t; r[2] = 0.0; break;
    default: r[0] = s; r[1] = t; r[2] = 1.0; break;
  }
}

static const int opposite_local_face[6] = {1, 0, 3, 2, 5, 4};

/* Pre-evaluated basis tables for order p with an n-point-per-direction Gauss
 * rule (basis.cpp:248-432). */
typedef struct
{
  int p, n1, nbf, nvq, nfq;
  real *x1, *w1;       /* 1D rule */
  real *vw, *fw;       /* 3D / 2D weights */
  real *vpos, *fpos;   /* positions */
  real* veval;         /* [q][bf] */
  real* vgrad;         /* [bf][q][3] */
  real* feval;         /* [f][q][bf] */
  real* fgrad;         /* [f][bf][q][3] */
} orc_basis;

static void basis_init(orc_basis* b, int p, int quad_order)
{
  b->p   = p;
  b->n1  = orc_quad_npoints(quad_order);
  int n1 = b->n1, n = p + 1;
  b->nbf = n * n * n;
  b->nvq = n1 * n1 * n1;
  b->nfq = n1 * n1;
  b->x1  = (real*)zalloc(sizeof(real) * n1);
  b->w1  = (real*)zalloc(sizeof(real) * n1);
  orc_gauss_legendre(0.0, 1.0, n1, b->x1, b->w1);
  b->vw   = (real*)zalloc(sizeof(real) * b->nvq);
  b->vpos = (real*)zalloc(sizeof(real) * b->nvq * 3);
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i)
      {
        int q = n1 * n1 * k + n1 * j + i;
        b->vpos[3 * q + 0] = b->x1[i];
        b->vpos[3 * q + 1] = b->x1[j];
        b->vpos[3 * q + 2] = b->x1[k];
        b->vw[q] = b->w1[i] * b->w1[j] * b->w1[k];
      }
  b->fw   = (real*)zalloc(sizeof(real) * b->nfq);
  b->fpos = (real*)zalloc(sizeof(real) * b->nfq * 2);
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i)
    {
      int q = n1 * j + i;
      b->fpos[2 * q + 0] = b->x1[i];
      b->fpos[2 * q + 1] = b->x1[j];
      b->fw[q] = b->w1[i] * b->w1[j];
    }
  int nbf = b->nbf;
  b->veval = (real*)zalloc(sizeof(real) * nbf * b->nvq);
  b->vgrad = (real*)zalloc(sizeof(real) * nbf * b->nvq * 3);
  b->feval = (real*)zalloc(sizeof(real) * nbf * b->nfq * 6);
  b->fgrad = (real*)zalloc(sizeof(real) * nbf * b->nfq * 6 * 3);
  for (int bf = 0; bf < nbf; ++bf)
    for (int q = 0; q < b->nvq; ++q)
    {
      b->veval[nbf * q + bf] = lagrange3d(bf, p, b->vpos + 3 * q);
      lagrange3d_grad(bf, p, b->vpos + 3 * q, b->vgrad + 3 * (b->nvq * bf + q));
    }
  for (int f = 0; f < 6; ++f)
    for (int bf = 0; bf < nbf; ++bf)
      for (int q = 0; q < b->nfq; ++q)
      {
        real r[3];
        orc_surf2vol(f, b->fpos[2 * q], b->fpos[2 * q + 1], r);
        b->feval[nbf * b->nfq * f + nbf * q + bf] = lagrange3d(bf, p, r);
        lagrange3d_grad(bf, p, r, b->fgrad + 3 * (nbf * b->nfq * f + b->nfq * bf + q));
      }
}

static void basis_free(orc_basis* b)
{
  free(b->x1); free(b->w1); free(b->vw); free(b->fw); free(b->vpos);
  free(b->fpos); free(b->veval); free(b->vgrad); free(b->feval); free(b->fgrad);
}

/* ------------------------------------------------------------------------ */
/* dense small linear algebra (dgmath.cpp:39-371)                            */
/* ------------------------------------------------------------------------ */

real orc_det3x3(const real* m)
{
  return m[0] * m[4] * m[8] + m[1] * m[5] * m[6] + m[2] * m[3] * m[7] -
         m[2] * m[4] * m[6] - m[1] * m[3] * m[8] - m[0] * m[5] * m[7];
}

static real det2(real a, real b, real c, real d)
{
  return a * d - b * c;
}

void orc_inv3x3(const real* m, real* inv)
{
  real id = 1.0 / orc_det3x3(m);
  inv[0] = id * det2(m[4], m[5], m[7], m[8]);
  inv[1] = id * det2(m[2], m[1], m[8], m[7]);
  inv[2] = id * det2(m[1], m[2], m[4], m[5]);
  inv[3] = id * det2(m[5], m[3], m[8], m[6]);
  inv[4] = id * det2(m[0], m[2], m[6], m[8]);
  inv[5] = id * det2(m[2], m[0], m[5], m[3]);
  inv[6] = id * det2(m[3], m[4], m[6], m[7]);
  inv[7] = id * det2(m[1], m[0], m[7], m[6]);
  inv[8] = id * det2(m[0], m[1], m[3], m[4]);
}

/* C = alpha*A*B + beta*C, row-major, A n x m, B m x p */
void orc_matmul(int n, int m, int p, real alpha, const real* A,
                const real* B, real beta, real* C)
{
  for (int ir = 0; ir < n; ++ir)
    for (int ic = 0; ic < p; ++ic)
    {
      int iC     = p * ir + ic;
      real tmp = 0.;
      C[iC] *= beta;
      for (int ia = 0; ia < m; ++ia)
        tmp += alpha * A[m * ir + ia] * B[ia * p + ic];
      C[iC] += tmp;
    }
}

/* r = c * M * v, M packed upper triangle (dgmath.cpp:143-161) */
void orc_symmstore_mvmul(int n, const real* M, const real* v, real c,
                         real* r)
{
  memset(r, 0, n * sizeof(*r));
  int iM = 0;
  for (int ir = 0; ir < n; ++ir)
  {
    ++iM;
    real tmp  = 0.;
    real vrow = c * v[ir];
    for (int ic = ir + 1; ic < n; ++ic)
    {
      tmp += M[iM] * c * v[ic];
      r[ic] += M[iM] * vrow;
      ++iM;
    }
    r[ir] += M[iM - (n - ir)] * vrow + tmp;
  }
}

/* r = c * M * V for p column vectors of length n stored with stride n
 * (dgmath.cpp:190-216) */
void orc_spmm(int n, int p, const real* M, const real* v, real c,
              real* r)
{
  memset(r, 0, (size_t)n * p * sizeof(*r));
  int iM = 0;
  for (int ir = 0; ir < n; ++ir)
  {
    ++iM;
    int iMs = iM;
    for (int iv = 0; iv < p; ++iv)
    {
      iM          = iMs;
      real tmp  = 0.;
      real vrow = c * v[n * iv + ir];
      for (int ic = ir + 1; ic < n; ++ic)
      {
        real tm = M[iM];
        tmp += tm * c * v[n * iv + ic];
        r[n * iv + ic] += tm * vrow;
        ++iM;
      }
      r[n * iv + ir] += M[iM - (n - ir)] * vrow + tmp;
    }
  }
}

/* Gauss-Jordan inverse with partial pivoting, in place (dgmath.cpp:221-371).
 * Returns 0 on success, 1 if the eliminated block is not the identity. */
int orc_invert(real* mat, int n)
{
  int nc = 2 * n;
  real* A = (real*)zalloc(sizeof(real) * 2 * n * n);
#define I2(r, c) (nc * (r) + (c))
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c)
    {
      A[I2(r, c)]     = mat[n * r + c];
      A[I2(r, c + n)] = (r == c) ? 1.0 : 0.0;
    }
  /* forward elimination */
  for (int pr = 0; pr < n - 1; ++pr)
  {
    int mr = 0;
    real ma = -ORC_MAXV;
    for (int r = pr; r < n; ++r)
    {
      real a = fabs(A[I2(r, pr)]);
      if (a > ma)
      {
        ma = a;
        mr = r;
      }
    }
    if (mr != pr)
      for (int c = pr; c < nc; ++c)
      {
        real t     = A[I2(pr, c)];
        A[I2(pr, c)] = A[I2(mr, c)];
        A[I2(mr, c)] = t;
      }
    for (int r = pr + 1; r < n; ++r)
    {
      real f    = -A[I2(r, pr)] / A[I2(pr, pr)];
      A[I2(r, pr)] = 0.0;
      for (int c = pr + 1; c < nc; ++c)
        A[I2(r, c)] += f * A[I2(pr, c)];
    }
  }
  /* back substitution to reduced row echelon form */
  for (int pr = n - 1; pr >= 0; --pr)
  {
    int pc = pr;
    for (int c = pr; c < nc; ++c)
      if (A[I2(pr, c)] != 0.0)
      {
        pc = c;
        break;
      }
    real op    = A[I2(pr, pc)];
    A[I2(pr, pc)] = 1.0;
    for (int c = pc + 1; c < nc; ++c)
      A[I2(pr, c)] /= op;
    for (int r = pr - 1; r >= 0; --r)
    {
      real f = -A[I2(r, pc)] / A[I2(pr, pc)];
      for (int c = r; c < nc; ++c)
      {
        if (c == pc)
          A[I2(r, c)] = 0.0;
        else
          A[I2(r, c)] += f * A[I2(pr, c)];
      }
    }
  }
  int bad = 0;
  for (int r = 0; r < n && !bad; ++r)
    for (int c = 0; c < n; ++c)
      if (A[I2(r, c)] != ((r == c) ? 1.0 : 0.0))
      {
        bad = 1;
        break;
      }
  if (!bad)
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c)
        mat[n * r + c] = A[I2(r, c + n)];
#undef I2
  free(A);
  return bad;
}

/* ------------------------------------------------------------------------ */
/* pointwise physics (flux.cpp:24-202, bound_states.cpp:24-130)              */
/* ------------------------------------------------------------------------ */

void orc_analytical_flux(const real* U, real gamma, real* Fx,
                         real* Fy, real* Fz)
{
  real rho = U[0];
  real u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
  real p = (gamma - 1.) * (U[4] - 0.5 * rho * (u * u + v * v + w * w));
  real H = (U[4] + p) / rho;
  Fx[0] = rho * u;     Fx[1] = rho * u * u + p; Fx[2] = rho * u * v;
  Fx[3] = rho * u * w; Fx[4] = rho * u * H;
  Fy[0] = rho * v;     Fy[1] = rho * u * v;     Fy[2] = rho * v * v + p;
  Fy[3] = rho * v * w; Fy[4] = rho * v * H;
  Fz[0] = rho * w;     Fz[1] = rho * u * w;     Fz[2] = rho * v * w;
  Fz[3] = rho * w * w + p; Fz[4] = rho * w * H;
}

/* Roe flux with the 5%-of-c entropy fix (flux.cpp:56-154) */
void orc_roe(const real* Ul, const real* Ur, const real* n, real g,
             real* f, real* smax)
{
  real gm1 = g - 1.0;
  real rl = Ul[0], ul = Ul[1] / rl, vl = Ul[2] / rl, wl = Ul[3] / rl;
  real El = Ul[4] / rl;
  real pl = gm1 * (Ul[4] - 0.5 * rl * (ul * ul + vl * vl + wl * wl));
  real Hl = El + (pl / rl);
  real rr = Ur[0], ur = Ur[1] / rr, vr = Ur[2] / rr, wr = Ur[3] / rr;
  real Er = Ur[4] / rr;
  real pr = gm1 * (Ur[4] - 0.5 * rr * (ur * ur + vr * vr + wr * wr));
  real Hr = Er + (pr / rr);
  real sl = sqrt(rl), sr = sqrt(rr);
  real den = sl + sr;
  real u = (sl * ul + sr * ur) / den;
  real v = (sl * vl + sr * vr) / den;
  real w = (sl * wl + sr * wr) / den;
  real H = (sl * Hl + sr * Hr) / den;
  real dr = rr - rl, dxm = Ur[1] - Ul[1], dym = Ur[2] - Ul[2];
  real dzm = Ur[3] - Ul[3], drE = Ur[4] - Ul[4];
  real qsq = u * u + v * v + w * w;
  real c   = sqrt(gm1 * (H - 0.5 * qsq));
  real un  = u * n[0] + v * n[1] + w * n[2];
  real l1 = un + c, l2 = un - c, l3 = un;
  real al1 = fabs(l1), al2 = fabs(l2), al3 = fabs(l3);
  real eps = 0.05 * c;
  if (al1 < eps) al1 = fabs((eps * eps + l1 * l1) / (2.0 * eps));
  if (al2 < eps) al2 = fabs((eps * eps + l2 * l2) / (2.0 * eps));
  if (al3 < eps) al3 = fabs((eps * eps + l3 * l3) / (2.0 * eps));
  real s1 = 0.5 * (al1 + al2);
  real s2 = 0.5 * (al1 - al2);
  real G1 = gm1 * (0.5 * qsq * dr - (u * dxm + v * dym + w * dzm) + drE);
  real G2 = -un * dr + (dxm * n[0] + dym * n[1] + dzm * n[2]);
  real C1 = (G1 / (c * c)) * (s1 - al3) + (G2 / c) * s2;
  real C2 = (G1 / c) * s2 + (s1 - al3) * G2;
  f[0] = 0.5 * ((Ul[1] * n[0] + Ul[2] * n[1] + Ul[3] * n[2]) +
                (Ur[1] * n[0] + Ur[2] * n[1] + Ur[3] * n[2])) -
         0.5 * (al3 * dr + C1);
  f[1] = 0.5 * (((rl * ul * ul + pl) * n[0] + (rl * ul * vl) * n[1] +
                 (rl * ul * wl) * n[2]) +
                ((rr * ur * ur + pr) * n[0] + (rr * ur * vr) * n[1] +
                 (rr * ur * wr) * n[2])) -
         0.5 * (al3 * dxm + C1 * u + C2 * n[0]);
  f[2] = 0.5 * (((rl * ul * vl) * n[0] + (rl * vl * vl + pl) * n[1] +
                 (rl * vl * wl) * n[2]) +
                ((rr * ur * vr) * n[0] + (rr * vr * vr + pr) * n[1] +
                 (rr * vr * wr) * n[2])) -
         0.5 * (al3 * dym + C1 * v + C2 * n[1]);
  f[3] = 0.5 * (((rl * ul * wl) * n[0] + (rl * vl * wl) * n[1] +
                 (rl * wl * wl + pl) * n[2]) +
                ((rr * ur * wr) * n[0] + (rr * vr * wr) * n[1] +
                 (rr * wr * wr + pr) * n[2])) -
         0.5 * (al3 * dzm + C1 * w + C2 * n[2]);
  f[4] = 0.5 * (((rl * ul * Hl) * n[0] + (rl * vl * Hl) * n[1] +
                 (rl * wl * Hl) * n[2]) +
                ((rr * ur * Hr) * n[0] + (rr * vr * Hr) * n[1] +
                 (rr * wr * Hr) * n[2])) -
         0.5 * (al3 * drE + C1 * H + C2 * un);
  real sm = al1;
  if (al2 > sm) sm = al2;
  if (al3 > sm) sm = al3;
  if (smax) *smax = sm;
}

/* Newtonian viscous flux, no heat conduction (flux.cpp:156-202) */
void orc_A(const real* s, const real* sx, const real* sy,
           const real* sz, real mu, real* Qx, real* Qy, real* Qz)
{
  real lam = -((real)2 / 3) * mu;
  real rho = s[0];
  real u = s[1] / rho, v = s[2] / rho, w = s[3] / rho;
  real ux = (sx[1] - sx[0] * u) / rho;
  real uy = (sy[1] - sy[0] * u) / rho;
  real uz = (sz[1] - sz[0] * u) / rho;
  real vx = (sx[2] - sx[0] * v) / rho;
  real vy = (sy[2] - sy[0] * v) / rho;
  real vz = (sz[2] - sz[0] * v) / rho;
  real wx = (sx[3] - sx[0] * w) / rho;
  real wy = (sy[3] - sy[0] * w) / rho;
  real wz = (sz[3] - sz[0] * w) / rho;
  real dv  = ux + vy + wz;
  real txx = lam * dv + 2. * mu * ux;
  real tyy = lam * dv + 2. * mu * vy;
  real tzz = lam * dv + 2. * mu * wz;
  real txy = mu * (vx + uy);
  real txz = mu * (uz + wx);
  real tyz = mu * (wy + vz);
  Qx[0] = 0.; Qx[1] = txx; Qx[2] = txy; Qx[3] = txz;
  Qx[4] = u * txx + v * txy + w * txz;
  Qy[0] = 0.; Qy[1] = txy; Qy[2] = tyy; Qy[3] = tyz;
  Qy[4] = u * txy + v * tyy + w * tyz;
  Qz[0] = 0.; Qz[1] = txz; Qz[2] = tyz; Qz[3] = tzz;
  Qz[4] = u * txz + v * tyz + w * tzz;
}

/* boundary states (bound_states.cpp:24-130).  Pt, Tt only matter for
 * subsonic inflow, P only for subsonic outflow.  Returns nonzero when the
 * inflow Mach-number quadratic has no non-negative root. */
int orc_bound_state(s64 bt, const real* UL, const real* n,
                    const orc_params* pr, real Pt, real Tt, real* Ub)
{
  real gamma = pr->gamma;
  real gm1   = gamma - 1.;
  switch (bt)
  {
    case ORC_FREESTREAM:
      for (int i = 0; i < 5; ++i)
        Ub[i] = pr->Ufr[i];
      return 0;
    case ORC_INVISCID_WALL:
    {
      real up = UL[1] / UL[0], vp = UL[2] / UL[0], wp = UL[3] / UL[0];
      real vdn = up * n[0] + vp * n[1] + wp * n[2];
      Ub[0] = UL[0];
      Ub[1] = UL[0] * (up - vdn * n[0]);
      Ub[2] = UL[0] * (vp - vdn * n[1]);
      Ub[3] = UL[0] * (wp - vdn * n[2]);
      Ub[4] = UL[4];
      return 0;
    }
    case ORC_NO_SLIP_WALL:
      Ub[0] = UL[0]; Ub[1] = 0.; Ub[2] = 0.; Ub[3] = 0.; Ub[4] = UL[4];
      return 0;
    case ORC_SUBSONIC_INFLOW:
    {
      real R    = pr->R;
      real up = UL[1] / UL[0], vp = UL[2] / UL[0], wp = UL[3] / UL[0];
      real Vsqp = up * up + vp * vp + wp * wp;
      real pp   = gm1 * (UL[4] - 0.5 * UL[0] * Vsqp);
      real unp  = up * n[0] + vp * n[1] + wp * n[2];
      real cp   = sqrt(gamma * pp / UL[0]);
      real Jp   = unp + ((2. * cp) / gm1);
      real a    = gamma * R * Tt - 0.5 * gm1 * Jp * Jp;
      real b    = (4. * gamma * R * Tt * -1.) / gm1;
      real c    = (4. * gamma * R * Tt) / (gm1 * gm1) - Jp * Jp;
      real s0   = (-b + sqrt(b * b - 4. * a * c)) / (2. * a);
      real s1   = (-b - sqrt(b * b - 4. * a * c)) / (2. * a);
      real Mb   = 0.;
      int err     = 0;
      if (s0 >= 0. && s1 >= 0.)
        Mb = (s0 < s1) ? s0 : s1;
      else if (s0 >= 0.)
        Mb = s0;
      else if (s1 >= 0.)
        Mb = s1;
      else
        err = 1;
      real Tb   = Tt / (1. + 0.5 * gm1 * Mb * Mb);
      real Pb   = Pt * pow(Tb / Tt, gamma / gm1);
      real rhob = Pb / (R * Tb);
      real cb   = sqrt((gamma * Pb) / rhob);
      real ub = Mb * cb * -n[0], vb = Mb * cb * -n[1], wb = Mb * cb * -n[2];
      real rEb = (Pb / gm1) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
      Ub[0] = rhob; Ub[1] = rhob * ub; Ub[2] = rhob * vb; Ub[3] = rhob * wb;
      Ub[4] = rEb;
      return err;
    }
    case ORC_SUBSONIC_OUTFLOW:
    {
      real P    = pr->P;
      real rhop = UL[0];
      real up = UL[1] / rhop, vp = UL[2] / rhop, wp = UL[3] / rhop;
      real Vsqp = up * up + vp * vp + wp * wp;
      real pp   = gm1 * (UL[4] - 0.5 * rhop * Vsqp);
      real Sp   = pp / pow(rhop, gamma);
      real rhob = pow(P / Sp, 1. / gamma);
      real cb   = sqrt((gamma * P) / rhob);
      real unp  = up * n[0] + vp * n[1] + wp * n[2];
      real cp   = sqrt(gamma * pp / rhop);
      real Jp   = unp + ((2. * cp) / gm1);
      real unb  = Jp - ((2. * cb) / (gamma - 1.));
      real vpdn = (up * n[0] + vp * n[1] + wp * n[2]);
      real ub   = up - vpdn * n[0] + unb * n[0];
      real vb   = vp - vpdn * n[1] + unb * n[1];
      real wb   = wp - vpdn * n[2] + unb * n[2];
      real rEb  = (P / (gamma - 1.)) + 0.5 * rhob * (ub * ub + vb * vb + wb * wb);
      Ub[0] = rhob; Ub[1] = rhob * ub; Ub[2] = rhob * vb; Ub[3] = rhob * wb;
      Ub[4] = rEb;
      return 0;
    }
  }
  return 2;
}

/* ------------------------------------------------------------------------ */
/* geometry (geometry_common.cpp, geometry_shared.cpp)                       */
/* ------------------------------------------------------------------------ */

typedef struct
{
  int p, q;
  u64 nx, ny, nz, nelem;
  s64 bounds[6];
  orc_basis refp;     /* basis(p, 2p+1) */
  orc_basis refq;     /* basis(q, 2p+1) */
  orc_basis refp_2p;  /* basis(p, 2p) for the mass matrix */
  orc_basis refq_2p;  /* basis(q, 2p) */
  real* node;       /* [e][(q+1)^3][3] */
  real* Minv;       /* packed upper triangle per element */
  real* vJ;         /* [e][q] */
  real* vgrad;      /* [e][q][bf][3] */
  real* fgrad;      /* [e][lf][q][bf][3] */
  u64 nface, nifc, nbfc;
  s64* ifl;           /* [ifi][5]: gfn, eL, eR, lfL, lfR */
  s64* bfl;           /* [bfi][4]: gfn, eL, bt, lfL */
  real* n_;         /* [gfn][q][3] */
  real* fJ_;        /* [gfn][q] */
  real* ih;         /* [ifi][2] */
  real* bh;         /* [bfi] */
} orc_geom;

static u64 nentry_minv(const orc_geom* g)
{
  u64 nbf = g->refp.nbf;
  return nbf * (nbf + 1) / 2;
}

/* J[3*i+j] = d x_i / d xi_j at a volume (f < 0) or face quad point
 * (geometry_common.cpp:264-290) */
static void jacobian(const orc_geom* g, u64 e, int qn, int f,
                     const orc_basis* rq, real* J)
{
  for (int i = 0; i < 9; ++i)
    J[i] = 0.;
  for (int gn = 0; gn < rq->nbf; ++gn)
  {
    const real* nd = g->node + ((u64)rq->nbf * e + gn) * 3;
    const real* rg = (f >= 0) ? rq->fgrad + 3 * (rq->nbf * rq->nfq * f + rq->nfq * gn + qn)
                                : rq->vgrad + 3 * (rq->nvq * gn + qn);
    J[0] += nd[0] * rg[0]; J[1] += nd[0] * rg[1]; J[2] += nd[0] * rg[2];
    J[3] += nd[1] * rg[0]; J[4] += nd[1] * rg[1]; J[5] += nd[1] * rg[2];
    J[6] += nd[2] * rg[0]; J[7] += nd[2] * rg[1]; J[8] += nd[2] * rg[2];
  }
}

/* per-element mass-matrix inverse, detJ, physical basis gradients
 * (geometry_common.cpp:292-395) */
static int geom_precompute(orc_geom* g)
{
  int nbf = g->refp.nbf;
  real* detJ = (real*)zalloc(sizeof(real) * g->refq_2p.nvq);
  real* M    = (real*)zalloc(sizeof(real) * nbf * nbf);
  int status   = 0;
  for (u64 e = 0; e < g->nelem; ++e)
  {
    real jac[9];
    for (int qn = 0; qn < g->refq_2p.nvq; ++qn)
    {
      jacobian(g, e, qn, -1, &g->refq_2p, jac);
      detJ[qn] = orc_det3x3(jac);
    }
    for (int t = 0; t < nbf; ++t)
      for (int b = t; b < nbf; ++b)
      {
        real acc = 0.;
        for (int qn = 0; qn < g->refp_2p.nvq; ++qn)
          acc += detJ[qn] * g->refp_2p.vw[qn] * g->refp_2p.veval[nbf * qn + t] *
                 g->refp_2p.veval[nbf * qn + b];
        M[nbf * t + b] = acc;
        M[nbf * b + t] = acc;
      }
    if (orc_invert(M, nbf) != 0)
    {
      status = 1;
      break;
    }
    real* Ms = g->Minv + e * nentry_minv(g);
    u64 iM = 0;
    for (int r = 0; r < nbf; ++r)
      for (int c = r; c < nbf; ++c)
        Ms[iM++] = M[nbf * r + c];
  }
  free(detJ);
  free(M);
  if (status)
    return status;

  for (u64 e = 0; e < g->nelem; ++e)
    for (int qn = 0; qn < g->refp.nvq; ++qn)
    {
      real J[9], Ji[9];
      jacobian(g, e, qn, -1, &g->refq, J);
      g->vJ[e * g->refp.nvq + qn] = orc_det3x3(J);
      orc_inv3x3(J, Ji);
      for (int bf = 0; bf < nbf; ++bf)
        orc_matmul(1, 3, 3, 1., g->refp.vgrad + 3 * (g->refp.nvq * bf + qn), Ji, 0.,
                   g->vgrad + ((u64)nbf * (g->refp.nvq * e + qn) + bf) * 3);
    }
  for (u64 e = 0; e < g->nelem; ++e)
    for (int f = 0; f < 6; ++f)
      for (int qn = 0; qn < g->refp.nfq; ++qn)
      {
        real J[9], Ji[9];
        jacobian(g, e, qn, f, &g->refq, J);
        orc_inv3x3(J, Ji);
        for (int bf = 0; bf < nbf; ++bf)
          orc_matmul(1, 3, 3, 1., g->refp.fgrad + 3 * (nbf * g->refp.nfq * f + g->refp.nfq * bf + qn),
                     Ji, 0., g->fgrad + ((u64)nbf * (g->refp.nfq * (6 * e + f) + qn) + bf) * 3);
      }
  return 0;
}

/* boundary index of a local face's neighbour (geometry_common.cpp:397-414) */
static s64 find_boundary_index(u64 i, u64 j, u64 k, int lfn, u64 nx, u64 ny,
                               u64 nz, s64* ia, s64* ja, s64* ka)
{
  static const s64 mi[6] = {-1, 1, 0, 0, 0, 0};
  static const s64 mj[6] = {0, 0, -1, 1, 0, 0};
  static const s64 mk[6] = {0, 0, 0, 0, -1, 1};
  *ia = (s64)i + mi[lfn];
  *ja = (s64)j + mj[lfn];
  *ka = (s64)k + mk[lfn];
  s64 b = -1;
  b += ((*ia < 0) * 1) + ((*ia > (s64)nx - 1) * 2) + ((*ja < 0) * 3) +
       ((*ja > (s64)ny - 1) * 4) + ((*ka < 0) * 5) + ((*ka > (s64)nz - 1) * 6);
  return b;
}

/* global position of reference point r on element e (geometry_common.cpp:235) */
static void ref2glo(const orc_geom* g, u64 e, const real* r, real* x)
{
  x[0] = x[1] = x[2] = 0.;
  for (int b = 0; b < g->refq.nbf; ++b)
  {
    real phi = lagrange3d(b, g->q, r);
    const real* nd = g->node + ((u64)g->refq.nbf * e + b) * 3;
    x[0] += nd[0] * phi;
    x[1] += nd[1] * phi;
    x[2] += nd[2] * phi;
  }
}

/* unit normal and surface element at a face quad point
 * (geometry_common.cpp:416-454) */
static void surface_geometry(const orc_geom* g, u64 e, int lfn, int qn,
                             s64 ia, s64 ja, s64 ka, real* nrm, real* sJ)
{
  real ds[3] = {0, 0, 0}, dt[3] = {0, 0, 0};
  int qq = g->q, nq1 = qq + 1;
  for (int b = 0; b < nq1 * nq1; ++b)
  {
    int bx = b % nq1, by = b / nq1;
    real r[3], x[3];
    orc_surf2vol(lfn, orc_lagrange_node1d(bx, qq), orc_lagrange_node1d(by, qq), r);
    ref2glo(g, e, r, x);
    real s = g->refq.fpos[2 * qn], t = g->refq.fpos[2 * qn + 1];
    real gx = orc_lagrange1d_deriv(bx, qq, s) * orc_lagrange1d(by, qq, t);
    real gy = orc_lagrange1d(bx, qq, s) * orc_lagrange1d_deriv(by, qq, t);
    ds[0] += x[0] * gx; ds[1] += x[1] * gx; ds[2] += x[2] * gx;
    dt[0] += x[0] * gy; dt[1] += x[1] * gy; dt[2] += x[2] * gy;
  }
  nrm[0] = ds[1] * dt[2] - ds[2] * dt[1];
  nrm[1] = ds[2] * dt[0] - ds[0] * dt[2];
  nrm[2] = ds[0] * dt[1] - ds[1] * dt[0];
  *sJ = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  nrm[0] /= *sJ;
  nrm[1] /= *sJ;
  nrm[2] /= *sJ;
  if (ia < 0 || ja < 0 || ka < 0)
  {
    nrm[0] *= -1.;
    nrm[1] *= -1.;
    nrm[2] *= -1.;
  }
}

/* single-domain face enumeration (geometry_shared.cpp:65-83, 156-267) */
static void face_precompute(orc_geom* g)
{
  u64 nx = g->nx, ny = g->ny, nz = g->nz;
  u64 gfn = 0, ifn = 0, bfn = 0;
  int nfq = g->refp.nfq;
  for (u64 k = 0; k < nz; ++k)
    for (u64 j = 0; j < ny; ++j)
      for (u64 i = 0; i < nx; ++i)
      {
        u64 e = nx * ny * k + nx * j + i;
        int todo[6] = {1, 3, 5, -1, -1, -1};
        if (i == 0) todo[3] = 0;
        if (j == 0) todo[4] = 2;
        if (k == 0) todo[5] = 4;
        for (int c = 0; c < 6; ++c)
        {
          if (todo[c] == -1)
            continue;
          int lfn = todo[c];
          if (g->bounds[1] == ORC_PERIODIC && i == nx - 1 && lfn == 1) continue;
          if (g->bounds[3] == ORC_PERIODIC && j == ny - 1 && lfn == 3) continue;
          if (g->bounds[5] == ORC_PERIODIC && k == nz - 1 && lfn == 5) continue;
          s64 ia, ja, ka;
          s64 bi = find_boundary_index(i, j, k, lfn, nx, ny, nz, &ia, &ja, &ka);
          int isb = 0;
          u64 eR  = 0;
          if (bi == -1 || g->bounds[bi] == ORC_PERIODIC)
          {
            eR = nx * ny * (u64)((ka + (s64)nz) % (s64)nz) +
                 nx * (u64)((ja + (s64)ny) % (s64)ny) + (u64)((ia + (s64)nx) % (s64)nx);
            g->ifl[5 * ifn + 0] = (s64)gfn;
            g->ifl[5 * ifn + 1] = (s64)e;
            g->ifl[5 * ifn + 2] = (s64)eR;
            g->ifl[5 * ifn + 3] = lfn;
            g->ifl[5 * ifn + 4] = opposite_local_face[lfn];
            ++ifn;
          }
          else
          {
            isb = 1;
            g->bfl[4 * bfn + 0] = (s64)gfn;
            g->bfl[4 * bfn + 1] = (s64)e;
            g->bfl[4 * bfn + 2] = g->bounds[bi];
            g->bfl[4 * bfn + 3] = lfn;
            ++bfn;
          }
          real area = 0.;
          for (int qn = 0; qn < nfq; ++qn)
          {
            real* nrm = g->n_ + 3 * ((u64)nfq * gfn + qn);
            real* sJ  = g->fJ_ + ((u64)nfq * gfn + qn);
            surface_geometry(g, e, lfn, qn, ia, ja, ka, nrm, sJ);
            area += *sJ * g->refp.fw[qn];
          }
          real vL = 0., vR = 0.;
          for (int qi = 0; qi < g->refp.nvq; ++qi)
          {
            vL += g->vJ[e * g->refp.nvq + qi] * g->refp.vw[qi];
            if (!isb)
              vR += g->vJ[eR * g->refp.nvq + qi] * g->refp.vw[qi];
          }
          if (!isb)
          {
            g->ih[2 * (ifn - 1) + 0] = vL / area;
            g->ih[2 * (ifn - 1) + 1] = vR / area;
          }
          else
            g->bh[bfn - 1] = vL / area;
          ++gfn;
        }
      }
}

void orc_geom_destroy(orc_geom* g);

/* Builds the geometry of a structured nx*ny*nz hex mesh of solution order p
 * and geometry order q from element nodes, including the reference's full
 * per-element precompute.  Returns NULL on invalid input or a singular mass
 * matrix. */
orc_geom* orc_geom_create(int p, int q, u64 nx, u64 ny, u64 nz,
                          const s64* bounds, const real* nodes)
{
  if (p < 0 || p > 8 || q < 1 || q > 8 || nx == 0 || ny == 0 || nz == 0)
    return NULL;
  for (int d = 0; d < 3; ++d)
    if ((bounds[2 * d] == ORC_PERIODIC) != (bounds[2 * d + 1] == ORC_PERIODIC))
      return NULL;
  orc_geom* g = (orc_geom*)calloc(1, sizeof(orc_geom));
  g->p = p; g->q = q; g->nx = nx; g->ny = ny; g->nz = nz;
  g->nelem = nx * ny * nz;
  for (int i = 0; i < 6; ++i)
    g->bounds[i] = bounds[i];
  basis_init(&g->refp, p, 2 * p + 1);
  basis_init(&g->refq, q, 2 * p + 1);
  basis_init(&g->refp_2p, p, 2 * p);
  basis_init(&g->refq_2p, q, 2 * p);
  u64 nn = g->nelem * g->refq.nbf * 3;
  g->node = (real*)zalloc(sizeof(real) * nn);
  memcpy(g->node, nodes, sizeof(real) * nn);
  u64 nbf = g->refp.nbf;
  g->Minv  = (real*)zalloc(sizeof(real) * g->nelem * nentry_minv(g));
  g->vJ    = (real*)zalloc(sizeof(real) * g->nelem * g->refp.nvq);
  g->vgrad = (real*)zalloc(sizeof(real) * g->nelem * nbf * g->refp.nvq * 3);
  g->fgrad = (real*)zalloc(sizeof(real) * g->nelem * nbf * g->refp.nfq * 6 * 3);
  if (geom_precompute(g) != 0)
  {
    orc_geom_destroy(g);
    return NULL;
  }
  g->nbfc = ny * nz * (bounds[0] < -2) + ny * nz * (bounds[1] < -2) +
            nx * nz * (bounds[2] < -2) + nx * nz * (bounds[3] < -2) +
            nx * ny * (bounds[4] < -2) + nx * ny * (bounds[5] < -2);
  g->nface = g->nelem * 3 + (bounds[0] != ORC_PERIODIC) * ny * nz +
             (bounds[2] != ORC_PERIODIC) * nz * nx +
             (bounds[4] != ORC_PERIODIC) * nx * ny;
  g->nifc = g->nface - g->nbfc;
  g->ifl = (s64*)zalloc(sizeof(s64) * (g->nifc * 5 + 1));
  g->bfl = (s64*)zalloc(sizeof(s64) * (g->nbfc * 4 + 1));
  g->n_  = (real*)zalloc(sizeof(real) * g->nface * g->refp.nfq * 3);
  g->fJ_ = (real*)zalloc(sizeof(real) * g->nface * g->refp.nfq);
  g->ih  = (real*)zalloc(sizeof(real) * (g->nifc * 2 + 1));
  g->bh  = (real*)zalloc(sizeof(real) * (g->nbfc + 1));
  face_precompute(g);
  return g;
}

void orc_geom_destroy(orc_geom* g)
{
  if (!g)
    return;
  basis_free(&g->refp); basis_free(&g->refq);
  basis_free(&g->refp_2p); basis_free(&g->refq_2p);
  free(g->node); free(g->Minv); free(g->vJ); free(g->vgrad); free(g->fgrad);
  free(g->ifl); free(g->bfl); free(g->n_); free(g->fJ_); free(g->ih); free(g->bh);
  free(g);
}

/* accessors for tests */
void orc_geom_sizes(const orc_geom* g, u64* out)
{
  out[0] = g->nelem; out[1] = (u64)g->refp.nbf; out[2] = (u64)g->refp.nvq;
  out[3] = (u64)g->refp.nfq; out[4] = g->nface; out[5] = g->nifc; out[6] = g->nbfc;
  out[7] = nentry_minv(g);
}
const real* orc_geom_vJ(const orc_geom* g) { return g->vJ; }
const real* orc_geom_Minv(const orc_geom* g) { return g->Minv; }
const real* orc_geom_vgrad(const orc_geom* g) { return g->vgrad; }
const real* orc_geom_fgrad(const orc_geom* g) { return g->fgrad; }
const real* orc_geom_n(const orc_geom* g) { return g->n_; }
const real* orc_geom_fJ(const orc_geom* g) { return g->fJ_; }
const real* orc_geom_ih(const orc_geom* g) { return g->ih; }
const real* orc_geom_bh(const orc_geom* g) { return g->bh; }
const s64* orc_geom_ifl(const orc_geom* g) { return g->ifl; }
const s64* orc_geom_bfl(const orc_geom* g) { return g->bfl; }
const real* orc_geom_vw(const orc_geom* g) { return g->refp.vw; }
const real* orc_geom_fw(const orc_geom* g) { return g->refp.fw; }
const real* orc_geom_veval(const orc_geom* g) { return g->refp.veval; }
const real* orc_geom_feval(const orc_geom* g) { return g->refp.feval; }

/* ------------------------------------------------------------------------ */
/* residual (solve.cpp:150-807) on the single-domain face lists              */
/* ------------------------------------------------------------------------ */

#define SIDX(e, r, b) (((u64)(e) * 5 + (r)) * nbf + (b))

static void eval_state_face(const orc_geom* g, const real* U, u64 e, int lf,
                            int qi, real* S)
{
  int nbf = g->refp.nbf;
  for (int r = 0; r < 5; ++r)
    S[r] = 0.;
  for (int b = 0; b < nbf; ++b)
  {
    real be = g->refp.feval[nbf * g->refp.nfq * lf + nbf * qi + b];
    for (int r = 0; r < 5; ++r)
      S[r] += U[SIDX(e, r, b)] * be;
  }
}

static void eval_grad_face(const orc_geom* g, const real* U, u64 e, int lf,
                           int qi, real* Sx, real* Sy, real* Sz)
{
  int nbf = g->refp.nbf;
  for (int r = 0; r < 5; ++r)
    Sx[r] = Sy[r] = Sz[r] = 0.;
  for (int b = 0; b < nbf; ++b)
  {
    const real* gr = g->fgrad + ((u64)nbf * (g->refp.nfq * (6 * e + lf) + qi) + b) * 3;
    for (int r = 0; r < 5; ++r)
    {
      Sx[r] += U[SIDX(e, r, b)] * gr[0];
      Sy[r] += U[SIDX(e, r, b)] * gr[1];
      Sz[r] += U[SIDX(e, r, b)] * gr[2];
    }
  }
}

/* A(s, d (x) n): the viscous flux of a jump d spread along normal nsign*n */
static void A_jump(const real* s, const real* d, const real* n,
                   real nsign, real mu, real* Qx, real* Qy, real* Qz)
{
  real dx[5], dy[5], dz[5];
  for (int i = 0; i < 5; ++i)
  {
    dx[i] = d[i] * (nsign * n[0]);
    dy[i] = d[i] * (nsign * n[1]);
    dz[i] = d[i] * (nsign * n[2]);
  }
  orc_A(s, dx, dy, dz, mu, Qx, Qy, Qz);
}

/* R = volume - source + face terms, f = -Minv R.  Returns nonzero if a
 * boundary state could not be formed. */
int orc_residual(const orc_geom* g, const orc_params* prm, const real* U,
                 real* R, real* f)
{
  real mu = prm->mu, gamma = prm->gamma, eta = prm->eta;
  int nbf = g->refp.nbf, nvq = g->refp.nvq, nfq = g->refp.nfq;
  u64 nelem = g->nelem;
  real Tt, Pt;
  {
    real u = prm->Ufr[1] / prm->Ufr[0], v = prm->Ufr[2] / prm->Ufr[0];
    real w = prm->Ufr[3] / prm->Ufr[0];
    real gm1 = gamma - 1.;
    real M = sqrt(u * u + v * v + w * w) / prm->c;
    Tt = prm->T * (1. + 0.5 * gm1 * M * M);
    Pt = prm->P * pow(Tt / prm->T, gamma / gm1);
  }
  int status = 0;
  memset(R, 0, sizeof(real) * nelem * 5 * nbf);

  /* face traces (solve.cpp:184-256) */
  u64 nface = g->nface;
  real* ULst = (real*)zalloc(sizeof(real) * nface * nfq * 5);
  real* URst = (real*)zalloc(sizeof(real) * nface * nfq * 5);
  real* Ubst = (real*)zalloc(sizeof(real) * nface * nfq * 5);
  real* ULg  = (real*)zalloc(sizeof(real) * nface * nfq * 15);
  real* URg  = (real*)zalloc(sizeof(real) * nface * nfq * 15);
#define ST(a, fi, qi, r) a[((u64)nfq * nface) * (r) + (u64)nfq * (fi) + (qi)]
  for (u64 ifi = 0; ifi < g->nifc; ++ifi)
  {
    s64 fi = g->ifl[5 * ifi], eL = g->ifl[5 * ifi + 1], eR = g->ifl[5 * ifi + 2];
    int lfL = (int)g->ifl[5 * ifi + 3], lfR = (int)g->ifl[5 * ifi + 4];
    for (int qi = 0; qi < nfq; ++qi)
    {
      real S[5], Sx[5], Sy[5], Sz[5];
      eval_state_face(g, U, eL, lfL, qi, S);
      eval_grad_face(g, U, eL, lfL, qi, Sx, Sy, Sz);
      for (int r = 0; r < 5; ++r)
      {
        ST(ULst, fi, qi, r) = S[r];
        ST(ULg, fi, qi, r) = Sx[r];
        ST(ULg, fi, qi, 5 + r) = Sy[r];
        ST(ULg, fi, qi, 10 + r) = Sz[r];
      }
      eval_state_face(g, U, eR, lfR, qi, S);
      eval_grad_face(g, U, eR, lfR, qi, Sx, Sy, Sz);
      for (int r = 0; r < 5; ++r)
      {
        ST(URst, fi, qi, r) = S[r];
        ST(URg, fi, qi, r) = Sx[r];
        ST(URg, fi, qi, 5 + r) = Sy[r];
        ST(URg, fi, qi, 10 + r) = Sz[r];
      }
    }
  }
  for (u64 bfi = 0; bfi < g->nbfc; ++bfi)
  {
    s64 fi = g->bfl[4 * bfi], eL = g->bfl[4 * bfi + 1], bt = g->bfl[4 * bfi + 2];
    int lfL = (int)g->bfl[4 * bfi + 3];
    for (int qi = 0; qi < nfq; ++qi)
    {
      real S[5], Sx[5], Sy[5], Sz[5], Sb[5];
      const real* nrm = g->n_ + 3 * ((u64)nfq * fi + qi);
      eval_state_face(g, U, eL, lfL, qi, S);
      eval_grad_face(g, U, eL, lfL, qi, Sx, Sy, Sz);
      status |= orc_bound_state(bt, S, nrm, prm, Pt, Tt, Sb);
      for (int r = 0; r < 5; ++r)
      {
        ST(ULst, fi, qi, r) = S[r];
        ST(ULg, fi, qi, r) = Sx[r];
        ST(ULg, fi, qi, 5 + r) = Sy[r];
        ST(ULg, fi, qi, 10 + r) = Sz[r];
        ST(Ubst, fi, qi, r) = Sb[r];
      }
    }
  }

  /* volume term (solve.cpp:276-341) */
  for (u64 e = 0; e < nelem; ++e)
    for (int qi = 0; qi < nvq; ++qi)
    {
      real Fx[5], Fy[5], Fz[5], Qx[5], Qy[5], Qz[5];
      real S[5] = {0}, Sx[5] = {0}, Sy[5] = {0}, Sz[5] = {0};
      real qw = g->refp.vw[qi];
      real Jd = g->vJ[e * nvq + qi];
      for (int b = 0; b < nbf; ++b)
      {
        real be = g->refp.veval[nbf * qi + b];
        const real* gr = g->vgrad + ((u64)nbf * (nvq * e + qi) + b) * 3;
        for (int r = 0; r < 5; ++r)
        {
          real u = U[SIDX(e, r, b)];
          S[r] += u * be;
          Sx[r] += u * gr[0];
          Sy[r] += u * gr[1];
          Sz[r] += u * gr[2];
        }
      }
      orc_analytical_flux(S, gamma, Fx, Fy, Fz);
      orc_A(S, Sx, Sy, Sz, mu, Qx, Qy, Qz);
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          const real* tg = g->vgrad + ((u64)nbf * (nvq * e + qi) + t) * 3;
          R[SIDX(e, r, t)] += ((tg[0] * Qx[r] + tg[1] * Qy[r] + tg[2] * Qz[r]) -
                               (tg[0] * Fx[r] + tg[1] * Fy[r] + tg[2] * Fz[r])) *
                              Jd * qw;
        }
    }

  /* source term (solve.cpp:345-360) */
  for (u64 e = 0; e < nelem; ++e)
    for (int qi = 0; qi < nvq; ++qi)
    {
      real Jd = g->vJ[e * nvq + qi];
      real qw = g->refp.vw[qi];
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
          R[SIDX(e, r, t)] -= g->refp.veval[nbf * qi + t] * prm->source[r] * Jd * qw;
    }

  real* RL = (real*)zalloc(sizeof(real) * 15 * nbf);
  real* RR = (real*)zalloc(sizeof(real) * 15 * nbf);
  real* DL = (real*)zalloc(sizeof(real) * 15 * nbf);
  real* DR = (real*)zalloc(sizeof(real) * 15 * nbf);
  int dimos = nbf * 5;

  /* interior faces (solve.cpp:393-626); every face here has both sides */
  for (u64 ifi = 0; ifi < g->nifc; ++ifi)
  {
    s64 fi = g->ifl[5 * ifi], eL = g->ifl[5 * ifi + 1], eR = g->ifl[5 * ifi + 2];
    int lfL = (int)g->ifl[5 * ifi + 3], lfR = (int)g->ifl[5 * ifi + 4];
    const real* fevL = g->refp.feval + (u64)nbf * nfq * lfL;
    const real* fevR = g->refp.feval + (u64)nbf * nfq * lfR;
    memset(RL, 0, sizeof(real) * 15 * nbf);
    memset(RR, 0, sizeof(real) * 15 * nbf);
    for (int qi = 0; qi < nfq; ++qi)
    {
      const real* n = g->n_ + 3 * ((u64)nfq * fi + qi);
      real se = g->fJ_[(u64)nfq * fi + qi], qw = g->refp.fw[qi];
      real SL[5], SR[5], dL[5], dR[5];
      real QLx[5], QLy[5], QLz[5], QRx[5], QRy[5], QRz[5];
      for (int r = 0; r < 5; ++r)
      {
        SL[r] = ST(ULst, fi, qi, r);
        SR[r] = ST(URst, fi, qi, r);
      }
      for (int i = 0; i < 5; ++i)
      {
        dL[i] = SL[i] - SR[i];
        dR[i] = SR[i] - SL[i];
      }
      A_jump(SL, dL, n, 1.0, mu, QLx, QLy, QLz);
      A_jump(SR, dR, n, -1.0, mu, QRx, QRy, QRz);
      const real hf = (g_plant == ORC_PLANT_BR2_INT_HALF) ? 1.0 : 0.5;
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          real tL = fevL[nbf * qi + t], tR = fevR[nbf * qi + t];
          int ts = nbf * r + t;
          RL[dimos * 0 + ts] += (hf * tL * QLx[r]) * se * qw;
          RL[dimos * 1 + ts] += (hf * tL * QLy[r]) * se * qw;
          RL[dimos * 2 + ts] += (hf * tL * QLz[r]) * se * qw;
          RR[dimos * 0 + ts] += (hf * tR * QRx[r]) * se * qw;
          RR[dimos * 1 + ts] += (hf * tR * QRy[r]) * se * qw;
          RR[dimos * 2 + ts] += (hf * tR * QRz[r]) * se * qw;
        }
    }
    orc_spmm(nbf, 15, g->Minv + eL * nentry_minv(g), RL, 1., DL);
    orc_spmm(nbf, 15, g->Minv + eR * nentry_minv(g), RR, 1., DR);

    for (int qi = 0; qi < nfq; ++qi)
    {
      const real* n = g->n_ + 3 * ((u64)nfq * fi + qi);
      real se = g->fJ_[(u64)nfq * fi + qi], qw = g->refp.fw[qi];
      real UL[5], UR[5], SLx[5], SLy[5], SLz[5], SRx[5], SRy[5], SRz[5];
      for (int r = 0; r < 5; ++r)
      {
        UL[r]  = ST(ULst, fi, qi, r);
        UR[r]  = ST(URst, fi, qi, r);
        SLx[r] = ST(ULg, fi, qi, r);
        SLy[r] = ST(ULg, fi, qi, 5 + r);
        SLz[r] = ST(ULg, fi, qi, 10 + r);
        SRx[r] = ST(URg, fi, qi, r);
        SRy[r] = ST(URg, fi, qi, 5 + r);
        SRz[r] = ST(URg, fi, qi, 10 + r);
      }
      real F[5];
      orc_roe(UL, UR, n, gamma, F, NULL);

      real dcQLx[5], dcQLy[5], dcQLz[5], dcQRx[5], dcQRy[5], dcQRz[5];
      {
        real Uh[5], dfL[5], dfR[5];
        for (int i = 0; i < 5; ++i)
          Uh[i] = 0.5 * (UL[i] + UR[i]);
        for (int i = 0; i < 5; ++i)
        {
          dfL[i] = UL[i] - Uh[i];
          dfR[i] = UR[i] - Uh[i];
        }
        A_jump(UL, dfL, n, 1.0, mu, dcQLx, dcQLy, dcQLz);
        A_jump(UR, dfR, n, -1.0, mu, dcQRx, dcQRy, dcQRz);
      }

      real QH[5];
      {
        real QLx[5], QLy[5], QLz[5], QRx[5], QRy[5], QRz[5];
        real dLx[5] = {0}, dLy[5] = {0}, dLz[5] = {0};
        real dRx[5] = {0}, dRy[5] = {0}, dRz[5] = {0};
        for (int r = 0; r < 5; ++r)
          for (int b = 0; b < nbf; ++b)
          {
            int ts = nbf * r + b;
            real pL = fevL[nbf * qi + b], pR = fevR[nbf * qi + b];
            dLx[r] += pL * DL[dimos * 0 + ts];
            dLy[r] += pL * DL[dimos * 1 + ts];
            dLz[r] += pL * DL[dimos * 2 + ts];
            dRx[r] += pR * DR[dimos * 0 + ts];
            dRy[r] += pR * DR[dimos * 1 + ts];
            dRz[r] += pR * DR[dimos * 2 + ts];
          }
        orc_A(UL, SLx, SLy, SLz, mu, QLx, QLy, QLz);
        orc_A(UR, SRx, SRy, SRz, mu, QRx, QRy, QRz);
        for (int r = 0; r < 5; ++r)
        {
          const real eh = (g_plant == ORC_PLANT_BR2_ETA_HALF) ? 1.0 : 0.5;
          real QHx = 0.5 * (QLx[r] + QRx[r]) - eta * eh * (dLx[r] + dRx[r]);
          real QHy = 0.5 * (QLy[r] + QRy[r]) - eta * eh * (dLy[r] + dRy[r]);
          real QHz = 0.5 * (QLz[r] + QRz[r]) - eta * eh * (dLz[r] + dRz[r]);
          QH[r] = QHx * n[0] + QHy * n[1] + QHz * n[2];
        }
      }

      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          real tL = fevL[nbf * qi + t];
          const real* gL = g->fgrad + ((u64)nbf * (nfq * (6 * eL + lfL) + qi) + t) * 3;
          R[SIDX(eL, r, t)] += ((tL * F[r]) -
                                (gL[0] * dcQLx[r] + gL[1] * dcQLy[r] + gL[2] * dcQLz[r]) -
                                (tL * QH[r])) *
                               se * qw;
        }
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          real tR = fevR[nbf * qi + t];
          const real* gR = g->fgrad + ((u64)nbf * (nfq * (6 * eR + lfR) + qi) + t) * 3;
          real dcR = gR[0] * dcQRx[r] + gR[1] * dcQRy[r] + gR[2] * dcQRz[r];
          if (g_plant == ORC_PLANT_DC_SIGN)
            dcR = -dcR;
          real FR = (g_plant == ORC_PLANT_DROP_RFLUX) ? 0.0 : tR * F[r];
          R[SIDX(eR, r, t)] += (-FR - dcR + (tR * QH[r])) * se * qw;
        }
    }
  }

  /* boundary faces (solve.cpp:628-794) */
  for (u64 bfi = 0; bfi < g->nbfc; ++bfi)
  {
    s64 fi = g->bfl[4 * bfi], eL = g->bfl[4 * bfi + 1];
    int lfL = (int)g->bfl[4 * bfi + 3];
    const real* fevL = g->refp.feval + (u64)nbf * nfq * lfL;
    memset(RL, 0, sizeof(real) * 15 * nbf);
    for (int qi = 0; qi < nfq; ++qi)
    {
      const real* n = g->n_ + 3 * ((u64)nfq * fi + qi);
      real se = g->fJ_[(u64)nfq * fi + qi], qw = g->refp.fw[qi];
      real SL[5], Sb[5], dL[5], QLx[5], QLy[5], QLz[5];
      for (int r = 0; r < 5; ++r)
      {
        SL[r] = ST(ULst, fi, qi, r);
        Sb[r] = ST(Ubst, fi, qi, r);
      }
      for (int i = 0; i < 5; ++i)
        dL[i] = SL[i] - Sb[i];
      A_jump((g_plant == ORC_PLANT_BR2_WALL_SL) ? SL : Sb, dL, n, 1.0, mu, QLx, QLy, QLz);
      const real bh = (g_plant == ORC_PLANT_BR2_BND_HALF) ? 0.5 : 1.0;
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          real tL = fevL[nbf * qi + t];
          int ts    = nbf * r + t;
          RL[dimos * 0 + ts] += (bh * tL * QLx[r]) * se * qw;
          RL[dimos * 1 + ts] += (bh * tL * QLy[r]) * se * qw;
          RL[dimos * 2 + ts] += (bh * tL * QLz[r]) * se * qw;
        }
    }
    for (int en = 0; en < 15; ++en)
      orc_symmstore_mvmul(nbf, g->Minv + eL * nentry_minv(g), RL + nbf * en, 1., DL + nbf * en);

    for (int qi = 0; qi < nfq; ++qi)
    {
      const real* n = g->n_ + 3 * ((u64)nfq * fi + qi);
      real se = g->fJ_[(u64)nfq * fi + qi], qw = g->refp.fw[qi];
      real UL[5], Ub[5], SLx[5], SLy[5], SLz[5];
      for (int r = 0; r < 5; ++r)
      {
        UL[r]  = ST(ULst, fi, qi, r);
        Ub[r]  = ST(Ubst, fi, qi, r);
        SLx[r] = ST(ULg, fi, qi, r);
        SLy[r] = ST(ULg, fi, qi, 5 + r);
        SLz[r] = ST(ULg, fi, qi, 10 + r);
      }
      real F[5];
      {
        real Fx[5], Fy[5], Fz[5];
        orc_analytical_flux(Ub, gamma, Fx, Fy, Fz);
        for (int r = 0; r < 5; ++r)
          F[r] = Fx[r] * n[0] + Fy[r] * n[1] + Fz[r] * n[2];
      }
      real dcQLx[5], dcQLy[5], dcQLz[5];
      {
        real dfL[5];
        for (int i = 0; i < 5; ++i)
          dfL[i] = UL[i] - Ub[i];
        A_jump(UL, dfL, n, 1.0, mu, dcQLx, dcQLy, dcQLz);
      }
      real QH[5];
      {
        real QLx[5], QLy[5], QLz[5];
        real dLx[5] = {0}, dLy[5] = {0}, dLz[5] = {0};
        for (int r = 0; r < 5; ++r)
          for (int b = 0; b < nbf; ++b)
          {
            int ts = nbf * r + b;
            real pL = fevL[nbf * qi + b];
            dLx[r] += pL * DL[dimos * 0 + ts];
            dLy[r] += pL * DL[dimos * 1 + ts];
            dLz[r] += pL * DL[dimos * 2 + ts];
          }
        orc_A(UL, SLx, SLy, SLz, mu, QLx, QLy, QLz);
        for (int r = 0; r < 5; ++r)
        {
          real QHx = QLx[r] - eta * dLx[r];
          real QHy = QLy[r] - eta * dLy[r];
          real QHz = QLz[r] - eta * dLz[r];
          QH[r] = QHx * n[0] + QHy * n[1] + QHz * n[2];
        }
      }
      for (int r = 0; r < 5; ++r)
        for (int t = 0; t < nbf; ++t)
        {
          real tL = fevL[nbf * qi + t];
          const real* gL = g->fgrad + ((u64)nbf * (nfq * (6 * eL + lfL) + qi) + t) * 3;
          R[SIDX(eL, r, t)] += ((tL * F[r]) -
                                (gL[0] * dcQLx[r] + gL[1] * dcQLy[r] + gL[2] * dcQLz[r]) -
                                (tL * QH[r])) *
                               se * qw;
        }
    }
  }
#undef ST

  /* f = -Minv R (solve.cpp:800-804) */
  for (u64 e = 0; e < nelem; ++e)
    orc_spmm(nbf, 5, g->Minv + e * nentry_minv(g), R + e * 5 * nbf, -1., f + e * 5 * nbf);

  free(RL); free(RR); free(DL); free(DR);
  free(ULst); free(URst); free(Ubst); free(ULg); free(URg);
  return status;
}
#undef SIDX

/* ------------------------------------------------------------------------ */
/* explicit time stepping (time_distributed.cpp:23-102)                      */
/* ------------------------------------------------------------------------ */

/* RMS of R over every entry (time_distributed.cpp:23-44, one rank) */
real orc_residual_norm(const real* R, u64 n)
{
  real s = 0.;
  for (u64 i = 0; i < n; ++i)
    s += R[i] * R[i];
  return sqrt((1. / ((real)n)) * s);
}

enum { ORC_RK4 = 0, ORC_TVDRK3 = 1 };

/* One step of classic RK4 (scheme 0, time_distributed.cpp:46-77) or
 * Shu-Osher TVD-RK3 (scheme 1, time_distributed.cpp:79-102).  norm_out
 * receives RMS(R) of the first stage.  Returns nonzero on a boundary-state
 * failure or a NaN norm (the reference aborts there). */
int orc_step(const orc_geom* g, const orc_params* prm, int scheme, real dt,
             real* U, real* norm_out)
{
  u64 N = g->nelem * 5 * (u64)g->refp.nbf;
  int st = 0;
  real* R = (real*)zalloc(sizeof(real) * N);
  if (scheme == ORC_RK4)
  {
    real* U1 = (real*)zalloc(sizeof(real) * N * 7);
    real *U2 = U1 + N, *U3 = U1 + 2 * N, *f1 = U1 + 3 * N, *f2 = U1 + 4 * N;
    real *f3 = U1 + 5 * N, *f4 = U1 + 6 * N;
    st |= orc_residual(g, prm, U, R, f1);
    real nrm = orc_residual_norm(R, N);
    if (norm_out) *norm_out = nrm;
    if (isnan(nrm)) st |= 4;
    for (u64 i = 0; i < N; ++i) U1[i] = U[i] + dt * f1[i] * 0.5;
    st |= orc_residual(g, prm, U1, R, f2);
    for (u64 i = 0; i < N; ++i) U2[i] = U[i] + dt * f2[i] * 0.5;
    st |= orc_residual(g, prm, U2, R, f3);
    {
      real c3 = (g_plant == ORC_PLANT_RK4_HALF) ? 0.5 : 1.0;
      for (u64 i = 0; i < N; ++i) U3[i] = U[i] + dt * f3[i] * c3;
    }
    st |= orc_residual(g, prm, U3, R, f4);
    for (u64 i = 0; i < N; ++i)
      U[i] = U[i] + ((real)1 / 6) * (f1[i] + 2. * f2[i] + 2. * f3[i] + f4[i]) * dt;
    free(U1);
  }
  else
  {
    real* U1 = (real*)zalloc(sizeof(real) * N * 3);
    real *U2 = U1 + N, *f = U1 + 2 * N;
    st |= orc_residual(g, prm, U, R, f);
    real nrm = orc_residual_norm(R, N);
    if (norm_out) *norm_out = nrm;
    if (isnan(nrm)) st |= 4;
    for (u64 i = 0; i < N; ++i) U1[i] = U[i] + dt * f[i];
    st |= orc_residual(g, prm, U1, R, f);
    for (u64 i = 0; i < N; ++i) U2[i] = 0.75 * U[i] + 0.25 * U1[i] + 0.25 * dt * f[i];
    st |= orc_residual(g, prm, U2, R, f);
    for (u64 i = 0; i < N; ++i)
      U[i] = ((real)1 / 3) * U[i] + ((real)2 / 3) * U2[i] + ((real)2 / 3) * dt * f[i];
    free(U1);
  }
  free(R);
  return st;
}

/* The same RK combinations driven by a caller-supplied right-hand side, so
 * the stage algebra can be checked on y' = lambda*y. */
typedef void (*orc_rhs_fn)(const real* U, real* f, void* ctx);

void orc_step_generic(int scheme, real dt, u64 N, real* U, orc_rhs_fn rhs,
                      void* ctx)
{
  if (scheme == ORC_RK4)
  {
    real* U1 = (real*)zalloc(sizeof(real) * N * 7);
    real *U2 = U1 + N, *U3 = U1 + 2 * N, *f1 = U1 + 3 * N, *f2 = U1 + 4 * N;
    real *f3 = U1 + 5 * N, *f4 = U1 + 6 * N;
    rhs(U, f1, ctx);
    for (u64 i = 0; i < N; ++i) U1[i] = U[i] + dt * f1[i] * 0.5;
    rhs(U1, f2, ctx);
    for (u64 i = 0; i < N; ++i) U2[i] = U[i] + dt * f2[i] * 0.5;
    rhs(U2, f3, ctx);
    {
      real c3 = (g_plant == ORC_PLANT_RK4_HALF) ? 0.5 : 1.0;
      for (u64 i = 0; i < N; ++i) U3[i] = U[i] + dt * f3[i] * c3;
    }
    rhs(U3, f4, ctx);
    for (u64 i = 0; i < N; ++i)
      U[i] = U[i] + ((real)1 / 6) * (f1[i] + 2. * f2[i] + 2. * f3[i] + f4[i]) * dt;
    free(U1);
  }
  else
  {
    real* U1 = (real*)zalloc(sizeof(real) * N * 3);
    real *U2 = U1 + N, *f = U1 + 2 * N;
    rhs(U, f, ctx);
    for (u64 i = 0; i < N; ++i) U1[i] = U[i] + dt * f[i];
    rhs(U1, f, ctx);
    for (u64 i = 0; i < N; ++i) U2[i] = 0.75 * U[i] + 0.25 * U1[i] + 0.25 * dt * f[i];
    rhs(U2, f, ctx);
    for (u64 i = 0; i < N; ++i)
      U[i] = ((real)1 / 3) * U[i] + ((real)2 / 3) * U2[i] + ((real)2 / 3) * dt * f[i];
    free(U1);
  }
}

/* ---- derived outputs ------------------------------------------------------ */

/* eval_output (render/vrtxgen.cpp:25-59) on the (N+1)^3 equispaced sub-grid
 * of every element (reference points (i/N, j/N, k/N), sub-node index
 * (N+1)^2 k + (N+1) j + i, vrtxgen.cpp:294-330): the state is evaluated
 * from its equispaced-Lagrange coefficients and var selects mach (0), rho
 * (1), u (2), v (3), w (4).  The Mach number keeps the reference's pressure
 * (gamma - 1)(rho E - |u|^2 / 2), without the density factor. */
void orc_eval_output(int p, u64 nelem, const real* U, int N, int var, real gamma,
                     real* out)
{
  int n = p + 1, nbf = n * n * n, np1 = N + 1;
  real runit = 1.0 / (real)N;
  real* phi = (real*)zalloc(sizeof(real) * (size_t)np1 * np1 * np1 * nbf);
  for (int k = 0; k < np1; ++k)
    for (int j = 0; j < np1; ++j)
      for (int i = 0; i < np1; ++i)
      {
        real r[3] = {(real)i * runit, (real)j * runit, (real)k * runit};
        int node = np1 * np1 * k + np1 * j + i;
        for (int b = 0; b < nbf; ++b)
          phi[(size_t)nbf * node + b] = lagrange3d(b, p, r);
      }
  for (u64 e = 0; e < nelem; ++e)
    for (int node = 0; node < np1 * np1 * np1; ++node)
    {
      real s[5] = {0, 0, 0, 0, 0};
      for (int b = 0; b < nbf; ++b)
      {
        real ev = phi[(size_t)nbf * node + b];
        for (int r = 0; r < 5; ++r)
          s[r] += U[((size_t)e * 5 + r) * nbf + b] * ev;
      }
      real v = 0.0;
      switch (var)
      {
        case 0: {
          real u = s[1] / s[0], vv = s[2] / s[0], w = s[3] / s[0];
          real sp = sqrt(u * u + vv * vv + w * w);
          real pr = (gamma - 1.0) * (s[4] - 0.5 * sp * sp);
          real c = sqrt(gamma * pr / s[0]);
          v = sp / c;
        }
        break;
        case 1: v = s[0]; break;
        case 2: v = s[1] / s[0]; break;
        case 3: v = s[2] / s[0]; break;
        case 4: v = s[3] / s[0]; break;
      }
      out[(size_t)e * np1 * np1 * np1 + node] = v;
    }
  free(phi);
}

/* update_mean (output.cpp:39-49): running mean of the state since mean_start. */
void orc_update_mean(const real* state, real* mean, u64 n, u64 time_step, u64 mean_start)
{
  real old_range = (real)(time_step - mean_start);
  real new_range = (real)(time_step + 1 - mean_start);
  for (u64 i = 0; i < n; ++i)
    mean[i] = (mean[i] * old_range + state[i]) / new_range;
}

/* turbulent (solver_init.cpp:30-73): free-stream density and energy; each
 * velocity component i gets amp * sin(2pi(x-0.3i)/L) sin(2pi(y-0.3i)/L)
 * sin(2pi(z-0.3i)/L) for L = l/2, l/4, l/8, l/16, amp = |u_fr|/10, at the
 * order-p Lagrange nodes. */
void orc_turbulent(const orc_geom* g, const orc_params* prm, real l, real* U)
{
  const real twopi = 2. * ORC_PI, offst = 0.3;
  const real rho = prm->Ufr[0];
  const real u = prm->Ufr[1] / rho, v = prm->Ufr[2] / rho, w = prm->Ufr[3] / rho;
  const real amp = sqrt(u * u + v * v + w * w) / 10.;
  const real scale[4] = {0.5, 0.25, 0.125, 0.0625};
  const int n = g->p + 1, nbf = n * n * n;
  for (u64 e = 0; e < g->nelem; ++e)
    for (int b = 0; b < nbf; ++b)
    {
      real r[3] = {orc_lagrange_node1d(b % n, g->p), orc_lagrange_node1d((b / n) % n, g->p),
                     orc_lagrange_node1d(b / (n * n), g->p)};
      real x[3];
      ref2glo(g, e, r, x);
      real* Ue = U + e * 5 * (u64)nbf;
      Ue[b] = prm->Ufr[0];
      Ue[4 * nbf + b] = prm->Ufr[4];
      for (int i = 0; i < 3; ++i)
      {
        real vel = prm->Ufr[1 + i] / rho;
        for (int k = 0; k < 4; ++k)
        {
          real L = scale[k] * l;
          vel += amp * (sin(twopi * ((x[0] - i * offst) / L)) *
                        sin(twopi * ((x[1] - i * offst) / L)) *
                        sin(twopi * ((x[2] - i * offst) / L)));
        }
        Ue[(1 + i) * nbf + b] = vel * rho;
      }
    }
}
