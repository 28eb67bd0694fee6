// dgx_general.hip -- the kernels of every order outside the tuned set: the
// residual + RK stage, the face trace records, the BR2 lifting coefficients,
// the basis conversions, the residual output / norm partials and the
// sampled outputs for solution order 0 <= p <= 8 with geometry order
// 1 <= q <= 8 (the reference's MAX_ORDER = 8, constants.cpp:21,
// geometry_distributed.cpp:819-830).
//
// Same method, record layout and stage algebra as the tuned kernels
// (dgx_stage.hpp, which describes them); the mapping differs:
//   * one Gauss node per lane, workgroup barriers between the phases: an
//     element's (p+1)^3 nodes on round_up((p+1)^3, 64) lanes (64 up to
//     p = 3, 128 at p = 4, 256 at p = 5, ..., 768 at p = 8), or at p <= 2,
//     where an element has at most 27 nodes, up to 64 / (p+1)^3 elements
//     per 64-lane workgroup (fewer when their geometry nodes would take
//     more than 32 KB of LDS);
//   * faces in passes of whole faces, one face point per lane;
//   * the geometry order is a run-time loop bound, the elements' nodes (up
//     to 9^3 x 3 doubles each) staged in dynamic LDS;
//   * the 1D tables of every (p, q) sit in one device array (built on the
//     host in long double) and are copied to LDS by each workgroup.
// The solution order is a template parameter (register arrays and loops
// over a line of nodes are sized by it).
#include "dgx_stage.hpp"
#include "dgx_dd.hpp"

namespace dgx
{

static __device__ GenTab g_gtab[9 * 8];  // [p * 8 + q - 1]

hipError_t upload_gen_tables(const GenTab* gt)
{
  return hipMemcpyToSymbol(HIP_SYMBOL(g_gtab), gt, sizeof(GenTab) * 72);
}

template <int P>
struct GDims
{
  static constexpr int N   = P + 1;
  static constexpr int NB  = N * N * N;
  static constexpr int NF  = N * N;
  // elements per workgroup at most (several when an element has <= 32 nodes)
  static constexpr int EPB = NB <= 32 ? 64 / NB : 1;
  static constexpr int BT  = NB <= 64 ? 64 : ((NB + 63) / 64) * 64;  // lanes per workgroup
  // whole faces per pass, one face point per lane (with one element per
  // workgroup at most its six faces)
  static constexpr int FPP = (EPB > 1 || BT / NF < 6) ? BT / NF : 6;
  static constexpr int NPASS1 = (6 + FPP - 1) / FPP;  // face passes of one element
  static constexpr int FS  = 17;  // doubles per face point in a face pass
  static constexpr int TS  = 9;   // doubles per face point while producing traces
  static constexpr int RG0 = 5 * NB * EPB;  // volume divergence of every element
  static constexpr int RG  = (RG0 > FS * FPP * NF) ? RG0 : FS * FPP * NF;  // exchange region
};

// Elements per workgroup for (p, q): the most that fit GDims::EPB and 32 KB
// of geometry nodes in LDS.
static int gen_epb_of(int epb_max, int q)
{
  const int n3 = (q + 1) * (q + 1) * (q + 1) * 3;
  int e = 32768 / (n3 * (int)sizeof(double));
  e = e < 1 ? 1 : e;
  return e < epb_max ? e : epb_max;
}

__device__ __forceinline__ int pick3i(int k, int a, int b, int c)
{
  return (k == 0) ? a : ((k == 1) ? b : c);
}

// J[d][c] = d x_d / d xi_c from the element's geometry nodes X (nq per
// direction, x fastest) and the 1D geometry-basis values v* / derivatives d*
// of the point in each direction.
__device__ __forceinline__ void gen_jacobian(const double* X, int nq, const double* vx,
                                             const double* dx, const double* vy,
                                             const double* dy, const double* vz,
                                             const double* dz, double J[3][3])
{
#pragma unroll
  for (int d = 0; d < 3; ++d)
    J[d][0] = J[d][1] = J[d][2] = 0.0;
  for (int cz = 0; cz < nq; ++cz)
    for (int cy = 0; cy < nq; ++cy)
      for (int cx = 0; cx < nq; ++cx)
      {
        const double* x = X + 3 * ((cz * nq + cy) * nq + cx);
        const double w0 = dx[cx] * vy[cy] * vz[cz];
        const double w1 = vx[cx] * dy[cy] * vz[cz];
        const double w2 = vx[cx] * vy[cy] * dz[cz];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
          J[d][0] += x[d] * w0;
          J[d][1] += x[d] * w1;
          J[d][2] += x[d] * w2;
        }
      }
}

// J at volume node (i, j, k) (nd < 0) or at the point of local face (nd, e)
// with Gauss indices (i, j, k) in the other two directions.
__device__ __forceinline__ void gen_jac_at(const GenTab& t, const double* X, int nq, int nd,
                                           int e, int i, int j, int k, double J[3][3])
{
  const double* vx = (nd == 0) ? t.Be[e] : t.Bq[i];
  const double* dx = (nd == 0) ? t.Bde[e] : t.Bqd[i];
  const double* vy = (nd == 1) ? t.Be[e] : t.Bq[j];
  const double* dy = (nd == 1) ? t.Bde[e] : t.Bqd[j];
  const double* vz = (nd == 2) ? t.Be[e] : t.Bq[k];
  const double* dz = (nd == 2) ? t.Bde[e] : t.Bqd[k];
  gen_jacobian(X, nq, vx, dx, vy, dy, vz, dz, J);
}

// J^-1, outward unit normal and surface element at a point of local face
// (nd, hi).
__device__ __forceinline__ double gen_face_geometry(const GenTab& t, const double* X, int nq,
                                                    int nd, int hi, int a, int b,
                                                    double Ji[3][3], double* nv)
{
  const int gi0 = (nd == 0) ? 0 : a;
  const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
  const int gi2 = (nd == 2) ? 0 : b;
  double J[3][3];
  gen_jac_at(t, X, nq, nd, hi, gi0, gi1, gi2, J);
  inv3(J, det3(J), Ji);
  return face_normal(J, nd, hi, nv);
}

// Face table of the workgroup's epb elements from e0: sNb[l][f] the
// neighbour across face f, or -1 - (boundary face index) with the boundary
// code in sBt[l][f].
__device__ __forceinline__ void gen_face_table(const MeshArgs& mesh, int e0, int epb,
                                               int (*sNb)[6], int (*sBt)[6])
{
  for (int x = threadIdx.x; x < epb * 6; x += blockDim.x)
  {
    const int l = x / 6, f = x % 6, e = e0 + l;
    int nb = 0, bt = -1;
    if (e < mesh.nelem)
    {
      const int ci = e % mesh.nx, cj = (e / mesh.nx) % mesh.ny, ck = e / (mesh.nx * mesh.ny);
      nb = neighbour(mesh, ci, cj, ck, f, &bt);
      if (nb < 0)
        nb = -1 - bface_index(mesh, ci, cj, ck, f);
    }
    sNb[l][f] = nb;
    sBt[l][f] = bt;
  }
}

// Copies the tables of (p, q) and the geometry nodes of the epb elements
// from e0 to LDS (sX[l * (q+1)^3 * 3 ...]).
template <int P>
__device__ __forceinline__ void gen_load_tables_nodes(const MeshArgs& mesh, int q, int e0,
                                                      int epb, GenTab* sT, double* sX)
{
  constexpr int BT = GDims<P>::BT;
  const double* src = reinterpret_cast<const double*>(&g_gtab[P * 8 + q - 1]);
  double* dst = reinterpret_cast<double*>(sT);
  for (int x = threadIdx.x; x < (int)(sizeof(GenTab) / sizeof(double)); x += BT)
    dst[x] = src[x];
  const int nq = q + 1, n3 = nq * nq * nq * 3;
  for (int x = threadIdx.x; x < epb * n3; x += BT)
  {
    const int l = x / n3, e = e0 + l;
    sX[x] = (e < mesh.nelem) ? mesh.nodes[(size_t)e * n3 + (x - l * n3)] : 0.0;
  }
}

// Face trace records of the epb elements from e0 whose states are in sV (see
// produce_traces in dgx_stage.hpp for the layout): the trace U (5) and the
// gradient traction (3) at every face point, and the boundary state on
// domain-boundary faces; in passes of FPP whole faces, one face point per
// lane.  Every lane calls.
template <int P, typename TT>
__device__ __forceinline__ void gen_produce_traces(const MeshArgs& mesh, const Phys& ph,
                                                   const GenTab& t,
                                                   const double (*sV)[5][GDims<P>::NB],
                                                   const double* sX, int nq, const int (*sNb)[6],
                                                   const int (*sBt)[6], double* sR, int e0,
                                                   int epb, TT* __restrict__ Tout, int* status)
{
  using D = GDims<P>;
  constexpr int N = D::N, NF = D::NF, FPP = D::FPP, TS = D::TS;
  const int n3 = nq * nq * nq * 3;
  // with one element per workgroup the pass count is a compile-time constant
  const int nfaces = (D::EPB == 1) ? 6 : epb * 6;
  const int npass  = (D::EPB == 1) ? D::NPASS1 : (nfaces + FPP - 1) / FPP;
  const int tid = threadIdx.x, slot = tid / NF, fpt = tid % NF;
  const int a = fpt % N, b = fpt / N, fbase = slot * NF;
  TT* Tb = Tout + (size_t)mesh.nelem * 6 * 8 * NF;
  bool ok = true;
  for (int pass = 0; pass < npass; ++pass)
  {
    const int g = pass * FPP + slot, l = g / 6, f = g % 6, e = e0 + l;
    const bool fact = (slot < FPP) && (g < nfaces) && (e < mesh.nelem);
    const int nd = f >> 1, hi = f & 1;
    // traces (5) and normal derivatives (4) of this pass's face points
    if (fact)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double v = 0.0, dv = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          const double x = sV[l][r][line_node<N>(nd, m, a, b)];
          v += t.L[hi][m] * x;
          if (r < 4)
            dv += t.Ld[hi][m] * x;
        }
        sR[tid * TS + r] = v;
        if (r < 4)
          sR[tid * TS + 5 + r] = dv;
      }
    }
    __syncthreads();
    if (fact)
    {
      double U[5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
        U[r] = sR[tid * TS + r];
      double Ji[3][3], nv[3];
      gen_face_geometry(t, sX + l * n3, nq, nd, hi, a, b, Ji, nv);
      double gr[3][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          d1 += t.Dg[a][m] * sR[(fbase + m + N * b) * TS + r];
          d2 += t.Dg[b][m] * sR[(fbase + a + N * m) * TS + r];
        }
        const double dr = sR[tid * TS + 5 + r];
        gr[0][r] = pick3(nd, dr, d1, d1);
        gr[1][r] = pick3(nd, d1, dr, d2);
        gr[2][r] = pick3(nd, d2, d2, dr);
      }
      double vg[3];
      grad_traction(gr, U, Ji, nv, ph.mu, vg);
      TT* T = Tout + ((size_t)e * 6 + f) * 8 * NF + fpt;
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

// Trace records of a Gauss-nodal state V: epb elements per workgroup.
template <int P, typename TT>
__global__ __launch_bounds__(GDims<P>::BT) void dgx_gen_trace_kernel(MeshArgs mesh, Phys ph,
                                                                    const double* __restrict__ V,
                                                                    TT* __restrict__ Tout,
                                                                    int* status, int q, int epb)
{
  using D = GDims<P>;
  constexpr int NB = D::NB, EPB = D::EPB;
  __shared__ double sV[EPB][5][NB];
  __shared__ double sR[D::RG];
  __shared__ GenTab sT;
  __shared__ int sNb[EPB][6], sBt[EPB][6];
  extern __shared__ double sX[];  // [epb][(q+1)^3 * 3]
  const int ep  = (EPB == 1) ? 1 : epb;
  const int tid = threadIdx.x, e0 = blockIdx.x * ep;
  const int le  = (EPB == 1) ? 0 : tid / NB;
  const int nd0 = (EPB == 1) ? tid : tid % NB;
  const int e   = e0 + le;
  gen_load_tables_nodes<P>(mesh, q, e0, ep, &sT, sX);
  gen_face_table(mesh, e0, ep, sNb, sBt);
  if ((EPB == 1 ? tid < NB : le < ep) && e < mesh.nelem)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = V[((size_t)e * 5 + r) * NB + nd0];
  __syncthreads();
  gen_produce_traces<P, TT>(mesh, ph, sT, sV, sX, q + 1, sNb, sBt, sR, e0, ep, Tout, status);
}

// One right-hand-side evaluation f_g of the epb elements from blockIdx.x *
// epb, fused with the RK stage combination of StageArgs (see stage_body_lds
// in dgx_stage.hpp: the same phases and face-point physics): volume terms,
// face terms from the trace records Tin, the epilogue, and the trace records
// of Vout into Tout.
template <int P, typename TT>
__global__ __launch_bounds__(GDims<P>::BT) void dgx_gen_stage_kernel(MeshArgs mesh, Phys ph,
                                                                    StageArgs sa,
                                                                    const double* __restrict__ lam,
                                                                    int* status, int q, int epb)
{
  using D = GDims<P>;
  constexpr int N = D::N, NB = D::NB, NF = D::NF, FS = D::FS, FPP = D::FPP, EPB = D::EPB;
  __shared__ double sV[EPB][5][NB];
  __shared__ double sR[D::RG];
  __shared__ GenTab sT;
  __shared__ int sNb[EPB][6], sBt[EPB][6];
  extern __shared__ double sX[];  // [epb][(q+1)^3 * 3]
  const int tid = threadIdx.x, nq = q + 1, n3 = nq * nq * nq * 3;
  // this lane's node: node nd0 of element le (le < epb <= EPB); with one
  // element per workgroup every index below is the compile-time form
  const int ep  = (EPB == 1) ? 1 : epb;
  const int e0  = blockIdx.x * ep;
  const int le  = (EPB == 1) ? 0 : tid / NB;
  const int nd0 = (EPB == 1) ? tid : tid % NB;
  const int e   = e0 + le;
  const TT* __restrict__ Tin = reinterpret_cast<const TT*>(sa.Tin);
  gen_load_tables_nodes<P>(mesh, q, e0, ep, &sT, sX);
  gen_face_table(mesh, e0, ep, sNb, sBt);
  const bool vact = (EPB == 1) ? (tid < NB && e0 < mesh.nelem) : (le < ep && e < mesh.nelem);
  if (vact)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = sa.Vin[((size_t)e * 5 + r) * NB + nd0];
  __syncthreads();
  const GenTab& t = sT;

  // ---- volume: gradient, fluxes ----
  const int i0 = nd0 % N, i1 = (nd0 / N) % N, i2 = nd0 / (N * N);
  double wJ = 1.0;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double Ji[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  double H[3][5];
  if (vact)
  {
    double J[3][3];
    gen_jac_at(t, sX + le * n3, nq, -1, 0, i0, i1, i2, J);
    const double det = det3(J);
    inv3(J, det, Ji);
    wJ = t.w[i0] * t.w[i1] * t.w[i2] * det;
    double Vq[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
      Vq[r] = sV[le][r][nd0];
    double gr[3][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
      double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        g0 += t.Dg[i0][m] * sV[le][r][nid<N>(m, i1, i2)];
        g1 += t.Dg[i1][m] * sV[le][r][nid<N>(i0, m, i2)];
        g2 += t.Dg[i2][m] * sV[le][r][nid<N>(i0, i1, m)];
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
#pragma unroll
      for (int r = 0; r < 5; ++r)
        sR[(le * 5 + r) * NB + nd0] =
            wJ * (Ji[c][0] * H[0][r] + Ji[c][1] * H[1][r] + Ji[c][2] * H[2][r]);
    __syncthreads();
    if (vact)
    {
      const int ic = pick3i(c, i0, i1, i2);
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          const int nn = (c == 0) ? nid<N>(m, i1, i2) : ((c == 1) ? nid<N>(i0, m, i2) : nid<N>(i0, i1, m));
          s += t.Dg[m][ic] * sR[(le * 5 + r) * NB + nn];
        }
        acc[r] += s;
      }
    }
    __syncthreads();
  }

  // ---- faces, in passes of FPP whole faces of the workgroup's elements
  // (face point layout in sR as in the tuned kernel: G 0-4, normal
  // contravariant 5-8, tangential 9-16) ----
  const int nfaces = ep * 6;
  const int npass  = (EPB == 1) ? D::NPASS1 : (nfaces + FPP - 1) / FPP;
  for (int pass = 0; pass < npass; ++pass)
  {
    const int slot = tid / NF;
    const int g    = pass * FPP + slot;  // face g % 6 of element g / 6
    const int fpt  = tid % NF;
    const int a = fpt % N, b = fpt / N;
    const int l = g / 6, f = g % 6, ee = e0 + l;
    const bool fact = (slot < FPP) && (g < nfaces) && (ee < mesh.nelem);
    const int nd = f >> 1, hi = f & 1;
    const int fbase = slot * NF;
    FacePoint fpo;
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
      double JiL[3][3], nv[3];
      const double fJ = gen_face_geometry(t, sX + l * n3, nq, nd, hi, a, b, JiL, nv);
      const double lamL = lam[((size_t)ee * 6 + f) * NF + fpt];
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
      const double sc = fJ * t.w[a] * t.w[b];
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
      for (int r = 0; r < 5; ++r)
      {
        fpo.G[r] *= sc;
        sR[tid * FS + r] = fpo.G[r];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        sR[tid * FS + 5 + r]  = pick3(nd, hc[0][r], hc[1][r], hc[2][r]);
        sR[tid * FS + 9 + r]  = pick3(nd, hc[1][r], hc[0][r], hc[0][r]);
        sR[tid * FS + 13 + r] = pick3(nd, hc[2][r], hc[2][r], hc[1][r]);
      }
    }
    __syncthreads();
    // T = G - Dg^T(t1) H_t1 - Dg^T(t2) H_t2 (slots 1-4; others read 9-16)
    if (fact)
    {
#pragma unroll
      for (int r = 1; r < 5; ++r)
      {
        double tr = fpo.G[r];
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          tr -= t.Dg[m][a] * sR[(fbase + m + N * b) * FS + 8 + r];
          tr -= t.Dg[m][b] * sR[(fbase + a + N * m) * FS + 12 + r];
        }
        sR[tid * FS + r] = tr;
      }
    }
    __syncthreads();
    // gather this pass's faces at the nodes
    if (vact)
    {
#pragma unroll
      for (int ff = 0; ff < 6; ++ff)
      {
        const int gg = le * 6 + ff;
        if (gg / FPP != pass)
          continue;
        const int fnd = ff >> 1, fhi = ff & 1;
        const int m  = pick3i(fnd, i0, i1, i2);
        const int t1 = (fnd == 0) ? i1 : i0;                   // first tangential index
        const int t2 = (fnd == 2) ? i1 : i2;                   // second tangential index
        const int fp = (gg % FPP) * NF + t2 * N + t1;
        acc[0] += t.L[fhi][m] * sR[fp * FS];
#pragma unroll
        for (int r = 1; r < 5; ++r)
          acc[r] += t.L[fhi][m] * sR[fp * FS + r] - t.Ld[fhi][m] * sR[fp * FS + 4 + r];
      }
    }
    __syncthreads();
  }

  // ---- form f_g and the RK epilogue ----
  if (vact)
  {
    const double iwJ = rcp(wJ);
    const size_t base = (size_t)e * 5 * NB + nd0;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      const size_t ix = base + (size_t)r * NB;
      const double fv = ph.source[r] - acc[r] * iwJ;
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
    gen_produce_traces<P, TT>(mesh, ph, t, sV, sX, nq, sNb, sBt, sR, e0, ep,
                              reinterpret_cast<TT*>(sa.Tout), status);
  }
}

// BR2 lifting coefficient of each element side of each face point (see
// dgx_lift_coef_kernel): fJ * sum_m L_hi[m]^2 / (w_m detJ(line node m)).
__global__ __launch_bounds__(256) void dgx_gen_lift_coef_kernel(MeshArgs mesh, double* lam, int p,
                                                               int q)
{
  const GenTab& t = g_gtab[p * 8 + q - 1];
  const int N = p + 1, NF = N * N, nq = q + 1, n3 = nq * nq * nq * 3;
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)mesh.nelem * 6 * NF)
    return;
  const int fpt = (int)(x % NF);
  const int f   = (int)((x / NF) % 6);
  const size_t e = x / (6 * NF);
  const int a = fpt % N, b = fpt / N, nd = f >> 1, hi = f & 1;
  const double* X = mesh.nodes + e * n3;
  const int gi0 = (nd == 0) ? 0 : a;
  const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
  const int gi2 = (nd == 2) ? 0 : b;
  double J[3][3];
  gen_jac_at(t, X, nq, nd, hi, gi0, gi1, gi2, J);
  const int c1 = (nd + 1) % 3, c2 = (nd + 2) % 3;
  const double n0 = J[1][c1] * J[2][c2] - J[2][c1] * J[1][c2];
  const double n1 = J[2][c1] * J[0][c2] - J[0][c1] * J[2][c2];
  const double n2 = J[0][c1] * J[1][c2] - J[1][c1] * J[0][c2];
  const double fJ = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  double s = 0.0;
  for (int m = 0; m < N; ++m)
  {
    const int k0 = (nd == 0) ? m : gi0, k1 = (nd == 1) ? m : gi1, k2 = (nd == 2) ? m : gi2;
    double Jm[3][3];
    gen_jac_at(t, X, nq, -1, 0, k0, k1, k2, Jm);
    s += t.L[hi][m] * t.L[hi][m] / (t.w[m] * det3(Jm));
  }
  lam[x] = fJ * s;
}

// Applies a 1D matrix along direction dir (M^T with trans) to the five
// variables of the lane's element s (node = this lane's node); every lane
// calls (two barriers).
template <int N>
__device__ __forceinline__ void gen_apply_dir(double (*s)[N * N * N], const double (*M)[GMAXN],
                                              int dir, bool trans, int node, bool act)
{
  double out[5];
  const int i0 = node % N, i1 = (node / N) % N, i2 = node / (N * N);
  const int id = pick3i(dir, i0, i1, i2);
  if (act)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        const int nn = (dir == 0) ? nid<N>(m, i1, i2) : ((dir == 1) ? nid<N>(i0, m, i2) : nid<N>(i0, i1, m));
        acc += (trans ? M[m][id] : M[id][m]) * s[r][nn];
      }
      out[r] = acc;
    }
  }
  __syncthreads();
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      s[r][node] = out[r];
  __syncthreads();
}

// Basis conversions of one element per workgroup (modes as dgx_basis_kernel:
// 0 V = B3 U, 1 U = B3^-1 V, 2 G = B3^T g, 3 g = B3^-T G), in double-double.
template <int P>
__global__ __launch_bounds__(GDims<P>::BT) void dgx_gen_basis_kernel(const double* in, double* out,
                                                                    int mode)
{
  constexpr int N = P + 1, NB = N * N * N;
  const GenTab& t = g_gtab[P * 8];
  __shared__ double sh[5][NB];
  __shared__ double sl[5][NB];
  const int node = threadIdx.x;
  const size_t e = blockIdx.x;
  const bool act = node < NB;
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      sh[r][node] = in[(e * 5 + r) * NB + node];
      sl[r][node] = 0.0;
    }
  __syncthreads();
  const bool inv = (mode == 1 || mode == 3);
  const double(*M)[GMAXN]  = inv ? t.Binv : t.B;
  const double(*Ml)[GMAXN] = inv ? t.Binvlo : t.Blo;
  const bool tr = (mode >= 2);
  const int i0 = node % N, i1 = (node / N) % N, i2 = node / (N * N);
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    DD o[5];
    if (act)
    {
      const int id = pick3i(d, i0, i1, i2);
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        DD acc = {0.0, 0.0};
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          const int src = (d == 0) ? nid<N>(m, i1, i2) : ((d == 1) ? nid<N>(i0, m, i2) : nid<N>(i0, i1, m));
          const double c  = tr ? M[m][id] : M[id][m];
          const double cl = tr ? Ml[m][id] : Ml[id][m];
          acc = dd_fma({sh[r][src], sl[r][src]}, c, cl, acc);
        }
        o[r] = acc;
      }
    }
    __syncthreads();
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        sh[r][node] = o[r].hi;
        sl[r][node] = o[r].lo;
      }
    __syncthreads();
  }
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      out[(e * 5 + r) * NB + node] = sh[r][node] + sl[r][node];
}

// From f_g of the epb elements from blockIdx.x * epb: f = B3^-1 f_g,
// R = -B3^T W f_g (either may be null) and, with partial,
// partial[blockIdx.x] = sum of R^2 over those elements (waves summed in a
// fixed order: deterministic).
template <int P>
__global__ __launch_bounds__(GDims<P>::BT) void dgx_gen_resid_out_kernel(MeshArgs mesh,
                                                                        const double* fg,
                                                                        double* f, double* R,
                                                                        double* partial, int q,
                                                                        int epb)
{
  using D = GDims<P>;
  constexpr int N = D::N, NB = D::NB, BT = D::BT, EPB = D::EPB;
  const GenTab& t = g_gtab[P * 8 + q - 1];
  __shared__ double s[EPB][5][NB];
  __shared__ double red[BT / 64];
  extern __shared__ double sX[];  // [epb][(q+1)^3 * 3]
  const int tid = threadIdx.x, nq = q + 1, n3 = nq * nq * nq * 3;
  const int e0 = blockIdx.x * epb;
  const int le = tid / NB, node = tid % NB;
  const size_t e = (size_t)e0 + le;
  const bool act = le < epb && e < (size_t)mesh.nelem;
  double (*se)[NB] = s[act ? le : 0];
  for (int x = tid; x < epb * n3; x += BT)
  {
    const int l = x / n3;
    sX[x] = (e0 + l < mesh.nelem) ? mesh.nodes[(size_t)(e0 + l) * n3 + (x - l * n3)] : 0.0;
  }
  double fgv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      fgv[r] = fg[(e * 5 + r) * NB + node];
      se[r][node] = fgv[r];
    }
  __syncthreads();
  if (f)
  {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gen_apply_dir<N>(se, t.Binv, d, false, node, act);
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
        f[(e * 5 + r) * NB + node] = se[r][node];
    __syncthreads();
  }
  double sq = 0.0;
  if (R || partial)
  {
    const int i0 = node % N, i1 = (node / N) % N, i2 = node / (N * N);
    if (act)
    {
      double J[3][3];
      gen_jac_at(t, sX + le * n3, nq, -1, 0, i0, i1, i2, J);
      const double wJ = t.w[i0] * t.w[i1] * t.w[i2] * det3(J);
#pragma unroll
      for (int r = 0; r < 5; ++r)
        se[r][node] = -wJ * fgv[r];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gen_apply_dir<N>(se, t.B, d, true, node, act);
    if (act)
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        const double v = se[r][node];
        if (R)
          R[(e * 5 + r) * NB + node] = v;
        sq += v * v;
      }
  }
  if (partial)
  {
#pragma unroll
    for (int w = 32; w > 0; w >>= 1)
      sq += __shfl_xor(sq, w, 64);
    if ((tid & 63) == 0)
      red[tid >> 6] = sq;
    __syncthreads();
    if (tid == 0)
    {
      double tot = 0.0;
      for (int w = 0; w < BT / 64; ++w)
        tot += red[w];
      partial[blockIdx.x] = tot;
    }
  }
}

// eval_output on the (Ns+1)^3 equispaced sub-grid of every element (see
// dgx_sample_kernel): one thread per sample point.
template <int P>
__global__ __launch_bounds__(256) void dgx_gen_sample_kernel(const double* __restrict__ V,
                                                            int nelem, int Ns, int var,
                                                            double gamma,
                                                            double* __restrict__ out)
{
  constexpr int N = P + 1, NB = N * N * N;
  const GenTab& t = g_gtab[P * 8];
  const int np1 = Ns + 1, nsub = np1 * np1 * np1;
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)nelem * nsub)
    return;
  const int node = (int)(x % nsub);
  const size_t e = x / nsub;
  const int ijk[3] = {node % np1, (node / np1) % np1, node / (np1 * np1)};
  double l[3][N];
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    const double xc = (double)ijk[c] / (double)Ns;
#pragma unroll
    for (int m = 0; m < N; ++m)
    {
      double v = 1.0;
#pragma unroll
      for (int o = 0; o < N; ++o)
        if (o != m)
          v *= (xc - t.g[o]) / (t.g[m] - t.g[o]);
      l[c][m] = v;
    }
  }
  double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double* Ve = V + e * 5 * NB;
  for (int kk = 0; kk < N; ++kk)
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
  out[x] = v;
}

// ---- launch helpers (host) -------------------------------------------------

#define DGX_GEN_DISPATCH_P(p, CALL)                                                         \
  switch (p)                                                                                \
  {                                                                                         \
    case 0: return CALL(0);                                                                 \
    case 1: return CALL(1);                                                                 \
    case 2: return CALL(2);                                                                 \
    case 3: return CALL(3);                                                                 \
    case 4: return CALL(4);                                                                 \
    case 5: return CALL(5);                                                                 \
    case 6: return CALL(6);                                                                 \
    case 7: return CALL(7);                                                                 \
    case 8: return CALL(8);                                                                 \
  }                                                                                         \
  return hipErrorInvalidValue;

static bool gen_orders_ok(int p, int q)
{
  return p >= 0 && p <= 8 && q >= 1 && q <= 8;
}

// dynamic LDS of epb elements' geometry nodes
static size_t gen_node_lds(int q, int epb)
{
  return (size_t)epb * (q + 1) * (q + 1) * (q + 1) * 3 * sizeof(double);
}

static const int g_gen_epb_max[9] = {GDims<0>::EPB, GDims<1>::EPB, GDims<2>::EPB,
                                     GDims<3>::EPB, GDims<4>::EPB, GDims<5>::EPB,
                                     GDims<6>::EPB, GDims<7>::EPB, GDims<8>::EPB};

int gen_elements_per_block(int p, int q)
{
  if (!gen_orders_ok(p, q))
    return 1;
  return gen_epb_of(g_gen_epb_max[p], q);
}

template <int P>
static hipError_t launch_gen_rhs(int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                                 const double* lam, int* status, hipStream_t st)
{
  const int epb = gen_epb_of(GDims<P>::EPB, q);
  const dim3 grid((m.nelem + epb - 1) / epb), block(GDims<P>::BT);
  if (sa.trace32)
    hipLaunchKernelGGL((dgx_gen_stage_kernel<P, float>), grid, block, gen_node_lds(q, epb), st,
                       m, ph, sa, lam, status, q, epb);
  else
    hipLaunchKernelGGL((dgx_gen_stage_kernel<P, double>), grid, block, gen_node_lds(q, epb), st,
                       m, ph, sa, lam, status, q, epb);
  return hipGetLastError();
}

hipError_t gen_rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                            const double* lam, int* status, hipStream_t st)
{
  if (!gen_orders_ok(p, q) || m.nelem < 1)
    return hipErrorInvalidValue;
#define CALL_GRHS(P) launch_gen_rhs<P>(q, m, ph, sa, lam, status, st)
  DGX_GEN_DISPATCH_P(p, CALL_GRHS)
#undef CALL_GRHS
}

template <int P>
static hipError_t launch_gen_trace(int q, const MeshArgs& m, const Phys& ph, const double* V,
                                   double* Tout, int* status, hipStream_t st, bool trace32)
{
  const int epb = gen_epb_of(GDims<P>::EPB, q);
  const dim3 grid((m.nelem + epb - 1) / epb), block(GDims<P>::BT);
  if (trace32)
    hipLaunchKernelGGL((dgx_gen_trace_kernel<P, float>), grid, block, gen_node_lds(q, epb), st,
                       m, ph, V, reinterpret_cast<float*>(Tout), status, q, epb);
  else
    hipLaunchKernelGGL((dgx_gen_trace_kernel<P, double>), grid, block, gen_node_lds(q, epb), st,
                       m, ph, V, Tout, status, q, epb);
  return hipGetLastError();
}

hipError_t gen_trace_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                              double* Tout, int* status, hipStream_t st, bool trace32)
{
  if (!gen_orders_ok(p, q) || m.nelem < 1)
    return hipErrorInvalidValue;
#define CALL_GTR(P) launch_gen_trace<P>(q, m, ph, V, Tout, status, st, trace32)
  DGX_GEN_DISPATCH_P(p, CALL_GTR)
#undef CALL_GTR
}

hipError_t gen_lift_coef_dispatch(int p, int q, const MeshArgs& m, double* lam, hipStream_t st)
{
  if (!gen_orders_ok(p, q) || m.nelem < 1)
    return hipErrorInvalidValue;
  const size_t total = (size_t)m.nelem * 6 * (p + 1) * (p + 1);
  hipLaunchKernelGGL(dgx_gen_lift_coef_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, st, m, lam, p, q);
  return hipGetLastError();
}

template <int P>
static hipError_t launch_gen_resid_out(int q, const MeshArgs& m, const double* fg, double* f,
                                       double* R, double* partial, hipStream_t st)
{
  const int epb = gen_epb_of(GDims<P>::EPB, q);
  hipLaunchKernelGGL((dgx_gen_resid_out_kernel<P>), dim3((m.nelem + epb - 1) / epb),
                     dim3(GDims<P>::BT), gen_node_lds(q, epb), st, m, fg, f, R, partial, q, epb);
  return hipGetLastError();
}

hipError_t gen_resid_out_dispatch(int p, int q, const MeshArgs& m, const double* fg, double* f,
                                  double* R, double* partial, hipStream_t st)
{
  if (!gen_orders_ok(p, q) || m.nelem < 1)
    return hipErrorInvalidValue;
#define CALL_GRO(P) launch_gen_resid_out<P>(q, m, fg, f, R, partial, st)
  DGX_GEN_DISPATCH_P(p, CALL_GRO)
#undef CALL_GRO
}

template <int P>
static hipError_t launch_gen_basis(const double* in, double* out, int nelem, int mode,
                                   hipStream_t st)
{
  hipLaunchKernelGGL((dgx_gen_basis_kernel<P>), dim3(nelem), dim3(GDims<P>::BT), 0, st, in, out,
                     mode);
  return hipGetLastError();
}

hipError_t gen_basis_dispatch(int p, const double* in, double* out, int nelem, int mode,
                              hipStream_t st)
{
  if (nelem < 1 || mode < 0 || mode > 3)
    return hipErrorInvalidValue;
#define CALL_GB(P) launch_gen_basis<P>(in, out, nelem, mode, st)
  DGX_GEN_DISPATCH_P(p, CALL_GB)
#undef CALL_GB
}

template <int P>
static hipError_t launch_gen_sample(const double* V, int nelem, int Ns, int var, double gamma,
                                    double* out, hipStream_t st)
{
  const size_t total = (size_t)nelem * (Ns + 1) * (Ns + 1) * (Ns + 1);
  hipLaunchKernelGGL((dgx_gen_sample_kernel<P>), dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, st, V, nelem, Ns, var, gamma, out);
  return hipGetLastError();
}

hipError_t gen_sample_dispatch(int p, const double* V, int nelem, int Ns, int var, double gamma,
                               double* out, hipStream_t st)
{
  if (nelem < 1 || Ns < 1)
    return hipErrorInvalidValue;
#define CALL_GS(P) launch_gen_sample<P>(V, nelem, Ns, var, gamma, out, st)
  DGX_GEN_DISPATCH_P(p, CALL_GS)
#undef CALL_GS
}

}  // namespace dgx
