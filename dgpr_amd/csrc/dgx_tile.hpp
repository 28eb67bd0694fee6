// dgx_tile.hpp -- the tiled stage kernel: R rows of whole elements per
// 64-lane wavefront, faces processed one reference direction at a time.
//
// Same method, same arithmetic per face point and per node as stage_body
// (dgx_kernels.hip), organised for lane occupancy and wave-uniform control:
//   * a tile is R rows; a row is the 64 / (p+1)^3 elements whose Gauss nodes
//     fill one wavefront (one element at p = 3), so each lane owns R nodes;
//   * the face work of a tile is done per reference direction d: the 2E faces
//     of direction d (low and high face of each of the tile's E elements) in
//     subpasses of 64 / (p+1)^2 whole faces, one face point per lane.  The
//     direction is a compile-time constant in every subpass, so geometry
//     selection, normals, the contravariant split of the dual-consistency
//     flux and every table index fold at compile time.  At p = 3 with R = 2
//     each subpass is exactly the 4 faces (64 points) of one direction of the
//     two elements: 3 full subpasses per 2 elements where the one-element
//     kernel needs 2 passes of 48 points per element;
//   * the trace records of the new state are produced the same way.
// LDS per face point: the tangential dual-consistency fluxes (8 doubles),
// then the projected face term and the normal flux (9 doubles), in one
// 64 x 9 region shared with the volume divergence and the trace exchange.
// Included by dgx_stage.hpp inside namespace dgx, after stage_body and its
// helpers.
#pragma once

#include "dgx_mfma.hpp"

// rows per tile of the tiled stage kernel at each order (0: none compiled):
// p = 3: two one-element rows, so each face subpass is the 64 points of one
// direction of two elements; p = 1: one row of 8 elements (16 faces of 4
// points per direction); p = 2: three rows of 2 elements (12 faces of 9
// points per direction in two subpasses)
constexpr int tile_rows(int p)
{
  return p == 3 ? 2 : (p == 2 ? 3 : (p == 1 ? 1 : 0));
}

template <int P, int R>
struct TileDims
{
  static constexpr int N    = P + 1;
  static constexpr int NB   = N * N * N;
  static constexpr int NF   = N * N;
  static constexpr int EPR  = 64 / NB;            // elements per row
  static constexpr int E    = EPR * R;            // elements per tile
  static constexpr int FPS  = 64 / NF;            // whole faces per subpass
  static constexpr int NSUB = (2 * E + FPS - 1) / FPS;  // subpasses per direction
  static constexpr int FX   = 9;                  // doubles per face point exchanged
  // exchange region: a face subpass (FX per lane); the residual norm uses
  // 6 per lane (one row's W f_g and the lane's running sum of R^2)
  static constexpr int SR   = FX * 64;
  static_assert(FX >= 6, "the residual norm needs 6 doubles per lane");
};

// tangential directions of faces normal to direction d (in increasing order)
template <int D>
struct Tang
{
  static constexpr int T1 = (D == 0) ? 1 : 0;
  static constexpr int T2 = (D == 2) ? 1 : 2;
};

// Face table of the tile's E elements (any E): sNb[l][f] as in face_table.
template <int E>
__device__ __forceinline__ void tile_face_table(const MeshArgs& mesh, int e0, int (*sNb)[6],
                                                int (*sBt)[6])
{
  for (int x = threadIdx.x; x < E * 6; x += 64)
  {
    const int l = x / 6, f = x % 6, ee = e0 + l;
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

// One subpass of direction D: face point physics of whole faces, the
// tangential (T) phase and the gather into the nodes' accumulators.
template <int P, int Q, int R, int D, typename TT, bool AFF>
__device__ __forceinline__ void tile_face_subpass(
    const MeshArgs& mesh, const Phys& ph, const PTab& pt, const QTab& qt, const TT* __restrict__ Tin,
    const double* __restrict__ lam, const double (*sX)[(Q + 1) * (Q + 1) * (Q + 1) * 3],
    const double (*sG)[GREC], const int (*sNb)[6], double* sR, int e0, int sp,
    double (&acc)[R][5], bool& ok)
{
  using TD = TileDims<P, R>;
  constexpr int N = TD::N, NB = TD::NB, NF = TD::NF, EPR = TD::EPR, FPS = TD::FPS;
  constexpr int T1 = Tang<D>::T1, T2 = Tang<D>::T2;
  const int tid   = threadIdx.x;
  const int slot  = tid / NF;
  const int g     = sp * FPS + slot;  // face of direction D: element g/2, side g%2
  const int fpt   = tid % NF;
  const int a = fpt % N, b = fpt / N;
  const int l = g >> 1, hi = g & 1, f = 2 * D + hi;
  const int ee = e0 + l;
  const bool fact = (slot < FPS) && (g < 2 * TD::E) && (ee < mesh.nelem);
  const int fbase = slot * NF;

  double G[5], hn[4];
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
    int gi[3];
    gi[D]  = 0;
    gi[T1] = a;
    gi[T2] = b;
    double JiL[3][3], nv[3];
    const double fJ = face_geometry<Q, AFF>(qt, sX[l], sG[l], D, hi, gi[0], gi[1], gi[2], JiL, nv);
    // affine meshes: the face's constant coefficient (48 B per element
    // instead of 6 (p+1)^2 doubles, read for the own and the neighbour side)
    const double lamL =
        AFF ? mesh.lamf[(size_t)ee * 6 + f] : lam[((size_t)ee * 6 + f) * NF + fpt];
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
    FacePoint fpo;
    face_point_physics(en >= 0, UL, Ub, vgL, vgR, nv, lamL, lamR, ph, fpo);

    // contravariant dual-consistency flux h_c = Ji[c] . dcQ (rows 1-4)
    const double irL  = rcp(UL[0]);
    const double uL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
    const double sc   = fJ * pt.w[a] * pt.w[b];
    const double* S   = fpo.S;
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
      G[r] = fpo.G[r] * sc;
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
      hn[r]              = hc[D][r];
      sR[tid * 8 + r]     = hc[T1][r];
      sR[tid * 8 + 4 + r] = hc[T2][r];
    }
  }
  __syncthreads();
  // T = G - Dg^T(t1) h_t1 - Dg^T(t2) h_t2 (the mass row has no viscous part)
  if (fact)
  {
#pragma unroll
    for (int r = 1; r < 5; ++r)
    {
      double t = G[r];
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        t -= pt.Dg[m][a] * sR[(fbase + m + N * b) * 8 + (r - 1)];
        t -= pt.Dg[m][b] * sR[(fbase + a + N * m) * 8 + 4 + (r - 1)];
      }
      G[r] = t;
    }
  }
  __syncthreads();
  if (fact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sR[tid * 9 + r] = G[r];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      sR[tid * 9 + 5 + r] = hn[r];
  }
  __syncthreads();
  // gather: each node takes its element's low and high face of direction D
  // when they belong to this subpass
#pragma unroll
  for (int rr = 0; rr < R; ++rr)
  {
    const int le = rr * EPR + tid / NB;
    const int nd0 = tid % NB;
    if (tid < EPR * NB && e0 + le < mesh.nelem)
    {
      const int idx[3] = {nd0 % N, (nd0 / N) % N, nd0 / (N * N)};
      const int m = idx[D];
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
        const int gg = 2 * le + h;
        if (gg / FPS != sp)
          continue;
        const int fp = (gg % FPS) * NF + idx[T2] * N + idx[T1];
        acc[rr][0] += pt.L[h][m] * sR[fp * 9];
#pragma unroll
        for (int r = 1; r < 5; ++r)
          acc[rr][r] += pt.L[h][m] * sR[fp * 9 + r] - pt.Ld[h][m] * sR[fp * 9 + 4 + r];
      }
    }
  }
  __syncthreads();
}

// Trace records (and boundary states) of faces of direction D in one
// subpass, from the tile's Gauss-nodal states in sV (layout [E][5][NB]).
template <int P, int Q, int R, int D, typename TT, bool AFF>
__device__ __forceinline__ void tile_trace_subpass(
    const MeshArgs& mesh, const Phys& ph, const PTab& pt, const QTab& qt,
    const double (*sV)[5][TileDims<P, R>::NB],
    const double (*sX)[(Q + 1) * (Q + 1) * (Q + 1) * 3], const double (*sG)[GREC],
    const int (*sNb)[6], const int (*sBt)[6], double* sR, int e0, int sp, TT* __restrict__ Tout,
    bool& ok)
{
  using TD = TileDims<P, R>;
  constexpr int N = TD::N, NF = TD::NF, FPS = TD::FPS;
  constexpr int T1 = Tang<D>::T1, T2 = Tang<D>::T2;
  TT* Tb = Tout + (size_t)mesh.nelem * 6 * 8 * NF;
  const int tid  = threadIdx.x;
  const int slot = tid / NF;
  const int g    = sp * FPS + slot;
  const int fpt  = tid % NF;
  const int a = fpt % N, b = fpt / N;
  const int l = g >> 1, hi = g & 1, f = 2 * D + hi;
  const int ee = e0 + l;
  const bool fact = (slot < FPS) && (g < 2 * TD::E) && (ee < mesh.nelem);
  const int fbase = slot * NF;

  double U[5], dn[4];
  if (fact)
  {
    int id[3];
    id[T1] = a;
    id[T2] = b;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double v = 0.0, dv = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        id[D] = m;
        const double x = sV[l][r][(id[2] * N + id[1]) * N + id[0]];
        v += pt.L[hi][m] * x;
        if (r < 4)
          dv += pt.Ld[hi][m] * x;
      }
      U[r] = v;
      if (r < 4)
      {
        dn[r]           = dv;
        sR[tid * 4 + r] = v;
      }
    }
  }
  __syncthreads();
  if (fact)
  {
    int gi[3];
    gi[D]  = 0;
    gi[T1] = a;
    gi[T2] = b;
    double Ji[3][3], nv[3];
    face_geometry<Q, AFF>(qt, sX[l], sG[l], D, hi, gi[0], gi[1], gi[2], Ji, nv);
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
      gr[D][r]  = dn[r];
      gr[T1][r] = d1;
      gr[T2][r] = d2;
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

template <int P, int Q, int R, int D, typename TT, bool AFF>
__device__ __forceinline__ void tile_face_direction(
    const MeshArgs& mesh, const Phys& ph, const PTab& pt, const QTab& qt, const TT* Tin,
    const double* lam, const double (*sX)[(Q + 1) * (Q + 1) * (Q + 1) * 3],
    const double (*sG)[GREC], const int (*sNb)[6], double* sR, int e0, double (&acc)[R][5],
    bool& ok)
{
#pragma unroll
  for (int sp = 0; sp < TileDims<P, R>::NSUB; ++sp)
    tile_face_subpass<P, Q, R, D, TT, AFF>(mesh, ph, pt, qt, Tin, lam, sX, sG, sNb, sR, e0, sp,
                                           acc, ok);
}

template <int P, int Q, int R, int D, typename TT, bool AFF>
__device__ __forceinline__ void tile_trace_direction(
    const MeshArgs& mesh, const Phys& ph, const PTab& pt, const QTab& qt,
    const double (*sV)[5][TileDims<P, R>::NB],
    const double (*sX)[(Q + 1) * (Q + 1) * (Q + 1) * 3], const double (*sG)[GREC],
    const int (*sNb)[6], const int (*sBt)[6], double* sR, int e0, TT* Tout, bool& ok)
{
#pragma unroll
  for (int sp = 0; sp < TileDims<P, R>::NSUB; ++sp)
    tile_trace_subpass<P, Q, R, D, TT, AFF>(mesh, ph, pt, qt, sV, sX, sG, sNb, sBt, sR, e0, sp,
                                            Tout, ok);
}

// One pass of R = -B3^T W f_g over one row of a tile: s[r][node] <-
// sum_m B[m][i_c] s[r][node with i_c = m] along direction c within each
// element (s holds 5 x 64 doubles, node = lane); every lane calls (two
// barriers).
template <int N, int NB>
__device__ __forceinline__ void tile_modal_pass(double* s, const double (*B)[MAXN], int c,
                                                const int (&idx)[3], bool act)
{
  const int tid  = threadIdx.x;
  const int base = (tid / NB) * NB;
  double out[5];
  __syncthreads();
  if (act)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double a = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        const int nn = (c == 0) ? nid<N>(m, idx[1], idx[2])
                     : ((c == 1) ? nid<N>(idx[0], m, idx[2]) : nid<N>(idx[0], idx[1], m));
        a += B[m][idx[c]] * s[r * 64 + base + nn];
      }
      out[r] = a;
    }
  }
  __syncthreads();
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      s[r * 64 + tid] = out[r];
}

// LDS of one tile, carved from a byte buffer (see StageLds).
template <int P, int Q, int R>
struct TileLds
{
  using TD = TileDims<P, R>;
  static constexpr int E = TD::E, NB = TD::NB, NG = (Q + 1) * (Q + 1) * (Q + 1);
  static constexpr size_t oV  = 0;
  static constexpr size_t oX  = oV + sizeof(double) * E * 5 * NB;
  static constexpr size_t oG  = oX + sizeof(double) * E * NG * 3;
  static constexpr size_t oR  = oG + sizeof(double) * E * GREC;
  static constexpr size_t oPT = oR + sizeof(double) * TD::SR;
  static constexpr size_t oQT = oPT + sizeof(PTab);
  static constexpr size_t oNb = oQT + sizeof(QTab);
  static constexpr size_t oBt = oNb + sizeof(int) * E * 6;
  static constexpr size_t bytes = oBt + sizeof(int) * E * 6;
};

// The tiled stage: f_g of the tile's E elements (tile number `block`)
// fused with the RK stage of StageArgs, then the trace records of the new
// state (see stage_body_lds for the meaning of every argument; the
// arithmetic per point is the same), in the LDS region lds
// (TileLds<P, Q, R>::bytes, 16-byte aligned).
// CT selects how the volume contractions are done: 0 in FP64 on the vector
// ALUs; at p = 3, 1 in FP32 and 2 in split BF16 on the matrix cores
// (dgx_mfma.hpp), the flux physics between them staying FP64.  NORM: the
// instance that also writes the tile's sum of R^2 to sa.normp (the first
// stage of a step with its residual norm); kept apart so that the other
// stages run code without it.
template <int P, int Q, int R, typename TT, bool AFF, int CT = 0, bool NORM = false>
__device__ __forceinline__ void tile_body_lds(const MeshArgs& mesh, const Phys& ph,
                                              const StageArgs& sa,
                                              const double* __restrict__ lam, int* status,
                                              int block, unsigned char* lds)
{
  static_assert(CT == 0 || P == 3, "matrix-core contractions need p = 3 (4-point lines)");
  using TD = TileDims<P, R>;
  using TL = TileLds<P, Q, R>;
  constexpr int N = TD::N, NB = TD::NB, EPR = TD::EPR, E = TD::E;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);

  auto sV  = reinterpret_cast<double(*)[5][NB]>(lds + TL::oV);
  auto sX  = reinterpret_cast<double(*)[NG * 3]>(lds + TL::oX);
  auto sG  = reinterpret_cast<double(*)[GREC]>(lds + TL::oG);
  double* sR = reinterpret_cast<double*>(lds + TL::oR);
  PTab& sPT = *reinterpret_cast<PTab*>(lds + TL::oPT);
  QTab& sQT = *reinterpret_cast<QTab*>(lds + TL::oQT);
  auto sNb = reinterpret_cast<int(*)[6]>(lds + TL::oNb);
  auto sBt = reinterpret_cast<int(*)[6]>(lds + TL::oBt);
  stage_tables<P, Q>(&sPT, &sQT);
  const PTab& pt = sPT;
  const QTab& qt = sQT;

  const int tid = threadIdx.x;
  const int e0  = block * E;
  const int nd0 = tid % NB;
  const TT* __restrict__ Tin = reinterpret_cast<const TT*>(sa.Tin);
  tile_face_table<E>(mesh, e0, sNb, sBt);

  // ---- load the tile's states and element geometry ----
#pragma unroll
  for (int rr = 0; rr < R; ++rr)
  {
    const int le = rr * EPR + tid / NB;
    if (tid < EPR * NB && e0 + le < mesh.nelem)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
        sV[le][r][nd0] = sa.Vin[((size_t)(e0 + le) * 5 + r) * NB + nd0];
    }
  }
  for (int x = tid; x < E * NG * 3; x += 64)
  {
    const int l = x / (NG * 3), ee = e0 + l;
    sX[l][x % (NG * 3)] = (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  if (AFF)
    for (int x = tid; x < E; x += 64)
      affine_record(mesh.nodes, e0 + x, mesh.nelem, sG[x]);
  __syncthreads();

  const int idx[3] = {nd0 % N, (nd0 / N) % N, nd0 / (N * N)};
  double acc[R][5];
  double wJ[R];

  // ---- volume, one row at a time: velocity gradient, flux, divergence ----
#pragma unroll
  for (int rr = 0; rr < R; ++rr)
  {
    const int le = rr * EPR + tid / NB;
    const bool vact = (tid < EPR * NB) && (e0 + le < mesh.nelem);
#pragma unroll
    for (int r = 0; r < 5; ++r)
      acc[rr][r] = 0.0;
    wJ[rr] = 1.0;
    double Ji[3][3];
    double H[3][5];
    if constexpr (CT != 0)
    {
      // every lane takes part in the matrix-core contractions
      double gr[3][4];
      mfma_grad<0, CT>(sV[le], pt, gr[0]);
      mfma_grad<1, CT>(sV[le], pt, gr[1]);
      mfma_grad<2, CT>(sV[le], pt, gr[2]);
      if (vact)
      {
        double J[3][3], det;
        geometry_at<Q, AFF>(qt, sX[le], sG[le], -1, 0, idx[0], idx[1], idx[2], J, Ji, det);
        wJ[rr] = pt.w[idx[0]] * pt.w[idx[1]] * pt.w[idx[2]] * det;
        double Vq[5];
#pragma unroll
        for (int r = 0; r < 5; ++r)
          Vq[r] = sV[le][r][nd0];
        volume_flux(Vq, gr, Ji, ph.gamma, ph.mu, H);
      }
      auto sF = reinterpret_cast<float(*)[64]>(sR);
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
#pragma unroll
        for (int r = 0; r < 5; ++r)
          sF[r][tid] = vact ? (float)(wJ[rr] * (Ji[c][0] * H[0][r] + Ji[c][1] * H[1][r] +
                                                 Ji[c][2] * H[2][r]))
                            : 0.0f;
        __syncthreads();
        if (c == 0)
          mfma_project<0, CT>(sF, pt, acc[rr]);
        else if (c == 1)
          mfma_project<1, CT>(sF, pt, acc[rr]);
        else
          mfma_project<2, CT>(sF, pt, acc[rr]);
        __syncthreads();
      }
      continue;
    }
    if (vact)
    {
      double J[3][3], det;
      geometry_at<Q, AFF>(qt, sX[le], sG[le], -1, 0, idx[0], idx[1], idx[2], J, Ji, det);
      wJ[rr] = pt.w[idx[0]] * pt.w[idx[1]] * pt.w[idx[2]] * det;
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
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
      if (vact)
      {
#pragma unroll
        for (int r = 0; r < 5; ++r)
          sR[r * 64 + tid] =
              wJ[rr] * (Ji[c][0] * H[0][r] + Ji[c][1] * H[1][r] + Ji[c][2] * H[2][r]);
      }
      __syncthreads();
      if (vact)
      {
        const int base = (tid / NB) * NB;
#pragma unroll
        for (int r = 0; r < 5; ++r)
        {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < N; ++m)
          {
            int id2[3] = {idx[0], idx[1], idx[2]};
            id2[c] = m;
            s += pt.Dg[m][idx[c]] * sR[r * 64 + base + nid<N>(id2[0], id2[1], id2[2])];
          }
          acc[rr][r] += s;
        }
      }
      __syncthreads();
    }
  }

  // ---- faces, one reference direction at a time ----
  bool ok = true;
  tile_face_direction<P, Q, R, 0, TT, AFF>(mesh, ph, pt, qt, Tin, lam, sX, sG, sNb, sR, e0, acc,
                                           ok);
  tile_face_direction<P, Q, R, 1, TT, AFF>(mesh, ph, pt, qt, Tin, lam, sX, sG, sNb, sR, e0, acc,
                                           ok);
  tile_face_direction<P, Q, R, 2, TT, AFF>(mesh, ph, pt, qt, Tin, lam, sX, sG, sNb, sR, e0, acc,
                                           ok);

  // ---- form f_g and the RK epilogue, one row at a time; with NORM, each
  // row's R = -B3^T W f_g in sR (three passes) and the lane's running sum of
  // R^2 in sR[5 * 64 + lane] ----
#pragma unroll
  for (int rr = 0; rr < R; ++rr)
  {
    const int le = rr * EPR + tid / NB;
    const int e  = e0 + le;
    const bool act = tid < EPR * NB && e < mesh.nelem;
    if (act)
    {
      const double iwJ = rcp(wJ[rr]);
      const size_t base = (size_t)e * 5 * NB + nd0;
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        const size_t ix = base + (size_t)r * NB;
        const double fv = ph.source[r] - acc[rr][r] * iwJ;
        if (NORM)
          sR[r * 64 + tid] = -wJ[rr] * fv;
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
    if constexpr (NORM)
    {
      // the modal residual whose RMS the step reports, one direction per
      // pass; only this lane reads and writes its running-sum slot
#pragma unroll
      for (int c = 0; c < 3; ++c)
        tile_modal_pass<N, NB>(sR, pt.B, c, idx, act);
      double s2 = 0.0;
      if (act)
#pragma unroll
        for (int r = 0; r < 5; ++r)
          s2 += sR[r * 64 + tid] * sR[r * 64 + tid];
      sR[5 * 64 + tid] = (rr == 0) ? s2 : sR[5 * 64 + tid] + s2;
      __syncthreads();
    }
  }
  if constexpr (NORM)
  {
    // the tile's sum of R^2: a butterfly over the wavefront (fixed order:
    // deterministic)
    double sq = sR[5 * 64 + tid];
#pragma unroll
    for (int w = 32; w > 0; w >>= 1)
      sq += __shfl_xor(sq, w, 64);
    if (tid == 0)
      sa.normp[block] = sq;
  }

  // ---- face traces of the new state for the next stage ----
  if (sa.Tout)
  {
    __syncthreads();
    TT* Tout = reinterpret_cast<TT*>(sa.Tout);
    tile_trace_direction<P, Q, R, 0, TT, AFF>(mesh, ph, pt, qt, sV, sX, sG, sNb, sBt, sR, e0,
                                              Tout, ok);
    tile_trace_direction<P, Q, R, 1, TT, AFF>(mesh, ph, pt, qt, sV, sX, sG, sNb, sBt, sR, e0,
                                              Tout, ok);
    tile_trace_direction<P, Q, R, 2, TT, AFF>(mesh, ph, pt, qt, sV, sX, sG, sNb, sBt, sR, e0,
                                              Tout, ok);
  }
  if (!ok)
    atomicOr(status, 1);
}

template <int P, int Q, int R, int W, typename TT, bool AFF, int CT = 0, bool NORM = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_tile_stage_kernel(MeshArgs mesh, Phys ph, StageArgs sa, const double* __restrict__ lam,
                           int* status)
{
  __shared__ __attribute__((aligned(16))) unsigned char lds[TileLds<P, Q, R>::bytes];
  tile_body_lds<P, Q, R, TT, AFF, CT, NORM>(mesh, ph, sa, lam, status, blockIdx.x, lds);
}
