// dgx_adjoint_kernels.hpp -- reverse (adjoint) kernels of the residual,
// compiled in the translation unit of dgx_kernels.hip (they share its
// constant tables and helpers; included from there inside namespace dgx).
//
// Given psi, the adjoint of the Gauss-nodal right-hand side f_g(V), the
// kernels compute psi_V = (d f_g / d V)^T psi (every boundary type:
// domain-boundary face points are finished by dgx_adj_boundary_kernel).  The
// forward evaluation of one element (volume terms, its six faces evaluated
// from its own side with both traces from the trace records, gather,
// f_g = s - acc / wJ) is reversed step by step:
//   A: volume terms (transposed contractions, local Jacobian of the volume
//      flux) into pvol, then, in a second launch, each face point's physics
//      reversed into the adjoint of the element's own trace record (aown) and
//      of its neighbour's record (anbr, written into the neighbour's slot:
//      every slot has one writer);
//   B: both adjoint records of each face point reversed through the trace
//      production (trace interpolation, tangential derivatives, gradient
//      traction), added to pvol.
// The local Jacobians of the pointwise physics are the reverse sweeps of
// dgx_vjp.hpp and roe_flux_vjp (dgx_device.hpp).

// PART 0: the volume terms (into pvol); PART 1: the face terms (into aown,
// anbr).  Separate launches: the volume part needs about half the registers
// of the face part and runs at a higher occupancy on its own.
// Occupancy target: at q = 1 the face-term part holds 256 VGPRs + 14-24
// AGPRs, one wave per SIMD; asked for two it spills 10 VGPRs (44 B/lane at
// p = 2) and runs the state-gradient adjoint 6 % faster (same results bit for
// bit).  q > 1 spills already at one wave and keeps it.
template <int P, int Q, int PART, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(Q == 1 ? 2 : 1)))
void dgx_adj_main_kernel(MeshArgs mesh, Phys ph, const double* __restrict__ V,
                         const double* __restrict__ Tin, const double* __restrict__ lam,
                         const double* __restrict__ psi, double* __restrict__ pvol,
                         double* __restrict__ aown, double* __restrict__ anbr)
{
  using FD = FaceDims<P>;
  constexpr int N = FD::N, NB = FD::NB, NF = FD::NF, EPW = FD::EPW;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);
  __shared__ double sV[EPW][5][NB];
  __shared__ double sA[EPW][5][NB];  // adjoint of acc
  __shared__ double sX[EPW][NG * 3];
  __shared__ double sG[EPW][GREC];  // affine-element geometry records
  __shared__ double sR[(4 * NB * EPW > 4 * 64) ? 4 * NB * EPW : 4 * 64];
  __shared__ PTab sPT;
  __shared__ QTab sQT;
  __shared__ int sNb[EPW][6], sBt[EPW][6];
  stage_tables<P, Q>(&sPT, &sQT);
  const PTab& pt = sPT;
  const QTab& qt = sQT;
  const int tid = threadIdx.x, le = tid / NB, nd0 = tid % NB;
  const int e0 = blockIdx.x * EPW, e = e0 + le;
  const bool vact = (le < EPW) && (e < mesh.nelem);
  if constexpr (PART == 1)
    face_table<EPW>(mesh, e0, sNb, sBt);
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int ee = e0 + x / (NG * 3);
    sX[x / (NG * 3)][x % (NG * 3)] =
        (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  if (AFF && tid < EPW)
    affine_record(mesh.nodes, e0 + tid, mesh.nelem, sG[tid]);
  const int idx[3] = {nd0 % N, (nd0 / N) % N, nd0 / (N * N)};
  double Ji[3][3], wJ = 1.0;
  if (PART == 0 && vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sV[le][r][nd0] = V[((size_t)e * 5 + r) * NB + nd0];
  }
  __syncthreads();
  if (vact)
  {
    double J[3][3], det;
    geometry_at<Q, AFF>(qt, sX[le], sG[le], -1, 0, idx[0], idx[1], idx[2], J, Ji, det);
    wJ = pt.w[idx[0]] * pt.w[idx[1]] * pt.w[idx[2]] * det;
    const double iwJ = rcp(wJ);
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sA[le][r][nd0] = -psi[((size_t)e * 5 + r) * NB + nd0] * iwJ;
  }
  __syncthreads();

  if constexpr (PART == 0)
  {
  // ---- volume terms ----
  double pv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double pgr[3][4];
  if (vact)
  {
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
    // acc(n) = sum_c sum_{n' on the c-line of n} Dg[i_c(n')][i_c(n)] G_c(n')
    double pG[3][5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m)
      {
        g0 += pt.Dg[idx[0]][m] * sA[le][r][nid<N>(m, idx[1], idx[2])];
        g1 += pt.Dg[idx[1]][m] * sA[le][r][nid<N>(idx[0], m, idx[2])];
        g2 += pt.Dg[idx[2]][m] * sA[le][r][nid<N>(idx[0], idx[1], m)];
      }
      pG[0][r] = g0;
      pG[1][r] = g1;
      pG[2][r] = g2;
    }
    // G_c = wJ sum_d Ji[c][d] H[d]
    double pH[3][5];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int r = 0; r < 5; ++r)
        pH[d][r] = wJ * (Ji[0][d] * pG[0][r] + Ji[1][d] * pG[1][r] + Ji[2][d] * pG[2][r]);
    double gx[17];
    volume_flux_vjp(Vq, gr, Ji, ph.gamma, ph.mu, pH, gx);
#pragma unroll
    for (int r = 0; r < 5; ++r)
      pv[r] = gx[r];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        pgr[c][r] = gx[5 + 4 * c + r];
  }
  // gr[c](n) = sum_m Dg[i_c(n)][m] V(n with i_c = m): transpose, one direction
  // at a time through LDS
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    if (vact)
    {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sR[(le * 4 + r) * NB + nd0] = pgr[c][r];
    }
    __syncthreads();
    if (vact)
    {
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          int id2[3] = {idx[0], idx[1], idx[2]};
          id2[c] = m;
          s += pt.Dg[m][idx[c]] * sR[(le * 4 + r) * NB + nid<N>(id2[0], id2[1], id2[2])];
        }
        pv[r] += s;
      }
    }
    __syncthreads();
  }
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      pvol[((size_t)e * 5 + r) * NB + nd0] = pv[r];
  }
  return;
  }

  // ---- faces ----
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
    // gather, reversed: T = sum_m L[hi][m] acc(line m); hn = -sum_m Ld[hi][m] acc
    double pT[5], phn[4];
    if (fact)
    {
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        double t = 0.0, h = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          double x = sA[l][r][line_node<N>(nd, m, a, b)];
          t += pt.L[hi][m] * x;
          h += pt.Ld[hi][m] * x;
        }
        pT[r] = t;
        if (r > 0)
          phn[r - 1] = -h;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        sR[tid * 4 + k] = pT[1 + k];
    }
    __syncthreads();
    if (fact)
    {
      // T(a',b) -= sum_m Dg[m][a'] h1(m,b) + ...: reversed at this point as m
      double ph1[4], ph2[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
      {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          s1 += pt.Dg[a][m] * sR[(fbase + m + N * b) * 4 + k];
          s2 += pt.Dg[b][m] * sR[(fbase + a + N * m) * 4 + k];
        }
        ph1[k] = -s1;
        ph2[k] = -s2;
      }
      // forward quantities of this face point; on a domain-boundary face UR
      // is the boundary state stored with the records
      const int en = sNb[l][f];
      const bool interior = en >= 0;
      const double* To = Tin + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      double UL[5], UR[5], vgL[3], vgR[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < 5; ++r)
        UL[r] = To[r * NF];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        vgL[i] = To[(5 + i) * NF];
      if (interior)
      {
        const double* Tn = Tin + ((size_t)en * 6 + (f ^ 1)) * 8 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          UR[r] = Tn[r * NF];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          vgR[i] = -Tn[(5 + i) * NF];
      }
      else
      {
        const double* B =
            Tin + (size_t)mesh.nelem * 6 * 8 * NF + (size_t)(-1 - en) * 5 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          UR[r] = B[r * NF];
      }
      const int gi0 = (nd == 0) ? 0 : a;
      const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
      const int gi2 = (nd == 2) ? 0 : b;
      double J[3][3], JiL[3][3], detL;
      geometry_at<Q, AFF>(qt, sX[l], sG[l], nd, hi, gi0, gi1, gi2, J, JiL, detL);
      double nv[3];
      const double fJ = face_normal(J, nd, hi, nv);
      const double lamL = lam[((size_t)ee * 6 + f) * NF + fpt];
      const double lamR = interior ? lam[((size_t)en * 6 + (f ^ 1)) * NF + fpt] : 0.0;
      double S[6];
      if (interior)
        interior_dc_stress(UL, UR, nv, ph.mu, S);
      else
        boundary_dc_stress(UL, UR, nv, ph.mu, S);
      const double irL = rcp(UL[0]);
      const double uL[3] = {UL[1] * irL, UL[2] * irL, UL[3] * irL};
      const double sc = fJ * pt.w[a] * pt.w[b];
      const double Sm[3][3] = {{S[0], S[3], S[4]}, {S[3], S[1], S[5]}, {S[4], S[5], S[2]}};
      double hc[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          hc[c][i] = sc * (JiL[c][0] * Sm[0][i] + JiL[c][1] * Sm[1][i] + JiL[c][2] * Sm[2][i]);
      // adjoint of hc[c][0..3] from the normal / tangential parts
      double phc[3][4];
      // the remaining direction (not nd, not t1) takes ph2
      const int t1 = (nd == 0) ? 1 : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          phc[c][k] = (c == nd) ? phn[k] : ((c == t1) ? ph1[k] : ph2[k]);
      // hc[c][3] = u_L . hc[c][0..2]
      double puL[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
          puL[i] += hc[c][i] * phc[c][3];
          phc[c][i] += uL[i] * phc[c][3];
        }
      // hc[c][i] = sc sum_k JiL[c][k] Sm[k][i]
      double pSm[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          pSm[k][i] = sc * (JiL[0][k] * phc[0][i] + JiL[1][k] * phc[1][i] + JiL[2][k] * phc[2][i]);
      const double pS[6] = {pSm[0][0], pSm[1][1], pSm[2][2], pSm[0][1] + pSm[1][0],
                            pSm[0][2] + pSm[2][0], pSm[1][2] + pSm[2][1]};
      double pG[5];
#pragma unroll
      for (int r = 0; r < 5; ++r)
        pG[r] = sc * pT[r];
      if (!interior)
      {
        // boundary face: the adjoints of G (into the boundary section of
        // anbr) and S (into the element's own anbr slot) go to
        // dgx_adj_boundary_kernel, which reverses the boundary physics; the
        // u_L part of the contravariant energy flux is the trace's already
        double* An = anbr + ((size_t)ee * 6 + f) * 8 * NF + fpt;
        double* Bg = anbr + (size_t)mesh.nelem * 6 * 8 * NF + (size_t)(-1 - en) * 5 * NF + fpt;
        double* Ao = aown + ((size_t)ee * 6 + f) * 8 * NF + fpt;
#pragma unroll
        for (int r = 0; r < 5; ++r)
          Bg[r * NF] = pG[r];
#pragma unroll
        for (int r = 0; r < 6; ++r)
          An[r * NF] = pS[r];
        Ao[0] = -(puL[0] * uL[0] + puL[1] * uL[1] + puL[2] * uL[2]) * irL;
#pragma unroll
        for (int i = 0; i < 3; ++i)
          Ao[(1 + i) * NF] = puL[i] * irL;
#pragma unroll
        for (int r = 4; r < 8; ++r)
          Ao[r * NF] = 0.0;
      }
      else
      {
      double gx[16];
      face_physics_vjp(UL, UR, vgL, vgR, nv, lamL, lamR, ph, pG, pS, gx);
      // u_L = m_L / rho_L
      gx[0] -= (puL[0] * uL[0] + puL[1] * uL[1] + puL[2] * uL[2]) * irL;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        gx[1 + i] += puL[i] * irL;
      double* Ao = aown + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      double* An = anbr + ((size_t)en * 6 + (f ^ 1)) * 8 * NF + fpt;
#pragma unroll
      for (int r = 0; r < 5; ++r)
      {
        Ao[r * NF] = gx[r];
        An[r * NF] = gx[5 + r];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
      {
        Ao[(5 + i) * NF] = gx[10 + i];
        An[(5 + i) * NF] = -gx[13 + i];  // the neighbour stored its traction with -n
      }
      }
    }
    __syncthreads();
  }
}

// Boundary faces (walls, inflow, outflow, free stream), one thread per
// boundary face point.  The face-term kernel left there the adjoint of the
// face flux G (5, in the boundary section of anbr) and of the
// dual-consistency stress S (6, in the element's own anbr slot), and wrote
// into aown the part of the trace adjoint that passes through u_L in the
// contravariant energy flux.  This kernel reverses the boundary face physics
// (face_point_physics(false, ...)), including the boundary state's
// dependence on the trace (boundary_state, bound_states.cpp:24-130), with
// forward-mode dual numbers, one pass per trace input (5 state, 3 traction),
// adds the result to aown and clears the element's anbr slot (a boundary
// face has no neighbour adjoint).
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_adj_boundary_kernel(MeshArgs mesh, Phys ph,
                                                              const double* __restrict__ Tin,
                                                              const double* __restrict__ lam,
                                                              double* __restrict__ aown,
                                                              double* __restrict__ anbr)
{
  constexpr int N = P + 1, NF = N * N, NG = (Q + 1) * (Q + 1) * (Q + 1);
  const int nx = mesh.nx, ny = mesh.ny, nz = mesh.nz;
  const int sx = ny * nz, sy = nx * nz, sz = nx * ny;
  const size_t t = (size_t)blockIdx.x * 64 + threadIdx.x;
  if (t >= (size_t)2 * (sx + sy + sz) * NF)
    return;
  const int bf = (int)(t / NF), fpt = (int)(t % NF);
  // side f and element of boundary face bf (the inverse of bface_index)
  int f, in;
  if (bf < 2 * sx)
  {
    f  = bf / sx;
    in = bf % sx;
  }
  else if (bf < 2 * sx + 2 * sy)
  {
    f  = 2 + (bf - 2 * sx) / sy;
    in = (bf - 2 * sx) % sy;
  }
  else
  {
    f  = 4 + (bf - 2 * sx - 2 * sy) / sz;
    in = (bf - 2 * sx - 2 * sy) % sz;
  }
  const int bt = mesh.bounds[f];
  if (bt == -2)
    return;  // periodic side: no boundary faces
  const int nd = f >> 1, hi = f & 1;
  int ci, cj, ck;
  if (nd == 0)
  {
    ci = hi ? nx - 1 : 0;
    cj = in % ny;
    ck = in / ny;
  }
  else if (nd == 1)
  {
    cj = hi ? ny - 1 : 0;
    ci = in % nx;
    ck = in / nx;
  }
  else
  {
    ck = hi ? nz - 1 : 0;
    ci = in % nx;
    cj = in / nx;
  }
  const int e = (ck * ny + cj) * nx + ci;
  const int a = fpt % N, b = fpt / N;
  const int gi0 = (nd == 0) ? 0 : a;
  const int gi1 = (nd == 1) ? 0 : ((nd == 0) ? a : b);
  const int gi2 = (nd == 2) ? 0 : b;
  double J[3][3];
  jac_at<Q>(c_qtab[P - 1][Q - 1], mesh.nodes + (size_t)e * NG * 3, nd, hi, gi0, gi1, gi2, J);
  double nv[3];
  face_normal(J, nd, hi, nv);
  const double* To = Tin + ((size_t)e * 6 + f) * 8 * NF + fpt;
  double* Ao = aown + ((size_t)e * 6 + f) * 8 * NF + fpt;
  double* An = anbr + ((size_t)e * 6 + f) * 8 * NF + fpt;
  const double* Bg = anbr + (size_t)mesh.nelem * 6 * 8 * NF + (size_t)bf * 5 * NF + fpt;
  const double lamL = lam[((size_t)e * 6 + f) * NF + fpt];
  double x[8], pG[5], pS[6];
#pragma unroll
  for (int r = 0; r < 8; ++r)
    x[r] = To[r * NF];
#pragma unroll
  for (int r = 0; r < 5; ++r)
    pG[r] = Bg[r * NF];
#pragma unroll
  for (int r = 0; r < 6; ++r)
    pS[r] = An[r * NF];
  for (int d = 0; d < 8; ++d)
  {
    Dual<1> UL[5], vgL[3], vgR[3], Ub[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      UL[r]      = Dual<1>(x[r]);
      UL[r].d[0] = (d == r) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
      vgL[i]      = Dual<1>(x[5 + i]);
      vgL[i].d[0] = (d == 5 + i) ? 1.0 : 0.0;
      vgR[i]      = Dual<1>(0.0);
    }
    (void)boundary_state(bt, UL, nv, ph, Ub);
    FacePointT<Dual<1>> o;
    face_point_physics(false, UL, Ub, vgL, vgR, nv, lamL, 0.0, ph, o);
    double g = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r)
      g += pG[r] * o.G[r].d[0];
#pragma unroll
    for (int r = 0; r < 6; ++r)
      g += pS[r] * o.S[r].d[0];
    Ao[d * NF] += g;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r)
    An[r * NF] = 0.0;
}

template <int P, int Q, bool AFF>
__global__ __launch_bounds__(64) void dgx_adj_trace_kernel(MeshArgs mesh, Phys ph,
                                                           const double* __restrict__ V,
                                                           const double* __restrict__ aown,
                                                           const double* __restrict__ anbr,
                                                           const double* __restrict__ pvol,
                                                           double* __restrict__ psiV)
{
  using FD = FaceDims<P>;
  constexpr int N = FD::N, NB = FD::NB, NF = FD::NF, EPW = FD::EPW;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);
  __shared__ double sV[EPW][5][NB];
  __shared__ double sX[EPW][NG * 3];
  __shared__ double sG[EPW][GREC];  // affine-element geometry records
  __shared__ double sU[4 * 64];  // trace values (rho, m) of the pass
  __shared__ double sQ[8 * 64];  // adjoint tangential derivatives
  __shared__ double sT[9 * 64];  // adjoint trace values and normal derivatives
  __shared__ PTab sPT;
  __shared__ QTab sQT;
  stage_tables<P, Q>(&sPT, &sQT);
  const PTab& pt = sPT;
  const QTab& qt = sQT;
  const int tid = threadIdx.x, le = tid / NB, nd0 = tid % NB;
  const int e0 = blockIdx.x * EPW, e = e0 + le;
  const bool vact = (le < EPW) && (e < mesh.nelem);
  for (int x = tid; x < EPW * NG * 3; x += 64)
  {
    int ee = e0 + x / (NG * 3);
    sX[x / (NG * 3)][x % (NG * 3)] =
        (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  if (AFF && tid < EPW)
    affine_record(mesh.nodes, e0 + tid, mesh.nelem, sG[tid]);
  double pv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
      sV[le][r][nd0] = V[((size_t)e * 5 + r) * NB + nd0];
      pv[r] = pvol[((size_t)e * 5 + r) * NB + nd0];
    }
  }
  const int idx[3] = {nd0 % N, (nd0 / N) % N, nd0 / (N * N)};
  __syncthreads();
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
          dv += pt.Ld[hi][m] * x;
        }
        U[r] = v;
        if (r < 4)
        {
          dn[r] = dv;
          sU[tid * 4 + r] = v;
        }
      }
    }
    __syncthreads();
    double pU[5], pdn[4];
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
          d1 += pt.Dg[a][m] * sU[(fbase + m + N * b) * 4 + r];
          d2 += pt.Dg[b][m] * sU[(fbase + a + N * m) * 4 + r];
        }
        const double dr = dn[r];
        gr[0][r] = pick3(nd, dr, d1, d1);
        gr[1][r] = pick3(nd, d1, dr, d2);
        gr[2][r] = pick3(nd, d2, d2, dr);
      }
      // adjoint of this point's trace record: own evaluation + neighbour's
      const double* Ao = aown + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      const double* An = anbr + ((size_t)ee * 6 + f) * 8 * NF + fpt;
      double prec[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        prec[k] = Ao[k * NF] + An[k * NF];
      double gx[16];
      traction_vjp(gr, U, Ji, nv, ph.mu, prec + 5, gx);
#pragma unroll
      for (int r = 0; r < 5; ++r)
        pU[r] = prec[r] + ((r < 4) ? gx[12 + r] : 0.0);
      const int t1 = (nd == 0) ? 1 : 0, t2 = (nd == 2) ? 1 : 2;
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        pdn[r] = pick3(nd, gx[r], gx[4 + r], gx[8 + r]);
        sQ[tid * 8 + r]     = gx[4 * t1 + r];
        sQ[tid * 8 + 4 + r] = gx[4 * t2 + r];
      }
    }
    __syncthreads();
    if (fact)
    {
      // d1(a',b) = sum_m Dg[a'][m] U(m,b): reversed at this point as m
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m)
        {
          s += pt.Dg[m][a] * sQ[(fbase + m + N * b) * 8 + r];
          s += pt.Dg[m][b] * sQ[(fbase + a + N * m) * 8 + 4 + r];
        }
        pU[r] += s;
      }
#pragma unroll
      for (int r = 0; r < 5; ++r)
        sT[tid * 9 + r] = pU[r];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sT[tid * 9 + 5 + r] = pdn[r];
    }
    __syncthreads();
    // U = sum_m L[hi][m] V(line m), dn = sum_m Ld[hi][m] V(line m): gathered
    // at the nodes of the faces of this pass
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
#pragma unroll
        for (int r = 0; r < 5; ++r)
          pv[r] += pt.L[fhi][m] * sT[fp * 9 + r];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          pv[r] += pt.Ld[fhi][m] * sT[fp * 9 + 5 + r];
      }
    }
    __syncthreads();
  }
  if (vact)
  {
#pragma unroll
    for (int r = 0; r < 5; ++r)
      psiV[((size_t)e * 5 + r) * NB + nd0] = pv[r];
  }
}

// Kinetic energy J = sum_e sum_nodes wJ |m|^2 / (2 rho) (the (p+1)^3 Gauss
// rule of the mass matrix) and its gradient dJ/dV = wJ (-|u|^2/2, u, 0) at
// every node; per-block partial sums of J.
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_kinetic_energy_kernel(MeshArgs mesh,
                                                                const double* __restrict__ V,
                                                                double* __restrict__ grad,
                                                                double* __restrict__ partial)
{
  constexpr int N = P + 1, NB = N * N * N, EPW = 64 / NB;
  constexpr int NG = (Q + 1) * (Q + 1) * (Q + 1);
  const PTab& pt = c_ptab[P - 1];
  const QTab& qt = c_qtab[P - 1][Q - 1];
  __shared__ double red[64];
  const int tid = threadIdx.x, le = tid / NB, nd0 = tid % NB;
  const int e = blockIdx.x * EPW + le;
  double k = 0.0;
  if (le < EPW && e < mesh.nelem)
  {
    const int i = nd0 % N, j = (nd0 / N) % N, kk = nd0 / (N * N);
    double J[3][3];
    jac_at<Q>(qt, mesh.nodes + (size_t)e * NG * 3, -1, 0, i, j, kk, J);
    const double wJ = pt.w[i] * pt.w[j] * pt.w[kk] * det3(J);
    double s[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
      s[r] = V[((size_t)e * 5 + r) * NB + nd0];
    const double ir = 1.0 / s[0];
    const double u = s[1] * ir, v = s[2] * ir, w = s[3] * ir;
    k = wJ * 0.5 * (s[1] * u + s[2] * v + s[3] * w);
    double* g = grad + (size_t)e * 5 * NB + nd0;
    g[0]      = -0.5 * wJ * (u * u + v * v + w * w);
    g[NB]     = wJ * u;
    g[2 * NB] = wJ * v;
    g[3 * NB] = wJ * w;
    g[4 * NB] = 0.0;
  }
  red[tid] = k;
  __syncthreads();
  for (int o = 32; o > 0; o >>= 1)
  {
    if (tid < o)
      red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0)
    partial[blockIdx.x] = red[0];
}

// Tail of one reverse RK stage, elementwise in one pass: with the stage's
// adjoint a, sum_out = sum_in + a (sum_in may be sum_out) and the next
// stage's input lk = wa a + wl l; without a, lk = wl l.  Null outputs are
// skipped.
__global__ __launch_bounds__(256) void dgx_adj_tail_kernel(const double* __restrict__ l,
                                                           const double* __restrict__ a,
                                                           const double* sum_in, double* sum_out,
                                                           double* __restrict__ lk, double wl,
                                                           double wa, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  if (a)
  {
    const double ai = a[i];
    if (sum_out)
      sum_out[i] = sum_in[i] + ai;
    if (lk)
      lk[i] = wa * ai + wl * l[i];
  }
  else if (lk)
    lk[i] = wl * l[i];
}

// partial[b] = scale * sum over workgroup b's grid-stride share of
// x[i] * (y[i] - z[i]) (z may be NULL: x . y); wave butterflies, fixed order
__global__ __launch_bounds__(256) void dgx_dot_kernel(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z, double scale,
                                                      size_t n, double* __restrict__ partial)
{
  __shared__ double red[4];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    s += x[i] * (z ? y[i] - z[i] : y[i]);
#pragma unroll
  for (int w = 32; w > 0; w >>= 1)
    s += __shfl_xor(s, w, 64);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = scale * ((red[0] + red[1]) + (red[2] + red[3]));
}
