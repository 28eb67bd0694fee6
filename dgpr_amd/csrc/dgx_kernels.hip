// dgx_kernels.hip -- gfx950 kernels of the FP64 DG residual + RK stage, the
// basis conversions, residual output, derived outputs and the discrete
// adjoint, with their host launch helpers.  The device code shared with the
// ensemble stage kernel lives in dgx_stage.hpp (method described there).
#include "dgx_stage.hpp"
#include "dgx_dual.hpp"
#include "dgx_dd.hpp"

#include <type_traits>

namespace dgx
{

// W is the occupancy target (waves per SIMD) handed to the register
// allocator; the launch picks one of the compiled targets per (P, Q).  TT is
// the type of the trace records read and written (float: mixed precision).
template <int P, int Q, int W, typename TT, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_rhs_stage_kernel(MeshArgs mesh, Phys ph, StageArgs sa, const double* __restrict__ lam,
                          int* status)
{
  stage_body<P, Q, TT, AFF>(mesh, ph, sa, lam, status, blockIdx.x);
}

// The inviscid right-hand side of the one-row body (INV): f(V; mu = 0) from
// the state and its trace records, whose tractions are not read.
template <int P, int Q, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void dgx_inviscid_rhs_kernel(MeshArgs mesh, Phys ph, StageArgs sa, const double* __restrict__ lam,
                             int* status)
{
  __shared__ __attribute__((aligned(16))) unsigned char lds[StageLds<P, Q>::bytes];
  stage_body_lds<P, Q, double, AFF, true>(mesh, ph, sa, lam, status, blockIdx.x, lds);
}

// Batched form: the blocks of several independent meshes of the same (P, Q)
// in one launch.  items[] is sorted by first block; each block finds its
// member by a uniform binary search.
template <int P, int Q, int W, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_rhs_batch_kernel(const BatchItem* __restrict__ items, int nitems, int* status)
{
  const int b = blockIdx.x;
  int lo = 0, hi = nitems - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const BatchItem& it = items[lo];
  stage_body<P, Q, double, AFF>(it.mesh, it.ph, it.sa, it.lam, status, b - it.block0);
}


// BR2 lifting coefficient of each element's side of each of its faces:
//   lam[e][f][pt] = fJ(pt) * sum_m L_hi[m]^2 / (w_m detJ(line node m))
// (the pointwise form of Phi_f M_e^-1 Phi_f^T W_f; geometry only).
template <int P, int Q>
__global__ __launch_bounds__(64) void dgx_lift_coef_kernel(MeshArgs mesh, double* lam)
{
  constexpr int N = P + 1, NF = N * N, NG = (Q + 1) * (Q + 1) * (Q + 1);
  const PTab& pt = c_ptab[P - 1];
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)mesh.nelem * 6 * NF;
  if (t >= total)
    return;
  int fpt = (int)(t % NF);
  int f   = (int)((t / NF) % 6);
  size_t e = t / (6 * NF);
  int a = fpt % N, b = fpt / N;
  int nd = f >> 1, hi = f & 1;
  int t1 = (nd == 0) ? 1 : 0;
  int t2 = (nd == 2) ? 1 : 2;
  const double* X = mesh.nodes + e * NG * 3;
  int gi[3];
  gi[nd] = 0;
  gi[t1] = a;
  gi[t2] = b;
  double J[3][3];
  jac_at<Q>(c_qtab[P - 1][Q - 1], X, nd, hi, gi[0], gi[1], gi[2], J);
  int c1 = (nd + 1) % 3, c2 = (nd + 2) % 3;
  double n0 = J[1][c1] * J[2][c2] - J[2][c1] * J[1][c2];
  double n1 = J[2][c1] * J[0][c2] - J[0][c1] * J[2][c2];
  double n2 = J[0][c1] * J[1][c2] - J[1][c1] * J[0][c2];
  double fJ = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  double s = 0.0;
  for (int m = 0; m < N; ++m)
  {
    int gm[3];
    gm[nd] = m;
    gm[t1] = a;
    gm[t2] = b;
    double Jm[3][3];
    jac_at<Q>(c_qtab[P - 1][Q - 1], X, -1, 0, gm[0], gm[1], gm[2], Jm);
    s += pt.L[hi][m] * pt.L[hi][m] / (pt.w[m] * det3(Jm));
  }
  lam[t] = fJ * s;
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

// mode 0: V = B3 U      (reference basis -> Gauss-nodal)
// mode 1: U = B3^-1 V   (Gauss-nodal -> reference basis)
// mode 2: G = B3^T g    (a gradient with respect to V -> with respect to U)
// mode 3: g = B3^-T G   (a gradient with respect to U -> with respect to V)
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
  const bool inv = (mode == 1 || mode == 3);
  const double(*M)[MAXN]  = inv ? pt.Binv : pt.B;
  const double(*Ml)[MAXN] = inv ? pt.Binvlo : pt.Blo;
  const bool tr = (mode >= 2);
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
          const double c  = tr ? M[m][idx[d]] : M[idx[d]][m];
          const double cl = tr ? Ml[m][idx[d]] : Ml[idx[d]][m];
          acc = dd_fma({sh[le][r][src], sl[le][r][src]}, c, cl, acc);
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
      jac_at<Q>(c_qtab[P - 1][Q - 1], sX[le], -1, 0, ni, nj, nk, J);
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
    // butterfly over the wavefront (fixed order: deterministic)
#pragma unroll
    for (int w = 32; w > 0; w >>= 1)
      sq += __shfl_xor(sq, w, 64);
    if (tid == 0)
      partial[blockIdx.x] = sq;
  }
}

// Per-workgroup sums of R^2, R = -B3^T W f_g (the modal residual of
// dgx_resid_out_kernel), for the per-step residual norm: 4 x EPW elements
// per 256-lane workgroup, the 1D tables copied to LDS (lane-dependent table
// reads would otherwise be vector loads from constant memory), one butterfly
// per wavefront.  partial[blockIdx.x] of norm_blocks(p, q, nelem) workgroups.
template <int P, int Q>
__global__ __launch_bounds__(256) void dgx_norm_kernel(MeshArgs mesh, const double* __restrict__ fg,
                                                       double* __restrict__ partial)
{
  using Dm = Dims<P, Q>;
  constexpr int N = Dm::N, NB = Dm::NB, NG = Dm::NG, EPB = 4 * Dm::EPW;
  __shared__ double s[EPB][5][NB];
  __shared__ double sX[EPB][NG * 3];
  __shared__ PTab sPT;
  __shared__ QTab sQT;
  __shared__ double red[4];
  const int tid = threadIdx.x, le = tid / NB, node = tid % NB;
  const int e0 = blockIdx.x * EPB, e = e0 + le;
  const bool act = le < EPB && e < mesh.nelem;
  {
    const double* a = reinterpret_cast<const double*>(&c_ptab[P - 1]);
    double* b       = reinterpret_cast<double*>(&sPT);
    for (int x = tid; x < (int)(sizeof(PTab) / 8); x += 256)
      b[x] = a[x];
    const double* c = reinterpret_cast<const double*>(&c_qtab[P - 1][Q - 1]);
    double* d       = reinterpret_cast<double*>(&sQT);
    for (int x = tid; x < (int)(sizeof(QTab) / 8); x += 256)
      d[x] = c[x];
  }
  for (int x = tid; x < EPB * NG * 3; x += 256)
  {
    const int l = x / (NG * 3), ee = e0 + l;
    sX[l][x % (NG * 3)] = (ee < mesh.nelem) ? mesh.nodes[(size_t)ee * NG * 3 + x % (NG * 3)] : 0.0;
  }
  double fgv[5];
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      fgv[r] = fg[((size_t)e * 5 + r) * NB + node];
  __syncthreads();
  const int ni = node % N, nj = (node / N) % N, nk = node / (N * N);
  if (act)
  {
    double J[3][3];
    jac_at<Q>(sQT, sX[le], -1, 0, ni, nj, nk, J);
    const double wJ = sPT.w[ni] * sPT.w[nj] * sPT.w[nk] * det3(J);
#pragma unroll
    for (int r = 0; r < 5; ++r)
      s[le][r][node] = -wJ * fgv[r];
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d)
    apply_dir<N, NB, EPB>(s, sPT.B, d, true, le, node, act);
  double sq = 0.0;
  if (act)
#pragma unroll
    for (int r = 0; r < 5; ++r)
      sq += s[le][r][node] * s[le][r][node];
#pragma unroll
  for (int w = 32; w > 0; w >>= 1)
    sq += __shfl_xor(sq, w, 64);
  if ((tid & 63) == 0)
    red[tid >> 6] = sq;
  __syncthreads();
  if (tid == 0)
    partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Deterministic sum of the per-block partial sums into out[0]; with
// scale > 0, out[0] = sqrt(scale * sum) instead (the residual norm of
// residual_norm, time_distributed.cpp:23-44, with scale = 1 / state length).
// One workgroup: for up to SUM_ONE_BLOCK partials.
__global__ __launch_bounds__(256) void dgx_sum_kernel(const double* partial, int n, double* out,
                                                      double scale)
{
  __shared__ double red[4];
  double t = 0.0;
  for (int x = threadIdx.x; x < n; x += 256)
    t += partial[x];
#pragma unroll
  for (int w = 32; w > 0; w >>= 1)
    t += __shfl_xor(t, w, 64);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    const double s = (red[0] + red[1]) + (red[2] + red[3]);
    out[0] = (scale > 0.0) ? sqrt(scale * s) : s;
  }
}

// First level of a long sum: workgroup b adds partial[b*SUM_CHUNK ..
// (b+1)*SUM_CHUNK) into chunk[b] (eight independent loads per lane).
__global__ __launch_bounds__(256) void dgx_chunk_sum_kernel(const double* __restrict__ partial,
                                                            int n, double* __restrict__ chunk)
{
  __shared__ double red[4];
  const int base = blockIdx.x * SUM_CHUNK + threadIdx.x;
  double v[SUM_CHUNK / 256];
#pragma unroll
  for (int k = 0; k < SUM_CHUNK / 256; ++k)
  {
    const int x = base + 256 * k;
    v[k] = (x < n) ? partial[x] : 0.0;
  }
  double t = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
#pragma unroll
  for (int w = 32; w > 0; w >>= 1)
    t += __shfl_xor(t, w, 64);
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0)
    chunk[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Plain streaming copy (16 B per lane per iteration).
__global__ __launch_bounds__(256) void dgx_copy_kernel(const double2* __restrict__ src,
                                                      double2* __restrict__ dst, size_t n2)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    dst[i] = src[i];
}

hipError_t copy_dispatch(const void* src, void* dst, size_t bytes, hipStream_t st)
{
  size_t n2 = bytes / 16;
  int nblk = (int)((n2 + 255) / 256);
  if (nblk > 8192)
    nblk = 8192;
  if (nblk < 1)
    nblk = 1;
  hipLaunchKernelGGL(dgx_copy_kernel, dim3(nblk), dim3(256), 0, st, (const double2*)src,
                     (double2*)dst, n2);
  return hipGetLastError();
}

// ---- derived outputs -------------------------------------------------------

// eval_output (render/vrtxgen.cpp:25-59) on the (Ns+1)^3 equispaced sub-grid
// of every element, reference points (i/Ns, j/Ns, k/Ns) with sub-node index
// (Ns+1)^2 k + (Ns+1) j + i (vrtxgen.cpp:294-330).  The Gauss-nodal state is
// interpolated with the Gauss-node Lagrange basis, which spans the same
// polynomials as the reference's equispaced basis.  var: mach (0), rho (1),
// u (2), v (3), w (4); the Mach number keeps the reference's pressure
// (gamma - 1)(rho E - |u|^2 / 2).  One thread per sample point.
template <int P>
__global__ __launch_bounds__(256) void dgx_sample_kernel(const double* __restrict__ V,
                                                         int nelem, int Ns, int var,
                                                         double gamma,
                                                         double* __restrict__ out)
{
  constexpr int N = P + 1, NB = N * N * N;
  const PTab& pt = c_ptab[P - 1];
  const int np1 = Ns + 1, nsub = np1 * np1 * np1;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nelem * nsub)
    return;
  const int node = (int)(t % nsub);
  const size_t e = t / nsub;
  const int ijk[3] = {node % np1, (node / np1) % np1, node / (np1 * np1)};
  double l[3][N];
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    const double x = (double)ijk[c] / (double)Ns;
#pragma unroll
    for (int m = 0; m < N; ++m)
    {
      double v = 1.0;
#pragma unroll
      for (int o = 0; o < N; ++o)
        if (o != m)
          v *= (x - pt.g[o]) / (pt.g[m] - pt.g[o]);
      l[c][m] = v;
    }
  }
  double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double* Ve = V + e * 5 * NB;
#pragma unroll
  for (int kk = 0; kk < N; ++kk)
#pragma unroll
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
  out[t] = v;
}

// update_mean (output.cpp:39-49): mean = (mean * old_range + U) / new_range.
__global__ __launch_bounds__(256) void dgx_mean_kernel(const double* __restrict__ U,
                                                       double* __restrict__ mean, size_t n,
                                                       double old_range, double new_range)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    mean[i] = (mean[i] * old_range + U[i]) / new_range;
}

// ---- discrete adjoint --------------------------------------------------------
#include "dgx_adjoint_kernels.hpp"

// ---- launch helpers (host) -----------------------------------------------

template <int P, int Q, typename TT, bool AFF, int CT>
hipError_t launch_tile_ct(const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                          const double* lam, int* status, int waves, hipStream_t st)
{
  constexpr int R = tile_rows(P);
  constexpr int E = TileDims<P, R>::E;
  int nblk = (m.nelem + E - 1) / E;
  const bool norm = sa.normp != nullptr;
  if (waves == 2 && norm)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 2, TT, AFF, CT, true>), dim3(nblk),
                       dim3(64), 0, st, m, ph, sa, lam, status);
  else if (waves == 2)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 2, TT, AFF, CT>), dim3(nblk), dim3(64), 0,
                       st, m, ph, sa, lam, status);
  else if (waves == 3 && norm)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 3, TT, AFF, CT, true>), dim3(nblk),
                       dim3(64), 0, st, m, ph, sa, lam, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_tile_stage_kernel<P, Q, R, 3, TT, AFF, CT>), dim3(nblk), dim3(64), 0,
                       st, m, ph, sa, lam, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// the matrix-core contraction forms exist for p = 3 with FP32 trace records
template <int P, int Q, typename TT, bool AFF>
hipError_t launch_tile_t(const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                         const double* lam, int* status, int waves, hipStream_t st)
{
  if constexpr (P == 3 && std::is_same<TT, float>::value)
  {
    if (sa.ctype == 1)
      return launch_tile_ct<P, Q, TT, AFF, 1>(m, ph, sa, lam, status, waves, st);
    if (sa.ctype == 2)
      return launch_tile_ct<P, Q, TT, AFF, 2>(m, ph, sa, lam, status, waves, st);
  }
  return launch_tile_ct<P, Q, TT, AFF, 0>(m, ph, sa, lam, status, waves, st);
}

template <int P, int Q, typename TT, bool AFF>
hipError_t launch_rhs_t(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                        int* status, int waves, int tile, hipStream_t st)
{
  if constexpr (Q == 1)
    if (tile)
      return launch_tile_t<P, Q, TT, AFF>(m, ph, sa, lam, status, waves, st);
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  if (waves == 1)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 1, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else if (waves == 2)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 2, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_rhs_stage_kernel<P, Q, 3, TT, AFF>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// affine q = 1 meshes take the constant-Jacobian instantiation
template <int P, int Q, typename TT>
hipError_t launch_rhs_a(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                        int* status, int waves, int tile, hipStream_t st)
{
  if constexpr (Q == 1)
    if (m.affine)
      return launch_rhs_t<P, Q, TT, true>(m, ph, sa, lam, status, waves, tile, st);
  return launch_rhs_t<P, Q, TT, false>(m, ph, sa, lam, status, waves, tile, st);
}

template <int P, int Q>
hipError_t launch_rhs(const MeshArgs& m, const Phys& ph, const StageArgs& sa, const double* lam,
                      int* status, int waves, int tile, hipStream_t st)
{
  if (sa.trace32)
    return launch_rhs_a<P, Q, float>(m, ph, sa, lam, status, waves, tile, st);
  return launch_rhs_a<P, Q, double>(m, ph, sa, lam, status, waves, tile, st);
}

template <int P, int Q, bool AFF>
hipError_t launch_rhs_batch_t(const BatchItem* items, int nitems, int nblk, int* status,
                              int waves, hipStream_t st)
{
  if (waves == 1)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 1, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else if (waves == 2)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 2, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else if (waves == 3)
    hipLaunchKernelGGL((dgx_rhs_batch_kernel<P, Q, 3, AFF>), dim3(nblk), dim3(64), 0, st, items,
                       nitems, status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// groups whose members are all affine q = 1 meshes take the constant-
// Jacobian instantiation
template <int P, int Q>
hipError_t launch_rhs_batch(const BatchItem* items, int nitems, int nblk, int* status, int waves,
                            bool affine, hipStream_t st)
{
  if constexpr (Q == 1)
    if (affine)
      return launch_rhs_batch_t<P, Q, true>(items, nitems, nblk, status, waves, st);
  return launch_rhs_batch_t<P, Q, false>(items, nitems, nblk, status, waves, st);
}

template <int P, int Q>
hipError_t launch_trace(const MeshArgs& m, const Phys& ph, const double* V, double* Tout,
                        int* status, hipStream_t st, bool trace32)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  const bool aff = (Q == 1) && m.affine;
  if (trace32 && aff)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, float, Q == 1>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, reinterpret_cast<float*>(Tout), status);
  else if (trace32)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, float, false>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, reinterpret_cast<float*>(Tout), status);
  else if (aff)
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, double, Q == 1>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, Tout, status);
  else
    hipLaunchKernelGGL((dgx_trace_kernel<P, Q, double, false>), dim3(nblk), dim3(64), 0, st, m,
                       ph, V, Tout, status);
  return hipGetLastError();
}

template <int P, int Q>
hipError_t launch_lift_coef(const MeshArgs& m, double* lam, hipStream_t st)
{
  size_t total = (size_t)m.nelem * 6 * (P + 1) * (P + 1);
  int nblk = (int)((total + 63) / 64);
  hipLaunchKernelGGL((dgx_lift_coef_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, lam);
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

size_t trace_doubles(int p, int nx, int ny, int nz)
{
  size_t nf = (size_t)(p + 1) * (p + 1);
  return (size_t)nx * ny * nz * 6 * 8 * nf + boundary_faces(nx, ny, nz) * 5 * nf;
}

// general orders: one partial per workgroup of gen_elements_per_block
// elements
int resid_out_blocks(int p, int q, int nelem)
{
  if (general_order(p, q))
  {
    const int epb = gen_elements_per_block(p, q);
    return (nelem + epb - 1) / epb;
  }
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

// the tiled kernels are compiled for q = 1 (with higher-order geometry the
// extra rows of accumulators spill); tile 2, the general-order kernel, runs
// every order (the only one outside 1 <= p, q <= 3; no occupancy targets)
bool tile_valid(int p, int q, int tile, int waves)
{
  if (general_order(p, q))
    return tile == 2;
  return tile == 0 || tile == 2 ||
         (tile == 1 && q == 1 && tile_rows(p) > 0 && (waves == 2 || waves == 3));
}

hipError_t rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const StageArgs& sa,
                        const double* lam, int* status, int waves, hipStream_t st, int tile)
{
  if (!tile_valid(p, q, tile, waves))
    return hipErrorInvalidValue;
  if (tile == 2)
    return gen_rhs_dispatch(p, q, m, ph, sa, lam, status, st);
#define CALL_RHS(P, Q) launch_rhs<P, Q>(m, ph, sa, lam, status, waves, tile, st)
  DGX_DISPATCH_PQ(p, q, CALL_RHS)
#undef CALL_RHS
}

hipError_t rhs_batch_dispatch(int p, int q, const BatchItem* items_dev, int nitems, int nblocks,
                              int* status, int waves, bool affine, hipStream_t st)
{
#define CALL_RB(P, Q) launch_rhs_batch<P, Q>(items_dev, nitems, nblocks, status, waves, affine, st)
  DGX_DISPATCH_PQ(p, q, CALL_RB)
#undef CALL_RB
}

template <int P, int Q>
static hipError_t launch_inviscid(const MeshArgs& m, const Phys& ph, const double* V,
                                  const double* Tin, double* fout, const double* lam, int* status,
                                  hipStream_t st)
{
  StageArgs sa{};
  sa.Vin     = V;
  sa.Tin     = Tin;
  sa.fout    = fout;
  sa.ca      = 1.0;
  sa.accmode = 0;
  const int nblk = stage_blocks(P, m.nelem);
  if (Q == 1 && m.affine)
    hipLaunchKernelGGL((dgx_inviscid_rhs_kernel<P, Q, Q == 1>), dim3(nblk), dim3(64), 0, st, m,
                       ph, sa, lam, status);
  else
    hipLaunchKernelGGL((dgx_inviscid_rhs_kernel<P, Q, false>), dim3(nblk), dim3(64), 0, st, m, ph,
                       sa, lam, status);
  return hipGetLastError();
}

hipError_t inviscid_rhs_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                                 const double* Tin, double* fout, const double* lam, int* status,
                                 hipStream_t st)
{
#define CALL_INV(P, Q) launch_inviscid<P, Q>(m, ph, V, Tin, fout, lam, status, st)
  DGX_DISPATCH_PQ(p, q, CALL_INV)
#undef CALL_INV
}

int tile_blocks(int p, int nelem)
{
  const int rows = tile_rows(p);
  const int e = rows > 0 ? rows * (64 / ((p + 1) * (p + 1) * (p + 1))) : 1;
  return (nelem + e - 1) / e;
}

int stage_blocks(int p, int nelem)
{
  int epw = 64 / ((p + 1) * (p + 1) * (p + 1));
  return (nelem + epw - 1) / epw;
}

hipError_t trace_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                          double* Tout, int* status, hipStream_t st, bool trace32)
{
  if (general_order(p, q))
    return gen_trace_dispatch(p, q, m, ph, V, Tout, status, st, trace32);
#define CALL_TR(P, Q) launch_trace<P, Q>(m, ph, V, Tout, status, st, trace32)
  DGX_DISPATCH_PQ(p, q, CALL_TR)
#undef CALL_TR
}

__global__ __launch_bounds__(256) void dgx_lift_face_kernel(const double* __restrict__ lam, int nf,
                                                           size_t n, double* __restrict__ lamf)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    lamf[i] = lam[i * nf];
}

hipError_t lift_face_dispatch(int p, const double* lam, int nelem, double* lamf, hipStream_t st)
{
  const size_t n = (size_t)nelem * 6;
  hipLaunchKernelGGL(dgx_lift_face_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lam,
                     (p + 1) * (p + 1), n, lamf);
  return hipGetLastError();
}

hipError_t lift_coef_dispatch(int p, int q, const MeshArgs& m, double* lam, hipStream_t st)
{
  if (general_order(p, q))
    return gen_lift_coef_dispatch(p, q, m, lam, st);
#define CALL_LC(P, Q) launch_lift_coef<P, Q>(m, lam, st)
  DGX_DISPATCH_PQ(p, q, CALL_LC)
#undef CALL_LC
}

int norm_blocks(int p, int q, int nelem)
{
  if (general_order(p, q))
    return resid_out_blocks(p, q, nelem);
  const int epb = 4 * (64 / ((p + 1) * (p + 1) * (p + 1)));
  return (nelem + epb - 1) / epb;
}

template <int P, int Q>
static hipError_t launch_norm(const MeshArgs& m, const double* fg, double* partial,
                              hipStream_t st)
{
  hipLaunchKernelGGL((dgx_norm_kernel<P, Q>), dim3(norm_blocks(P, Q, m.nelem)), dim3(256), 0, st, m,
                     fg, partial);
  return hipGetLastError();
}

hipError_t norm_partials_dispatch(int p, int q, const MeshArgs& m, const double* fg,
                                  double* partial, hipStream_t st)
{
  if (general_order(p, q))
    return gen_resid_out_dispatch(p, q, m, fg, nullptr, nullptr, partial, st);
#define CALL_NM(P, Q) launch_norm<P, Q>(m, fg, partial, st)
  DGX_DISPATCH_PQ(p, q, CALL_NM)
#undef CALL_NM
}

hipError_t resid_out_dispatch(int p, int q, const MeshArgs& m, const double* fg, double* f,
                              double* R, double* partial, hipStream_t st)
{
  if (general_order(p, q))
    return gen_resid_out_dispatch(p, q, m, fg, f, R, partial, st);
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
  return gen_basis_dispatch(p, in, out, nelem, mode, st);
}

hipError_t sample_dispatch(int p, const double* V, int nelem, int Ns, int var, double gamma,
                           double* out, hipStream_t st)
{
  size_t total = (size_t)nelem * (Ns + 1) * (Ns + 1) * (Ns + 1);
  int nblk = (int)((total + 255) / 256);
  switch (p)
  {
    case 1:
      hipLaunchKernelGGL(dgx_sample_kernel<1>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    case 2:
      hipLaunchKernelGGL(dgx_sample_kernel<2>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    case 3:
      hipLaunchKernelGGL(dgx_sample_kernel<3>, dim3(nblk), dim3(256), 0, st, V, nelem, Ns, var,
                         gamma, out);
      break;
    default: return gen_sample_dispatch(p, V, nelem, Ns, var, gamma, out, st);
  }
  return hipGetLastError();
}

hipError_t mean_dispatch(const double* U, double* mean, size_t n, double old_range,
                         double new_range, hipStream_t st)
{
  int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(dgx_mean_kernel, dim3(nblk), dim3(256), 0, st, U, mean, n, old_range,
                     new_range);
  return hipGetLastError();
}

template <int P, int Q, bool AFF>
hipError_t launch_adjoint_t(const MeshArgs& m, const Phys& ph, const double* V, const double* Tin,
                            const double* lam, const double* psi, double* pvol, double* aown,
                            double* anbr, double* psiV, hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  // the volume part keeps the general geometry: its affine form holds 20
  // more VGPRs (146 vs 126 at p = 2), one wave/SIMD less, and runs slower
  hipLaunchKernelGGL((dgx_adj_main_kernel<P, Q, 0, false>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     Tin, lam, psi, pvol, aown, anbr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL((dgx_adj_main_kernel<P, Q, 1, AFF>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     Tin, lam, psi, pvol, aown, anbr);
  e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  bool walls = false;
  for (int i = 0; i < 6; ++i)
    walls = walls || m.bounds[i] != -2;
  if (walls)
  {
    const size_t npts = boundary_faces(m.nx, m.ny, m.nz) * (size_t)((P + 1) * (P + 1));
    hipLaunchKernelGGL((dgx_adj_boundary_kernel<P, Q>), dim3((unsigned)((npts + 63) / 64)),
                       dim3(64), 0, st, m, ph, Tin, lam, aown, anbr);
    e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL((dgx_adj_trace_kernel<P, Q, AFF>), dim3(nblk), dim3(64), 0, st, m, ph, V,
                     aown, anbr, pvol, psiV);
  return hipGetLastError();
}

// affine q = 1 meshes take the constant-Jacobian instantiation
template <int P, int Q>
hipError_t launch_adjoint(const MeshArgs& m, const Phys& ph, const double* V, const double* Tin,
                          const double* lam, const double* psi, double* pvol, double* aown,
                          double* anbr, double* psiV, hipStream_t st)
{
  if constexpr (Q == 1)
    if (m.affine)
      return launch_adjoint_t<P, Q, true>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st);
  return launch_adjoint_t<P, Q, false>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st);
}

hipError_t adjoint_dispatch(int p, int q, const MeshArgs& m, const Phys& ph, const double* V,
                            const double* Tin, const double* lam, const double* psi,
                            double* pvol, double* aown, double* anbr, double* psiV,
                            hipStream_t st)
{
#define CALL_ADJ(P, Q) launch_adjoint<P, Q>(m, ph, V, Tin, lam, psi, pvol, aown, anbr, psiV, st)
  DGX_DISPATCH_PQ(p, q, CALL_ADJ)
#undef CALL_ADJ
}

template <int P, int Q>
hipError_t launch_kinetic_energy(const MeshArgs& m, const double* V, double* grad,
                                 double* partial, hipStream_t st)
{
  constexpr int EPW = Dims<P, Q>::EPW;
  int nblk = (m.nelem + EPW - 1) / EPW;
  hipLaunchKernelGGL((dgx_kinetic_energy_kernel<P, Q>), dim3(nblk), dim3(64), 0, st, m, V, grad,
                     partial);
  return hipGetLastError();
}

hipError_t kinetic_energy_dispatch(int p, int q, const MeshArgs& m, const double* V,
                                   double* grad, double* partial, hipStream_t st)
{
#define CALL_KE(P, Q) launch_kinetic_energy<P, Q>(m, V, grad, partial, st)
  DGX_DISPATCH_PQ(p, q, CALL_KE)
#undef CALL_KE
}

hipError_t adj_tail_dispatch(const double* l, const double* a, const double* sum_in,
                             double* sum_out, double* lk, double wl, double wa, size_t n,
                             hipStream_t st)
{
  int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(dgx_adj_tail_kernel, dim3(nblk), dim3(256), 0, st, l, a, sum_in, sum_out,
                     lk, wl, wa, n);
  return hipGetLastError();
}

hipError_t dot_dispatch(const double* x, const double* y, const double* z, double scale,
                        size_t n, double* partial, int nblocks, double* out, hipStream_t st)
{
  hipLaunchKernelGGL(dgx_dot_kernel, dim3(nblocks), dim3(256), 0, st, x, y, z, scale, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  return sum_dispatch(partial, nblocks, out, st);
}

size_t sum_scratch(int n)
{
  return n > SUM_ONE_BLOCK ? (size_t)(n + SUM_CHUNK - 1) / SUM_CHUNK : 0;
}

// n partials at partial[0..n), then sum_scratch(n) doubles of scratch at
// partial[n..): one pass of one workgroup, or chunk sums first
static hipError_t reduce_dispatch(const double* partial, int n, double scale, double* out,
                                  hipStream_t st)
{
  if (n <= SUM_ONE_BLOCK)
  {
    hipLaunchKernelGGL(dgx_sum_kernel, dim3(1), dim3(256), 0, st, partial, n, out, scale);
    return hipGetLastError();
  }
  const int nc = (int)sum_scratch(n);
  double* chunk = const_cast<double*>(partial) + n;
  hipLaunchKernelGGL(dgx_chunk_sum_kernel, dim3(nc), dim3(256), 0, st, partial, n, chunk);
  hipLaunchKernelGGL(dgx_sum_kernel, dim3(1), dim3(256), 0, st, chunk, nc, out, scale);
  return hipGetLastError();
}

hipError_t sum_dispatch(const double* partial, int n, double* out, hipStream_t st)
{
  return reduce_dispatch(partial, n, 0.0, out, st);
}

hipError_t norm_dispatch(const double* partial, int n, double inv_count, double* out,
                         hipStream_t st)
{
  return reduce_dispatch(partial, n, inv_count, out, st);
}

hipError_t upload_tables(const PTab* pt, const QTab* qt)
{
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ptab), pt, sizeof(PTab) * 3);
  if (e != hipSuccess)
    return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_qtab), qt, sizeof(QTab) * 9);
}

}  // namespace dgx
