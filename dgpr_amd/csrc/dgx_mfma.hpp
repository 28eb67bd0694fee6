// dgx_mfma.hpp -- the volume contractions of one p = 3 element on the matrix
// cores (the mixed-precision forms of the tiled stage kernel).
//
// At p = 3 an element's 64 Gauss nodes are one wavefront and its 16 lines in
// any reference direction are the 16 blocks of the 4x4 MFMA forms, so one
// MFMA sequence contracts a 4x4 table with 4 variables along every line of
// the element at once:
//   gradient  dV[var]/dxi_d (n) = sum_m Dg[n_d][m] V[var](line, m)
//   projection  acc[r](n)      += sum_m Dg[m][n_d] G[r](line, m)
// with the block's 4 rows the variables and its 4 columns the output points
// of the line.  Block layouts (measured on gfx950 with an in-bounds probe):
//   v_mfma_f32_4x4x1_16b_f32:  lane 4b+i holds A[b][i] and B[b][i] (K = 1);
//   v_mfma_f32_4x4x4_16b_bf16: lane 4b+i holds row i of A[b] and column i
//                              of B[b] (K = 4, four bf16 per operand);
//   both: lane 4b+j, result register i = D[b][i][j].
// Lane l = 16v + 4u + t.  Block b = l/4 is line (u, v) of direction d:
// x-lines (j, k) = (u, v), y-lines (i, k) = (u, v), z-lines (i, j) = (u, v);
// t is the variable (A rows) or the output point (B columns, results).  The
// result for x lands in the lane that owns the node (node index = lane); for
// y and z it lands in a lane whose 2-bit index fields are permuted, and one
// ds_bpermute per variable brings it to the owner.
//
// Two forms:
//   CT = 1  FP32 operands, FP32 accumulation (4 K = 1 steps);
//   CT = 2  split BF16: every FP32 operand x = x0 + x1 + x2 (three bf16 terms,
//           each the rounding of what the previous ones leave), the six
//           products x_a y_b with a + b <= 2 accumulated in FP32 (K = 4 in
//           one instruction each): FP32-level accuracy on BF16 matrix cores.
// Included by dgx_tile.hpp inside namespace dgx.
#pragma once

typedef float dgx_f32x4 __attribute__((ext_vector_type(4)));
typedef short dgx_bf16x4 __attribute__((ext_vector_type(4)));

// node index of point m of line (u, v) in direction D (i fastest)
template <int D>
__device__ __forceinline__ int mfma_node(int m, int u, int v)
{
  return D == 0 ? 16 * v + 4 * u + m : (D == 1 ? 16 * v + 4 * m + u : 16 * m + 4 * v + u);
}

// lane whose result belongs to node (= lane) l, for direction D
template <int D>
__device__ __forceinline__ int mfma_src_lane(int l)
{
  const int a = l >> 4, b = (l >> 2) & 3, c = l & 3;  // l = (k, j, i)
  if constexpr (D == 0)
    return l;
  else if constexpr (D == 1)
    return 16 * a + 4 * c + b;  // result at (k, i, j)
  else
    return 16 * b + 4 * c + a;  // result at (j, i, k)
}

__device__ __forceinline__ float lane_pull(float v, int src)
{
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

// round to nearest even bf16 (finite inputs)
__device__ __forceinline__ short bf16_bits(float x)
{
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_value(short h)
{
  return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}

// x = x0 + x1 + x2 in bf16 terms, component-wise over 4 values
__device__ __forceinline__ void split3(const float x[4], dgx_bf16x4& x0, dgx_bf16x4& x1,
                                       dgx_bf16x4& x2)
{
#pragma unroll
  for (int k = 0; k < 4; ++k)
  {
    const short h0 = bf16_bits(x[k]);
    const float r1 = x[k] - bf16_value(h0);
    const short h1 = bf16_bits(r1);
    const float r2 = r1 - bf16_value(h1);
    x0[k] = h0;
    x1[k] = h1;
    x2[k] = bf16_bits(r2);
  }
}

// D[b][t][n] += sum_m A[b][t][m] Bm[b][m][n]: a[m] is this lane's row
// (A[b][t][m], m = 0..3), bcol[m] its column (B[b][m][t]).
template <int CT>
__device__ __forceinline__ dgx_f32x4 mfma_line4(const float a[4], const float bcol[4],
                                                dgx_f32x4 d)
{
  if constexpr (CT == 1)
  {
#pragma unroll
    for (int m = 0; m < 4; ++m)
      d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[m], bcol[m], d, 0, 0, 0);
    return d;
  }
  else
  {
    dgx_bf16x4 a0, a1, a2, b0, b1, b2;
    split3(a, a0, a1, a2);
    split3(bcol, b0, b1, b2);
    // smallest terms first
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a2, b0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a1, b1, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a0, b2, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a1, b0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a0, b1, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a0, b0, d, 0, 0, 0);
    return d;
  }
}

// Reference-space gradient of (rho, m1, m2, m3) in direction D at this
// lane's node (every lane of the wavefront must call: the MFMA needs all 64
// lanes).  sV holds the element's Gauss-nodal state [5][64] in FP64.
template <int D, int CT>
__device__ __forceinline__ void mfma_grad(const double (*sV)[64], const PTab& pt, double g[4])
{
  const int l = threadIdx.x, t = l & 3, u = (l >> 2) & 3, v = l >> 4;
  float a[4], bcol[4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
  {
    a[m]    = (float)sV[t][mfma_node<D>(m, u, v)];  // row t: variable t along the line
    bcol[m] = (float)pt.Dg[t][m];                   // column t: output point t
  }
  dgx_f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = mfma_line4<CT>(a, bcol, d);
  const int src = mfma_src_lane<D>(l);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    g[r] = (double)(D == 0 ? d[r] : lane_pull(d[r], src));
}

// Projection of the contravariant flux of direction D, held as FP32 in
// sF[5][64] (node = lane), onto this lane's node: acc[r] += sum_m Dg[m][n_D]
// F[r](line, m).  Every lane must call.
template <int D, int CT>
__device__ __forceinline__ void mfma_project(const float (*sF)[64], const PTab& pt,
                                             double acc[5])
{
  const int l = threadIdx.x, t = l & 3, u = (l >> 2) & 3, v = l >> 4;
  float a[4], a4[4], bcol[4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
  {
    const int n = mfma_node<D>(m, u, v);
    a[m]    = sF[t][n];                     // rows 0..3: variables 0..3
    a4[m]   = (t == 0) ? sF[4][n] : 0.0f;   // row 0 of the second product: variable 4
    bcol[m] = (float)pt.Dg[m][t];
  }
  dgx_f32x4 d = {0.f, 0.f, 0.f, 0.f}, d4 = {0.f, 0.f, 0.f, 0.f};
  d  = mfma_line4<CT>(a, bcol, d);
  d4 = mfma_line4<CT>(a4, bcol, d4);
  const int src = mfma_src_lane<D>(l);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    acc[r] += (double)(D == 0 ? d[r] : lane_pull(d[r], src));
  acc[4] += (double)(D == 0 ? d4[0] : lane_pull(d4[0], src));
}
