// dgx_ensemble_kernels.hip -- the persistent stage kernel of batched
// ensembles (no counterpart in the reference, which runs one case per
// process, dg.cpp / cudg.cu).
//
// One launch per RK stage serves every member of a q = 1 group whatever its
// order: the kernel holds the stage bodies of p = 1, 2 and 3 (one-row or
// tiled, a template choice) in one LDS region sized for the largest, and
// each 64-lane workgroup loops taking the next work unit (one row, or one
// tile, of one member's elements) from an atomic counter until every unit
// is done.  Units are ordered by member, members by decreasing work, so the
// large high-order members start first and the small ones fill the tail.
// Every unit runs the arithmetic of the single-case kernels, so results are
// those of dgx_step on each member to rounding.
#include "dgx_stage.hpp"

namespace dgx
{

template <int P, int TILE>
constexpr size_t ens_body_bytes()
{
  if constexpr (TILE == 1)
    return TileLds<P, 1, tile_rows(P)>::bytes;
  else
    return StageLds<P, 1>::bytes;
}

template <int TILE>
constexpr size_t ens_lds_bytes()
{
  constexpr size_t a = ens_body_bytes<1, TILE>(), b = ens_body_bytes<2, TILE>(),
                   c = ens_body_bytes<3, TILE>();
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

template <int P, int TILE, bool AFF>
__device__ __forceinline__ void ens_body(const BatchItem& it, int unit, int* status,
                                         unsigned char* lds)
{
  if constexpr (TILE == 1)
    tile_body_lds<P, 1, tile_rows(P), double, AFF>(it.mesh, it.ph, it.sa, it.lam, status,
                                                       unit, lds);
  else
    stage_body_lds<P, 1, double, AFF>(it.mesh, it.ph, it.sa, it.lam, status, unit, lds);
}

template <int W, int TILE, bool AFF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_ens_stage_kernel(const BatchItem* __restrict__ items, int nitems, int nunits,
                          int* __restrict__ counter, int* status)
{
  __shared__ __attribute__((aligned(16))) unsigned char lds[ens_lds_bytes<TILE>()];
  __shared__ int s_unit;
  for (;;)
  {
    if (threadIdx.x == 0)
      s_unit = atomicAdd(counter, 1);
    __syncthreads();
    const int u = s_unit;
    __syncthreads();  // every lane has read s_unit before the next take
    if (u >= nunits)
      break;          // the counter only grows: every workgroup gets here
    int lo = 0, hi = nitems - 1;
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].block0 <= u)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BatchItem& it = items[lo];
    const int unit = u - it.block0;
    switch (it.p)
    {
      case 1: ens_body<1, TILE, AFF>(it, unit, status, lds); break;
      case 2: ens_body<2, TILE, AFF>(it, unit, status, lds); break;
      case 3: ens_body<3, TILE, AFF>(it, unit, status, lds); break;
    }
    __syncthreads();  // the next unit reuses the LDS region
  }
}

__global__ __launch_bounds__(64) void dgx_ens_reset_kernel(int* counters, int n)
{
  for (int i = threadIdx.x; i < n; i += 64)
    counters[i] = 0;
}

int ens_units(int p, int nelem, int tile)
{
  int per = 64 / ((p + 1) * (p + 1) * (p + 1));
  if (tile)
    per *= tile_rows(p);
  return (nelem + per - 1) / per;
}

template <int W, int TILE, bool AFF>
static hipError_t launch_ens(const BatchItem* items, int nitems, int nunits, int* counter,
                             int* status, int grid, hipStream_t st)
{
  hipLaunchKernelGGL((dgx_ens_stage_kernel<W, TILE, AFF>), dim3(grid), dim3(64), 0, st, items,
                     nitems, nunits, counter, status);
  return hipGetLastError();
}

template <int TILE, bool AFF>
static hipError_t launch_ens_w(const BatchItem* items, int nitems, int nunits, int* counter,
                               int* status, int waves, int grid, hipStream_t st)
{
  if (waves == 2)
    return launch_ens<2, TILE, AFF>(items, nitems, nunits, counter, status, grid, st);
  if (waves == 3)
    return launch_ens<3, TILE, AFF>(items, nitems, nunits, counter, status, grid, st);
  return hipErrorInvalidValue;
}

hipError_t ens_stage_dispatch(const BatchItem* items_dev, int nitems, int nunits, int* counter,
                              int* status, int waves, int tile, bool affine, int grid,
                              hipStream_t st)
{
  if (grid < 1 || nunits < 1)
    return hipErrorInvalidValue;
  if (tile)
    return affine ? launch_ens_w<1, true>(items_dev, nitems, nunits, counter, status, waves,
                                          grid, st)
                  : launch_ens_w<1, false>(items_dev, nitems, nunits, counter, status, waves,
                                           grid, st);
  return affine ? launch_ens_w<0, true>(items_dev, nitems, nunits, counter, status, waves, grid,
                                        st)
                : launch_ens_w<0, false>(items_dev, nitems, nunits, counter, status, waves, grid,
                                         st);
}

template <int W, int TILE, bool AFF>
static int occ()
{
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dgx_ens_stage_kernel<W, TILE, AFF>, 64,
                                                   0) != hipSuccess)
    return 0;
  return n;
}

int ens_blocks_per_cu(int waves, int tile, bool affine)
{
  if (waves == 2)
    return tile ? (affine ? occ<2, 1, true>() : occ<2, 1, false>())
                : (affine ? occ<2, 0, true>() : occ<2, 0, false>());
  if (waves == 3)
    return tile ? (affine ? occ<3, 1, true>() : occ<3, 1, false>())
                : (affine ? occ<3, 0, true>() : occ<3, 0, false>());
  return 0;
}

hipError_t ens_reset_dispatch(int* counters, int n, hipStream_t st)
{
  hipLaunchKernelGGL(dgx_ens_reset_kernel, dim3(1), dim3(64), 0, st, counters, n);
  return hipGetLastError();
}

hipError_t upload_tables_ensemble(const PTab* pt, const QTab* qt)
{
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ptab), pt, sizeof(PTab) * 3);
  if (e != hipSuccess)
    return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_qtab), qt, sizeof(QTab) * 9);
}

}  // namespace dgx
