// dgx_ensemble_kernels.hip -- the ensemble stage kernel of the one-launch
// layouts (no counterpart in the reference, which runs one case per process,
// dg.cpp / cudg.cu).
//
// One launch per RK stage serves every member of a q = 1 group whatever its
// order (PS = 0: the kernel holds the stage bodies of p = 1, 2 and 3, one-row
// or tiled, a template choice, in one LDS region sized for the largest), or
// every member of one order (PS = p: that order's body and LDS only).  The
// grid is the concatenation of the members' work units (one row, or one
// tile, of one member's elements), members in order of decreasing work, so
// the in-order dispatch starts the large high-order members first and the
// small ones fill the tail; each workgroup finds its member by a uniform
// search of the item table.  Each workgroup runs one unit: a loop over units
// (a persistent grid) keeps the body's uniform values live across
// iterations and spills 50-70 scalar and ~100 vector registers at p = 3,
// where the single pass needs 152 VGPRs and no spill.  Every unit runs the
// arithmetic of the single-case kernels, so results are those of dgx_step
// on each member to rounding.
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

// PS = 0: the bodies of every order in one kernel (LDS for the largest);
// PS = 1..3: members of order PS only.  Block b runs unit b.
template <int W, int TILE, bool AFF, int PS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W)))
void dgx_ens_stage_kernel(const BatchItem* __restrict__ items, int nitems, int* status)
{
  constexpr size_t LDS = PS ? ens_body_bytes<(PS ? PS : 1), TILE>() : ens_lds_bytes<TILE>();
  __shared__ __attribute__((aligned(16))) unsigned char lds[LDS];
  const int u = blockIdx.x;
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
  if constexpr (PS != 0)
    ens_body<PS, TILE, AFF>(it, unit, status, lds);
  else
    switch (it.p)
    {
      case 1: ens_body<1, TILE, AFF>(it, unit, status, lds); break;
      case 2: ens_body<2, TILE, AFF>(it, unit, status, lds); break;
      case 3: ens_body<3, TILE, AFF>(it, unit, status, lds); break;
    }
}

int ens_units(int p, int nelem, int tile)
{
  int per = 64 / ((p + 1) * (p + 1) * (p + 1));
  if (tile)
    per *= tile_rows(p);
  return (nelem + per - 1) / per;
}

template <int W, int TILE, bool AFF, int PS>
static hipError_t launch_ens(const BatchItem* items, int nitems, int nunits, int* status,
                             hipStream_t st)
{
  hipLaunchKernelGGL((dgx_ens_stage_kernel<W, TILE, AFF, PS>), dim3(nunits), dim3(64), 0, st,
                     items, nitems, status);
  return hipGetLastError();
}

template <int W, int TILE, bool AFF>
static hipError_t launch_ens_p(const BatchItem* items, int nitems, int nunits, int* status,
                               int psel, hipStream_t st)
{
  switch (psel)
  {
    case 0: return launch_ens<W, TILE, AFF, 0>(items, nitems, nunits, status, st);
    case 1: return launch_ens<W, TILE, AFF, 1>(items, nitems, nunits, status, st);
    case 2: return launch_ens<W, TILE, AFF, 2>(items, nitems, nunits, status, st);
    case 3: return launch_ens<W, TILE, AFF, 3>(items, nitems, nunits, status, st);
  }
  return hipErrorInvalidValue;
}

template <int TILE, bool AFF>
static hipError_t launch_ens_w(const BatchItem* items, int nitems, int nunits, int* status,
                               int waves, int psel, hipStream_t st)
{
  if (waves == 2)
    return launch_ens_p<2, TILE, AFF>(items, nitems, nunits, status, psel, st);
  if (waves == 3)
    return launch_ens_p<3, TILE, AFF>(items, nitems, nunits, status, psel, st);
  return hipErrorInvalidValue;
}

hipError_t ens_stage_dispatch(const BatchItem* items_dev, int nitems, int nunits, int* status,
                              int waves, int tile, bool affine, int psel, hipStream_t st)
{
  if (nunits < 1)
    return hipErrorInvalidValue;
  if (tile)
    return affine ? launch_ens_w<1, true>(items_dev, nitems, nunits, status, waves, psel, st)
                  : launch_ens_w<1, false>(items_dev, nitems, nunits, status, waves, psel, st);
  return affine ? launch_ens_w<0, true>(items_dev, nitems, nunits, status, waves, psel, st)
                : launch_ens_w<0, false>(items_dev, nitems, nunits, status, waves, psel, st);
}

hipError_t upload_tables_ensemble(const PTab* pt, const QTab* qt)
{
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_ptab), pt, sizeof(PTab) * 3);
  if (e != hipSuccess)
    return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_qtab), qt, sizeof(QTab) * 9);
}

}  // namespace dgx
