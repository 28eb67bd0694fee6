// dgx_ctx.hpp -- the device context behind the dgx_ctx handle, shared by
// the C ABI translation units of libdgx (single cases and ensembles).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "../../include/dgx.h"
#include "dgx_launch.hpp"

struct dgx_ctx
{
  int device;
  int p, q;
  int nx, ny, nz, nelem, nbf;
  int bounds[6];
  size_t nstate;  // nelem*5*nbf
  dgx_params params;
  dgx::Phys phys;
  hipStream_t stream;
  double* nodes;    // [e][(q+1)^3][3]
  double* U;        // state, Gauss-nodal basis
  double* Ua;       // stage buffer / staging
  double* Ub;       // stage buffer
  double* acc;      // RK4 accumulator
  double* fs;       // f_g scratch
  double* partial;  // per-block sums of R^2
  double* sum;      // reduced sum
  double* lam;      // BR2 lifting coefficients [e][6][(p+1)^2]
  double* lamf;     // affine meshes: one coefficient per element face [e][6]
  double* TU;       // face trace records of U, Ua, Ub [e][6][8][(p+1)^2]
  double* TUa;
  double* TUb;
  bool tu_valid;    // TU holds the traces of the current U
  int stage_waves;  // occupancy target of the stage kernel (1..3)
  int stage_tile;   // 0: one-row stage kernel, 1: tiled stage kernel
  int prec;         // DGX_PREC_* (float trace records in the mixed forms)
  int affine;       // q = 1 and every element a parallelepiped
  double* mean;     // running mean (Gauss-nodal), allocated on first use
  double* adj_store;       // stored stages of the last dgx_adjoint_rk4 (kept:
  size_t adj_store_bytes;  // freeing tens of GB can take seconds)
  size_t adj_store_limit;  // most bytes that store may take (0: no limit)
  int* status;      // kernel-side error flags
  int npartial;
  size_t bytes;
  char err[512];
};

namespace dgx
{
// MeshArgs of the context's mesh
MeshArgs mesh_args(const dgx_ctx* c);
// trace records of U, recomputed when U was set from outside
hipError_t ensure_traces(dgx_ctx* c);
// StageArgs of the stages of one step without a norm (returns the number of
// stages: 4 for RK4, 3 for TVD-RK3); the last stage writes U and its traces
int stage_plan(const dgx_ctx* c, int scheme, double dt, StageArgs sa[4]);
int default_stage_waves(int p, int q);
}  // namespace dgx
