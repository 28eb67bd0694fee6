// dgx_ensemble.cpp -- batched ensembles of independent cases.
//
// The reference runs one case per process (dg.cpp / cudg.cu).  An ensemble
// keeps one ordinary context per member (mesh, state, stage buffers, trace
// records, lifting coefficients) and advances all members together.  Per RK
// stage:
//  * one launch for all orders (layout 1, the default): per affine class,
//    all q = 1 members of every order run as ONE launch of the ensemble
//    stage kernel (dgx_ensemble_kernels.hip) over the concatenation of
//    their work units (rows or tiles of one member's elements), largest
//    members first; q > 1 members run one batched launch per (p, q);
//  * one launch per order (layout 2): the same kernel, one launch per
//    (order, affine) class of the q = 1 members (each launch holds one
//    order's body only);
//  * batched (layout 0): one launch of the batched one-row stage kernel per
//    (p, q, affine) group (each grid block finds its member in a sorted
//    item table).
// Members of an order outside 1 <= p, q <= 3 run one launch of the
// general-order stage kernel each, in every layout.
// One RK step of the whole ensemble can be captured in a hipGraph and
// replayed (kernels only).  The autotuner picks, per launch group, the
// occupancy target and, for the ensemble kernel, the tile (one-row units or
// tiled units: elements per workgroup and nodes per lane), then the layout,
// by timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_ctx.hpp"
#include "dgx_launch.hpp"

using namespace dgx;

namespace
{
struct EnsGroup
{
  int p, q;      // p = 0 for a one-launch group of q = 1 members of any order
  int affine;    // every member an affine q = 1 mesh
  bool unit_grid;  // the ensemble stage kernel (layouts 1, 2), else batched
  std::vector<int> members;
  int nblocks;   // grid blocks (batched) or work units (ensemble kernel)
  int waves;
  int tile;      // ensemble kernel: 0 one-row units, 1 tiled units
  bool general;  // members of an order outside 1 <= p, q <= 3: one launch of
                 // the general-order kernel per member (host items), in
                 // every layout
  BatchItem* items[4];  // device item tables, one per stage
  std::vector<BatchItem> host[4];  // the same on the host (general groups)
};
}  // namespace

struct dgx_ensemble
{
  int device;
  hipStream_t stream;
  std::vector<dgx_ctx*> ctx;
  std::vector<EnsGroup> groups;
  int* status;
  // stage items are built for one (scheme, dt[]) at a time
  bool planned;
  int scheme;
  int nstages;
  std::vector<double> dt;
  std::vector<Phys> phys;  // each member's physics when the items were built
  int layout;              // 0 batched, 1 one launch for all orders, 2 one per order
  bool graph_ready;
  hipGraph_t graph;
  hipGraphExec_t exec;
  char err[512];
};

static thread_local char g_ens_err[512] = "";

static dgx_status efail(dgx_ensemble* en, dgx_status st, const char* fmt, ...)
{
  char* buf = en ? en->err : g_ens_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, 512, fmt, ap);
  va_end(ap);
  return st;
}

#define EHIP(en, call)                                                                     \
  do                                                                                       \
  {                                                                                        \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return efail((en), DGX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

static void drop_graph(dgx_ensemble* en)
{
  if (en->graph_ready)
  {
    (void)hipGraphExecDestroy(en->exec);
    (void)hipGraphDestroy(en->graph);
    en->graph_ready = false;
  }
}

// Whether group g runs in the current layout.
static bool active(const dgx_ensemble* en, const EnsGroup& g)
{
  if (g.general)
    return true;
  if (en->layout == 0)
    return !g.unit_grid;
  if (!g.unit_grid)
    return g.q != 1;
  return (en->layout == 1) == (g.p == 0);
}

// True when some member's physics changed (dgx_set_params on a member
// context) since the item tables were built: the tables and a captured graph
// carry a copy of it.
static bool phys_changed(const dgx_ensemble* en)
{
  for (size_t i = 0; i < en->ctx.size(); ++i)
    if (std::memcmp(&en->phys[i], &en->ctx[i]->phys, sizeof(Phys)) != 0)
      return true;
  return false;
}

// Builds the per-stage item tables of every group for (scheme, dt[]) and
// the members' current physics.
static dgx_status plan(dgx_ensemble* en, int scheme, const double* dt)
{
  const int n = (int)en->ctx.size();
  if (en->planned && en->scheme == scheme &&
      std::memcmp(en->dt.data(), dt, sizeof(double) * n) == 0 && !phys_changed(en))
    return DGX_OK;
  drop_graph(en);
  int ns = 0;
  for (EnsGroup& g : en->groups)
  {
    std::vector<BatchItem> host[4];
    int block0 = 0;
    for (int m : g.members)
    {
      dgx_ctx* c = en->ctx[m];
      StageArgs sa[4];
      ns = stage_plan(c, scheme, dt[m], sa);
      for (int s = 0; s < ns; ++s)
      {
        BatchItem it;
        it.mesh = mesh_args(c);
        it.ph = c->phys;
        it.sa = sa[s];
        it.lam = c->lam;
        it.block0 = block0;
        it.p = c->p;
        host[s].push_back(it);
      }
      block0 += g.general ? c->nelem
                          : (g.unit_grid ? ens_units(c->p, c->nelem, g.tile)
                                         : stage_blocks(c->p, c->nelem));
    }
    g.nblocks = block0;
    for (int s = 0; s < 4; ++s)
      g.host[s] = g.general ? host[s] : std::vector<BatchItem>();
    EHIP(en, hipStreamSynchronize(en->stream));  // no launch may still read the tables
    for (int s = 0; s < ns; ++s)
      EHIP(en, hipMemcpy(g.items[s], host[s].data(), sizeof(BatchItem) * host[s].size(),
                         hipMemcpyHostToDevice));
  }
  en->planned = true;
  en->scheme = scheme;
  en->nstages = ns;
  en->dt.assign(dt, dt + n);
  en->phys.resize(n);
  for (int i = 0; i < n; ++i)
    en->phys[i] = en->ctx[i]->phys;
  return DGX_OK;
}

static hipError_t enqueue_stage(dgx_ensemble* en, int s, int only_group = -1)
{
  for (int gi = 0; gi < (int)en->groups.size(); ++gi)
  {
    if (only_group >= 0 && gi != only_group)
      continue;
    EnsGroup& g = en->groups[gi];
    if (!active(en, g))
      continue;
    if (g.general)
    {
      for (const BatchItem& it : g.host[s])
      {
        hipError_t e = gen_rhs_dispatch(it.p, g.q, it.mesh, it.ph, it.sa, it.lam, en->status,
                                        en->stream);
        if (e != hipSuccess)
          return e;
      }
      continue;
    }
    hipError_t e = g.unit_grid
                       ? ens_stage_dispatch(g.items[s], (int)g.members.size(), g.nblocks,
                                            en->status, g.waves, g.tile, g.affine != 0, g.p,
                                            en->stream)
                       : rhs_batch_dispatch(g.p, g.q, g.items[s], (int)g.members.size(),
                                            g.nblocks, en->status, g.waves, g.affine != 0,
                                            en->stream);
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

static hipError_t enqueue_step(dgx_ensemble* en)
{
  for (int s = 0; s < en->nstages; ++s)
  {
    hipError_t e = enqueue_stage(en, s);
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

// Trace records of members whose state was set from outside; every member
// stream is idle afterwards.
static dgx_status prepare_members(dgx_ensemble* en)
{
  for (dgx_ctx* c : en->ctx)
  {
    EHIP(en, ensure_traces(c));
    EHIP(en, hipStreamSynchronize(c->stream));
  }
  return DGX_OK;
}

extern "C" {

dgx_status dgx_ensemble_create(int ncase, const dgx_mesh* meshes, const dgx_params* params,
                               dgx_ensemble** out)
{
  if (ncase < 1 || !meshes || !params || !out)
    return efail(nullptr, DGX_ERR_INVALID, "invalid ensemble arguments");
  *out = nullptr;
  dgx_ensemble* en = new dgx_ensemble();
  en->planned = false;
  en->graph_ready = false;
  en->status = nullptr;
  en->stream = nullptr;
  en->err[0] = 0;
  for (int i = 0; i < ncase; ++i)
  {
    dgx_ctx* c = nullptr;
    dgx_status st = dgx_create(&meshes[i], &params[i], &c);
    if (st != DGX_OK)
    {
      efail(nullptr, st, "member %d: %s", i, dgx_last_error(nullptr));
      dgx_ensemble_destroy(en);
      return st;
    }
    en->ctx.push_back(c);
  }
  en->device = en->ctx[0]->device;
  en->layout = 1;
  auto group_of = [&](bool unit_grid, int p, int q, int affine) -> EnsGroup& {
    for (EnsGroup& x : en->groups)
      if (x.unit_grid == unit_grid && x.p == p && x.q == q && x.affine == affine)
        return x;
    EnsGroup ng;
    ng.p = p;
    ng.q = q;
    ng.affine = affine;
    ng.unit_grid = unit_grid;
    ng.nblocks = 0;
    ng.waves = unit_grid ? 3 : default_stage_waves(p, q);
    ng.tile = 0;
    ng.general = !unit_grid && general_order(p, q);
    for (int s = 0; s < 4; ++s)
      ng.items[s] = nullptr;
    en->groups.push_back(ng);
    return en->groups.back();
  };
  // batched groups per (p, q, affine); one-launch groups per affine class of
  // the q = 1 members, and per (order, affine class), largest members (most
  // elements x nodes) first
  for (int i = 0; i < ncase; ++i)
  {
    dgx_ctx* c = en->ctx[i];
    group_of(false, c->p, c->q, c->affine).members.push_back(i);
  }
  std::vector<int> order(ncase);
  for (int i = 0; i < ncase; ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (size_t)en->ctx[a]->nelem * en->ctx[a]->nbf > (size_t)en->ctx[b]->nelem * en->ctx[b]->nbf;
  });
  for (int i : order)
  {
    dgx_ctx* c = en->ctx[i];
    if (c->q == 1 && !general_order(c->p, c->q))
    {
      group_of(true, 0, 1, c->affine).members.push_back(i);
      group_of(true, c->p, 1, c->affine).members.push_back(i);
    }
  }
  auto bail = [&](const char* what, hipError_t e) {
    snprintf(g_ens_err, sizeof(g_ens_err), "%s failed: %s", what, hipGetErrorString(e));
    dgx_ensemble_destroy(en);
    return DGX_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&en->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail("hipStreamCreate", e);
  for (EnsGroup& g : en->groups)
    for (int s = 0; s < 4; ++s)
      if ((e = hipMalloc(&g.items[s], sizeof(BatchItem) * g.members.size())) != hipSuccess)
        return bail("hipMalloc(items)", e);
  if ((e = hipMalloc(&en->status, sizeof(int))) != hipSuccess)
    return bail("hipMalloc(status)", e);
  if ((e = hipMemsetAsync(en->status, 0, sizeof(int), en->stream)) != hipSuccess)
    return bail("hipMemset(status)", e);
  if ((e = hipStreamSynchronize(en->stream)) != hipSuccess)
    return bail("hipStreamSynchronize", e);
  *out = en;
  return DGX_OK;
}

dgx_status dgx_ensemble_destroy(dgx_ensemble* en)
{
  if (!en)
    return DGX_OK;
  if (en->stream)
    (void)hipStreamSynchronize(en->stream);
  drop_graph(en);
  for (EnsGroup& g : en->groups)
    for (int s = 0; s < 4; ++s)
      if (g.items[s])
        (void)hipFree(g.items[s]);
  if (en->status)
    (void)hipFree(en->status);
  for (dgx_ctx* c : en->ctx)
    dgx_destroy(c);
  if (en->stream)
    (void)hipStreamDestroy(en->stream);
  delete en;
  return DGX_OK;
}

dgx_status dgx_ensemble_member(dgx_ensemble* en, int i, dgx_ctx** ctx)
{
  if (!en || !ctx || i < 0 || i >= (int)en->ctx.size())
    return DGX_ERR_INVALID;
  *ctx = en->ctx[i];
  return DGX_OK;
}

dgx_status dgx_ensemble_step(dgx_ensemble* en, int scheme, const double* dt, uint64_t nsteps,
                             int use_graph)
{
  if (!en || !dt || (scheme != DGX_RK4 && scheme != DGX_TVDRK3))
    return DGX_ERR_INVALID;
  for (const dgx_ctx* c : en->ctx)
    if (c->prec != DGX_PREC_FP64)
      return efail(en, DGX_ERR_UNSUPPORTED, "ensemble steps need DGX_PREC_FP64 members");
  dgx_status st = plan(en, scheme, dt);
  if (st != DGX_OK)
    return st;
  if ((st = prepare_members(en)) != DGX_OK)
    return st;
  for (dgx_ctx* c : en->ctx)
    c->tu_valid = false;
  if (use_graph && !en->graph_ready && nsteps > 0)
  {
    EHIP(en, hipStreamBeginCapture(en->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_step(en);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(en->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess)
    {
      if (g)
        (void)hipGraphDestroy(g);
      return efail(en, DGX_ERR_HIP, "graph capture failed: %s",
                   hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipGraphExec_t x;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
      (void)hipGraphDestroy(g);
      return efail(en, DGX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    en->graph = g;
    en->exec = x;
    en->graph_ready = true;
  }
  for (uint64_t k = 0; k < nsteps; ++k)
  {
    if (use_graph)
      EHIP(en, hipGraphLaunch(en->exec, en->stream));
    else
      EHIP(en, enqueue_step(en));
  }
  for (dgx_ctx* c : en->ctx)
    c->tu_valid = true;
  return DGX_OK;
}

dgx_status dgx_ensemble_sync(dgx_ensemble* en)
{
  if (!en)
    return DGX_ERR_INVALID;
  EHIP(en, hipStreamSynchronize(en->stream));
  int h = 0;
  EHIP(en, hipMemcpy(&h, en->status, sizeof(int), hipMemcpyDeviceToHost));
  if (h)
  {
    EHIP(en, hipMemset(en->status, 0, sizeof(int)));
    return efail(en, DGX_ERR_INVALID, "negative inflow boundary Mach number in a member");
  }
  return DGX_OK;
}

dgx_status dgx_ensemble_get_stream(dgx_ensemble* en, void** stream)
{
  if (!en || !stream)
    return DGX_ERR_INVALID;
  *stream = (void*)en->stream;
  return DGX_OK;
}

dgx_status dgx_ensemble_autotune(dgx_ensemble* en, int scheme, const double* dt, int reps,
                                 int* waves_out)
{
  if (!en || !dt || reps < 1 || (scheme != DGX_RK4 && scheme != DGX_TVDRK3))
    return DGX_ERR_INVALID;
  for (const dgx_ctx* c : en->ctx)
    if (c->prec != DGX_PREC_FP64)
      return efail(en, DGX_ERR_UNSUPPORTED, "ensemble steps need DGX_PREC_FP64 members");
  dgx_status st = plan(en, scheme, dt);
  if (st != DGX_OK)
    return st;
  if ((st = prepare_members(en)) != DGX_OK)
    return st;
  drop_graph(en);
  hipEvent_t e0, e1;
  EHIP(en, hipEventCreate(&e0));
  EHIP(en, hipEventCreate(&e1));
  // stage 1 reads U and its records and writes only stage buffers
  auto one = [&](int gi) -> hipError_t { return enqueue_stage(en, 0, gi); };
  auto timed = [&](int gi, float& ms) -> dgx_status {
    EHIP(en, one(gi));
    EHIP(en, hipEventRecord(e0, en->stream));
    for (int r = 0; r < reps; ++r)
      EHIP(en, one(gi));
    EHIP(en, hipEventRecord(e1, en->stream));
    EHIP(en, hipEventSynchronize(e1));
    EHIP(en, hipEventElapsedTime(&ms, e0, e1));
    return DGX_OK;
  };
  auto replan = [&]() -> dgx_status {
    en->planned = false;  // units and grid depend on the tile and target
    return plan(en, scheme, dt);
  };
  // per launch group of the current layout: occupancy target and tile
  auto tune_groups = [&]() -> dgx_status {
    for (int gi = 0; gi < (int)en->groups.size(); ++gi)
    {
      EnsGroup& g = en->groups[gi];
      if (!active(en, g) || g.general)
        continue;
      int best_w = g.waves, best_t = g.tile;
      float best_ms = 0.0f;
      bool have = false;
      for (int t = 0; t <= (g.unit_grid ? 1 : 0); ++t)
        for (int w = g.unit_grid ? 2 : 1; w <= 3; ++w)
        {
          g.waves = w;
          if (g.unit_grid)
          {
            g.tile = t;
            dgx_status ps = replan();
            if (ps != DGX_OK)
              return ps;
          }
          float ms = 0.0f;
          dgx_status ts = timed(gi, ms);
          if (ts != DGX_OK)
            return ts;
          if (!have || ms < best_ms)
          {
            best_ms = ms;
            best_w = w;
            best_t = t;
            have = true;
          }
        }
      g.waves = best_w;
      g.tile  = best_t;
      if (g.unit_grid)
      {
        dgx_status ps = replan();
        if (ps != DGX_OK)
          return ps;
      }
    }
    return DGX_OK;
  };
  // then the layout: the first stage of every active group, per layout
  int best_layout = en->layout;
  float best_ms = 0.0f;
  bool have = false;
  for (int layout : {1, 2, 0})
  {
    en->layout = layout;
    dgx_status st2 = replan();
    if (st2 == DGX_OK)
      st2 = tune_groups();
    float ms = 0.0f;
    if (st2 == DGX_OK)
      st2 = timed(-1, ms);
    if (st2 != DGX_OK)
      return st2;
    if (!have || ms < best_ms)
    {
      best_ms = ms;
      best_layout = layout;
      have = true;
    }
  }
  en->layout = best_layout;
  if ((st = replan()) != DGX_OK)
    return st;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (waves_out)
    for (const EnsGroup& g : en->groups)
      if (active(en, g))
        for (int m : g.members)
          waves_out[m] = g.waves;
  return dgx_ensemble_sync(en);
}

dgx_status dgx_ensemble_set_layout(dgx_ensemble* en, int layout)
{
  if (!en || layout < 0 || layout > 2)
    return DGX_ERR_INVALID;
  if (layout != en->layout)
  {
    EHIP(en, hipStreamSynchronize(en->stream));
    drop_graph(en);
    en->layout = layout;
  }
  return DGX_OK;
}

dgx_status dgx_ensemble_member_tuning(const dgx_ensemble* en, int i, int* layout, int* tile,
                                      int* waves)
{
  if (!en || i < 0 || i >= (int)en->ctx.size())
    return DGX_ERR_INVALID;
  for (const EnsGroup& g : en->groups)
    if (active(en, g) && std::find(g.members.begin(), g.members.end(), i) != g.members.end())
    {
      if (layout)
        *layout = g.unit_grid ? en->layout : 0;
      if (tile)
        *tile = g.general ? 2 : (g.unit_grid ? g.tile : 0);
      if (waves)
        *waves = g.waves;
      return DGX_OK;
    }
  return DGX_ERR_INVALID;
}

const char* dgx_ensemble_last_error(const dgx_ensemble* en)
{
  return en ? en->err : g_ens_err;
}

}  // extern "C"
