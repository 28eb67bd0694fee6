// dgx_adjoint.cpp -- discrete adjoint of the residual and of classic RK4
// (no counterpart in the reference, which has no adjoint).
//
// Everything is the exact transpose of what the device computes (the same
// Gauss-nodal residual f_g, the stage algebra of stage_plan), so the
// gradients are those of the discrete forward map, on meshes with any
// boundary conditions (domain-boundary faces are reversed through their
// boundary states by a kernel of their own).
//
// RK4 step (time_distributed.cpp:46-76): S1 = V + dt/2 k1, S2 = V + dt/2 k2,
// S3 = V + dt k3, V' = V + dt/6 (k1 + 2 k2 + 2 k3 + k4), k_i = f(S_(i-1)).
// Reverse, with l the adjoint of V' and a_i = f_V(S_i)^T (adjoint of k_(i+1)):
//   l4 = dt/6 l,            a3 = f_V(S3)^T l4
//   l3 = dt/3 l + dt a3,    a2 = f_V(S2)^T l3
//   l2 = dt/3 l + dt/2 a2,  a1 = f_V(S1)^T l2
//   l1 = dt/6 l + dt/2 a1,  a0 = f_V(V)^T l1
//   adjoint of V = l + a0 + a1 + a2 + a3,
// and dJ/dmu gains sum_i l_i . df/dmu(S_(i-1)), df/dmu = (f(S; mu) - f(S; 0))/mu
// because every viscous and BR2 term is linear in mu; f(S; 0) is the
// inviscid right-hand side, evaluated from S and its stored records.
// Schedules: with memory to spare, the forward sweep writes every stage
// state and its trace records (and f) into one store and the reverse sweep
// only applies the four residual adjoints per step; otherwise checkpoints
// every `interval` steps (default ceil(sqrt(N))), each segment replayed
// once into a store of its steps' stages (or, when even that does not fit,
// its step-start states recomputed and every step's stages rebuilt).  All
// run the same kernels on the same inputs, so their gradients agree bit for
// bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_ctx.hpp"
#include "dgx_launch.hpp"

using namespace dgx;

namespace
{
struct DevBuf
{
  double* p = nullptr;
  ~DevBuf()
  {
    if (p)
      (void)hipFree(p);
  }
};

dgx_status afail(dgx_ctx* c, dgx_status st, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c->err, sizeof(c->err), fmt, ap);
  va_end(ap);
  return st;
}

#define AHIP(c, call)                                                                      \
  do                                                                                       \
  {                                                                                        \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return afail((c), DGX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
  } while (0)

// Workspace of the residual adjoint
struct AdjScratch
{
  DevBuf pvol, aown, anbr;
};

hipError_t alloc(DevBuf& b, size_t ndouble)
{
  return hipMalloc(&b.p, ndouble * sizeof(double));
}

hipError_t alloc_scratch(const dgx_ctx* c, AdjScratch& s)
{
  size_t tr = trace_doubles(c->p, c->nx, c->ny, c->nz);
  hipError_t e = alloc(s.pvol, c->nstate);
  if (e == hipSuccess)
    e = alloc(s.aown, tr);
  if (e == hipSuccess)
    e = alloc(s.anbr, tr);
  return e;
}

// one stage kernel launch with explicit physics: Vout = Vbase + cf f(Vin)
// (Vout may be null), fout = f(Vin) (may be null)
hipError_t stage(dgx_ctx* c, const Phys& ph, const double* Vin, const double* Tin,
                 const double* Vbase, double* Vout, double* Tout, double* fout, double cf)
{
  StageArgs sa;
  sa.Vin = Vin;
  sa.Tin = Tin;
  sa.Vbase = Vbase;
  sa.acc = nullptr;
  sa.Vout = Vout;
  sa.Tout = Tout;
  sa.fout = fout;
  sa.ca = 1.0;
  sa.cb = 0.0;
  sa.cf = cf;
  sa.accmode = 0;
  sa.trace32 = 0;
  sa.ctype   = 0;
  return rhs_dispatch(c->p, c->q, mesh_args(c), ph, sa, c->lam, c->status, c->stage_waves,
                      c->stream, c->stage_tile);
}

// one RK4 step of the context's state (the dgx_step path)
hipError_t step_ctx(dgx_ctx* c, double dt)
{
  hipError_t e = ensure_traces(c);
  if (e != hipSuccess)
    return e;
  c->tu_valid = false;
  StageArgs sa[4];
  int ns = stage_plan(c, DGX_RK4, dt, sa);
  for (int k = 0; k < ns && e == hipSuccess; ++k)
    e = rhs_dispatch(c->p, c->q, mesh_args(c), c->phys, sa[k], c->lam, c->status,
                     c->stage_waves, c->stream, c->stage_tile);
  if (e == hipSuccess)
    c->tu_valid = true;
  return e;
}

hipError_t copy_state(dgx_ctx* c, double* dst, const double* src)
{
  return hipMemcpyAsync(dst, src, c->nstate * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream);
}

}  // namespace

// ---- allocation-free form of the residual adjoint -------------------------
// dgx_residual_vjp_device_noalloc takes its workspace (the basis-converted
// cotangent, the volume-term partial adjoint, the two adjoint trace-record
// sets and the Gauss-nodal result: 3 states + 2 record sets) from this
// file-scope device array instead of hipMalloc.  Each device holds its own
// copy for the life of the process; every call on a device shares it.
constexpr size_t kScratchDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace
{
std::mutex g_scratch_mutex;
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on device dev, which must be current; looked
// up once per device (the host-side name of the array is not its address).
static hipError_t scratch_address(int dev, double** out)
{
  if (dev < 0 || dev >= kMaxDevices)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev])
  {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

extern "C" {

dgx_status dgx_kinetic_energy(dgx_ctx* c, double* J)
{
  if (!c || !J)
    return DGX_ERR_INVALID;
  if (general_order(c->p, c->q))
    return afail(c, DGX_ERR_UNSUPPORTED, "the adjoint covers 1 <= p, q <= 3");
  DevBuf g;
  AHIP(c, alloc(g, c->nstate));
  AHIP(c, kinetic_energy_dispatch(c->p, c->q, mesh_args(c), c->U, g.p, c->partial, c->stream));
  AHIP(c, sum_dispatch(c->partial, c->npartial, c->sum, c->stream));
  AHIP(c, hipMemcpyAsync(J, c->sum, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  AHIP(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_residual_vjp(dgx_ctx* c, const double* w_host, double* g_host)
{
  if (!c || !w_host || !g_host)
    return DGX_ERR_INVALID;
  if (general_order(c->p, c->q))
    return afail(c, DGX_ERR_UNSUPPORTED, "the adjoint covers 1 <= p, q <= 3");
  if (c->prec != DGX_PREC_FP64)
    return afail(c, DGX_ERR_UNSUPPORTED, "the residual adjoint needs DGX_PREC_FP64");
  AdjScratch s;
  DevBuf a, b;
  AHIP(c, alloc_scratch(c, s));
  AHIP(c, alloc(a, c->nstate));
  AHIP(c, alloc(b, c->nstate));
  AHIP(c, ensure_traces(c));
  // f = B3^-1 f_g(B3 U): (df/dU)^T w = B3^T (df_g/dV)^T B3^-T w
  AHIP(c, hipMemcpyAsync(a.p, w_host, c->nstate * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  AHIP(c, basis_dispatch(c->p, a.p, b.p, c->nelem, 3, c->stream));
  AHIP(c, adjoint_dispatch(c->p, c->q, mesh_args(c), c->phys, c->U, c->TU, c->lam, b.p,
                           s.pvol.p, s.aown.p, s.anbr.p, a.p, c->stream));
  AHIP(c, basis_dispatch(c->p, a.p, b.p, c->nelem, 2, c->stream));
  AHIP(c, hipMemcpyAsync(g_host, b.p, c->nstate * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  AHIP(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_residual_vjp_device_noalloc(dgx_ctx* c, const double* w_dev, double* g_dev)
{
  if (!c || !w_dev || !g_dev)
    return DGX_ERR_INVALID;
  if (general_order(c->p, c->q))
    return afail(c, DGX_ERR_UNSUPPORTED, "the adjoint covers 1 <= p, q <= 3");
  if (c->prec != DGX_PREC_FP64)
    return afail(c, DGX_ERR_UNSUPPORTED, "the residual adjoint needs DGX_PREC_FP64");
  const size_t n = c->nstate;
  const size_t tr = trace_doubles(c->p, c->nx, c->ny, c->nz);
  const size_t need = 3 * n + 2 * tr;
  if (need > kScratchDoubles)
    return afail(c, DGX_ERR_INVALID,
                 "mesh too large for the allocation-free residual adjoint: it needs %zu bytes "
                 "of workspace, the fixed scratch holds %zu (use dgx_residual_vjp)",
                 need * sizeof(double), kScratchDoubles * sizeof(double));
  int prev = -1;
  AHIP(c, hipGetDevice(&prev));
  if (prev != c->device)
    AHIP(c, hipSetDevice(c->device));
  double* base = nullptr;
  hipError_t e = scratch_address(c->device, &base);
  double* b    = base;          // B3^-T w
  double* pvol = base + n;      // volume-term partial adjoint
  double* a    = base + 2 * n;  // (df_g/dV)^T B3^-T w
  double* aown = base + 3 * n;  // adjoint trace records, own and neighbour side
  double* anbr = aown + tr;
  // f = B3^-1 f_g(B3 U): (df/dU)^T w = B3^T (df_g/dV)^T B3^-T w
  if (e == hipSuccess)
    e = ensure_traces(c);
  if (e == hipSuccess)
    e = basis_dispatch(c->p, w_dev, b, c->nelem, 3, c->stream);
  if (e == hipSuccess)
    e = adjoint_dispatch(c->p, c->q, mesh_args(c), c->phys, c->U, c->TU, c->lam, b, pvol, aown,
                         anbr, a, c->stream);
  if (e == hipSuccess)
    e = basis_dispatch(c->p, a, g_dev, c->nelem, 2, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  if (prev != c->device)
    (void)hipSetDevice(prev);
  if (e != hipSuccess)
    return afail(c, DGX_ERR_HIP, "residual adjoint (allocation-free): %s", hipGetErrorString(e));
  return DGX_OK;
}

dgx_status dgx_adjoint_rk4(dgx_ctx* c, double dt, uint64_t nsteps, uint64_t interval, double* J,
                           double* dJdU0_host, double* dJdmu)
{
  if (!c || nsteps == 0 || !(dt > 0.0))
    return DGX_ERR_INVALID;
  if (general_order(c->p, c->q))
    return afail(c, DGX_ERR_UNSUPPORTED, "the adjoint covers 1 <= p, q <= 3");
  if (c->prec != DGX_PREC_FP64)
    return afail(c, DGX_ERR_UNSUPPORTED, "the discrete adjoint needs DGX_PREC_FP64");
  if (dJdmu && !(c->phys.mu > 0.0))
    return afail(c, DGX_ERR_INVALID, "dJ/dmu needs mu > 0");
  const size_t n = c->nstate;
  const size_t ntr = trace_doubles(c->p, c->nx, c->ny, c->nz);

  DevBuf lam, lk, acur, asum, f0, dpart, dsum, final_state;
  AHIP(c, alloc(lam, n));
  AHIP(c, alloc(lk, n));
  AHIP(c, alloc(acur, n));
  AHIP(c, alloc(asum, n));
  AHIP(c, alloc(final_state, n));
  if (dJdmu)
    AHIP(c, alloc(f0, n));
  const int ndot = 1024;
  AHIP(c, alloc(dpart, ndot));
  // one dot product per reverse stage for dJ/dmu, read back once at the end
  AHIP(c, alloc(dsum, dJdmu ? 4 * nsteps : 1));
  AdjScratch scr;
  AHIP(c, alloc_scratch(c, scr));

  // Schedule.  interval == 0: keep every stage state with its trace records
  // (and f for dJ/dmu) from the forward sweep when that fits in three
  // quarters of the free device memory, so the reverse sweep recomputes
  // nothing; otherwise (or for interval >= 1) checkpoint every K steps, and
  // per segment replay its K steps once into a store of K steps' stage
  // states, which the reverse sweep then reads as it reads the full store
  // (when even that does not fit, rebuild each step's stages from the
  // segment's step-start states).  The store stays with the context for
  // later calls (an optimisation loop calls this repeatedly; freeing tens of
  // GB has taken seconds) until dgx_adjoint_release or dgx_destroy, and is
  // replaced only when too small.
  const size_t per_step = 4 * (n + ntr) + (dJdmu ? 4 * n : 0);
  const uint64_t K = interval ? interval : (uint64_t)std::ceil(std::sqrt((double)nsteps));
  auto acquire_store = [&](uint64_t steps) -> double* {
    const size_t need = (size_t)steps * per_step * sizeof(double);
    if (c->adj_store_limit && need > c->adj_store_limit)
      return nullptr;
    if (c->adj_store && c->adj_store_bytes >= need)
      return c->adj_store;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess ||
        (double)need > 0.75 * ((double)fr + (double)c->adj_store_bytes))
      return nullptr;
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return nullptr;
    if (c->adj_store)
      (void)hipFree(c->adj_store);
    c->adj_store       = nullptr;
    c->adj_store_bytes = 0;
    double* p          = nullptr;
    if (hipMalloc(&p, need) != hipSuccess)
    {
      (void)hipGetLastError();  // fall back to the next schedule
      return nullptr;
    }
    c->adj_store       = p;
    c->adj_store_bytes = need;
    return p;
  };
  double* store = (interval == 0) ? acquire_store(nsteps) : nullptr;
  const bool store_all = store != nullptr;
  uint64_t slots = store_all ? nsteps : 0;  // steps the store holds
  if (!store_all)
  {
    if (c->adj_store && c->adj_store_bytes < (size_t)K * per_step * sizeof(double))
    {
      // too small for a segment: it would only crowd the checkpoint
      // buffers out of memory
      AHIP(c, hipStreamSynchronize(c->stream));
      AHIP(c, hipFree(c->adj_store));
      c->adj_store       = nullptr;
      c->adj_store_bytes = 0;
    }
    // checkpoints are allocated first, the segment store from what is left
  }
  auto Sp = [&](uint64_t s, int i) { return store + (s * 4 + i) * (n + ntr); };
  auto Tp = [&](uint64_t s, int i) { return Sp(s, i) + n; };
  auto Kp = [&](uint64_t s, int i) { return store + slots * 4 * (n + ntr) + (s * 4 + i) * n; };

  // reverse of one step: stage states Si with records Ti and (dJ/dmu only)
  // their right-hand sides Ki = f(S_i); lam, the adjoint of the step's end
  // state, becomes the adjoint of its start state
  size_t ndmu = 0;
  // weights of stage k in the adjoint of V': l_k = w_l(k) l + w_a(k) a_(k+1)
  auto wl_of = [&](int k) { return (k == 3 || k == 0) ? dt / 6.0 : dt / 3.0; };
  auto wa_of = [&](int k) { return (k == 2) ? dt : 0.5 * dt; };
  auto reverse_step = [&](double* const* Si, double* const* Ti, double* const* Ki) -> dgx_status {
    // lk of stage 3 = w_l l
    AHIP(c, adj_tail_dispatch(lam.p, nullptr, nullptr, nullptr, lk.p, wl_of(3), 0.0, n,
                              c->stream));
    for (int i = 3; i >= 0; --i)
    {
      if (dJdmu)
      {
        // df/dmu at S_i: (k_(i+1) - f(S_i; 0)) / mu, f(S_i; 0) the inviscid
        // right-hand side from S_i and its records
        AHIP(c, inviscid_rhs_dispatch(c->p, c->q, mesh_args(c), c->phys, Si[i], Ti[i], f0.p,
                                      c->lam, c->status, c->stream));
        AHIP(c, dot_dispatch(lk.p, Ki[i], f0.p, 1.0 / c->phys.mu, n, dpart.p, ndot,
                             dsum.p + ndmu, c->stream));
        ++ndmu;
      }
      AHIP(c, adjoint_dispatch(c->p, c->q, mesh_args(c), c->phys, Si[i], Ti[i], c->lam, lk.p,
                               scr.pvol.p, scr.aown.p, scr.anbr.p, acur.p, c->stream));
      // asum = l + a_3 + ... + a_i; lk of stage i-1 = w_l l + w_a a_i
      AHIP(c, adj_tail_dispatch(lam.p, acur.p, (i == 3) ? lam.p : asum.p, asum.p,
                                (i > 0) ? lk.p : nullptr, (i > 0) ? wl_of(i - 1) : 0.0,
                                (i > 0) ? wa_of(i - 1) : 0.0, n, c->stream));
    }
    std::swap(lam.p, asum.p);  // the adjoint of the step's start state becomes l
    return DGX_OK;
  };

  const double cstage[4] = {0.5 * dt, 0.5 * dt, dt, dt};
  const uint64_t nchk = store_all ? 0 : (nsteps + K - 1) / K;
  std::vector<DevBuf> chk(nchk);
  for (auto& b : chk)
    AHIP(c, alloc(b, n));
  bool seg_store = false;
  if (!store_all)
  {
    store     = acquire_store(K);
    seg_store = store != nullptr;
    slots     = seg_store ? K : 0;
  }

  // ns steps from the state V0, every stage state and its records (and f
  // for dJ/dmu) written into store slots 0..ns-1, the end state to Vend:
  // the stage arguments of stage_plan (RK4) with other output buffers
  auto forward_store = [&](const double* V0, uint64_t ns, double* Vend) -> dgx_status {
    AHIP(c, copy_state(c, Sp(0, 0), V0));
    AHIP(c, trace_dispatch(c->p, c->q, mesh_args(c), c->phys, Sp(0, 0), Tp(0, 0), c->status,
                           c->stream));
    const int accmode[4] = {1, 2, 2, 3};
    for (uint64_t s = 0; s < ns; ++s)
      for (int i = 0; i < 4; ++i)
      {
        StageArgs sa;
        sa.Vin   = Sp(s, i);
        sa.Tin   = Tp(s, i);
        sa.Vbase = Sp(s, 0);
        sa.acc   = c->acc;
        if (i < 3)
        {
          sa.Vout = Sp(s, i + 1);
          sa.Tout = Tp(s, i + 1);
        }
        else
        {
          sa.Vout = (s + 1 < ns) ? Sp(s + 1, 0) : Vend;
          sa.Tout = (s + 1 < ns) ? Tp(s + 1, 0) : nullptr;
        }
        sa.fout    = dJdmu ? Kp(s, i) : nullptr;
        sa.ca      = (i == 0) ? 0.0 : 1.0;  // stage 1: its base is its input (in LDS)
        sa.cb      = (i == 0) ? 1.0 : 0.0;
        sa.cf      = cstage[i];
        sa.accmode = accmode[i];
        sa.trace32 = 0;
        sa.ctype   = 0;
        AHIP(c, rhs_dispatch(c->p, c->q, mesh_args(c), c->phys, sa, c->lam, c->status,
                             c->stage_waves, c->stream, c->stage_tile));
      }
    return DGX_OK;
  };
  // reverse of store slots ns-1..0
  auto reverse_store = [&](uint64_t ns) -> dgx_status {
    for (uint64_t s = ns; s-- > 0;)
    {
      double* const Si[4] = {Sp(s, 0), Sp(s, 1), Sp(s, 2), Sp(s, 3)};
      double* const Ti[4] = {Tp(s, 0), Tp(s, 1), Tp(s, 2), Tp(s, 3)};
      double* const Ki[4] = {dJdmu ? Kp(s, 0) : nullptr, dJdmu ? Kp(s, 1) : nullptr,
                             dJdmu ? Kp(s, 2) : nullptr, dJdmu ? Kp(s, 3) : nullptr};
      dgx_status st = reverse_step(Si, Ti, Ki);
      if (st != DGX_OK)
        return st;
    }
    return DGX_OK;
  };

  // forward sweep
  if (store_all)
  {
    dgx_status st = forward_store(c->U, nsteps, final_state.p);
    if (st != DGX_OK)
      return st;
    AHIP(c, copy_state(c, c->U, final_state.p));
    c->tu_valid = false;
  }
  else
  {
    // the last segment goes straight into the segment store (the reverse
    // sweep starts with it and need not replay it)
    const uint64_t s_last = (nchk - 1) * K;
    for (uint64_t s = 0; s < nsteps; ++s)
    {
      if (s % K == 0)
        AHIP(c, copy_state(c, chk[s / K].p, c->U));
      if (seg_store && s == s_last)
      {
        dgx_status st = forward_store(c->U, nsteps - s_last, final_state.p);
        if (st != DGX_OK)
          return st;
        AHIP(c, copy_state(c, c->U, final_state.p));
        c->tu_valid = false;
        break;
      }
      AHIP(c, step_ctx(c, dt));
    }
    if (!seg_store)
      AHIP(c, copy_state(c, final_state.p, c->U));
  }

  // objective and its gradient at the final state
  AHIP(c, kinetic_energy_dispatch(c->p, c->q, mesh_args(c), final_state.p, lam.p, c->partial,
                                  c->stream));
  AHIP(c, sum_dispatch(c->partial, c->npartial, c->sum, c->stream));
  double Jv = 0.0;
  AHIP(c, hipMemcpyAsync(&Jv, c->sum, sizeof(double), hipMemcpyDeviceToHost, c->stream));

  // reverse sweep
  if (store_all)
  {
    dgx_status st = reverse_store(nsteps);
    if (st != DGX_OK)
      return st;
  }
  else if (seg_store)
  {
    // per segment: replay its steps once into the store (the last one is
    // there from the forward sweep), reverse them
    for (uint64_t sg = nchk; sg-- > 0;)
    {
      const uint64_t s0 = sg * K, s1 = (s0 + K < nsteps) ? s0 + K : nsteps;
      // the segment's end state is not needed again: acur is free here
      dgx_status st = (sg + 1 < nchk) ? forward_store(chk[sg].p, s1 - s0, acur.p) : DGX_OK;
      if (st == DGX_OK)
        st = reverse_store(s1 - s0);
      if (st != DGX_OK)
        return st;
    }
  }
  else
  {
    // per segment: recompute its step-start states, then per step rebuild
    // the stage states S0..S3 with their records and rhs k1..k4, reverse it
    std::vector<DevBuf> seg(K);
    for (auto& b : seg)
      AHIP(c, alloc(b, n));
    DevBuf S[4], T[4], kf[4];
    for (int i = 0; i < 4; ++i)
    {
      AHIP(c, alloc(S[i], n));
      AHIP(c, alloc(T[i], ntr));
      AHIP(c, alloc(kf[i], n));
    }
    double* const Si[4] = {S[0].p, S[1].p, S[2].p, S[3].p};
    double* const Ti[4] = {T[0].p, T[1].p, T[2].p, T[3].p};
    double* const Ki[4] = {kf[0].p, kf[1].p, kf[2].p, kf[3].p};
    for (uint64_t sg = nchk; sg-- > 0;)
    {
      const uint64_t s0 = sg * K, s1 = (s0 + K < nsteps) ? s0 + K : nsteps;
      AHIP(c, copy_state(c, c->U, chk[sg].p));
      c->tu_valid = false;
      for (uint64_t s = s0; s < s1; ++s)
      {
        AHIP(c, copy_state(c, seg[s - s0].p, c->U));
        if (s + 1 < s1)
          AHIP(c, step_ctx(c, dt));
      }
      for (uint64_t s = s1; s-- > s0;)
      {
        AHIP(c, copy_state(c, S[0].p, seg[s - s0].p));
        AHIP(c, trace_dispatch(c->p, c->q, mesh_args(c), c->phys, S[0].p, T[0].p, c->status,
                               c->stream));
        for (int i = 0; i < 3; ++i)
          AHIP(c, stage(c, c->phys, S[i].p, T[i].p, S[0].p, S[i + 1].p, T[i + 1].p, kf[i].p,
                        cstage[i]));
        if (dJdmu)  // k4 only enters df/dmu
          AHIP(c, stage(c, c->phys, S[3].p, T[3].p, nullptr, nullptr, nullptr, kf[3].p, 0.0));
        dgx_status st = reverse_step(Si, Ti, Ki);
        if (st != DGX_OK)
          return st;
      }
    }
  }
  double dmu_total = 0.0;
  if (ndmu)
  {
    std::vector<double> parts(ndmu);
    AHIP(c, hipMemcpyAsync(parts.data(), dsum.p, ndmu * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    AHIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = ndmu; i-- > 0;)
      dmu_total += parts[i];
  }
  // gradient with respect to the initial state in the reference layout
  if (dJdU0_host)
  {
    AHIP(c, basis_dispatch(c->p, lam.p, asum.p, c->nelem, 2, c->stream));
    AHIP(c, hipMemcpyAsync(dJdU0_host, asum.p, n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  }
  // the context ends at the final state, as after dgx_step
  AHIP(c, copy_state(c, c->U, final_state.p));
  c->tu_valid = false;
  AHIP(c, hipStreamSynchronize(c->stream));
  int h = 0;
  AHIP(c, hipMemcpy(&h, c->status, sizeof(int), hipMemcpyDeviceToHost));
  if (h)
  {
    AHIP(c, hipMemset(c->status, 0, sizeof(int)));
    return afail(c, DGX_ERR_INVALID, "negative inflow boundary Mach number");
  }
  if (J)
    *J = Jv;
  if (dJdmu)
    *dJdmu = dmu_total;
  return DGX_OK;
}

dgx_status dgx_adjoint_set_store_limit(dgx_ctx* c, uint64_t bytes)
{
  if (!c)
    return DGX_ERR_INVALID;
  c->adj_store_limit = (size_t)bytes;
  if (bytes && c->adj_store && c->adj_store_bytes > bytes)
    return dgx_adjoint_release(c);
  return DGX_OK;
}

dgx_status dgx_adjoint_release(dgx_ctx* c)
{
  if (!c)
    return DGX_ERR_INVALID;
  if (c->adj_store)
  {
    AHIP(c, hipStreamSynchronize(c->stream));
    AHIP(c, hipFree(c->adj_store));
    c->adj_store = nullptr;
    c->adj_store_bytes = 0;
  }
  return DGX_OK;
}

}  // extern "C"
