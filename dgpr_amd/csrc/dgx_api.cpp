// dgx_api.cpp -- C ABI of libdgx: device context, state transfer, residual
// evaluation and RK stage sequencing on one HIP stream.
//
// Stage algebra follows the CPU drivers RK4_distributed and
// tvdRK3_distributed (time_distributed.cpp:46-102), including classic RK4
// (stage 3 uses U + dt f3) and the RMS residual norm of the first stage
// (time_distributed.cpp:23-44).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_ctx.hpp"
#include "dgx_launch.hpp"

using namespace dgx;


static thread_local char g_create_err[512] = "";
static std::mutex g_tab_mutex;
static std::vector<int> g_tab_devices;

static dgx_status fail(dgx_ctx* c, dgx_status st, const char* fmt, ...)
{
  char* buf = c ? c->err : g_create_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, 512, fmt, ap);
  va_end(ap);
  return st;
}

#define HIPCHK(ctx, call)                                                                  \
  do                                                                                       \
  {                                                                                        \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((ctx), DGX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));      \
  } while (0)

// Occupancy target per (p, q): 3 waves/SIMD where the stage kernel fits the
// 168-register budget without spilling (p=3, q=1), 2 elsewhere.
int dgx::default_stage_waves(int p, int q)
{
  return (p == 3 && q == 1) ? 3 : 2;
}

static void derive_phys(const dgx_params& p, Phys& ph)
{
  // total conditions of the reference state (solve.cpp:167-175)
  double gm1 = p.gamma - 1.0;
  double u = p.Ufr[1] / p.Ufr[0], v = p.Ufr[2] / p.Ufr[0], w = p.Ufr[3] / p.Ufr[0];
  double M = sqrt(u * u + v * v + w * w) / p.c;
  ph.gamma = p.gamma;
  ph.mu    = p.mu;
  ph.eta   = p.eta;
  ph.Rgas  = p.R;
  ph.P     = p.P;
  ph.Tt    = p.T * (1.0 + 0.5 * gm1 * M * M);
  ph.Pt    = p.P * pow(ph.Tt / p.T, p.gamma / gm1);
  for (int i = 0; i < 5; ++i)
  {
    ph.Ufr[i]    = p.Ufr[i];
    ph.source[i] = p.source[i];
  }
}

MeshArgs dgx::mesh_args(const dgx_ctx* c)
{
  MeshArgs m;
  m.nodes = c->nodes;
  m.nx = c->nx;
  m.ny = c->ny;
  m.nz = c->nz;
  m.nelem = c->nelem;
  for (int i = 0; i < 6; ++i)
    m.bounds[i] = c->bounds[i];
  m.affine = c->affine;
  m.lamf   = c->lamf;
  return m;
}

// True when q = 1 and every element's bilinear and trilinear node terms
// vanish to rounding (a parallelepiped, constant Jacobian); the stage
// kernels then use one Jacobian per element instead of the geometry basis.
// The twist terms are measured against the element's own size (its longest
// edge-vector component), plus the rounding of coordinates far from the
// origin (a few ulps of the largest coordinate), so a small element far from
// the origin is not classified affine by its position alone.
static bool mesh_is_affine(const dgx_mesh* m)
{
  if (m->q != 1)
    return false;
  const uint64_t ne = m->nx * m->ny * m->nz;
  for (uint64_t e = 0; e < ne; ++e)
  {
    const double* X = m->nodes + e * 24;
    double twist = 0.0, size = 0.0, mag = 0.0;
    for (int d = 0; d < 3; ++d)
    {
      const double x000 = X[d], x100 = X[3 + d], x010 = X[6 + d], x110 = X[9 + d];
      const double x001 = X[12 + d], x101 = X[15 + d], x011 = X[18 + d], x111 = X[21 + d];
      twist = std::fmax(twist, std::fabs((x110 - x100) - (x010 - x000)));
      twist = std::fmax(twist, std::fabs((x101 - x100) - (x001 - x000)));
      twist = std::fmax(twist, std::fabs((x011 - x010) - (x001 - x000)));
      twist = std::fmax(twist,
                        std::fabs(((x111 - x011) - (x101 - x001)) - ((x110 - x010) - (x100 - x000))));
      size = std::fmax(size, std::fmax(std::fabs(x100 - x000),
                                       std::fmax(std::fabs(x010 - x000), std::fabs(x001 - x000))));
      for (int c = 0; c < 8; ++c)
        mag = std::fmax(mag, std::fabs(X[3 * c + d]));
    }
    if (!(twist <= 1e-14 * size + 8.0 * 2.220446049250313e-16 * mag))
      return false;
  }
  return true;
}

extern "C" dgx_status dgx_mesh_is_affine(const dgx_mesh* mesh, int* affine)
{
  if (!mesh || !affine || !mesh->nodes || mesh->q < 1)
    return DGX_ERR_INVALID;
  *affine = mesh_is_affine(mesh) ? 1 : 0;
  return DGX_OK;
}

static dgx_status ensure_tables(int device)
{
  std::lock_guard<std::mutex> lk(g_tab_mutex);
  for (int d : g_tab_devices)
    if (d == device)
      return DGX_OK;
  static PTab pt[3];
  static QTab qt[9];
  static GenTab gt[72];
  static bool built = false;
  if (!built)
  {
    host::build_tables(pt, qt);
    host::build_gen_tables(gt);
    built = true;
  }
  hipError_t e = upload_tables(pt, qt);
  if (e == hipSuccess)
    e = upload_tables_ensemble(pt, qt);
  if (e == hipSuccess)
    e = upload_gen_tables(gt);
  if (e != hipSuccess)
    return fail(nullptr, DGX_ERR_HIP, "table upload failed: %s", hipGetErrorString(e));
  g_tab_devices.push_back(device);
  return DGX_OK;
}

extern "C" {

const char* dgx_last_error(const dgx_ctx* ctx)
{
  return ctx ? ctx->err : g_create_err;
}

dgx_status dgx_make_reference_state(double gamma, double R, double mu, double c, double rho,
                                    double u, double v, double w, double eta,
                                    const double* source, dgx_params* out)
{
  if (!out || rho <= 0.0 || gamma <= 1.0 || R <= 0.0)
    return DGX_ERR_INVALID;
  double P = c * c * (rho / gamma);
  out->gamma = gamma;
  out->R = R;
  out->mu = mu;
  out->c = c;
  out->T = P / (rho * R);
  out->P = P;
  out->eta = eta;
  out->Ufr[0] = rho;
  out->Ufr[1] = rho * u;
  out->Ufr[2] = rho * v;
  out->Ufr[3] = rho * w;
  out->Ufr[4] = (P / (gamma - 1.0)) + 0.5 * rho * (u * u + v * v + w * w);
  for (int i = 0; i < 5; ++i)
    out->source[i] = source ? source[i] : 0.0;
  return DGX_OK;
}

dgx_status dgx_create(const dgx_mesh* mesh, const dgx_params* params, dgx_ctx** out)
{
  if (!mesh || !params || !out || !mesh->nodes)
    return fail(nullptr, DGX_ERR_INVALID, "null argument");
  *out = nullptr;
  // the reference's range (MAX_ORDER, constants.cpp:21; geometry order >= 1)
  if (mesh->p > 8)
    return fail(nullptr, DGX_ERR_UNSUPPORTED, "solution order p=%llu outside 0..8",
                (unsigned long long)mesh->p);
  if (mesh->q < 1 || mesh->q > 8)
    return fail(nullptr, DGX_ERR_UNSUPPORTED, "geometry order q=%llu outside 1..8",
                (unsigned long long)mesh->q);
  if (mesh->nx == 0 || mesh->ny == 0 || mesh->nz == 0)
    return fail(nullptr, DGX_ERR_INVALID, "empty mesh");
  unsigned long long ne = (unsigned long long)mesh->nx * mesh->ny * mesh->nz;
  if (ne > (1ull << 30))
    return fail(nullptr, DGX_ERR_UNSUPPORTED, "mesh of %llu elements too large", ne);
  for (int d = 0; d < 3; ++d)
  {
    if ((mesh->bounds[2 * d] == DGX_BT_PERIODIC) != (mesh->bounds[2 * d + 1] == DGX_BT_PERIODIC))
      return fail(nullptr, DGX_ERR_INVALID, "periodic boundaries are not matching");
    for (int s = 0; s < 2; ++s)
    {
      int64_t b = mesh->bounds[2 * d + s];
      if (b > -2 || b < -7)
        return fail(nullptr, DGX_ERR_INVALID, "invalid boundary code %lld", (long long)b);
    }
  }
  if (params->Ufr[0] <= 0.0 || params->gamma <= 1.0)
    return fail(nullptr, DGX_ERR_INVALID, "invalid reference state");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, DGX_ERR_NODEVICE, "no HIP device visible");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return fail(nullptr, DGX_ERR_HIP, "hipGetDevice failed");
  dgx_status st = ensure_tables(dev);
  if (st != DGX_OK)
    return st;

  dgx_ctx* c = new dgx_ctx();
  memset(c, 0, sizeof(*c));
  c->device = dev;
  c->p = (int)mesh->p;
  c->q = (int)mesh->q;
  c->affine = mesh_is_affine(mesh) ? 1 : 0;
  c->nx = (int)mesh->nx;
  c->ny = (int)mesh->ny;
  c->nz = (int)mesh->nz;
  c->nelem = (int)ne;
  c->nbf = (c->p + 1) * (c->p + 1) * (c->p + 1);
  for (int i = 0; i < 6; ++i)
    c->bounds[i] = (int)mesh->bounds[i];
  c->nstate = (size_t)c->nelem * 5 * c->nbf;
  c->params = *params;
  derive_phys(*params, c->phys);
  size_t nn = (size_t)c->nelem * (c->q + 1) * (c->q + 1) * (c->q + 1) * 3;
  c->npartial = resid_out_blocks(c->p, c->q, c->nelem);

  auto bail = [&](const char* what, hipError_t e) {
    snprintf(g_create_err, sizeof(g_create_err), "%s failed: %s", what, hipGetErrorString(e));
    dgx_destroy(c);
    return DGX_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail("hipStreamCreate", e);
  size_t sb = c->nstate * sizeof(double);
  double** bufs[5] = {&c->U, &c->Ua, &c->Ub, &c->acc, &c->fs};
  for (int i = 0; i < 5; ++i)
    if ((e = hipMalloc(bufs[i], sb)) != hipSuccess)
      return bail("hipMalloc(state)", e);
  if ((e = hipMalloc(&c->nodes, nn * sizeof(double))) != hipSuccess)
    return bail("hipMalloc(nodes)", e);
  if ((e = hipMalloc(&c->partial, ((size_t)c->npartial + sum_scratch(c->npartial)) *
                                       sizeof(double))) != hipSuccess)
    return bail("hipMalloc(partial)", e);
  if ((e = hipMalloc(&c->sum, sizeof(double))) != hipSuccess)
    return bail("hipMalloc(sum)", e);
  size_t nlam = (size_t)c->nelem * 6 * (c->p + 1) * (c->p + 1);
  if ((e = hipMalloc(&c->lam, nlam * sizeof(double))) != hipSuccess)
    return bail("hipMalloc(lam)", e);
  if ((e = hipMalloc(&c->status, sizeof(int))) != hipSuccess)
    return bail("hipMalloc(status)", e);
  size_t tb = trace_doubles(c->p, c->nx, c->ny, c->nz) * sizeof(double);
  double** tbufs[3] = {&c->TU, &c->TUa, &c->TUb};
  for (int i = 0; i < 3; ++i)
    if ((e = hipMalloc(tbufs[i], tb)) != hipSuccess)
      return bail("hipMalloc(traces)", e);
  c->tu_valid = false;
  // the tiled kernel at p = 3, q = 1 (measured 14 % faster than the one-row
  // kernel on the 64^3 cube at 3 waves per SIMD, 7 % on the curved channel
  // at 2, where 3 spills); the one-row kernel at the other orders of
  // 1..3; the general-order kernel (tile 2) outside them
  c->stage_tile  = general_order(c->p, c->q) ? 2 : ((c->p == 3 && c->q == 1) ? 1 : 0);
  c->stage_waves = c->stage_tile == 1 ? (c->affine ? 3 : 2) : default_stage_waves(c->p, c->q);
  c->prec = DGX_PREC_FP64;
  c->bytes = 5 * sb + 3 * tb +
             (nn + nlam + (c->affine ? (size_t)c->nelem * 6 : 0) + (size_t)c->npartial +
              sum_scratch(c->npartial)) * sizeof(double) + 16;
  if ((e = hipMemcpyAsync(c->nodes, mesh->nodes, nn * sizeof(double), hipMemcpyHostToDevice,
                          c->stream)) != hipSuccess)
    return bail("hipMemcpy(nodes)", e);
  if ((e = hipMemsetAsync(c->U, 0, sb, c->stream)) != hipSuccess)
    return bail("hipMemset(U)", e);
  if ((e = lift_coef_dispatch(c->p, c->q, mesh_args(c), c->lam, c->stream)) != hipSuccess)
    return bail("lifting coefficients", e);
  if (c->affine)
  {
    if ((e = hipMalloc(&c->lamf, (size_t)c->nelem * 6 * sizeof(double))) != hipSuccess)
      return bail("hipMalloc(lamf)", e);
    if ((e = lift_face_dispatch(c->p, c->lam, c->nelem, c->lamf, c->stream)) != hipSuccess)
      return bail("face lifting coefficients", e);
  }
  if ((e = hipMemsetAsync(c->status, 0, sizeof(int), c->stream)) != hipSuccess)
    return bail("hipMemset(status)", e);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return bail("hipStreamSynchronize", e);
  *out = c;
  return DGX_OK;
}

dgx_status dgx_destroy(dgx_ctx* c)
{
  if (!c)
    return DGX_OK;
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  double* bufs[14] = {c->U,   c->Ua,  c->Ub,  c->acc, c->fs,  c->nodes, c->partial,
                      c->sum, c->lam, c->lamf, c->TU, c->TUa, c->TUb, c->mean};
  for (double* b : bufs)
    if (b)
      (void)hipFree(b);
  if (c->adj_store)
    (void)hipFree(c->adj_store);
  if (c->status)
    (void)hipFree(c->status);
  if (c->stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
  return DGX_OK;
}

dgx_status dgx_set_params(dgx_ctx* c, const dgx_params* params)
{
  if (!c || !params)
    return DGX_ERR_INVALID;
  if (params->Ufr[0] <= 0.0 || params->gamma <= 1.0)
    return fail(c, DGX_ERR_INVALID, "invalid reference state");
  c->params = *params;
  derive_phys(*params, c->phys);
  c->tu_valid = false;  // the gradient tractions scale with mu
  return DGX_OK;
}

dgx_status dgx_get_sizes(const dgx_ctx* c, uint64_t out[4])
{
  if (!c || !out)
    return DGX_ERR_INVALID;
  out[0] = (uint64_t)c->nelem;
  out[1] = (uint64_t)c->nbf;
  out[2] = (uint64_t)c->nstate;
  out[3] = (uint64_t)c->bytes;
  return DGX_OK;
}

dgx_status dgx_get_stream(dgx_ctx* c, void** stream)
{
  if (!c || !stream)
    return DGX_ERR_INVALID;
  *stream = (void*)c->stream;
  return DGX_OK;
}

dgx_status dgx_sync(dgx_ctx* c)
{
  if (!c)
    return DGX_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_set_state_device(dgx_ctx* c, const double* U_dev)
{
  if (!c || !U_dev)
    return DGX_ERR_INVALID;
  HIPCHK(c, basis_dispatch(c->p, U_dev, c->U, c->nelem, 0, c->stream));
  c->tu_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_get_state_device(dgx_ctx* c, double* U_dev)
{
  if (!c || !U_dev)
    return DGX_ERR_INVALID;
  HIPCHK(c, basis_dispatch(c->p, c->U, U_dev, c->nelem, 1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_set_state(dgx_ctx* c, const double* U_host)
{
  if (!c || !U_host)
    return DGX_ERR_INVALID;
  HIPCHK(c, hipMemcpyAsync(c->Ua, U_host, c->nstate * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, basis_dispatch(c->p, c->Ua, c->U, c->nelem, 0, c->stream));
  c->tu_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_get_state(dgx_ctx* c, double* U_host)
{
  if (!c || !U_host)
    return DGX_ERR_INVALID;
  HIPCHK(c, basis_dispatch(c->p, c->U, c->Ua, c->nelem, 1, c->stream));
  HIPCHK(c, hipMemcpyAsync(U_host, c->Ua, c->nstate * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

// One stage: Vin with its trace records Tin; Tout (with Vout) receives the
// trace records of Vout.
static hipError_t rhs(dgx_ctx* c, const double* Vin, const double* Tin, const double* Vbase,
                      double* acc, double* Vout, double* Tout, double* fout, double ca, double cb,
                      double cf, int accmode)
{
  StageArgs sa;
  sa.Vin = Vin;
  sa.Vbase = Vbase;
  sa.acc = acc;
  sa.Vout = Vout;
  sa.fout = fout;
  sa.Tin = Tin;
  sa.Tout = Tout;
  sa.ca = ca;
  sa.cb = cb;
  sa.cf = cf;
  sa.accmode = accmode;
  sa.trace32 = c->prec != DGX_PREC_FP64;
  sa.ctype   = c->prec == DGX_PREC_MIXED ? 1 : (c->prec == DGX_PREC_MIXED_BF16 ? 2 : 0);
  return rhs_dispatch(c->p, c->q, mesh_args(c), c->phys, sa, c->lam, c->status, c->stage_waves,
                      c->stream, c->stage_tile);
}

}  // extern "C"

// Trace records of U, recomputed when U was set from outside.
hipError_t dgx::ensure_traces(dgx_ctx* c)
{
  if (c->tu_valid)
    return hipSuccess;
  hipError_t e = trace_dispatch(c->p, c->q, mesh_args(c), c->phys, c->U, c->TU, c->status,
                                c->stream, c->prec != DGX_PREC_FP64);
  if (e == hipSuccess)
    c->tu_valid = true;
  return e;
}

extern "C" {

static dgx_status check_status(dgx_ctx* c)
{
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h)
  {
    (void)hipMemsetAsync(c->status, 0, sizeof(int), c->stream);
    return fail(c, DGX_ERR_INVALID, "negative inflow boundary Mach number");
  }
  return DGX_OK;
}

dgx_status dgx_residual_device(dgx_ctx* c, const double* U_dev, double* R_dev, double* f_dev)
{
  if (!c || !U_dev || !f_dev)
    return DGX_ERR_INVALID;
  HIPCHK(c, basis_dispatch(c->p, U_dev, c->Ua, c->nelem, 0, c->stream));
  HIPCHK(c, trace_dispatch(c->p, c->q, mesh_args(c), c->phys, c->Ua, c->TUa, c->status,
                           c->stream, c->prec != DGX_PREC_FP64));
  HIPCHK(c, rhs(c, c->Ua, c->TUa, nullptr, nullptr, nullptr, nullptr, c->fs, 0, 0, 0, 0));
  HIPCHK(c, resid_out_dispatch(c->p, c->q, mesh_args(c), c->fs, f_dev, R_dev, nullptr,
                               c->stream));
  return check_status(c);
}

dgx_status dgx_residual(dgx_ctx* c, double* R_host, double* f_host)
{
  if (!c || !f_host)
    return DGX_ERR_INVALID;
  HIPCHK(c, ensure_traces(c));
  HIPCHK(c, rhs(c, c->U, c->TU, nullptr, nullptr, nullptr, nullptr, c->fs, 0, 0, 0, 0));
  HIPCHK(c, resid_out_dispatch(c->p, c->q, mesh_args(c), c->fs, c->Ua, R_host ? c->Ub : nullptr,
                               nullptr, c->stream));
  size_t sb = c->nstate * sizeof(double);
  HIPCHK(c, hipMemcpyAsync(f_host, c->Ua, sb, hipMemcpyDeviceToHost, c->stream));
  if (R_host)
    HIPCHK(c, hipMemcpyAsync(R_host, c->Ub, sb, hipMemcpyDeviceToHost, c->stream));
  return check_status(c);
}

}  // extern "C"

// Stage arguments of one step (time_distributed.cpp:46-102, classic RK4 with
// an accumulator, Shu-Osher TVD-RK3).  Every stage writes the trace records
// of its output state, so the next stage (and the next step) reads them
// instead of recomputing.
int dgx::stage_plan(const dgx_ctx* c, int scheme, double dt, StageArgs sa[4])
{
  const int t32 = c->prec != DGX_PREC_FP64;
  const int ct  = c->prec == DGX_PREC_MIXED ? 1 : (c->prec == DGX_PREC_MIXED_BF16 ? 2 : 0);
  auto set = [t32, ct](StageArgs& a, const double* Vin, const double* Tin, const double* Vbase,
                   double* acc, double* Vout, double* Tout, double ca, double cb, double cf,
                   int accmode) {
    a.Vin = Vin;
    a.Tin = Tin;
    a.Vbase = Vbase;
    a.acc = acc;
    a.Vout = Vout;
    a.Tout = Tout;
    a.fout = nullptr;
    a.ca = ca;
    a.cb = cb;
    a.cf = cf;
    a.accmode = accmode;
    a.trace32 = t32;
    a.ctype   = ct;
  };
  // the first stage's base is its own input: cb = 1 takes it from the LDS
  // copy the kernel already holds instead of reading U from HBM again
  if (scheme == DGX_RK4)
  {
    set(sa[0], c->U, c->TU, c->U, c->acc, c->Ua, c->TUa, 0.0, 1.0, 0.5 * dt, 1);  // U1, acc = f1
    set(sa[1], c->Ua, c->TUa, c->U, c->acc, c->Ub, c->TUb, 1.0, 0.0, 0.5 * dt, 2);  // U2, acc += 2 f2
    set(sa[2], c->Ub, c->TUb, c->U, c->acc, c->Ua, c->TUa, 1.0, 0.0, dt, 2);  // U3, acc += 2 f3
    set(sa[3], c->Ua, c->TUa, c->U, c->acc, c->U, c->TU, 1.0, 0.0, dt, 3);  // U += dt/6 (acc + f4)
    return 4;
  }
  set(sa[0], c->U, c->TU, c->U, nullptr, c->Ua, c->TUa, 0.0, 1.0, dt, 0);  // U1 = U + dt f
  set(sa[1], c->Ua, c->TUa, c->U, nullptr, c->Ub, c->TUb, 0.75, 0.25, 0.25 * dt, 0);
  set(sa[2], c->Ub, c->TUb, c->U, nullptr, c->U, c->TU, 1.0 / 3.0, 2.0 / 3.0, (2.0 / 3.0) * dt,
      0);
  return 3;
}

extern "C" {

// Launches one step; if want_norm, stores sum(R^2) of stage 1 in c->sum
// (the tiled kernel computes the per-block sums in its epilogue; otherwise
// they come from the accumulator, which holds f1 in RK4, or from f written by
// the first stage in TVD-RK3), or, with norm_dev, the step's RMS(R) in
// *norm_dev.
static hipError_t launch_step(dgx_ctx* c, int scheme, double dt, bool want_norm,
                              double* norm_dev = nullptr)
{
  hipError_t e;
  if ((e = ensure_traces(c)) != hipSuccess)
    return e;
  c->tu_valid = false;  // until the last stage has rewritten them
  StageArgs sa[4];
  int ns = stage_plan(c, scheme, dt, sa);
  const bool fused = want_norm && c->stage_tile == 1;
  for (int k = 0; k < ns; ++k)
  {
    if (k == 0 && fused)
      sa[0].normp = c->partial;
    else if (k == 0 && want_norm && scheme != DGX_RK4)
      sa[0].fout = c->fs;
    if ((e = rhs_dispatch(c->p, c->q, mesh_args(c), c->phys, sa[k], c->lam, c->status,
                          c->stage_waves, c->stream, c->stage_tile)) != hipSuccess)
      return e;
    if (k == 0 && want_norm)
    {
      const double* f1 = (scheme == DGX_RK4) ? c->acc : c->fs;
      if (!fused && (e = norm_partials_dispatch(c->p, c->q, mesh_args(c), f1, c->partial,
                                                c->stream)) != hipSuccess)
        return e;
      const int nb = fused ? tile_blocks(c->p, c->nelem) : norm_blocks(c->p, c->q, c->nelem);
      e = norm_dev ? norm_dispatch(c->partial, nb, 1.0 / (double)c->nstate, norm_dev, c->stream)
                   : sum_dispatch(c->partial, nb, c->sum, c->stream);
      if (e != hipSuccess)
        return e;
    }
  }
  c->tu_valid = true;
  return hipSuccess;
}

dgx_status dgx_device_copy(dgx_ctx* c, void* dst, const void* src, uint64_t bytes)
{
  if (!c || !dst || !src || (bytes % 16) != 0)
    return DGX_ERR_INVALID;
  HIPCHK(c, copy_dispatch(src, dst, (size_t)bytes, c->stream));
  return DGX_OK;
}

dgx_status dgx_step_async(dgx_ctx* c, int scheme, double dt, uint64_t nsteps)
{
  if (!c || (scheme != DGX_RK4 && scheme != DGX_TVDRK3))
    return DGX_ERR_INVALID;
  for (uint64_t s = 0; s < nsteps; ++s)
    HIPCHK(c, launch_step(c, scheme, dt, false));
  return DGX_OK;
}

dgx_status dgx_step_norms_async(dgx_ctx* c, int scheme, double dt, uint64_t nsteps,
                                double* norms_dev)
{
  if (!c || !norms_dev || (scheme != DGX_RK4 && scheme != DGX_TVDRK3))
    return DGX_ERR_INVALID;
  for (uint64_t s = 0; s < nsteps; ++s)
    HIPCHK(c, launch_step(c, scheme, dt, true, norms_dev + s));
  return DGX_OK;
}

dgx_status dgx_step(dgx_ctx* c, int scheme, double dt, uint64_t nsteps, double* norms_host)
{
  if (!c || (scheme != DGX_RK4 && scheme != DGX_TVDRK3))
    return DGX_ERR_INVALID;
  double n = (double)c->nstate;
  for (uint64_t s = 0; s < nsteps; ++s)
  {
    HIPCHK(c, launch_step(c, scheme, dt, norms_host != nullptr));
    if (norms_host)
    {
      double sq = 0.0;
      HIPCHK(c, hipMemcpyAsync(&sq, c->sum, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      double nrm = sqrt((1.0 / n) * sq);
      norms_host[s] = nrm;
      if (std::isnan(nrm))
        return fail(c, DGX_ERR_NAN, "looks like things went unstable! (step %llu)",
                    (unsigned long long)s);
    }
  }
  return check_status(c);
}

}  // extern "C"

dgx_status dgx_set_precision(dgx_ctx* c, int precision)
{
  if (!c || (precision != DGX_PREC_FP64 && precision != DGX_PREC_MIXED &&
             precision != DGX_PREC_MIXED_BF16 && precision != DGX_PREC_MIXED_RECORDS))
    return DGX_ERR_INVALID;
  if (precision != c->prec)
  {
    c->prec = precision;
    c->tu_valid = false;
  }
  return DGX_OK;
}

dgx_status dgx_get_precision(const dgx_ctx* c, int* precision)
{
  if (!c || !precision)
    return DGX_ERR_INVALID;
  *precision = c->prec;
  return DGX_OK;
}

dgx_status dgx_set_tuning(dgx_ctx* c, int stage_waves)
{
  if (!c || stage_waves < 1 || stage_waves > 3 || !tile_valid(c->p, c->q, c->stage_tile, stage_waves))
    return DGX_ERR_INVALID;
  c->stage_waves = stage_waves;
  return DGX_OK;
}

dgx_status dgx_get_tuning(const dgx_ctx* c, int* stage_waves)
{
  if (!c || !stage_waves)
    return DGX_ERR_INVALID;
  *stage_waves = c->stage_waves;
  return DGX_OK;
}

dgx_status dgx_set_tile(dgx_ctx* c, int tile)
{
  if (!c || tile < 0 || tile > 2)
    return DGX_ERR_INVALID;
  if (!tile_valid(c->p, c->q, tile, 2))
    return fail(c, DGX_ERR_UNSUPPORTED, "no stage kernel of tile %d for p = %d, q = %d", tile,
                c->p, c->q);
  c->stage_tile = tile;
  if (!tile_valid(c->p, c->q, tile, c->stage_waves))
    c->stage_waves = 2;  // the tiled kernels are compiled for 2 and 3 waves per SIMD
  return DGX_OK;
}

dgx_status dgx_get_tile(const dgx_ctx* c, int* tile)
{
  if (!c || !tile)
    return DGX_ERR_INVALID;
  *tile = c->stage_tile;
  return DGX_OK;
}

// Times the stage kernel at each occupancy target of the current tile on
// the context's own mesh: U (with its traces) -> Ub, TUb, which are stage
// scratch between steps.  Targets the tile does not compile read 0.
dgx_status dgx_autotune(dgx_ctx* c, int reps, int* chosen, double us_out[3])
{
  if (!c || reps < 1)
    return DGX_ERR_INVALID;
  HIPCHK(c, ensure_traces(c));
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int keep = c->stage_waves, best = keep;
  double best_us = 0.0;
  bool have = false;
  dgx_status st = DGX_OK;
  for (int w = 1; w <= 3 && st == DGX_OK; ++w)
  {
    if (!tile_valid(c->p, c->q, c->stage_tile, w))
    {
      if (us_out)
        us_out[w - 1] = 0.0;
      continue;
    }
    c->stage_waves = w;
    hipError_t e = rhs(c, c->U, c->TU, c->U, nullptr, c->Ub, c->TUb, nullptr, 1.0, 0.0, 0.0, 0);
    if (e == hipSuccess)
      e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r)
      e = rhs(c, c->U, c->TU, c->U, nullptr, c->Ub, c->TUb, nullptr, 1.0, 0.0, 0.0, 0);
    if (e == hipSuccess)
      e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess)
      e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess)
      e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
    {
      st = fail(c, DGX_ERR_HIP, "autotune: %s", hipGetErrorString(e));
      break;
    }
    double us = 1e3 * (double)ms / reps;
    if (us_out)
      us_out[w - 1] = us;
    if (!have || us < best_us)
    {
      best_us = us;
      best = w;
      have = true;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->stage_waves = (st == DGX_OK) ? best : keep;
  if (chosen)
    *chosen = c->stage_waves;
  if (st != DGX_OK)
    return st;
  return check_status(c);
}

dgx_status dgx_autotune_tiles(dgx_ctx* c, int reps, int* tile_out, int* waves_out,
                              double us_out[6])
{
  if (!c || reps < 1)
    return DGX_ERR_INVALID;
  const int keep_tile = c->stage_tile, keep_waves = c->stage_waves;
  int best_tile = keep_tile, best_waves = keep_waves;
  double best_us = 0.0;
  bool have = false;
  for (int t = 0; t <= 1; ++t)
  {
    double us[3] = {0.0, 0.0, 0.0};
    if (tile_valid(c->p, c->q, t, 2))
    {
      c->stage_tile = t;
      if (!tile_valid(c->p, c->q, t, c->stage_waves))
        c->stage_waves = 2;
      int w = 0;
      dgx_status st = dgx_autotune(c, reps, &w, us);
      if (st != DGX_OK)
      {
        c->stage_tile  = keep_tile;
        c->stage_waves = keep_waves;
        return st;
      }
      for (int k = 0; k < 3; ++k)
        if (us[k] > 0.0 && (!have || us[k] < best_us))
        {
          best_us    = us[k];
          best_tile  = t;
          best_waves = k + 1;
          have       = true;
        }
    }
    if (us_out)
      for (int k = 0; k < 3; ++k)
        us_out[3 * t + k] = us[k];
  }
  c->stage_tile  = best_tile;
  c->stage_waves = best_waves;
  if (tile_out)
    *tile_out = best_tile;
  if (waves_out)
    *waves_out = best_waves;
  return DGX_OK;
}

dgx_status dgx_sample_output_device(dgx_ctx* c, int var, int N, double* out_dev)
{
  if (!c || !out_dev || var < DGX_OUT_MACH || var > DGX_OUT_W || N < 1 || N > 64)
    return DGX_ERR_INVALID;
  HIPCHK(c, sample_dispatch(c->p, c->U, c->nelem, N, var, c->params.gamma, out_dev, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_sample_output(dgx_ctx* c, int var, int N, double* out_host)
{
  if (!c || !out_host || var < DGX_OUT_MACH || var > DGX_OUT_W || N < 1 || N > 64)
    return DGX_ERR_INVALID;
  size_t n = (size_t)c->nelem * (N + 1) * (N + 1) * (N + 1);
  double* d = nullptr;
  HIPCHK(c, hipMalloc(&d, n * sizeof(double)));
  hipError_t e = sample_dispatch(c->p, c->U, c->nelem, N, var, c->params.gamma, d, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(c, DGX_ERR_HIP, "sample_output: %s", hipGetErrorString(e));
  return DGX_OK;
}

dgx_status dgx_update_mean(dgx_ctx* c, uint64_t time_step, uint64_t mean_start)
{
  if (!c || time_step < mean_start)
    return DGX_ERR_INVALID;
  size_t sb = c->nstate * sizeof(double);
  if (!c->mean)
  {
    HIPCHK(c, hipMalloc(&c->mean, sb));
    c->bytes += sb;
    HIPCHK(c, hipMemsetAsync(c->mean, 0, sb, c->stream));
  }
  else if (time_step == mean_start)
    HIPCHK(c, hipMemsetAsync(c->mean, 0, sb, c->stream));
  double old_range = (double)(time_step - mean_start);
  double new_range = (double)(time_step + 1 - mean_start);
  HIPCHK(c, mean_dispatch(c->U, c->mean, c->nstate, old_range, new_range, c->stream));
  return DGX_OK;
}

dgx_status dgx_set_mean(dgx_ctx* c, const double* mean_host)
{
  if (!c || !mean_host)
    return DGX_ERR_INVALID;
  size_t sb = c->nstate * sizeof(double);
  if (!c->mean)
  {
    HIPCHK(c, hipMalloc(&c->mean, sb));
    c->bytes += sb;
  }
  HIPCHK(c, hipMemcpyAsync(c->Ua, mean_host, sb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, basis_dispatch(c->p, c->Ua, c->mean, c->nelem, 0, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}

dgx_status dgx_get_mean(dgx_ctx* c, double* mean_host)
{
  if (!c || !mean_host)
    return DGX_ERR_INVALID;
  if (!c->mean)
    return fail(c, DGX_ERR_INVALID, "no running mean yet (call dgx_update_mean first)");
  HIPCHK(c, basis_dispatch(c->p, c->mean, c->Ua, c->nelem, 1, c->stream));
  HIPCHK(c, hipMemcpyAsync(mean_host, c->Ua, c->nstate * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DGX_OK;
}
