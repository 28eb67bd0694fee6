// dgx_solver.hpp -- C++ host API over the C ABI of libdgx, shaped like the
// reference solver's host side so that a driver reads like solver/cudg.cu:
//   solver_inputs / parse_solver_inputs  input.cpp:21-125, io.cpp:183-300
//   parameters / make_reference_state   solver_params.cpp:22-70
//   core_geometry / bump_shared          geometry_common.cpp:25-83, 494-560
//   simstate / uniform                   geometry_common.cpp:131-145, solver_init.cpp:22-28
//   read_state / write_state             io.cpp:33-159 (the "state" output)
//   device_geometry                      cuda_device_geometry + cuda_residual_workspace
//                                        (solve_cuda.cu:171-330)
//   RK4_gpu / tvdRK3_gpu                 the step function pointer of cudg.cu:75-91
//                                        (time_cuda.cu:126-185), printing the
//                                        residual norm as residual_norm does
//                                        (time_distributed.cpp:23-44)
// Header-only; needs libdgx (include/dgx.h).  Errors throw dgx::error.
#pragma once

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgx.h"

namespace dgx
{

using u64 = uint64_t;
using s64 = int64_t;

struct error : std::runtime_error
{
  dgx_status status;
  error(dgx_status st, const std::string& what) : std::runtime_error(what), status(st) {}
};

inline void check(dgx_status st, const dgx_ctx* ctx = nullptr)
{
  if (st != DGX_OK)
    throw error(st, dgx_last_error(ctx));
}

// parameters (solver_params.cpp:22-36)
using parameters = dgx_params;

// make_reference_state (solver_params.cpp:40-70)
inline parameters make_reference_state(double gamma, double R, double mu, double c, double rho,
                                       double u, double v, double w, double eta,
                                       const double source[5] = nullptr)
{
  parameters p;
  check(dgx_make_reference_state(gamma, R, mu, c, rho, u, v, w, eta, source, &p));
  return p;
}

// Driver options with the reference's names and defaults (input.cpp:21-66).
// The reference also reads -npx/-npy/-npz (MPI process grid of the CPU
// driver); they are accepted here and have no effect on one device.
struct solver_inputs
{
  u64 sol_order = 1;
  u64 geo_order = 2;

  u64 nx = 6, ny = 4, nz = 4;
  double lx = 2., ly = 1., lz = 1.;
  u64 npx = 1, npy = 1, npz = 1;
  double stretch = 0.;
  double bmp_height = 0.1;
  double bmp_center = 1.0;
  double bmp_variance = 0.05;
  s64 bnd_mx = DGX_BT_SUBSONIC_INFLOW;
  s64 bnd_px = DGX_BT_SUBSONIC_OUTFLOW;
  s64 bnd_my = DGX_BT_INVISCID_WALL;
  s64 bnd_py = DGX_BT_INVISCID_WALL;
  s64 bnd_mz = DGX_BT_PERIODIC;
  s64 bnd_pz = DGX_BT_PERIODIC;

  double gamma = 1.4, gcnst = 287., mu = 1e-3, c = 1.0, rho = 1.0;
  double u = 0.1, v = 0.0, w = 0.0;

  char ofile_prefix[512] = "output";
  char ifile_prefix[512] = "";

  double dt = 1e-4;
  u64 niter = 1000;
  u64 chkint = 1000;
  u64 mean_start = (u64)1 << 63;
  int scheme = DGX_TVDRK3;
  double eta = 6;
  double source[5] = {};
  bool init_turbulent = false;
};

// make_reference_state(solver_inputs) (solver_params.cpp:40-70)
inline parameters make_reference_state(const solver_inputs& in)
{
  return make_reference_state(in.gamma, in.gcnst, in.mu, in.c, in.rho, in.u, in.v, in.w, in.eta,
                              in.source);
}

// Name tables of boundaries.cpp:43-50 and time.cpp:29-31
inline bool btype_from_name(const char* s, s64& out)
{
  static const struct { const char* name; s64 code; } tab[] = {
    {"periodic", DGX_BT_PERIODIC},
    {"freestream", DGX_BT_FREESTREAM},
    {"inviscid_wall", DGX_BT_INVISCID_WALL},
    {"subsonic_inflow", DGX_BT_SUBSONIC_INFLOW},
    {"subsonic_outflow", DGX_BT_SUBSONIC_OUTFLOW},
    {"no_slip_wall", DGX_BT_NO_SLIP_WALL}};
  for (const auto& t : tab)
    if (std::strcmp(s, t.name) == 0)
    {
      out = t.code;
      return true;
    }
  return false;
}

inline bool scheme_from_name(const char* s, int& out)
{
  if (std::strcmp(s, "tvdRK3") == 0)
    out = DGX_TVDRK3;
  else if (std::strcmp(s, "RK4") == 0)
    out = DGX_RK4;
  else
    return false;
  return true;
}

// parse_solver_inputs (input.cpp:68-125) with optparse's rules
// (io.cpp:194-300): "-name value" pairs, "-turb" alone, "-h" as the first
// argument prints the option list and fails; unknown options, unknown
// boundary or scheme names and stray words are reported and skipped.
// Returns false only for "-h".
inline bool parse_solver_inputs(int argc, char** argv, solver_inputs& in)
{
  enum kind { k_u64, k_double, k_string, k_bool, k_btype, k_scheme };
  struct opt { const char* name; const char* description; kind type; void* data; };
  const opt opts[] = {
    {"p", "solution order", k_u64, &in.sol_order},
    {"q", "geometry order", k_u64, &in.geo_order},
    {"dt", "time step", k_double, &in.dt},
    {"niter", "iteration count", k_u64, &in.niter},
    {"chkint", "checkpoint interval", k_u64, &in.chkint},
    {"scheme", "time stepping scheme", k_scheme, &in.scheme},
    {"eta", "BR2 stabilization parameter", k_double, &in.eta},
    {"turb", "initialize turbulent instead of uniform", k_bool, &in.init_turbulent},
    {"mean_start", "iteration at which to start averaging", k_u64, &in.mean_start},
    {"gamma", "ratio of specific heats", k_double, &in.gamma},
    {"mu", "dynamic viscosity", k_double, &in.mu},
    {"c", "speed of sound", k_double, &in.c},
    {"rho", "reference density", k_double, &in.rho},
    {"u", "reference x velocity", k_double, &in.u},
    {"v", "reference y velocity", k_double, &in.v},
    {"w", "reference z velocity", k_double, &in.w},
    {"fx", "force per unit volume in x", k_double, &in.source[1]},
    {"nx", "num elems in x", k_u64, &in.nx},
    {"ny", "num elems in y", k_u64, &in.ny},
    {"nz", "num elems in z", k_u64, &in.nz},
    {"lx", "domain length in x", k_double, &in.lx},
    {"ly", "domain length in y", k_double, &in.ly},
    {"lz", "domain length in z", k_double, &in.lz},
    {"a", "wall normal stretching parameter (0, 1)", k_double, &in.stretch},
    {"bmx", "-x boundary type", k_btype, &in.bnd_mx},
    {"bpx", "+x boundary type", k_btype, &in.bnd_px},
    {"bmy", "-y boundary type", k_btype, &in.bnd_my},
    {"bpy", "+y boundary type", k_btype, &in.bnd_py},
    {"bmz", "-z boundary type", k_btype, &in.bnd_mz},
    {"bpz", "+z boundary type", k_btype, &in.bnd_pz},
    {"bmp_h", "bump height", k_double, &in.bmp_height},
    {"bmp_c", "bump center in x", k_double, &in.bmp_center},
    {"bmp_v", "bump variance", k_double, &in.bmp_variance},
    {"npx", "processor count in x", k_u64, &in.npx},
    {"npy", "processor count in y", k_u64, &in.npy},
    {"npz", "processor count in z", k_u64, &in.npz},
    {"ofile", "output file prefix", k_string, in.ofile_prefix},
    {"ifile", "input file prefix", k_string, in.ifile_prefix},
  };
  const int optc = (int)(sizeof(opts) / sizeof(opts[0]));

  if (argc > 1 && std::strcmp(argv[1], "-h") == 0)
  {
    std::printf("Options:\n");
    for (int oi = 0; oi < optc; ++oi)
      std::printf("  %10s: %s\n", opts[oi].name, opts[oi].description);
    return false;
  }
  int head = 1;
  while (head < argc)
  {
    if (argv[head][0] != '-')
    {
      std::printf("ignoring non-option argument %s\n", argv[head]);
      ++head;
      continue;
    }
    int oi = 0;
    for (; oi < optc; ++oi)
      if (std::strcmp(opts[oi].name, argv[head] + 1) == 0)
        break;
    if (oi == optc)
    {
      std::printf("failed to find option number %d : %s\n", head, argv[head] + 1);
      ++head;
      continue;
    }
    const opt& o = opts[oi];
    if (o.type == k_bool)
    {
      *(bool*)o.data = true;
      head += 1;
      continue;
    }
    if (head + 1 >= argc)
    {
      std::printf("no argument provided for \"%s\" at end of argument list\n", argv[head] + 1);
      ++head;
      continue;
    }
    const char* val = argv[head + 1];
    switch (o.type)
    {
      case k_u64: *(u64*)o.data = (u64)std::strtoull(val, nullptr, 10); break;
      case k_double: *(double*)o.data = std::strtod(val, nullptr); break;
      case k_string:
        std::snprintf((char*)o.data, sizeof(in.ofile_prefix), "%s", val);
        break;
      case k_btype:
        if (!btype_from_name(val, *(s64*)o.data))
          std::printf("unknown boundary type \"%s\"\n", val);
        break;
      case k_scheme:
        if (!scheme_from_name(val, *(int*)o.data))
          std::printf("unknown time scheme \"%s\"\n", val);
        break;
      default: break;
    }
    head += 2;
  }
  return true;
}

// The persistent part of core_geometry: orders, element counts, boundary
// codes per side (-x +x -y +y -z +z) and element nodes [e][(q+1)^3][3].
struct core_geometry
{
  u64 p = 1, q = 1, nx = 0, ny = 0, nz = 0;
  s64 bounds[6] = {DGX_BT_PERIODIC, DGX_BT_PERIODIC, DGX_BT_PERIODIC,
                   DGX_BT_PERIODIC, DGX_BT_PERIODIC, DGX_BT_PERIODIC};
  std::vector<double> node;

  u64 nelem() const { return nx * ny * nz; }
  u64 nbf() const { return (p + 1) * (p + 1) * (p + 1); }
  dgx_mesh mesh() const
  {
    dgx_mesh m;
    m.p = p;
    m.q = q;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    for (int i = 0; i < 6; ++i)
      m.bounds[i] = bounds[i];
    m.nodes = node.data();
    return m;
  }
};

// bump_shared (geometry_common.cpp:494-560): the bump channel mesh
inline core_geometry bump_shared(u64 p, u64 q, u64 nx, u64 ny, u64 nz, double lx, double ly,
                                 double lz, double a, double bmp_height, double bmp_center,
                                 double bmp_variance, const s64 bounds[6])
{
  core_geometry g;
  g.p = p;
  g.q = q;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  for (int i = 0; i < 6; ++i)
    g.bounds[i] = bounds[i];
  g.node.resize(nx * ny * nz * (q + 1) * (q + 1) * (q + 1) * 3);
  check(dgx_bump_nodes(q, nx, ny, nz, lx, ly, lz, a, bmp_height, bmp_center, bmp_variance,
                       g.node.data()));
  return g;
}

// simstate (geometry_common.cpp:131-145): U[e][r][bf]
struct simstate
{
  u64 nelem = 0, nbfnc = 0;
  std::vector<double> U;
  simstate() = default;
  explicit simstate(const core_geometry& g) : nelem(g.nelem()), nbfnc(g.nbf()), U(nelem * 5 * nbfnc) {}
  u64 size() const { return U.size(); }
  double& operator()(u64 e, u64 r, u64 b) { return U[(e * 5 + r) * nbfnc + b]; }
};

// uniform (solver_init.cpp:22-28)
inline void uniform(const parameters& prm, simstate& U)
{
  for (u64 e = 0; e < U.nelem; ++e)
    for (u64 r = 0; r < 5; ++r)
      for (u64 b = 0; b < U.nbfnc; ++b)
        U(e, r, b) = prm.Ufr[r];
}

// turbulent (solver_init.cpp:30-73): free stream plus four sine-product
// modes of wavelength l/2 .. l/16 on each velocity component
inline void turbulent(const core_geometry& g, const parameters& prm, double l, simstate& U)
{
  dgx_mesh m = g.mesh();
  check(dgx_turbulent_state(&m, &prm, l, U.U.data()));
}

// read_state / write_state (io.cpp:33-159) with the single output "state"
inline void read_state(const std::string& path, u64& tstep, core_geometry& g, simstate& U)
{
  dgx_mesh m;
  check(dgx_read_dg_header(path.c_str(), &tstep, &m));
  g.p = m.p;
  g.q = m.q;
  g.nx = m.nx;
  g.ny = m.ny;
  g.nz = m.nz;
  for (int i = 0; i < 6; ++i)
    g.bounds[i] = m.bounds[i];
  g.node.resize(g.nelem() * (g.q + 1) * (g.q + 1) * (g.q + 1) * 3);
  U = simstate(g);
  check(dgx_read_dg(path.c_str(), g.node.data(), U.U.data()));
}

inline void write_state(const std::string& path, u64 tstep, const core_geometry& g,
                        const simstate& U)
{
  dgx_mesh m = g.mesh();
  check(dgx_write_dg(path.c_str(), tstep, &m, U.U.data()));
}

// map<simstate> of the drivers (dg.cpp:44, cudg.cu:44): the named outputs a
// .dg file carries ("state", and "mean" once the CPU driver averages)
struct outputs_map
{
  std::vector<std::pair<std::string, simstate>> items;

  bool has(const std::string& key) const
  {
    for (const auto& it : items)
      if (it.first == key)
        return true;
    return false;
  }
  // map::add: replaces an existing output of that key
  simstate& add(const std::string& key, const simstate& s)
  {
    for (auto& it : items)
      if (it.first == key)
        return it.second = s;
    items.emplace_back(key, s);
    return items.back().second;
  }
  simstate& get(const std::string& key)
  {
    for (auto& it : items)
      if (it.first == key)
        return it.second;
    throw error(DGX_ERR_INVALID, "no output named \"" + key + "\"");
  }
};

// read_state / write_state (io.cpp:33-159) of every output
inline void read_state(const std::string& path, u64& tstep, core_geometry& g, outputs_map& out)
{
  simstate U;
  read_state(path, tstep, g, U);
  out.items.clear();
  u64 nout = 0;
  char key[1025];
  check(dgx_read_dg_output_key(path.c_str(), 0, key, &nout));
  for (u64 i = 0; i < nout; ++i)
  {
    check(dgx_read_dg_output_key(path.c_str(), i, key, &nout));
    simstate& s = out.add(key, simstate(g));
    check(dgx_read_dg_output(path.c_str(), key, s.U.data()));
  }
}

inline void write_state(const std::string& path, u64 tstep, const core_geometry& g,
                        const outputs_map& out)
{
  dgx_mesh m = g.mesh();
  std::vector<const char*> keys;
  std::vector<const double*> states;
  for (const auto& it : out.items)
  {
    keys.push_back(it.first.c_str());
    states.push_back(it.second.U.data());
  }
  check(dgx_write_dg_outputs(path.c_str(), tstep, &m, keys.size(), keys.data(), states.data()));
}

// cuda_device_geometry + cuda_residual_workspace: the mesh, the state and
// the stage buffers on the current HIP device
class device_geometry
{
 public:
  device_geometry(const core_geometry& g, const parameters& prm)
  {
    dgx_mesh m = g.mesh();
    check(dgx_create(&m, &prm, &ctx_));
  }
  ~device_geometry() { dgx_destroy(ctx_); }
  device_geometry(const device_geometry&) = delete;
  device_geometry& operator=(const device_geometry&) = delete;

  dgx_ctx* ctx() const { return ctx_; }
  void set_params(const parameters& prm) { check(dgx_set_params(ctx_, &prm), ctx_); }
  // DGX_PREC_FP64 (default) or DGX_PREC_MIXED (FP32 face trace records)
  void set_precision(int prec) { check(dgx_set_precision(ctx_, prec), ctx_); }
  // shuffle_state + H2D (cudg.cu:97-106)
  void upload(const simstate& U) { check(dgx_set_state(ctx_, U.U.data()), ctx_); }
  // D2H + unshuffle_state (cudg.cu:131-136)
  void download(simstate& U) const { check(dgx_get_state(ctx_, U.U.data()), ctx_); }
  // update_mean (output.cpp:39-49) of the device state, kept on the device;
  // restarts from zero when tstep == mean_start
  void update_mean(u64 tstep, u64 mean_start)
  {
    check(dgx_update_mean(ctx_, tstep, mean_start), ctx_);
  }
  void upload_mean(const simstate& M) { check(dgx_set_mean(ctx_, M.U.data()), ctx_); }
  void download_mean(simstate& M) const { check(dgx_get_mean(ctx_, M.U.data()), ctx_); }
  // one step; norm receives the RMS residual norm of the first stage
  // (NaN when the step went unstable, reported as DGX_ERR_NAN)
  dgx_status step(int scheme, double dt, double& norm)
  {
    norm = 0.0;
    return dgx_step(ctx_, scheme, dt, 1, &norm);
  }

 private:
  dgx_ctx* ctx_ = nullptr;
};

// The step functions of cudg.cu:75-91: one step of the device state, the
// residual norm printed as "tstep : norm" and a NaN norm an error, as
// residual_norm does (time_distributed.cpp:23-44)
using step_fn = void (*)(u64 tstep, double dt, device_geometry& geom);

inline void step_and_report(int scheme, u64 tstep, double dt, device_geometry& geom)
{
  double norm;
  dgx_status st = geom.step(scheme, dt, norm);
  if (st == DGX_OK || st == DGX_ERR_NAN)
    std::printf("%" PRIu64 " : %.17e\n", tstep, norm);
  if (st == DGX_ERR_NAN)
    throw error(st, "looks like things went unstable!");
  check(st, geom.ctx());
}

inline void RK4_gpu(u64 tstep, double dt, device_geometry& geom)
{
  step_and_report(DGX_RK4, tstep, dt, geom);
}

inline void tvdRK3_gpu(u64 tstep, double dt, device_geometry& geom)
{
  step_and_report(DGX_TVDRK3, tstep, dt, geom);
}

}  // namespace dgx
