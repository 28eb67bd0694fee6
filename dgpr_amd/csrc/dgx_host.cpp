// dgx_host.cpp -- host-side mesh/state generators and the .dg checkpoint
// format of libdgx (no device work).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/dgx.h"

namespace
{

const double kPi = 3.14159265358979323846;

// counter-based uniform [0,1) stream: splitmix64 of (seed, index)
inline uint64_t splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline double uniform01(uint64_t seed, uint64_t idx)
{
  uint64_t h = splitmix64(splitmix64(seed) ^ (idx * 0xD1B54A32D192ED03ull));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// equispaced Lagrange basis of order p on [0,1] (basis.cpp:91-104)
inline double lag(int i, int p, double x)
{
  if (p == 0)
    return 1.0;
  double xi = (double)i / p, v = 1.0;
  for (int j = 0; j <= p; ++j)
    if (j != i)
      v *= (x - (double)j / p) / (xi - (double)j / p);
  return v;
}

// global position of the equispaced order-p node bf of element e
// (ref2glo of lagrange_node3d, geometry_common.cpp:235-249; the single p = 0
// node sits at the reference origin, basis.cpp:64-70)
void node_position(const dgx_mesh* m, uint64_t e, int bf, double* x)
{
  int p = (int)m->p, q = (int)m->q, n = p + 1, nq = q + 1;
  const double dp = p ? (double)p : 1.0;  // p = 0: the only index is 0
  double r[3] = {(double)(bf % n) / dp, (double)((bf / n) % n) / dp, (double)(bf / (n * n)) / dp};
  x[0] = x[1] = x[2] = 0.0;
  const double* X = m->nodes + e * (uint64_t)nq * nq * nq * 3;
  for (int cz = 0; cz < nq; ++cz)
    for (int cy = 0; cy < nq; ++cy)
      for (int cx = 0; cx < nq; ++cx)
      {
        double w = lag(cx, q, r[0]) * lag(cy, q, r[1]) * lag(cz, q, r[2]);
        const double* xn = X + 3 * ((cz * nq + cy) * nq + cx);
        x[0] += xn[0] * w;
        x[1] += xn[1] * w;
        x[2] += xn[2] * w;
      }
}

bool mesh_ok(const dgx_mesh* m)
{
  return m && m->nodes && m->p <= 8 && m->q >= 1 && m->q <= 8 && m->nx > 0 &&
         m->ny > 0 && m->nz > 0;
}

}  // namespace

extern "C" {

dgx_status dgx_bump_nodes(uint64_t q, uint64_t nx, uint64_t ny, uint64_t nz, double lx,
                          double ly, double lz, double a, double bmp_height, double bmp_center,
                          double bmp_variance, double* nodes)
{
  if (!nodes || q < 1 || nx == 0 || ny == 0 || nz == 0)
    return DGX_ERR_INVALID;
  // The arithmetic below is the reference's node formula, expression for
  // expression, so the nodes (and every result built on them) agree bit for
  // bit with generate_bump_nodes (geometry_common.cpp:494-576):
  //  * element edges are uniform in x and z; in y they are uniform (a = 0)
  //    or tanh-clustered towards both walls, edge k of ny at
  //    ly ((tanh((2k/ny - 1) atanh a) / 2) (2/a) + 1) / 2;
  //  * the q + 1 nodes per direction split each element uniformly;
  //  * the straight y is then blended linearly from the Gaussian bump
  //    b(x) = h exp(-(x - c)^2 / (2 v)) at y = 0 to the flat wall at y = ly.
  const double hx = lx / (double)nx, hy = ly / (double)ny, hz = lz / (double)nz;
  std::vector<double> yedge(ny + 1);
  for (uint64_t k = 0; k <= ny; ++k)
  {
    if (a != 0.0)
    {
      const double t = -1. + 2. * ((double)k / ny);
      const double c = 0.5 * tanh(t * atanh(a));
      yedge[k] = ly * ((c * (2. / a) + 1.) / 2.);
    }
    else
      yedge[k] = k * hy;
  }
  const uint64_t m = q + 1, npe = m * m * m;
  const double dx = hx / (double)q, dz = hz / (double)q;
  for (uint64_t e = 0; e < nx * ny * nz; ++e)
  {
    const uint64_t ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    const double x0 = ex * hx, y0 = yedge[ey], z0 = ez * hz;
    const double dy = (yedge[ey + 1] - y0) / (double)q;
    double* X = nodes + e * npe * 3;
    for (uint64_t n = 0; n < npe; ++n)
    {
      const uint64_t i = n % m, j = (n / m) % m, k = n / (m * m);
      const double x = x0 + dx * (double)i;
      const double ystraight = y0 + dy * (double)j;
      const double bump =
          bmp_height * exp(-((x - bmp_center) * (x - bmp_center)) / (2. * bmp_variance));
      X[3 * n + 0] = x;
      X[3 * n + 1] = ((ystraight - 0.) / (ly - 0.)) * (ly - bump) + bump;
      X[3 * n + 2] = z0 + dz * (double)k;
    }
  }
  return DGX_OK;
}

dgx_status dgx_box_nodes(uint64_t q, uint64_t nx, uint64_t ny, uint64_t nz, double x0,
                         double y0, double z0, double lx, double ly, double lz, double amp,
                         uint64_t seed, double* nodes)
{
  if (!nodes || q < 1 || nx == 0 || ny == 0 || nz == 0)
    return DGX_ERR_INVALID;
  uint64_t qp1 = q + 1;
  double h[3] = {lx / nx, ly / ny, lz / nz};
  // phases of the displacement field, one pair per component
  double ph[3][2];
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      ph[d][s] = uniform01(seed, 1000 + 2 * d + s);
  for (uint64_t ez = 0; ez < nz; ++ez)
    for (uint64_t ey = 0; ey < ny; ++ey)
      for (uint64_t ex = 0; ex < nx; ++ex)
      {
        uint64_t e = nx * ny * ez + nx * ey + ex;
        for (uint64_t k = 0; k < qp1; ++k)
          for (uint64_t j = 0; j < qp1; ++j)
            for (uint64_t i = 0; i < qp1; ++i)
            {
              // global grid coordinates in [0,1]
              double sx = ((double)ex + (double)i / q) / nx;
              double sy = ((double)ey + (double)j / q) / ny;
              double sz = ((double)ez + (double)k / q) / nz;
              double* nd = nodes + (e * qp1 * qp1 * qp1 + qp1 * qp1 * k + qp1 * j + i) * 3;
              double base[3] = {x0 + lx * sx, y0 + ly * sy, z0 + lz * sz};
              for (int d = 0; d < 3; ++d)
              {
                double disp = 0.0;
                if (amp != 0.0)
                  disp = amp * h[d] * sin(2.0 * kPi * (sx + ph[d][0])) * sin(kPi * sy) *
                         sin(2.0 * kPi * (sz + ph[d][1]));
                nd[d] = base[d] + disp;
              }
            }
      }
  return DGX_OK;
}

dgx_status dgx_tgv_state(const dgx_mesh* m, double gamma, double rho0, double V0, double P0,
                         double noise, uint64_t seed, double* U)
{
  if (!mesh_ok(m) || !U || rho0 <= 0.0 || P0 <= 0.0)
    return DGX_ERR_INVALID;
  int n = (int)m->p + 1, nbf = n * n * n;
  uint64_t nelem = m->nx * m->ny * m->nz;
  for (uint64_t e = 0; e < nelem; ++e)
    for (int bf = 0; bf < nbf; ++bf)
    {
      double x[3];
      node_position(m, e, bf, x);
      double u = V0 * sin(x[0]) * cos(x[1]) * cos(x[2]);
      double v = -V0 * cos(x[0]) * sin(x[1]) * cos(x[2]);
      double w = 0.0;
      double p = P0 + (rho0 * V0 * V0 / 16.0) * (cos(2.0 * x[0]) + cos(2.0 * x[1])) *
                          (cos(2.0 * x[2]) + 2.0);
      double rho = rho0 * p / P0;
      uint64_t idx = (e * (uint64_t)nbf + bf) * 3;
      u += noise * (2.0 * uniform01(seed, idx + 0) - 1.0);
      v += noise * (2.0 * uniform01(seed, idx + 1) - 1.0);
      w += noise * (2.0 * uniform01(seed, idx + 2) - 1.0);
      double* Ue = U + e * 5 * (uint64_t)nbf;
      Ue[0 * nbf + bf] = rho;
      Ue[1 * nbf + bf] = rho * u;
      Ue[2 * nbf + bf] = rho * v;
      Ue[3 * nbf + bf] = rho * w;
      Ue[4 * nbf + bf] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    }
  return DGX_OK;
}

// turbulent (solver_init.cpp:30-73): the free stream with, on each velocity
// component i, amp * sum over wavelengths L = l/2, l/4, l/8, l/16 of
// prod_j sin(2 pi (x_j - 0.3 i) / L), amp = |u_fr| / 10; density and total
// energy stay at the free-stream values.
dgx_status dgx_turbulent_state(const dgx_mesh* m, const dgx_params* prm, double l, double* U)
{
  if (!mesh_ok(m) || !prm || !U || !(l > 0.0) || !(prm->Ufr[0] > 0.0))
    return DGX_ERR_INVALID;
  int n = (int)m->p + 1, nbf = n * n * n;
  uint64_t nelem = m->nx * m->ny * m->nz;
  const double rho = prm->Ufr[0];
  const double vfr[3] = {prm->Ufr[1] / rho, prm->Ufr[2] / rho, prm->Ufr[3] / rho};
  const double amp = sqrt(vfr[0] * vfr[0] + vfr[1] * vfr[1] + vfr[2] * vfr[2]) / 10.0;
  for (uint64_t e = 0; e < nelem; ++e)
    for (int bf = 0; bf < nbf; ++bf)
    {
      double x[3];
      node_position(m, e, bf, x);
      double* Ue = U + e * 5 * (uint64_t)nbf;
      Ue[0 * nbf + bf] = prm->Ufr[0];
      Ue[4 * nbf + bf] = prm->Ufr[4];
      for (int i = 0; i < 3; ++i)
      {
        double vel = vfr[i];
        for (int k = 1; k <= 4; ++k)
        {
          const double L = l / (double)(1 << k);
          double prod = 1.0;
          for (int j = 0; j < 3; ++j)
            prod *= sin(2.0 * kPi * ((x[j] - 0.3 * i) / L));
          vel += amp * prod;
        }
        Ue[(1 + i) * nbf + bf] = vel * rho;
      }
    }
  return DGX_OK;
}

dgx_status dgx_channel_state(const dgx_mesh* m, double gamma, double rho0, double Uc, double P0,
                             double frac, int kmax, double lx, double lz, uint64_t seed,
                             double* U)
{
  if (!mesh_ok(m) || !U || rho0 <= 0.0 || P0 <= 0.0 || kmax < 1 || kmax > 16)
    return DGX_ERR_INVALID;
  struct Mode
  {
    double kx, ky, kz, a[3], th[3], ps[3];
  };
  std::vector<Mode> modes;
  uint64_t cnt = 0;
  for (int mx = -kmax; mx <= kmax; ++mx)
    for (int my = 1; my <= kmax; ++my)
      for (int mz = -kmax; mz <= kmax; ++mz)
      {
        if (mx * mx + my * my + mz * mz > kmax * kmax)
          continue;
        Mode md;
        md.kx = 2.0 * kPi * mx / lx;
        md.ky = 0.5 * kPi * my;
        md.kz = 2.0 * kPi * mz / lz;
        for (int c = 0; c < 3; ++c)
        {
          md.a[c]  = 2.0 * uniform01(seed, 10 * cnt + c) - 1.0;
          md.th[c] = 2.0 * kPi * uniform01(seed, 10 * cnt + 3 + c);
          md.ps[c] = 2.0 * kPi * uniform01(seed, 10 * cnt + 6 + c);
        }
        modes.push_back(md);
        ++cnt;
      }
  int n = (int)m->p + 1, nbf = n * n * n;
  uint64_t nelem = m->nx * m->ny * m->nz;
  std::vector<double> pert(nelem * nbf * 3);
  double sumsq = 0.0;
  for (uint64_t e = 0; e < nelem; ++e)
    for (int bf = 0; bf < nbf; ++bf)
    {
      double x[3];
      node_position(m, e, bf, x);
      double y = x[1];
      double g = (1.0 - y * y) * (1.0 - y * y), gp = -4.0 * y * (1.0 - y * y);
      // gradient of each vector-potential component
      double dpsi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (const Mode& md : modes)
        for (int c = 0; c < 3; ++c)
        {
          double th = md.kx * x[0] + md.kz * x[2] + md.th[c];
          double ky = md.ky * y + md.ps[c];
          double C = cos(th), S = sin(th), K = cos(ky), Ks = sin(ky);
          dpsi[c][0] += md.a[c] * g * (-md.kx * S) * K;
          dpsi[c][1] += md.a[c] * C * (gp * K - g * md.ky * Ks);
          dpsi[c][2] += md.a[c] * g * (-md.kz * S) * K;
        }
      double* pv = &pert[(e * nbf + bf) * 3];
      pv[0] = dpsi[2][1] - dpsi[1][2];
      pv[1] = dpsi[0][2] - dpsi[2][0];
      pv[2] = dpsi[1][0] - dpsi[0][1];
      sumsq += pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2];
    }
  double rms = sqrt(sumsq / (double)(nelem * nbf));
  double scale = (rms > 0.0) ? frac * Uc / rms : 0.0;
  for (uint64_t e = 0; e < nelem; ++e)
    for (int bf = 0; bf < nbf; ++bf)
    {
      double x[3];
      node_position(m, e, bf, x);
      const double* pv = &pert[(e * nbf + bf) * 3];
      double u = Uc * (1.0 - x[1] * x[1]) + scale * pv[0];
      double v = scale * pv[1];
      double w = scale * pv[2];
      double* Ue = U + e * 5 * (uint64_t)nbf;
      Ue[0 * nbf + bf] = rho0;
      Ue[1 * nbf + bf] = rho0 * u;
      Ue[2 * nbf + bf] = rho0 * v;
      Ue[3 * nbf + bf] = rho0 * w;
      Ue[4 * nbf + bf] = P0 / (gamma - 1.0) + 0.5 * rho0 * (u * u + v * v + w * w);
    }
  return DGX_OK;
}

// ---- .dg files (io.cpp:33-159) ---------------------------------------------

// Bucket of an output key in the reference's map<simstate> (djb2 hash mod
// 256 table entries, data_structures.cpp:537-544, 670-673): write_state
// visits the outputs in increasing bucket order, and within a bucket in the
// order they were added (list<T>::add appends, data_structures.cpp:444-463).
static uint32_t output_bucket(const char* key)
{
  uint32_t h = 5381;
  for (const char* c = key; *c; ++c)
    h = h * 33u + (uint32_t)(unsigned char)*c;
  return h % 256u;
}

dgx_status dgx_write_dg_outputs(const char* path, uint64_t tstep, const dgx_mesh* m,
                                uint64_t nout, const char* const* keys,
                                const double* const* states)
{
  if (!path || !mesh_ok(m) || nout == 0 || !keys || !states)
    return DGX_ERR_INVALID;
  std::vector<uint64_t> order;
  for (uint64_t o = 0; o < nout; ++o)
  {
    if (!keys[o] || !states[o] || strlen(keys[o]) == 0 || strlen(keys[o]) >= 1024)
      return DGX_ERR_INVALID;
    for (uint64_t x = 0; x < o; ++x)
      if (strcmp(keys[x], keys[o]) == 0)
        return DGX_ERR_INVALID;  // the reference's map holds one output per key
    order.push_back(o);
  }
  std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
    return output_bucket(keys[a]) < output_bucket(keys[b]);
  });
  FILE* f = fopen(path, "wb");
  if (!f)
    return DGX_ERR_IO;
  bool ok = true;
  auto w64 = [&](const void* v) { ok = ok && fwrite(v, 8, 1, f) == 1; };
  w64(&tstep);
  w64(&m->nx);
  w64(&m->ny);
  w64(&m->nz);
  w64(&m->p);
  w64(&m->q);
  for (int i = 0; i < 6; ++i)
    w64(&m->bounds[i]);
  w64(&nout);
  uint64_t nelem = m->nx * m->ny * m->nz;
  uint64_t nq = m->q + 1, np = m->p + 1;
  uint64_t nn = nelem * nq * nq * nq * 3;
  ok = ok && fwrite(m->nodes, sizeof(double), nn, f) == nn;
  uint64_t ns = nelem * 5 * np * np * np;
  for (uint64_t o : order)
  {
    uint64_t klen = strlen(keys[o]);
    w64(&klen);
    ok = ok && fwrite(keys[o], 1, klen, f) == klen;
    ok = ok && fwrite(states[o], sizeof(double), ns, f) == ns;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? DGX_OK : DGX_ERR_IO;
}

dgx_status dgx_write_dg(const char* path, uint64_t tstep, const dgx_mesh* m, const double* U)
{
  const char* key = "state";
  return dgx_write_dg_outputs(path, tstep, m, 1, &key, &U);
}

static dgx_status read_header(FILE* f, uint64_t* tstep, dgx_mesh* m, uint64_t* nout)
{
  bool ok = true;
  auto r64 = [&](void* v) { ok = ok && fread(v, 8, 1, f) == 1; };
  r64(tstep);
  r64(&m->nx);
  r64(&m->ny);
  r64(&m->nz);
  r64(&m->p);
  r64(&m->q);
  for (int i = 0; i < 6; ++i)
    r64(&m->bounds[i]);
  r64(nout);
  m->nodes = nullptr;
  if (!ok || m->p > 8 || m->q < 1 || m->q > 8 || m->nx == 0 || m->ny == 0 || m->nz == 0)
    return DGX_ERR_IO;
  return DGX_OK;
}

dgx_status dgx_read_dg_header(const char* path, uint64_t* tstep, dgx_mesh* m)
{
  if (!path || !tstep || !m)
    return DGX_ERR_INVALID;
  FILE* f = fopen(path, "rb");
  if (!f)
    return DGX_ERR_IO;
  uint64_t nout = 0;
  dgx_status st = read_header(f, tstep, m, &nout);
  fclose(f);
  return st;
}

}  // extern "C"

// Visits the outputs of a .dg file in file order: fn(index, key, f) with the
// file positioned at the output's data, which fn reads or skips; returns
// the header's output count in *nout.
template <typename Fn>
static dgx_status for_each_output(const char* path, double* nodes_out, uint64_t* nout, Fn fn)
{
  FILE* f = fopen(path, "rb");
  if (!f)
    return DGX_ERR_IO;
  uint64_t tstep = 0;
  dgx_mesh m;
  dgx_status st = read_header(f, &tstep, &m, nout);
  if (st != DGX_OK)
  {
    fclose(f);
    return st;
  }
  uint64_t nelem = m.nx * m.ny * m.nz;
  uint64_t nq = m.q + 1, np = m.p + 1;
  uint64_t nn = nelem * nq * nq * nq * 3;
  uint64_t ns = nelem * 5 * np * np * np;
  bool ok = nodes_out ? fread(nodes_out, sizeof(double), nn, f) == nn
                      : fseek(f, (long)(nn * sizeof(double)), SEEK_CUR) == 0;
  for (uint64_t o = 0; ok && o < *nout; ++o)
  {
    uint64_t klen = 0;
    char key[1025];
    ok = fread(&klen, 8, 1, f) == 1 && klen < 1024;
    ok = ok && fread(key, 1, klen, f) == klen;
    if (!ok)
      break;
    key[klen] = 0;
    ok = fn(o, key, f, ns);
  }
  fclose(f);
  return ok ? DGX_OK : DGX_ERR_IO;
}

extern "C" {

dgx_status dgx_read_dg_output_key(const char* path, uint64_t index, char* key_out,
                                  uint64_t* nout)
{
  if (!path || !key_out || !nout)
    return DGX_ERR_INVALID;
  bool found = false;
  dgx_status st = for_each_output(path, nullptr, nout, [&](uint64_t o, const char* key, FILE* f,
                                                           uint64_t ns) {
    if (o == index)
    {
      strcpy(key_out, key);
      found = true;
    }
    return fseek(f, (long)(ns * sizeof(double)), SEEK_CUR) == 0;
  });
  if (st != DGX_OK)
    return st;
  return found ? DGX_OK : DGX_ERR_INVALID;
}

dgx_status dgx_read_dg_output(const char* path, const char* key, double* out)
{
  if (!path || !key || !out)
    return DGX_ERR_INVALID;
  uint64_t nout = 0;
  bool found = false;
  dgx_status st = for_each_output(path, nullptr, &nout, [&](uint64_t, const char* k, FILE* f,
                                                            uint64_t ns) {
    if (!found && strcmp(k, key) == 0)
    {
      found = true;
      return fread(out, sizeof(double), ns, f) == ns;
    }
    return fseek(f, (long)(ns * sizeof(double)), SEEK_CUR) == 0;
  });
  if (st != DGX_OK)
    return st;
  return found ? DGX_OK : DGX_ERR_IO;
}

dgx_status dgx_read_dg(const char* path, double* nodes_out, double* U_out)
{
  if (!path || !nodes_out || !U_out)
    return DGX_ERR_INVALID;
  FILE* f = fopen(path, "rb");
  if (!f)
    return DGX_ERR_IO;
  uint64_t tstep = 0, nout = 0;
  dgx_mesh m;
  dgx_status st = read_header(f, &tstep, &m, &nout);
  if (st != DGX_OK)
  {
    fclose(f);
    return st;
  }
  uint64_t nelem = m.nx * m.ny * m.nz;
  uint64_t nq = m.q + 1, np = m.p + 1;
  uint64_t nn = nelem * nq * nq * nq * 3;
  uint64_t ns = nelem * 5 * np * np * np;
  bool ok = fread(nodes_out, sizeof(double), nn, f) == nn;
  bool found = false;
  std::vector<double> scratch;
  for (uint64_t o = 0; ok && o < nout; ++o)
  {
    uint64_t klen = 0;
    char key[1025];
    ok = fread(&klen, 8, 1, f) == 1 && klen < 1024;
    ok = ok && fread(key, 1, klen, f) == klen;
    if (!ok)
      break;
    key[klen] = 0;
    if (!found && strcmp(key, "state") == 0)
    {
      ok = fread(U_out, sizeof(double), ns, f) == ns;
      found = ok;
    }
    else
    {
      scratch.resize(ns);
      ok = fread(scratch.data(), sizeof(double), ns, f) == ns;
    }
  }
  fclose(f);
  return (ok && found) ? DGX_OK : DGX_ERR_IO;
}

}  // extern "C"
