"""The C++ host API (include/dgx_solver.hpp) and the dgx_drive program, on
the CPU: option parsing as parse_solver_inputs / optparse do it
(input.cpp:68-125, io.cpp:194-300), the turbulent initial state
(solver_init.cpp:30-73) against the oracle and against closed-form values,
and the mesh/state/.dg helpers through a small program compiled against the
header.  Device runs of the driver are in test_gpu_driver.py."""
import os
import subprocess

import numpy as np
import pytest

from dgpr_amd import build, dgx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

OPTION_NAMES = ["p", "q", "dt", "niter", "chkint", "scheme", "eta", "turb", "mean_start",
                "gamma", "mu", "c", "rho", "u", "v", "w", "fx", "nx", "ny", "nz", "lx", "ly",
                "lz", "a", "bmx", "bpx", "bmy", "bpy", "bmz", "bpz", "bmp_h", "bmp_c", "bmp_v",
                "npx", "npy", "npz", "ofile", "ifile"]


@pytest.fixture(scope="module")
def driver():
    dgx.lib()
    if not os.path.exists(build.DRIVER):
        build.build_driver()
    return build.DRIVER


def test_driver_help_lists_reference_options(driver):
    r = subprocess.run([driver, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    lines = r.stdout.splitlines()
    assert lines[0] == "Options:"
    names = [ln.split(":")[0].strip() for ln in lines[1:1 + len(OPTION_NAMES)]]
    assert names == OPTION_NAMES
    assert "solver argument parsing failed!" in r.stdout


PROBE = r"""
#include <cstdio>
#include "dgx_solver.hpp"
using namespace dgx;
int main(int argc, char** argv)
{
  solver_inputs in;
  if (!parse_solver_inputs(argc, argv, in)) return 3;
  std::printf("p=%llu q=%llu nx=%llu ny=%llu nz=%llu dt=%.17g niter=%llu chkint=%llu\n",
              (unsigned long long)in.sol_order, (unsigned long long)in.geo_order,
              (unsigned long long)in.nx, (unsigned long long)in.ny, (unsigned long long)in.nz,
              in.dt, (unsigned long long)in.niter, (unsigned long long)in.chkint);
  std::printf("scheme=%d turb=%d fx=%.17g u=%.17g lx=%.17g a=%.17g\n", in.scheme,
              (int)in.init_turbulent, in.source[1], in.u, in.lx, in.stretch);
  std::printf("bounds=%lld %lld %lld %lld %lld %lld\n", (long long)in.bnd_mx,
              (long long)in.bnd_px, (long long)in.bnd_my, (long long)in.bnd_py,
              (long long)in.bnd_mz, (long long)in.bnd_pz);
  std::printf("ofile=%s ifile=%s\n", in.ofile_prefix, in.ifile_prefix);
  parameters prm = make_reference_state(in);
  std::printf("Ufr=%.17g %.17g %.17g %.17g %.17g src1=%.17g eta=%.17g\n", prm.Ufr[0], prm.Ufr[1],
              prm.Ufr[2], prm.Ufr[3], prm.Ufr[4], prm.source[1], prm.eta);
  s64 bounds[6] = {in.bnd_mx, in.bnd_px, in.bnd_my, in.bnd_py, in.bnd_mz, in.bnd_pz};
  core_geometry g = bump_shared(in.sol_order, in.geo_order, in.nx, in.ny, in.nz, in.lx, in.ly,
                                in.lz, in.stretch, in.bmp_height, in.bmp_center,
                                in.bmp_variance, bounds);
  simstate U(g);
  if (in.init_turbulent) turbulent(g, prm, in.lx, U); else uniform(prm, U);
  write_state(in.ofile_prefix, 41, g, U);
  u64 t = 0;
  core_geometry g2;
  simstate U2;
  read_state(in.ofile_prefix, t, g2, U2);
  bool same = t == 41 && g2.p == g.p && g2.q == g.q && g2.nelem() == g.nelem() &&
              g2.node == g.node && U2.U == U.U;
  for (int i = 0; i < 6; ++i) same = same && g2.bounds[i] == g.bounds[i];
  std::printf("roundtrip=%d nelem=%llu size=%llu\n", (int)same,
              (unsigned long long)g.nelem(), (unsigned long long)U.size());
  try { read_state("/nonexistent/x.dg", t, g2, U2); }
  catch (const error& e) { std::printf("io_error=%d\n", (int)e.status); }
  return 0;
}
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory, driver):
    d = tmp_path_factory.mktemp("probe")
    src = d / "probe.cpp"
    src.write_text(PROBE)
    exe = d / "probe"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L" + os.path.dirname(build.OUT), "-ldgx",
                    "-Wl,-rpath," + os.path.dirname(build.OUT)], check=True, timeout=240)
    return str(exe)


def run_probe(probe, args):
    r = subprocess.run([probe] + args, capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_parse_defaults_are_the_reference_defaults(probe, tmp_path):
    """input.cpp:21-66: p 1, q 2, 6x4x4 elements on 2x1x1, dt 1e-4, 1000
    iterations, checkpoint every 1000, tvdRK3, inflow/outflow in x,
    inviscid walls in y, periodic in z."""
    rc, out = run_probe(probe, ["-ofile", str(tmp_path / "d.dg")])
    assert rc == 0, out
    assert "p=1 q=2 nx=6 ny=4 nz=4 dt=0.0001 niter=1000 chkint=1000" in out
    assert ("scheme=%d turb=0 fx=0 u=0.10000000000000001 lx=2 a=0" % dgx.SCHEMES["tvdRK3"]
            in out)
    assert "bounds=-5 -6 -4 -4 -2 -2" in out
    assert "roundtrip=1 nelem=96 size=3840" in out
    assert "io_error=%d" % dgx.ERR_IO in out


def test_parse_every_kind_of_option(probe, tmp_path):
    args = ["-p", "3", "-q", "1", "-nx", "5", "-ny", "2", "-nz", "3", "-dt", "2.5e-4",
            "-niter", "17", "-chkint", "4", "-scheme", "RK4", "-turb", "-fx", "0.125",
            "-u", "0.2", "-lx", "3.5", "-a", "0.5", "-bmx", "periodic", "-bpx", "periodic",
            "-bmy", "no_slip_wall", "-bpy", "freestream", "-bmz", "inviscid_wall",
            "-bpz", "subsonic_outflow", "-ofile", str(tmp_path / "e.dg"), "-ifile", "abc"]
    rc, out = run_probe(probe, args)
    assert rc == 0, out
    assert "p=3 q=1 nx=5 ny=2 nz=3 dt=0.00025000000000000001 niter=17 chkint=4" in out
    assert ("scheme=%d turb=1 fx=0.125 u=0.20000000000000001 lx=3.5 a=0.5" % dgx.SCHEMES["RK4"]
            in out)
    assert "bounds=-2 -2 -7 -3 -4 -6" in out
    assert "ifile=abc" in out
    assert "src1=0.125 eta=6" in out
    assert "roundtrip=1 nelem=30 size=9600" in out


def test_parse_reports_and_skips_bad_options(probe, tmp_path):
    """optparse reports unknown options, names and stray words and goes on;
    a value-taking option at the end of the list is reported."""
    args = ["stray", "-bogus", "-bmx", "lava", "-scheme", "euler", "-nx", "3",
            "-ofile", str(tmp_path / "f.dg"), "-dt"]
    rc, out = run_probe(probe, args)
    assert rc == 0, out
    assert "ignoring non-option argument stray" in out
    assert "failed to find option number 2 : bogus" in out
    assert 'unknown boundary type "lava"' in out
    assert 'unknown time scheme "euler"' in out
    assert 'no argument provided for "dt" at end of argument list' in out
    assert "nx=3" in out and "dt=0.0001" in out
    assert "bounds=-5 -6 -4 -4 -2 -2" in out


def test_turbulent_state_matches_oracle(orc):
    """turbulent on a curved bump mesh (q 2, stretched) equals the oracle's
    restatement at every Lagrange node."""
    p, q, nx, ny, nz = 3, 2, 5, 3, 2
    nodes = dgx.bump_nodes(q, nx, ny, nz, 2.0, 1.0, 1.0, 0.4, 0.1, 1.0, 0.05)
    bounds = [dgx.BTYPES["subsonic_inflow"], dgx.BTYPES["subsonic_outflow"], dgx.BTYPES["inviscid_wall"],
              dgx.BTYPES["inviscid_wall"], dgx.BTYPES["periodic"], dgx.BTYPES["periodic"]]
    mesh = dgx.Mesh(p, q, nx, ny, nz, bounds, nodes)
    prm = dgx.make_reference_state(u=0.2, v=-0.05, w=0.1)
    U = dgx.turbulent_state(mesh, prm, 2.0)
    geo = orc.Geometry(p, q, nx, ny, nz, bounds, nodes)
    U_o = geo.turbulent(orc.params_from(prm), 2.0)
    assert np.max(np.abs(U - U_o)) < 1e-14
    Ur = U.reshape(mesh.nelem, 5, mesh.nbf)
    assert np.all(Ur[:, 0] == prm.Ufr[0]) and np.all(Ur[:, 4] == prm.Ufr[4])
    assert np.std(Ur[:, 1]) > 1e-3


def test_turbulent_state_closed_form():
    """On the affine box [0, 1.6]^3 of 8^3 p=1 elements (nodes every 0.2,
    l = 1.6): velocity component i is offset by 0.3 i, so component 0 keeps
    the free stream on the x = 0, y = 0 and z = 0 planes and component 2 on
    the x = 0.6, y = 0.6 and z = 0.6 planes (every mode has a sin 0 factor);
    at (0.2, 0.2, 0.2) = (l/8, l/8, l/8) component 0 gains exactly
    amp = |u_fr|/10 (the l/2 mode is 1 there, the shorter ones vanish)."""
    n, l = 8, 1.6
    nodes = dgx.box_nodes(1, n, n, n, 0.0, 0.0, 0.0, l, l, l)
    mesh = dgx.Mesh(1, 1, n, n, n, [dgx.BTYPES["periodic"]] * 6, nodes)
    prm = dgx.make_reference_state(u=0.3, v=0.1, w=-0.2)
    U = dgx.turbulent_state(mesh, prm, l).reshape(mesh.nelem, 5, 8)
    rho = prm.Ufr[0]
    amp = np.sqrt(0.3 ** 2 + 0.1 ** 2 + 0.2 ** 2) / 10.0
    X = nodes.reshape(mesh.nelem, 8, 3)
    vel = U[:, 1:4, :] / rho
    for comp, plane in ((0, 0.0), (2, 0.6)):
        for axis in range(3):
            on = np.abs(X[:, :, axis] - plane) < 1e-12
            assert on.sum() > 0
            d = vel[:, comp, :][on] - prm.Ufr[1 + comp] / rho
            assert np.max(np.abs(d)) < 1e-14, (comp, axis)
    at = np.all(np.abs(X - 0.2) < 1e-12, axis=2)
    assert at.sum() == 8
    d0 = vel[:, 0, :][at] - prm.Ufr[1] / rho
    assert np.allclose(d0, amp, rtol=0, atol=1e-14)
