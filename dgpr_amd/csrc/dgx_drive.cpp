// dgx_drive -- the solver's time-stepping program on one HIP device, the
// counterpart of the reference's GPU driver (solver/cudg.cu:24-151), built
// on the C++ host API of include/dgx_solver.hpp:
//   parse the options, make the reference state, build the bump channel and a
//   uniform (or -turb) state, or read "<ifile>.dg" and continue from the step
//   after the one it holds; step with tvdRK3 or RK4 printing "tstep : norm";
//   write "<ofile><tstep>.dg" whenever (tstep + 1) % chkint == 0 before that
//   step, and "<ofile>.dg" at the end, with every output the run holds: the
//   outputs read from "<ifile>.dg" are carried through (as cudg.cu's
//   outputs map is), and from step -mean_start on the running mean of the
//   state is kept on the device and written as "mean" (the CPU driver's
//   averaging, dg.cpp:122-131; a "mean" read from the input file is
//   continued).
// Errors print their message and exit with status 1, as errout does
// (error.cpp:34-57).
#include <cstdio>

#include "dgx_solver.hpp"

using namespace dgx;

static int fail(const char* what)
{
  std::printf("AN ERROR OCCURRED!!! (in main of dgx_drive):\n%s\n\n", what);
  std::fflush(stdout);
  return 1;
}

int main(int argc, char** argv)
{
  try
  {
    solver_inputs inputs;
    if (!parse_solver_inputs(argc, argv, inputs))
      return fail("solver argument parsing failed!");
    if (inputs.chkint == 0)
      return fail("checkpoint interval must be positive!");

    parameters sim_params = make_reference_state(inputs);

    /* initialize solution */

    u64 tstep = 0;
    core_geometry geom;
    outputs_map outputs;

    if (inputs.ifile_prefix[0] == '\0')
    {
      s64 bounds[6] = {inputs.bnd_mx, inputs.bnd_px, inputs.bnd_my,
                       inputs.bnd_py, inputs.bnd_mz, inputs.bnd_pz};
      geom = bump_shared(inputs.sol_order, inputs.geo_order, inputs.nx, inputs.ny, inputs.nz,
                         inputs.lx, inputs.ly, inputs.lz, inputs.stretch, inputs.bmp_height,
                         inputs.bmp_center, inputs.bmp_variance, bounds);
      simstate& U = outputs.add("state", simstate(geom));
      if (inputs.init_turbulent)
        turbulent(geom, sim_params, inputs.lx, U);
      else
        uniform(sim_params, U);
    }
    else
    {
      char ifile[1024];
      std::snprintf(ifile, sizeof(ifile), "%s.dg", inputs.ifile_prefix);
      read_state(ifile, tstep, geom, outputs);
      tstep += 1;  // starting on the step after the one that was read
    }
    simstate& U = outputs.get("state");

    /* data prep */

    step_fn step = nullptr;
    switch (inputs.scheme)
    {
      case DGX_TVDRK3: step = tvdRK3_gpu; break;
      case DGX_RK4: step = RK4_gpu; break;
      default: return fail("requested time scheme not implemented!");
    }

    device_geometry cugeom(geom, sim_params);
    cugeom.upload(U);
    if (outputs.has("mean"))
      cugeom.upload_mean(outputs.get("mean"));

    // outputs held on the device, brought to the host before each write
    auto download = [&]() {
      cugeom.download(outputs.get("state"));
      if (outputs.has("mean"))
        cugeom.download_mean(outputs.get("mean"));
    };

    /* time stepping */

    char ofile[1024 + 32];
    u64 tfinal = tstep + inputs.niter;
    for (; tstep < tfinal; ++tstep)
    {
      if ((tstep + 1) % inputs.chkint == 0)
      {
        download();
        std::snprintf(ofile, sizeof(ofile), "%s%" PRIu64 ".dg", inputs.ofile_prefix, tstep);
        write_state(ofile, tstep, geom, outputs);
      }
      step(tstep, inputs.dt, cugeom);

      // (re-)set the mean when averaging starts, then average
      if (tstep == inputs.mean_start)
        outputs.add("mean", simstate(geom));
      if (tstep >= inputs.mean_start)
      {
        outputs.get("mean");  // a restart past mean_start needs the mean it read
        cugeom.update_mean(tstep, inputs.mean_start);
      }
    }

    download();
    std::snprintf(ofile, sizeof(ofile), "%s.dg", inputs.ofile_prefix);
    write_state(ofile, tstep - 1, geom, outputs);
  }
  catch (const error& e)
  {
    return fail(e.what());
  }
  return 0;
}
