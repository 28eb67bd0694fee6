/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.  Drives the reference's own
 * CPU residual and explicit RK steps, compiled from the reference sources
 * where they lie (oracle/ref_build.py passes their directory with -I), to
 * write golden vectors for the tests.  Nothing in the product links or
 * loads it; its binary lives in oracle/_ref/ and is never committed.
 *
 *   ref_harness <in.dg> <out.bin> <scheme RK4|tvdRK3> <dt> <nsteps>
 *               <gamma> <mu> <c> <rho> <u> <v> <w> <eta> <s0> <s1> <s2> <s3> <s4>
 *
 * 1. reads the "state" output and the mesh of <in.dg> with the reference's
 *    read_state_distributed on one rank (io_distributed.cpp:37-54) and runs
 *    distributed_geometry::precompute, as the restart path of dg.cpp:64-75;
 * 2. builds the reference state with make_reference_state
 *    (solver_params.cpp:40-70) from the given free parameters and the
 *    reference's default gas constant, then sets the constant source term;
 * 3. evaluates residual() (solve.cpp:150-807) once at the file's state;
 * 4. takes <nsteps> steps with RK4_distributed / tvdRK3_distributed
 *    (time_distributed.cpp:46-102) and the auxiliary state count of
 *    dg.cpp:82-103; residual_norm prints "tstep : norm" on stdout.
 *
 * <out.bin>: u64 state size n, u64 nsteps, then R[n], f[n], U_final[n]
 * as doubles in the reference's simstate order [e][5][nbf].
 */
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "io_distributed.cpp"
#include "solver_init.cpp"
#include "time_distributed.cpp"

static double num(const char* s)
{
  char* end = nullptr;
  double v  = strtod(s, &end);
  if (end == s || *end != '\0')
  {
    fprintf(stderr, "ref_harness: bad number \"%s\"\n", s);
    exit(2);
  }
  return v;
}

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  if (argc != 19)
  {
    fprintf(stderr,
            "usage: ref_harness in.dg out.bin scheme dt nsteps gamma mu c rho u v w "
            "eta s0 s1 s2 s3 s4\n");
    MPI_Finalize();
    return 2;
  }

  solver_inputs in;
  in.gamma = num(argv[6]);
  in.mu    = num(argv[7]);
  in.c     = num(argv[8]);
  in.rho   = num(argv[9]);
  in.u     = num(argv[10]);
  in.v     = num(argv[11]);
  in.w     = num(argv[12]);
  in.eta   = num(argv[13]);
  for (int r = 0; r < 5; ++r)
    in.source[r] = num(argv[14 + r]);
  parameters prm = make_reference_state(in);

  const double dt   = num(argv[4]);
  const u64 nsteps  = (u64)strtoull(argv[5], nullptr, 10);
  const bool is_rk4 = strcmp(argv[3], "RK4") == 0;
  if (!is_rk4 && strcmp(argv[3], "tvdRK3") != 0)
  {
    fprintf(stderr, "ref_harness: unknown scheme %s\n", argv[3]);
    MPI_Finalize();
    return 2;
  }

  distributed_geometry geom;
  map<simstate> outputs;
  u64 tstep = 0;
  read_state_distributed(argv[1], MPI_COMM_WORLD, 1, 1, 1, 1, tstep, geom, outputs);
  geom.precompute();

  simstate& U = outputs.get("state");
  residual_workspace wsp(geom);
  simstate R(geom.core), f(geom.core);
  residual(U, geom, prm, wsp, R, f);

  FILE* fo = fopen(argv[2], "wb");
  if (!fo)
  {
    fprintf(stderr, "ref_harness: cannot write %s\n", argv[2]);
    MPI_Finalize();
    return 2;
  }
  u64 hdr[2] = {(u64)U.size(), nsteps};
  fwrite(hdr, sizeof(u64), 2, fo);
  fwrite(&R[0], sizeof(double), U.size(), fo);
  fwrite(&f[0], sizeof(double), U.size(), fo);

  u64 naux = is_rk4 ? 8 : 4;
  simstate* Uaux = new simstate[naux];
  for (u64 i = 0; i < naux; ++i)
    Uaux[i] = simstate(geom.core);
  for (u64 t = 0; t < nsteps; ++t)
  {
    if (is_rk4)
      RK4_distributed(t, dt, U, Uaux, geom, prm, wsp);
    else
      tvdRK3_distributed(t, dt, U, Uaux, geom, prm, wsp);
  }
  fflush(stdout);
  fwrite(&U[0], sizeof(double), U.size(), fo);
  fclose(fo);
  delete[] Uaux;
  MPI_Finalize();
  return 0;
}
