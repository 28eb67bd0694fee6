#!/usr/bin/env python3
"""Throughput benchmark of the gfx950 DG residual + explicit RK path.

Workload (default): the 64^3 periodic Taylor-Green cube, p = 3 (64 DOF per
element, 16.8 M DOF), affine q = 1 elements, Mach 0.1, Re 1600, classic RK4,
seeded initial state.  One "step" is one RK4 time step (four fused residual
+ stage kernels) over the whole mesh, with the state resident in HBM, plus
the RMS(R) of its first stage computed and kept on the device, as the
reference's step functions compute it every step.

Prints one JSON line:
  value      = DOF * stages * steps / time / 1e9            [GDOF-stage/s]
  roofline   = algorithmic HBM bytes per stage-kernel launch / average launch
               time (HIP events on the library's stream, steps without the
               norm) vs 8 TB/s
  mixed, mixed_bf16, mixed_records = the same workload in the three
               mixed-precision forms (FP32 trace records; volume
               contractions on the matrix cores in FP32, or in 3-term split
               BF16, or in FP64): throughput, speedup over
               FP64, and after the workload's 100 RK4 steps from the initial
               state per-field L2/Linf errors and the kinetic-energy
               dissipation rate against the FP64 run
  cpu_baseline = the oracle (CPU restatement of the reference's dense
               residual + RK4) on a bounded sample of the same workload,
               one 8^3 block per host thread

--workload config0/config1 run the 8^3 p=2 cube and the curved channel the
same way; config3 the 64-case ensemble, config4 the adjoint; orders times
every solution order 0..8 (periodic Taylor-Green cubes of 1.3-2.1 M DOF, the
general-order kernel and, at p = 1..3, the tuned one).  Only one GPU is
used: --gpus N > 1 exits with a message.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dgpr_amd import cases, dgx  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def hip():
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipEventSynchronize.argtypes = [C.c_void_p]
    lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    lib.hipEventDestroy.argtypes = [C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


def alg_bytes_per_dof_stage(case, affine):
    """Compulsory HBM bytes per DOF per stage (FP64, 5 variables).

    Each full-state array costs 40 B/DOF.  TVD-RK3 fused stages read/write
    (U->U1), (U1,U->U2), (U2,U->U): 8 arrays over 3 stages.  Classic RK4 with
    an accumulator: (U->U1,acc), (U1,U,acc->U2,acc), (U2,U,acc->U3,acc),
    (U3,U,acc->U): 17 arrays over 4 stages.  Curved elements add their nodes
    (24 doubles per q=1 element) once per stage; affine meshes have a
    constant metric per element and add nothing."""
    arrays = 8.0 / 3.0 if case.scheme == "tvdRK3" else 17.0 / 4.0
    nq = case.mesh.q + 1
    geom = 0.0 if affine else 8.0 * 3 * nq ** 3 / case.mesh.nbf
    return 40.0 * arrays + geom


def cpu_baseline(seconds_hint=20.0, threads=None):
    """The oracle (the CPU restatement of the reference's dense residual and
    RK4) on a bounded sample of the 64^3 p=3 Taylor-Green workload: one
    independent 8^3 p=3 periodic Taylor-Green block (same per-element work,
    same physics) per host thread, RK4 steps at the workload's dt, all
    threads concurrently (the oracle's C calls release the GIL).  Setup (the
    oracle's per-element mass-matrix inversion) is timed apart."""
    import threading

    from oracle import oracle as orc
    nthr = threads or min(os.cpu_count() or 1, 16)
    c = cases.tgv_case(n=8, p=3, seed=2, dt=2.5e-4, nsteps=100, scheme="RK4")
    m = c.mesh
    prm = orc.params_from(c.params)
    t0 = time.perf_counter()
    geos = [orc.Geometry(m.p, m.q, m.nx, m.ny, m.nz, m.bounds, m.nodes) for _ in range(nthr)]
    setup = (time.perf_counter() - t0) / nthr
    steps = [0] * nthr
    stop = [False]

    def work(i):
        U = c.U0.copy()
        while not stop[0]:
            U, _ = geos[i].step(prm, U, c.dt, "RK4", 1)
            steps[i] += 1

    ths = [threading.Thread(target=work, args=(i,)) for i in range(nthr)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    time.sleep(seconds_hint)
    stop[0] = True
    for t in ths:
        t.join()
    el = time.perf_counter() - t0
    total_steps = sum(steps)
    val = c.dof * 4 * total_steps / el / 1e9
    return {"value": val, "unit": "GDOF-stage/s", "cores": nthr, "kind": "port",
            "sample": "%d RK4 steps in total of %d independent 8^3 p=3 periodic Taylor-Green "
                      "blocks (512 elements, 32768 DOF each, dt 2.5e-4), one per host thread, "
                      "oracle (CPU restatement of the reference), %.1f s wall; setup %.2f s "
                      "per block not included" % (total_steps, nthr, el, setup)}


def run_orders(a, h):
    """Every solution order the reference accepts, p = 0..8, on periodic
    Taylor-Green cubes of 1.3-2.1 M DOF (Mach 0.2, q = 1, RK4 dt 1e-4): the
    general-order kernel (stage tile 2) at every order and the context's
    default (tuned) kernel at p = 1..3, 10 timed steps each after 2 untimed
    ones, HIP events on the library's stream, no per-step norm."""
    steps = a.steps if a.steps_given else 10
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    sizes = {0: 128, 1: 64, 2: 40, 3: 32, 4: 24, 5: 20, 6: 16, 7: 14, 8: 12}
    rows = []
    for p, n in sizes.items():
        case = cases.tgv_case(n=n, p=p, seed=5, M=0.2, dt=1e-4)
        with dgx.Solver(case.mesh, case.params) as s:
            stream = C.c_void_p(s.stream())
            tiles = [None, 2] if p in (1, 2, 3) else [2]
            for t in tiles:
                if t is not None:
                    s.tile = t
                s.set_state(case.U0)
                s.step_async(case.dt, "RK4", 2)
                s.sync()
                sec = timed_steps(h, s, case, steps, stream, ev0, ev1)
                rows.append({"p": p, "mesh": "%d^3" % n, "dof": case.dof, "tile": s.tile,
                             "kernel": s.stage_kernel_name(),
                             "us_per_stage": 1e6 * sec / (steps * 4),
                             "gdof_stage_s": case.dof * 4 * steps / sec / 1e9})
    p4 = [r for r in rows if r["p"] == 4][0]
    print(json.dumps({
        "metric": "GDOF-stage-updates/s per solution order", "value": p4["gdof_stage_s"],
        "unit": "GDOF-stage/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": 4e-3 * p4["us_per_stage"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded initial states)",
        "config": {"workload": "orders_p0_to_p8", "value_is": "p = 4 (general-order kernel)",
                   "parallelism": "single GPU"},
        "orders": rows}), flush=True)


def run_ensemble(a, h):
    """64 independent periodic cubes (sizes, p, Mach, Re from seed 3) advanced
    together, one RK step of the whole ensemble replayed from a captured
    hipGraph.  The autotuner picks, per launch group, the occupancy target
    and tile (elements per workgroup, nodes per lane), then the layout: ONE
    launch per RK stage of the ensemble stage kernel for all 64 members
    (p = 1..3 in one kernel, one workgroup per work unit, largest members
    first), one such launch per order, or one batched launch per (p, q)
    group.  Every layout is timed; `value` is the chosen one's.  Also
    timed: the same cases stepped one after the other with the single-case
    stepper ("64 sequential plain launches").  The step count is the
    workload's (50); --steps overrides it."""
    cs = cases.config3()
    steps = a.steps if a.steps_given else cs[0].nsteps
    dts = [c.dt for c in cs]
    dof = sum(c.dof for c in cs)
    stream = None
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    res = {}
    with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
        stream = C.c_void_p(en.stream())
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
        waves = en.autotune("RK4", dts, reps=2) if a.autotune else None
        chosen = en.member_tuning(0)[0]
        # the chosen layout last: its states are checked
        for mode in [x for x in (0, 2, 1) if x != chosen] + [chosen]:
            en.set_layout(mode)
            for m, c in zip(en.members, cs):
                m.set_state(c.U0)
            en.step("RK4", dts, max(a.warmup, 1), graph=True)
            for m, c in zip(en.members, cs):
                m.set_state(c.U0)
            en.step("RK4", dts, 0, graph=True)  # traces of the reset states, untimed
            h.hipEventRecord(ev0, stream)
            en.step("RK4", dts, steps, graph=True, sync=False)
            h.hipEventRecord(ev1, stream)
            h.hipEventSynchronize(ev1)
            en.sync()
            ms = C.c_float()
            h.hipEventElapsedTime(C.byref(ms), ev0, ev1)
            tuning = sorted(set(en.member_tuning(i) for i in range(len(cs))))
            res[mode] = (ms.value * 1e-3, tuning)
        finite = all(bool(dgx.np.isfinite(m.get_state()).all()) for m in en.members)
        # the same cases one after the other, single-case launches
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
            m.step_async(c.dt, "RK4", 1)
            m.sync()
        h.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for m, c in zip(en.members, cs):
            m.step_async(c.dt, "RK4", steps)
            m.sync()
        h.hipDeviceSynchronize()
        t_seq = time.perf_counter() - t0
    t_ens, tuning = res[chosen]
    work = dof * 4 * steps
    layout_names = {1: "one launch per stage for all orders",
                    2: "one launch per order and stage",
                    0: "one batched launch per (p, q) group and stage"}
    value = work / t_ens / 1e9
    bpd = 170.0  # RK4 with accumulator, affine meshes (no geometry bytes)
    out = {
        "metric": "GDOF-stage-updates/s, mixed vs FP64 GPU at stated error; adjoint/forward cost",
        "value": value,
        "unit": "GDOF-stage/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_ens / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded meshes and initial states)",
        "state_finite": finite,
        "config": {
            "workload": "ensemble_64",
            "description": "64 periodic Taylor-Green cubes, n in {8,12,16,24,32}^3, p in {1,2,3}, "
                           "Mach U[0.05,0.3], Re log-U[400,5000] (seed 3), RK4",
            "cases": len(cs), "dof": dof, "scheme": "RK4",
            "groups": sorted(set((c.mesh.p, c.mesh.q) for c in cs)),
            "stage_waves_per_member": waves,
            "launch_tuning": [{"layout": p_, "tile": t_, "waves": w_} for p_, t_, w_ in tuning],
            "layout": layout_names[chosen] + " (autotuned), hipGraph",
            "parallelism": "single GPU",
        },
        "layouts": {
            layout_names[m]: {"value": work / res[m][0] / 1e9, "unit": "GDOF-stage/s",
                              "ms_per_step": 1e3 * res[m][0] / steps,
                              "launch_tuning": [{"layout": p_, "tile": t_, "waves": w_}
                                                for p_, t_, w_ in res[m][1]]}
            for m in res},
        "sequential": {
            "value": work / t_seq / 1e9,
            "unit": "GDOF-stage/s",
            "ms_per_step": 1e3 * t_seq / steps,
            "what": "the same 64 cases stepped one after the other with dgx_step_async",
        },
        "speedup_vs_sequential": t_seq / t_ens,
        "roofline": {
            "bound": "hbm",
            "kernel": ("dgx_rhs_batch_kernel<P,Q,W,AFF>" if chosen == 0 else
                       "dgx_ens_stage_kernel<W,TILE,AFF,PS>") + " (whole ensemble step)",
            "achieved": bpd * work / t_ens / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bpd * work / t_ens / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_dof_stage": bpd,
        },
    }
    print(json.dumps(out), flush=True)


def run_adjoint(a, h):
    """Discrete adjoint on the 32^3 p=2 periodic random-smooth-field case:
    forward time of N RK4 steps, time of dgx_adjoint_rk4 with the
    workload's schedule (checkpoints every ceil(sqrt(N)) steps; gradients
    with respect to the initial state and to mu), their ratio, the default
    schedule (every stage state and its trace records kept from the forward
    sweep when they fit in HBM) and checkpoints at every step for
    comparison, and
    the checks: central differences (eps = 1e-6) of the final kinetic energy
    along 20 seeded random directions, central differences in mu, and a
    residual-level dot-product test of the tangent-linear residual against
    its adjoint.  N is the workload's (100); --steps overrides it."""
    c = cases.config4()
    N = a.steps if a.steps_given else c.nsteps
    U0 = c.U0
    with dgx.Solver(c.mesh, c.params) as s:
        s.set_state(U0)
        s.step_async(c.dt, "RK4", max(a.warmup, 1))
        s.sync()
        s.set_state(U0)
        h.hipDeviceSynchronize()
        t0 = time.perf_counter()
        s.step_async(c.dt, "RK4", N)
        s.sync()
        t_fwd = time.perf_counter() - t0
        J_fwd = s.kinetic_energy()
        # the default schedule keeps every stage when this fits in 3/4 of the
        # free HBM (the library's test; no stored stages are held yet)
        free, total = C.c_size_t(), C.c_size_t()
        h.hipMemGetInfo(C.byref(free), C.byref(total))
        ntr = c.mesh.nelem * 6 * 8 * (c.mesh.p + 1) ** 2
        need = N * (4 * (s.nstate + ntr) + 4 * s.nstate) * 8
        stored = need <= 0.75 * free.value
        s.set_state(U0)
        h.hipDeviceSynchronize()
        t0 = time.perf_counter()
        J, g, dmu = s.adjoint_rk4(c.dt, N)
        t_adj = time.perf_counter() - t0
        s.set_state(U0)
        t0 = time.perf_counter()
        _, g0, _ = s.adjoint_rk4(c.dt, N, want_mu=False)
        t_adj_state = time.perf_counter() - t0
        # checkpointed schedules: every ceil(sqrt(N)) steps (the workload's
        # schedule, with and without dJ/dmu), every step
        K = int(dgx.np.ceil(dgx.np.sqrt(N)))
        s.set_state(U0)
        t0 = time.perf_counter()
        Jk, gkm, dmuk = s.adjoint_rk4(c.dt, N, interval=K)
        t_adj_sqrt = time.perf_counter() - t0
        s.set_state(U0)
        t0 = time.perf_counter()
        _, gk, _ = s.adjoint_rk4(c.dt, N, interval=K, want_mu=False)
        t_adj_chk = time.perf_counter() - t0
        s.set_state(U0)
        t0 = time.perf_counter()
        _, g1, _ = s.adjoint_rk4(c.dt, N, interval=1, want_mu=False)
        t_adj_all = time.perf_counter() - t0
        same = float(max(dgx.np.max(dgx.np.abs(g1 - g)), dgx.np.max(dgx.np.abs(gk - g)),
                         dgx.np.max(dgx.np.abs(g0 - g)), dgx.np.max(dgx.np.abs(gkm - g)),
                         abs(dmuk - dmu), abs(Jk - J)))

        def energy(U, mu=None):
            if mu is not None:
                prm = c.params.copy()
                prm.mu = mu
                s.set_params(prm)
            s.set_state(U)
            s.step_async(c.dt, "RK4", N)
            s.sync()
            e = s.kinetic_energy()
            if mu is not None:
                s.set_params(c.params)
            return e

        rng = dgx.np.random.default_rng(4)
        eps, errs = 1e-6, []
        for k in range(20):
            d = U0 * rng.uniform(-1.0, 1.0, U0.shape)
            fd = (energy(U0 + eps * d) - energy(U0 - eps * d)) / (2 * eps)
            ad = float(g @ d)
            errs.append(abs(fd - ad) / abs(ad))
        mu = c.params.mu
        fdmu = (energy(U0, mu * (1 + 1e-4)) - energy(U0, mu * (1 - 1e-4))) / (2e-4 * mu)
        # residual-level dot-product test at the initial state
        v = U0 * rng.uniform(-1.0, 1.0, U0.shape)
        w = rng.uniform(-1.0, 1.0, U0.shape)
        s.set_state(U0 + eps * v)
        _, fp = s.residual()
        s.set_state(U0 - eps * v)
        _, fm = s.residual()
        s.set_state(U0)
        gv = s.residual_vjp(w)
        lhs, rhs = float(w @ ((fp - fm) / (2 * eps))), float(v @ gv)
    dof = c.dof
    value = dof * 4 * N / t_fwd / 1e9
    out = {
        "metric": "GDOF-stage-updates/s, mixed vs FP64 GPU at stated error; adjoint/forward cost",
        "value": value,
        "unit": "GDOF-stage/s",
        "n_gpus": 1,
        "steps": N,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_fwd / N,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random smooth velocity field)",
        "config": {
            "workload": c.name, "description": "32^3 periodic p=2, |k|<=3 unit-RMS velocity, "
            "M=0.1, Re=1600, RK4 dt=1e-3", "nelem": c.mesh.nelem, "p": c.mesh.p, "dof": dof,
            "scheme": "RK4", "parallelism": "single GPU",
        },
        "adjoint": {
            "objective": "final kinetic energy (Gauss rule of the mass matrix)",
            "J": J, "J_forward_run": J_fwd,
            "forward_s": t_fwd,
            "adjoint_s": t_adj_sqrt,
            "adjoint_over_forward": t_adj_sqrt / t_fwd,
            "schedule": "checkpoints every ceil(sqrt(N)) steps; per segment one replay into "
                        "a store of its stage states, then the reverse sweep (state and mu "
                        "gradients)",
            "checkpointed_sqrtN_state_only_over_forward": t_adj_chk / t_fwd,
            "default_schedule": "stored stages" if stored else "checkpoints every ceil(sqrt(N))",
            "default_s": t_adj,
            "default_over_forward": t_adj / t_fwd,
            "stored_stage_bytes": need,
            "default_state_only_s": t_adj_state,
            "default_state_only_over_forward": t_adj_state / t_fwd,
            "checkpointed_every_step_state_only_over_forward": t_adj_all / t_fwd,
            "schedules_max_abs_grad_difference": same,
            "checkpoint_interval": K,
            "fd_directions": 20, "fd_eps": eps,
            "fd_max_rel_err": max(errs), "fd_median_rel_err": float(dgx.np.median(errs)),
            "dJdmu": dmu, "dJdmu_fd": fdmu, "dJdmu_rel_err": abs(dmu - fdmu) / abs(fdmu),
            "residual_dot_test_rel_err": abs(lhs - rhs) / abs(rhs),
        },
    }
    print(json.dumps(out), flush=True)


def kinetic_rate(s, case):
    """Kinetic energy of the current state and the dissipation rate -dK/dt
    over one more step (central in the step: (K_n - K_n+1) / dt)."""
    k0 = s.kinetic_energy()
    s.step(case.dt, case.scheme, 1, norms=False)
    k1 = s.kinetic_energy()
    return k0, (k0 - k1) / case.dt


def timed_steps(h, s, case, nsteps, stream, ev0, ev1, norms_dev=None):
    """GPU seconds of nsteps steps enqueued back to back (HIP events on the
    library's stream); with norms_dev, every step also stores its RMS(R) on
    the device as the reference's step functions compute it."""
    h.hipDeviceSynchronize()
    h.hipEventRecord(ev0, stream)
    if norms_dev is not None:
        s.step_norms_async(case.dt, case.scheme, nsteps, norms_dev.value)
    else:
        s.step_async(case.dt, case.scheme, nsteps)
    h.hipEventRecord(ev1, stream)
    h.hipEventSynchronize(ev1)
    s.sync()
    ms = C.c_float()
    h.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    return ms.value * 1e-3


def read_norms(h, norms_dev, n):
    out = np.zeros(n)
    h.hipMemcpy(out.ctypes.data_as(C.c_void_p), norms_dev, 8 * n, 2)  # device to host
    return out


def field_errors(Um, U64, case, nsteps, k64, eps64, kmx, epsmx):
    A = Um.reshape(case.mesh.nelem, 5, -1)
    B = U64.reshape(case.mesh.nelem, 5, -1)
    fields = {}
    for r, nm in enumerate(["rho", "rho_u", "rho_v", "rho_w", "rho_E"]):
        d = A[:, r] - B[:, r]
        fields[nm] = {
            "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(B[:, r])),
            "linf_rel": float(np.abs(d).max() / np.abs(B[:, r]).max()),
            "l2_rel_to_state": float(np.linalg.norm(d) / np.linalg.norm(U64)),
        }
    return {
        "steps": nsteps,
        "scheme": case.scheme,
        "state_l2_rel": float(np.linalg.norm(Um - U64) / np.linalg.norm(U64)),
        "fields": fields,
        "kinetic_energy_rel": abs(kmx - k64) / abs(k64),
        "ke_dissipation_rate_fp64": eps64,
        "ke_dissipation_rate_mixed": epsmx,
        "ke_dissipation_rate_rel": abs(epsmx - eps64) / abs(eps64),
    }


def fp64_reference(s, case):
    """The FP64 state after the workload's step count from the initial state
    and its kinetic energy and dissipation rate (the error legs' reference)."""
    s.precision = "fp64"
    s.set_state(case.U0)
    s.step(case.dt, case.scheme, case.nsteps, norms=False)
    U64 = s.get_state()
    k64, eps64 = kinetic_rate(s, case)
    return U64, k64, eps64


def run_mixed(a, h, s, case, value64, stream, ev0, ev1, norms_dev, form, ref):
    """A mixed-precision form on the same workload: the same untimed warmup
    and tuning and the same timed steps (with per-step norms) from the
    initial state; then the error leg: the workload's own step count (100
    RK4 steps for the 64^3 case) from the initial state, per-field L2/Linf
    errors of the state against the FP64 run's and the kinetic-energy
    dissipation rate of both over one more step."""
    s.set_state(case.U0)
    s.precision = form
    s.step_async(case.dt, case.scheme, a.warmup)
    s.sync()
    if a.autotune:
        tile, waves, _ = s.autotune_tiles(reps=3)
    else:
        tile, waves = s.tile, s.stage_waves
    s.set_state(case.U0)
    gpu_s = timed_steps(h, s, case, a.steps, stream, ev0, ev1, norms_dev)
    value = case.dof * case.stages * a.steps / gpu_s / 1e9
    gpu_s_nn = timed_steps(h, s, case, a.steps, stream, ev0, ev1)
    nerr = case.nsteps
    s.set_state(case.U0)
    s.step(case.dt, case.scheme, nerr, norms=False)
    Um = s.get_state()
    kmx, epsmx = kinetic_rate(s, case)
    s.precision = "fp64"
    U64, k64, eps64 = ref
    return {
        "value": value,
        "unit": "GDOF-stage/s",
        "speedup_vs_fp64": value / value64,
        "ms_per_step": 1e3 * gpu_s / a.steps,
        "avg_launch_us": gpu_s_nn / (a.steps * case.stages) * 1e6,
        "stage_tile": tile,
        "stage_waves_per_simd": waves,
        "form": s.MIXED_FORMS[form],
        "error_vs_fp64": field_errors(Um, U64, case, nerr, k64, eps64, kmx, epsmx),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="config2",
                    choices=["config0", "config1", "config2", "config3", "config4", "orders"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default occupancy target instead of timing all of them")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision run of the same workload")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one copy kernel of this size (traffic calibration)")
    a = ap.parse_args()
    a.steps_given = a.steps is not None
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build runs the path on a single "
              "MI355X; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)

    h = hip()
    if a.workload == "config3":
        return run_ensemble(a, h)
    if a.workload == "config4":
        return run_adjoint(a, h)
    if a.workload == "orders":
        return run_orders(a, h)
    case = {"config0": cases.config0, "config1": cases.config1, "config2": cases.config2}[a.workload]()
    if a.steps is None:
        a.steps = case.nsteps
    affine = case.mesh.affine
    s = dgx.Solver(case.mesh, case.params)
    s.set_state(case.U0)
    stream = C.c_void_p(s.stream())
    s.step_async(case.dt, case.scheme, a.warmup)
    s.sync()
    # launch tuning (untimed): each compiled (tile, occupancy target) pair
    # of the stage kernel on this mesh
    if a.autotune:
        tile, waves, tune_us = s.autotune_tiles(reps=3)
    else:
        tile, waves, tune_us = s.tile, s.stage_waves, None
    norms_dev = C.c_void_p()
    h.hipMalloc(C.byref(norms_dev), 8 * a.steps)
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))

    # the timed run: steps from the initial state, each with its RMS(R)
    s.set_state(case.U0)
    t0 = time.perf_counter()
    gpu_s = timed_steps(h, s, case, a.steps, stream, ev0, ev1, norms_dev)
    wall = time.perf_counter() - t0
    norms = read_norms(h, norms_dev, a.steps)
    U = s.get_state()
    finite = bool(np.isfinite(U).all()) and bool(np.isfinite(norms).all())
    value = case.dof * case.stages * a.steps / gpu_s / 1e9
    # the same steps without the norm: stage-kernel launches only, which
    # gives the stage kernel's average launch time
    gpu_s_nn = timed_steps(h, s, case, a.steps, stream, ev0, ev1)
    launches = a.steps * case.stages
    avg_launch_s = gpu_s_nn / launches
    bpd = alg_bytes_per_dof_stage(case, affine)
    bytes_per_launch = bpd * case.dof
    achieved = bytes_per_launch / avg_launch_s / 1e9

    if a.calibrate_copy_mib:
        nb = a.calibrate_copy_mib << 20
        src, dst = C.c_void_p(), C.c_void_p()
        h.hipMalloc(C.byref(src), nb)
        h.hipMalloc(C.byref(dst), nb)
        dgx.lib().dgx_device_copy(s.ctx, dst, src, nb)
        s.sync()
        h.hipFree(src)
        h.hipFree(dst)

    kernel_name = s.stage_kernel_name(waves, tile)
    mixed = {}
    if not a.no_mixed:
        ref = fp64_reference(s, case)
        for form in ("mixed", "mixed_bf16", "mixed_records"):
            mixed[form] = run_mixed(a, h, s, case, value, stream, ev0, ev1, norms_dev, form, ref)
    h.hipFree(norms_dev)
    s.close()

    # HBM bytes and executed FP64 flops per stage-kernel launch, from the
    # committed rocprofv3 counter summary of the same kernel and workload
    kname = kernel_name
    traffic = flops = None
    cf = os.path.join(ROOT, "profiles", "counters_%s.json" % a.workload)
    if os.path.exists(cf):
        cnt = json.load(open(cf))
        if cnt.get("kernel") == kname:
            traffic = cnt.get("hbm_bytes_per_launch")
            flops = cnt.get("fp64_flops_per_launch")

    out = {
        "metric": "GDOF-stage-updates/s, mixed vs FP64 GPU at stated error; adjoint/forward cost",
        "value": value,
        "unit": "GDOF-stage/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * gpu_s / a.steps,
        "wall_ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh and initial state)",
        "state_finite": finite,
        "norm_first_last": [float(norms[0]), float(norms[-1])],
        "config": {
            "workload": case.name,
            "description": {
                "config0": "8^3 periodic Taylor-Green p=2, RK4",
                "config1": "48x24x24 curved channel p=3 q=1, no-slip walls, TVD-RK3 CFL 0.3",
                "config2": "64^3 periodic Taylor-Green p=3, RK4",
            }[a.workload],
            "nelem": case.mesh.nelem, "p": case.mesh.p, "q": case.mesh.q,
            "dof": case.dof, "scheme": case.scheme, "dt": case.dt, "affine": affine,
            "timed": "%d steps from the initial state, each with the RMS(R) of its first "
                     "stage kept on the device" % a.steps,
            "parallelism": "single GPU",
            "stage_tile": tile,
            "stage_waves_per_simd": waves,
            "autotune_us": None if tune_us is None else
            {"tile%d" % t: [round(float(x), 2) for x in tune_us[t]] for t in range(2)},
        },
        "without_norm": {
            "value": case.dof * case.stages * a.steps / gpu_s_nn / 1e9,
            "ms_per_step": 1e3 * gpu_s_nn / a.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_dof_stage": bpd,
            "avg_launch_us": avg_launch_s * 1e6,
        },
        # FP64 utilisation of the same kernel: executed flops (rocprofv3
        # FP64 VALU counters) over the live average launch time
        "fp64": None if flops is None else {
            "achieved": flops / avg_launch_s / 1e12,
            "peak": FP64_PEAK_TFS,
            "unit": "TFLOP/s",
            "frac": flops / avg_launch_s / 1e12 / FP64_PEAK_TFS,
            "flops_per_launch": flops,
            "flops_per_dof_stage": flops / case.dof,
        },
    }
    out.update(mixed)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds, a.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
