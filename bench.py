#!/usr/bin/env python3
"""Throughput of the MI355X mutual-polarization solve (PB-SAM's hot path).

Workload: two synthetic molecules of 60 spheres each (radii U[2,6] A, rejection-packed in
a 28.8 A ball, 10-40 charges of +-1 e per sphere), pole order p = 30, 1800 spiral points
per sphere, FP64, centres 50 A apart with random orientations, 0.05 M salt, eps 4 / 78,
inter-molecular polarization cutoff 100 A (every pair polarized).  Each molecule type is
self-polarized once (untimed set-up, the reference's `pbsam spol` stage).

A step is one force+torque evaluation of the system, the query Brownian dynamics makes
every step (CMolecule::computeForces, src/moldynamics.cpp:554-587): pair lists, mutual
polarization from the self values, the gradient phase and the forces and torques, in the
reference's Gauss-Seidel order with its stock stopping rules.  Inputs stay resident in
HBM; the timed region is bracketed by stream synchronizations.  `solves_per_s` is the
mutual solve alone (CMolecule::polarize_mutual, src/molecule.cpp:1989-2075), timed after.
`batch` is the batched path: 256 Brownian-dynamics snapshots of the same two molecules
(centre distance uniform in [50, 80] A) evaluated together through pbsam_batch_create,
beside the plain one-system-per-call path timed on a sample of the same snapshots.
`bd` is batched Brownian dynamics (CBDnam trajectories, one per system of a batch): 1024
trajectories of two 100-sphere molecules at p=10, forces every step, reported as
trajectory-steps/s beside the oracle's trajectories on the host cores.  `large` is the
large-system workload: eight distinct 500-sphere molecules at p=30 with 2000 points per
sphere in one system (set-up, mutual solves and force evaluations; operators past the
context's store cap are made per use).  `exact` is the exact derivative on the main system:
its fixed point by GMRES (pbsam_solve_exact; the reference's rounds do not converge on these
molecules, whose spheres nearly touch) and both molecules' forces and torques from the
adjoint (pbsam_adjoint_forces).

Prints one JSON line.  `roofline` is the kernel with the largest share of device time,
timed with HIP events around each of its launches on the library's stream, with the
algorithmic flops/bytes its launches processed (counted on the device from the
reference's loop bounds at the truncation orders actually used).  The event pairs add
host calls and queue packets between dependent kernels (about a fifth of an evaluation),
so they time a repeat of the K evaluations right after the timed region, not the timed
region itself; the batched leg's per-kernel figures likewise come from one extra call.  `cpu_baseline` is the
CPU oracle (a scalar restatement of the reference, 1 thread) on the same system.
"""
import argparse
import faulthandler
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pb-sam_amd"))

METRIC = "mutual-polarization solves/s & force+torque evals/s, p=30, FP64 rel err ≤1e-8"
N_SPH, P, NPTS, DIST = 60, 30, 1800, 50.0
P_ORD = P
SEEDS, POSE_SEED, BATCH_SEED = (101, 102), 7, 5000
PEAK_FP64 = 78.6e12      # MI355X FP64 (vector and matrix), flop/s
PEAK_HBM = 8.0e12        # MI355X HBM3E, B/s
# HBM bytes per launch of each instrumented kernel group at this workload: two rocprofv3
# counter passes over one evaluation (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, the
# dispatches from the system's pair lists on), each scaled by the same counter for one
# 256 MiB device-to-device copy (FETCH_SIZE reads half the bytes: x2.00; WRITE_SIZE x1.00),
# averaged over the group's dispatches in one evaluation.
TRAFFIC = {"solve": 1325971, "xform": 2352671, "surfgrad": 13555872, "p2l": 36803948, "surface": 23240738,
           "xgrad": 161750981}


def roctx():
    """ROCm's marker library (a no-op unless a profiler records markers), or None."""
    import ctypes
    for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        lib.roctxRangePushA.restype = ctypes.c_int
        lib.roctxRangePop.restype = ctypes.c_int
        return lib
    return None


def build_system(pbsam_amd, synth, ctx):
    mols = [synth.molecule(N_SPH, seed=s) for s in SEEDS]
    mts = []
    for m in mols:
        mt = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], P, eps_in=synth.EPS_IN,
                               rcen=m["rcen"], npts=[NPTS] * N_SPH)
        mt.selfpol(tol=0, max_rounds=0)
        mt.values()
        mts.append(mt)
    pos, quat = synth.place_pair(mols[0], mols[1], DIST, seed=POSE_SEED)
    system = pbsam_amd.System(ctx, mts, pos, quat)
    system.set_cutoffs(100.0, 100.0)
    return mols, mts, pos, quat, system


def batch_leg(pbsam_amd, synth, ctx, mols, mts, nsys, reps, nplain):
    """`nsys` Brownian-dynamics snapshots of the same two molecules (centre distance uniform
    in [50, 80] A, random orientations, seeded) as one batch (pbsam_batch_create): complete
    force+torque evaluations of every snapshot, timed over `reps` batched calls after one
    warm-up.  The plain path, one pbsam_compute_forces per system, is timed on the first
    `nplain` snapshots, whose forces must equal the batch's."""
    import numpy as np
    rng = np.random.default_rng(BATCH_SEED)
    P, Q = [], []
    for b in range(nsys):
        p, q = synth.place_pair(mols[0], mols[1], rng.uniform(50.0, 80.0), seed=BATCH_SEED + 1 + b)
        P.append(p)
        Q.append(q)
    P, Q = np.vstack(P), np.vstack(Q)
    batch = pbsam_amd.System(ctx, [mts[0], mts[1]] * nsys, P, Q, groups=[2] * nsys)
    batch.set_cutoffs(100.0, 100.0)
    batch.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"batch_region")
    t0 = time.perf_counter()
    for _ in range(reps):
        _, rc, fb, tb = batch.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    el = time.perf_counter() - t0
    if mark:
        mark.roctxRangePop()
    st, rounds, passes = batch.status()
    # the plain path on a sample of the same systems
    singles = [pbsam_amd.System(ctx, [mts[0], mts[1]], P[2 * b:2 * b + 2], Q[2 * b:2 * b + 2])
               for b in range(nplain)]
    for s in singles:
        s.set_cutoffs(100.0, 100.0)
    singles[0].compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    err = 0.0
    t0 = time.perf_counter()
    for b, s in enumerate(singles):
        _, _, f1, t1 = s.compute_forces(tol=0, max_rounds=0)
        scale = max(np.abs(f1).max(), 1e-300)
        err = max(err, float(np.abs(fb[2 * b:2 * b + 2] - f1).max() / scale))
    ctx.sync()
    plain = nplain / (time.perf_counter() - t0)
    value = reps * nsys / el
    return dict(systems=nsys, workload="%d snapshots of the two %d-sphere molecules, p=%d, %d points/sphere, "
                                      "centre distance U[50,80] A, random orientations" % (nsys, N_SPH, P_ORD, NPTS),
                evals_per_s=value, ms_per_batch=1e3 * el / reps, reps=reps,
                plain_loop_evals_per_s=plain, plain_loop_sample=nplain, speedup_vs_plain_loop=value / plain,
                max_rel_err_force_batch_vs_plain=err, groups_ok=int((st == 0).sum()),
                mutual_rounds_max=int(rounds.max()), grad_passes_mean=float(passes.mean()),
                snapshots=(P, Q))


def exact_leg(pbsam_amd, ctx, mts, pos, quat, reps):
    """The exact derivative of the solved energy on the same system: its fixed point by GMRES
    (pbsam_solve_exact) and every molecule's force and torque from the adjoint
    (pbsam_adjoint_forces), both to a squared relative residual of 1e-16, timed over `reps`
    evaluations after a 1e-24 evaluation (warm-up and the reference the timed forces are
    compared with); the poses are set again before each, so every timed evaluation remakes its
    pair lists, rotations and translation matrices as at new poses.  The reference's own rounds do not converge on these molecules (nearly
    touching spheres), so no forward-mode exact evaluation exists to time beside it."""
    import numpy as np
    S = pbsam_amd.System(ctx, mts, pos, quat)
    S.set_cutoffs(100.0, 100.0)
    S.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    S.solve_exact(tol=1e-24, max_rounds=2000)
    f0, t0_, _ = S.adjoint_forces(tol=1e-24, max_rounds=2000)
    ctx.sync()
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"exact_region")
    ts = ta = 0.0
    for _ in range(reps):
        for i in range(len(mts)):  # poses set again: pair lists, rotations and matrices are remade
            S.set_pose(i, pos[i], quat[i])
        t0 = time.perf_counter()
        napp, rc = S.solve_exact(tol=1e-16, max_rounds=1000)
        t1 = time.perf_counter()
        f, t, nadj = S.adjoint_forces(tol=1e-16, max_rounds=1000)
        t2 = time.perf_counter()
        ts += t1 - t0
        ta += t2 - t1
    if mark:
        mark.roctxRangePop()
    S.close()
    return dict(evals_per_s=reps / (ts + ta), solve_s=ts / reps, adjoint_s=ta / reps, reps=reps, tol=1e-16,
                applications_solve=napp, applications_adjoint=nadj, status=rc,
                max_rel_err_force_vs_1e24=float(np.abs(f - f0).max() / np.abs(f0).max()),
                max_rel_err_torque_vs_1e24=float(np.abs(t - t0_).max() / np.abs(t0_).max()),
                workload="the same 2 x 60-sphere p=30 system: exact-mode fixed point by GMRES + adjoint forces and "
                         "torques of both molecules (-dE/dr, -dE/domega, lab frame)")


BD_SPH, BD_P, BD_NPTS, BD_SEEDS = 100, 10, 400, (301, 302)
MIX_POOL, MIX_SPH, MIX_SEED = 4, (20, 60), 6000


def mixed_leg(pbsam_amd, synth, ctx, nsys, reps, nplain):
    """Independent two-molecule systems of mixed pole orders in one batch handle (the
    mixed-order batch workload, reduced in size): a pool of MIX_POOL molecule types at p = 10
    (400 points per sphere) and as many at p = 30 (1800 points), sphere counts uniform in
    MIX_SPH; each system draws its order, then two types of that order, a centre distance
    uniform in [30, 80] A (at least the two packing radii plus 4 A) and random orientations.
    Complete force+torque evaluations of every system, timed over `reps` batched calls after
    one warm-up, beside the plain one-system-per-call loop on the first `nplain` systems."""
    import numpy as np
    rng = np.random.default_rng(MIX_SEED)
    pool = {10: [], 30: []}
    for p, npts in ((10, 400), (30, 1800)):
        for t in range(MIX_POOL):
            n = int(rng.integers(MIX_SPH[0], MIX_SPH[1] + 1))
            m = synth.molecule(n, seed=MIX_SEED + 100 * p + t)
            g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], p, eps_in=synth.EPS_IN,
                                  rcen=m["rcen"], npts=[npts] * n)
            g.selfpol(tol=0, max_rounds=0)
            g.values()
            pool[p].append((m, g))
    types, P, Q, orders = [], [], [], []
    for b in range(nsys):
        p = 10 if rng.random() < 0.5 else 30
        a, c = pool[p][rng.integers(MIX_POOL)], pool[p][rng.integers(MIX_POOL)]
        dmin = synth.ball_radius(len(a[0]["rad"])) + synth.ball_radius(len(c[0]["rad"])) + 4.0
        pp, qq = synth.place_pair(a[0], c[0], max(rng.uniform(30.0, 80.0), dmin), seed=MIX_SEED + 1000 + b)
        types.append((a[1], c[1]))
        P.append(pp)
        Q.append(qq)
        orders.append(p)
    batch = pbsam_amd.System(ctx, [g for t in types for g in t], np.vstack(P), np.vstack(Q), groups=[2] * nsys)
    batch.set_cutoffs(100.0, 100.0)
    batch.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"mixed_region")
    t0 = time.perf_counter()
    for _ in range(reps):
        _, rc, fb, tb = batch.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    el = time.perf_counter() - t0
    if mark:
        mark.roctxRangePop()
    st, _, _ = batch.status()
    singles = [pbsam_amd.System(ctx, list(types[b]), P[b], Q[b]) for b in range(nplain)]
    for s1 in singles:
        s1.set_cutoffs(100.0, 100.0)
    singles[0].compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    err = 0.0
    t0 = time.perf_counter()
    for b, s1 in enumerate(singles):
        _, _, f1, _ = s1.compute_forces(tol=0, max_rounds=0)
        err = max(err, float(np.abs(fb[2 * b:2 * b + 2] - f1).max() / max(np.abs(f1).max(), 1e-300)))
    ctx.sync()
    plain = nplain / (time.perf_counter() - t0)
    value = reps * nsys / el
    return dict(systems=nsys, systems_p10=int(sum(o == 10 for o in orders)), systems_p30=int(sum(o == 30 for o in orders)),
                workload="%d two-molecule systems, p in {10, 30} per system (400 / 1800 points per sphere), %d-%d "
                         "spheres per molecule from a pool of %d types per order, centre distance U[30,80] A, random "
                         "orientations (reduced from 40-300 spheres per molecule)" % (nsys, MIX_SPH[0], MIX_SPH[1],
                                                                                       MIX_POOL),
                evals_per_s=value, ms_per_batch=1e3 * el / reps, reps=reps, plain_loop_evals_per_s=plain,
                plain_loop_sample=nplain, speedup_vs_plain_loop=value / plain, max_rel_err_force_batch_vs_plain=err,
                groups_ok=int((st == 0).sum()))


def bd_leg(pbsam_amd, synth, ctx, kappa, ntraj, nsteps, cpu_budget):
    """Batched Brownian dynamics (pbsam_bd_*): `ntraj` trajectories of two 100-sphere molecules
    (p=10, 400 points/sphere), D_t=0.01 A^2/ps, D_r=1e-4 rad^2/ps, fixed 2 ps steps, forces
    and torques every step (f_dist beyond reach), CBDnam's cutoffs (10 A polarization, 100 A
    interaction), starting near contact (centre distance 2R+6 A + U[0,10] A, random
    orientations).  One warm-up step, then `nsteps` timed steps of every trajectory.  The
    oracle runs trajectory 0 from the same start and seed on one host core for about
    `cpu_budget` seconds; its poses after the first steps are compared with the device's.
    Then the same trajectories with every step's forces and torques from the adjoint
    (PBSAM_BD_ADJOINT, cutoffs 10/10 A), timed the same way ("adjoint")."""
    import numpy as np
    rng = np.random.default_rng(BATCH_SEED + 7)
    mols = [synth.molecule(BD_SPH, seed=s) for s in BD_SEEDS]
    mts = []
    for m in mols:
        mt = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], BD_P, eps_in=synth.EPS_IN,
                               rcen=m["rcen"], npts=[BD_NPTS] * BD_SPH)
        mt.selfpol(tol=0, max_rounds=0)
        mt.values()
        mts.append(mt)
    d0 = 2 * synth.ball_radius(BD_SPH) + 6.0
    P, Q = [], []
    for g in range(ntraj):
        p, q = synth.place_pair(mols[0], mols[1], d0 + rng.uniform(0.0, 10.0), seed=BATCH_SEED + 100 + g)
        P.append(p)
        Q.append(q)
    P, Q = np.vstack(P), np.vstack(Q)
    batch = pbsam_amd.System(ctx, [mts[0], mts[1]] * ntraj, P, Q, groups=[2] * ntraj)
    batch.set_cutoffs(10.0, 100.0)
    prm = pbsam_amd.System.bd_params(dtr=0.01, dr=(1e-4, 1e-4), dt_fixed=2.0, f_dist=1e9)
    seeds = np.arange(ntraj, dtype=np.int64) + 424242
    batch.bd_start(seeds, place=False, params=prm)
    batch.bd_run(1)
    ctx.sync()
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"bd_region")
    t0 = time.perf_counter()
    rc, st, steps = batch.bd_run(nsteps)
    ctx.sync()
    el = time.perf_counter() - t0
    if mark:
        mark.roctxRangePop()
    running = int((st == pbsam_amd.BD_RUNNING).sum())
    out = dict(trajectories=ntraj, steps_timed=nsteps, traj_steps_per_s=ntraj * nsteps / el, s_per_step=el / nsteps,
               running_after=running, status_counts=np.bincount(st, minlength=5).tolist(),
               workload="%d BD trajectories x 2 molecules x %d spheres, p=%d, %d points/sphere, D_t=0.01, "
                        "D_r=1e-4, dt=2 ps, forces every step, cutoffs 10/100 A, start 2R+6+U[0,10] A" %
                        (ntraj, BD_SPH, BD_P, BD_NPTS))
    # the same trajectories with every step's forces and torques from the adjoint
    # (PBSAM_BD_ADJOINT: the exact-mode fixed point by GMRES, then the adjoint), interactions cut
    # at the 10 A polarization cutoff (which the adjoint needs)
    # (the reference's schedule with the same cutoffs beside it: the 10-100 A interaction-only
    # pairs, most of the work of the 10/100 A run, are not part of either)
    def timed_bd(flags):
        b = pbsam_amd.System(ctx, [mts[0], mts[1]] * ntraj, P, Q, groups=[2] * ntraj)
        b.set_cutoffs(10.0, 10.0)
        b.set_flags(flags)
        b.bd_start(seeds, place=False, params=prm)
        b.bd_run(1)
        ctx.sync()
        t0 = time.perf_counter()
        rc_b, st_b, _ = b.bd_run(nsteps)
        ctx.sync()
        el_b = time.perf_counter() - t0
        b.close()
        return dict(traj_steps_per_s=ntraj * nsteps / el_b, s_per_step=el_b / nsteps, status=int(rc_b),
                    running_after=int((st_b == pbsam_amd.BD_RUNNING).sum()))

    out["adjoint"] = timed_bd(pbsam_amd.EXACT_DERIVATIVE | pbsam_amd.BD_ADJOINT)
    out["adjoint"]["workload"] = ("the same trajectories, forces and torques of every step from the adjoint (exact "
                                  "fixed point by GMRES + adjoint, squared relative residual 1e-16), cutoffs 10/10 A")
    out["reference_10_10"] = timed_bd(0)
    out["reference_10_10"]["workload"] = "the same trajectories, the reference's force evaluation, cutoffs 10/10 A"
    print("Brownian dynamics, cutoffs 10/10 A: adjoint forces %.1f, reference schedule %.1f trajectory-steps/s"
          % (out["adjoint"]["traj_steps_per_s"], out["reference_10_10"]["traj_steps_per_s"]), file=sys.stderr,
          flush=True)
    if cpu_budget > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        orc = pyoracle.Oracle(kappa)
        omts = []
        for m, g in zip(mols, mts):
            o = pyoracle.MolType(orc, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], BD_P, eps_in=synth.EPS_IN,
                                 rcen=m["rcen"], npts=[BD_NPTS] * BD_SPH)
            o.set_self(*g.FH())
            o.values()
            omts.append(o)
        osys = pyoracle.System(orc, omts, P[0:2], Q[0:2])
        osys.set_cutoffs(10.0, 100.0)
        oprm = [prm.dtr, prm.dr[0], prm.dr[1], prm.b_dist, prm.q_dist, prm.f_dist, prm.fact_tr, prm.ikbt,
                prm.coul_k, prm.dt_fixed, prm.max_tries]
        b = pyoracle.BDTrajectory([(m["cen"] - m["rcen"], m["rad"]) for m in mols], int(seeds[0]), False, P[0:2],
                                  Q[0:2], prm=oprm)
        b.attach(osys)
        err = None
        for _ in range(1 + nsteps):
            b.step()
        err = float(np.abs(b.state()["pos"] - batch.bd_get(0)["pos"]).max())
        # all host cores: one trajectory of the batch per worker thread (the trajectories are
        # independent), stepped for about cpu_budget seconds
        import threading
        workers = host_workers()
        counts = [0] * workers
        FH = [g.FH() for g in mts]
        barrier = threading.Barrier(workers)
        t_start = []

        def work(w):
            _one_omp_thread()
            o2 = pyoracle.Oracle(kappa)
            om = []
            for m, (F, H) in zip(mols, FH):
                o = pyoracle.MolType(o2, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], BD_P, eps_in=synth.EPS_IN,
                                     rcen=m["rcen"], npts=[BD_NPTS] * BD_SPH)
                o.set_self(F, H)
                o.values()
                om.append(o)
            g = w % ntraj
            s2 = pyoracle.System(o2, om, P[2 * g:2 * g + 2], Q[2 * g:2 * g + 2])
            s2.set_cutoffs(10.0, 100.0)
            tr = pyoracle.BDTrajectory([(m["cen"] - m["rcen"], m["rad"]) for m in mols], int(seeds[g]), False,
                                       P[2 * g:2 * g + 2], Q[2 * g:2 * g + 2], prm=oprm)
            tr.attach(s2)
            barrier.wait()
            if w == 0:
                t_start.append(time.perf_counter())
            while not t_start:
                time.sleep(0.001)
            while time.perf_counter() - t_start[0] < cpu_budget:
                if tr.step() != 2:
                    break
                counts[w] += 1

        threads = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        tc = time.perf_counter() - t_start[0]
        n = sum(counts)
        out["cpu_baseline"] = dict(value=n / tc, unit="trajectory-steps/s", cores=workers, kind="port",
                                   sample="trajectories 0..%d of the same batch, one per host thread, %d steps in %.1f s "
                                          "on the oracle" % (workers - 1, n, tc))
        out["max_abs_pos_diff_traj0_device_vs_oracle_A"] = err
    return out


LARGE_MOLS, LARGE_P, LARGE_NPTS, LARGE_SEED = 8, 30, 2000, 7000
EXACT_TOL = 1e-16  # squared relative residual of the exact path's GMRES solves


def device_memory():
    """(free, total) bytes of the current HIP device"""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
    hip.hipMemGetInfo(ctypes.byref(fr), ctypes.byref(tot))
    return fr.value, tot.value


def large_leg(pbsam_amd, synth, ctx, nsph, solves, evals):
    """The large-system workload: eight distinct synthetic molecules of `nsph` spheres at
    p = 30 with 2000 points per sphere, centred on a 2 x 2 x 2 grid whose spacing leaves 10 A
    between the molecules' packing balls (a 200 A cube cannot hold eight 500-sphere balls
    of radius 58.5 A: the grid spacing is 2R + 10 A), random orientations; spheres of
    different molecules closer than 20 A are polarized (the reference's 10 A would polarize
    none across 10 A gaps), interactions to 100 A.  Set-up (types, self-polarization, frozen
    values) is timed apart; then `solves` mutual solves (stock tolerance) and `evals` complete
    force+torque evaluations.  Operators beyond the context's store cap are made per use.
    Then the exact-derivative path on the same poses with interactions cut at 20 A: the fixed
    point by GMRES and the adjoint gradient of every molecule's energy with respect to all 48
    rigid-body coordinates ("exact")."""
    import numpy as np
    rng = np.random.default_rng(LARGE_SEED)
    R = synth.ball_radius(nsph)
    a = 2 * R + 10.0
    t0 = time.perf_counter()
    mols, mts = [], []
    for i in range(LARGE_MOLS):
        m = synth.molecule(nsph, seed=LARGE_SEED + 1 + i)
        g = pbsam_amd.MolType(ctx, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], LARGE_P, eps_in=synth.EPS_IN,
                              rcen=m["rcen"], npts=[LARGE_NPTS] * nsph)
        g.selfpol(tol=0, max_rounds=0)
        g.values()
        mols.append(m)
        mts.append(g)
        print("large system: molecule type %d set up at %.1f s" % (i, time.perf_counter() - t0), file=sys.stderr,
              flush=True)
    t_setup = time.perf_counter() - t0
    pos = np.array([[a * (i & 1), a * ((i >> 1) & 1), a * ((i >> 2) & 1)] for i in range(LARGE_MOLS)], dtype=float)
    quat = np.array([synth.random_quat(rng) for _ in range(LARGE_MOLS)])
    S = pbsam_amd.System(ctx, mts, pos, quat)
    S.set_cutoffs(20.0, 100.0)
    t0 = time.perf_counter()
    r0, _ = S.solve(tol=0, max_rounds=0)
    ctx.sync()
    t_first = time.perf_counter() - t0
    n_inter, n_pol = S.counts()
    t0 = time.perf_counter()
    for _ in range(solves):
        rounds, _ = S.solve(tol=0, max_rounds=0)
    ctx.sync()
    t_solve = (time.perf_counter() - t0) / max(solves, 1)
    t0 = time.perf_counter()
    _, _, f, t = S.compute_forces(tol=0, max_rounds=0)  # first evaluation: stored operators built
    ctx.sync()
    t_first_eval = time.perf_counter() - t0
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"large_region")
    t0 = time.perf_counter()
    for _ in range(evals):
        _, _, f, t = S.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    t_eval = (time.perf_counter() - t0) / max(evals, 1)
    if mark:
        mark.roctxRangePop()
    st, passes = S.status()[0], S.status()[2]
    stored, limit = ctx.stored()
    free, total = device_memory()
    S.close()
    # the exact-derivative path on the same poses, interactions cut at the polarization cutoff:
    # the fixed point by GMRES and the adjoint gradient of every molecule's energy with respect
    # to all 48 rigid-body coordinates (one transposed solve per molecule); the second of two
    # evaluations is timed (the first makes the inter pairs' rotations)
    E = pbsam_amd.System(ctx, mts, pos, quat)
    E.set_cutoffs(20.0, 20.0)
    E.set_flags(pbsam_amd.EXACT_DERIVATIVE)
    ex = {}
    for rep in range(2):
        for i in range(LARGE_MOLS):  # poses set again: pair lists, rotations and matrices are remade
            E.set_pose(i, pos[i], quat[i])
        mark = roctx() if rep else None
        if mark:
            mark.roctxRangePushA(b"large_exact_region")
        t0 = time.perf_counter()
        n_s, rc_s = E.solve_exact(tol=EXACT_TOL, max_rounds=3000)
        ctx.sync()
        t1 = time.perf_counter()
        fe, te, n_a = E.adjoint_forces(tol=EXACT_TOL, max_rounds=5000)
        ctx.sync()
        t2 = time.perf_counter()
        if mark:
            mark.roctxRangePop()
        ex = dict(evals_per_s=1.0 / (t2 - t0), solve_s=t1 - t0, adjoint_s=t2 - t1, applications_solve=n_s,
                  applications_adjoint=n_a, status=int(rc_s), tol=EXACT_TOL, interact_rcutoff_A=20.0,
                  force_norm_max=float(np.abs(fe).max()), coordinates=6 * LARGE_MOLS)
    E.close()
    print("large system, exact path: %.3f evals/s (GMRES %d applications, adjoint %d over %d solves)"
          % (ex["evals_per_s"], ex["applications_solve"], ex["applications_adjoint"], LARGE_MOLS), file=sys.stderr,
          flush=True)
    out = dict(molecules=LARGE_MOLS, spheres_per_molecule=nsph, p=LARGE_P, points_per_sphere=LARGE_NPTS,
               grid_spacing_A=a, inter_rcutoff_A=20.0, interact_rcutoff_A=100.0,
               setup_s=t_setup, first_solve_s=t_first, solves_per_s=1.0 / t_solve, mutual_rounds=rounds,
               first_eval_s=t_first_eval, evals_per_s=1.0 / t_eval, grad_passes=int(np.sum(passes)),
               inter_pairs=n_inter, polarized_spheres=n_pol, status=int(st[0]),
               force_norm_max=float(np.abs(f).max()),
               stored_GB=stored / 1e9, store_limit_GB=limit / 1e9, device_used_GB=(total - free) / 1e9,
               device_total_GB=total / 1e9, exact=ex)
    for g in mts:
        g.close()
    return out


def cpu_info():
    """host cores this process may use and the CPU model"""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), usable=usable, model=model)


def host_workers():
    """worker threads for the parallel CPU legs: the usable cores, at most 16 (the host's CPU
    share for one GPU)"""
    return max(1, min(16, cpu_info()["usable"]))


def _one_omp_thread():
    """the oracle's OpenMP regions run one thread inside each worker thread"""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)
    except OSError:
        pass


def cpu_batch_baseline(kappa, mols, mts, P, Q, budget_s, workers):
    """The oracle on independent systems of the batched workload, one system per worker
    thread at a time (the oracle releases the GIL inside its calls; one OpenMP thread per
    worker): complete force+torque evaluations finished within about `budget_s` seconds
    (at least one round of `workers` systems)."""
    import threading
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from pbsam_amd import synth
    FH = [g.FH() for g in mts]
    nsys = len(P) // 2
    done, lock = [0], threading.Lock()
    t_start = [None]
    barrier = threading.Barrier(workers)

    def work(w):
        _one_omp_thread()
        orc = pyoracle.Oracle(kappa)
        omts = []
        for m, (F, H) in zip(mols, FH):
            o = pyoracle.MolType(orc, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], mts[0].p, eps_in=synth.EPS_IN,
                                 rcen=m["rcen"], npts=[NPTS] * len(m["rad"]))
            o.set_self(F, H)
            o.values()
            omts.append(o)
        barrier.wait()
        with lock:
            if t_start[0] is None:
                t_start[0] = time.perf_counter()
        b = w
        while True:
            S = pyoracle.System(orc, omts, P[2 * (b % nsys):2 * (b % nsys) + 2], Q[2 * (b % nsys):2 * (b % nsys) + 2])
            S.set_cutoffs(100.0, 100.0)
            S.compute_forces(0, 0.0, 0)
            with lock:
                done[0] += 1
            b += workers
            if time.perf_counter() - t_start[0] >= budget_s:
                break

    threads = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    el = time.perf_counter() - t_start[0]
    return dict(value=done[0] / el, unit="force+torque evals/s", cores=workers, kind="port",
                sample="%d complete force+torque evaluations of the batch's snapshots on %d host threads in %.1f s "
                       "(one system per thread at a time)" % (done[0], workers, el))


def cpu_baseline(kappa, mols, mts, pos, quat, gpu_force, gpu_torque, budget_s):
    """The oracle (a scalar restatement of the reference, 1 thread) on the same molecules,
    self values and pose: complete force+torque evaluations (pair lists, mutual solve,
    gradient phase, forces) repeated until `budget_s` seconds have passed (at least one),
    then complete mutual solves for a few seconds.  Also returns the largest relative
    difference of the device's forces and torques from the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from pbsam_amd import synth
    orc = pyoracle.Oracle(kappa)
    omts = []
    for m, g in zip(mols, mts):
        o = pyoracle.MolType(orc, m["cen"], m["rad"], m["chg_xyz"], m["chg_q"], P, eps_in=synth.EPS_IN,
                             rcen=m["rcen"], npts=[NPTS] * N_SPH)
        o.set_self(*g.FH())
        o.values()
        omts.append(o)
    osys = pyoracle.System(orc, omts, pos, quat)
    osys.set_cutoffs(100.0, 100.0)
    reps, t0 = 0, time.perf_counter()
    while True:
        rc, f, t = osys.compute_forces(0, 0.0, 0)
        reps += 1
        tf = time.perf_counter() - t0
        if tf >= budget_s:
            break
    err_f = float(np.abs(gpu_force - f).max() / np.abs(f).max())
    err_t = float(np.abs(gpu_torque - t).max() / np.abs(t).max())
    sreps, t0 = 0, time.perf_counter()
    while True:
        osys.lists()
        osys.mutual(0, 0.0, 0)
        sreps += 1
        ts = time.perf_counter() - t0
        if ts >= min(3.0, budget_s):
            break
    return dict(value=reps / tf, unit="force+torque evals/s", cores=1, kind="port",
                sample="%d complete force+torque evaluations of the same system in %.1f s (pair lists, mutual "
                       "solve, gradient phase, forces and torques)" % (reps, tf),
                solves_per_s=sreps / ts,
                max_rel_err_force_device_vs_oracle=err_f, max_rel_err_torque_device_vs_oracle=err_t)


def copy_calibration(mib):
    """One device-to-device hipMemcpy of `mib` MiB (the HIP runtime's copy kernel), for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters: the copy reads and writes exactly
    that many bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "rc": rc}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--autotune", type=int, default=2, metavar="REPS",
                    help="tune the evaluation's launch shapes first (pbsam_autotune, REPS timed calls per setting; "
                         "0: the defaults)")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of oracle force evaluations to time (at least one runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--batch", type=int, default=256, help="systems in the batched leg (0: skip it)")
    ap.add_argument("--batch-reps", type=int, default=2, help="timed batched calls")
    ap.add_argument("--plain-sample", type=int, default=8, help="systems timed one by one for the ratio")
    ap.add_argument("--solves", type=int, default=20, help="mutual solves timed for solves_per_s")
    ap.add_argument("--mixed", type=int, default=256, help="systems in the mixed-order batch leg (0: skip it)")
    ap.add_argument("--tight-solves", type=int, default=1, help="mutual solves timed at the tight tolerance")
    ap.add_argument("--exact", type=int, default=2, metavar="REPS",
                    help="exact evaluations (GMRES solve + adjoint) timed on the same system (0: skip)")
    ap.add_argument("--bd", type=int, default=1024, help="trajectories in the Brownian-dynamics leg (0: skip it)")
    ap.add_argument("--bd-steps", type=int, default=3, help="timed steps of the Brownian-dynamics leg")
    ap.add_argument("--large", type=int, default=500, metavar="NSPH",
                    help="the large-system workload: 8 molecules of NSPH spheres at p=30 (500: full size; 0: skip)")
    ap.add_argument("--large-evals", type=int, default=1, help="timed evaluations of the large system")
    ap.add_argument("--pot-grid", type=int, default=0, metavar="N",
                    help="also time pbsam_potential on an N^3 grid around the solved system (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    faulthandler.enable()  # a crash prints the Python stack
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the solve of one system runs on one device; "
              "independent systems would shard across GPUs with no exchange.  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pbsam_amd
    from pbsam_amd import synth
    kappa = pbsam_amd.kappa_from_salt(synth.SALT_M, synth.EPS_OUT, synth.TEMP)
    ctx = pbsam_amd.Context(kappa, synth.EPS_OUT)
    t_setup = time.perf_counter()
    mols, mts, pos, quat, system = build_system(pbsam_amd, synth, ctx)
    t_setup = time.perf_counter() - t_setup
    print("set-up (self-polarization of both molecule types) %.1f s" % t_setup, file=sys.stderr, flush=True)

    tuned = None
    if args.autotune > 0:  # set-up: launch shapes for this system's shape (pbsam_autotune)
        t_tune = time.perf_counter()
        tms, tuned = system.autotune(reps=args.autotune)
        print("autotuned launch shapes in %.1f s (%.1f ms per evaluation): %s" % (time.perf_counter() - t_tune, tms,
                                                                                  tuned), file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        system.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    mark = roctx()
    if mark:
        mark.roctxRangePushA(b"timed_region")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rounds, _, force, torque = system.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    if mark:
        mark.roctxRangePop()
    ms_step = 1e3 * elapsed / args.steps
    value = args.steps / elapsed
    print("timed %d force+torque evaluations in %.3f s (mutual rounds %d)" % (args.steps, elapsed, rounds),
          file=sys.stderr, flush=True)
    # Per-kernel HIP-event timings: the same K evaluations again with an event pair around
    # every instrumented launch, on the stream it is launched on.  The event records add host
    # calls and queue packets between dependent kernels, so they stay out of the timed region.
    ctx.profile(True)
    if mark:
        mark.roctxRangePushA(b"kernel_timing_region")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        system.compute_forces(tol=0, max_rounds=0)
    ctx.sync()
    elapsed_ev = time.perf_counter() - t0
    if mark:
        mark.roctxRangePop()
    prof = {k: ctx.profile_read(k) for k in pbsam_amd.KERNELS}
    ctx.profile(False)
    # the solve alone (outside the timed region above)
    nsol = args.solves
    solves_per_s = None
    if nsol > 0:
        t0 = time.perf_counter()
        for _ in range(nsol):
            system.solve(tol=0, max_rounds=0)
        ctx.sync()
        solves_per_s = nsol / (time.perf_counter() - t0)
    # the solve converged to a tighter fixed point: mean squared relative change of the
    # coefficients <= 1e-8 (relative change ~1e-4 per round against the stock 1e-2), the
    # reference's order; the Gauss-Seidel schedule converges slowly, so a few hundred rounds
    tight = None
    if args.tight_solves > 0:
        if mark:
            mark.roctxRangePushA(b"tight_region")
        t0 = time.perf_counter()
        for _ in range(args.tight_solves):
            trounds, trc = system.solve(tol=1e-8, max_rounds=1000)
        ctx.sync()
        tight = dict(solves_per_s=args.tight_solves / (time.perf_counter() - t0), tol=1e-8, mutual_rounds=trounds,
                     status=trc)
        if mark:
            mark.roctxRangePop()

    exact = None
    if args.exact > 0:
        exact = exact_leg(pbsam_amd, ctx, mts, pos, quat, args.exact)
        print("exact derivative (GMRES solve %d applications + adjoint %d): %.2f evals/s" %
              (exact["applications_solve"], exact["applications_adjoint"], exact["evals_per_s"]), file=sys.stderr,
              flush=True)

    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    avg_s = d["ms"] * 1e-3 / max(d["launches"], 1)
    compute_bound = dom != "solve" or d["bytes"] < 16 * d["flops"] / 4
    if compute_bound:
        achieved = d["flops"] / max(d["ms"] * 1e-3, 1e-30)
        roof = dict(bound="mfma", pipe="fp64 (VALU; the FP64 matrix peak is the same 78.6 TFLOP/s)",
                    achieved=achieved / 1e12, peak=PEAK_FP64 / 1e12, unit="TFLOP/s", frac=achieved / PEAK_FP64)
    else:
        achieved = d["bytes"] / max(d["ms"] * 1e-3, 1e-30)
        roof = dict(bound="hbm", achieved=achieved / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s", frac=achieved / PEAK_HBM)
    roof.update(kernel=dom, launches=d["launches"], avg_launch_us=avg_s * 1e6,
                timing="HIP events on the launching stream over %d instrumented evaluations run after the timed "
                       "region (%.1f ms each with the events)" % (args.steps, 1e3 * elapsed_ev / args.steps),
                algorithmic_per_launch=(d["flops"] if compute_bound else d["bytes"]) / max(d["launches"], 1),
                traffic=TRAFFIC.get(dom), traffic_unit="HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                                        "copy-calibrated)")
    # the kernel group that streams HBM: the stored-map products of the gradient phase
    h = prof["p2l"]
    roof_hbm = None
    if h["launches"] and h["ms"] > 0:
        bw = h["bytes"] / (h["ms"] * 1e-3)
        roof_hbm = dict(bound="hbm", kernel="p2l (stored-map GEMV k_p2lgemv of the gradient phase, with the solve's "
                                          "few k_p2l launches)", achieved=bw / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s",
                        frac=bw / PEAK_HBM, launches=h["launches"], avg_launch_us=1e3 * h["ms"] / h["launches"],
                        algorithmic_per_launch=h["bytes"] / h["launches"], traffic=TRAFFIC.get("p2l"))
    kernels = {k: dict(launches=v["launches"], ms_per_step=v["ms"] / args.steps,
                       gflop_per_step=v["flops"] / args.steps / 1e9) for k, v in prof.items()}

    batch = None
    if args.batch > 0:
        batch = batch_leg(pbsam_amd, synth, ctx, mols, mts, args.batch, args.batch_reps,
                          min(args.plain_sample, args.batch))
        print("batch of %d systems: %.1f evals/s (plain loop %.1f evals/s)" %
              (args.batch, batch["evals_per_s"], batch["plain_loop_evals_per_s"]), file=sys.stderr, flush=True)

    mixed = None
    if args.mixed > 0:
        mixed = mixed_leg(pbsam_amd, synth, ctx, args.mixed, args.batch_reps, min(args.plain_sample, args.mixed))
        print("mixed-order batch of %d systems: %.1f evals/s (plain loop %.1f evals/s)" %
              (args.mixed, mixed["evals_per_s"], mixed["plain_loop_evals_per_s"]), file=sys.stderr, flush=True)

    potential = None
    if args.pot_grid > 0:
        # potential map of the solved two-molecule system (computePotInSpace at every grid point)
        system.solve(tol=0, max_rounds=0)
        g = np.linspace(-40.0, 90.0, args.pot_grid)
        pts = np.stack(np.meshgrid(g, g - 25.0, g - 25.0, indexing="ij"), -1).reshape(-1, 3)
        system.potential(pts[:1024])
        ctx.sync()
        t0 = time.perf_counter()
        system.potential(pts)
        ctx.sync()
        el = time.perf_counter() - t0
        potential = dict(points=len(pts), seconds=el, points_per_s=len(pts) / el,
                         grid="%d^3 points spanning 130 A around the two molecules" % args.pot_grid)

    large = None
    if args.large > 0:
        large = large_leg(pbsam_amd, synth, ctx, args.large, 1, args.large_evals)
        print("large system (8 x %d spheres): %.3f solves/s, %.3f evals/s" %
              (args.large, large["solves_per_s"], large["evals_per_s"]), file=sys.stderr, flush=True)

    bd = None
    if args.bd > 0:
        bd = bd_leg(pbsam_amd, synth, ctx, kappa, args.bd, args.bd_steps, 0.0 if args.no_cpu else 5.0)
        print("Brownian dynamics, %d trajectories: %.1f trajectory-steps/s" % (args.bd, bd["traj_steps_per_s"]),
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(kappa, mols, mts, pos, quat, force, torque, args.cpu_budget)
        cpu["host"] = cpu_info()
        if batch is not None:
            P_, Q_ = batch["snapshots"]
            batch["cpu_baseline"] = cpu_batch_baseline(kappa, mols, mts, P_, Q_, args.cpu_budget, host_workers())
            print("CPU oracle on %d threads: %.2f evals/s" % (batch["cpu_baseline"]["cores"],
                                                              batch["cpu_baseline"]["value"]), file=sys.stderr, flush=True)
    if batch is not None:
        batch.pop("snapshots", None)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "force+torque evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded molecules, no dataset)",
        "config": {
            "workload": "force+torque evaluation (pair lists, mutual polarization, gradient phase, forces), "
                        "2 molecules x 60 spheres, p=30, 1800 points/sphere, centres 50 A apart, random "
                        "orientations, reference Gauss-Seidel schedule",
            "molecules": 2, "spheres_per_molecule": N_SPH, "p": P, "points_per_sphere": NPTS,
            "separation_A": DIST, "inter_rcutoff_A": 100.0, "mutual_rounds": rounds, "launch_shapes": tuned,
        },
        "solves_per_s": solves_per_s,
        "tight_solve": tight,
        "exact": exact,
        "batch": batch,
        "mixed_batch": mixed,
        "bd": bd,
        "large": large,
        "potential": potential,
        "roofline": roof,
        "roofline_hbm": roof_hbm,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
