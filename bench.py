#!/usr/bin/env python3
"""fix eph throughput on one MI355X.

Workload (default): fcc 80^3 cells = 2,048,000 atoms, a = 3.615 A, four
species assigned equiatomic at random, 300 K, one 10 keV primary knock-on
atom at the box centre, pair lj/cut (0.4 eV, 2.3 A, 5.5 A), synthetic .beta
with rho(r) = A exp(-r/0.9) (A = 1.6/1.8/2.0/2.2, cut 4.5 A) and
beta(rho) = B rho/(1 + 0.5 rho) (B = 0.015/0.018/0.02/0.025), 1000 knots,
T_e grid 80^3 with a 3000 K Gaussian hot spot (sigma 10 A), C_e = 3e-6,
kappa = 0.1, dt = 0.1 fs, FP64, full model (friction + random + FDM).

A step is one MD step; the timed quantity is the fix's own work per step
(initial_integrate + post_force + final_integrate + end_of_step, measured
with HIP events on the fix's stream), as the reference's "Modify" time is.
value = atoms x steps / fix seconds.  The pair force, ghost update and
neighbour rebuilds run in the same loop on the device (rebuilds through the
host) and are reported separately.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "user-eph_amd"))

METRIC = "fix-eph atom-steps/s, 2M atoms FP64 (% HBM roofline); ensemble replica-steps/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 vector


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def write_inputs(tmp, cells, a, grid, species):
    import numpy as np
    from eph_amd import inputs
    beta = os.path.join(tmp, "alloy.beta")
    names = inputs.workload_beta(beta, species)
    L = cells * a
    d = L / grid
    idx = np.indices((grid, grid, grid))[::-1].reshape(3, -1).T
    c = (idx + 0.5) * d - 0.5 * L
    T = 300.0 + 2700.0 * np.exp(-(c * c).sum(1) / (2 * 10.0 ** 2))
    tin = os.path.join(tmp, "hotspot.T")
    with open(tin, "w") as f:
        f.write("# T_e grid: 3000 K Gaussian hot spot, sigma 10 A\n#\n#\n")
        f.write(f"{grid} {grid} {grid} 1\n")
        for _ in range(3):
            f.write(f"0.0 {L!r}\n")
        f.write("NULL\n")
        f.write("\n".join(f"{i} {j} {k} {t!r} 0.0 1.0 3e-06 0.1 1 0"
                          for (i, j, k), t in zip(idx.tolist(), T.tolist())))
        f.write("\n")
    return beta, names, tin


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fix_eph_hbm_counters_fcc80_cascade.txt")


def counter_traffic(natoms):
    """HBM bytes per launch of each kernel of the 2,048,000-atom cascade, from
    the committed FETCH_SIZE / WRITE_SIZE counter summary (two rocprofv3 --pmc
    runs of this benchmark, scaled by a plain copy kernel); {} when the summary
    is missing or describes another system size."""
    out = {}
    if natoms != 2048000 or not os.path.exists(TRAFFIC_FILE):
        return out
    with open(TRAFFIC_FILE) as f:
        for line in f:
            t = line.split()
            if len(t) >= 3 and t[0] == "traffic":
                out[t[1]] = float(t[2])
    return out


SWEEP_COUNTER_FILE = os.path.join(ROOT, "profiles", "fix_eph_sweep_counters_fcc80_cascade.txt")
# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD 4
# cycles (16 lanes per cycle, the FP64 rate); 2.4 GHz peak engine clock
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def counter_valu(natoms):
    """VALU instructions per wave and the VALU-active share of SIMD cycles of
    each sweep of the 2,048,000-atom cascade, from the committed SQ counter
    summary (rocprofv3 --pmc SQ_INSTS_VALU / SQ_WAVES, SQ_ACTIVE_INST_VALU);
    {} when it is missing or describes another system size."""
    out = {}
    if natoms != 2048000 or not os.path.exists(SWEEP_COUNTER_FILE):
        return out
    with open(SWEEP_COUNTER_FILE) as f:
        for line in f:
            if line.startswith("#") or ":" not in line or "VALU instructions" not in line:
                continue
            name, rest = line.split(":", 1)
            try:
                per_wave = float(rest.split(" VALU instructions")[0].rsplit(",", 1)[1])
                active = float(rest.split("VALU active ")[1].split("%")[0]) / 100.0
            except (IndexError, ValueError):
                continue
            out[name.split()[0]] = {"valu_per_wave": per_wave, "valu_active": active}
    return out


def host_cpu_info():
    """CPU model (lscpu / /proc/cpuinfo) and online core count of this host."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_baseline(args_fix, species, snap, md, dt, nsteps, threads, runs=3, first=1):
    """The oracle (C++ restatement of FixEPH) on the same system; its per-atom loops
    run on `threads` host threads (OpenMP), the deposit and the FDM solve serially
    as in the reference.  One untimed warm-up step, then `runs` timed runs of
    `nsteps` steps each from the same starting state; returns the median rate."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import eph_oracle as eo
    from eph_amd import inputs
    used = eo.set_threads(threads)

    def one_run(s0, nst, tag):
        sn = copy_snapshot(s0)
        fix = eo.Fix(args_fix, species, sn.boxlo, sn.boxhi, dt, inputs.BOLTZ, inputs.FTM2V,
                     masses=[inputs.MASS_CU] * species)
        md.lj(sn)
        t_fix = 0.0
        for s in range(first, first + nst):
            t0 = time.perf_counter()
            fix.initial_integrate(sn)
            t_fix += time.perf_counter() - t0
            md.update_ghosts(sn)
            t0 = time.perf_counter()
            fix.post_force(sn, step=s)
            fix.final_integrate(sn)
            fix.end_of_step(sn, s)
            t_fix += time.perf_counter() - t0
        log(f"cpu baseline ({used} threads) {tag}: {nst} steps in {t_fix:.2f} s")
        return t_fix

    one_run(snap, 1, "warm-up")
    times = [one_run(snap, nsteps, f"run {r + 1}/{runs}") for r in range(runs)]
    t_med = float(np.median(times))
    return snap.nlocal * nsteps / t_med, t_med, used, times


def host_path_rate(fargs, species, snap, inf, dt, nsteps):
    """The same fix driven through the host-array hooks (x, v, f handed over
    from host memory every call, as a LAMMPS adaptor without device-resident
    atoms would): PCIe copies included.  Times only the four hook calls; the
    first step warms up."""
    import eph_amd as ea
    from eph_amd import inputs
    fix = ea.FixEPH(fargs, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU] * species)
    fix.init()
    fix.set_atoms(snap)
    # what a reneighbouring step adds: atoms, ghost map and the CSR list again
    t0 = time.perf_counter()
    fix.set_atoms(snap)
    t_reneigh = time.perf_counter() - t0
    snap.f[:] = 0.0
    t_fix = 0.0
    for n in range(1, nsteps + 2):
        t0 = time.perf_counter()
        fix.initial_integrate(snap)
        t1 = time.perf_counter()
        snap.x[snap.nlocal:] = snap.x[snap.ghost_owner] + snap.ghost_shift
        snap.v[snap.nlocal:] = snap.v[snap.ghost_owner]
        t2 = time.perf_counter()
        fix.post_force(snap, step=n)
        fix.final_integrate(snap)
        fix.end_of_step(snap, n)
        t3 = time.perf_counter()
        if n > 1:
            t_fix += (t1 - t0) + (t3 - t2)
    fix.close()
    return snap.nlocal * nsteps / t_fix, t_fix, t_reneigh


ENS_REPLICAS, ENS_STEPS = 64, 300


def copy_snapshot(s):
    import copy
    return copy.deepcopy(s)


def ensemble_bench(tmp, nrep, nsteps, warmup, sequential):
    """Batched ensemble: nrep replicas of n^3 fcc cells, n ~ logU[10, 25]
    (seeded), one 1 keV PKA each, own seed and n^3 T_e grid (19 sub-steps),
    advanced together; against the same replicas run one after another."""
    import eph_amd as ea
    from eph_amd import inputs
    ENS_DT = inputs.ENSEMBLE_DT
    reps, _ = inputs.ensemble_replicas(tmp, nrep)
    t0 = time.time()
    ens = ea.Ensemble()
    for kw, fargs in reps:
        ens.add(kw, fargs, 1, ENS_DT, inputs.BOLTZ, inputs.FTM2V, [inputs.MASS_CU])
    ens.setup(first_step=1)
    size = ens.size()
    sub = ens.info(0)["substeps"]
    log(f"ensemble: {nrep} replicas, {size['atoms']} atoms, {size['cells']} cells, "
        f"{sub} sub-steps ({time.time() - t0:.1f} s)")
    # kernel forms of the two neighbour sweeps, then the launch mode (recorded
    # graph or direct launches), each the faster on a short trial; every
    # choice gives the same bits
    tuned = ens.autotune()
    trial = {}
    for g in (True, False):
        trial[g] = ens.run(max(warmup, 2), graph=g, log=False)["ms"]
    use_graph = trial[True] <= trial[False]
    w0 = time.perf_counter()
    res = ens.run(nsteps, graph=use_graph, log=True)
    wall = time.perf_counter() - w0
    out = {
        "workload": f"ens{nrep}_fcc_logU10-25_pka1keV",
        "replicas": nrep, "atoms": size["atoms"], "Te_cells": size["cells"],
        "steps": nsteps, "dt_ps": ENS_DT, "substeps": sub, "pka_eV": 1000.0,
        "launch": "graph" if use_graph else "direct",
        "autotune": tuned,
        "trial_ms_per_step": {"graph": trial[True] / max(warmup, 2),
                              "direct": trial[False] / max(warmup, 2)},
        "ms_per_step": res["ms"] / nsteps,
        "wall_ms_per_step": wall * 1e3 / nsteps,
        "rebuild_steps": int(res["rebuilt"].sum()),
        "replica_steps_per_s": nrep * nsteps / wall,
        "atom_steps_per_s": size["atoms"] * nsteps / wall,
    }
    ens.close()
    if sequential:
        tot = 0.0
        t1 = time.time()
        for kw, fargs in reps:
            kw = dict(kw)
            n = kw.pop("cells")
            md = ea.MD(n, **kw)
            inf = md.info()
            fx = ea.FixEPH(fargs, 1, inf["boxlo"], inf["boxhi"], ENS_DT, inputs.BOLTZ,
                           inputs.FTM2V, masses=[inputs.MASS_CU])
            fx.init()
            md.attach(fx)
            md.run(fx, max(warmup, 2), first_step=1)
            w0 = time.perf_counter()
            md.run(fx, nsteps, first_step=1 + max(warmup, 2))
            tot += time.perf_counter() - w0
            md.close()
            fx.close()
        seq = nrep * nsteps / tot
        out["sequential_replica_steps_per_s"] = seq
        out["speedup_vs_sequential"] = out["replica_steps_per_s"] / seq
        log(f"sequential: {tot:.2f} s for {nrep} x {nsteps} steps ({time.time() - t1:.1f} s)")
    return out


def adjoint_bench(tmp, nsteps, every):
    """Trajectory gradient: fcc 20^3 (32,000 atoms), 600 K, 2 keV PKA, 64 beta knots,
    loss = Ee + sum (T_e - T_target)^2; forward and backward timed; the adjoint is
    checked against central differences along a random direction in beta (h = 1e-6
    and 1e-8 relative) and, on the first `every` steps, along one in v0."""
    import numpy as np
    from eph_amd.adjoint_worker import build_trajectory
    out = {}
    rng = np.random.default_rng(17)

    def fd(f, x, d, h):
        return (f(x + h * d) - f(x - h * d)) / (2 * h)

    for dt in (5e-4, 2.5e-4):
        spec = dict(cells=20, dt=dt, nsteps=nsteps, every=every, T=600.0, pka_eV=2000.0,
                    target_seed=1, n_beta=64, md_seed=2020, fix_seed=5150, tmpdir=tmp)
        md, fix, adj, _ = build_trajectory(spec)
        try:
            t0 = time.perf_counter()
            loss = adj.forward()
            t_fwd = time.perf_counter() - t0
        except Exception as e:  # a trajectory that cannot be integrated at this dt
            log(f"adjoint trajectory at dt {dt}: {e}")
            out[f"dt_{dt}"] = str(e)
            adj.close()
            continue
        t0 = time.perf_counter()
        gb, gv = adj.backward()
        t_bwd = time.perf_counter() - t0
        b0 = adj.beta0()
        db = rng.standard_normal(b0.shape) * b0
        ad_b = float((gb * db).sum())
        checks = {}
        for h in (1e-6, 1e-8):
            c = fd(lambda b: adj.forward(beta=b)[0], b0, db, h)
            checks[f"beta_h{h:.0e}"] = {"adjoint": ad_b, "central_difference": c,
                                        "rel_err": abs(ad_b - c) / max(abs(c), 1e-300)}
        adj.close()
        # v0 on the first checkpoint segment
        spec_s = dict(spec, nsteps=every)
        md2, fix2, adj2, _ = build_trajectory(spec_s)
        adj2.forward()
        _, gv2 = adj2.backward()
        v0 = adj2.v0()
        dv = rng.standard_normal(v0.shape) * np.abs(v0).mean()
        ad_v = float((gv2 * dv).sum())
        c = fd(lambda v: adj2.forward(v0=v)[0], v0, dv, 1e-6)
        checks[f"v0_first_{every}_steps_h1e-06"] = {"adjoint": ad_v, "central_difference": c,
                                                     "rel_err": abs(ad_v - c) / max(abs(c), 1e-300)}
        adj2.close()
        out.update({
            "workload": "fcc20_pka2keV_adjoint", "atoms": adj.nlocal, "Te_cells": adj.ncells,
            "steps": nsteps, "checkpoint_every": every, "dt_ps": dt, "beta_knots": adj.n_beta,
            "loss": loss[0], "Ee": loss[1], "grid_term": loss[2],
            "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
            "backward_over_forward": t_bwd / t_fwd, "fd_checks": checks,
        })
        break
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=80)
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="one-thread CPU baseline steps per timed run (median of 3 runs)")
    ap.add_argument("--cpu-all-steps", type=int, default=2,
                    help="all-thread CPU baseline steps per timed run (median of 3 runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="steps of the host-array (PCIe-inclusive) rate, 0 to skip")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default sweep forms instead of timing both")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy (k_copy_probe) of MIB MiB three times, to "
                         "scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    ap.add_argument("--ensemble-replicas", type=int, default=ENS_REPLICAS,
                    help="replicas of the batched-ensemble line (0 = skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=ENS_STEPS)
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the one-replica-at-a-time comparison of the ensemble line")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the trajectory-gradient line (0 = skip it)")
    ap.add_argument("--per-step", type=int, default=0, metavar="N",
                    help="after the timed steps, N more steps one call each, with each step's "
                         "sweep times (HIP events) and candidate-list statistics")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py: this build measures one GPU; the multi-GPU path (spatial domain "
                  "decomposition with halo exchange) is not built, so --gpus above 1 is not run.",
                  file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import eph_amd as ea
    from eph_amd import inputs

    cells, a, species, grid, dt = args.cells, 3.615, 4, args.cells, 1e-4
    tmp = tempfile.mkdtemp(prefix="ephbench")
    t0 = time.time()
    md = ea.MD(cells, a=a, ntypes=species, seed=2024, T=300.0, pka_eV=1.0e4)
    inf = md.info()
    N = inf["nlocal"]
    log(f"system: {N} atoms, {inf['nghost']} ghosts, {inf['nfix'] / N:.1f} list entries/atom "
        f"({time.time() - t0:.1f} s)")
    beta, names, tin = write_inputs(tmp, cells, a, grid, species)
    fargs = inputs.fix_args(beta, names, seed=4242, flags="7", grid=(grid, grid, grid),
                            T_infile=tin, C_e=3e-6, kappa=0.1)
    fix = ea.FixEPH(fargs, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU] * species)
    fix.init()
    cpu = cpu1 = None
    snap_host = None
    if args.host_steps > 0:
        snap_host = md.snapshot()  # for the host-array (PCIe-inclusive) rate
    if not args.no_cpu_baseline:
        # timed before the device run, while the harness lists still match snap0
        log("cpu baseline ...")
        snap0 = copy_snapshot(snap_host) if snap_host is not None else md.snapshot()
        # every core this process may use: the host's online cores, bounded by
        # the OMP_NUM_THREADS share the machine gives one GPU when it is set
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = max(1, min(share, os.cpu_count() or 1) if share > 0 else (os.cpu_count() or 1))
        if args.cpu_steps > 0:
            cpu1 = cpu_baseline(fargs, species, snap0, md, dt, args.cpu_steps, 1)
        if args.cpu_all_steps > 0:
            cpu = cpu_baseline(fargs, species, snap0, md, dt, args.cpu_all_steps, threads)
    md.attach(fix)
    g = fix.grid()
    T0max = g.get("T").max()
    log(f"attached ({time.time() - t0:.1f} s)")

    step = 1
    blowup_step = None

    def advance(n, timed):
        nonlocal step, blowup_step
        res = dict(fix_ms=0.0, total_ms=0.0, pair_ms=0.0, rebuild_ms=0.0, rebuilds=0)
        if timed and blowup_step is not None and n > 0:
            # one call for all n steps: each step's kernels are queued behind the
            # previous step's, as in a production run
            r = md.run(fix, n, first_step=step)
            step += n
            return {k: res[k] + r[k] for k in res}
        for _ in range(n):
            try:
                r = md.run(fix, 1, first_step=step)
            except ea.EphError as e:
                if e.code != -7:
                    raise
                r = dict(fix_ms=float("nan"), total_ms=0.0, pair_ms=0.0, rebuild_ms=0.0, rebuilds=0)
                if blowup_step is None:
                    blowup_step = step
            for k in res:
                res[k] += r[k]
            if blowup_step is None:
                # every cell starts at >= 300 K and only gains or diffuses heat; a
                # cell clamped at 0 K (eph_fdm.h:391-394) or a non-finite value is
                # the unstable explicit stencil (one sub-step at lambda 0.255)
                T = g.get("T")
                if not np.all(np.isfinite(T)) or T.min() <= 0.0:
                    blowup_step = step
            step += 1
        return res

    advance(args.warmup, False)
    log(f"warmup done ({time.time() - t0:.1f} s)")
    # both kernel forms of each sweep timed on the live system; the faster is kept
    tuned = fix.autotune() if not args.no_autotune else None
    log(f"sweep forms: {tuned}")
    # the value: K steps as the product runs them, events only at the hook
    # boundaries (an event between two kernels holds the second back ~10 us)
    md.checkpoint_save(fix)  # kept in host memory
    step0 = step
    res = advance(args.steps, True)
    log(f"timed {args.steps} steps ({time.time() - t0:.1f} s): {res}")
    # then the same K steps again from the checkpoint, bit for bit the same
    # trajectory, with an event pair around every kernel: the per-kernel
    # durations the roofline lines divide by
    md.checkpoint_load(fix)
    step = step0
    fix.timing(True)
    res_k = advance(args.steps, True)
    rep = fix.timing_report()
    fix.timing(False)
    log(f"kernel-timed rerun of the {args.steps} steps ({time.time() - t0:.1f} s): {res_k}")
    nact = fix.get("nactive").mean()
    nlist = fix.get("nlist").mean()
    fix_s = res["fix_ms"] / 1e3
    value = N * args.steps / fix_s
    # The two neighbour sweeps against their algorithmic bytes per atom-step
    # (FP64 SoA, int32 indices, K = full-list entries per atom, each per-atom
    # array read or written once, gathered neighbour data and tables cached):
    #   density sweep = calculate_environment fused with force_prl stage 1:
    #     x 24 + type 4 + count 4 + list 4K + v 24 + mask 4 + rho 8 + w 24 = 92 + 4K
    #   force sweep = force_prl stages 2 + 3 fused:
    #     x 24 + type 4 + count 4 + list 4K + rho 8 + w 24 + xi 24 + T_e 8
    #     + f in/out 48 + f_EPH 24 + f_RNG 24 = 156 + 4K (the K entries are
    #     the list the reference loops walk; the active list is this design's)
    #   force_prl as a whole (both sweeps) = 248 + 8K.
    # design_bytes_per_atom is what these kernels move by their own layout
    # (active-list write and read, the noise and alpha record slots, tags).
    prl_ms, prl_n = rep.get("prl_forces", (float("nan"), 1))
    dw_ms, dw_n = rep.get("density_w", (float("nan"), 1))
    # FP64 operations of the reference's expressions (+ - * / each one):
    # every list entry 8 (e, r^2); an active entry 7 more in the density
    # loop (spline, sum) and 36 in the w loop; a force pair 3 + 5 + 6 per
    # spline (two with several species) + 25 friction + 25 random
    fpair = 3 + 5 + 6 * (2 if species > 1 else 1) + 50
    sweeps = {
        "k_density_w": {"bytes_per_atom": 92.0 + 4.0 * nlist,
                        "design_bytes_per_atom": 132.0 + 4.0 * nlist + 4.0 * nact,
                        "flops_per_atom": 8.0 * nlist + 43.0 * nact,
                        "avg_launch_ms": dw_ms / max(dw_n, 1)},
        "k_prl_forces": {"bytes_per_atom": 156.0 + 4.0 * nlist,
                         "design_bytes_per_atom": 200.0 + 4.0 * nact,
                         "flops_per_atom": fpair * nact,
                         "avg_launch_ms": prl_ms / max(prl_n, 1)},
    }
    for v in sweeps.values():
        v["achieved_GBs"] = v["bytes_per_atom"] * N / (v["avg_launch_ms"] * 1e-3) / 1e9
        v["frac"] = v["achieved_GBs"] / HBM_PEAK_GBS
        v["design_GBs"] = v["design_bytes_per_atom"] * N / (v["avg_launch_ms"] * 1e-3) / 1e9
        v["fp64_TFLOPs"] = v["flops_per_atom"] * N / (v["avg_launch_ms"] * 1e-3) / 1e12
        v["fp64_frac"] = v["fp64_TFLOPs"] / FP64_PEAK_TFLOPS
    both_ms = sweeps["k_density_w"]["avg_launch_ms"] + sweeps["k_prl_forces"]["avg_launch_ms"]
    prl_bytes = 248.0 + 8.0 * nlist
    force_prl = {
        "kernels": ["k_density_w", "k_prl_forces"],
        "bytes_per_atom": prl_bytes, "ms": both_ms,
        "achieved_GBs": prl_bytes * N / (both_ms * 1e-3) / 1e9,
        "target_frac": 0.6,
        "atom_steps_per_s": N / (both_ms * 1e-3),
    }
    force_prl["frac"] = force_prl["achieved_GBs"] / HBM_PEAK_GBS
    traffic = counter_traffic(N)
    valu = counter_valu(N)
    for k, v in sweeps.items():
        v["traffic_bytes_per_launch"] = traffic.get(k)
        c = valu.get(k)
        if c:
            # the launch's VALU instructions issued back to back on every SIMD:
            # the time the arithmetic alone needs, beside the measured launch
            floor_ms = ((N + 63) // 64) * c["valu_per_wave"] * VALU_CYCLES / SIMDS / CLOCK_HZ * 1e3
            v["valu"] = {"instructions_per_wave": c["valu_per_wave"],
                         "floor_ms": floor_ms, "floor_frac": floor_ms / v["avg_launch_ms"],
                         "active_frac_counter": c["valu_active"],
                         "source": os.path.relpath(SWEEP_COUNTER_FILE, ROOT)}
    dom = max(sweeps, key=lambda k: sweeps[k]["avg_launch_ms"])
    achieved = sweeps[dom]["achieved_GBs"]
    kernels = {k: {"avg_ms": v[0] / max(v[1], 1), "launches": v[1]} for k, v in rep.items()}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fix_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"fcc{cells}_alloy{species}_cascade",
            "atoms": N, "ghosts": fix.nghost, "species": species, "pka_eV": 1.0e4,
            "list_entries_per_atom": nlist, "active_pairs_per_atom": nact,
            "Te_grid": [grid] * 3, "dt_ps": dt, "flags": 7, "model": 4,
            "md_ms_per_step": res["total_ms"] / args.steps,
            "pair_ms_per_step": res["pair_ms"] / args.steps,
            "list_rebuilds": res["rebuilds"], "list_check_and_rebuild_ms": res["rebuild_ms"],
            "fdm_blowup_step": blowup_step,
            "kernels": kernels,
            "kernels_note": (f"HIP events around every kernel, over the same {args.steps} steps "
                             "rerun from an in-memory checkpoint (same trajectory); their fix "
                             f"time {res_k['fix_ms'] / args.steps:.4f} ms per step includes the "
                             "gaps those events open"),
            "sweep_forms": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get(dom),
            "traffic_source": (os.path.relpath(TRAFFIC_FILE, ROOT) if traffic.get(dom) else None),
            "bytes_per_atom": sweeps[dom]["bytes_per_atom"],
            "design_bytes_per_atom": sweeps[dom]["design_bytes_per_atom"],
            "avg_launch_ms": sweeps[dom]["avg_launch_ms"],
            "force_prl": force_prl,
            "fp64": {"achieved": sweeps[dom]["fp64_TFLOPs"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": sweeps[dom]["fp64_frac"],
                     "flops_per_atom": sweeps[dom]["flops_per_atom"]},
            "valu": sweeps[dom].get("valu"),
            "sweeps": sweeps,
        },
        "cpu_baseline": None,
    }
    hostcpu = host_cpu_info() if (cpu is not None or cpu1 is not None) else None
    if cpu is not None:
        v_cpu, t_cpu, used, times = cpu
        out["cpu_baseline"] = {"value": v_cpu, "unit": "atom-steps/s", "cores": used, "kind": "port",
                               "host_cpu": hostcpu,
                               "run_seconds": times,
                               "sample": f"median of 3 runs of {args.cpu_all_steps} fix steps of the "
                                         f"same {N}-atom system after one warm-up step "
                                         f"({t_cpu:.2f} s per run); the oracle's per-atom loops on "
                                         f"{used} host threads, deposit and FDM serial"}
    if cpu1 is not None:
        v1, t1, _, times1 = cpu1
        out["cpu_baseline_1core"] = {"value": v1, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                                     "host_cpu": hostcpu, "run_seconds": times1,
                                     "sample": f"median of 3 runs of {args.cpu_steps} fix steps "
                                               f"after one warm-up step ({t1:.2f} s per run), "
                                               f"one thread"}
    if args.calibrate_copy > 0:
        # plain copy kernel (k_copy_probe) of a known size, in this same process
        nbytes = args.calibrate_copy * (1 << 20)
        ms = ea.copy_probe(nbytes, 3)
        out["calibration_copy"] = {"bytes_read": nbytes, "bytes_written": nbytes, "avg_ms": ms,
                                   "GBs": 2 * nbytes / (ms * 1e-3) / 1e9}
    if snap_host is not None:
        rate, t_h, t_rn = host_path_rate(fargs, species, snap_host, inf, dt, args.host_steps)
        out["host_array_path"] = {
            "atom_steps_per_s": rate, "steps": args.host_steps, "seconds": t_h,
            "set_atoms_and_list_ms": t_rn * 1e3,
            "note": "x, v, f in host memory, copied over PCIe at every hook (eph_post_force etc.); "
                    "not the value, which keeps the atoms resident in HBM"}
        log(f"host-array path: {rate:.3e} atom-steps/s")
        snap_host = None
    per_step = []
    for _ in range(args.per_step):
        # one step per call: the sweeps of this step alone, and which waves
        # walked their candidates (the rest scanned their full lists)
        fix.timing(True)
        try:
            r = advance(1, True)
        except ea.EphError as e:  # e.g. the unstable stencil's grid turned non-finite
            fix.timing(False)
            per_step.append({"step": step, "stopped": str(e)})
            break
        rp = fix.timing_report()
        fix.timing(False)
        cs = fix.candidate_stats()
        per_step.append({"step": step - 1, "rebuilt": bool(r["rebuilds"]),
                         "density_w_ms": rp.get("density_w", (float("nan"), 0))[0],
                         "prl_forces_ms": rp.get("prl_forces", (float("nan"), 0))[0],
                         "full_list_waves": cs["full_waves"], "waves": cs["waves"],
                         "moved_atoms": cs["moved"], "candidates_retaken": cs["retaken"]})
        log(f"step {step - 1}: {per_step[-1]}")
    if per_step:
        out["per_step"] = per_step
    if args.ensemble_replicas > 0:
        fix.close()
        md.close()
        try:
            out["ensemble"] = ensemble_bench(tmp, args.ensemble_replicas, args.ensemble_steps,
                                             args.warmup, not args.no_sequential)
        except ea.EphError as e:
            out["ensemble"] = {"error": str(e)}
    if args.adjoint_steps > 0:
        out["adjoint"] = adjoint_bench(tmp, args.adjoint_steps, 20)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
