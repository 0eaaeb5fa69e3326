"""GPU: the stated workloads at their own sizes against the oracle.

  ensemble  64 replicas of fcc n^3 cells, n = round(exp U[ln 10, ln 25]),
            1 keV primary atoms, per-replica T_e grids, graph-replayed steps:
            the smallest, median and largest replicas bit for bit against the
            oracle driven through the host harness at the logged rebuild steps
  adjoint   fcc 20^3 (32,000 atoms), 600 K, 2 keV primary atom, 64 beta knots,
            checkpoints every 20 steps: gradients against central differences
            of the device run and of the oracle at h = 1e-6
"""
import os

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
from drive import oracle_run

pytestmark = pytest.mark.gpu


def test_ensemble64_replicas_match_oracle(oracle, tmp_path):
    """fcc n^3, n ~ logU[10, 25] (seed 3), 64 replicas (1.3 M atoms), 12
    steps replayed from the recorded step graph; three replicas against the
    oracle."""
    reps, ns = inputs.ensemble_replicas(str(tmp_path), 64)
    print(f"ensemble64: {int((4 * ns ** 3).sum())} atoms, sizes {ns.min()}..{ns.max()}", flush=True)
    dt = inputs.ENSEMBLE_DT
    ens = ea.Ensemble()
    try:
        for kw, fargs in reps:
            ens.add(kw, fargs, 1, dt, inputs.BOLTZ, inputs.FTM2V, [inputs.MASS_CU])
        ens.setup(first_step=1)
        nsteps = 12
        res = ens.run(nsteps, graph=True)
        rebuilt = res["rebuilt"]
        assert rebuilt.sum() >= 1, "the primary atoms force at least one rebuild"
        order = np.argsort(ns, kind="stable")
        picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
        for r in picks:
            kw, fargs = reps[r]
            print(f"ensemble64: replica {r} (n = {ns[r]}) against the oracle", flush=True)
            s, ora = oracle_run(oracle, ea, kw, fargs, 1, dt, nsteps, rebuilt=rebuilt)
            n = s.nlocal
            assert np.array_equal(ens.get(r, "x"), s.x[:n]), f"replica {r}: x"
            assert np.array_equal(ens.get(r, "v"), s.v[:n]), f"replica {r}: v"
            assert np.array_equal(ens.get(r, "T"), ora.fdm().get("T")), f"replica {r}: T_e"
            E = ens.get(r, "E")
            assert abs(E[1] - ora.compute_vector(0)) <= 1e-10 * abs(ora.compute_vector(0)), r
    finally:
        ens.close()


# ------------------------------------------------------------------ adjoint
# The loss is a function of beta and v0 that is smooth only between the
# steps at which a pair crosses r_c (rho(r_c) != 0 cuts the density, friction
# and noise sums), a pair crosses the lj/cut cutoff (the pair force is not
# shifted), or an atom changes T_e cell (deposit and sqrt(T_e) lookup): fcc
# sites lie on the cell faces of this grid, so thermal motion makes the last
# frequent.  Central differences at h = 1e-6 agree with the adjoint on the
# steps before the first such event that differs between the +h and -h runs.
C4 = dict(cells=20, dt=2.5e-4, T=600.0, pka_eV=2000.0, target_seed=1, n_beta=64, md_seed=2020,
          fix_seed=5150, every=20)
H = 1e-6  # relative step of the central differences
RC, RC_LJ = 4.5, 5.5
NSCAN = 80  # steps searched for the first differing event


def _central(f, x, d, h):
    return (f(x + h * d) - f(x - h * d)) / (2.0 * h)


def _beta_file(meta, beta, path):
    """the trajectory's .beta with other beta knots (rho(r) lines unchanged)"""
    header = open(meta["beta_file"]).read().splitlines()
    n_rho = int(header[4].split()[0])
    lines = header[:6 + n_rho] + [repr(float(v)) for v in beta[0]]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    args = list(meta["args"])
    args[16] = path
    return args


def _events(spec, meta, args, v0, nsteps):
    """The device run (host-built lists), one step at a time; per step and
    local atom: neighbours inside r_c, neighbours inside the pair cutoff and
    the T_e cell (eph_fdm.h:494-509)."""
    md = ea.MD(spec["cells"], ntypes=1, seed=spec["md_seed"], T=spec["T"], pka_eV=spec["pka_eV"])
    md.set_device_rebuild(False)
    if v0 is not None:
        s = md.snapshot()
        s.v[:s.nlocal] = v0
        md.rebuild(s)
    inf = md.info()
    lo, hi = np.array(inf["boxlo"]), np.array(inf["boxhi"])
    g = spec["cells"]
    dx = (hi - lo) / g
    fix = ea.FixEPH(args, 1, inf["boxlo"], inf["boxhi"], spec["dt"], inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU])
    fix.init()
    md.attach(fix)
    out = []
    try:
        for n in range(1, nsteps + 1):
            md.run(fix, 1, first_step=n)
            s = md.snapshot()
            x, _, _ = fix.dev_atoms()
            i = np.repeat(np.arange(s.nlocal), np.diff(s.nbr_off))
            e = x[s.nbr_idx] - x[i]
            r2 = (e * e).sum(1)
            c = np.floor((x[:s.nlocal] - lo) / dx).astype(np.int64) % g
            out.append((np.bincount(i, weights=r2 < RC * RC, minlength=s.nlocal),
                        np.bincount(i, weights=r2 < RC_LJ * RC_LJ, minlength=s.nlocal),
                        c[:, 0] + g * (c[:, 1] + g * c[:, 2])))
    finally:
        fix.close()
        md.close()
    return out


def _first_event(plus, minus):
    for n, (p, m) in enumerate(zip(plus, minus), start=1):
        for kind, a, b in zip(("r_c", "pair cutoff", "T_e cell"), p, m):
            if not np.array_equal(a, b):
                return n, kind, int(np.count_nonzero(a != b))
    return None, None, 0


def _adjoint(spec, nsteps, tmp):
    from eph_amd.adjoint_worker import build_trajectory
    # at least two checkpoint segments on any length
    sp = dict(spec, nsteps=nsteps, every=max(1, min(spec["every"], (nsteps + 1) // 2)), tmpdir=tmp)
    os.makedirs(tmp, exist_ok=True)
    md, fix, adj, meta = build_trajectory(sp)
    return adj, meta, sp


def test_config4_beta_gradient_before_the_first_event(oracle, tmp_path, parity_report):
    """dL/dbeta of fcc 20^3, 2 keV: central differences at h = 1e-6 of the
    device run and of the oracle run (the CPU reference gradient, host
    harness) equal the adjoint on the trajectory that ends just before the
    first step at which the +h and -h runs differ in a cutoff or cell
    crossing; on the 40-step trajectory (two segments of 20) a step of
    h = 1e-8, too small to move any crossing, agrees too."""
    spec = dict(C4)
    adj, meta, _ = _adjoint(spec, 40, str(tmp_path / "a40"))
    try:
        b0 = adj.beta0()
        d = np.random.default_rng(47).standard_normal(b0.shape) * b0
        L = adj.forward()[0]
        gb, _ = adj.backward()
        ad = float((gb * d).sum())
        fd6 = _central(lambda b: adj.forward(beta=b)[0], b0, d, H)
        fd8 = _central(lambda b: adj.forward(beta=b)[0], b0, d, 1e-8)
        print(f"adjoint beta, 40 steps: adjoint {ad!r}, central differences h=1e-6 {fd6!r}, "
              f"h=1e-8 {fd8!r}", flush=True)
        assert abs(fd8 - ad) <= 1e-4 * abs(ad), (fd8, ad)
    finally:
        adj.close()
    plus = _events(spec, meta, _beta_file(meta, b0 + H * d, str(tmp_path / "p.beta")), None, NSCAN)
    minus = _events(spec, meta, _beta_file(meta, b0 - H * d, str(tmp_path / "m.beta")), None, NSCAN)
    first, kind, natoms = _first_event(plus, minus)
    print(f"adjoint beta: first differing event at step {first} ({kind}, {natoms} atoms)", flush=True)
    m = (first - 1) if first is not None else NSCAN
    assert m >= 2, f"an event already at step {first}"
    adj, _, sp = _adjoint(spec, m, str(tmp_path / "prefix"))
    try:
        adj.forward()
        gb, _ = adj.backward()
        ad = float((gb * d).sum())
        fd = _central(lambda b: adj.forward(beta=b)[0], b0, d, H)
    finally:
        adj.close()
    print(f"adjoint beta, {m} steps (checkpoints every {sp['every']}): adjoint {ad!r}, central "
          f"difference {fd!r}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    kw = dict(cells=spec["cells"], ntypes=1, seed=spec["md_seed"], T=spec["T"],
              pka_eV=spec["pka_eV"])

    def oracle_loss(beta, tag):
        args = _beta_file(meta, beta, str(tmp_path / f"o{tag}.beta"))
        s, ora = oracle_run(oracle, ea, kw, args, 1, spec["dt"], m)
        T = ora.fdm().get("T")
        return ora.compute_vector(0) + sum(((T - meta["target"]) ** 2).tolist())

    fdo = (oracle_loss(b0 + H * d, "p") - oracle_loss(b0 - H * d, "m")) / (2 * H)
    print(f"adjoint beta, {m} steps: oracle central difference {fdo!r}", flush=True)
    assert abs(fdo - ad) <= 1e-5 * abs(ad), (fdo, ad)
    parity_report.append(
        f"fcc20 adjoint dL/dbeta: {m} steps (no differing crossing in {NSCAN} at h=1e-6: first "
        f"{first}), checkpoints every {sp['every']}: rel err vs device central differences "
        f"{abs(fd - ad) / abs(ad):.1e}, vs oracle central differences {abs(fdo - ad) / abs(ad):.1e}")


def test_config4_v0_gradient_before_the_first_event(tmp_path, parity_report):
    """the same for dL/dv0 along a random direction"""
    spec = dict(C4)
    adj, meta, _ = _adjoint(spec, 2, str(tmp_path / "probe"))
    v0 = adj.v0()
    adj.close()
    dv = np.random.default_rng(53).standard_normal(v0.shape) * np.abs(v0).mean()
    args = meta["args"]
    plus = _events(spec, meta, args, v0 + H * dv, NSCAN)
    minus = _events(spec, meta, args, v0 - H * dv, NSCAN)
    first, kind, natoms = _first_event(plus, minus)
    print(f"adjoint v0: first differing event at step {first} ({kind}, {natoms} atoms)", flush=True)
    m = (first - 1) if first is not None else NSCAN
    assert m >= 2, f"an event already at step {first}"
    adj, _, sp = _adjoint(spec, m, str(tmp_path / "prefix"))
    try:
        adj.forward()
        _, gv = adj.backward()
        ad = float((gv * dv).sum())
        fd = _central(lambda v: adj.forward(v0=v)[0], v0, dv, H)
    finally:
        adj.close()
    print(f"adjoint v0, {m} steps (checkpoints every {sp['every']}): adjoint {ad!r}, central "
          f"difference {fd!r}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    parity_report.append(
        f"fcc20 adjoint dL/dv0: first crossing differing between the +-h runs at step {first} "
        f"({kind}, {natoms} atoms); on the {m} steps before it (checkpoints every {sp['every']}) "
        f"rel err vs device central differences {abs(fd - ad) / abs(ad):.1e}")


# ------------------------------------------------------- plain method runs
def _errs(p):
    A, B = p.sA, p.sB
    n = A.nlocal
    nall = n + A.nghost
    from drive import rel_err
    return {
        "rho": rel_err(p.gpu.get("rho")[:n], p.ora.get("rho", nall)[:n]),
        "f_EPH": rel_err(p.gpu.get("f_EPH"), p.ora.get("f_EPH", n)),
        "f_RNG": rel_err(p.gpu.get("f_RNG"), p.ora.get("f_RNG", n)),
        "x": rel_err(B.x[:n], A.x[:n]),
        "T_e": rel_err(p.gpu.grid().get("T"), p.ora.fdm().get("T")),
    }


def test_fcc40_200_steps_injected_noise(oracle, tmp_path, parity_report):
    """fcc 40^3 (256,000 atoms), 1000 K with 0.05 A lattice displacements,
    40^3 T_e grid, dt 1 fs, 200 steps, Philox noise injected identically into
    both fixes: rho, forces and T_e against the oracle every 25 steps."""
    from drive import Pair
    print("fcc40: 256,000 atoms, 200 steps", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=40, T=1000.0, disp=0.05, grid=40, dt=1e-3,
             inject=True)
    worst = {}
    for k in range(200):
        p.step()
        if k % 25 == 24:
            e = _errs(p)
            for q, v in e.items():
                worst[q] = max(worst.get(q, 0.0), v)
            assert all(v <= 1e-11 for v in e.values()), (k + 1, e)
            print(f"fcc40 step {k + 1}: {e}", flush=True)
    eg, eo = p.gpu.compute_vector(0), p.ora.compute_vector(0)
    tally = abs(eg - eo) / abs(eo)
    assert tally <= 1e-10
    parity_report.append("fcc40 (256,000 atoms, 200 steps, injected noise): max rel err "
                         + ", ".join(f"{q} {v:.1e}" for q, v in worst.items())
                         + f"; energy tally {tally:.1e}")


def test_cascade80_20_steps_t_e_until_unstable(oracle, tmp_path, parity_report):
    """The 2,048,000-atom four-species cascade (10 keV primary atom, 80^3 grid
    with a 3000 K hot spot, dt 0.1 fs) for 20 steps against the oracle.  Its
    FDM step is unstable (one sub-step at lambda = 0.255 > 1/6): per step the
    T_e error, the relative amplitude of the checkerboard mode the stencil
    amplifies, the first step with a cell clamped at 0 K and the first
    non-finite value are recorded.  The T_e error stays zero throughout."""
    from drive import Pair
    print("cascade80: 2,048,000 atoms, 20 steps", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=80, species=4, T=300.0, pka_eV=1.0e4, grid=80,
             dt=1e-4, hot_spot=(3000.0, 10.0))
    n = 80
    idx = np.indices((n, n, n))[::-1].reshape(3, -1).T
    sign = np.where(idx.sum(1) % 2 == 0, 1.0, -1.0)
    clamped = nonfinite = above = None
    rows = []
    for k in range(1, 21):
        p.step()
        Tg, To = p.gpu.grid().get("T"), p.ora.fdm().get("T")
        fin = bool(np.all(np.isfinite(To)))
        err = float(np.abs(Tg - To).max() / np.abs(To).max()) if fin else float("nan")
        mode = float(abs((To * sign).mean()) / To.mean()) if fin else float("nan")
        if clamped is None and fin and To.min() <= 0.0:
            clamped = k
        if nonfinite is None and not fin:
            nonfinite = k
        if above is None and fin and mode > 1e-6:
            above = k
        rows.append((k, err, mode))
        print(f"cascade80 step {k}: T_e rel err {err:.1e}, checkerboard amplitude {mode:.2e}",
              flush=True)
        assert np.array_equal(Tg, To, equal_nan=True), f"T_e differs at step {k}"
        if k in (1, 10, 20):
            e = _errs(p)
            assert all(v <= 1e-11 for q, v in e.items() if q != "T_e"), (k, e)
    parity_report.append(
        "cascade80 (2,048,000 atoms, 20 steps): T_e rel err "
        + ", ".join(f"{r[1]:.0e}" for r in rows)
        + " (bit-identical every step); checkerboard mode of the unstable stencil "
        + f"{rows[0][2]:.1e} of the mean after step 1 (above 1e-6 from step {above}), "
        + f"{rows[-1][2]:.1e} after step 20; first cell clamped at 0 K: step {clamped}; "
        + f"first non-finite: {nonfinite}")
