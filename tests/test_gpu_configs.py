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
C4 = dict(cells=20, dt=2.5e-4, T=600.0, pka_eV=2000.0, target_seed=1, n_beta=64, md_seed=2020,
          fix_seed=5150, every=20)
H = 1e-6  # relative step of the central differences
RC, RC_LJ = 4.5, 5.5


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


@pytest.fixture(scope="module")
def c4(tmp_path_factory):
    from eph_amd.adjoint_worker import build_trajectory
    spec = dict(C4, nsteps=40, tmpdir=str(tmp_path_factory.mktemp("c4")))
    print("adjoint: fcc 20^3, 2 keV, 40 steps in two checkpoint segments", flush=True)
    md, fix, adj, meta = build_trajectory(spec)
    loss = adj.forward()
    gb, gv = adj.backward()
    yield dict(spec=spec, meta=meta, adj=adj, loss=loss, gb=gb, gv=gv)
    adj.close()


def test_config4_gradients_match_device_central_differences(c4):
    """dL/dbeta and dL/dv0 along random directions against central
    differences of the device run at h = 1e-6, over two checkpoint segments"""
    adj, L0 = c4["adj"], c4["loss"][0]
    rng = np.random.default_rng(41)
    b0 = adj.beta0()
    d = rng.standard_normal(b0.shape) * b0
    fd = _central(lambda b: adj.forward(beta=b)[0], b0, d, H)
    ad = float((c4["gb"] * d).sum())
    print(f"adjoint beta: adjoint {ad!r} central difference {fd!r}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-9 * abs(L0), (fd, ad)
    v0 = adj.v0()
    dv = rng.standard_normal(v0.shape) * np.abs(v0).mean()
    fdv = _central(lambda v: adj.forward(v0=v)[0], v0, dv, H)
    adv = float((c4["gv"] * dv).sum())
    print(f"adjoint v0: adjoint {adv!r} central difference {fdv!r}", flush=True)
    assert abs(fdv - adv) <= 1e-5 * abs(adv) + 1e-9 * abs(L0), (fdv, adv)
    adj.forward()


def test_config4_beta_gradient_matches_oracle_central_differences(c4, oracle, tmp_path):
    """the CPU reference gradient: central differences of the oracle run
    (host harness) along the same kind of direction"""
    spec, meta, adj = c4["spec"], c4["meta"], c4["adj"]
    b0 = adj.beta0()
    d = np.random.default_rng(43).standard_normal(b0.shape) * b0
    kw = dict(cells=spec["cells"], ntypes=1, seed=spec["md_seed"], T=spec["T"],
              pka_eV=spec["pka_eV"])

    def oracle_loss(beta, tag):
        args = _beta_file(meta, beta, str(tmp_path / f"b{tag}.beta"))
        s, ora = oracle_run(oracle, ea, kw, args, 1, spec["dt"], spec["nsteps"])
        T = ora.fdm().get("T")
        return ora.compute_vector(0) + sum(((T - meta["target"]) ** 2).tolist())

    L0 = oracle_loss(b0, "0")
    assert abs(L0 - c4["loss"][0]) <= 1e-10 * abs(L0)  # the oracle run is the device run
    fd = (oracle_loss(b0 + H * d, "p") - oracle_loss(b0 - H * d, "m")) / (2 * H)
    ad = float((c4["gb"] * d).sum())
    print(f"adjoint beta vs oracle: adjoint {ad!r} oracle central difference {fd!r}", flush=True)
    assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-9 * abs(L0), (fd, ad)


def _stepwise_counts(spec, meta, beta, path, nsteps):
    """The device run with these beta knots, one step at a time (host-built
    lists); per step and local atom the neighbours inside r_c (the density
    and friction sums) and inside the pair cutoff (lj/cut)."""
    args = _beta_file(meta, beta, path)
    md = ea.MD(spec["cells"], ntypes=1, seed=spec["md_seed"], T=spec["T"], pka_eV=spec["pka_eV"])
    md.set_device_rebuild(False)
    inf = md.info()
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
            out.append((np.bincount(i, weights=r2 < RC * RC, minlength=s.nlocal),
                        np.bincount(i, weights=r2 < RC_LJ * RC_LJ, minlength=s.nlocal)))
    finally:
        fix.close()
        md.close()
    return out


def test_config4_central_differences_break_at_the_first_cutoff_crossing(c4, tmp_path):
    """Over the full 200 steps the central differences at h = 1e-6 stop
    agreeing with the adjoint.  The model is not smooth there: rho(r) is cut
    at r_c with rho(r_c) != 0 and lj/cut is not shifted, so a pair that
    crosses r_c or the pair cutoff in one of the two runs and not the other
    puts a jump into the difference.  The runs at beta +- h d are stepped
    side by side to find the first step at which any atom's count of
    neighbours inside either cutoff differs between them; on the trajectory
    that ends just before it, adjoint and central differences agree."""
    from eph_amd.adjoint_worker import build_trajectory
    spec, meta = c4["spec"], c4["meta"]
    b0 = c4["adj"].beta0()
    d = np.random.default_rng(47).standard_normal(b0.shape) * b0
    nfull = 200
    print("adjoint: stepping the +h and -h runs to find the first cutoff crossing", flush=True)
    plus = _stepwise_counts(spec, meta, b0 + H * d, str(tmp_path / "p.beta"), nfull)
    minus = _stepwise_counts(spec, meta, b0 - H * d, str(tmp_path / "m.beta"), nfull)
    first = None
    for n, ((rp, lp), (rm, lm)) in enumerate(zip(plus, minus), start=1):
        if not (np.array_equal(rp, rm) and np.array_equal(lp, lm)):
            first = n
            break
    print(f"adjoint: first step with a crossing in only one run: {first}", flush=True)
    assert first is not None, "no cutoff crossing differs between the runs in 200 steps"
    assert first > 2, f"crossing already at step {first}: no prefix to check"

    def check(nsteps):
        sp = dict(spec, nsteps=nsteps, tmpdir=str(tmp_path / f"t{nsteps}"))
        os.makedirs(sp["tmpdir"], exist_ok=True)
        _, _, adj, _ = build_trajectory(sp)
        try:
            L = adj.forward()[0]
            gb, _ = adj.backward()
            ad = float((gb * d).sum())
            fd = _central(lambda b: adj.forward(beta=b)[0], b0, d, H)
        finally:
            adj.close()
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        print(f"adjoint: {nsteps} steps: adjoint {ad!r} central difference {fd!r} rel {rel:.3e}",
              flush=True)
        return rel, ad, L

    rel, ad, L = check(first - 1)
    assert rel <= 1e-5 or abs(ad) <= 1e-9 * abs(L), rel
