"""GPU: batched ensembles against separate runs of each replica.

Every replica of an ensemble must follow, bit for bit, the run it would have
alone on the device path (harness + fix hooks), given the same neighbour
rebuild steps: the ensemble rebuilds all replicas together and logs the
steps, and the separate run is forced to rebuild at exactly those steps.
"""
import os

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
from drive import workload_files

pytestmark = pytest.mark.gpu

A = 3.615


def _replicas(tmp, specs, species, dt, flags="7", steps=1):
    """(md keywords, fix args) per replica; all share one .beta file."""
    out = []
    for k, sp in enumerate(specs):
        n = sp["cells"]
        beta, names, tin = workload_files(tmp, f"r{k}", species, n, A, n, hot_spot=(2000.0, 4.0),
                                          steps=steps)
        if k > 0:
            beta = out[0][1][16]  # the first replica's .beta file
        args = inputs.fix_args(beta, names, seed=sp["fix_seed"], flags=flags, grid=(n, n, n),
                               T_infile=tin)
        kw = dict(cells=n, a=A, ntypes=species, seed=sp["seed"], T=300.0, pka_eV=sp["pka"])
        out.append((kw, args))
    return out


def _ensemble(reps, species, dt):
    ens = ea.Ensemble()
    for kw, args in reps:
        ens.add(kw, args, species, dt, inputs.BOLTZ, inputs.FTM2V, [inputs.MASS_CU] * species)
    ens.setup(first_step=1)
    return ens


def _separate(kw, args, species, dt, rebuilt):
    """the replica alone: device-resident harness run, rebuilt at the logged steps"""
    kw = dict(kw)
    cells = kw.pop("cells")
    md = ea.MD(cells, **kw)
    inf = md.info()
    fix = ea.FixEPH(args, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU] * species)
    fix.init()
    md.attach(fix)
    for k, r in enumerate(rebuilt):
        if r:
            md.force_rebuild()
        md.run(fix, 1, first_step=1 + k)
    return md, fix


SPECS = [dict(cells=4, seed=11, fix_seed=101, pka=150.0),
         dict(cells=5, seed=12, fix_seed=102, pka=200.0),
         dict(cells=6, seed=13, fix_seed=103, pka=0.0),
         dict(cells=7, seed=14, fix_seed=104, pka=250.0)]


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_replicas_match_separate_runs(tmp_path, graph):
    species, dt, nsteps = 2, 5e-4, 30
    reps = _replicas(str(tmp_path), SPECS, species, dt)
    ens = _ensemble(reps, species, dt)
    res = ens.run(nsteps, graph=graph)
    rebuilt = res["rebuilt"]
    assert rebuilt.sum() >= 2, rebuilt
    for r, (kw, args) in enumerate(reps):
        md, fix = _separate(kw, args, species, dt, rebuilt)
        n = ens.info(r)["nlocal"]
        xs, vs, _ = fix.dev_atoms()
        assert np.array_equal(ens.get(r, "x"), xs[:n]), f"replica {r}: x"
        assert np.array_equal(ens.get(r, "v"), vs[:n]), f"replica {r}: v"
        assert np.array_equal(ens.get(r, "T"), fix.grid().get("T")), f"replica {r}: T_e"
        E = ens.get(r, "E")
        assert E[1] == fix.compute_vector(0), f"replica {r}: Ee {E[1]!r} vs {fix.compute_vector(0)!r}"
        assert E[0] == fix.last_energy(), f"replica {r}: E_local"


def test_ensemble_graph_replay_equals_direct_launches(tmp_path):
    """the recorded step, replayed, gives the launches' results (several calls)"""
    species, dt = 2, 5e-4
    reps = _replicas(str(tmp_path), SPECS[:3], species, dt)
    a = _ensemble(reps, species, dt)
    b = _ensemble(reps, species, dt)
    for n in (7, 1, 12):
        ra = a.run(n, graph=True)
        rb = b.run(n, graph=False)
        assert np.array_equal(ra["rebuilt"], rb["rebuilt"])
        for r in range(3):
            for q in ("x", "v", "T", "E"):
                assert np.array_equal(a.get(r, q), b.get(r, q)), (n, r, q)


def test_ensemble_fdm_substeps_differ_per_replica(tmp_path):
    """replicas with different sub-step counts (grid spacing differs) keep
    their own count: the ones with fewer sub-steps copy in the extra launches"""
    species, dt = 1, 1e-3
    specs = [dict(cells=4, seed=21, fix_seed=201, pka=0.0),
             dict(cells=6, seed=22, fix_seed=202, pka=0.0)]
    reps = []
    for k, sp in enumerate(specs):
        n = sp["cells"]
        g = 2 * n if k == 0 else n  # finer grid on replica 0 -> more sub-steps
        beta, names, tin = workload_files(str(tmp_path), f"s{k}", species, n, A, g, hot_spot=(2000.0, 4.0))
        if k > 0:
            beta = reps[0][1][16]
        args = inputs.fix_args(beta, names, seed=sp["fix_seed"], grid=(g, g, g), T_infile=tin)
        reps.append((dict(cells=n, a=A, ntypes=species, seed=sp["seed"], T=300.0, pka_eV=0.0), args))
    ens = _ensemble(reps, species, dt)
    s0, s1 = ens.info(0)["substeps"], ens.info(1)["substeps"]
    assert s0 != s1, (s0, s1)
    res = ens.run(5, graph=True)
    for r, (kw, args) in enumerate(reps):
        md, fix = _separate(kw, args, species, dt, res["rebuilt"])
        assert np.array_equal(ens.get(r, "T"), fix.grid().get("T")), r


def test_ensemble_sweep_forms_give_the_same_bits(tmp_path):
    """every form of the sweeps (chunk per block at any occupancy, persistent
    LDS-table blocks of 1024, 512 and 256 threads), mixed, and the autotuner's
    choice, all give the same trajectory"""
    species, dt = 2, 5e-4
    reps = _replicas(str(tmp_path), SPECS, species, dt)
    names = ea.Ensemble._FORM_NAMES
    runs = []
    for forms in [(f, f) for f in names] + [("chunk2", "lds256"), None]:
        print("ensemble forms:", forms, flush=True)
        ens = _ensemble(reps, species, dt)
        if forms is None:
            tuned = ens.autotune()
            assert set(tuned["density_ms"]) == set(names) == set(tuned["force_ms"]), tuned
            assert tuned["density"] in names and tuned["force"] in names
        else:
            ens.set_variants(*forms)
        ens.run(15, graph=True)
        runs.append([(ens.get(r, "x"), ens.get(r, "v"), ens.get(r, "T"), ens.get(r, "E"))
                     for r in range(len(reps))])
    for other in runs[1:]:
        for r in range(len(reps)):
            for q in range(4):
                assert np.array_equal(runs[0][r][q], other[r][q]), (r, q)


def _tdyn_replicas(tmp, specs, species):
    """replicas whose grid files make some cells temperature-dependent: C_e(T)
    and kappa(T) from one parameter file (eph_fdm.h:73-104), a hot spot so the
    stability extrema, and with them the sub-step counts, change from step to
    step; a replica with "tdyn": False keeps constant cells"""
    Tk = np.arange(2000) * 10.0
    pp = os.path.join(tmp, "tdyn.params")
    inputs.write_params(pp, 10.0, 3e-6 + 1e-9 * Tk, 0.1 + 2e-4 * Tk)
    out = []
    for k, sp in enumerate(specs):
        n = sp["cells"]
        beta, names, _ = workload_files(tmp, f"t{k}", species, n, A, n)
        if k > 0:
            beta = out[0][1][16]
        L = n * A
        idx = np.indices((n, n, n))[::-1].reshape(3, -1).T
        c = (idx + 0.5) * (L / n) - 0.5 * L
        T = 300.0 + 1700.0 * np.exp(-(c * c).sum(1) / 32.0)
        tdyn = (np.arange(n ** 3) % 3 != 0).astype(int) if sp["tdyn"] else 0
        tin = os.path.join(tmp, f"t{k}.Tin")
        inputs.write_grid(tin, n, n, n, (0.0, L, 0.0, L, 0.0, L), T, rho=1.0, C_e=3e-6, kappa=0.1,
                          tdyn=tdyn, params=pp if sp["tdyn"] else "NULL")
        args = inputs.fix_args(beta, names, seed=sp["fix_seed"], flags="7", grid=(n, n, n),
                               T_infile=tin)
        kw = dict(cells=n, a=A, ntypes=species, seed=sp["seed"], T=300.0, pka_eV=sp["pka"])
        out.append((kw, args))
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_temperature_dependent_grids_match_separate_runs(tmp_path, graph):
    """C_e(T), kappa(T) in an ensemble: each replica's sub-step count is found
    on the device every step (in the single-grid arithmetic) and every replica
    still follows its separate run bit for bit"""
    print(f"ensemble: T-dependent grids, graph={graph}", flush=True)
    species, dt = 1, 5e-4
    specs = [dict(cells=4, seed=31, fix_seed=301, pka=150.0, tdyn=True),
             dict(cells=5, seed=32, fix_seed=302, pka=0.0, tdyn=False),
             dict(cells=6, seed=33, fix_seed=303, pka=200.0, tdyn=True)]
    reps = _tdyn_replicas(str(tmp_path), specs, species)
    ens = _ensemble(reps, species, dt)
    counts = [[ens.info(r)["substeps"] for r in range(3)]]
    rebuilt = []
    for n in (1, 6, 13):
        res = ens.run(n, graph=graph)
        rebuilt.extend(res["rebuilt"].tolist())
        counts.append([ens.info(r)["substeps"] for r in range(3)])
    counts = np.array(counts)
    assert len(set(counts[:, 0])) > 1 and len(set(counts[:, 2])) > 1, counts  # T-dependent
    assert len(set(counts[:, 1])) == 1, counts  # constant cells
    for r, (kw, args) in enumerate(reps):
        md, fix = _separate(kw, args, species, dt, np.array(rebuilt))
        n = ens.info(r)["nlocal"]
        xs, vs, _ = fix.dev_atoms()
        assert np.array_equal(ens.get(r, "x"), xs[:n]), f"replica {r}: x"
        assert np.array_equal(ens.get(r, "v"), vs[:n]), f"replica {r}: v"
        assert np.array_equal(ens.get(r, "T"), fix.grid().get("T")), f"replica {r}: T_e"
        E = ens.get(r, "E")
        assert E[1] == fix.compute_vector(0), f"replica {r}: Ee"


def test_ensemble_candidates_follow_each_replicas_own_moved_atoms(tmp_path):
    """A 2 keV primary atom (0.078 A per 0.1 fs step) in a 4^3 replica (4
    waves) next to two 8^3 replicas at rest (32 waves each) that overlap it in
    space, it and its 26 ghost images.  From its third step it is a moved atom
    of replica 0: waves of replica 0 near it scan their full lists while every
    wave of the other replicas keeps walking its candidates (their lists never
    pair with replica 0's atoms).  Each replica bit for bit equal to its
    separate run."""
    species, dt, nsteps = 2, 1e-4, 12
    specs = [dict(cells=4, seed=31, fix_seed=301, pka=2000.0),
             dict(cells=8, seed=32, fix_seed=302, pka=0.0),
             dict(cells=8, seed=33, fix_seed=303, pka=0.0)]
    reps = _replicas(str(tmp_path), specs, species, dt)
    ens = _ensemble(reps, species, dt)
    ens.set_variants("chunk", "chunk")  # the forms that walk candidate lists
    waves0 = (ens.info(0)["nlocal"] + 63) // 64
    rebuilt = []
    seen = 0
    for k in range(nsteps):
        rebuilt.append(int(ens.run(1, graph=True)["rebuilt"][0]))
        st = ens.candidate_stats()
        assert st["waves"] == sum((ens.info(r)["nlocal"] + 63) // 64 for r in range(3)), st
        if st["moved"] > 0 and not st["retaken"]:
            seen += 1
            assert 0 < st["full_waves"] <= waves0, (k + 1, st)
    assert seen >= 5, f"the primary atom was a moved atom on only {seen} steps"
    for r, (kw, args) in enumerate(reps):
        md, fix = _separate(kw, args, species, dt, np.array(rebuilt))
        n = ens.info(r)["nlocal"]
        xs, vs, _ = fix.dev_atoms()
        assert np.array_equal(ens.get(r, "x"), xs[:n]), f"replica {r}: x"
        assert np.array_equal(ens.get(r, "v"), vs[:n]), f"replica {r}: v"
        assert np.array_equal(ens.get(r, "T"), fix.grid().get("T")), f"replica {r}: T_e"
