"""GPU: edge cases and the full-size workload's invariants.

* an empty system (no atoms) runs every hook and deposits nothing;
* an isolated atom (rho = 0) and an isolated dimer against the oracle;
* the 2,048,000-atom four-species cascade lattice (fcc 80^3, 10 keV PKA, the
  benchmark's system), too large for the oracle in a test, through
  size-independent properties of one post_force / end_of_step: momentum of
  f_EPH and of f_RNG (uniform T_e), exact dissipation sum f_EPH . v = -sum |w|^2,
  the deposit's energy E_local = -dt sum (f_EPH + f_RNG) . v and its grid sum.
"""
import types

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
import minisys

pytestmark = pytest.mark.gpu


def _empty_system():
    s = types.SimpleNamespace()
    s.nlocal = s.nghost = 0
    s.x = np.zeros((0, 3)); s.v = np.zeros((0, 3)); s.f = np.zeros((0, 3))
    s.type = np.zeros(0, np.int32); s.mask = np.zeros(0, np.int32); s.tag = np.zeros(0, np.int32)
    s.ghost_owner = np.zeros(0, np.int32); s.ghost_shift = np.zeros((0, 3))
    s.nbr_off = np.zeros(1, np.int32); s.nbr_idx = np.zeros(0, np.int32)
    return s


def test_empty_system_runs_every_hook(tmp_path):
    p = str(tmp_path / "e.beta")
    names = inputs.workload_beta(p, 1)
    fix = ea.FixEPH(inputs.fix_args(p, names, grid=(3, 3, 3)), 1, [0, 0, 0], [10.0] * 3, 1e-3,
                    inputs.BOLTZ, inputs.FTM2V, masses=[inputs.MASS_CU])
    fix.init()
    s = _empty_system()
    fix.set_atoms(s)
    T0 = fix.grid().get("T").copy()
    for n in (1, 2):
        fix.initial_integrate(s)
        fix.post_force(s, step=n)
        fix.final_integrate(s)
        fix.end_of_step(s, n)
    assert fix.compute_vector(0) == 0.0
    assert np.array_equal(fix.grid().get("T"), T0)  # uniform grid, nothing deposited


def test_isolated_atom_and_dimer_match_oracle(oracle, tmp_path):
    L = 40.0
    r = 2.7
    x = np.array([[10.0, 10.0, 10.0], [10.0 + r * 0.6, 10.0 + r * 0.8, 10.0], [25.0, 20.0, 22.0]])
    v = np.array([[1.0, -2.0, 0.5], [-0.3, 0.7, 2.0], [3.0, 1.0, -1.0]])
    sA = minisys.System(x, v, [1, 1, 1], [0, 0, 0], [L, L, L], 6.5, None)
    sB = minisys.System(x, v, [1, 1, 1], [0, 0, 0], [L, L, L], 6.5, None)
    p = str(tmp_path / "d.beta")
    names = inputs.workload_beta(p, 1)
    args = inputs.fix_args(p, names, grid=(2, 2, 2))
    ora = oracle.Fix(args, 1, [0, 0, 0], [L] * 3, 1e-3, inputs.BOLTZ, inputs.FTM2V,
                     masses=[inputs.MASS_CU])
    gpu = ea.FixEPH(args, 1, [0, 0, 0], [L] * 3, 1e-3, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU])
    gpu.init()
    gpu.set_atoms(sB)
    for n in (1, 2, 3):
        ora.post_force(sA, step=n)
        gpu.post_force(sB, step=n)
        ora.end_of_step(sA, n)
        gpu.end_of_step(sB, n)
        assert gpu.get("rho")[2] == 0.0 and np.all(gpu.get("f_EPH")[2] == 0.0)
        for q in ("rho", "f_EPH", "f_RNG"):
            g, o = gpu.get(q)[:3], ora.get(q, 3)
            assert np.array_equal(g, o), q
        assert np.array_equal(sB.f[:3], sA.f[:3])
        assert np.array_equal(gpu.grid().get("T"), ora.fdm().get("T"))


def test_full_size_cascade_invariants(tmp_path):
    cells, species = 80, 4
    md = ea.MD(cells, ntypes=species, seed=2024, T=300.0, pka_eV=1.0e4)
    inf = md.info()
    p = str(tmp_path / "a4.beta")
    names = inputs.workload_beta(p, species)
    dt = 1e-4
    args = inputs.fix_args(p, names, flags="3", grid=(cells,) * 3, seed=4242)
    fix = ea.FixEPH(args, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU] * species)
    fix.init()
    s = md.snapshot()
    n = s.nlocal
    assert n == 2_048_000
    fix.set_atoms(s)
    s.f[:] = 0.0
    fix.post_force(s, step=1)
    fE = fix.get("f_EPH")
    fR = fix.get("f_RNG")
    w = fix.get("w")[:n]
    v = s.v[:n]
    for f in (fE, fR):
        assert np.all(np.abs(f.sum(0)) <= 1e-11 * np.abs(f).sum(0)), f.sum(0)
    w2 = (w * w).sum()
    assert abs((fE * v).sum() + w2) <= 1e-10 * w2
    fix.end_of_step(s, 1)
    E_local = fix.last_energy()
    P = -dt * ((fE + fR) * v).sum()
    assert abs(E_local - P) <= 1e-10 * abs(P)
    g = fix.grid()
    dV = (inf["boxhi"][0] - inf["boxlo"][0]) ** 3 / cells ** 3
    assert abs(g.get("dT").sum() * dV * dt - E_local) <= 1e-10 * abs(E_local)
    assert np.isfinite(fix.compute_vector(0))
