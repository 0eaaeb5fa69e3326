"""GPU: the device-resident path (harness on the device) against the host path.

The device rebuild of ghosts and lists must reproduce the host harness build
exactly, and a device-resident trajectory (pair, rebuilds, fix hooks all on
the GPU) must follow the oracle-driven host loop bit for bit.
"""
import copy
import os
import tempfile

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
from drive import workload_files, rel_err

pytestmark = pytest.mark.gpu


def _fix(tmp, md, species=1, grid=6, flags="7", dt=1e-3, seed=99, hot_spot=None, cells=6):
    beta, names, tin = workload_files(tmp, f"d{cells}_{species}", species, cells, 3.615, grid, hot_spot)
    args = inputs.fix_args(beta, names, seed=seed, flags=flags, grid=(grid,) * 3, T_infile=tin)
    inf = md.info()
    return args, inf


def _compare_systems(a, b):
    assert a.nlocal == b.nlocal and a.nghost == b.nghost
    assert np.array_equal(a.x, b.x)
    # ghost velocities are refreshed by the next step's ghost update, not here
    assert np.array_equal(a.v[:a.nlocal], b.v[:b.nlocal])
    assert np.array_equal(a.ghost_owner, b.ghost_owner)
    assert np.array_equal(a.ghost_shift, b.ghost_shift)
    assert np.array_equal(a.type, b.type) and np.array_equal(a.tag, b.tag)
    assert np.array_equal(a.nbr_off, b.nbr_off)
    assert np.array_equal(a.nbr_idx, b.nbr_idx)


# skin 2 A: fix lists of about 90 entries; 3 A: about 145 (the device sort's
# 256-key LDS-tile form); 5 A: about 320 (512 keys, one wave per atom)
@pytest.mark.parametrize("cells,species,skin", [(6, 1, 2.0), (9, 4, 2.0), (7, 2, 3.0), (8, 1, 5.0)])
def test_device_rebuild_matches_host_build(tmp_path, cells, species, skin):
    print(f"device rebuild: fcc{cells}, {species} species, skin {skin}", flush=True)
    cfg = dict(a=3.615, ntypes=species, seed=11, T=2500.0, disp_sigma=0.08, skin=skin)
    md_d = ea.MD(cells, **cfg)
    args, inf = _fix(str(tmp_path), md_d, species=species, cells=cells, grid=cells)
    fix = ea.FixEPH(args, species, inf["boxlo"], inf["boxhi"], 1e-3, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU] * species)
    md_d.attach(fix)
    for _ in range(3):
        md_d.force_rebuild()
        md_d.run(fix, 1)
    md_d.pull(fix)
    sD = md_d.snapshot()
    # the host harness rebuilt from the same local atoms
    md_h = ea.MD(cells, **cfg)
    s = md_h.snapshot()
    s.x[:s.nlocal] = sD.x[:sD.nlocal]
    s.v[:s.nlocal] = sD.v[:sD.nlocal]
    md_h.rebuild(s)
    sH = md_h.snapshot()
    # the device rebuild remaps with the positions it had before the remap;
    # those equal sD's locals after remap, so a second remap is the identity
    _compare_systems(sD, sH)
    longest = int(np.diff(sD.nbr_off).max())
    print(f"longest fix list {longest}", flush=True)
    # list capacity = longest + longest / 4 + 16: skin 3 A within 256 keys, 5 A beyond
    assert {2.0: True, 3.0: 100 < longest <= 190, 5.0: 200 < longest <= 390}[skin], longest


def test_device_resident_run_matches_oracle(oracle, tmp_path):
    """fcc 6^3, 200 eV PKA at dt = 0.5 fs (the PKA triggers rebuilds), 40 steps:
    the device loop (pair, ghosts, device rebuilds, fix hooks, Philox noise)
    against the oracle driven through the host harness."""
    cells, grid, dt, nsteps = 6, 6, 5e-4, 40
    cfg = dict(a=3.615, ntypes=1, seed=5, T=300.0, pka_eV=200.0)
    md_d = ea.MD(cells, **cfg)
    args, inf = _fix(str(tmp_path), md_d, cells=cells, grid=grid)
    fix = ea.FixEPH(args, 1, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU])
    fix.init()
    md_d.attach(fix)
    res = md_d.run(fix, nsteps, first_step=1)
    assert res["rebuilds"] >= 2
    xg, vg, fg = fix.dev_atoms()
    # oracle through the host harness
    md_h = ea.MD(cells, **cfg)
    ora = oracle.Fix(args, 1, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ, inputs.FTM2V,
                     masses=[inputs.MASS_CU])
    s = md_h.snapshot()
    md_h.lj(s)
    for n in range(1, nsteps + 1):
        ora.initial_integrate(s)
        if md_h.needs_rebuild(s):
            md_h.rebuild(s)
            v = s.v[:s.nlocal].copy()
            s = md_h.snapshot()
            s.v[:s.nlocal] = v
            md_h.update_ghosts(s)
        else:
            md_h.update_ghosts(s)
        md_h.lj(s)
        ora.post_force(s, step=n)
        ora.final_integrate(s)
        ora.end_of_step(s, n)
    n = s.nlocal
    assert np.array_equal(xg[:n], s.x[:n]), rel_err(xg[:n], s.x[:n])
    assert np.array_equal(vg[:n], s.v[:n]), rel_err(vg[:n], s.v[:n])
    assert np.array_equal(fix.grid().get("T"), ora.fdm().get("T"))
    assert abs(fix.compute_vector(0) - ora.compute_vector(0)) <= 1e-10 * abs(ora.compute_vector(0))
