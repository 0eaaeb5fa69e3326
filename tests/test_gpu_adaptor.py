"""GPU: the LAMMPS adaptor (user-eph_amd/lammps/fix_eph_amd.cpp) in the
LAMMPS-shaped test harness (tests/lammps_mock) against the same steps driven
directly through the C ABI.

The harness owns the atoms in LAMMPS layout, hands the adaptor a NeighList and
calls its hooks in the order of LAMMPS's Verlet loop (no pair style).  The
adaptor re-sends atoms, ghost map and list after the list build and forwards
every hook; the direct run calls the same C ABI from Python.  Both must leave
x, v, f, the tally and the per-atom array bit-identical.
"""
import os
import subprocess

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
import minisys

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lammps_mock", "_build",
                      "mock_driver")


def _system(seed, species):
    a0 = 3.615
    x = minisys.fcc(4, a0)
    L = 4 * a0
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0.0, 0.05, x.shape)
    x = np.mod(x, L)
    v = rng.normal(0.0, 4.0, x.shape)
    types = 1 + np.arange(len(x)) % species
    return minisys.System(x, v, types, [0, 0, 0], [L] * 3, 6.5, None), L


def _write_input(path, s, ntypes, nsteps, dt, masses, L):
    nall = s.nlocal + s.nghost
    with open(path, "wb") as f:
        np.array([s.nlocal, s.nghost, ntypes, nsteps, len(s.nbr_idx)], np.int32).tofile(f)
        np.array([dt, inputs.BOLTZ, inputs.FTM2V], np.float64).tofile(f)
        np.array([0.0, 0.0, 0.0, L, L, L], np.float64).tofile(f)
        np.array([0.0] + list(masses), np.float64).tofile(f)
        np.ascontiguousarray(s.x[:nall], np.float64).tofile(f)
        np.ascontiguousarray(s.v[:nall], np.float64).tofile(f)
        for arr in (s.type, s.mask, s.tag, s.ghost_owner):
            np.ascontiguousarray(arr, np.int32).tofile(f)
        np.ascontiguousarray(s.ghost_shift, np.float64).tofile(f)
        np.ascontiguousarray(s.nbr_off, np.int32).tofile(f)
        np.ascontiguousarray(s.nbr_idx, np.int32).tofile(f)


@pytest.mark.parametrize("style,species,flags", [("eph", 1, "7"), ("eph", 2, "0x37"),
                                                 ("eph/coloured", 2, "7")])
def test_adaptor_matches_direct_c_abi(tmp_path, style, species, flags):
    assert os.path.exists(DRIVER), "build the harness first (__graft_entry__.build())"
    s, L = _system(8 + species, species)
    pb = str(tmp_path / "a.beta")
    names = inputs.workload_beta(pb, species)
    dt, nsteps = 1e-3, 6
    masses = [inputs.MASS_CU] * species
    args = inputs.fix_args(pb, names, seed=77, flags=flags, grid=(3, 3, 3), style=style,
                           T_out=str(tmp_path / "Te_out"))
    lmp_args = list(args)
    lmp_args[2] = style + "/amd"
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, s, species, nsteps, dt, masses, L)
    r = subprocess.run([DRIVER, fin, fout] + lmp_args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr

    fix = ea.FixEPH(args, species, [0, 0, 0], [L] * 3, dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=masses)
    fix.init()
    fix.set_atoms(s)
    for n in range(1, nsteps + 1):
        fix.initial_integrate(s)
        s.update_ghosts()
        s.f[:] = 0.0
        fix.post_force(s, step=n)
        fix.final_integrate(s)
        fix.end_of_step(s, n)
    n = s.nlocal
    out = np.fromfile(fout, np.float64)
    ncols = 8
    assert out.size == 9 * n + 2 + ncols * n
    x, v, f = out[:3 * n], out[3 * n:6 * n], out[6 * n:9 * n]
    assert np.array_equal(x, s.x[:n].ravel())
    assert np.array_equal(v, s.v[:n].ravel())
    assert np.array_equal(f, s.f[:n].ravel())
    assert out[9 * n] == fix.compute_vector(0)
    assert out[9 * n + 1] == fix.compute_vector(1)
    assert np.array_equal(out[9 * n + 2:].reshape(n, ncols), fix.peratom())
    assert not np.array_equal(x, _system(8 + species, species)[0].x[:n].ravel())  # it moved


def test_adaptor_rejects_styles_with_per_atom_state(tmp_path):
    """the exp-filtered coloured styles keep per-atom filter states the
    adaptor does not carry with migrating atoms: a LAMMPS error"""
    s, L = _system(3, 1)
    pb = str(tmp_path / "a.beta")
    names = inputs.workload_beta(pb, 1)
    args = inputs.fix_args(pb, names, style="eph/coloured/exp", tau0=0.002, grid=(3, 3, 3))
    args[2] = "eph/coloured/exp/amd"
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, s, 1, 1, 1e-3, [inputs.MASS_CU], L)
    r = subprocess.run([DRIVER, fin, fout] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "per-atom state" in r.stderr
