"""GPU: the LAMMPS adaptor (user-eph_amd/lammps/fix_eph_amd.cpp) in the
LAMMPS-shaped test harness (tests/lammps_mock) against the same steps driven
directly through the C ABI.

The harness owns the atoms in LAMMPS layout, hands the adaptor a NeighList and
calls its hooks in the order of LAMMPS's Verlet loop (no pair style).  Midway
it re-orders the local atoms as LAMMPS's atom sort does (pre_exchange, rows
moved through a scratch slot with copy_arrays, ghosts and list rebuilt).  The
direct run calls the same C ABI from Python and, at that step, carries the
per-atom state with the atoms itself.  Both must leave x, v, f, the tally and
the per-atom array bit-identical, for every style.
"""
import copy
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
    x = np.mod(x + rng.normal(0.0, 0.05, x.shape), L)
    v = rng.normal(0.0, 4.0, x.shape)
    types = 1 + np.arange(len(x)) % species
    return minisys.System(x, v, types, [0, 0, 0], [L] * 3, 6.5, None), L


def _permute(s, perm):
    """local slot k takes old local atom perm[k]; ghosts keep their slots and
    get their owners renumbered; lists follow their atoms, renumbered"""
    n = s.nlocal
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    for name in ("x", "v", "f", "type", "mask", "tag"):
        a = getattr(s, name)
        a[:n] = a[:n][perm]
    s.ghost_owner = inv[s.ghost_owner].astype(np.int32)
    off, idx = s.nbr_off, s.nbr_idx
    rows = []
    for o in perm:
        r = idx[off[o]:off[o + 1]].astype(np.int64)
        loc = r < n
        r[loc] = inv[r[loc]]
        rows.append(r)
    s.nbr_off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    s.nbr_idx = np.concatenate(rows).astype(np.int32)


def _write_input(path, s, ntypes, nsteps, dt, masses, L, sort_step=0, perm=None, sorted_s=None):
    nall = s.nlocal + s.nghost
    i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
    with open(path, "wb") as f:
        i32([s.nlocal, s.nghost, ntypes, nsteps, len(s.nbr_idx)]).tofile(f)
        np.array([dt, inputs.BOLTZ, inputs.FTM2V, 0.0, 0.0, 0.0, L, L, L], np.float64).tofile(f)
        np.array([0.0] + list(masses), np.float64).tofile(f)
        np.ascontiguousarray(s.x[:nall], np.float64).tofile(f)
        np.ascontiguousarray(s.v[:nall], np.float64).tofile(f)
        for arr in (s.type, s.mask, s.tag, s.ghost_owner):
            i32(arr).tofile(f)
        np.ascontiguousarray(s.ghost_shift, np.float64).tofile(f)
        i32(s.nbr_off).tofile(f)
        i32(s.nbr_idx).tofile(f)
        i32([sort_step]).tofile(f)
        if sort_step > 0:
            i32(perm).tofile(f)
            i32(sorted_s.ghost_owner).tofile(f)
            i32(sorted_s.nbr_off).tofile(f)
            i32([len(sorted_s.nbr_idx)]).tofile(f)
            i32(sorted_s.nbr_idx).tofile(f)


def _args(tmp_path, style, species, flags):
    pb = str(tmp_path / "a.beta")
    names = inputs.workload_beta(pb, species)
    if style == "eph/atomic":
        pk = str(tmp_path / "a.kappa")
        inputs.synthetic_kappa(pk, names=names)
        return inputs.atomic_args(pb, pk, names, seed=77, flags=flags, inner_loops=2)
    tau0 = 0.002 if style.startswith("eph/coloured/exp") else None
    return inputs.fix_args(pb, names, seed=77, flags=flags, grid=(3, 3, 3), style=style, tau0=tau0,
                           T_out=str(tmp_path / "Te_out"))


CASES = [("eph", 1, "7"), ("eph", 2, "0x37"), ("eph/coloured", 2, "7"),
         ("eph/coloured/exp", 2, "7"), ("eph/coloured/exp/v1", 1, "7"), ("eph/atomic", 1, "7")]


@pytest.mark.parametrize("style,species,flags", CASES)
def test_adaptor_matches_direct_c_abi(tmp_path, style, species, flags):
    assert os.path.exists(DRIVER), "build the harness first (__graft_entry__.build())"
    s, L = _system(8 + species, species)
    x_start = s.x[:s.nlocal].copy()
    args = _args(tmp_path, style, species, flags)
    dt, nsteps, sort_step = 1e-3, 7, 4
    masses = [inputs.MASS_CU] * species
    perm = np.random.default_rng(5).permutation(s.nlocal)
    sorted_s = copy.deepcopy(s)
    _permute(sorted_s, perm)

    lmp_args = list(args)
    lmp_args[2] = style + "/amd"
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, s, species, nsteps, dt, masses, L, sort_step, perm, sorted_s)
    r = subprocess.run([DRIVER, fin, fout] + lmp_args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr

    fix = ea.FixEPH(args, species, [0, 0, 0], [L] * 3, dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=masses)
    fix.init()
    fix.set_atoms(s)
    for n in range(1, nsteps + 1):
        fix.initial_integrate(s)
        if n == sort_step:
            if style == "eph/atomic":
                E = fix.atomic_energy()
            elif style.startswith("eph/coloured/exp"):
                F = [fix.filter_state(w) for w in (0, 1)]
            _permute(s, perm)
            s.update_ghosts()
            fix.set_atoms(s)
            if style == "eph/atomic":
                fix.set_atomic_energy(E[perm])
            elif style.startswith("eph/coloured/exp"):
                for w in (0, 1):
                    fix.set_filter_state(w, F[w][perm])
        s.update_ghosts()
        s.f[:] = 0.0
        fix.post_force(s, step=n)
        fix.final_integrate(s)
        fix.end_of_step(s, n)
    n = s.nlocal
    ncols = lib_cols = fix.peratom().shape[1]
    out = np.fromfile(fout, np.float64)
    assert out.size == 9 * n + 2 + ncols * n, (out.size, lib_cols)
    x, v, f = out[:3 * n], out[3 * n:6 * n], out[6 * n:9 * n]
    assert np.array_equal(x, s.x[:n].ravel())
    assert np.array_equal(v, s.v[:n].ravel())
    assert np.array_equal(f, s.f[:n].ravel())
    assert out[9 * n] == fix.compute_vector(0)
    assert out[9 * n + 1] == fix.compute_vector(1)
    assert np.array_equal(out[9 * n + 2:].reshape(n, ncols), fix.peratom())
    assert not np.array_equal(s.x[:n], x_start[perm])  # the atoms moved


def test_adaptor_state_travels_with_the_atoms(tmp_path):
    """without the adaptor's hand-over the sort would leave each filter state
    on the wrong atom: the direct run that skips it must differ"""
    style, species = "eph/coloured/exp", 2
    s, L = _system(10, species)
    args = _args(tmp_path, style, species, "7")
    dt, nsteps, sort_step = 1e-3, 5, 3
    perm = np.random.default_rng(5).permutation(s.nlocal)
    masses = [inputs.MASS_CU] * species
    fix = ea.FixEPH(args, species, [0, 0, 0], [L] * 3, dt, inputs.BOLTZ, inputs.FTM2V,
                    masses=masses)
    fix.init()
    fix.set_atoms(s)
    for n in range(1, nsteps + 1):
        fix.initial_integrate(s)
        if n == sort_step:
            _permute(s, perm)
            s.update_ghosts()
            fix.set_atoms(s)  # states left in the old order
        s.update_ghosts()
        s.f[:] = 0.0
        fix.post_force(s, step=n)
        fix.final_integrate(s)
        fix.end_of_step(s, n)
    sorted_s = copy.deepcopy(_system(10, species)[0])
    _permute(sorted_s, perm)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    s0, _ = _system(10, species)
    _write_input(fin, s0, species, nsteps, dt, masses, L, sort_step, perm, sorted_s)
    lmp_args = list(args)
    lmp_args[2] = style + "/amd"
    r = subprocess.run([DRIVER, fin, fout] + lmp_args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    n = s.nlocal
    f_adaptor = np.fromfile(fout, np.float64)[6 * n:9 * n]
    assert not np.array_equal(f_adaptor, s.f[:n].ravel())


def test_adaptor_rejects_unknown_styles(tmp_path):
    s, L = _system(3, 1)
    args = _args(tmp_path, "eph", 1, "7")
    args[2] = "eph/unknown/amd"
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, s, 1, 1, 1e-3, [inputs.MASS_CU], L)
    r = subprocess.run([DRIVER, fin, fout] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "unknown style" in r.stderr
