"""GPU parity: the HIP path against the oracle on identical inputs.

Bar: forces, rho_i and T_e within 1e-11 relative (normwise), the energy tally
within 1e-10.  The kernels evaluate the reference's expressions
in its order, so most checks here also hold bit for bit and say so.
"""
import copy
import math
import os

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
from drive import Pair, rel_err, bitwise_fraction

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-11
TOL_E = 1e-10


def _fix_state_close(p, bitwise=True):
    A, B = p.sA, p.sB
    n = A.nlocal
    nall = n + A.nghost
    checks = {
        "rho": (p.gpu.get("rho")[:n], p.ora.get("rho", nall)[:n]),
        "w": (p.gpu.get("w")[:n], p.ora.get("w", nall)[:n]),
        "f_EPH": (p.gpu.get("f_EPH"), p.ora.get("f_EPH", n)),
        "f_RNG": (p.gpu.get("f_RNG"), p.ora.get("f_RNG", n)),
        "x": (B.x[:n], A.x[:n]),
        "v": (B.v[:n], A.v[:n]),
        "f": (B.f[:n], A.f[:n]),
    }
    for k, (g, o) in checks.items():
        e = rel_err(g, o)
        assert e <= TOL, f"{k}: rel err {e:.3e}"
        if bitwise:
            assert bitwise_fraction(g, o) == 1.0, f"{k}: not bit-identical (rel err {e:.3e})"
    Tg = p.gpu.grid().get("T")
    To = p.ora.fdm().get("T")
    assert rel_err(Tg, To) <= TOL
    if bitwise:
        assert np.array_equal(Tg, To)
    eg, eo = p.gpu.compute_vector(0), p.ora.compute_vector(0)
    assert abs(eg - eo) <= TOL_E * max(abs(eo), 1e-300), (eg, eo)


def test_single_step_bitwise_injected_noise(oracle, tmp_path):
    p = Pair(oracle, ea, str(tmp_path), cells=6, T=600.0, disp=0.05, grid=6, inject=True)
    p.step()
    _fix_state_close(p)
    # the per-atom output array of fix_eph.cpp:406-428
    pa = p.gpu.peratom()
    po = p.ora.get("array", p.sA.nlocal)
    assert np.array_equal(pa, po)


def test_device_noise_stream_matches_oracle_stream(oracle, tmp_path):
    p = Pair(oracle, ea, str(tmp_path), cells=5, grid=5, inject=False)
    p.step()
    n = p.sA.nlocal
    xg = p.gpu.get("xi")[:n]
    xo = oracle.xi(p.seed, p.sA.tag[:n], 1)
    assert np.array_equal(xg, xo), f"max diff {np.abs(xg - xo).max():.3e}"
    _fix_state_close(p)


def test_trajectory_fcc10_100_steps(oracle, tmp_path):
    """fcc 10^3 cells (4,000 atoms), 300 K, full model, dt 1 fs, 100 steps."""
    p = Pair(oracle, ea, str(tmp_path), cells=10, T=300.0, grid=10, dt=1e-3)
    for k in range(100):
        p.step()
        if k % 25 == 24:
            _fix_state_close(p)
    _fix_state_close(p)
    assert abs(p.gpu.compute_vector(1) - p.ora.compute_vector(1)) <= 1e-12 * p.ora.compute_vector(1)


@pytest.mark.parametrize("delta", [0.0, 1e-3, 0.4, -1.0])
def test_candidate_margin_keeps_bits(oracle, tmp_path, delta):
    """The density sweep's rho loop over the candidate list gives the oracle's
    bits for any margin: none (full list every step), one so small that the
    candidates are retaken nearly every step, a wide one and the default, over
    a 60-step 900 K trajectory with list rebuilds."""
    p = Pair(oracle, ea, str(tmp_path), cells=6, T=900.0, grid=6, dt=1e-3)
    p.gpu.set_candidate_margin(delta)
    for k in range(60):
        p.step()
        if k % 20 == 19:
            _fix_state_close(p)


@pytest.mark.parametrize("forms", [("auto", "auto"), ("lds", "lds")])
def test_alloy_cascade_small(oracle, tmp_path, forms):
    """4 species, 1 keV PKA, hot spot, dt 0.1 fs: multi-element tables and
    rebuilds; by default (chunk forms at this size) and with the LDS-table
    forms forced on a small system."""
    p = Pair(oracle, ea, str(tmp_path), cells=8, species=4, T=300.0, pka_eV=1000.0, grid=8,
             dt=1e-4, hot_spot=(3000.0, 6.0))
    p.gpu.set_sweep_forms(*forms)
    for _ in range(30):
        p.step()
    _fix_state_close(p)


@pytest.mark.parametrize("forms", [("auto", "auto"), ("chunk", "chunk"), ("chunk", "lds"),
                                   ("lds", "chunk"), ("lds512", "lds256"), ("lds256", "lds512"),
                                   ("chunk3", "chunk2"), ("chunk2", "chunk4")])
def test_alloy_lds_table_path(oracle, tmp_path, forms):
    """fcc 21^3 (37,044 atoms), 4 species: large enough that the default is the
    persistent kernels staging all four rho(r^2) tables (128 KB) in LDS; every
    combination of sweep forms (and LDS block sizes 1024, 512, 256) gives the
    oracle's bits."""
    p = Pair(oracle, ea, str(tmp_path), cells=21, species=4, T=900.0, disp=0.05, grid=7, dt=5e-4)
    p.gpu.set_sweep_forms(*forms)
    for _ in range(3):
        p.step()
    _fix_state_close(p)


def test_autotune_keeps_forces_and_bits(oracle, tmp_path):
    """eph_autotune between steps changes neither f nor the trajectory"""
    p = Pair(oracle, ea, str(tmp_path), cells=21, species=4, T=900.0, disp=0.05, grid=7, dt=5e-4)
    p.step()
    before = p.gpu.dev_atoms()
    tuned = p.gpu.autotune()
    after = p.gpu.dev_atoms()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    forms = ("chunk", "lds", "lds512", "lds256", "chunk4", "chunk3", "chunk2")
    assert tuned["density"] in forms and tuned["force"] in forms
    assert set(tuned["density_ms"]) == set(forms) == set(tuned["force_ms"]), tuned
    assert all(v > 0 for v in tuned["density_ms"].values()), tuned
    for _ in range(2):
        p.step()
    _fix_state_close(p)


@pytest.mark.parametrize("flags", ["1", "2", "3", "5", "0x17", "0x27", "0xF"])
def test_flag_combinations(oracle, tmp_path, flags):
    p = Pair(oracle, ea, str(tmp_path), cells=5, T=800.0, disp=0.05, grid=5, flags=flags)
    for _ in range(4):
        p.step()
    _fix_state_close(p)


def test_group_mask_subset(oracle, tmp_path):
    p = Pair(oracle, ea, str(tmp_path), cells=5, T=800.0, grid=5)
    rng = np.random.default_rng(3)
    sel = rng.uniform(size=p.sA.nlocal) < 0.6
    for s in (p.sA, p.sB):
        s.mask[:s.nlocal] = np.where(sel, 3, 4)
        s.mask[s.nlocal:] = s.mask[s.ghost_owner]
    p.ora.set_groupbit(2)
    p.gpu.set_groupbit(2)
    p.gpu.set_atoms(p.sB)
    for _ in range(3):
        p.step()
    _fix_state_close(p)


def test_fcc40_config_parity(oracle, tmp_path):
    """fcc 40^3 (256,000 atoms), 1000 K, 0.05 A displacements, Te grid 40^3 (5 steps)."""
    p = Pair(oracle, ea, str(tmp_path), cells=40, T=1000.0, disp=0.05, grid=40, dt=1e-3)
    for _ in range(5):
        p.step()
    _fix_state_close(p)


# ---------------------------------------------------------------- T_e grid
def _grid_from_fixture(d, tmp_path, name):
    params = "NULL"
    if "params_dT" in d:
        params = str(tmp_path / (name + ".params"))
        inputs.write_params(params, float(d["params_dT"]), d["params_C"], d["params_kappa"])
    gp = str(tmp_path / (name + ".in"))
    inputs.write_grid(gp, int(d["nx"]), int(d["ny"]), int(d["nz"]), tuple(d["box"]), d["T"], d["S"],
                      d["rho"], d["C"], d["kappa"], d["flag"], d["tdyn"], steps=int(d["steps"]),
                      params=params)
    return gp


@pytest.mark.parametrize("name", ["fdm_periodic", "fdm_walls", "fdm_unstable", "fdm_tdyn"])
def test_grid_reference_goldens_bitwise(tmp_path, name):
    """EPH_FDM::solve golden runs written by the compiled reference headers."""
    d = np.load(os.path.join(G, name + ".npz"))
    g = ea.Grid.load(_grid_from_fixture(d, tmp_path, name))
    g.set_dt(float(d["dt"]))
    E = d["energies"] if "energies" in d else None
    for s in range(int(d["nsteps"])):
        if E is not None:
            g.insert_energy(E[:, 0], E[:, 1], E[:, 2], E[:, 3])
        g.solve()
        assert np.array_equal(g.get("T"), d["T_steps"][s]), f"step {s + 1}"


@pytest.mark.parametrize("dims", [(13, 11, 9), (8, 8, 4), (3, 9, 5), (17, 2, 6)])
def test_grid_mixed_cells_match_oracle(oracle, tmp_path, dims):
    """Random grids, thin ones included, with constant, dynamic, wall and
    temperature-dependent cells side by side and deposits every step, bitwise
    against the oracle's EPH_FDM::solve over 6 solves"""
    nx, ny, nz = dims
    ncell = nx * ny * nz
    rng = np.random.default_rng(100 * nx + 10 * ny + nz)
    flag = rng.choice([0, 1, 2], p=[0.1, 0.75, 0.15], size=ncell)
    tdyn = ((rng.uniform(size=ncell) < 0.3) & (flag == 1)).astype(int)
    d = np.load(os.path.join(G, "fdm_tdyn.npz"))
    params = str(tmp_path / "t.params")
    inputs.write_params(params, float(d["params_dT"]), d["params_C"], d["params_kappa"])
    box = (0.0, 1.5 * nx, -1.0, 1.2 * ny - 1.0, 2.0, 2.0 + 1.3 * nz)
    gp = str(tmp_path / "tiles.in")
    inputs.write_grid(gp, nx, ny, nz, box, rng.uniform(400.0, 1500.0, ncell), rng.normal(0, 1e-3, ncell),
                      rng.uniform(0.8, 1.2, ncell), rng.uniform(1.5e-5, 3e-5, ncell),
                      rng.uniform(0.04, 0.09, ncell), flag, tdyn, steps=2, params=params)
    g = ea.Grid.load(gp)
    o = oracle.FDM.load(gp)
    g.set_dt(1e-3)
    o.set_dt(1e-3)
    for s in range(6):
        n = 40
        x = rng.uniform(box[0], box[1], n)
        y = rng.uniform(box[2], box[3], n)
        z = rng.uniform(box[4], box[5], n)
        E = rng.normal(0, 1e-3, n)
        g.insert_energy(x, y, z, E)
        for k in range(n):
            o.insert_energy(x[k], y[k], z[k], E[k])
        ng, no = g.solve(), o.solve()
        assert ng == no, (s, ng, no)
        assert np.array_equal(g.get("T"), o.get("T")), f"solve {s + 1}"


def test_grid_out_ref_laser(oracle):
    """Tests/EPH_FDM/Test3/Out_Ref through the HIP stencil (variable C_e, kappa)."""
    d = np.load(os.path.join(G, "fdm_outref.npz"))
    nx, c_e, kappa_e, dt = 1000, 1.0, 1.0, 1e-4
    g = ea.Grid.const(nx, 1, 1, [-10.0, 10.0, -1.0, 1.0, -1.0, 1.0], 1.0, c_e, 1.0, kappa_e)
    g.set_dt(dt)
    C = np.array([2.0 * c_e + c_e * math.sin(i * 2.0 * math.pi / nx) for i in range(nx)])
    K = np.array([2.0 * kappa_e - kappa_e * math.sin(i * 2.0 * math.pi / nx) for i in range(nx)])
    g.set("C", C)
    g.set("kappa", K)
    r = dt * (1.0 / 0.02 / 0.02 + 0.25 + 0.25) / C.min() * K.max()
    g.set_steps(max(1, math.ceil(r / 0.4)))  # Out_Ref was written with the ceil count
    S = np.zeros(nx)
    S[500] = 10.0 / 0.08
    g.set("S", S)
    save = int(1.0 / dt)
    for i in range(2 * save + 1):
        if i == save:
            g.set("S", np.zeros(nx))
        g.solve()
        if i in (0, save, 2 * save):
            assert np.allclose(g.get("T"), d[f"test3_snap{i // save}"], rtol=6e-6, atol=0)


def test_grid_deposit_order_and_lookup(oracle):
    """insert_energy in array order (bitwise) and get_index/get_T."""
    rng = np.random.default_rng(4)
    box = [0.0, 7.0, -2.0, 3.0, 1.0, 9.0]
    g = ea.Grid.const(7, 5, 8, box, 300.0, 2e-4, 1.0, 0.1)
    o = oracle.FDM.const(7, 5, 8, box, 300.0, 2e-4, 1.0, 0.1)
    g.set_dt(1e-3); o.set_dt(1e-3)
    n = 5000
    x = rng.uniform(-20, 27, n); y = rng.uniform(-9, 11, n); z = rng.uniform(-15, 25, n)
    E = rng.normal(0, 1e-3, n)
    g.insert_energy(x, y, z, E)
    for k in range(n):
        o.insert_energy(x[k], y[k], z[k], E[k])
    assert np.array_equal(g.get("dT"), o.get("dT"))
    idx, _ = g.lookup(x, y, z)
    assert all(idx[k] == o.get_index(x[k], y[k], z[k]) for k in range(0, n, 37))
    g.solve(); o.solve()
    assert np.array_equal(g.get("T"), o.get("T"))


def test_grid_many_atoms_per_cell_deposit(oracle):
    """coarse grids: cells holding > 64 and > 8192 items keep the serial order"""
    rng = np.random.default_rng(8)
    for dims, n in (((2, 2, 1), 3000), ((1, 1, 1), 20000)):
        box = [0.0, 4.0, 0.0, 4.0, 0.0, 4.0]
        g = ea.Grid.const(*dims, box, 300.0, 2e-4, 1.0, 0.1)
        o = oracle.FDM.const(*dims, box, 300.0, 2e-4, 1.0, 0.1)
        x, y, z = (rng.uniform(0, 4, n) for _ in range(3))
        E = rng.normal(0, 1.0, n)
        g.insert_energy(x, y, z, E)
        for k in range(n):
            o.insert_energy(x[k], y[k], z[k], E[k])
        assert np.array_equal(g.get("dT"), o.get("dT")), dims


@pytest.mark.parametrize("model,species", [(1, 1), (1, 4), (2, 1), (2, 4)])
def test_models_ttm_and_prb(oracle, tmp_path, model, species):
    """fix eph models 1 (TTM) and 2 (PRB) against the oracle, with rebuilds and
    the T_e grid in the loop"""
    p = Pair(oracle, ea, str(tmp_path), cells=6, species=species, T=600.0, pka_eV=300.0, grid=6,
             dt=2.5e-4, hot_spot=(2000.0, 5.0), model=model)
    for _ in range(25):
        p.step()
    _fix_state_close(p)


@pytest.mark.parametrize("style", ["eph/coloured", "eph/coloured/exp", "eph/coloured/exp/v1"])
def test_coloured_styles(oracle, tmp_path, style):
    """the coloured styles' filter states persist over steps and rebuilds and
    give the oracle's bits"""
    p = Pair(oracle, ea, str(tmp_path), cells=6, species=2, T=600.0, pka_eV=300.0, grid=6,
             dt=2.5e-4, hot_spot=(2000.0, 5.0), style=style, tau0=1e-3)
    for _ in range(25):
        p.step()
    _fix_state_close(p)
    if style != "eph/coloured":
        for which in (0, 1):
            st = p.gpu.filter_state(which)
            assert np.abs(st).max() > 0, which
            p.gpu.set_filter_state(which, st * 0.5)
            assert np.array_equal(p.gpu.filter_state(which), st * 0.5)
