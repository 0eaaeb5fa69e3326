"""GPU: the T_e grid's file formats through the product.

The state file (T.restart, EPH_FDM::save_state, eph_fdm.h:226-265), the heat
map (save_temperature, :198-224) and the grid reader (:48-119) against the
compiled reference's own files (tests/golden/fdm_save.npz, fdm_legacy.npz):
written byte for byte, read back and resumed bit for bit, the older grid
layout read, and the fix's post_run restart file equal to the oracle's.
"""
import os

import numpy as np
import pytest

import eph_amd as ea
from drive import Pair

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case_files(d, name, where):
    with open(os.path.join(where, name + ".in"), "wb") as f:
        f.write(d[f"{name}_grid"].tobytes())
    if f"{name}_params" in d:
        with open(os.path.join(where, name + ".params"), "wb") as f:
            f.write(d[f"{name}_params"].tobytes())
    return d[f"{name}_energies"] if f"{name}_energies" in d else None


@pytest.mark.parametrize("name", ["periodic", "walls", "tdyn"])
def test_grid_writers_and_resume_match_reference(tmp_path, monkeypatch, name):
    print(f"grid files: {name}", flush=True)
    d = np.load(os.path.join(G, "fdm_save.npz"))
    monkeypatch.chdir(tmp_path)
    E = _case_files(d, name, str(tmp_path))
    dt = float(d[f"{name}_dt"])

    def run(g, n):
        g.set_dt(dt)
        for _ in range(n):
            if E is not None:
                g.insert_energy(E[:, 0], E[:, 1], E[:, 2], E[:, 3])
            g.solve()

    g = ea.Grid.load(str(tmp_path / (name + ".in")))
    try:
        run(g, int(d[f"{name}_nsteps"]))
        g.save_state("T.restart")
        g.save_temperature("Te_out", 7)
    finally:
        g.close()
    for fn, key in (("T.restart", "state"), ("Te_out_000007", "heat")):
        with open(fn, "rb") as f:
            got = f.read()
        assert got == d[f"{name}_{key}"].tobytes(), f"{fn} differs from the reference's"
    r = ea.Grid.load("T.restart")
    try:
        run(r, int(d[f"{name}_nmore"]))
        assert np.array_equal(r.get("T"), d[f"{name}_resume_T"])
    finally:
        r.close()


def test_grid_legacy_layout(tmp_path):
    print("grid files: legacy layout", flush=True)
    d = np.load(os.path.join(G, "fdm_legacy.npz"))
    gp = str(tmp_path / "T.in")
    with open(gp, "wb") as f:
        f.write(d["legacy_file"].tobytes())
    g = ea.Grid.load(gp)
    try:
        assert g.ntotal == 125
        g.set_dt(float(d["dt"]))
        E = d["energies"]
        for s in range(4):
            g.insert_energy(E[:, 0], E[:, 1], E[:, 2], E[:, 3])
            g.solve()
            assert np.array_equal(g.get("T"), d["T_steps"][s]), f"step {s + 1}"
    finally:
        g.close()
    bad = d["legacy_file"].tobytes().decode().split("\n")
    bad[7] = "0 0 0 300.0 0.0"
    bp = str(tmp_path / "bad.in")
    with open(bp, "w") as f:
        f.write("\n".join(bad))
    with pytest.raises(ea.EphError) as ei:
        ea.Grid.load(bp)
    assert ei.value.code == -3  # malformed file


def test_fix_post_run_restart_file_equals_oracle(oracle, tmp_path):
    """fix eph's post_run writes <T_infile>.restart (fix_eph.cpp:155-160,
    1019-1021); after 20 steps of a hot-spot run the product's file is the
    oracle's byte for byte, and a fix started from it continues on the
    oracle's bits."""
    print("grid files: fix post_run restart", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=6, T=600.0, grid=6, dt=1e-3, hot_spot=(2000.0, 4.0))
    for _ in range(20):
        p.step()
    restart = p.T_infile + ".restart"
    p.ora.post_run()
    with open(restart, "rb") as f:
        want = f.read()
    os.remove(restart)
    p.gpu.post_run()
    with open(restart, "rb") as f:
        got = f.read()
    assert got == want
    g = ea.Grid.load(restart)
    o = oracle.FDM.load(restart)
    try:
        assert np.array_equal(g.get("T"), o.get("T"))
        for k in range(3):
            for grid in (g, o):
                grid.set_dt(1e-3)
                grid.solve()
            assert np.array_equal(g.get("T"), o.get("T")), f"resumed solve {k + 1}"
    finally:
        g.close()
