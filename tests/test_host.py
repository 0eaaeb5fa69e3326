"""CPU checks of the product library's host side (no GPU needed).

The library must load and export every symbol the headers declare; its own
table builder must reproduce the reference's tables bit for bit; argument and
file errors must be reported the way FixEPH reports them; and the harness
that stands in for LAMMPS must produce exact ghosts, complete lists and
correct pair forces.
"""
import gzip
import glob
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
import minisys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def test_library_exports_every_declared_symbol():
    L = ea.lib()
    names = ea.exported_symbols(HEADERS)
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_product_spline_builder_matches_reference_goldens():
    d = np.load(os.path.join(G, "spline.npz"))
    assert np.array_equal(ea.spline_build(float(d["sin_dx"]), d["sin_y"]), d["sin_coeffs"])
    for k in range(4):
        assert np.array_equal(ea.spline_build(float(d[f"rand{k}_dx"]), d[f"rand{k}_y"]),
                              d[f"rand{k}_coeffs"])


def _beta_file(tmp_path, name):
    if name == "beta_synth1.npz":
        p = str(tmp_path / "s1.beta"); inputs.workload_beta(p, 1); return p
    if name == "beta_synth4.npz":
        p = str(tmp_path / "s4.beta"); inputs.workload_beta(p, 4); return p
    src = os.path.join(G, name.replace("beta_", "").replace(".npz", ".beta.gz"))
    p = str(tmp_path / os.path.basename(src)[:-3])
    with gzip.open(src, "rb") as fi, open(p, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return p


@pytest.mark.parametrize("name", ["beta_synth1.npz", "beta_synth4.npz", "beta_Ni.npz",
                                  "beta_Ni_model_4.npz"])
def test_product_beta_tables_match_reference_goldens(tmp_path, name):
    d = np.load(os.path.join(G, name))
    path = _beta_file(tmp_path, name)
    for e in range(int(d["n_elements"])):
        for w in range(4):
            rows, inv, rc, rho = ea.beta_table(path, e, w)
            key = f"e{e}_t{w}"
            assert inv == float(d[key + "_inv_dx"])
            assert rc == float(d["r_cutoff"]) and rho == float(d["rho_cutoff"])
            assert sha(rows) == str(d[key + "_sha256"]), key


def _args(tmp_path, **kw):
    p = str(tmp_path / "s1.beta")
    inputs.workload_beta(p, 1)
    return inputs.fix_args(p, ["Cu"], grid=(4, 4, 4), **kw)


@pytest.mark.parametrize("mutate,code", [
    (lambda a: a[:17], -1),                                   # too few arguments
    (lambda a: a[:16] + ["/nonexistent.beta", "Cu"], -2),     # missing .beta
    (lambda a: a[:17] + ["Xx"], -1),                          # element not in file
    (lambda a: a[:10] + ["0"] + a[11:], -1),                  # non-positive grid
    (lambda a: a[:5] + ["3"] + a[6:], -8),                    # PRLCM: undefined in the reference
    (lambda a: a[:2] + ["eph/other"] + a[3:], -1),            # unknown style
    (lambda a: a[:2] + ["eph/coloured/exp"] + a[3:5] + ["0"] + a[6:], -1),  # tau0 <= 0
    (lambda a: a[:5] + ["7"] + a[6:], -1),                    # invalid model
    (lambda a: a[:13] + ["/nonexistent.T"] + a[14:], -2),     # missing grid file
])
def test_fix_command_errors_like_the_reference(tmp_path, mutate, code):
    args = mutate(_args(tmp_path))
    with pytest.raises(ea.EphError) as e:
        ea.FixEPH(args, 1, [0, 0, 0], [10, 10, 10], 1e-3, inputs.BOLTZ, inputs.FTM2V)
    assert e.value.code == code


def _no_gpu():
    return not os.path.exists("/dev/kfd")


@pytest.mark.skipif(not _no_gpu(), reason="a GPU is present")
def test_valid_fix_needs_a_device_and_says_so(tmp_path):
    with pytest.raises(ea.EphError) as e:
        ea.FixEPH(_args(tmp_path), 1, [0, 0, 0], [10, 10, 10], 1e-3, inputs.BOLTZ, inputs.FTM2V)
    assert e.value.code == -9


@pytest.mark.skipif(not _no_gpu(), reason="a GPU is present")
def test_ensemble_needs_a_device_and_says_so():
    with pytest.raises(ea.EphError) as e:
        ea.Ensemble()
    assert e.value.code == -9


def test_grid_file_errors(tmp_path):
    p = str(tmp_path / "bad.T")
    with open(p, "w") as f:
        f.write("no comment\n")
    with pytest.raises(ea.EphError) as e:
        ea.Grid.load(p)
    assert e.value.code == -3
    # a temperature-dependent cell without a parameter file is rejected
    q = str(tmp_path / "t.T")
    inputs.write_grid(q, 2, 2, 2, (0, 1, 0, 1, 0, 1), 300.0, tdyn=1)
    with pytest.raises(ea.EphError) as e:
        ea.Grid.load(q)
    assert e.value.code == -3


def test_setmask_names_the_four_hooks():
    assert ea.lib().eph_setmask() == 0b1111


def test_harness_ghosts_lists_and_lattice():
    md = ea.MD(4, seed=3, T=300.0, disp_sigma=0.05)
    s = md.snapshot()
    nl = s.nlocal
    assert nl == 256
    # ghost images are owner + shift exactly, shifts are multiples of L
    assert np.array_equal(s.x[nl:], s.x[s.ghost_owner] + s.ghost_shift)
    L = s.boxhi - s.boxlo
    assert np.all(np.isin(np.round(s.ghost_shift / L), [-1, 0, 1]))
    # every image within the ghost cutoff of the box exists exactly once
    cut = 7.5
    want = 0
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if (sx, sy, sz) == (0, 0, 0):
                    continue
                p = s.x[:nl] + np.array([sx, sy, sz]) * L
                want += np.sum(np.all((p >= s.boxlo - cut) & (p < s.boxhi + cut), axis=1))
    assert s.nghost == want
    # fix list = all neighbours within r_cut + skin (brute force)
    c2 = 6.5 ** 2
    for i in range(0, nl, 7):
        d = s.x - s.x[i]
        js = set(np.nonzero((d * d).sum(1) < c2)[0].tolist()) - {i}
        got = s.nbr_idx[s.nbr_off[i]:s.nbr_off[i + 1]].tolist()
        assert len(got) == len(js) and set(got) == js
    # zero momentum, temperature 300 K with 3N - 3 dof
    v = s.v[:nl]
    assert np.abs(v.sum(0)).max() < 1e-9
    T = (v * v).sum() * inputs.MASS_CU * inputs.MVV2E / ((3 * nl - 3) * inputs.BOLTZ)
    assert abs(T - 300.0) < 1e-9


def test_harness_fix_lists_in_distance_order():
    """Each fix list of the harness is in ascending q(r^2) = r^2 x scale
    truncated, scale = 2^(23 - ceil(log2(cf^2))) (ties in stencil order)."""
    md = ea.MD(5, ntypes=2, seed=8, T=600.0, disp_sigma=0.07)
    s = md.snapshot()
    nl = s.nlocal
    c2 = 6.5 ** 2
    scale = 2.0 ** (23 - int(np.ceil(np.log2(c2))))
    for i in range(nl):
        js = s.nbr_idx[s.nbr_off[i]:s.nbr_off[i + 1]]
        d = s.x[js] - s.x[i]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        q = np.floor(r2 * scale)
        assert np.all(q < 2 ** 23)
        assert np.all(np.diff(q) >= 0), f"atom {i}: fix list not in distance order"


def test_harness_lj_forces_match_brute_force():
    md = ea.MD(3, seed=2, T=300.0, disp_sigma=0.1)
    s = md.snapshot()
    md.lj(s)
    eps, sig, rc = 0.4, 2.3, 5.5
    for i in range(0, s.nlocal, 5):
        d = s.x[i] - s.x
        r2 = (d * d).sum(1)
        m = (r2 < rc * rc) & (r2 > 0)
        sr6 = (sig * sig / r2[m]) ** 3
        fp = 24 * eps * (2 * sr6 * sr6 - sr6) / r2[m]
        f = (d[m] * fp[:, None]).sum(0)
        assert np.allclose(s.f[i], f, rtol=1e-12, atol=1e-12)
    # Newton's third law over the periodic system
    assert np.abs(s.f[:s.nlocal].sum(0)).max() < 1e-10


def test_harness_pka_and_species():
    md = ea.MD(6, ntypes=4, seed=9, T=300.0, pka_eV=1000.0)
    s = md.snapshot()
    nl = s.nlocal
    counts = np.bincount(s.type[:nl])[1:]
    assert list(counts) == [nl // 4] * 4
    ke = 0.5 * inputs.MASS_CU * inputs.MVV2E * (s.v[:nl] ** 2).sum(1)
    assert abs(ke.max() - 1000.0) < 1e-9


def test_neighbour_rebuild_after_large_moves():
    md = ea.MD(4, seed=4, T=300.0)
    s = md.snapshot()
    assert not md.needs_rebuild(s)
    s.x[5, 0] += 1.01
    assert md.needs_rebuild(s)
    s.x[7, 1] -= 20.0  # leaves the box: the rebuild remaps it
    md.rebuild(s)
    s2 = md.snapshot()
    assert np.all(s2.x[:s2.nlocal] >= s2.boxlo) and np.all(s2.x[:s2.nlocal] < s2.boxhi)


def test_spline_jacobian_matches_finite_differences():
    """d(table)/d(knots) of the Akima build, used by the beta-knot gradient"""
    rng = np.random.default_rng(21)
    for n, dx in ((8, 0.3), (64, 20.0 / 63)):
        x = np.arange(n) * dx
        y = np.sqrt(0.02 * x / (1 + 0.5 * x)) + 0.01 * rng.standard_normal(n)
        J = ea.spline_jacobian(dx, y)
        for m in range(n):
            h = 1e-6 * max(abs(y[m]), 1e-3)
            yp, ym = y.copy(), y.copy()
            yp[m] += h
            ym[m] -= h
            fd = (ea.spline_build(dx, yp) - ea.spline_build(dx, ym)) / (2 * h)
            scale = np.abs(J[:, :, m]).max()
            assert np.abs(fd - J[:, :, m]).max() <= 1e-6 * scale, m
    # the top knot's constant piece: only y[n-1], with weight 1
    assert J[-1, 0, -1] == 1.0 and np.count_nonzero(J[-1]) == 1


def _kappa_file(tmp_path, name):
    if name == "kappa_synth3.npz":
        p = str(tmp_path / "synth3.kappa")
        inputs.synthetic_kappa(p, names=("A", "B", "C"), n_r=120, n_T=150)
        return p
    p = str(tmp_path / "Cu.kappa")
    with gzip.open(os.path.join(G, "Cu.kappa.gz"), "rb") as fi, open(p, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return p


def kappa_golden_tables(d):
    """(key, which, index) of every table in a kappa fixture"""
    out = [(f"e{e}_t{w}", w, e) for e in range(int(d["n_elements"])) for w in range(3)]
    return out + [(f"p{p}", 3, p) for p in range(int(d["n_pairs"]))]


@pytest.mark.parametrize("name", ["kappa_synth3.npz", "kappa_Cu.npz"])
def test_product_kappa_tables_match_reference_goldens(tmp_path, name):
    """.kappa reader (eph_kappa.h:55-150) against EPH_kappa compiled from the
    reference's header: rho(r), rho(r^2), the cumulative E(T) and the K(T)
    slots, bit for bit"""
    d = np.load(os.path.join(G, name))
    path = _kappa_file(tmp_path, name)
    with open(path, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == str(d["file_sha256"])
    ne, npair, rc, tmax = ea.kappa_info(path)
    assert (ne, npair) == (int(d["n_elements"]), int(d["n_pairs"]))
    assert rc == float(d["r_cutoff"]) and tmax == float(d["T_max"])
    for key, which, index in kappa_golden_tables(d):
        rows, dx = ea.kappa_table(path, index, which)
        assert dx == float(d[key + "_dx"]), key
        assert sha(rows) == str(d[key + "_sha256"]), key


def test_lammps_adaptor_compiles_against_the_lammps_api():
    """user-eph_amd/lammps/fix_eph_amd.cpp against the LAMMPS-shaped headers of
    tests/lammps_mock (the Fix API as the reference's fix_eph.cpp uses it);
    the harness driver too"""
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    mpi_inc = "/opt/conda/include"  # MPICH's headers (the adaptor calls MPI_Allreduce)
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        pytest.skip("no mpi.h")
    root = ROOT
    mock = os.path.join(ROOT, "tests", "lammps_mock")
    for src in (os.path.join(root, "user-eph_amd", "lammps", "fix_eph_amd.cpp"),
                os.path.join(mock, "mock_driver.cpp"), os.path.join(mock, "mock_mpi_driver.cpp")):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                            "-I" + mock, "-I" + os.path.join(root, "include"),
                            "-I" + os.path.join(root, "user-eph_amd", "lammps"),
                            "-I" + mpi_inc, src],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_bench_reads_the_committed_counter_summaries():
    """bench.py's roofline line takes the HBM traffic and the VALU instruction
    counts of the sweeps from the rocprofv3 summaries under profiles/: both
    parse, for both sweeps, with plausible magnitudes"""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    traffic = bench.counter_traffic(2048000)
    valu = bench.counter_valu(2048000)
    for k in ("k_density_w", "k_prl_forces"):
        assert 1e8 < traffic[k] < 1e10, (k, traffic.get(k))
        assert 1e3 < valu[k]["valu_per_wave"] < 1e5, (k, valu.get(k))
        assert 0.0 < valu[k]["valu_active"] <= 1.0
    assert bench.counter_traffic(4000) == {} and bench.counter_valu(4000) == {}
