"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has the USER-EPH source tree (the reference is not
available where the tests later run):

    make -C oracle            # builds oracle/_ref/eph_ref_driver from the
                              # reference's own headers
    python tests/golden/make_golden.py

Every expected value here was produced by the reference's code
(eph_spline.h, eph_beta.h, eph_linear.h, eph_kappa.h, eph_fdm.h) through
oracle/_ref/eph_ref_driver, printed as exact hex floats.  Inputs are either
regenerated deterministically by the tests (synthetic .beta / grid files,
whose sha256 is stored) or are data files the reference ships
(Examples/Beta/*.beta, Tests/EPH_FDM/*/Out_Ref/*,
Tests/EPH_atomic_run/thermalisation/Cu.kappa), copied here gzipped as data.
"""
import gzip
import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "user-eph_amd"))

import eph_oracle as eo  # noqa: E402
from eph_amd import inputs  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def file_sha(p):
    with open(p, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def hexlines(text):
    return [[float.fromhex(t) for t in l.split()] for l in text.strip().split("\n") if l.strip()]


def write_values(path, vals):
    with open(path, "w") as f:
        for v in vals:
            f.write(repr(float(v)) + "\n")


def ref_spline(tmp, dx, y):
    p = os.path.join(tmp, "y.txt")
    write_values(p, y)
    out = eo.ref_run("spline", repr(dx), p).strip().split("\n")
    n, inv = out[0].split()
    return np.array(hexlines("\n".join(out[1:]))), float.fromhex(inv)


def ref_beta(path):
    lines = eo.ref_run("beta", path).strip().split("\n")
    head = lines[0].split()
    res = {"n_elements": int(head[0]), "r_cutoff": float.fromhex(head[1]),
           "r_cutoff_sq": float.fromhex(head[2]), "rho_cutoff": float.fromhex(head[3]),
           "tables": []}
    k = 1
    for e in range(res["n_elements"]):
        assert lines[k].startswith("ELEMENT")
        k += 1
        tabs = []
        for _ in range(4):
            n, inv = lines[k].split()
            n = int(n)
            rows = np.array(hexlines("\n".join(lines[k + 1:k + 1 + n])))
            tabs.append((rows, float.fromhex(inv)))
            k += 1 + n
        res["tables"].append(tabs)
    return res


def beta_fixture(name, beta_path, full, tmp, sample_rows=64):
    r = ref_beta(beta_path)
    d = {"n_elements": r["n_elements"], "r_cutoff": r["r_cutoff"], "rho_cutoff": r["rho_cutoff"],
         "file_sha256": file_sha(beta_path)}
    rng = np.random.default_rng(7)
    for e, tabs in enumerate(r["tables"]):
        for w, (rows, inv) in enumerate(tabs):
            key = f"e{e}_t{w}"
            d[key + "_inv_dx"] = inv
            d[key + "_sha256"] = sha(rows)
            if full:
                d[key] = rows
            else:
                idx = np.unique(np.concatenate([[0, 1, 2, len(rows) - 2, len(rows) - 1],
                                                rng.integers(0, len(rows), sample_rows)]))
                d[key + "_idx"] = idx
                d[key + "_rows"] = rows[idx]
    # public-API evaluations: get_rho_r_sq, get_beta, get_alpha (eph_beta.h:164-198)
    rc2 = r["r_cutoff_sq"]
    xr = np.sort(rng.uniform(0, rc2, 200))
    xr[-1] = np.nextafter(rc2, 0)
    xb = np.sort(np.concatenate([rng.uniform(0, r["rho_cutoff"], 200),
                                 [0.0, r["rho_cutoff"], r["rho_cutoff"] * 1.01]]))
    for e in range(r["n_elements"]):
        for w, xs in ((1, xr), (2, xb), (3, xb)):
            p = os.path.join(tmp, "x.txt")
            write_values(p, xs)
            ys = [float.fromhex(l) for l in eo.ref_run("beta_eval", beta_path, e, w, p).split("\n")
                  if l.startswith(("0x", "-0x"))]
            d[f"eval_e{e}_w{w}_x"] = xs
            d[f"eval_e{e}_w{w}_y"] = np.array(ys)
    np.savez_compressed(os.path.join(HERE, name), **d)
    print("wrote", name)


def ref_fdm(grid_path, dt, nsteps, every, energies=None, tmp=None):
    args = ["fdm", grid_path, repr(dt), nsteps, every]
    if energies is not None:
        p = os.path.join(tmp, "e.txt")
        write_values(p, np.asarray(energies).ravel())
        args.append(p)
    text = eo.ref_run(*args)
    snaps = []
    for block in text.split("STEP ")[1:]:
        lines = block.strip().split("\n")
        snaps.append([float.fromhex(l) for l in lines[1:]])
    return np.array(snaps)


def fdm_fixtures(tmp):
    rng = np.random.default_rng(11)
    cases = {}
    # (a) periodic 6x5x4, random T, constant parameters, energy every step
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    box = (0.0, 6.0, -1.0, 4.0, 2.0, 10.0)
    cases["fdm_periodic"] = dict(nx=nx, ny=ny, nz=nz, box=box, T=rng.uniform(100, 900, n),
                                 S=np.zeros(n), rho=np.full(n, 1.0), C=np.full(n, 3e-6 * 47),
                                 kappa=np.full(n, 0.1), flag=np.ones(n, int), tdyn=np.zeros(n, int),
                                 steps=1, dt=1e-3, nsteps=6, every=1,
                                 energies=np.column_stack([rng.uniform(0, 6, 20), rng.uniform(-1, 4, 20),
                                                           rng.uniform(2, 10, 20), rng.normal(0, 1e-3, 20)]))
    # (b) heterogeneous parameters, walls (flag 2), fixed cells (flag 0), sources
    nx, ny, nz = 7, 6, 5
    n = nx * ny * nz
    flag = np.ones(n, int)
    flag[rng.choice(n, 20, replace=False)] = 2
    flag[rng.choice(n, 10, replace=False)] = 0
    cases["fdm_walls"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 7.0, 0.0, 6.0, 0.0, 5.0),
                              T=rng.uniform(200, 2000, n), S=rng.normal(0, 1e-2, n),
                              rho=rng.uniform(0.5, 1.5, n), C=rng.uniform(1e-4, 3e-4, n),
                              kappa=rng.uniform(0.05, 0.2, n), flag=flag, tdyn=np.zeros(n, int),
                              steps=2, dt=5e-4, nsteps=5, every=1, energies=None)
    # (c) unstable window: r in (0.5, 0.8) gives one sub-step at lambda > 1/6
    nx = ny = nz = 8
    n = nx ** 3
    a = 3.615
    T0 = 300.0 + 2700.0 * np.exp(-((np.indices((nz, ny, nx))[::-1].reshape(3, -1).T + 0.5 - 4) ** 2)
                                  .sum(1) * a * a / (2 * 10.0 ** 2))
    cases["fdm_unstable"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 8 * a, 0.0, 8 * a, 0.0, 8 * a),
                                 T=T0, S=np.zeros(n), rho=np.ones(n), C=np.full(n, 3e-6),
                                 kappa=np.full(n, 0.1), flag=np.ones(n, int), tdyn=np.zeros(n, int),
                                 steps=1, dt=1e-4, nsteps=12, every=1, energies=None)
    # (d) temperature-dependent C_e(T), kappa(T) with E_e(T) energy stepping
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    Tgrid = np.arange(400) * 10.0
    Ce_T = 1e-5 + 2e-8 * Tgrid + 1e-6 * np.sin(Tgrid / 300.0) ** 2
    ka_T = 0.05 + 1e-5 * Tgrid
    tdyn = (rng.uniform(size=n) < 0.6).astype(int)
    cases["fdm_tdyn"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 5.0, 0.0, 4.0, 0.0, 3.0),
                             T=rng.uniform(300, 1500, n), S=rng.normal(0, 1e-3, n), rho=np.ones(n),
                             C=np.full(n, 2e-5), kappa=np.full(n, 0.06), flag=np.ones(n, int), tdyn=tdyn,
                             steps=1, dt=1e-3, nsteps=4, every=1, energies=None,
                             params=(10.0, Ce_T, ka_T))
    for name, c in cases.items():
        gp = os.path.join(tmp, name + ".in")
        params = "NULL"
        if "params" in c:
            pp = os.path.join(tmp, name + ".params")
            inputs.write_params(pp, *c["params"])
            params = pp
        inputs.write_grid(gp, c["nx"], c["ny"], c["nz"], c["box"], c["T"], c["S"], c["rho"], c["C"],
                          c["kappa"], c["flag"], c["tdyn"], steps=c["steps"], params=params)
        snaps = ref_fdm(gp, c["dt"], c["nsteps"], c["every"], c["energies"], tmp)
        d = {k: np.asarray(v) for k, v in c.items() if k not in ("params", "energies")}
        if c["energies"] is not None:
            d["energies"] = np.asarray(c["energies"])
        if "params" in c:
            d["params_dT"] = c["params"][0]
            d["params_C"] = c["params"][1]
            d["params_kappa"] = c["params"][2]
        d["T_steps"] = snaps
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, snaps.shape)


def fdm_save_fixtures(tmp):
    """save_state (T.restart) and save_temperature (heat map) files of the
    compiled EPH_FDM after a few solves, and its resume from the state file:
    fdm_save.npz holds each case's input files, the two written files byte for
    byte, and the T of the resumed grid.  The parameter file is named
    relative to the working directory, as fix eph users give it."""
    import subprocess
    rng = np.random.default_rng(23)
    d = {}
    cases = {}
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    cases["periodic"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 6.0, -1.0, 4.0, 2.0, 10.0),
                             T=rng.uniform(100, 900, n), S=np.zeros(n), rho=np.ones(n),
                             C=np.full(n, 1.41e-4), kappa=np.full(n, 0.1), flag=np.ones(n, int),
                             tdyn=np.zeros(n, int), steps=1, dt=1e-3, nsteps=5, nmore=3,
                             energies=np.column_stack([rng.uniform(0, 6, 16), rng.uniform(-1, 4, 16),
                                                       rng.uniform(2, 10, 16), rng.normal(0, 1e-3, 16)]))
    nx, ny, nz = 7, 6, 5
    n = nx * ny * nz
    flag = np.ones(n, int)
    flag[rng.choice(n, 20, replace=False)] = 2
    flag[rng.choice(n, 10, replace=False)] = 0
    cases["walls"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 7.0, 0.0, 6.0, 0.0, 5.0),
                          T=rng.uniform(200, 2000, n), S=rng.normal(0, 1e-2, n),
                          rho=rng.uniform(0.5, 1.5, n), C=rng.uniform(1e-4, 3e-4, n),
                          kappa=rng.uniform(0.05, 0.2, n), flag=flag, tdyn=np.zeros(n, int),
                          steps=2, dt=5e-4, nsteps=4, nmore=2, energies=None)
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    Tg = np.arange(400) * 10.0
    cases["tdyn"] = dict(nx=nx, ny=ny, nz=nz, box=(0.0, 5.0, 0.0, 4.0, 0.0, 3.0),
                         T=rng.uniform(300, 1500, n), S=rng.normal(0, 1e-3, n), rho=np.ones(n),
                         C=np.full(n, 2e-5), kappa=np.full(n, 0.06), flag=np.ones(n, int),
                         tdyn=(rng.uniform(size=n) < 0.6).astype(int), steps=1, dt=1e-3, nsteps=3,
                         nmore=2, energies=None,
                         params=(10.0, 1e-5 + 2e-8 * Tg + 1e-6 * np.sin(Tg / 300.0) ** 2,
                                 0.05 + 1e-5 * Tg))
    for name, c in cases.items():
        work = os.path.join(tmp, "save_" + name)
        os.makedirs(work)
        params = "NULL"
        if "params" in c:
            params = name + ".params"
            inputs.write_params(os.path.join(work, params), *c["params"])
            with open(os.path.join(work, params), "rb") as f:
                d[f"{name}_params"] = np.frombuffer(f.read(), np.uint8)
        inputs.write_grid(os.path.join(work, name + ".in"), c["nx"], c["ny"], c["nz"], c["box"],
                          c["T"], c["S"], c["rho"], c["C"], c["kappa"], c["flag"], c["tdyn"],
                          steps=c["steps"], params=params)
        efile = "NULL"
        if c["energies"] is not None:
            efile = "e.txt"
            write_values(os.path.join(work, efile), np.asarray(c["energies"]).ravel())
            d[f"{name}_energies"] = np.asarray(c["energies"])
        out = subprocess.run([eo.REF_DRIVER, "fdm_save", name + ".in", repr(c["dt"]), str(c["nsteps"]),
                              efile, "T.restart", "Te_out", "7", str(c["nmore"])],
                             cwd=work, check=True, capture_output=True, text=True).stdout
        resumed = [float.fromhex(l) for l in out.split("STEP ")[1].strip().split("\n")[1:]]
        for key, fn in (("grid", name + ".in"), ("state", "T.restart"), ("heat", "Te_out_000007")):
            with open(os.path.join(work, fn), "rb") as f:
                d[f"{name}_{key}"] = np.frombuffer(f.read(), np.uint8)
        d[f"{name}_dt"] = c["dt"]
        d[f"{name}_nsteps"] = c["nsteps"]
        d[f"{name}_nmore"] = c["nmore"]
        d[f"{name}_resume_T"] = np.array(resumed)
    d["cases"] = np.array(list(cases))
    np.savez_compressed(os.path.join(HERE, "fdm_save.npz"), **d)
    print("wrote fdm_save.npz")


def legacy_grid_fixture(tmp):
    """Examples/FDM/T.in is in the older grid layout (no parameter-file line,
    nine columns per cell).  The compiled reference reads only the current
    layout, so it runs the same grid converted to it (NULL parameter file,
    T_dyn 0): fdm_legacy.npz holds the original file and T after each of four
    solves with a few energies inserted before every solve."""
    src = os.path.join(REF, "Examples/FDM/T.in")
    with open(src) as f:
        lines = f.read().split("\n")
    conv = lines[:7] + ["NULL"] + [l + " 0" for l in lines[7:] if l.strip()]
    cp = os.path.join(tmp, "legacy_converted.in")
    with open(cp, "w") as f:
        f.write("\n".join(conv) + "\n")
    rng = np.random.default_rng(31)
    e = np.column_stack([rng.uniform(0, 17.6, 6), rng.uniform(0, 17.6, 6), rng.uniform(0, 17.6, 6),
                         rng.uniform(0.5, 2.0, 6)])
    snaps = ref_fdm(cp, 1.0, 4, 1, e, tmp)
    with open(src, "rb") as f:
        raw = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "fdm_legacy.npz"), legacy_file=raw, energies=e, dt=1.0,
                        T_steps=snaps)
    print("wrote fdm_legacy.npz", snaps.shape)


def outref_fixture():
    d = {}
    for t in (1, 2, 3, 4):
        base = os.path.join(REF, "Tests", "EPH_FDM", f"Test{t}", "Out_Ref")
        for snap in (0, 1, 2, 4, 8, 16, 32, 64):
            p = os.path.join(base, f"T_out_{snap:06d}")
            if not os.path.exists(p):
                continue
            rows = np.loadtxt(p, skiprows=1)
            d[f"test{t}_snap{snap}"] = rows[:, 3]
            if snap == 0:
                d[f"test{t}_x"] = rows[:, 0]
    np.savez_compressed(os.path.join(HERE, "fdm_outref.npz"), **d)
    print("wrote fdm_outref.npz")


def linear_fixture(tmp):
    rng = np.random.default_rng(5)
    dT = 10.0
    C = 1e-5 + 2e-8 * np.arange(300) * dT
    E = np.zeros_like(C)
    for i in range(1, len(C)):  # cumulative table as eph_fdm.h:98-103 builds it
        E[i] = E[i - 1] + C[i] * dT
    xs_fwd = np.sort(rng.uniform(0, (len(C) - 1.5) * dT, 300))
    ys = np.sort(rng.uniform(E[0], E[-1] * 0.999, 300))
    yp = os.path.join(tmp, "ly.txt"); write_values(yp, E)
    out = {}
    for tag, xs in (("fwd", xs_fwd), ("rev", ys)):
        xp = os.path.join(tmp, "lx.txt"); write_values(xp, xs)
        rows = np.array(hexlines(eo.ref_run("linear", repr(dT), yp, xp)))
        out[f"{tag}_x"] = xs
        out[f"{tag}_op"] = rows[:, 0]
        out[f"{tag}_rev"] = rows[:, 1]
    np.savez_compressed(os.path.join(HERE, "linear.npz"), dT=dT, y=E, **out)
    print("wrote linear.npz")


def ref_kappa(path):
    lines = eo.ref_run("kappa", path).strip().split("\n")
    head = lines[0].split()
    res = {"n_elements": int(head[0]), "n_pairs": int(head[1]), "r_cutoff": float.fromhex(head[2]),
           "r_cutoff_sq": float.fromhex(head[3]), "T_max": float.fromhex(head[4]),
           "elements": [], "pairs": []}
    k = 1
    for e in range(res["n_elements"]):
        assert lines[k].startswith("ELEMENT")
        k += 1
        tabs = []
        for _ in range(3):  # rho(r), rho(r^2) splines, E(T) linear table
            n, inv = lines[k].split()
            n = int(n)
            tabs.append((np.array(hexlines("\n".join(lines[k + 1:k + 1 + n]))), float.fromhex(inv)))
            k += 1 + n
        res["elements"].append(tabs)
    for p in range(res["n_pairs"]):
        assert lines[k].startswith("PAIR")
        n, dx = lines[k + 1].split()
        n = int(n)
        res["pairs"].append((np.array(hexlines("\n".join(lines[k + 2:k + 2 + n]))), float.fromhex(dx)))
        k += 2 + n
    return res


def kappa_fixture(name, path, full, sample_rows=64):
    r = ref_kappa(path)
    d = {"n_elements": r["n_elements"], "n_pairs": r["n_pairs"], "r_cutoff": r["r_cutoff"],
         "T_max": r["T_max"], "file_sha256": file_sha(path)}
    rng = np.random.default_rng(11)
    tables = [(f"e{e}_t{w}", rows, inv) for e, tabs in enumerate(r["elements"])
              for w, (rows, inv) in enumerate(tabs)]
    tables += [(f"p{p}", rows, dx) for p, (rows, dx) in enumerate(r["pairs"])]
    for key, rows, inv in tables:
        d[key + "_dx"] = inv
        d[key + "_sha256"] = sha(rows)
        if full:
            d[key] = rows
        else:
            idx = np.unique(np.concatenate([[0, 1, 2, len(rows) - 2, len(rows) - 1],
                                            rng.integers(0, len(rows), sample_rows)]))
            d[key + "_idx"] = idx
            d[key + "_rows"] = rows[idx]
    np.savez_compressed(os.path.join(HERE, name), **d)
    print("wrote", name)


def kappa_fixtures(tmp):
    p3 = os.path.join(tmp, "synth3.kappa")
    inputs.synthetic_kappa(p3, names=("A", "B", "C"), n_r=120, n_T=150)
    kappa_fixture("kappa_synth3.npz", p3, True)
    src = os.path.join(REF, "Tests/EPH_atomic_run/thermalisation/Cu.kappa")
    with open(src, "rb") as fi, gzip.open(os.path.join(HERE, "Cu.kappa.gz"), "wb", 9) as fo:
        shutil.copyfileobj(fi, fo)
    kappa_fixture("kappa_Cu.npz", src, False)


def main():
    if not eo.ref_available():
        sys.exit("oracle/_ref/eph_ref_driver missing: run `make -C oracle` where the reference exists")
    tmp = tempfile.mkdtemp()
    if "--only" in sys.argv:  # e.g. --only kappa: write that family alone
        try:
            {"kappa": kappa_fixtures, "fdm_save": fdm_save_fixtures,
             "fdm_legacy": legacy_grid_fixture}[
                sys.argv[sys.argv.index("--only") + 1]](tmp)
        finally:
            shutil.rmtree(tmp)
        return
    try:
        # splines: the sin case of Tests/EPH_Spline/test.cpp:13-25 and random data
        x = np.arange(1001) * 0.01
        rows, inv = ref_spline(tmp, 0.01, np.sin(x))
        d = {"sin_dx": 0.01, "sin_y": np.sin(x), "sin_coeffs": rows}
        rng = np.random.default_rng(3)
        for k in range(4):
            n = int(rng.integers(5, 120)); dx = float(rng.uniform(0.01, 2.0))
            y = rng.normal(size=n).cumsum()
            rows, _ = ref_spline(tmp, dx, y)
            d[f"rand{k}_dx"] = dx; d[f"rand{k}_y"] = y; d[f"rand{k}_coeffs"] = rows
        np.savez_compressed(os.path.join(HERE, "spline.npz"), **d)
        print("wrote spline.npz")

        # .beta tables: synthetic 1-element and 4-element files, and reference data
        p1 = os.path.join(tmp, "fcc1.beta"); inputs.workload_beta(p1, 1)
        beta_fixture("beta_synth1.npz", p1, True, tmp)
        p4 = os.path.join(tmp, "alloy4.beta"); inputs.workload_beta(p4, 4)
        beta_fixture("beta_synth4.npz", p4, False, tmp)
        for src in ("Examples/Beta/Ni.beta", "Examples/Beta/Ni_model_4.beta"):
            name = os.path.basename(src)
            gz = os.path.join(HERE, name + ".gz")
            with open(os.path.join(REF, src), "rb") as fi, gzip.open(gz, "wb", 9) as fo:
                shutil.copyfileobj(fi, fo)
            beta_fixture("beta_" + name.replace(".beta", "") + ".npz", os.path.join(REF, src), False, tmp)

        linear_fixture(tmp)
        kappa_fixtures(tmp)
        fdm_fixtures(tmp)
        fdm_save_fixtures(tmp)
        legacy_grid_fixture(tmp)
        outref_fixture()
        kat = {"source": "Random123 kat_vectors, philox4x32_10 (Salmon et al., SC'11)",
               "vectors": [
                   {"ctr": [0, 0, 0, 0], "key": [0, 0],
                    "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
                   {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
                    "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
                   {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                    "key": [0xa4093822, 0x299f31d0],
                    "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]}]}
        with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
            json.dump(kat, f, indent=1)
    finally:
        shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
