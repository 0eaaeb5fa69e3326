"""GPU: the LAMMPS adaptor over several MPI ranks (tests/lammps_mock's
mock_mpi_driver under MPICH's mpiexec, every rank on the one GPU) against the
same system run on one rank by mock_driver.

The system is cut into slabs along x.  Each rank holds the atoms of its slab
as locals and, as ghosts, every periodic image of another rank's atom (or of
its own, shifted) within the list cutoff of its sub-box, grouped by owning
rank as LAMMPS's Comm delivers them.  Its neighbour lists are the global lists
of its atoms in the same order, renumbered, so that each atom's sums run in the
single-rank order.  Per step the harness refreshes the ghosts' x and v from
their owners over MPI, and the adaptor exchanges xi, rho and w through
comm->forward_comm (the XI, RHO and WI halos of fix_eph.cpp:863-871) and sums
the grid's deposits and the energy tally with MPI_Allreduce (eph_fdm.cpp's
reduce_and_update and fix_eph.cpp:953-958).  What differs from one rank is the
order of each grid cell's deposit sum (the atoms of a cell split over ranks),
so the results agree to rounding, and on one rank through mpiexec bitwise.
"""
import itertools
import os
import shutil
import subprocess

import numpy as np
import pytest

from eph_amd import inputs
import test_gpu_adaptor as single

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MPI_DRIVER = os.path.join(HERE, "lammps_mock", "_build", "mock_mpi_driver")
MPIEXEC = "/opt/conda/bin/mpiexec"
# relative error allowed against the one-rank run (the deposit order of a cell
# split over ranks perturbs T_e by an ulp; 10 steps grow it to ~1e-14)
RTOL = 1e-11


def _decompose(s, L, nranks):
    """per rank: its locals (global indices, ascending), its ghosts as
    (owner rank, global atom, image shift) in rank-block order, and its
    renumbered neighbour lists"""
    n = s.nlocal
    x0 = s.x[:n]
    rank_of = np.minimum((x0[:, 0] // (L / nranks)).astype(np.int64), nranks - 1)
    locals_ = [np.nonzero(rank_of == r)[0] for r in range(nranks)]
    lidx = np.empty(n, np.int64)
    for r in range(nranks):
        lidx[locals_[r]] = np.arange(len(locals_[r]))
    Lv = np.array([L, L, L])
    ranks = []
    for r in range(nranks):
        lo = np.array([r * L / nranks, 0.0, 0.0]) - s.cut
        hi = np.array([(r + 1) * L / nranks if r < nranks - 1 else L, L, L]) + s.cut
        blocks = []
        for q in range(nranks):
            block = []
            for j in locals_[q]:
                for sh in itertools.product((-1, 0, 1), repeat=3):
                    if q == r and sh == (0, 0, 0):
                        continue
                    p = x0[j] + np.array(sh, float) * Lv
                    if np.all(p >= lo) and np.all(p < hi):
                        block.append((int(j), sh))
            blocks.append(block)
        nl = len(locals_[r])
        slot = {}
        for g, (j, sh) in enumerate(b for blk in blocks for b in blk):
            slot[(j, sh)] = nl + g
        off, idx = [0], []
        for gi in locals_[r]:
            for k in s.nbr_idx[s.nbr_off[gi]:s.nbr_off[gi + 1]]:
                if k < n:
                    key = (int(k), (0, 0, 0))
                else:
                    sh = tuple(int(t) for t in np.rint(s.ghost_shift[k - n] / L))
                    key = (int(s.ghost_owner[k - n]), sh)
                if key[1] == (0, 0, 0) and rank_of[key[0]] == r:
                    idx.append(int(lidx[key[0]]))
                else:
                    idx.append(slot[key])
            off.append(len(idx))
        ranks.append(dict(locals=locals_[r], blocks=blocks, off=np.array(off, np.int32),
                          idx=np.array(idx, np.int32)))
    for r in range(nranks):  # what rank r sends to rank q: its atoms in q's block from r
        ranks[r]["send"] = [np.array([lidx[j] for j, _ in ranks[q]["blocks"][r]], np.int32)
                            for q in range(nranks)]
    return ranks


def _write_rank_inputs(prefix, s, L, ranks, ntypes, nsteps, dt, masses):
    nranks = len(ranks)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
    Lv = np.array([L, L, L])
    for r, d in enumerate(ranks):
        ghosts = [b for blk in d["blocks"] for b in blk]
        own = np.concatenate([d["locals"], np.array([j for j, _ in ghosts], np.int64)])
        shift = np.array([np.array(sh, float) * Lv for _, sh in ghosts]).reshape(-1, 3)
        nl, ng = len(d["locals"]), len(ghosts)
        x = s.x[own].copy()
        x[nl:] = s.x[own[nl:]] + shift
        v = s.v[own]
        recv_n = np.array([len(blk) for blk in d["blocks"]], np.int32)
        recv_first = nl + np.concatenate([[0], np.cumsum(recv_n)[:-1]])
        with open(f"{prefix}.{r}.bin", "wb") as f:
            i32([nl, ng, ntypes, nsteps, len(d["idx"]), nranks]).tofile(f)
            np.array([dt, inputs.BOLTZ, inputs.FTM2V, 0.0, 0.0, 0.0, L, L, L],
                     np.float64).tofile(f)
            np.array([0.0] + list(masses), np.float64).tofile(f)
            np.ascontiguousarray(x, np.float64).tofile(f)
            np.ascontiguousarray(v, np.float64).tofile(f)
            for arr in (s.type, s.mask, s.tag):
                i32(arr[own]).tofile(f)
            i32(recv_first).tofile(f)
            i32(recv_n).tofile(f)
            i32([len(sl) for sl in d["send"]]).tofile(f)
            for sl in d["send"]:
                i32(sl).tofile(f)
            np.ascontiguousarray(shift, np.float64).tofile(f)
            i32(d["off"]).tofile(f)
            i32(d["idx"]).tofile(f)


def _read_rank_output(prefix, r, nl, ncols):
    raw = open(f"{prefix}.{r}.bin", "rb").read()
    tag = np.frombuffer(raw[:4 * nl], np.int32)
    d = np.frombuffer(raw[4 * nl:], np.float64)
    assert d.size == 9 * nl + 2 + ncols * nl, (d.size, nl, ncols)
    return (tag, d[:3 * nl].reshape(nl, 3), d[3 * nl:6 * nl].reshape(nl, 3),
            d[6 * nl:9 * nl].reshape(nl, 3), d[9 * nl], d[9 * nl + 1],
            d[9 * nl + 2:].reshape(nl, ncols))


def _mpiexec(nranks, fin, fout, lmp_args):
    cmd = ["timeout", "-k", "10", "170", MPIEXEC, "-launcher", "fork", "-hosts", "127.0.0.1",
           "-n", str(nranks), MPI_DRIVER, fin, fout] + lmp_args
    print("mpiexec:", nranks, "ranks", flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=200)


MPI_CASES = [("eph", 1, "7"), ("eph", 2, "0x37"), ("eph/coloured", 2, "7"),
             ("eph/coloured/exp", 2, "7")]


def _prerequisites():
    if not os.path.exists(MPIEXEC):
        pytest.skip("no MPICH mpiexec in this image")
    assert os.path.exists(MPI_DRIVER), "build the harness first (__graft_entry__.build())"
    assert os.path.exists(single.DRIVER), "build the harness first (__graft_entry__.build())"
    assert shutil.which("timeout")


@pytest.mark.parametrize("style,species,flags", MPI_CASES)
def test_adaptor_over_ranks_matches_one_rank(tmp_path, style, species, flags, parity_report):
    _prerequisites()
    s, L = single._system(20 + species, species)
    args = single._args(tmp_path, style, species, flags)
    lmp_args = list(args)
    lmp_args[2] = style + "/amd"
    dt, nsteps = 1e-3, 10
    masses = [inputs.MASS_CU] * species
    n = s.nlocal

    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    single._write_input(fin, s, species, nsteps, dt, masses, L)
    r = subprocess.run([single.DRIVER, fin, fout] + lmp_args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    ref = np.fromfile(fout, np.float64)
    ncols = (ref.size - 9 * n - 2) // n
    x1, v1, f1 = (ref[k * 3 * n:(k + 1) * 3 * n].reshape(n, 3) for k in range(3))
    Ee1, Te1 = ref[9 * n], ref[9 * n + 1]
    arr1 = ref[9 * n + 2:].reshape(n, ncols)

    for nranks in (1, 2, 3):
        ranks = _decompose(s, L, nranks)
        pin, pout = str(tmp_path / f"r{nranks}_in"), str(tmp_path / f"r{nranks}_out")
        _write_rank_inputs(pin, s, L, ranks, species, nsteps, dt, masses)
        r = _mpiexec(nranks, pin, pout, lmp_args)
        assert r.returncode == 0, r.stderr[-3000:]
        x, v, f, arr = (np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3)),
                        np.empty((n, ncols)))
        Ees, Tes = [], []
        for rk, d in enumerate(ranks):
            tag, xr, vr, fr, Ee, Te, ar = _read_rank_output(pout, rk, len(d["locals"]), ncols)
            assert np.array_equal(tag, s.tag[d["locals"]])
            g = tag - 1
            x[g], v[g], f[g], arr[g] = xr, vr, fr, ar
            Ees.append(Ee)
            Tes.append(Te)
        assert len(set(Ees)) == 1 and len(set(Tes)) == 1  # every rank solves the same grid
        if nranks == 1:
            for a, b in ((x, x1), (v, v1), (f, f1), (arr, arr1)):
                assert np.array_equal(a, b)
            assert Ees[0] == Ee1 and Tes[0] == Te1
            continue
        err = {}
        for name, a, b in (("x", x, x1), ("v", v, v1), ("f", f, f1), ("array", arr, arr1)):
            # (flags 0x37 leave f at zero)
            err[name] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
            assert err[name] <= RTOL, (name, err[name])
        err["Ee"] = abs(Ees[0] - Ee1) / abs(Ee1)
        err["Te"] = abs(Tes[0] - Te1) / abs(Te1)
        assert err["Ee"] <= RTOL and err["Te"] <= RTOL, err
        parity_report.append(f"adaptor {style} species={species} over {nranks} ranks, "
                             f"{nsteps} steps vs one rank: "
                             + ", ".join(f"{k} {e:.1e}" for k, e in err.items()))


def test_adaptor_over_ranks_rejects_the_styles_it_cannot_split(tmp_path):
    """eph/atomic and the model-1/2 forms of eph need halos the split entry
    points do not provide: the adaptor stops with a message"""
    _prerequisites()
    s, L = single._system(3, 1)
    ranks = _decompose(s, L, 2)
    pin, pout = str(tmp_path / "in"), str(tmp_path / "out")
    _write_rank_inputs(pin, s, L, ranks, 1, 1, 1e-3, [inputs.MASS_CU])
    pb = str(tmp_path / "a.beta")
    names = inputs.workload_beta(pb, 1)
    args = inputs.fix_args(pb, names, seed=77, flags="7", grid=(3, 3, 3), model=1,
                           T_out=str(tmp_path / "Te_out"))
    args[2] = "eph/amd"
    r = _mpiexec(2, pin, pout, args)
    assert r.returncode != 0
    assert "over several MPI ranks" in r.stderr
