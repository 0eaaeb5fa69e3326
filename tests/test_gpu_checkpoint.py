"""GPU: full-precision checkpoints (eph_checkpoint_save/load, eph_md_checkpoint_save/load).

The reference keeps only the T_e grid across runs, in T.restart with six
significant digits (eph_fdm.h:226-265), so its resumed run is not the run that
went on.  These binary checkpoints must give that run's bits: a run split at a
checkpoint, the second half in fresh contexts, ends where the uninterrupted run
ends (x, v, f, T_e, the tally, the per-atom state).  The per-atom state follows
the atoms' tags, and files of another style or grid are refused.
"""
import copy

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
import test_gpu_adaptor as adaptor
from drive import workload_files

pytestmark = pytest.mark.gpu

STYLES = [("eph", 2), ("eph/coloured/exp", 2), ("eph/atomic", 1)]


def _fix(args, species, L, dt):
    f = ea.FixEPH(args, species, [0, 0, 0], [L] * 3, dt, inputs.BOLTZ, inputs.FTM2V,
                  masses=[inputs.MASS_CU] * species)
    f.init()
    return f


def _steps(fix, s, first, last):
    for n in range(first, last + 1):
        fix.initial_integrate(s)
        s.update_ghosts()
        s.f[:] = 0.0
        fix.post_force(s, step=n)
        fix.final_integrate(s)
        fix.end_of_step(s, n)


def _outputs(fix, s, style):
    n = s.nlocal
    out = [s.x[:n].copy(), s.v[:n].copy(), s.f[:n].copy(), fix.peratom(),
           np.array([fix.compute_vector(0), fix.compute_vector(1)])]
    if style != "eph/atomic":
        out.append(fix.grid().get("T"))
    return out


@pytest.mark.parametrize("style,species", STYLES)
def test_fix_checkpoint_resumes_bit_for_bit(tmp_path, style, species):
    print(f"checkpoint: {style}", flush=True)
    s0, L = adaptor._system(31, species)
    args = adaptor._args(tmp_path, style, species, "7")
    dt, n1, n2 = 1e-3, 6, 6
    ck = str(tmp_path / "fix.ckpt")
    # uninterrupted
    sA = copy.deepcopy(s0)
    A = _fix(args, species, L, dt)
    A.set_atoms(sA)
    _steps(A, sA, 1, n1 + n2)
    want = _outputs(A, sA, style)
    # split: the second half in a fresh context that loads the checkpoint
    sB = copy.deepcopy(s0)
    B = _fix(args, species, L, dt)
    B.set_atoms(sB)
    _steps(B, sB, 1, n1)
    B.checkpoint_save(ck)
    B.close()
    C = _fix(args, species, L, dt)
    C.set_atoms(sB)
    C.checkpoint_load(ck)
    _steps(C, sB, n1 + 1, n1 + n2)
    got = _outputs(C, sB, style)
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), (style, k)


@pytest.mark.parametrize("style,species", [("eph/coloured/exp", 2), ("eph/atomic", 1)])
def test_checkpoint_state_follows_the_tags(tmp_path, style, species):
    """loaded into atoms in another order, each atom gets its own state back"""
    s0, L = adaptor._system(32, species)
    args = adaptor._args(tmp_path, style, species, "7")
    ck = str(tmp_path / "fix.ckpt")
    A = _fix(args, species, L, 1e-3)
    A.set_atoms(s0)
    _steps(A, s0, 1, 4)
    if style == "eph/atomic":
        before = [A.atomic_energy()]
    else:
        before = [A.filter_state(w) for w in (0, 1)]
    A.checkpoint_save(ck)
    perm = np.random.default_rng(9).permutation(s0.nlocal)
    sP = copy.deepcopy(s0)
    adaptor._permute(sP, perm)
    sP.update_ghosts()
    B = _fix(args, species, L, 1e-3)
    B.set_atoms(sP)
    B.checkpoint_load(ck)
    if style == "eph/atomic":
        after = [B.atomic_energy()]
    else:
        after = [B.filter_state(w) for w in (0, 1)]
    assert any(np.any(b != 0) for b in before)
    for a, b in zip(after, before):
        assert np.array_equal(a, b[perm])
    assert B.compute_vector(0) == A.compute_vector(0)


def test_checkpoint_refuses_other_files(tmp_path):
    s0, L = adaptor._system(33, 1)
    ck = str(tmp_path / "fix.ckpt")
    a = _fix(adaptor._args(tmp_path, "eph", 1, "7"), 1, L, 1e-3)
    a.set_atoms(s0)
    a.checkpoint_save(ck)
    b = _fix(adaptor._args(tmp_path, "eph/coloured/exp", 1, "7"), 1, L, 1e-3)
    b.set_atoms(s0)
    with pytest.raises(ea.EphError, match="another fix style"):
        b.checkpoint_load(ck)
    junk = tmp_path / "junk"
    junk.write_text("T.restart is text\n")
    with pytest.raises(ea.EphError, match="not an eph checkpoint"):
        a.checkpoint_load(str(junk))
    # a per-atom count beyond the file's end, and a file cut short inside the
    # per-atom block: refused before anything is allocated or changed
    raw = open(ck, "rb").read()
    dims = np.frombuffer(raw[56:72], np.int32)
    off = 72 + 8 * int(dims[0]) * int(dims[1]) * int(dims[2])
    assert int(np.frombuffer(raw[off:off + 4], np.int32)[0]) == s0.nlocal
    big = bytearray(raw)
    big[off:off + 4] = np.int32(1 << 30).tobytes()
    (tmp_path / "big").write_bytes(bytes(big))
    (tmp_path / "cut").write_bytes(raw[:off + 8 + 4 * (s0.nlocal // 2)])
    a.grid().set("T", a.grid().get("T") + 1.0)
    T_before = a.grid().get("T")
    for bad in ("big", "cut"):
        with pytest.raises(ea.EphError, match="ends early"):
            a.checkpoint_load(str(tmp_path / bad))
        assert np.array_equal(a.grid().get("T"), T_before), bad
    a.checkpoint_load(ck)
    assert np.array_equal(a.grid().get("T") + 1.0, T_before)


def test_harness_checkpoint_resumes_bit_for_bit(tmp_path):
    """device-resident run (pair forces, device rebuilds, fix hooks): 30 steps
    straight against 15 + checkpoint + 15 in a fresh harness and fix; the PKA
    makes the lists rebuild in both halves"""
    print("checkpoint: harness fcc 7^3", flush=True)
    cells, grid, dt, species = 7, 7, 5e-4, 2
    cfg = dict(a=3.615, ntypes=species, seed=17, T=300.0, pka_eV=300.0)
    beta, names, tin = workload_files(str(tmp_path), "ck", species, cells, 3.615, grid,
                                      hot_spot=(2000.0, 4.0))
    args = inputs.fix_args(beta, names, seed=71, flags="7", grid=(grid,) * 3, T_infile=tin)
    ck = str(tmp_path / "md.ckpt")

    def fresh():
        md = ea.MD(cells, **cfg)
        inf = md.info()
        fix = ea.FixEPH(args, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ,
                        inputs.FTM2V, masses=[inputs.MASS_CU] * species)
        fix.init()
        md.attach(fix)
        return md, fix

    def steps(md, fix, first, n):
        # one step per call, with the density sweep's candidate-list path of each
        rebuilds, paths = 0, []
        for s in range(first, first + n):
            rebuilds += md.run(fix, 1, first_step=s)["rebuilds"]
            paths.append(fix.candidate_stats())
        return rebuilds, paths

    mdA, fA = fresh()
    ra = mdA.run(fA, 15, first_step=1)["rebuilds"]
    rb, pathsA = steps(mdA, fA, 16, 15)
    mdB, fB = fresh()
    r1 = mdB.run(fB, 15, first_step=1)["rebuilds"]
    mdB.checkpoint_save(fB, ck)
    mdC, fC = fresh()
    mdC.checkpoint_load(fC, ck)
    r2, pathsC = steps(mdC, fC, 16, 15)
    assert r1 >= 1 and r2 >= 1, (r1, r2)
    assert r1 == ra and r2 == rb
    # the candidate lists travel with the checkpoint: the resumed run walks
    # them where the run that went on did (no full-list retake at the load)
    assert pathsC == pathsA, (pathsA, pathsC)
    for a, b in zip(fA.dev_atoms(), fC.dev_atoms()):
        assert np.array_equal(a, b)
    assert np.array_equal(fA.grid().get("T"), fC.grid().get("T"))
    assert fA.compute_vector(0) == fC.compute_vector(0)
