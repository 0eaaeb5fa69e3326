"""GPU: the density sweep's candidate list against the oracle's full scans.

The rho loop may walk an atom's candidates (the list entries within
rc + delta when they were taken) instead of the full list while no moved atom
(one 0.45 delta or farther from where it was at that time) lies near it.  The
decision is per wave.  These runs put a fast atom through a lattice and move
ghost images with a shrinking box between list uploads, and require the
oracle's bits throughout, with the per-wave choice visibly engaged.
"""

import numpy as np
import pytest

import eph_amd as ea
from drive import Pair, rel_err, bitwise_fraction

pytestmark = pytest.mark.gpu


def _same(p, tag):
    A, B = p.sA, p.sB
    n = A.nlocal
    nall = n + A.nghost
    for k, (g, o) in {
        "rho": (p.gpu.get("rho")[:n], p.ora.get("rho", nall)[:n]),
        "w": (p.gpu.get("w")[:n], p.ora.get("w", nall)[:n]),
        "f_EPH": (p.gpu.get("f_EPH"), p.ora.get("f_EPH", n)),
        "f_RNG": (p.gpu.get("f_RNG"), p.ora.get("f_RNG", n)),
        "x": (B.x[:n], A.x[:n]),
        "v": (B.v[:n], A.v[:n]),
    }.items():
        assert bitwise_fraction(g, o) == 1.0, f"{tag}: {k} differs (rel err {rel_err(g, o):.3e})"
    assert np.array_equal(p.gpu.grid().get("T"), p.ora.fdm().get("T")), f"{tag}: T_e differs"


def test_fast_atom_keeps_bits_with_per_wave_candidates(oracle, tmp_path):
    """fcc 12^3 (6,912 atoms, 108 waves), 4 species, a 10 keV primary atom at
    dt 0.1 fs (0.17 A per step): from its second step on it is a moved atom.
    Only the waves near it scan their full lists; the rest walk candidates.
    Bit for bit against the oracle over 60 steps."""
    print("candidates: fcc12 alloy, 10 keV PKA, 60 steps", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=12, species=4, T=300.0, pka_eV=1.0e4, grid=6,
             dt=1e-4, hot_spot=(3000.0, 6.0))
    mixed = 0
    for k in range(60):
        p.step()
        st = p.gpu.candidate_stats()
        assert st["waves"] == (p.sA.nlocal + 63) // 64
        if not st["retaken"] and st["moved"] > 0 and 0 < st["full_waves"] < st["waves"]:
            mixed += 1
        if k % 15 == 14:
            _same(p, f"step {k + 1}")
    # the per-wave choice was in effect on many steps (not the all-or-nothing
    # retake): some waves scanned the full list, the others walked candidates
    assert mixed >= 10, f"only {mixed} steps mixed full scans and candidate walks"


def test_ghosts_moving_with_the_box_keep_bits(oracle, tmp_path):
    """Between list uploads the box shrinks by 0.05 A per step along x (as
    under fix deform or npt): the ghost images across the x faces move while
    their owners do not, and pairs across the boundary come inside r_c
    without any local atom moving.  The ghosts' own displacements put them on
    the moved list, so the nearby waves scan their full lists; the result
    stays the oracle's, bit for bit.  The lattice starts at rest and there is
    no pair force, so the pairs that enter r_c are the box's doing: the
    fourth shell (5.11 A, outside rc + delta = 4.95 A) across the x faces
    reaches r_c after 0.61 A of shrinking."""
    print("candidates: fcc10, box shrinking along x, 24 steps", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=10, T=0.0, grid=5, dt=2e-4)
    A, B = p.sA, p.sB
    n = A.nlocal
    sh = A.ghost_shift.copy()
    Lx = p.md.info()["boxhi"][0] - p.md.info()["boxlo"][0]
    grew = moved = False
    first_active = None
    for step in range(1, 25):
        scale = (Lx - 0.05 * step) / Lx
        shs = sh.copy()
        shs[:, 0] *= scale
        p.ora.initial_integrate(A)
        p.gpu.initial_integrate(B)
        for s in (A, B):
            s.x[n:] = s.x[s.ghost_owner] + shs
            s.v[n:] = s.v[s.ghost_owner]
            s.f[:n] = 0.0
        p.ora.post_force(A, step=step)
        p.gpu.post_force(B, step=step)
        p.ora.final_integrate(A)
        p.gpu.final_integrate(B)
        p.ora.end_of_step(A, step)
        p.gpu.end_of_step(B, step)
        na = int(p.gpu.get("nactive").sum())
        if first_active is None:
            first_active = na
        grew = grew or na > first_active
        moved = moved or p.gpu.candidate_stats()["moved"] > 0
        if step % 6 == 0:
            _same(p, f"step {step}")
    # pairs across the x faces came inside r_c during the run, and the
    # ghosts' displacements were seen
    assert grew, "the shrinking box brought no new active pair"
    assert moved


def _step_to(p, step, x_local):
    """One Verlet step of both fixes with the local positions set to x_local
    after the first half-kick (ghost images follow their owners)."""
    A, B = p.sA, p.sB
    n = A.nlocal
    p.ora.initial_integrate(A)
    p.gpu.initial_integrate(B)
    for s in (A, B):
        s.x[:n] = x_local
        s.x[n:] = s.x[s.ghost_owner] + s.ghost_shift
        s.v[n:] = s.v[s.ghost_owner]
        s.f[:n] = 0.0
    p.ora.post_force(A, step=step)
    p.gpu.post_force(B, step=step)
    p.ora.final_integrate(A)
    p.gpu.final_integrate(B)
    p.ora.end_of_step(A, step)
    p.gpu.end_of_step(B, step)


def test_overflowing_moved_list_then_return_keeps_bits(oracle, tmp_path):
    """More than 512 atoms move at once (the moved list overflows), then all
    return to the lattice.  Atoms i and j, third-shell neighbours (4.43 A,
    inside r_c = 4.5 A), are pulled apart to 5.03 A (outside r_c + delta =
    4.95 A) on the overflowing step and come back on the next.  Back on the
    lattice no atom is 0.45 delta from the positions the candidates were last
    checked against, yet i's candidates must again hold j: the overflowing
    step may not leave candidates taken at its own positions behind.  Bit for
    bit against the oracle on the steps after the return."""
    print("candidates: fcc10, moved-list overflow and return, 4 steps", flush=True)
    p = Pair(oracle, ea, str(tmp_path), cells=10, T=0.0, grid=5, dt=2e-4)
    A = p.sA
    n = A.nlocal
    x0 = A.x[:n].copy()
    c = 0.5 * (x0.min(0) + x0.max(0))
    i = int(np.argmin(((x0 - c) ** 2).sum(1)))
    d = np.sqrt(((x0 - x0[i]) ** 2).sum(1))
    j = int(np.argmin(np.abs(d - 3.615 * np.sqrt(1.5))))
    assert abs(d[j] - 3.615 * np.sqrt(1.5)) < 1e-9
    u = (x0[j] - x0[i]) / d[j]
    x1 = x0.copy()
    x1[i] -= 0.3 * u
    x1[j] += 0.3 * u
    others = np.setdiff1d(np.arange(n), [i, j])[-600:]
    x1[others, 2] += 0.25
    assert np.sqrt(((x1[j] - x1[i]) ** 2).sum()) > 4.95 + 0.05
    _step_to(p, 1, x0)
    _step_to(p, 2, x1)
    st = p.gpu.candidate_stats()
    assert st["moved"] > 512 and st["full_waves"] == st["waves"], st
    _step_to(p, 3, x0)
    _same(p, "step 3 (back on the lattice)")
    assert p.gpu.candidate_stats()["retaken"]
    _step_to(p, 4, x0)
    _same(p, "step 4")
    st = p.gpu.candidate_stats()
    assert not st["retaken"] and st["full_waves"] == 0, st
