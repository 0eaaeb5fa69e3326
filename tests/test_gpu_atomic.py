"""GPU: the fix eph/atomic style against the oracle's FixEPHAtomic.

Both sides run the same harness trajectory (drive.Pair's step order); the
per-atom quantities (rho, rho_a, w, forces, the deposit dE, the energies E and
temperatures T after the heat diffusion) must agree bit for bit, the tallies
Ee and Te (sums over atoms, in a different order on the device) to 1e-10.
"""
import copy

import numpy as np
import pytest

import eph_amd as ea
from eph_amd import inputs
from drive import Pair, bitwise_fraction

pytestmark = pytest.mark.gpu

NAMES = {1: ["Cu"], 3: ["A", "B", "C"]}


class AtomicPair(Pair):
    """drive.Pair with both fixes in the eph/atomic style"""

    def __init__(self, oracle, tmpdir, cells=5, species=1, T=600.0, pka_eV=300.0, dt=2.5e-4,
                 flags="7", inner_loops=0, T_e=300.0, md_seed=7, seed=4242):
        self.ea = ea
        self.oracle = oracle
        self.md = ea.MD(cells, a=3.615, ntypes=species, seed=md_seed, T=T, pka_eV=pka_eV)
        names = NAMES[species]
        pb = f"{tmpdir}/at{species}.beta"
        inputs.synthetic_beta(pb, amplitudes=[2.0 + 0.1 * e for e in range(species)],
                              prefactors=[0.02 + 0.002 * e for e in range(species)], names=names)
        pk = f"{tmpdir}/at{species}.kappa"
        inputs.synthetic_kappa(pk, names=names)
        args = inputs.atomic_args(pb, pk, names, seed=seed, flags=flags, T_e=T_e,
                                  inner_loops=inner_loops)
        inf = self.md.info()
        masses = [inputs.MASS_CU] * species
        self.ora = oracle.Atomic(args, species, dt, inputs.BOLTZ, inputs.FTM2V, masses=masses)
        self.gpu = ea.FixEPH(args, species, inf["boxlo"], inf["boxhi"], dt, inputs.BOLTZ,
                             inputs.FTM2V, masses=masses)
        self.gpu.init()
        self.sA = self.md.snapshot()
        self.sB = copy.deepcopy(self.sA)
        self.gpu.set_atoms(self.sB)
        self.ora.initialise(self.sA)
        self.seed = seed
        self.inject = False
        self.step_no = 0


def _close(p):
    n = p.sA.nlocal
    nall = n + p.sA.nghost
    pairs = {
        "rho": (p.gpu.get("rho")[:n], p.ora.get("rho", n)),
        "rho_a": (p.gpu.get("rho_a")[:n], p.ora.get("rho_a", n)),
        "w": (p.gpu.get("w")[:n], p.ora.get("w", n)),
        "f_EPH": (p.gpu.get("f_EPH"), p.ora.get("f_EPH", n)),
        "f_RNG": (p.gpu.get("f_RNG"), p.ora.get("f_RNG", n)),
        "dE": (p.gpu.get("dE"), p.ora.get("dE", n)),
        "E": (p.gpu.get("E")[:n], p.ora.get("E", n)),
        "T": (p.gpu.get("T")[:n], p.ora.get("T", n)),
        "x": (p.sB.x[:n], p.sA.x[:n]),
        "v": (p.sB.v[:n], p.sA.v[:n]),
        "f": (p.sB.f[:n], p.sA.f[:n]),
        "array": (p.gpu.peratom(), p.ora.get("array", n)),
    }
    for k, (g, o) in pairs.items():
        assert bitwise_fraction(g, o) == 1.0, f"{k}: not bit-identical"
    for i in (0, 1):
        eg, eo = p.gpu.compute_vector(i), p.ora.compute_vector(i)
        assert abs(eg - eo) <= 1e-10 * abs(eo), (i, eg, eo)
    assert nall > n


@pytest.mark.parametrize("species,flags,loops", [(1, "7", 0), (3, "7", 3), (1, "0x37", 2),
                                                 (3, "3", 0)])
def test_atomic_trajectory_matches_oracle(oracle, tmp_path, species, flags, loops):
    p = AtomicPair(oracle, str(tmp_path), cells=6, species=species, flags=flags,
                   inner_loops=loops)
    E0 = p.gpu.atomic_energy()
    assert np.array_equal(E0, p.ora.get("E", p.sA.nlocal))  # set-up from T_e
    for _ in range(20):
        p.step()
    _close(p)


def test_atomic_energy_hand_over(oracle, tmp_path):
    """the energies read out and written back (as a caller that reorders atoms
    does) continue the same trajectory"""
    p = AtomicPair(oracle, str(tmp_path), cells=5, species=1)
    for _ in range(5):
        p.step()
    E = p.gpu.atomic_energy()
    p.gpu.set_atomic_energy(E)
    assert np.array_equal(p.gpu.atomic_energy(), E)
    for _ in range(5):
        p.step()
    _close(p)
