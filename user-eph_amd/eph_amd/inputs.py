"""Input-file writers for the `fix eph` formats and the benchmark workloads.

Formats (read by the library exactly like the reference reads them):
  .beta    three comment lines; "N name_1 .. name_N"; "Nrho dr Nbeta drho rcut";
           then per element: atomic number, Nrho values of rho(r), Nbeta values
           of beta(rho)  (USER-EPH eph_beta.h:46-125, Doc/Beta/input.beta)
  T grid   three comment lines; "nx ny nz steps"; three "lo hi" box lines;
           a parameter-file name or NULL; then one line per cell
           "i j k T S rho_e C_e kappa flag T_dyn"  (eph_fdm.h:48-119)
  params   three comment lines; "n dT"; n lines "C_e kappa"  (eph_fdm.h:76-104)
"""
import math
import os

import numpy as np

# LAMMPS metal-unit constants (LAMMPS src/update.cpp, units metal).
BOLTZ = 8.617343e-5          # eV/K
MVV2E = 1.0364269e-4         # (g/mol * A^2/ps^2) -> eV
FTM2V = 1.0 / 1.0364269e-4   # (eV/A)/(g/mol) -> A/ps^2
MASS_CU = 63.546             # g/mol; the workloads do not state a mass


def write_beta(path, names, numbers, n_rho, dr, n_beta, drho, r_cut, rho_knots, beta_knots,
               comment="synthetic"):
    """Write a .beta file.  rho_knots/beta_knots: one 1-D array per element."""
    with open(path, "w") as f:
        f.write(f"# {comment}\n# r -> in Ang\n# beta -> in eV * ps / Ang^2\n")
        f.write(f"{len(names)} {' '.join(names)}\n")
        f.write(f"{n_rho} {dr!r} {n_beta} {drho!r} {r_cut!r}\n")
        for z, rk, bk in zip(numbers, rho_knots, beta_knots):
            f.write(f"{z}\n")
            f.write("\n".join(repr(float(v)) for v in rk))
            f.write("\n")
            f.write("\n".join(repr(float(v)) for v in bk))
            f.write("\n")


def synthetic_beta(path, amplitudes=(2.0,), prefactors=(0.02,), names=None, n_knots=1000,
                   r_cut=4.5, r_decay=0.9, drho=0.02):
    """rho(r) = A exp(-r / 0.9 A), beta(rho) = B rho / (1 + 0.5 rho) per element.

    dr = r_cut / (n - 1) so that dr (n - 1) = r_cut as eph_beta.h:93,110 need;
    drho = 0.02 gives rho_cutoff = 19.98.
    """
    ne = len(amplitudes)
    names = names or (["Cu"] if ne == 1 else [f"E{k}" for k in range(ne)])
    dr = r_cut / (n_knots - 1)
    r = np.arange(n_knots) * dr
    rho_grid = np.arange(n_knots) * drho
    rho_knots = [a * np.exp(-r / r_decay) for a in amplitudes]
    beta_knots = [b * rho_grid / (1.0 + 0.5 * rho_grid) for b in prefactors]
    write_beta(path, names, [29] * ne, n_knots, dr, n_knots, drho, r_cut, rho_knots, beta_knots)
    return names


def adjoint_beta(path, n_beta=64, amplitude=2.0, prefactor=0.02, n_rho=1000, r_cut=4.5,
                 r_decay=0.9, rho_max=20.0):
    """One element: rho(r) as synthetic_beta (1000 knots), beta(rho) = B rho/(1 + 0.5 rho)
    on n_beta knots up to rho_max (the parameterisation the adjoint differentiates)."""
    dr = r_cut / (n_rho - 1)
    drho = rho_max / (n_beta - 1)
    r = np.arange(n_rho) * dr
    rho_grid = np.arange(n_beta) * drho
    write_beta(path, ["Cu"], [29], n_rho, dr, n_beta, drho, r_cut,
               [amplitude * np.exp(-r / r_decay)], [prefactor * rho_grid / (1.0 + 0.5 * rho_grid)],
               comment=f"synthetic, {n_beta} beta knots")
    return ["Cu"]


def write_params(path, dT, C_e, kappa):
    with open(path, "w") as f:
        f.write("#\n#\n#\n")
        f.write(f"{len(C_e)} {dT!r}\n")
        for c, k in zip(C_e, kappa):
            f.write(f"{float(c)!r} {float(k)!r}\n")


def write_grid(path, nx, ny, nz, box, T, S=None, rho=None, C_e=None, kappa=None, flag=None,
               tdyn=None, steps=1, params="NULL"):
    """Write a T_e grid file; per-cell arrays are flat in i + j nx + k nx ny order."""
    n = nx * ny * nz
    def arr(a, default):
        return np.full(n, default) if a is None else np.broadcast_to(np.asarray(a), (n,))
    T = arr(T, 300.0); S = arr(S, 0.0); rho = arr(rho, 1.0); C_e = arr(C_e, 1.0)
    kappa = arr(kappa, 1.0); flag = arr(flag, 1).astype(int); tdyn = arr(tdyn, 0).astype(int)
    with open(path, "w") as f:
        f.write("# T_e grid\n#\n#\n")
        f.write(f"{nx} {ny} {nz} {steps}\n")
        for a in range(3):
            f.write(f"{float(box[2 * a])!r} {float(box[2 * a + 1])!r}\n")
        f.write(f"{params}\n")
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    c = i + j * nx + k * nx * ny
                    f.write(f"{i} {j} {k} {float(T[c])!r} {float(S[c])!r} {float(rho[c])!r} "
                            f"{float(C_e[c])!r} {float(kappa[c])!r} {flag[c]} {tdyn[c]}\n")


def fix_args(beta_file, elements, seed=12345, flags="7", model=4, rho_e=1.0, C_e=3e-6,
             kappa=0.1, T_e=300.0, grid=(10, 10, 10), T_infile="NULL", T_freq=0,
             T_out="Te_out", style="eph", tau0=None):
    """`fix ID group <style> seed flags model|tau0 rho_e C_e kappa T_e nx ny nz T_infile freq
    T_out beta_file elem...` (USER-EPH README / fix_eph.cpp:36-58).  The styles
    eph/coloured/exp and eph/coloured/exp/v1 take tau0 (ps) in the model slot
    (fix_eph_coloured_exp.cpp:43)."""
    slot = repr(float(tau0)) if style.startswith("eph/coloured/exp") else str(model)
    return ["e", "all", style, str(seed), str(flags), slot, repr(rho_e), repr(C_e),
            repr(kappa), repr(T_e), str(grid[0]), str(grid[1]), str(grid[2]), T_infile,
            str(T_freq), T_out, beta_file] + list(elements)


def synthetic_kappa(path, names=("Cu",), numbers=None, n_r=500, r_cut=4.0, amplitudes=None,
                    r_decay=0.9, n_T=1000, dT=10.0, c0=3e-6, c1=1e-8, k0=5e-5, k1=5e-8):
    """.kappa file for fix eph/atomic (eph_kappa.h:55-150): per element the
    locality rho(r) = A exp(-r / r_decay) on n_r knots up to r_cut, the
    electronic heat capacity per atom C(T) = c0 + c1 T (eV/K) on n_T knots of
    dT; then n_pairs = (n + 1)(n - 1) / 2 (1 for one element) conduction tables
    K(T) = (k0 + k1 T) (1 + 0.1 k), k the slot (eV / (ps K))."""
    n = len(names)
    numbers = numbers or [29] * n
    amplitudes = amplitudes or [1.0 + 0.1 * e for e in range(n)]
    dr = r_cut / (n_r - 1)
    r = np.arange(n_r) * dr
    T = np.arange(n_T) * dT
    n_pairs = (n + 1) * (n - 1) // 2 if n > 1 else 1
    with open(path, "w") as f:
        f.write("# synthetic kappa file\n#\n#\n")
        f.write(f"{n} " + " ".join(names) + "\n")
        f.write(f"{n_r} {dr!r} {r_cut!r} {n_T} {dT!r} {(n_T - 1) * dT!r}\n")
        for e in range(n):
            f.write(f"{numbers[e]}\n")
            f.write("\n".join(repr(float(v)) for v in amplitudes[e] * np.exp(-r / r_decay)))
            f.write("\n")
            f.write("\n".join(repr(float(v)) for v in c0 + c1 * T))
            f.write("\n")
        for k in range(n_pairs):
            f.write("\n".join(repr(float(v)) for v in (k0 + k1 * T) * (1.0 + 0.1 * k)))
            f.write("\n")
    return list(names)


def atomic_args(beta_file, kappa_file, elements, seed=12345, flags="7", T_e=300.0,
                T_infile="NULL", inner_loops=0, out="NULL"):
    """`fix ID group eph/atomic seed flags T_e T_infile inner_loops out beta_file kappa_file
    elem...` (fix_eph_atomic.cpp:59-269)."""
    return ["e", "all", "eph/atomic", str(seed), str(flags), repr(float(T_e)), T_infile,
            str(inner_loops), out, beta_file, kappa_file] + list(elements)


def pka_speed(energy_eV, mass=MASS_CU):
    """|v| in A/ps of an atom of `mass` g/mol carrying `energy_eV` of kinetic energy."""
    return math.sqrt(2.0 * energy_eV / (mass * MVV2E))


# ---------------------------------------------------------------------------
# Workloads (the synthetic systems the parity tests and the benchmark use)
# ---------------------------------------------------------------------------
WORKLOADS = {
    "fcc10_plain": dict(cells=10, a=3.615, T=300.0, disp=0.0, species=1, pka_eV=0.0, dt=1e-3,
                        grid=10, steps=100),
    "fcc40_plain": dict(cells=40, a=3.615, T=1000.0, disp=0.05, species=1, pka_eV=0.0, dt=1e-3,
                        grid=40, steps=200),
    "fcc80_alloy4_cascade": dict(cells=80, a=3.615, T=300.0, disp=0.0, species=4, pka_eV=1.0e4,
                                 dt=1e-4, grid=80, steps=500, hot_spot=(3000.0, 10.0)),
}

ALLOY_AMPLITUDES = (1.6, 1.8, 2.0, 2.2)
ALLOY_PREFACTORS = (0.015, 0.018, 0.02, 0.025)


def workload_beta(path, species):
    if species == 1:
        return synthetic_beta(path)
    return synthetic_beta(path, ALLOY_AMPLITUDES[:species], ALLOY_PREFACTORS[:species],
                          names=["A", "B", "C", "D"][:species])


def ensure_dir(p):
    os.makedirs(p, exist_ok=True)
    return p


def write_uniform_grid(path, n, L, steps, T=300.0):
    """n^3 T_e grid over a cubic box [0, L), uniform T, the given initial
    sub-step count, constant C_e = 3e-6 and kappa = 0.1"""
    idx = np.indices((n, n, n))[::-1].reshape(3, -1).T
    with open(path, "w") as f:
        f.write(f"# T_e grid: uniform {T} K\n#\n#\n")
        f.write(f"{n} {n} {n} {steps}\n")
        for _ in range(3):
            f.write(f"0.0 {L!r}\n")
        f.write("NULL\n")
        f.write("\n".join(f"{i} {j} {k} {T!r} 0.0 1.0 3e-06 0.1 1 0" for i, j, k in idx.tolist()))
        f.write("\n")


ENSEMBLE_A = 3.615
ENSEMBLE_DT = 2.5e-4  # 1 fs integrates no 1 keV atom with this lj/cut: it overruns a neighbour


def ensemble_replicas(tmp, nrep=64, size_seed=3):
    """The batched-ensemble workload: nrep replicas of the fcc plain-method
    physics, n^3 cells with n = round(exp U[ln 10, ln 25]) (seeded), each with
    its own n^3 T_e grid (uniform 300 K, 19 sub-steps), fix seed 7000 + r,
    lattice seed 900 + r and a 1 keV primary atom.  Returns the list of
    (harness keywords, fix arguments) and the sizes."""
    rng = np.random.default_rng(size_seed)
    ns = np.rint(np.exp(rng.uniform(np.log(10.0), np.log(25.0), nrep))).astype(int)
    beta = os.path.join(tmp, "ens.beta")
    names = workload_beta(beta, 1)
    reps = []
    for r, n in enumerate(ns.tolist()):
        tin = os.path.join(tmp, f"ens{r}.T")
        write_uniform_grid(tin, n, n * ENSEMBLE_A, 19)
        fargs = fix_args(beta, names, seed=7000 + r, flags="7", grid=(n, n, n), T_infile=tin,
                         C_e=3e-6, kappa=0.1)
        kw = dict(cells=n, a=ENSEMBLE_A, ntypes=1, seed=900 + r, T=300.0, pka_eV=1000.0)
        reps.append((kw, fargs))
    return reps, ns

