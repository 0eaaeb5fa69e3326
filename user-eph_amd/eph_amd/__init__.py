"""Python binding of the MI355X-native `fix eph` library (libeph_amd.so).

The library is the product: every hook runs as gfx950 HIP kernels behind the
C ABI of include/eph_amd.h.  There is no CPU fallback -- if the shared library
is missing or no HIP device is present, the calls raise.

    FixEPH  -- the fix (FixEPH hooks, fix_eph.cpp), host-array or device-resident
    Grid    -- the electronic T_e grid (EPH_FDM, eph_fdm.h), also standalone
    MD      -- the minimal MD harness standing in for LAMMPS (include/eph_md.h)
    Ensemble -- many replicas advanced together (include/eph_ensemble.h)
    Adjoint -- trajectory gradients w.r.t. beta knots and v0 (include/eph_adjoint.h)
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libeph_amd.so")

EPH_OK = 0
STATUS = {0: "ok", -1: "bad argument", -2: "file", -3: "format", -4: "out of memory",
          -5: "HIP error", -6: "state", -7: "non-finite T_e", -8: "unsupported", -9: "no device"}
FRICTION, RANDOM, FDM, NOINT, NOFRICTION, NORANDOM = 0x01, 0x02, 0x04, 0x08, 0x10, 0x20

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


class EphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class MDConfig(C.Structure):
    _fields_ = [("cells", C.c_int), ("a", C.c_double), ("ntypes", C.c_int), ("seed", C.c_uint),
                ("T", C.c_double), ("disp_sigma", C.c_double), ("pka_eV", C.c_double),
                ("mass", C.c_double), ("cut_fix", C.c_double), ("cut_lj", C.c_double),
                ("skin", C.c_double), ("lj_eps", C.c_double), ("lj_sigma", C.c_double)]


_SIG = {
    "eph_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(C.c_char_p), C.c_int, dp, dp,
                             C.c_double, C.c_double, C.c_double, C.c_int]),
    "eph_destroy": (None, [vp]),
    "eph_last_error": (C.c_char_p, [vp]),
    "eph_init": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
    "eph_setmask": (C.c_int, []),
    "eph_reset_dt": (C.c_int, [vp, C.c_double]),
    "eph_set_masses": (C.c_int, [vp, dp, C.c_int]),
    "eph_set_groupbit": (C.c_int, [vp, C.c_int]),
    "eph_set_candidate_margin": (C.c_int, [vp, C.c_double]),
    "eph_get_params": (C.c_int, [vp, ip, ip, ip, dp, dp, dp, dp]),
    "eph_set_atoms": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, ip, ip, ip]),
    "eph_set_ghost_map": (C.c_int, [vp, C.c_int, ip, dp]),
    "eph_set_neighbors": (C.c_int, [vp, C.c_int, ip, ip, C.c_int]),
    "eph_set_noise": (C.c_int, [vp, dp]),
    "eph_set_timestep": (C.c_int, [vp, C.c_longlong]),
    "eph_initial_integrate": (C.c_int, [vp, dp, dp, dp]),
    "eph_post_force": (C.c_int, [vp, dp, dp, dp]),
    "eph_final_integrate": (C.c_int, [vp, dp, dp]),
    "eph_end_of_step": (C.c_int, [vp, dp, dp, C.c_longlong]),
    "eph_post_run": (C.c_int, [vp]),
    "eph_checkpoint_save": (C.c_int, [vp, C.c_char_p]),
    "eph_checkpoint_load": (C.c_int, [vp, C.c_char_p]),
    "eph_compute_vector": (C.c_int, [vp, C.c_int, dp]),
    "eph_get_peratom": (C.c_int, [vp, dp]),
    "eph_peratom_cols": (C.c_int, [vp]),
    "eph_get_atomic_energy": (C.c_int, [vp, dp]),
    "eph_set_atomic_energy": (C.c_int, [vp, dp]),
    "eph_get_state": (C.c_int, [vp, C.c_int, dp]),
    "eph_last_energy": (C.c_int, [vp, dp]),
    "eph_rho_cutoff_warning": (C.c_int, [vp]),
    "eph_dev_initial_integrate": (C.c_int, [vp]),
    "eph_dev_post_force": (C.c_int, [vp]),
    "eph_dev_final_integrate": (C.c_int, [vp]),
    "eph_dev_end_of_step": (C.c_int, [vp, C.c_longlong]),
    "eph_dev_get_atoms": (C.c_int, [vp, dp, dp, dp]),
    "eph_sync": (C.c_int, [vp]),
    "eph_stream": (vp, [vp]),
    "eph_timing_enable": (C.c_int, [vp, C.c_int]),
    "eph_set_sweep_forms": (C.c_int, [vp, C.c_int, C.c_int]),
    "eph_get_filter_state": (C.c_int, [vp, C.c_int, dp]),
    "eph_set_filter_state": (C.c_int, [vp, C.c_int, dp]),
    "eph_autotune": (C.c_int, [vp, ip, dp]),
    "eph_timing_report": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int, dp, C.POINTER(C.c_longlong)]),
    "eph_copy_probe": (C.c_int, [C.c_longlong, C.c_int, C.c_int, dp]),
    "eph_spline_build": (C.c_int, [C.c_double, dp, C.c_int, dp]),
    "eph_beta_table": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp]),
    "eph_kappa_info": (C.c_int, [C.c_char_p, ip, ip, dp, dp]),
    "eph_kappa_table": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, dp]),
    "eph_grid_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, dp, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int]),
    "eph_grid_load": (C.c_int, [C.POINTER(vp), C.c_char_p, C.c_int]),
    "eph_grid_destroy": (None, [vp]),
    "eph_grid_last_error": (C.c_char_p, [vp]),
    "eph_context_grid": (vp, [vp]),
    "eph_grid_set_dt": (C.c_int, [vp, C.c_double]),
    "eph_grid_set_steps": (C.c_int, [vp, C.c_int]),
    "eph_grid_set_cell": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]),
    "eph_grid_insert_energy": (C.c_int, [vp, C.c_int, dp, dp, dp, dp]),
    "eph_grid_insert_energy_noalloc": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
    "eph_noalloc_scratch_bytes": (C.c_longlong, []),
    "eph_grid_solve": (C.c_int, [vp, ip]),
    "eph_grid_get": (C.c_int, [vp, C.c_int, dp]),
    "eph_grid_set": (C.c_int, [vp, C.c_int, dp]),
    "eph_grid_lookup": (C.c_int, [vp, C.c_int, dp, dp, dp, ip, dp]),
    "eph_grid_info": (C.c_int, [vp, ip, dp, dp]),
    "eph_grid_T_mean": (C.c_int, [vp, dp]),
    "eph_grid_save_state": (C.c_int, [vp, C.c_char_p]),
    "eph_grid_save_temperature": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "eph_md_create_fcc": (C.c_int, [C.POINTER(vp), C.POINTER(MDConfig)]),
    "eph_md_destroy": (None, [vp]),
    "eph_md_last_error": (C.c_char_p, [vp]),
    "eph_md_info": (C.c_int, [vp, ip, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp, dp]),
    "eph_md_get": (C.c_int, [vp, dp, dp, ip, ip, ip, ip, dp, ip, ip]),
    "eph_md_update_ghosts": (C.c_int, [vp, dp, dp]),
    "eph_md_lj_host": (C.c_int, [vp, dp, dp]),
    "eph_md_needs_rebuild": (C.c_int, [vp, dp]),
    "eph_md_rebuild": (C.c_int, [vp, dp, dp]),
    "eph_md_attach": (C.c_int, [vp, vp]),
    "eph_md_run": (C.c_int, [vp, vp, C.c_int, C.c_longlong, dp, ip]),
    "eph_md_set_device_rebuild": (C.c_int, [vp, C.c_int]),
    "eph_md_force_rebuild": (C.c_int, [vp]),
    "eph_md_pull": (C.c_int, [vp, vp]),
    "eph_md_checkpoint_save": (C.c_int, [vp, vp, C.c_char_p]),
    "eph_md_checkpoint_load": (C.c_int, [vp, vp, C.c_char_p]),
    "eph_ensemble_create": (C.c_int, [C.POINTER(vp), C.c_int]),
    "eph_ensemble_destroy": (None, [vp]),
    "eph_ensemble_last_error": (C.c_char_p, [vp]),
    "eph_ensemble_add": (C.c_int, [vp, C.POINTER(MDConfig), C.c_int, C.POINTER(C.c_char_p), C.c_int,
                                   C.c_double, C.c_double, C.c_double, dp]),
    "eph_ensemble_setup": (C.c_int, [vp, C.c_longlong]),
    "eph_ensemble_run": (C.c_int, [vp, C.c_int, C.c_int, dp, ip]),
    "eph_ensemble_size": (C.c_int, [vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "eph_ensemble_autotune": (C.c_int, [vp, ip, dp]),
    "eph_ensemble_set_variants": (C.c_int, [vp, C.c_int, C.c_int]),
    "eph_ensemble_info": (C.c_int, [vp, C.c_int, ip, ip, ip]),
    "eph_ensemble_get": (C.c_int, [vp, C.c_int, C.c_int, dp]),
    "eph_adjoint_create": (C.c_int, [C.POINTER(vp), vp, vp, C.c_int, C.c_int, dp]),
    "eph_adjoint_destroy": (None, [vp]),
    "eph_adjoint_last_error": (C.c_char_p, [vp]),
    "eph_adjoint_info": (C.c_int, [vp, ip, ip, ip, ip]),
    "eph_adjoint_get_beta": (C.c_int, [vp, dp]),
    "eph_adjoint_get_v0": (C.c_int, [vp, dp]),
    "eph_adjoint_forward": (C.c_int, [vp, dp, dp, dp, dp]),
    "eph_adjoint_backward": (C.c_int, [vp, dp, dp]),
    "eph_spline_jacobian": (C.c_int, [C.c_double, dp, C.c_int, dp]),
}

_lib = None


def build():
    """Compile libeph_amd.so for gfx950 (hipcc)."""
    subprocess.check_call(["make", "-s", "-C", PKG, "-j8"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C user-eph_amd` "
                               "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def copy_probe(nbytes, reps=3, device=0):
    """Average ms of a plain device copy of nbytes (read) + nbytes (write)."""
    ms = C.c_double()
    rc = lib().eph_copy_probe(int(nbytes), reps, device, C.byref(ms))
    if rc != EPH_OK:
        raise EphError(rc, "copy probe failed")
    return ms.value


def spline_build(dx, y):
    """Host-built global-x Akima cubic (EPH_Spline) as [n][4] coefficients."""
    y = f64(y)
    out = np.zeros((len(y), 4))
    rc = lib().eph_spline_build(dx, _d(y), len(y), _d(out))
    if rc != EPH_OK:
        raise EphError(rc, lib().eph_last_error(None).decode())
    return out


def beta_table(path, elem, which):
    """Host-built table of a .beta file: (coeffs [n][4], inv_dx, r_cutoff, rho_cutoff)."""
    L = lib()
    n = L.eph_beta_table(path.encode(), elem, which, None, None, None, None)
    if n < 0:
        raise EphError(n, L.eph_last_error(None).decode())
    out = np.zeros((n, 4))
    inv, rc, rho = C.c_double(), C.c_double(), C.c_double()
    L.eph_beta_table(path.encode(), elem, which, _d(out), C.byref(inv), C.byref(rc), C.byref(rho))
    return out, inv.value, rc.value, rho.value


def kappa_info(path):
    """(n_elements, n_pairs, r_cutoff, T_max) of a .kappa file (host only)"""
    L = lib()
    ne, npair, rc, tm = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    rc_ = L.eph_kappa_info(path.encode(), C.byref(ne), C.byref(npair), C.byref(rc), C.byref(tm))
    if rc_ < 0:
        raise EphError(rc_, L.eph_last_error(None).decode())
    return ne.value, npair.value, rc.value, tm.value


def kappa_table(path, index, which):
    """Host-built table of a .kappa file: 0 rho(r), 1 rho(r^2) ([n][4], inv_dx);
    2 E(T) of an element, 3 K(T) of a pair slot ([n][2] = y, dy; dx)"""
    L = lib()
    n = L.eph_kappa_table(path.encode(), index, which, None, None)
    if n < 0:
        raise EphError(n, L.eph_last_error(None).decode())
    out = np.zeros((n, 4 if which < 2 else 2))
    dx = C.c_double()
    L.eph_kappa_table(path.encode(), index, which, _d(out), C.byref(dx))
    return out, dx.value


class FixEPH:
    """`fix ID group eph ...` on one GPU (FixEPH, fix_eph.cpp)."""

    def __init__(self, args, ntypes, boxlo, boxhi, dt, boltz, ftm2v, device=0, masses=None):
        L = lib()
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        h = vp()
        lo, hi = f64(boxlo), f64(boxhi)
        rc = L.eph_create(C.byref(h), len(args), arr, ntypes, _d(lo), _d(hi), dt, boltz, ftm2v, device)
        if rc != EPH_OK:
            raise EphError(rc, L.eph_last_error(None).decode())
        self.h = h
        self.ntypes = ntypes
        self.nlocal = self.nghost = 0
        if masses is not None:
            self.set_masses(masses)

    def _check(self, rc):
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            lib().eph_destroy(self.h)
            self.h = None

    __del__ = close

    def init(self, dimension=3, nonperiodic=0, triclinic=0):
        c = C.c_double()
        self._check(lib().eph_init(self.h, dimension, nonperiodic, triclinic, C.byref(c)))
        return c.value

    def params(self):
        fl, mo, se = C.c_int(), C.c_int(), C.c_int()
        rc, rho, eta, dtf = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self._check(lib().eph_get_params(self.h, C.byref(fl), C.byref(mo), C.byref(se), C.byref(rc),
                                         C.byref(rho), C.byref(eta), C.byref(dtf)))
        return dict(flags=fl.value, model=mo.value, seed=se.value, r_cutoff=rc.value,
                    rho_cutoff=rho.value, eta_factor=eta.value, dtf=dtf.value)

    def reset_dt(self, dt): self._check(lib().eph_reset_dt(self.h, dt))

    def set_masses(self, masses):
        m = f64(np.concatenate([[0.0], np.asarray(masses, dtype=np.float64)]))
        self._check(lib().eph_set_masses(self.h, _d(m), self.ntypes))

    def set_groupbit(self, gb): self._check(lib().eph_set_groupbit(self.h, gb))

    def set_candidate_margin(self, delta):
        """delta < 0: 0.1 r_cutoff (default); 0: the rho loop always scans the full list"""
        self._check(lib().eph_set_candidate_margin(self.h, C.c_double(delta)))

    def set_atoms(self, s):
        """s: object with nlocal, nghost, x, v ([nall][3]), type, mask, tag"""
        self.nlocal, self.nghost = s.nlocal, s.nghost
        self._check(lib().eph_set_atoms(self.h, s.nlocal, s.nghost, _d(f64(s.x)), _d(f64(s.v)),
                                        _i(i32(s.type)), _i(i32(s.mask)), _i(i32(s.tag))))
        sh = f64(s.ghost_shift) if s.nghost else np.zeros((0, 3))
        self._check(lib().eph_set_ghost_map(self.h, s.nghost, _i(i32(s.ghost_owner)), _d(sh)))
        self._check(lib().eph_set_neighbors(self.h, s.nlocal, _i(i32(s.nbr_off)), _i(i32(s.nbr_idx)), -1))

    def set_noise(self, xi):
        if xi is None:
            self._check(lib().eph_set_noise(self.h, None))
        else:
            xi = f64(xi)
            self._check(lib().eph_set_noise(self.h, _d(xi)))

    def set_timestep(self, n): self._check(lib().eph_set_timestep(self.h, n))

    # host-array hooks (arrays updated in place)
    def initial_integrate(self, s):
        self._check(lib().eph_initial_integrate(self.h, _d(s.x), _d(s.v), _d(s.f)))

    def post_force(self, s, xi=None, step=None):
        if step is not None:
            self.set_timestep(step)
        if xi is not None:
            self.set_noise(xi)
        self._check(lib().eph_post_force(self.h, _d(s.x), _d(s.v), _d(s.f)))

    def final_integrate(self, s):
        self._check(lib().eph_final_integrate(self.h, _d(s.v), _d(s.f)))

    def end_of_step(self, s, ntimestep=0):
        self._check(lib().eph_end_of_step(self.h, _d(s.x), _d(s.v), ntimestep))

    def post_run(self): self._check(lib().eph_post_run(self.h))

    def checkpoint_save(self, path):
        """full-precision binary checkpoint of the fix's state (T_e, tally, per-atom state)"""
        self._check(lib().eph_checkpoint_save(self.h, path.encode()))

    def checkpoint_load(self, path):
        self._check(lib().eph_checkpoint_load(self.h, path.encode()))

    def compute_vector(self, i):
        out = C.c_double()
        self._check(lib().eph_compute_vector(self.h, i, C.byref(out)))
        return out.value

    def last_energy(self):
        out = C.c_double()
        self._check(lib().eph_last_energy(self.h, C.byref(out)))
        return out.value

    def rho_cutoff_warning(self): return bool(lib().eph_rho_cutoff_warning(self.h))

    def peratom(self):
        out = np.zeros((self.nlocal, lib().eph_peratom_cols(self.h)))
        self._check(lib().eph_get_peratom(self.h, _d(out)))
        return out

    def atomic_energy(self):
        """eph/atomic: per-atom electronic energies [nlocal]"""
        out = np.zeros(self.nlocal)
        self._check(lib().eph_get_atomic_energy(self.h, _d(out)))
        return out

    def set_atomic_energy(self, E):
        a = np.ascontiguousarray(E, dtype=np.float64).reshape(self.nlocal)
        self._check(lib().eph_set_atomic_energy(self.h, _d(a)))

    def get(self, which):
        idx = {"rho": 0, "w": 1, "xi": 2, "f_EPH": 3, "f_RNG": 4, "alpha": 5, "nactive": 6,
               "nlist": 7, "rho_a": 8, "E": 9, "dE": 10, "T": 11}[which]
        nall = self.nlocal + self.nghost
        shape = {0: (nall,), 5: (nall,), 1: (nall, 3), 2: (nall, 3), 6: (self.nlocal,),
                 7: (self.nlocal,), 8: (nall,), 9: (nall,), 10: (self.nlocal,),
                 11: (nall,)}.get(idx, (self.nlocal, 3))
        out = np.zeros(shape)
        self._check(lib().eph_get_state(self.h, idx, _d(out)))
        return out

    def candidate_stats(self):
        """The density sweep's candidate list at its last run: waves that
        scanned the full list, waves in all, atoms found moved, retaken (bool)"""
        out = np.zeros(4)
        self._check(lib().eph_get_state(self.h, 12, _d(out)))
        return {"full_waves": int(out[0]), "waves": int(out[1]), "moved": int(out[2]),
                "retaken": bool(out[3])}

    # device-resident hooks
    def dev_initial_integrate(self): self._check(lib().eph_dev_initial_integrate(self.h))
    def dev_post_force(self): self._check(lib().eph_dev_post_force(self.h))
    def dev_final_integrate(self): self._check(lib().eph_dev_final_integrate(self.h))
    def dev_end_of_step(self, n): self._check(lib().eph_dev_end_of_step(self.h, n))
    def sync(self): self._check(lib().eph_sync(self.h))
    def stream(self): return lib().eph_stream(self.h)

    def dev_atoms(self):
        nall = self.nlocal + self.nghost
        x, v, f = np.zeros((nall, 3)), np.zeros((nall, 3)), np.zeros((self.nlocal, 3))
        self._check(lib().eph_dev_get_atoms(self.h, _d(x), _d(v), _d(f)))
        return x, v, f

    def timing(self, on=True): self._check(lib().eph_timing_enable(self.h, 1 if on else 0))

    def filter_state(self, which):
        """coloured styles: 0 f_dis / zv, 1 f_sto / zi, [nlocal][3]"""
        out = np.zeros((self.nlocal, 3))
        self._check(lib().eph_get_filter_state(self.h, which, _d(out)))
        return out

    def set_filter_state(self, which, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.nlocal, 3)
        self._check(lib().eph_set_filter_state(self.h, which, _d(a)))

    # kernel forms of the sweeps: tables through the caches, or in LDS with
    # 1024 / 512 / 256 threads per block; chunk4/3/2: tables through the
    # caches with at most 4, 3 or 2 blocks per CU
    _FORM_NAMES = ("chunk", "lds", "lds512", "lds256", "chunk4", "chunk3", "chunk2")
    _FORMS = {"chunk": 0, "lds": 1, "lds512": 2, "lds256": 3, "chunk4": 4, "chunk3": 5,
              "chunk2": 6, "auto": -1}

    def set_sweep_forms(self, density="auto", force="auto"):
        """kernel forms of the density and force sweeps: 'chunk', 'lds' or 'auto'"""
        self._check(lib().eph_set_sweep_forms(self.h, self._FORMS[density], self._FORMS[force]))

    def autotune(self):
        """time every form of each sweep on the resident atoms, keep the fastest"""
        nf = len(self._FORM_NAMES)
        forms = (C.c_int * 2)()
        ms = (C.c_double * (2 * nf))()
        self._check(lib().eph_autotune(self.h, forms, ms))
        name = self._FORM_NAMES
        return {"density": name[forms[0]], "force": name[forms[1]],
                "density_ms": {name[v]: ms[v] for v in range(nf) if ms[v] > 0},
                "force_ms": {name[v]: ms[nf + v] for v in range(nf) if ms[nf + v] > 0}}

    def timing_report(self):
        L = lib()
        n = L.eph_timing_report(self.h, -1, None, 0, None, None)
        if n < 0:
            self._check(n)
        res = {}
        for k in range(n):
            name = C.create_string_buffer(64)
            ms = C.c_double()
            cnt = C.c_longlong()
            self._check(L.eph_timing_report(self.h, k, name, 64, C.byref(ms), C.byref(cnt)))
            res[name.value.decode()] = (ms.value, cnt.value)
        return res

    def grid(self):
        return Grid(lib().eph_context_grid(self.h), owner=self)


class Grid:
    """Electronic temperature grid (EPH_FDM, eph_fdm.h)."""

    def __init__(self, h, owner=None):
        self.h = h
        self.owner = owner  # keeps the fix alive when the grid belongs to it
        dims = (C.c_int * 3)()
        lib().eph_grid_info(self.h, dims, None, None)
        self.dims = (dims[0], dims[1], dims[2])

    @classmethod
    def const(cls, nx, ny, nz, box, T, C_e, rho, kappa, device=0):
        h = vp()
        b = f64(box)
        rc = lib().eph_grid_create(C.byref(h), nx, ny, nz, _d(b), T, C_e, rho, kappa, device)
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_grid_last_error(None).decode())
        return cls(h)

    @classmethod
    def load(cls, path, device=0):
        h = vp()
        rc = lib().eph_grid_load(C.byref(h), path.encode(), device)
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_grid_last_error(None).decode())
        return cls(h)

    def _check(self, rc):
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_grid_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.owner is None:
            lib().eph_grid_destroy(self.h)
        self.h = None

    __del__ = close

    @property
    def ntotal(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def set_dt(self, dt): self._check(lib().eph_grid_set_dt(self.h, dt))
    def set_steps(self, s): self._check(lib().eph_grid_set_steps(self.h, s))

    def set_cell(self, i, j, k, T, S, rho, Ce, kappa, flag, tdyn):
        self._check(lib().eph_grid_set_cell(self.h, i, j, k, T, S, rho, Ce, kappa, flag, tdyn))

    def insert_energy(self, x, y, z, E):
        x, y, z, E = (f64(np.atleast_1d(a)) for a in (x, y, z, E))
        self._check(lib().eph_grid_insert_energy(self.h, len(x), _d(x), _d(y), _d(z), _d(E)))

    def insert_energy_noalloc(self, n, x, y, z, E):
        """insert_energy for n points already in device memory (x, y, z, E are
        device addresses); allocates nothing, returns once the work is queued"""
        self._check(lib().eph_grid_insert_energy_noalloc(self.h, int(n), x, y, z, E))

    def solve(self):
        n = C.c_int()
        self._check(lib().eph_grid_solve(self.h, C.byref(n)))
        return n.value

    def last_r(self):
        r = C.c_double()
        lib().eph_grid_info(self.h, None, None, C.byref(r))
        return r.value

    def get(self, which="T"):
        idx = {"T": 0, "dT": 1, "C": 2, "rho": 3, "kappa": 4, "S": 5}[which]
        out = np.zeros(self.ntotal)
        self._check(lib().eph_grid_get(self.h, idx, _d(out)))
        return out

    def set(self, which, values):
        idx = {"T": 0, "dT": 1, "C": 2, "rho": 3, "kappa": 4, "S": 5}[which]
        v = f64(values)
        self._check(lib().eph_grid_set(self.h, idx, _d(v)))

    def lookup(self, x, y, z):
        x, y, z = (f64(np.atleast_1d(a)) for a in (x, y, z))
        idx = np.zeros(len(x), np.int32)
        T = np.zeros(len(x))
        self._check(lib().eph_grid_lookup(self.h, len(x), _d(x), _d(y), _d(z), _i(idx), _d(T)))
        return idx, T

    def T_mean(self):
        out = C.c_double()
        self._check(lib().eph_grid_T_mean(self.h, C.byref(out)))
        return out.value

    def save_state(self, path): self._check(lib().eph_grid_save_state(self.h, path.encode()))

    def save_temperature(self, base, n):
        self._check(lib().eph_grid_save_temperature(self.h, base.encode(), n))


class MDSystem:
    """Host snapshot of the harness system in LAMMPS layout."""
    pass


class MD:
    """Minimal MD harness (include/eph_md.h): fcc lattice, ghosts, lists, LJ."""

    def __init__(self, cells, a=3.615, ntypes=1, seed=1, T=300.0, disp_sigma=0.0, pka_eV=0.0,
                 mass=63.546, cut_fix=4.5, cut_lj=5.5, skin=2.0, lj_eps=0.4, lj_sigma=2.3):
        self.cfg = MDConfig(cells, a, ntypes, seed & 0xFFFFFFFF, T, disp_sigma, pka_eV, mass,
                            cut_fix, cut_lj, skin, lj_eps, lj_sigma)
        h = vp()
        rc = lib().eph_md_create_fcc(C.byref(h), C.byref(self.cfg))
        self.h = h
        if rc != EPH_OK:
            msg = lib().eph_md_last_error(h).decode() if h else ""
            raise EphError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            lib().eph_md_destroy(self.h)
            self.h = None

    __del__ = close

    def info(self):
        nl, ng = C.c_int(), C.c_int()
        nf, nj = C.c_longlong(), C.c_longlong()
        lo, hi = np.zeros(3), np.zeros(3)
        lib().eph_md_info(self.h, C.byref(nl), C.byref(ng), C.byref(nf), C.byref(nj), _d(lo), _d(hi))
        return dict(nlocal=nl.value, nghost=ng.value, nfix=nf.value, nlj=nj.value, boxlo=lo, boxhi=hi)

    def snapshot(self):
        inf = self.info()
        nl, ng = inf["nlocal"], inf["nghost"]
        nall = nl + ng
        s = MDSystem()
        s.nlocal, s.nghost = nl, ng
        s.boxlo, s.boxhi = inf["boxlo"], inf["boxhi"]
        s.x, s.v = np.zeros((nall, 3)), np.zeros((nall, 3))
        s.type, s.mask, s.tag = (np.zeros(nall, np.int32) for _ in range(3))
        s.ghost_owner = np.zeros(ng, np.int32)
        s.ghost_shift = np.zeros((ng, 3))
        s.nbr_off = np.zeros(nl + 1, np.int32)
        s.nbr_idx = np.zeros(inf["nfix"], np.int32)
        lib().eph_md_get(self.h, _d(s.x), _d(s.v), _i(s.type), _i(s.mask), _i(s.tag),
                         _i(s.ghost_owner), _d(s.ghost_shift), _i(s.nbr_off), _i(s.nbr_idx))
        s.f = np.zeros((nall, 3))
        return s

    def update_ghosts(self, s):
        lib().eph_md_update_ghosts(self.h, _d(s.x), _d(s.v))

    def lj(self, s):
        """pair forces into s.f (all local rows overwritten)"""
        f = np.zeros((s.nlocal, 3))
        lib().eph_md_lj_host(self.h, _d(s.x), _d(f))
        s.f[:s.nlocal] = f

    def needs_rebuild(self, s):
        return bool(lib().eph_md_needs_rebuild(self.h, _d(f64(s.x[:s.nlocal]))))

    def rebuild(self, s):
        x = f64(s.x[:s.nlocal]); v = f64(s.v[:s.nlocal])
        rc = lib().eph_md_rebuild(self.h, _d(x), _d(v))
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_md_last_error(self.h).decode())

    def attach(self, fix):
        rc = lib().eph_md_attach(self.h, fix.h)
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_md_last_error(self.h).decode())
        inf = self.info()
        fix.nlocal, fix.nghost = inf["nlocal"], inf["nghost"]

    def set_device_rebuild(self, on=True):
        lib().eph_md_set_device_rebuild(self.h, 1 if on else 0)

    def force_rebuild(self):
        lib().eph_md_force_rebuild(self.h)

    def pull(self, fix):
        """copy the device-resident system back so snapshot() describes it"""
        rc = lib().eph_md_pull(self.h, fix.h)
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_md_last_error(self.h).decode())

    def checkpoint_save(self, fix, path=None):
        """full-precision binary checkpoint of the device-resident run (atoms,
        forces, ghosts, lists and the fix's state); path None keeps it in the
        harness's host memory"""
        fix._check(lib().eph_md_checkpoint_save(self.h, fix.h, None if path is None else path.encode()))

    def checkpoint_load(self, fix, path=None):
        fix._check(lib().eph_md_checkpoint_load(self.h, fix.h, None if path is None else path.encode()))
        inf = self.info()
        fix.nlocal, fix.nghost = inf["nlocal"], inf["nghost"]

    def run(self, fix, nsteps, first_step=1):
        t = np.zeros(4)
        nreb = C.c_int()
        rc = lib().eph_md_run(self.h, fix.h, nsteps, first_step, _d(t), C.byref(nreb))
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_md_last_error(self.h).decode())
        inf = self.info()
        fix.nlocal, fix.nghost = inf["nlocal"], inf["nghost"]
        return dict(fix_ms=t[0], total_ms=t[1], pair_ms=t[2], rebuild_ms=t[3], rebuilds=nreb.value)


class Ensemble:
    """Replicas of (harness system, fix) advanced together on one GPU
    (include/eph_ensemble.h).  Each replica keeps its own seed, T_e grid and
    energy tally; results equal separate runs rebuilt at the same steps."""

    def __init__(self, device=0):
        h = vp()
        rc = lib().eph_ensemble_create(C.byref(h), device)
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_ensemble_last_error(None).decode())
        self.h = h
        self.nrep = 0

    def close(self):
        if getattr(self, "h", None):
            lib().eph_ensemble_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_ensemble_last_error(self.h).decode())

    def add(self, md_kwargs, args, ntypes, dt, boltz, ftm2v, masses):
        """md_kwargs: MD(...) keywords (cells=..., seed=..., pka_eV=...);
        args: the fix command words (box taken from the replica's system)."""
        kw = dict(a=3.615, ntypes=1, seed=1, T=300.0, disp_sigma=0.0, pka_eV=0.0, mass=63.546,
                  cut_fix=4.5, cut_lj=5.5, skin=2.0, lj_eps=0.4, lj_sigma=2.3)
        kw.update(md_kwargs)
        cfg = MDConfig(kw["cells"], kw["a"], kw["ntypes"], kw["seed"] & 0xFFFFFFFF, kw["T"],
                       kw["disp_sigma"], kw["pka_eV"], kw["mass"], kw["cut_fix"], kw["cut_lj"],
                       kw["skin"], kw["lj_eps"], kw["lj_sigma"])
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        m = f64(np.concatenate([[0.0], np.asarray(masses, dtype=np.float64)]))
        if len(m) != ntypes + 1:
            raise ValueError("one mass per atom type")
        self._check(lib().eph_ensemble_add(self.h, C.byref(cfg), len(args), arr, ntypes, dt, boltz,
                                           ftm2v, _d(m)))
        self.nrep += 1

    def setup(self, first_step=1):
        self._check(lib().eph_ensemble_setup(self.h, first_step))

    def run(self, nsteps, graph=True, log=True):
        ms = C.c_double()
        rebuilt = np.zeros(max(nsteps, 1), np.int32)
        self._check(lib().eph_ensemble_run(self.h, nsteps, 1 if graph else 0, C.byref(ms),
                                           _i(rebuilt) if log else None))
        return dict(ms=ms.value, rebuilt=rebuilt[:nsteps] if log else None)

    # the sweep forms of include/eph_ensemble.h, named as FixEPH names them
    _FORM_NAMES = FixEPH._FORM_NAMES

    def autotune(self):
        """time every form of the density and force sweeps and keep the fastest
        of each (same bits either way); forms that cannot run here are left out"""
        nf = len(self._FORM_NAMES)
        v = np.zeros(2, np.int32)
        ms = np.zeros(2 * nf)
        self._check(lib().eph_ensemble_autotune(self.h, _i(v), _d(ms)))
        names = self._FORM_NAMES
        return dict(density=names[v[0]], force=names[v[1]],
                    density_ms={names[k]: ms[k] for k in range(nf) if ms[k] >= 0},
                    force_ms={names[k]: ms[nf + k] for k in range(nf) if ms[nf + k] >= 0})

    def set_variants(self, density, force):
        """pin the sweep forms by name: chunk, lds, lds512, lds256, chunk4, chunk3, chunk2"""
        f = {n: k for k, n in enumerate(self._FORM_NAMES)}
        self._check(lib().eph_ensemble_set_variants(self.h, f[density], f[force]))

    def size(self):
        n, na, nc = C.c_int(), C.c_longlong(), C.c_longlong()
        self._check(lib().eph_ensemble_size(self.h, C.byref(n), C.byref(na), C.byref(nc)))
        return dict(replicas=n.value, atoms=na.value, cells=nc.value)

    def info(self, r):
        nl, nc, ns = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().eph_ensemble_info(self.h, r, C.byref(nl), C.byref(nc), C.byref(ns)))
        return dict(nlocal=nl.value, cells=nc.value, substeps=ns.value)

    def get(self, r, which):
        """which: 'x', 'v', 'f' ([nlocal][3]), 'T' ([cells]), 'E' ((E_local, Ee))"""
        k = {"x": 0, "v": 1, "f": 2, "T": 3, "E": 4}[which]
        inf = self.info(r)
        out = np.zeros((inf["nlocal"], 3)) if k < 3 else np.zeros(inf["cells"] if k == 3 else 2)
        self._check(lib().eph_ensemble_get(self.h, r, k, _d(out)))
        return out

    def candidate_stats(self):
        """The batch density sweep's candidate lists at its last chunk-form run:
        waves that scanned the full list, waves in all, moved atoms, retaken"""
        out = np.zeros(4)
        self._check(lib().eph_ensemble_get(self.h, 0, 5, _d(out)))
        return {"full_waves": int(out[0]), "waves": int(out[1]), "moved": int(out[2]),
                "retaken": bool(out[3])}


def spline_jacobian(dx, y):
    """d(table coefficients)/d(knot values) of spline_build: shape (n, 4, n)."""
    y = f64(y)
    n = len(y)
    out = np.zeros((n, 4, n))
    rc = lib().eph_spline_jacobian(dx, _d(y), n, _d(out))
    if rc != EPH_OK:
        raise EphError(rc, "spline Jacobian")
    return out


class Adjoint:
    """Loss L = Ee(after nsteps) + sum_c (T_e,c - T_target,c)^2 of the device run of
    `md` with `fix` (attached), and its gradient w.r.t. the beta knots and the initial
    velocities (include/eph_adjoint.h).  Every forward starts from the state at creation."""

    def __init__(self, md, fix, nsteps, every=20, T_target=None):
        h = vp()
        tt = None if T_target is None else f64(T_target)
        rc = lib().eph_adjoint_create(C.byref(h), md.h, fix.h, nsteps, every,
                                      None if tt is None else _d(tt))
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_adjoint_last_error(None).decode())
        self.h = h
        self._keep = (md, fix)
        ne, nb, nl, nc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().eph_adjoint_info(h, C.byref(ne), C.byref(nb), C.byref(nl), C.byref(nc))
        self.n_elem, self.n_beta, self.nlocal, self.ncells = ne.value, nb.value, nl.value, nc.value

    def close(self):
        if getattr(self, "h", None):
            lib().eph_adjoint_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != EPH_OK:
            raise EphError(rc, lib().eph_adjoint_last_error(self.h).decode())

    def beta0(self):
        out = np.zeros((self.n_elem, self.n_beta))
        self._check(lib().eph_adjoint_get_beta(self.h, _d(out)))
        return out

    def v0(self):
        out = np.zeros((self.nlocal, 3))
        self._check(lib().eph_adjoint_get_v0(self.h, _d(out)))
        return out

    def forward(self, beta=None, v0=None):
        """(loss, Ee, grid term) of the trajectory with these beta knots / start velocities"""
        b = None if beta is None else f64(beta)
        v = None if v0 is None else f64(v0)
        loss, parts = C.c_double(), np.zeros(2)
        self._check(lib().eph_adjoint_forward(self.h, None if b is None else _d(b),
                                              None if v is None else _d(v), C.byref(loss), _d(parts)))
        return loss.value, parts[0], parts[1]

    def backward(self):
        """(dL/dbeta [n_elem, n_beta], dL/dv0 [nlocal, 3]) of the last forward"""
        gb = np.zeros((self.n_elem, self.n_beta))
        gv = np.zeros((self.nlocal, 3))
        self._check(lib().eph_adjoint_backward(self.h, _d(gb), _d(gv)))
        return gb, gv


def exported_symbols(header_paths):
    """Function names declared in the C headers (for the export check)."""
    import re
    names = []
    for p in header_paths:
        text = open(p).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(eph_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))
