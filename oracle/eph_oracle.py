"""ctypes wrapper of the CPU parity oracle (oracle/libeph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  See
eph_oracle.cpp for what each function restates (USER-EPH file:line) and how
each part of the oracle is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeph_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "eph_ref_driver")

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
u32p = C.POINTER(C.c_uint32)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "ora_last_error": (C.c_char_p, []),
            "ora_spline_build": (C.c_int, [C.c_double, dp, C.c_int, dp]),
            "ora_spline_eval": (None, [dp, C.c_int, C.c_double, dp, C.c_int, dp]),
            "ora_linear_eval": (None, [C.c_double, dp, C.c_int, dp, C.c_int, dp, dp]),
            "ora_beta_load": (C.c_void_p, [C.c_char_p]),
            "ora_beta_free": (None, [C.c_void_p]),
            "ora_beta_info": (None, [C.c_void_p, ip, dp, dp, ip, ip]),
            "ora_beta_table": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp]),
            "ora_beta_name": (C.c_char_p, [C.c_void_p, C.c_int]),
            "ora_fdm_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, dp, C.c_double,
                                            C.c_double, C.c_double, C.c_double]),
            "ora_fdm_load": (C.c_void_p, [C.c_char_p]),
            "ora_fdm_free": (None, [C.c_void_p]),
            "ora_fdm_set_dt": (None, [C.c_void_p, C.c_double]),
            "ora_fdm_set_steps": (None, [C.c_void_p, C.c_int]),
            "ora_fdm_set_cell": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.c_int]),
            "ora_fdm_set_source": (None, [C.c_void_p, C.c_int, C.c_double]),
            "ora_fdm_insert_energy": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                             C.c_double]),
            "ora_fdm_add_power": (None, [C.c_void_p, dp, C.c_int]),
            "ora_fdm_solve": (C.c_int, [C.c_void_p]),
            "ora_fdm_last_r": (C.c_double, [C.c_void_p]),
            "ora_fdm_get_T": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
            "ora_fdm_get_index": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
            "ora_fdm_T_total": (C.c_double, [C.c_void_p]),
            "ora_fdm_dims": (None, [C.c_void_p, ip]),
            "ora_fdm_get": (None, [C.c_void_p, C.c_int, dp]),
            "ora_fdm_set_T": (None, [C.c_void_p, dp]),
            "ora_fdm_save_state": (C.c_int, [C.c_void_p, C.c_char_p]),
            "ora_fdm_save_temperature": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
            "ora_philox": (None, [u32p, u32p, u32p]),
            "ora_xi_batch": (None, [C.c_uint32, ip, C.c_int, C.c_long, dp]),
            "ora_log_fn": (C.c_double, [C.c_double]),
            "ora_sincos2pi_fn": (None, [C.c_double, dp, dp]),
            "ora_fix_create": (C.c_void_p, [C.c_int, C.POINTER(C.c_char_p), C.c_int, dp, dp,
                                            C.c_double, C.c_double, C.c_double]),
            "ora_fix_free": (None, [C.c_void_p]),
            "ora_fix_set_mass": (None, [C.c_void_p, C.c_int, dp]),
            "ora_fix_set_groupbit": (None, [C.c_void_p, C.c_int]),
            "ora_fix_reset_dt": (None, [C.c_void_p, C.c_double]),
            "ora_fix_initial_integrate": (None, [C.c_void_p, C.c_int, dp, dp, dp, ip, ip]),
            "ora_fix_final_integrate": (None, [C.c_void_p, C.c_int, dp, dp, ip, ip]),
            "ora_fix_post_force": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, ip, ip,
                                             ip, ip, ip, dp, ip, C.c_long]),
            "ora_fix_end_of_step": (None, [C.c_void_p, C.c_int, dp, dp, ip, ip, C.c_long]),
            "ora_fix_compute_vector": (C.c_double, [C.c_void_p, C.c_int]),
            "ora_fix_last_E_local": (C.c_double, [C.c_void_p]),
            "ora_fix_get": (None, [C.c_void_p, C.c_int, C.c_int, dp]),
            "ora_fix_fdm": (C.c_void_p, [C.c_void_p]),
            "ora_fix_post_run": (C.c_int, [C.c_void_p]),
            "ora_fix_alpha": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_set_threads": (C.c_int, [C.c_int]),
            "ora_fix_rho_r_sq": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_fix_rho_r": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_fix_beta": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_fix_zeta": (C.c_double, [C.c_void_p]),
            "ora_kappa_load": (C.c_void_p, [C.c_char_p]),
            "ora_kappa_free": (None, [C.c_void_p]),
            "ora_kappa_info": (None, [C.c_void_p, ip, ip, dp, dp]),
            "ora_kappa_table": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp]),
            "ora_kappa_T_of_E": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_kappa_E_of_T": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_kappa_K": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_kappa_rho_r_sq": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "ora_atomic_create": (C.c_void_p, [C.c_int, C.POINTER(C.c_char_p), C.c_int,
                                               C.c_double, C.c_double, C.c_double]),
            "ora_atomic_free": (None, [C.c_void_p]),
            "ora_atomic_set_mass": (None, [C.c_void_p, C.c_int, dp]),
            "ora_atomic_set_groupbit": (None, [C.c_void_p, C.c_int]),
            "ora_atomic_initial_integrate": (None, [C.c_void_p, C.c_int, dp, dp, dp, ip, ip]),
            "ora_atomic_final_integrate": (None, [C.c_void_p, C.c_int, dp, dp, ip, ip]),
            "ora_atomic_post_force": (C.c_int, [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, ip, ip,
                                                ip, ip, ip, dp, ip, C.c_long]),
            "ora_atomic_end_of_step": (None, [C.c_void_p, C.c_int, C.c_int, dp, ip, ip, ip, ip, ip]),
            "ora_atomic_initialise": (None, [C.c_void_p, C.c_int, ip, ip]),
            "ora_atomic_compute_vector": (C.c_double, [C.c_void_p, C.c_int]),
            "ora_atomic_set_E": (None, [C.c_void_p, C.c_int, dp]),
            "ora_atomic_get": (None, [C.c_void_p, C.c_int, C.c_int, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --- splines -----------------------------------------------------------------
def spline_build(dx, y):
    y = _f64(y)
    out = np.zeros((len(y), 4))
    tail = lib().ora_spline_build(dx, _d(y), len(y), _d(out))
    return out, bool(tail)


def spline_eval(coeffs, inv_dx, x):
    coeffs = _f64(coeffs)
    x = _f64(np.atleast_1d(x))
    out = np.zeros_like(x)
    lib().ora_spline_eval(_d(coeffs), len(coeffs), inv_dx, _d(x), len(x), _d(out))
    return out


def linear_eval(dx, y, x):
    y = _f64(y)
    x = _f64(np.atleast_1d(x))
    fwd = np.zeros_like(x)
    rev = np.zeros_like(x)
    lib().ora_linear_eval(dx, _d(y), len(y), _d(x), len(x), _d(fwd), _d(rev))
    return fwd, rev


class Beta:
    def __init__(self, path):
        self.h = lib().ora_beta_load(path.encode())
        if not self.h:
            raise RuntimeError(lib().ora_last_error().decode())
        n = C.c_int(); rc = C.c_double(); rhoc = C.c_double(); nr = C.c_int(); nb = C.c_int()
        lib().ora_beta_info(self.h, C.byref(n), C.byref(rc), C.byref(rhoc), C.byref(nr), C.byref(nb))
        self.n_elements, self.r_cutoff, self.rho_cutoff = n.value, rc.value, rhoc.value
        self.n_rho, self.n_beta = nr.value, nb.value
        self.names = [lib().ora_beta_name(self.h, e).decode() for e in range(self.n_elements)]

    def table(self, elem, which):
        n = self.n_rho if which in (0, 1) else self.n_beta
        out = np.zeros((n, 4))
        inv = C.c_double()
        lib().ora_beta_table(self.h, elem, which, _d(out), C.byref(inv))
        return out, inv.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_beta_free(self.h)
            self.h = None


class FDM:
    """EPH_FDM restatement (eph_fdm.h)."""

    def __init__(self, h, owned=True):
        self.h = h
        self.owned = owned
        n = (C.c_int * 3)()
        lib().ora_fdm_dims(self.h, n)
        self.dims = (n[0], n[1], n[2])

    @classmethod
    def const(cls, nx, ny, nz, box, T, C_e, rho, kappa):
        box = _f64(box)
        return cls(lib().ora_fdm_create(nx, ny, nz, _d(box), T, C_e, rho, kappa))

    @classmethod
    def load(cls, path):
        h = lib().ora_fdm_load(path.encode())
        if not h:
            raise RuntimeError(lib().ora_last_error().decode())
        return cls(h)

    @property
    def ntotal(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def set_dt(self, dt): lib().ora_fdm_set_dt(self.h, dt)
    def set_steps(self, s): lib().ora_fdm_set_steps(self.h, s)

    def set_cell(self, i, j, k, T, S, rho, Ce, kappa, flag, tdyn):
        lib().ora_fdm_set_cell(self.h, i, j, k, T, S, rho, Ce, kappa, flag, tdyn)

    def insert_energy(self, x, y, z, E): lib().ora_fdm_insert_energy(self.h, x, y, z, E)
    def set_source(self, idx, S): lib().ora_fdm_set_source(self.h, idx, S)

    def add_power(self, p):
        p = _f64(p)
        lib().ora_fdm_add_power(self.h, _d(p), len(p))

    def solve(self): return lib().ora_fdm_solve(self.h)
    def last_r(self): return lib().ora_fdm_last_r(self.h)
    def get_T_at(self, x, y, z): return lib().ora_fdm_get_T(self.h, x, y, z)
    def get_index(self, x, y, z): return lib().ora_fdm_get_index(self.h, x, y, z)
    def T_total(self): return lib().ora_fdm_T_total(self.h)

    def get(self, which="T"):
        idx = {"T": 0, "dT": 1, "C": 2, "rho": 3, "kappa": 4, "S": 5}[which]
        out = np.zeros(self.ntotal)
        lib().ora_fdm_get(self.h, idx, _d(out))
        return out

    def set_T(self, T):
        T = _f64(T)
        assert T.size == self.ntotal
        lib().ora_fdm_set_T(self.h, _d(T))

    def save_state(self, path): return lib().ora_fdm_save_state(self.h, path.encode())
    def save_temperature(self, base, n): return lib().ora_fdm_save_temperature(self.h, base.encode(), n)

    def __del__(self):
        if getattr(self, "h", None) and self.owned:
            lib().ora_fdm_free(self.h)
            self.h = None


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().ora_philox(c.ctypes.data_as(u32p), k.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
    return out


def xi(seed, tags, step):
    tags = _i32(tags)
    out = np.zeros((len(tags), 3))
    lib().ora_xi_batch(seed & 0xFFFFFFFF, _i(tags), len(tags), step, _d(out))
    return out


def log_fn(x): return lib().ora_log_fn(x)


def sincos2pi(u):
    s = C.c_double(); c = C.c_double()
    lib().ora_sincos2pi_fn(u, C.byref(s), C.byref(c))
    return s.value, c.value


class Fix:
    """FixEPH restatement (fix_eph.cpp: models 1, 2, 4 and TESTING) and the
    eph/coloured, eph/coloured/exp and eph/coloured/exp/v1 styles; one rank."""

    def __init__(self, args, ntypes, boxlo, boxhi, dt, boltz, ftm2v, masses=None):
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        lo = _f64(boxlo); hi = _f64(boxhi)
        self.h = lib().ora_fix_create(len(args), arr, ntypes, _d(lo), _d(hi), dt, boltz, ftm2v)
        if not self.h:
            raise RuntimeError(lib().ora_last_error().decode())
        self.ntypes = ntypes
        if masses is not None:
            self.set_mass(masses)

    def set_mass(self, masses):
        m = _f64(np.concatenate([[0.0], np.asarray(masses, dtype=np.float64)]))
        lib().ora_fix_set_mass(self.h, self.ntypes, _d(m))

    def set_groupbit(self, gb): lib().ora_fix_set_groupbit(self.h, gb)
    def reset_dt(self, dt): lib().ora_fix_reset_dt(self.h, dt)

    def initial_integrate(self, s):
        lib().ora_fix_initial_integrate(self.h, s.nlocal, _d(s.x), _d(s.v), _d(s.f), _i(s.type), _i(s.mask))

    def final_integrate(self, s):
        lib().ora_fix_final_integrate(self.h, s.nlocal, _d(s.v), _d(s.f), _i(s.type), _i(s.mask))

    def post_force(self, s, xi=None, step=0):
        xip = None
        if xi is not None:
            xi = _f64(xi)
            xip = _d(xi)
        rc = lib().ora_fix_post_force(self.h, s.nlocal, s.nghost, _d(s.x), _d(s.v), _d(s.f),
                                      _i(s.type), _i(s.mask), _i(s.nbr_off), _i(s.nbr_idx),
                                      _i(s.ghost_owner), xip, _i(s.tag), step)
        if rc != 0:
            raise RuntimeError(lib().ora_last_error().decode())

    def end_of_step(self, s, ntimestep=0):
        lib().ora_fix_end_of_step(self.h, s.nlocal, _d(s.x), _d(s.v), _i(s.type), _i(s.mask), ntimestep)

    def compute_vector(self, i): return lib().ora_fix_compute_vector(self.h, i)
    def last_E_local(self): return lib().ora_fix_last_E_local(self.h)

    def get(self, which, n):
        idx = {"rho": 0, "w": 1, "xi": 2, "f_EPH": 3, "f_RNG": 4, "array": 5}[which]
        width = 1 if idx == 0 else 8 if idx == 5 else 3
        out = np.zeros(n * width)
        lib().ora_fix_get(self.h, idx, n, _d(out))
        return out if width == 1 else out.reshape(n, width)

    def fdm(self):
        return FDM(lib().ora_fix_fdm(self.h), owned=False)

    def post_run(self): return lib().ora_fix_post_run(self.h)
    def alpha(self, elem, rho): return lib().ora_fix_alpha(self.h, elem, rho)
    def rho_r_sq(self, elem, r2): return lib().ora_fix_rho_r_sq(self.h, elem, r2)
    def rho_r(self, elem, r): return lib().ora_fix_rho_r(self.h, elem, r)
    def beta(self, elem, rho): return lib().ora_fix_beta(self.h, elem, rho)
    def zeta(self): return lib().ora_fix_zeta(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_fix_free(self.h)
            self.h = None


class Kappa:
    """EPH_kappa restatement (eph_kappa.h:55-150)."""

    def __init__(self, path):
        self.h = lib().ora_kappa_load(path.encode())
        if not self.h:
            raise RuntimeError(lib().ora_last_error().decode())
        ne, npair = C.c_int(), C.c_int()
        rc, tm = C.c_double(), C.c_double()
        lib().ora_kappa_info(self.h, C.byref(ne), C.byref(npair), C.byref(rc), C.byref(tm))
        self.n_elements, self.n_pairs, self.r_cutoff, self.T_max = ne.value, npair.value, rc.value, tm.value

    def table(self, index, which):
        n = lib().ora_kappa_table(self.h, index, which, None, None)
        w = 4 if which < 2 else 2
        out = np.zeros(n * w)
        dx = C.c_double()
        lib().ora_kappa_table(self.h, index, which, _d(out), C.byref(dx))
        return out.reshape(n, w), dx.value

    def T_of_E(self, e, E): return lib().ora_kappa_T_of_E(self.h, e, E)
    def E_of_T(self, e, T): return lib().ora_kappa_E_of_T(self.h, e, T)
    def K(self, slot, T): return lib().ora_kappa_K(self.h, slot, T)
    def rho_r_sq(self, e, r2): return lib().ora_kappa_rho_r_sq(self.h, e, r2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_kappa_free(self.h)
            self.h = None


class Atomic:
    """FixEPHAtomic restatement (fix_eph_atomic.cpp), one rank; the same hook
    methods as Fix."""

    def __init__(self, args, ntypes, dt, boltz, ftm2v, masses=None):
        arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
        self.h = lib().ora_atomic_create(len(args), arr, ntypes, dt, boltz, ftm2v)
        if not self.h:
            raise RuntimeError(lib().ora_last_error().decode())
        self.ntypes = ntypes
        if masses is not None:
            m = _f64(np.concatenate([[0.0], np.asarray(masses, dtype=np.float64)]))
            lib().ora_atomic_set_mass(self.h, ntypes, _d(m))

    def set_groupbit(self, gb): lib().ora_atomic_set_groupbit(self.h, gb)

    def initialise(self, s):
        lib().ora_atomic_initialise(self.h, s.nlocal, _i(s.type), _i(s.mask))

    def initial_integrate(self, s):
        lib().ora_atomic_initial_integrate(self.h, s.nlocal, _d(s.x), _d(s.v), _d(s.f), _i(s.type),
                                           _i(s.mask))

    def final_integrate(self, s):
        lib().ora_atomic_final_integrate(self.h, s.nlocal, _d(s.v), _d(s.f), _i(s.type), _i(s.mask))

    def post_force(self, s, xi=None, step=0):
        xip = None
        if xi is not None:
            xi = _f64(xi)
            xip = _d(xi)
        rc = lib().ora_atomic_post_force(self.h, s.nlocal, s.nghost, _d(s.x), _d(s.v), _d(s.f),
                                         _i(s.type), _i(s.mask), _i(s.nbr_off), _i(s.nbr_idx),
                                         _i(s.ghost_owner), xip, _i(s.tag), step)
        if rc != 0:
            raise RuntimeError(lib().ora_last_error().decode())

    def end_of_step(self, s, ntimestep=0):
        lib().ora_atomic_end_of_step(self.h, s.nlocal, s.nghost, _d(s.x), _i(s.type), _i(s.mask),
                                     _i(s.nbr_off), _i(s.nbr_idx), _i(s.ghost_owner))

    def compute_vector(self, i): return lib().ora_atomic_compute_vector(self.h, i)

    def set_E(self, E):
        E = _f64(E)
        lib().ora_atomic_set_E(self.h, len(E), _d(E))

    _WHICH = {"rho": (0, 1), "w": (1, 3), "xi": (2, 3), "f_EPH": (3, 3), "f_RNG": (4, 3),
              "array": (5, 12), "rho_a": (6, 1), "E": (7, 1), "dE": (8, 1), "T": (9, 1)}

    def get(self, which, n):
        idx, width = self._WHICH[which]
        out = np.zeros(n * width)
        lib().ora_atomic_get(self.h, idx, n, _d(out))
        return out if width == 1 else out.reshape(n, width)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_atomic_free(self.h)
            self.h = None


# --- compiled-reference driver (golden generation only) -------------------------
def ref_available():
    return os.path.exists(REF_DRIVER)


def ref_run(*args):
    out = subprocess.run([REF_DRIVER, *map(str, args)], check=True, capture_output=True, text=True)
    return out.stdout


def set_threads(n):
    """threads of the per-atom loops (0: leave as is); returns the count in effect"""
    return lib().ora_set_threads(int(n))
