"""ctypes binding of the oracle (liblsqt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the reference results to compare
with (or the CPU baseline),
never as the thing measured or shipped.  See lsqt_oracle.h for what each
function restates in the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsqt_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


class Model(C.Structure):
    _fields_ = [
        ("use_lattice_model", C.c_int),
        ("calculate_vac0", C.c_int), ("calculate_vac", C.c_int), ("calculate_msd", C.c_int),
        ("calculate_spin", C.c_int), ("calculate_ldos", C.c_int),
        ("has_anderson", C.c_int), ("anderson_W", C.c_double),
        ("has_charge", C.c_int), ("charge_Ni", C.c_int),
        ("charge_W", C.c_double), ("charge_xi", C.c_double),
        ("has_vacancy", C.c_int), ("number_of_vacancies", C.c_int),
        ("number_of_random_vectors", C.c_int), ("number_of_moments", C.c_int),
        ("energy_max", C.c_double), ("requires_time", C.c_int),
        ("number_of_atoms", C.c_int), ("max_neighbor", C.c_int), ("number_of_pairs", C.c_int),
        ("neighbor_number", _ip), ("neighbor_list", _ip), ("potential", _dp),
        ("hopping_real", _dp), ("hopping_imag", _dp), ("xx", _dp), ("volume", C.c_double),
        ("number_of_energy_points", C.c_int), ("energy", _dp),
        ("number_of_steps_correlation", C.c_int), ("time_step", _dp),
        ("number_of_local_orbitals", C.c_int), ("local_orbitals", _ip),
        ("input_dir", C.c_char * 4096), ("generator", MT), ("error", C.c_char * 1024),
    ]


class Ham(C.Structure):
    _fields_ = [("n", C.c_int), ("max_neighbor", C.c_int), ("energy_max", C.c_double),
                ("neighbor_number", _ip), ("neighbor_list", _ip), ("potential", _dp),
                ("hopping_real", _dp), ("hopping_imag", _dp), ("xx", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_model_load.argtypes = [C.c_char_p, C.POINTER(Model)]
        L.orc_model_free.argtypes = [C.POINTER(Model)]
        L.orc_initialize_state.argtypes = [C.POINTER(Model), C.c_int, _dp, _dp]
        L.orc_ham_from_model.argtypes = [C.POINTER(Model), C.POINTER(Ham)]
        L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
        L.orc_mt_next.argtypes = [C.POINTER(MT)]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_uniform_real.argtypes = [C.POINTER(MT), C.c_double, C.c_double]
        L.orc_uniform_real.restype = C.c_double
        L.orc_uniform_int_0_intmax.argtypes = [C.POINTER(MT)]
        L.orc_uniform_int_0_intmax.restype = C.c_int
        for name in ("orc_apply_hamiltonian", "orc_apply_commutator", "orc_apply_current"):
            getattr(L, name).argtypes = [C.POINTER(Ham), _dp, _dp, _dp, _dp]
        L.orc_kernel_polynomial.argtypes = [C.POINTER(Ham)] + [_dp] * 6
        L.orc_inner_product.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_find_moments.argtypes = [C.POINTER(Ham), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_apply_damping.argtypes = [C.c_int, _dp, _dp]
        L.orc_apply_damping_lorentz.argtypes = [C.c_int, _dp, _dp]
        L.orc_chebyshev_summation.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_double,
                                              C.c_double, _dp]
        L.orc_bessel_stop.argtypes = [C.c_double]
        L.orc_bessel_stop.restype = C.c_int
        L.orc_evolve.argtypes = [C.POINTER(Ham), C.c_int, C.c_double, _dp, _dp]
        L.orc_evolvex.argtypes = [C.POINTER(Ham), C.c_int, C.c_double, _dp, _dp]
        for name in ("orc_find_dos", "orc_find_vac0", "orc_find_vac", "orc_find_msd",
                     "orc_find_spin"):
            getattr(L, name).argtypes = [C.POINTER(Model), C.POINTER(Ham), _dp, _dp, _dp, _dp]
        L.orc_lsqt.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _ipp(a):
    return a.ctypes.data_as(_ip)


class OracleError(RuntimeError):
    pass


class OracleModel:
    """A model built by the oracle from an input directory."""

    def __init__(self, dirname: str):
        self._m = Model()
        if lib().orc_model_load(dirname.encode(), C.byref(self._m)):
            msg = self._m.error.decode()
            lib().orc_model_free(C.byref(self._m))
            raise OracleError(msg)
        self._h = Ham()
        lib().orc_ham_from_model(C.byref(self._m), C.byref(self._h))

    def __del__(self):
        try:
            lib().orc_model_free(C.byref(self._m))
        except Exception:
            pass

    # -- model fields --
    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return getattr(self._m, name)

    @property
    def n(self):
        return self._m.number_of_atoms

    def arrays(self):
        """Row-major ELL arrays (copies)."""
        m = self._m
        N, K = m.number_of_atoms, m.max_neighbor
        P = N * K
        nn = np.ctypeslib.as_array(m.neighbor_number, (N,)).copy()
        nl = np.ctypeslib.as_array(m.neighbor_list, (P,)).copy()
        pot = np.ctypeslib.as_array(m.potential, (N,)).copy()
        hr = np.ctypeslib.as_array(m.hopping_real, (P,)).copy()
        hi = np.ctypeslib.as_array(m.hopping_imag, (P,)).copy()
        xx = np.ctypeslib.as_array(m.xx, (P,)).copy()
        return dict(neighbor_number=nn, neighbor_list=nl, potential=pot, hopping_real=hr,
                    hopping_imag=hi, xx=xx)

    def valid_pairs(self):
        """Arrays of the valid pairs only, in (site, slot) order."""
        a = self.arrays()
        K = self._m.max_neighbor
        nn = a["neighbor_number"]
        mask = (np.arange(K)[None, :] < nn[:, None]).reshape(-1)
        return {k: a[k][mask] for k in ("neighbor_list", "hopping_real", "hopping_imag", "xx")}

    @property
    def energy(self):
        return np.ctypeslib.as_array(self._m.energy, (self._m.number_of_energy_points,)).copy()

    @property
    def time_step(self):
        n = self._m.number_of_steps_correlation
        return np.ctypeslib.as_array(self._m.time_step, (n,)).copy() if n else np.zeros(0)

    @property
    def local_orbitals(self):
        n = self._m.number_of_local_orbitals
        return np.ctypeslib.as_array(self._m.local_orbitals, (n,)).copy() if n else []

    # -- numerics --
    def random_state(self, orbital: int = -1):
        N = self.n
        re, im = np.zeros(N), np.zeros(N)
        lib().orc_initialize_state(C.byref(self._m), orbital, _p(re), _p(im))
        return re, im

    def _op(self, name, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        o_re, o_im = np.zeros(self.n), np.zeros(self.n)
        getattr(lib(), name)(C.byref(self._h), _p(re), _p(im), _p(o_re), _p(o_im))
        return o_re, o_im

    def apply(self, re, im):
        return self._op("orc_apply_hamiltonian", re, im)

    def apply_commutator(self, re, im):
        return self._op("orc_apply_commutator", re, im)

    def apply_current(self, re, im):
        return self._op("orc_apply_current", re, im)

    def kernel_polynomial(self, s0, s1):
        out_r, out_i = np.zeros(self.n), np.zeros(self.n)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (*s0, *s1)]
        lib().orc_kernel_polynomial(C.byref(self._h), *(_p(x) for x in a), _p(out_r), _p(out_i))
        return out_r, out_i

    def moments(self, L, R, M=None):
        M = M or self._m.number_of_moments
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (*L, *R)]
        mr, mi = np.zeros(M), np.zeros(M)
        lib().orc_find_moments(C.byref(self._h), M, *(_p(x) for x in a), _p(mr), _p(mi))
        return mr, mi

    def evolve(self, direction, t_scaled, re, im):
        re, im = np.array(re, dtype=np.float64), np.array(im, dtype=np.float64)
        lib().orc_evolve(C.byref(self._h), direction, t_scaled, _p(re), _p(im))
        return re, im

    def evolvex(self, direction, t_scaled, re, im):
        re, im = np.array(re, dtype=np.float64), np.array(im, dtype=np.float64)
        lib().orc_evolvex(C.byref(self._h), direction, t_scaled, _p(re), _p(im))
        return re, im

    def _driver(self, name, r, nrows):
        M, NE = self._m.number_of_moments, self._m.number_of_energy_points
        rows = np.zeros((nrows, NE))
        mu = np.zeros((nrows, 2, M))
        rr, ri = (np.ascontiguousarray(x, dtype=np.float64) for x in r)
        getattr(lib(), name)(C.byref(self._m), C.byref(self._h), _p(rr), _p(ri), _p(rows), _p(mu))
        return rows, mu[:, 0, :] + 1j * mu[:, 1, :]

    def find_dos(self, r):
        return self._driver("orc_find_dos", r, 1)

    def find_vac0(self, r):
        return self._driver("orc_find_vac0", r, 1)

    def find_vac(self, r):
        return self._driver("orc_find_vac", r, self._m.number_of_steps_correlation)

    def find_msd(self, r):
        return self._driver("orc_find_msd", r, self._m.number_of_steps_correlation)

    def find_spin(self, r):
        return self._driver("orc_find_spin", r, self._m.number_of_steps_correlation)


class OracleHam:
    """The oracle's operators on Hamiltonian arrays given directly (row-major
    ELL, as a Model builds them), for systems that no input directory
    describes: the same functions as OracleModel's numerics."""

    def __init__(self, arrays, n, max_neighbor, energy_max):
        self._keep = {
            "neighbor_number": np.ascontiguousarray(arrays["neighbor_number"], dtype=np.int32),
            "neighbor_list": np.ascontiguousarray(arrays["neighbor_list"], dtype=np.int32),
        }
        for k in ("potential", "hopping_real", "hopping_imag", "xx"):
            self._keep[k] = np.ascontiguousarray(arrays[k], dtype=np.float64)
        assert self._keep["neighbor_number"].size == n
        assert self._keep["neighbor_list"].size == n * max_neighbor
        h = Ham()
        h.n, h.max_neighbor, h.energy_max = n, max_neighbor, float(energy_max)
        h.neighbor_number = _ipp(self._keep["neighbor_number"])
        h.neighbor_list = _ipp(self._keep["neighbor_list"])
        for k in ("potential", "hopping_real", "hopping_imag", "xx"):
            setattr(h, k, _p(self._keep[k]))
        self._h = h
        self.n = n

    apply = OracleModel.apply
    apply_commutator = OracleModel.apply_commutator
    apply_current = OracleModel.apply_current
    kernel_polynomial = OracleModel.kernel_polynomial
    _op = OracleModel._op
    evolve = OracleModel.evolve
    evolvex = OracleModel.evolvex

    def moments(self, L, R, M):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (*L, *R)]
        mr, mi = np.zeros(M), np.zeros(M)
        lib().orc_find_moments(C.byref(self._h), M, *(_p(x) for x in a), _p(mr), _p(mi))
        return mr, mi


def damping(mu_re, mu_im=None, lorentz=False):
    mu_re = np.array(mu_re, dtype=np.float64)
    mu_im = np.zeros_like(mu_re) if mu_im is None else np.array(mu_im, dtype=np.float64)
    fn = lib().orc_apply_damping_lorentz if lorentz else lib().orc_apply_damping
    fn(len(mu_re), _p(mu_re), _p(mu_im))
    return mu_re, mu_im


def summation(mu_re, energies, energy_max, volume):
    mu_re = np.ascontiguousarray(mu_re, dtype=np.float64)
    e = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros(len(e))
    lib().orc_chebyshev_summation(len(mu_re), _p(mu_re), len(e), _p(e), energy_max, volume,
                                  _p(out))
    return out


def bessel_stop(t_scaled):
    return lib().orc_bessel_stop(t_scaled)


def lsqt(dirname: str):
    err = C.create_string_buffer(1024)
    rc = lib().orc_lsqt(dirname.encode(), err, 1024)
    if rc:
        raise OracleError(err.value.decode())


def set_threads(n: int):
    lib().orc_set_num_threads(n)


def num_threads() -> int:
    return lib().orc_num_threads()


def mt_first_draws(seed: int, count: int):
    g = MT()
    lib().orc_mt_seed(C.byref(g), seed)
    return np.array([lib().orc_mt_next(C.byref(g)) for _ in range(count)], dtype=np.uint64)


def uniform_real_draws(seed: int, a: float, b: float, count: int):
    g = MT()
    lib().orc_mt_seed(C.byref(g), seed)
    return np.array([lib().orc_uniform_real(C.byref(g), a, b) for _ in range(count)])


def uniform_int_draws(seed: int, count: int):
    g = MT()
    lib().orc_mt_seed(C.byref(g), seed)
    return np.array([lib().orc_uniform_int_0_intmax(C.byref(g)) for _ in range(count)])


def reference_binary() -> str | None:
    """The reference CPU program compiled from source by `make -C oracle ref`."""
    p = os.path.join(REF_DIR, "lsqt_cpu")
    return p if os.path.exists(p) else None


def run_reference(driver_file: str, timeout: float | None = None):
    p = reference_binary()
    if p is None:
        raise RuntimeError("reference program not built (make -C oracle ref)")
    return subprocess.run([p, driver_file], capture_output=True, text=True, timeout=timeout)
