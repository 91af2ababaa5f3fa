"""ctypes binding of liblsqt_host.so (include/lsqt_host.h): the drop-in's
Model and driver, for tests and tools.  Model construction touches no GPU;
running a directory does."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsqt_host.so")
CLI_PATH = os.path.join(HERE, "lsqt_gpu")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

SIGNATURES = {
    "lsqt_host_main": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
    "lsqt_host_run": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int]),
    "lsqt_host_set_output_precision": (None, [C.c_int]),
    "lsqt_host_damped_sum": (C.c_int, [C.c_int, _dp, _dp, C.c_int, _dp, C.c_double,
                                       C.c_double, _dp]),
    "lsqt_host_model_load": (C.c_int, [C.c_char_p, C.POINTER(_vp), C.c_char_p, C.c_int]),
    "lsqt_host_model_load_ex": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_vp), C.c_char_p,
                                          C.c_int]),
    "lsqt_host_model_lattice": (C.c_int, [_vp, C.c_void_p, C.POINTER(_ip), _ip]),
    "lsqt_host_model_potential": (C.c_int, [_vp, _dp]),
    "lsqt_host_model_expand_lattice": (C.c_int, [_vp]),
    "lsqt_host_model_free": (None, [_vp]),
    "lsqt_host_model_info": (C.c_int, [_vp, _ip, _dp]),
    "lsqt_host_model_arrays": (C.c_int, [_vp, _ip, _ip, _dp, _dp, _dp, _dp]),
    "lsqt_host_model_inputs": (C.c_int, [_vp, _dp, _dp, _ip]),
    "lsqt_host_model_random_state": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "lsqt_host_model_evolve": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp]),
    "lsqt_host_model_moments": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]),
    "lsqt_host_conductivity": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int,
                                         _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_char_p,
                                         C.c_int]),
    "lsqt_host_postprocess": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int]),
    "lsqt_host_write_general": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _ip, _ip, _dp,
                                          C.c_double, C.c_double, _dp, _dp, _dp, C.c_char_p,
                                          C.c_int]),
}

_lib = None


class InputError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"host library not built: {LIB_PATH}")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


LATTICE_ON_DEVICE = 1


class HostModel:
    """The drop-in's Model built from an input directory.  lattice_on_device:
    a lattice model keeps its description (lattice()) instead of expanding
    its neighbour arrays, as the drop-in does before building the
    Hamiltonian on the GPU."""

    def __init__(self, dirname: str, lattice_on_device: bool = False):
        self.h = _vp()
        err = C.create_string_buffer(4096)
        flags = LATTICE_ON_DEVICE if lattice_on_device else 0
        if load().lsqt_host_model_load_ex(dirname.encode(), flags, C.byref(self.h), err, 4096):
            raise InputError(err.value.decode())
        ints = (C.c_int * 8)()
        reals = (C.c_double * 2)()
        load().lsqt_host_model_info(self.h, ints, reals)
        (self.n, self.max_neighbor, self.number_of_moments, self.number_of_random_vectors,
         self.number_of_energy_points, self.number_of_steps_correlation,
         self.number_of_local_orbitals, flags) = list(ints)
        self.calculate_vac0 = bool(flags & 1)
        self.calculate_vac = bool(flags & 2)
        self.calculate_msd = bool(flags & 4)
        self.calculate_spin = bool(flags & 8)
        self.calculate_ldos = bool(flags & 16)
        self.energy_max, self.volume = reals[0], reals[1]

    def arrays(self):
        N, P = self.n, self.n * self.max_neighbor
        nn = np.zeros(N, np.int32)
        nl = np.zeros(P, np.int32)
        pot, hr, hi, xx = np.zeros(N), np.zeros(P), np.zeros(P), np.zeros(P)
        if load().lsqt_host_model_arrays(self.h, nn.ctypes.data_as(_ip), nl.ctypes.data_as(_ip),
                                         *(a.ctypes.data_as(_dp) for a in (pot, hr, hi, xx))):
            raise InputError("the model holds no neighbour arrays (lattice kept for the "
                             "device: expand_lattice() first)")
        return dict(neighbor_number=nn, neighbor_list=nl, potential=pot, hopping_real=hr,
                    hopping_imag=hi, xx=xx)

    def lattice(self):
        """(lattice dict as gpuqt_amd.inputs.read_lattice gives it, site index
        after vacancies or None) of a model loaded with lattice_on_device;
        None for a model without a lattice description."""
        from .native import Lattice
        L = Lattice()
        idx = _ip()
        npr = C.c_int()
        if load().lsqt_host_model_lattice(self.h, C.byref(L), C.byref(idx), C.byref(npr)):
            return None
        No, K = L.orbitals, L.max_neighbor
        arr = lambda p, n: np.ctypeslib.as_array(p, (n,)).copy()
        lat = {"cells": list(L.cells), "pbc": list(L.pbc), "transport": L.transport,
               "lattice_constant": list(L.lattice_constant), "orbitals": No,
               "max_neighbor": K, "orbital_x": arr(L.orbital_x, No),
               "hop_count": arr(L.hop_count, No),
               "hop_cell": arr(L.hop_cell, No * K * 3).reshape(No, K, 3),
               "hop_orbital": arr(L.hop_orbital, No * K).reshape(No, K),
               "hop_re": arr(L.hop_re, No * K).reshape(No, K),
               "hop_im": arr(L.hop_im, No * K).reshape(No, K)}
        site_index = arr(idx, npr.value) if idx else None
        return lat, site_index

    def expand_lattice(self):
        """Expands a lattice_on_device model's neighbour arrays on the host."""
        if load().lsqt_host_model_expand_lattice(self.h):
            raise InputError("expanding the lattice failed")

    def potential(self):
        pot = np.zeros(self.n)
        if load().lsqt_host_model_potential(self.h, pot.ctypes.data_as(_dp)):
            raise InputError("the model holds no potential")
        return pot

    def valid_pairs(self):
        a = self.arrays()
        mask = (np.arange(self.max_neighbor)[None, :] < a["neighbor_number"][:, None]).reshape(-1)
        return {k: a[k][mask] for k in ("neighbor_list", "hopping_real", "hopping_imag", "xx")}

    def inputs(self):
        e = np.zeros(max(self.number_of_energy_points, 1))
        t = np.zeros(max(self.number_of_steps_correlation, 1))
        o = np.zeros(max(self.number_of_local_orbitals, 1), np.int32)
        load().lsqt_host_model_inputs(self.h, e.ctypes.data_as(_dp), t.ctypes.data_as(_dp),
                                      o.ctypes.data_as(_ip))
        return (e[:self.number_of_energy_points], t[:self.number_of_steps_correlation],
                o[:self.number_of_local_orbitals])

    def random_state(self, orbital: int = -1):
        re, im = np.zeros(self.n), np.zeros(self.n)
        if load().lsqt_host_model_random_state(self.h, orbital, re.ctypes.data_as(_dp),
                                               im.ctypes.data_as(_dp)):
            raise InputError("bad orbital")
        return re, im

    def evolve(self, kind, direction, t_scaled, re, im, by_terms=True):
        """The host classes' evolve (kind 0) or evolvex (kind 1) of (re, im)
        on the model's Hamiltonian (built on first use, which releases the
        model's arrays); by_terms sequences it one chebyshev_2(x) per term."""
        re = np.array(re, dtype=np.float64)
        im = np.array(im, dtype=np.float64)
        if load().lsqt_host_model_evolve(self.h, int(kind), int(direction), float(t_scaled),
                                         1 if by_terms else 0, re.ctypes.data_as(_dp),
                                         im.ctypes.data_as(_dp)):
            raise InputError("lsqt_host_model_evolve failed")
        return re, im

    def moments(self, left, right, by_terms=True):
        """find_moments_chebyshev of the host classes into a device Vector;
        by_terms: kernel_polynomial + Vector::inner_product_1/2 per moment."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (*left, *right)]
        M = self.number_of_moments
        mr, mi = np.zeros(M), np.zeros(M)
        if load().lsqt_host_model_moments(self.h, *(x.ctypes.data_as(_dp) for x in a),
                                          1 if by_terms else 0, mr.ctypes.data_as(_dp),
                                          mi.ctypes.data_as(_dp)):
            raise InputError("lsqt_host_model_moments failed")
        return mr + 1j * mi

    def close(self):
        if self.h:
            load().lsqt_host_model_free(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def correlation_row(mu, energies, energy_max, volume):
    """Jackson damping + Chebyshev summation of one moment set, by the
    drop-in's host code (src/sigma.cu:76-110)."""
    mu = np.asarray(mu)
    re = np.ascontiguousarray(mu.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(mu.imag, dtype=np.float64).copy()
    e = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros(len(e))
    if load().lsqt_host_damped_sum(len(re), re.ctypes.data_as(_dp), im.ctypes.data_as(_dp),
                                   len(e), e.ctypes.data_as(_dp), float(energy_max),
                                   float(volume), out.ctypes.data_as(_dp)):
        raise ValueError("invalid moment set")
    return out


def conductivity(time_step, dos=None, vac=None, msd=None):
    """sigma from the VAC and MSD rows by the host code
    (lsqt_host_conductivity): returns dict of [Nt, Ne] arrays sigma_vac,
    sigma_msd, vac, msd (those that the inputs allow)."""
    ts = np.ascontiguousarray(time_step, dtype=np.float64)
    arrs = {}
    for k, a in (("dos", dos), ("vac", vac), ("msd", msd)):
        arrs[k] = None if a is None else np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
    Ne = next(a.shape[1] for a in arrs.values() if a is not None)
    Nt = len(ts)
    out = {k: np.zeros((Nt, Ne)) for k in ("sigma_vac", "sigma_msd", "vac", "msd")}
    err = C.create_string_buffer(4096)

    def ptr(a):
        return None if a is None else a.ctypes.data_as(_dp)

    def rows(a):
        return 0 if a is None else a.shape[0]

    rc = load().lsqt_host_conductivity(Ne, Nt, ptr(ts), ptr(arrs["dos"]), rows(arrs["dos"]),
                                       ptr(arrs["vac"]), rows(arrs["vac"]), ptr(arrs["msd"]),
                                       rows(arrs["msd"]), ptr(out["sigma_vac"]),
                                       ptr(out["sigma_msd"]), ptr(out["vac"]), ptr(out["msd"]),
                                       err, len(err))
    if rc:
        raise InputError(err.value.decode())
    keep = {}
    if vac is not None:
        keep["sigma_vac"] = out["sigma_vac"]
        if dos is not None:
            keep["vac"] = out["vac"]
    if msd is not None:
        keep["sigma_msd"] = out["sigma_msd"]
        if dos is not None:
            keep["msd"] = out["msd"]
    return keep


def write_general(dirname: str, arrays, n, max_neighbor, x, box, volume, complex_hoppings=True,
                  potential=True):
    """Writes the general-model input files from row-major ELL arrays
    (lsqt_host_write_general)."""
    nn = np.ascontiguousarray(arrays["neighbor_number"], dtype=np.int32)
    nl = np.ascontiguousarray(arrays["neighbor_list"], dtype=np.int32)
    xs = np.ascontiguousarray(x, dtype=np.float64)
    hr = np.ascontiguousarray(arrays["hopping_real"], dtype=np.float64)
    hi = np.ascontiguousarray(arrays["hopping_imag"], dtype=np.float64)
    pot = np.ascontiguousarray(arrays["potential"], dtype=np.float64)
    assert nn.shape == (n,) and nl.shape == (n * max_neighbor,) and xs.shape == (n,)
    err = C.create_string_buffer(4096)
    ip = lambda a: a.ctypes.data_as(_ip)
    dp = lambda a: a.ctypes.data_as(_dp)
    if load().lsqt_host_write_general(dirname.encode(), int(n), int(max_neighbor), ip(nn), ip(nl),
                                      dp(xs), float(box), float(volume),
                                      dp(pot) if potential else None, dp(hr),
                                      dp(hi) if complex_hoppings else None, err, len(err)):
        raise InputError(err.value.decode())


def postprocess(dirname: str):
    """Writes dirname/sigma_vac.out and sigma_msd.out (lsqt_host_postprocess)."""
    err = C.create_string_buffer(4096)
    if load().lsqt_host_postprocess(dirname.encode(), err, len(err)):
        raise InputError(err.value.decode())


def run_cli(driver_file: str, digits: int | None = None, timeout: float | None = None,
            env=None, sigma: bool = False):
    """Runs the lsqt_gpu program on a driver file (sigma: the --sigma
    post-processing mode)."""
    e = dict(os.environ if env is None else env)
    if digits is not None:
        e["LSQT_OUTPUT_DIGITS"] = str(digits)
    argv = [CLI_PATH, "--sigma", driver_file] if sigma else [CLI_PATH, driver_file]
    return subprocess.run(argv, capture_output=True, text=True, timeout=timeout, env=e)
