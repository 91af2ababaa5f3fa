"""ctypes binding of liblsqt_hip.so (the C ABI of include/lsqt_hip.h).

This is the Python face of the product path: every call goes through the
HIP library.  There is no fallback -- if the library is missing or the GPU
cannot be used, calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsqt_hip.so")

LAYOUT_AUTO, LAYOUT_FULL, LAYOUT_PACKED = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class HamInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("max_neighbor", C.c_int), ("layout", C.c_int),
                ("classes", C.c_int), ("index_bits", C.c_int), ("energy_max", C.c_double),
                ("bytes_per_site_f64", C.c_double), ("slab", C.c_int), ("row", C.c_int),
                ("pattern_period", C.c_int), ("pattern_irregular", C.c_double),
                ("sweep_slab", C.c_int), ("sweep_halo", C.c_int), ("sweep_tiles", C.c_int),
                ("sweep_tiles_f32", C.c_int), ("bytes_per_site_sweep_f64", C.c_double),
                ("bytes_per_site_sweep_f32", C.c_double), ("block_sweep_systems", C.c_int),
                ("sweep_steps", C.c_int), ("sweep_steps_f32", C.c_int),
                ("sweep_tiles_deep", C.c_int), ("block_sweep_steps", C.c_int)]


class Lattice(C.Structure):
    """lsqt_lattice: a lattice model for construction on the device"""
    _fields_ = [("cells", C.c_int * 3), ("pbc", C.c_int * 3), ("transport", C.c_int),
                ("lattice_constant", C.c_double * 3), ("orbitals", C.c_int),
                ("orbital_x", _dp), ("max_neighbor", C.c_int), ("hop_count", _ip),
                ("hop_cell", _ip), ("hop_orbital", _ip), ("hop_re", _dp), ("hop_im", _dp)]


# every symbol declared in include/lsqt_hip.h, with (restype, argtypes)
SIGNATURES = {
    "lsqt_strerror": (C.c_char_p, [C.c_int]),
    "lsqt_last_error": (C.c_char_p, []),
    "lsqt_version": (C.c_char_p, []),
    "lsqt_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "lsqt_ctx_destroy": (C.c_int, [_vp]),
    "lsqt_ctx_sync": (C.c_int, [_vp]),
    "lsqt_ctx_stream": (_vp, [_vp]),
    "lsqt_device_count": (C.c_int, [_ip]),
    "lsqt_ham_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp,
                                  C.c_double, C.c_int, C.POINTER(_vp)]),
    "lsqt_ham_destroy": (C.c_int, [_vp]),
    "lsqt_ham_create_lattice": (C.c_int, [_vp, C.POINTER(Lattice), _dp, C.c_double,
                                          C.POINTER(_vp)]),
    "lsqt_ham_create_lattices": (C.c_int, [_vp, C.c_int, C.POINTER(Lattice),
                                           C.POINTER(_ip), _dp, C.c_double, C.POINTER(_vp)]),
    "lsqt_ham_create_segmented": (C.c_int, [_vp, C.c_int, _ip, C.c_int, C.c_int, _ip, _ip, _dp,
                                            _dp, _dp, _dp, C.c_double, C.POINTER(_vp)]),
    "lsqt_block_set_graphs": (C.c_int, [_vp, C.c_int]),
    "lsqt_block_apply": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "lsqt_block_vac": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]),
    "lsqt_block_msd": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]),
    "lsqt_block_spin": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]),
    "lsqt_ham_get_info": (C.c_int, [_vp, C.POINTER(HamInfo)]),
    "lsqt_vec_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "lsqt_vec_clone": (C.c_int, [_vp, C.POINTER(_vp)]),
    "lsqt_vec_destroy": (C.c_int, [_vp]),
    "lsqt_vec_size": (C.c_int, [_vp, _ip]),
    "lsqt_vec_from_host": (C.c_int, [_vp, _dp, _dp]),
    "lsqt_vec_to_host": (C.c_int, [_vp, _dp, _dp]),
    "lsqt_vec_copy": (C.c_int, [_vp, _vp]),
    "lsqt_vec_add": (C.c_int, [_vp, _vp]),
    "lsqt_vec_apply_sz": (C.c_int, [_vp, _vp]),
    "lsqt_vec_swap": (C.c_int, [_vp, _vp]),
    "lsqt_vec_dot": (C.c_int, [_vp, _vp, _dp, _dp]),
    "lsqt_vec_inner_product_1": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "lsqt_vec_inner_product_2": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "lsqt_vec_device_data": (C.c_int, [_vp, C.POINTER(_vp)]),
    "lsqt_moments_to_vec": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "lsqt_moments_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_apply": (C.c_int, [_vp, _vp, _vp]),
    "lsqt_apply_commutator": (C.c_int, [_vp, _vp, _vp]),
    "lsqt_apply_current": (C.c_int, [_vp, _vp, _vp]),
    "lsqt_kernel_polynomial": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lsqt_kernel_polynomial_pair": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lsqt_chebyshev_01": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int]),
    "lsqt_chebyshev_2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int]),
    "lsqt_chebyshev_1x": (C.c_int, [_vp, _vp, _vp, C.c_double]),
    "lsqt_chebyshev_2x": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                    C.c_int]),
    "lsqt_moments": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_evolve": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "lsqt_evolvex": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "lsqt_bessel_terms": (C.c_int, [C.c_double, _ip]),
    "lsqt_moments_f32": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_evolve_f32": (C.c_int, [_vp, C.c_int, C.c_double, _vp, C.c_int]),
    "lsqt_evolvex_f32": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "lsqt_block_create": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lsqt_block_destroy": (C.c_int, [_vp]),
    "lsqt_block_from_host": (C.c_int, [_vp, _dp, _dp]),
    "lsqt_block_to_host": (C.c_int, [_vp, _dp, _dp]),
    "lsqt_moments_block": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_moments_block_self": (C.c_int, [_vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_moments_block_pair": (C.c_int, [_vp, _vp, _vp, C.c_int, _dp, _dp]),
    "lsqt_block_tune": (C.c_int, [_vp, _vp, _vp, C.c_int, _ip, _ip, _ip, _ip, _dp]),
    "lsqt_block_set_config": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
}

_lib = None


class LsqtError(RuntimeError):
    pass


def load():
    """Loads liblsqt_hip.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LsqtError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        lib = load()
        raise LsqtError(f"{lib.lsqt_strerror(rc).decode()}: {lib.lsqt_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


# Live handles, closed at interpreter exit in dependency order (vectors and
# blocks, then Hamiltonians, then contexts) while this module is intact, so
# that no handle outlives the context whose stream and device it uses.
_live = {"data": weakref.WeakSet(), "ham": weakref.WeakSet(), "ctx": weakref.WeakSet()}


def close_all():
    for kind in ("data", "ham", "ctx"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(close_all)


def device_count() -> int:
    n = C.c_int(0)
    rc = load().lsqt_device_count(C.byref(n))
    return n.value if rc == 0 else 0


class Context:
    def __init__(self, device: int = 0):
        self.h = _vp()
        _check(load().lsqt_ctx_create(device, C.byref(self.h)))
        _live["ctx"].add(self)

    @property
    def stream(self):
        return load().lsqt_ctx_stream(self.h)

    def sync(self):
        _check(load().lsqt_ctx_sync(self.h))

    def close(self):
        if self.h:
            load().lsqt_ctx_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Vector:
    """Vector (src/vector.h:23-47) on the device."""

    def __init__(self, ctx: Context, n: int, _handle=None):
        self.ctx = ctx
        self.n = n
        if _handle is not None:
            self.h = _handle
        else:
            self.h = _vp()
            _check(load().lsqt_vec_create(ctx.h, n, C.byref(self.h)))
        _live["data"].add(self)

    @classmethod
    def from_host(cls, ctx, re, im):
        v = cls(ctx, len(re))
        v.upload(re, im)
        return v

    def upload(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        assert re.shape == (self.n,) and im.shape == (self.n,)
        _check(load().lsqt_vec_from_host(self.h, _dptr(re), _dptr(im)))

    def download(self):
        re, im = np.empty(self.n), np.empty(self.n)
        _check(load().lsqt_vec_to_host(self.h, _dptr(re), _dptr(im)))
        return re, im

    def clone(self):
        h = _vp()
        _check(load().lsqt_vec_clone(self.h, C.byref(h)))
        return Vector(self.ctx, self.n, _handle=h)

    def copy_from(self, other):
        _check(load().lsqt_vec_copy(self.h, other.h))

    def add(self, other):
        _check(load().lsqt_vec_add(self.h, other.h))

    def apply_sz(self, other):
        _check(load().lsqt_vec_apply_sz(self.h, other.h))

    def swap(self, other):
        _check(load().lsqt_vec_swap(self.h, other.h))

    def dot(self, psi):
        re, im = C.c_double(), C.c_double()
        _check(load().lsqt_vec_dot(self.h, psi.h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def inner_product_1(self, number_of_atoms, other, target, offset):
        """target[offset + b] = sum over 512-site block b of self conj(other)"""
        _check(load().lsqt_vec_inner_product_1(self.h, int(number_of_atoms), other.h, target.h,
                                               int(offset)))

    def inner_product_2(self, number_of_atoms, number_of_moments, target):
        """target[m] = sum of the block partials of moment m"""
        _check(load().lsqt_vec_inner_product_2(self.h, int(number_of_atoms),
                                               int(number_of_moments), target.h))

    def device_data(self):
        p = _vp()
        _check(load().lsqt_vec_device_data(self.h, C.byref(p)))
        return p.value

    def close(self):
        if self.h:
            load().lsqt_vec_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Block:
    """R complex vectors of n sites stored together on the device (FP32)."""

    def __init__(self, ctx: Context, n: int, R: int):
        self.ctx, self.n, self.R = ctx, n, R
        self.h = _vp()
        _check(load().lsqt_block_create(ctx.h, n, R, C.byref(self.h)))
        _live["data"].add(self)

    def upload(self, re, im):
        """re, im: arrays of shape (R, n)."""
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        assert re.shape == (self.R, self.n) and im.shape == (self.R, self.n)
        _check(load().lsqt_block_from_host(self.h, _dptr(re), _dptr(im)))

    def download(self):
        re, im = np.empty((self.R, self.n)), np.empty((self.R, self.n))
        _check(load().lsqt_block_to_host(self.h, _dptr(re), _dptr(im)))
        return re, im

    def close(self):
        if self.h:
            load().lsqt_block_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Hamiltonian:
    """Hamiltonian (src/hamiltonian.h:25-54) on the device, from row-major ELL arrays."""

    def __init__(self, ctx: Context, n, max_neighbor, neighbor_number, neighbor_list, potential,
                 hopping_real, hopping_imag, xx, energy_max, layout=LAYOUT_AUTO,
                 segment_starts=None):
        self.ctx = ctx
        self.n = int(n)
        nn = np.ascontiguousarray(neighbor_number, dtype=np.int32)
        nl = np.ascontiguousarray(neighbor_list, dtype=np.int32)
        arrs = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (potential, hopping_real, hopping_imag, xx)]
        assert nn.shape == (n,) and nl.shape == (n * max_neighbor,)
        self.h = _vp()
        self.segments = 1
        if segment_starts is None:
            _check(load().lsqt_ham_create(ctx.h, int(n), int(max_neighbor), _iptr(nn),
                                          _iptr(nl), *(_dptr(a) for a in arrs),
                                          float(energy_max), int(layout), C.byref(self.h)))
        else:
            st = np.ascontiguousarray(segment_starts, dtype=np.int32)
            self.segments = len(st) - 1
            _check(load().lsqt_ham_create_segmented(ctx.h, self.segments, _iptr(st), int(n),
                                                    int(max_neighbor), _iptr(nn), _iptr(nl),
                                                    *(_dptr(a) for a in arrs),
                                                    float(energy_max), C.byref(self.h)))
        _live["ham"].add(self)

    @staticmethod
    def _lattice_struct(lattice: dict, keep: list):
        """lsqt_lattice for a dict as returned by gpuqt_amd.inputs.read_lattice;
        the arrays it points to are appended to keep."""
        No, K = lattice["orbitals"], lattice["max_neighbor"]
        arr = {
            "x": np.ascontiguousarray(lattice["orbital_x"], dtype=np.float64),
            "count": np.ascontiguousarray(lattice["hop_count"], dtype=np.int32),
            "cell": np.ascontiguousarray(lattice["hop_cell"], dtype=np.int32).reshape(-1),
            "orb": np.ascontiguousarray(lattice["hop_orbital"], dtype=np.int32).reshape(-1),
            "re": np.ascontiguousarray(lattice["hop_re"], dtype=np.float64).reshape(-1),
            "im": np.ascontiguousarray(lattice["hop_im"], dtype=np.float64).reshape(-1),
        }
        assert arr["count"].shape == (No,) and arr["orb"].shape == (No * K,)
        keep.append(arr)
        L = Lattice()
        L.cells[:] = [int(c) for c in lattice["cells"]]
        L.pbc[:] = [int(c) for c in lattice["pbc"]]
        L.transport = int(lattice["transport"])
        L.lattice_constant[:] = [float(a) for a in lattice["lattice_constant"]]
        L.orbitals, L.max_neighbor = int(No), int(K)
        L.orbital_x, L.hop_count = _dptr(arr["x"]), _iptr(arr["count"])
        L.hop_cell, L.hop_orbital = _iptr(arr["cell"]), _iptr(arr["orb"])
        L.hop_re, L.hop_im = _dptr(arr["re"]), _dptr(arr["im"])
        return L

    @classmethod
    def from_lattice(cls, ctx, lattice: dict, potential, energy_max):
        """Builds a pristine lattice model on the device (lsqt_ham_create_lattice);
        `lattice` as returned by gpuqt_amd.inputs.read_lattice."""
        keep = []
        L = cls._lattice_struct(lattice, keep)
        pot = np.ascontiguousarray(potential, dtype=np.float64)
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = int(np.prod(lattice["cells"]) * lattice["orbitals"])
        assert pot.shape == (self.n,)
        self.segments = 1
        self.h = _vp()
        _check(load().lsqt_ham_create_lattice(ctx.h, C.byref(L), _dptr(pot), float(energy_max),
                                              C.byref(self.h)))
        _live["ham"].add(self)
        return self

    @classmethod
    def from_lattices(cls, ctx, lattices, site_indices, potential, energy_max):
        """Lattice models (with vacancies: site_indices[k] maps pristine sites
        to their index, -1 for vacancies; None for a pristine lattice) built
        on the device, side by side as a segmented Hamiltonian when there are
        several (lsqt_ham_create_lattices)."""
        keep = []
        S = len(lattices)
        arr = (Lattice * S)(*[cls._lattice_struct(l, keep) for l in lattices])
        idx = [None if si is None else np.ascontiguousarray(si, dtype=np.int32)
               for si in (site_indices or [None] * S)]
        ptrs = (_ip * S)(*[(_iptr(i) if i is not None else _ip()) for i in idx])
        n = 0
        for l, i in zip(lattices, idx):
            n += int(np.prod(l["cells"]) * l["orbitals"]) if i is None else int((i >= 0).sum())
        pot = np.ascontiguousarray(potential, dtype=np.float64)
        assert pot.shape == (n,)
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = n
        self.segments = S
        self.h = _vp()
        _check(load().lsqt_ham_create_lattices(ctx.h, S, arr, ptrs, _dptr(pot),
                                               float(energy_max), C.byref(self.h)))
        _live["ham"].add(self)
        return self

    @classmethod
    def from_arrays(cls, ctx, arrays, n, max_neighbor, energy_max, layout=LAYOUT_AUTO,
                    segment_starts=None):
        return cls(ctx, n, max_neighbor, arrays["neighbor_number"], arrays["neighbor_list"],
                   arrays["potential"], arrays["hopping_real"], arrays["hopping_imag"],
                   arrays["xx"], energy_max, layout, segment_starts)

    def info(self) -> HamInfo:
        inf = HamInfo()
        _check(load().lsqt_ham_get_info(self.h, C.byref(inf)))
        return inf

    def vector(self, re=None, im=None):
        v = Vector(self.ctx, self.n)
        if re is not None:
            v.upload(re, im)
        return v

    def apply(self, vin, vout):
        _check(load().lsqt_apply(self.h, vin.h, vout.h))

    def apply_commutator(self, vin, vout):
        _check(load().lsqt_apply_commutator(self.h, vin.h, vout.h))

    def apply_current(self, vin, vout):
        _check(load().lsqt_apply_current(self.h, vin.h, vout.h))

    def kernel_polynomial(self, s0, s1, s2):
        _check(load().lsqt_kernel_polynomial(self.h, s0.h, s1.h, s2.h))

    def kernel_polynomial_pair(self, s0, s1, s2, s3):
        _check(load().lsqt_kernel_polynomial_pair(self.h, s0.h, s1.h, s2.h, s3.h))

    def chebyshev_01(self, s0, s1, s, b0, b1, direction):
        _check(load().lsqt_chebyshev_01(self.h, s0.h, s1.h, s.h, b0, b1, direction))

    def chebyshev_2(self, s0, s1, s2, s, bm, label):
        _check(load().lsqt_chebyshev_2(self.h, s0.h, s1.h, s2.h, s.h, bm, label))

    def chebyshev_1x(self, s1x, s, b1):
        _check(load().lsqt_chebyshev_1x(self.h, s1x.h, s.h, b1))

    def chebyshev_2x(self, s0, s0x, s1, s1x, s2, s2x, s, bm, label):
        _check(load().lsqt_chebyshev_2x(self.h, s0.h, s0x.h, s1.h, s1x.h, s2.h, s2x.h, s.h, bm,
                                        label))

    def moments(self, left, right, M):
        re, im = np.empty(M), np.empty(M)
        _check(load().lsqt_moments(self.h, left.h, right.h, int(M), _dptr(re), _dptr(im)))
        return re + 1j * im

    def moments_noalloc(self, left, right, M):
        """lsqt_moments_noalloc: the moments of lsqt_moments from the fixed
        per-GPU scratch (no device allocation)"""
        re, im = np.empty(M), np.empty(M)
        _check(load().lsqt_moments_noalloc(self.h, left.h, right.h, int(M), _dptr(re),
                                           _dptr(im)))
        return re + 1j * im

    def moments_to_vec(self, left, right, M, out):
        """find_moments_chebyshev into a device vector: out[m] = mu_m"""
        _check(load().lsqt_moments_to_vec(self.h, left.h, right.h, int(M), out.h))

    def evolve(self, direction, t_scaled, state):
        _check(load().lsqt_evolve(self.h, int(direction), float(t_scaled), state.h))

    def evolvex(self, direction, t_scaled, state):
        _check(load().lsqt_evolvex(self.h, int(direction), float(t_scaled), state.h))

    def moments_f32(self, left, right, M):
        re, im = np.empty(M), np.empty(M)
        _check(load().lsqt_moments_f32(self.h, left.h, right.h, int(M), _dptr(re), _dptr(im)))
        return re + 1j * im

    def evolve_f32(self, direction, t_scaled, state, renormalise=True):
        _check(load().lsqt_evolve_f32(self.h, int(direction), float(t_scaled), state.h,
                                      1 if renormalise else 0))

    def evolvex_f32(self, direction, t_scaled, state):
        _check(load().lsqt_evolvex_f32(self.h, int(direction), float(t_scaled), state.h))

    def moments_block(self, left, right, M):
        """complex array (R, M) of <left_r|T_m|right_r> ((segments, R, M) on a
        segmented Hamiltonian)"""
        R = left.R
        S = self.segments
        re, im = np.empty(S * R * M), np.empty(S * R * M)
        _check(load().lsqt_moments_block(self.h, left.h, right.h, int(M), _dptr(re), _dptr(im)))
        mu = (re + 1j * im).reshape(S, R, M)
        return mu[0] if S == 1 else mu

    def moments_block_self(self, v, M):
        """complex array (R, M) of <v_r|T_m|v_r> by the product identities
        (M/2 steps)"""
        R = v.R
        S = self.segments
        re, im = np.empty(S * R * M), np.empty(S * R * M)
        _check(load().lsqt_moments_block_self(self.h, v.h, int(M), _dptr(re), _dptr(im)))
        mu = (re + 1j * im).reshape(S, R, M)
        return mu[0] if S == 1 else mu

    def moments_block_pair(self, left, right, M):
        """complex array (R, M) of <left_r|T_m|right_r> from the left and
        right recursions run side by side (product identities, M/2 steps)"""
        R = left.R
        S = self.segments
        re, im = np.empty(S * R * M), np.empty(S * R * M)
        _check(load().lsqt_moments_block_pair(self.h, left.h, right.h, int(M), _dptr(re),
                                              _dptr(im)))
        mu = (re + 1j * im).reshape(S, R, M)
        return mu[0] if S == 1 else mu

    def _block_driver(self, fn, re, im, M, time_steps_scaled):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        R = re.shape[0]
        assert re.shape == (R, self.n) and im.shape == (R, self.n)
        ts = np.ascontiguousarray(time_steps_scaled, dtype=np.float64)
        Nt = len(ts)
        size = Nt * self.segments * R * M
        mr, mi = np.empty(size), np.empty(size)
        _check(fn(self.h, R, _dptr(re), _dptr(im), int(M), Nt, _dptr(ts), _dptr(mr), _dptr(mi)))
        return (mr + 1j * mi).reshape(Nt, self.segments, R, M)

    def block_apply(self, op: int, bin, bout):
        """bout = op(bin): 0 H~, 1 [X, H~], 2 V"""
        _check(load().lsqt_block_apply(self.h, int(op), bin.h, bout.h))

    def block_set_graphs(self, enable: bool):
        _check(load().lsqt_block_set_graphs(self.h, 1 if enable else 0))

    def block_vac(self, re, im, M, time_steps_scaled):
        """complex (Nt, segments, R, M): find_vac's moment sets for R vectors"""
        return self._block_driver(load().lsqt_block_vac, re, im, M, time_steps_scaled)

    def block_msd(self, re, im, M, time_steps_scaled):
        """complex (Nt, segments, R, M): find_msd's moment sets for R vectors"""
        return self._block_driver(load().lsqt_block_msd, re, im, M, time_steps_scaled)

    def block_spin(self, re, im, M, time_steps_scaled):
        """complex (Nt, 1, R, M): find_spin_polarization's moment sets for R vectors"""
        return self._block_driver(load().lsqt_block_spin, re, im, M, time_steps_scaled)

    def block_tune(self, left, right, force=False):
        """(W, rows, blocked, us per step) of the autotuned block kernel shape
        (the depth is in block_depth after the call)"""
        W, rows, blocked, depth, us = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(load().lsqt_block_tune(self.h, left.h, right.h, 1 if force else 0, C.byref(W),
                                      C.byref(rows), C.byref(blocked), C.byref(depth),
                                      C.byref(us)))
        self.block_depth = depth.value
        return W.value, rows.value, blocked.value, us.value

    def block_set_config(self, R, W=0, rows=0, blocked=0, depth=1):
        """blocked: 0 natural tiles, 1 (or True) blocked, 2 blocked with the
        column walk, 3 the sliding window (rows = lattice rows per tile); depth: passes in flight per wave (1 or 2)"""
        _check(load().lsqt_block_set_config(self.h, int(R), int(W), int(rows), int(blocked),
                                            int(depth)))

    def close(self):
        if self.h:
            load().lsqt_ham_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bessel_last_term(t_scaled: float) -> int:
    m = C.c_int()
    _check(load().lsqt_bessel_terms(float(t_scaled), C.byref(m)))
    return m.value
