"""Python binding of libmcstar_hip.so (the MI355X PFC time step) over its C ABI.

The product is the C ABI in include/mcstar_hip.h and the C++ host headers
that keep the reference's API (include/cuda/*.hpp).  This module binds the
same C ABI with ctypes for tests and benchmarks.  It has no fallback: if the
shared library is missing, loading fails loudly.

Arrays on the device are DeviceBuffer objects (mcs_malloc'd); numpy arrays
move in and out with upload()/download().  Shapes follow numpy order
(nz, ny, nx); the C ABI takes ext = {nx, ny, nz}.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmcstar_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "mcstar_hip.h")

OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_NOMEM, ERR_CAPACITY = 0, 1, 2, 3, 4, 5
H2D, D2H, D2D, H2H = 1, 2, 3, 4
R2C, C2R = 0, 1
SPLIT_CUBIC, SPLIT_REFERENCE = 0, 1
FORM_SPECTRAL, FORM_MIXED = 0, 1

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class McsError(RuntimeError):
    def __init__(self, code, where, detail):
        super().__init__(f"{where}: {detail} (code {code})")
        self.code = code


class PfcParams(ctypes.Structure):
    _fields_ = [("ext", ctypes.c_int * 3), ("h", ctypes.c_double * 3), ("eps", ctypes.c_double),
                ("dt", ctypes.c_double), ("temp", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("split", ctypes.c_int), ("prec", ctypes.c_int)]


class ReplicaParams(ctypes.Structure):
    _fields_ = [("ext", ctypes.c_int * 3), ("h", ctypes.c_double * 3), ("eps", ctypes.c_double),
                ("dt", ctypes.c_double), ("temp", ctypes.c_double), ("psi0", ctypes.c_double),
                ("amp", ctypes.c_double), ("init_seed", ctypes.c_uint64), ("noise_seed", ctypes.c_uint64)]


ENS_LOOP, ENS_BATCHED, ENS_GRAPH = 0, 1, 2

_SIGS = {
    "mcs_error_string": ([ctypes.c_int], ctypes.c_char_p),
    "mcs_last_error": ([], ctypes.c_char_p),
    "mcs_init": ([ctypes.c_int], ctypes.c_int),
    "mcs_sync": ([], ctypes.c_int),
    "mcs_malloc": ([ctypes.POINTER(_vp), ctypes.c_size_t], ctypes.c_int),
    "mcs_free": ([_vp], ctypes.c_int),
    "mcs_host_alloc": ([ctypes.POINTER(_vp), ctypes.c_size_t], ctypes.c_int),
    "mcs_host_free": ([_vp], ctypes.c_int),
    "mcs_memcpy": ([_vp, _vp, ctypes.c_size_t, ctypes.c_int], ctypes.c_int),
    "mcs_memset": ([_vp, ctypes.c_int, ctypes.c_size_t], ctypes.c_int),
    "mcs_fft_plan_create": ([ctypes.POINTER(_vp), _ip, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_fft_exec": ([_vp, _vp, _vp], ctypes.c_int),
    "mcs_fft_plan_destroy": ([_vp], ctypes.c_int),
    "mcs_rng_create": ([ctypes.POINTER(_vp), ctypes.c_uint64], ctypes.c_int),
    "mcs_rng_destroy": ([_vp], ctypes.c_int),
    "mcs_rng_set_draw": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "mcs_rng_get_draw": ([_vp, _u32p], ctypes.c_int),
    "mcs_rng_fill_normal": ([_vp, _vp, ctypes.c_size_t, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "mcs_rng_fill_kspace": ([_vp, _vp, _ip, ctypes.c_int], ctypes.c_int),
    "mcs_philox_blocks": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t,
                           _vp], ctypes.c_int),
    "mcs_psi3_split": ([_vp, _vp, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int], ctypes.c_int),
    "mcs_norm": ([_vp, _ip, ctypes.c_int], ctypes.c_int),
    "mcs_init_eom_op": ([_vp, _ip, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int], ctypes.c_int),
    "mcs_init_el_op": ([_vp, _ip, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int], ctypes.c_int),
    "mcs_update_eom_etd": ([_vp, _vp, _vp, _vp, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, _dp,
                            ctypes.c_int], ctypes.c_int),
    "mcs_update_el_etd": ([_vp, _vp, _vp, _ip, ctypes.c_int, _dp], ctypes.c_int),
    "mcs_update_eom_hpfc": ([_vp, _vp, _vp, _vp, _vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                             ctypes.c_int], ctypes.c_int),
    "mcs_update_eom_poly": ([_vp, _vp, _vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                             ctypes.c_int], ctypes.c_int),
    "mcs_update_eom_with_fen": ([_vp, _vp, _vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                 ctypes.c_int], ctypes.c_int),
    "mcs_calc_fen": ([_vp, _vp, _ip, _dp, ctypes.c_int, _dp], ctypes.c_int),
    "mcs_update_el": ([_vp, _vp, _ip, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, _dp], ctypes.c_int),
    "mcs_update_eom": ([_vp, _vp, _vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                        ctypes.c_int], ctypes.c_int),
    "mcs_update_pfc": ([_vp, _vp, _ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                        ctypes.c_uint32, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_init_field": ([_vp, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int], ctypes.c_int),
    "mcs_field_stats": ([_vp, _ip, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "mcs_stepper_create": ([ctypes.POINTER(_vp), ctypes.POINTER(PfcParams)], ctypes.c_int),
    "mcs_stepper_set_psi": ([_vp, _vp], ctypes.c_int),
    "mcs_stepper_get_psi": ([_vp, _vp], ctypes.c_int),
    "mcs_stepper_step": ([_vp, ctypes.c_int], ctypes.c_int),
    "mcs_stepper_step_async": ([_vp, ctypes.c_int], ctypes.c_int),
    "mcs_stepper_get_step": ([_vp, _i64p], ctypes.c_int),
    "mcs_stepper_set_step": ([_vp, ctypes.c_int64], ctypes.c_int),
    "mcs_stepper_mass": ([_vp, _dp], ctypes.c_int),
    "mcs_stepper_destroy": ([_vp], ctypes.c_int),
    "mcs_stepper_pass_times": ([_vp, ctypes.c_int, _dp, ctypes.c_int, _ip], ctypes.c_int),
    "mcs_stepper_plan_info": ([_vp, _ip], ctypes.c_int),
    "mcs_set_step_form": ([ctypes.c_int], ctypes.c_int),
    "mcs_set_lane_shuffle": ([ctypes.c_int], ctypes.c_int),
    "mcs_get_lane_shuffle": ([_ip], ctypes.c_int),
    "mcs_get_step_form": ([_ip], ctypes.c_int),
    "mcs_stepper_set_form": ([_vp, ctypes.c_int], ctypes.c_int),
    "mcs_stepper_get_form": ([_vp, _ip], ctypes.c_int),
    "mcs_stepper_get_state": ([_vp, ctypes.c_void_p, _dp, _i64p], ctypes.c_int),
    "mcs_stepper_set_state": ([_vp, ctypes.c_void_p, ctypes.c_double, ctypes.c_int64], ctypes.c_int),
    "mcs_set_autotune": ([ctypes.c_int], ctypes.c_int),
    "mcs_get_autotune": ([_ip], ctypes.c_int),
    "mcs_set_fixed_kernels": ([ctypes.c_int], ctypes.c_int),
    "mcs_get_fixed_kernels": ([_ip], ctypes.c_int),
    "mcs_ensemble_run": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _i64p, _i64p, _i64p],
                         ctypes.c_int),
    "mcs_stepper_create_planned": ([ctypes.POINTER(_vp), ctypes.c_void_p, _ip], ctypes.c_int),
    "mcs_plan_options": ([ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _ip], ctypes.c_int),
    "mcs_tile_options": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _ip],
                         ctypes.c_int),
    "mcs_timer_start": ([], ctypes.c_int),
    "mcs_timer_stop": ([_dp], ctypes.c_int),
    "mcs_ensemble_create": ([ctypes.POINTER(_vp), ctypes.POINTER(ReplicaParams), ctypes.c_int, ctypes.c_int,
                             ctypes.c_int], ctypes.c_int),
    "mcs_ensemble_init": ([_vp], ctypes.c_int),
    "mcs_ensemble_set_psi": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
    "mcs_ensemble_get_psi": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
    "mcs_ensemble_step": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_ensemble_step_async": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_ensemble_analyze": ([_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _i64p, _i64p, _i64p], ctypes.c_int),
    "mcs_ensemble_destroy": ([_vp], ctypes.c_int),
    "mcs_ensemble_launch_info": ([_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "mcs_ensemble_set_launch_form": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_ensemble_graph_branches": ([_vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "mcs_ensemble_set_graph_branches": ([_vp, ctypes.c_int], ctypes.c_int),
    "mcs_findpeaks": ([_vp, _ip, ctypes.c_double, _i64p, _dp, ctypes.c_int64, _i64p, ctypes.c_int], ctypes.c_int),
    "mcs_bond_order_cutoff": ([_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip, _dp], ctypes.c_int),
    "mcs_bond_order_knn": ([_ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _dp, _i64p,
                            _i64p, ctypes.c_int], ctypes.c_int),
    "mcs_gaussian_init": ([_vp, _ip, _dp, _dp, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "mcs_kmult": ([_vp, _vp, _ip, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mcs_square": ([_vp, _ip, ctypes.c_int], ctypes.c_int),
    "mcs_pair_distances": ([_dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int64, _i64p], ctypes.c_int),
    "mcs_pair_histogram": ([_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _i64p],
                           ctypes.c_int),
}

_LIB = None


def lib():
    """Load libmcstar_hip.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make lib` (no fallback path exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def declared_symbols():
    """Function names declared in include/mcstar_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(mcs_\w+)\s*\(", txt, re.M)))


def check(code, where=""):
    if code != OK:
        L = lib()
        raise McsError(code, where or L.mcs_error_string(code).decode(), L.mcs_last_error().decode())
    return code


def _ext(shape_zyx):
    s = list(shape_zyx)[::-1] + [1] * (3 - len(shape_zyx))
    return (ctypes.c_int * 3)(*s)


def _prec(dtype):
    return 64 if np.dtype(dtype) in (np.float64, np.complex128) else 32


def half_shape(shape_zyx):
    s = list(shape_zyx)
    s[-1] = s[-1] // 2 + 1
    return tuple(s)


def buffer_bytes(shape_zyx, dtype):
    """Bytes of a GPUArray for a real shape: (nx/2+1)*ny*nz complex values."""
    return int(np.prod(half_shape(shape_zyx))) * 2 * np.dtype(dtype).itemsize


class DeviceBuffer:
    """Owning device allocation (mcs_malloc / mcs_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().mcs_malloc(ctypes.byref(p), max(self.nbytes, 1)), "mcs_malloc")
        self.ptr = p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes, "upload larger than buffer"
        check(lib().mcs_memcpy(self.ptr, a.ctypes.data_as(_vp), a.nbytes, H2D), "mcs_memcpy H2D")
        return self

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes, "download larger than buffer"
        check(lib().mcs_memcpy(out.ctypes.data_as(_vp), self.ptr, out.nbytes, D2H), "mcs_memcpy D2H")
        return out

    def zero(self):
        check(lib().mcs_memset(self.ptr, 0, self.nbytes), "mcs_memset")
        return self

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().mcs_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(arr, nbytes=None):
    a = np.ascontiguousarray(arr)
    return DeviceBuffer(nbytes or a.nbytes).upload(a)


class FFTPlan:
    """FFT::CufftPlan<F, R2C|C2R> over the C ABI."""

    def __init__(self, shape_zyx, dtype, direction):
        self.shape = tuple(shape_zyx)
        self.prec = _prec(dtype)
        p = _vp()
        check(lib().mcs_fft_plan_create(ctypes.byref(p), _ext(shape_zyx), self.prec, direction),
              "mcs_fft_plan_create")
        self.ptr = p

    def execute(self, din, dout):
        check(lib().mcs_fft_exec(self.ptr, din.ptr, dout.ptr), "mcs_fft_exec")

    def __del__(self):
        try:
            lib().mcs_fft_plan_destroy(self.ptr)
        except Exception:
            pass


def rfftn(x):
    """Unnormalised R2C of a real numpy array on the GPU (reference layout)."""
    x = np.ascontiguousarray(x)
    ct = np.complex128 if x.dtype == np.float64 else np.complex64
    nb = buffer_bytes(x.shape, x.dtype)
    din = DeviceBuffer(nb).upload(x)
    dout = DeviceBuffer(nb)
    FFTPlan(x.shape, x.dtype, R2C).execute(din, dout)
    return dout.download(half_shape(x.shape), ct)


def irfftn(k, shape_zyx):
    """Unnormalised C2R of a half spectrum on the GPU."""
    k = np.ascontiguousarray(k)
    rt = np.float64 if k.dtype == np.complex128 else np.float32
    nb = buffer_bytes(shape_zyx, rt)
    din = DeviceBuffer(nb).upload(k)
    dout = DeviceBuffer(nb)
    FFTPlan(shape_zyx, rt, C2R).execute(din, dout)
    return dout.download(tuple(shape_zyx), rt)


def update_eom(kpsi, knl, knoise, shape_zyx, namp, dt, eps, lens):
    prec = _prec(kpsi.dtype)
    nb = buffer_bytes(shape_zyx, np.float64 if prec == 64 else np.float32)
    dk, dn = to_device(kpsi, nb), to_device(knl, nb)
    dz = to_device(knoise, nb) if knoise is not None else None
    ln = (ctypes.c_double * 3)(*lens)
    check(lib().mcs_update_eom(dk.ptr, dn.ptr, dz.ptr if dz else None, _ext(shape_zyx), namp, dt, eps, ln, prec),
          "mcs_update_eom")
    return dk.download(kpsi.shape, kpsi.dtype)


def update_eom_with_fen(kpsi, knl, knoise, shape_zyx, namp, dt, eps, lens):
    """returns (updated kpsi, lk * old kpsi) (call_kernel_update_eom_with_fen)"""
    prec = _prec(kpsi.dtype)
    nb = buffer_bytes(shape_zyx, np.float64 if prec == 64 else np.float32)
    dk, dn, dz = to_device(kpsi, nb), to_device(knl, nb), to_device(knoise, nb)
    ln = (ctypes.c_double * 3)(*lens)
    check(lib().mcs_update_eom_with_fen(dk.ptr, dn.ptr, dz.ptr, _ext(shape_zyx), namp, dt, eps, ln, prec),
          "mcs_update_eom_with_fen")
    return dk.download(kpsi.shape, kpsi.dtype), dz.download(knoise.shape, knoise.dtype)


def calc_fen(psi, lpsi, h):
    """returns (total free energy, per-cell density) (call_kernel_calc_fen)"""
    psi = np.ascontiguousarray(psi)
    d1, d2 = to_device(psi), to_device(np.ascontiguousarray(lpsi, dtype=psi.dtype))
    f = ctypes.c_double()
    check(lib().mcs_calc_fen(d1.ptr, d2.ptr, _ext(psi.shape), (ctypes.c_double * 3)(*h), _prec(psi.dtype),
                             ctypes.byref(f)), "mcs_calc_fen")
    return f.value, d2.download(psi.shape, psi.dtype)


def update_el(kpsi, knl, shape_zyx, lens, dt, eps):
    """returns (updated kpsi, max change of a mode) (call_kernel_update_el)"""
    prec = _prec(kpsi.dtype)
    nb = buffer_bytes(shape_zyx, np.float64 if prec == 64 else np.float32)
    dk, dn = to_device(kpsi, nb), to_device(knl, nb)
    m = ctypes.c_double()
    check(lib().mcs_update_el(dk.ptr, dn.ptr, _ext(shape_zyx), dt, eps, (ctypes.c_double * 3)(*lens), prec,
                              ctypes.byref(m)), "mcs_update_el")
    return dk.download(kpsi.shape, kpsi.dtype), m.value


def _spec_nb(shape_zyx, ct):
    return buffer_bytes(shape_zyx, np.float64 if np.dtype(ct) == np.complex128 else np.float32)


def etd_op(shape_zyx, lens, dt, eps, el, dtype=np.float32):
    """(E, S) table of exponential time differencing, conserved (el False) or not (cuda/kernels/*-exp.hpp)"""
    ct = np.complex128 if np.dtype(dtype) == np.float64 else np.complex64
    d = DeviceBuffer(_spec_nb(shape_zyx, ct))
    fn = lib().mcs_init_el_op if el else lib().mcs_init_eom_op
    check(fn(d.ptr, _ext(shape_zyx), dt, eps, (ctypes.c_double * 3)(*lens), _prec(dtype)), "mcs_init_*_op")
    return d.download(half_shape(shape_zyx), ct)


def update_eom_etd(kop, kpsi, knl, knoise, shape_zyx, lens, namp, eps=0.0, with_fen=False):
    nb = _spec_nb(shape_zyx, kpsi.dtype)
    do, dk, dn = to_device(kop, nb), to_device(kpsi, nb), to_device(knl, nb)
    dz = to_device(knoise, nb) if knoise is not None else None
    check(lib().mcs_update_eom_etd(do.ptr, dk.ptr, dn.ptr, dz.ptr if dz else None, 1 if with_fen else 0,
                                   _ext(shape_zyx), namp, eps, (ctypes.c_double * 3)(*lens), _prec(kpsi.dtype)),
          "mcs_update_eom_etd")
    out = dk.download(kpsi.shape, kpsi.dtype)
    return (out, dz.download(knoise.shape, knoise.dtype)) if with_fen else out


def update_el_etd(kop, kpsi, knl, shape_zyx):
    nb = _spec_nb(shape_zyx, kpsi.dtype)
    do, dk, dn = to_device(kop, nb), to_device(kpsi, nb), to_device(knl, nb)
    m = ctypes.c_double()
    check(lib().mcs_update_el_etd(do.ptr, dk.ptr, dn.ptr, _ext(shape_zyx), _prec(kpsi.dtype), ctypes.byref(m)),
          "mcs_update_el_etd")
    return dk.download(kpsi.shape, kpsi.dtype), m.value


def update_eom_hpfc(kpsi, knl, x1, x2, x3, shape_zyx, lens, namp, dt, eps):
    nb = _spec_nb(shape_zyx, kpsi.dtype)
    dk, dn = to_device(kpsi, nb), to_device(knl, nb)
    d1, d2, d3 = to_device(x1, nb), to_device(x2, nb), to_device(x3, nb)
    check(lib().mcs_update_eom_hpfc(dk.ptr, dn.ptr, d1.ptr, d2.ptr, d3.ptr, _ext(shape_zyx), namp, dt, eps,
                                    (ctypes.c_double * 3)(*lens), _prec(kpsi.dtype)), "mcs_update_eom_hpfc")
    return dk.download(kpsi.shape, kpsi.dtype)


def update_eom_poly(kpsi, knl, knoise, shape_zyx, lens, namp, dt, eps):
    nb = _spec_nb(shape_zyx, kpsi.dtype)
    dk, dn = to_device(kpsi, nb), to_device(knl, nb)
    dz = to_device(knoise, nb) if knoise is not None else None
    check(lib().mcs_update_eom_poly(dk.ptr, dn.ptr, dz.ptr if dz else None, _ext(shape_zyx), namp, dt, eps,
                                    (ctypes.c_double * 3)(*lens), _prec(kpsi.dtype)), "mcs_update_eom_poly")
    return dk.download(kpsi.shape, kpsi.dtype)


def update_pfc(kpsi, knl, shape_zyx, lens, dt, eps, temp, seed, draw, split):
    prec = _prec(kpsi.dtype)
    nb = buffer_bytes(shape_zyx, np.float64 if prec == 64 else np.float32)
    dk, dn = to_device(kpsi, nb), to_device(knl, nb)
    ln = (ctypes.c_double * 3)(*lens)
    check(lib().mcs_update_pfc(dk.ptr, dn.ptr, _ext(shape_zyx), ln, dt, eps, temp, seed, draw, split, prec),
          "mcs_update_pfc")
    return dk.download(kpsi.shape, kpsi.dtype)


def psi3_split(a, mu=0.0, norm=False, pot=None):
    prec = _prec(a.dtype)
    d = to_device(a)
    dp = to_device(pot) if pot is not None else None
    check(lib().mcs_psi3_split(d.ptr, dp.ptr if dp else None, mu, int(norm), _ext(a.shape), prec), "mcs_psi3_split")
    return d.download(a.shape, a.dtype)


def norm(a):
    d = to_device(a)
    check(lib().mcs_norm(d.ptr, _ext(a.shape), _prec(a.dtype)), "mcs_norm")
    return d.download(a.shape, a.dtype)


def kspace_noise(shape_zyx, seed, draw, dtype=np.float32):
    prec = _prec(dtype)
    ct = np.complex128 if prec == 64 else np.complex64
    d = DeviceBuffer(buffer_bytes(shape_zyx, dtype))
    r = _vp()
    check(lib().mcs_rng_create(ctypes.byref(r), seed), "mcs_rng_create")
    try:
        check(lib().mcs_rng_set_draw(r, draw), "mcs_rng_set_draw")
        check(lib().mcs_rng_fill_kspace(r, d.ptr, _ext(shape_zyx), prec), "mcs_rng_fill_kspace")
    finally:
        lib().mcs_rng_destroy(r)
    return d.download(half_shape(shape_zyx), ct)


def fill_normal(n, sigma, seed, draw, dtype=np.float32):
    d = DeviceBuffer(n * np.dtype(dtype).itemsize)
    r = _vp()
    check(lib().mcs_rng_create(ctypes.byref(r), seed), "mcs_rng_create")
    try:
        check(lib().mcs_rng_set_draw(r, draw), "mcs_rng_set_draw")
        check(lib().mcs_rng_fill_normal(r, d.ptr, n, sigma, _prec(dtype)), "mcs_rng_fill_normal")
    finally:
        lib().mcs_rng_destroy(r)
    return d.download((n,), dtype)


def philox_blocks(seed, c0, c2, c3, n):
    d = DeviceBuffer(16 * n)
    check(lib().mcs_philox_blocks(seed, c0, c2, c3, n, d.ptr), "mcs_philox_blocks")
    return d.download((4 * n,), np.uint32)


def init_field(shape_zyx, psi0, amp, seed, dtype=np.float32):
    d = DeviceBuffer(int(np.prod(shape_zyx)) * np.dtype(dtype).itemsize)
    check(lib().mcs_init_field(d.ptr, _ext(shape_zyx), psi0, amp, seed, _prec(dtype)), "mcs_init_field")
    return d.download(tuple(shape_zyx), dtype)


def field_stats(psi, h, eps):
    d = to_device(psi)
    hh = (ctypes.c_double * 3)(*h)
    mean, fen = ctypes.c_double(), ctypes.c_double()
    check(lib().mcs_field_stats(d.ptr, _ext(psi.shape), hh, eps, _prec(psi.dtype), ctypes.byref(mean),
                                ctypes.byref(fen)), "mcs_field_stats")
    return mean.value, fen.value


def findpeaks(field, ths, cap=None, with_values=False):
    field = np.ascontiguousarray(field)
    d = to_device(field)
    cap = field.size if cap is None else cap
    idx = np.empty(max(cap, 1), dtype=np.int64)
    vals = np.empty(max(cap, 1), dtype=np.float64) if with_values else None
    cnt = ctypes.c_int64()
    rc = lib().mcs_findpeaks(d.ptr, _ext(field.shape), ths, idx.ctypes.data_as(_i64p),
                             vals.ctypes.data_as(_dp) if with_values else None, cap, ctypes.byref(cnt),
                             _prec(field.dtype))
    if rc == ERR_CAPACITY:
        raise McsError(rc, "mcs_findpeaks", f"capacity {cap} < count {cnt.value}")
    check(rc, "mcs_findpeaks")
    n = cnt.value
    return (idx[:n].copy(), vals[:n].copy()) if with_values else idx[:n].copy()


def bond_order_cutoff(xyz, box, rlim):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    P = xyz.shape[0]
    nb = np.zeros(max(P, 1), dtype=np.int32)
    out = np.zeros((max(P, 1), 9), dtype=np.float64)
    check(lib().mcs_bond_order_cutoff(xyz.ctypes.data_as(_dp), P, box, rlim, nb.ctypes.data_as(_ip),
                                      out.ctypes.data_as(_dp)), "mcs_bond_order_cutoff")
    return nb[:P], out[:P]


def bond_order_knn(ijk, n, h, k=12, nbins=64):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    P = ijk.shape[0]
    nbr = np.zeros((max(P, 1), k), dtype=np.int32)
    out = np.zeros((max(P, 1), 4), dtype=np.float64)
    h4 = np.zeros(nbins, dtype=np.int64)
    h6 = np.zeros(nbins, dtype=np.int64)
    check(lib().mcs_bond_order_knn(ijk.ctypes.data_as(_ip), P, n, h, k, nbr.ctypes.data_as(_ip),
                                   out.ctypes.data_as(_dp), h4.ctypes.data_as(_i64p), h6.ctypes.data_as(_i64p),
                                   nbins), "mcs_bond_order_knn")
    return nbr[:P], out[:P], h4, h6



# ---------------------------------------------------------------- post-processing

def gaussian_init(shape_zyx, lens, h, divs_per_sigma, dtype=np.float32):
    """(e, 1 - e) per half-spectrum mode (call_kernel_gaussian_init, cuda/kernels/gaussian.hpp:23-40)"""
    ct = np.complex128 if np.dtype(dtype) == np.float64 else np.complex64
    d = DeviceBuffer(buffer_bytes(shape_zyx, dtype))
    check(lib().mcs_gaussian_init(d.ptr, _ext(shape_zyx), (ctypes.c_double * 3)(*lens), (ctypes.c_double * 3)(*h),
                                  divs_per_sigma, _prec(dtype)), "mcs_gaussian_init")
    return d.download(half_shape(shape_zyx), ct)


def kmult(k1, k2, shape_zyx, part):
    """k1 * Re k2 (part 0) or k1 * Im k2 (part 1) (call_kernel_kmult_real / _imag)"""
    nb = _spec_nb(shape_zyx, k1.dtype)
    d1, d2 = to_device(k1, nb), to_device(k2, nb)
    check(lib().mcs_kmult(d1.ptr, d2.ptr, _ext(shape_zyx), part, _prec(k1.dtype)), "mcs_kmult")
    return d1.download(k1.shape, k1.dtype)


def square(a):
    d = to_device(a)
    check(lib().mcs_square(d.ptr, _ext(a.shape), _prec(a.dtype)), "mcs_square")
    return d.download(a.shape, a.dtype)


def amplitude_lowpass(psi, lens, h, divs_per_sigma):
    """calc_amplitude_lowpass (cuda/kernels/lowpass.hpp:79-93) of a 2D field, all on the GPU:
    returns (low-passed spectrum of the squared high-passed field, its real-space amplitude)"""
    psi = np.ascontiguousarray(psi)
    shp = psi.shape
    if len(shp) != 2 or min(shp) < 2:
        raise McsError(ERR_INVALID, "amplitude_lowpass", "field is not 2d")
    prec = _prec(psi.dtype)
    ct = np.complex128 if prec == 64 else np.complex64
    nb = buffer_bytes(shp, psi.dtype)
    spec, aux, kg = DeviceBuffer(nb).upload(psi), DeviceBuffer(nb), DeviceBuffer(nb)
    ext = _ext(shp)
    r2c, c2r = FFTPlan(shp, psi.dtype, R2C), FFTPlan(shp, psi.dtype, C2R)
    L = lib()
    check(L.mcs_gaussian_init(kg.ptr, ext, (ctypes.c_double * 3)(*lens), (ctypes.c_double * 3)(*h), divs_per_sigma,
                              prec), "mcs_gaussian_init")
    r2c.execute(spec, aux)  # aux = spectrum of psi
    check(L.mcs_kmult(aux.ptr, kg.ptr, ext, 1, prec), "mcs_kmult")
    c2r.execute(aux, spec)
    check(L.mcs_norm(spec.ptr, ext, prec), "mcs_norm")
    check(L.mcs_square(spec.ptr, ext, prec), "mcs_square")
    r2c.execute(spec, aux)
    check(L.mcs_kmult(aux.ptr, kg.ptr, ext, 0, prec), "mcs_kmult")
    k = aux.download(half_shape(shp), ct)
    c2r.execute(aux, spec)
    check(L.mcs_norm(spec.ptr, ext, prec), "mcs_norm")
    return k, spec.download(shp, psi.dtype)


def write_lowpass_2d(spec, shape_zyx, ll, fn):
    """gzip text of a 2D low-passed half spectrum in the format of write_lowpass_2d
    (cuda/kernels/lowpass.hpp:54-78): "LOWPASS", the header length, a header naming
    n0 n1 l0 l1, then "i0 i1 re im" (C's %g) for modes with re^2 + im^2 > 1e-6, computed
    in the spectrum's precision; i0 runs over the half axis in the outer loop."""
    import gzip
    ny, nx = shape_zyx
    head = "(S'n0'\nI%d\nS'n1'\nI%d\nS'l0'\nF%f\nS'l1'\nF%f\nd.\n" % (nx, ny, ll[0], ll[1])
    sp = np.asarray(spec)
    rt = np.float64 if sp.dtype == np.complex128 else np.float32
    re, im = sp.real.astype(rt).T, sp.imag.astype(rt).T  # (i0, i1)
    keep = (re * re + im * im).astype(np.float64) > 1e-6
    i0, i1 = np.nonzero(keep)  # row-major over (i0, i1): the file's loop order
    body = "".join("%d %d %g %g\n" % (a, b, float(re[a, b]), float(im[a, b])) for a, b in zip(i0, i1))
    with gzip.open(fn, "wb") as f:
        f.write(("LOWPASS\n%d\n" % len(head) + head + body).encode())


def pair_distances(xyz, box):
    """minimum-image distances of all ordered pairs i != j, i-major (particlepairs/main.cpp:8-53)"""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    P = xyz.shape[0]
    total = P * (P - 1)
    out = np.empty(max(total, 1), dtype=np.float64)
    cnt = ctypes.c_int64()
    check(lib().mcs_pair_distances(xyz.ctypes.data_as(_dp), P, box, out.ctypes.data_as(_dp), total,
                                   ctypes.byref(cnt)), "mcs_pair_distances")
    return out[:cnt.value].copy()


def pair_histogram(xyz, box, rmax, nbins):
    """counts of unordered pairs per distance bin over [0, rmax)"""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    hist = np.zeros(nbins, dtype=np.int64)
    check(lib().mcs_pair_histogram(xyz.ctypes.data_as(_dp), xyz.shape[0], box, rmax, nbins,
                                   hist.ctypes.data_as(_i64p)), "mcs_pair_histogram")
    return hist

class Stepper:
    """The fused PFC stepper (mcs_stepper_*)."""

    def __init__(self, shape_zyx, h, eps, dt, temp=0.0, seed=0, split=SPLIT_CUBIC, dtype=np.float32,
                 plan=None, form=None):
        """plan: optional explicit [P_x, pairs/wg, P_y, lines/wg, P_z, lines/wg] instead of the tuned plans;
        form: FORM_SPECTRAL or FORM_MIXED (default: the library's default form, mcs_set_step_form)"""
        self.shape = tuple(shape_zyx)
        self.dtype = np.dtype(dtype)
        p = PfcParams()
        p.ext = _ext(shape_zyx)
        p.h = (ctypes.c_double * 3)(*h)
        p.eps, p.dt, p.temp, p.seed, p.split, p.prec = eps, dt, temp, seed, split, _prec(dtype)
        self.params = p
        ptr = _vp()
        if plan is None:
            check(lib().mcs_stepper_create(ctypes.byref(ptr), ctypes.byref(p)), "mcs_stepper_create")
        else:
            pl = (ctypes.c_int * 6)(*(list(plan) + [0] * (6 - len(plan))))
            check(lib().mcs_stepper_create_planned(ctypes.byref(ptr), ctypes.addressof(p), pl),
                  "mcs_stepper_create_planned")
        self.ptr = ptr
        if form is not None:
            self.set_form(form)
        self.buf = DeviceBuffer(int(np.prod(shape_zyx)) * self.dtype.itemsize)

    def set_form(self, form):
        """switch the state's form (converted in place when a field is loaded)"""
        check(lib().mcs_stepper_set_form(self.ptr, int(form)), "mcs_stepper_set_form")

    def form(self):
        v = ctypes.c_int()
        check(lib().mcs_stepper_get_form(self.ptr, ctypes.byref(v)), "mcs_stepper_get_form")
        return v.value

    def set_psi(self, psi):
        self.buf.upload(np.ascontiguousarray(psi, dtype=self.dtype))
        check(lib().mcs_stepper_set_psi(self.ptr, self.buf.ptr), "mcs_stepper_set_psi")

    def set_psi_device(self, dbuf):
        check(lib().mcs_stepper_set_psi(self.ptr, dbuf.ptr), "mcs_stepper_set_psi")

    def get_psi(self):
        check(lib().mcs_stepper_get_psi(self.ptr, self.buf.ptr), "mcs_stepper_get_psi")
        return self.buf.download(self.shape, self.dtype)

    def step(self, n=1):
        check(lib().mcs_stepper_step(self.ptr, n), "mcs_stepper_step")

    def step_async(self, n=1):
        check(lib().mcs_stepper_step_async(self.ptr, n), "mcs_stepper_step_async")

    def mass(self):
        m = ctypes.c_double()
        check(lib().mcs_stepper_mass(self.ptr, ctypes.byref(m)), "mcs_stepper_mass")
        return m.value

    def steps_taken(self):
        s = ctypes.c_int64()
        check(lib().mcs_stepper_get_step(self.ptr, ctypes.byref(s)), "mcs_stepper_get_step")
        return s.value

    def pass_times(self, n):
        """mean milliseconds per pass over n steps (HIP events between passes)"""
        ms = (ctypes.c_double * 8)()
        npass = ctypes.c_int()
        check(lib().mcs_stepper_pass_times(self.ptr, n, ms, 8, ctypes.byref(npass)), "mcs_stepper_pass_times")
        names = ["x_update", "y_inverse", "z_nonlinear", "y_forward"] if len(self.shape) == 3 else \
            ["x_update", "y_nonlinear"]
        return dict(zip(names, [ms[i] for i in range(npass.value)]))

    def get_state(self):
        """(state [nz][ny][nx/2] complex: the packed spectrum U in the spectral form, IFFT_x(U)/(nx/2) in the
        mixed form; FP64 mass, step): an exact checkpoint for a stepper of the same plans and form"""
        ct = np.complex128 if self.dtype == np.float64 else np.complex64
        shp = self.shape[:-1] + (self.shape[-1] // 2,)
        spec = np.empty(shp, ct)
        mass, step = ctypes.c_double(), ctypes.c_int64()
        check(lib().mcs_stepper_get_state(self.ptr, spec.ctypes.data, ctypes.byref(mass), ctypes.byref(step)),
              "mcs_stepper_get_state")
        return spec, mass.value, step.value

    def set_state(self, spec, mass, step):
        ct = np.complex128 if self.dtype == np.float64 else np.complex64
        spec = np.ascontiguousarray(spec, dtype=ct)
        check(lib().mcs_stepper_set_state(self.ptr, spec.ctypes.data, mass, step), "mcs_stepper_set_state")

    def plan_info(self):
        """points per thread and lines (row pairs) per workgroup of each pass"""
        info = (ctypes.c_int * 6)()
        check(lib().mcs_stepper_plan_info(self.ptr, info), "mcs_stepper_plan_info")
        names = ["x", "y", "z"] if len(self.shape) == 3 else ["x", "y"]
        return {n: {"P": info[2 * i], "tile": info[2 * i + 1]} for i, n in enumerate(names)}

    def __del__(self):
        try:
            lib().mcs_stepper_destroy(self.ptr)
        except Exception:
            pass


def set_autotune(on):
    """switch the per-shape plan and tile autotuner on (default) or off (fixed default plans)"""
    check(lib().mcs_set_autotune(1 if on else 0), "mcs_set_autotune")


def plan_options(L, dtype=np.float32):
    """register-plan points per thread available for an axis of length L (default first)"""
    buf = (ctypes.c_int * 8)()
    n = ctypes.c_int()
    check(lib().mcs_plan_options(L, _prec(dtype), buf, 8, ctypes.byref(n)), "mcs_plan_options")
    return [buf[i] for i in range(min(n.value, 8))]


def tile_options(L, dtype, P, xpass):
    buf = (ctypes.c_int * 16)()
    n = ctypes.c_int()
    check(lib().mcs_tile_options(L, _prec(dtype), P, 1 if xpass else 0, buf, 16, ctypes.byref(n)),
          "mcs_tile_options")
    return [buf[i] for i in range(min(n.value, 16))]


def set_step_form(form):
    """form of steppers created afterwards: FORM_MIXED (the default; 9 sweeps per 3D step) or FORM_SPECTRAL"""
    check(lib().mcs_set_step_form(int(form)), "mcs_set_step_form")


def set_lane_shuffle(mode):
    """lane shuffles in the mixed form's fixed-plan x pass: 0 never, 1 where the tuner timed them faster
    (default), 2 wherever a plan has them (results are identical either way)"""
    check(lib().mcs_set_lane_shuffle(int(mode)), "mcs_set_lane_shuffle")


def get_lane_shuffle():
    v = ctypes.c_int()
    check(lib().mcs_get_lane_shuffle(ctypes.byref(v)), "mcs_get_lane_shuffle")
    return v.value


def get_step_form():
    v = ctypes.c_int()
    check(lib().mcs_get_step_form(ctypes.byref(v)), "mcs_get_step_form")
    return v.value


def get_autotune():
    v = ctypes.c_int()
    check(lib().mcs_get_autotune(ctypes.byref(v)), "mcs_get_autotune")
    return bool(v.value)


FIXED_COLUMNS, FIXED_XPASS = 1, 2


def set_fixed_kernels(mask):
    """fixed-plan kernels for the passes in mask (FIXED_COLUMNS | FIXED_XPASS, the default) where a plan
    matches one; the other passes run the runtime-plan kernels"""
    check(lib().mcs_set_fixed_kernels(int(mask)), "mcs_set_fixed_kernels")


def get_fixed_kernels():
    v = ctypes.c_int()
    check(lib().mcs_get_fixed_kernels(ctypes.byref(v)), "mcs_get_fixed_kernels")
    return v.value


def sync():
    check(lib().mcs_sync(), "mcs_sync")


class Timer:
    """HIP events on the library stream around a region."""

    def __enter__(self):
        check(lib().mcs_timer_start(), "mcs_timer_start")
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_double()
        check(lib().mcs_timer_stop(ctypes.byref(ms)), "mcs_timer_stop")
        self.ms = ms.value
        return False


class Ensemble:
    """Batched replicas (mcs_ensemble_*).  Each replica is a dict with keys
    shape (numpy order), h, eps, dt, temp, psi0, amp, init_seed, noise_seed."""

    def __init__(self, replicas, split=SPLIT_CUBIC, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.shapes = [tuple(r["shape"]) for r in replicas]
        arr = (ReplicaParams * len(replicas))()
        for a, r in zip(arr, replicas):
            a.ext = _ext(r["shape"])
            h = r.get("h", [np.pi / 4] * 3)
            a.h = (ctypes.c_double * 3)(*(list(h) + [h[-1]] * (3 - len(h))))
            a.eps, a.dt, a.temp = r["eps"], r["dt"], r.get("temp", 0.0)
            a.psi0, a.amp = r["psi0"], r["amp"]
            a.init_seed, a.noise_seed = r["init_seed"], r.get("noise_seed", 0)
        p = _vp()
        check(lib().mcs_ensemble_create(ctypes.byref(p), arr, len(replicas), split, _prec(dtype)),
              "mcs_ensemble_create")
        self.ptr = p
        self.buf = DeviceBuffer(max(int(np.prod(s)) for s in self.shapes) * self.dtype.itemsize)

    def init(self):
        check(lib().mcs_ensemble_init(self.ptr), "mcs_ensemble_init")

    def set_psi(self, r, psi):
        self.buf.upload(np.ascontiguousarray(psi, dtype=self.dtype))
        check(lib().mcs_ensemble_set_psi(self.ptr, r, self.buf.ptr), "mcs_ensemble_set_psi")

    def get_psi(self, r):
        check(lib().mcs_ensemble_get_psi(self.ptr, r, self.buf.ptr), "mcs_ensemble_get_psi")
        return self.buf.download(self.shapes[r], self.dtype)

    def step(self, n, mode=ENS_GRAPH):
        check(lib().mcs_ensemble_step(self.ptr, n, mode), "mcs_ensemble_step")

    def step_async(self, n, mode=ENS_GRAPH):
        check(lib().mcs_ensemble_step_async(self.ptr, n, mode), "mcs_ensemble_step_async")

    def analyze(self, ths, k=12, nbins=64):
        nr = len(self.shapes)
        counts = np.zeros(nr, np.int64)
        h4 = np.zeros((nr, nbins), np.int64)
        h6 = np.zeros((nr, nbins), np.int64)
        check(lib().mcs_ensemble_analyze(self.ptr, ths, k, nbins, counts.ctypes.data_as(_i64p),
                                         h4.ctypes.data_as(_i64p), h6.ctypes.data_as(_i64p)), "mcs_ensemble_analyze")
        return counts, h4, h6

    def launch_info(self):
        """size groups, plan-class launches per step, and whether the step graph replays the classes"""
        g, c, bc, gc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 4)()
        check(lib().mcs_ensemble_launch_info(self.ptr, ctypes.byref(g), ctypes.byref(c), ctypes.byref(bc),
                                             ctypes.byref(gc), ms), "mcs_ensemble_launch_info")
        names = ["batched_per_group", "batched_classes", "graph_per_group", "graph_classes"]
        nb = ctypes.c_int()
        check(lib().mcs_ensemble_graph_branches(self.ptr, ctypes.byref(nb)), "mcs_ensemble_graph_branches")
        return {"size_groups": g.value, "class_launches_per_step": c.value, "batched_uses_classes": bool(bc.value),
                "graph_replays_classes": bool(gc.value), "graph_branches": nb.value,
                "ms_per_step_at_create": {n: round(ms[i], 4) for i, n in enumerate(names)}}

    def set_launch_form(self, batched_classes, graph_classes):
        check(lib().mcs_ensemble_set_launch_form(self.ptr, int(batched_classes), int(graph_classes)),
              "mcs_ensemble_set_launch_form")

    def set_graph_branches(self, branches):
        """per-group step graph on `branches` packed branches (0: one per size group)"""
        check(lib().mcs_ensemble_set_graph_branches(self.ptr, int(branches)), "mcs_ensemble_set_graph_branches")

    def run(self, nsteps, ths, k=12, nbins=64):
        """nsteps graph-replayed steps with peaks and q4/q6 histograms captured after the last one"""
        nr = len(self.shapes)
        counts = np.zeros(nr, np.int64)
        h4 = np.zeros((nr, nbins), np.int64)
        h6 = np.zeros((nr, nbins), np.int64)
        check(lib().mcs_ensemble_run(self.ptr, nsteps, ths, k, nbins, counts.ctypes.data_as(_i64p),
                                     h4.ctypes.data_as(_i64p), h6.ctypes.data_as(_i64p)), "mcs_ensemble_run")
        return counts, h4, h6

    def __del__(self):
        try:
            lib().mcs_ensemble_destroy(self.ptr)
        except Exception:
            pass


def ensemble_config(kind, nrep=None, seed=2024):
    """Replica lists of the two ensemble workloads, drawn from a seeded generator:
    '2d': sizes {128,192,320,384,640,768,1024}^2, eps U[0.1,0.4], psi0 U[-0.35,-0.15], T=5e-4, dt=0.5
    '3d': sizes {64,96,128,160,192}^3, eps U[0.2,0.35], psi0 U[-0.3,-0.18], T=1e-3, dt=0.25"""
    rng = np.random.default_rng(seed)
    out = []
    if kind == "2d":
        sizes, n = [128, 192, 320, 384, 640, 768, 1024], nrep or 64
        for i in range(n):
            s = int(rng.choice(sizes))
            out.append(dict(shape=(s, s), eps=float(rng.uniform(0.1, 0.4)), psi0=float(rng.uniform(-0.35, -0.15)),
                            amp=0.05, temp=5e-4, dt=0.5, init_seed=1000 + i, noise_seed=5000 + i))
    else:
        sizes, n = [64, 96, 128, 160, 192], nrep or 32
        for i in range(n):
            s = int(rng.choice(sizes))
            out.append(dict(shape=(s, s, s), eps=float(rng.uniform(0.2, 0.35)), psi0=float(rng.uniform(-0.3, -0.18)),
                            amp=0.05, temp=1e-3, dt=0.25, init_seed=2000 + i, noise_seed=6000 + i))
    return out
