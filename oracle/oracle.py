"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u32 = ctypes.POINTER(ctypes.c_uint32)

SPLIT_CUBIC = 0
SPLIT_REFERENCE = 1


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_fft_c2c.argtypes = [_d, ctypes.c_int, ctypes.c_int]
        L.orc_r2c.argtypes = [_d, _d, _i]
        L.orc_c2r.argtypes = [_d, _d, _i]
        L.orc_philox4x32_10.argtypes = [_u32, _u32, _u32]
        L.orc_philox_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_int64, _u32]
        L.orc_kspace_noise.argtypes = [_d, _i, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_stepper_noise.argtypes = [_d, _i, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_init_field.argtypes = [_d, _i, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        L.orc_update.argtypes = [_d, _d, _d, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int]
        L.orc_pfc_run.argtypes = [_d, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.orc_update_eom_fen.argtypes = [_d, _d, _d, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_calc_fen.argtypes = [_d, _d, _i, _d]
        L.orc_calc_fen.restype = ctypes.c_double
        L.orc_update_el.argtypes = [_d, _d, _i, _d, ctypes.c_double, ctypes.c_double]
        L.orc_update_el.restype = ctypes.c_double
        L.orc_etd_op.argtypes = [_d, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_update_eom_etd.argtypes = [_d, _d, _d, _d, _i, _d, ctypes.c_double]
        L.orc_update_el_etd.argtypes = [_d, _d, _d, _i]
        L.orc_update_el_etd.restype = ctypes.c_double
        L.orc_update_eom_hpfc.argtypes = [_d, _d, _d, _d, _d, _i, _d, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double]
        L.orc_update_eom_poly.argtypes = [_d, _d, _d, _i, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _d, ctypes.c_int]
        L.orc_poly_coefficients.argtypes = [_d]
        L.orc_free_energy.argtypes = [_d, _i, _d, ctypes.c_double]
        L.orc_free_energy.restype = ctypes.c_double
        L.orc_findpeaks.argtypes = [_d, _i, ctypes.c_double, _i64, ctypes.c_int64]
        L.orc_findpeaks.restype = ctypes.c_int64
        L.orc_bond_order_cutoff.argtypes = [_d, ctypes.c_int, ctypes.c_double, ctypes.c_double, _i, _d]
        L.orc_bond_order_knn.argtypes = [_i, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _i, _d]
        L.orc_bond_order_knn_points.argtypes = [_i, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _i,
                                                ctypes.c_int, _i, _d]
        L.orc_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _d, _d]
        L.orc_wigner3j.argtypes = [ctypes.c_int] * 6
        L.orc_wigner3j.restype = ctypes.c_double
        L.orc_gaussian_init.argtypes = [_d, _i, _d, _d, ctypes.c_double]
        L.orc_pair_distances.argtypes = [_d, ctypes.c_int, ctypes.c_double, _d]
        L.orc_pair_distances.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _dims(shape_zyx):
    """numpy shape (nz, ny, nx) / (ny, nx) / (nx,) -> int[3] {nx, ny, nz}"""
    s = list(shape_zyx)[::-1] + [1] * (3 - len(shape_zyx))
    return np.array(s, dtype=np.int32)


def half_shape(shape_zyx):
    s = list(shape_zyx)
    s[-1] = s[-1] // 2 + 1
    return tuple(s)


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(half_shape(x.shape), dtype=np.complex128)
    lib().orc_r2c(_p(x), _p(out), _p(_dims(x.shape), _i))
    return out


def c2r(k, shape_zyx):
    k = np.ascontiguousarray(k, dtype=np.complex128)
    out = np.empty(tuple(shape_zyx), dtype=np.float64)
    lib().orc_c2r(_p(k), _p(out), _p(_dims(shape_zyx), _i))
    return out


def fft_c2c(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft_c2c(_p(a), a.size, sign)
    return a


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32), _p(k, _u32), _p(o, _u32))
    return o


def philox_stream(seed, c0, c2, c3, n):
    o = np.empty(4 * n, dtype=np.uint32)
    lib().orc_philox_stream(seed, c0, c2, c3, n, _p(o, _u32))
    return o


def kspace_noise(shape_zyx, seed, step):
    out = np.empty(half_shape(shape_zyx), dtype=np.complex128)
    lib().orc_kspace_noise(_p(out), _p(_dims(shape_zyx), _i), seed, step)
    return out


def stepper_noise(shape_zyx, seed, step):
    out = np.empty(half_shape(shape_zyx), dtype=np.complex128)
    lib().orc_stepper_noise(_p(out), _p(_dims(shape_zyx), _i), seed, step)
    return out


def init_field(shape_zyx, psi0, amp, seed):
    out = np.empty(tuple(shape_zyx), dtype=np.float64)
    lib().orc_init_field(_p(out), _p(_dims(shape_zyx), _i), psi0, amp, seed)
    return out


def update(kpsi, knl, xi, shape_zyx, lens, dt, eps, noise_amp, split):
    k = np.ascontiguousarray(kpsi, dtype=np.complex128).copy()
    nl = np.ascontiguousarray(knl, dtype=np.complex128)
    xp = None if xi is None else _p(np.ascontiguousarray(xi, dtype=np.complex128))
    ln = np.ascontiguousarray(lens, dtype=np.float64)
    lib().orc_update(_p(k), _p(nl), xp, _p(_dims(shape_zyx), _i), _p(ln), dt, eps, noise_amp, split)
    return k


def pfc_run(psi, h, eps, dt, temp, seed, step0, nsteps, split=SPLIT_CUBIC, want_mass=False):
    p = np.ascontiguousarray(psi, dtype=np.float64).copy()
    hh = np.ascontiguousarray(h, dtype=np.float64)
    mass = np.empty(nsteps + 1, dtype=np.float64) if want_mass else None
    lib().orc_pfc_run(_p(p), _p(_dims(p.shape), _i), _p(hh), eps, dt, temp, seed, step0, nsteps, split,
                      None if mass is None else _p(mass))
    return (p, mass) if want_mass else p


def update_eom_fen(kpsi, knl, knoise, shape_zyx, lens, namp, dt, eps):
    """returns (updated kpsi, lk * old kpsi)"""
    kp = np.ascontiguousarray(kpsi, dtype=np.complex128).copy()
    kn = np.ascontiguousarray(knoise, dtype=np.complex128).copy()
    lib().orc_update_eom_fen(_p(kp), _p(np.ascontiguousarray(knl, dtype=np.complex128)), _p(kn),
                             _p(_dims(shape_zyx), _i), _p(np.asarray(lens, dtype=np.float64)), namp, dt, eps)
    return kp, kn


def calc_fen(psi, lpsi, h):
    """returns (total, per-cell density)"""
    lp = np.ascontiguousarray(lpsi, dtype=np.float64).copy()
    f = lib().orc_calc_fen(_p(np.ascontiguousarray(psi, dtype=np.float64)), _p(lp), _p(_dims(np.shape(psi)), _i),
                           _p(np.asarray(h, dtype=np.float64)))
    return f, lp


def update_el(kpsi, knl, shape_zyx, lens, dt, eps):
    """returns (updated kpsi, max change)"""
    kp = np.ascontiguousarray(kpsi, dtype=np.complex128).copy()
    m = lib().orc_update_el(_p(kp), _p(np.ascontiguousarray(knl, dtype=np.complex128)), _p(_dims(shape_zyx), _i),
                            _p(np.asarray(lens, dtype=np.float64)), dt, eps)
    return kp, m


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def etd_op(shape_zyx, lens, dt, eps, el):
    out = np.empty(half_shape(shape_zyx), dtype=np.complex128)  # (E, S) pairs
    lib().orc_etd_op(_p(out), _p(_dims(shape_zyx), _i), _p(np.asarray(lens, dtype=np.float64)), dt, eps, int(el))
    return out


def update_eom_etd(kop, kpsi, knl, knoise, shape_zyx, lens, namp):
    kp = _c(kpsi).copy()
    kz = None if knoise is None else _c(knoise)
    lib().orc_update_eom_etd(_p(_c(kop)), _p(kp), _p(_c(knl)), None if kz is None else _p(kz),
                             _p(_dims(shape_zyx), _i), _p(np.asarray(lens, dtype=np.float64)), namp)
    return kp


def update_el_etd(kop, kpsi, knl, shape_zyx):
    kp = _c(kpsi).copy()
    m = lib().orc_update_el_etd(_p(_c(kop)), _p(kp), _p(_c(knl)), _p(_dims(shape_zyx), _i))
    return kp, m


def update_eom_hpfc(kpsi, knl, x1, x2, x3, shape_zyx, lens, namp, dt, eps):
    kp = _c(kpsi).copy()
    lib().orc_update_eom_hpfc(_p(kp), _p(_c(knl)), _p(_c(x1)), _p(_c(x2)), _p(_c(x3)), _p(_dims(shape_zyx), _i),
                              _p(np.asarray(lens, dtype=np.float64)), namp, dt, eps)
    return kp


def update_eom_poly(kpsi, knl, knoise, shape_zyx, lens, namp, dt, eps, coef=None):
    """coef: P(q) = sum_j coef[j] q^j; None takes the reference's fitted coefficients"""
    kp = _c(kpsi).copy()
    kz = None if knoise is None else _c(knoise)
    cf = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    lib().orc_update_eom_poly(_p(kp), _p(_c(knl)), None if kz is None else _p(kz), _p(_dims(shape_zyx), _i),
                              _p(np.asarray(lens, dtype=np.float64)), namp, dt, eps, None if cf is None else _p(cf),
                              0 if cf is None else cf.size)
    return kp


def poly_coefficients():
    out = np.empty(7, dtype=np.float64)
    lib().orc_poly_coefficients(_p(out))
    return out


def free_energy(psi, h, eps):
    p = np.ascontiguousarray(psi, dtype=np.float64)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    return lib().orc_free_energy(_p(p), _p(_dims(p.shape), _i), _p(hh), eps)


def findpeaks(v, ths):
    v = np.ascontiguousarray(v, dtype=np.float64)
    cap = v.size
    out = np.empty(max(cap, 1), dtype=np.int64)
    n = lib().orc_findpeaks(_p(v), _p(_dims(v.shape), _i), ths, _p(out, _i64), cap)
    return out[:n].copy()


def bond_order_cutoff(xyz, box, rlim):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    P = xyz.shape[0]
    nb = np.zeros(max(P, 1), dtype=np.int32)
    out = np.zeros((max(P, 1), 9), dtype=np.float64)
    lib().orc_bond_order_cutoff(_p(xyz), P, box, rlim, _p(nb, _i), _p(out))
    return nb[:P], out[:P]


def bond_order_knn(ijk, n, h, k=12):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    P = ijk.shape[0]
    nbr = np.zeros((max(P, 1), k), dtype=np.int32)
    out = np.zeros((max(P, 1), 4), dtype=np.float64)
    lib().orc_bond_order_knn(_p(ijk, _i), P, n, h, k, _p(nbr, _i), _p(out))
    return nbr[:P], out[:P]


def bond_order_knn_points(ijk, n, h, queries, k=12):
    """bond_order_knn for the points `queries` only (all points are candidates): rows follow `queries`"""
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    qi = np.ascontiguousarray(queries, dtype=np.int32)
    nq = qi.size
    nbr = np.zeros((max(nq, 1), k), dtype=np.int32)
    out = np.zeros((max(nq, 1), 4), dtype=np.float64)
    lib().orc_bond_order_knn_points(_p(ijk, _i), ijk.shape[0], n, h, k, _p(qi, _i), nq, _p(nbr, _i), _p(out))
    return nbr[:nq], out[:nq]


def ylm(l, m, ct, phi):
    re = ctypes.c_double()
    im = ctypes.c_double()
    lib().orc_ylm(l, m, ct, phi, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().orc_wigner3j(j1, j2, j3, m1, m2, m3)


def ref_bond_order_binary():
    """Path of the reference's own bond-order analysis built from its sources
    (bond-order-analysis/standalone.cpp + bond-order.cpp), or None."""
    p = os.path.join(_HERE, "_ref", "bond_order_standalone")
    return p if os.path.exists(p) else None


# ---------------------------------------------------------------- post-processing

def gaussian_init(shape_zyx, lens, h, divs):
    """(e, 1 - e) per half-spectrum mode (cuda/kernels/gaussian.hpp:23-33)"""
    out = np.empty(half_shape(shape_zyx), dtype=np.complex128)
    lib().orc_gaussian_init(_p(out), _p(_dims(shape_zyx), _i), _p(np.asarray(lens, dtype=np.float64)),
                            _p(np.asarray(h, dtype=np.float64)), divs)
    return out


def kmult(k1, k2, part):
    """k1 * Re k2 (part 0) or k1 * Im k2 (part 1) (gaussian.hpp:42-67, lowpass.hpp:3-35)"""
    f = k2.real if part == 0 else k2.imag
    return k1 * f


def amplitude_lowpass(psi, lens, h, divs):
    """calc_amplitude_lowpass (cuda/kernels/lowpass.hpp:79-93) of a 2D field in FP64:
    returns (low-passed spectrum written to the text file, amplitude field)"""
    shp = psi.shape
    N = psi.size
    kg = gaussian_init(shp, lens, h, divs)
    hi = c2r(kmult(r2c(psi), kg, 1), shp) / N
    spec = kmult(r2c(hi * hi), kg, 0)
    return spec, c2r(spec, shp) / N


def lowpass_text(spec, shape_zyx, ll, real_type=np.float64):
    """the text of write_lowpass_2d (lowpass.hpp:54-78) for a 2D half spectrum
    (ny, nx/2+1), values as real_type: C's %g (the stream's default 6 digits)"""
    ny, nx = shape_zyx
    head = "(S'n0'\nI%d\nS'n1'\nI%d\nS'l0'\nF%f\nS'l1'\nF%f\nd.\n" % (nx, ny, ll[0], ll[1])
    lines = ["LOWPASS\n%d\n" % len(head), head]
    sp = np.asarray(spec)
    re = sp.real.astype(real_type)
    im = sp.imag.astype(real_type)
    p = re * re + im * im  # in real_type, compared in double
    for i0 in range(nx // 2 + 1):
        for i1 in range(ny):
            if float(p[i1, i0]) > 1e-6:
                lines.append("%d %d %g %g\n" % (i0, i1, float(re[i1, i0]), float(im[i1, i0])))
    return "".join(lines)


def pair_distances(xyz, box):
    """minimum-image distances of all ordered pairs, i-major (particlepairs/main.cpp:8-53)"""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    P = xyz.shape[0]
    out = np.empty(max(P * (P - 1), 1), dtype=np.float64)
    n = lib().orc_pair_distances(_p(xyz), P, box, _p(out))
    return out[:n].copy()


def pair_histogram(xyz, box, rmax, nbins):
    """counts of unordered pairs i < j per bin floor(d / rmax * nbins) for d < rmax"""
    P = len(xyz)
    d = pair_distances(xyz, box).reshape(P, P - 1) if P > 1 else np.zeros((P, 0))
    iu = [d[i, i:] for i in range(P)]  # row i holds j = 0..i-1, i+1..; j > i are columns i..
    dd = np.concatenate(iu) if iu else np.zeros(0)
    dd = dd[dd < rmax]
    b = np.minimum((dd / rmax * nbins).astype(np.int64), nbins - 1)
    return np.bincount(b, minlength=nbins).astype(np.int64)


def ref_particlepairs_binary():
    """Path of the reference's particlepairs tool built from particlepairs/main.cpp, or None."""
    p = os.path.join(_HERE, "_ref", "particlepairs")
    return p if os.path.exists(p) else None
