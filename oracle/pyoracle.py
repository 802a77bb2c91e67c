"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the arma_amd product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class MtParams(C.Structure):
    """Mirror of orc_mt_params (the dcmt mt_struct fields, src/parallel_mt.hh:109-123)."""
    _fields_ = [
        ("aaa", C.c_uint32), ("mm", C.c_int32), ("nn", C.c_int32), ("rr", C.c_int32),
        ("ww", C.c_int32), ("wmask", C.c_uint32), ("umask", C.c_uint32),
        ("lmask", C.c_uint32), ("shift0", C.c_int32), ("shift1", C.c_int32),
        ("shiftB", C.c_int32), ("shiftC", C.c_int32), ("maskB", C.c_uint32),
        ("maskC", C.c_uint32),
    ]


def mt19937_params():
    """The MT19937 parameter set written as a dynamic-creator mt_struct."""
    return MtParams(aaa=0x9908B0DF, mm=397, nn=624, rr=31, ww=32, wmask=0xFFFFFFFF,
                    umask=0x80000000, lmask=0x7FFFFFFF, shift0=11, shift1=18,
                    shiftB=7, shiftC=15, maskB=0x9D2C5680, maskC=0xEFC60000)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        ip3 = C.c_int * 3
        L.orc_mt_fill.argtypes = [C.POINTER(MtParams), C.c_uint32, up, C.c_size_t]
        L.orc_noise_sequential_d.argtypes = [C.POINTER(MtParams), C.c_uint32, C.c_double, dp, C.c_size_t]
        L.orc_noise_sequential_f.argtypes = [C.POINTER(MtParams), C.c_uint32, C.c_float, fp, C.c_size_t]
        L.orc_noise_partitioned_d.argtypes = [C.POINTER(MtParams), C.c_void_p, C.c_int, ip3, ip3,
                                              C.c_double, dp]
        L.orc_noise_partitioned_d_omp.argtypes = L.orc_noise_partitioned_d.argtypes + [C.c_int]
        L.orc_acf.argtypes = [C.c_int, C.c_double * 3, ip3, C.c_double, C.c_double,
                              C.c_double * 3, C.c_double * 2, dp]
        L.orc_yw_choi.argtypes = [dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                  dp, ip3, C.POINTER(C.c_double)]
        L.orc_yw_gauss.argtypes = [dp, ip3, ip3, dp]
        L.orc_chop_right.argtypes = [dp, ip3, C.c_double, ip3]
        L.orc_white_noise_variance.argtypes = [dp, ip3, dp, ip3]
        L.orc_white_noise_variance.restype = C.c_double
        L.orc_partition_shape.argtypes = [ip3, ip3, ip3, C.c_int, ip3]
        L.orc_ar_generate.argtypes = [dp, ip3, dp, ip3, ip3, ip3]
        L.orc_ar_generate_f.argtypes = [fp, ip3, fp, ip3, ip3, ip3]
        L.orc_ar_generate_dag.argtypes = [dp, ip3, dp, ip3, ip3]
        L.orc_ar_generate_dag_omp.argtypes = [dp, ip3, dp, ip3, ip3, C.c_int]
        L.orc_gsl_variance_d.argtypes = [dp, C.c_size_t]
        L.orc_gsl_variance_d.restype = C.c_double
        L.orc_gsl_variance_f.argtypes = [fp, C.c_size_t]
        L.orc_gsl_variance_f.restype = C.c_float
        L.orc_rescale_factor.argtypes = [dp, ip3, C.c_double]
        L.orc_rescale_factor.restype = C.c_double
        L.orc_rescale_factor_f.argtypes = [fp, ip3, C.c_double]
        L.orc_rescale_factor_f.restype = C.c_double
        L.orc_derivative_t.argtypes = [dp, ip3, C.c_double, C.c_int, dp]
        L.orc_factor_waves.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double * 4]
        L.orc_window.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.orc_fft2.argtypes = [dp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_interp_edges.argtypes = [dp, C.c_int, C.c_int]
        L.orc_linear_potential.argtypes = [dp, ip3, C.c_double * 3, C.c_double, C.c_double * 2,
                                           C.c_double * 2, C.c_int * 2, dp, C.c_void_p]
        L.orc_derivative_xy.argtypes = [dp, ip3, C.c_double, C.c_int, C.c_int, dp]
        L.orc_high_amplitude_potential.argtypes = L.orc_linear_potential.argtypes
        L.orc_linear_potential_omp.argtypes = [dp, ip3, C.c_double * 3, C.c_double,
                                               C.c_double * 2, C.c_double * 2, C.c_int * 2, dp,
                                               C.c_int]
        _lib = L
    return _lib


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def params_array(plist):
    arr = (MtParams * len(plist))()
    for i, p in enumerate(plist):
        arr[i] = p
    return arr


def mt_fill(params, seed, n):
    out = np.empty(n, np.uint32)
    rc = lib().orc_mt_fill(C.byref(params), seed, out, n)
    assert rc == 0
    return out


def noise_sequential(params, seed, stddev, n, dtype=np.float64):
    out = np.empty(n, dtype)
    if dtype == np.float64:
        rc = lib().orc_noise_sequential_d(C.byref(params), seed, stddev, out, n)
    else:
        rc = lib().orc_noise_sequential_f(C.byref(params), seed, stddev, out, n)
    assert rc == 0
    return out


def noise_partitioned(plist, seeds, shape, part, stddev, threads=1):
    """Partitioned noise; threads > 1 spreads the partitions over OpenMP
    threads (same result)."""
    out = np.empty(tuple(shape), np.float64)
    arr = params_array(plist)
    sarr = None
    if seeds is not None:
        sarr = np.ascontiguousarray(seeds, dtype=np.uint32)
    args = (arr, None if sarr is None else sarr.ctypes.data, len(plist), _i3(shape), _i3(part),
            stddev, out.reshape(-1))
    if threads > 1:
        rc = lib().orc_noise_partitioned_d_omp(*args, threads)
    else:
        rc = lib().orc_noise_partitioned_d(*args)
    if rc != 0:
        raise RuntimeError(f"orc_noise_partitioned_d failed: {rc}")
    return out


STANDING, PROPAGATING = 0, 1


def acf(func, shape, delta=(1.0, 1.0, 1.0), amplitude=1.0, velocity=0.8,
        alpha=(0.06, 0.06, 0.06), beta=(0.8, 0.0)):
    out = np.empty(tuple(shape), np.float64)
    lib().orc_acf(func, (C.c_double * 3)(*delta), _i3(shape), amplitude, velocity,
                  (C.c_double * 3)(*alpha), (C.c_double * 2)(*beta), out.reshape(-1))
    return out


def yw_choi(acf_cube, max_order, var_eps=1e-6, determine_order=True, chop=True, chop_eps=1e-10):
    n = acf_cube.shape[0]
    acf_cube = np.ascontiguousarray(acf_cube, np.float64)
    phi = np.zeros(max(max_order, 2) ** 3 + n ** 3, np.float64)
    order = (C.c_int * 3)()
    var = C.c_double()
    rc = lib().orc_yw_choi(acf_cube.reshape(-1), n, max_order, var_eps, int(determine_order),
                           int(chop), chop_eps, phi, order, C.byref(var))
    if rc != 0:
        raise RuntimeError(f"orc_yw_choi failed: {rc}")
    shp = tuple(order)
    return phi[: shp[0] * shp[1] * shp[2]].reshape(shp).copy(), var.value


def yw_gauss(acf_cube, order):
    acf_cube = np.ascontiguousarray(acf_cube, np.float64)
    phi = np.zeros(int(np.prod(order)), np.float64)
    rc = lib().orc_yw_gauss(acf_cube.reshape(-1), _i3(acf_cube.shape), _i3(order), phi)
    if rc != 0:
        raise RuntimeError(f"orc_yw_gauss failed: {rc}")
    return phi.reshape(tuple(order))


def white_noise_variance(acf_cube, phi):
    return lib().orc_white_noise_variance(np.ascontiguousarray(acf_cube).reshape(-1),
                                          _i3(acf_cube.shape),
                                          np.ascontiguousarray(phi).reshape(-1), _i3(phi.shape))


def partition_shape(partition, grid, order, parallelism):
    out = (C.c_int * 3)()
    lib().orc_partition_shape(_i3(partition), _i3(grid), _i3(order), parallelism, out)
    return tuple(out)


def ar_generate(eps, phi, lo=None, hi=None):
    """Applies the recurrence in place to a copy of eps and returns it."""
    if eps.dtype == np.float32:
        z = np.ascontiguousarray(eps, np.float32).copy()
        ph = np.ascontiguousarray(phi, np.float32)
        fn = lib().orc_ar_generate_f
    else:
        z = np.ascontiguousarray(eps, np.float64).copy()
        ph = np.ascontiguousarray(phi, np.float64)
        fn = lib().orc_ar_generate
    lo = (0, 0, 0) if lo is None else lo
    hi = tuple(s - 1 for s in z.shape) if hi is None else hi
    fn(z.reshape(-1), _i3(z.shape), ph.reshape(-1), _i3(ph.shape), _i3(lo), _i3(hi))
    return z


def ar_generate_dag(eps, phi, part, threads=1):
    """The recurrence over the partition DAG; threads > 1 runs each
    anti-diagonal level of partitions in parallel (same result)."""
    z = np.ascontiguousarray(eps, np.float64).copy()
    ph = np.ascontiguousarray(phi, np.float64)
    if threads > 1:
        lib().orc_ar_generate_dag_omp(z.reshape(-1), _i3(z.shape), ph.reshape(-1), _i3(ph.shape),
                                      _i3(part), threads)
    else:
        lib().orc_ar_generate_dag(z.reshape(-1), _i3(z.shape), ph.reshape(-1), _i3(ph.shape),
                                  _i3(part))
    return z


def gsl_variance(data):
    data = np.ascontiguousarray(data).reshape(-1)
    if data.dtype == np.float32:
        return float(lib().orc_gsl_variance_f(data, data.size))
    return float(lib().orc_gsl_variance_d(np.ascontiguousarray(data, np.float64), data.size))


def rescale_factor(zeta, gamma0):
    z = np.ascontiguousarray(zeta)
    if z.dtype == np.float32:
        return float(lib().orc_rescale_factor_f(z.reshape(-1), _i3(z.shape), gamma0))
    return float(lib().orc_rescale_factor(np.ascontiguousarray(z, np.float64).reshape(-1),
                                          _i3(z.shape), gamma0))


def derivative_t(zeta, dt, idx_t):
    z = np.ascontiguousarray(zeta, np.float64)
    out = np.empty(z.shape[1:], np.float64)
    lib().orc_derivative_t(z.reshape(-1), _i3(z.shape), dt, idx_t, out.reshape(-1))
    return out


def factor_waves(slice2d, t, dt):
    s = np.ascontiguousarray(slice2d, np.float64)
    out = (C.c_double * 4)()
    lib().orc_factor_waves(s.reshape(-1), s.shape[0], s.shape[1], int(t), dt, out)
    return tuple(out)


def window(nx, ny, wnmax_x, wnmax_y, z, depth):
    out = np.empty((nx, ny), np.float64)
    lib().orc_window(nx, ny, wnmax_x, wnmax_y, z, depth, out.reshape(-1))
    return out


def fft2(x, sign):
    re = np.ascontiguousarray(x.real, np.float64).copy()
    im = np.ascontiguousarray(x.imag, np.float64).copy()
    lib().orc_fft2(re.reshape(-1), im.reshape(-1), x.shape[0], x.shape[1], sign)
    return re + 1j * im


def interp_edges(res):
    r = np.ascontiguousarray(res, np.float64).copy()
    lib().orc_interp_edges(r.reshape(-1), r.shape[0], r.shape[1])
    return r


def derivative_xy(zeta, d, dim, idx_t):
    """d zeta / dx (dim 1) or d zeta / dy (dim 2) at slice idx_t."""
    z = np.ascontiguousarray(zeta, np.float64)
    out = np.empty(z.shape[1:], np.float64)
    lib().orc_derivative_xy(z.reshape(-1), _i3(z.shape), d, dim, idx_t, out.reshape(-1))
    return out


def linear_potential(zeta, delta, depth, lbound, ubound, npoints, return_wnmax=False,
                     solver="linear"):
    z = np.ascontiguousarray(zeta, np.float64)
    out = np.empty((npoints[0], npoints[1], z.shape[1], z.shape[2]), np.float64)
    wn = np.empty(2 * npoints[0], np.float64)
    fn = {"linear": lib().orc_linear_potential,
          "high_amplitude": lib().orc_high_amplitude_potential}[solver]
    rc = fn(z.reshape(-1), _i3(z.shape), (C.c_double * 3)(*delta), depth,
            (C.c_double * 2)(*lbound), (C.c_double * 2)(*ubound),
            (C.c_int * 2)(*npoints), out.reshape(-1), wn.ctypes.data)
    if rc != 0:
        raise ValueError("min(z) is less than -h")
    if return_wnmax:
        return out, wn.reshape(-1, 2)
    return out


def linear_potential_omp(zeta, delta, depth, lbound, ubound, npoints, threads):
    """The linear potential with the levels of each slice spread over OpenMP
    threads (same result as linear_potential)."""
    z = np.ascontiguousarray(zeta, np.float64)
    out = np.empty((npoints[0], npoints[1], z.shape[1], z.shape[2]), np.float64)
    rc = lib().orc_linear_potential_omp(z.reshape(-1), _i3(z.shape), (C.c_double * 3)(*delta),
                                        depth, (C.c_double * 2)(*lbound),
                                        (C.c_double * 2)(*ubound), (C.c_int * 2)(*npoints),
                                        out.reshape(-1), threads)
    if rc != 0:
        raise ValueError("min(z) is less than -h")
    return out


def high_amplitude_potential(zeta, delta, depth, lbound, ubound, npoints, return_wnmax=False):
    return linear_potential(zeta, delta, depth, lbound, ubound, npoints, return_wnmax,
                            solver="high_amplitude")


# ---- non-linear inertialess transform of the realisation ----

GRAM_CHARLIER, SKEW_NORMAL = 0, 1


def _nit_lib():
    L = lib()
    if not getattr(L, "_nit_set", False):
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_gauss_legendre_half.argtypes = [C.c_int, dp, dp]
        L.orc_owen_t.argtypes = [C.c_double, C.c_double]
        L.orc_owen_t.restype = C.c_double
        L.orc_gram_charlier_cdf.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_gram_charlier_cdf.restype = C.c_double
        L.orc_skew_normal_cdf.argtypes = [C.c_double] * 4
        L.orc_skew_normal_cdf.restype = C.c_double
        L.orc_nit_realisation.argtypes = [dp, C.c_size_t, C.c_double, C.c_int, C.c_double * 5,
                                          C.c_double, C.c_double, C.c_double, C.c_int]
        L._nit_set = True
    return L


def gauss_legendre_half(m):
    x = np.empty(m)
    w = np.empty(m)
    _nit_lib().orc_gauss_legendre_half(m, x, w)
    return x, w


def owen_t(h, a):
    return _nit_lib().orc_owen_t(h, a)


def gram_charlier_cdf(x, skewness, kurtosis):
    return _nit_lib().orc_gram_charlier_cdf(x, skewness, kurtosis)


def skew_normal_cdf(x, mean, sigma, alpha):
    return _nit_lib().orc_skew_normal_cdf(x, mean, sigma, alpha)


def nit_realisation(data, stdev, dist, skewness=0.0, kurtosis=3.0, mean=0.0, sigma=1.0,
                    alpha=0.0, interval=(0.0, 1.0), eps=1e-6, max_iter=100):
    """transform_realisation of the reference on a copy of data (FP64)."""
    out = np.ascontiguousarray(data, np.float64).copy()
    p = (C.c_double * 5)(skewness, kurtosis, mean, sigma, alpha)
    _nit_lib().orc_nit_realisation(out.reshape(-1), out.size, stdev, dist, p, interval[0],
                                   interval[1], eps, max_iter)
    return out


# ---- ACF_generator ----

WAVE_SIN, WAVE_COS, WAVE_STOKES, WAVE_STANDING = 0, 1, 2, 3


def acf_generate(func, amplitude=1.0, velocity=1.0, wavenum=(0.8, 0.0), alpha=(0.06,) * 3,
                 nwaves=(1.5,) * 3, shape=(16, 16, 16)):
    L = lib()
    if not getattr(L, "_acfgen_set", False):
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_acf_generate.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double * 2,
                                       C.c_double * 3, C.c_double * 3, C.c_int * 3, dp,
                                       C.c_size_t, C.c_int * 3]
        L._acfgen_set = True
    cap = 1 << 20
    out = np.empty(cap)
    so = (C.c_int * 3)()
    rc = L.orc_acf_generate(func, amplitude, velocity, (C.c_double * 2)(*wavenum),
                            (C.c_double * 3)(*alpha), (C.c_double * 3)(*nwaves),
                            (C.c_int * 3)(*shape), out, cap, so)
    if rc != 0:
        raise ValueError("ACF too large")
    s = tuple(so)
    return out[:int(np.prod(s))].reshape(s).copy()
