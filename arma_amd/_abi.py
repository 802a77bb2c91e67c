"""ctypes binding of libarma_hip.so (the C ABI declared in include/arma_hip.h).

This is the same binding a Python caller of the reference would add around
its C++ generator and solver; the compute runs in the HIP library, never in
Python.  Device buffers are torch tensors on a ROCm device (torch is used for
allocation and streams only).  The library is loaded from the package
directory; if it is missing, importing fails loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarma_hip.so")

ARMA_HIP_OK = 0
ARMA_HIP_EINVAL = -1
ARMA_HIP_EPRNG = -2
ARMA_HIP_ERUNTIME = -3
ARMA_HIP_EDEVICE = -4
ARMA_HIP_ENOMEM = -5

FP64, FP32 = 0, 1
NOISE_SEQUENTIAL, NOISE_PARTITIONED = 0, 1
NOISE_DOUBLE, NOISE_FLOAT = 0, 1
AR_PLAIN, AR_WAVEFRONT = 0, 1
BATCH_AUTO, BATCH_SURFACE_MAJOR, BATCH_INTERLEAVED = 0, 1, 2
REQ_SURFACE, REQ_STREAM_CHUNK = 0, 1

EXPORTS = [
    "arma_hip_init", "arma_hip_destroy", "arma_hip_last_error", "arma_hip_set_stream",
    "arma_hip_get_stream", "arma_hip_synchronize", "arma_hip_white_noise",
    "arma_hip_ar_recurrence", "arma_hip_ar_generate", "arma_hip_rescale",
    "arma_hip_linear_potential", "arma_hip_high_amplitude_potential", "arma_hip_mt_raw", "arma_hip_read_mt_config",
    "arma_hip_write_mt_config", "arma_hip_dcmt_create", "arma_hip_dcmt_create_many",
    "arma_hip_acf", "arma_hip_yule_walker_choi", "arma_hip_yule_walker_gauss",
    "arma_hip_white_noise_variance", "arma_hip_partition_shape", "arma_hip_profile",
    "arma_hip_profile_report", "arma_hip_batch_create", "arma_hip_batch_destroy",
    "arma_hip_batch_layout", "arma_hip_batch_generate", "arma_hip_batch_capture",
    "arma_hip_batch_replay", "arma_hip_batch_tune", "arma_hip_batch_get_config",
    "arma_hip_batch_set_config", "arma_hip_batch_factors", "arma_hip_server_clock",
    "arma_hip_server_create", "arma_hip_server_destroy", "arma_hip_server_open_stream",
    "arma_hip_server_close_stream", "arma_hip_server_submit", "arma_hip_server_poll",
    "arma_hip_server_release", "arma_hip_server_stats", "arma_hip_server_copy",
    "arma_hip_yule_walker_choi_device", "arma_hip_white_noise_variance_device",
    "arma_hip_validate_process", "arma_hip_nit_realisation", "arma_hip_acf_generate", "arma_hip_writer_open", "arma_hip_writer_append",
    "arma_hip_writer_close", "arma_hip_ar_generate_noalloc", "arma_hip_acf_generate_device",
    "arma_hip_yule_walker_choi_device_many",
]


class MtParams(C.Structure):
    _fields_ = [
        ("aaa", C.c_uint32), ("mm", C.c_int32), ("nn", C.c_int32), ("rr", C.c_int32),
        ("ww", C.c_int32), ("wmask", C.c_uint32), ("umask", C.c_uint32),
        ("lmask", C.c_uint32), ("shift0", C.c_int32), ("shift1", C.c_int32),
        ("shiftB", C.c_int32), ("shiftC", C.c_int32), ("maskB", C.c_uint32),
        ("maskC", C.c_uint32),
    ]

    def as_tuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


class ArDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int32 * 3), ("phi", C.POINTER(C.c_double)), ("phi_shape", C.c_int32 * 3),
        ("varwn", C.c_double), ("noise_mode", C.c_int32), ("noise_semantics", C.c_int32),
        ("partition", C.c_int32 * 3), ("mts", C.POINTER(MtParams)), ("n_mts", C.c_int32),
        ("seeds", C.POINTER(C.c_uint32)), ("precision", C.c_int32), ("algorithm", C.c_int32),
        ("tile", C.c_int32 * 3),
    ]


class PotDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int32 * 3), ("delta", C.c_double * 3), ("depth", C.c_double),
        ("lbound", C.c_double * 2), ("ubound", C.c_double * 2), ("npoints", C.c_int32 * 2),
        ("precision", C.c_int32),
    ]


class AcfGenDesc(C.Structure):
    """Mirror of arma_acf_gen_desc (the ACF_generator input block)."""
    _fields_ = [
        ("func", C.c_int32), ("amplitude", C.c_double), ("velocity", C.c_double),
        ("beta", C.c_double * 2), ("alpha", C.c_double * 3), ("nwaves", C.c_double * 3),
        ("shape", C.c_int32 * 3),
    ]


WAVE_SIN, WAVE_COS, WAVE_STOKES, WAVE_STANDING = 0, 1, 2, 3


def acf_generate(func=WAVE_COS, amplitude=1.0, velocity=1.0, beta=(0.8, 0.0),
                 alpha=(0.06, 0.06, 0.06), nwaves=(1.5, 1.5, 1.5), shape=(16, 16, 16),
                 ctx=None):
    """ACF_generator (src/generator/acf_generator.cc:130-215): (acf, delta).
    Defaults are the reference's (src/generator/acf_generator.hh:25-34).
    With a Context the auto-covariance runs on its GPU
    (arma_hip_acf_generate_device), else on the host."""
    d = AcfGenDesc()
    d.func, d.amplitude, d.velocity = func, amplitude, velocity
    d.beta[:] = list(beta)
    d.alpha[:] = list(alpha)
    d.nwaves[:] = list(nwaves)
    d.shape[:] = [int(v) for v in shape]
    cap = 1 << 21
    out = np.empty(cap, np.float64)
    so = (C.c_int32 * 3)()
    de = (C.c_double * 3)()
    if ctx is None:
        check(lib().arma_hip_acf_generate(C.byref(d), out, cap, so, de))
    else:
        check(lib().arma_hip_acf_generate_device(ctx.h, C.byref(d), out, cap, so, de))
    s = tuple(so)
    return out[:int(np.prod(s))].reshape(s).copy(), tuple(de)


class NitDesc(C.Structure):
    """Mirror of arma_nit_desc (transform = nit { ... } of the reference)."""
    _fields_ = [
        ("distribution", C.c_int32), ("skewness", C.c_double), ("kurtosis", C.c_double),
        ("mean", C.c_double), ("stdev", C.c_double), ("alpha", C.c_double),
        ("interval", C.c_double * 2), ("absolute_error", C.c_double),
        ("max_iterations", C.c_int32),
    ]


NIT_GRAM_CHARLIER, NIT_SKEW_NORMAL = 0, 1


def nit_desc(distribution, skewness=0.0, kurtosis=3.0, mean=0.0, stdev=1.0, alpha=0.0,
             interval=(0.0, 1.0), absolute_error=1e-6, max_iterations=100):
    """Defaults are the reference's (src/nonlinear/nit_transform.hh:48-66)."""
    d = NitDesc()
    d.distribution = distribution
    d.skewness, d.kurtosis = skewness, kurtosis
    d.mean, d.stdev, d.alpha = mean, stdev, alpha
    d.interval[:] = [float(interval[0]), float(interval[1])]
    d.absolute_error = absolute_error
    d.max_iterations = max_iterations
    return d


class BatchConfig(C.Structure):
    _fields_ = [("tile_x", C.c_int32), ("schedule", C.c_int32), ("slots_per_cu", C.c_int32)]

    def as_tuple(self):
        return (self.tile_x, self.schedule, self.slots_per_cu)


class BatchDesc(C.Structure):
    _fields_ = [("n_surfaces", C.c_int32), ("surfaces", C.POINTER(ArDesc)),
                ("gamma0", C.POINTER(C.c_double)), ("history", C.POINTER(C.c_int32)),
                ("precision", C.c_int32),
                ("config", BatchConfig)]


class BatchTrial(C.Structure):
    _fields_ = [("config", BatchConfig), ("grid", C.c_int32), ("ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("max_batch", C.c_int32), ("hbm_budget", C.c_int64),
                ("precision", C.c_int32), ("batch", BatchConfig)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("surface", ArDesc), ("gamma0", C.c_double),
                ("want_potential", C.c_int32), ("potential", PotDesc), ("stream", C.c_int64),
                ("arrival", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("kind", C.c_int32),
                ("stream", C.c_int64), ("chunk", C.c_int64), ("zeta", C.c_void_p),
                ("phi", C.c_void_p), ("shape", C.c_int32 * 3), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("batch_size", C.c_int32),
                ("lane", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("in_use", C.c_int64), ("peak", C.c_int64), ("budget", C.c_int64),
                ("batches", C.c_int64), ("batched_requests", C.c_int64),
                ("completed", C.c_int64), ("pending", C.c_int64), ("busy_lanes", C.c_int32)]


def mt19937_params():
    """MT19937 written as a dynamic-creator parameter record."""
    return MtParams(aaa=0x9908B0DF, mm=397, nn=624, rr=31, ww=32, wmask=0xFFFFFFFF,
                    umask=0x80000000, lmask=0x7FFFFFFF, shift0=11, shift1=18,
                    shiftB=7, shiftC=15, maskB=0x9D2C5680, maskC=0xEFC60000)


class ArmaHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class PRNGError(ArmaHipError):
    """arma::PRNG_error: fewer MT records than streams."""


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.arma_hip_init.argtypes = [C.c_int, C.POINTER(vp)]
        L.arma_hip_destroy.argtypes = [vp]
        L.arma_hip_destroy.restype = None
        L.arma_hip_last_error.restype = C.c_char_p
        L.arma_hip_set_stream.argtypes = [vp, vp]
        L.arma_hip_get_stream.argtypes = [vp]
        L.arma_hip_get_stream.restype = vp
        L.arma_hip_synchronize.argtypes = [vp]
        L.arma_hip_white_noise.argtypes = [vp, C.POINTER(ArDesc), vp]
        L.arma_hip_ar_recurrence.argtypes = [vp, C.POINTER(ArDesc), vp]
        L.arma_hip_ar_generate.argtypes = [vp, C.POINTER(ArDesc), vp, C.c_int]
        L.arma_hip_ar_generate_noalloc.argtypes = [vp, C.POINTER(ArDesc), vp]
        L.arma_hip_rescale.argtypes = [vp, vp, C.POINTER(C.c_int32), C.c_int32, C.c_double,
                                       C.POINTER(C.c_double)]
        L.arma_hip_linear_potential.argtypes = [vp, C.POINTER(PotDesc), vp, vp, C.c_int, vp]
        L.arma_hip_high_amplitude_potential.argtypes = L.arma_hip_linear_potential.argtypes
        L.arma_hip_mt_raw.argtypes = [vp, C.POINTER(MtParams), vp, C.c_int32, C.c_int32, vp]
        L.arma_hip_read_mt_config.argtypes = [C.c_char_p, C.c_int32, C.POINTER(MtParams),
                                              C.POINTER(C.c_int32)]
        L.arma_hip_write_mt_config.argtypes = [C.c_char_p, C.POINTER(MtParams), C.c_int32]
        L.arma_hip_dcmt_create.argtypes = [C.c_int32, C.c_int32, C.c_uint32, C.POINTER(MtParams)]
        L.arma_hip_dcmt_create_many.argtypes = [C.c_int32, C.c_int32, C.c_uint32, C.c_int32,
                                                C.POINTER(MtParams)]
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i3 = C.c_int32 * 3
        L.arma_hip_acf.argtypes = [C.c_int32, C.c_double * 3, i3, C.c_double, C.c_double,
                                   C.c_double * 3, C.c_double * 2, dp]
        L.arma_hip_yule_walker_choi.argtypes = [dp, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                                C.c_int32, C.c_double, dp, i3,
                                                C.POINTER(C.c_double)]
        L.arma_hip_yule_walker_gauss.argtypes = [dp, i3, i3, dp]
        L.arma_hip_white_noise_variance.argtypes = [dp, i3, dp, i3]
        L.arma_hip_white_noise_variance.restype = C.c_double
        L.arma_hip_partition_shape.argtypes = [i3, i3, i3, C.c_int32, i3]
        L.arma_hip_validate_process.argtypes = [dp, i3, C.POINTER(C.c_int32), vp, C.c_int32]
        L.arma_hip_acf_generate.argtypes = [C.POINTER(AcfGenDesc), dp, C.c_size_t,
                                            C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.arma_hip_acf_generate_device.argtypes = [vp, C.POINTER(AcfGenDesc), dp, C.c_size_t,
                                                   C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.arma_hip_nit_realisation.argtypes = [vp, vp, i3, C.c_int32, C.c_double,
                                                C.POINTER(NitDesc)]
        L.arma_hip_writer_open.argtypes = [vp, C.c_char_p, C.c_int32, C.c_size_t, C.POINTER(vp)]
        L.arma_hip_writer_append.argtypes = [vp, vp, C.c_size_t]
        L.arma_hip_writer_close.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        L.arma_hip_partition_shape.restype = None
        L.arma_hip_batch_create.argtypes = [vp, C.POINTER(BatchDesc), C.POINTER(vp)]
        L.arma_hip_batch_destroy.argtypes = [vp]
        L.arma_hip_batch_destroy.restype = None
        L.arma_hip_batch_layout.argtypes = [vp, C.POINTER(C.c_int64)]
        L.arma_hip_batch_generate.argtypes = [vp, vp, vp]
        L.arma_hip_batch_capture.argtypes = [vp, vp, vp]
        L.arma_hip_batch_replay.argtypes = [vp, vp]
        L.arma_hip_batch_tune.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(BatchTrial),
                                          C.c_int32, C.POINTER(C.c_int32)]
        L.arma_hip_batch_get_config.argtypes = [vp, C.POINTER(BatchConfig),
                                                C.POINTER(C.c_int32)]
        L.arma_hip_batch_set_config.argtypes = [vp, vp, C.POINTER(BatchConfig)]
        L.arma_hip_batch_factors.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.arma_hip_server_clock.argtypes = []
        L.arma_hip_server_clock.restype = C.c_double
        L.arma_hip_server_create.argtypes = [C.c_int, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.arma_hip_server_destroy.argtypes = [vp]
        L.arma_hip_server_destroy.restype = None
        L.arma_hip_server_open_stream.argtypes = [vp, C.POINTER(ArDesc), C.c_int32,
                                                  C.POINTER(C.c_int64)]
        L.arma_hip_server_close_stream.argtypes = [vp, C.c_int64]
        L.arma_hip_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
        L.arma_hip_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_int32,
                                           C.POINTER(C.c_int32), C.c_int32]
        L.arma_hip_server_release.argtypes = [vp, C.c_int64]
        L.arma_hip_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.arma_hip_server_copy.argtypes = [vp, vp, vp, C.c_size_t]
        L.arma_hip_yule_walker_choi_device.argtypes = [vp, dp, C.c_int32, C.c_int32, C.c_double,
                                                       C.c_int32, C.c_int32, C.c_double, dp,
                                                       i3, C.POINTER(C.c_double)]
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.arma_hip_yule_walker_choi_device_many.argtypes = [
            vp, C.c_int32, dp, i64p, i32p, i32p, C.c_double, C.c_int32, C.c_int32, C.c_double,
            dp, i64p, i32p, dp, i32p]
        L.arma_hip_white_noise_variance_device.argtypes = [vp, dp, i3, dp, i3,
                                                           C.POINTER(C.c_double)]
        L.arma_hip_profile.argtypes = [vp, C.c_int32]
        L.arma_hip_profile_report.argtypes = [vp, C.c_char_p, C.c_size_t, C.c_int32]
        _lib = L
    return _lib


def check(rc):
    if rc != ARMA_HIP_OK:
        msg = lib().arma_hip_last_error().decode(errors="replace")
        if rc == ARMA_HIP_EPRNG:
            raise PRNGError(rc, msg)
        if rc == ARMA_HIP_EINVAL:
            raise ValueError(f"{msg} (code {rc})")
        raise ArmaHipError(rc, msg)


def params_array(plist):
    arr = (MtParams * len(plist))()
    for i, p in enumerate(plist):
        arr[i] = p
    return arr


def dcmt_create(p, id_, seed):
    out = MtParams()
    check(lib().arma_hip_dcmt_create(p, id_, seed, C.byref(out)))
    return out


def dcmt_create_many(p, n, seed, nthreads=None):
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    arr = (MtParams * n)()
    check(lib().arma_hip_dcmt_create_many(p, n, seed, nthreads, arr))
    return arr


DCMT_TABLE = os.path.join(HERE, "data", "dcmt_p521_aaa_seed20240611.npy")
DCMT_TABLE_SEED = 20240611


def dcmt_p521_params(n, nthreads=None):
    """n dynamic-creator records for p = 521 (ids 0..n-1, creator seed
    20240611).  The first 4096 come from the shipped table of aaa values
    (every other field is fixed for p = 521, w = 32); more are created."""
    table = np.load(DCMT_TABLE) if os.path.exists(DCMT_TABLE) else np.zeros(0, np.uint32)
    if n > table.size:
        return dcmt_create_many(521, n, DCMT_TABLE_SEED, nthreads)
    tmpl = dcmt_create(521, 0, DCMT_TABLE_SEED)
    arr = (MtParams * n)()
    for i in range(n):
        arr[i] = tmpl
        arr[i].aaa = int(table[i])
    return arr


def read_mt_config(path, n):
    arr = (MtParams * n)()
    got = C.c_int32()
    check(lib().arma_hip_read_mt_config(path.encode(), n, arr, C.byref(got)))
    return arr


def write_mt_config(path, arr):
    check(lib().arma_hip_write_mt_config(path.encode(), arr, len(arr)))


ACF_STANDING_WAVE, ACF_PROPAGATING_WAVE = 0, 1


def acf(func, shape, delta=(1.0, 1.0, 1.0), amplitude=1.0, velocity=1.0,
        alpha=(0.06, 0.06, 0.06), beta=(0.8, 0.0)):
    """Closed-form ACF (defaults of src/bits/acf_wrapper.hh:23-28)."""
    out = np.empty(tuple(int(s) for s in shape), np.float64)
    check(lib().arma_hip_acf(func, (C.c_double * 3)(*delta), (C.c_int32 * 3)(*shape),
                             amplitude, velocity, (C.c_double * 3)(*alpha),
                             (C.c_double * 2)(*beta), out.reshape(-1)))
    return out


def yule_walker_choi(acf_cube, max_order, var_eps=1e-6, determine_order=True, chop=True,
                     chop_eps=1e-10):
    """Choi coefficients; the AR model calls this with var_eps = 1e-6
    (src/generator/ar_model.cc:142-150).  Returns (phi, solver variance)."""
    a = np.ascontiguousarray(acf_cube, np.float64)
    n = a.shape[0]
    phi = np.zeros(max(max_order, 2) ** 3 + n ** 3, np.float64)
    order = (C.c_int32 * 3)()
    var = C.c_double()
    check(lib().arma_hip_yule_walker_choi(a.reshape(-1), n, max_order, var_eps,
                                          int(determine_order), int(chop), chop_eps, phi, order,
                                          C.byref(var)))
    shp = tuple(order)
    return phi[: int(np.prod(shp))].reshape(shp).copy(), var.value


def yule_walker_gauss(acf_cube, order):
    a = np.ascontiguousarray(acf_cube, np.float64)
    phi = np.zeros(int(np.prod(order)), np.float64)
    check(lib().arma_hip_yule_walker_gauss(a.reshape(-1), (C.c_int32 * 3)(*a.shape),
                                           (C.c_int32 * 3)(*order), phi))
    return phi.reshape(tuple(order))


def white_noise_variance(acf_cube, phi):
    a = np.ascontiguousarray(acf_cube, np.float64)
    p = np.ascontiguousarray(phi, np.float64)
    return lib().arma_hip_white_noise_variance(a.reshape(-1), (C.c_int32 * 3)(*a.shape),
                                               p.reshape(-1), (C.c_int32 * 3)(*p.shape))


def validate_process(phi):
    """validate_process (src/arma.cc:15-70): (number of roots of
    1 - phi_1 x - ... inside or on the unit circle, the roots).  Raises
    ArmaHipError (runtime error) as the reference does when that number is
    not zero or the roots cannot be found."""
    phi = np.ascontiguousarray(phi, np.float64)
    n = phi.size
    out = np.zeros(2 * max(n, 1), np.float64)
    nbad = C.c_int32()
    rc = lib().arma_hip_validate_process(phi.reshape(-1), (C.c_int32 * 3)(*phi.shape),
                                         C.byref(nbad), out.ctypes.data, out.size)
    roots = out[:2 * max(n - 1, 0)].reshape(-1, 2)
    roots = roots[:, 0] + 1j * roots[:, 1]
    if rc != ARMA_HIP_OK:
        err = ArmaHipError(rc, lib().arma_hip_last_error().decode(errors="replace"))
        err.n_bad = nbad.value
        err.roots = roots
        raise err
    return nbad.value, roots


class Writer:
    """Streams device arrays to a file in network byte order
    (arma_hip_writer_*): append() returns at once, the bytes go to the file
    while the context's stream goes on; close() waits and reports
    (bytes written, seconds from the data being ready to the last write)."""

    def __init__(self, ctx, path, element_size, slab_bytes=0):
        self.h = C.c_void_p()
        check(lib().arma_hip_writer_open(ctx.h, os.fsencode(path), element_size, slab_bytes,
                                         C.byref(self.h)))
        self._keep = []

    def append(self, data):
        """data: a device tensor that stays unchanged until close()."""
        self._keep.append(data)
        check(lib().arma_hip_writer_append(self.h, C.c_void_p(_ptr(data)),
                                           data.numel() * data.element_size()))

    def close(self):
        if not self.h:
            return None
        n, t = C.c_uint64(), C.c_double()
        h, self.h = self.h, C.c_void_p()
        try:
            check(lib().arma_hip_writer_close(h, C.byref(n), C.byref(t)))
        finally:
            self._keep = []
        return n.value, t.value


def partition_shape(partition, grid, order, parallelism):
    out = (C.c_int32 * 3)()
    lib().arma_hip_partition_shape((C.c_int32 * 3)(*partition), (C.c_int32 * 3)(*grid),
                                   (C.c_int32 * 3)(*order), parallelism, out)
    return tuple(out)


def _ptr(t):
    """Device or host address of a torch tensor / numpy array."""
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


class Context:
    """One arma_hip context: a device, a HIP stream and the device buffers."""

    def __init__(self, device=0, torch_stream=True):
        self.h = C.c_void_p()
        check(lib().arma_hip_init(device, C.byref(self.h)))
        self.device = device
        if torch_stream:
            # order the library's kernels with torch's work on this device
            import torch
            with torch.cuda.device(device):
                self.set_stream(torch.cuda.current_stream().cuda_stream)

    def close(self):
        if self.h:
            lib().arma_hip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        check(lib().arma_hip_set_stream(self.h, C.c_void_p(stream_handle)))

    def synchronize(self):
        check(lib().arma_hip_synchronize(self.h))

    @staticmethod
    def ar_desc(shape, phi, varwn, mts, seeds=None, noise_mode=NOISE_SEQUENTIAL,
                partition=(0, 0, 0), precision=FP64, algorithm=AR_PLAIN, tile=(0, 0, 0),
                noise_semantics=NOISE_DOUBLE):
        phi = np.ascontiguousarray(phi, np.float64)
        d = ArDesc()
        d.shape[:] = [int(s) for s in shape]
        d.phi = phi.ctypes.data_as(C.POINTER(C.c_double))
        d.phi_shape[:] = [int(s) for s in phi.shape]
        d.varwn = float(varwn)
        d.noise_mode = noise_mode
        d.noise_semantics = noise_semantics
        d.partition[:] = [int(s) for s in partition]
        marr = mts if isinstance(mts, C.Array) else params_array(list(mts))
        d.mts = C.cast(marr, C.POINTER(MtParams))
        d.n_mts = len(marr)
        sarr = None
        if seeds is not None:
            sarr = np.ascontiguousarray(seeds, np.uint32)
            d.seeds = sarr.ctypes.data_as(C.POINTER(C.c_uint32))
        d.precision = precision
        d.algorithm = algorithm
        d.tile[:] = [int(s) for s in tile]
        d._keep = (phi, marr, sarr)  # keep host buffers alive with the descriptor
        return d

    def white_noise(self, desc, eps):
        check(lib().arma_hip_white_noise(self.h, C.byref(desc), C.c_void_p(_ptr(eps))))

    def ar_recurrence(self, desc, zeta):
        check(lib().arma_hip_ar_recurrence(self.h, C.byref(desc), C.c_void_p(_ptr(zeta))))

    def ar_generate(self, desc, zeta, on_device=True):
        check(lib().arma_hip_ar_generate(self.h, C.byref(desc), C.c_void_p(_ptr(zeta)),
                                         1 if on_device else 0))

    def ar_generate_noalloc(self, desc, zeta):
        """arma_hip_ar_generate_noalloc: zeta is a device tensor; the working
        buffers come from the library's static device scratch."""
        check(lib().arma_hip_ar_generate_noalloc(self.h, C.byref(desc), C.c_void_p(_ptr(zeta))))

    def rescale(self, zeta, shape, precision, gamma0):
        f = C.c_double()
        s = (C.c_int32 * 3)(*[int(v) for v in shape])
        check(lib().arma_hip_rescale(self.h, C.c_void_p(_ptr(zeta)), s, precision, gamma0,
                                     C.byref(f)))
        return f.value

    @staticmethod
    def pot_desc(shape, delta, depth, lbound, ubound, npoints, precision=FP64):
        d = PotDesc()
        d.shape[:] = [int(s) for s in shape]
        d.delta[:] = [float(v) for v in delta]
        d.depth = float(depth)
        d.lbound[:] = [float(v) for v in lbound]
        d.ubound[:] = [float(v) for v in ubound]
        d.npoints[:] = [int(v) for v in npoints]
        d.precision = precision
        return d

    def linear_potential(self, desc, zeta, phi, on_device=True, want_wnmax=False,
                         solver="linear"):
        """Velocity potential of zeta into phi (nt_out, nz, nx, ny); solver is
        "linear" or "high_amplitude" (the reference's registry keys)."""
        fn = {"linear": lib().arma_hip_linear_potential,
              "high_amplitude": lib().arma_hip_high_amplitude_potential}[solver]
        wn = None
        if want_wnmax:
            wn = np.empty(2 * desc.npoints[0], np.float64)
        check(fn(self.h, C.byref(desc), C.c_void_p(_ptr(zeta)), C.c_void_p(_ptr(phi)),
                 1 if on_device else 0, None if wn is None else wn.ctypes.data))
        return None if wn is None else wn.reshape(-1, 2)

    def high_amplitude_potential(self, desc, zeta, phi, on_device=True, want_wnmax=False):
        return self.linear_potential(desc, zeta, phi, on_device, want_wnmax,
                                     solver="high_amplitude")

    def nit_realisation(self, zeta, shape, precision, gamma0, desc):
        """transform = nit on the device surface zeta, in place."""
        check(lib().arma_hip_nit_realisation(self.h, C.c_void_p(_ptr(zeta)),
                                             (C.c_int32 * 3)(*shape), precision, gamma0,
                                             C.byref(desc)))

    def mt_raw(self, mts, seeds, draws, out):
        marr = mts if isinstance(mts, C.Array) else params_array(list(mts))
        sarr = None if seeds is None else np.ascontiguousarray(seeds, np.uint32)
        check(lib().arma_hip_mt_raw(self.h, C.cast(marr, C.POINTER(MtParams)),
                                    None if sarr is None else sarr.ctypes.data, len(marr), draws,
                                    C.c_void_p(_ptr(out))))

    def yule_walker_choi(self, acf_cube, max_order, var_eps=1e-6, determine_order=True,
                         chop=True, chop_eps=1e-10):
        """Choi coefficients in FP64 on this context's GPU: (phi, solver variance)."""
        acf_cube = np.ascontiguousarray(acf_cube, np.float64)
        n = acf_cube.shape[0]
        out = np.zeros(max(max_order, 2) ** 3 + n ** 3, np.float64)
        order = (C.c_int32 * 3)()
        var = C.c_double()
        check(lib().arma_hip_yule_walker_choi_device(
            self.h, acf_cube.reshape(-1), n, max_order, var_eps, 1 if determine_order else 0,
            1 if chop else 0, chop_eps, out, order, C.byref(var)))
        shape = tuple(order)
        return out[:int(np.prod(shape))].reshape(shape).copy(), var.value

    def yule_walker_choi_many(self, acfs, max_orders, var_eps=1e-6, determine_order=True,
                              chop=True, chop_eps=1e-10):
        """Choi coefficients of several models in one GPU launch (one
        workgroup per model): a list of (phi, solver variance), or None for a
        model whose recursion failed."""
        cubes = [np.ascontiguousarray(a, np.float64) for a in acfs]
        ns = np.array([c.shape[0] for c in cubes], np.int32)
        mo = np.array([int(m) for m in max_orders], np.int32)
        acf_off = np.cumsum([0] + [c.size for c in cubes[:-1]]).astype(np.int64)
        cap = [max(int(m), 2) ** 3 + int(n) ** 3 for m, n in zip(mo, ns)]
        phi_off = np.cumsum([0] + cap[:-1]).astype(np.int64)
        flat = np.concatenate([c.reshape(-1) for c in cubes]) if cubes else np.zeros(1)
        out = np.zeros(max(1, sum(cap)), np.float64)
        order = np.zeros(3 * len(cubes), np.int32)
        var = np.zeros(len(cubes), np.float64)
        status = np.zeros(len(cubes), np.int32)
        check(lib().arma_hip_yule_walker_choi_device_many(
            self.h, len(cubes), flat, acf_off, ns, mo, var_eps, 1 if determine_order else 0,
            1 if chop else 0, chop_eps, out, phi_off, order, var, status))
        res = []
        for i in range(len(cubes)):
            if status[i] != 0:
                res.append(None)
                continue
            shp = tuple(int(v) for v in order[3 * i:3 * i + 3])
            o = int(phi_off[i])
            res.append((out[o:o + int(np.prod(shp))].reshape(shp).copy(), float(var[i])))
        return res

    def white_noise_variance(self, acf_cube, phi):
        """AR white-noise variance in FP64 on this context's GPU."""
        acf_cube = np.ascontiguousarray(acf_cube, np.float64)
        phi = np.ascontiguousarray(phi, np.float64)
        out = C.c_double()
        check(lib().arma_hip_white_noise_variance_device(
            self.h, acf_cube.reshape(-1), (C.c_int32 * 3)(*acf_cube.shape), phi.reshape(-1),
            (C.c_int32 * 3)(*phi.shape), C.byref(out)))
        return out.value

    def batch(self, descs, gamma0=None, precision=FP32, config=(0, 0, 0), history=None):
        return Batch(self, descs, gamma0, precision, config, history)


class Batch:
    """Many surfaces generated by one persistent recurrence launch
    (arma_hip_batch_*).  descs are ar_desc() descriptors; surface i lives at
    arena[offsets[i] : offsets[i] + prod(shape_i)]."""

    def __init__(self, ctx, descs, gamma0=None, precision=FP32, config=(0, 0, 0),
                 history=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._descs = list(descs)
        arr = (ArDesc * len(self._descs))(*self._descs)
        g = None
        if gamma0 is not None:
            g = np.ascontiguousarray(gamma0, np.float64)
        bd = BatchDesc()
        bd.n_surfaces = len(self._descs)
        bd.surfaces = C.cast(arr, C.POINTER(ArDesc))
        bd.gamma0 = None if g is None else g.ctypes.data_as(C.POINTER(C.c_double))
        h = None
        if history is not None:
            h = np.ascontiguousarray(history, np.int32)
            bd.history = h.ctypes.data_as(C.POINTER(C.c_int32))
        bd.precision = precision
        bd.config = BatchConfig(*config)
        check(lib().arma_hip_batch_create(ctx.h, C.byref(bd), C.byref(self.h)))
        self.precision = precision
        self.shapes = [tuple(d.shape) for d in self._descs]
        offs = (C.c_int64 * (len(self._descs) + 1))()
        check(lib().arma_hip_batch_layout(self.h, offs))
        self.offsets = [int(v) for v in offs]

    @property
    def arena_elements(self):
        return self.offsets[-1]

    def surface(self, arena, i):
        n = int(np.prod(self.shapes[i]))
        return arena[self.offsets[i]:self.offsets[i] + n].view(self.shapes[i])

    def generate(self, arena):
        check(lib().arma_hip_batch_generate(self.ctx.h, self.h, C.c_void_p(_ptr(arena))))

    def capture(self, arena):
        check(lib().arma_hip_batch_capture(self.ctx.h, self.h, C.c_void_p(_ptr(arena))))

    def replay(self):
        check(lib().arma_hip_batch_replay(self.ctx.h, self.h))

    def tune(self, arena, reps=2):
        trials = (BatchTrial * 64)()
        n = C.c_int32()
        check(lib().arma_hip_batch_tune(self.ctx.h, self.h, C.c_void_p(_ptr(arena)), reps,
                                        trials, 64, C.byref(n)))
        return [(trials[i].config.as_tuple(), trials[i].grid, trials[i].ms)
                for i in range(n.value)]

    def config(self):
        c = BatchConfig()
        g = C.c_int32()
        check(lib().arma_hip_batch_get_config(self.h, C.byref(c), C.byref(g)))
        return c.as_tuple(), g.value

    def set_config(self, config):
        c = BatchConfig(*config)
        check(lib().arma_hip_batch_set_config(self.ctx.h, self.h, C.byref(c)))

    def factors(self):
        out = np.zeros(len(self._descs), np.float64)
        check(lib().arma_hip_batch_factors(self.ctx.h, self.h,
                                           out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def close(self):
        if self.h:
            lib().arma_hip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (arma_hip_server_*)."""

    def __init__(self, device=0, lanes=2, budget=1 << 30, precision=FP32, max_batch=0,
                 config=(0, 0, 0)):
        self.h = C.c_void_p()
        cfg = ServerConfig(lanes, max_batch, int(budget), precision, BatchConfig(*config))
        check(lib().arma_hip_server_create(device, C.byref(cfg), C.byref(self.h)))
        self.precision = precision

    @staticmethod
    def clock():
        return lib().arma_hip_server_clock()

    def open_stream(self, desc, chunk):
        sid = C.c_int64()
        check(lib().arma_hip_server_open_stream(self.h, C.byref(desc), chunk, C.byref(sid)))
        return sid.value

    def close_stream(self, sid):
        check(lib().arma_hip_server_close_stream(self.h, sid))

    def submit_surface(self, desc, gamma0=0.0, pot=None, arrival=0.0):
        r = Request()
        r.kind = REQ_SURFACE
        r.surface = desc
        r.gamma0 = gamma0
        if pot is not None:
            r.want_potential = 1
            r.potential = pot
        r.arrival = arrival
        rid = C.c_int64()
        check(lib().arma_hip_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def submit_chunk(self, sid, arrival=0.0):
        r = Request()
        r.kind = REQ_STREAM_CHUNK
        r.stream = sid
        r.arrival = arrival
        rid = C.c_int64()
        check(lib().arma_hip_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def poll(self, max_n=64, wait_ms=-1):
        out = (Completion * max_n)()
        n = C.c_int32()
        check(lib().arma_hip_server_poll(self.h, out, max_n, C.byref(n), wait_ms))
        return [out[i] for i in range(n.value)]

    def release(self, rid):
        check(lib().arma_hip_server_release(self.h, rid))

    def copy_out(self, ptr, shape, dtype=np.float32):
        """Host copy (numpy) of a device result the server owns."""
        out = np.empty([int(v) for v in shape], dtype)
        check(lib().arma_hip_server_copy(self.h, out.ctypes.data, C.c_void_p(ptr), out.nbytes))
        return out

    def stats(self):
        st = ServerStats()
        check(lib().arma_hip_server_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().arma_hip_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
