"""MI355X-native ARMA hot path: white noise, 3-D AR surface recurrence and
the linear and high-amplitude FFT velocity potentials, as hand-written HIP kernels behind the C
ABI of include/arma_hip.h (libarma_hip.so).

The C++ drop-in for the reference's generator and solver registry lives in
arma_amd/host/; this Python package is a thin ctypes binding used by tests
and benchmarks.
"""
from ._abi import (AR_PLAIN, AR_WAVEFRONT, BATCH_AUTO, BATCH_INTERLEAVED,  # noqa: F401
                   BATCH_SURFACE_MAJOR, FP32, FP64, NOISE_DOUBLE, NOISE_FLOAT,
                   NIT_GRAM_CHARLIER, NIT_SKEW_NORMAL, NOISE_PARTITIONED, NOISE_SEQUENTIAL,
                   ArmaHipError, nit_desc, Batch, Context, MtParams,
                   Server, Writer,
                   PRNGError, dcmt_create, dcmt_create_many, lib, mt19937_params,
                   read_mt_config, write_mt_config)
