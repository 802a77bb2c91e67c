"""Regenerates the committed golden fixtures under tests/golden/.

* libstdcxx_random.npz -- std::mt19937 and std::normal_distribution
  vectors from the g++ toolchain (gen_libstdcxx.cc).
* numpy_fft.npz -- numpy.fft.fft2 / ifft2 of seeded complex inputs, the
  independent check of the GSL transform conventions (forward
  exp(-2 pi i jk/n), backward unnormalised exp(+2 pi i jk/n)).

Run:  python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build_libstdcxx():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "gen")
        subprocess.check_call(["g++", "-O2", "-std=c++14", "-o", exe,
                               os.path.join(HERE, "gen_libstdcxx.cc")])
        blob = subprocess.check_output([exe])
    off = 0

    def take(dtype):
        nonlocal off
        (n,) = struct.unpack_from("<I", blob, off)
        off += 4
        arr = np.frombuffer(blob, dtype=dtype, count=n, offset=off).copy()
        off += n * np.dtype(dtype).itemsize
        return arr

    seeds = [0, 1, 42, 5489]
    out = {"seeds": np.array(seeds, dtype=np.uint32), "sigmas": np.array([1.0, 0.37])}
    for s in seeds:
        out[f"mt19937_{s}"] = take(np.uint32)
    out["mt19937_default_10000th"] = take(np.uint32)
    for s in seeds:
        for si, _ in enumerate([1.0, 0.37]):
            out[f"normal_d_{s}_{si}"] = take(np.float64)
            out[f"normal_f_{s}_{si}"] = take(np.float32)
    assert off == len(blob)
    np.savez_compressed(os.path.join(HERE, "libstdcxx_random.npz"), **out)


def build_fft():
    rng = np.random.default_rng(20240611)
    out = {}
    for shape in [(16, 16), (8, 12), (5, 7)]:
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = f"{shape[0]}x{shape[1]}"
        out[f"in_{key}"] = x
        out[f"fwd_{key}"] = np.fft.fft2(x)
        out[f"bwd_{key}"] = np.fft.ifft2(x) * x.size  # unnormalised backward
    np.savez_compressed(os.path.join(HERE, "numpy_fft.npz"), **out)


if __name__ == "__main__":
    build_libstdcxx()
    build_fft()
    print("golden fixtures written to", HERE)
