"""Streaming writer (arma_hip_writer_*): device arrays to a file in network
byte order, the zeta.bin format of io::Binary_stream
(src/io/binary_stream.hh:43-70).

Bar: the file equals the appended arrays, bit-exact, read back as big-endian
values, for both element sizes, slabs smaller than, equal to and larger than
an append, several appends in order, and with other GPU work enqueued on the
stream right after each append (it must not disturb what is written).
"""
import numpy as np
import pytest
import torch

import arma_amd

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("dtype,be", [(torch.float64, ">f8"), (torch.float32, ">f4")])
@pytest.mark.parametrize("slab", [0, 4096, 1 << 20, 24])
def test_writer_round_trip(hip_ctx, tmp_path, dtype, be, slab):
    print(f"writer {dtype} slab {slab}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    parts = [torch.randn(n, generator=g, dtype=torch.float64).to(DEV, dtype)
             for n in (100003, 1, 0, 65536, 7)]
    path = tmp_path / "out.bin"
    w = arma_amd.Writer(hip_ctx, str(path), parts[0].element_size(), slab)
    for p in parts:
        w.append(p)
        # later work on the same stream into another buffer
        scratch = torch.empty(1 << 20, dtype=dtype, device=DEV)
        scratch.fill_(1.5)
    nbytes, secs = w.close()
    expect = np.concatenate([p.cpu().numpy() for p in parts])
    assert nbytes == expect.nbytes and secs >= 0.0
    got = np.fromfile(path, dtype=be)
    assert got.tobytes() == expect.astype(be).tobytes()
    np.testing.assert_array_equal(got.astype(expect.dtype), expect)


def test_writer_waits_for_the_producer(hip_ctx, tmp_path):
    """append() takes effect after the work enqueued before it: a surface
    generated on the same stream and appended at once is written whole."""
    shape = (64, 48, 40)
    acf_phi = np.zeros((2, 2, 2))
    acf_phi[1, 0, 0], acf_phi[0, 1, 0], acf_phi[0, 0, 1] = 0.3, 0.2, 0.1
    d = hip_ctx.ar_desc(shape, acf_phi, 1.0, [arma_amd.mt19937_params()], seeds=[5],
                        algorithm=arma_amd.AR_WAVEFRONT)
    z = torch.empty(shape, dtype=torch.float64, device=DEV)
    path = tmp_path / "zeta.bin"
    w = arma_amd.Writer(hip_ctx, str(path), 8, 1 << 16)
    hip_ctx.ar_generate(d, z)
    w.append(z)
    w.close()
    np.testing.assert_array_equal(np.fromfile(path, dtype=">f8").reshape(shape), z.cpu().numpy())


def test_writer_errors(hip_ctx, tmp_path):
    with pytest.raises(ValueError):
        arma_amd.Writer(hip_ctx, str(tmp_path / "x.bin"), 2)
    with pytest.raises(arma_amd.ArmaHipError):
        arma_amd.Writer(hip_ctx, str(tmp_path / "no_such_dir" / "x.bin"), 8)
    w = arma_amd.Writer(hip_ctx, str(tmp_path / "y.bin"), 8)
    with pytest.raises(ValueError):
        w.append(torch.zeros(3, dtype=torch.float32, device=DEV))   # 12 bytes, not 8-byte elements
    assert w.close()[0] == 0
    assert w.close() is None
