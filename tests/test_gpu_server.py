"""Serving: continuous batching of surface, potential and realtime-stream
requests (arma_hip_server_*).

Bar: every result of a mixed, overlapping request stream is bit-identical to
the same request made alone through the single-surface entry points
(surfaces and their rescale through arma_hip_ar_generate + arma_hip_rescale,
potentials through arma_hip_linear_potential), and the chunks of a realtime
stream concatenate bit-exactly to the one-shot surface with the equivalent
partition and seeds.  Admission never exceeds the HBM budget.
"""
import math
import time

import numpy as np
import pytest
import torch

import arma_amd
from arma_amd import _abi
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
PREC = arma_amd.FP32


def model(kind, vel, beta, alpha, order):
    acf = O.acf(kind, (10, 10, 10), velocity=vel, beta=beta, alpha=(alpha,) * 3)
    phi, _ = O.yw_choi(acf, order)
    return acf, phi, O.white_noise_variance(acf, phi)


@pytest.fixture(scope="module")
def setup():
    params = _abi.dcmt_p521_params(64)
    models = [model(O.STANDING, 0.8, (0.8, 0.8), 0.06, 3),
              model(O.STANDING, 0.6, (0.9, 0.7), 0.08, 5),
              model(O.PROPAGATING, 1.1, (-1.1, 0.0), 0.06, 7)]
    return params, models


# surface requests: shape, model index, noise partition, potential (nz or 0)
SURFACES = [
    ((40, 37, 50), 0, (16, 16, 16), 0),
    ((64, 20, 33), 1, (32, 16, 16), 0),
    ((17, 70, 16), 2, (8, 32, 8), 0),
    ((33, 32, 32), 0, (33, 32, 32), 3),
    ((50, 45, 41), 1, (25, 45, 41), 0),
    ((9, 8, 70), 2, (9, 8, 35), 0),
    ((24, 64, 64), 1, (12, 32, 32), 1),
    ((70, 16, 32), 0, (35, 16, 16), 2),
]
# realtime streams: (nx, ny), model index, column partition (px, py), chunks
STREAMS = [((24, 40), 2, (12, 20), 3), ((32, 16), 1, (32, 16), 3)]
CHUNK = 16


def surface_desc(ctx, params, models, shape, mi, part, seed):
    acf, phi, varwn = models[mi]
    n = math.prod(-(-shape[k] // part[k]) for k in range(3))
    return ctx.ar_desc(shape, phi, varwn, [params[j] for j in range(n)],
                       seeds=np.full(n, seed, np.uint32), noise_mode=arma_amd.NOISE_PARTITIONED,
                       partition=part, precision=PREC, algorithm=arma_amd.AR_WAVEFRONT)


def pot_desc(ctx, shape, nz):
    return ctx.pot_desc(shape, (2 * math.pi,) * 3, 12.0, (0.0, -10.0),
                        (float(shape[0] - 1), 0.0), (shape[0], nz), PREC)


def stream_desc(ctx, params, models, nxy, mi, cols, seed):
    acf, phi, varwn = models[mi]
    ncol = (-(-nxy[0] // cols[0])) * (-(-nxy[1] // cols[1]))
    return ctx.ar_desc((1,) + tuple(nxy), phi, varwn, [params[j] for j in range(ncol)],
                       seeds=[seed], noise_mode=arma_amd.NOISE_PARTITIONED,
                       partition=(CHUNK,) + tuple(cols), precision=PREC,
                       algorithm=arma_amd.AR_WAVEFRONT), ncol


def test_server_mix_equals_alone(hip_ctx, setup):
    params, models = setup
    srv = arma_amd.Server(0, lanes=3, budget=2 << 30, precision=PREC)
    try:
        sids, ids = [], {}
        for si, (nxy, mi, cols, nchunks) in enumerate(STREAMS):
            d, _ = stream_desc(hip_ctx, params, models, nxy, mi, cols, 1000 + 100 * si)
            sids.append(srv.open_stream(d, CHUNK))
        # everything arrives at once, streams interleaved with surfaces
        for i, (shape, mi, part, nz) in enumerate(SURFACES):
            d = surface_desc(hip_ctx, params, models, shape, mi, part, 7 + i)
            pd = pot_desc(hip_ctx, shape, nz) if nz else None
            ids[srv.submit_surface(d, float(models[mi][0][0, 0, 0]), pd)] = ("s", i)
            if i < 6:
                si = i % len(STREAMS)
                ids[srv.submit_chunk(sids[si])] = ("c", si)
        print("submitted", len(ids), flush=True)
        got, sizes = {}, []
        while len(got) < len(ids):
            for c in srv.poll(64, wait_ms=60000):
                assert c.status == 0, c.status
                shape = tuple(c.shape)
                z = srv.copy_out(c.zeta, shape)
                ph = None
                if c.phi:
                    nz = SURFACES[ids[c.id][1]][3]
                    ph = srv.copy_out(c.phi, (shape[0], nz, shape[1], shape[2]))
                got[c.id] = (z, ph, c.chunk)
                sizes.append(c.batch_size)
                assert c.t_arrival <= c.t_admit <= c.t_done
                srv.release(c.id)
        st = srv.stats()
        print("batch sizes", sizes, "stats", st, flush=True)
        assert st["peak"] <= st["budget"]
        assert max(sizes) > 1  # requests shared batches
        for sid in sids:
            srv.close_stream(sid)
    finally:
        srv.close()
    # alone: every surface (and potential) through the single-surface path
    for rid, (kind, i) in ids.items():
        if kind != "s":
            continue
        shape, mi, part, nz = SURFACES[i]
        d = surface_desc(hip_ctx, params, models, shape, mi, part, 7 + i)
        z = torch.empty(shape, dtype=torch.float32, device=DEV)
        hip_ctx.ar_generate(d, z)
        hip_ctx.rescale(z, shape, PREC, float(models[mi][0][0, 0, 0]))
        np.testing.assert_array_equal(got[rid][0], z.cpu().numpy(), err_msg=f"surface {i}")
        if nz:
            ph = torch.empty((shape[0], nz) + shape[1:], dtype=torch.float32, device=DEV)
            hip_ctx.linear_potential(pot_desc(hip_ctx, shape, nz), z, ph)
            np.testing.assert_array_equal(got[rid][1], ph.cpu().numpy(),
                                          err_msg=f"potential {i}")
    # streams: the chunks, in order, are the one-shot surface
    for si, (nxy, mi, cols, _) in enumerate(STREAMS):
        chunks = sorted((v[2], v[0]) for rid, v in got.items() if ids[rid] == ("c", si))
        assert [k for k, _ in chunks] == list(range(len(chunks)))
        whole = np.concatenate([z for _, z in chunks])
        nt = whole.shape[0]
        acf, phi, varwn = models[mi]
        _, ncol = stream_desc(hip_ctx, params, models, nxy, mi, cols, 0)
        nparts = (nt // CHUNK) * ncol
        d = hip_ctx.ar_desc((nt,) + tuple(nxy), phi, varwn,
                            [params[p % ncol] for p in range(nparts)],
                            seeds=np.arange(nparts, dtype=np.uint32) + 1000 + 100 * si,
                            noise_mode=arma_amd.NOISE_PARTITIONED,
                            partition=(CHUNK,) + tuple(cols), precision=PREC,
                            algorithm=arma_amd.AR_WAVEFRONT)
        z = torch.empty((nt,) + tuple(nxy), dtype=torch.float32, device=DEV)
        hip_ctx.ar_generate(d, z)
        np.testing.assert_array_equal(whole, z.cpu().numpy(), err_msg=f"stream {si}")


def test_server_budget_admission(hip_ctx, setup):
    params, models = setup
    shape = (48, 64, 64)
    d = surface_desc(hip_ctx, params, models, shape, 1, (24, 32, 32), 3)
    # room for about two requests at a time (each needs its 0.8 MB surface
    # and about as much working set while in flight)
    srv = arma_amd.Server(0, lanes=4, budget=4 << 20, precision=PREC)
    try:
        ids = [srv.submit_surface(d, 1.0) for _ in range(10)]
        seen = []
        while len(seen) < len(ids):
            for c in srv.poll(16, wait_ms=60000):
                assert c.status == 0
                seen.append(srv.copy_out(c.zeta, tuple(c.shape)))
                srv.release(c.id)
        st = srv.stats()
        assert st["peak"] <= st["budget"]
        assert st["completed"] == len(ids)
        # released memory returns to the budget once its lane has freed it
        deadline = time.time() + 10
        while srv.stats()["in_use"] and time.time() < deadline:
            time.sleep(0.01)
        assert srv.stats()["in_use"] == 0
        for z in seen[1:]:
            np.testing.assert_array_equal(z, seen[0])
        with pytest.raises(arma_amd.ArmaHipError):
            srv.submit_surface(surface_desc(hip_ctx, params, models, (512, 256, 256), 1,
                                            (512, 256, 256), 3), 1.0)
    finally:
        srv.close()


def test_server_stream_rules(hip_ctx, setup):
    params, models = setup
    srv = arma_amd.Server(0, lanes=1, budget=1 << 30, precision=PREC)
    try:
        d, _ = stream_desc(hip_ctx, params, models, (16, 16), 1, (16, 16), 5)
        sid = srv.open_stream(d, CHUNK)
        rid = srv.submit_chunk(sid)
        (c,) = srv.poll(1, wait_ms=60000)
        assert c.id == rid and c.status == 0 and tuple(c.shape) == (CHUNK, 16, 16)
        srv.release(rid)
        srv.close_stream(sid)
        with pytest.raises(ValueError):
            srv.submit_chunk(sid)
        d64 = hip_ctx.ar_desc((8, 8, 8), models[0][1], models[0][2], [params[0]],
                              precision=arma_amd.FP64)
        with pytest.raises(ValueError):
            srv.submit_surface(d64, 1.0)
        # the potential needs power-of-two (x, y) sizes: refused at submit
        d = surface_desc(hip_ctx, params, models, (20, 20, 33), 1, (20, 20, 33), 5)
        with pytest.raises(ValueError):
            srv.submit_surface(d, 1.0, pot_desc(hip_ctx, (20, 20, 33), 1))
    finally:
        srv.close()


def test_server_lanes_beyond_residency(hip_ctx, setup):
    """Lanes whose batches each hold more tiles than one recurrence launch
    keeps resident: the launches run one after another on the device (two
    partly resident persistent launches could wait on each other's
    unscheduled workgroups), and every result equals the surface made alone."""
    params, models = setup
    shape, part = (24, 640, 640), (24, 160, 160)
    descs = [surface_desc(hip_ctx, params, models, shape, 1, part, 50 + i) for i in range(4)]
    gamma0 = float(models[1][0][0, 0, 0])
    probe = arma_amd.Batch(hip_ctx, descs[:1], [gamma0], PREC, (16, 1, 0))
    try:
        _, grid = probe.config()
    finally:
        probe.close()
    tiles = (640 // 16) ** 2
    print(f"tiles per batch {tiles}, resident workgroups per launch {grid}", flush=True)
    assert grid < tiles
    srv = arma_amd.Server(0, lanes=3, budget=8 << 30, precision=PREC, max_batch=1)
    try:
        ids = {srv.submit_surface(d, gamma0): i for i, d in enumerate(descs)}
        got = {}
        while len(got) < len(ids):
            for c in srv.poll(8, wait_ms=60000):
                assert c.status == 0, c.status
                assert c.batch_size == 1
                got[ids[c.id]] = srv.copy_out(c.zeta, tuple(c.shape))
                srv.release(c.id)
    finally:
        srv.close()
    for i, d in enumerate(descs):
        z = torch.empty(shape, dtype=torch.float32, device=DEV)
        hip_ctx.ar_generate(d, z)
        hip_ctx.rescale(z, shape, PREC, gamma0)
        np.testing.assert_array_equal(got[i], z.cpu().numpy(), err_msg=f"surface {i}")
