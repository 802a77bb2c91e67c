"""Batching: many surfaces in one persistent recurrence launch (arma_hip_batch_*).

Bar: every member of a batch is bit-identical to the same surface made
alone through the single-surface entry points (partitioned noise, wavefront
recurrence, rescale), for every tile shape, schedule and residency, and for
graph replays; the single-surface path is pinned to the oracle in
test_gpu_noise_ar.py, and one member is checked against the oracle here too.
"""
import numpy as np
import pytest
import torch

import arma_amd
from arma_amd import _abi
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

# shape, ACF (kind, velocity, beta, alpha), requested order, noise partition
MEMBERS = [
    ((40, 37, 50), (O.STANDING, 0.8, (0.8, 0.8), 0.06), 3, (16, 16, 16)),
    ((64, 20, 33), (O.STANDING, 0.6, (0.9, 0.7), 0.08), 5, (32, 16, 16)),
    ((17, 70, 16), (O.PROPAGATING, 1.1, (-1.1, 0.0), 0.06), 7, (8, 32, 8)),
    ((30, 16, 16), (O.STANDING, 1.0, (0.5, 1.2), 0.05), 9, (30, 16, 16)),
    ((50, 45, 41), (O.PROPAGATING, 0.5, (-0.5, 0.0), 0.09), 4, (25, 45, 41)),
    ((9, 8, 70), (O.STANDING, 0.8, (0.8, 0.8), 0.06), 2, (9, 8, 35)),
]


def model(acf_spec, order):
    kind, vel, beta, alpha = acf_spec
    acf = O.acf(kind, (10, 10, 10), velocity=vel, beta=beta, alpha=(alpha,) * 3)
    phi, _ = O.yw_choi(acf, order)
    return acf, phi, O.white_noise_variance(acf, phi)


@pytest.fixture(scope="module")
def members():
    params = _abi.dcmt_p521_params(64)
    out = []
    for i, (shape, spec, order, part) in enumerate(MEMBERS):
        acf, phi, varwn = model(spec, order)
        nparts = int(np.prod([-(-shape[k] // part[k]) for k in range(3)]))
        seeds = np.full(nparts, 42 + i, np.uint32)
        out.append(dict(shape=shape, phi=phi, varwn=varwn, gamma0=float(acf[0, 0, 0]),
                        part=part, plist=[params[j] for j in range(nparts)], seeds=seeds))
    return out


def descs(ctx, members, prec):
    return [ctx.ar_desc(m["shape"], m["phi"], m["varwn"], m["plist"], seeds=m["seeds"],
                        noise_mode=arma_amd.NOISE_PARTITIONED, partition=m["part"],
                        precision=prec, algorithm=arma_amd.AR_WAVEFRONT) for m in members]


def singles(ctx, members, prec):
    dt = torch.float32 if prec == arma_amd.FP32 else torch.float64
    zs, fs = [], []
    for d, m in zip(descs(ctx, members, prec), members):
        z = torch.empty(m["shape"], dtype=dt, device=DEV)
        ctx.ar_generate(d, z)
        fs.append(ctx.rescale(z, m["shape"], prec, m["gamma0"]))
        zs.append(z.cpu().numpy())
    return zs, fs


def assert_batch_equal(b, arena, zs, fs=None):
    torch.cuda.synchronize()
    b.ctx.synchronize()
    for i, z in enumerate(zs):
        got = b.surface(arena, i).cpu().numpy()
        np.testing.assert_array_equal(got, z, err_msg=f"member {i}")
    if fs is not None:
        np.testing.assert_array_equal(b.factors(), np.array(fs))


@pytest.mark.parametrize("prec", [arma_amd.FP32, arma_amd.FP64], ids=["fp32", "fp64"])
def test_batch_matches_single_surfaces(hip_ctx, members, prec):
    zs, fs = singles(hip_ctx, members, prec)
    dt = torch.float32 if prec == arma_amd.FP32 else torch.float64
    b = hip_ctx.batch(descs(hip_ctx, members, prec), [m["gamma0"] for m in members], prec)
    assert b.config()[0] == (16, arma_amd.BATCH_SURFACE_MAJOR, 0)
    arena = torch.full((b.arena_elements,), float("nan"), dtype=dt, device=DEV)
    torch.cuda.synchronize()
    print(f"batch {prec} default config, {len(members)} members", flush=True)
    b.generate(arena)
    assert_batch_equal(b, arena, zs, fs)
    for cfg in [(8, arma_amd.BATCH_INTERLEAVED, 0), (32, arma_amd.BATCH_SURFACE_MAJOR, 1),
                (16, arma_amd.BATCH_INTERLEAVED, 2), (8, arma_amd.BATCH_SURFACE_MAJOR, 1),
                (32, arma_amd.BATCH_INTERLEAVED, 0)]:
        print(f"batch {prec} config {cfg}", flush=True)
        b.set_config(cfg)
        assert b.config()[0] == cfg
        arena.fill_(float("nan"))
        torch.cuda.synchronize()
        b.generate(arena)
        assert_batch_equal(b, arena, zs, fs)
    b.close()


def test_batch_matches_single_surfaces_quad_tile(hip_ctx, members, monkeypatch):
    """ARMA_HIP_QUAD_TILE=1 puts FP32 surfaces, single and batched, on the
    four-lanes-per-column tile: members still equal their single surfaces,
    and both stay within the FP32 bar of the one-lane tile's result."""
    prec = arma_amd.FP32
    ref, _ = singles(hip_ctx, members, prec)
    monkeypatch.setenv("ARMA_HIP_QUAD_TILE", "1")
    zs, fs = singles(hip_ctx, members, prec)
    b = hip_ctx.batch(descs(hip_ctx, members, prec), [m["gamma0"] for m in members], prec)
    arena = torch.full((b.arena_elements,), float("nan"), dtype=torch.float32, device=DEV)
    torch.cuda.synchronize()
    print(f"batch quad tile, {len(members)} members", flush=True)
    b.generate(arena)
    assert_batch_equal(b, arena, zs, fs)
    b.close()
    for z, r in zip(zs, ref):
        scale = float(np.max(np.abs(r)))
        err = np.abs(z.astype(np.float64) - r.astype(np.float64)) / scale
        assert float(np.sqrt(np.mean(err ** 2))) <= 1e-6 and float(np.max(err)) <= 2e-5


def test_batch_member_matches_oracle(hip_ctx, members):
    """Member 1 (64x20x33, order 5) against the FP64 oracle directly."""
    m = members[1]
    b = hip_ctx.batch(descs(hip_ctx, members, arma_amd.FP64), [x["gamma0"] for x in members],
                      arma_amd.FP64, config=(32, arma_amd.BATCH_INTERLEAVED, 0))
    arena = torch.empty(b.arena_elements, dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    b.generate(arena)
    hip_ctx.synchronize()
    got = b.surface(arena, 1).cpu().numpy()
    eps = O.noise_partitioned([O.MtParams(*p.as_tuple()) for p in m["plist"]], m["seeds"],
                              m["shape"], m["part"], np.sqrt(m["varwn"]))
    ref = O.ar_generate(eps, m["phi"])
    f = O.rescale_factor(ref, m["gamma0"])
    assert b.factors()[1] == pytest.approx(f, rel=1e-12)
    err = float(np.max(np.abs(got - ref * f)) / np.max(np.abs(ref * f)))
    assert err <= 1e-12, err
    b.close()


def test_batch_graph_replay(members):
    """Capture generate into a HIP graph and replay it three times with other
    work on the stream in between; every replay equals the eager result."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = arma_amd.Context(0, torch_stream=False)
    try:
        prec = arma_amd.FP32
        b = ctx.batch(descs(ctx, members, prec), [m["gamma0"] for m in members], prec)
        eager = torch.empty(b.arena_elements, dtype=torch.float32, device=DEV)
        torch.cuda.synchronize()
        b.generate(eager)
        ctx.synchronize()
        ref = [b.surface(eager, i).cpu().numpy() for i in range(len(members))]
        fref = b.factors()
        arena = torch.full((b.arena_elements,), float("nan"), dtype=torch.float32, device=DEV)
        torch.cuda.synchronize()
        b.capture(arena)
        for r in range(3):
            print(f"graph replay {r}", flush=True)
            b.replay()
            ctx.synchronize()
            for i, z in enumerate(ref):
                np.testing.assert_array_equal(b.surface(arena, i).cpu().numpy(), z,
                                              err_msg=f"replay {r} member {i}")
            np.testing.assert_array_equal(b.factors(), fref)
            # eager work between replays: a plain generate into another arena
            # and a torch kernel on the arena itself
            other = torch.empty_like(eager)
            torch.cuda.synchronize()
            b.generate(other)
            ctx.synchronize()
            arena.mul_(0.0)
            torch.cuda.synchronize()
        b.close()
    finally:
        ctx.close()


def test_batch_tune_and_cache(members):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = arma_amd.Context(0, torch_stream=False)
    try:
        prec = arma_amd.FP32
        ds = descs(ctx, members, prec)
        g0 = [m["gamma0"] for m in members]
        b = ctx.batch(ds, g0, prec)
        arena = torch.empty(b.arena_elements, dtype=torch.float32, device=DEV)
        torch.cuda.synchronize()
        b.generate(arena)
        ctx.synchronize()
        ref = [b.surface(arena, i).cpu().numpy() for i in range(len(members))]
        trials = b.tune(arena, reps=1)
        print("tune trials", trials, flush=True)
        assert len(trials) >= 6
        cfg, grid = b.config()
        assert cfg in [t[0] for t in trials]
        best = min(trials, key=lambda t: t[2])
        assert cfg == best[0]
        assert grid > 0 and all(t[2] > 0 for t in trials)
        b.generate(arena)
        ctx.synchronize()
        for i, z in enumerate(ref):
            np.testing.assert_array_equal(b.surface(arena, i).cpu().numpy(), z)
        # an identical batch picks the tuned configuration up from the context
        b2 = ctx.batch(ds, g0, prec)
        assert b2.config()[0] == cfg
        b2.close()
        b.close()
    finally:
        ctx.close()


def test_batch_rejects_bad_members(hip_ctx, members):
    m = members[0]
    big = np.zeros((11, 2, 2))
    big[0, 0, 1] = 0.1
    d = hip_ctx.ar_desc(m["shape"], big, 1.0, m["plist"], seeds=m["seeds"],
                        noise_mode=arma_amd.NOISE_PARTITIONED, partition=m["part"],
                        precision=arma_amd.FP32)
    with pytest.raises(ValueError):
        hip_ctx.batch([d], None, arma_amd.FP32)
    few = hip_ctx.ar_desc(m["shape"], m["phi"], m["varwn"], m["plist"][:2], seeds=m["seeds"][:2],
                          noise_mode=arma_amd.NOISE_PARTITIONED, partition=m["part"],
                          precision=arma_amd.FP32)
    with pytest.raises(arma_amd.PRNGError):
        hip_ctx.batch([few], None, arma_amd.FP32)
    with pytest.raises(ValueError):
        hip_ctx.batch(descs(hip_ctx, members[:1], arma_amd.FP32), None, arma_amd.FP32,
                      config=(12, 0, 0))
