"""The C++ drop-in driver (arma_amd/host/arma-hip): the reference's .arma
input syntax, registry keys "AR"/"linear"/"high_amplitude", model setup, and (on a GPU) the
whole generate -> potential -> write path against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "arma_amd", "host", "arma-hip")

INPUT = """# small standing-wave case
model = AR {
	out_grid = (%(nt)d,%(nx)d,%(ny)d)
	acf = {
		func = standing_wave
		grid = (10,10,10)
		velocity = 0.8
		beta = (0.8,0.8)
	}
	order = (7,7,7)
	no_seed = 1
	output = surface,%(fmt)s
	validate = 0
%(transform)s}
velocity_potential_solver = %(solver)s {
	depth = 12
	domain = from (0,-10) to (%(tlast)d,0) npoints (%(nt)d,%(nz)d)
}
"""


NIT = """	transform = nit {
		distribution = gram_charlier {
			skewness=0.4
			kurtosis=3.6
		}
		interpolation_nodes = 100
		max_interpolation_order = 10
		max_expansion_order = 20
		cdf_solver = {
			interval = [-5,5]
		}
		acf_solver = {
			interval = [-10,10]
		}
	}
"""


def write_input(tmp_path, nt=40, nx=32, ny=32, nz=4, fmt="binary", solver="linear",
                transform=""):
    p = tmp_path / "in.arma"
    p.write_text(INPUT % dict(nt=nt, nx=nx, ny=ny, nz=nz, fmt=fmt, tlast=nt - 1, solver=solver,
                              transform=transform))
    return str(p)


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_help():
    r = run(["-h"], ROOT)
    assert r.returncode == 0 and "usage" in r.stdout


def test_dry_run_model_setup(tmp_path):
    r = run(["-n", write_input(tmp_path)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    out = dict(line.split("=", 1) for line in r.stdout.splitlines())
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, 7)
    assert out["order"] == "(%d,%d,%d)" % phi.shape
    got = np.array([float(v) for v in out["phi"].split(",")]).reshape(phi.shape)
    # independent FP64 implementations: matrix products and inverses round differently
    np.testing.assert_allclose(got, phi, rtol=1e-9, atol=1e-10)
    assert float(out["varwn"]) == pytest.approx(O.white_noise_variance(acf, phi), rel=1e-12)
    # out_grid length rescaled to acf.delta * patches * 2 pi
    lengths = [float(v) for v in out["grid"].split(":")[1].strip("()").split(",")]
    np.testing.assert_allclose(lengths, [39 * 2 * np.pi, 31 * 2 * np.pi, 31 * 2 * np.pi])


@pytest.mark.parametrize("bad,msg", [
    ("order = (7,7,7)", "Unknown parameter"),
    ("func = standing_wave", "bad ACF function"),
])
def test_input_errors(tmp_path, bad, msg):
    text = open(write_input(tmp_path)).read()
    if msg == "Unknown parameter":
        text = text.replace(bad, "ordre = (7,7,7)")
    else:
        text = text.replace(bad, "func = triangle")
    (tmp_path / "in.arma").write_text(text)
    r = run(["-n", str(tmp_path / "in.arma")], str(tmp_path))
    assert r.returncode != 0 and msg in r.stderr


def test_unknown_solver(tmp_path):
    text = open(write_input(tmp_path)).read().replace("= linear {", "= small_amplitude {")
    (tmp_path / "in.arma").write_text(text)
    r = run(["-n", str(tmp_path / "in.arma")], str(tmp_path))
    assert r.returncode != 0 and "Invalid object" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["linear", "high_amplitude"])
def test_end_to_end_surface_and_potential(tmp_path, solver):
    """arma-hip on a 40x32x32 standing-wave case; zeta.bin is big-endian."""
    import arma_amd
    arma_amd.write_mt_config(str(tmp_path / "mt.dat"),
                             arma_amd._abi.params_array([arma_amd.mt19937_params()]))
    r = run([write_input(tmp_path, fmt="binary,csv", solver=solver)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    zeta = np.fromfile(tmp_path / "zeta.bin", dtype=">f8").reshape(40, 32, 32)
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, 7)
    varwn = O.white_noise_variance(acf, phi)
    eps = O.noise_sequential(O.mt19937_params(), 0, np.sqrt(varwn), 40 * 32 * 32).reshape(40, 32, 32)
    ref = O.ar_generate(eps, phi)
    ref *= O.rescale_factor(ref, acf[0, 0, 0])
    err = float(np.max(np.abs(zeta - ref)) / np.max(np.abs(ref)))
    assert err < 1e-10, err
    pot = np.loadtxt(tmp_path / "phi.csv", delimiter=",", skiprows=1)[:, 4].reshape(40, 4, 32, 32)
    delta = (2 * np.pi,) * 3
    pref = O.linear_potential(zeta, delta, 12.0, (0.0, -10.0), (39.0, 0.0), (40, 4),
                              solver=solver)
    perr = float(np.max(np.abs(pot - pref)) / np.max(np.abs(pref)))
    assert perr < 1e-9, perr


def test_nit_input_parses(tmp_path):
    r = run(["-n", write_input(tmp_path, transform=NIT)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    bad = NIT.replace("interval = [-5,5]", "interval = [5,-5]")
    r = run(["-n", write_input(tmp_path, transform=bad)], str(tmp_path))
    assert r.returncode != 0 and "bad parameter" in r.stderr


@pytest.mark.gpu
def test_end_to_end_nit(tmp_path):
    """transform = nit: the rescaled surface, then the Gram-Charlier transform."""
    import arma_amd
    arma_amd.write_mt_config(str(tmp_path / "mt.dat"),
                             arma_amd._abi.params_array([arma_amd.mt19937_params()]))
    r = run([write_input(tmp_path, transform=NIT)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    zeta = np.fromfile(tmp_path / "zeta.bin", dtype=">f8").reshape(40, 32, 32)
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, 7)
    varwn = O.white_noise_variance(acf, phi)
    eps = O.noise_sequential(O.mt19937_params(), 0, np.sqrt(varwn), 40 * 32 * 32).reshape(40, 32, 32)
    ref = O.ar_generate(eps, phi)
    ref *= O.rescale_factor(ref, acf[0, 0, 0])
    ref = O.nit_realisation(ref, np.sqrt(acf[0, 0, 0]), O.GRAM_CHARLIER, skewness=0.4,
                            kurtosis=3.6, interval=(-5.0, 5.0))
    assert np.mean(np.abs(zeta - ref) <= 1e-9) >= 0.99
    assert np.max(np.abs(zeta - ref)) <= 1e-4


HEAD_INPUT = """model = AR {
	out_grid = (40,32,32)
	acf = {
		func = standing_wave
		amplitude = 3
		alpha = (2,0.2,1)
		velocity = 0.50
		beta = (0.0625,0)
		nwaves = (1.85,16,1)
		shape = (8,8,8)
	}
	algorithm = choi_recursive
	order = (7,7,7)
	output = surface
	validate = 0
}
velocity_potential_solver = linear {
	depth = 12
	domain = from (10,-12) to (10,3) npoints (1,128)
}
"""


def test_dry_run_acf_generator(tmp_path):
    """An acf block without grid = selects ACF_generator (the reference's HEAD
    syntax, input/standing_wave.arma with a smaller shape)."""
    p = tmp_path / "head.arma"
    p.write_text(HEAD_INPUT)
    r = run(["-n", str(p)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    out = dict(line.split("=", 1) for line in r.stdout.splitlines())
    acf = O.acf_generate(O.WAVE_STANDING, amplitude=3.0, velocity=0.5, wavenum=(0.0625, 0.0),
                         alpha=(2.0, 0.2, 1.0), nwaves=(1.85, 16.0, 1.0), shape=(8, 8, 8))
    phi, _ = O.yw_choi(acf, 7)
    assert out["order"] == "(%d,%d,%d)" % phi.shape
    got = np.array([float(v) for v in out["phi"].split(",")]).reshape(phi.shape)
    # the two ACFs agree to 1e-11 (independent DFTs); the order-7 Choi solve
    # amplifies that to about 1e-9 on this ACF
    np.testing.assert_allclose(got, phi, rtol=1e-7, atol=1e-8)
    # out_grid length = acf delta * patches * 2 pi (src/generator/basic_arma_model.cc:61-66),
    # acf delta = nwaves / (8 - 1)
    lengths = [float(v) for v in out["grid"].split(":")[1].strip("()").split(",")]
    np.testing.assert_allclose(lengths, [1.85 / 7 * 39 * 2 * np.pi, 16 / 7 * 31 * 2 * np.pi,
                                         1 / 7 * 31 * 2 * np.pi], rtol=1e-12)
