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
    # phi.csv holds 6 significant digits (the stream default the reference
    # writes with): every value within half a unit of its 6th digit, plus the
    # GPU's 1e-12-of-max distance from the oracle
    assert np.all(np.abs(pot - pref) <= 5e-6 * np.abs(pref) + 1e-12 * np.abs(pref).max())


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


# ---- text outputs in the reference's formats, and the Gauss/validate path ----

GAUSS_INPUT = """model = AR {
	out_grid = (%(nt)d,%(nx)d,%(ny)d)
	acf = {
		func = %(func)s
		grid = (10,10,10)
		velocity = 0.8
		beta = %(beta)s
	}
	order = (7,7,7)
	algorithm = gauss_elimination
	no_seed = 1
	validate = 1
	output = surface,binary,csv,blitz
}
velocity_potential_solver = linear {
	depth = 12
	domain = from (0,-10) to (%(tlast)d,0) npoints (%(nt)d,%(nz)d)
}
"""


def fmt(v):
    """A double as std::ostream writes it by default: %g, 6 significant digits."""
    return "%g" % v


def blitz_text(a):
    """Blitz++'s operator<< for an array of rank >= 2 (the reference writes
    `std::ofstream(name) << array`, src/arma_driver.cc:33,54)."""
    out = [" x ".join("(0,%d)" % (n - 1) for n in a.shape), "\n[ "]
    for p, v in enumerate(a.reshape(-1), 1):
        out.append("%9s " % fmt(v))
        if p % 7 == 0:
            out.append("\n  ")
    out.append("]\n")
    return "".join(out)


def surface_csv(z, dx, dy):
    """src/bits/write_csv.hh:13-42."""
    nt, nx, ny = z.shape
    lines = ["t,x,y,z"]
    for i in range(nt):
        for j in range(nx):
            for k in range(ny):
                lines.append("%d,%s,%s,%s" % (i, fmt(dx * j), fmt(dy * k), fmt(z[i, j, k])))
    return "\n".join(lines) + "\n"


def potential_csv(phi, lb, ub, dx, dy):
    """src/bits/write_csv.hh:44-80 with the domain point lbound + delta * i."""
    nt, nz, nx, ny = phi.shape
    dt = (ub[0] - lb[0]) / max(nt - 1, 1)
    dz = (ub[1] - lb[1]) / max(nz - 1, 1)
    lines = ["t,z,x,y,phi"]
    for i in range(nt):
        for j in range(nz):
            head = "%s,%s," % (fmt(lb[0] + dt * i), fmt(lb[1] + dz * j))
            for k in range(nx):
                for l in range(ny):
                    lines.append(head + "%s,%s,%s" % (fmt(dx * k), fmt(dy * l),
                                                      fmt(phi[i, j, k, l])))
    return "\n".join(lines) + "\n"


def same_text_up_to_last_digit(got, want, rel=1e-5):
    """Byte-equal, except numbers whose 6-digit forms differ in the last digit
    because the two values (GPU and oracle) differ by rounding; at most
    0.1 % of the tokens may do so."""
    if got == want:
        return
    gt, wt = got.split(), want.split()
    assert len(gt) == len(wt)
    diff = 0
    for a, b in zip(gt, wt):
        if a == b:
            continue
        diff += 1
        fa = [float(x) for x in a.split(",")]
        fb = [float(x) for x in b.split(",")]
        for x, y in zip(fa, fb):
            assert abs(x - y) <= rel * abs(y) + 1e-300, (a, b)
    assert diff <= max(1, len(wt) // 1000), diff


@pytest.mark.gpu
def test_end_to_end_gauss_validate_text_outputs(tmp_path):
    """algorithm = gauss_elimination and validate = 1 (src/generator/ar_model.cc:97-135,
    src/arma.cc:15-70) on the standing wave, whose Gauss coefficients pass the
    root check; every output format of the surface and the potential."""
    import arma_amd
    nt, nx, ny, nz = 24, 16, 16, 3
    arma_amd.write_mt_config(str(tmp_path / "mt.dat"),
                             arma_amd._abi.params_array([arma_amd.mt19937_params()]))
    (tmp_path / "g.arma").write_text(GAUSS_INPUT % dict(
        nt=nt, nx=nx, ny=ny, nz=nz, tlast=nt - 1, func="standing_wave", beta="(0.8,0.8)"))
    r = run([str(tmp_path / "g.arma")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    zeta = np.fromfile(tmp_path / "zeta.bin", dtype=">f8").reshape(nt, nx, ny).astype(np.float64)
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi = O.yw_gauss(acf, (7, 7, 7))
    varwn = O.white_noise_variance(acf, phi)
    eps = O.noise_sequential(O.mt19937_params(), 0, np.sqrt(varwn), nt * nx * ny)
    ref = O.ar_generate(eps.reshape(nt, nx, ny), phi)
    ref *= O.rescale_factor(ref, acf[0, 0, 0])
    err = float(np.max(np.abs(zeta - ref)) / np.max(np.abs(ref)))
    assert err < 1e-10, err
    # the surface's text forms are those of its exact values
    d = 2 * np.pi  # unit ACF grid: delta = 1 * patches * 2 pi / patches
    dx = (1.0 * (nx - 1) * 2.0 * np.pi) / (nx - 1)
    dy = (1.0 * (ny - 1) * 2.0 * np.pi) / (ny - 1)
    assert abs(dx - d) < 1e-15 * d
    assert (tmp_path / "zeta.csv").read_text() == surface_csv(zeta, dx, dy)
    assert (tmp_path / "zeta").read_text() == blitz_text(zeta)
    # the potential's, against the oracle's potential of that surface
    pref = O.linear_potential(zeta, (2 * np.pi,) * 3, 12.0, (0.0, -10.0), (nt - 1.0, 0.0),
                              (nt, nz))
    same_text_up_to_last_digit((tmp_path / "phi").read_text(), blitz_text(pref))
    same_text_up_to_last_digit((tmp_path / "phi.csv").read_text(),
                               potential_csv(pref, (0.0, -10.0), (nt - 1.0, 0.0), dx, dy))


@pytest.mark.gpu
def test_end_to_end_validate_rejects(tmp_path):
    """validate = 1 with coefficients that have roots inside the unit circle
    (the propagating wave's Gauss fit): the reference's runtime error."""
    import arma_amd
    arma_amd.write_mt_config(str(tmp_path / "mt.dat"),
                             arma_amd._abi.params_array([arma_amd.mt19937_params()]))
    (tmp_path / "g.arma").write_text(GAUSS_INPUT % dict(
        nt=16, nx=16, ny=16, nz=2, tlast=15, func="propagating_wave", beta="(-0.8,0)"))
    acf = O.acf(O.PROPAGATING, (10, 10, 10), velocity=0.8, beta=(-0.8, 0.0))
    # the decision the driver must reproduce, from numpy's roots of
    # 1 - phi_1 x - phi_2 x^2 - ... (flattened coefficients)
    c = -O.yw_gauss(acf, (7, 7, 7)).reshape(-1)
    c[0] = 1.0
    assert np.min(np.abs(np.roots(c[::-1]))) <= 1.0
    r = run([str(tmp_path / "g.arma")], str(tmp_path))
    assert r.returncode != 0
    assert "not stationary" in r.stderr
    assert not (tmp_path / "zeta.bin").exists()


@pytest.mark.gpu
def test_end_to_end_acf_generator(tmp_path):
    """HEAD's acf syntax (no grid =) through a whole run: ACF_generator with
    the auto-covariance on the GPU, then Choi, noise, recurrence, rescale,
    against the oracle's chain.  The ACFs agree to ~1e-13 and the order-7
    Choi solve amplifies that to ~1e-9 in phi; the bar is 1e-6 of max."""
    import arma_amd
    arma_amd.write_mt_config(str(tmp_path / "mt.dat"),
                             arma_amd._abi.params_array([arma_amd.mt19937_params()]))
    p = tmp_path / "head.arma"
    p.write_text(HEAD_INPUT.replace("output = surface", "no_seed = 1\n\toutput = surface,binary")
                 .replace("domain = from (10,-12) to (10,3) npoints (1,128)",
                          "domain = from (10,-12) to (10,0) npoints (1,4)"))
    r = run([str(p)], str(tmp_path))
    assert r.returncode == 0, r.stderr
    zeta = np.fromfile(tmp_path / "zeta.bin", dtype=">f8").reshape(40, 32, 32)
    acf = O.acf_generate(O.WAVE_STANDING, amplitude=3.0, velocity=0.5, wavenum=(0.0625, 0.0),
                         alpha=(2.0, 0.2, 1.0), nwaves=(1.85, 16.0, 1.0), shape=(8, 8, 8))
    phi, _ = O.yw_choi(acf, 7)
    varwn = O.white_noise_variance(acf, phi)
    eps = O.noise_sequential(O.mt19937_params(), 0, np.sqrt(varwn), 40 * 32 * 32)
    ref = O.ar_generate(eps.reshape(40, 32, 32), phi)
    ref *= O.rescale_factor(ref, acf[0, 0, 0])
    err = float(np.max(np.abs(zeta - ref)) / np.max(np.abs(ref)))
    assert err < 1e-6, err
