// FP64 model setup that the AR generator runs before the GPU work:
// closed-form ACFs, Yule-Walker coefficient solvers, white-noise variance
// and the partition shape.  These are host computations in the reference
// too (src/generator/basic_arma_model.cc:51-101); they are small (at most a
// few 271 x 271 matrices) and stay in FP64.
//
//   acf .................. src/acf.hh:18-62, src/bits/acf_wrapper.cc:41-70
//   Choi order recursion . src/yule_walker.cc:420-539 (blocks :41-353)
//   Gauss elimination .... src/generator/ar_model.cc:97-135,
//                          src/generator/voodoo.cc:58-104
//   chop ................. src/chop.cc:4-26
//   white_noise_variance . src/generator/ar_model.cc:55-61
//   partition shape ...... src/generator/ar_model.cc:23-50
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "arma_hip.h"
#include "common.hpp"

namespace {

// Dense row-major FP64 matrix.
struct Matrix {
	int rows = 0, cols = 0;
	std::vector<double> v;
	Matrix() = default;
	Matrix(int r, int c) : rows(r), cols(c), v(size_t(r) * c, 0.0) {}
	double& operator()(int i, int j) { return v[size_t(i) * cols + j]; }
	double operator()(int i, int j) const { return v[size_t(i) * cols + j]; }
};

// c(i, j) = sum over k in order of a(i, k) b(k, j), each term rounded and then
// added (no fused multiply-add: the rounding is the reference's)
Matrix operator*(const Matrix& a, const Matrix& b) {
	Matrix c(a.rows, b.cols);
	for (int i = 0; i < a.rows; ++i)
		for (int k = 0; k < a.cols; ++k) {
			const double aik = a(i, k);
			const double* brow = &b.v[size_t(k) * b.cols];
			double* crow = &c.v[size_t(i) * c.cols];
			for (int j = 0; j < b.cols; ++j) crow[j] += aik * brow[j];
		}
	return c;
}

Matrix operator-(Matrix a, const Matrix& b) {
	for (size_t i = 0; i < a.v.size(); ++i) a.v[i] -= b.v[i];
	return a;
}

double dot(const Matrix& a, const Matrix& b) {
	double s = 0.0;
	for (size_t i = 0; i < a.v.size(); ++i) s += a.v[i] * b.v[i];
	return s;
}

// Inverse by LU with partial pivoting (getrf + getri).
Matrix inverse(const Matrix& m) {
	const int n = m.rows;
	Matrix lu = m;
	std::vector<int> piv(n);
	for (int k = 0; k < n; ++k) {
		int p = k;
		for (int i = k + 1; i < n; ++i)
			if (std::fabs(lu(i, k)) > std::fabs(lu(p, k))) p = i;
		if (lu(p, k) == 0.0) throw arma_hip::Error(ARMA_HIP_ERUNTIME, "getrf error");
		piv[k] = p;
		if (p != k)
			for (int j = 0; j < n; ++j) std::swap(lu(k, j), lu(p, j));
		for (int i = k + 1; i < n; ++i) {
			const double f = lu(i, k) / lu(k, k);
			lu(i, k) = f;
			for (int j = k + 1; j < n; ++j) lu(i, j) -= f * lu(k, j);
		}
	}
	Matrix inv(n, n);
	std::vector<double> col(n);
	for (int c = 0; c < n; ++c) {
		std::fill(col.begin(), col.end(), 0.0);
		col[c] = 1.0;
		for (int k = 0; k < n; ++k) std::swap(col[k], col[piv[k]]);
		for (int i = 0; i < n; ++i)
			for (int j = 0; j < i; ++j) col[i] -= lu(i, j) * col[j];
		// back substitution with each row's terms in descending column order
		// (the column-oriented order of a triangular solve, which the device
		// solver blocks by rows)
		for (int i = n - 1; i >= 0; --i) {
			for (int j = n - 1; j > i; --j) col[i] -= lu(i, j) * col[j];
			col[i] /= lu(i, i);
		}
		for (int i = 0; i < n; ++i) inv(i, c) = col[i];
	}
	return inv;
}

int shell(int a) { return 3 * a * a + 3 * a + 1; }

// Offsets of the shell points of radius d around the corner (c, c, c) in the
// order of the Choi block vectors: the corner, then for e = 1..d six runs of
// e points each.
struct Pt {
	int i, j, k;
};

std::vector<Pt> shell_points(int d) {
	std::vector<Pt> p;
	p.reserve(shell(d));
	p.push_back({0, 0, 0});
	for (int e = 1; e <= d; ++e) {
		for (int q = 0; q < e; ++q) p.push_back({e, q, 0});
		for (int q = 0; q < e; ++q) p.push_back({e - q, e, 0});
		for (int q = 0; q < e; ++q) p.push_back({0, e, q});
		for (int q = 0; q < e; ++q) p.push_back({0, e - q, e});
		for (int q = 0; q < e; ++q) p.push_back({q, 0, e});
		for (int q = 0; q < e; ++q) p.push_back({e, 0, e - q});
	}
	return p;
}

// Row points of R_{a,b}: (a,a,a) followed by six runs per e, mirrored.
std::vector<Pt> row_points(int a) {
	std::vector<Pt> p;
	p.reserve(shell(a));
	p.push_back({a, a, a});
	for (int e = 1; e <= a; ++e) {
		for (int q = 0; q < e; ++q) p.push_back({a - e, a - q, a});
		for (int q = 0; q < e; ++q) p.push_back({a - e + q, a - e, a});
		for (int q = 0; q < e; ++q) p.push_back({a, a - e, a - q});
		for (int q = 0; q < e; ++q) p.push_back({a, a - e + q, a - e});
		for (int q = 0; q < e; ++q) p.push_back({a - q, a, a - e});
		for (int q = 0; q < e; ++q) p.push_back({a - e, a, a - e + q});
	}
	return p;
}

class ChoiSolver {
public:
	ChoiSolver(const double* gamma, int n) : n_(n), rho_(size_t(n) * n * n) {
		for (size_t q = 0; q < rho_.size(); ++q) rho_[q] = gamma[q] / gamma[0];
		variance_ = gamma[0];
	}

	double rho(int i, int j, int k) const {
		return rho_[(size_t(std::abs(i)) * n_ + std::abs(j)) * n_ + std::abs(k)];
	}

	// R_{a,b}: rows are rho vectors of the row points of a, of depth b.
	Matrix R(int a, int b) const {
		const auto rows = row_points(a);
		const auto cols = shell_points(b);
		Matrix m(int(rows.size()), int(cols.size()));
		for (size_t r = 0; r < rows.size(); ++r)
			for (size_t c = 0; c < cols.size(); ++c)
				m(int(r), int(c)) = rho(rows[r].i - b + cols[c].i, rows[r].j - b + cols[c].j,
				                        rows[r].k - b + cols[c].k);
		return m;
	}

	Matrix R_b0(int a) const {
		const auto rows = row_points(a);
		Matrix m(int(rows.size()), 1);
		for (size_t r = 0; r < rows.size(); ++r) m(int(r), 0) = rho(rows[r].i, rows[r].j, rows[r].k);
		return m;
	}

	Matrix R_a0(int b) const {
		const auto cols = shell_points(b);
		Matrix m(int(cols.size()), 1);
		for (size_t c = 0; c < cols.size(); ++c)
			m(int(c), 0) = rho(-b + cols[c].i, -b + cols[c].j, -b + cols[c].k);
		return m;
	}

	// Returns phi as an l x l x l array (before chopping) and the order l.
	std::vector<double> solve(int max_order, double var_eps, bool determine, int& l_out,
	                          double& varwn) {
		const int M = std::min(max_order, n_ - 1);
		if (M < 2) throw arma_hip::Error(ARMA_HIP_EINVAL, "AR order below 2 for Choi solver");
		std::map<std::tuple<int, int, int>, Matrix> Phi;
		std::vector<Matrix> Pi_l(M + 2), Pi_lp1(M + 2), Theta(M + 2);
		const Matrix Rsup11 = inverse(R(1, 1));
		Pi_l[1] = Rsup11 * R_b0(1);
		double lambda = 1.0 - dot(R_a0(1), Pi_l[1]);
		double var0 = variance_;
		double var = variance_ * lambda;
		Phi[{2, 1, 0}] = Rsup11 * R(1, 2);
		bool changed = false;
		int l = 1;
		do {
			++l;
			var0 = var;
			Matrix theta = R(l, l);
			Matrix h = R_b0(l);
			for (int m = 1; m <= l - 1; ++m) {
				const Matrix Rlm = R(l, m);
				theta = theta - Rlm * Phi[{l, m, 0}];
				h = h - Rlm * Pi_l[m];
			}
			Theta[l] = inverse(theta);
			Pi_lp1[l] = Theta[l] * h;
			for (int a = 1; a <= l - 1; ++a) Pi_lp1[a] = Pi_l[a] - Phi[{l, a, 0}] * Pi_lp1[l];
			lambda -= dot(h, Pi_lp1[l]);
			var = variance_ * lambda;
			if (l < M) {
				Phi[{2, 1, l - 1}] = Rsup11 * R(1, l + 1);
				for (int m = 2; m <= l; ++m) {
					Matrix rhs = R(m, l + 1);
					for (int q = 1; q <= m - 1; ++q) rhs = rhs - R(m, q) * Phi[{m, q, l - m + 1}];
					Phi[{m + 1, m, l - m}] = Theta[m] * rhs;
					for (int a = 1; a <= m - 1; ++a)
						Phi[{m + 1, a, l - m}] = Phi[{m, a, l - m + 1}] - Phi[{m, a, 0}] * Phi[{m + 1, m, l - m}];
				}
			}
			std::swap(Pi_l, Pi_lp1);
			changed = !(determine && std::fabs(var - var0) < var_eps);
		} while (l < M && changed);
		const std::vector<Matrix>& beta = changed ? Pi_l : Pi_lp1;
		varwn = changed ? var : var0;
		std::vector<double> out(size_t(l) * l * l, 0.0);
		for (int d = 1; d < l; ++d) {
			const double* src = beta[d].v.data();
			auto put = [&](int i, int j, int k) { out[(size_t(i) * l + j) * l + k] = *src++; };
			put(d, d, d);
			for (int e = 1; e <= d; ++e) {
				for (int q = 0; q < e; ++q) put(d - e, d - q, d);
				for (int q = 0; q < e; ++q) put(d - e + q, d - e, d);
				for (int q = 0; q < e; ++q) put(d, d - e, d - q);
				for (int q = 0; q < e; ++q) put(d, d - e + q, d - e);
				for (int q = 0; q < e; ++q) put(d - q, d, d - e);
				for (int q = 0; q < e; ++q) put(d - e, d, d - e + q);
			}
		}
		l_out = l;
		return out;
	}

private:
	int n_;
	std::vector<double> rho_;
	double variance_ = 0.0;
};

void chop(const std::vector<double>& phi, const int shape[3], double eps, int out[3]) {
	const int n0 = shape[0], n1 = shape[1], n2 = shape[2];
	auto at = [&](int a, int b, int c) { return phi[(size_t(a) * n1 + b) * n2 + c]; };
	auto small_plane = [&](int a0, int a1, int b0, int b1, int c0, int c1) {
		for (int a = a0; a <= a1; ++a)
			for (int b = b0; b <= b1; ++b)
				for (int c = c0; c <= c1; ++c)
					if (!(std::fabs(at(a, b, c)) < eps)) return false;
		return true;
	};
	int k = n2 - 1;
	while (k >= 1 && small_plane(0, n0 - 1, 0, n1 - 1, k, k)) --k;
	int j = n1 - 1;
	while (j >= 1 && small_plane(0, n0 - 1, j, j, 0, k)) --j;
	int i = n0 - 1;
	while (i >= 1 && small_plane(i, i, 0, j, 0, k)) --i;
	out[0] = i + 1;
	out[1] = j + 1;
	out[2] = k + 1;
}

// ---- roots of 1 - phi_1 x - phi_2 x^2 - ... (validate_process) ----

// Eigenvalues of the upper Hessenberg matrix h (n x n, row-major, destroyed)
// by the Francis double-shift QR iteration with deflation on small
// subdiagonal entries; wr/wi receive real and imaginary parts.  Returns
// false when an eigenvalue does not converge in 60 iterations.
bool hessenberg_eigenvalues(std::vector<double>& h, int n, std::vector<double>& wr,
                            std::vector<double>& wi) {
	auto H = [&](int i, int j) -> double& { return h[size_t(i) * n + j]; };
	wr.assign(n, 0.0);
	wi.assign(n, 0.0);
	double norm = 0.0;
	for (int i = 0; i < n; ++i)
		for (int j = std::max(i - 1, 0); j < n; ++j) norm += std::fabs(H(i, j));
	const double eps = std::numeric_limits<double>::epsilon();
	int hi = n - 1;
	int its = 0;
	double exshift = 0.0;  // shifts accumulated by exceptional steps
	while (hi >= 0) {
		// smallest l with a negligible subdiagonal entry H(l, l-1)
		int l = hi;
		for (; l > 0; --l) {
			double s = std::fabs(H(l - 1, l - 1)) + std::fabs(H(l, l));
			if (s == 0.0) s = norm;
			if (std::fabs(H(l, l - 1)) <= eps * s) {
				H(l, l - 1) = 0.0;
				break;
			}
		}
		const double x = H(hi, hi);
		if (l == hi) {  // one real eigenvalue
			wr[hi] = x + exshift;
			--hi;
			its = 0;
			continue;
		}
		const double y = H(hi - 1, hi - 1);
		const double w = H(hi, hi - 1) * H(hi - 1, hi);
		if (l == hi - 1) {  // a 2 x 2 block: two eigenvalues
			const double p = 0.5 * (y - x);
			const double q = p * p + w;
			const double z = std::sqrt(std::fabs(q));
			const double xs = x + exshift;
			if (q >= 0.0) {
				const double zz = p + (p >= 0.0 ? z : -z);
				wr[hi - 1] = wr[hi] = xs + zz;
				if (zz != 0.0) wr[hi] = xs - w / zz;
			} else {
				wr[hi - 1] = wr[hi] = xs + p;
				wi[hi - 1] = z;
				wi[hi] = -z;
			}
			hi -= 2;
			its = 0;
			continue;
		}
		if (its == 60) return false;
		double sx = x, sy = y, sw = w;
		if (its == 10 || its == 20) {
			// exceptional shift
			exshift += x;
			for (int i = 0; i <= hi; ++i) H(i, i) -= x;
			const double s = std::fabs(H(hi, hi - 1)) + std::fabs(H(hi - 1, hi - 2));
			sx = sy = 0.75 * s;
			sw = -0.4375 * s * s;
		}
		++its;
		// double-shift step on rows/columns l..hi: first column of
		// (H - s1)(H - s2) from the trailing 2 x 2 block
		int m = hi - 2;
		double p = 0, q = 0, r = 0;
		for (; m >= l; --m) {
			const double z = H(m, m);
			const double rr = sx - z, ss = sy - z;
			p = (rr * ss - sw) / H(m + 1, m) + H(m, m + 1);
			q = H(m + 1, m + 1) - z - rr - ss;
			r = H(m + 2, m + 1);
			const double s = std::fabs(p) + std::fabs(q) + std::fabs(r);
			p /= s;
			q /= s;
			r /= s;
			if (m == l) break;
			const double u = std::fabs(H(m, m - 1)) * (std::fabs(q) + std::fabs(r));
			const double v = std::fabs(p) * (std::fabs(H(m - 1, m - 1)) + std::fabs(z) +
			                                 std::fabs(H(m + 1, m + 1)));
			if (u <= eps * v) break;
		}
		for (int i = m + 2; i <= hi; ++i) {
			H(i, i - 2) = 0.0;
			if (i != m + 2) H(i, i - 3) = 0.0;
		}
		// chase the bulge with 3 x 3 (last: 2 x 2) Householder reflectors
		// P = I - v v^T * 2 / v^T v, v = (p + s, q, r), s = sign(p) |(p, q, r)|
		for (int k = m; k <= hi - 1; ++k) {
			const bool three = k != hi - 1;
			double xn = 1.0;
			if (k != m) {
				p = H(k, k - 1);
				q = H(k + 1, k - 1);
				r = three ? H(k + 2, k - 1) : 0.0;
				xn = std::fabs(p) + std::fabs(q) + std::fabs(r);
				if (xn == 0.0) continue;
				p /= xn;
				q /= xn;
				r /= xn;
			}
			double sn = std::sqrt(p * p + q * q + r * r);
			if (p < 0.0) sn = -sn;
			if (k == m) {
				if (l != m) H(k, k - 1) = -H(k, k - 1);
			} else {
				H(k, k - 1) = -sn * xn;
			}
			const double ps = p + sn;
			const double vx = ps / sn, vy = q / sn, vz = r / sn;
			const double q2 = q / ps, r2 = r / ps;
			for (int c = k; c <= hi; ++c) {  // rows k..k+2
				double t = H(k, c) + q2 * H(k + 1, c);
				if (three) {
					t += r2 * H(k + 2, c);
					H(k + 2, c) -= t * vz;
				}
				H(k + 1, c) -= t * vy;
				H(k, c) -= t * vx;
			}
			const int imax = std::min(k + 3, hi);
			for (int row = l; row <= imax; ++row) {  // columns k..k+2
				double t = vx * H(row, k) + vy * H(row, k + 1);
				if (three) {
					t += vz * H(row, k + 2);
					H(row, k + 2) -= t * r2;
				}
				H(row, k + 1) -= t * q2;
				H(row, k) -= t;
			}
		}
	}
	return true;
}

}  // namespace

int arma_hip::polynomial_roots(const std::vector<double>& c, std::vector<double>& re,
                               std::vector<double>& im) {
	// companion matrix of c[0] + c[1] x + ... + c[n-1] x^(n-1), balanced by
	// powers of two (exact) row/column scalings before the QR iteration
	const int n = int(c.size()) - 1;
	if (n < 1) return ARMA_HIP_OK;
	if (c[n] == 0.0) return ARMA_HIP_ERUNTIME;  // leading term must be non-zero
	std::vector<double> h(size_t(n) * n, 0.0);
	for (int j = 0; j < n; ++j) h[j] = -c[n - 1 - j] / c[n];
	for (int i = 1; i < n; ++i) h[size_t(i) * n + i - 1] = 1.0;
	for (bool done = false; !done;) {
		done = true;
		for (int i = 0; i < n; ++i) {
			double rn = 0.0, cn = 0.0;
			for (int j = 0; j < n; ++j)
				if (j != i) {
					rn += std::fabs(h[size_t(i) * n + j]);
					cn += std::fabs(h[size_t(j) * n + i]);
				}
			if (rn == 0.0 || cn == 0.0) continue;
			double f = 1.0;
			const double sum = rn + cn;
			while (cn < rn / 2.0) {
				cn *= 2.0;
				rn /= 2.0;
				f *= 2.0;
			}
			while (cn >= rn * 2.0) {
				cn /= 2.0;
				rn *= 2.0;
				f /= 2.0;
			}
			if ((rn + cn) < 0.95 * sum * 1.0 && f != 1.0) {
				done = false;
				for (int j = 0; j < n; ++j) h[size_t(i) * n + j] /= f;
				for (int j = 0; j < n; ++j) h[size_t(j) * n + i] *= f;
			}
		}
	}
	if (!hessenberg_eigenvalues(h, n, re, im)) return ARMA_HIP_ERUNTIME;
	return ARMA_HIP_OK;
}

void arma_hip::chop_shape(const std::vector<double>& phi, const int shape[3], double eps,
                          int out[3]) {
	chop(phi, shape, eps, out);
}

namespace {

// gsl_stats_variance: running mean and running mean of squared deviations in
// long double, times n/(n-1).
double gsl_style_variance(const std::vector<double>& v) {
	long double mean = 0.0L;
	for (size_t i = 0; i < v.size(); ++i) mean += (v[i] - mean) / (long double)(i + 1);
	const double m = double(mean);
	long double var = 0.0L;
	for (size_t i = 0; i < v.size(); ++i) {
		const long double d = v[i] - m;
		var += (d * d - var) / (long double)(i + 1);
	}
	return double(var) * (double(v.size()) / double(v.size() - 1));
}

double wave_profile(int func, double amp, double kx, double ky, double vel, double x, double y,
                    double t) {
	const double two_pi = 2.0 * M_PI;  // T(2) * pi (src/physical_constants.hh:15)
	const double theta = two_pi * (kx * x + ky * y - vel * t);
	switch (func) {
	case ARMA_WAVE_SIN: return amp * std::sin(theta + 0.0);
	case ARMA_WAVE_COS: return amp * std::sin(theta + (0.0 + 0.5 * M_PI));
	case ARMA_WAVE_STOKES: {
		const double st = two_pi * amp * std::sqrt(kx * kx + ky * ky);
		return amp * (std::cos(theta + 0.0) + 0.5 * st * std::cos(2.0 * (theta + 0.0)) +
		              (3.0 / 8.0) * st * st * std::cos(3.0 * (theta + 0.0)));
	}
	default:
		return amp * std::cos(two_pi * (kx * x + ky * y) + 0.0) * std::cos(two_pi * (vel * t));
	}
}

struct Wave {
	int s[3];
	std::vector<double> v;
};

// The plain wave on Domain(-r, r, ws + 1) (src/generator/acf_generator.hh:156-167).
Wave make_wave(const arma_acf_gen_desc& d, const int ws[3]) {
	Wave w;
	double step[3];
	for (int q = 0; q < 3; ++q) {
		w.s[q] = ws[q] + 1;
		step[q] = (d.nwaves[q] - (-d.nwaves[q])) / ws[q];
	}
	w.v.resize(size_t(w.s[0]) * w.s[1] * w.s[2]);
	size_t q = 0;
	for (int i = 0; i < w.s[0]; ++i)
		for (int j = 0; j < w.s[1]; ++j)
			for (int k = 0; k < w.s[2]; ++k)
				w.v[q++] = wave_profile(d.func, d.amplitude, d.beta[0], d.beta[1], d.velocity,
				                        -d.nwaves[1] + step[1] * j, -d.nwaves[2] + step[2] * k,
				                        -d.nwaves[0] + step[0] * i);
	return w;
}

// In-place DFT of every line along axis ax of a (s0, s1, s2) complex array,
// e^{sign 2 pi i jk/n} from a table of the n roots of unity.
void dft_axis(std::vector<std::complex<double>>& a, const int s[3], int ax, int sign) {
	const int n = s[ax];
	std::vector<std::complex<double>> root(n), line(n), res(n);
	for (int m = 0; m < n; ++m) {
		const double ang = 2.0 * M_PI * double(m) / double(n);
		root[m] = std::complex<double>(std::cos(ang), sign * std::sin(ang));
	}
	const size_t stride = ax == 0 ? size_t(s[1]) * s[2] : ax == 1 ? size_t(s[2]) : 1;
	const int o1 = ax == 0 ? s[1] : s[0], o2 = ax == 2 ? s[1] : s[2];
	for (int p = 0; p < o1; ++p)
		for (int r = 0; r < o2; ++r) {
			size_t base;
			if (ax == 0) base = size_t(p) * s[2] + r;
			else if (ax == 1) base = size_t(p) * s[1] * s[2] + r;
			else base = (size_t(p) * s[1] + r) * s[2];
			for (int m = 0; m < n; ++m) line[m] = a[base + m * stride];
			for (int k = 0; k < n; ++k) {
				std::complex<double> acc(0.0, 0.0);
				int idx = 0;
				for (int m = 0; m < n; ++m) {
					acc += line[m] * root[idx];
					idx += k;
					if (idx >= n) idx -= n;
				}
				res[k] = acc;
			}
			for (int m = 0; m < n; ++m) a[base + m * stride] = res[m];
		}
}

}  // namespace

// ACF_generator<T>::generate up to the auto-covariance
// (src/generator/acf_generator.cc:130-215): the plain wave on
// [-nwaves, nwaves] with 2*shape+1 points per dimension (or, without a shape,
// doubled from 2 until its variance settles), times exp(-sum |alpha (t,x,y)|),
// rescaled to the plain wave's variance.  s receives its extents.
void arma_hip::acf_gen_wave(const arma_acf_gen_desc& d, std::vector<double>& we, int s[3]) {
	int ws[3] = {2, 2, 2};
	Wave w;
	if (d.shape[0] > 0 && d.shape[1] > 0 && d.shape[2] > 0) {
		for (int q = 0; q < 3; ++q) ws[q] = 2 * d.shape[q];
		w = make_wave(d, ws);
	} else {
		// generate_optimal_wavy_surface (src/generator/acf_generator.cc:158-193)
		double var0 = -1.0, var = -1.0;
		while (ws[0] < 64 && ws[1] < 64 && ws[2] < 64) {
			w = make_wave(d, ws);
			var0 = var;
			var = gsl_style_variance(w.v);
			if (!(var0 < 0.0) && std::fabs(var - var0) < 1e-3) break;
			for (int q = 0; q < 3; ++q) ws[q] *= 2;
		}
	}
	// add_exponential_decay (src/generator/acf_generator.cc:196-215)
	double step[3];
	for (int q = 0; q < 3; ++q) step[q] = (d.nwaves[q] - (-d.nwaves[q])) / (w.s[q] - 1);
	we.resize(w.v.size());
	size_t q = 0;
	for (int i = 0; i < w.s[0]; ++i)
		for (int j = 0; j < w.s[1]; ++j)
			for (int k = 0; k < w.s[2]; ++k, ++q) {
				const double t = -d.nwaves[0] + step[0] * i, x = -d.nwaves[1] + step[1] * j,
				             y = -d.nwaves[2] + step[2] * k;
				we[q] = w.v[q] * std::exp(-(std::fabs(t * d.alpha[0]) + std::fabs(x * d.alpha[1]) +
				                            std::fabs(y * d.alpha[2])));
			}
	const double f = std::sqrt(gsl_style_variance(w.v) / gsl_style_variance(we));
	for (double& v : we) v *= f;
	for (int r = 0; r < 3; ++r) s[r] = w.s[r];
}

// The first half of the auto-covariance in every dimension and its grid
// (Grid{shape, nwaves}: spacing nwaves / (shape - 1)).
void arma_hip::acf_gen_crop(const arma_acf_gen_desc& d, const int s[3], int h[3],
                            int32_t shape_out[3], double delta_out[3]) {
	for (int r = 0; r < 3; ++r) {
		h[r] = s[r] / 2;
		shape_out[r] = h[r];
		delta_out[r] = h[r] > 1 ? d.nwaves[r] / double(h[r] - 1) : 0.0;
	}
}

extern "C" {

int arma_hip_acf(int32_t func, const double delta[3], const int32_t shape[3], double amplitude,
                 double velocity, const double alpha[3], const double beta[2], double* out) {
	if (!delta || !shape || !alpha || !beta || !out) return ARMA_HIP_EINVAL;
	if (func != ARMA_ACF_STANDING_WAVE && func != ARMA_ACF_PROPAGATING_WAVE) return ARMA_HIP_EINVAL;
	for (int i = 0; i < shape[0]; ++i)
		for (int j = 0; j < shape[1]; ++j)
			for (int k = 0; k < shape[2]; ++k) {
				const double decay =
					std::exp(-(alpha[0] * i * delta[0] + alpha[1] * j * delta[1] + alpha[2] * k * delta[2]));
				const double wave =
					func == ARMA_ACF_STANDING_WAVE
						? std::cos(velocity * i * delta[0]) * std::cos(beta[0] * j * delta[1]) *
						      std::cos(beta[1] * k * delta[2])
						: std::cos(velocity * i * delta[0] + beta[0] * j * delta[1] +
						           beta[1] * k * delta[2]);
				out[(size_t(i) * shape[1] + j) * shape[2] + k] = amplitude * decay * wave;
			}
	return ARMA_HIP_OK;
}


int arma_hip_acf_generate(const arma_acf_gen_desc* desc, double* out, size_t cap,
                          int32_t shape_out[3], double delta_out[3]) {
	if (!desc || !out || !shape_out || !delta_out) return ARMA_HIP_EINVAL;
	if (desc->func < ARMA_WAVE_SIN || desc->func > ARMA_WAVE_STANDING) return ARMA_HIP_EINVAL;
	try {
		std::vector<double> we;
		int s[3];
		arma_hip::acf_gen_wave(*desc, we, s);
		// auto_covariance (src/arma.cc:81-100)
		std::vector<std::complex<double>> a(we.begin(), we.end());
		for (int ax = 0; ax < 3; ++ax) dft_axis(a, s, ax, -1);
		for (auto& c : a) {
			const double m = std::abs(c);
			c = std::complex<double>(m * m, 0.0);
		}
		for (int ax = 0; ax < 3; ++ax) dft_axis(a, s, ax, +1);
		const double n = double(a.size());
		int h[3];
		arma_hip::acf_gen_crop(*desc, s, h, shape_out, delta_out);
		if (size_t(h[0]) * h[1] * h[2] > cap) {
			arma_hip::set_last_error("ACF output buffer too small");
			return ARMA_HIP_EINVAL;
		}
		for (int i = 0; i < h[0]; ++i)
			for (int j = 0; j < h[1]; ++j)
				for (int k = 0; k < h[2]; ++k)
					out[(size_t(i) * h[1] + j) * h[2] + k] =
						a[(size_t(i) * s[1] + j) * s[2] + k].real() / n / n;
		return ARMA_HIP_OK;
	} catch (const std::exception& e) {
		arma_hip::set_last_error(e.what());
		return ARMA_HIP_ERUNTIME;
	}
}

int arma_hip_yule_walker_choi(const double* acf, int32_t n, int32_t max_order, double var_eps,
                              int32_t determine_order, int32_t do_chop, double chop_eps,
                              double* phi_out, int32_t order_out[3], double* varwn_solver) {
	if (!acf || !phi_out || !order_out || n < 3) return ARMA_HIP_EINVAL;
	try {
		ChoiSolver solver(acf, n);
		int l = 0;
		double var = 0.0;
		std::vector<double> phi = solver.solve(max_order, var_eps, determine_order != 0, l, var);
		int shp[3] = {l, l, l};
		if (do_chop) chop(phi, shp, chop_eps, shp);
		for (int a = 0; a < shp[0]; ++a)
			for (int b = 0; b < shp[1]; ++b)
				for (int c = 0; c < shp[2]; ++c)
					phi_out[(size_t(a) * shp[1] + b) * shp[2] + c] = phi[(size_t(a) * l + b) * l + c];
		order_out[0] = shp[0];
		order_out[1] = shp[1];
		order_out[2] = shp[2];
		if (varwn_solver) *varwn_solver = var;
		return ARMA_HIP_OK;
	} catch (const arma_hip::Error& e) {
		return e.code;
	} catch (...) {
		return ARMA_HIP_ERUNTIME;
	}
}

int arma_hip_yule_walker_gauss(const double* acf, const int32_t acf_shape[3],
                               const int32_t order[3], double* phi_out) {
	if (!acf || !acf_shape || !order || !phi_out) return ARMA_HIP_EINVAL;
	for (int d = 0; d < 3; ++d)
		if (order[d] < 1 || order[d] > acf_shape[d]) return ARMA_HIP_EINVAL;
	const int P = order[0] * order[1] * order[2];
	const int m = P - 1;
	phi_out[0] = 0.0;
	if (m == 0) return ARMA_HIP_OK;
	auto gam = [&](int a, int b, int c) {
		return acf[(size_t(std::abs(a)) * acf_shape[1] + std::abs(b)) * acf_shape[2] + std::abs(c)];
	};
	auto coord = [&](int r, int& i, int& j, int& k) {
		i = r / (order[1] * order[2]);
		j = (r / order[2]) % order[1];
		k = r % order[2];
	};
	Matrix A(m, m);
	std::vector<double> b(m);
	for (int r = 1; r < P; ++r) {
		int ri, rj, rk;
		coord(r, ri, rj, rk);
		b[r - 1] = gam(ri, rj, rk);
		for (int c = 1; c < P; ++c) {
			int ci, cj, ck;
			coord(c, ci, cj, ck);
			A(r - 1, c - 1) = gam(ri - ci, rj - cj, rk - ck);
		}
	}
	// The block-Toeplitz matrix is symmetric but not always definite (the
	// |lag| symmetrisation of a propagating ACF makes it indefinite), so it
	// is solved as the reference's dsysv does for any nonsingular symmetric
	// matrix, here by Gaussian elimination with partial pivoting.
	for (int k = 0; k < m; ++k) {
		int p = k;
		for (int i = k + 1; i < m; ++i)
			if (std::fabs(A(i, k)) > std::fabs(A(p, k))) p = i;
		if (A(p, k) == 0.0) {
			arma_hip::set_last_error("singular Yule-Walker system (dsysv info > 0)");
			return ARMA_HIP_ERUNTIME;
		}
		if (p != k) {
			for (int j = 0; j < m; ++j) std::swap(A(k, j), A(p, j));
			std::swap(b[k], b[p]);
		}
		for (int i = k + 1; i < m; ++i) {
			const double f = A(i, k) / A(k, k);
			if (f == 0.0) continue;
			for (int j = k + 1; j < m; ++j) A(i, j) -= f * A(k, j);
			b[i] -= f * b[k];
		}
	}
	for (int i = m - 1; i >= 0; --i) {
		double s = b[i];
		for (int j = i + 1; j < m; ++j) s -= A(i, j) * b[j];
		b[i] = s / A(i, i);
	}
	for (int i = 0; i < m; ++i) phi_out[i + 1] = b[i];
	return ARMA_HIP_OK;
}

double arma_hip_white_noise_variance(const double* acf, const int32_t acf_shape[3],
                                     const double* phi, const int32_t phi_shape[3]) {
	double s = 0.0;
	for (int a = 0; a < phi_shape[0]; ++a)
		for (int b = 0; b < phi_shape[1]; ++b)
			for (int c = 0; c < phi_shape[2]; ++c)
				s += phi[(size_t(a) * phi_shape[1] + b) * phi_shape[2] + c] *
				     acf[(size_t(a) * acf_shape[1] + b) * acf_shape[2] + c];
	return acf[0] - s;
}

int arma_hip_validate_process(const double* phi, const int32_t shape[3], int32_t* n_bad,
                              double* roots, int32_t roots_len) {
	try {
		if (!phi || !shape || !n_bad) return ARMA_HIP_EINVAL;
		*n_bad = 0;
		const size_t n = size_t(shape[0]) * shape[1] * shape[2];
		if (n <= 1) return ARMA_HIP_OK;
		// 1 - phi_1 x - phi_2 x^2 - ... over the flattened coefficients
		std::vector<double> c(n);
		for (size_t i = 0; i < n; ++i) c[i] = -phi[i];
		c[0] = 1.0;
		std::vector<double> re, im;
		if (arma_hip::polynomial_roots(c, re, im) != ARMA_HIP_OK) {
			arma_hip::set_last_error(
				"Can not find roots of the polynomial to verify AR/MA model "
				"stationarity/invertibility.");
			return ARMA_HIP_ERUNTIME;
		}
		int bad = 0;
		for (size_t i = 0; i < re.size(); ++i) {
			if (!(std::hypot(re[i], im[i]) > 1.0)) ++bad;
			if (roots && int32_t(2 * i + 1) < roots_len) {
				roots[2 * i] = re[i];
				roots[2 * i + 1] = im[i];
			}
		}
		*n_bad = bad;
		if (bad > 0) {
			arma_hip::set_last_error(
				"AR/MA process is not stationary/invertible: some roots lie inside unit circle "
				"or on its borderline.");
			return ARMA_HIP_ERUNTIME;
		}
		return ARMA_HIP_OK;
	} catch (...) {
		return ARMA_HIP_ERUNTIME;
	}
}

void arma_hip_partition_shape(const int32_t partition[3], const int32_t grid[3],
                              const int32_t order[3], int32_t parallelism, int32_t out[3]) {
	if (partition[0] * partition[1] * partition[2] > 0) {
		for (int d = 0; d < 3; ++d) out[d] = partition[d];
		return;
	}
	const int npar = std::max(1, 7 * int(std::cbrt(double(parallelism))));
	for (int d = 0; d < 3; ++d) {
		const int g1 = std::max(2 * order[d], 10);
		const int g2 = grid[d] / npar + (grid[d] % npar ? 1 : 0);
		out[d] = std::min(g1, g2) + std::abs(g1 - g2) / 2;
	}
}

}  // extern "C"
