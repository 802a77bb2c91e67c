/*
 * arma_oracle.c -- CPU FP64 restatement of the ARMA hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  Build: oracle/Makefile (-O2 -ffp-contract=off, no fast-math).
 *
 * Parity status (what pins each function; reference paths are relative to
 * the igankevich/arma source tree):
 *   orc_mt_*            pinned: MT19937 parameter set encoded as an mt_struct
 *                       reproduces std::mt19937 (known answer 4123659995 at
 *                       the 10000th draw of seed 5489, plus g++ golden
 *                       vectors); p=521 parameter sets pinned by
 *                       Berlekamp-Massey linear complexity 521 and the
 *                       primitivity test x^(2^521) = x mod f.
 *   orc_canonical_*, orc_normal_next_*
 *                       pinned: bit-exact against g++ 11 libstdc++
 *                       std::normal_distribution<double/float> golden vectors.
 *   orc_noise_*         pinned by the two above plus index bookkeeping tests.
 *   orc_acf             pinned: closed form checked pointwise in tests.
 *   orc_yw_choi         pinned: reference yule-walker-test cases (Choi equals
 *                       Gauss to 1e-4 / 1e-2, order (2,2,2), chop (2,2,2),
 *                       real case extent 2) and the Yule-Walker residual.
 *   orc_yw_gauss        pinned: residual of the normal equations, brute force.
 *   orc_ar_generate     pinned: inverse-filter identity, impulse response,
 *                       chunked-equals-one-shot, brute-force tiny grids.
 *   orc_gsl_variance_*  pinned: exact small cases and a long-double two-pass
 *                       brute force (GSL itself is absent: parity against GSL
 *                       bit patterns unpinned, tolerance 1e-15 relative).
 *   orc_derivative_t    pinned: exact on quadratics in t.
 *   orc_factor_waves    pinned: hand-worked known answers on built lines
 *                       (event detection, parabola vertex, run
 *                       representative = first largest |elevation|, the
 *                       last run unscored, runs continuing across columns,
 *                       no events), plus a run-by-run Python restatement on
 *                       random slices.
 *   orc_interp_edges    pinned: hand-worked values on a non-planar integer
 *                       field (exact), including res(0,1) reading the fixed
 *                       res(1,0) and res(2,0); misread orders fail them.
 *   orc_fft*            pinned: numpy.fft golden vectors.
 *   orc_window          pinned: closed form cosh(l(z+h))/cosh(lh) and the
 *                       rotate index map.
 *   orc_linear_potential pinned: single-Fourier-mode analytic solution and
 *                       numpy restatement on 16x16 grids.
 *   orc_gauss_legendre_half, orc_owen_t
 *                       pinned: numpy's leggauss(20) nodes and weights;
 *                       T(0, a) = atan(a) / (2 pi), T(h, 1) =
 *                       Phi(h) (1 - Phi(h)) / 2, adaptive quadrature, and
 *                       the known answers of src/test/owen-t-test.cc.
 *   orc_gram_charlier_cdf, orc_skew_normal_cdf
 *                       pinned: Gram-Charlier with skewness 0 and kurtosis 3
 *                       is Phi; its derivative is the Gram-Charlier A density
 *                       phi (1 + s He3 / 6 + (k - 3) He4 / 24); skew normal
 *                       with alpha 0 is the normal CDF, its derivative the
 *                       skew normal density 2 phi(z) Phi(alpha z) / sigma;
 *                       skew normal at the nine known answers of
 *                       src/test/skew-normal-test.cc.
 *   orc_nit_realisation pinned: brute-force Python restatement of the
 *                       bisection (src/linalg.hh:199-213), and the identity
 *                       transforms (Gram-Charlier (0, 3) maps x to x/stdev,
 *                       skew normal alpha 0 to mean + sigma x / stdev).
 *   orc_acf_generate    pinned: numpy restatement (numpy.fft auto-covariance,
 *                       ddof=1 variance) and the circular auto-covariance by
 *                       brute force on a tiny grid; acf(0,0,0) is the mean
 *                       square of the decayed wave.
 *   orc_derivative_xy   pinned: exact on quadratics in x and y, raw slice for
 *                       extents below 3.
 *   orc_high_amplitude_potential
 *                       pinned: closed form for a plane surface
 *                       zeta = a x + b y + c t (phi = 4 pi Re(c / D) with the
 *                       constant denominator D), the small-slope limit equal
 *                       to the linear solver to O(slope^2), and a numpy
 *                       restatement with complex FFTs.  The sign of Im D
 *                       enters only through the odd part of the rotated
 *                       window; nothing but the restatement of
 *                       src/velocity/high_amplitude_solver.cc:20-22 fixes it
 *                       (parity unpinned for that sign).
 */
#include "arma_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Mersenne Twister, dynamic-creator parameterisation.                       */
/* Reference call sites: src/parallel_mt.hh:236 (genrand_mt), :259           */
/* (sgenrand_mt).  Algorithm: Matsumoto & Nishimura, "Dynamic creation of    */
/* pseudorandom number generators" (libdcmt genmtrand.c).                    */
/* ------------------------------------------------------------------------ */

int orc_mt_seed(orc_mt* g, const orc_mt_params* p, uint32_t seed) {
	if (p->nn <= 0 || p->nn > ORC_MT_MAXN || p->mm <= 0 || p->mm >= p->nn) {
		return -1;
	}
	g->p = *p;
	/* Knuth multiplier initialisation, then masking to w bits. */
	uint32_t s = seed;
	for (int i = 0; i < p->nn; ++i) {
		g->state[i] = s;
		s = 1812433253u * (s ^ (s >> 30)) + (uint32_t)(i + 1);
	}
	for (int i = 0; i < p->nn; ++i) {
		g->state[i] &= p->wmask;
	}
	g->i = p->nn;
	return 0;
}

static inline uint32_t orc_mt_twist_word(uint32_t hi, uint32_t lo, uint32_t src,
                                         const orc_mt_params* p) {
	const uint32_t x = (hi & p->umask) | (lo & p->lmask);
	return src ^ (x >> 1) ^ ((x & 1u) ? p->aaa : 0u);
}

uint32_t orc_mt_next(orc_mt* g) {
	const orc_mt_params* p = &g->p;
	uint32_t* st = g->state;
	if (g->i >= p->nn) {
		const int n = p->nn, m = p->mm;
		int k = 0;
		for (; k < n - m; ++k) st[k] = orc_mt_twist_word(st[k], st[k + 1], st[k + m], p);
		for (; k < n - 1; ++k) st[k] = orc_mt_twist_word(st[k], st[k + 1], st[k + m - n], p);
		st[n - 1] = orc_mt_twist_word(st[n - 1], st[0], st[m - 1], p);
		g->i = 0;
	}
	uint32_t x = st[g->i++];
	x ^= x >> p->shift0;
	x ^= (x << p->shiftB) & p->maskB;
	x ^= (x << p->shiftC) & p->maskC;
	x ^= x >> p->shift1;
	return x;
}

int orc_mt_fill(const orc_mt_params* p, uint32_t seed, uint32_t* out, size_t n) {
	orc_mt* g = (orc_mt*)malloc(sizeof(orc_mt));
	if (!g) return -2;
	if (orc_mt_seed(g, p, seed)) { free(g); return -1; }
	for (size_t i = 0; i < n; ++i) out[i] = orc_mt_next(g);
	free(g);
	return 0;
}

/* ------------------------------------------------------------------------ */
/* libstdc++ <random> semantics (g++ 11):                                    */
/*   generate_canonical: /usr/include/c++/11/bits/random.tcc:3348-3380       */
/*   normal_distribution (Marsaglia polar): random.tcc:1802-1836             */
/* Used by src/generator/ar_model_wn.cc:10-16 and src/white_noise.hh:54.     */
/* ------------------------------------------------------------------------ */

double orc_canonical_d(orc_mt* g) {
	/* 53 bits from a 32-bit engine: two draws, low word first. */
	double sum = 0.0;
	double tmp = 1.0;
	for (int k = 0; k < 2; ++k) {
		sum += (double)orc_mt_next(g) * tmp;
		tmp *= 4294967296.0;
	}
	double ret = sum / tmp;
	if (ret >= 1.0) ret = nextafter(1.0, 0.0);
	return ret;
}

float orc_canonical_f(orc_mt* g) {
	float sum = (float)orc_mt_next(g) * 1.0f;
	float ret = sum / 4294967296.0f;
	if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
	return ret;
}

double orc_normal_next_d(orc_mt* g, orc_normal_d* st, double mean, double stddev) {
	double ret;
	if (st->saved_available) {
		st->saved_available = 0;
		ret = st->saved;
	} else {
		double x, y, r2;
		do {
			x = 2.0 * orc_canonical_d(g) - 1.0;
			y = 2.0 * orc_canonical_d(g) - 1.0;
			r2 = x * x + y * y;
		} while (r2 > 1.0 || r2 == 0.0);
		const double mult = sqrt(-2 * log(r2) / r2);
		st->saved = x * mult;
		st->saved_available = 1;
		ret = y * mult;
	}
	return ret * stddev + mean;
}

float orc_normal_next_f(orc_mt* g, orc_normal_f* st, float mean, float stddev) {
	float ret;
	if (st->saved_available) {
		st->saved_available = 0;
		ret = st->saved;
	} else {
		float x, y, r2;
		do {
			/* float * float, then minus a double literal: exact in double,
			   rounded once to float -- identical to a float subtraction. */
			x = (float)((double)(2.0f * orc_canonical_f(g)) - 1.0);
			y = (float)((double)(2.0f * orc_canonical_f(g)) - 1.0);
			r2 = x * x + y * y;
		} while ((double)r2 > 1.0 || (double)r2 == 0.0);
		const float mult = sqrtf(-2 * logf(r2) / r2);
		st->saved = x * mult;
		st->saved_available = 1;
		ret = y * mult;
	}
	return ret * stddev + mean;
}

/* src/white_noise.hh:23-58 with nthreads=1 (framework=none), and
   src/generator/basic_arma_model.cc:36-48 for sigma = sqrt(var_wn). */
int orc_noise_sequential_d(const orc_mt_params* p, uint32_t seed, double stddev,
                           double* out, size_t n) {
	orc_mt* g = (orc_mt*)malloc(sizeof(orc_mt));
	if (!g) return -2;
	if (orc_mt_seed(g, p, seed)) { free(g); return -1; }
	orc_normal_d st = {0, 0.0};
	for (size_t i = 0; i < n; ++i) out[i] = orc_normal_next_d(g, &st, 0.0, stddev);
	free(g);
	return 0;
}

int orc_noise_sequential_f(const orc_mt_params* p, uint32_t seed, float stddev,
                           float* out, size_t n) {
	orc_mt* g = (orc_mt*)malloc(sizeof(orc_mt));
	if (!g) return -2;
	if (orc_mt_seed(g, p, seed)) { free(g); return -1; }
	orc_normal_f st = {0, 0.0f};
	for (size_t i = 0; i < n; ++i) out[i] = orc_normal_next_f(g, &st, 0.0f, stddev);
	free(g);
	return 0;
}

/* Partition grid: src/generator/ar_model_parallel.cc:51-72 (row-major
   partition order, bounds clipped to the grid); per-partition noise with a
   fresh normal_distribution: src/generator/ar_model_wn.cc:3-18.  Stream p is
   assigned to partition p (deterministic variant of :153-186). */
static int noise_partitioned(const orc_mt_params* p, const uint32_t* seeds, int n_mts,
                             const int shape[3], const int part[3], double stddev,
                             double* out, int nthreads) {
	int np[3];
	for (int d = 0; d < 3; ++d) np[d] = shape[d] / part[d] + (shape[d] % part[d] ? 1 : 0);
	const int ntotal = np[0] * np[1] * np[2];
	if (n_mts < ntotal) return -1;
	int rc = 0;
	/* partitions are independent (one stream each): the OpenMP variant of
	   src/generator/ar_model_parallel.cc:181-186 with a fixed assignment */
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
	{
		orc_mt* g = (orc_mt*)malloc(sizeof(orc_mt));
#pragma omp for schedule(dynamic)
		for (int pid = 0; pid < ntotal; ++pid) {
			const int ijk[3] = {pid / (np[1] * np[2]), (pid / np[2]) % np[1], pid % np[2]};
			int lo[3], hi[3];
			for (int d = 0; d < 3; ++d) {
				lo[d] = ijk[d] * part[d];
				int u = (ijk[d] + 1) * part[d];
				hi[d] = (u < shape[d] ? u : shape[d]);
			}
			if (!g || orc_mt_seed(g, &p[pid], seeds ? seeds[pid] : 0u)) {
#pragma omp atomic write
				rc = -3;
				continue;
			}
			orc_normal_d st = {0, 0.0};
			for (int t = lo[0]; t < hi[0]; ++t)
			for (int x = lo[1]; x < hi[1]; ++x)
			for (int y = lo[2]; y < hi[2]; ++y) {
				out[((size_t)t * shape[1] + x) * shape[2] + y] =
					orc_normal_next_d(g, &st, 0.0, stddev);
			}
		}
		free(g);
	}
	return rc;
}

int orc_noise_partitioned_d(const orc_mt_params* p, const uint32_t* seeds, int n_mts,
                            const int shape[3], const int part[3], double stddev,
                            double* out) {
	return noise_partitioned(p, seeds, n_mts, shape, part, stddev, out, 1);
}

int orc_noise_partitioned_d_omp(const orc_mt_params* p, const uint32_t* seeds, int n_mts,
                                const int shape[3], const int part[3], double stddev,
                                double* out, int nthreads) {
	return noise_partitioned(p, seeds, n_mts, shape, part, stddev, out, nthreads);
}

/* ------------------------------------------------------------------------ */
/* Closed-form ACFs: src/acf.hh:18-62 (selected by src/bits/acf_wrapper.cc). */
/* ------------------------------------------------------------------------ */

void orc_acf(int func, const double delta[3], const int shape[3], double amplitude,
             double velocity, const double alpha[3], const double beta[2], double* out) {
	for (int i = 0; i < shape[0]; ++i)
	for (int j = 0; j < shape[1]; ++j)
	for (int k = 0; k < shape[2]; ++k) {
		const double decay = exp(-(alpha[0] * i * delta[0] + alpha[1] * j * delta[1] +
		                           alpha[2] * k * delta[2]));
		double v;
		if (func == ORC_ACF_STANDING) {
			v = amplitude * decay * cos(velocity * i * delta[0]) *
			    cos(beta[0] * j * delta[1]) * cos(beta[1] * k * delta[2]);
		} else {
			v = amplitude * decay *
			    cos(velocity * i * delta[0] + beta[0] * j * delta[1] + beta[1] * k * delta[2]);
		}
		out[((size_t)i * shape[1] + j) * shape[2] + k] = v;
	}
}

/* ------------------------------------------------------------------------ */
/* Dense FP64 helpers for Yule-Walker (LAPACK/BLAS semantics of             */
/* src/linalg.hh:56-146 and src/linalg.cc:81-123, written out directly).    */
/* ------------------------------------------------------------------------ */

typedef struct { int r, c; double* a; } mat;

static mat* mat_new(int r, int c) {
	mat* m = (mat*)malloc(sizeof(mat));
	m->r = r; m->c = c;
	m->a = (double*)calloc((size_t)r * c > 0 ? (size_t)r * c : 1, sizeof(double));
	return m;
}
static void mat_free(mat* m) { if (m) { free(m->a); free(m); } }
#define MAT(m, i, j) ((m)->a[(size_t)(i) * (m)->c + (j)])

static mat* mat_mul(const mat* A, const mat* B) {
	mat* C = mat_new(A->r, B->c);
	for (int i = 0; i < A->r; ++i)
	for (int k = 0; k < A->c; ++k) {
		const double aik = MAT(A, i, k);
		for (int j = 0; j < B->c; ++j) MAT(C, i, j) += aik * MAT(B, k, j);
	}
	return C;
}

/* In-place inverse via LU with partial pivoting (getrf + getri). */
static int mat_inverse(mat* A) {
	const int n = A->r;
	int* piv = (int*)malloc(sizeof(int) * n);
	double* inv = (double*)calloc((size_t)n * n, sizeof(double));
	double* lu = A->a;
	for (int k = 0; k < n; ++k) {
		int p = k;
		double best = fabs(lu[(size_t)k * n + k]);
		for (int i = k + 1; i < n; ++i) {
			double v = fabs(lu[(size_t)i * n + k]);
			if (v > best) { best = v; p = i; }
		}
		if (best == 0.0) { free(piv); free(inv); return -1; }
		piv[k] = p;
		if (p != k) {
			for (int j = 0; j < n; ++j) {
				double t = lu[(size_t)k * n + j];
				lu[(size_t)k * n + j] = lu[(size_t)p * n + j];
				lu[(size_t)p * n + j] = t;
			}
		}
		const double d = lu[(size_t)k * n + k];
		for (int i = k + 1; i < n; ++i) {
			const double f = lu[(size_t)i * n + k] / d;
			lu[(size_t)i * n + k] = f;
			for (int j = k + 1; j < n; ++j) lu[(size_t)i * n + j] -= f * lu[(size_t)k * n + j];
		}
	}
	/* Solve LU X = P I column by column. */
	double* col = (double*)malloc(sizeof(double) * n);
	for (int c = 0; c < n; ++c) {
		for (int i = 0; i < n; ++i) col[i] = (i == c) ? 1.0 : 0.0;
		for (int k = 0; k < n; ++k) {
			if (piv[k] != k) { double t = col[k]; col[k] = col[piv[k]]; col[piv[k]] = t; }
		}
		for (int i = 0; i < n; ++i) {
			double s = col[i];
			for (int j = 0; j < i; ++j) s -= lu[(size_t)i * n + j] * col[j];
			col[i] = s;
		}
		for (int i = n - 1; i >= 0; --i) {
			double s = col[i];
			for (int j = i + 1; j < n; ++j) s -= lu[(size_t)i * n + j] * col[j];
			col[i] = s / lu[(size_t)i * n + i];
		}
		for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = col[i];
	}
	memcpy(A->a, inv, sizeof(double) * (size_t)n * n);
	free(col); free(inv); free(piv);
	return 0;
}

/* Symmetric solve A x = b (dsysv semantics; A is SPD here). */
static int solve_spd(double* A, double* b, int n) {
	/* Gaussian elimination with partial pivoting; result in b. */
	for (int k = 0; k < n; ++k) {
		int p = k;
		double best = fabs(A[(size_t)k * n + k]);
		for (int i = k + 1; i < n; ++i) {
			double v = fabs(A[(size_t)i * n + k]);
			if (v > best) { best = v; p = i; }
		}
		if (best == 0.0) return -1;
		if (p != k) {
			for (int j = 0; j < n; ++j) {
				double t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)p * n + j];
				A[(size_t)p * n + j] = t;
			}
			double t = b[k]; b[k] = b[p]; b[p] = t;
		}
		for (int i = k + 1; i < n; ++i) {
			const double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
			for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
			b[i] -= f * b[k];
		}
	}
	for (int i = n - 1; i >= 0; --i) {
		double s = b[i];
		for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * b[j];
		b[i] = s / A[(size_t)i * n + i];
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Choi order-recursive Yule-Walker: src/yule_walker.cc:16-539.             */
/* Blocks rho_{i,j,k;d}, R_{a,b}, R_{a,0}, R_{0,b} and the result mapping    */
/* follow :41-122 / :146-224 / :229-309 / :323-353; the recursion :426-539. */
/* ------------------------------------------------------------------------ */

static int yw_len(int a) { return 3 * a * a + 3 * a + 1; }

typedef struct { const double* rho; int n; } acf_view;

static double acf_at(const acf_view* v, int x, int y, int z) {
	x = x < 0 ? -x : x; y = y < 0 ? -y : y; z = z < 0 ? -z : z;
	return v->rho[((size_t)x * v->n + y) * v->n + z];
}

/* Shell ordering shared by the rho vector and the result mapping: the
   first element is (i-d,j-d,k-d); then for e = 1..d six runs of e points
   each.  cb(ctx, x, y, z) receives offsets relative to (i,j,k)-(d,d,d). */
static void rho_vector(const acf_view* v, int i, int j, int k, int d, double* out) {
	int o = 0;
	out[o++] = acf_at(v, i - d, j - d, k - d);
	for (int e = 1; e <= d; ++e) {
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d + e, j - d + q, k - d);
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d + e - q, j - d + e, k - d);
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d, j - d + e, k - d + q);
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d, j - d + e - q, k - d + e);
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d + q, j - d, k - d + e);
		for (int q = 0; q < e; ++q) out[o++] = acf_at(v, i - d + e, j - d, k - d + e - q);
	}
}

/* Row generators of R_{a,b}: row 0 is rho_{a,a,a;b}, then six runs per e. */
static void r_rows(int a, int* I, int* J, int* K) {
	int o = 0;
	I[o] = a; J[o] = a; K[o] = a; ++o;
	for (int e = 1; e <= a; ++e) {
		for (int q = 0; q < e; ++q) { I[o] = a - e;     J[o] = a - q;     K[o] = a;         ++o; }
		for (int q = 0; q < e; ++q) { I[o] = a - e + q; J[o] = a - e;     K[o] = a;         ++o; }
		for (int q = 0; q < e; ++q) { I[o] = a;         J[o] = a - e;     K[o] = a - q;     ++o; }
		for (int q = 0; q < e; ++q) { I[o] = a;         J[o] = a - e + q; K[o] = a - e;     ++o; }
		for (int q = 0; q < e; ++q) { I[o] = a - q;     J[o] = a;         K[o] = a - e;     ++o; }
		for (int q = 0; q < e; ++q) { I[o] = a - e;     J[o] = a;         K[o] = a - e + q; ++o; }
	}
}

static mat* R_matrix(const acf_view* v, int a, int b) {
	const int nr = yw_len(a), nc = yw_len(b);
	mat* R = mat_new(nr, nc);
	int* I = (int*)malloc(sizeof(int) * nr * 3);
	int* J = I + nr;
	int* K = J + nr;
	r_rows(a, I, J, K);
	for (int r = 0; r < nr; ++r) rho_vector(v, I[r], J[r], K[r], b, &MAT(R, r, 0));
	free(I);
	return R;
}

static mat* R_vector_b0(const acf_view* v, int a) {
	const int nr = yw_len(a);
	mat* R = mat_new(nr, 1);
	int* I = (int*)malloc(sizeof(int) * nr * 3);
	int* J = I + nr;
	int* K = J + nr;
	r_rows(a, I, J, K);
	for (int r = 0; r < nr; ++r) MAT(R, r, 0) = acf_at(v, I[r], J[r], K[r]);
	free(I);
	return R;
}

static mat* R_vector_a0(const acf_view* v, int b) {
	mat* R = mat_new(yw_len(b), 1);
	rho_vector(v, 0, 0, 0, b, R->a);
	return R;
}

static double vdot(const mat* a, const mat* b) {
	double s = 0.0;
	for (int i = 0; i < a->r * a->c; ++i) s += a->a[i] * b->a[i];
	return s;
}

/* result_array: src/yule_walker.cc:323-353. beta[d] for d = 1..order-1. */
static void result_array(mat** beta, int order, double* out) {
	const size_t o2 = (size_t)order * order;
	memset(out, 0, sizeof(double) * o2 * order);
#define RES(a, b, c) out[(size_t)(a) * o2 + (size_t)(b) * order + (c)]
	for (int d = 1; d < order; ++d) {
		const double* p = beta[d]->a;
		RES(d, d, d) = *p++;
		for (int e = 1; e <= d; ++e) {
			for (int q = 0; q < e; ++q) RES(d - e, d - q, d) = *p++;
			for (int q = 0; q < e; ++q) RES(d - e + q, d - e, d) = *p++;
			for (int q = 0; q < e; ++q) RES(d, d - e, d - q) = *p++;
			for (int q = 0; q < e; ++q) RES(d, d - e + q, d - e) = *p++;
			for (int q = 0; q < e; ++q) RES(d - q, d, d - e) = *p++;
			for (int q = 0; q < e; ++q) RES(d - e, d, d - e + q) = *p++;
		}
	}
#undef RES
}

/* src/chop.cc:4-26 */
void orc_chop_right(const double* phi, const int shape[3], double eps, int out[3]) {
	const int n0 = shape[0], n1 = shape[1], n2 = shape[2];
#define PH(a, b, c) phi[((size_t)(a) * n1 + (b)) * n2 + (c)]
	int k = n2 - 1;
	for (; k >= 1; --k) {
		int all_small = 1;
		for (int a = 0; a < n0 && all_small; ++a)
			for (int b = 0; b < n1; ++b) if (!(fabs(PH(a, b, k)) < eps)) { all_small = 0; break; }
		if (!all_small) break;
	}
	int j = n1 - 1;
	for (; j >= 1; --j) {
		int all_small = 1;
		for (int a = 0; a < n0 && all_small; ++a)
			for (int c = 0; c <= k; ++c) if (!(fabs(PH(a, j, c)) < eps)) { all_small = 0; break; }
		if (!all_small) break;
	}
	int i = n0 - 1;
	for (; i >= 1; --i) {
		int all_small = 1;
		for (int b = 0; b <= j && all_small; ++b)
			for (int c = 0; c <= k; ++c) if (!(fabs(PH(i, b, c)) < eps)) { all_small = 0; break; }
		if (!all_small) break;
	}
#undef PH
	out[0] = i + 1; out[1] = j + 1; out[2] = k + 1;
}

int orc_yw_choi(const double* acf, int n, int max_order, double var_eps,
                int determine_order, int chop, double chop_eps,
                double* phi_out, int order_out[3], double* varwn_solver) {
	/* Constructor: rho = gamma / gamma000, variance = gamma000
	   (src/yule_walker.cc:403-417); max_order clipped to n-1 (:415-418). */
	const size_t n3 = (size_t)n * n * n;
	double* rho = (double*)malloc(sizeof(double) * n3);
	for (size_t i = 0; i < n3; ++i) rho[i] = acf[i] / acf[0];
	const double variance = acf[0];
	int M = max_order < n - 1 ? max_order : n - 1;
	if (M < 2) { free(rho); return -1; }
	acf_view v = {rho, n};
	const int MP = M + 3;
#define PHI(i, j, k) phi_map[((size_t)(i) * MP + (j)) * MP + (k)]
	mat** phi_map = (mat**)calloc((size_t)MP * MP * MP, sizeof(mat*));
	mat** Pi_l = (mat**)calloc(MP, sizeof(mat*));
	mat** Pi_lp1 = (mat**)calloc(MP, sizeof(mat*));
	mat** Theta = (mat**)calloc(MP, sizeof(mat*));
	int rc = 0;

	mat* R11 = R_matrix(&v, 1, 1);
	mat* R10 = R_vector_b0(&v, 1);
	mat* R01 = R_vector_a0(&v, 1);
	mat* Rsup11 = mat_new(R11->r, R11->c);
	memcpy(Rsup11->a, R11->a, sizeof(double) * R11->r * R11->c);
	if (mat_inverse(Rsup11)) { rc = -2; goto done; }
	Pi_l[1] = mat_mul(Rsup11, R10);
	double lambda = 1.0 - vdot(R01, Pi_l[1]);
	double var0 = variance;
	double var = variance * lambda;
	{
		mat* R12 = R_matrix(&v, 1, 2);
		PHI(2, 1, 0) = mat_mul(Rsup11, R12);
		mat_free(R12);
	}
	int changed = 0;
	int l = 1;
	do {
		++l;
		var0 = var;
		mat* Rll = R_matrix(&v, l, l);
		mat* Rl0 = R_vector_b0(&v, l);
		for (int m = 1; m <= l - 1; ++m) {
			mat* Rlm = R_matrix(&v, l, m);
			mat* s1 = mat_mul(Rlm, PHI(l, m, 0));
			mat* s2 = mat_mul(Rlm, Pi_l[m]);
			for (int q = 0; q < Rll->r * Rll->c; ++q) Rll->a[q] -= s1->a[q];
			for (int q = 0; q < Rl0->r; ++q) Rl0->a[q] -= s2->a[q];
			mat_free(s1); mat_free(s2); mat_free(Rlm);
		}
		/* Rll now holds Theta_l, Rl0 holds h_l. */
		if (mat_inverse(Rll)) { mat_free(Rll); mat_free(Rl0); rc = -2; goto done; }
		mat_free(Theta[l]);
		Theta[l] = Rll;
		mat_free(Pi_lp1[l]);
		Pi_lp1[l] = mat_mul(Rll, Rl0);
		for (int a = 1; a <= l - 1; ++a) {
			mat* t = mat_mul(PHI(l, a, 0), Pi_lp1[l]);
			mat* nv = mat_new(Pi_l[a]->r, 1);
			for (int q = 0; q < nv->r; ++q) nv->a[q] = Pi_l[a]->a[q] - t->a[q];
			mat_free(t);
			mat_free(Pi_lp1[a]);
			Pi_lp1[a] = nv;
		}
		lambda -= vdot(Rl0, Pi_lp1[l]);
		mat_free(Rl0);
		var = variance * lambda;
		if (l < M) {
			mat* R1 = R_matrix(&v, 1, l + 1);
			mat_free(PHI(2, 1, l - 1));
			PHI(2, 1, l - 1) = mat_mul(Rsup11, R1);
			mat_free(R1);
			for (int m = 2; m <= l; ++m) {
				mat* Rm = R_matrix(&v, m, l + 1);
				for (int q = 1; q <= m - 1; ++q) {
					mat* Rmq = R_matrix(&v, m, q);
					mat* s3 = mat_mul(Rmq, PHI(m, q, l - m + 1));
					for (int z = 0; z < Rm->r * Rm->c; ++z) Rm->a[z] -= s3->a[z];
					mat_free(s3); mat_free(Rmq);
				}
				mat_free(PHI(m + 1, m, l - m));
				PHI(m + 1, m, l - m) = mat_mul(Theta[m], Rm);
				mat_free(Rm);
				for (int a = 1; a <= m - 1; ++a) {
					mat* t = mat_mul(PHI(m, a, 0), PHI(m + 1, m, l - m));
					const mat* src = PHI(m, a, l - m + 1);
					mat* nv = mat_new(src->r, src->c);
					for (int z = 0; z < nv->r * nv->c; ++z) nv->a[z] = src->a[z] - t->a[z];
					mat_free(t);
					mat_free(PHI(m + 1, a, l - m));
					PHI(m + 1, a, l - m) = nv;
				}
			}
		}
		/* blitz::cycleArrays(Pi_l, Pi_lp1) */
		mat** tmp = Pi_l; Pi_l = Pi_lp1; Pi_lp1 = tmp;
		const int not_changed = determine_order && fabs(var - var0) < var_eps;
		changed = !not_changed;
	} while (l < M && changed);
	{
		double* res = (double*)malloc(sizeof(double) * (size_t)l * l * l);
		if (changed) {
			*varwn_solver = var;
			result_array(Pi_l, l, res);
		} else {
			*varwn_solver = var0;
			result_array(Pi_lp1, l, res);
		}
		int shp[3] = {l, l, l};
		if (chop) {
			orc_chop_right(res, shp, chop_eps, shp);
		}
		for (int a = 0; a < shp[0]; ++a)
		for (int b = 0; b < shp[1]; ++b)
		for (int c = 0; c < shp[2]; ++c)
			phi_out[((size_t)a * shp[1] + b) * shp[2] + c] = res[((size_t)a * l + b) * l + c];
		order_out[0] = shp[0]; order_out[1] = shp[1]; order_out[2] = shp[2];
		free(res);
	}
done:
	for (size_t q = 0; q < (size_t)MP * MP * MP; ++q) mat_free(phi_map[q]);
	for (int q = 0; q < MP; ++q) { mat_free(Pi_l[q]); mat_free(Pi_lp1[q]); mat_free(Theta[q]); }
	free(phi_map); free(Pi_l); free(Pi_lp1); free(Theta);
	mat_free(R11); mat_free(R10); mat_free(R01); mat_free(Rsup11);
	free(rho);
#undef PHI
	return rc;
}

/* Gauss elimination path: src/generator/ar_model.cc:97-135 with the AC
   matrix of src/generator/voodoo.cc:58-104 (element
   [(i,j,k),(i',j',k')] = gamma(|i-i'|,|j-j'|,|k-k'|)). */
int orc_yw_gauss(const double* acf, const int acf_shape[3], const int order[3],
                 double* phi_out) {
	for (int d = 0; d < 3; ++d) if (order[d] > acf_shape[d]) return -1;
	const int P = order[0] * order[1] * order[2];
	const int m = P - 1;
	if (m <= 0) { phi_out[0] = 0.0; return 0; }
	double* A = (double*)malloc(sizeof(double) * (size_t)m * m);
	double* b = (double*)malloc(sizeof(double) * (size_t)m);
#define G(a, c, e) acf[((size_t)(a) * acf_shape[1] + (c)) * acf_shape[2] + (e)]
	for (int r = 0; r < P; ++r) {
		const int ri = r / (order[1] * order[2]), rj = (r / order[2]) % order[1], rk = r % order[2];
		for (int c = 0; c < P; ++c) {
			const int ci = c / (order[1] * order[2]), cj = (c / order[2]) % order[1], ck = c % order[2];
			const double val = G(abs(ri - ci), abs(rj - cj), abs(rk - ck));
			if (r >= 1 && c >= 1) A[(size_t)(r - 1) * m + (c - 1)] = val;
			if (r >= 1 && c == 0) b[r - 1] = val;
		}
	}
#undef G
	const int rc = solve_spd(A, b, m);
	phi_out[0] = 0.0;
	for (int i = 0; i < m; ++i) phi_out[i + 1] = b[i];
	free(A); free(b);
	return rc;
}

/* src/generator/ar_model.cc:55-61 */
double orc_white_noise_variance(const double* acf, const int acf_shape[3],
                                const double* phi, const int phi_shape[3]) {
	double s = 0.0;
	for (int a = 0; a < phi_shape[0]; ++a)
	for (int b = 0; b < phi_shape[1]; ++b)
	for (int c = 0; c < phi_shape[2]; ++c)
		s += phi[((size_t)a * phi_shape[1] + b) * phi_shape[2] + c] *
		     acf[((size_t)a * acf_shape[1] + b) * acf_shape[2] + c];
	return acf[0] - s;
}

/* src/generator/ar_model.cc:23-50 */
void orc_partition_shape(const int partition[3], const int grid[3], const int order[3],
                         int parallelism, int out[3]) {
	if (partition[0] * partition[1] * partition[2] > 0) {
		for (int d = 0; d < 3; ++d) out[d] = partition[d];
		return;
	}
	int npar = 7 * (int)cbrt((double)parallelism);
	if (npar < 1) npar = 1;
	for (int d = 0; d < 3; ++d) {
		int g1 = 2 * order[d] > 10 ? 2 * order[d] : 10;
		int g2 = grid[d] / npar + (grid[d] % npar ? 1 : 0);
		int mn = g1 < g2 ? g1 : g2;
		out[d] = mn + abs(g1 - g2) / 2;
	}
}

/* ------------------------------------------------------------------------ */
/* AR recurrence: src/generator/ar_model_surf.cc:3-36 (same loop body as    */
/* src/kernels/ar_generate_surface.cl:6-49).                                 */
/* ------------------------------------------------------------------------ */

void orc_ar_generate(double* zeta, const int shape[3], const double* phi,
                     const int pshape[3], const int lo[3], const int hi[3]) {
	const int n1 = shape[1], n2 = shape[2];
	const int p0 = pshape[0], p1 = pshape[1], p2 = pshape[2];
	for (int t = lo[0]; t <= hi[0]; ++t)
	for (int x = lo[1]; x <= hi[1]; ++x)
	for (int y = lo[2]; y <= hi[2]; ++y) {
		const int m1 = t + 1 < p0 ? t + 1 : p0;
		const int m2 = x + 1 < p1 ? x + 1 : p1;
		const int m3 = y + 1 < p2 ? y + 1 : p2;
		double sum = 0;
		for (int k = 0; k < m1; ++k)
		for (int i = 0; i < m2; ++i)
		for (int j = 0; j < m3; ++j)
			sum += phi[((size_t)k * p1 + i) * p2 + j] *
			       zeta[((size_t)(t - k) * n1 + (x - i)) * n2 + (y - j)];
		zeta[((size_t)t * n1 + x) * n2 + y] += sum;
	}
}

void orc_ar_generate_f(float* zeta, const int shape[3], const float* phi,
                       const int pshape[3], const int lo[3], const int hi[3]) {
	const int n1 = shape[1], n2 = shape[2];
	const int p0 = pshape[0], p1 = pshape[1], p2 = pshape[2];
	for (int t = lo[0]; t <= hi[0]; ++t)
	for (int x = lo[1]; x <= hi[1]; ++x)
	for (int y = lo[2]; y <= hi[2]; ++y) {
		const int m1 = t + 1 < p0 ? t + 1 : p0;
		const int m2 = x + 1 < p1 ? x + 1 : p1;
		const int m3 = y + 1 < p2 ? y + 1 : p2;
		float sum = 0;
		for (int k = 0; k < m1; ++k)
		for (int i = 0; i < m2; ++i)
		for (int j = 0; j < m3; ++j)
			sum += phi[((size_t)k * p1 + i) * p2 + j] *
			       zeta[((size_t)(t - k) * n1 + (x - i)) * n2 + (y - j)];
		zeta[((size_t)t * n1 + x) * n2 + y] += sum;
	}
}

/* Partition DAG order (src/generator/ar_model_parallel.cc:151-196): any
   topological order of the 7-predecessor DAG; row-major is one. */
void orc_ar_generate_dag(double* zeta, const int shape[3], const double* phi,
                         const int pshape[3], const int part[3]) {
	int np[3];
	for (int d = 0; d < 3; ++d) np[d] = shape[d] / part[d] + (shape[d] % part[d] ? 1 : 0);
	for (int a = 0; a < np[0]; ++a)
	for (int b = 0; b < np[1]; ++b)
	for (int c = 0; c < np[2]; ++c) {
		const int ijk[3] = {a, b, c};
		int lo[3], hi[3];
		for (int d = 0; d < 3; ++d) {
			lo[d] = ijk[d] * part[d] < shape[d] ? ijk[d] * part[d] : shape[d];
			int u = (ijk[d] + 1) * part[d];
			hi[d] = (u < shape[d] ? u : shape[d]) - 1;
		}
		orc_ar_generate(zeta, shape, phi, pshape, lo, hi);
	}
}

/* The partition DAG with OpenMP (src/generator/ar_model_parallel.cc:151-196
   runs every partition whose 7 predecessors are done): partitions of one
   anti-diagonal level a + b + c = L depend only on lower levels, so each
   level runs as one parallel loop.  Every point is computed by the same
   loop body, so the result equals orc_ar_generate_dag bit for bit. */
void orc_ar_generate_dag_omp(double* zeta, const int shape[3], const double* phi,
                             const int pshape[3], const int part[3], int nthreads) {
	int np[3];
	for (int d = 0; d < 3; ++d) np[d] = shape[d] / part[d] + (shape[d] % part[d] ? 1 : 0);
	const int ntotal = np[0] * np[1] * np[2];
	for (int L = 0; L <= np[0] + np[1] + np[2] - 3; ++L) {
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
		for (int pid = 0; pid < ntotal; ++pid) {
			const int ijk[3] = {pid / (np[1] * np[2]), (pid / np[2]) % np[1], pid % np[2]};
			if (ijk[0] + ijk[1] + ijk[2] != L) continue;
			int lo[3], hi[3];
			for (int d = 0; d < 3; ++d) {
				lo[d] = ijk[d] * part[d];
				int u = (ijk[d] + 1) * part[d];
				hi[d] = (u < shape[d] ? u : shape[d]) - 1;
			}
			orc_ar_generate(zeta, shape, phi, pshape, lo, hi);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* GSL gsl_stats_variance (statistics/mean_source.c, variance_source.c):    */
/* running long-double mean, then running long-double mean of squared        */
/* deviations, times n/(n-1).  Called by src/stats/statistics.hh:35-44.      */
/* ------------------------------------------------------------------------ */

double orc_gsl_variance_d(const double* data, size_t n) {
	long double mean = 0;
	for (size_t i = 0; i < n; ++i) mean += (data[i] - mean) / (i + 1);
	const double m = (double)mean;
	long double var = 0;
	for (size_t i = 0; i < n; ++i) {
		const long double delta = (data[i] - m);
		var += (delta * delta - var) / (i + 1);
	}
	const double v = (double)var;
	return v * ((double)n / (double)(n - 1));
}

float orc_gsl_variance_f(const float* data, size_t n) {
	long double mean = 0;
	for (size_t i = 0; i < n; ++i) mean += (data[i] - mean) / (i + 1);
	const double m = (double)mean;
	long double var = 0;
	for (size_t i = 0; i < n; ++i) {
		const long double delta = ((double)data[i] - m);
		var += (delta * delta - var) / (i + 1);
	}
	const double v = (double)var;
	return (float)(v * ((double)n / (double)(n - 1)));
}

/* Post-rescale factor of src/generator/basic_arma_model.cc:102-106: the
   variance reads a CONTIGUOUS run of ceil(nt/2)ceil(nx/2)ceil(ny/2)
   elements starting at zeta(nt/2, nx/2, ny/2). */
double orc_rescale_factor(const double* zeta, const int shape[3], double gamma0) {
	const size_t start = ((size_t)(shape[0] / 2) * shape[1] + shape[1] / 2) * shape[2] + shape[2] / 2;
	const size_t cnt = (size_t)(shape[0] - shape[0] / 2) * (shape[1] - shape[1] / 2) *
	                   (shape[2] - shape[2] / 2);
	return sqrt(gamma0 / orc_gsl_variance_d(zeta + start, cnt));
}

double orc_rescale_factor_f(const float* zeta, const int shape[3], double gamma0) {
	const size_t start = ((size_t)(shape[0] / 2) * shape[1] + shape[1] / 2) * shape[2] + shape[2] / 2;
	const size_t cnt = (size_t)(shape[0] - shape[0] / 2) * (shape[1] - shape[1] / 2) *
	                   (shape[2] - shape[2] / 2);
	return sqrtf((float)gamma0 / orc_gsl_variance_f(zeta + start, cnt));
}

/* ------------------------------------------------------------------------ */
/* Velocity potential, linear solver.                                       */
/* ------------------------------------------------------------------------ */

/* src/derivative.hh:8-73 along t. */
void orc_derivative_t(const double* zeta, const int shape[3], double dt, int idx_t,
                      double* out) {
	const int nt = shape[0], nx = shape[1], ny = shape[2];
	const size_t s = (size_t)nx * ny;
	const double* f0 = zeta + (size_t)idx_t * s;
	if (nt - 1 < 2) {
		memcpy(out, f0, sizeof(double) * s);
		return;
	}
	for (size_t q = 0; q < s; ++q) {
		double r;
		if (idx_t == 0) {
			r = 0.5 * (-f0[q + 2 * s] + 4.0 * f0[q + s] - 3.0 * f0[q]) / dt;
		} else if (idx_t == nt - 1) {
			r = 0.5 * (3.0 * f0[q] - 4.0 * f0[q - s] + f0[q - 2 * s]) / dt;
		} else {
			r = 0.5 * (f0[q + s] - f0[q - s]) / dt;
		}
		out[q] = r;
	}
}

/* src/factor_waves.cc:8-89; z(x,y) = slice[x*ny+y] scanned along x, or
   along y when transposed.  g = 9.82 (src/physical_constants.hh:27). */
static void factor_waves_dim(const double* slice, int nx, int ny, int transpose,
                             int t, double dt, double* kmin, double* kmax) {
	const double g = 9.82;
	const double two_pi = 2.0 * M_PI;
	double min_k = DBL_MAX, max_k = DBL_MIN;
	/* dimension sizes of the (possibly transposed) view */
	const int n0 = transpose ? ny : nx;
	const int n1 = transpose ? nx : ny;
#define Z(a, b) (transpose ? slice[(size_t)(b) * ny + (a)] : slice[(size_t)(a) * ny + (b)])
	int have1 = 0, j = 0;
	double e1_period = 0, e1_elev = 0, e2_period = 0, e2_elev = 0;
	for (int y = 0; y < n1; y++) {
		for (int x = 1; x < n0 - 1; x++) {
			const double z0 = Z(x, y);
			const double z1 = Z(x - 1, y);
			const double z2 = Z(x + 1, y);
			const double dw1 = z0 - z1;
			const double dw2 = z0 - z2;
			if (!((dw1 > 0.0) ^ (dw2 > 0.0))) {
				const double a = -0.5 * (dw1 + dw2) / (dt * dt);
				const double b = dw1 / dt - a * dt * (2.0 * t - 1.0);
				const double c = z0 - t * dt * (a * t * dt + b);
				const double Tex = -0.5 * b / a;
				const double Wex = c + Tex * (b + a * Tex);
				if (!have1) {
					e1_period = Tex; e1_elev = Wex; have1 = 1;
					continue;
				}
				if (!((e1_elev > 0.0) ^ (Wex > 0.0))) {
					if (fabs(e1_elev) < fabs(Wex)) { e1_period = Tex; e1_elev = Wex; }
				} else {
					if (j > 0) {
						const double w = two_pi / fabs((e1_period - e2_period) * 2.0);
						const double k = w * w / g;
						if (k < min_k) min_k = k;
						if (k > max_k) max_k = k;
					}
					e2_period = e1_period; e2_elev = e1_elev;
					e1_period = Tex; e1_elev = Wex;
					j++;
				}
			}
		}
	}
#undef Z
	(void)e2_elev;
	*kmin = min_k;
	*kmax = max_k;
}

void orc_factor_waves(const double* slice, int nx, int ny, int t, double dt, double out[4]) {
	factor_waves_dim(slice, nx, ny, 0, t, dt, &out[0], &out[2]);
	factor_waves_dim(slice, nx, ny, 1, t, dt, &out[1], &out[3]);
}

/* src/interpolate.cc:3-29 and src/interpolate.hh:17-33 (plane through three
   grid points evaluated at a fourth). */
static double interp3(const double* f, int ny, int a0, int a1, int b0, int b1, int c0,
                      int c1, int x0, int x1) {
	const double fa = f[(size_t)a0 * ny + a1];
	const double fb = f[(size_t)b0 * ny + b1];
	const double fc = f[(size_t)c0 * ny + c1];
	const double den = (double)(b0 * a1 - c0 * a1 - a0 * b1 + c0 * b1 + a0 * c1 - b0 * c1);
	const double p0 = (fb * a1 - fc * a1 - fa * b1 + fc * b1 + fa * c1 - fb * c1) / den;
	const double p1 = (fc * (-a0 + b0) + fb * (a0 - c0) + fa * (-b0 + c0)) /
	                  (double)(c0 * (a1 - b1) + a0 * (b1 - c1) + b0 * (-a1 + c1));
	const double p2 = (fc * b0 * a1 - fb * c0 * a1 - fc * a0 * b1 + fa * c0 * b1 +
	                   fb * a0 * c1 - fa * b0 * c1) / den;
	return p0 * x0 + p1 * x1 + p2;
}

/* src/velocity/linear_solver.cc:123-184 (window), src/blitz.hh:54-71
   (rotate by shape/2 via swaps). */
void orc_window(int nx, int ny, double wnmax_x, double wnmax_y, double z, double depth,
                double* out) {
	const double two_pi = 2.0 * M_PI;
	const double h = depth;
	const double du = wnmax_x / (double)(nx - 1 > 1 ? nx - 1 : 1);
	const double dv = wnmax_y / (double)(ny - 1 > 1 ? ny - 1 : 1);
	double* w = (double*)malloc(sizeof(double) * (size_t)nx * ny);
	for (int i = 0; i < nx; ++i) {
		for (int j = 0; j < ny; ++j) {
			const double u = 0.0 + du * i;
			const double v = 0.0 + dv * j;
			const double l = two_pi * sqrt(u * u + v * v);
			const double explz = exp(-l * z);
			const double exp2lh = exp(-2.0 * l * h);
			const double numerator = 1.0 + explz * explz * exp2lh;
			const double denominator = explz * (1.0 + exp2lh);
			w[(size_t)i * ny + j] = two_pi * 2.0 * numerator / denominator;
		}
	}
	{
		double acc = 0.0;
		int cnt = 0;
		if (isfinite(w[(size_t)1 * ny + 1])) { acc += w[(size_t)1 * ny + 1]; ++cnt; }
		w[0] = acc / cnt;
	}
	for (int i = 1; i < nx; ++i) {
		if (!isfinite(w[(size_t)i * ny])) {
			w[(size_t)i * ny] = interp3(w, ny, i - 1, 1, i, 1, i - 1, 2, i, 0);
		}
	}
	for (int j = 1; j < ny; ++j) {
		if (!isfinite(w[j])) {
			w[j] = interp3(w, ny, 1, j - 1, 1, j, 2, j - 1, 0, j);
		}
	}
	for (size_t q = 0; q < (size_t)nx * ny; ++q) if (!isfinite(w[q])) w[q] = 0.0;
	/* rotate: first swap columns j <-> (j+ny/2)%ny for j < ny/2, then rows. */
	const int mr = nx / 2, mc = ny / 2;
	for (int i = 0; i < nx; ++i)
		for (int j = 0; j < mc; ++j) {
			double tmp = w[(size_t)i * ny + j];
			w[(size_t)i * ny + j] = w[(size_t)i * ny + (j + mc) % ny];
			w[(size_t)i * ny + (j + mc) % ny] = tmp;
		}
	for (int i = 0; i < mr; ++i)
		for (int j = 0; j < ny; ++j) {
			double tmp = w[(size_t)i * ny + j];
			w[(size_t)i * ny + j] = w[(size_t)((i + mr) % nx) * ny + j];
			w[(size_t)((i + mr) % nx) * ny + j] = tmp;
		}
	memcpy(out, w, sizeof(double) * (size_t)nx * ny);
	free(w);
}

/* Complex DFT along one strided line; sign -1 = GSL forward, +1 = backward
   (unnormalised).  Radix-2 when n is a power of two, direct sum otherwise. */
void orc_fft_1d(double* re, double* im, int n, int stride, int sign) {
	double* xr = (double*)malloc(sizeof(double) * n);
	double* xi = (double*)malloc(sizeof(double) * n);
	for (int q = 0; q < n; ++q) { xr[q] = re[(size_t)q * stride]; xi[q] = im[(size_t)q * stride]; }
	if (n > 1 && (n & (n - 1)) == 0) {
		/* iterative radix-2, bit-reversed input */
		for (int i = 1, j = 0; i < n; ++i) {
			int bit = n >> 1;
			for (; j & bit; bit >>= 1) j ^= bit;
			j ^= bit;
			if (i < j) {
				double t = xr[i]; xr[i] = xr[j]; xr[j] = t;
				t = xi[i]; xi[i] = xi[j]; xi[j] = t;
			}
		}
		for (int len = 2; len <= n; len <<= 1) {
			const int half = len >> 1;
			for (int k = 0; k < half; ++k) {
				const double ang = sign * 2.0 * M_PI * (double)k / (double)len;
				const double wr = cos(ang), wi = sin(ang);
				for (int s = 0; s < n; s += len) {
					const int a = s + k, b = s + k + half;
					const double tr = xr[b] * wr - xi[b] * wi;
					const double ti = xr[b] * wi + xi[b] * wr;
					xr[b] = xr[a] - tr; xi[b] = xi[a] - ti;
					xr[a] += tr; xi[a] += ti;
				}
			}
		}
		for (int q = 0; q < n; ++q) { re[(size_t)q * stride] = xr[q]; im[(size_t)q * stride] = xi[q]; }
	} else {
		for (int k = 0; k < n; ++k) {
			double sr = 0.0, si = 0.0;
			for (int j = 0; j < n; ++j) {
				const long long jk = ((long long)j * k) % n;
				const double ang = sign * 2.0 * M_PI * (double)jk / (double)n;
				const double c = cos(ang), s = sin(ang);
				sr += xr[j] * c - xi[j] * s;
				si += xr[j] * s + xi[j] * c;
			}
			re[(size_t)k * stride] = sr;
			im[(size_t)k * stride] = si;
		}
	}
	free(xr); free(xi);
}

/* 2-D transform as strided 1-D transforms along each dimension
   (src/apmath/fourier_gsl.hh:140-184). */
void orc_fft2(double* re, double* im, int nx, int ny, int sign) {
	for (int j = 0; j < ny; ++j) orc_fft_1d(re + j, im + j, nx, ny, sign);
	for (int i = 0; i < nx; ++i) orc_fft_1d(re + (size_t)i * ny, im + (size_t)i * ny, ny, 1, sign);
}

/* Edge fix-up of src/velocity/basic_solver.cc:61-68. */
void orc_interp_edges(double* res, int nx, int ny) {
	res[0] = interp3(res, ny, 1, 1, 1, 2, 2, 1, 0, 0);
	for (int k = 1; k < nx; ++k)
		res[(size_t)k * ny] = interp3(res, ny, k - 1, 1, k, 1, k - 1, 2, k, 0);
	for (int l = 1; l < ny; ++l)
		res[l] = interp3(res, ny, 1, l - 1, 1, l, 2, l - 1, 0, l);
}

/* src/derivative.hh:8-73 along x (dim 1) or y (dim 2) at slice idx_t:
   central differences inside, second-order one-sided at the borders, the raw
   slice when the extent is below 3. */
void orc_derivative_xy(const double* zeta, const int shape[3], double d, int dim, int idx_t,
                       double* out) {
	const int nx = shape[1], ny = shape[2];
	const size_t s = (size_t)nx * ny;
	const double* f = zeta + (size_t)idx_t * s;
	const int n = dim == 1 ? nx : ny;
	const size_t step = dim == 1 ? (size_t)ny : 1;
	if (n - 1 < 2) {
		memcpy(out, f, sizeof(double) * s);
		return;
	}
	for (int i = 0; i < nx; ++i) {
		for (int j = 0; j < ny; ++j) {
			const size_t q = (size_t)i * ny + j;
			const int c = dim == 1 ? i : j;
			double r;
			if (c == 0) {
				r = 0.5 * (-f[q + 2 * step] + 4.0 * f[q + step] - 3.0 * f[q]) / d;
			} else if (c == n - 1) {
				r = 0.5 * (3.0 * f[q] - 4.0 * f[q - step] + f[q - 2 * step]) / d;
			} else {
				r = 0.5 * (f[q + step] - f[q - step]) / d;
			}
			out[q] = r;
		}
	}
}

/* The multiplier of High_amplitude_solver::precompute
   (src/velocity/high_amplitude_solver.cc:14-22) at one point:
   F = zeta_t / (i ((zx + zy)/s - zx - zy) - 1/s),  s = sqrt(1 + zx^2 + zy^2),
   with zeta_t = +d zeta/dt.  The division is C's complex division (the same
   library routine std::complex<double> division calls under g++). */
static void ha_multiplier(double zt, double zx, double zy, double* re, double* im) {
	const double s = sqrt(1.0 + zx * zx + zy * zy);
	const double e = (zx + zy) / s - zx - zy;
	const _Complex double den = __builtin_complex(0.0 - 1.0 / s, e);
	const _Complex double f = __builtin_complex(zt, 0.0) / den;
	*re = __real__ f;
	*im = __imag__ f;
}

/* Velocity_potential_solver::operator() (src/velocity/basic_solver.cc:39-77)
   with Linear_solver (src/velocity/linear_solver.cc:20-121) or, for
   solver 1, High_amplitude_solver (src/velocity/high_amplitude_solver.cc:6-23,
   the same transforms of a complex multiplier), and the wave number range of
   src/velocity/basic_solver.cc:95-112.  Output layout (nt_domain, nz, nx, ny)
   row-major.  wnmax_out (optional) receives the per-slice
   (1/kmin_x, 1/kmin_y). */
static int potential(const double* zeta, const int shape[3], const double delta[3],
                     double depth, const double lbound[2], const double ubound[2],
                     const int npoints[2], double* out, double* wnmax_out, int solver,
                     int nthreads) {
	const int nx = shape[1], ny = shape[2];
	const size_t s = (size_t)nx * ny;
	if (lbound[1] < -depth) return -1;
	const double dtd = (ubound[0] - lbound[0]) / (double)(npoints[0] - 1 > 1 ? npoints[0] - 1 : 1);
	const double dzd = (ubound[1] - lbound[1]) / (double)(npoints[1] - 1 > 1 ? npoints[1] - 1 : 1);
	double* zt = (double*)malloc(sizeof(double) * s);
	double* zx = (double*)malloc(sizeof(double) * s);
	double* zy = (double*)malloc(sizeof(double) * s);
	double* fr = (double*)malloc(sizeof(double) * s);
	double* fi = (double*)malloc(sizeof(double) * s);
	for (int i = 0; i < npoints[0]; ++i) {
		const double tval = lbound[0] + dtd * i;
		const int idx_t = (int)tval;
		orc_derivative_t(zeta, shape, delta[0], idx_t, zt);
		if (solver == 1) {
			orc_derivative_xy(zeta, shape, delta[1], 1, idx_t, zx);
			orc_derivative_xy(zeta, shape, delta[2], 2, idx_t, zy);
			for (size_t q = 0; q < s; ++q) ha_multiplier(zt[q], zx[q], zy[q], fr + q, fi + q);
		} else {
			for (size_t q = 0; q < s; ++q) { fr[q] = -zt[q]; fi[q] = 0.0; }
		}
		double kr[4];
		const double tgrid = delta[0] * idx_t;
		orc_factor_waves(zeta + (size_t)idx_t * s, nx, ny, (int)tgrid, delta[0], kr);
		const double wx = 1.0 / kr[0], wy = 1.0 / kr[1];
		if (wnmax_out) { wnmax_out[2 * i] = wx; wnmax_out[2 * i + 1] = wy; }
		orc_fft2(fr, fi, nx, ny, -1);
		/* the levels of one slice are independent: OpenMP over z as in
		   src/velocity/basic_solver.cc:55-57 */
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
		{
			double* wr = (double*)malloc(sizeof(double) * s);
			double* wi = (double*)malloc(sizeof(double) * s);
			double* win = (double*)malloc(sizeof(double) * s);
#pragma omp for schedule(static)
			for (int j = 0; j < npoints[1]; ++j) {
				const double z = lbound[1] + dzd * j;
				orc_window(nx, ny, wx, wy, z, depth, win);
				for (size_t q = 0; q < s; ++q) { wr[q] = fr[q] * win[q]; wi[q] = fi[q] * win[q]; }
				orc_fft2(wr, wi, nx, ny, +1);
				double* res = out + ((size_t)i * npoints[1] + j) * s;
				for (size_t q = 0; q < s; ++q) res[q] = wr[q] / (double)s;
				orc_interp_edges(res, nx, ny);
			}
			free(wr); free(wi); free(win);
		}
	}
	free(zt); free(zx); free(zy); free(fr); free(fi);
	return 0;
}

int orc_linear_potential(const double* zeta, const int shape[3], const double delta[3],
                         double depth, const double lbound[2], const double ubound[2],
                         const int npoints[2], double* out, double* wnmax_out) {
	return potential(zeta, shape, delta, depth, lbound, ubound, npoints, out, wnmax_out, 0, 1);
}

int orc_linear_potential_omp(const double* zeta, const int shape[3], const double delta[3],
                             double depth, const double lbound[2], const double ubound[2],
                             const int npoints[2], double* out, int nthreads) {
	return potential(zeta, shape, delta, depth, lbound, ubound, npoints, out, NULL, 0, nthreads);
}

int orc_high_amplitude_potential(const double* zeta, const int shape[3], const double delta[3],
                                 double depth, const double lbound[2], const double ubound[2],
                                 const int npoints[2], double* out, double* wnmax_out) {
	return potential(zeta, shape, delta, depth, lbound, ubound, npoints, out, wnmax_out, 1, 1);
}

/* ------------------------------------------------------------------------ */
/* Non-linear inertialess transform of the realisation                      */
/* (NIT_transform::transform_realisation, src/nonlinear/nit_transform.cc:    */
/* 127-164, transform_data src/nonlinear/transforms.hh:58-75).               */
/* ------------------------------------------------------------------------ */

/* Positive nodes and weights of the 2m-point Gauss-Legendre rule on [-1, 1]
   (roots of the Legendre polynomial P_2m by Newton's method from the
   Chebyshev-like initial guesses; w = 2 / ((1 - x^2) P'(x)^2)). */
void orc_gauss_legendre_half(int m, double* x, double* w) {
	const int n = 2 * m;
	for (int i = 0; i < m; ++i) {
		/* the i-th root from the top, ordered ascending at the end */
		double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
		for (int it = 0; it < 100; ++it) {
			double p0 = 1.0, p1 = z;
			for (int k = 2; k <= n; ++k) {
				const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
				p0 = p1;
				p1 = p2;
			}
			dp = n * (z * p1 - p0) / (z * z - 1.0);
			const double dz = p1 / dp;
			z -= dz;
			if (fabs(dz) < 1e-15) break;  /* converged to within a few ulp */
		}
		x[m - 1 - i] = z;
		w[m - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
	}
}

/* Owen's T(h, a) by the symmetric 20-point Gauss-Legendre rule
   (src/apmath/owen_t.cc:38-55). */
double orc_owen_t(double h, double alpha) {
	static double x[10], w[10];
	static int have = 0;
	if (!have) {
		orc_gauss_legendre_half(10, x, w);
		have = 1;
	}
	const double a2 = alpha * alpha, h2 = -0.5 * h * h;
	double r = 0.0;
	for (int i = 0; i < 10; ++i) {
		const double term = 1.0 + a2 * x[i] * x[i];
		r += w[i] * exp(h2 * term) / term;
	}
	return r * alpha / (2.0 * M_PI);
}

/* Gram-Charlier CDF (src/stats/gram_charlier.hh:40-45), with the
   reference's 13-digit constants sqrt(2 pi) = 2.506628274631 and
   sqrt(2) = 1.4142135623731 (src/physical_constants.hh:21-24). */
double orc_gram_charlier_cdf(double x, double skewness, double kurtosis) {
	const double sqrt2pi = 2.506628274631, sqrt2 = 1.4142135623731;
	return exp(-0.5 * x * x) *
	           (kurtosis * (3.0 * x - x * x * x) + skewness * (4.0 - 4.0 * x * x) +
	            3.0 * x * x * x - 9.0 * x) /
	           (24.0 * sqrt2pi) +
	       0.5 * erf(x / sqrt2) + 0.5;
}

/* Normal CDF with mean mu and deviation sigma (gsl_cdf_gaussian_P). */
static double normal_cdf(double x, double mu, double sigma) {
	return 0.5 * erfc(-((x - mu) / sigma) / 1.4142135623730951);
}

/* Skew normal CDF (src/stats/skew_normal.hh:38-42). */
double orc_skew_normal_cdf(double x, double mean, double sigma, double alpha) {
	return normal_cdf(x, mean, sigma) - 2.0 * orc_owen_t((x - mean) / sigma, alpha);
}

static double target_cdf(int dist, const double p[5], double x) {
	return dist == 0 ? orc_gram_charlier_cdf(x, p[0], p[1])
	                 : orc_skew_normal_cdf(x, p[2], p[3], p[4]);
}

/* Every point x -> the root in [a, b] of F(c) = Phi(x / stdev) by the
   reference's bisection (src/linalg.hh:199-213): at most max_iter steps,
   stopping once b - a <= eps or |F(c) - Phi| <= eps; the midpoint of the
   last step is the result.  dist 0: Gram-Charlier (p[0] skewness, p[1]
   kurtosis); 1: skew normal (p[2] mean, p[3] stdev, p[4] alpha). */
void orc_nit_realisation(double* data, size_t n, double stdev, int dist, const double p[5],
                         double a0, double b0, double eps, int max_iter) {
	for (size_t q = 0; q < n; ++q) {
		const double y = normal_cdf(data[q], 0.0, stdev);
		double a = a0, b = b0, c, fc;
		int i = 0;
		do {
			c = 0.5 * (a + b);
			fc = target_cdf(dist, p, c) - y;
			if ((target_cdf(dist, p, a) - y) * fc < 0.0) b = c;
			if ((target_cdf(dist, p, b) - y) * fc < 0.0) a = c;
			++i;
		} while (i < max_iter && (b - a) > eps && fabs(fc) > eps);
		data[q] = c;
	}
}

/* ------------------------------------------------------------------------ */
/* ACF_generator (src/generator/acf_generator.cc:130-215): the ACF as the    */
/* auto-covariance of a synthetic plain wave with exponential decay.         */
/* ------------------------------------------------------------------------ */

/* Plain wave profiles (src/generator/plain_wave.hh): 0 sin, 1 cos, 2 stokes,
   3 standing_wave; 2 pi written as the reference's T(2) * pi. */
static double plain_wave(int func, double amp, double kx, double ky, double vel, double phase,
                         double x, double y, double t) {
	const double two_pi = 2.0 * M_PI;
	switch (func) {
	case 0: return amp * sin(two_pi * (kx * x + ky * y - vel * t) + phase);
	case 1: return amp * sin(two_pi * (kx * x + ky * y - vel * t) + (phase + 0.5 * M_PI));
	case 2: {
		const double theta = two_pi * (kx * x + ky * y - vel * t) + phase;
		const double st = two_pi * amp * sqrt(kx * kx + ky * ky);
		return amp * (cos(theta) + 0.5 * st * cos(2.0 * theta) +
		              (3.0 / 8.0) * st * st * cos(3.0 * theta));
	}
	default: return amp * cos(two_pi * (kx * x + ky * y) + phase) * cos(two_pi * (vel * t));
	}
}

/* The wave on Domain(-r, r, ws + 1) (src/generator/acf_generator.hh:156-167),
   t = -r0 + i (2 r0 / ws0) and so on. */
static void wave_on_grid(int func, double amp, double vel, const double k[2], const double r[3],
                         const int ws[3], double* out) {
	const int n0 = ws[0] + 1, n1 = ws[1] + 1, n2 = ws[2] + 1;
	double d[3];
	for (int q = 0; q < 3; ++q) d[q] = ws[q] == 0 ? 0.0 : (r[q] - (-r[q])) / ws[q];
	for (int i = 0; i < n0; ++i)
		for (int j = 0; j < n1; ++j)
			for (int l = 0; l < n2; ++l)
				out[((size_t)i * n1 + j) * n2 + l] =
					plain_wave(func, amp, k[0], k[1], vel, 0.0, -r[1] + d[1] * j, -r[2] + d[2] * l,
					           -r[0] + d[0] * i);
}

/* auto_covariance (src/arma.cc:81-100): real(IFFT(|FFT(w)|^2)) / n / n. */
static void auto_covariance(const double* w, const int s[3], double* out) {
	const size_t n = (size_t)s[0] * s[1] * s[2];
	double* re = (double*)malloc(sizeof(double) * n);
	double* im = (double*)calloc(n, sizeof(double));
	memcpy(re, w, sizeof(double) * n);
	for (int sign = -1; sign <= 1; sign += 2) {
		for (int j = 0; j < s[1]; ++j)
			for (int l = 0; l < s[2]; ++l)
				orc_fft_1d(re + (size_t)j * s[2] + l, im + (size_t)j * s[2] + l, s[0],
				           s[1] * s[2], sign);
		for (int i = 0; i < s[0]; ++i)
			for (int l = 0; l < s[2]; ++l)
				orc_fft_1d(re + (size_t)i * s[1] * s[2] + l, im + (size_t)i * s[1] * s[2] + l,
				           s[1], s[2], sign);
		for (int i = 0; i < s[0]; ++i)
			for (int j = 0; j < s[1]; ++j)
				orc_fft_1d(re + ((size_t)i * s[1] + j) * s[2], im + ((size_t)i * s[1] + j) * s[2],
				           s[2], 1, sign);
		if (sign < 0)
			for (size_t q = 0; q < n; ++q) {
				const double a = sqrt(re[q] * re[q] + im[q] * im[q]);
				re[q] = a * a;
				im[q] = 0.0;
			}
	}
	for (size_t q = 0; q < n; ++q) out[q] = re[q] / (double)n / (double)n;
	free(re);
	free(im);
}

/* ACF_generator::generate().  func: 0 sin, 1 cos, 2 stokes, 3 standing_wave.
   shape all > 0: the wave has 2 shape + 1 points per dimension; otherwise the
   wave shape doubles from (2,2,2) while all extents are below 64 until the
   variance changes by less than 1e-3.  out receives the first
   (wave points / 2) per dimension of the auto-covariance (capacity cap);
   shape_out that shape.  Returns 0, or -1 when cap is too small. */
int orc_acf_generate(int func, double amplitude, double velocity, const double wavenum[2],
                     const double alpha[3], const double nwaves[3], const int shape[3],
                     double* out, size_t cap, int shape_out[3]) {
	int ws[3] = {2, 2, 2};
	double* wave = NULL;
	size_t n = 0;
	int s[3];
	if (shape[0] > 0 && shape[1] > 0 && shape[2] > 0) {
		for (int q = 0; q < 3; ++q) ws[q] = 2 * shape[q];
		for (int q = 0; q < 3; ++q) s[q] = ws[q] + 1;
		n = (size_t)s[0] * s[1] * s[2];
		wave = (double*)malloc(sizeof(double) * n);
		wave_on_grid(func, amplitude, velocity, wavenum, nwaves, ws, wave);
	} else {
		double var0 = -1.0, var = -1.0;
		while (ws[0] < 64 && ws[1] < 64 && ws[2] < 64) {
			for (int q = 0; q < 3; ++q) s[q] = ws[q] + 1;
			n = (size_t)s[0] * s[1] * s[2];
			free(wave);
			wave = (double*)malloc(sizeof(double) * n);
			wave_on_grid(func, amplitude, velocity, wavenum, nwaves, ws, wave);
			var0 = var;
			var = orc_gsl_variance_d(wave, n);
			if (!(var0 < 0.0) && fabs(var - var0) < 1e-3) break;
			for (int q = 0; q < 3; ++q) ws[q] *= 2;
		}
		for (int q = 0; q < 3; ++q) ws[q] = s[q] - 1;
	}
	/* add_exponential_decay (src/generator/acf_generator.cc:196-215) */
	double d[3];
	for (int q = 0; q < 3; ++q) d[q] = ws[q] == 0 ? 0.0 : (nwaves[q] - (-nwaves[q])) / ws[q];
	double* we = (double*)malloc(sizeof(double) * n);
	for (int i = 0; i < s[0]; ++i)
		for (int j = 0; j < s[1]; ++j)
			for (int l = 0; l < s[2]; ++l) {
				const double t = -nwaves[0] + d[0] * i, x = -nwaves[1] + d[1] * j,
				             y = -nwaves[2] + d[2] * l;
				const size_t q = ((size_t)i * s[1] + j) * s[2] + l;
				we[q] = wave[q] * exp(-(fabs(t * alpha[0]) + fabs(x * alpha[1]) +
				                        fabs(y * alpha[2])));
			}
	const double f = sqrt(orc_gsl_variance_d(wave, n) / orc_gsl_variance_d(we, n));
	for (size_t q = 0; q < n; ++q) we[q] *= f;
	double* acf = (double*)malloc(sizeof(double) * n);
	auto_covariance(we, s, acf);
	int h[3];
	for (int q = 0; q < 3; ++q) h[q] = s[q] / 2;
	for (int q = 0; q < 3; ++q) shape_out[q] = h[q];
	int rc = 0;
	if ((size_t)h[0] * h[1] * h[2] > cap) {
		rc = -1;
	} else {
		for (int i = 0; i < h[0]; ++i)
			for (int j = 0; j < h[1]; ++j)
				for (int l = 0; l < h[2]; ++l)
					out[((size_t)i * h[1] + j) * h[2] + l] = acf[((size_t)i * s[1] + j) * s[2] + l];
	}
	free(wave);
	free(we);
	free(acf);
	return rc;
}
