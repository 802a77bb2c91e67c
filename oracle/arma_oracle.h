/*
 * arma_oracle.h -- CPU FP64 restatement of the ARMA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product (arma_amd/)
 * never links, imports or calls it.
 *
 * Every function cites the reference file:line it restates.  Parity status
 * of each function (how it is pinned) is listed in the header of
 * arma_oracle.c.
 */
#ifndef ARMA_ORACLE_H
#define ARMA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dynamic-creator Mersenne Twister (libdcmt genrand_mt/sgenrand_mt) ---- */

typedef struct orc_mt_params {
	uint32_t aaa;
	int32_t mm, nn, rr, ww;
	uint32_t wmask, umask, lmask;
	int32_t shift0, shift1, shiftB, shiftC;
	uint32_t maskB, maskC;
} orc_mt_params;

#define ORC_MT_MAXN 1392

typedef struct orc_mt {
	orc_mt_params p;
	int32_t i;
	uint32_t state[ORC_MT_MAXN];
} orc_mt;

int orc_mt_seed(orc_mt* g, const orc_mt_params* p, uint32_t seed);
uint32_t orc_mt_next(orc_mt* g);
int orc_mt_fill(const orc_mt_params* p, uint32_t seed, uint32_t* out, size_t n);

/* ---- libstdc++ generate_canonical + Marsaglia polar normal ---- */

typedef struct orc_normal_d { int saved_available; double saved; } orc_normal_d;
typedef struct orc_normal_f { int saved_available; float saved; } orc_normal_f;

double orc_canonical_d(orc_mt* g);
float orc_canonical_f(orc_mt* g);
double orc_normal_next_d(orc_mt* g, orc_normal_d* st, double mean, double stddev);
float orc_normal_next_f(orc_mt* g, orc_normal_f* st, float mean, float stddev);

/* Whole-grid noise from stream 0 (framework=none). */
int orc_noise_sequential_d(const orc_mt_params* p, uint32_t seed, double stddev,
                           double* out, size_t n);
int orc_noise_sequential_f(const orc_mt_params* p, uint32_t seed, float stddev,
                           float* out, size_t n);
/* One stream per partition, partitions in row-major order, noise row-major
 * inside each partition.  Returns -1 when n_mts < number of partitions. */
int orc_noise_partitioned_d(const orc_mt_params* p, const uint32_t* seeds, int n_mts,
                            const int shape[3], const int part[3], double stddev,
                            double* out);
/* The same with the partitions spread over nthreads OpenMP threads (the
   all-core CPU baseline; the result does not depend on nthreads). */
int orc_noise_partitioned_d_omp(const orc_mt_params* p, const uint32_t* seeds, int n_mts,
                                const int shape[3], const int part[3], double stddev,
                                double* out, int nthreads);

/* ---- ACF, Yule-Walker, AR ---- */

enum { ORC_ACF_STANDING = 0, ORC_ACF_PROPAGATING = 1 };
void orc_acf(int func, const double delta[3], const int shape[3], double amplitude,
             double velocity, const double alpha[3], const double beta[2], double* out);

/* Choi order-recursive Yule-Walker.  acf is an n*n*n cube (gamma, not rho).
 * phi_out must hold max_order^3 doubles; order_out receives phi's shape. */
int orc_yw_choi(const double* acf, int n, int max_order, double var_eps,
                int determine_order, int chop, double chop_eps,
                double* phi_out, int order_out[3], double* varwn_solver);
int orc_yw_gauss(const double* acf, const int acf_shape[3], const int order[3],
                 double* phi_out);
void orc_chop_right(const double* phi, const int shape[3], double eps, int out_shape[3]);
double orc_white_noise_variance(const double* acf, const int acf_shape[3],
                                const double* phi, const int phi_shape[3]);
void orc_partition_shape(const int partition[3], const int grid[3], const int order[3],
                         int parallelism, int out[3]);

void orc_ar_generate(double* zeta, const int shape[3], const double* phi,
                     const int pshape[3], const int lo[3], const int hi[3]);
void orc_ar_generate_f(float* zeta, const int shape[3], const float* phi,
                       const int pshape[3], const int lo[3], const int hi[3]);
void orc_ar_generate_dag(double* zeta, const int shape[3], const double* phi,
                         const int pshape[3], const int part[3]);
void orc_ar_generate_dag_omp(double* zeta, const int shape[3], const double* phi,
                             const int pshape[3], const int part[3], int nthreads);

double orc_gsl_variance_d(const double* data, size_t n);
float orc_gsl_variance_f(const float* data, size_t n);
double orc_rescale_factor(const double* zeta, const int shape[3], double gamma0);
double orc_rescale_factor_f(const float* zeta, const int shape[3], double gamma0);

/* ---- velocity potential (linear solver) ---- */

void orc_derivative_t(const double* zeta, const int shape[3], double dt, int idx_t,
                      double* out);
void orc_factor_waves(const double* slice, int nx, int ny, int t, double dt,
                      double out[4]);
void orc_window(int nx, int ny, double wnmax_x, double wnmax_y, double z, double depth,
                double* out);
void orc_fft_1d(double* re, double* im, int n, int stride, int sign);
void orc_fft2(double* re, double* im, int nx, int ny, int sign);
void orc_interp_edges(double* res, int nx, int ny);
int orc_linear_potential(const double* zeta, const int shape[3], const double delta[3],
                         double depth, const double lbound[2], const double ubound[2],
                         const int npoints[2], double* out, double* wnmax_out);
int orc_linear_potential_omp(const double* zeta, const int shape[3], const double delta[3],
                             double depth, const double lbound[2], const double ubound[2],
                             const int npoints[2], double* out, int nthreads);

/* ---- velocity potential (high-amplitude solver) ---- */

void orc_derivative_xy(const double* zeta, const int shape[3], double d, int dim, int idx_t,
                       double* out);
int orc_high_amplitude_potential(const double* zeta, const int shape[3], const double delta[3],
                                 double depth, const double lbound[2], const double ubound[2],
                                 const int npoints[2], double* out, double* wnmax_out);

/* ---- non-linear inertialess transform of the realisation ---- */

void orc_gauss_legendre_half(int m, double* x, double* w);
double orc_owen_t(double h, double alpha);
double orc_gram_charlier_cdf(double x, double skewness, double kurtosis);
double orc_skew_normal_cdf(double x, double mean, double sigma, double alpha);
void orc_nit_realisation(double* data, size_t n, double stdev, int dist, const double p[5],
                         double a0, double b0, double eps, int max_iter);

/* ---- ACF_generator ---- */

int orc_acf_generate(int func, double amplitude, double velocity, const double wavenum[2],
                     const double alpha[3], const double nwaves[3], const int shape[3],
                     double* out, size_t cap, int shape_out[3]);

#ifdef __cplusplus
}
#endif

#endif
