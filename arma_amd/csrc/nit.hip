// Non-linear inertialess transform (NIT) of the realisation on the GPU.
//
// Reference: NIT_transform<T>::transform_realisation
// (src/nonlinear/nit_transform.cc:127-164) -> transform_data
// (src/nonlinear/transforms.hh:58-75): every point x of the surface is
// replaced by the root c of F(c) = Phi(x / stdev), stdev = sqrt(acf(0,0,0)),
// where F is the target CDF, Gram-Charlier (src/stats/gram_charlier.hh:40-45)
// or skew normal (src/stats/skew_normal.hh:38-42, Owen's T by the symmetric
// 20-point Gauss-Legendre rule of src/apmath/owen_t.cc:38-55), and the root
// comes from the reference's bisection (src/linalg.hh:199-213): at most
// max_iter halvings of [a, b], stopping once b - a <= eps or |F(c) - y| <= eps,
// the last midpoint being the result.
//
// One thread per point; the bisection runs in FP64 for both precisions.  The
// reference evaluates F at a and at b again in every step; F is a pure
// function, so the kernel keeps F(a) and F(b) from the step that set them and
// evaluates F once per step, with identical results.  The Gauss-Legendre
// nodes are computed on the host (Newton's method on P_20) and passed in.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"

namespace arma_hip {
namespace {

struct NitParams {
	double stdev;  // deviation of the data's normal law
	double a, b, eps;
	int max_iter, dist;
	double skew, kurt;          // Gram-Charlier
	double mean, sigma, alpha;  // skew normal
	double gx[10], gw[10];      // positive nodes and weights of the 20-point rule
};

// The reference's 13-digit constants (src/physical_constants.hh:21-24).
constexpr double kSqrt2Pi = 2.506628274631;
constexpr double kSqrt2 = 1.4142135623731;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// gsl_cdf_gaussian_P(x - mu, sigma)
__device__ __forceinline__ double normal_cdf(double x, double mu, double sigma) {
	return 0.5 * erfc(-((x - mu) / sigma) / 1.4142135623730951);
}

__device__ __forceinline__ double gram_charlier_cdf(double x, double s, double k) {
	return exp(-0.5 * x * x) * (k * (3.0 * x - x * x * x) + s * (4.0 - 4.0 * x * x) +
	                            3.0 * x * x * x - 9.0 * x) /
	           (24.0 * kSqrt2Pi) +
	       0.5 * erf(x / kSqrt2) + 0.5;
}

__device__ __forceinline__ double owen_t(const NitParams& p, double h, double alpha) {
	const double a2 = alpha * alpha, h2 = -0.5 * h * h;
	double r = 0.0;
#pragma unroll
	for (int i = 0; i < 10; ++i) {
		const double term = 1.0 + a2 * p.gx[i] * p.gx[i];
		r += p.gw[i] * exp(h2 * term) / term;
	}
	return r * alpha / kTwoPi;
}

__device__ __forceinline__ double target_cdf(const NitParams& p, double x) {
	if (p.dist == ARMA_NIT_GRAM_CHARLIER) return gram_charlier_cdf(x, p.skew, p.kurt);
	return normal_cdf(x, p.mean, p.sigma) - 2.0 * owen_t(p, (x - p.mean) / p.sigma, p.alpha);
}

template <class T>
__global__ __launch_bounds__(256) void nit_kernel(T* __restrict__ data, size_t n, NitParams p) {
	const size_t stride = size_t(gridDim.x) * blockDim.x;
	for (size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += stride) {
		const double y = normal_cdf(double(data[q]), 0.0, p.stdev);
		double a = p.a, b = p.b, c, fc;
		double fa = target_cdf(p, a) - y, fb = target_cdf(p, b) - y;
		int i = 0;
		do {
			c = 0.5 * (a + b);
			fc = target_cdf(p, c) - y;
			if (fa * fc < 0.0) {
				b = c;
				fb = fc;
			}
			if (fb * fc < 0.0) {
				a = c;
				fa = fc;
			}
			++i;
		} while (i < p.max_iter && (b - a) > p.eps && fabs(fc) > p.eps);
		data[q] = T(c);
	}
}

// Positive nodes and weights of the 2m-point Gauss-Legendre rule (Newton's
// method on P_2m; w = 2 / ((1 - x^2) P'(x)^2)), ascending.
void gauss_legendre_half(int m, double* x, double* w) {
	const int n = 2 * m;
	for (int i = 0; i < m; ++i) {
		double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
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
			if (std::fabs(dz) < 1e-15) break;  // converged to within a few ulp
		}
		x[m - 1 - i] = z;
		w[m - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
	}
}

}  // namespace

void launch_nit_realisation(arma_hip_ctx* ctx, void* zeta, size_t n, int precision,
                            double gamma0, const arma_nit_desc* d) {
	NitParams p;
	p.stdev = std::sqrt(gamma0);
	p.a = d->interval[0];
	p.b = d->interval[1];
	p.eps = d->absolute_error;
	p.max_iter = d->max_iterations;
	p.dist = d->distribution;
	p.skew = d->skewness;
	p.kurt = d->kurtosis;
	p.mean = d->mean;
	p.sigma = d->stdev;
	p.alpha = d->alpha;
	gauss_legendre_half(10, p.gx, p.gw);
	if (n == 0) return;
	const int grid = int(std::min<size_t>((n + 255) / 256, size_t(1) << 20));
	KernelTimer t(ctx, "nit_realisation");
	if (precision == ARMA_FP32)
		nit_kernel<float><<<grid, 256, 0, ctx->stream>>>(static_cast<float*>(zeta), n, p);
	else
		nit_kernel<double><<<grid, 256, 0, ctx->stream>>>(static_cast<double*>(zeta), n, p);
	ARMA_HIP_CHECK(hipGetLastError());
}

}  // namespace arma_hip
