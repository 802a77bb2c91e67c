// Post-rescale of the generated surface
// (src/generator/basic_arma_model.cc:102-106):
//     zeta *= sqrt(gamma000 / variance(zeta(RectDomain(shape/2, shape-1))))
// where stats::variance (src/stats/statistics.hh:35-44) hands the view's
// data pointer and element count to gsl_stats_variance, i.e. it reads the
// CONTIGUOUS run of ceil(nt/2)ceil(nx/2)ceil(ny/2) elements starting at
// zeta(nt/2, nx/2, ny/2).  GSL computes a long-double running mean and a
// long-double running mean of squared deviations times n/(n-1); here both
// passes are compensated (TwoSum) FP64 sums per workgroup, combined on the
// host in long double, which agrees with GSL to the last bits of a double.
#include "common.hpp"

#include <cmath>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
	s = a + b;
	const double bb = s - a;
	e = (a - (s - bb)) + (b - bb);
}

// PASS 0: sum of x; PASS 1: sum of (x - mean)^2 with the difference rounded
// to double before squaring, as GSL does.
template <class T, int PASS>
__global__ __launch_bounds__(kRedThreads) void moment_kernel(const T* __restrict__ x, size_t n,
                                                             double mean,
                                                             double* __restrict__ partials) {
	double s = 0.0, c = 0.0;
	for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
	     i += size_t(gridDim.x) * blockDim.x) {
		double v = double(x[i]);
		if (PASS == 1) {
			const double dlt = v - mean;
			v = dlt * dlt;
		}
		double e;
		two_sum(s, v, s, e);
		c += e;
	}
	__shared__ double ss[kRedThreads], cc[kRedThreads];
	ss[threadIdx.x] = s;
	cc[threadIdx.x] = c;
	__syncthreads();
	for (int w = kRedThreads / 2; w > 0; w >>= 1) {
		if (threadIdx.x < w) {
			double e;
			double a = ss[threadIdx.x], b = ss[threadIdx.x + w];
			double r;
			two_sum(a, b, r, e);
			ss[threadIdx.x] = r;
			cc[threadIdx.x] += cc[threadIdx.x + w] + e;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		partials[2 * blockIdx.x] = ss[0];
		partials[2 * blockIdx.x + 1] = cc[0];
	}
}

template <class T>
__global__ void scale_kernel(T* __restrict__ x, size_t n, T f) {
	for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
	     i += size_t(gridDim.x) * blockDim.x)
		x[i] = x[i] * f;
}

template <class T>
long double device_sum(arma_hip_ctx* ctx, const T* x, size_t n, double mean, int pass) {
	const int blocks = int(std::min<size_t>(kRedBlocks, (n + kRedThreads - 1) / kRedThreads));
	KernelTimer timer(ctx, "rescale_moment");
	double* part = (double*)ctx->partials.get(sizeof(double) * 2 * kRedBlocks);
	if (pass == 0)
		moment_kernel<T, 0><<<blocks, kRedThreads, 0, ctx->stream>>>(x, n, mean, part);
	else
		moment_kernel<T, 1><<<blocks, kRedThreads, 0, ctx->stream>>>(x, n, mean, part);
	ARMA_HIP_CHECK(hipGetLastError());
	std::vector<double> h(2 * blocks);
	ARMA_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * blocks,
	                              hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	long double s = 0.0L;
	for (int b = 0; b < blocks; ++b) s += (long double)h[2 * b] + (long double)h[2 * b + 1];
	return s;
}

}  // namespace

double launch_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int precision,
                      double gamma0) {
	const size_t n = numel3(shape);
	const size_t start = (size_t(shape[0] / 2) * shape[1] + shape[1] / 2) * size_t(shape[2]) +
	                     shape[2] / 2;
	const size_t cnt = size_t(shape[0] - shape[0] / 2) * size_t(shape[1] - shape[1] / 2) *
	                   size_t(shape[2] - shape[2] / 2);
	ARMA_HIP_REQUIRE(cnt >= 2, ARMA_HIP_EINVAL, "surface too small for the variance");
	double var;
	if (precision == ARMA_FP32) {
		const float* x = (const float*)zeta + start;
		const double mean = double(device_sum(ctx, x, cnt, 0.0, 0) / (long double)cnt);
		const long double ss = device_sum(ctx, x, cnt, mean, 1);
		var = double(ss / (long double)cnt) * (double(cnt) / double(cnt - 1));
	} else {
		const double* x = (const double*)zeta + start;
		const double mean = double(device_sum(ctx, x, cnt, 0.0, 0) / (long double)cnt);
		const long double ss = device_sum(ctx, x, cnt, mean, 1);
		var = double(ss / (long double)cnt) * (double(cnt) / double(cnt - 1));
	}
	const double factor = std::sqrt(gamma0 / var);
	const int blocks = int(std::min<size_t>(4096, (n + 255) / 256));
	KernelTimer timer(ctx, "rescale_scale");
	if (precision == ARMA_FP32)
		scale_kernel<float><<<blocks, 256, 0, ctx->stream>>>((float*)zeta, n, float(factor));
	else
		scale_kernel<double><<<blocks, 256, 0, ctx->stream>>>((double*)zeta, n, factor);
	ARMA_HIP_CHECK(hipGetLastError());
	return factor;
}

}  // namespace arma_hip
