// Post-rescale of the generated surface
// (src/generator/basic_arma_model.cc:102-106):
//     zeta *= sqrt(gamma000 / variance(zeta(RectDomain(shape/2, shape-1))))
// where stats::variance (src/stats/statistics.hh:35-44) hands the view's
// data pointer and element count to gsl_stats_variance, i.e. it reads the
// CONTIGUOUS run of ceil(nt/2)ceil(nx/2)ceil(ny/2) elements starting at
// zeta(nt/2, nx/2, ny/2).  GSL computes a long-double running mean and a
// long-double running mean of squared deviations times n/(n-1); here both
// passes are compensated (TwoSum) FP64 sums per workgroup, combined on the
// device as a double-double total per surface (on the device, so a batch of
// surfaces rescales without a host round trip), which agrees with GSL's
// long-double running mean to the last bits of a double.
#include "common.hpp"

#include <cmath>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;  // most workgroups per surface and pass

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
	s = a + b;
	const double bb = s - a;
	e = (a - (s - bb)) + (b - bb);
}

// The contiguous run one surface's variance reads, and its surface.
struct RescaleRec {
	long long start;  // arena element where the run starts
	long long cnt;    // run length
	long long zoff;   // first arena element of the surface
	long long n;      // surface length
	double gamma0;    // target variance; <= 0: no rescale
	int parts;        // workgroups summing the run (a function of cnt only, so a
	                  // surface's sums do not depend on the rest of a batch)
	int pad;
};

// Workgroup-wide compensated sum of (s, c) pairs; the result is on thread 0.
__device__ void block_two_sum(double& s, double& c) {
	__shared__ double ss[kRedThreads], cc[kRedThreads];
	ss[threadIdx.x] = s;
	cc[threadIdx.x] = c;
	__syncthreads();
	for (int w = kRedThreads / 2; w > 0; w >>= 1) {
		if (threadIdx.x < w) {
			double e, r;
			two_sum(ss[threadIdx.x], ss[threadIdx.x + w], r, e);
			ss[threadIdx.x] = r;
			cc[threadIdx.x] += cc[threadIdx.x + w] + e;
		}
		__syncthreads();
	}
	s = ss[0];
	c = cc[0];
}

// PASS 0: sum of x; PASS 1: sum of (x - mean)^2 with the difference rounded
// to double before squaring, as GSL does.  Workgroup (bx, s), bx < parts,
// sums a strided share of surface s's run into partials[s][bx].
template <class T, int PASS>
__global__ __launch_bounds__(kRedThreads) void moment_kernel(const T* __restrict__ arena,
                                                             const RescaleRec* __restrict__ recs,
                                                             const double* __restrict__ stats,
                                                             double* __restrict__ partials) {
	const RescaleRec r = recs[blockIdx.y];
	const double mean = PASS == 1 ? stats[2 * blockIdx.y] : 0.0;
	const T* x = arena + r.start;
	if ((int)blockIdx.x >= r.parts) return;
	double s = 0.0, c = 0.0;
	if (r.gamma0 > 0.0) {
		for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < r.cnt;
		     i += (long long)r.parts * blockDim.x) {
			double v = double(x[i]);
			if (PASS == 1) {
				const double dlt = v - mean;
				v = dlt * dlt;
			}
			double e;
			two_sum(s, v, s, e);
			c += e;
		}
	}
	block_two_sum(s, c);
	if (threadIdx.x == 0) {
		double* p = partials + 2 * (size_t(blockIdx.y) * kRedBlocks + blockIdx.x);
		p[0] = s;
		p[1] = c;
	}
}

// One workgroup per surface folds the partials into a double-double total
// and divides by the run length: PASS 0 stores the mean, PASS 1 the
// variance (times n/(n-1)) and the factor sqrt(gamma0 / var).
template <int PASS>
__global__ __launch_bounds__(kRedThreads) void finalize_kernel(const RescaleRec* __restrict__ recs,
                                                               const double* __restrict__ partials,
                                                               double* __restrict__ stats) {
	const RescaleRec r = recs[blockIdx.x];
	const int nparts = r.parts;
	const double* p = partials + 2 * size_t(blockIdx.x) * kRedBlocks;
	double s = 0.0, c = 0.0;
	for (int i = threadIdx.x; i < nparts; i += kRedThreads) {
		double e;
		two_sum(s, p[2 * i], s, e);
		c += e + p[2 * i + 1];
	}
	block_two_sum(s, c);
	if (threadIdx.x != 0) return;
	// (s + c) / cnt, correctly rounded up to the last bit of the quotient
	double hi, lo;
	two_sum(s, c, hi, lo);
	const double n = double(r.cnt);
	double q = hi / n;
	q += (fma(-q, n, hi) + lo) / n;
	if (PASS == 0) {
		stats[2 * blockIdx.x] = q;
	} else {
		const double var = q * (n / (n - 1.0));
		stats[2 * blockIdx.x + 1] = r.gamma0 > 0.0 ? sqrt(r.gamma0 / var) : 0.0;
	}
}

template <class T>
__global__ void scale_kernel(T* __restrict__ arena, const RescaleRec* __restrict__ recs,
                             const double* __restrict__ stats) {
	const RescaleRec r = recs[blockIdx.y];
	if (!(r.gamma0 > 0.0)) return;
	const T f = T(stats[2 * blockIdx.y + 1]);
	T* x = arena + r.zoff;
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < r.n;
	     i += (long long)gridDim.x * blockDim.x)
		x[i] = x[i] * f;
}

RescaleRec make_rec(size_t zoff, const int32_t shape[3], double gamma0) {
	RescaleRec r;
	r.zoff = (long long)zoff;
	r.n = (long long)numel3(shape);
	r.start = r.zoff + (long long)((size_t(shape[0] / 2) * shape[1] + shape[1] / 2) *
	                                   size_t(shape[2]) +
	                               shape[2] / 2);
	r.cnt = (long long)(size_t(shape[0] - shape[0] / 2) * size_t(shape[1] - shape[1] / 2) *
	                    size_t(shape[2] - shape[2] / 2));
	r.gamma0 = gamma0;
	r.parts = int(std::max<long long>(
		1, std::min<long long>(kRedBlocks, (r.cnt + kRedThreads - 1) / kRedThreads)));
	r.pad = 0;
	return r;
}

// moments, mean, second moments, factor, scale: five launches, no host wait
void rescale_core(arma_hip_ctx* ctx, int precision, void* arena, const RescaleRec* drecs,
                  int nsurf, long long max_cnt, long long max_n, double* partials,
                  double* stats) {
	const int bx = int(std::min<long long>(kRedBlocks, (max_cnt + kRedThreads - 1) / kRedThreads));
	const dim3 mg(std::max(bx, 1), nsurf);
	const int sx = int(std::min<long long>(4096, (max_n + 255) / 256));
	const dim3 sg(std::max(sx, 1), nsurf);
	{
		KernelTimer timer(ctx, "rescale_moment");
		if (precision == ARMA_FP32) {
			moment_kernel<float, 0><<<mg, kRedThreads, 0, ctx->stream>>>((const float*)arena,
			                                                              drecs, stats, partials);
			finalize_kernel<0><<<nsurf, kRedThreads, 0, ctx->stream>>>(drecs, partials, stats);
			moment_kernel<float, 1><<<mg, kRedThreads, 0, ctx->stream>>>((const float*)arena,
			                                                              drecs, stats, partials);
		} else {
			moment_kernel<double, 0><<<mg, kRedThreads, 0, ctx->stream>>>(
				(const double*)arena, drecs, stats, partials);
			finalize_kernel<0><<<nsurf, kRedThreads, 0, ctx->stream>>>(drecs, partials, stats);
			moment_kernel<double, 1><<<mg, kRedThreads, 0, ctx->stream>>>(
				(const double*)arena, drecs, stats, partials);
		}
		finalize_kernel<1><<<nsurf, kRedThreads, 0, ctx->stream>>>(drecs, partials, stats);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	KernelTimer timer(ctx, "rescale_scale");
	if (precision == ARMA_FP32)
		scale_kernel<float><<<sg, 256, 0, ctx->stream>>>((float*)arena, drecs, stats);
	else
		scale_kernel<double><<<sg, 256, 0, ctx->stream>>>((double*)arena, drecs, stats);
	ARMA_HIP_CHECK(hipGetLastError());
}

}  // namespace

double launch_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int precision,
                      double gamma0) {
	const RescaleRec r = make_rec(0, shape, gamma0);
	ARMA_HIP_REQUIRE(r.cnt >= 2, ARMA_HIP_EINVAL, "surface too small for the variance");
	ARMA_HIP_REQUIRE(gamma0 > 0.0, ARMA_HIP_EINVAL, "rescale target must be positive");
	// record, then partials, then (mean, factor)
	const size_t pbytes = sizeof(double) * 2 * kRedBlocks;
	unsigned char* buf = (unsigned char*)ctx->partials.get(256 + pbytes + 256);
	auto* drec = (RescaleRec*)buf;
	auto* partials = (double*)(buf + 256);
	auto* stats = (double*)(buf + 256 + pbytes);
	upload(ctx, drec, &r, sizeof(r));
	rescale_core(ctx, precision, zeta, drec, 1, r.cnt, r.n, partials, stats);
	double st[2];
	ARMA_HIP_CHECK(hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	return st[1];
}

struct RescaleBatch {
	int precision = 0;
	int n = 0;
	long long max_cnt = 0, max_n = 0;
	bool any = false;
	Scratch recs, partials, stats;
};

RescaleBatch* rescale_batch_create(arma_hip_ctx* ctx, int precision,
                                   const std::vector<BatchMember>& m) {
	auto* b = new RescaleBatch;
	try {
		rescale_batch_build(ctx, b, precision, m);
	} catch (...) {
		delete b;
		throw;
	}
	return b;
}

RescaleBatch* rescale_batch_new() { return new RescaleBatch; }

void rescale_batch_build(arma_hip_ctx* ctx, RescaleBatch* b, int precision,
                         const std::vector<BatchMember>& m) {
	{
		b->precision = precision;
		b->n = int(m.size());
		b->max_cnt = 0;
		b->max_n = 0;
		b->any = false;
		std::vector<RescaleRec> h;
		for (const BatchMember& x : m) {
			const RescaleRec r = make_rec(x.zoff, x.d->shape, x.gamma0);
			if (x.gamma0 > 0.0) {
				ARMA_HIP_REQUIRE(r.cnt >= 2, ARMA_HIP_EINVAL,
				                 "surface too small for the variance");
				b->any = true;
				b->max_cnt = std::max(b->max_cnt, r.cnt);
				b->max_n = std::max(b->max_n, r.n);
			}
			h.push_back(r);
		}
		b->recs.get(sizeof(RescaleRec) * h.size());
		b->partials.get(sizeof(double) * 2 * size_t(kRedBlocks) * h.size());
		b->stats.get(sizeof(double) * 2 * h.size());
		upload(ctx, b->recs.ptr, h.data(), sizeof(RescaleRec) * h.size());
		// factors stay 0 until the first launch
		std::vector<double> z(2 * h.size(), 0.0);
		upload(ctx, b->stats.ptr, z.data(), sizeof(double) * z.size());
	}
}

void rescale_batch_launch(arma_hip_ctx* ctx, RescaleBatch* b, void* arena) {
	if (!b->any) return;
	rescale_core(ctx, b->precision, arena, (const RescaleRec*)b->recs.ptr, b->n, b->max_cnt,
	             b->max_n, (double*)b->partials.ptr, (double*)b->stats.ptr);
}

void rescale_batch_factors(arma_hip_ctx* ctx, RescaleBatch* b, double* out) {
	std::vector<double> st(2 * size_t(b->n));
	ARMA_HIP_CHECK(hipMemcpyAsync(st.data(), b->stats.ptr, sizeof(double) * st.size(),
	                              hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	for (int i = 0; i < b->n; ++i) out[i] = st[2 * size_t(i) + 1];
}

void rescale_batch_destroy(RescaleBatch* b) { delete b; }

}  // namespace arma_hip
