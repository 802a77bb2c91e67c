// ACF_generator's auto-covariance on the GPU.
//
// The reference computes real(IFFT(|FFT(w)|^2)) / n^2 of the decayed wave w
// (src/arma.cc:81-100, called from src/generator/acf_generator.cc:216-233)
// and keeps the first half of every dimension.  With the unnormalised
// inverse that is the circular auto-covariance
//     acf(d) = (1/n) sum_x w(x) w((x - d) mod s),
// which this kernel sums directly in FP64 for the lags the crop keeps.  The
// extents are odd (2 shape + 1, at most 129), which the power-of-two
// transforms of the potential do not cover, and the direct sum has no
// transform rounding.
//
// Decomposition: a workgroup owns a 16 x 16 block of lags (d1, d2) at one d0
// and walks the planes x0 = 0 .. s0-1.  For each plane it stages the shifted
// plane w((x0 - d0) mod s0) in LDS; every lane then adds
// w(x0, x1, x2) * P((x1 - d1) mod s1, (x2 - d2) mod s2) over the plane, the
// first factor a uniform value read through the scalar cache.
#include "common.hpp"

#include <cstring>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kLagTile = 16;  // lags per workgroup along d1 and d2

__global__ __launch_bounds__(kLagTile* kLagTile) void autocov_kernel(
	const double* __restrict__ w, int s0, int s1, int s2, int h0, int h1, int h2,
	double inv_n, double* __restrict__ out) {
	extern __shared__ double plane[];
	const int d0 = blockIdx.y;
	const int tiles2 = (h2 + kLagTile - 1) / kLagTile;
	const int d1 = (blockIdx.x / tiles2) * kLagTile + threadIdx.x / kLagTile;
	const int d2 = (blockIdx.x % tiles2) * kLagTile + threadIdx.x % kLagTile;
	const int ps = s1 * s2;
	double acc = 0.0;
	for (int x0 = 0; x0 < s0; ++x0) {
		int y0 = x0 - d0;
		if (y0 < 0) y0 += s0;
		const double* src = w + size_t(y0) * ps;
		__syncthreads();  // the previous plane's readers are done
		for (int q = threadIdx.x; q < ps; q += blockDim.x) plane[q] = src[q];
		__syncthreads();
		const double* wx = w + size_t(x0) * ps;  // uniform across the workgroup
		int y1 = (s1 - d1 % s1) % s1;  // (x1 - d1) mod s1 at x1 = 0
		for (int x1 = 0; x1 < s1; ++x1) {
			const double* prow = plane + y1 * s2;
			const double* wrow = wx + x1 * s2;
			int y2 = (s2 - d2 % s2) % s2;
			for (int x2 = 0; x2 < s2; ++x2) {
				acc = fma(wrow[x2], prow[y2], acc);
				if (++y2 == s2) y2 = 0;
			}
			if (++y1 == s1) y1 = 0;
		}
	}
	if (d1 < h1 && d2 < h2) out[(size_t(d0) * h1 + d1) * h2 + d2] = acc * inv_n;
}

}  // namespace
}  // namespace arma_hip

extern "C" int arma_hip_acf_generate_device(arma_hip_ctx* ctx, const arma_acf_gen_desc* desc,
                                            double* out, size_t cap, int32_t shape_out[3],
                                            double delta_out[3]) {
	using arma_hip::Error;
	try {
		ARMA_HIP_REQUIRE(ctx != nullptr, ARMA_HIP_EINVAL, "null context");
		ARMA_HIP_REQUIRE(desc && out && shape_out && delta_out, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(desc->func >= ARMA_WAVE_SIN && desc->func <= ARMA_WAVE_STANDING,
		                 ARMA_HIP_EINVAL, "bad wave profile");
		ARMA_HIP_CHECK(hipSetDevice(ctx->device));
		std::vector<double> we;
		int s[3], h[3];
		arma_hip::acf_gen_wave(*desc, we, s);
		int32_t so[3];
		double de[3];
		arma_hip::acf_gen_crop(*desc, s, h, so, de);
		ARMA_HIP_REQUIRE(size_t(h[0]) * h[1] * h[2] <= cap, ARMA_HIP_EINVAL,
		                 "ACF output buffer too small");
		const size_t plane_bytes = sizeof(double) * size_t(s[1]) * s[2];
		ARMA_HIP_REQUIRE(plane_bytes <= 160 * 1024, ARMA_HIP_EINVAL,
		                 "wave plane larger than the LDS of a workgroup");
		const size_t nw = we.size(), nout = size_t(h[0]) * h[1] * h[2];
		double* dw = (double*)ctx->yw.get(sizeof(double) * (nw + nout));
		double* dout = dw + nw;
		arma_hip::upload(ctx, dw, we.data(), sizeof(double) * nw);
		if (plane_bytes > 65536)
			ARMA_HIP_CHECK(hipFuncSetAttribute(
				reinterpret_cast<const void*>(&arma_hip::autocov_kernel),
				hipFuncAttributeMaxDynamicSharedMemorySize, int(plane_bytes)));
		const int t1 = arma_hip::div_ceil(h[1], arma_hip::kLagTile);
		const int t2 = arma_hip::div_ceil(h[2], arma_hip::kLagTile);
		{
			arma_hip::KernelTimer timer(ctx, "acf_autocov");
			arma_hip::autocov_kernel<<<dim3(t1 * t2, h[0]), arma_hip::kLagTile * arma_hip::kLagTile,
			                           plane_bytes, ctx->stream>>>(
				dw, s[0], s[1], s[2], h[0], h[1], h[2], 1.0 / double(nw), dout);
			ARMA_HIP_CHECK(hipGetLastError());
		}
		ARMA_HIP_CHECK(hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost,
		                              ctx->stream));
		ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		for (int r = 0; r < 3; ++r) {
			shape_out[r] = so[r];
			delta_out[r] = de[r];
		}
		return ARMA_HIP_OK;
	} catch (const Error& e) {
		arma_hip::set_last_error(e.what());
		return e.code;
	} catch (const std::exception& e) {
		arma_hip::set_last_error(e.what());
		return ARMA_HIP_ERUNTIME;
	}
}
