// The plain GPU AR path: the same partition DAG as the reference OpenCL
// backend (src/generator/ar_model_opencl.cc:46-121), where each partition is
// computed by ONE work-item (src/kernels/ar_generate_surface.cl:6-49) once
// its seven predecessors are done.  Here every anti-diagonal level of the
// partition grid is one launch, which is the same schedule without the
// event wait lists.
//
// Each point's sum runs over (k, i, j) in the reference order with separate
// multiply and add roundings, so the FP64 result is bit-identical to the
// sequential CPU recurrence (src/generator/ar_model_surf.cc:20-35) for any
// valid partition shape.
#include "common.hpp"

#include <vector>

namespace arma_hip {
namespace {

template <class T>
__global__ void ar_plain_kernel(T* __restrict__ zeta, int3 shape, const T* __restrict__ phi,
                                int3 ps, int3 part, int3 np, const int* __restrict__ level_list,
                                int count) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= count) return;
	const int pid = level_list[q];
	const int a = pid / (np.y * np.z), b = (pid / np.z) % np.y, c = pid % np.z;
	const int t0 = a * part.x, x0 = b * part.y, y0 = c * part.z;
	const int t1 = min(t0 + part.x, shape.x), x1 = min(x0 + part.y, shape.y),
	          y1 = min(y0 + part.z, shape.z);
	const size_t sx = shape.z, st = size_t(shape.y) * shape.z;
	for (int t = t0; t < t1; ++t) {
		for (int x = x0; x < x1; ++x) {
			for (int y = y0; y < y1; ++y) {
				const int m1 = min(t + 1, ps.x);
				const int m2 = min(x + 1, ps.y);
				const int m3 = min(y + 1, ps.z);
				T sum = 0;
				for (int k = 0; k < m1; ++k)
					for (int i = 0; i < m2; ++i)
						for (int j = 0; j < m3; ++j)
							sum += phi[(k * ps.y + i) * ps.z + j] *
							       zeta[size_t(t - k) * st + size_t(x - i) * sx + (y - j)];
				zeta[size_t(t) * st + size_t(x) * sx + y] += sum;
			}
		}
	}
}

}  // namespace

void launch_ar_plain(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta) {
	const int3 shape{d->shape[0], d->shape[1], d->shape[2]};
	const int3 ps{d->phi_shape[0], d->phi_shape[1], d->phi_shape[2]};
	int3 part{d->partition[0], d->partition[1], d->partition[2]};
	if (part.x <= 0 || part.y <= 0 || part.z <= 0) part = shape;
	// The 7-predecessor DAG is only valid when partitions span order-1
	// points in every dimension (src/generator/ar_model_parallel.cc:158-171).
	ARMA_HIP_REQUIRE((part.x >= ps.x - 1 || part.x >= shape.x) &&
	                     (part.y >= ps.y - 1 || part.y >= shape.y) &&
	                     (part.z >= ps.z - 1 || part.z >= shape.z),
	                 ARMA_HIP_EINVAL, "partition is smaller than the AR order");
	const int3 np{div_ceil(shape.x, part.x), div_ceil(shape.y, part.y), div_ceil(shape.z, part.z)};
	const int nlevels = np.x + np.y + np.z - 2;
	std::vector<int> list;
	std::vector<int> offset(nlevels + 1, 0);
	list.reserve(size_t(np.x) * np.y * np.z);
	for (int L = 0; L < nlevels; ++L) {
		offset[L] = int(list.size());
		for (int a = 0; a < np.x; ++a)
			for (int b = 0; b < np.y; ++b) {
				const int c = L - a - b;
				if (c >= 0 && c < np.z) list.push_back((a * np.y + b) * np.z + c);
			}
	}
	offset[nlevels] = int(list.size());
	int* dlist = (int*)work_buffer(ctx, ctx->flags, sizeof(int) * list.size());
	upload(ctx, dlist, list.data(), sizeof(int) * list.size());
	const size_t nphi = size_t(ps.x) * ps.y * ps.z;
	const bool fp32 = d->precision == ARMA_FP32;
	void* dphi = work_buffer(ctx, ctx->phi, nphi * 8);
	if (fp32) {
		std::vector<float> pf(d->phi, d->phi + nphi);
		upload(ctx, dphi, pf.data(), nphi * 4);
	} else {
		upload(ctx, dphi, d->phi, nphi * 8);
	}
	KernelTimer timer(ctx, "ar_plain");
	for (int L = 0; L < nlevels; ++L) {
		const int count = offset[L + 1] - offset[L];
		if (count == 0) continue;
		const int tpb = 64;
		if (fp32)
			ar_plain_kernel<float><<<div_ceil(count, tpb), tpb, 0, ctx->stream>>>(
				(float*)zeta, shape, (const float*)dphi, ps, part, np, dlist + offset[L], count);
		else
			ar_plain_kernel<double><<<div_ceil(count, tpb), tpb, 0, ctx->stream>>>(
				(double*)zeta, shape, (const double*)dphi, ps, part, np, dlist + offset[L], count);
		ARMA_HIP_CHECK(hipGetLastError());
	}
}

}  // namespace arma_hip
