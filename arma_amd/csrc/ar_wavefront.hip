// Persistent wavefront kernel for the 3-D AR recurrence
//   zeta(t,x,y) = eps(t,x,y) + sum_{k,i,j} phi(k,i,j) zeta(t-k, x-i, y-j)
// (src/generator/ar_model_surf.cc:3-36; zero outside the grid, which is the
// reference's truncated causal boundary).
//
// Decomposition:
//  * Every workgroup owns a TX x TY tile of (x, y) columns and marches over
//    the whole t axis.  Lane (x, y) computes slice t = s - x - y at step s,
//    so every in-slice dependency (x-i, y-j) was finished at an earlier step:
//    the skew replaces the reference's partition DAG inside the tile.
//  * "Push" form of the sum: when lane (x, y) reaches slice t it reads the
//    P1*P2 values zeta(t, x-i, y-j) once from LDS and adds
//    sum_{i,j} phi(k,i,j) * zeta(t, x-i, y-j) to accumulator k, which holds
//    the partial sum of point (t+k, x, y).  Accumulator 0 is complete once
//    the in-slice terms are added, so P1*P2 LDS reads feed P0*P1*P2 FMAs.
//  * LDS holds a ring of R = P1+P2-1 slices of the tile plus its causal halo
//    (P1-1 rows above, P2-1 columns to the left): slot t mod R of position p
//    is rewritten only after every reader of slice t is done.
//  * Halo values come from the neighbouring tiles through HBM.  A tile
//    publishes the number of steps it has completed (agent-scope release
//    after all its stores); a consumer acquires its neighbours' progress
//    before issuing the halo loads, which are prefetched D steps ahead.
//    Tiles only wait on tiles with smaller workgroup ids, so the launch
//    cannot deadlock.
//  * FP32 mode keeps every accumulator as a compensated (Kahan/TwoSum) pair;
//    each step adds one P1*P2-term FP32 partial to it.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kTX = 16;
constexpr int kTY = 16;
constexpr int kThreads = kTX * kTY;
constexpr int kD = 4;        // prefetch distance in steps
constexpr int kB = 8;        // steps between progress publications / polls
constexpr int kMaxP12 = 10;  // largest P1, P2 handled

template <class T, bool KAHAN>
struct Acc;

template <class T>
struct Acc<T, false> {
	T s;
	__device__ void zero() { s = T(0); }
	__device__ void add(T v) { s += v; }
	__device__ T value() const { return s; }
};

template <class T>
struct Acc<T, true> {
	T s, c;
	__device__ void zero() {
		s = T(0);
		c = T(0);
	}
	__device__ void add(T v) {
		// TwoSum(s, v); the rounding error goes to the compensation term
		const T r = s + v;
		const T bb = r - s;
		c += (s - (r - bb)) + (v - bb);
		s = r;
	}
	__device__ T value() const { return s + c; }
};

__global__ void reset_progress(int* p, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

__device__ __forceinline__ int load_progress(const int* p) {
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, int P0, bool KAHAN>
__global__ __launch_bounds__(kThreads) void ar_wave_kernel(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ phi, int P1, int P2,
	int nY, int* __restrict__ progress, int total_steps) {
	extern __shared__ unsigned char smem_raw[];
	T* ring = reinterpret_cast<T*>(smem_raw);
	const int HX = P1 - 1, HY = P2 - 1;
	const int WX = kTX + HX, WY = kTY + HY;
	const int R = P1 + P2 - 1;
	const int slot = WX * WY;
	const int b = blockIdx.x;
	const int X = b / nY, Y = b - (b / nY) * nY;
	const int X0 = X * kTX, Y0 = Y * kTY;
	const int tid = threadIdx.x;
	const int lx = tid / kTY, ly = tid - (tid / kTY) * kTY;
	const bool active = (X0 + lx < nx) && (Y0 + ly < ny);
	const size_t sxy = size_t(nx) * ny;
	const size_t own_off = size_t(X0 + lx) * ny + (Y0 + ly);
	const int own_lds = (lx + HX) * WY + (ly + HY);

	// halo positions: top band (hx < 0, all hy), then left band (hx >= 0, hy < 0)
	const int H = WX * WY - kTX * kTY;
	int hsk[2], hlds[2];  // skew hx + hy and LDS position of each halo slot
	size_t hoff[2];
	bool hval[2], hin[2];
#pragma unroll
	for (int r = 0; r < 2; ++r) {
		const int h = tid + r * kThreads;
		int hx = 0, hy = 0;
		hin[r] = false;
		hval[r] = false;
		hsk[r] = 0;
		hlds[r] = 0;
		hoff[r] = 0;
		if (h < H) {
			if (h < HX * WY) {
				hx = -HX + h / WY;
				hy = -HY + (h - (h / WY) * WY);
			} else {
				const int q = h - HX * WY;
				hx = q / HY;
				hy = -HY + (q - (q / HY) * HY);
			}
			hval[r] = true;
			hsk[r] = hx + hy;  // step s carries slice s - hx - hy
			hlds[r] = (hx + HX) * WY + (hy + HY);
			const int gx = X0 + hx, gy = Y0 + hy;
			hin[r] = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
			if (hin[r]) hoff[r] = size_t(gx) * ny + gy;
		}
	}
	// neighbour tiles and how far ahead they must be (in their steps)
	const int nb_top = X > 0 ? b - nY : -1;
	const int nb_left = Y > 0 ? b - 1 : -1;
	const int nb_corner = (X > 0 && Y > 0) ? b - nY - 1 : -1;

	Acc<T, KAHAN> acc[P0];
#pragma unroll
	for (int k = 0; k < P0; ++k) acc[k].zero();
	const T phi000 = phi[0];

	// Step u of the loop is skew step s = u + s0; s0 < 0 lets the halo
	// positions (negative skew hx + hy) deliver their first slices.  Progress
	// counts completed u-steps; every tile uses the same s0.
	const int s0 = -(HX + HY);
	const int total_u = total_steps - s0;

	// the ring starts zeroed: halo positions outside the grid stay zero
	for (int q = tid; q < R * slot; q += kThreads) ring[q] = T(0);

	// prefetch queues: noise of the own point and halo values, kD steps deep
	T qe[kD], qh0[kD], qh1[kD];
#pragma unroll
	for (int d = 0; d < kD; ++d) {
		qe[d] = T(0);
		qh0[d] = T(0);
		qh1[d] = T(0);
	}
	int ready_to = 0;  // loads for u-steps < ready_to are covered by acquired progress

	auto wait_neighbours = [&](int last_u) {
		// make the halo values consumed at u-steps <= last_u readable
		if (tid == 0) {
			const int need_t = min(last_u + 1 + kTX, total_u);
			const int need_l = min(last_u + 1 + kTY, total_u);
			const int need_c = min(last_u + 1 + kTX + kTY, total_u);
			if (nb_top >= 0)
				while (load_progress(progress + nb_top) < need_t) __builtin_amdgcn_s_sleep(1);
			if (nb_left >= 0)
				while (load_progress(progress + nb_left) < need_l) __builtin_amdgcn_s_sleep(1);
			if (nb_corner >= 0)
				while (load_progress(progress + nb_corner) < need_c) __builtin_amdgcn_s_sleep(1);
		}
	};

	auto issue_loads = [&](int u2, T& e, T& h0, T& h1) {
		// loads that u-step u2 will consume
		const int s2 = u2 + s0;
		const int t = s2 - lx - ly;
		e = (active && t >= 0 && t < nt) ? zeta[size_t(t) * sxy + own_off] : T(0);
		const int t0 = s2 - hsk[0];
		h0 = (hval[0] && hin[0] && t0 >= 0 && t0 < nt) ? zeta[size_t(t0) * sxy + hoff[0]] : T(0);
		const int t1 = s2 - hsk[1];
		h1 = (hval[1] && hin[1] && t1 >= 0 && t1 < nt) ? zeta[size_t(t1) * sxy + hoff[1]] : T(0);
	};

	// prologue: loads for u-steps 0..kD-1
	wait_neighbours(kD - 1 + kB - 1);
	ready_to = kD + kB - 1;
	__syncthreads();
#pragma unroll
	for (int d = 0; d < kD; ++d) issue_loads(d, qe[d], qh0[d], qh1[d]);

	auto step = [&](int u, T& e_q, T& h0_q, T& h1_q) {
		const int s = u + s0;
		// 1. halo values of step s into the ring
		if (hval[0]) {
			const int t0 = s - hsk[0];
			if (t0 >= 0 && t0 < nt) ring[(t0 % R) * slot + hlds[0]] = h0_q;
		}
		if (hval[1]) {
			const int t1 = s - hsk[1];
			if (t1 >= 0 && t1 < nt) ring[(t1 % R) * slot + hlds[1]] = h1_q;
		}
		// 2. own point
		const int t = s - lx - ly;
		if (active && t >= 0 && t < nt) {
			const T* base = ring + (t % R) * slot + own_lds;
			T part[P0];
#pragma unroll
			for (int k = 0; k < P0; ++k) part[k] = T(0);
			for (int i = 0; i < P1; ++i) {
				for (int j = 0; j < P2; ++j) {
					if (i == 0 && j == 0) continue;
					const T v = base[-(i * WY + j)];
#pragma unroll
					for (int k = 0; k < P0; ++k) part[k] = fma(phi[(k * P1 + i) * P2 + j], v, part[k]);
				}
			}
			acc[0].add(part[0]);
			const T eps = e_q;
			const T z = eps + (acc[0].value() + phi000 * eps);
			ring[(t % R) * slot + own_lds] = z;
			zeta[size_t(t) * sxy + own_off] = z;
#pragma unroll
			for (int k = 1; k < P0; ++k) acc[k].add(fma(phi[k * P1 * P2], z, part[k]));
#pragma unroll
			for (int k = 0; k + 1 < P0; ++k) acc[k] = acc[k + 1];
			acc[P0 - 1].zero();
		}
		// 3. acquire neighbour progress for the next kB steps of loads
		const int u_load = u + kD;
		if (u_load >= ready_to && u_load < total_u) {
			wait_neighbours(u_load + kB - 1);
			ready_to = u_load + kB;
		}
		// 4. publish own progress
		const bool publish = ((u + 1) % kB == 0) || (u + 1 == total_u);
		if (publish) __threadfence();
		__syncthreads();
		if (publish && tid == 0)
			__hip_atomic_store(progress + b, u + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		// 5. loads for step u + kD into the queue slot just consumed
		if (u_load < total_u) issue_loads(u_load, e_q, h0_q, h1_q);
	};

	int u = 0;
	for (; u + kD <= total_u; u += kD) {
#pragma unroll
		for (int d = 0; d < kD; ++d) step(u + d, qe[d], qh0[d], qh1[d]);
	}
	for (int d = 0; u < total_u; ++u, ++d) {
		if (d == 0) step(u, qe[0], qh0[0], qh1[0]);
		else if (d == 1) step(u, qe[1], qh0[1], qh1[1]);
		else if (d == 2) step(u, qe[2], qh0[2], qh1[2]);
		else step(u, qe[3], qh0[3], qh1[3]);
	}
}

template <class T, bool KAHAN>
void launch_typed(arma_hip_ctx* ctx, const arma_ar_desc* d, T* zeta) {
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	ARMA_HIP_REQUIRE(P1 <= kMaxP12 && P2 <= kMaxP12 && P0 <= 10, ARMA_HIP_EINVAL,
	                 "AR order above 10 is not supported by the wavefront kernel");
	const size_t nphi = size_t(P0) * P1 * P2;
	std::vector<T> ph(d->phi, d->phi + nphi);
	T* dphi = (T*)ctx->phi.get(nphi * sizeof(T));
	upload(ctx, dphi, ph.data(), nphi * sizeof(T));
	const int nX = div_ceil(nx, kTX), nY = div_ceil(ny, kTY);
	const int ntiles = nX * nY;
	const int total_steps = nt + kTX + kTY - 2;
	int* prog = (int*)ctx->flags.get(sizeof(int) * size_t(ntiles));
	reset_progress<<<div_ceil(ntiles, 256), 256, 0, ctx->stream>>>(prog, ntiles);
	ARMA_HIP_CHECK(hipGetLastError());
	const int R = P1 + P2 - 1;
	const size_t lds = size_t(R) * (kTX + P1 - 1) * (kTY + P2 - 1) * sizeof(T);
	ARMA_HIP_REQUIRE(lds <= 160 * 1024, ARMA_HIP_EINVAL, "AR order too large for the LDS ring");
	KernelTimer timer(ctx, "ar_wavefront");
	// rings above 64 KiB (P1, P2 >= 9 in FP64) use the larger LDS of gfx950
#define ARMA_WAVE(P)                                                                          \
	case P:                                                                                   \
		if (lds > 65536)                                                                      \
			ARMA_HIP_CHECK(hipFuncSetAttribute(                                               \
				reinterpret_cast<const void*>(&ar_wave_kernel<T, P, KAHAN>),                  \
				hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));                       \
		ar_wave_kernel<T, P, KAHAN><<<ntiles, kThreads, lds, ctx->stream>>>(                  \
			zeta, nt, nx, ny, dphi, P1, P2, nY, prog, total_steps);                           \
		break;
	switch (P0) {
		ARMA_WAVE(1)
		ARMA_WAVE(2)
		ARMA_WAVE(3)
		ARMA_WAVE(4)
		ARMA_WAVE(5)
		ARMA_WAVE(6)
		ARMA_WAVE(7)
		ARMA_WAVE(8)
		ARMA_WAVE(9)
		ARMA_WAVE(10)
	default:
		throw Error(ARMA_HIP_EINVAL, "bad AR order");
	}
#undef ARMA_WAVE
	ARMA_HIP_CHECK(hipGetLastError());
}

}  // namespace

void launch_ar_wavefront(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta) {
	if (d->precision == ARMA_FP32)
		launch_typed<float, true>(ctx, d, (float*)zeta);
	else
		launch_typed<double, false>(ctx, d, (double*)zeta);
}

}  // namespace arma_hip
