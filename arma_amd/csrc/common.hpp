// Internal declarations shared by the HIP translation units of libarma_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "arma_hip.h"

namespace arma_hip {

// Error carrying one of the ARMA_HIP_* codes; converted to a return code at
// the C ABI boundary.
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

#define ARMA_HIP_CHECK(expr)                                                         \
	do {                                                                             \
		hipError_t e_ = (expr);                                                      \
		if (e_ != hipSuccess) {                                                      \
			throw ::arma_hip::Error(ARMA_HIP_EDEVICE,                                \
			    std::string(#expr) + ": " + hipGetErrorString(e_));                  \
		}                                                                            \
	} while (0)

#define ARMA_HIP_REQUIRE(cond, code, msg)                                            \
	do {                                                                             \
		if (!(cond)) throw ::arma_hip::Error((code), (msg));                         \
	} while (0)

// Device buffer that grows on demand and is freed with the context.
struct Scratch {
	void* ptr = nullptr;
	size_t bytes = 0;
	void* get(size_t n) {
		if (n > bytes) {
			if (ptr) ARMA_HIP_CHECK(hipFree(ptr));
			ptr = nullptr;
			bytes = 0;
			ARMA_HIP_CHECK(hipMalloc(&ptr, n));
			bytes = n;
		}
		return ptr;
	}
	void release() {
		if (ptr) (void)hipFree(ptr);
		ptr = nullptr;
		bytes = 0;
	}
};

}  // namespace arma_hip

namespace arma_hip {

// Per-kernel timing with HIP events on the context stream (the analogue of
// the reference's prfl counters, src/profile.hh:70-101).
struct ProfileRecord {
	const char* name;
	hipEvent_t start, stop;
};

struct ProfileTotal {
	double ms = 0.0;
	long long count = 0;
};

}  // namespace arma_hip

struct arma_hip_ctx {
	int device = 0;
	bool profiling = false;
	std::vector<arma_hip::ProfileRecord> pending;
	std::map<std::string, arma_hip::ProfileTotal> totals;
	hipStream_t stream = nullptr;
	bool own_stream = false;
	int num_cus = 0;
	// named scratch areas
	arma_hip::Scratch params, seeds, state, phi, partials, flags, noise_tmp, mbox;
	arma_hip::Scratch pot_spec, pot_mid, pot_wn, pot_twiddle, pot_kmin, zeta_tmp, phi_tmp;
	int twiddle_n = 0;
};

namespace arma_hip {

inline size_t numel3(const int32_t s[3]) {
	return size_t(s[0]) * size_t(s[1]) * size_t(s[2]);
}

inline int div_ceil(int a, int b) { return a / b + (a % b ? 1 : 0); }

// Records HIP events around the kernels launched in its scope when the
// context has profiling enabled.
struct KernelTimer {
	arma_hip_ctx* ctx;
	ProfileRecord rec{};
	KernelTimer(arma_hip_ctx* c, const char* name) : ctx(c) {
		if (!ctx->profiling) return;
		rec.name = name;
		ARMA_HIP_CHECK(hipEventCreate(&rec.start));
		ARMA_HIP_CHECK(hipEventCreate(&rec.stop));
		ARMA_HIP_CHECK(hipEventRecord(rec.start, ctx->stream));
	}
	~KernelTimer() {
		if (!ctx->profiling) return;
		if (hipEventRecord(rec.stop, ctx->stream) == hipSuccess) ctx->pending.push_back(rec);
	}
};

// Host-to-device copy of a small, caller-owned host buffer, complete on return.
inline void upload(arma_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
	if (bytes == 0) return;
	ARMA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// Kernel launchers implemented in the .hip translation units.
void launch_white_noise(arma_hip_ctx* ctx, const arma_ar_desc* d, void* eps);
void launch_mt_raw(arma_hip_ctx* ctx, const arma_mt_params* mts, const uint32_t* seeds,
                   int n_streams, int draws, uint32_t* out);
void launch_ar_plain(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta);
void launch_ar_wavefront(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta);
double launch_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int precision,
                      double gamma0);
void launch_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta,
                             void* phi, double* wnmax_out);

}  // namespace arma_hip
