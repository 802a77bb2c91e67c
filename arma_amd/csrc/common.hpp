// Internal declarations shared by the HIP translation units of libarma_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
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
	Scratch() = default;
	Scratch(const Scratch&) = delete;
	Scratch& operator=(const Scratch&) = delete;
	~Scratch() { release(); }
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
	arma_hip::Scratch pot_spec, pot_mid, pot_wn, pot_twiddle, pot_kmin, pot_field, pot_edge,
	    zeta_tmp, phi_tmp;
	arma_hip::Scratch yw;
	int twiddle_n = 0;
	// a second stream (default priority) on which the potential's row pass of
	// one chunk overlaps the column pass of the next, and the events that
	// order the two over their ping-pong intermediate buffers
	hipStream_t aux_stream = nullptr;
	hipEvent_t pot_events[4] = {};
	// batch configurations chosen by the autotuner, by batch signature:
	// tile width, schedule, workgroups per CU
	std::map<std::string, std::vector<int>> batch_tuned;
	// slabs of the streaming writer (pinned host and device staging), kept
	// for the next writer on this context; writer_cache_busy while one uses them
	void* writer_host = nullptr;
	arma_hip::Scratch writer_stage;
	size_t writer_bytes = 0;
	bool writer_cache_busy = false;
	std::mutex writer_mutex;
	// While set, work_buffer() carves device buffers from this caller-given
	// region (bump order, 256-byte aligned) instead of growing the scratch
	// areas above: the no-allocation entry point installs the library's
	// static device scratch here for the duration of one call.
	unsigned char* borrow_base = nullptr;
	size_t borrow_bytes = 0, borrow_used = 0;
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
	hipStream_t stream;
	ProfileRecord rec{};
	KernelTimer(arma_hip_ctx* c, const char* name) : KernelTimer(c, name, c->stream) {}
	KernelTimer(arma_hip_ctx* c, const char* name, hipStream_t s) : ctx(c), stream(s) {
		if (!ctx->profiling) return;
		rec.name = name;
		ARMA_HIP_CHECK(hipEventCreate(&rec.start));
		ARMA_HIP_CHECK(hipEventCreate(&rec.stop));
		ARMA_HIP_CHECK(hipEventRecord(rec.start, stream));
	}
	~KernelTimer() {
		if (!ctx->profiling) return;
		if (hipEventRecord(rec.stop, stream) == hipSuccess) ctx->pending.push_back(rec);
	}
};

// A device buffer of n bytes for one launch sequence: the context's scratch
// area s (grown on demand), or, while a borrowed region is installed on the
// context, the next n bytes of that region (ARMA_HIP_EINVAL when it is full).
inline void* work_buffer(arma_hip_ctx* ctx, Scratch& s, size_t n) {
	if (!ctx->borrow_base) return s.get(n);
	const size_t at = (ctx->borrow_used + 255) / 256 * 256;
	ARMA_HIP_REQUIRE(at <= ctx->borrow_bytes && n <= ctx->borrow_bytes - at, ARMA_HIP_EINVAL,
	                 "the request needs more device scratch than the no-allocation entry "
	                 "point has (" + std::to_string(ctx->borrow_bytes) + " bytes)");
	ctx->borrow_used = at + n;
	return ctx->borrow_base + at;
}

// Host-to-device copy of a small, caller-owned host buffer, complete on return.
inline void upload(arma_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
	if (bytes == 0) return;
	ARMA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// Records msg as the calling thread's last error (arma_hip_last_error).
void set_last_error(const char* msg);

// Kernel launchers implemented in the .hip translation units.
void launch_white_noise(arma_hip_ctx* ctx, const arma_ar_desc* d, void* eps);
void launch_mt_raw(arma_hip_ctx* ctx, const arma_mt_params* mts, const uint32_t* seeds,
                   int n_streams, int draws, uint32_t* out);
void launch_ar_plain(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta);
void launch_ar_wavefront(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta);
double launch_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int precision,
                      double gamma0);

// Device bytes the buffers of launch_white_noise and launch_ar_wavefront take
// for d (each buffer 256-byte aligned, in the order the launchers take them).
size_t white_noise_work_bytes(const arma_ar_desc* d);
size_t ar_wavefront_work_bytes(const arma_ar_desc* d);
// Looks up this device's copy of the static scratch of the no-allocation
// entry point (once per device; the device must be current).
void noalloc_setup(int device);

// FP64 Yule-Walker (Choi) and white-noise variance on the GPU.  phi gets
// the l x l x l coefficients before chopping.
void yule_walker_choi_device(arma_hip_ctx* ctx, const double* acf, int n, int max_order,
                             double var_eps, bool determine, std::vector<double>& phi, int& l,
                             double& varwn);
// Several Choi recursions in one launch (one workgroup per model).
struct YwProblem {
	const double* acf;  // n^3 cube
	int n, max_order;
	double var_eps;
	bool determine;
};
struct YwSolution {
	int status = ARMA_HIP_ERUNTIME;
	int l = 0;
	double varwn = 0.0;
	std::vector<double> phi;  // l^3, before chopping
};
void yule_walker_choi_device_many(arma_hip_ctx* ctx, const std::vector<YwProblem>& probs,
                                  std::vector<YwSolution>& sols);
double white_noise_variance_device(arma_hip_ctx* ctx, const double* acf, const int32_t acf_shape[3],
                                   const double* phi, const int32_t phi_shape[3]);
// Roots of c[0] + c[1] x + ... + c[n-1] x^(n-1) (eigenvalues of the balanced
// companion matrix by Francis double-shift QR); ARMA_HIP_ERUNTIME when the
// leading coefficient is zero or the iteration does not converge.
int polynomial_roots(const std::vector<double>& c, std::vector<double>& re,
                     std::vector<double>& im);
// ACF_generator up to the auto-covariance (the decayed, rescaled plain wave,
// extents s) and the crop of the auto-covariance to its first half (h, and
// the ACF's shape and grid spacing).
void acf_gen_wave(const arma_acf_gen_desc& d, std::vector<double>& we, int s[3]);
void acf_gen_crop(const arma_acf_gen_desc& d, const int s[3], int h[3], int32_t shape_out[3],
                  double delta_out[3]);
// chop (src/chop.cc:4-26) of an n0 x n1 x n2 array: the trimmed shape.
void chop_shape(const std::vector<double>& phi, const int shape[3], double eps, int out[3]);

// ---- batches of surfaces in one arena ----

// The output block of one noise stream inside a batch arena.
struct NoiseBlockRec {
	long long base;  // arena element of the surface's (0, 0, 0)
	int sy, sz;      // surface extents along x and y
	int lo0, lo1, lo2;
	int e0, e1, e2;
	float sdf;
	double sd;
};

// One surface of a batch: its first arena element and its generator.
struct BatchMember {
	size_t zoff;
	const arma_ar_desc* d;
	double gamma0;  // post-rescale target variance; <= 0 skips the rescale
	int t0;         // leading slices given by the caller (noise and values start at t0)
};

struct BatchArConfig {
	int tile_x = 16;       // 8, 16 or 32; tiles are tile_x x (256 / tile_x)
	int schedule = ARMA_BATCH_SURFACE_MAJOR;
	int slots_per_cu = 0;  // resident workgroups per CU, 0 = as many as fit
};

struct NoiseBatch;
struct ArBatch;
struct RescaleBatch;
// Construction uploads every table and allocates every buffer; the launches
// only enqueue kernels on the context stream (no allocation, no host wait),
// so they can be captured into a graph.
NoiseBatch* noise_batch_create(arma_hip_ctx* ctx, int precision,
                               const std::vector<BatchMember>& m);
void noise_batch_launch(arma_hip_ctx* ctx, NoiseBatch* b, void* arena);
void noise_batch_destroy(NoiseBatch* b);
ArBatch* ar_batch_create(arma_hip_ctx* ctx, int precision, const std::vector<BatchMember>& m,
                         const BatchArConfig& cfg);
// Enqueues the persistent recurrence launch.  Launches on one device are
// chained (each waits for the previous one, whatever its stream) outside
// graph capture.
void ar_batch_launch(arma_hip_ctx* ctx, ArBatch* b, void* arena);
// Waits for the stream and throws ARMA_HIP_ERUNTIME when the last launch's
// bounded waits gave up.
void ar_batch_check(arma_hip_ctx* ctx, ArBatch* b);
void ar_batch_destroy(ArBatch* b);
int ar_batch_grid(const ArBatch* b);
bool ar_batch_fits(const std::vector<BatchMember>& m, int tile_x);
int ar_batch_max_slots(const ArBatch* b);
RescaleBatch* rescale_batch_create(arma_hip_ctx* ctx, int precision,
                                   const std::vector<BatchMember>& m);
void rescale_batch_launch(arma_hip_ctx* ctx, RescaleBatch* b, void* arena);
// Factors of the last launch, one per member (0 for members without a
// rescale); waits for the stream.
void rescale_batch_factors(arma_hip_ctx* ctx, RescaleBatch* b, double* out);
void rescale_batch_destroy(RescaleBatch* b);
// Empty objects and rebuilds in place: a rebuild reuses the device buffers
// of earlier builds (they only grow), so a serving lane allocates nothing
// once warm.
NoiseBatch* noise_batch_new();
void noise_batch_build(arma_hip_ctx* ctx, NoiseBatch* b, int precision,
                       const std::vector<BatchMember>& m);
ArBatch* ar_batch_new();
void ar_batch_build(arma_hip_ctx* ctx, ArBatch* b, int precision,
                    const std::vector<BatchMember>& m, const BatchArConfig& cfg);
RescaleBatch* rescale_batch_new();
void rescale_batch_build(arma_hip_ctx* ctx, RescaleBatch* b, int precision,
                         const std::vector<BatchMember>& m);
// Device bytes a member adds to a recurrence launch (mailboxes, sink,
// progress counters).
size_t ar_batch_member_bytes(const arma_ar_desc* d, int tile_x);
void launch_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta,
                             void* phi, double* wnmax_out);
void launch_nit_realisation(arma_hip_ctx* ctx, void* zeta, size_t n, int precision,
                            double gamma0, const arma_nit_desc* d);

// solver 0: Linear_solver, 1: High_amplitude_solver
void launch_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta, void* phi,
                      double* wnmax_out, int solver);

}  // namespace arma_hip
