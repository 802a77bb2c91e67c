// Surface generation that allocates no device memory.
//
// arma_hip_ar_generate_noalloc is arma_hip_ar_generate (white noise, then
// the wavefront recurrence, into a caller's device buffer) for callers that
// cannot call hipMalloc: tight serving loops, harnesses that forbid
// allocation.  (Its small host tables go up with a stream synchronisation, so
// it is not meant for graph capture.)  Every buffer its kernels use beyond the
// output (MT parameter records and seeds, coefficient and tap tables, tile
// progress counters, the zero word, the store sink and the halo mailboxes)
// is carved from `scratch`, a static device array of kNoAllocScratchBytes
// that every GPU the library runs on holds until the process ends.  The
// kernels are those of arma_hip_ar_generate, so the results are identical.
//
// Every call on a device shares that device's array: calls that overlap on
// one GPU (several host threads, or several streams) race and must be
// serialised by the caller.  What must start at zero (the progress counters
// and the zero word) is cleared by a kernel on the caller's stream at every
// call; everything else is written before it is read.
#include "common.hpp"

#include <mutex>

namespace arma_hip {
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Called by arma_hip_init with `device` current: the device address of this
// device's copy of the array.  The array's name is a host-side stand-in and
// is never passed to a kernel or a copy.
void noalloc_setup(int device) {
	ARMA_HIP_REQUIRE(device >= 0 && device < kMaxDevices, ARMA_HIP_EDEVICE,
	                 "device index beyond the no-allocation scratch table");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (g_scratch_ptr[device]) return;
	void* p = nullptr;
	ARMA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
	ARMA_HIP_REQUIRE(p != nullptr, ARMA_HIP_EDEVICE, "no device address for the static scratch");
	g_scratch_ptr[device] = p;
}

namespace {

void* scratch_of(int device) {
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Installs the static scratch as the context's borrowed region for one call
// and turns the event profiler off meanwhile (it creates events).
struct Borrow {
	arma_hip_ctx* ctx;
	bool prof;
	Borrow(arma_hip_ctx* c, void* base, size_t bytes) : ctx(c), prof(c->profiling) {
		ctx->borrow_base = (unsigned char*)base;
		ctx->borrow_bytes = bytes;
		ctx->borrow_used = 0;
		ctx->profiling = false;
	}
	~Borrow() {
		ctx->borrow_base = nullptr;
		ctx->borrow_bytes = 0;
		ctx->borrow_used = 0;
		ctx->profiling = prof;
	}
};

}  // namespace
}  // namespace arma_hip

extern "C" int arma_hip_ar_generate_noalloc(arma_hip_ctx* ctx, const arma_ar_desc* desc,
                                            void* zeta) {
	using arma_hip::Error;
	try {
		ARMA_HIP_REQUIRE(ctx != nullptr, ARMA_HIP_EINVAL, "null context");
		ARMA_HIP_CHECK(hipSetDevice(ctx->device));
		ARMA_HIP_REQUIRE(desc && zeta, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(desc->phi && desc->mts && desc->n_mts > 0, ARMA_HIP_EINVAL,
		                 "null generator argument");
		for (int k = 0; k < 3; ++k)
			ARMA_HIP_REQUIRE(desc->shape[k] > 0 && desc->phi_shape[k] > 0, ARMA_HIP_EINVAL,
			                 "bad shape");
		ARMA_HIP_REQUIRE(desc->precision == ARMA_FP64 || desc->precision == ARMA_FP32,
		                 ARMA_HIP_EINVAL, "bad precision");
		ARMA_HIP_REQUIRE(desc->algorithm == ARMA_AR_WAVEFRONT, ARMA_HIP_EINVAL,
		                 "the no-allocation entry point runs the wavefront recurrence only");
		void* base = arma_hip::scratch_of(ctx->device);
		ARMA_HIP_REQUIRE(base != nullptr, ARMA_HIP_EDEVICE,
		                 "no static scratch on this device (context not made by arma_hip_init)");
		// reject what does not fit before anything is enqueued
		auto al = [](size_t b) { return (b + 255) / 256 * 256; };
		const size_t need = al(arma_hip::white_noise_work_bytes(desc)) +
		                    arma_hip::ar_wavefront_work_bytes(desc);
		ARMA_HIP_REQUIRE(need <= arma_hip::kNoAllocScratchBytes, ARMA_HIP_EINVAL,
		                 "surface needs " + std::to_string(need) +
		                     " bytes of device scratch; the no-allocation entry point has " +
		                     std::to_string(arma_hip::kNoAllocScratchBytes));
		arma_hip::Borrow borrow(ctx, base, arma_hip::kNoAllocScratchBytes);
		arma_hip::launch_white_noise(ctx, desc, zeta);
		arma_hip::launch_ar_wavefront(ctx, desc, zeta);
		return ARMA_HIP_OK;
	} catch (const Error& e) {
		arma_hip::set_last_error(e.what());
		return e.code;
	} catch (const std::bad_alloc&) {
		arma_hip::set_last_error("out of host memory");
		return ARMA_HIP_ENOMEM;
	} catch (const std::exception& e) {
		arma_hip::set_last_error(e.what());
		return ARMA_HIP_ERUNTIME;
	}
}
