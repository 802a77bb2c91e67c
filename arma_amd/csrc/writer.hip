// Streaming writer: device arrays to a file in network byte order, the
// zeta.bin format of io::Binary_stream (src/io/binary_stream.hh:43-70), with
// the copies and the file writes overlapped with the caller's GPU work.
//
// The reference streams finished t-slabs of the surface from a writer thread
// while the OpenMP workers compute the remaining partitions
// (src/generator/ar_model_parallel.cc:112-145) and byte-swaps every element
// on the host (bits::Bytes<T>::to_network_format).  Here an append records an
// event on the context's stream.  A feeder thread waits for that event and
// then, slab by slab, runs on its own copy stream a byte-swap kernel into a
// device staging slot and a copy into the matching pinned host slot; I/O
// threads write finished slots to their offsets in the file (pwrite), so
// that several slabs are in the page cache copy at once.  kSlots slots go
// round; a slot is reused once its write has returned.  The context's stream
// is never blocked: the GPU goes on with later work (the velocity potential)
// while the surface goes to the file.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace arma_hip {
namespace {

constexpr int kSlots = 4;        // slabs in flight
constexpr int kIoThreads = 4;    // concurrent pwrite calls

// dst[i] = src[i] with its bytes reversed (W = uint32_t or uint64_t words).
template <class W>
__global__ __launch_bounds__(256) void bswap_kernel(const W* __restrict__ src,
                                                    W* __restrict__ dst, size_t n) {
	const size_t stride = size_t(gridDim.x) * blockDim.x;
	for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		if constexpr (sizeof(W) == 8)
			dst[i] = __builtin_bswap64(src[i]);
		else
			dst[i] = __builtin_bswap32(src[i]);
	}
}

struct Job {
	const unsigned char* src;
	size_t n;
	hipEvent_t ready;
};

struct IoTask {
	int slot;
	size_t len;
	off_t offset;
};

}  // namespace
}  // namespace arma_hip

struct arma_hip_writer {
	arma_hip_ctx* ctx = nullptr;
	int device = 0;
	int fd = -1;
	int esize = 8;
	size_t slab = 0;
	hipStream_t copy = nullptr;
	unsigned char* host = nullptr;   // kSlots pinned slabs
	unsigned char* stage = nullptr;  // kSlots device slabs
	bool cached = false;             // host/stage belong to the context's cache
	hipEvent_t copied[arma_hip::kSlots] = {};
	bool slot_free[arma_hip::kSlots] = {true, true, true, true};
	std::thread feeder;
	std::vector<std::thread> io;
	std::mutex m;
	std::condition_variable cv;
	std::deque<arma_hip::Job> jobs;
	std::deque<arma_hip::IoTask> tasks;
	bool closing = false;     // no more appends
	bool fed = false;         // the feeder has issued everything
	int error = ARMA_HIP_OK;
	std::string message;
	off_t offset = 0;         // file offset of the next appended byte
	uint64_t written = 0;
	std::chrono::steady_clock::time_point t_first, t_last;
	bool started = false;
};

namespace arma_hip {
namespace {

template <class F>
int wguard(F&& f) {
	try {
		f();
		return ARMA_HIP_OK;
	} catch (const Error& e) {
		set_last_error(e.what());
		return e.code;
	} catch (const std::bad_alloc&) {
		set_last_error("out of host memory");
		return ARMA_HIP_ENOMEM;
	} catch (const std::exception& e) {
		set_last_error(e.what());
		return ARMA_HIP_ERUNTIME;
	}
}

void fail(arma_hip_writer* w, int code, const std::string& msg) {
	std::lock_guard<std::mutex> lk(w->m);
	if (w->error == ARMA_HIP_OK) {
		w->error = code;
		w->message = msg;
	}
	w->cv.notify_all();
}

// Byte-swap bytes [off, off + len) of job j into staging slot b and copy
// them to pinned slot b.
void issue(arma_hip_writer* w, const Job& j, size_t off, size_t len, int b) {
	unsigned char* st = w->stage + size_t(b) * w->slab;
	const size_t words = len / size_t(w->esize);
	const int grid = int(std::min<size_t>((words + 255) / 256, 4096));
	if (w->esize == 8)
		bswap_kernel<uint64_t><<<grid, 256, 0, w->copy>>>(
			reinterpret_cast<const uint64_t*>(j.src + off), reinterpret_cast<uint64_t*>(st), words);
	else
		bswap_kernel<uint32_t><<<grid, 256, 0, w->copy>>>(
			reinterpret_cast<const uint32_t*>(j.src + off), reinterpret_cast<uint32_t*>(st), words);
	ARMA_HIP_CHECK(hipGetLastError());
	ARMA_HIP_CHECK(hipMemcpyAsync(w->host + size_t(b) * w->slab, st, len, hipMemcpyDeviceToHost,
	                              w->copy));
	ARMA_HIP_CHECK(hipEventRecord(w->copied[b], w->copy));
}

void feeder_main(arma_hip_writer* w) {
	(void)hipSetDevice(w->device);
	int next = 0;
	for (;;) {
		Job j;
		{
			std::unique_lock<std::mutex> lk(w->m);
			w->cv.wait(lk, [&] { return w->closing || !w->jobs.empty(); });
			if (w->jobs.empty()) {
				w->fed = true;
				w->cv.notify_all();
				return;
			}
			j = w->jobs.front();
			w->jobs.pop_front();
		}
		try {
			// the producer's work first; the time from here to the last write is
			// what the writer reports
			ARMA_HIP_CHECK(hipEventSynchronize(j.ready));
			{
				std::lock_guard<std::mutex> lk(w->m);
				if (!w->started) {
					w->started = true;
					w->t_first = std::chrono::steady_clock::now();
					w->t_last = w->t_first;
				}
			}
			for (size_t off = 0; off < j.n; off += w->slab) {
				const size_t len = std::min(w->slab, j.n - off);
				const int b = next;
				next = (next + 1) % kSlots;
				off_t at;
				{
					std::unique_lock<std::mutex> lk(w->m);
					w->cv.wait(lk, [&] { return w->slot_free[b] || w->error != ARMA_HIP_OK; });
					if (w->error != ARMA_HIP_OK) break;
					w->slot_free[b] = false;
					at = w->offset;
					w->offset += off_t(len);
				}
				issue(w, j, off, len, b);
				{
					std::lock_guard<std::mutex> lk(w->m);
					w->tasks.push_back(IoTask{b, len, at});
				}
				w->cv.notify_all();
			}
		} catch (const Error& e) {
			fail(w, e.code, e.what());
		} catch (const std::exception& e) {
			fail(w, ARMA_HIP_ERUNTIME, e.what());
		}
		(void)hipEventDestroy(j.ready);
	}
}

void io_main(arma_hip_writer* w) {
	(void)hipSetDevice(w->device);
	for (;;) {
		IoTask t;
		{
			std::unique_lock<std::mutex> lk(w->m);
			w->cv.wait(lk, [&] { return !w->tasks.empty() || w->fed; });
			if (w->tasks.empty()) return;
			t = w->tasks.front();
			w->tasks.pop_front();
		}
		int code = ARMA_HIP_OK;
		std::string msg;
		if (hipEventSynchronize(w->copied[t.slot]) != hipSuccess) {
			code = ARMA_HIP_EDEVICE;
			msg = "copy of a slab to the host failed";
		} else {
			const unsigned char* p = w->host + size_t(t.slot) * w->slab;
			size_t done = 0;
			while (done < t.len) {
				const ssize_t r = pwrite(w->fd, p + done, t.len - done, t.offset + off_t(done));
				if (r < 0 && errno == EINTR) continue;
				if (r <= 0) {
					code = ARMA_HIP_ERUNTIME;
					msg = std::string("write to the output file failed: ") + std::strerror(errno);
					break;
				}
				done += size_t(r);
			}
		}
		std::lock_guard<std::mutex> lk(w->m);
		if (code == ARMA_HIP_OK) {
			w->written += t.len;
			const auto now = std::chrono::steady_clock::now();
			if (now > w->t_last) w->t_last = now;
		} else if (w->error == ARMA_HIP_OK) {
			w->error = code;
			w->message = msg;
		}
		w->slot_free[t.slot] = true;
		w->cv.notify_all();
	}
}

void stop_threads(arma_hip_writer* w) {
	{
		std::lock_guard<std::mutex> lk(w->m);
		w->closing = true;
	}
	w->cv.notify_all();
	if (w->feeder.joinable()) w->feeder.join();
	{
		// without a feeder (failed open) the I/O threads are released here
		std::lock_guard<std::mutex> lk(w->m);
		w->fed = true;
	}
	w->cv.notify_all();
	for (auto& t : w->io)
		if (t.joinable()) t.join();
}

void release(arma_hip_writer* w) {
	stop_threads(w);
	for (auto& j : w->jobs) (void)hipEventDestroy(j.ready);
	w->jobs.clear();
	if (w->copy) (void)hipStreamSynchronize(w->copy);
	for (int b = 0; b < kSlots; ++b)
		if (w->copied[b]) (void)hipEventDestroy(w->copied[b]);
	if (w->copy) (void)hipStreamDestroy(w->copy);
	if (w->cached) {
		std::lock_guard<std::mutex> lk(w->ctx->writer_mutex);
		w->ctx->writer_cache_busy = false;
	} else {
		if (w->stage) (void)hipFree(w->stage);
		if (w->host) (void)hipHostFree(w->host);
	}
	if (w->fd >= 0) ::close(w->fd);
	delete w;
}

// Slot buffers: the context's cache when it is free (grown if needed), else
// buffers of the writer's own.
void get_buffers(arma_hip_writer* w) {
	const size_t bytes = size_t(kSlots) * w->slab;
	arma_hip_ctx* c = w->ctx;
	{
		std::lock_guard<std::mutex> lk(c->writer_mutex);
		if (!c->writer_cache_busy) {
			if (c->writer_bytes < bytes) {
				if (c->writer_host) ARMA_HIP_CHECK(hipHostFree(c->writer_host));
				c->writer_host = nullptr;
				c->writer_bytes = 0;
				ARMA_HIP_CHECK(hipHostMalloc(&c->writer_host, bytes, hipHostMallocDefault));
				c->writer_bytes = bytes;
			}
			w->host = static_cast<unsigned char*>(c->writer_host);
			w->stage = static_cast<unsigned char*>(c->writer_stage.get(bytes));
			c->writer_cache_busy = true;
			w->cached = true;
			return;
		}
	}
	void* h = nullptr;
	void* d = nullptr;
	ARMA_HIP_CHECK(hipHostMalloc(&h, bytes, hipHostMallocDefault));
	w->host = static_cast<unsigned char*>(h);
	ARMA_HIP_CHECK(hipMalloc(&d, bytes));
	w->stage = static_cast<unsigned char*>(d);
}

}  // namespace
}  // namespace arma_hip

using arma_hip::wguard;

extern "C" {

int arma_hip_writer_open(arma_hip_ctx* ctx, const char* path, int32_t element_size,
                         size_t slab_bytes, arma_hip_writer** out) {
	return wguard([&] {
		ARMA_HIP_REQUIRE(ctx && path && out, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(element_size == 4 || element_size == 8, ARMA_HIP_EINVAL,
		                 "element size must be 4 or 8 bytes");
		if (slab_bytes == 0) slab_bytes = size_t(32) << 20;
		slab_bytes -= slab_bytes % size_t(element_size);
		ARMA_HIP_REQUIRE(slab_bytes >= size_t(element_size), ARMA_HIP_EINVAL, "slab too small");
		*out = nullptr;
		auto* w = new arma_hip_writer;
		try {
			w->ctx = ctx;
			w->device = ctx->device;
			w->esize = element_size;
			w->slab = slab_bytes;
			ARMA_HIP_CHECK(hipSetDevice(ctx->device));
			w->fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
			ARMA_HIP_REQUIRE(w->fd >= 0, ARMA_HIP_ERUNTIME,
			                 std::string("cannot open ") + path + " for writing: " +
			                     std::strerror(errno));
			ARMA_HIP_CHECK(hipStreamCreateWithFlags(&w->copy, hipStreamNonBlocking));
			for (int b = 0; b < arma_hip::kSlots; ++b)
				ARMA_HIP_CHECK(hipEventCreateWithFlags(&w->copied[b], hipEventDisableTiming));
			arma_hip::get_buffers(w);
			w->feeder = std::thread(arma_hip::feeder_main, w);
			for (int t = 0; t < arma_hip::kIoThreads; ++t)
				w->io.emplace_back(arma_hip::io_main, w);
		} catch (...) {
			arma_hip::release(w);
			throw;
		}
		*out = w;
	});
}

int arma_hip_writer_append(arma_hip_writer* w, const void* data, size_t nbytes) {
	return wguard([&] {
		ARMA_HIP_REQUIRE(w && (data || nbytes == 0), ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(nbytes % size_t(w->esize) == 0 &&
		                     reinterpret_cast<uintptr_t>(data) % uintptr_t(w->esize) == 0,
		                 ARMA_HIP_EINVAL, "data is not a whole number of aligned elements");
		{
			std::lock_guard<std::mutex> lk(w->m);
			ARMA_HIP_REQUIRE(w->error == ARMA_HIP_OK, w->error, w->message);
		}
		if (nbytes == 0) return;
		ARMA_HIP_CHECK(hipSetDevice(w->device));
		hipEvent_t ev;
		ARMA_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		const hipError_t e = hipEventRecord(ev, w->ctx->stream);
		if (e != hipSuccess) {
			(void)hipEventDestroy(ev);
			ARMA_HIP_CHECK(e);
		}
		{
			std::lock_guard<std::mutex> lk(w->m);
			w->jobs.push_back(arma_hip::Job{static_cast<const unsigned char*>(data), nbytes, ev});
		}
		w->cv.notify_all();
	});
}

int arma_hip_writer_close(arma_hip_writer* w, uint64_t* bytes_out, double* seconds_out) {
	return wguard([&] {
		ARMA_HIP_REQUIRE(w != nullptr, ARMA_HIP_EINVAL, "null writer");
		(void)hipSetDevice(w->device);
		arma_hip::stop_threads(w);
		const int code = w->error;
		const std::string msg = w->message;
		if (bytes_out) *bytes_out = w->written;
		if (seconds_out)
			*seconds_out = w->started
			                   ? std::chrono::duration<double>(w->t_last - w->t_first).count()
			                   : 0.0;
		arma_hip::release(w);
		ARMA_HIP_REQUIRE(code == ARMA_HIP_OK, code, msg);
	});
}

}  // extern "C"
