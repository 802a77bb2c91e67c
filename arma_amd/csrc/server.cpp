// Serving: a request loop with continuous batching.
//
// Requests (surfaces, optionally with their linear potential, and chunks of
// realtime streams) queue as they arrive.  A scheduler thread admits them in
// arrival order while the device bytes they need fit the HBM budget, and
// hands every admitted group to a free lane as one batch: one persistent
// recurrence launch over all of its surfaces, then the potentials.  Each lane
// owns a context (a HIP stream and its scratch) and a host thread, so the
// batches of different lanes overlap on the device, and requests that arrive
// while lanes are busy are admitted as soon as a lane and memory free up.
//
// A realtime stream keeps its last order_t - 1 slices on the device.  Chunk k
// of a stream is generated as a surface of history + chunk slices whose
// first slices are that history (given values), with noise partitioned as
// (chunk, px, py) and stream seed base + k * columns + column, so the chunks
// of a stream concatenate to exactly the one-shot surface generated with
// partition (chunk, px, py) and seeds base + p.  No reference counterpart:
// the reference generates one surface per run (src/arma_driver.cc:29-87).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace arma_hip {
namespace {

double now_s() {
	using namespace std::chrono;
	return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t align64(size_t n) { return (n + 63) / 64 * 64; }

// Device memory of a batch's surfaces, freed when the last of them is
// released.  Only the lane that allocated it touches it (all its requests
// return to that lane to be freed), so the count needs no lock.
struct Arena {
	void* ptr = nullptr;
	int refs = 0;
};

struct Stream {
	int64_t id = 0;
	int nx = 0, ny = 0, chunk = 0, hist = 0;
	int px = 0, py = 0, columns = 0;
	arma_ar_desc desc{};
	std::vector<double> phi;
	std::vector<arma_mt_params> mts;
	uint32_t seed = 0;
	void* history = nullptr;  // hist x nx x ny values on the device
	int64_t next_chunk = 0;   // index of the next chunk submitted
	bool busy = false;        // a chunk is admitted and not yet done
	int pending = 0;          // chunks submitted and not yet done
};

struct Request {
	int64_t id = 0;
	int kind = ARMA_REQ_SURFACE;
	arma_ar_desc desc{};
	std::vector<double> phi;
	std::vector<arma_mt_params> mts;
	std::vector<uint32_t> seeds;
	double gamma0 = 0.0;
	bool want_potential = false;
	arma_pot_desc pot{};
	Stream* stream = nullptr;
	int64_t chunk = 0;
	int history = 0;
	size_t out_bytes = 0;    // surface block in the batch arena
	size_t pot_bytes = 0;
	size_t work_bytes = 0;   // recurrence and potential working set while in flight
	double t_arrival = 0, t_admit = 0, t_done = 0;
	int status = ARMA_HIP_OK;
	Arena* arena = nullptr;
	size_t zoff = 0;         // element offset of the surface in the arena
	void* phi_out = nullptr;
	int batch_size = 0, lane = -1;
};

struct Lane {
	int index = 0;
	arma_hip_ctx* ctx = nullptr;
	NoiseBatch* noise = nullptr;
	ArBatch* ar = nullptr;
	RescaleBatch* rescale = nullptr;
	std::vector<Request*> job;
	std::vector<Request*> frees;  // released requests whose memory this lane frees
	bool busy = false;
	std::thread th;
	std::condition_variable cv;
};

}  // namespace
}  // namespace arma_hip

struct arma_hip_server {
	int device = 0;
	int precision = ARMA_FP32;
	int max_batch = 64;
	int64_t budget = 0;
	arma_hip::BatchArConfig cfg;
	std::mutex mu;
	std::condition_variable cv_sched, cv_done;
	std::deque<arma_hip::Request*> pending;
	std::deque<arma_hip::Request*> done;
	std::map<int64_t, arma_hip::Request*> finished;  // polled, not yet released
	std::map<int64_t, std::unique_ptr<arma_hip::Stream>> streams;
	std::vector<std::unique_ptr<arma_hip::Lane>> lanes;
	std::thread scheduler;
	bool stop = false;
	// bumped on every event that can make a queued request admissible (a
	// submission, a finished batch, freed memory); the scheduler sleeps until
	// it changes
	uint64_t epoch = 0;
	int64_t next_id = 1, next_stream = 1;
	int64_t in_use = 0, peak = 0;
	int64_t batches = 0, batched_requests = 0, completed = 0;
	std::string lane_error;
};

namespace arma_hip {
namespace {

size_t elem(int precision) { return precision == ARMA_FP32 ? 4 : 8; }

// Potential working set: the half spectrum of every output slice, and the
// column-transformed chunk (at most 8 GiB).
size_t potential_work(const arma_pot_desc& p, size_t es) {
	const size_t nx = p.shape[1], ny = p.shape[2];
	const size_t spec = size_t(p.npoints[0]) * (ny / 2 + 1) * nx * 2 * es;
	const size_t all = size_t(p.npoints[0]) * p.npoints[1] * (ny / 2) * nx * 2 * es;
	return spec + std::min(all, size_t(8) << 30);
}

void run_job(arma_hip_server* s, Lane* L) {
	std::vector<Request*>& job = L->job;
	arma_hip_ctx* ctx = L->ctx;
	const size_t es = elem(s->precision);
	std::vector<BatchMember> mem;
	size_t total = 0;
	for (Request* r : job) {
		r->zoff = total;
		total += align64(numel3(r->desc.shape));
	}
	Arena* arena = new Arena;
	arena->refs = int(job.size());
	for (Request* r : job) r->arena = arena;
	ARMA_HIP_CHECK(hipMallocAsync(&arena->ptr, total * es, ctx->stream));
	for (Request* r : job) {
		mem.push_back(BatchMember{r->zoff, &r->desc, r->gamma0, r->history});
		if (r->history > 0) {
			const size_t hb = size_t(r->history) * r->desc.shape[1] * r->desc.shape[2] * es;
			ARMA_HIP_CHECK(hipMemcpyAsync((char*)arena->ptr + r->zoff * es, r->stream->history, hb,
			                              hipMemcpyDeviceToDevice, ctx->stream));
		}
	}
	noise_batch_build(ctx, L->noise, s->precision, mem);
	ar_batch_build(ctx, L->ar, s->precision, mem, s->cfg);
	rescale_batch_build(ctx, L->rescale, s->precision, mem);
	noise_batch_launch(ctx, L->noise, arena->ptr);
	ar_batch_launch(ctx, L->ar, arena->ptr);
	rescale_batch_launch(ctx, L->rescale, arena->ptr);
	for (Request* r : job) {
		char* z = (char*)arena->ptr + r->zoff * es;
		if (r->want_potential) {
			ARMA_HIP_CHECK(hipMallocAsync(&r->phi_out, r->pot_bytes, ctx->stream));
			launch_linear_potential(ctx, &r->pot, z, r->phi_out, nullptr);
		}
		if (r->stream) {
			// the stream's new history: the last slices of this chunk
			const size_t slice = size_t(r->desc.shape[1]) * r->desc.shape[2] * es;
			const size_t hb = size_t(r->stream->hist) * slice;
			ARMA_HIP_CHECK(hipMemcpyAsync(
				r->stream->history, z + size_t(r->desc.shape[0] - r->stream->hist) * slice, hb,
				hipMemcpyDeviceToDevice, ctx->stream));
		}
	}
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	ar_batch_check(ctx, L->ar);
}

// Frees a request's device memory on the stream of the lane that made it
// (called by that lane's thread, or after every lane thread has ended) and
// returns the bytes it still held in the budget.
int64_t free_request(hipStream_t stream, Request* r) {
	if (r->arena) {
		if (--r->arena->refs == 0) {
			if (r->arena->ptr) (void)hipFreeAsync(r->arena->ptr, stream);
			delete r->arena;
		}
		r->arena = nullptr;
	}
	if (r->phi_out) (void)hipFreeAsync(r->phi_out, stream);
	r->phi_out = nullptr;
	const int64_t held = r->status == ARMA_HIP_OK ? int64_t(r->out_bytes + r->pot_bytes) : 0;
	delete r;
	return held;
}

void lane_main(arma_hip_server* s, Lane* L) {
	(void)hipSetDevice(s->device);
	std::unique_lock<std::mutex> lk(s->mu);
	for (;;) {
		L->cv.wait(lk, [&] { return s->stop || !L->job.empty() || !L->frees.empty(); });
		if (!L->frees.empty()) {
			// HIP calls only outside the lock, on this lane's own stream
			std::vector<Request*> frees;
			frees.swap(L->frees);
			lk.unlock();
			int64_t back = 0;
			for (Request* r : frees) back += free_request(L->ctx->stream, r);
			lk.lock();
			s->in_use -= back;
			++s->epoch;
			s->cv_sched.notify_all();
			continue;
		}
		if (L->job.empty()) return;  // stop with nothing to do
		lk.unlock();
		int status = ARMA_HIP_OK;
		std::string err;
		try {
			run_job(s, L);
		} catch (const Error& e) {
			status = e.code;
			err = e.what();
		} catch (const std::exception& e) {
			status = ARMA_HIP_ERUNTIME;
			err = e.what();
		}
		const double t = now_s();
		lk.lock();
		if (status != ARMA_HIP_OK) s->lane_error = err;
		for (Request* r : L->job) {
			r->t_done = t;
			r->status = status;
			s->in_use -= int64_t(r->work_bytes);  // the working set is free again
			if (status != ARMA_HIP_OK) s->in_use -= int64_t(r->out_bytes + r->pot_bytes);
			if (r->stream) {
				r->stream->busy = false;
				--r->stream->pending;
			}
			s->done.push_back(r);
			++s->completed;
		}
		L->job.clear();
		L->busy = false;
		++s->epoch;
		s->cv_done.notify_all();
		s->cv_sched.notify_all();
	}
}

void scheduler_main(arma_hip_server* s) {
	std::unique_lock<std::mutex> lk(s->mu);
	uint64_t seen = ~uint64_t(0);
	for (;;) {
		// sleep (releasing the lock) until something changed
		s->cv_sched.wait(lk, [&] { return s->stop || s->epoch != seen; });
		if (s->stop) return;
		seen = s->epoch;
		for (auto& L : s->lanes) {
			if (L->busy || s->pending.empty()) continue;
			// admit in arrival order while the budget allows; a chunk whose
			// stream has a chunk in flight waits without blocking the others
			std::vector<Request*> take;
			for (auto it = s->pending.begin();
			     it != s->pending.end() && int(take.size()) < s->max_batch;) {
				Request* r = *it;
				if (r->stream && r->stream->busy) {
					++it;
					continue;
				}
				const int64_t need = int64_t(r->out_bytes + r->pot_bytes + r->work_bytes);
				if (s->in_use + need > s->budget) break;  // keep arrival order
				s->in_use += need;
				s->peak = std::max(s->peak, s->in_use);
				if (r->stream) r->stream->busy = true;
				r->t_admit = now_s();
				take.push_back(r);
				it = s->pending.erase(it);
			}
			if (take.empty()) break;
			for (Request* r : take) {
				r->batch_size = int(take.size());
				r->lane = L->index;
			}
			++s->batches;
			s->batched_requests += int64_t(take.size());
			L->job = std::move(take);
			L->busy = true;
			L->cv.notify_one();
		}
	}
}

void copy_desc(const arma_ar_desc& src, arma_ar_desc& d, std::vector<double>& phi,
               std::vector<arma_mt_params>& mts, std::vector<uint32_t>& seeds) {
	d = src;
	phi.assign(src.phi, src.phi + numel3(src.phi_shape));
	d.phi = phi.data();
	mts.assign(src.mts, src.mts + src.n_mts);
	d.mts = mts.data();
	if (src.seeds) {
		seeds.assign(src.seeds, src.seeds + src.n_mts);
		d.seeds = seeds.data();
	}
}

}  // namespace

}  // namespace arma_hip

using arma_hip::Error;

namespace {

template <class F>
int sguard(F&& f) {
	try {
		f();
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

void check_desc(const arma_ar_desc& d, int precision) {
	for (int k = 0; k < 3; ++k)
		ARMA_HIP_REQUIRE(d.shape[k] > 0 && d.phi_shape[k] > 0, ARMA_HIP_EINVAL, "bad shape");
	ARMA_HIP_REQUIRE(d.phi && d.mts && d.n_mts > 0, ARMA_HIP_EINVAL, "null generator argument");
	ARMA_HIP_REQUIRE(d.precision == precision, ARMA_HIP_EINVAL,
	                 "request precision differs from the server's");
}

}  // namespace

extern "C" {

double arma_hip_server_clock(void) { return arma_hip::now_s(); }

int arma_hip_server_create(int device, const arma_server_config* cfg, arma_hip_server** out) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(cfg && out, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(cfg->lanes > 0 && cfg->lanes <= 16, ARMA_HIP_EINVAL,
		                 "lanes must be 1..16");
		ARMA_HIP_REQUIRE(cfg->hbm_budget > 0, ARMA_HIP_EINVAL, "budget must be positive");
		ARMA_HIP_REQUIRE(cfg->precision == ARMA_FP32 || cfg->precision == ARMA_FP64,
		                 ARMA_HIP_EINVAL, "bad precision");
		std::unique_ptr<arma_hip_server> s(new arma_hip_server);
		s->device = device;
		s->precision = cfg->precision;
		s->max_batch = cfg->max_batch > 0 ? cfg->max_batch : 64;
		s->budget = cfg->hbm_budget;
		s->cfg.tile_x = cfg->batch.tile_x ? cfg->batch.tile_x : 16;
		s->cfg.schedule = cfg->batch.schedule ? cfg->batch.schedule : ARMA_BATCH_SURFACE_MAJOR;
		s->cfg.slots_per_cu = cfg->batch.slots_per_cu;
		for (int i = 0; i < cfg->lanes; ++i) {
			auto L = std::make_unique<arma_hip::Lane>();
			L->index = i;
			const int rc = arma_hip_init(device, &L->ctx);
			if (rc != ARMA_HIP_OK) throw Error(rc, arma_hip_last_error());
			L->noise = arma_hip::noise_batch_new();
			L->ar = arma_hip::ar_batch_new();
			L->rescale = arma_hip::rescale_batch_new();
			s->lanes.push_back(std::move(L));
		}
		// keep freed batch memory in the device's pool instead of returning
		// it to the driver at every synchronisation
		ARMA_HIP_CHECK(hipSetDevice(device));
		hipMemPool_t pool = nullptr;
		if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
			uint64_t keep = uint64_t(cfg->hbm_budget);
			(void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
		}
		arma_hip_server* sp = s.get();
		for (auto& L : s->lanes) L->th = std::thread(arma_hip::lane_main, sp, L.get());
		s->scheduler = std::thread(arma_hip::scheduler_main, sp);
		*out = s.release();
	});
}

int arma_hip_server_open_stream(arma_hip_server* s, const arma_ar_desc* desc, int32_t chunk,
                                int64_t* stream) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s && desc && stream, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(chunk > 0, ARMA_HIP_EINVAL, "bad chunk length");
		check_desc(*desc, s->precision);
		auto st = std::make_unique<arma_hip::Stream>();
		st->nx = desc->shape[1];
		st->ny = desc->shape[2];
		st->chunk = chunk;
		st->hist = desc->phi_shape[0] - 1;
		ARMA_HIP_REQUIRE(st->hist <= chunk, ARMA_HIP_EINVAL, "chunk shorter than the AR order in t");
		st->px = desc->noise_mode == ARMA_NOISE_PARTITIONED ? desc->partition[1] : st->nx;
		st->py = desc->noise_mode == ARMA_NOISE_PARTITIONED ? desc->partition[2] : st->ny;
		ARMA_HIP_REQUIRE(st->px > 0 && st->py > 0, ARMA_HIP_EINVAL, "bad partition shape");
		st->columns = arma_hip::div_ceil(st->nx, st->px) * arma_hip::div_ceil(st->ny, st->py);
		ARMA_HIP_REQUIRE(desc->n_mts >= st->columns, ARMA_HIP_EPRNG, "bad number of MT configs");
		arma_ar_desc d = *desc;
		d.seeds = nullptr;  // only seeds[0], the stream's base seed, is read
		d.n_mts = st->columns;
		std::vector<uint32_t> none;
		arma_hip::copy_desc(d, st->desc, st->phi, st->mts, none);
		st->seed = desc->seeds ? desc->seeds[0] : 0u;
		const size_t hb = std::max<size_t>(1, size_t(st->hist) * st->nx * st->ny) *
		                  arma_hip::elem(s->precision);
		ARMA_HIP_CHECK(hipSetDevice(s->device));
		ARMA_HIP_CHECK(hipMalloc(&st->history, hb));
		ARMA_HIP_CHECK(hipMemset(st->history, 0, hb));
		std::lock_guard<std::mutex> g(s->mu);
		st->id = s->next_stream++;
		*stream = st->id;
		s->in_use += int64_t(hb);
		s->peak = std::max(s->peak, s->in_use);
		s->streams[st->id] = std::move(st);
	});
}

int arma_hip_server_close_stream(arma_hip_server* s, int64_t stream) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s, ARMA_HIP_EINVAL, "null server");
		std::unique_ptr<arma_hip::Stream> st;
		{
			std::lock_guard<std::mutex> g(s->mu);
			auto it = s->streams.find(stream);
			ARMA_HIP_REQUIRE(it != s->streams.end(), ARMA_HIP_EINVAL, "no such stream");
			ARMA_HIP_REQUIRE(it->second->pending == 0, ARMA_HIP_EINVAL,
			                 "stream has chunks in flight");
			st = std::move(it->second);
			s->streams.erase(it);
			s->in_use -= int64_t(std::max<size_t>(1, size_t(st->hist) * st->nx * st->ny) *
			                     arma_hip::elem(s->precision));
		}
		(void)hipSetDevice(s->device);
		(void)hipFree(st->history);
	});
}

int arma_hip_server_submit(arma_hip_server* s, const arma_request* q, int64_t* id) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s && q && id, ARMA_HIP_EINVAL, "null argument");
		auto r = std::make_unique<arma_hip::Request>();
		const size_t es = arma_hip::elem(s->precision);
		r->kind = q->kind;
		r->t_arrival = q->arrival > 0 ? q->arrival : arma_hip::now_s();
		if (q->kind == ARMA_REQ_SURFACE) {
			check_desc(q->surface, s->precision);
			arma_hip::copy_desc(q->surface, r->desc, r->phi, r->mts, r->seeds);
			r->gamma0 = q->gamma0;
			r->want_potential = q->want_potential != 0;
			if (r->want_potential) {
				r->pot = q->potential;
				for (int k = 0; k < 3; ++k)
					ARMA_HIP_REQUIRE(r->pot.shape[k] == r->desc.shape[k], ARMA_HIP_EINVAL,
					                 "potential shape differs from the surface");
				ARMA_HIP_REQUIRE(r->pot.npoints[0] > 0 && r->pot.npoints[1] > 0,
				                 ARMA_HIP_EINVAL, "bad domain");
				ARMA_HIP_REQUIRE(!(r->pot.lbound[1] < -r->pot.depth), ARMA_HIP_EINVAL,
				                 "min(z) is less than -h");
				auto pow2 = [](int n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; };
				ARMA_HIP_REQUIRE(pow2(r->pot.shape[1]) && pow2(r->pot.shape[2]), ARMA_HIP_EINVAL,
				                 "surface extents along x and y must be powers of two in [4, 1024]");
				r->pot.precision = s->precision;
				r->pot_bytes = size_t(r->pot.npoints[0]) * r->pot.npoints[1] * r->pot.shape[1] *
				               r->pot.shape[2] * es;
			}
		} else {
			ARMA_HIP_REQUIRE(q->kind == ARMA_REQ_STREAM_CHUNK, ARMA_HIP_EINVAL,
			                 "bad request kind");
		}
		std::lock_guard<std::mutex> g(s->mu);
		ARMA_HIP_REQUIRE(!s->stop, ARMA_HIP_EINVAL, "server is stopping");
		if (q->kind == ARMA_REQ_STREAM_CHUNK) {
			auto it = s->streams.find(q->stream);
			ARMA_HIP_REQUIRE(it != s->streams.end(), ARMA_HIP_EINVAL, "no such stream");
			arma_hip::Stream* st = it->second.get();
			r->stream = st;
			r->chunk = st->next_chunk++;
			r->history = st->hist;
			r->desc = st->desc;
			r->desc.shape[0] = st->hist + st->chunk;
			r->desc.noise_mode = ARMA_NOISE_PARTITIONED;
			r->desc.partition[0] = st->chunk;
			r->desc.partition[1] = st->px;
			r->desc.partition[2] = st->py;
			r->desc.n_mts = st->columns;
			r->desc.mts = st->mts.data();
			r->seeds.resize(st->columns);
			for (int c = 0; c < st->columns; ++c)
				r->seeds[c] = st->seed + uint32_t(r->chunk * st->columns + c);
			r->desc.seeds = r->seeds.data();
			++st->pending;
		}
		r->out_bytes = arma_hip::align64(arma_hip::numel3(r->desc.shape)) * es;
		r->work_bytes = arma_hip::ar_batch_member_bytes(&r->desc, s->cfg.tile_x);
		if (r->want_potential) r->work_bytes += arma_hip::potential_work(r->pot, es);
		const int64_t need = int64_t(r->out_bytes + r->pot_bytes + r->work_bytes);
		if (need > s->budget) {
			if (r->stream) --r->stream->pending;
			throw Error(ARMA_HIP_ENOMEM, "request larger than the HBM budget");
		}
		r->id = s->next_id++;
		*id = r->id;
		s->pending.push_back(r.release());
		++s->epoch;
		s->cv_sched.notify_all();
	});
}

int arma_hip_server_poll(arma_hip_server* s, arma_completion* out, int32_t max, int32_t* n,
                         int32_t wait_ms) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s && out && n && max > 0, ARMA_HIP_EINVAL, "bad argument");
		std::unique_lock<std::mutex> lk(s->mu);
		if (s->done.empty() && wait_ms != 0) {
			auto pred = [&] { return !s->done.empty(); };
			if (wait_ms < 0)
				s->cv_done.wait(lk, pred);
			else
				s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
		}
		int k = 0;
		const size_t es = arma_hip::elem(s->precision);
		while (k < max && !s->done.empty()) {
			arma_hip::Request* r = s->done.front();
			s->done.pop_front();
			arma_completion& c = out[k++];
			std::memset(&c, 0, sizeof c);
			c.id = r->id;
			c.status = r->status;
			c.kind = r->kind;
			c.stream = r->stream ? r->stream->id : 0;
			c.chunk = r->chunk;
			if (r->status == ARMA_HIP_OK) {
				const size_t slice = size_t(r->desc.shape[1]) * r->desc.shape[2];
				c.zeta = (const char*)r->arena->ptr + (r->zoff + size_t(r->history) * slice) * es;
				c.phi = r->phi_out;
			}
			c.shape[0] = r->desc.shape[0] - r->history;
			c.shape[1] = r->desc.shape[1];
			c.shape[2] = r->desc.shape[2];
			c.t_arrival = r->t_arrival;
			c.t_admit = r->t_admit;
			c.t_done = r->t_done;
			c.batch_size = r->batch_size;
			c.lane = r->lane;
			s->finished[r->id] = r;
		}
		*n = k;
	});
}

int arma_hip_server_release(arma_hip_server* s, int64_t id) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s, ARMA_HIP_EINVAL, "null server");
		std::lock_guard<std::mutex> g(s->mu);
		auto it = s->finished.find(id);
		ARMA_HIP_REQUIRE(it != s->finished.end(), ARMA_HIP_EINVAL, "no such completed request");
		arma_hip::Request* r = it->second;
		s->finished.erase(it);
		// the lane that made it frees its memory (on its own thread and stream)
		arma_hip::Lane* L = s->lanes[std::max(r->lane, 0)].get();
		L->frees.push_back(r);
		L->cv.notify_one();
	});
}

int arma_hip_server_stats(arma_hip_server* s, arma_server_stats* st) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s && st, ARMA_HIP_EINVAL, "null argument");
		std::lock_guard<std::mutex> g(s->mu);
		std::memset(st, 0, sizeof *st);
		st->in_use = s->in_use;
		st->peak = s->peak;
		st->budget = s->budget;
		st->batches = s->batches;
		st->batched_requests = s->batched_requests;
		st->completed = s->completed;
		st->pending = int64_t(s->pending.size());
		int busy = 0;
		for (auto& L : s->lanes) busy += L->busy ? 1 : 0;
		st->busy_lanes = busy;
	});
}

int arma_hip_server_copy(arma_hip_server* s, void* dst, const void* src, size_t bytes) {
	return sguard([&] {
		ARMA_HIP_REQUIRE(s && dst && src, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_CHECK(hipSetDevice(s->device));
		ARMA_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
	});
}

void arma_hip_server_destroy(arma_hip_server* s) {
	if (!s) return;
	{
		// let admitted work finish; what is still queued is dropped
		std::unique_lock<std::mutex> lk(s->mu);
		s->cv_done.wait(lk, [&] {
			for (auto& L : s->lanes)
				if (L->busy) return false;
			return true;
		});
		s->stop = true;
		s->cv_sched.notify_all();
		for (auto& L : s->lanes) L->cv.notify_all();
	}
	if (s->scheduler.joinable()) s->scheduler.join();
	for (auto& L : s->lanes)
		if (L->th.joinable()) L->th.join();
	(void)hipSetDevice(s->device);
	for (arma_hip::Request* r : s->pending) {
		if (r->stream) --r->stream->pending;
		delete r;
	}
	auto lane_stream = [&](arma_hip::Request* r) { return s->lanes[std::max(r->lane, 0)]->ctx->stream; };
	for (arma_hip::Request* r : s->done) arma_hip::free_request(lane_stream(r), r);
	for (auto& kv : s->finished) arma_hip::free_request(lane_stream(kv.second), kv.second);
	for (auto& L : s->lanes)
		for (arma_hip::Request* r : L->frees) arma_hip::free_request(L->ctx->stream, r);
	for (auto& L : s->lanes) (void)hipStreamSynchronize(L->ctx->stream);
	for (auto& kv : s->streams) (void)hipFree(kv.second->history);
	for (auto& L : s->lanes) {
		arma_hip::noise_batch_destroy(L->noise);
		arma_hip::ar_batch_destroy(L->ar);
		arma_hip::rescale_batch_destroy(L->rescale);
		arma_hip_destroy(L->ctx);
	}
	delete s;
}

}  // extern "C"
