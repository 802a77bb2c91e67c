// C ABI of libarma_hip: context management, argument validation, host/device
// staging, mt.dat input/output.  Every entry point converts exceptions into
// the ARMA_HIP_* codes and keeps the message for arma_hip_last_error().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
	try {
		f();
		return ARMA_HIP_OK;
	} catch (const arma_hip::Error& e) {
		g_last_error = e.what();
		return e.code;
	} catch (const std::bad_alloc&) {
		g_last_error = "out of host memory";
		return ARMA_HIP_ENOMEM;
	} catch (const std::exception& e) {
		g_last_error = e.what();
		return ARMA_HIP_ERUNTIME;
	}
}

void use_device(arma_hip_ctx* ctx) {
	ARMA_HIP_REQUIRE(ctx != nullptr, ARMA_HIP_EINVAL, "null context");
	ARMA_HIP_CHECK(hipSetDevice(ctx->device));
}

size_t elem_size(int precision) { return precision == ARMA_FP32 ? 4 : 8; }

void check_ar_desc(const arma_ar_desc* d) {
	ARMA_HIP_REQUIRE(d != nullptr, ARMA_HIP_EINVAL, "null descriptor");
	for (int k = 0; k < 3; ++k) {
		ARMA_HIP_REQUIRE(d->shape[k] > 0, ARMA_HIP_EINVAL, "bad output grid shape");
		ARMA_HIP_REQUIRE(d->phi_shape[k] > 0 && d->phi_shape[k] <= 64, ARMA_HIP_EINVAL,
		                 "bad AR order");
	}
	ARMA_HIP_REQUIRE(d->phi != nullptr, ARMA_HIP_EINVAL, "null AR coefficients");
	ARMA_HIP_REQUIRE(d->precision == ARMA_FP64 || d->precision == ARMA_FP32, ARMA_HIP_EINVAL,
	                 "bad precision");
	ARMA_HIP_REQUIRE(arma_hip::numel3(d->shape) < (size_t(1) << 40), ARMA_HIP_EINVAL,
	                 "output grid too large");
	for (size_t i = 0; i < arma_hip::numel3(d->phi_shape); ++i)
		ARMA_HIP_REQUIRE(std::isfinite(d->phi[i]), ARMA_HIP_EINVAL, "non-finite AR coefficient");
}

void recurrence(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta) {
	if (d->algorithm == ARMA_AR_PLAIN) {
		arma_hip::launch_ar_plain(ctx, d, zeta);
	} else if (d->algorithm == ARMA_AR_WAVEFRONT) {
		arma_hip::launch_ar_wavefront(ctx, d, zeta);
	} else {
		throw arma_hip::Error(ARMA_HIP_EINVAL, "bad AR algorithm");
	}
}

void check_config(const arma_batch_config& c) {
	ARMA_HIP_REQUIRE(c.tile_x == 0 || c.tile_x == 8 || c.tile_x == 16 || c.tile_x == 32,
	                 ARMA_HIP_EINVAL, "batch tile width must be 8, 16 or 32");
	ARMA_HIP_REQUIRE(c.schedule >= ARMA_BATCH_AUTO && c.schedule <= ARMA_BATCH_INTERLEAVED,
	                 ARMA_HIP_EINVAL, "bad batch schedule");
	ARMA_HIP_REQUIRE(c.slots_per_cu >= 0, ARMA_HIP_EINVAL, "bad batch residency");
}

arma_hip::BatchArConfig to_internal(const arma_batch_config& c) {
	arma_hip::BatchArConfig r;
	r.tile_x = c.tile_x ? c.tile_x : 16;
	r.schedule = c.schedule ? c.schedule : ARMA_BATCH_SURFACE_MAJOR;
	r.slots_per_cu = c.slots_per_cu;
	return r;
}

// mt.dat record: the raw mt_struct (15 four-byte fields, 4 bytes of padding
// and the 8-byte state pointer on LP64) followed by nn state words.
constexpr size_t kRecordHeader = 72;

}  // namespace

void arma_hip::set_last_error(const char* msg) { g_last_error = msg; }

// A batch of surfaces: host copies of every descriptor, the arena layout,
// and the device tables of its noise, recurrence and rescale launches.
struct arma_hip_batch {
	int precision = 0;
	std::vector<arma_ar_desc> descs;
	std::vector<std::vector<double>> phi;
	std::vector<std::vector<arma_mt_params>> mts;
	std::vector<std::vector<uint32_t>> seeds;
	std::vector<arma_hip::BatchMember> members;
	std::vector<int64_t> offsets;
	std::string signature;
	arma_hip::BatchArConfig cfg;
	arma_hip::NoiseBatch* noise = nullptr;
	arma_hip::ArBatch* ar = nullptr;
	arma_hip::RescaleBatch* rescale = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	void drop_graph() {
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
	}
	~arma_hip_batch() {
		drop_graph();
		if (noise) arma_hip::noise_batch_destroy(noise);
		if (ar) arma_hip::ar_batch_destroy(ar);
		if (rescale) arma_hip::rescale_batch_destroy(rescale);
	}
};

namespace {

void batch_enqueue(arma_hip_ctx* ctx, arma_hip_batch* b, void* arena) {
	arma_hip::noise_batch_launch(ctx, b->noise, arena);
	arma_hip::ar_batch_launch(ctx, b->ar, arena);
	arma_hip::rescale_batch_launch(ctx, b->rescale, arena);
}

}  // namespace

extern "C" {

const char* arma_hip_last_error(void) { return g_last_error.c_str(); }

int arma_hip_init(int device, arma_hip_ctx** out) {
	return guarded([&] {
		ARMA_HIP_REQUIRE(out != nullptr, ARMA_HIP_EINVAL, "null output pointer");
		int n = 0;
		ARMA_HIP_CHECK(hipGetDeviceCount(&n));
		ARMA_HIP_REQUIRE(device >= 0 && device < n, ARMA_HIP_EDEVICE, "no such HIP device");
		ARMA_HIP_CHECK(hipSetDevice(device));
		arma_hip::noalloc_setup(device);
		auto* ctx = new arma_hip_ctx;
		ctx->device = device;
		hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
		if (e != hipSuccess) {
			delete ctx;
			throw arma_hip::Error(ARMA_HIP_EDEVICE, hipGetErrorString(e));
		}
		ctx->own_stream = true;
		(void)hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device);
		*out = ctx;
	});
}

void arma_hip_destroy(arma_hip_ctx* ctx) {
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	for (arma_hip::Scratch* s : {&ctx->yw, &ctx->params, &ctx->seeds, &ctx->state, &ctx->phi,
	                             &ctx->partials, &ctx->flags, &ctx->noise_tmp, &ctx->mbox, &ctx->pot_spec,
	                             &ctx->pot_mid, &ctx->pot_wn, &ctx->pot_twiddle, &ctx->pot_kmin, &ctx->pot_field,
	                             &ctx->pot_edge,
	                             &ctx->zeta_tmp, &ctx->phi_tmp, &ctx->writer_stage})
		s->release();
	if (ctx->writer_host) (void)hipHostFree(ctx->writer_host);
	if (ctx->aux_stream) {
		(void)hipStreamSynchronize(ctx->aux_stream);
		(void)hipStreamDestroy(ctx->aux_stream);
	}
	for (hipEvent_t e : ctx->pot_events)
		if (e) (void)hipEventDestroy(e);
	if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

int arma_hip_set_stream(arma_hip_ctx* ctx, void* stream) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		if (ctx->own_stream) ARMA_HIP_CHECK(hipStreamDestroy(ctx->stream));
		ctx->stream = (hipStream_t)stream;
		ctx->own_stream = false;
	});
}

void* arma_hip_get_stream(arma_hip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int arma_hip_synchronize(arma_hip_ctx* ctx) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	});
}

int arma_hip_white_noise(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* eps) {
	return guarded([&] {
		use_device(ctx);
		check_ar_desc(desc);
		ARMA_HIP_REQUIRE(eps != nullptr, ARMA_HIP_EINVAL, "null output buffer");
		arma_hip::launch_white_noise(ctx, desc, eps);
	});
}

int arma_hip_ar_recurrence(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* zeta) {
	return guarded([&] {
		use_device(ctx);
		check_ar_desc(desc);
		ARMA_HIP_REQUIRE(zeta != nullptr, ARMA_HIP_EINVAL, "null surface buffer");
		recurrence(ctx, desc, zeta);
	});
}

int arma_hip_ar_generate(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* zeta,
                         int zeta_on_device) {
	return guarded([&] {
		use_device(ctx);
		check_ar_desc(desc);
		ARMA_HIP_REQUIRE(zeta != nullptr, ARMA_HIP_EINVAL, "null surface buffer");
		const size_t bytes = arma_hip::numel3(desc->shape) * elem_size(desc->precision);
		void* dz = zeta;
		if (!zeta_on_device) dz = ctx->zeta_tmp.get(bytes);
		arma_hip::launch_white_noise(ctx, desc, dz);
		recurrence(ctx, desc, dz);
		if (!zeta_on_device) {
			ARMA_HIP_CHECK(hipMemcpyAsync(zeta, dz, bytes, hipMemcpyDeviceToHost, ctx->stream));
			ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		}
	});
}

int arma_hip_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int32_t precision,
                     double gamma0, double* factor_out) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(zeta && shape, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(shape[0] > 0 && shape[1] > 0 && shape[2] > 0, ARMA_HIP_EINVAL,
		                 "bad shape");
		ARMA_HIP_REQUIRE(precision == ARMA_FP64 || precision == ARMA_FP32, ARMA_HIP_EINVAL,
		                 "bad precision");
		const double f = arma_hip::launch_rescale(ctx, zeta, shape, precision, gamma0);
		if (factor_out) *factor_out = f;
	});
}

namespace {

// Velocity_potential_solver<T>::operator() (src/velocity/basic_solver.cc:21-77)
// for solver 0 (Linear_solver) or 1 (High_amplitude_solver).
int potential_entry(arma_hip_ctx* ctx, const arma_pot_desc* desc, const void* zeta, void* phi,
                    int on_device, double* wnmax_out, int solver) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(desc && zeta && phi, ARMA_HIP_EINVAL, "null argument");
		for (int k = 0; k < 3; ++k)
			ARMA_HIP_REQUIRE(desc->shape[k] > 0, ARMA_HIP_EINVAL, "bad surface shape");
		ARMA_HIP_REQUIRE(desc->npoints[0] > 0 && desc->npoints[1] > 0, ARMA_HIP_EINVAL,
		                 "bad domain");
		ARMA_HIP_REQUIRE(desc->precision == ARMA_FP64 || desc->precision == ARMA_FP32,
		                 ARMA_HIP_EINVAL, "bad precision");
		// src/velocity/basic_solver.cc:166-169
		ARMA_HIP_REQUIRE(!(desc->lbound[1] < -desc->depth), ARMA_HIP_EINVAL,
		                 "min(z) is less than -h");
		const size_t es = elem_size(desc->precision);
		const size_t zbytes = arma_hip::numel3(desc->shape) * es;
		const size_t pbytes = size_t(desc->npoints[0]) * desc->npoints[1] * desc->shape[1] *
		                      desc->shape[2] * es;
		const void* dz = zeta;
		void* dp = phi;
		if (!on_device) {
			void* tz = ctx->zeta_tmp.get(zbytes);
			ARMA_HIP_CHECK(hipMemcpyAsync(tz, zeta, zbytes, hipMemcpyHostToDevice, ctx->stream));
			dz = tz;
			dp = ctx->phi_tmp.get(pbytes);
		}
		arma_hip::launch_potential(ctx, desc, dz, dp, wnmax_out, solver);
		if (!on_device) {
			ARMA_HIP_CHECK(hipMemcpyAsync(phi, dp, pbytes, hipMemcpyDeviceToHost, ctx->stream));
			ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		}
	});
}

}  // namespace

int arma_hip_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* desc, const void* zeta,
                              void* phi, int on_device, double* wnmax_out) {
	return potential_entry(ctx, desc, zeta, phi, on_device, wnmax_out, 0);
}

int arma_hip_high_amplitude_potential(arma_hip_ctx* ctx, const arma_pot_desc* desc,
                                      const void* zeta, void* phi, int on_device,
                                      double* wnmax_out) {
	return potential_entry(ctx, desc, zeta, phi, on_device, wnmax_out, 1);
}

int arma_hip_nit_realisation(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3],
                             int32_t precision, double gamma0, const arma_nit_desc* desc) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(zeta && shape && desc, ARMA_HIP_EINVAL, "null argument");
		for (int k = 0; k < 3; ++k)
			ARMA_HIP_REQUIRE(shape[k] > 0, ARMA_HIP_EINVAL, "bad surface shape");
		ARMA_HIP_REQUIRE(precision == ARMA_FP64 || precision == ARMA_FP32, ARMA_HIP_EINVAL,
		                 "bad precision");
		ARMA_HIP_REQUIRE(desc->distribution == ARMA_NIT_GRAM_CHARLIER ||
		                     desc->distribution == ARMA_NIT_SKEW_NORMAL,
		                 ARMA_HIP_EINVAL, "bad distribution");
		// src/linalg.cc:246-254
		ARMA_HIP_REQUIRE(desc->interval[0] < desc->interval[1], ARMA_HIP_EINVAL,
		                 "bad parameter: bisection.interval");
		ARMA_HIP_REQUIRE(desc->absolute_error > 0.0, ARMA_HIP_EINVAL,
		                 "bad parameter: bisection.absolute_error");
		ARMA_HIP_REQUIRE(desc->max_iterations > 0, ARMA_HIP_EINVAL,
		                 "bad parameter: bisection.max_iterations");
		ARMA_HIP_REQUIRE(gamma0 >= 0.0, ARMA_HIP_EINVAL, "negative variance");
		arma_hip::launch_nit_realisation(ctx, zeta, arma_hip::numel3(shape), precision, gamma0,
		                                 desc);
	});
}

int arma_hip_profile(arma_hip_ctx* ctx, int32_t enable) {
	return guarded([&] {
		use_device(ctx);
		ctx->profiling = enable != 0;
	});
}

int arma_hip_profile_report(arma_hip_ctx* ctx, char* buf, size_t len, int32_t reset) {
	return guarded([&] {
		use_device(ctx);
		for (auto& r : ctx->pending) {
			ARMA_HIP_CHECK(hipEventSynchronize(r.stop));
			float ms = 0.0f;
			ARMA_HIP_CHECK(hipEventElapsedTime(&ms, r.start, r.stop));
			auto& tot = ctx->totals[r.name];
			tot.ms += ms;
			tot.count += 1;
			(void)hipEventDestroy(r.start);
			(void)hipEventDestroy(r.stop);
		}
		ctx->pending.clear();
		std::string out;
		for (const auto& kv : ctx->totals) {
			char line[256];
			std::snprintf(line, sizeof line, "prfl %s %.3fus %lld\n", kv.first.c_str(),
			              kv.second.ms * 1000.0, kv.second.count);
			out += line;
		}
		if (buf && len) {
			std::strncpy(buf, out.c_str(), len - 1);
			buf[len - 1] = 0;
		}
		if (reset) ctx->totals.clear();
	});
}

int arma_hip_mt_raw(arma_hip_ctx* ctx, const arma_mt_params* mts, const uint32_t* seeds,
                    int32_t n_streams, int32_t draws_per_stream, uint32_t* out) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(out != nullptr, ARMA_HIP_EINVAL, "null output");
		arma_hip::launch_mt_raw(ctx, mts, seeds, n_streams, draws_per_stream, out);
	});
}

int arma_hip_read_mt_config(const char* path, int32_t n, arma_mt_params* out, int32_t* n_read) {
	return guarded([&] {
		ARMA_HIP_REQUIRE(path && out && n >= 0, ARMA_HIP_EINVAL, "null argument");
		std::ifstream in(path, std::ios::binary);
		if (!in.is_open()) {
			throw arma_hip::Error(ARMA_HIP_ERUNTIME,
			                      std::string("Mersenne Twister configuration file not found: ") + path);
		}
		int32_t k = 0;
		unsigned char hdr[kRecordHeader];
		while (k < n && in.read((char*)hdr, kRecordHeader)) {
			uint32_t f[15];
			std::memcpy(f, hdr, sizeof f);
			arma_mt_params q;
			q.aaa = f[0];
			q.mm = int32_t(f[1]);
			q.nn = int32_t(f[2]);
			q.rr = int32_t(f[3]);
			q.ww = int32_t(f[4]);
			q.wmask = f[5];
			q.umask = f[6];
			q.lmask = f[7];
			q.shift0 = int32_t(f[8]);
			q.shift1 = int32_t(f[9]);
			q.shiftB = int32_t(f[10]);
			q.shiftC = int32_t(f[11]);
			q.maskB = f[12];
			q.maskC = f[13];
			ARMA_HIP_REQUIRE(q.nn > 0 && q.nn <= 1392, ARMA_HIP_ERUNTIME, "corrupt mt.dat record");
			in.ignore(std::streamsize(q.nn) * 4);
			if (!in) break;
			out[k++] = q;
		}
		if (n_read) *n_read = k;
		if (k < n) throw arma_hip::Error(ARMA_HIP_EPRNG, "bad number of MT configs");
	});
}

int arma_hip_write_mt_config(const char* path, const arma_mt_params* params, int32_t n) {
	return guarded([&] {
		ARMA_HIP_REQUIRE(path && params && n >= 0, ARMA_HIP_EINVAL, "null argument");
		std::ofstream out(path, std::ios::binary);
		ARMA_HIP_REQUIRE(out.is_open(), ARMA_HIP_ERUNTIME, "cannot open mt.dat for writing");
		for (int32_t k = 0; k < n; ++k) {
			const arma_mt_params& q = params[k];
			unsigned char hdr[kRecordHeader];
			std::memset(hdr, 0, sizeof hdr);
			const uint32_t f[15] = {q.aaa, uint32_t(q.mm), uint32_t(q.nn), uint32_t(q.rr),
			                        uint32_t(q.ww), q.wmask, q.umask, q.lmask,
			                        uint32_t(q.shift0), uint32_t(q.shift1), uint32_t(q.shiftB),
			                        uint32_t(q.shiftC), q.maskB, q.maskC, uint32_t(q.nn)};
			std::memcpy(hdr, f, sizeof f);
			out.write((const char*)hdr, kRecordHeader);
			// state words as sgenrand_mt(0) leaves them; readers re-seed anyway
			uint32_t s = 0;
			for (int i = 0; i < q.nn; ++i) {
				const uint32_t w = s & q.wmask;
				out.write((const char*)&w, 4);
				s = 1812433253u * (s ^ (s >> 30)) + uint32_t(i + 1);
			}
		}
		ARMA_HIP_REQUIRE(bool(out), ARMA_HIP_ERUNTIME, "short write to mt.dat");
	});
}

int arma_hip_batch_create(arma_hip_ctx* ctx, const arma_batch_desc* desc, arma_hip_batch** out) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(desc && out, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(desc->n_surfaces > 0 && desc->surfaces, ARMA_HIP_EINVAL,
		                 "empty batch");
		ARMA_HIP_REQUIRE(desc->precision == ARMA_FP64 || desc->precision == ARMA_FP32,
		                 ARMA_HIP_EINVAL, "bad precision");
		check_config(desc->config);
		std::unique_ptr<arma_hip_batch> b(new arma_hip_batch);
		const int n = desc->n_surfaces;
		b->precision = desc->precision;
		b->descs.assign(desc->surfaces, desc->surfaces + n);
		b->phi.resize(n);
		b->mts.resize(n);
		b->seeds.resize(n);
		b->offsets.resize(n + 1);
		int64_t off = 0;
		char sig[64];
		std::snprintf(sig, sizeof sig, "p%d n%d", desc->precision, n);
		b->signature = sig;
		for (int i = 0; i < n; ++i) {
			arma_ar_desc& d = b->descs[i];
			d.precision = desc->precision;
			check_ar_desc(&d);
			b->phi[i].assign(d.phi, d.phi + arma_hip::numel3(d.phi_shape));
			d.phi = b->phi[i].data();
			if (d.mts && d.n_mts > 0) {
				b->mts[i].assign(d.mts, d.mts + d.n_mts);
				d.mts = b->mts[i].data();
				if (d.seeds) {
					b->seeds[i].assign(d.seeds, d.seeds + d.n_mts);
					d.seeds = b->seeds[i].data();
				}
			}
			const double g0 = desc->gamma0 ? desc->gamma0[i] : 0.0;
			const int h = desc->history ? desc->history[i] : 0;
			ARMA_HIP_REQUIRE(h >= 0 && h < d.shape[0], ARMA_HIP_EINVAL, "bad history length");
			b->members.push_back(arma_hip::BatchMember{size_t(off), &b->descs[i], g0, h});
			b->offsets[i] = off;
			// every surface starts on a 64-element boundary
			off += int64_t((arma_hip::numel3(d.shape) + 63) / 64 * 64);
			std::snprintf(sig, sizeof sig, " %d,%d,%d:%d,%d,%d", d.shape[0], d.shape[1],
			              d.shape[2], d.phi_shape[0], d.phi_shape[1], d.phi_shape[2]);
			b->signature += sig;
		}
		b->offsets[n] = off;
		const arma_batch_config& c = desc->config;
		if (!c.tile_x && !c.schedule && !c.slots_per_cu &&
		    ctx->batch_tuned.count(b->signature)) {
			const std::vector<int>& t = ctx->batch_tuned[b->signature];
			b->cfg.tile_x = t[0];
			b->cfg.schedule = t[1];
			b->cfg.slots_per_cu = t[2];
		} else {
			b->cfg = to_internal(c);
		}
		b->noise = arma_hip::noise_batch_create(ctx, b->precision, b->members);
		b->ar = arma_hip::ar_batch_create(ctx, b->precision, b->members, b->cfg);
		b->rescale = arma_hip::rescale_batch_create(ctx, b->precision, b->members);
		*out = b.release();
	});
}

void arma_hip_batch_destroy(arma_hip_batch* batch) { delete batch; }

int arma_hip_batch_layout(const arma_hip_batch* batch, int64_t* offsets) {
	return guarded([&] {
		ARMA_HIP_REQUIRE(batch && offsets, ARMA_HIP_EINVAL, "null argument");
		std::memcpy(offsets, batch->offsets.data(), sizeof(int64_t) * batch->offsets.size());
	});
}

int arma_hip_batch_generate(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && arena, ARMA_HIP_EINVAL, "null argument");
		batch_enqueue(ctx, batch, arena);
	});
}

int arma_hip_batch_capture(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && arena, ARMA_HIP_EINVAL, "null argument");
		ARMA_HIP_REQUIRE(ctx->stream != nullptr, ARMA_HIP_EINVAL,
		                 "graph capture needs a stream other than the null stream");
		batch->drop_graph();
		const bool prof = ctx->profiling;
		ctx->profiling = false;  // no event records inside the graph
		ARMA_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
		hipGraph_t g = nullptr;
		try {
			batch_enqueue(ctx, batch, arena);
		} catch (...) {
			(void)hipStreamEndCapture(ctx->stream, &g);
			if (g) (void)hipGraphDestroy(g);
			ctx->profiling = prof;
			throw;
		}
		ctx->profiling = prof;
		ARMA_HIP_CHECK(hipStreamEndCapture(ctx->stream, &g));
		batch->graph = g;
		ARMA_HIP_CHECK(hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0));
	});
}

int arma_hip_batch_replay(arma_hip_ctx* ctx, arma_hip_batch* batch) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && batch->exec, ARMA_HIP_EINVAL, "no captured graph");
		arma_hip::KernelTimer timer(ctx, "batch_graph");
		ARMA_HIP_CHECK(hipGraphLaunch(batch->exec, ctx->stream));
	});
}

int arma_hip_batch_tune(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena, int32_t reps,
                        arma_batch_trial* trials, int32_t max_trials, int32_t* n_trials) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && arena, ARMA_HIP_EINVAL, "null argument");
		reps = std::max(reps, 1);
		hipEvent_t e0, e1;
		ARMA_HIP_CHECK(hipEventCreate(&e0));
		ARMA_HIP_CHECK(hipEventCreate(&e1));
		struct Ev {
			hipEvent_t a, b;
			~Ev() {
				(void)hipEventDestroy(a);
				(void)hipEventDestroy(b);
			}
		} ev{e0, e1};
		const bool prof = ctx->profiling;
		ctx->profiling = false;
		struct Restore {
			arma_hip_ctx* c;
			bool p;
			~Restore() { c->profiling = p; }
		} restore{ctx, prof};
		std::vector<arma_hip::BatchArConfig> cands;
		for (int tx : {16, 8, 32})
			for (int sc : {ARMA_BATCH_SURFACE_MAJOR, ARMA_BATCH_INTERLEAVED})
				for (int half = 0; half < 2; ++half) {
					arma_hip::BatchArConfig c;
					c.tile_x = tx;
					c.schedule = sc;
					c.slots_per_cu = half;  // resolved below: 0 = all that fit, 1 = half
					cands.push_back(c);
				}
		batch->drop_graph();
		arma_hip::ArBatch* best = nullptr;
		arma_hip::BatchArConfig best_cfg;
		double best_ms = 0.0;
		int count = 0;
		try {
			for (arma_hip::BatchArConfig c : cands) {
				if (!arma_hip::ar_batch_fits(batch->members, c.tile_x)) continue;
				int want = 0;
				if (c.slots_per_cu == 1) {
					// half residency: only a candidate when more than one fits
					arma_hip::BatchArConfig full = c;
					full.slots_per_cu = 0;
					std::unique_ptr<arma_hip::ArBatch, void (*)(arma_hip::ArBatch*)> probe(
						arma_hip::ar_batch_create(ctx, batch->precision, batch->members, full),
						arma_hip::ar_batch_destroy);
					const int mx = arma_hip::ar_batch_max_slots(probe.get());
					if (mx < 2) continue;
					want = mx / 2;
				}
				c.slots_per_cu = want;
				arma_hip::ArBatch* ab =
					arma_hip::ar_batch_create(ctx, batch->precision, batch->members, c);
				double ms = 1e300;
				for (int r = 0; r < reps; ++r) {
					arma_hip::noise_batch_launch(ctx, batch->noise, arena);
					ARMA_HIP_CHECK(hipEventRecord(e0, ctx->stream));
					arma_hip::ar_batch_launch(ctx, ab, arena);
					ARMA_HIP_CHECK(hipEventRecord(e1, ctx->stream));
					ARMA_HIP_CHECK(hipEventSynchronize(e1));
					arma_hip::ar_batch_check(ctx, ab);
					float t = 0.0f;
					ARMA_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
					ms = std::min(ms, double(t));
				}
				if (trials && count < max_trials) {
					arma_batch_trial& tr = trials[count];
					tr.config.tile_x = c.tile_x;
					tr.config.schedule = c.schedule;
					tr.config.slots_per_cu = c.slots_per_cu;
					tr.grid = arma_hip::ar_batch_grid(ab);
					tr.ms = ms;
				}
				++count;
				if (!best || ms < best_ms) {
					if (best) arma_hip::ar_batch_destroy(best);
					best = ab;
					best_cfg = c;
					best_ms = ms;
				} else {
					arma_hip::ar_batch_destroy(ab);
				}
			}
		} catch (...) {
			if (best) arma_hip::ar_batch_destroy(best);
			throw;
		}
		if (n_trials) *n_trials = count;
		ARMA_HIP_REQUIRE(best != nullptr, ARMA_HIP_ERUNTIME, "no batch configuration ran");
		arma_hip::ar_batch_destroy(batch->ar);
		batch->ar = best;
		batch->cfg = best_cfg;
		ctx->batch_tuned[batch->signature] = {best_cfg.tile_x, best_cfg.schedule,
		                                      best_cfg.slots_per_cu};
	});
}

int arma_hip_batch_get_config(const arma_hip_batch* batch, arma_batch_config* config,
                              int32_t* grid) {
	return guarded([&] {
		ARMA_HIP_REQUIRE(batch, ARMA_HIP_EINVAL, "null argument");
		if (config) {
			config->tile_x = batch->cfg.tile_x;
			config->schedule = batch->cfg.schedule;
			config->slots_per_cu = batch->cfg.slots_per_cu;
		}
		if (grid) *grid = arma_hip::ar_batch_grid(batch->ar);
	});
}

int arma_hip_batch_set_config(arma_hip_ctx* ctx, arma_hip_batch* batch,
                              const arma_batch_config* config) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && config, ARMA_HIP_EINVAL, "null argument");
		check_config(*config);
		const arma_hip::BatchArConfig c = to_internal(*config);
		arma_hip::ArBatch* ab = arma_hip::ar_batch_create(ctx, batch->precision, batch->members, c);
		batch->drop_graph();
		ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		arma_hip::ar_batch_destroy(batch->ar);
		batch->ar = ab;
		batch->cfg = c;
	});
}

int arma_hip_batch_factors(arma_hip_ctx* ctx, arma_hip_batch* batch, double* out) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(batch && out, ARMA_HIP_EINVAL, "null argument");
		arma_hip::rescale_batch_factors(ctx, batch->rescale, out);
		arma_hip::ar_batch_check(ctx, batch->ar);
	});
}

int arma_hip_yule_walker_choi_device(arma_hip_ctx* ctx, const double* acf, int32_t n,
                                     int32_t max_order, double var_eps, int32_t determine_order,
                                     int32_t do_chop, double chop_eps, double* phi_out,
                                     int32_t order_out[3], double* varwn_solver) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(acf && phi_out && order_out && n >= 3, ARMA_HIP_EINVAL,
		                 "bad Yule-Walker argument");
		std::vector<double> phi;
		int l = 0;
		double var = 0.0;
		arma_hip::yule_walker_choi_device(ctx, acf, n, max_order, var_eps, determine_order != 0,
		                                  phi, l, var);
		int shp[3] = {l, l, l};
		if (do_chop) arma_hip::chop_shape(phi, shp, chop_eps, shp);
		for (int a = 0; a < shp[0]; ++a)
			for (int b = 0; b < shp[1]; ++b)
				for (int c = 0; c < shp[2]; ++c)
					phi_out[(size_t(a) * shp[1] + b) * shp[2] + c] = phi[(size_t(a) * l + b) * l + c];
		for (int k = 0; k < 3; ++k) order_out[k] = shp[k];
		if (varwn_solver) *varwn_solver = var;
	});
}

int arma_hip_yule_walker_choi_device_many(arma_hip_ctx* ctx, int32_t count, const double* acf,
                                          const int64_t* acf_off, const int32_t* n,
                                          const int32_t* max_order, double var_eps,
                                          int32_t determine_order, int32_t do_chop,
                                          double chop_eps, double* phi_out,
                                          const int64_t* phi_off, int32_t* order_out,
                                          double* varwn_solver, int32_t* status) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(count >= 0, ARMA_HIP_EINVAL, "bad model count");
		if (count == 0) return;
		ARMA_HIP_REQUIRE(acf && acf_off && n && max_order && phi_out && phi_off && order_out &&
		                     status,
		                 ARMA_HIP_EINVAL, "null argument");
		std::vector<arma_hip::YwProblem> probs(count);
		for (int i = 0; i < count; ++i) {
			ARMA_HIP_REQUIRE(n[i] >= 3, ARMA_HIP_EINVAL, "bad Yule-Walker argument");
			probs[i] = arma_hip::YwProblem{acf + acf_off[i], n[i], max_order[i], var_eps,
			                               determine_order != 0};
		}
		std::vector<arma_hip::YwSolution> sols;
		arma_hip::yule_walker_choi_device_many(ctx, probs, sols);
		for (int i = 0; i < count; ++i) {
			const arma_hip::YwSolution& r = sols[i];
			status[i] = r.status;
			if (r.status != ARMA_HIP_OK) continue;
			const int l = r.l;
			int shp[3] = {l, l, l};
			if (do_chop) arma_hip::chop_shape(r.phi, shp, chop_eps, shp);
			double* out = phi_out + phi_off[i];
			for (int a = 0; a < shp[0]; ++a)
				for (int b = 0; b < shp[1]; ++b)
					for (int c = 0; c < shp[2]; ++c)
						out[(size_t(a) * shp[1] + b) * shp[2] + c] =
							r.phi[(size_t(a) * l + b) * l + c];
			for (int k = 0; k < 3; ++k) order_out[3 * i + k] = shp[k];
			if (varwn_solver) varwn_solver[i] = r.varwn;
		}
	});
}

int arma_hip_white_noise_variance_device(arma_hip_ctx* ctx, const double* acf,
                                         const int32_t acf_shape[3], const double* phi,
                                         const int32_t phi_shape[3], double* out) {
	return guarded([&] {
		use_device(ctx);
		ARMA_HIP_REQUIRE(acf && acf_shape && phi && phi_shape && out, ARMA_HIP_EINVAL,
		                 "null argument");
		for (int k = 0; k < 3; ++k)
			ARMA_HIP_REQUIRE(phi_shape[k] > 0 && phi_shape[k] <= acf_shape[k], ARMA_HIP_EINVAL,
			                 "AR order larger than the ACF grid");
		*out = arma_hip::white_noise_variance_device(ctx, acf, acf_shape, phi, phi_shape);
	});
}

}  // extern "C"
