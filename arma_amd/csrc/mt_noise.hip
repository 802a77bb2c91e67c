// White noise on the GPU: dynamic-creator Mersenne Twister streams feeding
// the libstdc++ polar normal transform.
//
// Reference semantics:
//   genrand_mt / sgenrand_mt ........ src/parallel_mt.hh:217-262 (libdcmt)
//   generate_canonical, polar normal  /usr/include/c++/11/bits/random.tcc:
//                                     3348-3380, 1802-1836
//   whole-grid noise (framework=none) src/white_noise.hh:23-58,
//                                     src/generator/basic_arma_model.cc:36-48
//   per-partition noise ............. src/generator/ar_model_wn.cc:3-18 with
//                                     partitions of ar_model_parallel.cc:51-72
//
// Every stream is sequential by definition (the polar method rejects a
// data-dependent number of draws), so the partitioned kernel runs one
// stream per thread with the generator state in LDS, interleaved across the
// workgroup so that every state access is bank-conflict free.  The
// whole-grid (single stream) mode runs one workgroup that twists the state
// in parallel chunks, evaluates 256 polar attempts at a time and compacts
// the accepted ones with a workgroup prefix sum.
//
// All arithmetic is compiled without FP contraction, so every draw is
// computed with the roundings the reference source writes.  The FP64 polar
// step's logarithm is log_dd (log_dd.hpp): correctly rounded, where the
// device library's log is within an ulp, so a draw differs from the host's
// only where the host's log is not correctly rounded.
#include "common.hpp"
#include "log_dd.hpp"

#include <vector>

namespace arma_hip {
namespace {

constexpr int kMaxNN = 1392;  // largest dcmt state (p = 44497)

// The log table staged in LDS by the whole workgroup (before any lane leaves).
__device__ __forceinline__ void stage_log_table(double* lt) {
	for (int i = threadIdx.x; i < kLogDdRows * 3; i += blockDim.x) lt[i] = kLogDdTable[i];
	__syncthreads();
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t x, const arma_mt_params& p) {
	x ^= x >> p.shift0;
	x ^= (x << p.shiftB) & p.maskB;
	x ^= (x << p.shiftC) & p.maskC;
	x ^= x >> p.shift1;
	return x;
}

__device__ __forceinline__ uint32_t mt_twist_word(uint32_t hi, uint32_t lo, uint32_t src,
                                                  const arma_mt_params& p) {
	const uint32_t x = (hi & p.umask) | (lo & p.lmask);
	return src ^ (x >> 1) ^ ((x & 1u) ? p.aaa : 0u);
}

// Generator whose nn state words live in LDS at st[k * stride].
struct LdsMt {
	uint32_t* st;
	int stride;
	int i;
	arma_mt_params p;

	__device__ __forceinline__ uint32_t get(int k) const { return st[k * stride]; }
	__device__ __forceinline__ void set(int k, uint32_t v) { st[k * stride] = v; }

	__device__ void seed(uint32_t s) {
		for (int k = 0; k < p.nn; ++k) {
			set(k, s & p.wmask);
			s = 1812433253u * (s ^ (s >> 30)) + uint32_t(k + 1);
		}
		i = p.nn;
	}

	__device__ uint32_t next() {
		if (i >= p.nn) {
			const int n = p.nn, m = p.mm;
			int k = 0;
			for (; k < n - m; ++k) set(k, mt_twist_word(get(k), get(k + 1), get(k + m), p));
			for (; k < n - 1; ++k) set(k, mt_twist_word(get(k), get(k + 1), get(k + m - n), p));
			set(n - 1, mt_twist_word(get(n - 1), get(0), get(m - 1), p));
			i = 0;
		}
		return mt_temper(get(i++), p);
	}
};

// generate_canonical<double,53> over a 32-bit engine: two draws.
__device__ __forceinline__ double canonical_from(uint32_t u0, uint32_t u1) {
	double sum = double(u0);
	sum = sum + double(u1) * 4294967296.0;
	double ret = sum / 18446744073709551616.0;
	if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;  // nextafter(1, 0)
	return ret;
}

// generate_canonical<float,24>: one draw.
__device__ __forceinline__ float canonical_from_f(uint32_t u0) {
	float sum = float(u0);
	float ret = sum / 4294967296.0f;
	if (ret >= 1.0f) ret = 0x1.fffffep-1f;
	return ret;
}

template <class G>
__device__ __forceinline__ double canonical_d(G& g) {
	const uint32_t u0 = g.next();
	const uint32_t u1 = g.next();
	return canonical_from(u0, u1);
}

// Polar normal generator with the libstdc++ cache of the second variate.
struct PolarD {
	bool saved_available = false;
	double saved = 0.0;

	template <class G>
	__device__ double next(G& g, double stddev, const double* ltab) {
		double ret;
		if (saved_available) {
			saved_available = false;
			ret = saved;
		} else {
			double x, y, r2;
			do {
				x = 2.0 * canonical_d(g) - 1.0;
				y = 2.0 * canonical_d(g) - 1.0;
				r2 = x * x + y * y;
			} while (r2 > 1.0 || r2 == 0.0);
			const double mult = sqrt(-2.0 * log_dd(r2, ltab) / r2);
			saved = x * mult;
			saved_available = true;
			ret = y * mult;
		}
		return ret * stddev + 0.0;
	}
};

struct PolarF {
	bool saved_available = false;
	float saved = 0.0f;

	template <class G>
	__device__ float next(G& g, float stddev) {
		float ret;
		if (saved_available) {
			saved_available = false;
			ret = saved;
		} else {
			float x, y, r2;
			do {
				x = 2.0f * canonical_from_f(g.next()) - 1.0f;
				y = 2.0f * canonical_from_f(g.next()) - 1.0f;
				r2 = x * x + y * y;
			} while (r2 > 1.0f || r2 == 0.0f);
			const float mult = sqrtf(-2.0f * logf(r2) / r2);
			saved = x * mult;
			saved_available = true;
			ret = y * mult;
		}
		return ret * stddev + 0.0f;
	}
};

// The block of the output one stream fills: its origin and extents inside a
// surface, the surface's row lengths and the standard deviation.
template <class OutT>
struct StreamBlock {
	OutT* out;   // element (0, 0, 0) of the surface
	int sy, sz;  // surface extents along x and y
	int lo0, lo1, lo2;
	int e0, e1, e2;
	double sd;
	float sdf;
};

// One surface: stream p fills partition p (row-major partition index).
template <class OutT>
struct GridBlocks {
	OutT* out;
	int3 shape, part, np;
	double sd;
	float sdf;
	__device__ __forceinline__ StreamBlock<OutT> operator()(int p) const {
		const int a = p / (np.y * np.z), b = (p / np.z) % np.y, c = p % np.z;
		StreamBlock<OutT> k;
		k.out = out;
		k.sy = shape.y;
		k.sz = shape.z;
		k.lo0 = a * part.x;
		k.lo1 = b * part.y;
		k.lo2 = c * part.z;
		k.e0 = min(k.lo0 + part.x, shape.x) - k.lo0;
		k.e1 = min(k.lo1 + part.y, shape.y) - k.lo1;
		k.e2 = min(k.lo2 + part.z, shape.z) - k.lo2;
		k.sd = sd;
		k.sdf = sdf;
		return k;
	}
};

// Many surfaces: one record per stream.
template <class OutT>
struct TableBlocks {
	OutT* arena;
	const NoiseBlockRec* recs;
	__device__ __forceinline__ StreamBlock<OutT> operator()(int p) const {
		const NoiseBlockRec r = recs[p];
		return StreamBlock<OutT>{arena + r.base, r.sy, r.sz, r.lo0, r.lo1, r.lo2,
		                         r.e0,           r.e1, r.e2, r.sd,  r.sdf};
	}
};

// One thread per partition; partitions in row-major order (stream p fills
// partition p), noise row-major inside each partition.
template <class OutT, int SEM, class Geo>
__global__ void noise_partitioned_kernel(const arma_mt_params* __restrict__ params,
                                         const uint32_t* __restrict__ seeds, Geo geo,
                                         int ntotal) {
	extern __shared__ uint32_t lds_state[];
	const int tid = threadIdx.x;
	const int p = blockIdx.x * blockDim.x + tid;
	if (p >= ntotal) return;
	LdsMt g;
	g.st = lds_state + tid;
	g.stride = blockDim.x;
	g.p = params[p];
	g.seed(seeds ? seeds[p] : 0u);
	const StreamBlock<OutT> k = geo(p);
	PolarD nd;
	PolarF nf;
	for (int t = 0; t < k.e0; ++t) {
		for (int x = 0; x < k.e1; ++x) {
			OutT* row = k.out + ((size_t)(k.lo0 + t) * k.sy + (k.lo1 + x)) * k.sz + k.lo2;
			for (int y = 0; y < k.e2; ++y) {
				if (SEM == ARMA_NOISE_DOUBLE) {
					row[y] = OutT(nd.next(g, k.sd, kLogDdTable));
				} else {
					row[y] = OutT(nf.next(g, k.sdf));
				}
			}
		}
	}
}

// Raw tempered draws, one thread per stream.
__global__ void mt_raw_kernel(const arma_mt_params* __restrict__ params,
                              const uint32_t* __restrict__ seeds, int n_streams, int draws,
                              uint32_t* __restrict__ out) {
	extern __shared__ uint32_t lds_state[];
	const int tid = threadIdx.x;
	const int s = blockIdx.x * blockDim.x + tid;
	if (s >= n_streams) return;
	LdsMt g;
	g.st = lds_state + tid;
	g.stride = blockDim.x;
	g.p = params[s];
	g.seed(seeds ? seeds[s] : 0u);
	for (int k = 0; k < draws; ++k) out[(size_t)s * draws + k] = g.next();
}

// ---- one stream per workgroup: cooperative twist, parallel polar attempts ----
//
// A workgroup owns one stream and the block of the grid it fills (the whole
// grid for framework=none, one partition otherwise).  The twist runs in
// chunks of c = nn - mm words (within a chunk every new word depends only on
// old words and on earlier chunks); each round evaluates THREADS polar
// attempts in parallel, compacts the accepted ones with a ballot prefix sum
// and writes their two variates at the positions the sequential generator
// would have given them.

template <int THREADS, int MAXNN>
__device__ void twist_parallel(uint32_t* st, const arma_mt_params& p, int tid) {
	constexpr int PER = (MAXNN + THREADS - 1) / THREADS;
	const int n = p.nn, m = p.mm, c = n - m;
	for (int k0 = 0; k0 < n; k0 += c) {
		const int k1 = min(k0 + c, n);
		uint32_t v[PER];
#pragma unroll
		for (int j = 0; j < PER; ++j) {
			const int k = k0 + tid + j * THREADS;
			if (k < k1) {
				if (k < n - m) v[j] = mt_twist_word(st[k], st[k + 1], st[k + m], p);
				else if (k < n - 1) v[j] = mt_twist_word(st[k], st[k + 1], st[k + m - n], p);
				else v[j] = mt_twist_word(st[n - 1], st[0], st[m - 1], p);
			}
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < PER; ++j) {
			const int k = k0 + tid + j * THREADS;
			if (k < k1) st[k] = v[j];
		}
		__syncthreads();
	}
}

template <class OutT, int SEM, int THREADS, int MAXNN, class Geo>
__global__ __launch_bounds__(THREADS) void noise_stream_kernel(
	const arma_mt_params* __restrict__ params, const uint32_t* __restrict__ seeds, Geo geo) {
	constexpr int words_per_attempt = SEM == ARMA_NOISE_DOUBLE ? 4 : 2;
	constexpr int round_words = THREADS * words_per_attempt;
	constexpr int KEEP = (MAXNN + THREADS - 1) / THREADS;
	__shared__ uint32_t st[MAXNN];
	__shared__ uint32_t buf[round_words + MAXNN];
	__shared__ int wave_tot[THREADS / 64];
	__shared__ double ltab[kLogDdRows * 3];
	stage_log_table(ltab);
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = tid >> 6;
	const int pid = blockIdx.x;
	const arma_mt_params p = params[pid];
	// the block of the grid this stream fills, row-major
	const StreamBlock<OutT> blk = geo(pid);
	OutT* const out = blk.out;
	const int sy = blk.sy, sz = blk.sz;
	const int lo0 = blk.lo0, lo1 = blk.lo1, lo2 = blk.lo2;
	const int e2 = blk.e2;
	const double stddev = blk.sd;
	const float stddev_f = blk.sdf;
	const long long e12 = (long long)blk.e1 * e2;
	const long long n = (long long)blk.e0 * e12;
	// position -> (t, x, y) inside the block; 32-bit quotients through a
	// double reciprocal (exact after one correction) while positions fit
	const double r12 = 1.0 / double(e12), r2 = 1.0 / double(e2);
	const bool small = n < (1ll << 31);
	auto put = [&](long long q, OutT v) {
		long long t;
		int x, y;
		if (small) {
			const uint32_t uq = uint32_t(q), u12 = uint32_t(e12), u2 = uint32_t(e2);
			uint32_t qt = uint32_t(double(uq) * r12);
			if (uint64_t(qt) * u12 > uq) --qt;
			else if (uint64_t(qt + 1) * u12 <= uq) ++qt;
			const uint32_t rr = uq - qt * u12;
			uint32_t qx = uint32_t(double(rr) * r2);
			if (uint64_t(qx) * u2 > rr) --qx;
			else if (uint64_t(qx + 1) * u2 <= rr) ++qx;
			t = qt;
			x = int(qx);
			y = int(rr - qx * u2);
		} else {
			t = q / e12;
			const long long r = q - t * e12;
			x = int(r / e2);
			y = int(r - (long long)x * e2);
		}
		out[(size_t(lo0 + t) * sy + (lo1 + x)) * sz + (lo2 + y)] = v;
	};
	if (tid == 0) {
		uint32_t s = seeds ? seeds[pid] : 0u;
		for (int k = 0; k < p.nn; ++k) {
			st[k] = s & p.wmask;
			s = 1812433253u * (s ^ (s >> 30)) + uint32_t(k + 1);
		}
	}
	__syncthreads();
	int avail = 0;
	long long produced = 0;
	while (produced < n) {
		while (avail < round_words) {
			twist_parallel<THREADS, MAXNN>(st, p, tid);
			for (int k = tid; k < p.nn; k += THREADS) buf[avail + k] = mt_temper(st[k], p);
			avail += p.nn;
			__syncthreads();
		}
		const uint32_t* w = buf + tid * words_per_attempt;
		bool accept;
		double xd = 0, yd = 0, r2d = 0;
		float xf = 0, yf = 0, r2f = 0;
		if (SEM == ARMA_NOISE_DOUBLE) {
			xd = 2.0 * canonical_from(w[0], w[1]) - 1.0;
			yd = 2.0 * canonical_from(w[2], w[3]) - 1.0;
			r2d = xd * xd + yd * yd;
			accept = !(r2d > 1.0 || r2d == 0.0);
		} else {
			xf = 2.0f * canonical_from_f(w[0]) - 1.0f;
			yf = 2.0f * canonical_from_f(w[1]) - 1.0f;
			r2f = xf * xf + yf * yf;
			accept = !(r2f > 1.0f || r2f == 0.0f);
		}
		const unsigned long long mask = __ballot(accept);
		const int below = __popcll(mask & ((1ull << lane) - 1ull));
		if (lane == 0) wave_tot[wave] = __popcll(mask);
		__syncthreads();
		int before = 0, total = 0;
		for (int q = 0; q < THREADS / 64; ++q) {
			if (q < wave) before += wave_tot[q];
			total += wave_tot[q];
		}
		if (accept) {
			const long long pos = produced + 2ll * (before + below);
			if (pos < n) {
				if (SEM == ARMA_NOISE_DOUBLE) {
					const double mult = sqrt(-2.0 * log_dd(r2d, ltab) / r2d);
					put(pos, OutT(yd * mult * stddev + 0.0));
					if (pos + 1 < n) put(pos + 1, OutT(xd * mult * stddev + 0.0));
				} else {
					const float mult = sqrtf(-2.0f * logf(r2f) / r2f);
					put(pos, OutT(yf * mult * stddev_f + 0.0f));
					if (pos + 1 < n) put(pos + 1, OutT(xf * mult * stddev_f + 0.0f));
				}
			}
		}
		produced += 2ll * total;
		__syncthreads();
		// drop the consumed words
		const int left = avail - round_words;
		uint32_t keep[KEEP];
#pragma unroll
		for (int j = 0; j < KEEP; ++j) {
			const int k = tid + j * THREADS;
			if (k < left) keep[j] = buf[round_words + k];
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < KEEP; ++j) {
			const int k = tid + j * THREADS;
			if (k < left) buf[k] = keep[j];
		}
		avail = left;
		__syncthreads();
	}
}

// ---- two streams per wave: one-round twist, 32 polar attempts per stream ----
//
// For small dcmt states (nn <= 32, 2*mm <= nn, e.g. p = 521, 607) a wave
// runs two streams on 32 lanes each.  Every twist is a single LDS round:
// lane k computes word k of the new state from old words only, evaluating
// the one or two new words it depends on (k + mm - nn, and 0 and mm - 1 for
// the last word) itself, which are themselves functions of old words when
// 2*mm <= nn.  Each round then tempers whole twists into a word buffer until
// 32 attempts' worth of words is there, evaluates the attempts in parallel
// and places the accepted pairs at the positions the sequential generator
// gives them (prefix count over the 32-lane half of the wave ballot).

constexpr int kSegLanes = 32;

__device__ __forceinline__ uint32_t twist_word_old(const uint32_t* st, int k,
                                                   const arma_mt_params& p) {
	// new word k < nn - mm: old words k, k+1, k+mm
	return mt_twist_word(st[k], st[k + 1], st[k + p.mm], p);
}

template <class OutT, int SEM, class Geo>
__global__ __launch_bounds__(64) void noise_seg_kernel(
	const arma_mt_params* __restrict__ params, const uint32_t* __restrict__ seeds, Geo geo,
	int nstreams) {
	constexpr int WPA = SEM == ARMA_NOISE_DOUBLE ? 4 : 2;  // words per attempt
	constexpr int RW = kSegLanes * WPA;                     // words per round
	__shared__ uint32_t st_all[2][32];
	__shared__ uint32_t buf_all[2][RW + 32];
	__shared__ double ltab[kLogDdRows * 3];
	stage_log_table(ltab);
	const int seg = threadIdx.x >> 5;
	const int lane = threadIdx.x & 31;
	const int pid = blockIdx.x * 2 + seg;
	const bool valid = pid < nstreams;
	uint32_t* st = st_all[seg];
	uint32_t* buf = buf_all[seg];
	arma_mt_params p = params[valid ? pid : 0];
	const int nn = p.nn;
	// the block of the grid this stream fills, row-major
	const StreamBlock<OutT> blk = geo(valid ? pid : 0);
	OutT* const out = blk.out;
	const int sy = blk.sy, sz = blk.sz;
	const int lo0 = blk.lo0, lo1 = blk.lo1, lo2 = blk.lo2;
	const int e2 = blk.e2;
	const double stddev = blk.sd;
	const float stddev_f = blk.sdf;
	const long long e12 = (long long)blk.e1 * e2;
	const long long n = valid ? (long long)blk.e0 * e12 : 0;
	const double r12 = 1.0 / double(e12), r2 = 1.0 / double(e2);
	// block extents that are powers of two (full partitions of the usual
	// shapes) split a position with shifts and masks
	const bool pow2 = (e2 & (e2 - 1)) == 0 && (e12 & (e12 - 1)) == 0;
	const int s2 = __ffs(unsigned(e2)) - 1, s12 = __ffsll((unsigned long long)e12) - 1;
	auto put = [&](long long q, OutT v) {
		// q < 2^31: 32-bit quotients
		const uint32_t uq = uint32_t(q), u12 = uint32_t(e12), u2 = uint32_t(e2);
		uint32_t qt, rr, qx;
		if (pow2) {
			qt = uq >> s12;
			rr = uq & (u12 - 1u);
			qx = rr >> s2;
		} else {
			// through a double reciprocal, corrected
			qt = uint32_t(double(uq) * r12);
			if (uint64_t(qt) * u12 > uq) --qt;
			else if (uint64_t(qt + 1) * u12 <= uq) ++qt;
			rr = uq - qt * u12;
			qx = uint32_t(double(rr) * r2);
			if (uint64_t(qx) * u2 > rr) --qx;
			else if (uint64_t(qx + 1) * u2 <= rr) ++qx;
		}
		out[(size_t(lo0 + qt) * sy + (lo1 + qx)) * sz + (lo2 + (rr - qx * u2))] = v;
	};
	if (lane == 0) {
		uint32_t sd = valid && seeds ? seeds[pid] : 0u;
		for (int k = 0; k < nn; ++k) {
			st[k] = sd & p.wmask;
			sd = 1812433253u * (sd ^ (sd >> 30)) + uint32_t(k + 1);
		}
	}
	__syncthreads();
	int avail = 0;
	long long produced = 0;
	// the workgroup is one wave: loop conditions are wave votes, and the
	// barriers below only order LDS accesses between lanes
	while (__any(produced < n)) {
		// twists until a round of words is buffered (avail is the same on
		// every lane of a segment)
		while (__any(avail < RW)) {
			const bool mine = avail < RW;
			uint32_t w = 0;
			if (mine && lane < nn) {
				const int k = lane, m = p.mm;
				if (k < nn - m) {
					w = twist_word_old(st, k, p);
				} else if (k < nn - 1) {
					w = mt_twist_word(st[k], st[k + 1], twist_word_old(st, k + m - nn, p), p);
				} else {
					w = mt_twist_word(st[nn - 1], twist_word_old(st, 0, p),
					                  twist_word_old(st, m - 1, p), p);
				}
			}
			__syncthreads();
			if (mine && lane < nn) {
				st[lane] = w;
				buf[avail + lane] = mt_temper(w, p);
			}
			if (mine) avail += nn;
			__syncthreads();
		}
		// one attempt per lane
		const uint32_t* wv = buf + lane * WPA;
		bool accept;
		double xd = 0, yd = 0, r2d = 0;
		float xf = 0, yf = 0, r2f = 0;
		if (SEM == ARMA_NOISE_DOUBLE) {
			xd = 2.0 * canonical_from(wv[0], wv[1]) - 1.0;
			yd = 2.0 * canonical_from(wv[2], wv[3]) - 1.0;
			r2d = xd * xd + yd * yd;
			accept = !(r2d > 1.0 || r2d == 0.0);
		} else {
			xf = 2.0f * canonical_from_f(wv[0]) - 1.0f;
			yf = 2.0f * canonical_from_f(wv[1]) - 1.0f;
			r2f = xf * xf + yf * yf;
			accept = !(r2f > 1.0f || r2f == 0.0f);
		}
		const unsigned long long mask = __ballot(accept);
		const uint32_t half = uint32_t(mask >> (32 * seg));
		const int below = __popc(half & ((1u << lane) - 1u));
		const int total = __popc(half);
		if (accept && produced < n) {
			const long long pos = produced + 2ll * below;
			if (pos < n) {
				if (SEM == ARMA_NOISE_DOUBLE) {
					const double mult = sqrt(-2.0 * log_dd(r2d, ltab) / r2d);
					put(pos, OutT(yd * mult * stddev + 0.0));
					if (pos + 1 < n) put(pos + 1, OutT(xd * mult * stddev + 0.0));
				} else {
					const float mult = sqrtf(-2.0f * logf(r2f) / r2f);
					put(pos, OutT(yf * mult * stddev_f + 0.0f));
					if (pos + 1 < n) put(pos + 1, OutT(xf * mult * stddev_f + 0.0f));
				}
			}
		}
		produced += 2ll * total;
		// drop the consumed words (fewer than 32 remain)
		const int left = avail - RW;
		const uint32_t keep = lane < left ? buf[RW + lane] : 0u;
		__syncthreads();
		if (lane < left) buf[lane] = keep;
		avail = left;
		__syncthreads();
	}
}

// ---- MT521 streams in registers: producer and consumer waves ----
//
// For the dynamic-creator records of p = 521 with 32-bit words (nn = 17,
// mm = 8, the records the partitioned configurations use), each 16-lane row
// of a producer wave runs one stream's twist: lane k of the row holds state
// word k (k < 16) and lane 15 also holds word 16.  A twist, in genrand_mt's
// sequential order (src/parallel_mt.hh:236), is
//   new[k]  = tw(old[k], old[k+1], old[k+8])     k = 0..8
//   new[k]  = tw(old[k], old[k+1], new[k-9])     k = 9..15
//   new[16] = tw(old[16], new[0], new[7])
// where tw(a, b, s) = s ^ y(a, b) and y depends on old words only, so a twist
// is one y per lane, the row shifts that bring old[k+1], old[k+8], new[k-9],
// new[0] and new[7] into place (DPP within each row, so the four streams of
// a wave never mix), and two XORs: 68 words for about the instructions the
// 17 words of one stream per wave took.
//
// The untempered words go into a 1024-word ring per stream in LDS.  A
// consumer wave takes the same four streams: lane k of a row evaluates
// attempts 16a + k (a < 4 in FP64 semantics, a < 8 in FP32) of its stream's
// round of 256 words, tempering the words it reads, with the attempts'
// chains interleaved, and stores the accepted pairs where
// the sequential generator puts them (prefix counts over the row's 16 bits of
// the wave ballots).  Both roles step through rounds together, one workgroup
// barrier per round: while the consumers use round r the producers twist
// round r + 1, so a twist's dependent chain and the attempts' transcendental
// chains issue side by side.  A workgroup is eight waves, four stream groups,
// with the roles ordered P0 C1 P2 C3 C0 P1 C2 P3 so that waves w and w + 4
// (the same SIMD when waves are placed round-robin) are one producer and one
// consumer.  The consumers publish per round whether their streams are
// complete (flags alternating by round parity, so a flag is rewritten only
// after every wave has read it); all waves leave after the same barrier.

constexpr int kWaveNN = 17, kWaveMM = 8;
constexpr int kRingWords = 1024;  // >= a round in use + a round + 17 written ahead
constexpr int kRoundWords = 256;  // words per round per stream
constexpr int kPcThreads = 512;

template <int CTRL>
__device__ __forceinline__ uint32_t row_move(uint32_t v) {
	return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xf, 0xf, true));
}
// as row_move, with lanes whose source falls outside the row keeping old
template <int CTRL>
__device__ __forceinline__ uint32_t row_move_or(uint32_t old, uint32_t v) {
	return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(v), CTRL, 0xf, 0xf, false));
}
constexpr int kRowShl1 = 0x101, kRowShl7 = 0x107, kRowShl8 = 0x108;  // lane k <- lane k + n
constexpr int kRowShr8 = 0x118, kRowShr9 = 0x119, kRowShr15 = 0x11f;  // lane k <- lane k - n

// y(a, b) = ((a & upper) | (b & lower)) >> 1, xor aaa when the low bit is set
__device__ __forceinline__ uint32_t mt_twist_y(uint32_t a, uint32_t b, const arma_mt_params& p) {
	const uint32_t x = (a & p.umask) | (b & p.lmask);
	return (x >> 1) ^ ((x & 1u) ? p.aaa : 0u);
}

template <class OutT, int SEM, class Geo>
__global__ __launch_bounds__(kPcThreads) void noise_wave_kernel(
	const arma_mt_params* __restrict__ params, const uint32_t* __restrict__ seeds, Geo geo,
	int nstreams) {
	constexpr int WPA = SEM == ARMA_NOISE_DOUBLE ? 4 : 2;  // words per attempt
	constexpr int A = kRoundWords / (16 * WPA);            // attempts per lane per round
	static_assert(A * 16 * WPA == kRoundWords, "a round is whole attempts");
	static_assert(3 * kRoundWords + kWaveNN <= kRingWords, "ring too small");
	// rows 16 words (16 banks) apart, so the four rows' stores of a twist and
	// reads of a round fall on different banks
	__shared__ __align__(16) uint32_t ring_all[4][4][kRingWords + 16];  // [group][row][word]
	__shared__ int gdone[2][4];
	__shared__ double ltab[kLogDdRows * 3];
	stage_log_table(ltab);
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int row = lane >> 4, k = lane & 15;
	const int g = wave & 3;
	const bool producer = ((wave >> 2) ^ (wave & 1)) == 0;
	const int pid = (int(blockIdx.x) * 4 + g) * 4 + row;
	const bool valid = pid < nstreams;
	uint32_t* ring = ring_all[g][row];
	const arma_mt_params p = params[valid ? pid : 0];

	// producer: the state (lane k: word k; word 16 on lane 15, held by all)
	uint32_t s = 0, t = 0;
	int wpos = 0;
	auto twist = [&]() {
		const uint32_t nx1 = row_move_or<kRowShl1>(t, s);  // old[k + 1]; lane 15: old[16]
		const uint32_t t8 = row_move<kRowShl7>(t);          // lane 8 <- old[16]
		const uint32_t nx8 = row_move_or<kRowShl8>(t8, s);  // old[k + 8]; lane 8: old[16]
		const uint32_t y = mt_twist_y(s, nx1, p);
		const uint32_t a = nx8 ^ y;                  // new[k], k <= 8
		const uint32_t a9 = row_move<kRowShr9>(a);   // new[k - 9]
		const uint32_t a0 = row_move<kRowShr15>(a);  // lane 15 <- new[0]
		const uint32_t a7 = row_move<kRowShr8>(a);   // lane 15 <- new[7]
		t = a7 ^ mt_twist_y(t, a0, p);               // new[16] (lane 15)
		s = k <= kWaveMM ? a : (a9 ^ y);
		ring[(wpos + k) & (kRingWords - 1)] = s;
		if (k == 15) ring[(wpos + 16) & (kRingWords - 1)] = t;
		wpos += kWaveNN;
	};

	// consumer: the block of the grid this row's stream fills
	const StreamBlock<OutT> blk = geo(valid ? pid : 0);
	OutT* const out = blk.out;
	const int sy = blk.sy, sz = blk.sz;
	const int lo0 = blk.lo0, lo1 = blk.lo1, lo2 = blk.lo2;
	const int e2 = blk.e2;
	const double stddev = blk.sd;
	const float stddev_f = blk.sdf;
	const long long e12 = (long long)blk.e1 * e2;
	const long long n = valid ? (long long)blk.e0 * e12 : 0;
	const double r12 = 1.0 / double(e12), r2 = 1.0 / double(e2);
	const bool pow2 = (e2 & (e2 - 1)) == 0 && (e12 & (e12 - 1)) == 0;
	const int s2 = __ffs(unsigned(e2)) - 1, s12 = __ffsll((unsigned long long)e12) - 1;
	auto index = [&](long long q) {
		// q < 2^31: 32-bit quotients
		const uint32_t uq = uint32_t(q), u12 = uint32_t(e12), u2 = uint32_t(e2);
		uint32_t qt, rr, qx;
		if (pow2) {
			qt = uq >> s12;
			rr = uq & (u12 - 1u);
			qx = rr >> s2;
		} else {
			qt = uint32_t(double(uq) * r12);
			if (uint64_t(qt) * u12 > uq) --qt;
			else if (uint64_t(qt + 1) * u12 <= uq) ++qt;
			rr = uq - qt * u12;
			qx = uint32_t(double(rr) * r2);
			if (uint64_t(qx) * u2 > rr) --qx;
			else if (uint64_t(qx + 1) * u2 <= rr) ++qx;
		}
		return (size_t(lo0 + qt) * sy + (lo1 + qx)) * sz + (lo2 + (rr - qx * u2));
	};
	// the pair at positions q (even) and q + 1: with an even row length the
	// second is the next element of the same row
	const bool even_rows = (e2 & 1) == 0;
	auto put_pair = [&](long long q, OutT v0, OutT v1) {
		const size_t i0 = index(q);
		out[i0] = v0;
		if (q + 1 < n) out[even_rows ? i0 + 1 : index(q + 1)] = v1;
	};
	long long produced = 0;
	auto consume = [&](int r) {
		// attempts a * 16 + k (a < A) of the row's stream in round r; the
		// transform is evaluated on every lane (a rejected attempt on a
		// stand-in radius) and only accepted pairs are stored
		const uint32_t* words = ring + ((r * kRoundWords) & (kRingWords - 1));
		bool acc[A];
		double xd[A], yd[A], r2d[A];
		float xf[A], yf[A], r2f[A];
#pragma unroll
		for (int a = 0; a < A; ++a) {
			const uint32_t* wv = words + (a * 16 + k) * WPA;
			if (SEM == ARMA_NOISE_DOUBLE) {
				const uint4 w = *reinterpret_cast<const uint4*>(wv);
				xd[a] = 2.0 * canonical_from(mt_temper(w.x, p), mt_temper(w.y, p)) - 1.0;
				yd[a] = 2.0 * canonical_from(mt_temper(w.z, p), mt_temper(w.w, p)) - 1.0;
				r2d[a] = xd[a] * xd[a] + yd[a] * yd[a];
				acc[a] = !(r2d[a] > 1.0 || r2d[a] == 0.0);
			} else {
				const uint2 w = *reinterpret_cast<const uint2*>(wv);
				xf[a] = 2.0f * canonical_from_f(mt_temper(w.x, p)) - 1.0f;
				yf[a] = 2.0f * canonical_from_f(mt_temper(w.y, p)) - 1.0f;
				r2f[a] = xf[a] * xf[a] + yf[a] * yf[a];
				acc[a] = !(r2f[a] > 1.0f || r2f[a] == 0.0f);
			}
		}
		int below[A];
		int total = 0;
#pragma unroll
		for (int a = 0; a < A; ++a) {
			const uint32_t bits = uint32_t(__ballot(acc[a]) >> (16 * row)) & 0xffffu;
			below[a] = total + __popc(bits & ((1u << k) - 1u));
			total += __popc(bits);
		}
#pragma unroll
		for (int a = 0; a < A; ++a) {
			const long long pos = produced + 2ll * below[a];
			if (SEM == ARMA_NOISE_DOUBLE) {
				const double rr2 = acc[a] ? r2d[a] : 0.5;
				const double mult = sqrt(-2.0 * log_dd(rr2, ltab) / rr2);
				if (acc[a] && pos < n)
					put_pair(pos, OutT(yd[a] * mult * stddev + 0.0), OutT(xd[a] * mult * stddev + 0.0));
			} else {
				const float rr2 = acc[a] ? r2f[a] : 0.5f;
				const float mult = sqrtf(-2.0f * logf(rr2) / rr2);
				if (acc[a] && pos < n)
					put_pair(pos, OutT(yf[a] * mult * stddev_f + 0.0f),
					         OutT(xf[a] * mult * stddev_f + 0.0f));
			}
		}
		produced += 2ll * total;
	};

	if (producer) {
		// sgenrand_mt: word w of the recurrence from the seed; lane k keeps
		// word k, every lane keeps word 16 (lane 15's copy is the one used)
		uint32_t x = valid && seeds ? seeds[pid] : 0u;
		for (int w = 0; w < kWaveNN; ++w) {
			const uint32_t v = x & p.wmask;
			if (w == k) s = v;
			if (w == kWaveNN - 1) t = v;
			x = 1812433253u * (x ^ (x >> 30)) + uint32_t(w + 1);
		}
		while (wpos < kRoundWords) twist();
	}
	__syncthreads();  // round 0 is in the rings
	for (int r = 0;; ++r) {
		if (producer) {
			while (wpos < (r + 2) * kRoundWords) twist();  // round r + 1
		} else {
			consume(r);
			const bool mine_done = !__any(produced < n);
			if (lane == 0) gdone[r & 1][g] = mine_done;
		}
		__syncthreads();
		if (gdone[r & 1][0] && gdone[r & 1][1] && gdone[r & 1][2] && gdone[r & 1][3]) break;
	}
}

bool wave_kernel_fits(const arma_mt_params* mts, int n) {
	for (int i = 0; i < n; ++i)
		if (mts[i].nn != kWaveNN || mts[i].mm != kWaveMM) return false;
	return true;
}

// The two-per-wave kernel applies to states of at most 32 words whose twist
// needs only old words for its redundant evaluations.
bool seg_kernel_fits(const arma_mt_params* mts, int n) {
	for (int i = 0; i < n; ++i)
		if (mts[i].nn > kSegLanes || 2 * mts[i].mm > mts[i].nn) return false;
	return true;
}

int threads_for_state(int nn) {
	// per-thread LDS state of nn words; at most 64 KiB per workgroup
	int t = 256;
	while (t > 1 && size_t(t) * nn * 4 > 65536) t >>= 1;
	return t;
}

void validate_params(const arma_mt_params* mts, int n) {
	for (int i = 0; i < n; ++i) {
		const arma_mt_params& q = mts[i];
		ARMA_HIP_REQUIRE(q.nn > 1 && q.nn <= kMaxNN && q.mm > 0 && q.mm < q.nn &&
		                     q.shift0 >= 0 && q.shift0 < 32 && q.shift1 >= 0 && q.shift1 < 32 &&
		                     q.shiftB >= 0 && q.shiftB < 32 && q.shiftC >= 0 && q.shiftC < 32,
		                 ARMA_HIP_EINVAL, "bad Mersenne Twister parameter record");
	}
}

// Launches the noise of nstreams streams whose output blocks geo describes,
// with the kernel their parameter sets allow: a whole workgroup for a single
// stream, two streams per wave for small dcmt states, one wave per stream up
// to 64 state words, one workgroup per stream for a few large states, else
// one thread per stream.
template <class OutT, int SEM, class Geo>
void launch_streams(arma_hip_ctx* ctx, const arma_mt_params* hparams,
                    const arma_mt_params* dparams, const uint32_t* dseeds, Geo geo, int nstreams,
                    bool small_blocks) {
	int nn_max = 0;
	for (int i = 0; i < nstreams; ++i) nn_max = std::max(nn_max, int(hparams[i].nn));
	if (nstreams == 1) {
		// the whole grid from one stream (framework=none)
		noise_stream_kernel<OutT, SEM, 256, kMaxNN, Geo>
			<<<1, 256, 0, ctx->stream>>>(dparams, dseeds, geo);
	} else if (small_blocks && wave_kernel_fits(hparams, nstreams)) {
		noise_wave_kernel<OutT, SEM, Geo>
			<<<div_ceil(nstreams, 16), kPcThreads, 0, ctx->stream>>>(dparams, dseeds, geo,
			                                                       nstreams);
	} else if (small_blocks && seg_kernel_fits(hparams, nstreams)) {
		noise_seg_kernel<OutT, SEM, Geo>
			<<<div_ceil(nstreams, 2), 64, 0, ctx->stream>>>(dparams, dseeds, geo, nstreams);
	} else if (nn_max <= 64) {
		noise_stream_kernel<OutT, SEM, 64, 64, Geo>
			<<<nstreams, 64, 0, ctx->stream>>>(dparams, dseeds, geo);
	} else if (nstreams <= 4 * std::max(ctx->num_cus, 1)) {
		noise_stream_kernel<OutT, SEM, 256, kMaxNN, Geo>
			<<<nstreams, 256, 0, ctx->stream>>>(dparams, dseeds, geo);
	} else {
		const int tpb = std::min(threads_for_state(nn_max), 64);
		noise_partitioned_kernel<OutT, SEM, Geo>
			<<<div_ceil(nstreams, tpb), tpb, size_t(tpb) * nn_max * 4, ctx->stream>>>(
				dparams, dseeds, geo, nstreams);
	}
	ARMA_HIP_CHECK(hipGetLastError());
}

template <class Geo2>
void launch_by_type(arma_hip_ctx* ctx, bool fp32, int sem, const arma_mt_params* hparams,
                    const arma_mt_params* dparams, const uint32_t* dseeds, void* out,
                    const Geo2& g, int nstreams, bool small_blocks) {
	if (fp32) {
		auto geo = g.template make<float>((float*)out);
		if (sem == ARMA_NOISE_DOUBLE)
			launch_streams<float, ARMA_NOISE_DOUBLE>(ctx, hparams, dparams, dseeds, geo, nstreams,
			                                         small_blocks);
		else
			launch_streams<float, ARMA_NOISE_FLOAT>(ctx, hparams, dparams, dseeds, geo, nstreams,
			                                        small_blocks);
	} else {
		auto geo = g.template make<double>((double*)out);
		if (sem == ARMA_NOISE_DOUBLE)
			launch_streams<double, ARMA_NOISE_DOUBLE>(ctx, hparams, dparams, dseeds, geo,
			                                          nstreams, small_blocks);
		else
			launch_streams<double, ARMA_NOISE_FLOAT>(ctx, hparams, dparams, dseeds, geo,
			                                         nstreams, small_blocks);
	}
}

struct GridSpec {
	int3 shape, part, np;
	double sd;
	float sdf;
	template <class OutT>
	GridBlocks<OutT> make(OutT* out) const {
		return GridBlocks<OutT>{out, shape, part, np, sd, sdf};
	}
};

struct TableSpec {
	const NoiseBlockRec* recs;
	template <class OutT>
	TableBlocks<OutT> make(OutT* arena) const {
		return TableBlocks<OutT>{arena, recs};
	}
};

}  // namespace

void launch_white_noise(arma_hip_ctx* ctx, const arma_ar_desc* d, void* eps) {
	ARMA_HIP_REQUIRE(d->varwn >= 0.0, ARMA_HIP_EINVAL, "variance is less than zero");
	ARMA_HIP_REQUIRE(d->shape[0] > 0 && d->shape[1] > 0 && d->shape[2] > 0, ARMA_HIP_EINVAL,
	                 "bad output grid shape");
	const bool fp32 = d->precision == ARMA_FP32;
	const int sem = d->noise_semantics;
	ARMA_HIP_REQUIRE(sem == ARMA_NOISE_DOUBLE || sem == ARMA_NOISE_FLOAT, ARMA_HIP_EINVAL,
	                 "bad noise semantics");
	GridSpec g;
	g.shape = int3{d->shape[0], d->shape[1], d->shape[2]};
	g.sd = std::sqrt(d->varwn);
	g.sdf = std::sqrt(float(d->varwn));
	if (d->noise_mode == ARMA_NOISE_SEQUENTIAL) {
		g.part = g.shape;
		g.np = int3{1, 1, 1};
	} else {
		ARMA_HIP_REQUIRE(d->noise_mode == ARMA_NOISE_PARTITIONED, ARMA_HIP_EINVAL,
		                 "bad noise mode");
		ARMA_HIP_REQUIRE(d->partition[0] > 0 && d->partition[1] > 0 && d->partition[2] > 0,
		                 ARMA_HIP_EINVAL, "bad partition shape");
		g.part = int3{d->partition[0], d->partition[1], d->partition[2]};
		g.np = int3{div_ceil(g.shape.x, g.part.x), div_ceil(g.shape.y, g.part.y),
		            div_ceil(g.shape.z, g.part.z)};
	}
	const int ntotal = g.np.x * g.np.y * g.np.z;
	ARMA_HIP_REQUIRE(d->n_mts >= ntotal && d->mts, ARMA_HIP_EPRNG, "bad number of MT configs");
	validate_params(d->mts, ntotal);
	auto* dparams = (arma_mt_params*)work_buffer(ctx, ctx->params, sizeof(arma_mt_params) * ntotal);
	upload(ctx, dparams, d->mts, sizeof(arma_mt_params) * ntotal);
	uint32_t* dseeds = nullptr;
	if (d->seeds) {
		dseeds = (uint32_t*)work_buffer(ctx, ctx->seeds, sizeof(uint32_t) * ntotal);
		upload(ctx, dseeds, d->seeds, sizeof(uint32_t) * ntotal);
	}
	// every stream's block must fit 32-bit positions for the fast kernels
	const bool small_blocks = (long long)g.part.x * g.part.y * g.part.z < (1ll << 31);
	KernelTimer timer(ctx, d->noise_mode == ARMA_NOISE_SEQUENTIAL ? "noise_sequential"
	                                                               : "noise_partitioned");
	launch_by_type(ctx, fp32, sem, d->mts, dparams, dseeds, eps, g, ntotal, small_blocks);
}

size_t white_noise_work_bytes(const arma_ar_desc* d) {
	size_t n = 1;
	if (d->noise_mode == ARMA_NOISE_PARTITIONED && d->partition[0] > 0 && d->partition[1] > 0 &&
	    d->partition[2] > 0)
		n = size_t(div_ceil(d->shape[0], d->partition[0])) * div_ceil(d->shape[1], d->partition[1]) *
		    div_ceil(d->shape[2], d->partition[2]);
	const size_t params = (sizeof(arma_mt_params) * n + 255) / 256 * 256;
	return params + (d->seeds ? sizeof(uint32_t) * n : 0);
}

void launch_mt_raw(arma_hip_ctx* ctx, const arma_mt_params* mts, const uint32_t* seeds,
                   int n_streams, int draws, uint32_t* out) {
	ARMA_HIP_REQUIRE(n_streams > 0 && draws >= 0 && mts, ARMA_HIP_EINVAL, "bad stream count");
	validate_params(mts, n_streams);
	int nn_max = 0;
	for (int i = 0; i < n_streams; ++i) nn_max = std::max(nn_max, int(mts[i].nn));
	auto* dparams = (arma_mt_params*)ctx->params.get(sizeof(arma_mt_params) * n_streams);
	upload(ctx, dparams, mts, sizeof(arma_mt_params) * n_streams);
	uint32_t* dseeds = nullptr;
	if (seeds) {
		dseeds = (uint32_t*)ctx->seeds.get(sizeof(uint32_t) * n_streams);
		upload(ctx, dseeds, seeds, sizeof(uint32_t) * n_streams);
	}
	const int tpb = std::min(threads_for_state(nn_max), 64);
	mt_raw_kernel<<<div_ceil(n_streams, tpb), tpb, size_t(tpb) * nn_max * 4, ctx->stream>>>(
		dparams, dseeds, n_streams, draws, out);
	ARMA_HIP_CHECK(hipGetLastError());
}

struct NoiseBatch {
	int precision = 0;
	int sem = ARMA_NOISE_DOUBLE;
	int nstreams = 0;
	bool small_blocks = true;
	std::vector<arma_mt_params> hparams;
	Scratch params, seeds, recs;
};

NoiseBatch* noise_batch_create(arma_hip_ctx* ctx, int precision,
                               const std::vector<BatchMember>& m) {
	auto* b = new NoiseBatch;
	try {
		noise_batch_build(ctx, b, precision, m);
	} catch (...) {
		delete b;
		throw;
	}
	return b;
}

NoiseBatch* noise_batch_new() { return new NoiseBatch; }

void noise_batch_build(arma_hip_ctx* ctx, NoiseBatch* b, int precision,
                       const std::vector<BatchMember>& m) {
	{
		b->precision = precision;
		b->small_blocks = true;
		b->hparams.clear();
		ARMA_HIP_REQUIRE(!m.empty(), ARMA_HIP_EINVAL, "empty batch");
		b->sem = m[0].d->noise_semantics;
		std::vector<uint32_t> seeds;
		std::vector<NoiseBlockRec> recs;
		for (const BatchMember& x : m) {
			const arma_ar_desc* d = x.d;
			ARMA_HIP_REQUIRE(d->varwn >= 0.0, ARMA_HIP_EINVAL, "variance is less than zero");
			ARMA_HIP_REQUIRE(d->noise_semantics == b->sem, ARMA_HIP_EINVAL,
			                 "every surface of a batch needs the same noise semantics");
			// the generated block: slices t0 .. nt-1
			const int gt = d->shape[0] - x.t0;
			ARMA_HIP_REQUIRE(x.t0 >= 0 && gt > 0, ARMA_HIP_EINVAL, "bad history length");
			int3 part{gt, d->shape[1], d->shape[2]};
			if (d->noise_mode == ARMA_NOISE_PARTITIONED) {
				ARMA_HIP_REQUIRE(d->partition[0] > 0 && d->partition[1] > 0 && d->partition[2] > 0,
				                 ARMA_HIP_EINVAL, "bad partition shape");
				part = int3{d->partition[0], d->partition[1], d->partition[2]};
			} else {
				ARMA_HIP_REQUIRE(d->noise_mode == ARMA_NOISE_SEQUENTIAL, ARMA_HIP_EINVAL,
				                 "bad noise mode");
			}
			const int np0 = div_ceil(gt, part.x), np1 = div_ceil(d->shape[1], part.y),
			          np2 = div_ceil(d->shape[2], part.z);
			const int n = np0 * np1 * np2;
			ARMA_HIP_REQUIRE(d->n_mts >= n && d->mts, ARMA_HIP_EPRNG, "bad number of MT configs");
			validate_params(d->mts, n);
			b->small_blocks = b->small_blocks && (long long)part.x * part.y * part.z < (1ll << 31);
			const double sd = std::sqrt(d->varwn);
			const float sdf = std::sqrt(float(d->varwn));
			for (int p = 0; p < n; ++p) {
				const int a = p / (np1 * np2), bb = (p / np2) % np1, c = p % np2;
				NoiseBlockRec r;
				r.base = (long long)(x.zoff + size_t(x.t0) * d->shape[1] * d->shape[2]);
				r.sy = d->shape[1];
				r.sz = d->shape[2];
				r.lo0 = a * part.x;
				r.lo1 = bb * part.y;
				r.lo2 = c * part.z;
				r.e0 = std::min(r.lo0 + part.x, gt) - r.lo0;
				r.e1 = std::min(r.lo1 + part.y, d->shape[1]) - r.lo1;
				r.e2 = std::min(r.lo2 + part.z, d->shape[2]) - r.lo2;
				r.sd = sd;
				r.sdf = sdf;
				recs.push_back(r);
				b->hparams.push_back(d->mts[p]);
				seeds.push_back(d->seeds ? d->seeds[p] : 0u);
			}
		}
		b->nstreams = int(recs.size());
		b->params.get(sizeof(arma_mt_params) * recs.size());
		b->seeds.get(sizeof(uint32_t) * recs.size());
		b->recs.get(sizeof(NoiseBlockRec) * recs.size());
		upload(ctx, b->params.ptr, b->hparams.data(), sizeof(arma_mt_params) * recs.size());
		upload(ctx, b->seeds.ptr, seeds.data(), sizeof(uint32_t) * recs.size());
		upload(ctx, b->recs.ptr, recs.data(), sizeof(NoiseBlockRec) * recs.size());
	}
}

void noise_batch_launch(arma_hip_ctx* ctx, NoiseBatch* b, void* arena) {
	KernelTimer timer(ctx, "noise_batch");
	TableSpec g{(const NoiseBlockRec*)b->recs.ptr};
	launch_by_type(ctx, b->precision == ARMA_FP32, b->sem, b->hparams.data(),
	               (const arma_mt_params*)b->params.ptr, (const uint32_t*)b->seeds.ptr, arena, g,
	               b->nstreams, b->small_blocks);
}

void noise_batch_destroy(NoiseBatch* b) { delete b; }

}  // namespace arma_hip
