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
//  * Halo values come from the neighbouring tiles through HBM.  Each tile
//    keeps a mailbox with one row per step of the values its boundary lanes
//    (the bottom P1-1 rows and right P2-1 columns) computed at that step,
//    stored device-coherently (agent-scope relaxed atomics, which write
//    through the per-XCD L2) as contiguous rows; the surface itself goes out
//    with ordinary write-back stores.  A neighbour computed the value at a
//    halo position at its own step u + 16 (edge) or u + 32 (corner), so a
//    consumer reads a fixed mailbox row per step.  A tile publishes the
//    number of steps whose mailbox rows are complete (each wave's stores are
//    older than the queue loads it has consumed, and one lane stores the
//    count after a barrier), and polls its neighbours' counters (relaxed
//    device-coherent loads, an acquire load every 64 polls) before issuing
//    the halo loads, which are prefetched D steps ahead.  Counters sit on
//    separate cache lines.  Tiles only wait on tiles with smaller workgroup
//    ids, so the launch cannot deadlock.
//  * In-slice taps go in chunks of 16: the chunk's ring values are read from
//    LDS into registers once, and the coefficients (uniform across the
//    workgroup) arrive through the scalar cache in rows of 16 and enter the
//    FMAs as scalar operands.  Per-lane LDS traffic is one value per tap.
//  * FP32 mode keeps every accumulator as a compensated (Kahan/TwoSum) pair;
//    each step adds one P1*P2-term FP32 partial to it.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace arma_hip {
namespace {

// Tiles are TX x TY columns with TX * TY = kThreads: 8 x 32, 16 x 16 or
// 32 x 8 (the halo of a 10 x 10 in-slice order then needs at most two slots
// per lane).
constexpr int kThreads = 256;
constexpr int kD = 4;  // prefetch distance of the noise and halo loads, in steps
constexpr int kMaxP12 = 10;  // largest P1, P2 handled
constexpr int kProgStride = 32;  // ints between tile progress counters: one per 128-byte line

template <class T, bool KAHAN>
struct Acc;

template <class T>
struct Acc<T, false> {
	T s;
	__device__ void zero() { s = T(0); }
	__device__ void add(T v) { s += v; }
	__device__ T value() const { return s; }
	__device__ void pin() { asm volatile("" : "+v"(s)); }
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
	__device__ void pin() { asm volatile("" : "+v"(s), "+v"(c)); }
};

__global__ void reset_progress(int* p, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

// Spin until *p >= need and return the value seen: relaxed device-coherent
// polls, with an acquire load every 64 polls.  With limit > 0 the wait is
// bounded: every 256 polls it gives up (returning need) once *abortw is set,
// or sets *abortw itself when it has spun for more than limit ticks of the
// constant-rate wall clock.  A launch that gives up still runs every tile to
// its end; the host reads *abortw afterwards and reports the launch as
// failed.
__device__ __forceinline__ int wait_progress(const int* p, int need, int* abortw,
                                             long long limit) {
	long long start = 0;
	for (int n = 1;; ++n) {
		const int v = (n & 63) ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
		                       : __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
		if (v >= need) return v;
		if (limit > 0 && (n & 255) == 0) {
			if (__hip_atomic_load(abortw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
				return need;
			const long long now = wall_clock64();
			if (start == 0) {
				start = now;
			} else if (now - start > limit) {
				__hip_atomic_store(abortw, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				return need;
			}
		}
		__builtin_amdgcn_s_sleep(1);
	}
}

template <class T>
__device__ __forceinline__ T coherent_load(const T* p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ void coherent_store(T* p, T v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent loads and stores through a buffer descriptor (the
// instruction form of the agent-scope relaxed atomics above: sc1).  An
// offset past the descriptor's byte range loads zero and stores nothing,
// without a memory access.  Word 3 of the descriptor: raw 32-bit data for
// gfx9-class buffer instructions.
constexpr int kBufferDword3 = 0x00020000;
constexpr int kCoherentSc1 = 16;
typedef unsigned int uint2v __attribute__((ext_vector_type(2)));

template <class T>
__device__ __forceinline__ T buffer_load_coherent(__amdgpu_buffer_rsrc_t r, unsigned off) {
	if constexpr (sizeof(T) == 8)
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kCoherentSc1));
	else
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kCoherentSc1));
}

template <class T>
__device__ __forceinline__ void buffer_store_coherent(__amdgpu_buffer_rsrc_t r, unsigned off, T v) {
	if constexpr (sizeof(T) == 8)
		__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(uint2v, v), r, off, 0, kCoherentSc1);
	else
		__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0,
		                                      kCoherentSc1);
}

// Checked launches (ARMA_HIP_CHECK_HALO set) fill every mailbox with this
// signalling NaN before the launch; arithmetic only produces quiet NaNs, so a
// consumer that reads it for a slice in range has read a mailbox row before
// its producer stored it, and raises the check word.
constexpr unsigned long long kHaloSentinel64 = 0x7ff4a5a5a5a5a5a5ull;
constexpr unsigned kHaloSentinel32 = 0x7fa5a5a5u;

__device__ __forceinline__ bool is_sentinel(double v) {
	return (unsigned long long)__double_as_longlong(v) == kHaloSentinel64;
}
__device__ __forceinline__ bool is_sentinel(float v) { return __float_as_uint(v) == kHaloSentinel32; }

template <class T>
__global__ void fill_sentinel(T* p, size_t n) {
	for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
	     i += size_t(gridDim.x) * blockDim.x) {
		if constexpr (sizeof(T) == 8)
			p[i] = __longlong_as_double((long long)kHaloSentinel64);
		else
			p[i] = __uint_as_float(kHaloSentinel32);
	}
}

// A consumed halo value of a slice in range: flag it when it is the sentinel.
template <class T>
__device__ __forceinline__ void check_halo(int* check, T v) {
	if (is_sentinel(v)) __hip_atomic_store(check, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int pmod(int a, int r) {
	const int m = a % r;
	return m < 0 ? m + r : m;
}

__device__ __forceinline__ int wrap_inc(int v, int r) { return v + 1 == r ? 0 : v + 1; }

// In-slice taps per chunk of the tap loop.  A chunk's kCH ring values are
// read from LDS into registers once; its coefficients come in k-rows of kCH
// through the scalar cache and enter the FMAs as scalar operands, so the
// per-lane LDS traffic is one value per tap.
constexpr int kCH = 16;

__host__ __device__ constexpr int tap_chunks(int p12) { return (p12 - 1 + kCH - 1) / kCH; }

// Mailbox column of tile-local lane (x, y), or -1 when no neighbour reads it:
// the bottom HX rows first (row-major), then the right HY columns of the
// rows above them.
template <int TX>
__host__ __device__ __forceinline__ int mbox_column(int x, int y, int HX, int HY) {
	constexpr int TY = kThreads / TX;
	if (x >= TX - HX) return (x - (TX - HX)) * TY + y;
	if (y >= TY - HY) return TY * HX + x * HY + (y - (TY - HY));
	return -1;
}

template <int TX>
__host__ __device__ constexpr int mbox_width(int HX, int HY) {
	return (kThreads / TX) * HX + (TX - HX) * HY;
}

inline int mbox_width_rt(int TX, int HX, int HY) { return (kThreads / TX) * HX + (TX - HX) * HY; }

// The march of one tile over all of t (P0 accumulators; a surface of lower
// order in t has zero coefficients past its own order).
template <class T, int TX, int P0, bool KAHAN, bool BUF = false, int KD = kD, int STG = 0>
__device__ __forceinline__ void wave_tile(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ phi0,
	const T* __restrict__ phic, const int* __restrict__ toff, int P1, int P2, int nY,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, int NB, int total_steps, int t0, int b, int* abortw, long long limit,
	int* check, unsigned char* smem_raw) {
	constexpr int kTX = TX, kTY = kThreads / TX;
	static_assert(kTX * kTY == kThreads && kTX >= 8 && kTY >= 8, "tile shape");
	const int P12 = P1 * P2;
	T* ring = reinterpret_cast<T*>(smem_raw);
	const int HX = P1 - 1, HY = P2 - 1;
	const int WX = kTX + HX, WY = kTY + HY;
	const int R = P1 + P2 - 1;
	const int slot = WX * WY;
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
	int hnb[2], hm[2], hdu[2];  // producing tile, its mailbox column, step offset
	bool hval[2], hin[2];
#pragma unroll
	for (int r = 0; r < 2; ++r) {
		const int h = tid + r * kThreads;
		int hx = 0, hy = 0;
		hin[r] = false;
		hval[r] = false;
		hsk[r] = 0;
		hlds[r] = 0;
		hnb[r] = 0;
		hm[r] = 0;
		hdu[r] = 0;
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
			// the producing tile computed this position at its lane (px, py),
			// at u-step u + 16 (edge neighbours) or u + 32 (corner)
			const int px = hx < 0 ? kTX + hx : hx, py = hy < 0 ? kTY + hy : hy;
			hnb[r] = b - (hx < 0 ? nY : 0) - (hy < 0 ? 1 : 0);
			hm[r] = mbox_column<TX>(px, py, HX, HY);
			hdu[r] = (hx < 0 ? kTX : 0) + (hy < 0 ? kTY : 0);
		}
	}
	// Neighbour tiles.  The corner tile needs no poll: the top tile had
	// observed it far enough ahead before publishing the progress we wait for.
	const int nb_top = X > 0 ? b - nY : -1;
	const int nb_left = Y > 0 ? b - 1 : -1;
	int* const my_prog = progress + size_t(b) * kProgStride;
	const int* const top_prog = progress + size_t(max(nb_top, 0)) * kProgStride;
	const int* const left_prog = progress + size_t(max(nb_left, 0)) * kProgStride;

	Acc<T, KAHAN> acc[P0];
#pragma unroll
	for (int k = 0; k < P0; ++k) acc[k].zero();
	const T phi000 = phi0[0];

	// Step u of the loop is skew step s = u + s0; s0 < 0 lets the halo
	// positions (negative skew hx + hy) deliver their first slices.  Progress
	// counts u-steps whose surface stores are complete; every tile uses the
	// same s0.
	const int s0 = -(HX + HY);
	const int total_u = total_steps - s0;
	// halo loads consumed at u-step u read points the neighbours computed at
	// their u-steps up to u + kTX (top) and u + kTY (left)
	auto need_top = [&](int u) { return min(u + kTX + 1, total_u); };
	auto need_left = [&](int u) { return min(u + kTY + 1, total_u); };

	// the ring starts zeroed: halo positions outside the grid stay zero
	for (int q = tid; q < R * slot; q += kThreads) ring[q] = T(0);
	// ring slots (slice t lives in slot t mod R), advanced every step
	int rs_own = pmod(s0 - lx - ly, R);
	int rs_h0 = pmod(s0 - hsk[0], R);
	int rs_h1 = pmod(s0 - hsk[1], R);
	// STG > 0: the surface stores of a tile row go to one slice per step, and
	// with STG = 2 the noise loads too.  Lane (lx, ly) computes slice t at
	// step t + lx + ly; its value is stored at step t + lx + kTY and (STG = 2)
	// its noise loaded for step t + lx, as for ly = 0, through a line of
	// kTY + 1 slots of its own in LDS after the ring: slot t mod (kTY + 1)
	// receives the noise at step t + lx, gives it up and takes the value at
	// step t + lx + ly, and gives that up at step t + lx + kTY, before the
	// noise of slice t + kTY + 1 arrives.  A wave's store or load then covers
	// whole row segments of four slices instead of touching 64 cache lines of
	// as many slices.  The line starts zeroed (the noise of slices before the
	// first is zero).
	constexpr int kDL = kTY + 1;
	T* const dline = ring + R * slot;
	int dl_d = pmod(s0 - lx, kDL);        // slot whose noise arrives at this step
	int dl_w = pmod(s0 - lx - ly, kDL);   // slot of the slice computed at this step
	int dl_r = pmod(s0 - lx - kTY, kDL);  // slot of the slice stored at this step
	if constexpr (STG > 0)
		for (int q = 0; q < kDL; ++q) dline[q * kThreads + tid] = T(0);

	// Every lane issues the same global loads and stores at every step
	// (out-of-range loads read a zero, out-of-range stores go to this tile's
	// scratch sink), so the vector memory counter waits for exactly the queue
	// slot being consumed and no load result needs a select at issue.
	T* const my_sink = sink + size_t(b) * kThreads + tid;
	// Boundary lanes (the bottom HX rows and the right HY columns, which the
	// tiles below and to the right read as halo) also store every step's
	// value into this tile's mailbox, row u, column my_m: whole rows per
	// step, so the device-coherent stores are contiguous.  The surface
	// itself is stored with ordinary (write-back) stores.
	const int my_m = mbox_column<TX>(lx, ly, HX, HY);
	T* const my_mbox = my_m >= 0 ? mbox + size_t(b) * total_u * NB + my_m : my_sink;
	const size_t mstep = my_m >= 0 ? size_t(NB) : 0;
	// BUF: the mailbox traffic goes through one buffer descriptor spanning the
	// mailboxes from the corner neighbour's to this tile's (the host chooses
	// it when that span fits 31 bits), as in quad_tile: a lane with no value
	// to store or load passes the offset just past the end, which the
	// hardware drops without a memory access.  Without it such lanes store
	// to their sink slot and load the zero word, device-coherently, every
	// step.
	const int bb0 = max(b - nY - 1, 0);
	const size_t tile_mb = size_t(total_u) * NB;  // elements of one tile's mailbox
	const unsigned mb_end = unsigned((size_t(b - bb0) + 1) * tile_mb * sizeof(T));
	const __amdgpu_buffer_rsrc_t mrsrc = __builtin_amdgcn_make_buffer_rsrc(
		mbox + size_t(bb0) * tile_mb, 0, int(mb_end), kBufferDword3);
	const unsigned moff0 =
		my_m >= 0 ? unsigned((size_t(b - bb0) * tile_mb + my_m) * sizeof(T)) : mb_end;
	const unsigned mstep_b = my_m >= 0 ? unsigned(NB * sizeof(T)) : 0u;
	auto halo_off = [&](int r, int u2) {
		return unsigned(((size_t(hnb[r] - bb0) * total_u + u2 + hdu[r]) * NB + hm[r]) * sizeof(T));
	};
	auto issue_loads = [&](int u2, T& e, T& h0, T& h1) {
		const int s2 = u2 + s0;
		const bool in = u2 < total_u;
		const int t = STG == 2 ? s2 - lx : s2 - lx - ly;  // row-aligned noise
		const bool ve = in && active && t >= 0 && t < nt;
		const int t0 = s2 - hsk[0];
		const bool v0 = in && hin[0] && t0 >= 0 && t0 < nt;
		const int t1 = s2 - hsk[1];
		const bool v1 = in && hin[1] && t1 >= 0 && t1 < nt;
		e = *(ve ? zeta + size_t(t) * sxy + own_off : zero);
		if constexpr (BUF) {
			h0 = buffer_load_coherent<T>(mrsrc, v0 ? halo_off(0, u2) : mb_end);
			h1 = buffer_load_coherent<T>(mrsrc, v1 ? halo_off(1, u2) : mb_end);
		} else {
			h0 = coherent_load(v0 ? mbox + (size_t(hnb[0]) * total_u + u2 + hdu[0]) * NB + hm[0]
			                      : zero);
			h1 = coherent_load(v1 ? mbox + (size_t(hnb[1]) * total_u + u2 + hdu[1]) * NB + hm[1]
			                      : zero);
		}
	};

	// neighbour progress as last observed by lane 0, and the flag loads in
	// flight (read KD steps after they are issued)
	int known_t = total_u, known_l = total_u;
	if (tid == 0) {
		known_t = nb_top >= 0 ? wait_progress(top_prog, need_top(KD - 1), abortw, limit) : total_u;
		known_l =
			nb_left >= 0 ? wait_progress(left_prog, need_left(KD - 1), abortw, limit) : total_u;
	}
	__syncthreads();
	T qe[KD], qh0[KD], qh1[KD];
	int qt[KD], ql[KD];
#pragma unroll
	for (int d = 0; d < KD; ++d) {
		issue_loads(d, qe[d], qh0[d], qh1[d]);
		qt[d] = known_t;
		ql[d] = known_l;
	}

	auto step = [&](int u, T& e_q, T& h0_q, T& h1_q, int& t_q, int& l_q) {
		const int s = u + s0;
		// 1. halo values of step s into the ring
		{
			const int t0 = s - hsk[0];
			if (hval[0] && t0 >= 0 && t0 < nt) {
				ring[rs_h0 * slot + hlds[0]] = h0_q;
				if (check && hin[0]) check_halo(check, h0_q);
			}
			const int t1 = s - hsk[1];
			if (hval[1] && t1 >= 0 && t1 < nt) {
				ring[rs_h1 * slot + hlds[1]] = h1_q;
				if (check && hin[1]) check_halo(check, h1_q);
			}
		}
		// 2. own point, slice t = s - lx - ly (computed on every lane; only
		//    live lanes write the ring and the surface)
		const int t = s - lx - ly;
		const bool live = active && t >= 0 && t < nt;
		{
			const T* base = ring + rs_own * slot + own_lds;
			T part[P0];
#pragma unroll
			for (int k = 0; k < P0; ++k) part[k] = T(0);
			// taps m = 1 .. P12-1 in chunks of kCH; toff holds each tap's
			// LDS byte distance back from the own position (taps past the
			// last one have zero coefficients and read the own position)
			const T* pc = phic;
			const int* oc = toff;
			const char* bb = reinterpret_cast<const char*>(base);
			for (int c0 = 1; c0 < P12; c0 += kCH, pc += P0 * kCH, oc += kCH) {
				T v[kCH];
#pragma unroll
				for (int q = 0; q < kCH; ++q) v[q] = *reinterpret_cast<const T*>(bb - oc[q]);
#pragma unroll
				for (int q = 0; q < kCH; ++q) {
#pragma unroll
					for (int k = 0; k < P0; ++k) part[k] = fma(pc[q * P0 + k], v[q], part[k]);
				}
			}
			acc[0].add(part[0]);
			// slices t < t0 are given (a realtime stream's history): they
			// feed the sums of later slices unchanged
			T eps = e_q;
			T zd = T(0);
			// the value due for storage; with STG = 2 the noise arriving
			// (slice s - lx) into its slot, then this lane's noise (slice t)
			if constexpr (STG > 0) zd = dline[dl_r * kThreads + tid];
			if constexpr (STG == 2) {
				dline[dl_d * kThreads + tid] = e_q;
				eps = dline[dl_w * kThreads + tid];
			}
			const T zc = eps + (acc[0].value() + phi000 * eps);
			const T z = t < t0 ? eps : zc;
			if (live) ring[rs_own * slot + own_lds] = z;
			if constexpr (STG > 0) {
				const int ts = s - lx - kTY;
				const bool st = active && ts >= 0 && ts < nt;
				*(st ? zeta + size_t(ts) * sxy + own_off : my_sink) = zd;
				dline[dl_w * kThreads + tid] = z;
			} else {
				*(live ? zeta + size_t(t) * sxy + own_off : my_sink) = z;
			}
			if constexpr (BUF)
				buffer_store_coherent(mrsrc, moff0 + unsigned(u) * mstep_b, z);
			else
				coherent_store(my_mbox + size_t(u) * mstep, z);
#pragma unroll
			for (int k = 1; k < P0; ++k) acc[k].add(fma(phi0[k], z, part[k]));
#pragma unroll
			for (int k = 0; k + 1 < P0; ++k) acc[k] = acc[k + 1];
			acc[P0 - 1].zero();
		}
		rs_own = wrap_inc(rs_own, R);
		rs_h0 = wrap_inc(rs_h0, R);
		rs_h1 = wrap_inc(rs_h1, R);
		if constexpr (STG > 0) {
			dl_d = wrap_inc(dl_d, kDL);
			dl_w = wrap_inc(dl_w, kDL);
			dl_r = wrap_inc(dl_r, kDL);
		}
		// 3. lane 0: may the loads of step u + KD go out?
		const int u2 = u + KD;
		if (tid == 0) {
			known_t = max(known_t, t_q);
			known_l = max(known_l, l_q);
			if (u2 < total_u) {
				if (known_t < need_top(u2))
					known_t = wait_progress(top_prog, need_top(u2), abortw, limit);
				if (known_l < need_left(u2))
					known_l = wait_progress(left_prog, need_left(u2), abortw, limit);
			}
			t_q = nb_top >= 0 ? coherent_load(top_prog) : total_u;
			l_q = nb_left >= 0 ? coherent_load(left_prog) : total_u;
		}
		__syncthreads();
		// 4. publish: the stores of step u - KD were issued before the queue
		//    loads consumed at this step, so every wave has seen them complete
		if (tid == 0 && u >= KD) coherent_store(my_prog, u - KD + 1);
		// 5. loads for step u + KD into the queue slot just consumed
		issue_loads(u2, e_q, h0_q, h1_q);
	};

	int u = 0;
	for (; u + KD <= total_u; u += KD) {
#pragma unroll
		for (int d = 0; d < KD; ++d) step(u + d, qe[d], qh0[d], qh1[d], qt[d], ql[d]);
	}
	// the last steps (fewer than KD), each on its own queue slot; the
	// condition is the same on every lane
#pragma unroll
	for (int d = 0; d < KD; ++d)
		if (u + d < total_u) step(u + d, qe[d], qh0[d], qh1[d], qt[d], ql[d]);
	// STG: the stores due at step total_steps (slice nt - 1 of the tile's
	// last row; every other row's last slice was stored inside the loop)
	if constexpr (STG > 0) {
		const int ts = total_steps - lx - kTY;
		if (active && ts >= 0 && ts < nt) zeta[size_t(ts) * sxy + own_off] = dline[dl_r * kThreads + tid];
	}
	// every store complete, then the final count
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	if (tid == 0) coherent_store(my_prog, total_u);
}

// ---- four lanes per column ----
//
// With one lane per column (wave_tile) every step streams all P0*P1*P2
// coefficients through the scalar cache, and a tile is one wave per SIMD:
// the step is a chain of scalar-load and LDS waits with little to overlap.
// For in-slice orders of at most 16 taps quad_tile gives each column a quad
// of four adjacent lanes.  Lane g of the quad owns the in-slice taps
// m = g + 4n (n < NT), keeps their coefficients for every k in registers and
// keeps its own partial accumulators; the quad's partial sums of the point
// being completed are added with two DPP butterfly stages, in a fixed lane
// order, so the value of a point does not depend on the tile shape or the
// schedule.  A tile of 256 columns is 1024 lanes, four waves per SIMD.
// Tap m = 0 (i = j = 0) sits on lane 0 of the quad: its k = 0 coefficient is
// zero in the table (phi(0,0,0) multiplies the noise of the completed point
// instead), and its k >= 1 terms take the completed value itself.
// Lane 0 of a quad loads the noise and stores the surface, lane 1 stores
// the mailbox, lanes 1 and 2 carry one halo slot each; the ring, mailbox,
// progress protocol and halo geometry are those of wave_tile.
// Measured on MI355X: one tile alone steps in 0.78 us against 1.13 us for
// wave_tile (1024 x 16 x 16, phi (7,7,2), FP64), but a full grid of tiles is
// set by the inter-tile fill and the memory traffic of 1024 lanes: wave_tile
// is faster on full grids (2048 x 512 x 512 FP32: 7.5 ms with its per-lane
// line; FP64 1024 x 256 x 256: 2.2 against 3.0 ms), so the quad tile runs
// only when ARMA_HIP_QUAD_TILE=1 asks for it (FP32 surfaces).
constexpr int kQuad = 4;
constexpr int kQThreads = kThreads * kQuad;
constexpr int kQD = 4;  // prefetch distance of the queue loads, in steps

// quad_perm DPP controls: lanes (1,0,3,2), lanes (2,3,0,1), all from lane 0
constexpr int kDppXor1 = 0xb1, kDppXor2 = 0x4e, kDppLane0 = 0x00;

template <int CTRL>
__device__ __forceinline__ double quad_dpp(double v) {
	const int lo = __double2loint(v), hi = __double2hiint(v);
	return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, false),
	                        __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ float quad_dpp(float v) {
	return __int_as_float(
		__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum over the quad, the same bits on all four lanes (each stage adds two
// values in either order, which rounds the same).
template <class T>
__device__ __forceinline__ T quad_sum(T s) {
	s = s + quad_dpp<kDppXor1>(s);
	return s + quad_dpp<kDppXor2>(s);
}

// Compensated sum over the quad of the pairs (s, c): TwoSum of the leading
// parts (its error term is exact, so either operand order gives the same
// bits), the compensations and the error added to c.
template <class T>
__device__ __forceinline__ T quad_sum2(T s, T c) {
	auto stage = [&](T s2, T c2) {
		const T r = s + s2;
		const T bb = r - s;
		const T e = (s - (r - bb)) + (s2 - bb);
		c = (c + c2) + e;
		s = r;
	};
	stage(quad_dpp<kDppXor1>(s), quad_dpp<kDppXor1>(c));
	stage(quad_dpp<kDppXor2>(s), quad_dpp<kDppXor2>(c));
	return s + c;
}

// Taps per lane of the quad tile for an in-slice order of P12 taps, 0 when
// the tile does not take it.
__host__ __device__ constexpr int quad_taps(int p12) { return p12 <= 8 ? 2 : p12 <= 16 ? 4 : 0; }

template <class T, int TX, int P0, int NT, bool KAHAN>
__device__ __forceinline__ void quad_tile(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ qcoef,
	const int* __restrict__ qoff, int P1, int P2, int nY, int* __restrict__ progress,
	const T* __restrict__ zero, T* __restrict__ sink, T* __restrict__ mbox, int NB,
	int total_steps, int t0, int b, int* abortw, long long limit, int* check,
	unsigned char* smem_raw) {
	constexpr int kTX = TX, kTY = kThreads / TX;
	static_assert(kTX * kTY == kThreads && kTX >= 8 && kTY >= 8, "tile shape");
	T* ring = reinterpret_cast<T*>(smem_raw);
	const int HX = P1 - 1, HY = P2 - 1;
	const int WX = kTX + HX, WY = kTY + HY;
	const int R = P1 + P2 - 1;
	const int slot = WX * WY;
	const int X = b / nY, Y = b - (b / nY) * nY;
	const int X0 = X * kTX, Y0 = Y * kTY;
	const int tid = threadIdx.x;
	const int g = tid & (kQuad - 1), col = tid / kQuad;
	const int lx = col / kTY, ly = col - (col / kTY) * kTY;
	const bool active = (X0 + lx < nx) && (Y0 + ly < ny);
	const size_t sxy = size_t(nx) * ny;
	const size_t own_off = size_t(X0 + lx) * ny + (Y0 + ly);
	const int own_lds = (lx + HX) * WY + (ly + HY);

	// the halo slot of quad lanes 1 and 2 (top band, then left band; at most
	// 2 * 256 positions for orders up to 10)
	const int H = WX * WY - kThreads;
	const int h = g == 1 ? col : (g == 2 ? col + kThreads : H);
	const bool hval = h < H;
	int hsk = 0, hlds = 0, hnb = 0, hm = 0, hdu = 0;
	bool hin = false;
	if (hval) {
		int hx, hy;
		if (h < HX * WY) {
			hx = -HX + h / WY;
			hy = -HY + (h - (h / WY) * WY);
		} else {
			const int q = h - HX * WY;
			hx = q / HY;
			hy = -HY + (q - (q / HY) * HY);
		}
		hsk = hx + hy;
		hlds = (hx + HX) * WY + (hy + HY);
		const int gx = X0 + hx, gy = Y0 + hy;
		hin = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
		const int px = hx < 0 ? kTX + hx : hx, py = hy < 0 ? kTY + hy : hy;
		hnb = b - (hx < 0 ? nY : 0) - (hy < 0 ? 1 : 0);
		hm = mbox_column<TX>(px, py, HX, HY);
		hdu = (hx < 0 ? kTX : 0) + (hy < 0 ? kTY : 0);
	}
	const int nb_top = X > 0 ? b - nY : -1;
	const int nb_left = Y > 0 ? b - 1 : -1;
	int* const my_prog = progress + size_t(b) * kProgStride;
	const int* const top_prog = progress + size_t(max(nb_top, 0)) * kProgStride;
	const int* const left_prog = progress + size_t(max(nb_left, 0)) * kProgStride;

	// this lane's taps: coefficients [n][k] and LDS byte distances
	T c[NT][P0];
	int toff[NT];
#pragma unroll
	for (int n = 0; n < NT; ++n) {
		toff[n] = qoff[g * NT + n];
#pragma unroll
		for (int k = 0; k < P0; ++k) c[n][k] = qcoef[(g * NT + n) * P0 + k];
	}
	const T phi000 = qcoef[kQuad * NT * P0];

	Acc<T, KAHAN> acc[P0];
#pragma unroll
	for (int k = 0; k < P0; ++k) acc[k].zero();

	const int s0 = -(HX + HY);
	const int total_u = total_steps - s0;
	auto need_top = [&](int u) { return min(u + kTX + 1, total_u); };
	auto need_left = [&](int u) { return min(u + kTY + 1, total_u); };

	for (int q = tid; q < R * slot; q += kQThreads) ring[q] = T(0);
	int rs_own = pmod(s0 - lx - ly, R);
	int rs_h = pmod(s0 - hsk, R);

	const int my_m = mbox_column<TX>(lx, ly, HX, HY);
	// Memory traffic per lane and step, the same four instructions on every
	// lane (so that the compiler's vector-memory counter waits exactly for
	// the queue slot consumed):
	//  * quad lane 0 loads the noise and stores the surface with ordinary
	//    accesses; the other lanes load and store their own slot of the
	//    tile's sink, which stays in the L2;
	//  * quad lanes 1 and 2 load a halo value and lane 1 of a boundary
	//    column stores the mailbox, device-coherently, through one buffer
	//    descriptor spanning the mailboxes from the corner neighbour's to this
	//    tile's; every other lane passes the offset just past its end, which
	//    the hardware drops without a memory access.
	T* const my_sink = sink + size_t(b) * kQThreads + tid;
	const int bb0 = max(b - nY - 1, 0);
	const size_t tile_mb = size_t(total_u) * NB;  // elements of one tile's mailbox
	const unsigned mb_end = unsigned((size_t(b - bb0) + 1) * tile_mb * sizeof(T));
	const __amdgpu_buffer_rsrc_t mrsrc = __builtin_amdgcn_make_buffer_rsrc(
		mbox + size_t(bb0) * tile_mb, 0, int(mb_end), kBufferDword3);
	const bool mstore = g == 1 && my_m >= 0;
	unsigned moff = mstore ? unsigned((size_t(b - bb0) * tile_mb + my_m) * sizeof(T)) : mb_end;
	const unsigned mstep = mstore ? unsigned(NB * sizeof(T)) : 0u;
	// the noise of slice te and the halo of slice th_ld are the next to load;
	// both advance by one slice per step (a slice in range implies a step in
	// range)
	const bool eps_lane = g == 0 && active;
	int te = s0 - lx - ly, th_ld = s0 - hsk;
	T* pe = zeta + own_off + ptrdiff_t(te) * ptrdiff_t(sxy);
	unsigned hoff = hin ? unsigned(((size_t(hnb - bb0) * total_u + hdu) * NB + hm) * sizeof(T)) : 0u;
	const unsigned hstep = unsigned(NB * sizeof(T));
	auto issue_load = [=, &te, &th_ld, &pe, &hoff](T& qe, T& qh) {
		const bool ve = eps_lane && unsigned(te) < unsigned(nt);
		const bool vh = hin && unsigned(th_ld) < unsigned(nt);
		qe = *(ve ? pe : my_sink);
		qh = buffer_load_coherent<T>(mrsrc, vh ? hoff : mb_end);
		++te;
		++th_ld;
		pe += sxy;
		hoff += hstep;
	};

	int known_t = total_u, known_l = total_u;
	if (tid == 0) {
		known_t = nb_top >= 0 ? wait_progress(top_prog, need_top(kQD - 1), abortw, limit) : total_u;
		known_l =
			nb_left >= 0 ? wait_progress(left_prog, need_left(kQD - 1), abortw, limit) : total_u;
	}
	__syncthreads();
	T qe[kQD], qh[kQD];
	int qt[kQD], ql[kQD];  // neighbour progress loaded kQD steps earlier (lane 0)
#pragma unroll
	for (int d = 0; d < kQD; ++d) {
		issue_load(qe[d], qh[d]);
		qt[d] = known_t;
		ql[d] = known_l;
	}

	auto step = [&](int u, T& qe_q, T& qh_q, int& t_q, int& l_q) {
		const int s = u + s0;
		// 1. the halo value of step s into the ring
		{
			const int th = s - hsk;
			if (hval && th >= 0 && th < nt) {
				ring[rs_h * slot + hlds] = qh_q;
				if (check && hin) check_halo(check, qh_q);
			}
		}
		// 2. own point, slice t = s - lx - ly
		const int t = s - lx - ly;
		const bool live = active && t >= 0 && t < nt;
		const char* bb = reinterpret_cast<const char*>(ring + rs_own * slot + own_lds);
		T v[NT];
#pragma unroll
		for (int n = 0; n < NT; ++n) v[n] = *reinterpret_cast<const T*>(bb - toff[n]);
		// zero outside the column's slices (what the idle loads returned is
		// whatever the sink held), so that the steps before a column's first
		// slice push exact zeros into its accumulators
		const T eps_in = quad_dpp<kDppLane0>(qe_q);
		const T eps = live ? eps_in : T(0);
		T sum;
		if constexpr (KAHAN) {
			T p = T(0);
#pragma unroll
			for (int n = 0; n < NT; ++n) p = fma(c[n][0], v[n], p);
			acc[0].add(p);
			sum = quad_sum2(acc[0].s, acc[0].c);
		} else {
#pragma unroll
			for (int n = 0; n < NT; ++n) acc[0].s = fma(c[n][0], v[n], acc[0].s);
			sum = quad_sum(acc[0].s);
		}
		// slices t < t0 are given: they feed later sums unchanged
		const T zc = eps + (sum + phi000 * eps);
		const T z = t < t0 ? eps : zc;
		const bool zlive = g == 0 && live;
		if (zlive) ring[rs_own * slot + own_lds] = z;
		// (the noise of slice t + kQD is the next to load)
		*(zlive ? pe - ptrdiff_t(kQD) * ptrdiff_t(sxy) : my_sink) = z;
		buffer_store_coherent(mrsrc, moff, z);
		moff += mstep;
		// the terms of points t+1 .. t+P0-1 (tap 0 of quad lane 0 is the
		// point itself)
		const T v0 = g == 0 ? z : v[0];
#pragma unroll
		for (int k = 1; k < P0; ++k) {
			if constexpr (KAHAN) {
				T p = T(0);
#pragma unroll
				for (int n = 1; n < NT; ++n) p = fma(c[n][k], v[n], p);
				acc[k].add(fma(c[0][k], v0, p));
			} else {
#pragma unroll
				for (int n = 1; n < NT; ++n) acc[k].s = fma(c[n][k], v[n], acc[k].s);
				acc[k].s = fma(c[0][k], v0, acc[k].s);
			}
		}
#pragma unroll
		for (int k = 0; k + 1 < P0; ++k) acc[k] = acc[k + 1];
		acc[P0 - 1].zero();
		rs_own = wrap_inc(rs_own, R);
		rs_h = wrap_inc(rs_h, R);
		// keep this step's sums in this step: left free, the compiler defers
		// the k >= 1 terms to the step that completes their point, holding
		// the ring values of up to P0 - 1 steps in registers
#pragma unroll
		for (int k = 0; k < P0; ++k) acc[k].pin();
		// 3. lane 0: may the loads of step u + kQD go out?
		const int u2 = u + kQD;
		if (tid == 0) {
			known_t = max(known_t, t_q);
			known_l = max(known_l, l_q);
			if (u2 < total_u) {
				if (known_t < need_top(u2))
					known_t = wait_progress(top_prog, need_top(u2), abortw, limit);
				if (known_l < need_left(u2))
					known_l = wait_progress(left_prog, need_left(u2), abortw, limit);
			}
			t_q = nb_top >= 0 ? coherent_load(top_prog) : total_u;
			l_q = nb_left >= 0 ? coherent_load(left_prog) : total_u;
		}
		__syncthreads();
		// 4. publish: every wave consumed at this step a load issued after
		//    its stores of step u - kQD
		if (tid == 0 && u >= kQD) coherent_store(my_prog, u - kQD + 1);
		// 5. the loads of step u + kQD into the queue slot just consumed
		issue_load(qe_q, qh_q);
	};

	int u = 0;
	for (; u + kQD <= total_u; u += kQD) {
#pragma unroll
		for (int d = 0; d < kQD; ++d) step(u + d, qe[d], qh[d], qt[d], ql[d]);
	}
#pragma unroll
	for (int d = 0; d < kQD; ++d)
		if (u + d < total_u) step(u + d, qe[d], qh[d], qt[d], ql[d]);
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	if (tid == 0) coherent_store(my_prog, total_u);
}

// Kernel arguments stay plain restrict-qualified pointers: the coefficient
// reads are then provably unclobbered by the kernel's stores and compile to
// scalar loads (a struct argument costs 30-60 VGPRs per lane).
template <class T, int TX, int P0, bool KAHAN, bool BUF, int KD, int STG = 0>
__global__ __launch_bounds__(kThreads) void ar_wave_kernel(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ phi0,
	const T* __restrict__ phic, const int* __restrict__ toff, int P1, int P2, int nY,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, int NB, int total_steps, int* check) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	// unbounded waits: tiles wait only on tiles with smaller workgroup ids,
	// which were dispatched before them
	wave_tile<T, TX, P0, KAHAN, BUF, KD, STG>(zeta, nt, nx, ny, phi0, phic, toff, P1, P2, nY,
	                                          progress, zero, sink, mbox, NB, total_steps, 0,
	                                          blockIdx.x, nullptr, 0, check, smem_raw);
}

template <class T, int P0, int NT, bool KAHAN>
__global__ __launch_bounds__(kQThreads) void ar_quad_kernel(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ qcoef,
	const int* __restrict__ qoff, int P1, int P2, int nY, int* __restrict__ progress,
	const T* __restrict__ zero, T* __restrict__ sink, T* __restrict__ mbox, int NB,
	int total_steps, int* check) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	quad_tile<T, 16, P0, NT, KAHAN>(zeta, nt, nx, ny, qcoef, qoff, P1, P2, nY, progress, zero,
	                                sink, mbox, NB, total_steps, 0, blockIdx.x, nullptr, 0,
	                                check, smem_raw);
}

// ---- many surfaces in one persistent launch ----
//
// Each surface keeps the layout of a single-surface launch (its own tiles,
// progress counters and mailboxes, at offsets in shared arenas).  A queue
// lists (surface, tile) pairs in an order where every tile comes after the
// tiles it waits on; each workgroup marches its share of the entries in
// queue order.  The coefficients of a surface are padded with
// zeros in k to the nearest compiled P0; the padded terms add exact zeros, so
// a surface's values do not depend on the padding, the tile shape or the
// schedule.

struct BatchSurface {
	long long zoff;  // first element of the surface in the arena
	long long moff;  // first element of its mailboxes
	int nt, nx, ny;
	int P1, P2;
	int bucket;  // index into kBatchP0
	int nY;      // tiles along y
	int coff;    // coefficient block (elements): phi0 block, then the chunks
	int n0;      // length of the phi0 block
	int toff;    // tap offset table (ints)
	int poff;    // first progress counter (tiles)
	int NB, total_steps;
	int t0;  // leading slices that are given, not generated
	int pad;
};

constexpr int kBatchP0[] = {2, 3, 5, 7, 9, 10};
constexpr int kBatchBuckets = 6;

int batch_bucket(int P0) {
	for (int i = 0; i < kBatchBuckets; ++i)
		if (P0 <= kBatchP0[i]) return i;
	return -1;
}

// 32-bit registers a lane of the quad tile may spend on its coefficients
// (four waves per SIMD leave 128 per lane in all).
constexpr int kQuadCoefRegs = 56;

__host__ __device__ constexpr bool quad_fits(int es, int P0, int nt) {
	return nt > 0 && P0 * nt * (es / 4) <= kQuadCoefRegs;
}

// Whether a surface runs on the quad tile: at most 16 in-slice taps, the
// coefficients of its batch bucket's P0 within the register budget, and, for
// every tile shape, the mailboxes from a tile's corner neighbour to itself
// within the reach of one buffer descriptor.  Single surfaces and batch
// members follow the same rule, so a batch member's values equal those of
// the same surface alone.
bool quad_eligible(const arma_ar_desc* d) {
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	const int b = batch_bucket(P0);
	const size_t es = d->precision == ARMA_FP32 ? 4 : 8;
	// The one-lane tile is the default in both precisions: with its mailbox
	// traffic behind a buffer descriptor it ran 2.38 ms against the quad
	// tile's 2.68 ms on 1024x256x256 FP32 and 16.7 against 17.2 ms on
	// 2048x512x512 FP32 (whose 262 144 columns fit at one lane each, four
	// workgroups per CU, but at four lanes each only a quarter at a time),
	// and 2.44 against 3.0 ms on 1024x256x256 FP64.  ARMA_HIP_QUAD_TILE=1
	// selects the quad tile for FP32 surfaces, single and batched alike (a
	// point's sum is added in another order there, so the choice must be the
	// same for a batch member and the single surface it is compared with).
	const char* e = std::getenv("ARMA_HIP_QUAD_TILE");
	if (!(e && e[0] == '1' && e[1] == 0)) return false;
	if (d->precision != ARMA_FP32) return false;
	if (b < 0 || P1 > kMaxP12 || P2 > kMaxP12 ||
	    !quad_fits(int(es), kBatchP0[b], quad_taps(P1 * P2)))
		return false;
	for (int TX : {8, 16, 32}) {
		const int TY = kThreads / TX;
		const size_t nY = size_t(div_ceil(d->shape[2], TY));
		const size_t total_u = size_t(d->shape[0]) + TX + TY - 2 + (P1 - 1) + (P2 - 1);
		const size_t NB = size_t(std::max(1, mbox_width_rt(TX, P1 - 1, P2 - 1)));
		if ((nY + 2) * total_u * NB * es >= (size_t(1) << 31) - 64) return false;
	}
	return true;
}

// Quad tile tables of one surface, appended to ph and offs: the
// coefficients [g][n][k] of tap m = g + 4n (k stride P0S, zero past P0 and
// past the last tap; zero for m = 0, k = 0), then phi(0,0,0); offs gets each
// (g, n) tap's LDS byte distance back from the own position for a ring row
// of WY slots (zero past the last tap).  Returns the taps per lane.
template <class T>
int append_quad_tables(const double* phi, int P0, int P1, int P2, int P0S, int WY,
                       std::vector<T>& ph, std::vector<int>& offs) {
	const int P12 = P1 * P2, NT = quad_taps(P12);
	const size_t base = ph.size();
	ph.resize(base + size_t(kQuad) * NT * P0S + 1, T(0));
	ph.resize((ph.size() + 15) / 16 * 16, T(0));
	const size_t ob = offs.size();
	offs.resize(ob + size_t(kQuad) * NT, 0);
	for (int g = 0; g < kQuad; ++g)
		for (int n = 0; n < NT; ++n) {
			const int m = g + kQuad * n;
			if (m >= P12) continue;
			for (int k = 0; k < P0; ++k)
				if (m > 0 || k > 0)
					ph[base + (size_t(g) * NT + n) * P0S + k] = T(phi[size_t(k) * P12 + m]);
			offs[ob + size_t(g) * NT + n] = ((m / P2) * WY + (m % P2)) * int(sizeof(T));
		}
	ph[base + size_t(kQuad) * NT * P0S] = T(phi[0]);
	return NT;
}

// Workgroup w takes queue entries w, w + G, w + 2G, ... in order (G
// workgroups).  When all G are resident, the earliest unfinished entry
// always belongs to a workgroup that has finished its earlier entries and is
// marching that one, and it only waits on earlier entries, which are done:
// the launch cannot deadlock.  Residency of the whole grid is the
// precondition: the host never lets two batch launches run at once on a
// device (they could each hold part of the CUs and wait on each other's
// unscheduled workgroups), and the waits are bounded (abortw, limit) so that
// a launch that still lost residency ends with an error instead of hanging.
// The loop control is uniform (no atomic fetch under a thread-0 branch),
// which keeps the compiled loop a plain scalar loop around the tile's
// barriers.
template <class T, int TX, bool KAHAN, bool BUF = false, int STG = 0>
__global__ __launch_bounds__(kThreads) void ar_batch_kernel(
	T* __restrict__ arena, const T* __restrict__ coef, const int* __restrict__ taps,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, const BatchSurface* __restrict__ surf, const int2* __restrict__ queue,
	int nqueue, int* __restrict__ abortw, long long limit) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	for (int item = blockIdx.x; item < nqueue; item += gridDim.x) {
		const int si = queue[item].x;
		const int b = queue[item].y;
		const BatchSurface& sf = surf[si];
		T* const z = arena + sf.zoff;
		const T* const c0 = coef + sf.coff;
		const T* const cc = c0 + sf.n0;
		const int* const tp = taps + sf.toff;
		int* const pr = progress + size_t(sf.poff) * kProgStride;
		T* const mb = mbox + sf.moff;
		switch (sf.bucket) {
#define ARMA_BATCH_BODY(I)                                                                    \
	case I:                                                                                   \
		wave_tile<T, TX, kBatchP0[I], KAHAN, BUF, kD, STG>(                                     \
			z, sf.nt, sf.nx, sf.ny, c0, cc, tp, sf.P1, sf.P2, sf.nY, pr, zero, sink, mb, sf.NB, \
			sf.total_steps, sf.t0, b, abortw, limit, nullptr, smem_raw);                        \
		break;
			ARMA_BATCH_BODY(0)
			ARMA_BATCH_BODY(1)
			ARMA_BATCH_BODY(2)
			ARMA_BATCH_BODY(3)
			ARMA_BATCH_BODY(4)
			ARMA_BATCH_BODY(5)
#undef ARMA_BATCH_BODY
		}
	}
}

// The same queue walk over surfaces on the quad tile (1024 lanes): the
// switch is over (P0 bucket, taps per lane), for the pairs whose
// coefficients fit the register budget.
template <class T, int TX, bool KAHAN>
__global__ __launch_bounds__(kQThreads) void ar_quad_batch_kernel(
	T* __restrict__ arena, const T* __restrict__ coef, const int* __restrict__ taps,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, const BatchSurface* __restrict__ surf, const int2* __restrict__ queue,
	int nqueue, int* __restrict__ abortw, long long limit) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	for (int item = blockIdx.x; item < nqueue; item += gridDim.x) {
		const int si = queue[item].x;
		const int b = queue[item].y;
		const BatchSurface& sf = surf[si];
		T* const z = arena + sf.zoff;
		const T* const qc = coef + sf.coff;
		const int* const qo = taps + sf.toff;
		int* const pr = progress + size_t(sf.poff) * kProgStride;
		T* const mb = mbox + sf.moff;
		const int wide = quad_taps(sf.P1 * sf.P2) == 4 ? 1 : 0;
		switch (2 * sf.bucket + wide) {
#define ARMA_QBATCH_BODY(I, N)                                                                   \
	case 2 * I + (N == 4 ? 1 : 0):                                                               \
		if constexpr (quad_fits(int(sizeof(T)), kBatchP0[I], N))                                 \
			quad_tile<T, TX, kBatchP0[I], N, KAHAN>(z, sf.nt, sf.nx, sf.ny, qc, qo, sf.P1, sf.P2,  \
			                                        sf.nY, pr, zero, sink, mb, sf.NB,              \
			                                        sf.total_steps, sf.t0, b, abortw, limit,       \
			                                        nullptr, smem_raw);                            \
		break;
			ARMA_QBATCH_BODY(0, 2)
			ARMA_QBATCH_BODY(0, 4)
			ARMA_QBATCH_BODY(1, 2)
			ARMA_QBATCH_BODY(1, 4)
			ARMA_QBATCH_BODY(2, 2)
			ARMA_QBATCH_BODY(2, 4)
			ARMA_QBATCH_BODY(3, 2)
			ARMA_QBATCH_BODY(3, 4)
			ARMA_QBATCH_BODY(4, 2)
			ARMA_QBATCH_BODY(4, 4)
			ARMA_QBATCH_BODY(5, 2)
			ARMA_QBATCH_BODY(5, 4)
#undef ARMA_QBATCH_BODY
		}
	}
}

// Coefficient and tap tables of one surface, appended to ph and offs.
// ph gets the phi0 block (n0 elements: phi(k,0,0), zero past P0) and then
// the in-slice taps as [chunk][q][k] with k stride P0S (the P0 the kernel is
// compiled for; zero for k >= P0): tap m = i*P2 + j = 1 + chunk*kCH + q.
// offs gets each tap's LDS byte distance back from the own position for a
// ring row of WY slots (zero, i.e. the own position, past the last tap).
template <class T>
int append_tables(const double* phi, int P0, int P1, int P2, int P0S, int WY,
                  std::vector<T>& ph, std::vector<int>& offs) {
	const int P12 = P1 * P2;
	const int nch = tap_chunks(P12);
	const int n0 = (P0S + 15) / 16 * 16;
	const size_t base = ph.size();
	ph.resize(base + n0 + size_t(nch) * P0S * kCH, T(0));
	for (int k = 0; k < P0; ++k) {
		const double* pk = phi + size_t(k) * P12;
		ph[base + k] = T(pk[0]);
		for (int m = 1; m < P12; ++m) {
			const int ch = (m - 1) / kCH, q = (m - 1) % kCH;
			ph[base + n0 + (size_t(ch) * kCH + q) * P0S + k] = T(pk[m]);
		}
	}
	// keep every surface's block 16-element aligned
	ph.resize((ph.size() + 15) / 16 * 16, T(0));
	const size_t ob = offs.size();
	offs.resize(ob + size_t(nch) * kCH, 0);
	for (int m = 1; m < P12; ++m) offs[ob + m - 1] = ((m / P2) * WY + (m % P2)) * int(sizeof(T));
	return n0;
}

// Checked launches of single surfaces: ARMA_HIP_CHECK_HALO set (and not
// "0") in the environment.  arm() fills the launch's mailboxes with the
// sentinel and points the kernel at a check word in the launch's cleared
// zero-word line; verify() reads the word back after the launch and fails
// the call if a consumer read a mailbox row before its producer stored it.
template <class T>
struct HaloCheck {
	int* word = nullptr;
	void arm(arma_hip_ctx* ctx, unsigned char* zero_line, T* mbox, size_t n) {
		const char* e = std::getenv("ARMA_HIP_CHECK_HALO");
		if (!e || !*e || (e[0] == '0' && e[1] == 0)) return;
		word = reinterpret_cast<int*>(zero_line + 128);
		const size_t blocks = std::min<size_t>(std::max<size_t>(div_ceil(n, size_t(256)), 1), 65536);
		fill_sentinel<T><<<unsigned(blocks), 256, 0, ctx->stream>>>(mbox, n);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	void verify(arma_hip_ctx* ctx) const {
		if (!word) return;
		int v = 0;
		ARMA_HIP_CHECK(hipMemcpyAsync(&v, word, sizeof v, hipMemcpyDeviceToHost, ctx->stream));
		ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
		ARMA_HIP_REQUIRE(v == 0, ARMA_HIP_ERUNTIME,
		                 "checked launch: a halo value was read before its producer stored it");
	}
};

// One surface on the quad tile (16 x 16 columns, 1024 lanes per tile).
template <class T, bool KAHAN>
void launch_quad(arma_hip_ctx* ctx, const arma_ar_desc* d, T* zeta) {
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	constexpr int kTX = 16, kTY = kThreads / kTX;
	std::vector<T> ph;
	std::vector<int> offs;
	const int NT = append_quad_tables(d->phi, P0, P1, P2, P0, kTY + P2 - 1, ph, offs);
	const size_t phi_bytes = (ph.size() * sizeof(T) + 255) / 256 * 256;
	unsigned char* dbuf =
		(unsigned char*)work_buffer(ctx, ctx->phi, phi_bytes + offs.size() * sizeof(int));
	T* dphi = (T*)dbuf;
	int* doffs = (int*)(dbuf + phi_bytes);
	upload(ctx, dphi, ph.data(), ph.size() * sizeof(T));
	upload(ctx, doffs, offs.data(), offs.size() * sizeof(int));
	const int nY = div_ceil(ny, kTY);
	const int ntiles = div_ceil(nx, kTX) * nY;
	const int total_steps = nt + kTX + kTY - 2;
	// flags buffer: progress counters, a zero word for idle loads and a
	// per-tile sink for idle stores
	const size_t prog_bytes = sizeof(int) * kProgStride * size_t(ntiles);
	const size_t zero_bytes = 256;
	unsigned char* fb = (unsigned char*)work_buffer(
		ctx, ctx->flags, prog_bytes + zero_bytes + size_t(ntiles) * kQThreads * sizeof(T));
	int* prog = (int*)fb;
	const T* zero = (const T*)(fb + prog_bytes);
	T* sink = (T*)(fb + prog_bytes + zero_bytes);
	const int NB = std::max(1, mbox_width<kTX>(P1 - 1, P2 - 1));
	const int total_u = total_steps + (P1 - 1) + (P2 - 1);
	T* mbox = (T*)work_buffer(ctx, ctx->mbox, sizeof(T) * size_t(ntiles) * total_u * NB);
	const int nclear = int((prog_bytes + zero_bytes) / sizeof(int));
	reset_progress<<<div_ceil(nclear, 256), 256, 0, ctx->stream>>>(prog, nclear);
	ARMA_HIP_CHECK(hipGetLastError());
	HaloCheck<T> hc;
	hc.arm(ctx, fb + prog_bytes, mbox, size_t(ntiles) * total_u * NB);
	const int R = P1 + P2 - 1;
	const size_t lds = size_t(R) * (kTX + P1 - 1) * (kTY + P2 - 1) * sizeof(T);
	ARMA_HIP_REQUIRE(lds <= 160 * 1024, ARMA_HIP_EINVAL, "AR order too large for the LDS ring");
	KernelTimer timer(ctx, "ar_wavefront");
	auto go = [&](auto p0c, auto ntc) {
		constexpr int P0C = decltype(p0c)::value, NTC = decltype(ntc)::value;
		if constexpr (quad_fits(int(sizeof(T)), P0C, NTC)) {
			auto* fn = &ar_quad_kernel<T, P0C, NTC, KAHAN>;
			if (lds > 65536)
				ARMA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
				                                   hipFuncAttributeMaxDynamicSharedMemorySize,
				                                   int(lds)));
			fn<<<ntiles, kQThreads, lds, ctx->stream>>>(zeta, nt, nx, ny, dphi, doffs, P1, P2, nY,
			                                            prog, zero, sink, mbox, NB, total_steps,
			                                            hc.word);
		} else {
			throw Error(ARMA_HIP_EINVAL, "AR order outside the quad tile");
		}
	};
	auto by_nt = [&](auto p0c) {
		if (NT == 2)
			go(p0c, std::integral_constant<int, 2>{});
		else
			go(p0c, std::integral_constant<int, 4>{});
	};
	switch (P0) {
#define ARMA_QUAD(P) \
	case P:          \
		by_nt(std::integral_constant<int, P>{}); \
		break;
		ARMA_QUAD(1)
		ARMA_QUAD(2)
		ARMA_QUAD(3)
		ARMA_QUAD(4)
		ARMA_QUAD(5)
		ARMA_QUAD(6)
		ARMA_QUAD(7)
		ARMA_QUAD(8)
		ARMA_QUAD(9)
		ARMA_QUAD(10)
#undef ARMA_QUAD
	default:
		throw Error(ARMA_HIP_EINVAL, "bad AR order");
	}
	ARMA_HIP_CHECK(hipGetLastError());
	hc.verify(ctx);
}

template <class T, bool KAHAN>
void launch_typed(arma_hip_ctx* ctx, const arma_ar_desc* d, T* zeta) {
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	ARMA_HIP_REQUIRE(P1 <= kMaxP12 && P2 <= kMaxP12 && P0 <= 10, ARMA_HIP_EINVAL,
	                 "AR order above 10 is not supported by the wavefront kernel");
	if constexpr (sizeof(T) == 4) {
		if (quad_eligible(d)) {
			launch_quad<T, KAHAN>(ctx, d, zeta);
			return;
		}
	}
	constexpr int kTX = 16, kTY = kThreads / kTX;
	std::vector<T> ph;
	std::vector<int> offs;
	const size_t n0 = append_tables(d->phi, P0, P1, P2, P0, kTY + P2 - 1, ph, offs);
	const size_t phi_bytes = (ph.size() * sizeof(T) + 255) / 256 * 256;
	unsigned char* dbuf =
		(unsigned char*)work_buffer(ctx, ctx->phi, phi_bytes + offs.size() * sizeof(int) + 4);
	T* dphi = (T*)dbuf;
	int* doffs = (int*)(dbuf + phi_bytes);
	upload(ctx, dphi, ph.data(), ph.size() * sizeof(T));
	if (!offs.empty()) upload(ctx, doffs, offs.data(), offs.size() * sizeof(int));
	const int nX = div_ceil(nx, kTX), nY = div_ceil(ny, kTY);
	const int ntiles = nX * nY;
	const int total_steps = nt + kTX + kTY - 2;
	// flags buffer: progress counters, a zero word for out-of-range loads,
	// and a per-tile sink for out-of-range stores
	const size_t prog_bytes = sizeof(int) * kProgStride * size_t(ntiles);
	const size_t zero_bytes = 256;
	unsigned char* fb = (unsigned char*)work_buffer(
		ctx, ctx->flags, prog_bytes + zero_bytes + size_t(ntiles) * kThreads * sizeof(T));
	int* prog = (int*)fb;
	const T* zero = (const T*)(fb + prog_bytes);
	T* sink = (T*)(fb + prog_bytes + zero_bytes);
	// mailboxes: one row of boundary values per tile and step
	const int NB = std::max(1, mbox_width<kTX>(P1 - 1, P2 - 1));
	const int total_u = total_steps + (P1 - 1) + (P2 - 1);
	T* mbox = (T*)work_buffer(ctx, ctx->mbox, sizeof(T) * size_t(ntiles) * total_u * NB);
	const int nclear = int((prog_bytes + zero_bytes) / sizeof(int));
	reset_progress<<<div_ceil(nclear, 256), 256, 0, ctx->stream>>>(prog, nclear);
	ARMA_HIP_CHECK(hipGetLastError());
	HaloCheck<T> hc;
	hc.arm(ctx, fb + prog_bytes, mbox, size_t(ntiles) * total_u * NB);
	const int R = P1 + P2 - 1;
	const size_t ring_lds = size_t(R) * (kTX + P1 - 1) * (kTY + P2 - 1) * sizeof(T);
	ARMA_HIP_REQUIRE(ring_lds <= 160 * 1024, ARMA_HIP_EINVAL,
	                 "AR order too large for the LDS ring");
	KernelTimer timer(ctx, "ar_wavefront");
	// rings above 64 KiB (P1, P2 >= 9 in FP64) use the larger LDS of gfx950
	// mailbox accesses through a buffer descriptor when the span from a
	// tile's corner neighbour's mailbox to the end of its own fits 31 bits
	const bool buf = (size_t(nY) + 2) * total_u * NB * sizeof(T) < (size_t(1) << 31) - 64;
	// prefetch distance: 3 steps when every tile has a CU to itself (one
	// wave per SIMD), 4 with several tiles per CU (1024x256x256 FP64: 2.34
	// against 2.44 ms; 512^3, four tiles per CU: 4.92 against 4.87 ms)
	const bool short_queue = buf && ntiles <= std::max(ctx->num_cus, 1);
	// surface stores (and, with several tiles per CU, noise loads) through
	// the per-lane line, whole row segments per wave, when its LDS leaves as
	// many tiles resident on a CU as before.  2048x512x512 FP32, four tiles
	// per CU: 16.65 ms, 9.97 with the stores, 7.50 with the loads too;
	// 1024x256x256 FP64, one tile per CU: 2.34, 2.16, 2.19 (the loads' LDS
	// round trip then lengthens the step's chain more than it saves)
	// It is taken too where it leaves fewer (but at least two) tiles resident
	// per CU: tiles wait only on tiles dispatched before them, and the
	// shorter steps outweigh the later start of the last tiles (2048x512x512
	// FP64: 19.15 -> 9.23 ms; FP32 with phi (7,7,7): 20.1 -> 13.95 ms; 512^3
	// FP64: 4.86 -> 3.98 ms; two or three tiles per CU instead of four).
	const int per_cu = div_ceil(ntiles, std::max(ctx->num_cus, 1));
	const size_t dline_lds = size_t(kTY + 1) * kThreads * sizeof(T);
	const bool stg = buf && (size_t(per_cu) * (ring_lds + dline_lds) <= 160 * 1024 ||
	                         (ring_lds + dline_lds) * 2 <= 160 * 1024);
	const size_t lds = ring_lds + (stg ? dline_lds : 0);
#define ARMA_WAVE_LAUNCH(P, B, Q, S)                                                          \
	{                                                                                         \
		if (lds > 65536)                                                                      \
			ARMA_HIP_CHECK(hipFuncSetAttribute(                                               \
				reinterpret_cast<const void*>(&ar_wave_kernel<T, kTX, P, KAHAN, B, Q, S>),    \
				hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));                       \
		ar_wave_kernel<T, kTX, P, KAHAN, B, Q, S><<<ntiles, kThreads, lds, ctx->stream>>>(    \
			zeta, nt, nx, ny, dphi, dphi + n0, doffs, P1, P2, nY, prog, zero, sink, mbox, NB, \
			total_steps, hc.word);                                                            \
	}
#define ARMA_WAVE(P)                                                                          \
	case P:                                                                                   \
		if (short_queue && stg)                                                               \
			ARMA_WAVE_LAUNCH(P, true, 3, 1)                                                   \
		else if (short_queue)                                                                 \
			ARMA_WAVE_LAUNCH(P, true, 3, 0)                                                   \
		else if (buf && stg)                                                                  \
			ARMA_WAVE_LAUNCH(P, true, kD, 2)                                                  \
		else if (buf)                                                                         \
			ARMA_WAVE_LAUNCH(P, true, kD, 0)                                                  \
		else                                                                                  \
			ARMA_WAVE_LAUNCH(P, false, kD, 0)                                                 \
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
#undef ARMA_WAVE_LAUNCH
	ARMA_HIP_CHECK(hipGetLastError());
	hc.verify(ctx);
}

// The kernel a batch part launches: the quad tile, the one-lane tile with
// the per-lane line (stg: row-aligned noise loads and surface stores, with
// mailboxes behind a buffer descriptor), or the plain one-lane tile.
template <class T, int TX, bool KAHAN>
struct BatchKernel {
	static const void* fn(bool quad, bool stg) {
		if constexpr (sizeof(T) == 4)
			if (quad) return reinterpret_cast<const void*>(&ar_quad_batch_kernel<T, TX, KAHAN>);
		if (stg) return reinterpret_cast<const void*>(&ar_batch_kernel<T, TX, KAHAN, true, 2>);
		return reinterpret_cast<const void*>(&ar_batch_kernel<T, TX, KAHAN>);
	}
};

const void* batch_kernel_fn(int precision, int tx, bool quad, bool stg = false) {
	if (precision == ARMA_FP32) {
		if (tx == 8) return BatchKernel<float, 8, true>::fn(quad, stg);
		if (tx == 32) return BatchKernel<float, 32, true>::fn(quad, stg);
		return BatchKernel<float, 16, true>::fn(quad, stg);
	}
	if (tx == 8) return BatchKernel<double, 8, false>::fn(quad, stg);
	if (tx == 32) return BatchKernel<double, 32, false>::fn(quad, stg);
	return BatchKernel<double, 16, false>::fn(quad, stg);
}

}  // namespace

namespace {

// A batch launch is complete when its event has fired.  Events are shared so
// that the chain below can outlive the batch that recorded them.
struct LaunchEvent {
	hipEvent_t ev = nullptr;
	LaunchEvent() { ARMA_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming)); }
	~LaunchEvent() {
		if (ev) (void)hipEventDestroy(ev);
	}
	LaunchEvent(const LaunchEvent&) = delete;
	LaunchEvent& operator=(const LaunchEvent&) = delete;
};

// The last batch launch enqueued on each device, on any stream.
std::mutex g_batch_mu;
std::map<int, std::shared_ptr<LaunchEvent>> g_batch_last;

// Bounded waits give up after this long without progress.
constexpr double kBatchWaitLimitS = 10.0;

}  // namespace

// One persistent launch over the members of one tile kind: the classic tile
// (256 lanes) or the quad tile (1024 lanes).  A batch has one part per kind
// its members need; the parts run one after the other.
struct ArBatchPart {
	bool quad = false;
	bool buf = true;  // every member's mailbox span fits a buffer descriptor
	bool stg = false;  // one-lane tiles through the per-lane line (implies buf)
	int nqueue = 0, grid = 0, max_slots = 0;
	size_t lds = 0;
	size_t nclear = 0;  // ints cleared before a launch: the abort word line, progress, zero word
	Scratch coef, taps, surf, queue, flags, mbox;
	int* clear_base = nullptr;  // start of the words cleared before a launch
	int* abortw = nullptr;      // set by a launch whose bounded waits gave up
	int* progress = nullptr;
	void* zero = nullptr;
	void* sink = nullptr;
};

struct ArBatch {
	int precision = 0;
	BatchArConfig cfg;
	ArBatchPart part[2];  // classic, quad
	long long limit = 0;  // wall-clock ticks of kBatchWaitLimitS
	std::shared_ptr<LaunchEvent> done;
};

namespace {

template <class T>
void build_part(arma_hip_ctx* ctx, const ArBatch* ab, ArBatchPart* pt,
                const std::vector<BatchMember>& all, const std::vector<int>& idx) {
	const int TX = ab->cfg.tile_x, TY = kThreads / TX;
	const int lanes = pt->quad ? kQThreads : kThreads;
	std::vector<T> coef;
	std::vector<int> taps;
	std::vector<BatchSurface> surf(idx.size());
	std::vector<double> work(idx.size());
	std::vector<int> ntiles(idx.size());
	long long tiles = 0, mbox_elems = 0;
	int max_tiles = 0;
	size_t lds = 0;
	pt->buf = true;
	for (size_t i = 0; i < idx.size(); ++i) {
		const arma_ar_desc* d = all[idx[i]].d;
		const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
		const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
		ARMA_HIP_REQUIRE(P1 <= kMaxP12 && P2 <= kMaxP12 && P0 <= 10, ARMA_HIP_EINVAL,
		                 "AR order above 10 is not supported by the wavefront kernel");
		// a tile's halo comes from its direct neighbours only
		ARMA_HIP_REQUIRE(P1 - 1 <= TX && P2 - 1 <= TY, ARMA_HIP_EINVAL,
		                 "AR order wider than the batch tile");
		const int bucket = batch_bucket(P0);
		BatchSurface& sf = surf[i];
		sf.zoff = (long long)all[idx[i]].zoff;
		sf.nt = nt;
		sf.nx = nx;
		sf.ny = ny;
		sf.P1 = P1;
		sf.P2 = P2;
		sf.bucket = bucket;
		sf.nY = div_ceil(ny, TY);
		sf.coff = int(coef.size());
		sf.toff = int(taps.size());
		sf.n0 = pt->quad
		            ? append_quad_tables(d->phi, P0, P1, P2, kBatchP0[bucket], TY + P2 - 1, coef, taps)
		            : append_tables(d->phi, P0, P1, P2, kBatchP0[bucket], TY + P2 - 1, coef, taps);
		ntiles[i] = div_ceil(nx, TX) * sf.nY;
		sf.poff = int(tiles);
		sf.total_steps = nt + TX + TY - 2;
		sf.t0 = all[idx[i]].t0;
		sf.pad = 0;
		sf.NB = std::max(1, mbox_width_rt(TX, P1 - 1, P2 - 1));
		const long long total_u = sf.total_steps + (P1 - 1) + (P2 - 1);
		pt->buf = pt->buf && (size_t(sf.nY) + 2) * size_t(total_u) * sf.NB * sizeof(T) <
		                         (size_t(1) << 31) - 64;
		sf.moff = mbox_elems;
		mbox_elems += (long long)ntiles[i] * total_u * sf.NB;
		mbox_elems = (mbox_elems + 63) / 64 * 64;
		tiles += ntiles[i];
		max_tiles = std::max(max_tiles, ntiles[i]);
		lds = std::max(lds, size_t(P1 + P2 - 1) * (TX + P1 - 1) * (TY + P2 - 1) * sizeof(T));
		work[i] = double(nt) * nx * ny * kBatchP0[bucket] * P1 * P2;
	}
	ARMA_HIP_REQUIRE(lds <= 160 * 1024, ARMA_HIP_EINVAL, "AR order too large for the LDS ring");
	ARMA_HIP_REQUIRE(tiles < (1ll << 31) / kProgStride && coef.size() < (1u << 31) &&
	                     taps.size() < (1u << 31),
	                 ARMA_HIP_EINVAL, "batch too large");
	// the queue: surfaces by decreasing work; tile rows in order inside a
	// surface, so every tile follows the tiles it waits on
	std::vector<int> order(idx.size());
	for (size_t i = 0; i < idx.size(); ++i) order[i] = int(i);
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
	std::vector<int2> queue;
	queue.reserve(size_t(tiles));
	if (ab->cfg.schedule == ARMA_BATCH_INTERLEAVED) {
		std::vector<int> done(idx.size(), 0);
		for (bool more = true; more;) {
			more = false;
			for (int i : order) {
				const int row = surf[i].nY;
				for (int k = 0; k < row && done[i] < ntiles[i]; ++k, ++done[i])
					queue.push_back(make_int2(i, done[i]));
				more = more || done[i] < ntiles[i];
			}
		}
	} else {
		for (int i : order)
			for (int b = 0; b < ntiles[i]; ++b) queue.push_back(make_int2(i, b));
	}
	pt->nqueue = int(queue.size());
	pt->lds = lds;
	// device tables
	pt->coef.get(std::max<size_t>(coef.size(), 16) * sizeof(T));
	pt->taps.get(std::max<size_t>(taps.size(), 16) * sizeof(int));
	pt->surf.get(surf.size() * sizeof(BatchSurface));
	pt->queue.get(std::max<size_t>(queue.size(), 1) * sizeof(int2));
	upload(ctx, pt->coef.ptr, coef.data(), coef.size() * sizeof(T));
	upload(ctx, pt->taps.ptr, taps.data(), taps.size() * sizeof(int));
	upload(ctx, pt->surf.ptr, surf.data(), surf.size() * sizeof(BatchSurface));
	upload(ctx, pt->queue.ptr, queue.data(), queue.size() * sizeof(int2));
	// flags: the abort word (its own line after a spare one), progress
	// counters, a zero word, and a sink for idle stores sized for the
	// largest surface
	const size_t head = 256, prog_bytes = sizeof(int) * kProgStride * size_t(tiles);
	const size_t zero_bytes = 256;
	unsigned char* fb = (unsigned char*)pt->flags.get(head + prog_bytes + zero_bytes +
	                                                  size_t(max_tiles) * lanes * sizeof(T));
	pt->clear_base = (int*)fb;
	pt->abortw = (int*)(fb + 128);
	pt->progress = (int*)(fb + head);
	pt->zero = fb + head + prog_bytes;
	pt->sink = fb + head + prog_bytes + zero_bytes;
	pt->nclear = (head + prog_bytes + zero_bytes) / sizeof(int);
	pt->mbox.get(std::max<size_t>(size_t(mbox_elems), 64) * sizeof(T));
	const void* fn = batch_kernel_fn(ab->precision, TX, pt->quad);
	if (lds > 65536)
		ARMA_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(lds)));
	int per_cu = 0;
	ARMA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, lanes, lds));
	// the per-lane line (TY + 1 slots per lane after the ring) when the
	// kernel that uses it keeps at least as many workgroups resident
	pt->stg = false;
	if (!pt->quad && pt->buf) {
		const size_t lds_s = lds + size_t(TY + 1) * kThreads * sizeof(T);
		const void* fs = batch_kernel_fn(ab->precision, TX, false, true);
		if (lds_s <= 160 * 1024) {
			if (lds_s > 65536)
				ARMA_HIP_CHECK(hipFuncSetAttribute(
					fs, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_s)));
			int per_cu_s = 0;
			ARMA_HIP_CHECK(
				hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s, fs, lanes, lds_s));
			if (per_cu_s >= per_cu) {
				pt->stg = true;
				pt->lds = lds = lds_s;
				per_cu = per_cu_s;
			}
		}
	}
	pt->max_slots = std::max(per_cu, 1);
	const int slots = ab->cfg.slots_per_cu > 0 ? std::min(ab->cfg.slots_per_cu, pt->max_slots)
	                                           : pt->max_slots;
	pt->grid = int(std::min<long long>(pt->nqueue, (long long)slots * std::max(ctx->num_cus, 1)));
}

// Members whose order fits the quad tile go to the quad part, the others to
// the classic part (the same rule as single surfaces).
template <class T>
void build_batch(arma_hip_ctx* ctx, ArBatch* ab, const std::vector<BatchMember>& m) {
	std::vector<int> idx[2];
	for (size_t i = 0; i < m.size(); ++i) {
		// the member as a single surface at the batch's precision
		arma_ar_desc d = *m[i].d;
		d.precision = ab->precision;
		idx[quad_eligible(&d) ? 1 : 0].push_back(int(i));
	}
	for (int k = 0; k < 2; ++k) {
		ab->part[k].quad = k == 1;
		ab->part[k].nqueue = 0;
		if (!idx[k].empty()) build_part<T>(ctx, ab, &ab->part[k], m, idx[k]);
	}
	int khz = 0;
	ARMA_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
	ab->limit = (long long)(double(std::max(khz, 1)) * 1e3 * kBatchWaitLimitS);
	if (!ab->done) ab->done = std::make_shared<LaunchEvent>();
}

template <class T, int TX, bool KAHAN>
void batch_go(arma_hip_ctx* ctx, const ArBatch* ab, const ArBatchPart* pt, T* arena) {
	if constexpr (sizeof(T) == 4) {
		if (pt->quad) {
			ar_quad_batch_kernel<T, TX, KAHAN><<<pt->grid, kQThreads, pt->lds, ctx->stream>>>(
				arena, (const T*)pt->coef.ptr, (const int*)pt->taps.ptr, pt->progress,
				(const T*)pt->zero, (T*)pt->sink, (T*)pt->mbox.ptr,
				(const BatchSurface*)pt->surf.ptr, (const int2*)pt->queue.ptr, pt->nqueue,
				pt->abortw, ab->limit);
			return;
		}
	}
	// mailboxes behind a buffer descriptor when every member's span fits (its
	// registers are fewer, so the residency computed for the other form holds)
	auto kernel = pt->stg  ? &ar_batch_kernel<T, TX, KAHAN, true, 2>
	              : pt->buf ? &ar_batch_kernel<T, TX, KAHAN, true, 0>
	                        : &ar_batch_kernel<T, TX, KAHAN, false, 0>;
	kernel<<<pt->grid, kThreads, pt->lds, ctx->stream>>>(
		arena, (const T*)pt->coef.ptr, (const int*)pt->taps.ptr, pt->progress,
		(const T*)pt->zero, (T*)pt->sink, (T*)pt->mbox.ptr, (const BatchSurface*)pt->surf.ptr,
		(const int2*)pt->queue.ptr, pt->nqueue, pt->abortw, ab->limit);
}

void batch_enqueue(arma_hip_ctx* ctx, ArBatch* ab, void* arena) {
	KernelTimer timer(ctx, "ar_batch");
	const int tx = ab->cfg.tile_x;
	for (const ArBatchPart& pt : ab->part) {
		if (pt.nqueue == 0) continue;
		const int nclear = int(pt.nclear);
		reset_progress<<<div_ceil(nclear, 256), 256, 0, ctx->stream>>>(pt.clear_base, nclear);
		ARMA_HIP_CHECK(hipGetLastError());
		if (ab->precision == ARMA_FP32) {
			if (tx == 8) batch_go<float, 8, true>(ctx, ab, &pt, (float*)arena);
			else if (tx == 32) batch_go<float, 32, true>(ctx, ab, &pt, (float*)arena);
			else batch_go<float, 16, true>(ctx, ab, &pt, (float*)arena);
		} else {
			if (tx == 8) batch_go<double, 8, false>(ctx, ab, &pt, (double*)arena);
			else if (tx == 32) batch_go<double, 32, false>(ctx, ab, &pt, (double*)arena);
			else batch_go<double, 16, false>(ctx, ab, &pt, (double*)arena);
		}
		ARMA_HIP_CHECK(hipGetLastError());
	}
}

}  // namespace

void launch_ar_wavefront(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta) {
	if (d->precision == ARMA_FP32)
		launch_typed<float, true>(ctx, d, (float*)zeta);
	else
		launch_typed<double, false>(ctx, d, (double*)zeta);
}

size_t ar_wavefront_work_bytes(const arma_ar_desc* d) {
	// the buffers launch_quad or launch_typed takes, in their order, each
	// 256-byte aligned
	const size_t es = d->precision == ARMA_FP32 ? 4 : 8;
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	constexpr int kTX = 16, kTY = kThreads / kTX;
	auto al = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t ntiles = size_t(div_ceil(nx, kTX)) * div_ceil(ny, kTY);
	const size_t NB = size_t(std::max(1, mbox_width<kTX>(P1 - 1, P2 - 1)));
	const size_t total_u = size_t(nt) + kTX + kTY - 2 + (P1 - 1) + (P2 - 1);
	const size_t mbox = es * ntiles * total_u * NB;
	const size_t prog = sizeof(int) * kProgStride * ntiles + 256;
	if (quad_eligible(d)) {
		const size_t nq = size_t(kQuad) * quad_taps(P1 * P2);
		const size_t tables = al((nq * P0 + 1 + 15) / 16 * 16 * es) + nq * sizeof(int);
		return al(tables) + al(prog + ntiles * kQThreads * es) + mbox;
	}
	const size_t nch = size_t(tap_chunks(P1 * P2));
	const size_t nph = (size_t((P0 + 15) / 16 * 16) + nch * P0 * kCH + 15) / 16 * 16;
	const size_t tables = al(nph * es) + nch * kCH * sizeof(int) + 4;
	return al(tables) + al(prog + ntiles * kThreads * es) + mbox;
}

bool ar_batch_fits(const std::vector<BatchMember>& m, int tile_x) {
	for (const BatchMember& x : m)
		if (x.d->phi_shape[1] - 1 > tile_x || x.d->phi_shape[2] - 1 > kThreads / tile_x)
			return false;
	return true;
}

void ar_batch_build(arma_hip_ctx* ctx, ArBatch* ab, int precision,
                    const std::vector<BatchMember>& m, const BatchArConfig& cfg) {
	ARMA_HIP_REQUIRE(cfg.tile_x == 8 || cfg.tile_x == 16 || cfg.tile_x == 32, ARMA_HIP_EINVAL,
	                 "batch tile width must be 8, 16 or 32");
	ARMA_HIP_REQUIRE(cfg.schedule == ARMA_BATCH_SURFACE_MAJOR ||
	                     cfg.schedule == ARMA_BATCH_INTERLEAVED,
	                 ARMA_HIP_EINVAL, "bad batch schedule");
	ARMA_HIP_REQUIRE(!m.empty(), ARMA_HIP_EINVAL, "empty batch");
	ab->precision = precision;
	ab->cfg = cfg;
	if (precision == ARMA_FP32)
		build_batch<float>(ctx, ab, m);
	else
		build_batch<double>(ctx, ab, m);
}

ArBatch* ar_batch_create(arma_hip_ctx* ctx, int precision, const std::vector<BatchMember>& m,
                         const BatchArConfig& cfg) {
	auto* ab = new ArBatch;
	try {
		ar_batch_build(ctx, ab, precision, m, cfg);
	} catch (...) {
		delete ab;
		throw;
	}
	return ab;
}

ArBatch* ar_batch_new() { return new ArBatch; }

size_t ar_batch_member_bytes(const arma_ar_desc* d, int tile_x) {
	const int TX = tile_x, TY = kThreads / TX;
	const size_t es = d->precision == ARMA_FP32 ? 4 : 8;
	const int P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	const size_t tiles = size_t(div_ceil(d->shape[1], TX)) * div_ceil(d->shape[2], TY);
	const size_t total_u = size_t(d->shape[0]) + TX + TY - 2 + (P1 - 1) + (P2 - 1);
	const size_t NB = std::max(1, mbox_width_rt(TX, P1 - 1, P2 - 1));
	const size_t lanes = quad_eligible(d) ? kQThreads : kThreads;
	return es * (tiles * total_u * NB + tiles * lanes) + sizeof(int) * kProgStride * tiles;
}

// Batch launches on one device run one after another, whatever stream they
// are enqueued on: each waits for the previous one's event.  Under graph
// capture the chain is left out (an event recorded outside the capture cannot
// be waited on inside it); the caller then keeps replays from overlapping
// other batch launches.
void ar_batch_launch(arma_hip_ctx* ctx, ArBatch* ab, void* arena) {
	hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
	ARMA_HIP_CHECK(hipStreamIsCapturing(ctx->stream, &cap));
	if (cap != hipStreamCaptureStatusNone) {
		batch_enqueue(ctx, ab, arena);
		return;
	}
	std::lock_guard<std::mutex> lock(g_batch_mu);
	std::shared_ptr<LaunchEvent>& last = g_batch_last[ctx->device];
	if (last && last != ab->done) ARMA_HIP_CHECK(hipStreamWaitEvent(ctx->stream, last->ev, 0));
	batch_enqueue(ctx, ab, arena);
	ARMA_HIP_CHECK(hipEventRecord(ab->done->ev, ctx->stream));
	last = ab->done;
}

void ar_batch_check(arma_hip_ctx* ctx, ArBatch* ab) {
	int stalled[2] = {0, 0};
	for (int k = 0; k < 2; ++k)
		if (ab->part[k].nqueue > 0)
			ARMA_HIP_CHECK(hipMemcpyAsync(&stalled[k], ab->part[k].abortw, sizeof(int),
			                              hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	ARMA_HIP_REQUIRE(stalled[0] == 0 && stalled[1] == 0, ARMA_HIP_ERUNTIME,
	                 "batch recurrence launch stalled (its workgroups were not all resident); "
	                 "its surfaces are invalid");
}

void ar_batch_destroy(ArBatch* ab) { delete ab; }

// Workgroups of the batch's launches (the largest part's grid).
int ar_batch_grid(const ArBatch* ab) { return std::max(ab->part[0].grid, ab->part[1].grid); }

// Resident workgroups per CU of the batch's tile kinds (the most of its parts).
int ar_batch_max_slots(const ArBatch* ab) {
	int m = 0;
	for (const ArBatchPart& pt : ab->part)
		if (pt.nqueue > 0) m = std::max(m, pt.max_slots);
	return std::max(m, 1);
}

}  // namespace arma_hip
