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
template <class T, int TX, int P0, bool KAHAN>
__device__ __forceinline__ void wave_tile(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ phi0,
	const T* __restrict__ phic, const int* __restrict__ toff, int P1, int P2, int nY,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, int NB, int total_steps, int t0, int b, int* abortw, long long limit,
	unsigned char* smem_raw) {
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
	auto issue_loads = [&](int u2, T& e, T& h0, T& h1) {
		const int s2 = u2 + s0;
		const bool in = u2 < total_u;
		const int t = s2 - lx - ly;
		const bool ve = in && active && t >= 0 && t < nt;
		const int t0 = s2 - hsk[0];
		const bool v0 = in && hin[0] && t0 >= 0 && t0 < nt;
		const int t1 = s2 - hsk[1];
		const bool v1 = in && hin[1] && t1 >= 0 && t1 < nt;
		e = *(ve ? zeta + size_t(t) * sxy + own_off : zero);
		h0 = coherent_load(v0 ? mbox + (size_t(hnb[0]) * total_u + u2 + hdu[0]) * NB + hm[0]
		                      : zero);
		h1 = coherent_load(v1 ? mbox + (size_t(hnb[1]) * total_u + u2 + hdu[1]) * NB + hm[1]
		                      : zero);
	};

	// neighbour progress as last observed by lane 0, and the flag loads in
	// flight (read kD steps after they are issued)
	int known_t = total_u, known_l = total_u;
	if (tid == 0) {
		known_t = nb_top >= 0 ? wait_progress(top_prog, need_top(kD - 1), abortw, limit) : total_u;
		known_l =
			nb_left >= 0 ? wait_progress(left_prog, need_left(kD - 1), abortw, limit) : total_u;
	}
	__syncthreads();
	T qe[kD], qh0[kD], qh1[kD];
	int qt[kD], ql[kD];
#pragma unroll
	for (int d = 0; d < kD; ++d) {
		issue_loads(d, qe[d], qh0[d], qh1[d]);
		qt[d] = known_t;
		ql[d] = known_l;
	}

	auto step = [&](int u, T& e_q, T& h0_q, T& h1_q, int& t_q, int& l_q) {
		const int s = u + s0;
		// 1. halo values of step s into the ring
		{
			const int t0 = s - hsk[0];
			if (hval[0] && t0 >= 0 && t0 < nt) ring[rs_h0 * slot + hlds[0]] = h0_q;
			const int t1 = s - hsk[1];
			if (hval[1] && t1 >= 0 && t1 < nt) ring[rs_h1 * slot + hlds[1]] = h1_q;
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
			const T eps = e_q;
			const T zc = eps + (acc[0].value() + phi000 * eps);
			const T z = t < t0 ? eps : zc;
			if (live) ring[rs_own * slot + own_lds] = z;
			*(live ? zeta + size_t(t) * sxy + own_off : my_sink) = z;
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
		// 3. lane 0: may the loads of step u + kD go out?
		const int u2 = u + kD;
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
		// 4. publish: the stores of step u - kD were issued before the queue
		//    loads consumed at this step, so every wave has seen them complete
		if (tid == 0 && u >= kD) coherent_store(my_prog, u - kD + 1);
		// 5. loads for step u + kD into the queue slot just consumed
		issue_loads(u2, e_q, h0_q, h1_q);
	};

	int u = 0;
	for (; u + kD <= total_u; u += kD) {
#pragma unroll
		for (int d = 0; d < kD; ++d) step(u + d, qe[d], qh0[d], qh1[d], qt[d], ql[d]);
	}
	// the last steps (fewer than kD), each on its own queue slot; the
	// condition is the same on every lane
#pragma unroll
	for (int d = 0; d < kD; ++d)
		if (u + d < total_u) step(u + d, qe[d], qh0[d], qh1[d], qt[d], ql[d]);
	// every store complete, then the final count
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	if (tid == 0) coherent_store(my_prog, total_u);
}

// Kernel arguments stay plain restrict-qualified pointers: the coefficient
// reads are then provably unclobbered by the kernel's stores and compile to
// scalar loads (a struct argument costs 30-60 VGPRs per lane).
template <class T, int TX, int P0, bool KAHAN>
__global__ __launch_bounds__(kThreads) void ar_wave_kernel(
	T* __restrict__ zeta, int nt, int nx, int ny, const T* __restrict__ phi0,
	const T* __restrict__ phic, const int* __restrict__ toff, int P1, int P2, int nY,
	int* __restrict__ progress, const T* __restrict__ zero, T* __restrict__ sink,
	T* __restrict__ mbox, int NB, int total_steps) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	// unbounded waits: tiles wait only on tiles with smaller workgroup ids,
	// which were dispatched before them
	wave_tile<T, TX, P0, KAHAN>(zeta, nt, nx, ny, phi0, phic, toff, P1, P2, nY, progress, zero, sink,
	                            mbox, NB, total_steps, 0, blockIdx.x, nullptr, 0, smem_raw);
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
template <class T, int TX, bool KAHAN>
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
		wave_tile<T, TX, kBatchP0[I], KAHAN>(z, sf.nt, sf.nx, sf.ny, c0, cc, tp, sf.P1, sf.P2,  \
		                                     sf.nY, pr, zero, sink, mb, sf.NB, sf.total_steps, \
		                                     sf.t0, b, abortw, limit, smem_raw);               \
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

template <class T, bool KAHAN>
void launch_typed(arma_hip_ctx* ctx, const arma_ar_desc* d, T* zeta) {
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	ARMA_HIP_REQUIRE(P1 <= kMaxP12 && P2 <= kMaxP12 && P0 <= 10, ARMA_HIP_EINVAL,
	                 "AR order above 10 is not supported by the wavefront kernel");
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
	const int R = P1 + P2 - 1;
	const size_t lds = size_t(R) * (kTX + P1 - 1) * (kTY + P2 - 1) * sizeof(T);
	ARMA_HIP_REQUIRE(lds <= 160 * 1024, ARMA_HIP_EINVAL, "AR order too large for the LDS ring");
	KernelTimer timer(ctx, "ar_wavefront");
	// rings above 64 KiB (P1, P2 >= 9 in FP64) use the larger LDS of gfx950
#define ARMA_WAVE(P)                                                                          \
	case P:                                                                                   \
		if (lds > 65536)                                                                      \
			ARMA_HIP_CHECK(hipFuncSetAttribute(                                               \
				reinterpret_cast<const void*>(&ar_wave_kernel<T, kTX, P, KAHAN>),             \
				hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));                       \
		ar_wave_kernel<T, kTX, P, KAHAN><<<ntiles, kThreads, lds, ctx->stream>>>(             \
			zeta, nt, nx, ny, dphi, dphi + n0, doffs, P1, P2, nY, prog, zero, sink, mbox, NB, \
			total_steps);                                                                     \
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

template <class T, int TX, bool KAHAN>
struct BatchKernel {
	static const void* fn() { return reinterpret_cast<const void*>(&ar_batch_kernel<T, TX, KAHAN>); }
};

const void* batch_kernel_fn(int precision, int tx) {
	if (precision == ARMA_FP32) {
		if (tx == 8) return BatchKernel<float, 8, true>::fn();
		if (tx == 32) return BatchKernel<float, 32, true>::fn();
		return BatchKernel<float, 16, true>::fn();
	}
	if (tx == 8) return BatchKernel<double, 8, false>::fn();
	if (tx == 32) return BatchKernel<double, 32, false>::fn();
	return BatchKernel<double, 16, false>::fn();
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

struct ArBatch {
	int precision = 0;
	BatchArConfig cfg;
	int nqueue = 0, grid = 0, max_slots = 0;
	size_t lds = 0;
	size_t nclear = 0;  // ints cleared before a launch: queue counter, abort word, progress, zero word
	Scratch coef, taps, surf, queue, flags, mbox;
	int* clear_base = nullptr;  // start of the words cleared before a launch
	int* abortw = nullptr;      // set by a launch whose bounded waits gave up
	int* progress = nullptr;
	void* zero = nullptr;
	void* sink = nullptr;
	long long limit = 0;  // wall-clock ticks of kBatchWaitLimitS
	std::shared_ptr<LaunchEvent> done;
};

namespace {

template <class T>
void build_batch(arma_hip_ctx* ctx, ArBatch* ab, const std::vector<BatchMember>& m) {
	const int TX = ab->cfg.tile_x, TY = kThreads / TX;
	std::vector<T> coef;
	std::vector<int> taps;
	std::vector<BatchSurface> surf(m.size());
	std::vector<double> work(m.size());
	std::vector<int> ntiles(m.size());
	long long tiles = 0, mbox_elems = 0;
	int max_tiles = 0;
	size_t lds = 0;
	for (size_t i = 0; i < m.size(); ++i) {
		const arma_ar_desc* d = m[i].d;
		const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
		const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
		ARMA_HIP_REQUIRE(P1 <= kMaxP12 && P2 <= kMaxP12 && P0 <= 10, ARMA_HIP_EINVAL,
		                 "AR order above 10 is not supported by the wavefront kernel");
		// a tile's halo comes from its direct neighbours only
		ARMA_HIP_REQUIRE(P1 - 1 <= TX && P2 - 1 <= TY, ARMA_HIP_EINVAL,
		                 "AR order wider than the batch tile");
		const int bucket = batch_bucket(P0);
		BatchSurface& sf = surf[i];
		sf.zoff = (long long)m[i].zoff;
		sf.nt = nt;
		sf.nx = nx;
		sf.ny = ny;
		sf.P1 = P1;
		sf.P2 = P2;
		sf.bucket = bucket;
		sf.nY = div_ceil(ny, TY);
		sf.coff = int(coef.size());
		sf.toff = int(taps.size());
		sf.n0 = append_tables(d->phi, P0, P1, P2, kBatchP0[bucket], TY + P2 - 1, coef, taps);
		ntiles[i] = div_ceil(nx, TX) * sf.nY;
		sf.poff = int(tiles);
		sf.total_steps = nt + TX + TY - 2;
		sf.t0 = m[i].t0;
		sf.pad = 0;
		sf.NB = std::max(1, mbox_width_rt(TX, P1 - 1, P2 - 1));
		const long long total_u = sf.total_steps + (P1 - 1) + (P2 - 1);
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
	std::vector<int> order(m.size());
	for (size_t i = 0; i < m.size(); ++i) order[i] = int(i);
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
	std::vector<int2> queue;
	queue.reserve(size_t(tiles));
	if (ab->cfg.schedule == ARMA_BATCH_INTERLEAVED) {
		std::vector<int> done(m.size(), 0);
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
	ab->nqueue = int(queue.size());
	ab->lds = lds;
	// device tables
	ab->coef.get(std::max<size_t>(coef.size(), 16) * sizeof(T));
	ab->taps.get(std::max<size_t>(taps.size(), 16) * sizeof(int));
	ab->surf.get(surf.size() * sizeof(BatchSurface));
	ab->queue.get(std::max<size_t>(queue.size(), 1) * sizeof(int2));
	upload(ctx, ab->coef.ptr, coef.data(), coef.size() * sizeof(T));
	upload(ctx, ab->taps.ptr, taps.data(), taps.size() * sizeof(int));
	upload(ctx, ab->surf.ptr, surf.data(), surf.size() * sizeof(BatchSurface));
	upload(ctx, ab->queue.ptr, queue.data(), queue.size() * sizeof(int2));
	// flags: the queue counter and the abort word (one line each), progress
	// counters, a zero word, and a sink for out-of-range stores sized for the
	// largest surface
	const size_t head = 256, prog_bytes = sizeof(int) * kProgStride * size_t(tiles);
	const size_t zero_bytes = 256;
	unsigned char* fb = (unsigned char*)ab->flags.get(head + prog_bytes + zero_bytes +
	                                                  size_t(max_tiles) * kThreads * sizeof(T));
	ab->clear_base = (int*)fb;
	ab->abortw = (int*)(fb + 128);
	ab->progress = (int*)(fb + head);
	ab->zero = fb + head + prog_bytes;
	ab->sink = fb + head + prog_bytes + zero_bytes;
	ab->nclear = (head + prog_bytes + zero_bytes) / sizeof(int);
	ab->mbox.get(std::max<size_t>(size_t(mbox_elems), 64) * sizeof(T));
	const void* fn = batch_kernel_fn(ab->precision, TX);
	if (lds > 65536)
		ARMA_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
		                                   int(lds)));
	int per_cu = 0;
	ARMA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, lds));
	ab->max_slots = std::max(per_cu, 1);
	const int slots = ab->cfg.slots_per_cu > 0 ? std::min(ab->cfg.slots_per_cu, ab->max_slots)
	                                           : ab->max_slots;
	ab->grid = int(std::min<long long>(ab->nqueue, (long long)slots * std::max(ctx->num_cus, 1)));
	int khz = 0;
	ARMA_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
	ab->limit = (long long)(double(std::max(khz, 1)) * 1e3 * kBatchWaitLimitS);
	if (!ab->done) ab->done = std::make_shared<LaunchEvent>();
}

template <class T, int TX, bool KAHAN>
void batch_go(arma_hip_ctx* ctx, ArBatch* ab, T* arena) {
	ar_batch_kernel<T, TX, KAHAN><<<ab->grid, kThreads, ab->lds, ctx->stream>>>(
		arena, (const T*)ab->coef.ptr, (const int*)ab->taps.ptr, ab->progress,
		(const T*)ab->zero, (T*)ab->sink, (T*)ab->mbox.ptr, (const BatchSurface*)ab->surf.ptr,
		(const int2*)ab->queue.ptr, ab->nqueue, ab->abortw, ab->limit);
}

void batch_enqueue(arma_hip_ctx* ctx, ArBatch* ab, void* arena) {
	const int nclear = int(ab->nclear);
	reset_progress<<<div_ceil(nclear, 256), 256, 0, ctx->stream>>>(ab->clear_base, nclear);
	ARMA_HIP_CHECK(hipGetLastError());
	KernelTimer timer(ctx, "ar_batch");
	const int tx = ab->cfg.tile_x;
	if (ab->precision == ARMA_FP32) {
		if (tx == 8) batch_go<float, 8, true>(ctx, ab, (float*)arena);
		else if (tx == 32) batch_go<float, 32, true>(ctx, ab, (float*)arena);
		else batch_go<float, 16, true>(ctx, ab, (float*)arena);
	} else {
		if (tx == 8) batch_go<double, 8, false>(ctx, ab, (double*)arena);
		else if (tx == 32) batch_go<double, 32, false>(ctx, ab, (double*)arena);
		else batch_go<double, 16, false>(ctx, ab, (double*)arena);
	}
	ARMA_HIP_CHECK(hipGetLastError());
}

}  // namespace

void launch_ar_wavefront(arma_hip_ctx* ctx, const arma_ar_desc* d, void* zeta) {
	if (d->precision == ARMA_FP32)
		launch_typed<float, true>(ctx, d, (float*)zeta);
	else
		launch_typed<double, false>(ctx, d, (double*)zeta);
}

size_t ar_wavefront_work_bytes(const arma_ar_desc* d) {
	// the buffers launch_typed takes, in its order, each 256-byte aligned
	const size_t es = d->precision == ARMA_FP32 ? 4 : 8;
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int P0 = d->phi_shape[0], P1 = d->phi_shape[1], P2 = d->phi_shape[2];
	constexpr int kTX = 16, kTY = kThreads / kTX;
	auto al = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t nch = size_t(tap_chunks(P1 * P2));
	const size_t nph = (size_t((P0 + 15) / 16 * 16) + nch * P0 * kCH + 15) / 16 * 16;
	const size_t tables = al(nph * es) + nch * kCH * sizeof(int) + 4;
	const size_t ntiles = size_t(div_ceil(nx, kTX)) * div_ceil(ny, kTY);
	const size_t flags = sizeof(int) * kProgStride * ntiles + 256 + ntiles * kThreads * es;
	const size_t NB = size_t(std::max(1, mbox_width<kTX>(P1 - 1, P2 - 1)));
	const size_t total_u = size_t(nt) + kTX + kTY - 2 + (P1 - 1) + (P2 - 1);
	return al(tables) + al(flags) + es * ntiles * total_u * NB;
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
	return es * (tiles * total_u * NB + tiles * kThreads) + sizeof(int) * kProgStride * tiles;
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
	int stalled = 0;
	ARMA_HIP_CHECK(
		hipMemcpyAsync(&stalled, ab->abortw, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	ARMA_HIP_REQUIRE(stalled == 0, ARMA_HIP_ERUNTIME,
	                 "batch recurrence launch stalled (its workgroups were not all resident); "
	                 "its surfaces are invalid");
}

void ar_batch_destroy(ArBatch* ab) { delete ab; }

int ar_batch_grid(const ArBatch* ab) { return ab->grid; }

int ar_batch_max_slots(const ArBatch* ab) { return ab->max_slots; }

}  // namespace arma_hip
