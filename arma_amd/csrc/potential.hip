// Linear and high-amplitude velocity potentials on the GPU.
//
// Reference: Velocity_potential_solver<T>::operator()
// (src/velocity/basic_solver.cc:39-77) with Linear_solver<T>
// (src/velocity/linear_solver.cc:20-184): for every output time slice
//   zeta_t = -d zeta/dt                        (src/derivative.hh:8-73)
//   wnmax  = 1/kmin from factor_waves          (src/velocity/basic_solver.cc:95-112,
//                                                src/factor_waves.cc:8-89)
// and for every depth z
//   phi(t,z) = Re(IFFT2(FFT2(zeta_t) * W_rot(z))) / (nx*ny),  then edge fix-up
//   W(u,v)   = 4 pi cosh(l(z+h)) / cosh(l h),  l = 2 pi |(u,v)|, W(0,0) := W(1,1),
//   W_rot    = W cyclically half-shifted        (src/blitz.hh:54-71).
//
// GPU formulation:
//  * zeta_t is real, so F = FFT2(zeta_t) is Hermitian and
//    Re(IFFT2(F W_rot)) = C2R(F W_sym) with W_sym(k) = (W_rot(k) + W_rot(-k))/2.
//    Only the half spectrum ky = 0..M (M = ny/2) is formed.  After the
//    inverse along x the columns ky = 0 and ky = M are real; they travel
//    packed as one complex column, so every slice is M complex columns.
//  * The forward spectrum of a slice is computed once and reused for every z
//    (the reference recomputes it per z).
//  * W is evaluated as 4 pi (e^{lz} + e^{-l(z+2h)}) / (1 + e^{-2lh}), which
//    equals the reference's exp-based expression and never overflows; along
//    the equally spaced z levels both exponentials are stepped by the factor
//    e^{l dz}.  Where the reference's expression overflows to NaN (and is
//    then zeroed) the true W is below 4 pi e^{-354}; this formulation returns
//    that tiny value instead of 0.
//  * Every 1-D transform is a four-step transform of register-resident
//    short DFTs with one LDS exchange (fft_reg.hpp).  Rows of length ny are
//    real: they go through a complex transform of length M with the usual
//    even/odd split.  Twiddles are generated in FP64 on the GPU and rounded
//    to the working precision.
//  * Intermediate layouts are column-major ([column][x]) so that both the
//    x-transforms and the row transforms read and write whole cache lines.
//  * High_amplitude_solver (src/velocity/high_amplitude_solver.cc:6-23) feeds
//    the same transforms the complex multiplier F = A + iB instead of the
//    real -zeta_t.  A and B are real, so with W = Ws + Wa (even and odd parts
//    of W_rot in k) Re(IFFT2((FFT2 A + i FFT2 B) W)) = C2R(FFT2(A) Ws +
//    i FFT2(B) Wa): the product is Hermitian again, the half spectra of A
//    and B are formed by the real-row path, and only the window step of
//    the column pass differs.
//  * factor_waves is a sequential run scan per slice; it runs as a chunked
//    scan where each thread summarises a contiguous chunk of events and the
//    summaries are concatenated as a tree (exactly: representatives are
//    first-maximum picks, pairs are scored only between adjacent closed
//    runs, and the minimum and maximum of a set do not depend on the order).
#include "common.hpp"
#include "fft_reg.hpp"

#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kThreads = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kFourPi = 12.566370614359172953850573533118;

int ilog2_exact(int n) {
	int l = 0;
	while ((1 << l) < n) ++l;
	return (1 << l) == n ? l : -1;
}

__global__ void twiddle_kernel(double2* tw, int n) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q < n) {
		double s, c;
		sincospi(2.0 * double(q) / double(n), &s, &c);
		tw[q] = make_double2(c, -s);  // e^{-2 pi i q / n}
	}
}

// -dzeta/dt at (it, x, y) (src/derivative.hh:19-70), evaluated in double.
template <class T>
__device__ __forceinline__ double neg_dzdt(const T* __restrict__ zeta, int nt, size_t slice,
                                           int it, size_t off, double dt) {
	const T* f = zeta + size_t(it) * slice + off;
	double r;
	if (nt - 1 < 2) {
		r = double(f[0]);
	} else if (it == 0) {
		r = 0.5 * (-double(f[2 * slice]) + 4.0 * double(f[slice]) - 3.0 * double(f[0])) / dt;
	} else if (it == nt - 1) {
		r = 0.5 * (3.0 * double(f[0]) - 4.0 * double(f[-(ptrdiff_t)slice]) +
		           double(f[-2 * (ptrdiff_t)slice])) / dt;
	} else {
		r = 0.5 * (double(f[slice]) - double(f[-(ptrdiff_t)slice])) / dt;
	}
	return -r;
}

// d zeta/dx (step = ny) or d zeta/dy (step = 1) at f = &zeta(it, x, y), c the
// position along that dimension of extent n (src/derivative.hh:19-70).
template <class T>
__device__ __forceinline__ double dzeta_dxy(const T* __restrict__ f, int c, int n,
                                            ptrdiff_t step, double d) {
	if (n - 1 < 2) return double(f[0]);
	if (c == 0)
		return 0.5 * (-double(f[2 * step]) + 4.0 * double(f[step]) - 3.0 * double(f[0])) / d;
	if (c == n - 1)
		return 0.5 * (3.0 * double(f[0]) - 4.0 * double(f[-step]) + double(f[-2 * step])) / d;
	return 0.5 * (double(f[step]) - double(f[-step])) / d;
}

// The multiplier of src/velocity/high_amplitude_solver.cc:14-22 at one point:
// F = zt / (i ((zx + zy)/s - zx - zy) - 1/s), s = sqrt(1 + zx^2 + zy^2).
// The complex division is Smith's algorithm (R. L. Smith, CACM 5(8), 1962),
// the method of the compiler runtime's complex division on the host.
__device__ __forceinline__ cpx<double> ha_multiplier(double zt, double zx, double zy) {
	const double s = sqrt(1.0 + zx * zx + zy * zy);
	const double c = 0.0 - 1.0 / s;             // real part of the denominator
	const double d = (zx + zy) / s - zx - zy;   // imaginary part
	if (fabs(c) >= fabs(d)) {
		const double r = d / c, den = d * r + c;
		return cpx<double>{zt / den, (0.0 - zt * r) / den};
	}
	const double r = c / d, den = c * r + d;
	return cpx<double>{(zt * r) / den, (0.0 - zt) / den};
}

// The high-amplitude multiplier of every output slice as two real planes,
// field[0][i][x][y] = Re F and field[1][i][x][y] = Im F, rounded to T (the
// forward transform reads them in place of -zeta_t).  One thread per point,
// consecutive threads along y.
template <class T>
__global__ __launch_bounds__(256) void ha_field_kernel(const T* __restrict__ zeta, int nt,
                                                       int nx, int ny, int nto,
                                                       const int* __restrict__ idx_t, double dt,
                                                       double dx, double dy,
                                                       T* __restrict__ field) {
	const size_t slice = size_t(nx) * ny;
	const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const int i = blockIdx.y;
	if (q >= slice) return;
	const int x = int(q / ny), y = int(q - size_t(x) * ny);
	const int it = idx_t[i];
	const T* f = zeta + size_t(it) * slice + q;
	const cpx<double> m = ha_multiplier(-neg_dzdt(zeta, nt, slice, it, q, dt),
	                                    dzeta_dxy(f, x, nx, ny, dx), dzeta_dxy(f, y, ny, 1, dy));
	field[size_t(i) * slice + q] = T(m.x);
	field[(size_t(nto) + i) * slice + q] = T(m.y);
}

// Value of the transformed real field at (it, x, y): -zeta_t of slice it
// (the linear solver), or (src != nullptr) the plane value src[i][x][y] of
// output slice i.
template <class T>
__device__ __forceinline__ double field_value(const T* __restrict__ zeta, int nt, int nx, int ny,
                                              int it, int x, int y, double dt,
                                              const T* __restrict__ src, int i) {
	const size_t slice = size_t(nx) * ny;
	const size_t off = size_t(x) * ny + y;
	if (src) return double(src[size_t(i) * slice + off]);
	return neg_dzdt(zeta, nt, slice, it, off, dt);
}

// Exchange buffer index of the four-step transform: sequence s, output k2 of
// the first DFTs, input n1 of the second.  Rows are padded so that both the
// writes (n1 or s varying across lanes) and the reads (k2 varying) spread
// over the LDS banks.
template <int N1, int N2>
__device__ __forceinline__ int xidx(int s, int k2, int n1) {
	return s * (N2 * (N1 + 1) + 1) + k2 * (N1 + 1) + n1;
}
template <int N1, int N2>
constexpr int xsize(int nseq) {
	return nseq * (N2 * (N1 + 1) + 1);
}

// A complex value in LDS, aligned to its size so that it moves as one 8- or
// 16-byte access (two 8-byte halves would each cover only half the banks).
template <class R>
struct alignas(2 * sizeof(R)) lds_cpx {
	R x, y;
	__device__ __forceinline__ lds_cpx& operator=(cpx<R> a) {
		x = a.x;
		y = a.y;
		return *this;
	}
	__device__ __forceinline__ operator cpx<R>() const { return cpx<R>{x, y}; }
};

// In-place exchange index within one sequence's LDS row of the four-step
// transform: output k2 of the first DFTs, input n1 of the second.  n1 is
// XOR-swizzled with k2 so that both the writes (n1 varying across lanes) and
// the reads (k2 varying) fall in distinct banks of a 256-byte bank row.
template <int N1, class T>
__device__ __forceinline__ int swz(int k2, int n1) {
	constexpr int E = 256 / int(sizeof(cpx<T>));  // elements per bank row
	constexpr int D = E / N1 > 1 ? E / N1 : 1;
	return k2 * N1 + (n1 ^ ((k2 / D) & (N1 - 1)));
}

// Orders a wave's LDS stores before its later LDS loads of what other lanes
// of the same wave stored (LDS operations of one wave complete in order;
// this only keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- forward spectrum -------------------------------------------------------

// The field values at (x, 2n) and (x, 2n + 1), each as field_value computes
// it; with VEC (zeta and src aligned to two elements) every slice is read
// with one two-element load.
template <class T, bool VEC>
__device__ __forceinline__ cpx<T> field_pair(const T* __restrict__ zeta, int nt, int nx, int ny,
                                             int it, int x, int n, double dt,
                                             const T* __restrict__ src, int i) {
	if constexpr (!VEC) {
		return cpx<T>{T(field_value(zeta, nt, nx, ny, it, x, 2 * n, dt, src, i)),
		              T(field_value(zeta, nt, nx, ny, it, x, 2 * n + 1, dt, src, i))};
	} else {
		using V = std::conditional_t<sizeof(T) == 8, double2, float2>;
		const size_t slice = size_t(nx) * ny;
		const size_t off = size_t(x) * ny + 2 * size_t(n);
		auto ld = [](const T* p) { return *reinterpret_cast<const V*>(p); };
		if (src) {
			const V v = ld(src + size_t(i) * slice + off);
			return cpx<T>{v.x, v.y};
		}
		// neg_dzdt on both components
		const T* f = zeta + size_t(it) * slice + off;
		double rx, ry;
		if (nt - 1 < 2) {
			const V a = ld(f);
			rx = double(a.x);
			ry = double(a.y);
		} else if (it == 0) {
			const V c2 = ld(f + 2 * slice), c1 = ld(f + slice), c0 = ld(f);
			rx = 0.5 * (-double(c2.x) + 4.0 * double(c1.x) - 3.0 * double(c0.x)) / dt;
			ry = 0.5 * (-double(c2.y) + 4.0 * double(c1.y) - 3.0 * double(c0.y)) / dt;
		} else if (it == nt - 1) {
			const V c0 = ld(f), m1 = ld(f - slice), m2 = ld(f - 2 * slice);
			rx = 0.5 * (3.0 * double(c0.x) - 4.0 * double(m1.x) + double(m2.x)) / dt;
			ry = 0.5 * (3.0 * double(c0.y) - 4.0 * double(m1.y) + double(m2.y)) / dt;
		} else {
			const V p1 = ld(f + slice), m1 = ld(f - slice);
			rx = 0.5 * (double(p1.x) - double(m1.x)) / dt;
			ry = 0.5 * (double(p1.y) - double(m1.y)) / dt;
		}
		return cpx<T>{T(-rx), T(-ry)};
	}
}

// Rows of the real field of field_value (-zeta_t, or the plane src), real
// transform of length ny = 2M along y via a complex transform of length
// M = N1*N2.  Writes the half spectrum k = 0..M column-major: spec[i][k][x].
// R rows per block, R*N1 <= 256.
template <class T, int N1, int N2, bool VEC>
__global__ __launch_bounds__(kThreads) void fwd_rows_kernel(
	const T* __restrict__ zeta, int nt, int nx, int ny, const int* __restrict__ idx_t, double dt,
	const T* __restrict__ src, const double2* __restrict__ twy, cpx<T>* __restrict__ spec,
	int R) {
	constexpr int M = N1 * N2;
	extern __shared__ __align__(16) unsigned char smem_raw[];
	lds_cpx<T>* X = reinterpret_cast<lds_cpx<T>*>(smem_raw);
	const int i = blockIdx.y;
	const int x0 = blockIdx.x * R;
	const int it = idx_t[i];
	const int tid = threadIdx.x;
	if (tid < R * N1) {
		const int n1 = tid % N1, r = tid / N1;
		const int x = x0 + r;
		cpx<T> v[N2];
#pragma unroll
		for (int n2 = 0; n2 < N2; ++n2) {
			const int n = n1 + N1 * n2;
			v[n2] = cpx<T>{T(0), T(0)};
			if (x < nx) {
				v[n2] = field_pair<T, VEC>(zeta, nt, nx, ny, it, x, n, dt, src, i);
			}
		}
		dft<N2, -1>(v);
		twiddle_out<N2, -1>(v, twy, 2 * n1,
		                    [&](int k2, cpx<T> a) { X[xidx<N1, N2>(r, k2, n1)] = a; });
	}
	__syncthreads();
	// second DFTs (at most two tasks per lane: R*N2 <= 2*R*N1 <= 512)
	cpx<T> w[2][N1];
	int task[2];
#pragma unroll
	for (int q = 0; q < 2; ++q) {
		task[q] = tid + q * kThreads;
		if (task[q] < R * N2) {
			const int k2 = task[q] % N2, r = task[q] / N2;
#pragma unroll
			for (int n1 = 0; n1 < N1; ++n1) w[q][n1] = X[xidx<N1, N2>(r, k2, n1)];
		}
	}
	__syncthreads();
	// Z[r][k] reuses the exchange buffer, row stride M + 1
#pragma unroll
	for (int q = 0; q < 2; ++q) {
		if (task[q] < R * N2) {
			const int k2 = task[q] % N2, r = task[q] / N2;
			dft<N1, -1>(w[q]);
#pragma unroll
			for (int m = 0; m < N1; ++m) X[r * (M + 1) + k2 + N2 * brev(m, N1)] = w[q][m];
		}
	}
	__syncthreads();
	// H[k] = (Z[k] + conj Z[M-k])/2 - (i/2) e^{-i pi k/M} (Z[k] - conj Z[M-k])
	for (int t = tid; t < R * (M + 1); t += kThreads) {
		const int r = t % R, k = t / R;
		const int x = x0 + r;
		if (x >= nx) continue;
		const cpx<T> zk = X[r * (M + 1) + (k == M ? 0 : k)];
		const cpx<T> zm = conj(cpx<T>(X[r * (M + 1) + (k == 0 ? 0 : M - k)]));
		const cpx<T> e = zk + zm;
		const cpx<T> wd = twiddle<-1>(zk - zm, twy, k);
		spec[(size_t(i) * (M + 1) + k) * nx + x] =
			cpx<T>{T(0.5) * (e.x + wd.y), T(0.5) * (e.y - wd.x)};
	}
}

// Columns k = 0..M of the half spectrum, forward transform of length
// nx = N1*N2 along x, in place in spec[i][k][x].  S columns per block.
template <class T, int N1, int N2>
__global__ __launch_bounds__(kThreads) void fwd_cols_kernel(int nx, int M,
                                                             const double2* __restrict__ twx,
                                                             cpx<T>* __restrict__ spec, int S) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	lds_cpx<T>* X = reinterpret_cast<lds_cpx<T>*>(smem_raw);
	const int i = blockIdx.y;
	const int c0 = blockIdx.x * S;
	cpx<T>* base = spec + size_t(i) * (M + 1) * nx;
	const int tid = threadIdx.x;
	if (tid < S * N1) {
		const int n1 = tid % N1, s = tid / N1;
		const int col = c0 + s;
		cpx<T> v[N2];
#pragma unroll
		for (int n2 = 0; n2 < N2; ++n2)
			v[n2] = col <= M ? base[size_t(col) * nx + n1 + N1 * n2] : cpx<T>{T(0), T(0)};
		dft<N2, -1>(v);
		twiddle_out<N2, -1>(v, twx, n1,
		                    [&](int k2, cpx<T> a) { X[xidx<N1, N2>(s, k2, n1)] = a; });
	}
	__syncthreads();
	for (int t = tid; t < S * N2; t += kThreads) {
		const int k2 = t % N2, s = t / N2;
		const int col = c0 + s;
		cpx<T> w[N1];
#pragma unroll
		for (int n1 = 0; n1 < N1; ++n1) w[n1] = X[xidx<N1, N2>(s, k2, n1)];
		dft<N1, -1>(w);
		if (col <= M) {
#pragma unroll
			for (int m = 0; m < N1; ++m) base[size_t(col) * nx + k2 + N2 * brev(m, N1)] = w[m];
		}
	}
}

// ---- factor_waves (src/factor_waves.cc:103-158) as a chunked scan ----

struct Ev {
	double period, elev;
};

struct Summary {
	int runs;          // number of runs of same-sign elevation in the chunk
	int first_pos;     // class of the first run (elev > 0)
	int last_pos;      // class of the last run
	Ev first;          // representative of run 1 (complete only if runs > 1)
	Ev second;         // representative of run 2 (complete when runs > 2)
	Ev prev_last;      // representative of run runs-1
	Ev last;           // representative of the last run
	double kmin, kmax; // pairs of adjacent closed runs inside the chunk (3..runs-1)
};

__device__ __forceinline__ double pair_k(const Ev& e1, const Ev& e2) {
	const double w = kTwoPi / fabs((e1.period - e2.period) * 2.0);
	return w * w / 9.82;
}

__device__ __forceinline__ Ev first_max(const Ev& cur, const Ev& e) {
	return fabs(cur.elev) < fabs(e.elev) ? e : cur;
}

// Summary of the concatenation of chunk a and the chunk b after it.  A
// summary scores the pairs (r, r-1) of adjacent runs for 3 <= r < runs: its
// last run is still open, and run 1 may merge with the run before the chunk,
// so the pair (2, 1) is left to the caller.  The runs of a and b join where
// a's last and b's first run have the same class (the representative is the
// first maximum of the two), and the pairs that involve a's last or b's first
// two runs are scored here when their index in the concatenation is at
// least 3.  The scored pairs, and so the minimum and maximum, are those of a
// sequential scan of the concatenation.
__device__ __forceinline__ Summary concat(const Summary& a, const Summary& b) {
	if (a.runs == 0) return b;
	if (b.runs == 0) return a;
	const int na = a.runs, nb = b.runs;
	const bool merge = a.last_pos == b.first_pos;
	Summary r;
	r.runs = na + nb - (merge ? 1 : 0);
	r.first_pos = a.first_pos;
	r.last_pos = b.last_pos;
	r.kmin = fmin(a.kmin, b.kmin);
	r.kmax = fmax(a.kmax, b.kmax);
	auto score = [&](double p1, double p2) {
		const double k = pair_k(Ev{p1, 0.0}, Ev{p2, 0.0});
		r.kmin = fmin(r.kmin, k);
		r.kmax = fmax(r.kmax, k);
	};
	if (merge) {
		// runs: a_1 .. a_(na-1), J, b_2 .. b_nb
		const bool jb = fabs(a.last.elev) < fabs(b.first.elev);  // J = first_max(a.last, b.first)
		const double jp = jb ? b.first.period : a.last.period;
		const double je = jb ? b.first.elev : a.last.elev;
		if (na >= 3 && nb >= 2) score(jp, a.prev_last.period);  // pair at index na
		if (na >= 2 && nb >= 3) score(b.second.period, jp);     // pair at index na + 1
		if (na >= 2) {
			r.first = a.first;
		} else {
			r.first.period = jp;
			r.first.elev = je;
		}
		if (na >= 3) {
			r.second = a.second;
		} else if (na == 2) {
			r.second.period = jp;
			r.second.elev = je;
		} else {
			r.second = b.second;
		}
		if (nb >= 2) {
			r.last = b.last;
		} else {
			r.last.period = jp;
			r.last.elev = je;
		}
		if (nb >= 3) {
			r.prev_last = b.prev_last;
		} else if (nb == 2) {
			r.prev_last.period = jp;
			r.prev_last.elev = je;
		} else {
			r.prev_last = a.prev_last;
		}
	} else {
		// runs: a_1 .. a_na, b_1 .. b_nb
		if (na >= 3) score(a.last.period, a.prev_last.period);     // index na
		if (na >= 2 && nb >= 2) score(b.first.period, a.last.period);  // index na + 1
		if (nb >= 3) score(b.second.period, b.first.period);       // index na + 2
		r.first = a.first;
		if (na >= 2)
			r.second = a.second;
		else
			r.second = b.first;
		r.last = b.last;
		if (nb >= 2)
			r.prev_last = b.prev_last;
		else
			r.prev_last = a.last;
	}
	return r;
}

// Positions whose loads a column scan issues together.
constexpr int kFwAhead = 16;

// Columns of the row tile staged in LDS by the row scan.
__host__ __device__ constexpr int fw_tile(int lines) { return lines > 512 ? 8 : 16; }

// One block per (slice, direction) with one thread per scanned line (the
// block size, a power of two, is the larger line count): thread b scans line b (dir 0: column y = b along x; dir 1: row x = b along y), so
// its positions are exactly the contiguous chunk [b*inner, (b+1)*inner) of
// the sequential scan order.  Column scans read coalesced rows of the slice;
// row scans go through LDS tiles of the slice's columns.
template <class T>
__global__ __launch_bounds__(1024) void factor_waves_kernel(const T* __restrict__ zeta, int nx,
                                                           int ny, const int* __restrict__ idx_t,
                                                           const int* __restrict__ t_int,
                                                           double dt, double* __restrict__ kout) {
	extern __shared__ __align__(16) unsigned char fw_smem[];
	Summary* sm = reinterpret_cast<Summary*>(fw_smem);  // after the scan
	T* tile = reinterpret_cast<T*>(fw_smem);  // during a row scan, in the input's type
	const int i = blockIdx.x;
	const int dir = blockIdx.y;
	const T* z = zeta + size_t(idx_t[i]) * nx * ny;
	const double t = double(t_int[i]);
	const int n0 = dir == 0 ? nx : ny;  // scanned extent
	const int n1 = dir == 0 ? ny : nx;  // lines, one per thread
	const int b = threadIdx.x;
	const bool has_line = b < n1;  // the block spans the larger line count
	Summary s;
	s.runs = 0;
	s.first_pos = s.last_pos = 0;
	s.first = s.second = s.prev_last = s.last = Ev{0, 0};
	s.kmin = 1.7976931348623157e308;
	s.kmax = 2.2250738585072014e-308;
	Ev cur{0, 0}, prev{0, 0};
	int cur_pos = 0;
	// the event test and run update at one position (z1, z0, z2 = a-1, a, a+1)
	auto visit = [&](double z1, double z0, double z2) {
		const double dw1 = z0 - z1;
		const double dw2 = z0 - z2;
		if ((dw1 > 0.0) != (dw2 > 0.0)) return;
		const double aa = -0.5 * (dw1 + dw2) / (dt * dt);
		const double bb = dw1 / dt - aa * dt * (2.0 * t - 1.0);
		const double cc = z0 - t * dt * (aa * t * dt + bb);
		const double tex = -0.5 * bb / aa;
		const Ev e{tex, cc + tex * (bb + aa * tex)};
		const int pos = e.elev > 0.0;
		if (s.runs == 0) {
			s.runs = 1;
			s.first_pos = pos;
			cur = e;
			cur_pos = pos;
			return;
		}
		if (pos == cur_pos) {
			cur = first_max(cur, e);
		} else {
			if (s.runs == 1) {
				s.first = cur;
			} else if (s.runs == 2) {
				s.second = cur;
			} else {
				const double k = pair_k(cur, prev);
				s.kmin = fmin(s.kmin, k);
				s.kmax = fmax(s.kmax, k);
			}
			prev = cur;
			cur = e;
			cur_pos = pos;
			++s.runs;
		}
	};
	if (dir == 0) {
		if (n0 >= 3 && has_line) {
			double zm = double(z[b]), zc = double(z[size_t(ny) + b]);
			// the loads of kFwAhead positions are issued together, ahead of
			// their (sequential) run updates
			for (int a0 = 1; a0 + 1 < n0; a0 += kFwAhead) {
				double zp[kFwAhead];
#pragma unroll
				for (int u = 0; u < kFwAhead; ++u) {
					const int a = min(a0 + u + 1, n0 - 1);
					zp[u] = double(z[size_t(a) * ny + b]);
				}
#pragma unroll
				for (int u = 0; u < kFwAhead; ++u) {
					if (a0 + u + 1 < n0) {
						visit(zm, zc, zp[u]);
						zm = zc;
						zc = zp[u];
					}
				}
			}
		}
	} else {
		const int TW = fw_tile(n1);
		double zm = 0.0, zc = 0.0;
		for (int c0 = 0; c0 < n0; c0 += TW) {
			__syncthreads();
			// a thread's loads of the tile (at most TW <= 16: n1 <= blockDim.x)
			// are issued eight at a time, then stored
#pragma unroll 1
			for (int u0 = 0; u0 < 16; u0 += 8) {
				T tv[8];
#pragma unroll
				for (int u = 0; u < 8; ++u) {
					const int e = b + (u0 + u) * int(blockDim.x);
					const int row = e / TW, col = e - row * TW;
					tv[u] = e < n1 * TW && c0 + col < n0 ? z[size_t(row) * ny + c0 + col] : T(0);
				}
#pragma unroll
				for (int u = 0; u < 8; ++u) {
					const int e = b + (u0 + u) * int(blockDim.x);
					const int row = e / TW, col = e - row * TW;
					if (e < n1 * TW && c0 + col < n0) tile[row * (TW + 1) + col] = tv[u];
				}
			}
			__syncthreads();
			const int cw = has_line ? min(TW, n0 - c0) : 0;
			if (cw == 16) {
				// a whole tile row: its LDS reads are issued together
				double zr[16];
#pragma unroll
				for (int col = 0; col < 16; ++col) zr[col] = double(tile[b * (TW + 1) + col]);
#pragma unroll
				for (int col = 0; col < 16; ++col) {
					if (c0 + col >= 2) visit(zm, zc, zr[col]);
					zm = zc;
					zc = zr[col];
				}
			} else {
				for (int col = 0; col < cw; ++col) {
					const double zp = double(tile[b * (TW + 1) + col]);
					const int c = c0 + col;  // position of zp along the row
					if (c >= 2) visit(zm, zc, zp);
					zm = zc;
					zc = zp;
				}
			}
		}
		__syncthreads();  // the tile area becomes the summary array
	}
	if (s.runs == 1) s.first = cur;
	if (s.runs == 2) s.second = cur;
	s.last = cur;
	s.last_pos = cur_pos;
	s.prev_last = prev;
	// the summaries of the lines, in scan order, combined as a tree
	sm[threadIdx.x] = s;
	for (int w = 1; w < int(blockDim.x); w *= 2) {
		__syncthreads();
		const int t = threadIdx.x;
		if ((t & (2 * w - 1)) == 0 && t + w < int(blockDim.x)) sm[t] = concat(sm[t], sm[t + w]);
	}
	if (threadIdx.x == 0) {
		const Summary& q = sm[0];
		double kmin = q.kmin, kmax = q.kmax;
		// the pair (run 2, run 1), which a summary leaves to its caller
		if (q.runs >= 3) {
			const double k = pair_k(q.second, q.first);
			kmin = fmin(kmin, k);
			kmax = fmax(kmax, k);
		}
		kout[4 * i + dir] = kmin;
		kout[4 * i + 2 + dir] = kmax;
	}
}

// ---- window -----------------------------------------------------------------

// Rotate map of blitz::rotate(W, shape/2): swaps j <-> j + n/2 for j < n/2.
__device__ __forceinline__ int rot_index(int j, int n) {
	const int m = n / 2;
	if (j < m) return j + m;
	if (j < 2 * m) return j - m;
	return j;
}

// W at one wave number along nz equally spaced levels z_j = zs + j dz, in
// the working precision R (FP64, or FP32 as the reference's FP32 build
// evaluates it): W_j = scale (a_j + c_j), a_j = e^{l z_j},
// c_j = e^{-l (z_j + 2h)}.
template <class R>
struct WinParts {
	// exponent below which a_j is taken as zero (e^x under 1e-304 in FP64,
	// 1e-35 in FP32), and the step l |dz| above which every level is
	// evaluated directly
	static constexpr double kLow = sizeof(R) == 4 ? -80.0 : -700.0;
	static constexpr double kDirect = sizeof(R) == 4 ? 60.0 : 600.0;
	R a, c, rho, irho, scale, l;
	int jstart;   // a_j = 0 for j < jstart
	bool direct;  // step too large: evaluate the exponentials directly

	__device__ __forceinline__ void init(int px, int py, double du, double dv, double zs,
	                                     double dz, double h, int nz) {
		if (px == 0 && py == 0) {
			px = 1;
			py = 1;
		}
		const R u = R(du * px), v = R(dv * py);
		l = R(kTwoPi) * sqrt(u * u + v * v);
		const R e2h = exp(R(-2) * l * R(h));
		scale = R(kFourPi) / (R(1) + e2h);
		direct = double(l) * fabs(dz) > kDirect;
		rho = exp(l * R(dz));
		irho = R(1) / rho;
		// first level where l z > kLow
		int js = 0;
		if (double(l) * zs < kLow && dz > 0.0) {
			js = int(ceil((kLow / double(l) - zs) / dz));
			if (js < 0) js = 0;
		}
		jstart = js;
		a = js < nz ? exp(l * R(zs + dz * js)) : R(0);
		// c_0 = e^{-l (zs + 2h)} = e^{-2lh} / e^{l zs} while both factors are
		// in range (a few ulps from the direct exponential), else directly
		const bool split = js == 0 && !direct && double(l) * fabs(zs) < kDirect &&
		                   double(l) * 2.0 * h < kDirect;
		c = split ? e2h / a : exp(-l * R(zs + 2.0 * h));
	}
	// W_j for j = 0, 1, 2, ... in order
	__device__ __forceinline__ R next(int j, double zs, double dz, double h) {
		R av, cv;
		if (direct) {
			const R zj = R(zs + dz * j);
			av = exp(l * zj);
			cv = exp(-l * (zj + R(2.0 * h)));
		} else {
			av = j < jstart ? R(0) : a;
			cv = c;
			if (j >= jstart) a *= rho;
			c *= irho;
		}
		return scale * (av + cv);
	}
	// Whether every level from j on takes the recurrence with a_j nonzero,
	// and W_j then, for j = j0, j0 + 1, ... in order once it does: the same
	// operations as next, without its per-level tests.
	__device__ __forceinline__ bool plain_from(int j0) const { return !direct && jstart <= j0; }
	__device__ __forceinline__ R step() {
		const R v = scale * (a + c);
		a *= rho;
		c *= irho;
		return v;
	}
};

// The reference replaces every non-finite window value by zero
// (src/velocity/linear_solver.cc:176-183).  Here W only overflows, where
// z > 0 and l z is large: above ~708 in FP64, where the reference's FP64
// expression is non-finite too, and above ~88 in FP32, where the value does
// not fit the working type (as the reference's FP32 expression).
template <class T>
__device__ __forceinline__ double finite_or_zero(double w) {
	return isfinite(T(w)) ? w : 0.0;
}

// Symmetric and antisymmetric parts of the window at one level from
// W_rot(k) = va and W_rot(-k) = vb.  CHECKED (a level range where W may
// exceed the working type): non-finite values become zero first and the
// halves are added (the sum of two values near the top of the range would
// overflow).  Otherwise 0.5 (va +- vb): the same bits for any pair of
// values below the top of the range.
template <class T, bool CHECKED>
__device__ __forceinline__ void win_pair(double va, double vb, T& w, T& wd) {
	if constexpr (CHECKED) {
		va = finite_or_zero<T>(va);
		vb = finite_or_zero<T>(vb);
		w = T(0.5 * va + 0.5 * vb);
		wd = T(0.5 * va - 0.5 * vb);
	} else {
		w = T(0.5 * (va + vb));
		wd = T(0.5 * (va - vb));
	}
}

// Whether W may exceed the working type at some level z in [zlo, zhi]:
// W is below 4 pi e^{l z} + 4 pi, so l z <= 80 keeps it under FLT_MAX and
// l z <= 700 keeps it (and a sum of two) under DBL_MAX.
template <class T>
__device__ __forceinline__ bool window_may_overflow(double l, double zlo, double zhi) {
	const double zmax = fmax(zlo, zhi);
	return zmax > 0.0 && l * zmax > (sizeof(T) == 4 ? 80.0 : 700.0);
}

// ---- inverse transforms -------------------------------------------------------

// Levels per seed of the window recurrence: W at level j is reached from the
// exponentials at the last multiple of kWinSpan by at most kWinSpan - 1 steps.
constexpr int kWinSpan = 16;

// Column task: packed columns c0..c0+CB-1 of slice i (column 0 carries
// ky = 0 and ky = M), times W_sym(z) for every level, inverse transform of
// length nx = N1*N2 along x.  Output Gs[z][c][x].  Levels go in groups of ZB
// through LDS; ZB*CB*N1 <= 256, and ZB divides kWinSpan or covers all levels.
// For transforms of up to 2*256 points the recurrence state of the window runs
// on from one group to the next and is seeded at every multiple of kWinSpan,
// so the values do not depend on ZB; longer transforms seed every group.
template <class T, int N1, int N2>
__device__ __forceinline__ void col_task(lds_cpx<T>* X, int ny, int M, int nz, int i, int c0,
                                         const double2* __restrict__ twx,
                                         const cpx<T>* __restrict__ spec,
                                         const cpx<T>* __restrict__ specb,
                                         const double* __restrict__ kmin, double z0, double dz,
                                         double h, cpx<T>* Gs, int CB, int ZB) {
	constexpr int N = N1 * N2;
	constexpr int EPT = N > kThreads ? N / kThreads : 1;  // CB*N <= max(256, N)
	const int tid = threadIdx.x;
	const int nel = CB * N;
	const size_t sbase = size_t(i) * (M + 1) * N;
	const bool ha = specb != nullptr;
	// the window in the working precision (FP32 for FP32 transforms)
	using WR = std::conditional_t<sizeof(T) == 4, float, double>;
	cpx<T> f0[EPT], f1[EPT], g0[EPT], g1[EPT];
#pragma unroll
	for (int e = 0; e < EPT; ++e) {
		const int q = tid + e * kThreads;
		f0[e] = f1[e] = g0[e] = g1[e] = cpx<T>{T(0), T(0)};
		if (q < nel) {
			const int c = q / N, kx = q % N;
			f0[e] = spec[sbase + size_t(c0 + c) * N + kx];
			if (c0 + c == 0) f1[e] = spec[sbase + size_t(M) * N + kx];
			if (ha) {
				g0[e] = specb[sbase + size_t(c0 + c) * N + kx];
				if (c0 + c == 0) g1[e] = specb[sbase + size_t(M) * N + kx];
			}
		}
	}
	const double wx = 1.0 / kmin[4 * i + 0];
	const double wy = 1.0 / kmin[4 * i + 1];
	const double du = wx / double(N - 1 > 1 ? N - 1 : 1);
	const double dv = wy / double(ny - 1 > 1 ? ny - 1 : 1);
	// every wave number of the slice is at most 2 pi |(wx, wy)|
	const double lmax = kTwoPi * sqrt(wx * wx + wy * wy);
	WinParts<WR> wpa[EPT], wpb[EPT];  // the main column's window, per element
	for (int j0 = 0; j0 < nz; j0 += ZB) {
		const int zb = min(ZB, nz - j0);
		const double zs = z0 + dz * j0;
		// the seed of this group's levels (with more than two elements per
		// lane the carried state would not fit the registers: every group is
		// seeded at its first level)
		constexpr bool kCarry = EPT <= 2;
		const int jseed = kCarry ? j0 - j0 % kWinSpan : j0;
		const int span = kCarry ? min(kWinSpan, nz - jseed) : zb;
		const double zseed = z0 + dz * jseed;
		// the same for the whole workgroup: one code path per level group
		const bool checked = window_may_overflow<T>(lmax, zs, zs + dz * (zb - 1));
		// F * W_sym for the levels of this group
#pragma unroll
		for (int e = 0; e < EPT; ++e) {
			const int q = tid + e * kThreads;
			if (q < nel) {
				const int c = q / N, kx = q % N;
				const int ky = c0 + c;
				const int kxm = kx == 0 ? 0 : N - kx;
				lds_cpx<T>* dst = X + c * N + kx;  // level zz at dst[zz * CB * N]
				{
					WinParts<WR>& wa = wpa[e];
					WinParts<WR>& wb = wpb[e];
					if (j0 == jseed) {
						wa.init(rot_index(kx, N), rot_index(ky, ny), du, dv, zseed, dz, h, span);
						wb.init(rot_index(kxm, N), rot_index(ky == 0 ? 0 : ny - ky, ny), du, dv,
						        zseed, dz, h, span);
					}
					const int jz = j0 - jseed;
					auto put = [&](int zz, T w, T wd) {
						if (ha)
							dst[zz * CB * N] =
								cpx<T>{f0[e].x * w - g0[e].y * wd, f0[e].y * w + g0[e].x * wd};
						else
							dst[zz * CB * N] = cpx<T>{f0[e].x * w, f0[e].y * w};
					};
					auto levels = [&](auto checked) {
						// the usual case, decided per wave: both windows step by
						// the plain recurrence at every level of the group
						if (__all(wa.plain_from(jz) && wb.plain_from(jz))) {
							for (int zz = 0; zz < zb; ++zz) {
								T w, wd;
								win_pair<T, decltype(checked)::value>(wa.step(), wb.step(), w, wd);
								put(zz, w, wd);
							}
							return;
						}
						for (int zz = 0; zz < zb; ++zz) {
							T w, wd;
							win_pair<T, decltype(checked)::value>(
								wa.next(jz + zz, zseed, dz, h), wb.next(jz + zz, zseed, dz, h), w,
								wd);
							put(zz, w, wd);
						}
					};
					if (checked)
						levels(std::true_type{});
					else
						levels(std::false_type{});
				}
				if (ky == 0) {
					// packed column: add i * F(kx, M) W_sym(kx, M); its window is
					// seeded again in every group and stepped to the group's first
					// level (one workgroup in M / CB has this column)
					WinParts<WR> pa, pb;
					pa.init(rot_index(kx, N), rot_index(M, ny), du, dv, zseed, dz, h, span);
					pb.init(rot_index(kxm, N), rot_index(M, ny), du, dv, zseed, dz, h, span);
					const int jz = j0 - jseed;
					for (int zz = 0; zz < jz; ++zz) {
						pa.next(zz, zseed, dz, h);
						pb.next(zz, zseed, dz, h);
					}
					auto levels = [&](auto checked) {
						for (int zz = 0; zz < zb; ++zz) {
							T w, wd;
							win_pair<T, decltype(checked)::value>(
								pa.next(jz + zz, zseed, dz, h), pb.next(jz + zz, zseed, dz, h), w,
								wd);
							const cpx<T> a = dst[zz * CB * N];
							if (ha) {
								const cpx<T> hm{f1[e].x * w - g1[e].y * wd,
								                f1[e].y * w + g1[e].x * wd};
								dst[zz * CB * N] = cpx<T>{a.x - hm.y, a.y + hm.x};
							} else {
								dst[zz * CB * N] = cpx<T>{a.x - f1[e].y * w, a.y + f1[e].x * w};
							}
						}
					};
					if (checked)
						levels(std::true_type{});
					else
						levels(std::false_type{});
				}
			}
		}
		__syncthreads();
		// Sequence s = tid / N1 lies in the LDS row X[s N .. s N + N - 1] and
		// its N1 lanes belong to one wave (N1 divides 64), so from here on
		// each wave transforms its own sequences in place, without a
		// workgroup barrier.  The lanes' loads of a sequence are complete
		// before any of them stores into it: every store depends on all of
		// the loaded values.
		const int nseq = zb * CB;
		const bool act = tid < nseq * N1;
		const int n1 = tid % N1, s = tid / N1;
		if (act) {
			cpx<T> v[N2];
#pragma unroll
			for (int n2 = 0; n2 < N2; ++n2) v[n2] = X[s * N + n1 + N1 * n2];
			dft<N2, +1>(v);
			twiddle_out<N2, +1>(v, twx, n1,
			                    [&](int k2, cpx<T> a) { X[s * N + swz<N1, T>(k2, n1)] = a; });
		}
		wave_lds_sync();
		// the same wave's sequences, one (sequence, k2) per lane
		constexpr int SPW = 64 / N1;  // sequences per wave
		const int lane = tid % 64, seq0 = (tid / 64) * SPW;
#pragma unroll
		for (int it = 0; it < SPW * N2 / 64; ++it) {
			const int u = lane + 64 * it;
			const int sq = seq0 + u / N2, k2 = u % N2;
			if (sq < nseq) {
				const int zz = sq / CB, c = sq % CB;
				cpx<T> w[N1];
#pragma unroll
				for (int m = 0; m < N1; ++m) w[m] = X[sq * N + swz<N1, T>(k2, m)];
				dft<N1, +1>(w);
				cpx<T>* out = Gs + (size_t(j0 + zz) * M + c0 + c) * N;
#pragma unroll
				for (int m = 0; m < N1; ++m) out[k2 + N2 * brev(m, N1)] = w[m];
			}
		}
		// the next group's window step overwrites every row
		if (j0 + ZB < nz) __syncthreads();
	}
}

// Row task: rows x0..x0+R-1 of one (slice, level): real output of length
// ny = 2M from the M packed columns g[k][x], through a complex transform of
// length M = N1*N2, scaled by 1/(nx ny).  R*N1 <= 256.
template <class T, int N1, int N2>
__device__ __forceinline__ void row_task(lds_cpx<T>* X, int nx, int ny, int x0,
                                         const double2* __restrict__ twy, const cpx<T>* g,
                                         T* __restrict__ out, int R, T* __restrict__ E) {
	constexpr int M = N1 * N2;
	const int tid = threadIdx.x;
	const bool act = tid < R * N1;
	const int r = tid % R, n1 = tid / R;
	// rows past nx (none for the launch shapes used) transform a copy of the
	// last row and are not stored
	const int x = min(x0 + r, nx - 1);
	cpx<T> hk[N2];
	double2 tk[N2];
	if (act) {
		// every load is issued before any arithmetic, with no branch between
		// them, so that all of them are in flight at once
#pragma unroll
		for (int n2 = 0; n2 < N2; ++n2) {
			const int k = n1 + N1 * n2;
			hk[n2] = g[size_t(k) * nx + x];
			tk[n2] = twy[k];
		}
		// the rows are staged in LDS: the partner H[M-k] of a lane's H[k] is
		// held by another lane and is read from there, not from memory again
#pragma unroll
		for (int n2 = 0; n2 < N2; ++n2) X[r * (M + 1) + n1 + N1 * n2] = hk[n2];
	}
	__syncthreads();
	cpx<T> v[N2];
	if (act) {
#pragma unroll
		for (int n2 = 0; n2 < N2; ++n2) {
			const int k = n1 + N1 * n2;
			// Z[k] = (H[k] + conj H[M-k]) + i e^{i pi k/M} (H[k] - conj H[M-k])
			const cpx<T> a = hk[n2], b = conj(cpx<T>(X[r * (M + 1) + (k == 0 ? 0 : M - k)]));
			const cpx<T> sum = a + b;
			const cpx<T> wd = cmul(a - b, T(tk[n2].x), T(-tk[n2].y));
			v[n2] = cpx<T>{sum.x - wd.y, sum.y + wd.x};
		}
		// packed column k = 0: H[0] = real part, H[M] = imaginary part
		if (n1 == 0) v[0] = cpx<T>{hk[0].x + hk[0].y, hk[0].x - hk[0].y};
	}
	__syncthreads();  // the exchange below overwrites the staged rows
	if (act) {
		dft<N2, +1>(v);
		twiddle_out<N2, +1>(v, twy, 2 * n1,
		                    [&](int k2, cpx<T> a) { X[xidx<N1, N2>(r, k2, n1)] = a; });
	}
	__syncthreads();
	const T inv = T(1.0 / (double(nx) * double(ny)));
	for (int t = tid; t < R * N2; t += kThreads) {
		const int k2 = t % N2, r = t / N2;
		const int x = x0 + r;
		cpx<T> w[N1];
#pragma unroll
		for (int n1 = 0; n1 < N1; ++n1) w[n1] = X[xidx<N1, N2>(r, k2, n1)];
		dft<N1, +1>(w);
		if (x < nx) {
			// columns 1 and 2 of the row also go to the edge fix-up's array
			if (k2 == 0) E[2 * size_t(x)] = w[0].y * inv;
			if (k2 == 1) E[2 * size_t(x) + 1] = w[0].x * inv;
#pragma unroll
			for (int m = 0; m < N1; ++m) {
				const int n = k2 + N2 * brev(m, N1);
				*reinterpret_cast<cpx<T>*>(out + size_t(x) * ny + 2 * n) =
					cpx<T>{w[m].x * inv, w[m].y * inv};
			}
		}
	}
}

// Plane through three grid points evaluated at a fourth
// (src/interpolate.cc:3-29, src/interpolate.hh:17-33).
__device__ double interp3v(double fa, double fb, double fc, int a0, int a1, int b0, int b1,
                           int c0, int c1, int x0, int x1) {
	const double den = double(b0 * a1 - c0 * a1 - a0 * b1 + c0 * b1 + a0 * c1 - b0 * c1);
	const double p0 = (fb * a1 - fc * a1 - fa * b1 + fc * b1 + fa * c1 - fb * c1) / den;
	const double p1 = (fc * (-a0 + b0) + fb * (a0 - c0) + fa * (-b0 + c0)) /
	                  double(c0 * (a1 - b1) + a0 * (b1 - c1) + b0 * (-a1 + c1));
	const double p2 = (fc * b0 * a1 - fb * c0 * a1 - fc * a0 * b1 + fa * c0 * b1 +
	                   fb * a0 * c1 - fa * b0 * c1) / den;
	return p0 * x0 + p1 * x1 + p2;
}

template <class T>
__device__ double interp3(const T* f, int ny, int a0, int a1, int b0, int b1, int c0, int c1,
                          int x0, int x1) {
	return interp3v(double(f[size_t(a0) * ny + a1]), double(f[size_t(b0) * ny + b1]),
	                double(f[size_t(c0) * ny + c1]), a0, a1, b0, b1, c0, c1, x0, x1);
}

// Edge fix-up of src/velocity/basic_solver.cc:61-68 on one (slice, level)
// plane f: (0,0), then column 0 (row k from columns 1 and 2 of rows k-1 and
// k), then row 0 (from rows 1 and 2, and for l = 1 the fixed column values),
// each point a plane through three neighbours.  Columns 1.. of rows 1.. never
// change.  The row pass writes the transformed plane in full and, beside it,
// columns 1 and 2 of every row contiguously (E[x][0..1]); this kernel (one
// thread per point: column 0 at k = 1..nx-1, (0,0), row 0 at l = 1..ny-1)
// takes column 1 and 2 values from E only, so the column points read no
// value the kernel writes, and row 0 at l = 1 evaluates the fixed (1,0) and
// (2,0) itself.
template <class T>
__global__ void edge_kernel(int nx, int ny, T* __restrict__ phi, const T* __restrict__ edge,
                            int nplanes) {
	const int per = nx + ny - 1;  // points per plane
	const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= (long long)nplanes * per) return;
	const int plane = int(q / per), u = int(q % per);
	T* f = phi + size_t(plane) * nx * ny;
	const T* E = edge + size_t(plane) * nx * 2;
	// column 0 at row k from E
	auto col0 = [&](int k) {
		return interp3v(double(E[2 * (k - 1)]), double(E[2 * k]), double(E[2 * (k - 1) + 1]),
		                k - 1, 1, k, 1, k - 1, 2, k, 0);
	};
	if (u < nx - 1) {
		const int k = u + 1;
		f[size_t(k) * ny] = T(col0(k));
	} else if (u == nx - 1) {
		f[0] = T(interp3v(double(E[2]), double(E[3]), double(E[4]), 1, 1, 1, 2, 2, 1, 0, 0));
	} else {
		const int l = u - (nx - 1);
		if (l == 1) {
			const double f10 = double(T(col0(1))), f20 = double(T(col0(2)));
			f[1] = T(interp3v(f10, double(E[2]), f20, 1, 0, 1, 1, 2, 0, 0, 1));
		} else {
			f[l] = T(interp3(f, ny, 1, l - 1, 1, l, 2, l - 1, 0, l));
		}
	}
}

// Column pass over a chunk of output slices: block (cb, il) runs the column
// task of slice i0 + il, columns cb*CB...; G holds [il][z][c][x].
template <class T, int N1, int N2>
__global__ __launch_bounds__(kThreads) void inv_cols_kernel(
	int ny, int M, int nz, int i0, const double2* __restrict__ twx,
	const cpx<T>* __restrict__ spec, const cpx<T>* __restrict__ specb,
	const double* __restrict__ kmin, double z0, double dz, double h, cpx<T>* __restrict__ G,
	int CB, int ZB) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	lds_cpx<T>* X = reinterpret_cast<lds_cpx<T>*>(smem_raw);
	const int il = blockIdx.y;
	col_task<T, N1, N2>(X, ny, M, nz, i0 + il, blockIdx.x * CB, twx, spec, specb, kmin, z0, dz,
	                    h, G + size_t(il) * nz * M * (N1 * N2), CB, ZB);
}

// Row pass over a chunk: block (rb, item) runs the row task of
// (slice i0 + item / nz, level item % nz), rows rb*R...
template <class T, int N1, int N2>
__global__ __launch_bounds__(kThreads) void inv_rows_kernel(int nx, int ny, int nz, int i0,
                                                             const double2* __restrict__ twy,
                                                             const cpx<T>* __restrict__ G,
                                                             T* __restrict__ phi,
                                                             T* __restrict__ edge, int R) {
	extern __shared__ __align__(16) unsigned char smem_raw[];
	lds_cpx<T>* X = reinterpret_cast<lds_cpx<T>*>(smem_raw);
	const int item = blockIdx.y;
	const int il = item / nz, j = item - il * nz;
	const int x0 = blockIdx.x * R;
	const size_t plane = size_t(i0 + il) * nz + j;
	row_task<T, N1, N2>(X, nx, ny, x0, twy, G + size_t(item) * (N1 * N2) * nx,
	                    phi + plane * nx * ny, R, edge + plane * nx * 2);
}

// ---- launch plumbing ------------------------------------------------------------

constexpr int split1(int n) {
	int l = 0;
	while ((1 << l) < n) ++l;
	return 1 << (l / 2);
}

template <class K>
void allow_lds(K kernel, size_t lds) {
	if (lds > 65536)
		ARMA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
}

struct PotPlan {
	int nt, nx, ny, M, nto, nz;
	double dt, z0, dz, h;
	double dx, dy;    // grid spacing along x and y (derivatives of the high-amplitude solver)
	int solver;       // 0 linear, 1 high amplitude
	int nfields() const { return solver == 1 ? 2 : 1; }
};

// Rows per row-pass workgroup for nx rows of ny = 2M: R * split1(M) lanes.
int rows_per_task(int nx, int M) { return std::min(nx, kThreads / split1(M)); }

// Length-dependent launches, instantiated for every power of two.
template <class T, int NX>
struct ColLaunch {
	static constexpr int N1 = split1(NX), N2 = NX / N1;

	static void forward(arma_hip_ctx* ctx, const PotPlan& p, const double2* twx, cpx<T>* spec) {
		const int S = std::min(p.M + 1, kThreads / N1);
		const size_t lds = sizeof(cpx<T>) * size_t(xsize<N1, N2>(S));
		allow_lds(&fwd_cols_kernel<T, N1, N2>, lds);
		// the half spectra of every field are contiguous: one launch covers them
		fwd_cols_kernel<T, N1, N2><<<dim3(div_ceil(p.M + 1, S), p.nto * p.nfields()), kThreads,
		                             lds, ctx->stream>>>(NX, p.M, twx, spec, S);
		ARMA_HIP_CHECK(hipGetLastError());
	}

};

template <class T, int M>
struct RowLaunch {
	static constexpr int N1 = split1(M), N2 = M / N1;

	static int rows(const PotPlan& p) { return std::min(p.nx, kThreads / N1); }

	static void forward(arma_hip_ctx* ctx, const PotPlan& p, const T* zeta, const int* didx,
	                    const double2* twy, cpx<T>* spec, const T* field) {
		const int R = rows(p);
		const size_t lds = sizeof(cpx<T>) * size_t(std::max(xsize<N1, N2>(R), R * (M + 1)));
		// two-element loads where the field's base allows them (a surface may
		// start at any element of a caller's buffer)
		const bool vec = reinterpret_cast<uintptr_t>(zeta) % (2 * sizeof(T)) == 0 &&
		                 reinterpret_cast<uintptr_t>(field) % (2 * sizeof(T)) == 0;
		auto kernel = vec ? &fwd_rows_kernel<T, N1, N2, true> : &fwd_rows_kernel<T, N1, N2, false>;
		allow_lds(kernel, lds);
		// linear: -zeta_t; high amplitude: the planes Re F, then Im F, into
		// consecutive blocks of spec
		const size_t sblock = size_t(p.nto) * (M + 1) * p.nx;
		const size_t fblock = size_t(p.nto) * p.nx * p.ny;
		for (int f = 0; f < p.nfields(); ++f) {
			kernel<<<dim3(div_ceil(p.nx, R), p.nto), kThreads, lds, ctx->stream>>>(
				zeta, p.nt, p.nx, p.ny, didx, p.dt, field ? field + f * fblock : nullptr, twy,
				spec + f * sblock, R);
			ARMA_HIP_CHECK(hipGetLastError());
		}
	}

};

// Bytes of the column-transformed intermediate per chunk of output slices
// (ARMA_HIP_POT_CHUNK_MB overrides).
size_t chunk_bytes() {
	if (const char* e = std::getenv("ARMA_HIP_POT_CHUNK_MB")) {
		const long v = std::atol(e);
		if (v > 0) return size_t(v) << 20;
	}
	return size_t(8) << 30;
}

// LDS bytes of a column-pass workgroup at most.  Fewer levels per group
// (less LDS, more resident workgroups) measured slower: the transform steps
// then leave lanes idle.
constexpr size_t kColLdsCap = 81920;

// Column and row passes for nx = NX and ny = 2M over chunks of output slices.
template <class T, int NX, int M>
void launch_inverse(arma_hip_ctx* ctx, const PotPlan& p, const double2* twx, const double2* twy,
                    const cpx<T>* spec, const double* dk, T* phi, T* edge) {
	constexpr int CN1 = split1(NX), CN2 = NX / CN1, RN1 = split1(M), RN2 = M / RN1;
	const int CB = std::max(1, std::min(M, kThreads / NX));
	static_assert(CN1 <= 64 && CN2 >= CN1, "a wave holds whole sequences");
	const size_t per_seq = sizeof(cpx<T>) * size_t(NX);
	// a power of two that divides kWinSpan, or all the levels
	int ZB = std::min(kWinSpan, std::max(1, kThreads / CN1 / CB));
	while (ZB > 1 && size_t(ZB) * CB * per_seq > kColLdsCap) ZB /= 2;
	if (p.nz <= ZB) ZB = p.nz;
	const size_t lds_c = size_t(ZB) * CB * per_seq;
	const int R = rows_per_task(NX, M);
	const size_t lds_r = sizeof(cpx<T>) * size_t(xsize<RN1, RN2>(R));
	allow_lds(&inv_cols_kernel<T, CN1, CN2>, lds_c);
	allow_lds(&inv_rows_kernel<T, RN1, RN2>, lds_r);
	// the column-transformed intermediate of a chunk of output slices goes
	// from the column pass to the row pass through device memory.  With more
	// than one chunk, two intermediate buffers alternate and the row pass of
	// chunk k (on the context's second stream) overlaps the column pass of
	// chunk k + 1 (on its main stream)
	const size_t per_slice = sizeof(cpx<T>) * size_t(p.nz) * M * NX;
	const int chunk = std::min(p.nto, int(std::max<size_t>(1, chunk_bytes() / per_slice)));
	const int nchunks = div_ceil(p.nto, chunk);
	const size_t gsize = per_slice * chunk;
	unsigned char* gbase = (unsigned char*)ctx->pot_mid.get(gsize * (nchunks > 1 ? 2 : 1));
	hipStream_t rows_stream = ctx->stream;
	hipEvent_t* ev = ctx->pot_events;  // [0, 1]: columns done, [2, 3]: rows done
	if (nchunks > 1) {
		if (!ctx->aux_stream)
			ARMA_HIP_CHECK(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
		for (int e = 0; e < 4; ++e)
			if (!ev[e]) ARMA_HIP_CHECK(hipEventCreateWithFlags(&ev[e], hipEventDisableTiming));
		rows_stream = ctx->aux_stream;
	}
	for (int k = 0; k < nchunks; ++k) {
		const int i0 = k * chunk, cn = std::min(chunk, p.nto - i0), b = k & 1;
		cpx<T>* G = (cpx<T>*)(gbase + gsize * (nchunks > 1 ? b : 0));
		// the column pass reuses buffer b once the rows of chunk k - 2 are done
		if (k >= 2) ARMA_HIP_CHECK(hipStreamWaitEvent(ctx->stream, ev[2 + b], 0));
		{
			KernelTimer t(ctx, "pot_inv_cols");
			inv_cols_kernel<T, CN1, CN2><<<dim3(M / CB, cn), kThreads, lds_c, ctx->stream>>>(
				p.ny, M, p.nz, i0, twx, spec,
				p.solver == 1 ? spec + size_t(p.nto) * (M + 1) * NX : nullptr, dk, p.z0, p.dz, p.h,
				G, CB, ZB);
			ARMA_HIP_CHECK(hipGetLastError());
		}
		if (nchunks > 1) {
			ARMA_HIP_CHECK(hipEventRecord(ev[b], ctx->stream));
			ARMA_HIP_CHECK(hipStreamWaitEvent(rows_stream, ev[b], 0));
		}
		{
			KernelTimer t(ctx, "pot_inv_rows", rows_stream);
			inv_rows_kernel<T, RN1, RN2><<<dim3(div_ceil(NX, R), cn * p.nz), kThreads, lds_r,
			                               rows_stream>>>(NX, 2 * M, p.nz, i0, twy, G, phi, edge,
			                                              R);
			ARMA_HIP_CHECK(hipGetLastError());
		}
		if (nchunks > 1) ARMA_HIP_CHECK(hipEventRecord(ev[2 + b], rows_stream));
	}
	// what follows on the main stream reads every row pass's output
	if (nchunks > 1)
		ARMA_HIP_CHECK(hipStreamWaitEvent(ctx->stream, ev[2 + ((nchunks - 1) & 1)], 0));
}

template <class T, int NX, int M = 2>
void inverse_by_m(int m, arma_hip_ctx* ctx, const PotPlan& p, const double2* twx,
                  const double2* twy, const cpx<T>* spec, const double* dk, T* phi, T* edge) {
	if constexpr (M <= 512) {
		if (m == M) {
			launch_inverse<T, NX, M>(ctx, p, twx, twy, spec, dk, phi, edge);
			return;
		}
		inverse_by_m<T, NX, M * 2>(m, ctx, p, twx, twy, spec, dk, phi, edge);
	} else {
		ARMA_HIP_REQUIRE(false, ARMA_HIP_EINVAL, "transform length must be a power of two");
	}
}

template <class T, int NX = 4>
void inverse_dispatch(int nx, int m, arma_hip_ctx* ctx, const PotPlan& p, const double2* twx,
                      const double2* twy, const cpx<T>* spec, const double* dk, T* phi,
                      T* edge) {
	if constexpr (NX <= 1024) {
		if (nx == NX) {
			inverse_by_m<T, NX>(m, ctx, p, twx, twy, spec, dk, phi, edge);
			return;
		}
		inverse_dispatch<T, NX * 2>(nx, m, ctx, p, twx, twy, spec, dk, phi, edge);
	} else {
		ARMA_HIP_REQUIRE(false, ARMA_HIP_EINVAL, "transform length must be a power of two");
	}
}

// Calls F<L>::run(args...) for the power of two L = n in [2, 1024].
template <template <int> class F, int L = 2, class... A>
void by_length(int n, A&&... args) {
	if constexpr (L <= 1024) {
		if (n == L) {
			F<L>::run(std::forward<A>(args)...);
			return;
		}
		by_length<F, L * 2>(n, std::forward<A>(args)...);
	} else {
		ARMA_HIP_REQUIRE(false, ARMA_HIP_EINVAL, "transform length must be a power of two");
	}
}

template <class T>
struct Ops {
	template <int L>
	struct ColFwd {
		static void run(arma_hip_ctx* c, const PotPlan& p, const double2* tw, cpx<T>* s) {
			if constexpr (L >= 4) ColLaunch<T, L>::forward(c, p, tw, s);
		}
	};
	template <int L>
	struct RowFwd {
		static void run(arma_hip_ctx* c, const PotPlan& p, const T* z, const int* di,
		                const double2* tw, cpx<T>* s, const T* field) {
			if constexpr (L <= 512) RowLaunch<T, L>::forward(c, p, z, di, tw, s, field);
		}
	};
};

template <class T>
void run_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const T* zeta, T* phi,
                   double* wnmax_out, int solver) {
	PotPlan p;
	p.solver = solver;
	p.dx = d->delta[1];
	p.dy = d->delta[2];
	p.nt = d->shape[0];
	p.nx = d->shape[1];
	p.ny = d->shape[2];
	const int lx = ilog2_exact(p.nx), ly = ilog2_exact(p.ny);
	ARMA_HIP_REQUIRE(lx >= 2 && ly >= 2 && p.nx <= 1024 && p.ny <= 1024, ARMA_HIP_EINVAL,
	                 "surface extents along x and y must be powers of two in [4, 1024]");
	p.M = p.ny / 2;
	p.nto = d->npoints[0];
	p.nz = d->npoints[1];
	p.dt = d->delta[0];
	p.z0 = d->lbound[1];
	p.h = d->depth;
	const double dtd = (d->ubound[0] - d->lbound[0]) / double(std::max(p.nto - 1, 1));
	p.dz = (d->ubound[1] - d->lbound[1]) / double(std::max(p.nz - 1, 1));
	std::vector<int> idx(p.nto), tint(p.nto);
	for (int i = 0; i < p.nto; ++i) {
		const double tv = d->lbound[0] + dtd * i;
		idx[i] = int(tv);
		ARMA_HIP_REQUIRE(idx[i] >= 0 && idx[i] < p.nt, ARMA_HIP_EINVAL,
		                 "potential domain time index outside the surface");
		tint[i] = int(d->delta[0] * idx[i]);
	}
	const hipStream_t st = ctx->stream;
	// twiddles (FP64, generated on the device)
	double2* twx = (double2*)ctx->pot_twiddle.get(sizeof(double2) * size_t(p.nx + p.ny));
	double2* twy = twx + p.nx;
	twiddle_kernel<<<div_ceil(p.nx, 256), 256, 0, st>>>(twx, p.nx);
	twiddle_kernel<<<div_ceil(p.ny, 256), 256, 0, st>>>(twy, p.ny);
	int* didx = (int*)ctx->pot_wn.get(sizeof(int) * 2 * size_t(p.nto));
	int* dtint = didx + p.nto;
	std::vector<int> both(idx);
	both.insert(both.end(), tint.begin(), tint.end());
	upload(ctx, didx, both.data(), sizeof(int) * 2 * size_t(p.nto));
	double* dk = (double*)ctx->pot_kmin.get(sizeof(double) * 4 * size_t(p.nto));
	{
		KernelTimer t_fw(ctx, "pot_factor_waves");
		const int lines = std::max(p.nx, p.ny);
		const size_t fw_lds = std::max(sizeof(Summary) * size_t(lines),
		                               sizeof(T) * size_t(p.nx) * (fw_tile(p.nx) + 1));
		allow_lds(&factor_waves_kernel<T>, fw_lds);
		// both directions in one launch: the block size is the larger line count and
		// threads past a direction's line count only take part in barriers
		factor_waves_kernel<T><<<dim3(p.nto, 2), lines, fw_lds, st>>>(zeta, p.nx, p.ny, didx, dtint,
		                                                       d->delta[0], dk);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	// forward half spectra of every output slice: spec[f][i][k][kx], k = 0..M,
	// one field f for the linear solver, two (A, B) for the high-amplitude one
	cpx<T>* spec = (cpx<T>*)ctx->pot_spec.get(sizeof(cpx<T>) * size_t(p.nfields()) * p.nto *
	                                          (p.M + 1) * p.nx);
	{
		KernelTimer t_fwd(ctx, "pot_forward");
		T* field = nullptr;
		if (p.solver == 1) {
			const size_t slice = size_t(p.nx) * p.ny;
			field = (T*)ctx->pot_field.get(sizeof(T) * 2 * p.nto * slice);
			ha_field_kernel<T><<<dim3(div_ceil(int(slice), 256), p.nto), 256, 0, st>>>(
				zeta, p.nt, p.nx, p.ny, p.nto, didx, p.dt, p.dx, p.dy, field);
			ARMA_HIP_CHECK(hipGetLastError());
		}
		by_length<Ops<T>::template RowFwd>(p.M, ctx, p, zeta, (const int*)didx,
		                                   (const double2*)twy, spec, (const T*)field);
		by_length<Ops<T>::template ColFwd>(p.nx, ctx, p, (const double2*)twx, spec);
	}
	// every (slice, level): window, inverse along x, inverse along y
	// columns 1 and 2 of every row of every plane, for the edge fix-up
	T* edge = (T*)ctx->pot_edge.get(sizeof(T) * 2 * size_t(p.nto) * p.nz * p.nx);
	inverse_dispatch<T>(p.nx, p.M, ctx, p, (const double2*)twx, (const double2*)twy,
	                    (const cpx<T>*)spec, (const double*)dk, phi, edge);
	{
		KernelTimer t_ed(ctx, "pot_edge");
		const long long n = (long long)p.nto * p.nz * (p.nx + p.ny - 1);
		edge_kernel<T><<<int((n + 255) / 256), 256, 0, st>>>(p.nx, p.ny, phi, edge, p.nto * p.nz);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	if (wnmax_out) {
		std::vector<double> hk(4 * size_t(p.nto));
		ARMA_HIP_CHECK(hipMemcpyAsync(hk.data(), dk, sizeof(double) * 4 * p.nto,
		                              hipMemcpyDeviceToHost, st));
		ARMA_HIP_CHECK(hipStreamSynchronize(st));
		for (int i = 0; i < p.nto; ++i) {
			wnmax_out[2 * i] = 1.0 / hk[4 * i];
			wnmax_out[2 * i + 1] = 1.0 / hk[4 * i + 1];
		}
	}
}

}  // namespace

void launch_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta, void* phi,
                      double* wnmax_out, int solver) {
	if (d->precision == ARMA_FP32)
		run_potential<float>(ctx, d, (const float*)zeta, (float*)phi, wnmax_out, solver);
	else
		run_potential<double>(ctx, d, (const double*)zeta, (double*)phi, wnmax_out, solver);
}

void launch_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta,
                             void* phi, double* wnmax_out) {
	launch_potential(ctx, d, zeta, phi, wnmax_out, 0);
}

}  // namespace arma_hip
