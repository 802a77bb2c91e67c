// Linear velocity potential on the GPU.
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
//  * zeta_t is real, so FFT2(zeta_t) is Hermitian and
//    Re(IFFT2(F W_rot)) = C2R(F W_sym) with W_sym(k) = (W_rot(k) + W_rot(-k))/2.
//    Only the half spectrum ky = 0..ny/2 is transformed and stored.
//  * The forward spectrum of a slice is computed once and reused for every z
//    (the reference recomputes it per z).
//  * W is evaluated as 4 pi (e^{lz} + e^{-l(z+2h)}) / (1 + e^{-2lh}), which
//    equals the reference's exp-based expression and never overflows; along
//    the equally spaced z levels both exponentials are stepped by the factor
//    e^{l dz}.  Where the reference's expression overflows to NaN (and is
//    then zeroed) the true W is below 4 pi e^{-354}; this formulation returns
//    that tiny value instead of 0.
//  * 1-D transforms are radix-2 Stockham passes in LDS over batches of rows
//    or columns; twiddles are generated in FP64 on the GPU and rounded to the
//    working precision.
//  * factor_waves is a sequential run scan per slice; it runs as a chunked
//    scan where each thread summarises a contiguous chunk of events and one
//    thread folds the summaries in order (the fold is exact: representatives
//    are first-maximum picks, pairs are scored only between adjacent closed
//    runs).
#include "common.hpp"

#include <cmath>
#include <vector>

namespace arma_hip {
namespace {

template <class R>
struct cpx {
	R x, y;
};

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

template <class R>
__device__ __forceinline__ cpx<R> twiddled(const cpx<R>& a, const double2* __restrict__ tw, int idx,
                                           int sign) {
	const double2 w = tw[idx];
	const R wr = R(w.x);
	const R wi = sign < 0 ? R(w.y) : R(-w.y);
	return cpx<R>{a.x * wr - a.y * wi, a.x * wi + a.y * wr};
}

// Stockham transforms (radix-4 passes, one radix-2 pass when log2 n is odd)
// of nb sequences of length n = 2^log2n held in LDS.  sign < 0: forward
// e^{-i}, sign > 0: backward e^{+i} (unnormalised), as
// gsl_fft_complex_transform.  tw[q] = e^{-2 pi i q / n}, q < n.  Returns the
// buffer that holds the result.
template <class R>
__device__ cpx<R>* fft_lds(cpx<R>* x, cpx<R>* y, int n, int log2n, int nb,
                           const double2* __restrict__ tw, int sign) {
	int ls = 0;  // log2 of the current sub-transform length Ns
	while (log2n - ls >= 2) {
		const int Ns = 1 << ls;
		const int quarter = n >> 2;
		const int total = nb * quarter;
		const int tstep = log2n - 2 - ls;  // twiddle stride n / (4 Ns)
		for (int q = threadIdx.x; q < total; q += blockDim.x) {
			const int b = q >> (log2n - 2);
			const int j = q & (quarter - 1);
			const cpx<R>* src = x + b * n + j;
			cpx<R> a0 = src[0];
			cpx<R> a1 = src[quarter];
			cpx<R> a2 = src[2 * quarter];
			cpx<R> a3 = src[3 * quarter];
			const int k = j & (Ns - 1);
			if (k) {
				a1 = twiddled(a1, tw, k << tstep, sign);
				a2 = twiddled(a2, tw, (2 * k) << tstep, sign);
				a3 = twiddled(a3, tw, (3 * k) << tstep, sign);
			}
			const cpx<R> b0{a0.x + a2.x, a0.y + a2.y};
			const cpx<R> b1{a0.x - a2.x, a0.y - a2.y};
			const cpx<R> b2{a1.x + a3.x, a1.y + a3.y};
			// (a1 - a3) times -i (forward) or +i (backward)
			const cpx<R> d13{a1.x - a3.x, a1.y - a3.y};
			const cpx<R> b3 = sign < 0 ? cpx<R>{d13.y, -d13.x} : cpx<R>{-d13.y, d13.x};
			cpx<R>* dst = y + b * n + ((j >> ls) << (ls + 2)) + k;
			dst[0] = cpx<R>{b0.x + b2.x, b0.y + b2.y};
			dst[Ns] = cpx<R>{b1.x + b3.x, b1.y + b3.y};
			dst[2 * Ns] = cpx<R>{b0.x - b2.x, b0.y - b2.y};
			dst[3 * Ns] = cpx<R>{b1.x - b3.x, b1.y - b3.y};
		}
		__syncthreads();
		cpx<R>* t = x;
		x = y;
		y = t;
		ls += 2;
	}
	if (ls < log2n) {
		const int Ns = 1 << ls;
		const int half = n >> 1;
		const int total = nb * half;
		for (int q = threadIdx.x; q < total; q += blockDim.x) {
			const int b = q >> (log2n - 1);
			const int j = q & (half - 1);
			const cpx<R> a0 = x[b * n + j];
			const int k = j & (Ns - 1);
			const cpx<R> a1 = twiddled(x[b * n + j + half], tw, k, sign);
			const int d = ((j >> ls) << (ls + 1)) + k;
			y[b * n + d] = cpx<R>{a0.x + a1.x, a0.y + a1.y};
			y[b * n + d + Ns] = cpx<R>{a0.x - a1.x, a0.y - a1.y};
		}
		__syncthreads();
		cpx<R>* t = x;
		x = y;
		y = t;
	}
	return x;
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

// Rows of -zeta_t, transformed along y; the half row ky = 0..ny/2 is kept.
template <class T>
__global__ void fwd_rows_kernel(const T* __restrict__ zeta, int nt, int nx, int ny, int log2ny,
                                const int* __restrict__ idx_t, double dt,
                                const double2* __restrict__ twy, cpx<T>* __restrict__ spec,
                                int rows) {
	extern __shared__ unsigned char smem_raw[];
	cpx<T>* b0 = reinterpret_cast<cpx<T>*>(smem_raw);
	cpx<T>* b1 = b0 + rows * ny;
	const int i = blockIdx.y;
	const int x0 = blockIdx.x * rows;
	const int it = idx_t[i];
	const size_t slice = size_t(nx) * ny;
	for (int q = threadIdx.x; q < rows * ny; q += blockDim.x) {
		const int r = q / ny, y = q - r * ny;
		const int x = x0 + r;
		T v = 0;
		if (x < nx) v = T(neg_dzdt(zeta, nt, slice, it, size_t(x) * ny + y, dt));
		b0[q] = cpx<T>{v, T(0)};
	}
	__syncthreads();
	cpx<T>* res = fft_lds(b0, b1, ny, log2ny, rows, twy, -1);
	const int nyh = ny / 2 + 1;
	for (int q = threadIdx.x; q < rows * nyh; q += blockDim.x) {
		const int r = q / nyh, ky = q - r * nyh;
		const int x = x0 + r;
		if (x < nx) spec[(size_t(i) * nx + x) * nyh + ky] = res[r * ny + ky];
	}
}

// Columns of the half spectrum, transformed along x in place.
template <class T>
__global__ void fwd_cols_kernel(int nx, int log2nx, int nyh, const double2* __restrict__ twx,
                                cpx<T>* __restrict__ spec, int cols) {
	extern __shared__ unsigned char smem_raw[];
	cpx<T>* b0 = reinterpret_cast<cpx<T>*>(smem_raw);
	cpx<T>* b1 = b0 + cols * nx;
	const int i = blockIdx.y;
	const int c0 = blockIdx.x * cols;
	cpx<T>* base = spec + size_t(i) * nx * nyh;
	for (int q = threadIdx.x; q < cols * nx; q += blockDim.x) {
		const int kx = q / cols, c = q - kx * cols;
		const int ky = c0 + c;
		b0[c * nx + kx] = ky < nyh ? base[size_t(kx) * nyh + ky] : cpx<T>{T(0), T(0)};
	}
	__syncthreads();
	cpx<T>* res = fft_lds(b0, b1, nx, log2nx, cols, twx, -1);
	for (int q = threadIdx.x; q < cols * nx; q += blockDim.x) {
		const int kx = q / cols, c = q - kx * cols;
		const int ky = c0 + c;
		if (ky < nyh) base[size_t(kx) * nyh + ky] = res[c * nx + kx];
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

template <class T>
__global__ __launch_bounds__(256) void factor_waves_kernel(const T* __restrict__ zeta, int nx,
                                                          int ny, const int* __restrict__ idx_t,
                                                          const int* __restrict__ t_int,
                                                          double dt, double* __restrict__ kout) {
	__shared__ Summary sm[256];
	const int i = blockIdx.x;
	const int dir = blockIdx.y;
	const T* z = zeta + size_t(idx_t[i]) * nx * ny;
	const double t = double(t_int[i]);
	// view: Z(a, b) with a scanned (1..n0-2) inside b (0..n1-1)
	const int n0 = dir == 0 ? nx : ny;
	const int n1 = dir == 0 ? ny : nx;
	const int inner = n0 - 2;
	const long long P = inner > 0 ? (long long)inner * n1 : 0;
	const long long L = (P + blockDim.x - 1) / blockDim.x;
	const long long p0 = (long long)threadIdx.x * L;
	const long long p1 = min(P, p0 + L);
	Summary s;
	s.runs = 0;
	s.kmin = 1.7976931348623157e308;
	s.kmax = 2.2250738585072014e-308;
	Ev cur{0, 0}, prev{0, 0};
	int cur_pos = 0;
	for (long long p = p0; p < p1; ++p) {
		const int b = int(p / inner);
		const int a = int(p - (long long)b * inner) + 1;
		double z0, z1, z2;
		if (dir == 0) {
			z0 = double(z[size_t(a) * ny + b]);
			z1 = double(z[size_t(a - 1) * ny + b]);
			z2 = double(z[size_t(a + 1) * ny + b]);
		} else {
			z0 = double(z[size_t(b) * ny + a]);
			z1 = double(z[size_t(b) * ny + a - 1]);
			z2 = double(z[size_t(b) * ny + a + 1]);
		}
		const double dw1 = z0 - z1;
		const double dw2 = z0 - z2;
		if ((dw1 > 0.0) != (dw2 > 0.0)) continue;
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
			continue;
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
	}
	if (s.runs == 1) s.first = cur;
	if (s.runs == 2) s.second = cur;
	s.last = cur;
	s.last_pos = cur_pos;
	s.prev_last = prev;
	sm[threadIdx.x] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		bool have = false;
		int cls = 0, j = 0;
		Ev e1{0, 0}, e2{0, 0};
		double kmin = 1.7976931348623157e308, kmax = 2.2250738585072014e-308;
		auto transition = [&](const Ev& nxt, int ncls) {
			if (j > 0) {
				const double k = pair_k(e1, e2);
				kmin = fmin(kmin, k);
				kmax = fmax(kmax, k);
			}
			e2 = e1;
			e1 = nxt;
			cls = ncls;
			++j;
		};
		for (int c = 0; c < blockDim.x; ++c) {
			const Summary& q = sm[c];
			if (q.runs == 0) continue;
			if (!have) {
				have = true;
				cls = q.first_pos;
				e1 = q.first;
			} else if (q.first_pos == cls) {
				e1 = first_max(e1, q.first);
			} else {
				transition(q.first, q.first_pos);
			}
			if (q.runs == 2) {
				transition(q.last, q.last_pos);
			} else if (q.runs >= 3) {
				transition(q.second, 1 - cls);
				if (q.runs == 3) {
					transition(q.last, q.last_pos);
				} else {
					// entering run 3 scores (run 2, merged run 1); the pairs
					// (3,2) .. (runs-1, runs-2) were scored inside the chunk
					const double k = pair_k(e1, e2);
					kmin = fmin(fmin(kmin, k), q.kmin);
					kmax = fmax(fmax(kmax, k), q.kmax);
					e2 = q.prev_last;
					e1 = q.last;
					cls = q.last_pos;
					j += q.runs - 2;
				}
			}
		}
		kout[4 * i + dir] = kmin;
		kout[4 * i + 2 + dir] = kmax;
	}
}

// Rotate map of blitz::rotate(W, shape/2): swaps j <-> j + n/2 for j < n/2.
__device__ __forceinline__ int rot_index(int j, int n) {
	const int m = n / 2;
	if (j < m) return j + m;
	if (j < 2 * m) return j - m;
	return j;
}

struct WinParts {
	double a, c, rho, irho, scale;  // W_j = scale * (a_j + c_j)
	int jstart;                     // a_j = 0 for j < jstart (e^{l z} below 1e-304)
	bool direct;                    // step too large: evaluate exponentials directly
	double l;
};

__device__ __forceinline__ WinParts win_parts(int px, int py, double du, double dv, double z0,
                                              double dz, double h, int nz) {
	if (px == 0 && py == 0) {
		px = 1;
		py = 1;
	}
	const double u = du * px, v = dv * py;
	WinParts w;
	w.l = kTwoPi * sqrt(u * u + v * v);
	const double l = w.l;
	w.scale = kFourPi / (1.0 + exp(-2.0 * l * h));
	w.direct = l * fabs(dz) > 600.0;
	w.c = exp(-l * (z0 + 2.0 * h));
	w.rho = exp(l * dz);
	w.irho = exp(-l * dz);
	// first level where l z > -700
	int js = 0;
	if (l * z0 < -700.0 && dz > 0.0) {
		js = int(ceil((-700.0 / l - z0) / dz));
		if (js < 0) js = 0;
	}
	w.jstart = js;
	w.a = js < nz ? exp(l * (z0 + dz * js)) : 0.0;
	return w;
}

// Per element of a column block: multiply by W_sym(z) for every z and
// transform along x (backward).  Output G[(i, z)][x][ky].
template <class T, int EPT>
__global__ __launch_bounds__(256) void inv_cols_kernel(
	int nx, int log2nx, int ny, int nyh, int nz, int i0, const double2* __restrict__ twx,
	const cpx<T>* __restrict__ spec, const double* __restrict__ kmin, double z0, double dz,
	double h, cpx<T>* __restrict__ G, int cols) {
	extern __shared__ unsigned char smem_raw[];
	cpx<T>* fin = reinterpret_cast<cpx<T>*>(smem_raw);
	cpx<T>* b0 = fin + cols * nx;
	cpx<T>* b1 = b0 + cols * nx;
	const int il = blockIdx.y;
	const int i = i0 + il;
	const int c0 = blockIdx.x * cols;
	const cpx<T>* base = spec + size_t(i) * nx * nyh;
	for (int q = threadIdx.x; q < cols * nx; q += blockDim.x) {
		const int kx = q / cols, c = q - kx * cols;
		const int ky = c0 + c;
		fin[c * nx + kx] = ky < nyh ? base[size_t(kx) * nyh + ky] : cpx<T>{T(0), T(0)};
	}
	const double wx = 1.0 / kmin[4 * i + 0];
	const double wy = 1.0 / kmin[4 * i + 1];
	const double du = wx / double(nx - 1 > 1 ? nx - 1 : 1);
	const double dv = wy / double(ny - 1 > 1 ? ny - 1 : 1);
	WinParts wp[EPT][2];
	const int nel = cols * nx;
#pragma unroll
	for (int e = 0; e < EPT; ++e) {
		const int q = threadIdx.x + e * blockDim.x;
		if (q < nel) {
			const int c = q / nx, kx = q - c * nx;
			const int ky = c0 + c;
			const int kxm = kx == 0 ? 0 : nx - kx;
			const int kym = ky == 0 ? 0 : ny - ky;
			wp[e][0] = win_parts(rot_index(kx, nx), rot_index(ky, ny), du, dv, z0, dz, h, nz);
			wp[e][1] = win_parts(rot_index(kxm, nx), rot_index(kym % ny, ny), du, dv, z0, dz, h, nz);
		}
	}
	__syncthreads();
	for (int j = 0; j < nz; ++j) {
#pragma unroll
		for (int e = 0; e < EPT; ++e) {
			const int q = threadIdx.x + e * blockDim.x;
			if (q < nel) {
				double wsum = 0.0;
#pragma unroll
				for (int s = 0; s < 2; ++s) {
					WinParts& w = wp[e][s];
					double av, cv;
					if (w.direct) {
						const double zj = z0 + dz * j;
						av = exp(w.l * zj);
						cv = exp(-w.l * (zj + 2.0 * h));
					} else {
						av = j < w.jstart ? 0.0 : w.a;
						cv = w.c;
						if (j >= w.jstart) w.a *= w.rho;
						w.c *= w.irho;
					}
					wsum += w.scale * (av + cv);
				}
				const T ws = T(0.5 * wsum);
				const cpx<T> f = fin[q];
				b0[q] = cpx<T>{f.x * ws, f.y * ws};
			}
		}
		__syncthreads();
		cpx<T>* res = fft_lds(b0, b1, nx, log2nx, cols, twx, +1);
		cpx<T>* out = G + (size_t(il) * nz + j) * nx * nyh;
		for (int q = threadIdx.x; q < nel; q += blockDim.x) {
			const int kx = q / cols, c = q - kx * cols;
			const int ky = c0 + c;
			if (ky < nyh) out[size_t(kx) * nyh + ky] = res[c * nx + kx];
		}
		__syncthreads();
	}
}

// Rows: complex-to-real backward transform along y, scaled by 1/(nx ny).
template <class T>
__global__ void inv_rows_kernel(int nx, int ny, int log2ny, int nyh, int nz, int i0,
                                const double2* __restrict__ twy, const cpx<T>* __restrict__ G,
                                T* __restrict__ phi, int rows) {
	extern __shared__ unsigned char smem_raw[];
	cpx<T>* b0 = reinterpret_cast<cpx<T>*>(smem_raw);
	cpx<T>* b1 = b0 + rows * ny;
	const int item = blockIdx.y;  // (il, z) pair
	const int x0 = blockIdx.x * rows;
	const cpx<T>* g = G + size_t(item) * nx * nyh;
	for (int q = threadIdx.x; q < rows * ny; q += blockDim.x) {
		const int r = q / ny, y = q - r * ny;
		const int x = x0 + r;
		cpx<T> v{T(0), T(0)};
		if (x < nx) {
			if (y < nyh) {
				v = g[size_t(x) * nyh + y];
			} else {
				const cpx<T> u = g[size_t(x) * nyh + (ny - y)];
				v = cpx<T>{u.x, -u.y};
			}
		}
		b0[q] = v;
	}
	__syncthreads();
	cpx<T>* res = fft_lds(b0, b1, ny, log2ny, rows, twy, +1);
	const int il = item / nz, j = item - il * nz;
	T* out = phi + ((size_t(i0 + il) * nz + j) * nx) * ny;
	const double inv = 1.0 / (double(nx) * double(ny));
	for (int q = threadIdx.x; q < rows * ny; q += blockDim.x) {
		const int r = q / ny, y = q - r * ny;
		const int x = x0 + r;
		if (x < nx) out[size_t(x) * ny + y] = T(double(res[q].x) * inv);
	}
}

// Plane through three grid points evaluated at a fourth
// (src/interpolate.cc:3-29, src/interpolate.hh:17-33).
template <class T>
__device__ double interp3(const T* f, int ny, int a0, int a1, int b0, int b1, int c0, int c1,
                          int x0, int x1) {
	const double fa = double(f[size_t(a0) * ny + a1]);
	const double fb = double(f[size_t(b0) * ny + b1]);
	const double fc = double(f[size_t(c0) * ny + c1]);
	const double den = double(b0 * a1 - c0 * a1 - a0 * b1 + c0 * b1 + a0 * c1 - b0 * c1);
	const double p0 = (fb * a1 - fc * a1 - fa * b1 + fc * b1 + fa * c1 - fb * c1) / den;
	const double p1 = (fc * (-a0 + b0) + fb * (a0 - c0) + fa * (-b0 + c0)) /
	                  double(c0 * (a1 - b1) + a0 * (b1 - c1) + b0 * (-a1 + c1));
	const double p2 = (fc * b0 * a1 - fb * c0 * a1 - fc * a0 * b1 + fa * c0 * b1 +
	                   fb * a0 * c1 - fa * b0 * c1) / den;
	return p0 * x0 + p1 * x1 + p2;
}

// Edge fix-up of src/velocity/basic_solver.cc:61-68: (0,0), then column 0
// (reads columns 1 and 2 only), then row 0 (reads rows 1 and 2, and for
// l = 1 the fixed column values).
template <class T>
__global__ void edge_kernel(int nx, int ny, T* __restrict__ phi, int i0, int nz) {
	const int item = blockIdx.x;
	const int il = item / nz, j = item - il * nz;
	T* f = phi + ((size_t(i0 + il) * nz + j) * nx) * ny;
	if (threadIdx.x == 0) f[0] = T(interp3(f, ny, 1, 1, 1, 2, 2, 1, 0, 0));
	for (int k = 1 + threadIdx.x; k < nx; k += blockDim.x)
		f[size_t(k) * ny] = T(interp3(f, ny, k - 1, 1, k, 1, k - 1, 2, k, 0));
	__syncthreads();
	for (int l = 1 + threadIdx.x; l < ny; l += blockDim.x)
		f[l] = T(interp3(f, ny, 1, l - 1, 1, l, 2, l - 1, 0, l));
}

// Largest batch of length-n transforms whose LDS buffers fit in 64 KiB.
int lds_batch(int n, size_t elem, int buffers, int cap) {
	const size_t per = size_t(buffers) * elem * size_t(n);
	int b = int(65536 / per);
	return std::max(1, std::min(cap, b));
}

template <class T>
void run_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const T* zeta, T* phi,
                   double* wnmax_out) {
	const int nt = d->shape[0], nx = d->shape[1], ny = d->shape[2];
	const int lx = ilog2_exact(nx), ly = ilog2_exact(ny);
	ARMA_HIP_REQUIRE(lx >= 2 && ly >= 2 && nx <= 1024 && ny <= 1024, ARMA_HIP_EINVAL,
	                 "surface extents along x and y must be powers of two in [4, 1024]");
	const int nto = d->npoints[0], nz = d->npoints[1];
	const int nyh = ny / 2 + 1;
	const double dtd = (d->ubound[0] - d->lbound[0]) / double(std::max(nto - 1, 1));
	const double dz = (d->ubound[1] - d->lbound[1]) / double(std::max(nz - 1, 1));
	std::vector<int> idx(nto), tint(nto);
	for (int i = 0; i < nto; ++i) {
		const double tv = d->lbound[0] + dtd * i;
		idx[i] = int(tv);
		ARMA_HIP_REQUIRE(idx[i] >= 0 && idx[i] < nt, ARMA_HIP_EINVAL,
		                 "potential domain time index outside the surface");
		tint[i] = int(d->delta[0] * idx[i]);
	}
	const hipStream_t st = ctx->stream;
	// twiddles (FP64, generated on the device)
	double2* twx = (double2*)ctx->pot_twiddle.get(sizeof(double2) * size_t(nx + ny));
	double2* twy = twx + nx;
	twiddle_kernel<<<div_ceil(nx, 256), 256, 0, st>>>(twx, nx);
	twiddle_kernel<<<div_ceil(ny, 256), 256, 0, st>>>(twy, ny);
	int* didx = (int*)ctx->pot_wn.get(sizeof(int) * 2 * size_t(nto));
	int* dtint = didx + nto;
	std::vector<int> both(idx);
	both.insert(both.end(), tint.begin(), tint.end());
	upload(ctx, didx, both.data(), sizeof(int) * 2 * size_t(nto));
	double* dk = (double*)ctx->pot_kmin.get(sizeof(double) * 4 * size_t(nto));
	KernelTimer t_fw(ctx, "pot_factor_waves");
	factor_waves_kernel<T><<<dim3(nto, 2), 256, 0, st>>>(zeta, nx, ny, didx, dtint, d->delta[0],
	                                                     dk);
	ARMA_HIP_CHECK(hipGetLastError());
	// forward half spectra of every output slice
	cpx<T>* spec = (cpx<T>*)ctx->pot_spec.get(sizeof(cpx<T>) * size_t(nto) * nx * nyh);
	{
		const int rows = lds_batch(ny, sizeof(cpx<T>), 2, 8);
		const size_t lds = 2 * sizeof(cpx<T>) * size_t(rows) * ny;
		KernelTimer t_fwd(ctx, "pot_forward");
		fwd_rows_kernel<T><<<dim3(div_ceil(nx, rows), nto), 256, lds, st>>>(
			zeta, nt, nx, ny, ly, didx, d->delta[0], twy, spec, rows);
		ARMA_HIP_CHECK(hipGetLastError());
		const int cols = lds_batch(nx, sizeof(cpx<T>), 2, 8);
		const size_t lds2 = 2 * sizeof(cpx<T>) * size_t(cols) * nx;
		fwd_cols_kernel<T><<<dim3(div_ceil(nyh, cols), nto), 256, lds2, st>>>(nx, lx, nyh, twx,
		                                                                       spec, cols);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	// inverse transforms in chunks of output slices
	const size_t per_slice = size_t(nz) * nx * nyh * sizeof(cpx<T>);
	int chunk = int(std::max<size_t>(1, (size_t(1) << 30) / per_slice));
	chunk = std::min(chunk, nto);
	cpx<T>* G = (cpx<T>*)ctx->pot_mid.get(per_slice * chunk);
	const int cols = lds_batch(nx, sizeof(cpx<T>), 3, std::max(1, 1024 / nx));
	const int ept = div_ceil(cols * nx, 256);
	ARMA_HIP_REQUIRE(ept <= 4, ARMA_HIP_ERUNTIME, "column batch too large");
	const size_t lds_c = 3 * sizeof(cpx<T>) * size_t(cols) * nx;
	const int rows = lds_batch(ny, sizeof(cpx<T>), 2, 8);
	const size_t lds_r = 2 * sizeof(cpx<T>) * size_t(rows) * ny;
	for (int i0 = 0; i0 < nto; i0 += chunk) {
		const int cn = std::min(chunk, nto - i0);
		const dim3 gc(div_ceil(nyh, cols), cn);
		{
		KernelTimer t_ic(ctx, "pot_inv_cols");
		switch (ept) {
		case 1:
			inv_cols_kernel<T, 1><<<gc, 256, lds_c, st>>>(nx, lx, ny, nyh, nz, i0, twx, spec, dk,
			                                             d->lbound[1], dz, d->depth, G, cols);
			break;
		case 2:
			inv_cols_kernel<T, 2><<<gc, 256, lds_c, st>>>(nx, lx, ny, nyh, nz, i0, twx, spec, dk,
			                                             d->lbound[1], dz, d->depth, G, cols);
			break;
		default:
			inv_cols_kernel<T, 4><<<gc, 256, lds_c, st>>>(nx, lx, ny, nyh, nz, i0, twx, spec, dk,
			                                             d->lbound[1], dz, d->depth, G, cols);
			break;
		}
		ARMA_HIP_CHECK(hipGetLastError());
		}
		{
		KernelTimer t_ir(ctx, "pot_inv_rows");
		inv_rows_kernel<T><<<dim3(div_ceil(nx, rows), cn * nz), 256, lds_r, st>>>(
			nx, ny, ly, nyh, nz, i0, twy, G, phi, rows);
		ARMA_HIP_CHECK(hipGetLastError());
		}
		KernelTimer t_ed(ctx, "pot_edge");
		edge_kernel<T><<<cn * nz, 256, 0, st>>>(nx, ny, phi, i0, nz);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	if (wnmax_out) {
		std::vector<double> hk(4 * size_t(nto));
		ARMA_HIP_CHECK(hipMemcpyAsync(hk.data(), dk, sizeof(double) * 4 * nto,
		                              hipMemcpyDeviceToHost, st));
		ARMA_HIP_CHECK(hipStreamSynchronize(st));
		for (int i = 0; i < nto; ++i) {
			wnmax_out[2 * i] = 1.0 / hk[4 * i];
			wnmax_out[2 * i + 1] = 1.0 / hk[4 * i + 1];
		}
	}
}

}  // namespace

void launch_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* d, const void* zeta,
                             void* phi, double* wnmax_out) {
	if (d->precision == ARMA_FP32)
		run_potential<float>(ctx, d, (const float*)zeta, (float*)phi, wnmax_out);
	else
		run_potential<double>(ctx, d, (const double*)zeta, (double*)phi, wnmax_out);
}

}  // namespace arma_hip
