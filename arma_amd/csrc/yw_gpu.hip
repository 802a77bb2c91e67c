// Yule-Walker coefficients by Choi's order recursion, and the white-noise
// variance, in FP64 on the GPU (src/yule_walker.cc:420-539,
// src/generator/ar_model.cc:55-61).
//
// The recursion is a chain of small dense products and inverses (matrices of
// at most shell(l) = 3l^2 + 3l + 1 rows: 169 for l = 7, 271 for l = 9) with a
// data-dependent stop, so it runs as one workgroup that walks the whole
// recursion with the matrices in device memory and every product, inverse
// and update spread over its lanes.  Every operation keeps the order of the
// host solver's arithmetic (products summed over k in order, LU with the
// first largest pivot, substitutions column by column, dot products in
// order) and no contraction into FMAs, so the GPU result is the host result
// to the last bit.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace arma_hip {
namespace {

constexpr int kYwThreads = 1024;

__host__ __device__ inline int shell(int a) { return 3 * a * a + 3 * a + 1; }

struct YwArgs {
	const double* rho;  // rho(i,j,k) = gamma / gamma000 on an n^3 cube
	int n;
	int M;              // largest order tried
	double var_eps;
	int determine;
	double variance;    // gamma000
	// point tables: row points of a and shell points of b (i, j, k triples)
	const int* rows;
	const int* rows_off;
	const int* cols;
	const int* cols_off;
	// storage
	double* phi;            // Phi blocks
	const long long* phi_off;  // by key ((p * (M + 2)) + a) * (M + 2) + j
	double* pi[2];          // Pi_l / Pi_{l+1}, blocks per m at pi_off[m]
	const long long* pi_off;
	double* theta;          // Theta[l] at theta_off[l]
	const long long* theta_off;
	double* rsup11;
	double* t0;             // scratch matrices of shell(M+1)^2
	double* t1;
	double* t2;
	double* lu;
	int* piv;
	double* out;            // l, varwn, then the l^3 coefficients
	int* status;
};

struct DMat {
	double* p;
	int r, c;
};

__device__ inline double rho_at(const YwArgs& a, int i, int j, int k) {
	i = i < 0 ? -i : i;
	j = j < 0 ? -j : j;
	k = k < 0 ? -k : k;
	return a.rho[(size_t(i) * a.n + j) * a.n + k];
}

// R_{a,b} into dst (shell(a) x shell(b)).
__device__ DMat build_R(const YwArgs& y, double* dst, int a, int b) {
	const int nr = shell(a), nc = shell(b);
	const int* rp = y.rows + 3 * y.rows_off[a];
	const int* cp = y.cols + 3 * y.cols_off[b];
	for (int e = threadIdx.x; e < nr * nc; e += blockDim.x) {
		const int r = e / nc, c = e - (e / nc) * nc;
		dst[e] = rho_at(y, rp[3 * r] - b + cp[3 * c], rp[3 * r + 1] - b + cp[3 * c + 1],
		                rp[3 * r + 2] - b + cp[3 * c + 2]);
	}
	__syncthreads();
	return DMat{dst, nr, nc};
}

__device__ DMat build_Rb0(const YwArgs& y, double* dst, int a) {
	const int nr = shell(a);
	const int* rp = y.rows + 3 * y.rows_off[a];
	for (int r = threadIdx.x; r < nr; r += blockDim.x)
		dst[r] = rho_at(y, rp[3 * r], rp[3 * r + 1], rp[3 * r + 2]);
	__syncthreads();
	return DMat{dst, nr, 1};
}

__device__ DMat build_Ra0(const YwArgs& y, double* dst, int b) {
	const int nc = shell(b);
	const int* cp = y.cols + 3 * y.cols_off[b];
	for (int c = threadIdx.x; c < nc; c += blockDim.x)
		dst[c] = rho_at(y, -b + cp[3 * c], -b + cp[3 * c + 1], -b + cp[3 * c + 2]);
	__syncthreads();
	return DMat{dst, nc, 1};
}

// Elements (i, j0 .. j0+3) of A B by one lane (cnt of them: fewer at the
// end of a row): element (i, j) is row i of A times column j of B, summed
// over k in order (the host's loop).  One load of A's factor serves the four
// sums, whose B terms are contiguous, and the four sums are independent, so
// the loads of later terms go out while earlier terms are summed: each sum
// is the host's to the bit.  d(j, s) receives the product of column j.
template <class D>
__device__ __forceinline__ void products_row4(const double* __restrict__ A,
                                              const double* __restrict__ B, int k, int n, int i,
                                              int j0, int cnt, D&& d) {
	const double* ar = A + size_t(i) * k;
	const double* bc = B + j0;
	double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
	for (int q = 0; q < k; ++q) {
		const double aik = ar[q];
		const double* bq = bc + size_t(q) * n;
#pragma unroll
		for (int r = 0; r < 4; ++r) s[r] = s[r] + aik * bq[r < cnt ? r : 0];
	}
#pragma unroll
	for (int r = 0; r < 4; ++r)
		if (r < cnt) d(j0 + r, s[r]);
}

// Products with more than a few columns go through LDS: 64 x 64 tiles of the
// result, k in chunks of 32, a 2 x 2 block of the tile per lane (1024
// lanes).  Each element still sums its terms over k in order, so the tiling
// changes no bit; it cuts the L2 reads of one workgroup by the tile's reuse.
constexpr int kGT = 64, kGK = 32;
constexpr int kLuLdsN = 127;  // LU factorisations up to this order run in LDS (shell(6))
constexpr int kPoolDoubles = kLuLdsN * kLuLdsN > kGT * (kGK + 1) + kGK * (kGT + 1)
                                 ? kLuLdsN * kLuLdsN
                                 : kGT * (kGK + 1) + kGK * (kGT + 1);

// The workgroup's one LDS buffer, shared by the product tiles, the LU
// factorisations that fit and the substitution's row blocks (a
// function-scope __shared__ array is a single allocation however many
// callers inline it).
__device__ __forceinline__ double* lds_pool() {
	__shared__ double pool[kPoolDoubles];
	return pool;
}

// the products of a tile, handed to d(i, j, s) for elements in range
template <class D>
__device__ void tiled_products(const double* __restrict__ A, const double* __restrict__ B,
                               int m, int k, int n, D&& d) {
	double(*sa)[kGK + 1] = reinterpret_cast<double(*)[kGK + 1]>(lds_pool());
	double(*sb)[kGT + 1] = reinterpret_cast<double(*)[kGT + 1]>(lds_pool() + kGT * (kGK + 1));
	const int tid = threadIdx.x;
	const int tx = tid % (kGT / 2), ty = tid / (kGT / 2);  // ty < 32 for 1024 lanes
	for (int ti = 0; ti < m; ti += kGT)
		for (int tj = 0; tj < n; tj += kGT) {
			double s[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
			for (int kc = 0; kc < k; kc += kGK) {
				const int kn = min(kGK, k - kc);
				for (int e = tid; e < kGT * kGK; e += blockDim.x) {
					const int r = e / kGK, q = e - (e / kGK) * kGK;
					sa[r][q] = ti + r < m && q < kn ? A[size_t(ti + r) * k + kc + q] : 0.0;
				}
				for (int e = tid; e < kGK * kGT; e += blockDim.x) {
					const int q = e / kGT, c = e - (e / kGT) * kGT;
					sb[q][c] = q < kn && tj + c < n ? B[size_t(kc + q) * n + tj + c] : 0.0;
				}
				__syncthreads();
				for (int q = 0; q < kn; ++q) {
					const double a0 = sa[2 * ty][q], a1 = sa[2 * ty + 1][q];
					const double b0 = sb[q][2 * tx], b1 = sb[q][2 * tx + 1];
					s[0][0] = s[0][0] + a0 * b0;
					s[0][1] = s[0][1] + a0 * b1;
					s[1][0] = s[1][0] + a1 * b0;
					s[1][1] = s[1][1] + a1 * b1;
				}
				__syncthreads();
			}
			for (int a = 0; a < 2; ++a)
				for (int b = 0; b < 2; ++b) {
					const int i = ti + 2 * ty + a, j = tj + 2 * tx + b;
					if (i < m && j < n) d(i, j, s[a][b]);
				}
		}
}

// C = A B (SUB = false) or C = C - A B (SUB = true); C may not alias A or B.
template <bool SUB>
__device__ void gemm(DMat C, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c, nb = (n + 3) / 4;
	if (n >= 8 && blockDim.x == 1024) {
		tiled_products(A.p, B.p, m, k, n, [&](int i, int j, double s) {
			double* c = C.p + size_t(i) * n + j;
			*c = SUB ? *c - s : s;
		});
	} else {
		for (int e = threadIdx.x; e < m * nb; e += blockDim.x) {
			const int i = e / nb, j0 = 4 * (e - (e / nb) * nb);
			products_row4(A.p, B.p, k, n, i, j0, min(4, n - j0), [&](int j, double s) {
				double* c = C.p + size_t(i) * n + j;
				*c = SUB ? *c - s : s;
			});
		}
	}
	__syncthreads();
}

// C = D - A B (C may alias D).
__device__ void gemm_from(DMat C, DMat D, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c, nb = (n + 3) / 4;
	if (n >= 8 && blockDim.x == 1024) {
		tiled_products(A.p, B.p, m, k, n, [&](int i, int j, double s) {
			C.p[size_t(i) * n + j] = D.p[size_t(i) * n + j] - s;
		});
	} else {
		for (int e = threadIdx.x; e < m * nb; e += blockDim.x) {
			const int i = e / nb, j0 = 4 * (e - (e / nb) * nb);
			products_row4(A.p, B.p, k, n, i, j0, min(4, n - j0), [&](int j, double s) {
				C.p[size_t(i) * n + j] = D.p[size_t(i) * n + j] - s;
			});
		}
	}
	__syncthreads();
}

// Sum of a[i] b[i] in index order (one lane; every lane gets the result).
__device__ double dot(const double* a, const double* b, int n) {
	__shared__ double res;
	if (threadIdx.x == 0) {
		double s = 0.0;
		for (int i = 0; i < n; ++i) s += a[i] * b[i];
		res = s;
	}
	__syncthreads();
	const double v = res;
	__syncthreads();
	return v;
}

// The two triangular solves of the inverse, one column of it per lane, kept
// in place in inv: column c starts as e_c through the row interchanges
// (tracked as the position of its one), then
//   forward  x_i = b_i - sum_{j<i} L_ij x_j      (j ascending)
//   backward x_i = (y_i - sum_{j>i} U_ij x_j) / U_ii  (j descending),
// each row's terms subtracted one at a time in that order, as the host
// solver does.  Rows go in blocks of kSubRows: the terms of a block's rows
// with j outside the block come first in either order, so the block's rows
// take them together (one load of x_j serves kSubRows sums in registers)
// before the in-block terms are taken row by row.  The block's rows of
// L / U are staged through LDS in chunks of kSubCols columns (aligned, so a
// block's own columns share one chunk); every lane reads the same element,
// a broadcast.
constexpr int kSubRows = 16, kSubCols = 256;
static_assert(kSubRows * kSubCols <= kPoolDoubles, "substitution chunk exceeds the LDS pool");
static_assert(kSubCols % kSubRows == 0, "row blocks must not straddle column chunks");

constexpr int kSubX = 4;  // x_j loads in flight

// v[r] -= L / U(row r, column j) * x_j for the block's rows, in registers
__device__ __forceinline__ void sub_column(double (&v)[kSubRows], const double* b, double x) {
#pragma unroll
	for (int r = 0; r < kSubRows; ++r) v[r] -= b[r] * x;
}

__device__ __forceinline__ void stage_rows(double* blk, const double* lu, int n, int I, int nb, int j0, int jn) {
	__syncthreads();  // the previous chunk's readers are done
	for (int e = threadIdx.x; e < kSubRows * kSubCols; e += blockDim.x) {
		const int q = e / kSubRows, r = e - q * kSubRows;  // blk[q][r]: column-major
		blk[e] = r < nb && q < jn ? lu[size_t(I + r) * n + j0 + q] : 0.0;
	}
	__syncthreads();
}

__device__ __forceinline__ void substitute(const YwArgs& y, DMat inv) {
	const int n = inv.r;
	const double* __restrict__ lu = y.lu;
	double* blk = lds_pool();
	for (int c0 = 0; c0 < n; c0 += blockDim.x) {
		const int c = c0 + threadIdx.x;
		const bool act = c < n;
		double* __restrict__ col = inv.p + (act ? c : 0);  // stride n
		int pos = c;
		if (act)
			for (int k = 0; k < n; ++k) {
				const int p = y.piv[k];
				pos = pos == k ? p : (pos == p ? k : pos);
			}
		// forward, blocks top down
		for (int I = 0; I < n; I += kSubRows) {
			const int nb = min(kSubRows, n - I);
			double v[kSubRows];
#pragma unroll
			for (int r = 0; r < kSubRows; ++r) v[r] = I + r == pos ? 1.0 : 0.0;
			for (int j0 = 0; j0 < I + nb; j0 += kSubCols) {
				const int jn = min(kSubCols, n - j0);
				stage_rows(blk, lu, n, I, nb, j0, jn);
				if (!act) continue;
				const int je = min(jn, I - j0);  // outer columns of this chunk
				int q = 0;
				for (; q + kSubX <= je; q += kSubX) {
					double x[kSubX];  // the loads go out together
#pragma unroll
					for (int u = 0; u < kSubX; ++u) x[u] = col[size_t(j0 + q + u) * n];
#pragma unroll
					for (int u = 0; u < kSubX; ++u) sub_column(v, blk + (q + u) * kSubRows, x[u]);
				}
				for (; q < je; ++q) sub_column(v, blk + q * kSubRows, col[size_t(j0 + q) * n]);
				if (I < j0 + jn) {  // the block's own columns: unit lower triangle
					const double* b = blk + (I - j0) * kSubRows;
#pragma unroll
					for (int r = 1; r < kSubRows; ++r)
#pragma unroll
						for (int t = 0; t < r; ++t) v[r] -= b[t * kSubRows + r] * v[t];
				}
			}
			if (act)
#pragma unroll
				for (int r = 0; r < kSubRows; ++r)
					if (r < nb) col[size_t(I + r) * n] = v[r];
		}
		// backward, blocks bottom up; chunks from the last column down
		for (int I = (n - 1) / kSubRows * kSubRows; I >= 0; I -= kSubRows) {
			const int nb = min(kSubRows, n - I);
			double v[kSubRows];
#pragma unroll
			for (int r = 0; r < kSubRows; ++r) v[r] = act && r < nb ? col[size_t(I + r) * n] : 0.0;
			for (int j0 = (n - 1) / kSubCols * kSubCols; j0 >= I / kSubCols * kSubCols;
			     j0 -= kSubCols) {
				const int jn = min(kSubCols, n - j0);
				stage_rows(blk, lu, n, I, nb, j0, jn);
				if (!act) continue;
				const int qlo = max(0, I + nb - j0);  // outer columns are j >= I + nb
				int q = jn - 1;
				for (; q - kSubX + 1 >= qlo; q -= kSubX) {
					double x[kSubX];
#pragma unroll
					for (int u = 0; u < kSubX; ++u) x[u] = col[size_t(j0 + q - u) * n];
#pragma unroll
					for (int u = 0; u < kSubX; ++u) sub_column(v, blk + (q - u) * kSubRows, x[u]);
				}
				for (; q >= qlo; --q) sub_column(v, blk + q * kSubRows, col[size_t(j0 + q) * n]);
				if (j0 <= I) {  // the block's own columns: upper triangle, then the pivot
					const double* b = blk + (I - j0) * kSubRows;
#pragma unroll
					for (int r = kSubRows - 1; r >= 0; --r) {
						if (r >= nb) continue;
#pragma unroll
						for (int t = kSubRows - 1; t > r; --t)
							if (t < nb) v[r] -= b[t * kSubRows + r] * v[t];
						v[r] = v[r] / b[r * kSubRows + r];
					}
				}
			}
			if (act)
#pragma unroll
				for (int r = 0; r < kSubRows; ++r)
					if (r < nb) col[size_t(I + r) * n] = v[r];
		}
	}
	__syncthreads();
}

// LU factorisation with partial pivoting of the n x n matrix at lu (in LDS
// or device memory; inlined per call site, so each copy knows which), the
// pivots to y.piv.  Returns false when a pivot is exactly zero.
__device__ __forceinline__ bool lu_factor(const YwArgs& y, double* lu, int n) {
	__shared__ int pk;
	__shared__ int bad;
	__shared__ double wbest[kYwThreads / 64];
	__shared__ int wbidx[kYwThreads / 64];
	if (threadIdx.x == 0) bad = 0;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int k = 0; k < n; ++k) {
		// pivot: the first row i >= k of largest |lu(i, k)| (the host's
		// sequential scan keeps a row only when a later one is strictly
		// larger), as a lane scan, a wave reduction and a fold of the waves,
		// each preferring the smaller row on ties
		double bv = -1.0;
		int bi = n;
		for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
			const double v = fabs(lu[size_t(i) * n + k]);
			if (v > bv) {
				bv = v;
				bi = i;
			}
		}
		for (int off = 32; off > 0; off >>= 1) {
			const double ov = __shfl_xor(bv, off);
			const int oi = __shfl_xor(bi, off);
			if (ov > bv || (ov == bv && oi < bi)) {
				bv = ov;
				bi = oi;
			}
		}
		if (lane == 0) {
			wbest[wave] = bv;
			wbidx[wave] = bi;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			int p = wbidx[0];
			double pv = wbest[0];
			for (int q = 1; q < int(blockDim.x) / 64; ++q)
				if (wbest[q] > pv || (wbest[q] == pv && wbidx[q] < p)) {
					pv = wbest[q];
					p = wbidx[q];
				}
			// the scan never leaves a NaN diagonal (no value compares larger)
			// and never picks a NaN row
			if (p >= n || isnan(lu[size_t(k) * n + k])) p = k;
			if (lu[size_t(p) * n + k] == 0.0) bad = 1;
			pk = p;
			y.piv[k] = p;
		}
		__syncthreads();
		if (bad) return false;
		const int p = pk;
		if (p != k)
			for (int j = threadIdx.x; j < n; j += blockDim.x) {
				const double t = lu[size_t(k) * n + j];
				lu[size_t(k) * n + j] = lu[size_t(p) * n + j];
				lu[size_t(p) * n + j] = t;
			}
		__syncthreads();
		const double d = lu[size_t(k) * n + k];
		for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) lu[size_t(i) * n + k] /= d;
		__syncthreads();
		// the rank-1 update, four consecutive elements of a row per lane:
		// every operand is loaded before any of the four is stored (an updated
		// element (i > k, j > k) is never an operand of another: those are
		// column k, row k and the element itself)
		const int w = n - k - 1, wb = (w + 3) / 4;
		for (int e = threadIdx.x; e < w * wb; e += blockDim.x) {
			const int i = k + 1 + e / wb, j0 = k + 1 + 4 * (e - (e / wb) * wb);
			const int cnt = min(4, n - j0);
			double* row = lu + size_t(i) * n;
			const double* rk = lu + size_t(k) * n;
			const double l = row[k];
			double a[4], u[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int j = r < cnt ? j0 + r : j0;
				a[r] = row[j];
				u[r] = rk[j];
			}
#pragma unroll
			for (int r = 0; r < 4; ++r)
				if (r < cnt) row[j0 + r] = a[r] - l * u[r];
		}
		__syncthreads();
	}
	return true;
}

// inv = src^{-1} by LU with partial pivoting and one substitution per column
// (getrf + getri in the order of the host solver).  Returns false when a
// pivot is exactly zero.  A matrix that fits the LDS pool is factorised
// there (each of its n column steps is four barriers apart, so LDS instead
// of L2 round trips is most of the gain) and copied to y.lu for the
// substitution; the arithmetic is the same either way.
__device__ __forceinline__ bool inverse(const YwArgs& y, DMat inv, DMat src) {
	const int n = src.r;
	if (n <= kLuLdsN) {
		double* lu = lds_pool();
		for (int e = threadIdx.x; e < n * n; e += blockDim.x) lu[e] = src.p[e];
		__syncthreads();
		if (!lu_factor(y, lu, n)) return false;
		for (int e = threadIdx.x; e < n * n; e += blockDim.x) y.lu[e] = lu[e];
		__syncthreads();
	} else {
		for (int e = threadIdx.x; e < n * n; e += blockDim.x) y.lu[e] = src.p[e];
		__syncthreads();
		if (!lu_factor(y, y.lu, n)) return false;
	}
	substitute(y, inv);
	return true;
}

__device__ DMat phi_block(const YwArgs& y, int p, int a, int j) {
	const long long off = y.phi_off[(size_t(p) * (y.M + 2) + a) * (y.M + 2) + j];
	return DMat{y.phi + off, shell(a), shell(p + j)};
}

__device__ DMat pi_block(const YwArgs& y, double* base, int m) {
	return DMat{base + y.pi_off[m], shell(m), 1};
}

__device__ void choi_body(const YwArgs& y) {
	const int M = y.M;
	double* Pl = y.pi[0];
	double* Plp1 = y.pi[1];
	DMat R11 = build_R(y, y.t0, 1, 1);
	DMat Rsup11{y.rsup11, R11.r, R11.c};
	if (!inverse(y, Rsup11, R11)) {
		if (threadIdx.x == 0) *y.status = ARMA_HIP_ERUNTIME;
		return;
	}
	DMat rb0 = build_Rb0(y, y.t1, 1);
	gemm<false>(pi_block(y, Pl, 1), Rsup11, rb0);
	DMat ra0 = build_Ra0(y, y.t1, 1);
	double lambda = 1.0 - dot(ra0.p, pi_block(y, Pl, 1).p, ra0.r);
	double var0 = y.variance;
	double var = y.variance * lambda;
	gemm<false>(phi_block(y, 2, 1, 0), Rsup11, build_R(y, y.t0, 1, 2));
	bool changed = false;
	int l = 1;
	do {
		++l;
		var0 = var;
		// theta = R(l,l) - sum R(l,m) Phi[l,m,0]; h = R_b0(l) - sum R(l,m) Pi_l[m]
		DMat theta = build_R(y, y.t1, l, l);
		DMat h = build_Rb0(y, y.t2, l);
		for (int m = 1; m <= l - 1; ++m) {
			DMat Rlm = build_R(y, y.t0, l, m);
			gemm<true>(theta, Rlm, phi_block(y, l, m, 0));
			gemm<true>(h, Rlm, pi_block(y, Pl, m));
		}
		DMat Th{y.theta + y.theta_off[l], theta.r, theta.c};
		if (!inverse(y, Th, theta)) {
			if (threadIdx.x == 0) *y.status = ARMA_HIP_ERUNTIME;
			return;
		}
		gemm<false>(pi_block(y, Plp1, l), Th, h);
		for (int a = 1; a <= l - 1; ++a)
			gemm_from(pi_block(y, Plp1, a), pi_block(y, Pl, a), phi_block(y, l, a, 0),
			          pi_block(y, Plp1, l));
		lambda -= dot(h.p, pi_block(y, Plp1, l).p, h.r);
		var = y.variance * lambda;
		if (l < M) {
			gemm<false>(phi_block(y, 2, 1, l - 1), Rsup11, build_R(y, y.t0, 1, l + 1));
			for (int m = 2; m <= l; ++m) {
				DMat rhs = build_R(y, y.t1, m, l + 1);
				for (int q = 1; q <= m - 1; ++q)
					gemm<true>(rhs, build_R(y, y.t0, m, q), phi_block(y, m, q, l - m + 1));
				DMat Thm{y.theta + y.theta_off[m], shell(m), shell(m)};
				gemm<false>(phi_block(y, m + 1, m, l - m), Thm, rhs);
				for (int a = 1; a <= m - 1; ++a)
					gemm_from(phi_block(y, m + 1, a, l - m), phi_block(y, m, a, l - m + 1),
					          phi_block(y, m, a, 0), phi_block(y, m + 1, m, l - m));
			}
		}
		double* t = Pl;
		Pl = Plp1;
		Plp1 = t;
		changed = !(y.determine && fabs(var - var0) < y.var_eps);
	} while (l < M && changed);
	double* beta = changed ? Pl : Plp1;
	const double varwn = changed ? var : var0;
	// phi(l x l x l) from the block vectors, in the order of the row points
	const size_t L3 = size_t(l) * l * l;
	double* coef = y.out + 2;
	for (size_t e = threadIdx.x; e < L3; e += blockDim.x) coef[e] = 0.0;
	__syncthreads();
	for (int d = 1 + threadIdx.x; d < l; d += blockDim.x) {
		const double* src = beta + y.pi_off[d];
		const int* rp = y.rows + 3 * y.rows_off[d];
		for (int r = 0; r < shell(d); ++r)
			coef[(size_t(rp[3 * r]) * l + rp[3 * r + 1]) * l + rp[3 * r + 2]] = src[r];
	}
	if (threadIdx.x == 0) {
		y.out[0] = double(l);
		y.out[1] = varwn;
		*y.status = ARMA_HIP_OK;
	}
}

// One workgroup per model: the recursions of a batch's models run side by
// side, each exactly as a single solve.
__global__ __launch_bounds__(kYwThreads) void choi_kernel(const YwArgs* __restrict__ ys) {
	const YwArgs y = ys[blockIdx.x];
	choi_body(y);
}

// sigma^2 = gamma000 - sum phi * gamma over phi's shape, summed in the host
// order (one lane).
__global__ void varwn_kernel(const double* acf, int a1, int a2, const double* phi, int p0, int p1,
                             int p2, double* out) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	double s = 0.0;
	for (int a = 0; a < p0; ++a)
		for (int b = 0; b < p1; ++b)
			for (int c = 0; c < p2; ++c)
				s += phi[(size_t(a) * p1 + b) * p2 + c] * acf[(size_t(a) * a1 + b) * a2 + c];
	*out = acf[0] - s;
}

void push_points(std::vector<int>& v, int i, int j, int k) {
	v.push_back(i);
	v.push_back(j);
	v.push_back(k);
}

}  // namespace

namespace {

// Host tables and device layout of one model's recursion.
struct YwPlan {
	int n = 0, M = 0, MP = 0;
	double var_eps = 0.0, variance = 0.0;
	bool determine = true;
	std::vector<int> rows, cols, rows_off, cols_off;
	std::vector<long long> phi_off, pi_off, theta_off;
	std::vector<double> rho;
	long long nphi = 0, npi = 0, nth = 0, S = 0;
	size_t nd = 0, nll = 0, ni = 0;  // doubles, 8-byte offsets, ints
	size_t bytes() const { return (nd * 8 + nll * 8 + ni * 4 + 255) / 256 * 256; }
	size_t out_len() const { return size_t(M + 1) * (M + 1) * (M + 1) + 2; }
};

YwPlan plan_model(const double* acf, int n, int max_order, double var_eps, bool determine) {
	YwPlan q;
	q.n = n;
	q.M = std::min(max_order, n - 1);
	ARMA_HIP_REQUIRE(q.M >= 2, ARMA_HIP_EINVAL, "AR order below 2 for Choi solver");
	ARMA_HIP_REQUIRE(acf[0] != 0.0, ARMA_HIP_EINVAL, "zero variance in the ACF");
	q.var_eps = var_eps;
	q.determine = determine;
	q.variance = acf[0];
	const int MP = q.MP = q.M + 2;
	// point tables up to M + 1
	q.rows_off.resize(MP);
	q.cols_off.resize(MP);
	for (int a = 0; a < MP; ++a) {
		q.rows_off[a] = int(q.rows.size() / 3);
		push_points(q.rows, a, a, a);
		for (int e = 1; e <= a; ++e) {
			for (int r = 0; r < e; ++r) push_points(q.rows, a - e, a - r, a);
			for (int r = 0; r < e; ++r) push_points(q.rows, a - e + r, a - e, a);
			for (int r = 0; r < e; ++r) push_points(q.rows, a, a - e, a - r);
			for (int r = 0; r < e; ++r) push_points(q.rows, a, a - e + r, a - e);
			for (int r = 0; r < e; ++r) push_points(q.rows, a - r, a, a - e);
			for (int r = 0; r < e; ++r) push_points(q.rows, a - e, a, a - e + r);
		}
		q.cols_off[a] = int(q.cols.size() / 3);
		push_points(q.cols, 0, 0, 0);
		for (int e = 1; e <= a; ++e) {
			for (int r = 0; r < e; ++r) push_points(q.cols, e, r, 0);
			for (int r = 0; r < e; ++r) push_points(q.cols, e - r, e, 0);
			for (int r = 0; r < e; ++r) push_points(q.cols, 0, e, r);
			for (int r = 0; r < e; ++r) push_points(q.cols, 0, e - r, e);
			for (int r = 0; r < e; ++r) push_points(q.cols, r, 0, e);
			for (int r = 0; r < e; ++r) push_points(q.cols, e, 0, e - r);
		}
	}
	// block offsets (in doubles)
	q.phi_off.assign(size_t(MP) * MP * MP, -1);
	q.pi_off.resize(MP);
	q.theta_off.resize(MP);
	for (int p = 2; p < MP; ++p)
		for (int a = 1; a < p; ++a)
			for (int j = 0; p + j < MP; ++j) {
				q.phi_off[(size_t(p) * MP + a) * MP + j] = q.nphi;
				q.nphi += (long long)shell(a) * shell(p + j);
			}
	for (int m = 0; m < MP; ++m) {
		q.pi_off[m] = q.npi;
		q.npi += shell(m);
		q.theta_off[m] = q.nth;
		q.nth += (long long)shell(m) * shell(m);
	}
	q.S = shell(q.M + 1);
	const size_t n3 = size_t(n) * n * n;
	q.rho.resize(n3);
	for (size_t e = 0; e < n3; ++e) q.rho[e] = acf[e] / acf[0];
	q.nd = n3 + size_t(q.nphi) + 2 * size_t(q.npi) + size_t(q.nth) + size_t(q.S) * q.S * 5 +
	       q.out_len();
	q.nll = q.phi_off.size() + q.pi_off.size() + q.theta_off.size();
	q.ni = q.rows.size() + q.cols.size() + 2 * MP + size_t(q.S) + 1;
	return q;
}

// Lays the model out at device address dev (host image at img, same
// offsets): doubles first, then the 8-byte index tables, then ints.  The
// status word starts at 1 (not finished).
YwArgs place_model(const YwPlan& q, unsigned char* dev, unsigned char* img) {
	YwArgs y{};
	y.n = q.n;
	y.M = q.M;
	y.var_eps = q.var_eps;
	y.determine = q.determine ? 1 : 0;
	y.variance = q.variance;
	size_t at = 0;  // bytes from dev
	auto take = [&](size_t bytes) {
		const size_t r = at;
		at += bytes;
		return r;
	};
	const size_t n3 = q.rho.size();
	const size_t o_rho = take(n3 * 8);
	std::memcpy(img + o_rho, q.rho.data(), n3 * 8);
	y.rho = (const double*)(dev + o_rho);
	y.phi = (double*)(dev + take(size_t(q.nphi) * 8));
	y.pi[0] = (double*)(dev + take(size_t(q.npi) * 8));
	y.pi[1] = (double*)(dev + take(size_t(q.npi) * 8));
	y.theta = (double*)(dev + take(size_t(q.nth) * 8));
	const size_t SS = size_t(q.S) * q.S * 8;
	y.rsup11 = (double*)(dev + take(SS));
	y.t0 = (double*)(dev + take(SS));
	y.t1 = (double*)(dev + take(SS));
	y.t2 = (double*)(dev + take(SS));
	y.lu = (double*)(dev + take(SS));
	y.out = (double*)(dev + take(q.out_len() * 8));
	auto put_ll = [&](const std::vector<long long>& v) {
		const size_t o = take(v.size() * 8);
		std::memcpy(img + o, v.data(), v.size() * 8);
		return (const long long*)(dev + o);
	};
	y.phi_off = put_ll(q.phi_off);
	y.pi_off = put_ll(q.pi_off);
	y.theta_off = put_ll(q.theta_off);
	auto put_i = [&](const std::vector<int>& v) {
		const size_t o = take(v.size() * 4);
		std::memcpy(img + o, v.data(), v.size() * 4);
		return (const int*)(dev + o);
	};
	y.rows = put_i(q.rows);
	y.cols = put_i(q.cols);
	y.rows_off = put_i(q.rows_off);
	y.cols_off = put_i(q.cols_off);
	y.piv = (int*)(dev + take(size_t(q.S) * 4));
	const size_t o_st = take(4);
	const int unset = 1;
	std::memcpy(img + o_st, &unset, 4);
	y.status = (int*)(dev + o_st);
	return y;
}

}  // namespace

void yule_walker_choi_device_many(arma_hip_ctx* ctx, const std::vector<YwProblem>& probs,
                                  std::vector<YwSolution>& sols) {
	const size_t nm = probs.size();
	sols.assign(nm, YwSolution{});
	if (nm == 0) return;
	std::vector<YwPlan> plans;
	plans.reserve(nm);
	std::vector<size_t> base(nm);
	size_t total = 0;
	for (size_t i = 0; i < nm; ++i) {
		const YwProblem& pr = probs[i];
		plans.push_back(plan_model(pr.acf, pr.n, pr.max_order, pr.var_eps, pr.determine));
		base[i] = total;
		total += plans.back().bytes();
	}
	const size_t args_at = total;
	total += sizeof(YwArgs) * nm;
	unsigned char* dev = (unsigned char*)ctx->yw.get(total);
	std::vector<unsigned char> img(total, 0);
	std::vector<YwArgs> args(nm);
	for (size_t i = 0; i < nm; ++i) args[i] = place_model(plans[i], dev + base[i], img.data() + base[i]);
	std::memcpy(img.data() + args_at, args.data(), sizeof(YwArgs) * nm);
	upload(ctx, dev, img.data(), total);
	{
		KernelTimer timer(ctx, "yule_walker_choi");
		choi_kernel<<<int(nm), kYwThreads, 0, ctx->stream>>>((const YwArgs*)(dev + args_at));
		ARMA_HIP_CHECK(hipGetLastError());
	}
	// every model's status and results come back with one copy of the
	// whole buffer's result words
	std::vector<int> status(nm);
	std::vector<std::vector<double>> tails(nm);
	for (size_t i = 0; i < nm; ++i) {
		tails[i].resize(plans[i].out_len());
		ARMA_HIP_CHECK(hipMemcpyAsync(&status[i], args[i].status, 4, hipMemcpyDeviceToHost,
		                              ctx->stream));
		ARMA_HIP_CHECK(hipMemcpyAsync(tails[i].data(), args[i].out, tails[i].size() * 8,
		                              hipMemcpyDeviceToHost, ctx->stream));
	}
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	for (size_t i = 0; i < nm; ++i) {
		YwSolution& r = sols[i];
		if (status[i] != ARMA_HIP_OK) {
			r.status = ARMA_HIP_ERUNTIME;  // getrf: a zero pivot
			continue;
		}
		r.l = int(tails[i][0]);
		if (!(r.l >= 2 && r.l <= plans[i].M)) {
			r.status = ARMA_HIP_ERUNTIME;
			continue;
		}
		r.status = ARMA_HIP_OK;
		r.varwn = tails[i][1];
		const size_t L3 = size_t(r.l) * r.l * r.l;
		r.phi.assign(tails[i].begin() + 2, tails[i].begin() + 2 + L3);
	}
}

void yule_walker_choi_device(arma_hip_ctx* ctx, const double* acf, int n, int max_order,
                             double var_eps, bool determine, std::vector<double>& phi, int& l,
                             double& varwn) {
	std::vector<YwSolution> sol;
	yule_walker_choi_device_many(ctx, {YwProblem{acf, n, max_order, var_eps, determine}}, sol);
	ARMA_HIP_REQUIRE(sol[0].status == ARMA_HIP_OK, ARMA_HIP_ERUNTIME, "getrf error");
	l = sol[0].l;
	varwn = sol[0].varwn;
	phi = std::move(sol[0].phi);
}

double white_noise_variance_device(arma_hip_ctx* ctx, const double* acf, const int32_t acf_shape[3],
                                   const double* phi, const int32_t phi_shape[3]) {
	const size_t na = numel3(acf_shape), np = numel3(phi_shape);
	double* d = (double*)ctx->yw.get((na + np + 1) * 8);
	upload(ctx, d, acf, na * 8);
	upload(ctx, d + na, phi, np * 8);
	varwn_kernel<<<1, 64, 0, ctx->stream>>>(d, acf_shape[1], acf_shape[2], d + na, phi_shape[0],
	                                        phi_shape[1], phi_shape[2], d + na + np);
	ARMA_HIP_CHECK(hipGetLastError());
	double v = 0.0;
	ARMA_HIP_CHECK(hipMemcpyAsync(&v, d + na + np, 8, hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	return v;
}

}  // namespace arma_hip
