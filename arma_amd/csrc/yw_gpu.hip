// Yule-Walker coefficients by Choi's order recursion, and the white-noise
// variance, in FP64 on the GPU (src/yule_walker.cc:420-539,
// src/generator/ar_model.cc:55-61).
//
// The recursion is a chain of small dense products and inverses (matrices of
// at most shell(l) = 3l^2 + 3l + 1 rows: 169 for l = 7, 271 for l = 9) with a
// data-dependent stop, so it runs as one workgroup that walks the whole
// recursion with the matrices in device memory and every product, inverse
// and update spread over its lanes.  Every operation keeps the order of the
// host solver's arithmetic (products summed over k in order with zero
// factors skipped, LU with the first largest pivot, substitutions column by
// column, dot products in order) and no contraction into FMAs, so the GPU
// result is the host result to the last bit.
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

// Row i of A times column j of B, summed over k in order with zero factors
// of A skipped (the host's loop).  The skip is a select rather than a
// branch, so the loads of later terms go out while earlier terms are summed:
// the sum is the same to the bit.
__device__ __forceinline__ double row_col(const double* __restrict__ ar,
                                          const double* __restrict__ bc, int k, int n) {
	double s = 0.0;
#pragma unroll 8
	for (int q = 0; q < k; ++q) {
		const double aik = ar[q];
		const double t = s + aik * bc[size_t(q) * n];
		s = aik == 0.0 ? s : t;
	}
	return s;
}

// C = A B (SUB = false) or C = C - A B (SUB = true); C may not alias A or B.
template <bool SUB>
__device__ void gemm(DMat C, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c;
	for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
		const int i = e / n, j = e - (e / n) * n;
		const double s = row_col(A.p + size_t(i) * k, B.p + j, k, n);
		C.p[e] = SUB ? C.p[e] - s : s;
	}
	__syncthreads();
}

// C = D - A B (C may alias D).
__device__ void gemm_from(DMat C, DMat D, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c;
	for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
		const int i = e / n, j = e - (e / n) * n;
		const double s = row_col(A.p + size_t(i) * k, B.p + j, k, n);
		C.p[e] = D.p[e] - s;
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

// inv = src^{-1} by LU with partial pivoting and one substitution per column
// (getrf + getri in the order of the host solver).  Returns false when a
// pivot is exactly zero.
__device__ bool inverse(const YwArgs& y, DMat inv, DMat src) {
	const int n = src.r;
	double* lu = y.lu;
	for (int e = threadIdx.x; e < n * n; e += blockDim.x) lu[e] = src.p[e];
	__syncthreads();
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
		const int w = n - k - 1;
		for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
			const int i = k + 1 + e / w, j = k + 1 + (e - (e / w) * w);
			lu[size_t(i) * n + j] -= lu[size_t(i) * n + k] * lu[size_t(k) * n + j];
		}
		__syncthreads();
	}
	// column c of the inverse by lane c, kept in place in inv
	for (int c = threadIdx.x; c < n; c += blockDim.x) {
		double* col = inv.p + c;  // stride n
		for (int i = 0; i < n; ++i) col[size_t(i) * n] = i == c ? 1.0 : 0.0;
		for (int k = 0; k < n; ++k) {
			const int p = y.piv[k];
			const double t = col[size_t(k) * n];
			col[size_t(k) * n] = col[size_t(p) * n];
			col[size_t(p) * n] = t;
		}
		for (int i = 0; i < n; ++i) {
			double v = col[size_t(i) * n];
			for (int j = 0; j < i; ++j) v -= lu[size_t(i) * n + j] * col[size_t(j) * n];
			col[size_t(i) * n] = v;
		}
		for (int i = n - 1; i >= 0; --i) {
			double v = col[size_t(i) * n];
			for (int j = i + 1; j < n; ++j) v -= lu[size_t(i) * n + j] * col[size_t(j) * n];
			col[size_t(i) * n] = v / lu[size_t(i) * n + i];
		}
	}
	__syncthreads();
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
