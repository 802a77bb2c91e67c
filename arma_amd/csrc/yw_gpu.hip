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

// C = A B (SUB = false) or C = C - A B (SUB = true); C may not alias A or B.
template <bool SUB>
__device__ void gemm(DMat C, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c;
	for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
		const int i = e / n, j = e - (e / n) * n;
		const double* ar = A.p + size_t(i) * k;
		double s = 0.0;
		for (int q = 0; q < k; ++q) {
			const double aik = ar[q];
			if (aik == 0.0) continue;
			s += aik * B.p[size_t(q) * n + j];
		}
		C.p[e] = SUB ? C.p[e] - s : s;
	}
	__syncthreads();
}

// C = D - A B (C may alias D).
__device__ void gemm_from(DMat C, DMat D, DMat A, DMat B) {
	const int m = A.r, k = A.c, n = B.c;
	for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
		const int i = e / n, j = e - (e / n) * n;
		const double* ar = A.p + size_t(i) * k;
		double s = 0.0;
		for (int q = 0; q < k; ++q) {
			const double aik = ar[q];
			if (aik == 0.0) continue;
			s += aik * B.p[size_t(q) * n + j];
		}
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
	if (threadIdx.x == 0) bad = 0;
	for (int k = 0; k < n; ++k) {
		if (threadIdx.x == 0) {
			int p = k;
			for (int i = k + 1; i < n; ++i)
				if (fabs(lu[size_t(i) * n + k]) > fabs(lu[size_t(p) * n + k])) p = i;
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

__global__ __launch_bounds__(kYwThreads) void choi_kernel(YwArgs y) {
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

void yule_walker_choi_device(arma_hip_ctx* ctx, const double* acf, int n, int max_order,
                             double var_eps, bool determine, std::vector<double>& phi, int& l,
                             double& varwn) {
	const int M = std::min(max_order, n - 1);
	ARMA_HIP_REQUIRE(M >= 2, ARMA_HIP_EINVAL, "AR order below 2 for Choi solver");
	ARMA_HIP_REQUIRE(acf[0] != 0.0, ARMA_HIP_EINVAL, "zero variance in the ACF");
	const int MP = M + 2;
	// point tables up to M + 1
	std::vector<int> rows, cols, rows_off(MP), cols_off(MP);
	for (int a = 0; a < MP; ++a) {
		rows_off[a] = int(rows.size() / 3);
		push_points(rows, a, a, a);
		for (int e = 1; e <= a; ++e) {
			for (int q = 0; q < e; ++q) push_points(rows, a - e, a - q, a);
			for (int q = 0; q < e; ++q) push_points(rows, a - e + q, a - e, a);
			for (int q = 0; q < e; ++q) push_points(rows, a, a - e, a - q);
			for (int q = 0; q < e; ++q) push_points(rows, a, a - e + q, a - e);
			for (int q = 0; q < e; ++q) push_points(rows, a - q, a, a - e);
			for (int q = 0; q < e; ++q) push_points(rows, a - e, a, a - e + q);
		}
		cols_off[a] = int(cols.size() / 3);
		push_points(cols, 0, 0, 0);
		for (int e = 1; e <= a; ++e) {
			for (int q = 0; q < e; ++q) push_points(cols, e, q, 0);
			for (int q = 0; q < e; ++q) push_points(cols, e - q, e, 0);
			for (int q = 0; q < e; ++q) push_points(cols, 0, e, q);
			for (int q = 0; q < e; ++q) push_points(cols, 0, e - q, e);
			for (int q = 0; q < e; ++q) push_points(cols, q, 0, e);
			for (int q = 0; q < e; ++q) push_points(cols, e, 0, e - q);
		}
	}
	// block offsets (in doubles)
	std::vector<long long> phi_off(size_t(MP) * MP * MP, -1), pi_off(MP), theta_off(MP);
	long long nphi = 0, npi = 0, nth = 0;
	for (int p = 2; p < MP; ++p)
		for (int a = 1; a < p; ++a)
			for (int j = 0; p + j < MP; ++j) {
				phi_off[(size_t(p) * MP + a) * MP + j] = nphi;
				nphi += (long long)shell(a) * shell(p + j);
			}
	for (int m = 0; m < MP; ++m) {
		pi_off[m] = npi;
		npi += shell(m);
		theta_off[m] = nth;
		nth += (long long)shell(m) * shell(m);
	}
	const long long S = shell(M + 1);
	const size_t n3 = size_t(n) * n * n;
	std::vector<double> rho(n3);
	for (size_t q = 0; q < n3; ++q) rho[q] = acf[q] / acf[0];
	// one device buffer: doubles first, then 8-byte index tables, then ints
	const size_t nd = n3 + size_t(nphi) + 2 * size_t(npi) + size_t(nth) + size_t(S) * S * 5 +
	                  size_t(M + 1) * (M + 1) * (M + 1) + 2;
	const size_t nll = phi_off.size() + pi_off.size() + theta_off.size();
	const size_t ni = rows.size() + cols.size() + 2 * MP + size_t(S) + 1;
	unsigned char* buf = (unsigned char*)ctx->yw.get(nd * 8 + nll * 8 + ni * 4 + 64);
	double* d = (double*)buf;
	YwArgs y{};
	y.n = n;
	y.M = M;
	y.var_eps = var_eps;
	y.determine = determine ? 1 : 0;
	y.variance = acf[0];
	double* rho_d = d;
	d += n3;
	y.rho = rho_d;
	y.phi = d;
	d += nphi;
	y.pi[0] = d;
	d += npi;
	y.pi[1] = d;
	d += npi;
	y.theta = d;
	d += nth;
	y.rsup11 = d;
	d += S * S;
	y.t0 = d;
	d += S * S;
	y.t1 = d;
	d += S * S;
	y.t2 = d;
	d += S * S;
	y.lu = d;
	d += S * S;
	y.out = d;
	d += size_t(M + 1) * (M + 1) * (M + 1) + 2;
	long long* ll = (long long*)d;
	long long* phi_off_d = ll;
	ll += phi_off.size();
	long long* pi_off_d = ll;
	ll += pi_off.size();
	long long* theta_off_d = ll;
	ll += theta_off.size();
	int* ip = (int*)ll;
	int* rows_d = ip;
	ip += rows.size();
	int* cols_d = ip;
	ip += cols.size();
	int* rows_off_d = ip;
	ip += MP;
	int* cols_off_d = ip;
	ip += MP;
	y.piv = ip;
	ip += S;
	y.status = ip;
	y.phi_off = phi_off_d;
	y.pi_off = pi_off_d;
	y.theta_off = theta_off_d;
	y.rows = rows_d;
	y.cols = cols_d;
	y.rows_off = rows_off_d;
	y.cols_off = cols_off_d;
	upload(ctx, rho_d, rho.data(), n3 * 8);
	upload(ctx, phi_off_d, phi_off.data(), phi_off.size() * 8);
	upload(ctx, pi_off_d, pi_off.data(), pi_off.size() * 8);
	upload(ctx, theta_off_d, theta_off.data(), theta_off.size() * 8);
	upload(ctx, rows_d, rows.data(), rows.size() * 4);
	upload(ctx, cols_d, cols.data(), cols.size() * 4);
	upload(ctx, rows_off_d, rows_off.data(), MP * 4);
	upload(ctx, cols_off_d, cols_off.data(), MP * 4);
	const int unset = 1;
	upload(ctx, y.status, &unset, 4);
	{
		KernelTimer timer(ctx, "yule_walker_choi");
		choi_kernel<<<1, kYwThreads, 0, ctx->stream>>>(y);
		ARMA_HIP_CHECK(hipGetLastError());
	}
	int status = 0;
	ARMA_HIP_CHECK(hipMemcpyAsync(&status, y.status, 4, hipMemcpyDeviceToHost, ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	ARMA_HIP_REQUIRE(status == ARMA_HIP_OK, ARMA_HIP_ERUNTIME, "getrf error");
	// l, varwn, then the coefficients
	std::vector<double> tail(size_t(M + 1) * (M + 1) * (M + 1) + 2);
	ARMA_HIP_CHECK(hipMemcpyAsync(tail.data(), y.out, tail.size() * 8, hipMemcpyDeviceToHost,
	                              ctx->stream));
	ARMA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
	l = int(tail[0]);
	ARMA_HIP_REQUIRE(l >= 2 && l <= M, ARMA_HIP_ERUNTIME, "Choi solver returned no order");
	const size_t L3 = size_t(l) * l * l;
	varwn = tail[1];
	phi.assign(tail.begin() + 2, tail.begin() + 2 + L3);
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
