// Register-resident DFT building blocks for the batched 2-D transforms of
// the velocity potential.
//
// A length-N transform (N = N1 * N2, both powers of two, N2 <= 32) runs as
// a four-step transform: N1 DFTs of length N2 over the stride-N1
// subsequences, a twiddle by w_N^(n1 k2), then N2 DFTs of length N1.  Each
// short DFT is held entirely in registers of one lane (fully unrolled radix-2
// decimation in frequency with compile-time twiddles), so one transform costs
// one exchange through LDS instead of one LDS round trip per radix stage.
#pragma once

#include <hip/hip_runtime.h>

namespace arma_hip {

template <class R>
struct cpx {
	R x, y;
};

template <class R>
__host__ __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
	return cpx<R>{a.x + b.x, a.y + b.y};
}
template <class R>
__host__ __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
	return cpx<R>{a.x - b.x, a.y - b.y};
}
template <class R>
__host__ __device__ __forceinline__ cpx<R> conj(cpx<R> a) {
	return cpx<R>{a.x, -a.y};
}
// a * (c + i s)
template <class R>
__host__ __device__ __forceinline__ cpx<R> cmul(cpx<R> a, R c, R s) {
	return cpx<R>{R(fma(a.x, c, -(a.y * s))), R(fma(a.x, s, a.y * c))};
}

// cos(2 pi q / 32), q = 0..8, correctly rounded to double.
__host__ __device__ constexpr double cos32(int q) {
	return q == 0   ? 1.0
	       : q == 1 ? 0.9807852804032304
	       : q == 2 ? 0.9238795325112867
	       : q == 3 ? 0.8314696123025452
	       : q == 4 ? 0.7071067811865476
	       : q == 5 ? 0.5555702330196022
	       : q == 6 ? 0.3826834323650898
	       : q == 7 ? 0.19509032201612828
	                : 0.0;
}
// cos and sin of 2 pi q / 32 for q = 0..15.
__host__ __device__ constexpr double cq(int q) { return q <= 8 ? cos32(q) : -cos32(16 - q); }
__host__ __device__ constexpr double sq(int q) { return q <= 8 ? cos32(8 - q) : cos32(q - 8); }

// a * e^{SIGN 2 pi i Q / 32}, Q in [0, 16), with the trivial angles folded.
template <int Q, int SIGN, class R>
__host__ __device__ __forceinline__ cpx<R> rot32(cpx<R> a) {
	if constexpr (Q == 0) {
		return a;
	} else if constexpr (Q == 8) {
		// times SIGN * i
		return SIGN > 0 ? cpx<R>{-a.y, a.x} : cpx<R>{a.y, -a.x};
	} else if constexpr (Q == 4 || Q == 12) {
		constexpr R h = R(0.7071067811865476);
		// (1 + SIGN i)/sqrt2 for Q = 4, (-1 + SIGN i)/sqrt2 for Q = 12
		const R re = Q == 4 ? (SIGN > 0 ? a.x - a.y : a.x + a.y)
		                    : (SIGN > 0 ? -a.x - a.y : -a.x + a.y);
		const R im = Q == 4 ? (SIGN > 0 ? a.x + a.y : a.y - a.x)
		                    : (SIGN > 0 ? a.x - a.y : -a.x - a.y);
		return cpx<R>{re * h, im * h};
	} else {
		constexpr R c = R(cq(Q));
		constexpr R s = R(SIGN > 0 ? sq(Q) : -sq(Q));
		return cmul(a, c, s);
	}
}

__host__ __device__ constexpr int brev(int k, int n) {
	int r = 0;
	for (int b = 1; b < n; b <<= 1) {
		r = (r << 1) | (k & 1);
		k >>= 1;
	}
	return r;
}

// Butterflies J.. of block B.. of one radix-2 DIF stage of span LEN.
template <int N, int LEN, int SIGN, int B, int J, class R>
__host__ __device__ __forceinline__ void dif_bfly(cpx<R> (&v)[N]) {
	if constexpr (B < N) {
		if constexpr (J < LEN / 2) {
			const cpx<R> a = v[B + J], c = v[B + J + LEN / 2];
			v[B + J] = a + c;
			v[B + J + LEN / 2] = rot32<J * (32 / LEN), SIGN>(a - c);
			dif_bfly<N, LEN, SIGN, B, J + 1>(v);
		} else {
			dif_bfly<N, LEN, SIGN, B + LEN, 0>(v);
		}
	}
}

template <int N, int LEN, int SIGN, class R>
__host__ __device__ __forceinline__ void dif_stage(cpx<R> (&v)[N]) {
	dif_bfly<N, LEN, SIGN, 0, 0>(v);
	if constexpr (LEN > 2) dif_stage<N, LEN / 2, SIGN>(v);
}

// In-register DFT of length N (a power of two, at most 32):
// v[brev(k)] <- sum_n v[n] e^{SIGN 2 pi i n k / N}.
template <int N, int SIGN, class R>
__host__ __device__ __forceinline__ void dft(cpx<R> (&v)[N]) {
	if constexpr (N > 1) dif_stage<N, N, SIGN>(v);
}

// w_N^(SIGN q) from the table tw[q] = e^{-2 pi i q / N}; stride selects
// w_{N/stride} from the same table.
template <int SIGN, class R>
__device__ __forceinline__ cpx<R> twiddle(cpx<R> a, const double2* __restrict__ tw, int q) {
	const double2 w = tw[q];
	return cmul(a, R(w.x), R(SIGN < 0 ? w.y : -w.y));
}

// Hands the outputs of a lane's first DFT (v[brev(k2)] holds output k2)
// to store(k2, value) after the twiddle w^(SIGN n1 k2), k2 = 0..N2-1.  Only
// w^(n1) = tw[base] is read; its powers are formed by repeated
// multiplication in double precision (relative error below N2 ulps).
template <int N2, int SIGN, class R, class Store>
__device__ __forceinline__ void twiddle_out(const cpx<R> (&v)[N2], const double2* __restrict__ tw,
                                            int base, Store store) {
	const double2 b = tw[base];
	const double br = b.x, bi = SIGN < 0 ? b.y : -b.y;
	double wr = 1.0, wi = 0.0;
#pragma unroll
	for (int k2 = 0; k2 < N2; ++k2) {
		const cpx<R> a = v[brev(k2, N2)];
		store(k2, k2 == 0 ? a : cmul(a, R(wr), R(wi)));
		const double nr = fma(wr, br, -(wi * bi));
		wi = fma(wr, bi, wi * br);
		wr = nr;
	}
}

}  // namespace arma_hip
