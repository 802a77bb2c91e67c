// Natural logarithm of a positive normal double, evaluated in double-double
// arithmetic and rounded once, for the polar normal transform.
//
// Tang's table-driven reduction (P. T. P. Tang, "Table-driven
// implementation of the logarithm function in IEEE floating-point
// arithmetic", ACM TOMS 16(4), 1990): x = 2^e m with m in [sqrt(1/2),
// sqrt(2)), c = k/128 the table point nearest m, m = c (1 + f), and
//   ln x = e ln 2 + ln c + ln(1 + f).
// m - c is exact (c/2 <= m <= 2c); f = (m - c)/c is carried as a head
// (m - c) * (1/c) and a tail from the exact FMA residual; ln(1 + f) =
// f - f^2/2 + f^3 (1/3 - f/4 + ... - f^7/10) with f^2 as an exact product
// and |f| < 2^-7.5, so the series is truncated below 2^-86.  The three parts
// are summed with error-free two-sums (e * ln2 head is exact: the head has
// 42 significant bits).  The sum before the final rounding is within about
// 2^-72 of ln x relative to |ln x|, so the result is the correctly rounded
// ln x unless ln x lies that close to a midpoint between two doubles.
//
// Host and device evaluate the same operations in the same order (explicit
// FMAs, no other contraction), so they agree to the bit.
#pragma once

#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>

#include "log_dd_table.hpp"

namespace arma_hip {

// s + err = a + b exactly, for |a| >= |b| or a == 0 (Dekker's Fast2Sum)
__host__ __device__ inline void fast_two_sum(double a, double b, double& s, double& err) {
	s = a + b;
	err = b - (s - a);
}

// tab: kLogDdTable (in any address space the caller chooses)
__host__ __device__ inline double log_dd(double x, const double* tab) {
	uint64_t ix;
	memcpy(&ix, &x, sizeof ix);
	int e = int(ix >> 52) - 1022;  // x = m 2^e, m in [1/2, 1)
	const uint64_t im = (ix & 0x000fffffffffffffull) | 0x3fe0000000000000ull;
	double m;
	memcpy(&m, &im, sizeof m);
	if (m < 0x1.6a09e667f3bcdp-1) {
		m = m * 2.0;
		e -= 1;
	}
	const double kd = __builtin_rint(m * 128.0);
	const double* row = tab + 3 * (int(kd) - kLogDdFirst);
	const double c = kd * 0x1p-7;
	const double d = m - c;
	const double fh = d * row[2];
	const double res = __builtin_fma(-fh, c, d);  // d - fh c, exact
	const double fl = res * row[2];
	// f^2 = sh + sl, f^2 / 2 = hh + hl
	const double sh = fh * fh;
	const double sl = __builtin_fma(fh, fh, -sh) + 2.0 * fh * fl;
	const double hh = 0.5 * sh, hl = 0.5 * sl;
	double p = -1.0 / 10;
	p = p * fh + 1.0 / 9;
	p = p * fh - 1.0 / 8;
	p = p * fh + 1.0 / 7;
	p = p * fh - 1.0 / 6;
	p = p * fh + 1.0 / 5;
	p = p * fh - 1.0 / 4;
	p = p * fh + 1.0 / 3;
	const double t3 = sh * fh * p + sh * fl;  // f^3 (...) with the tail's f^2 fl term
	double ah, ae;
	fast_two_sum(fh, -hh, ah, ae);
	const double lp_lo = ae + fl - hl + t3;
	const double ed = double(e);
	double s1, e1, s2, e2;
	// |e ln2| >= 0.69 > |ln c| unless e = 0; then |ln c| >= ln(129/128) >
	// |f| ~ |ah| unless c = 1, where ln c = 0: each sum's first operand is
	// zero or the larger, so the fast two-sums are exact
	fast_two_sum(ed * kLn2Hi, row[0], s1, e1);
	fast_two_sum(s1, ah, s2, e2);
	const double lo = e1 + e2 + ed * kLn2Lo + row[1] + lp_lo;
	return s2 + lo;
}

}  // namespace arma_hip
