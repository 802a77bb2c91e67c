// Dynamic creation of Mersenne Twister parameter sets (the arma-dcmt tool,
// src/dcmt.cc:15-84 via parallel_mt_seq, src/parallel_mt.hh:177-210).
//
// For a Mersenne exponent p and w = 32 the generator has n = p/w + 1 words,
// m = n/2 (or n-1 when n/2 < 2) and r = n*w - p lower bits masked out of the
// first word, as in Matsumoto & Nishimura's dynamic creator.  Candidates
// aaa carry the top bit and the 16-bit stream id in their low bits; a
// candidate is accepted when the characteristic polynomial of its state
// transition is primitive.  Because 2^p - 1 is prime, primitivity of a
// degree-p polynomial f is equivalent to f(0) = f(1) = 1 and
// x^(2^p) = x (mod f).  The polynomial is the minimal polynomial of one
// output bit, found by Berlekamp-Massey over 2p bits; small factors are
// screened out first by gcd(f, x^(2^k) - x) for k <= 12.
//
// Tempering: the shifts are the dynamic creator's (12, 18, 7, 15); the
// masks are fixed to 0x9d2c5680 / 0xefc60000 instead of being searched for
// equidistribution.  Tempering is a bijection on the output word and does
// not affect the period or the characteristic polynomial.
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "arma_hip.h"

namespace {

typedef std::vector<uint64_t> poly;  // GF(2)[x], bit i = coefficient of x^i

int degree(const poly& a) {
	for (int w = int(a.size()) - 1; w >= 0; --w)
		if (a[w]) return w * 64 + 63 - __builtin_clzll(a[w]);
	return -1;
}

inline bool bit(const poly& a, int i) { return (a[i >> 6] >> (i & 63)) & 1u; }
inline void flip(poly& a, int i) { a[i >> 6] ^= uint64_t(1) << (i & 63); }

// a ^= b << s
void xor_shifted(poly& a, const poly& b, int s) {
	const int ws = s >> 6, bs = s & 63;
	for (size_t i = 0; i < b.size(); ++i) {
		if (!b[i]) continue;
		const size_t j = i + ws;
		if (j < a.size()) a[j] ^= b[i] << bs;
		if (bs && j + 1 < a.size()) a[j + 1] ^= b[i] >> (64 - bs);
	}
}

// Reduction modulo f (degree L) with 64 pre-shifted copies of f.
struct Modulus {
	int L;
	size_t words;
	std::vector<poly> fs;  // f << s, s = 0..63, sized 2*words+1
	explicit Modulus(const poly& f) : L(degree(f)) {
		words = size_t(L / 64 + 1);
		fs.resize(64);
		for (int s = 0; s < 64; ++s) {
			fs[s].assign(2 * words + 2, 0);
			xor_shifted(fs[s], f, s);
		}
	}
	void reduce(poly& a) const {
		for (int i = int(a.size()) * 64 - 1; i >= L; --i) {
			if (!bit(a, i)) continue;
			const int sh = i - L;
			const poly& g = fs[sh & 63];
			const int ws = sh >> 6;
			for (size_t k = 0; k + ws < a.size() && k < g.size(); ++k) a[k + ws] ^= g[k];
		}
		a.resize(words);
	}
	poly square(const poly& a) const {
		static const std::vector<uint16_t> spread = [] {
			std::vector<uint16_t> t(256);
			for (int v = 0; v < 256; ++v) {
				uint16_t r = 0;
				for (int b = 0; b < 8; ++b) r |= uint16_t(((v >> b) & 1) << (2 * b));
				t[v] = r;
			}
			return t;
		}();
		poly r(2 * words + 2, 0);
		for (size_t i = 0; i < a.size(); ++i) {
			const uint64_t v = a[i];
			uint64_t lo = 0, hi = 0;
			for (int b = 0; b < 4; ++b) {
				lo |= uint64_t(spread[(v >> (8 * b)) & 0xff]) << (16 * b);
				hi |= uint64_t(spread[(v >> (8 * b + 32)) & 0xff]) << (16 * b);
			}
			r[2 * i] ^= lo;
			r[2 * i + 1] ^= hi;
		}
		reduce(r);
		return r;
	}
};

poly gcd(poly a, poly b) {
	int da = degree(a), db = degree(b);
	while (db >= 0) {
		while (da >= db) {
			xor_shifted(a, b, da - db);
			da = degree(a);
		}
		std::swap(a, b);
		std::swap(da, db);
	}
	return a;
}

struct Layout {
	int n, m, r;
	uint32_t umask, lmask;
};

Layout layout(int p) {
	Layout l;
	l.n = p / 32 + 1;
	l.m = l.n / 2;
	if (l.m < 2) l.m = l.n - 1;
	l.r = l.n * 32 - p;
	l.umask = uint32_t(0xffffffffu << l.r);
	l.lmask = ~l.umask;
	return l;
}

// Berlekamp-Massey connection polynomial of the bit sequence s (N bits).
// R holds the reversed window R[i] = s[n-i], so each discrepancy is the
// parity of C & R over whole words.
poly berlekamp_massey(const std::vector<uint8_t>& s, int& Lout) {
	const int N = int(s.size());
	const size_t W = size_t(N / 64 + 2);
	poly C(W, 0), B(W, 0), R(W, 0);
	C[0] = B[0] = 1;
	int L = 0, mshift = 1;
	for (int n = 0; n < N; ++n) {
		int d = s[n];
		uint64_t acc = 0;
		const size_t lw = size_t(L / 64 + 1);
		for (size_t w = 0; w < lw && w < W; ++w) acc ^= C[w] & R[w];
		d ^= __builtin_parityll(acc);  // bit 0 of R is always clear
		if (!d) {
			++mshift;
		} else if (2 * L <= n) {
			poly T = C;
			xor_shifted(C, B, mshift);
			L = n + 1 - L;
			B = T;
			mshift = 1;
		} else {
			xor_shifted(C, B, mshift);
			++mshift;
		}
		// R <- R << 1 with R[1] = s[n]
		for (size_t w = W - 1; w > 0; --w) R[w] = (R[w] << 1) | (R[w - 1] >> 63);
		R[0] = (R[0] << 1) | (uint64_t(s[n]) << 1);
		R[0] &= ~uint64_t(1);
	}
	Lout = L;
	return C;
}

bool primitive_candidate(uint32_t aaa, const Layout& l, int p) {
	// untempered MT recurrence; the sequence is bit 31 of successive words
	std::vector<uint32_t> st(l.n);
	uint32_t seed = 4357u;
	for (int k = 0; k < l.n; ++k) {
		st[k] = seed;
		seed = 1812433253u * (seed ^ (seed >> 30)) + uint32_t(k + 1);
	}
	const int N = 2 * p;
	std::vector<uint8_t> s(N);
	int idx = l.n;
	for (int q = 0; q < N; ++q) {
		if (idx >= l.n) {
			int k = 0;
			auto tw = [&](uint32_t hi, uint32_t lo, uint32_t src) {
				const uint32_t x = (hi & l.umask) | (lo & l.lmask);
				return src ^ (x >> 1) ^ ((x & 1u) ? aaa : 0u);
			};
			for (; k < l.n - l.m; ++k) st[k] = tw(st[k], st[k + 1], st[k + l.m]);
			for (; k < l.n - 1; ++k) st[k] = tw(st[k], st[k + 1], st[k + l.m - l.n]);
			st[l.n - 1] = tw(st[l.n - 1], st[0], st[l.m - 1]);
			idx = 0;
		}
		s[q] = uint8_t(st[idx++] >> 31);
	}
	int L = 0;
	poly f = berlekamp_massey(s, L);
	if (L != p || !bit(f, 0) || !bit(f, L)) return false;
	// f(1) = parity of the number of terms must be 1
	int terms = 0;
	for (uint64_t w : f) terms += __builtin_popcountll(w);
	if (!(terms & 1)) return false;
	f.resize(size_t(L / 64 + 1));
	Modulus mod(f);
	poly xpow(mod.words, 0);
	xpow[0] = 2;  // x
	poly x1 = xpow;
	for (int k = 1; k <= p; ++k) {
		xpow = mod.square(xpow);
		if (k <= 12 && k < p) {
			poly g = xpow;
			for (size_t i = 0; i < x1.size(); ++i) g[i] ^= x1[i];
			poly ff = f;
			if (degree(gcd(ff, g)) > 0) return false;
		}
	}
	for (size_t i = 0; i < xpow.size(); ++i)
		if (xpow[i] != x1[i]) return false;
	return true;
}

bool mersenne_exponent(int p) {
	static const int known[] = {521, 607, 1279, 2203, 2281, 3217, 4253, 4423, 9689, 9941,
	                            11213, 19937, 21701, 23209, 44497};
	for (int k : known)
		if (k == p) return true;
	return false;
}

int create_one(int p, int id, uint32_t seed, arma_mt_params* out) {
	if (!mersenne_exponent(p) || id < 0 || id > 0xffff) return ARMA_HIP_EINVAL;
	const Layout l = layout(p);
	std::mt19937 org(seed ^ (uint32_t(id) * 0x9e3779b9u));
	for (int tries = 0; tries < 200000; ++tries) {
		uint32_t a = (org() & 0xffff0000u) | 0x80000000u | uint32_t(id);
		if (primitive_candidate(a, l, p)) {
			arma_mt_params q;
			std::memset(&q, 0, sizeof q);
			q.aaa = a;
			q.mm = l.m;
			q.nn = l.n;
			q.rr = l.r;
			q.ww = 32;
			q.wmask = 0xffffffffu;
			q.umask = l.umask;
			q.lmask = l.lmask;
			q.shift0 = 12;
			q.shift1 = 18;
			q.shiftB = 7;
			q.shiftC = 15;
			q.maskB = 0x9d2c5680u;
			q.maskC = 0xefc60000u;
			*out = q;
			return ARMA_HIP_OK;
		}
	}
	return ARMA_HIP_ERUNTIME;
}

}  // namespace

extern "C" int arma_hip_dcmt_create(int32_t p, int32_t id, uint32_t seed, arma_mt_params* out) {
	try {
		return create_one(p, id, seed, out);
	} catch (...) {
		return ARMA_HIP_ERUNTIME;
	}
}

extern "C" int arma_hip_dcmt_create_many(int32_t p, int32_t n, uint32_t seed, int32_t nthreads,
                                         arma_mt_params* out) {
	if (n < 0 || n > 65536) return ARMA_HIP_EINVAL;
	if (nthreads < 1) nthreads = 1;
	std::vector<int> rc(nthreads, ARMA_HIP_OK);
	std::vector<std::thread> pool;
	for (int t = 0; t < nthreads; ++t) {
		pool.emplace_back([&, t]() {
			for (int id = t; id < n; id += nthreads) {
				int r = arma_hip_dcmt_create(p, id, seed, &out[id]);
				if (r != ARMA_HIP_OK) rc[t] = r;
			}
		});
	}
	for (auto& th : pool) th.join();
	for (int r : rc)
		if (r != ARMA_HIP_OK) return r;
	return ARMA_HIP_OK;
}
