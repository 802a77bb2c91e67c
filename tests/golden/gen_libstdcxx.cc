// Writes known-answer vectors of the toolchain's <random> (g++ libstdc++):
// std::mt19937 raw draws and std::normal_distribution<double/float>
// variates.  The reference draws its white noise through the same
// libstdc++ templates (std::normal_distribution over a 32-bit engine,
// src/generator/ar_model_wn.cc:10-16), so these vectors pin the oracle's
// restatement of generate_canonical + the polar method bit for bit.
// Output: little-endian binary records on stdout, read by make_golden.py.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static void put_u32(const std::vector<uint32_t>& v) {
	uint32_t n = v.size();
	fwrite(&n, 4, 1, stdout);
	fwrite(v.data(), 4, v.size(), stdout);
}

template <class T>
static void put_real(const std::vector<T>& v) {
	uint32_t n = v.size();
	fwrite(&n, 4, 1, stdout);
	fwrite(v.data(), sizeof(T), v.size(), stdout);
}

int main() {
	const uint32_t seeds[] = {0u, 1u, 42u, 5489u};
	const int ndraw = 1024;
	const int nnorm = 512;
	for (uint32_t seed : seeds) {
		std::mt19937 mt(seed);
		std::vector<uint32_t> raw(ndraw);
		for (auto& x : raw) x = mt();
		put_u32(raw);
	}
	{
		std::mt19937 mt;  // default seed 5489
		uint32_t x = 0;
		for (int i = 0; i < 10000; ++i) x = mt();
		put_u32(std::vector<uint32_t>{x});
	}
	const double sigmas[] = {1.0, 0.37};
	for (uint32_t seed : seeds) {
		for (double s : sigmas) {
			std::mt19937 mt(seed);
			std::normal_distribution<double> nd(0.0, s);
			std::vector<double> v(nnorm);
			for (auto& x : v) x = nd(mt);
			put_real(v);
			std::mt19937 mt2(seed);
			std::normal_distribution<float> nf(0.0f, float(s));
			std::vector<float> w(nnorm);
			for (auto& x : w) x = nf(mt2);
			put_real(w);
		}
	}
	return 0;
}
