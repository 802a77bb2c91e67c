// Implementation of the HIP-framework drop-in driver (see arma_hip_driver.hh).
#include "arma_hip_driver.hh"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <random>
#include <sstream>

namespace arma {

// ---- errors: the reference's exception types (src/errors.hh:10-33) ----

struct PRNG_error : std::runtime_error {
	using std::runtime_error::runtime_error;
};

void check(int rc) {
	if (rc == ARMA_HIP_OK) return;
	const std::string msg = arma_hip_last_error();
	switch (rc) {
	case ARMA_HIP_EINVAL: throw std::invalid_argument(msg);
	case ARMA_HIP_EPRNG: throw PRNG_error(msg);
	default: throw std::runtime_error(msg);
	}
}

arma_hip_ctx* Hip_context::get() {
	if (!ctx_) check(arma_hip_init(device_, &ctx_));
	return ctx_;
}

Hip_context::~Hip_context() {
	if (ctx_) arma_hip_destroy(ctx_);
}

// ---- input syntax (src/params.cc:9-57 and the operator>> of each type) ----

namespace {

std::istream& skip_comments(std::istream& in) {
	char ch = 0;
	while (in >> std::ws >> ch && ch == '#') {
		in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
		ch = 0;
	}
	if (ch != 0) in.putback(ch);
	return in;
}

bool expect(std::istream& in, char c) {
	char ch = 0;
	if (in >> std::ws >> ch && ch == c) return true;
	if (ch) in.putback(ch);
	return false;
}

// Blitz TinyVector syntax "(a, b, c)".
template <class V, size_t N>
void read_vec(std::istream& in, std::array<V, N>& v, const std::string& name) {
	if (!expect(in, '(')) throw std::runtime_error("bad vector for \"" + name + "\"");
	for (size_t i = 0; i < N; ++i) {
		if (!(in >> v[i])) throw std::runtime_error("bad vector for \"" + name + "\"");
		if (i + 1 < N && !expect(in, ',')) throw std::runtime_error("bad vector for \"" + name + "\"");
	}
	if (!expect(in, ')')) throw std::runtime_error("bad vector for \"" + name + "\"");
}

// "(n0,n1,n2)" or "(n0,n1,n2) : (l0,l1,l2)"; the length defaults to n - 1.
void read_grid(std::istream& in, Grid<3>& g, const std::string& name) {
	read_vec(in, g.npoints, name);
	if (expect(in, ':')) {
		read_vec(in, g.length, name);
	} else {
		for (int d = 0; d < 3; ++d) g.length[d] = g.npoints[d] - 1;
	}
	for (int d = 0; d < 3; ++d)
		if (g.npoints[d] <= 0 || !(g.length[d] > 0) || !std::isfinite(g.length[d]))
			throw std::invalid_argument("bad grid \"" + name + "\"");
}

// "from (a,b) to (c,d) npoints (n,m)" in any order.
void read_domain(std::istream& in, Domain<2>& d, const std::string& name) {
	d.lbound = {0, 0};
	d.ubound = {0, 0};
	d.npoints = {1, 1};
	for (int k = 0; k < 3; ++k) {
		std::string word;
		if (!(in >> std::ws >> word)) break;
		if (word == "from") read_vec(in, d.lbound, name);
		else if (word == "to") read_vec(in, d.ubound, name);
		else if (word == "npoints") read_vec(in, d.npoints, name);
		else throw std::runtime_error("bad prefix: " + word);
	}
	if (d.npoints[0] <= 0 || d.npoints[1] <= 0) throw std::invalid_argument("bad domain \"" + name + "\"");
}

void read_flags(std::istream& in, Output_flags& f) {
	std::string line;
	std::getline(in, line);
	std::stringstream ss(line);
	std::string name;
	while (std::getline(ss, name, ',')) {
		name.erase(0, name.find_first_not_of(" \t"));
		name.erase(name.find_last_not_of(" \t\r") + 1);
		if (name == "none") f.bits = 0;
		else if (name == "surface") f.bits |= Output_flags::Surface;
		else if (name == "acf") f.bits |= Output_flags::ACF;
		else if (name == "csv") f.bits |= Output_flags::CSV;
		else if (name == "blitz") f.bits |= Output_flags::Blitz;
		else if (name == "binary") f.bits |= Output_flags::Binary;
		else if (name == "summary" || name == "qq" || name == "waves") f.bits |= Output_flags::Summary;
		else throw std::runtime_error("bad output flag: " + name);
	}
	// default format (src/output_flags.cc:39-51)
	if (f.isset(Output_flags::Surface) && !f.isset(Output_flags::Blitz) &&
	    !f.isset(Output_flags::CSV) && !f.isset(Output_flags::Binary))
		f.bits |= Output_flags::Blitz;
}

template <class T>
void read_value(std::istream& in, T& v, const std::string& name) {
	if (!(in >> v)) throw std::runtime_error("bad value for \"" + name + "\"");
}

}  // namespace

void parameter_map::read(std::istream& in) {
	if (parens_ && !expect(in, '{')) throw std::runtime_error("Expecting \"{\".");
	std::string name;
	while (skip_comments(in)) {
		if (parens_ && expect(in, '}')) break;
		if (!(in >> std::ws && std::getline(in, name, '='))) break;
		while (!name.empty() && name.back() <= ' ') name.pop_back();
		auto it = r_.find(name);
		if (it == r_.end()) throw std::runtime_error("Unknown parameter: \"" + name + "\".");
		it->second(in, name);
	}
}

// ---- AR model ----

namespace generator {

namespace {

// closed_interval input "[a,b]" (src/apmath/closed_interval.cc:12-28)
void read_interval(std::istream& in, double (&v)[2], const std::string& name) {
	if (!expect(in, '[')) throw std::runtime_error("Expecting \"[\" in " + name);
	in >> v[0];
	if (!expect(in, ',')) throw std::runtime_error("Expecting \",\" in " + name);
	in >> v[1];
	if (!expect(in, ']')) throw std::runtime_error("Expecting \"]\" in " + name);
	if (!in) throw std::runtime_error("bad interval in " + name);
}

// Bisection input (src/linalg.cc:238-256)
void read_bisection(std::istream& in, double (&interval)[2], double& eps, int32_t& niter) {
	parameter_map p(
		{
			{"interval", [&](std::istream& s, const std::string& n) { read_interval(s, interval, n); }},
			{"absolute_error", [&](std::istream& s, const std::string& n) { read_value(s, eps, n); }},
			{"max_iterations", [&](std::istream& s, const std::string& n) { read_value(s, niter, n); }},
		},
		true);
	p.read(in);
	if (!(interval[0] < interval[1])) throw std::runtime_error("bad parameter");
	if (!(eps > 0)) throw std::runtime_error("bad parameter");
	if (!(niter > 0)) throw std::runtime_error("bad parameter");
}

}  // namespace

template <class T>
void AR_model<T>::read_nit(std::istream& in) {
	// NIT_transform input (src/nonlinear/nit_transform.cc:206-263) with the
	// defaults of its constructor (src/nonlinear/nit_transform.hh:48-66)
	std::memset(&nit_, 0, sizeof nit_);
	nit_.distribution = ARMA_NIT_GRAM_CHARLIER;
	nit_.interval[0] = 0.0;
	nit_.interval[1] = 1.0;
	nit_.absolute_error = 1e-6;
	nit_.max_iterations = 100;
	double acf_interval[2] = {0.0, 1.0};
	double acf_eps = 1e-6;
	int32_t acf_niter = 100;
	long nodes = 100, max_int = 12, max_exp = 10;
	auto positive = [](long v) {
		if (!(v > 0)) throw std::runtime_error("bad parameter");
	};
	parameter_map p(
		{
			{"distribution",
			 [&](std::istream& s, const std::string& n) {
				 std::string name;
				 read_value(s, name, n);
				 if (name == "gram_charlier") {
					 nit_.distribution = ARMA_NIT_GRAM_CHARLIER;
					 parameter_map d(
						 {
							 {"skewness", [&](std::istream& q, const std::string& m) { read_value(q, nit_.skewness, m); }},
							 {"kurtosis", [&](std::istream& q, const std::string& m) { read_value(q, nit_.kurtosis, m); }},
						 },
						 true);
					 d.read(s);
				 } else if (name == "skew_normal") {
					 nit_.distribution = ARMA_NIT_SKEW_NORMAL;
					 parameter_map d(
						 {
							 {"mean", [&](std::istream& q, const std::string& m) { read_value(q, nit_.mean, m); }},
							 {"stdev", [&](std::istream& q, const std::string& m) { read_value(q, nit_.stdev, m); }},
							 {"alpha", [&](std::istream& q, const std::string& m) { read_value(q, nit_.alpha, m); }},
						 },
						 true);
					 d.read(s);
				 } else {
					 throw std::runtime_error("bad distribution");
				 }
			 }},
			{"max_interpolation_order", [&](std::istream& s, const std::string& n) { read_value(s, max_int, n); positive(max_int); }},
			{"interpolation_nodes", [&](std::istream& s, const std::string& n) { read_value(s, nodes, n); positive(nodes); }},
			{"max_expansion_order", [&](std::istream& s, const std::string& n) { read_value(s, max_exp, n); positive(max_exp); }},
			{"cdf_solver",
			 [&](std::istream& s, const std::string&) {
				 read_bisection(s, nit_.interval, nit_.absolute_error, nit_.max_iterations);
			 }},
			{"acf_solver",
			 [&](std::istream& s, const std::string&) {
				 read_bisection(s, acf_interval, acf_eps, acf_niter);
			 }},
		},
		true);
	p.read(in);
}

template <class T>
void AR_model<T>::read(std::istream& in) {
	std::string transform = "none";
	bool have_acf = false;
	parameter_map params(
		{
			{"algorithm",
			 [&](std::istream& s, const std::string& n) {
				 std::string a;
				 read_value(s, a, n);
				 if (a == "gauss_elimination") algorithm_ = AR_algorithm::Gauss_elimination;
				 else if (a == "choi_recursive") algorithm_ = AR_algorithm::Choi;
				 else throw std::runtime_error("bad algorithm");
			 }},
			{"partition", [&](std::istream& s, const std::string& n) { read_vec(s, partition_, n); }},
			{"out_grid", [&](std::istream& s, const std::string& n) { read_grid(s, this->outgrid_, n); }},
			{"no_seed", [&](std::istream& s, const std::string& n) { read_value(s, this->noseed_, n); }},
			{"transform",
			 [&](std::istream& s, const std::string& n) {
				 // src/bits/transform_wrapper.hh:25-39
				 read_value(s, transform, n);
				 if (transform == "nit") {
					 read_nit(s);
					 linear_ = false;
				 } else if (transform == "none") {
					 linear_ = true;
				 } else {
					 throw std::runtime_error("bad transform");
				 }
			 }},
			{"acf",
			 [&](std::istream& s, const std::string&) {
				 // with grid = ...: the closed-form ACF (src/bits/acf_wrapper.cc:41-70);
				 // without: ACF_generator (src/generator/acf_generator.cc:216-233),
				 // the form the reference's AR model parses at
				 // src/generator/basic_arma_model.cc:26,57
				 std::string func;
				 double amplitude = 1.0, velocity = 1.0;
				 std::array<double, 3> alpha{0.06, 0.06, 0.06};
				 std::array<double, 2> beta{0.8, 0.0};
				 std::array<double, 3> nwaves{1.5, 1.5, 1.5};
				 std::array<int, 3> gshape{16, 16, 16};
				 parameter_map p(
					 {
						 {"grid", [&](std::istream& q, const std::string& n) { read_grid(q, acf_grid_, n); }},
						 {"func", [&](std::istream& q, const std::string& n) { read_value(q, func, n); }},
						 {"amplitude", [&](std::istream& q, const std::string& n) { read_value(q, amplitude, n); }},
						 {"velocity", [&](std::istream& q, const std::string& n) { read_value(q, velocity, n); }},
						 {"alpha", [&](std::istream& q, const std::string& n) { read_vec(q, alpha, n); }},
						 {"beta", [&](std::istream& q, const std::string& n) { read_vec(q, beta, n); }},
						 {"nwaves", [&](std::istream& q, const std::string& n) { read_vec(q, nwaves, n); }},
						 {"shape", [&](std::istream& q, const std::string& n) { read_vec(q, gshape, n); }},
					 },
					 true);
				 p.read(s);
				 if (acf_grid_.npoints[0] <= 0) {
					 arma_acf_gen_desc g;
					 std::memset(&g, 0, sizeof g);
					 // src/generator/plain_wave_profile.cc:7-24
					 if (func.empty() || func == "cos") g.func = ARMA_WAVE_COS;
					 else if (func == "sin") g.func = ARMA_WAVE_SIN;
					 else if (func == "stokes") g.func = ARMA_WAVE_STOKES;
					 else if (func == "standing_wave") g.func = ARMA_WAVE_STANDING;
					 else throw std::runtime_error("bad function");
					 g.amplitude = amplitude;
					 g.velocity = velocity;
					 for (int k = 0; k < 3; ++k) {
						 g.alpha[k] = alpha[k];
						 g.nwaves[k] = nwaves[k];
						 g.shape[k] = gshape[k];
					 }
					 g.beta[0] = beta[0];
					 g.beta[1] = beta[1];
					 // generated in prepare(): its auto-covariance runs on the GPU
					 // when a generation follows, on the host for a dry run
					 acf_gen_ = g;
					 have_acf_gen_ = true;
					 acf_.clear();
					 have_acf = true;
					 return;
				 }
				 int f;
				 if (func == "standing_wave") f = ARMA_ACF_STANDING_WAVE;
				 else if (func == "propagating_wave") f = ARMA_ACF_PROPAGATING_WAVE;
				 else throw std::runtime_error("bad ACF function name");
				 if (acf_grid_.npoints[0] <= 0)
					 throw std::runtime_error("the ACF block needs grid = (n0,n1,n2) : (l0,l1,l2)");
				 const double delta[3] = {acf_grid_.delta(0), acf_grid_.delta(1), acf_grid_.delta(2)};
				 acf_shape_ = acf_grid_.npoints;
				 acf_.assign(size_t(acf_shape_[0]) * acf_shape_[1] * acf_shape_[2], 0.0);
				 check(arma_hip_acf(f, delta, acf_shape_.data(), amplitude, velocity, alpha.data(),
				                    beta.data(), acf_.data()));
				 have_acf = true;
			 }},
			{"output", [&](std::istream& s, const std::string&) { read_flags(s, this->oflags_); }},
			{"order", [&](std::istream& s, const std::string& n) { read_vec(s, order_, n); }},
			{"validate", [&](std::istream& s, const std::string& n) { read_value(s, validate_, n); }},
		},
		true);
	params.read(in);
	if (!have_acf) throw std::runtime_error("the AR model needs an acf block");
}

template <class T>
void AR_model<T>::determine_coefficients(bool on_device) {
	// src/generator/ar_model.cc:81-154; FP64 on the GPU when a generation
	// follows (bit-identical to the host solver, which the dry run uses)
	if (algorithm_ == AR_algorithm::Choi) {
		const int n = acf_shape_[0];
		if (acf_shape_[1] != n || acf_shape_[2] != n) throw std::invalid_argument("ACF is not square");
		int max_order = n - 1;
		if (order_[0] > 0 && order_[1] > 0 && order_[2] > 0)
			max_order = *std::max_element(order_.begin(), order_.end());
		phi_.assign(size_t(std::max(max_order, 2)) * std::max(max_order, 2) * std::max(max_order, 2) +
		                size_t(n) * n * n,
		            0.0);
		double var = 0.0;
		if (on_device)
			check(arma_hip_yule_walker_choi_device(ctx_->get(), acf_.data(), n, max_order, 1e-6, 1,
			                                       1, 1e-10, phi_.data(), phi_shape_.data(),
			                                       &var));
		else
			check(arma_hip_yule_walker_choi(acf_.data(), n, max_order, 1e-6, 1, 1, 1e-10,
			                                phi_.data(), phi_shape_.data(), &var));
		phi_.resize(size_t(phi_shape_[0]) * phi_shape_[1] * phi_shape_[2]);
		order_ = phi_shape_;
		std::clog << "New AR model order = (" << order_[0] << ',' << order_[1] << ','
		          << order_[2] << ")\n";
	} else {
		phi_shape_ = order_;
		phi_.assign(size_t(order_[0]) * order_[1] * order_[2], 0.0);
		check(arma_hip_yule_walker_gauss(acf_.data(), acf_shape_.data(), order_.data(), phi_.data()));
	}
	if (on_device)
		check(arma_hip_white_noise_variance_device(ctx_->get(), acf_.data(), acf_shape_.data(),
		                                           phi_.data(), phi_shape_.data(), &varwn_));
	else
		varwn_ = arma_hip_white_noise_variance(acf_.data(), acf_shape_.data(), phi_.data(),
		                                       phi_shape_.data());
}

namespace {

std::vector<arma_mt_params> read_mts(size_t n) {
	// MT_CONFIG_FILE (src/config.hh.in:4)
	std::vector<arma_mt_params> mts(n);
	int32_t got = 0;
	check(arma_hip_read_mt_config("mt.dat", int32_t(n), mts.data(), &got));
	return mts;
}

std::vector<uint32_t> make_seeds(size_t n, bool noseed) {
	// src/parallel_mt.cc:37-48
	std::vector<uint32_t> seeds(n, 0u);
	if (!noseed) {
		std::random_device dev;
		std::seed_seq seq{uint32_t(std::chrono::high_resolution_clock::now().time_since_epoch().count()),
		                  uint32_t(dev())};
		seq.generate(seeds.begin(), seeds.end());
	}
	return seeds;
}

// Device buffer owned by the host code.
struct Device_buffer {
	void* p = nullptr;
	explicit Device_buffer(size_t bytes) {
		if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
	}
	~Device_buffer() {
		if (p) (void)hipFree(p);
	}
};

}  // namespace

template <class T>
Array3D<T> AR_model<T>::do_generate() {
	// the HIP body of AR_model<T>::do_generate (src/generator/ar_model.cc:189-197)
	if (varwn_ < 0) throw std::invalid_argument("variance is less than zero");
	const Shape3D shape = this->outgrid_.npoints;
	const bool partitioned = partition_[0] > 0 && partition_[1] > 0 && partition_[2] > 0;
	size_t nstreams = 1;
	if (partitioned)
		for (int d = 0; d < 3; ++d) nstreams *= size_t((shape[d] + partition_[d] - 1) / partition_[d]);
	const std::vector<arma_mt_params> mts = read_mts(nstreams);
	const std::vector<uint32_t> seeds = make_seeds(nstreams, this->noseed_);
	arma_ar_desc d;
	std::memset(&d, 0, sizeof d);
	for (int k = 0; k < 3; ++k) {
		d.shape[k] = shape[k];
		d.phi_shape[k] = phi_shape_[k];
		d.partition[k] = partition_[k];
	}
	d.phi = phi_.data();
	d.varwn = varwn_;
	d.noise_mode = partitioned ? ARMA_NOISE_PARTITIONED : ARMA_NOISE_SEQUENTIAL;
	d.noise_semantics = ARMA_NOISE_DOUBLE;
	d.mts = mts.data();
	d.n_mts = int32_t(mts.size());
	d.seeds = seeds.data();
	d.precision = sizeof(T) == 4 ? ARMA_FP32 : ARMA_FP64;
	const char* kern = std::getenv("ARMA_AR_KERNEL");
	d.algorithm = (kern && std::string(kern) == "plain") ? ARMA_AR_PLAIN : ARMA_AR_WAVEFRONT;
	Array3D<T> zeta(shape);
	const size_t bytes = zeta.size() * sizeof(T);
	std::shared_ptr<void> dev(new Device_buffer(bytes), [](void* b) {
		delete static_cast<Device_buffer*>(b);
	});
	void* dz = static_cast<Device_buffer*>(dev.get())->p;
	check(arma_hip_ar_generate(ctx_->get(), &d, dz, 1));
	// post-rescale of Basic_ARMA_model::generate (src/generator/basic_arma_model.cc:102-106)
	double factor = 0.0;
	check(arma_hip_rescale(ctx_->get(), dz, shape.data(), d.precision, acf_[0], &factor));
	// src/generator/basic_arma_model.cc:107-111
	if (!linear_)
		check(arma_hip_nit_realisation(ctx_->get(), dz, shape.data(), d.precision, acf_[0], &nit_));
	if (writes_in_parallel()) {
		// src/generator/ar_model_parallel.cc:112-145: zeta.bin is written while
		// the run goes on.  The writer starts once the surface is final (after
		// the rescale) and streams it while the GPU computes the velocity
		// potential; the device copy lives until the writer is done.
		finish_writes();
		check(arma_hip_writer_open(ctx_->get(), "zeta.bin", int32_t(sizeof(T)), 0, &writer_));
		check(arma_hip_writer_append(writer_, dz, bytes));
		surface_dev_ = dev;
	}
	check(arma_hip_synchronize(ctx_->get()));
	if (hipMemcpy(zeta.data.data(), dz, bytes, hipMemcpyDeviceToHost) != hipSuccess)
		throw std::runtime_error("hipMemcpy failed");
	return zeta;
}

template <class T>
AR_model<T>::~AR_model() {
	if (writer_) (void)arma_hip_writer_close(writer_, nullptr, nullptr);
}

template <class T>
bool AR_model<T>::writes_in_parallel() const {
	return this->oflags_.isset(Output_flags::Binary);
}

template <class T>
void AR_model<T>::finish_writes() {
	if (!writer_) return;
	arma_hip_writer* w = writer_;
	writer_ = nullptr;
	uint64_t nbytes = 0;
	double secs = 0.0;
	const int rc = arma_hip_writer_close(w, &nbytes, &secs);
	surface_dev_.reset();
	check(rc);
	std::clog << "Wrote zeta.bin: " << nbytes << " bytes, writer busy " << secs << " s\n";
}

template <class T>
void AR_model<T>::prepare() {
	prepare(false);
}

template <class T>
void AR_model<T>::prepare(bool on_device) {
	// src/generator/basic_arma_model.cc:51-97 with the closed-form ACF, or
	// ACF_generator's (src/generator/basic_arma_model.cc:57)
	if (have_acf_gen_ && acf_.empty()) {
		std::vector<double> buf(size_t(1) << 21);
		int32_t so[3];
		double de[3];
		if (on_device)
			check(arma_hip_acf_generate_device(ctx_->get(), &acf_gen_, buf.data(), buf.size(), so,
			                                   de));
		else
			check(arma_hip_acf_generate(&acf_gen_, buf.data(), buf.size(), so, de));
		for (int k = 0; k < 3; ++k) {
			acf_shape_[k] = so[k];
			acf_grid_.npoints[k] = so[k];
			acf_grid_.length[k] = acf_gen_.nwaves[k];
		}
		acf_.assign(buf.begin(), buf.begin() + size_t(so[0]) * so[1] * so[2]);
	}
	if (!(order_[0] > 0 && order_[1] > 0 && order_[2] > 0)) order_ = acf_shape_;
	for (int d = 0; d < 3; ++d)
		this->outgrid_.length[d] = acf_grid_.delta(d) * this->outgrid_.num_patches(d) * 2.0 * M_PI;
	determine_coefficients(on_device);
	std::clog << "White noise variance = " << varwn_ << '\n';
	// src/generator/basic_arma_model.cc:93-97 -> src/arma.cc:15-70
	if (validate_) {
		int32_t nbad = 0;
		check(arma_hip_validate_process(phi_.data(), phi_shape_.data(), &nbad, nullptr, 0));
	}
}

template <class T>
void AR_model<T>::describe(std::ostream& out) const {
	out.precision(17);
	out << "grid=(" << this->outgrid_.npoints[0] << ',' << this->outgrid_.npoints[1] << ','
	    << this->outgrid_.npoints[2] << "):(" << this->outgrid_.length[0] << ','
	    << this->outgrid_.length[1] << ',' << this->outgrid_.length[2] << ")\n";
	out << "order=(" << order_[0] << ',' << order_[1] << ',' << order_[2] << ")\n";
	out << "varwn=" << varwn_ << "\n";
	out << "phi=";
	for (size_t i = 0; i < phi_.size(); ++i) out << (i ? "," : "") << phi_[i];
	out << "\n";
}

template <class T>
Array3D<T> AR_model<T>::generate() {
	prepare(true);
	return do_generate();
}

}  // namespace generator

// ---- velocity potential ----

namespace velocity {

template <class T>
void Velocity_potential_solver<T>::read(std::istream& in) {
	parameter_map params(
		{
			{"wnmax", [&](std::istream& s, const std::string& n) { read_domain(s, wnmax_, n); }},
			{"depth", [&](std::istream& s, const std::string& n) { read_value(s, depth_, n); }},
			{"domain", [&](std::istream& s, const std::string& n) { read_domain(s, domain_, n); }},
		},
		true);
	params.read(in);
	if (domain_.lbound[1] < -depth_) throw std::invalid_argument("min(z) is less than -h");
}

template <class T>
Array4D<T> Linear_solver<T>::operator()(const Array3D<T>& zeta, const Grid<3>& grid) {
	arma_pot_desc d;
	std::memset(&d, 0, sizeof d);
	for (int k = 0; k < 3; ++k) {
		d.shape[k] = zeta.shape[k];
		d.delta[k] = grid.delta(k);
	}
	d.depth = this->depth_;
	for (int k = 0; k < 2; ++k) {
		d.lbound[k] = this->domain_.lbound[k];
		d.ubound[k] = this->domain_.ubound[k];
		d.npoints[k] = this->domain_.npoints[k];
	}
	d.precision = sizeof(T) == 4 ? ARMA_FP32 : ARMA_FP64;
	Array4D<T> out({this->domain_.npoints[0], this->domain_.npoints[1], zeta.shape[1], zeta.shape[2]});
	auto entry = high_amplitude() ? arma_hip_high_amplitude_potential : arma_hip_linear_potential;
	check(entry(ctx_->get(), &d, zeta.data.data(), out.data.data(), 0, nullptr));
	return out;
}

}  // namespace velocity

// ---- driver ----

template <class T>
void ARMA_driver<T>::open(const std::string& filename) {
	std::ifstream cfg(filename);
	if (!cfg.is_open()) throw std::runtime_error("Cannot open input file \"" + filename + "\".");
	std::clog << "Input file = " << filename << '\n';
	read(cfg);
}

template <class T>
void ARMA_driver<T>::read(std::istream& in) {
	// src/arma_driver.cc:133-178 with Object_wrapper (src/bits/object_wrapper.hh:31-46)
	parameter_map params(
		{
			{"model",
			 [&](std::istream& s, const std::string&) {
				 std::string name;
				 s >> std::ws >> name;
				 auto it = models_.find(name);
				 if (it == models_.end()) throw std::runtime_error("Invalid object: " + name);
				 model_ = it->second();
				 model_->read(s);
			 }},
			{"velocity_potential_solver",
			 [&](std::istream& s, const std::string&) {
				 std::string name;
				 s >> std::ws >> name;
				 auto it = solvers_.find(name);
				 if (it == solvers_.end()) throw std::runtime_error("Invalid object: " + name);
				 solver_ = it->second();
				 solver_->read(s);
			 }},
		},
		false);
	params.read(in);
	if (!solver_) throw std::runtime_error("bad solver");
	if (!model_) throw std::runtime_error("bad generator");
}

template <class T>
void ARMA_driver<T>::dry_run(std::ostream& out) {
	model_->prepare();
	model_->describe(out);
	const auto& d = solver_->domain();
	out << "domain=from (" << d.lbound[0] << ',' << d.lbound[1] << ") to (" << d.ubound[0] << ','
	    << d.ubound[1] << ") npoints (" << d.npoints[0] << ',' << d.npoints[1] << ")\n";
}

template <class T>
void ARMA_driver<T>::generate_wavy_surface() {
	zeta_ = model_->generate();
}

template <class T>
void ARMA_driver<T>::compute_velocity_potentials() {
	vpotentials_ = (*solver_)(zeta_, model_->grid());
}

namespace {

// network byte order (src/io/binary_stream.hh:43-70)
template <class T>
void write_big_endian(std::ostream& out, const std::vector<T>& v) {
	std::vector<unsigned char> buf(v.size() * sizeof(T));
	for (size_t i = 0; i < v.size(); ++i) {
		unsigned char b[sizeof(T)];
		std::memcpy(b, &v[i], sizeof(T));
		for (size_t k = 0; k < sizeof(T); ++k) buf[i * sizeof(T) + k] = b[sizeof(T) - 1 - k];
	}
	out.write((const char*)buf.data(), std::streamsize(buf.size()));
}

// Blitz++ text form of an array of rank >= 2, as `std::ofstream(name) << a`
// writes it (src/arma_driver.cc:33,54; Blitz++ operator<< for Array<T,N>):
// the index ranges "(0,n0-1) x (0,n1-1) x ...", a newline, "[ ", every
// element right-aligned in 9 columns followed by a space, a newline and two
// spaces after every 7 elements, then "]" and a newline.  Stream defaults
// otherwise (6 significant digits).
template <class T, int N>
void write_blitz(std::ostream& out, const Array<T, N>& a) {
	for (int d = 0; d < N; ++d) {
		out << '(' << 0 << ',' << a.shape[d] - 1 << ')';
		if (d != N - 1) out << " x ";
	}
	out << '\n' << "[ ";
	size_t p = 0;
	for (const T& v : a.data) {
		out << std::setw(9) << v << ' ';
		if (++p % 7 == 0) out << '\n' << "  ";
	}
	out << ']' << '\n';
}

}  // namespace

template <class T>
void ARMA_driver<T>::write_all() {
	// src/arma_driver.cc:29-69,121-130
	const Output_flags f = model_->oflags();
	model_->finish_writes();
	if (!f.isset(Output_flags::Surface)) return;
	const Grid<3>& g = model_->grid();
	const int nt = zeta_.shape[0], nx = zeta_.shape[1], ny = zeta_.shape[2];
	if (f.isset(Output_flags::Binary) && !model_->writes_in_parallel()) {
		std::ofstream out("zeta.bin", std::ios::binary);
		write_big_endian(out, zeta_.data);
	}
	// text outputs keep the stream's default format (6 significant digits),
	// as the reference's writers do (src/bits/write_csv.hh:13-80)
	if (f.isset(Output_flags::Blitz)) {
		std::ofstream out("zeta");
		write_blitz(out, zeta_);
	}
	if (f.isset(Output_flags::CSV)) {
		std::ofstream out("zeta.csv");
		out << "t,x,y,z\n";
		for (int i = 0; i < nt; ++i)
			for (int j = 0; j < nx; ++j)
				for (int k = 0; k < ny; ++k)
					out << i << ',' << g.delta(1) * j << ',' << g.delta(2) * k << ','
					    << zeta_.data[(size_t(i) * nx + j) * ny + k] << '\n';
	}
	// velocity potentials (src/arma_driver.cc:50-69; no binary form)
	if (f.isset(Output_flags::Blitz)) {
		std::ofstream out("phi");
		write_blitz(out, vpotentials_);
	}
	if (f.isset(Output_flags::CSV)) {
		const auto& dom = solver_->domain();
		const int ntp = vpotentials_.shape[0], nz = vpotentials_.shape[1];
		std::ofstream po("phi.csv");
		po << "t,z,x,y,phi\n";
		for (int i = 0; i < ntp; ++i)
			for (int j = 0; j < nz; ++j)
				for (int k = 0; k < nx; ++k)
					for (int l = 0; l < ny; ++l)
						po << dom(i, 0) << ',' << dom(j, 1) << ',' << g.delta(1) * k << ','
						   << g.delta(2) * l << ','
						   << vpotentials_.data[((size_t(i) * nz + j) * nx + k) * ny + l] << '\n';
	}
}

template class generator::AR_model<double>;
template class generator::AR_model<float>;
template class velocity::Velocity_potential_solver<double>;
template class velocity::Velocity_potential_solver<float>;
template class velocity::Linear_solver<double>;
template class velocity::Linear_solver<float>;
template class velocity::High_amplitude_solver<double>;
template class velocity::High_amplitude_solver<float>;
template class ARMA_driver<double>;
template class ARMA_driver<float>;

}  // namespace arma
