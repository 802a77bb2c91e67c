// C++ drop-in for the reference's generator/solver registry and driver API
// with the HIP framework: AR_model<T>, Linear_solver<T> and
// High_amplitude_solver<T> keep their registry keys ("AR", "linear",
// "high_amplitude") and .arma input syntax, and their compute bodies call
// libarma_hip through the C ABI (include/arma_hip.h).
//
// Reference interfaces mirrored here:
//   ARMA_driver<T> ............ src/arma_driver.hh:43-150, src/arma_driver.cc
//   register_all_models/solvers  src/register_all.hh:47-68
//   parameter_map ............. src/params.hh:80-170, src/params.cc:30-57
//   Object_wrapper ............ src/bits/object_wrapper.hh:15-53
//   Grid / Domain input ....... src/grid.hh:181-193, src/domain.hh:195-217
//   AR_model<T> ............... src/generator/ar_model.hh:27-140
//   Basic_ARMA_model::generate  src/generator/basic_arma_model.cc:51-113
//   Linear_solver<T> .......... src/velocity/linear_solver.hh:18-70
//   High_amplitude_solver<T> .. src/velocity/high_amplitude_solver.hh:12-45
//   Velocity_potential_solver . src/velocity/basic_solver.hh:32-129
//   Output_flags .............. src/output_flags.cc:27-87
//   Binary_stream (zeta.bin) .. src/io/binary_stream.hh:19-71
#pragma once

#include <array>
#include <functional>
#include <iosfwd>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "arma_hip.h"

namespace arma {

typedef std::array<int, 3> Shape3D;
typedef std::array<int, 2> Shape2D;

// Row-major dense array with the Blitz layout the reference uses.
template <class T, int N>
struct Array {
	std::array<int, N> shape{};
	std::vector<T> data;
	Array() = default;
	explicit Array(const std::array<int, N>& s) : shape(s) {
		size_t n = 1;
		for (int d = 0; d < N; ++d) n *= size_t(s[d]);
		data.assign(n, T(0));
	}
	size_t size() const { return data.size(); }
};

template <class T>
using Array3D = Array<T, 3>;
template <class T>
using Array4D = Array<T, 4>;

// Number of points and length per dimension (src/grid.hh:42-210).
template <int N>
struct Grid {
	std::array<int, N> npoints{};
	std::array<double, N> length{};
	int num_patches(int i) const { return npoints[i] - 1 > 1 ? npoints[i] - 1 : 1; }
	double delta(int i) const { return length[i] / num_patches(i); }
};

// Bounds and number of points per dimension (src/domain.hh:28-235).
template <int N>
struct Domain {
	std::array<double, N> lbound{}, ubound{};
	std::array<int, N> npoints{};
	double delta(int i) const {
		const int np = npoints[i] - 1 > 1 ? npoints[i] - 1 : 1;
		return (ubound[i] - lbound[i]) / np;
	}
	double operator()(int idx, int dim) const { return lbound[dim] + delta(dim) * idx; }
};

// Output flags (src/output_flags.hh).
struct Output_flags {
	enum Flag { Surface = 1, ACF = 2, CSV = 4, Blitz = 8, Binary = 16, Summary = 32 };
	unsigned bits = 0;
	bool isset(Flag f) const { return (bits & f) != 0; }
};

// key = value configuration reader with '#' comments and {} blocks.
class parameter_map {
public:
	typedef std::function<void(std::istream&, const std::string&)> reader;
	parameter_map(std::map<std::string, reader> r, bool parens) : r_(std::move(r)), parens_(parens) {}
	void read(std::istream& in);

private:
	std::map<std::string, reader> r_;
	bool parens_;
};

enum class Precision { FP64, FP32 };

// Device and stream shared by the models and solvers of one driver; the
// device is opened on first use, so parsing and model setup need no GPU.
class Hip_context {
public:
	explicit Hip_context(int device = 0) : device_(device) {}
	~Hip_context();
	Hip_context(const Hip_context&) = delete;
	Hip_context& operator=(const Hip_context&) = delete;
	arma_hip_ctx* get();

private:
	int device_;
	arma_hip_ctx* ctx_ = nullptr;
};

// Throws the reference's exception types for the C ABI's error codes.
void check(int rc);

namespace generator {

template <class T>
class Basic_model {
public:
	virtual ~Basic_model() = default;
	virtual void read(std::istream& in) = 0;
	virtual Array3D<T> generate() = 0;
	// model setup without the GPU part (ACF, coefficients), for dry runs
	virtual void prepare() = 0;
	virtual void describe(std::ostream& out) const = 0;
	const Grid<3>& grid() const { return outgrid_; }
	Output_flags oflags() const { return oflags_; }
	// true when the model writes the binary surface itself while the run goes
	// on (src/generator/basic_model.hh:101-103, src/generator/ar_model.hh:60-62)
	virtual bool writes_in_parallel() const { return false; }
	// waits until the surface the model writes is in its file
	virtual void finish_writes() {}

protected:
	Grid<3> outgrid_;
	Output_flags oflags_;
	bool noseed_ = false;
};

enum class AR_algorithm { Gauss_elimination, Choi };

// AR_model<T> with the HIP do_generate body.
template <class T>
class AR_model : public Basic_model<T> {
public:
	explicit AR_model(std::shared_ptr<Hip_context> ctx) : ctx_(std::move(ctx)) {}
	~AR_model() override;
	void read(std::istream& in) override;
	Array3D<T> generate() override;
	void prepare() override;
	void prepare(bool on_device);
	void describe(std::ostream& out) const override;
	const std::vector<double>& coefficients() const { return phi_; }
	Shape3D order() const { return order_; }
	double white_noise_variance() const { return varwn_; }
	// device-resident surface of the last generate() (Array3D on the host)
	void determine_coefficients(bool on_device);
	Array3D<T> do_generate();
	bool writes_in_parallel() const override;
	void finish_writes() override;

private:
	void read_nit(std::istream& in);
	// transform = nit (src/generator/basic_arma_model.cc:22-25)
	bool linear_ = true;
	arma_nit_desc nit_{};
	std::shared_ptr<Hip_context> ctx_;
	// zeta.bin streaming: the writer and the device surface it reads
	arma_hip_writer* writer_ = nullptr;
	std::shared_ptr<void> surface_dev_;
	std::vector<double> acf_;
	Shape3D acf_shape_{0, 0, 0};
	Grid<3> acf_grid_;
	// an acf block without grid = : ACF_generator's parameters, the ACF made in prepare()
	arma_acf_gen_desc acf_gen_{};
	bool have_acf_gen_ = false;
	Shape3D order_{0, 0, 0};
	Shape3D partition_{0, 0, 0};
	AR_algorithm algorithm_ = AR_algorithm::Choi;
	bool validate_ = true;
	std::vector<double> phi_;
	Shape3D phi_shape_{0, 0, 0};
	double varwn_ = 0.0;
};

}  // namespace generator

namespace velocity {

template <class T>
class Velocity_potential_solver {
public:
	virtual ~Velocity_potential_solver() = default;
	virtual void read(std::istream& in);
	virtual Array4D<T> operator()(const Array3D<T>& zeta, const Grid<3>& grid) = 0;
	const Domain<2>& domain() const { return domain_; }

protected:
	Domain<2> wnmax_;
	double depth_ = 0.0;
	Domain<2> domain_;
};

template <class T>
class Linear_solver : public Velocity_potential_solver<T> {
public:
	explicit Linear_solver(std::shared_ptr<Hip_context> ctx) : ctx_(std::move(ctx)) {}
	Array4D<T> operator()(const Array3D<T>& zeta, const Grid<3>& grid) override;

protected:
	// the multiplier whose transform is the potential: -zeta_t here
	virtual bool high_amplitude() const { return false; }

private:
	std::shared_ptr<Hip_context> ctx_;
};

// The same transforms of zeta_t / (i((zeta_x + zeta_y)/s - zeta_x - zeta_y) - 1/s)
// (src/velocity/high_amplitude_solver.cc:6-23).
template <class T>
class High_amplitude_solver : public Linear_solver<T> {
public:
	using Linear_solver<T>::Linear_solver;

protected:
	bool high_amplitude() const override { return true; }
};

}  // namespace velocity

template <class T>
class ARMA_driver {
public:
	typedef velocity::Velocity_potential_solver<T> vpsolver_type;
	typedef generator::Basic_model<T> model_type;

	explicit ARMA_driver(std::shared_ptr<Hip_context> ctx) : ctx_(std::move(ctx)) {}

	template <class Type>
	void register_model(const std::string& key) {
		auto c = ctx_;
		models_[key] = [c]() { return std::unique_ptr<model_type>(new Type(c)); };
	}

	template <class Type>
	void register_solver(const std::string& key) {
		auto c = ctx_;
		solvers_[key] = [c]() { return std::unique_ptr<vpsolver_type>(new Type(c)); };
	}

	void open(const std::string& filename);
	void read(std::istream& in);
	// parse + model setup only; prints the model (no GPU needed)
	void dry_run(std::ostream& out);
	void generate_wavy_surface();
	void compute_velocity_potentials();
	void write_all();
	const Array3D<T>& wavy_surface() const { return zeta_; }
	const Array4D<T>& velocity_potentials() const { return vpotentials_; }

private:
	std::shared_ptr<Hip_context> ctx_;
	std::unordered_map<std::string, std::function<std::unique_ptr<model_type>()>> models_;
	std::unordered_map<std::string, std::function<std::unique_ptr<vpsolver_type>()>> solvers_;
	std::unique_ptr<model_type> model_;
	std::unique_ptr<vpsolver_type> solver_;
	Array3D<T> zeta_;
	Array4D<T> vpotentials_;
};

// register_all_models / register_all_solvers for the HIP build.
template <class T>
void register_all(ARMA_driver<T>& drv) {
	drv.template register_model<generator::AR_model<T>>("AR");
	drv.template register_solver<velocity::Linear_solver<T>>("linear");
	drv.template register_solver<velocity::High_amplitude_solver<T>>("high_amplitude");
}

}  // namespace arma
