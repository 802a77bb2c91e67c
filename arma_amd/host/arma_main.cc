// arma-hip: the reference's command-line driver (src/arma_main.cc:90-124,
// src/common_main.hh:77-121) built with the HIP framework.  Usage:
//   arma-hip [-h] [-n] [-f float|double] INPUTFILE
// Reads the .arma input file, generates the wavy surface on the GPU,
// computes the velocity potentials and writes the selected outputs.
#include <unistd.h>

#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "arma_hip_driver.hh"

namespace {

void usage() {
	std::cout << "usage: arma-hip [-h] [-n] [-f float|double] INPUTFILE\n"
	             "  -h            show this help\n"
	             "  -n            parse the input and print the model, no GPU work\n"
	             "  -f TYPE       real type (default double)\n";
}

template <class T>
int run_arma(const std::string& input, bool dry) {
	auto ctx = std::make_shared<arma::Hip_context>(0);
	arma::ARMA_driver<T> driver(ctx);
	arma::register_all(driver);
	driver.open(input);
	if (dry) {
		driver.dry_run(std::cout);
		return 0;
	}
	driver.generate_wavy_surface();
	driver.compute_velocity_potentials();
	driver.write_all();
	return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
	std::string type = "double";
	bool dry = false;
	int opt;
	while ((opt = ::getopt(argc, argv, "hnf:")) != -1) {
		if (opt == 'h') {
			usage();
			return 0;
		}
		if (opt == 'n') dry = true;
		else if (opt == 'f') type = optarg;
		else {
			usage();
			return 1;
		}
	}
	if (optind >= argc) {
		usage();
		return 1;
	}
	const std::string input = argv[optind];
	try {
		return type == "float" ? run_arma<float>(input, dry) : run_arma<double>(input, dry);
	} catch (const std::exception& e) {
		std::cerr << "arma-hip: " << e.what() << std::endl;
		return 1;
	}
}
