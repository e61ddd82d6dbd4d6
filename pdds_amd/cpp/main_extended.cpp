// main_extended.cpp -- the plain PDDSparse driver of the reference (main_EXTENDED.cpp:1-22)
// on the MI355X engine: Monegros Poisson problem (BVPs/EDPs/Monegros_Poisson_CPP.hpp), Dirichlet
// data g = u, configuration from configuration.txt, outputs in the reference's CSV shapes.
//
// usage: pdds_main_extended [configuration.txt] [output directory]
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>

#include "pdds/pddsparse.hpp"

static double monegros_u(double x, double y) {
    const double k1 = 1.0, k2 = 0.01, k3 = 0.02, k4 = 0.12, k5 = 0.05, k6 = 0.05, k7 = -0.12, sf = 0.333;
    return 3.0 + sf * (std::sin(std::sqrt(k1 + k2 * x * x + k3 * y * y)) +
                       std::tanh(std::sin(k4 * x + k5 * y) + std::sin(k6 * x + k7 * y)));
}

int main(int argc, char *argv[]) {
    const std::string config = argc > 1 ? argv[1] : "configuration.txt";
    const std::string out = argc > 2 ? argv[2] : "Output";
    try {
        std::filesystem::create_directories(out + "/Debug");
        pdds::Device dev(0);
        pdds::bvp boundvalprob;
        boundvalprob.pde = PDDS_PDE_MONEGROS;
        boundvalprob.u = monegros_u;
        pdds::PDDSparseGM PDDS(argc, argv, config, dev);
        PDDS.out_dir = out;
        PDDS.Solve(boundvalprob);
        std::printf("knots %d  nnz %zu  path-steps %lld  kernel %.1f ms  BiCGSTAB %d its (res %.2e)\n",
                    PDDS.nNodes(), PDDS.cols.size(), (long long)PDDS.counters.path_steps, PDDS.counters.kernel_ms,
                    PDDS.bicgstab_iters, PDDS.bicgstab_relres);
    } catch (const pdds::Error &e) {
        std::fprintf(stderr, "pdds error: %s\n", e.what());
        return 1;
    }
    return 0;
}
