// emu_main -- the reference's executable (main.cpp:229-262) on this build:
//
//   emu_main <inputs_file> [key=value ...] [--path plain|fused-fp64|fused-fp32]
//            [--dump <file>] [--quiet] [--no-plotfiles] [--print-params]
//
// reads the deck like TestParams::Initialize, runs evolve_flavor and prints
// the figure of merit (particles advanced per microsecond).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "Emu.hpp"

int main(int argc, char* argv[])
{
    if (argc < 2) {
        std::fprintf(stderr,
                     "usage: %s <inputs_file> [key=value ...] [--path plain|fused-fp64|fused-fp32] [--dump file] "
                     "[--quiet] [--no-plotfiles] [--print-params]\n",
                     argv[0]);
        return 2;
    }
    std::vector<std::string> overrides;
    emu::EvolveOptions opt;
    bool print_params = false;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--path" && i + 1 < argc) opt.path = argv[++i];
        else if (a == "--dump" && i + 1 < argc) opt.dump_path = argv[++i];
        else if (a == "--quiet") opt.quiet = true;
        else if (a == "--no-plotfiles") opt.write_plotfiles = false;
        else if (a == "--print-params") print_params = true;
        else if (a.find('=') != std::string::npos) overrides.push_back(a);
        else {
            std::fprintf(stderr, "unknown argument '%s'\n", a.c_str());
            return 2;
        }
    }
    emu::TestParams parms;
    parms.Initialize(argv[1], overrides);
    if (print_params) {
        const emu_params p = parms.c_params();
        std::printf("{\"nf\": %d, \"shape_order\": %d, \"ncell\": [%d, %d, %d], \"phi\": [%.17g, %.17g, %.17g], "
                    "\"mass\": [%.17g, %.17g, %.17g], \"theta12\": %.17g, \"theta13\": %.17g, \"theta23\": %.17g, "
                    "\"deltaCP\": %.17g, \"maxError\": %.17g, \"cfl_factor\": %.17g, \"flavor_cfl_factor\": %.17g}\n",
                    p.nf, p.shape_order, p.ncell[0], p.ncell[1], p.ncell[2], p.phi[0], p.phi[1], p.phi[2], p.mass[0],
                    p.mass[1], p.mass[2], p.theta12, p.theta13, p.theta23, p.deltaCP, p.maxError, p.cfl_factor,
                    p.flavor_cfl_factor);
        return 0;
    }
    emu::evolve_flavor(&parms, opt);
    return 0;
}
