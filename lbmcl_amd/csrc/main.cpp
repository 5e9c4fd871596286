// main.cpp -- the lbmcl command line (main.cpp:1-51 of the reference): parse
// the options, run the cavity in float or double, print the timings and the
// statistics line.
//
// Portions restate the user interface of LBMCL (blackwut/LBMCL, GPL-3.0):
// the run sequence and the timing lines of main.cpp:7-51.  This file is distributed under the GNU General Public License
// version 3; see LICENSE at the repository root.
#include <iostream>

#include "lbm_options.hpp"
#include "lbmcl.hpp"

template <typename T>
void performSimulation(const lbm_options& opts) {
    LBMCL<T> lbmcl(opts.dim, T(opts.viscosity), T(opts.velocity), opts.iterations, opts.every, opts.vtk_path,
                   opts.lwx, opts.lwy, opts.lwz, opts.stride, opts.optimize, opts.dump_path, opts.dump_map,
                   opts.dump_f);
    lbmcl.setStorage(opts.storage);
    if (opts.arith >= 0) lbmcl.setArithmetic(opts.arith);
    lbmcl.setupSimulation(opts.platformID, opts.deviceID);
    lbmcl.printConfiguration();
    lbmcl.performSimulation();
    lbmcl.waitCompletion();

    std::cout << "   Total time: " << lbmcl.totalTimeMS() << " ms" << std::endl;
    std::cout << " Kernels time: " << lbmcl.kernelsTimeMS() << " ms" << std::endl;
    std::cout << "  Total MLUPS: " << lbmcl.MLUPS() << " MLUPS" << std::endl;
    std::cout << "Kernels MLUPS: " << lbmcl.kernelsMLUPS() << " MLUPS" << std::endl;

    std::cerr << lbmcl.statistics(';');
}

int main(int argc, char* argv[]) {
    lbm_options opts;
    opts.process_args(argc, argv);
    if (opts.use_double) {
        performSimulation<double>(opts);
    } else {
        performSimulation<float>(opts);
    }
    return 0;
}
