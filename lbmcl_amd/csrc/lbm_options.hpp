// lbm_options.hpp -- command line of lbmcl.
//
// Same options, defaults and validation as the reference
// (libs/lbm_options.hpp:13-196): -P -D -d -n -u -i -e -v -w -s -F -o -p -m -f -h.
// Two long-only additions: --storage native|fp16|bf16 selects the 16-bit
// shifted population storage (FP32 arithmetic), which the reference lacks, and
// --arith opencl|exact the collision arithmetic.  Without -o and with native
// storage the default is opencl: the reference's kernels.cl evaluated as
// ROCm's OpenCL compiler does, so the outputs are those of the reference on
// the same GPU, bit for bit; exact evaluates each operation as written.
//
// Portions restate the user interface of LBMCL (blackwut/LBMCL, GPL-3.0):
// the option table, help text, defaults and error messages of
// libs/lbm_options.hpp:13-196.  This file is distributed under the GNU General Public License
// version 3; see LICENSE at the repository root.
#pragma once

#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>

#include "lbm_hip.h"

#define RESULTS_FOLDER "./results"

struct lbm_options {
    int platformID = -1;
    int deviceID = -1;
    size_t dim = 8;
    double viscosity = 0.0089;
    double velocity = 0.05;
    size_t iterations = 10;
    size_t every = 1;
    std::string vtk_path = RESULTS_FOLDER;
    size_t lwx = 8;
    size_t lwy = 8;
    size_t lwz = 8;
    size_t stride = 32;
    bool use_double = false;
    bool optimize = false;
    std::string dump_path = RESULTS_FOLDER;
    bool dump_map = false;
    bool dump_f = false;
    int storage = LBM_STORE_NATIVE;
    int arith = -1;  // LBM_ARITH_*; -1: LBM_ARITH_FUSED with -o or 16-bit storage

    [[noreturn]] void print_help() {
        std::cout << "-P  --platform            Use the specified platform                     \n"
                     "-D  --device              Use the specified device                       \n"
                     "-d  --dim                 Set the lattice cube dimension                 \n"
                     "-n  --viscosity           Set the fluid viscosity                        \n"
                     "-u  --velocity            Set the x velocity of the moving wall          \n"
                     "-i  --iterations          Specify the number of iterations               \n"
                     "-e  --every               Save simulation results every N iterations     \n"
                     "-w  --work_group_size     Specify the work group size \"x,y,z\"          \n"
                     "-s  --stride              Specify the stride used in CSoA memory layout  \n"
                     "-F  --use_double          Make use of \"double\" type                    \n"
                     "-o  --optimize            Request fast math (fused collision arithmetic) \n"
                     "-v  --vtk_path            Specify where store VTI files                  \n"
                     "-p  --dump_path           Specify where store dumps                      \n"
                     "-m  --dump_map            Dump the lattice map                           \n"
                     "-f  --dump_f              Dump the lattice \"f\" for each iteration      \n"
                     "    --storage             Population storage: native, fp16 or bf16       \n"
                     "    --arith               Collision arithmetic: opencl (default, the     \n"
                     "                          reference's own rounding on this GPU) or exact \n"
                     "-h  --help                Show this help message and exit                \n";
        std::exit(1);
    }

    static int to_int(const char* s) {
        try {
            return std::stoi(s);
        } catch (...) {
            std::cerr << "invalid number: " << s << std::endl;
            std::exit(1);
        }
    }

    void process_args(int argc, char* argv[]) {
        opterr = 0;
        const char* const short_opts = "P:D:d:n:u:i:e:v:w:s:Fop:mfh";
        const option long_opts[] = {
            {"platform", required_argument, nullptr, 'P'},
            {"device", required_argument, nullptr, 'D'},
            {"dim", required_argument, nullptr, 'd'},
            {"viscosity", required_argument, nullptr, 'n'},
            {"velocity", required_argument, nullptr, 'u'},
            {"iterations", required_argument, nullptr, 'i'},
            {"every", optional_argument, nullptr, 'e'},
            {"work_group_size", optional_argument, nullptr, 'w'},
            {"stride", optional_argument, nullptr, 's'},
            {"use_double", no_argument, nullptr, 'F'},
            {"optimize", no_argument, nullptr, 'o'},
            {"vtk_path", optional_argument, nullptr, 'v'},
            {"dump_path", optional_argument, nullptr, 'p'},
            {"dump_map", no_argument, nullptr, 'm'},
            {"dump_f", no_argument, nullptr, 'f'},
            {"storage", required_argument, nullptr, 'S'},
            {"arith", required_argument, nullptr, 'A'},
            {"help", no_argument, nullptr, 'h'},
            {nullptr, no_argument, nullptr, 0},
        };
        while (true) {
            const int opt = getopt_long(argc, argv, short_opts, long_opts, nullptr);
            if (opt < 0) break;
            const char* arg = optarg ? optarg : "";
            int v = 0;
            double r = 0.0;
            switch (opt) {
                case 'P':
                    if ((v = to_int(arg)) < 0) { std::cerr << "Please enter a valid platform" << std::endl; std::exit(1); }
                    platformID = v;
                    break;
                case 'D':
                    if ((v = to_int(arg)) < 0) { std::cerr << "Please enter a valid device" << std::endl; std::exit(1); }
                    deviceID = v;
                    break;
                case 'd':
                    if ((v = to_int(arg)) < 0) { std::cerr << "Please enter a valid lattice dimension" << std::endl; std::exit(1); }
                    dim = size_t(v);
                    break;
                case 'n':
                    if ((r = std::atof(arg)) < 0) { std::cerr << "Please enter a valid viscosity value" << std::endl; std::exit(1); }
                    viscosity = r;
                    break;
                case 'u':
                    velocity = std::atof(arg);
                    break;
                case 'i':
                    if ((v = to_int(arg)) < 0) { std::cerr << "Please enter a valid number of iterations" << std::endl; std::exit(1); }
                    iterations = size_t(v);
                    break;
                case 'e':
                    if ((v = to_int(arg)) < 0) {
                        std::cerr << "Please enter a valid number for save simulation results every N iterations" << std::endl;
                        std::exit(1);
                    }
                    every = size_t(v);
                    break;
                case 'w':
                    if (std::sscanf(arg, "%zu,%zu,%zu", &lwx, &lwy, &lwz) < 1) {}
                    break;
                case 's':
                    if ((v = to_int(arg)) < 0) { std::cerr << "Please enter a valid number for stride value" << std::endl; std::exit(1); }
                    stride = size_t(v);
                    break;
                case 'F':
                    use_double = true;
                    break;
                case 'o':
                    optimize = true;
                    break;
                case 'v':
                    vtk_path = arg;
                    if (vtk_path.empty()) {
                        vtk_path = RESULTS_FOLDER;
                        std::cout << "VTI files will be stored in:" << vtk_path << std::endl;
                    }
                    break;
                case 'p':
                    dump_path = arg;
                    if (dump_path.empty()) {
                        dump_path = RESULTS_FOLDER;
                        std::cout << "dump files will be stored in: " << dump_path << std::endl;
                    }
                    break;
                case 'm':
                    dump_map = true;
                    break;
                case 'f':
                    dump_f = true;
                    break;
                case 'S':
                    if (!std::strcmp(arg, "native")) storage = LBM_STORE_NATIVE;
                    else if (!std::strcmp(arg, "fp16")) storage = LBM_STORE_FP16_SHIFT;
                    else if (!std::strcmp(arg, "bf16")) storage = LBM_STORE_BF16_SHIFT;
                    else { std::cerr << "Please enter a valid storage (native, fp16, bf16)" << std::endl; std::exit(1); }
                    break;
                case 'A':
                    if (!std::strcmp(arg, "exact")) arith = LBM_ARITH_EXACT;
                    else if (!std::strcmp(arg, "opencl")) arith = LBM_ARITH_OPENCL;
                    else { std::cerr << "Please enter a valid arithmetic (exact, opencl)" << std::endl; std::exit(1); }
                    break;
                case 'h':
                case '?':
                default:
                    print_help();
            }
        }
        if (arith >= 0 && optimize) {
            std::cerr << "-o selects the fused arithmetic; drop --arith or -o" << std::endl;
            std::exit(1);
        }
        if (arith == LBM_ARITH_OPENCL && storage != LBM_STORE_NATIVE) {
            std::cerr << "--arith opencl needs native storage" << std::endl;
            std::exit(1);
        }
        if (storage != LBM_STORE_NATIVE && use_double) {
            std::cerr << "16-bit shifted storage computes in single precision; drop -F" << std::endl;
            std::exit(1);
        }
        if (arith < 0 && !optimize && storage == LBM_STORE_NATIVE) arith = LBM_ARITH_OPENCL;
    }
};
