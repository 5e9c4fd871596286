// lbmcl.hpp -- the LBMCL<T> simulation class over the HIP C ABI.
//
// Same public interface, argument meaning, output files, timing figures and
// error behaviour as the reference class (lbmcl.hpp:33-679); only the device
// layer underneath changed: every OpenCL object and call is replaced by the
// C ABI of include/lbm_hip.h.  A nonzero status from the library prints
// `file:line what(code) - message` to stderr and exits with the status, as
// CLUCheckErrorExit does (libs/CLUtil.hpp:101-117).
//
// Portions restate the user interface of LBMCL (blackwut/LBMCL, GPL-3.0):
// the class interface, the constructor's checks and messages, the
// printed configuration and the statistics line of lbmcl.hpp:33-679.  This file is distributed under the GNU General Public License
// version 3; see LICENSE at the repository root.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lbm_hip.h"
#include "lbm_io.hpp"
#include "lbm_output.hpp"

#define LBM_CHECK_EXIT(call, what) lbm_check_exit((call), (what), __FILE__, __LINE__)

inline void lbm_check_exit(int status, const std::string& what, const char* file, int line) {
    if (status != LBM_OK) {
        std::cerr << file << ":" << line << " " << what << "(" << status << ") - " << lbm_last_error() << std::endl;
        std::exit(status);
    }
}

#define INITIALIZE_KERNEL_NAME "initialize"
#define COMPUTE_KERNEL_NAME "compute"
#define READ_MAP_NAME "read_map"
#define READ_F_NAME "read_f"
#define READ_RHO_NAME "read_rho"
#define READ_U_NAME "read_u"

template <typename T>
class LBMCL {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "Only float or double data type is valid.");

   private:
    size_t dim;
    T viscosity;
    T velocity;
    size_t iterations;
    size_t every;
    std::string vtk_path;
    size_t lws[3];
    size_t stride;
    bool optimize;
    std::string dump_path;
    bool dump_map;
    bool dump_f;
    bool dump_data = false;
    int storage = LBM_STORE_NATIVE;
    int arith = -1;  // LBM_ARITH_*; -1: from optimize and the storage

    int device_id = -1;
    std::string device_name;
    lbm_ctx* ctx = nullptr;

    // files are written by a background thread from asynchronous snapshots
    std::unique_ptr<lbm_out::Writer> writer;

    size_t f_dim() const { return dim * dim * dim * lbm_io::Q; }
    size_t u_dim() const { return dim * dim * dim * 3; }
    size_t rho_dim() const { return dim * dim * dim; }
    size_t map_dim() const { return dim * dim * dim; }
    size_t wet_dim() const { return (dim - 2) * (dim - 2) * (dim - 2); }

    size_t f_size() const { return f_dim() * sizeof(T); }
    size_t u_size() const { return u_dim() * sizeof(T); }
    size_t rho_size() const { return rho_dim() * sizeof(T); }
    size_t map_size() const { return map_dim() * sizeof(int); }

    size_t device_memory_size_b() const {
        const size_t lattice = (storage == LBM_STORE_NATIVE) ? f_size() : f_dim() * 2;
        return lattice * 2 + u_size() + rho_size() + (dump_map ? map_size() : 0);  // the map only when dumped
    }
    size_t device_memory_size_k() const { return device_memory_size_b() / (1 << 10); }
    size_t device_memory_size_m() const { return device_memory_size_b() / (1 << 20); }

    // powers of two through the bit width of the value
    static int bit_width(size_t x) { return x ? 64 - __builtin_clzll(uint64_t(x)) : 0; }
    static bool is_power_of_two(size_t x) { return x != 0 && (x & (x - 1)) == 0; }
    static size_t log2i(size_t x) { return x > 1 ? size_t(bit_width(x) - 1) : 0; }
    // the largest power of two <= x (0 for 0)
    static size_t previous_power_of_two(size_t x) { return x ? size_t(1) << (bit_width(x) - 1) : 0; }

    // What the reference passes to its JIT build (lbmcl.hpp:131-156); here the
    // same values configure the precompiled kernels.
    std::string kernelOptionsStr() const {
        std::stringstream b;
        b << "-DDIM=" << dim << " ";
        b << "-DLWS=" << lws[0] << " ";
        b << "-DSTRIDE_DIV=" << log2i(stride) << " ";
        b << "-DSTRIDE_MOD=" << (stride - 1) << " ";
        b << "-DVISCOSITY=" << viscosity << " ";
        b << "-DVELOCITY=" << velocity << " ";
        b << (std::is_same<T, float>::value ? "-DFP_SINGLE " : "-DFP_DOUBLE ");
        if (storage == LBM_STORE_FP16_SHIFT) b << "-DSTORAGE_FP16_SHIFT ";
        if (storage == LBM_STORE_BF16_SHIFT) b << "-DSTORAGE_BF16_SHIFT ";
        if (optimize) b << "-DOPTIMIZE ";
        return b.str();
    }

    // The dumps: a snapshot is enqueued behind the work enqueued so far and the
    // background writer formats the file while the device keeps computing.
    int64_t snapshot(int32_t what, int32_t buffer, const char* name) {
        int64_t ticket = 0;
        LBM_CHECK_EXIT(lbm_snapshot(ctx, what, buffer, &ticket), name);
        return ticket;
    }

    void storeMap() {
        const int64_t t = snapshot(LBM_SNAPSHOT_MAP, 0, READ_MAP_NAME);
        writer->push(lbm_out::Writer::MAP_DUMP, t, dump_path + "/map.dump");
    }

    // buffer 0 is the reference's f_collide, 1 its f_stream
    void storeF(int buffer, size_t iteration) {
        const int64_t t = snapshot(LBM_SNAPSHOT_F, buffer, READ_F_NAME);
        writer->push(lbm_out::Writer::F_DUMP, t, lbm_io::numbered_path(dump_path, "f_", iteration, iterations, ".dump"));
    }

    void storeData(size_t iteration) {
        const int64_t t = snapshot(LBM_SNAPSHOT_MACRO, 0, READ_RHO_NAME);
        writer->push(lbm_out::Writer::VTI, t, lbm_io::numbered_path(vtk_path, "lbmcl.", iteration, iterations, ".vti"));
    }

    int selectDevice(int deviceID) {
        int32_t count = 0;
        LBM_CHECK_EXIT(lbm_get_device_count(&count), "lbm_get_device_count");
        // An invalid id lists the devices and asks on stdin, as the reference
        // does (libs/CLUtil.hpp:145-178); end of input is an error instead of a
        // retry loop.
        while (deviceID < 0 || deviceID >= count) {
            for (int d = 0; d < count; ++d) {
                char name[256];
                LBM_CHECK_EXIT(lbm_get_device_name(d, name, sizeof name), "lbm_get_device_name");
                std::cout << "#" << d << " [GPU] " << name << std::endl;
            }
            std::cout << std::endl << "Select a device: ";
            if (!(std::cin >> deviceID)) {
                std::cerr << "no device selected" << std::endl;
                std::exit(LBM_ERR_DEVICE);
            }
            std::cout << std::endl;
        }
        return deviceID;
    }

   public:
    LBMCL(size_t dim, T viscosity, T velocity, size_t iterations, size_t every, std::string vtk_path = "",
          size_t lwx = 1, size_t lwy = 1, size_t lwz = 1, size_t stride = 32, bool optimize = true,
          std::string dump_path = "", bool dump_map = false, bool dump_f = false)
        : dim(dim),
          viscosity(viscosity),
          velocity(velocity),
          iterations(iterations),
          every(every),
          vtk_path(vtk_path),
          stride(stride),
          optimize(optimize),
          dump_path(dump_path),
          dump_map(dump_map),
          dump_f(dump_f) {
        dump_data = (every != 0);
        if (!is_power_of_two(dim)) {
            this->dim = previous_power_of_two(dim);
            std::cout << "dim is rounded to the previous power of 2: " << this->dim << std::endl;
        }
        if (lwx == 0) lwx = 1;
        if (lwy == 0) lwy = 1;
        if (lwz == 0) lwz = 1;
        if ((lwx * lwy * lwz) > (this->dim * this->dim * this->dim)) {
            std::cerr << "Please enter a good work_group_size to run the simulation" << std::endl;
            std::exit(-1);
        }
        lws[0] = lwx;
        lws[1] = lwy;
        lws[2] = lwz;
        if (!is_power_of_two(this->stride)) {
            this->stride = previous_power_of_two(this->stride);
            std::cout << "stride is rounded to the previous power of 2: " << this->stride << std::endl;
        }
    }

    LBMCL(const LBMCL&) = delete;
    LBMCL& operator=(const LBMCL&) = delete;

    // Population storage (LBM_STORE_*); call before setupSimulation.  The
    // 16-bit shifted forms exist for float only and have no counterpart in the
    // reference.
    void setStorage(int s) { storage = s; }

    // Collision arithmetic (LBM_ARITH_*); call before setupSimulation.  By
    // default the constructor's `optimize` selects LBM_ARITH_FUSED, and
    // otherwise LBM_ARITH_OPENCL, the reference's own rounding on this GPU
    // (16-bit storage always uses the deviation form); LBM_ARITH_EXACT evaluates
    // each operation as written.
    void setArithmetic(int a) { arith = a; }

    // Select the device and allocate all device memory (lbmcl.hpp:392-482).
    // platformID is accepted for interface compatibility: HIP exposes one
    // platform, and any id other than 0 or -1 is reported and treated as 0.
    void setupSimulation(int platformID, int deviceID) {
        if (platformID > 0) std::cout << "HIP has a single platform; using platform 0" << std::endl;
        device_id = selectDevice(deviceID);
        char name[256];
        LBM_CHECK_EXIT(lbm_get_device_name(device_id, name, sizeof name), "lbm_get_device_name");
        device_name = name;

        lbm_config cfg{};
        cfg.struct_size = sizeof(lbm_config);
        cfg.dim = int32_t(dim);
        cfg.precision = std::is_same<T, float>::value ? LBM_F32 : LBM_F64;
        cfg.storage = storage;
        cfg.device = device_id;
        cfg.stride = int64_t(stride);
        cfg.viscosity = double(viscosity);
        cfg.velocity = double(velocity);
        cfg.lws[0] = int32_t(lws[0]);
        cfg.lws[1] = int32_t(lws[1]);
        cfg.lws[2] = int32_t(lws[2]);
        cfg.optimize = arith >= 0                                         ? arith
                       : (optimize || storage != LBM_STORE_NATIVE) ? LBM_ARITH_FUSED
                                                                     : LBM_ARITH_OPENCL;
        cfg.kernel = LBM_KERNEL_AUTO;
        cfg.profile = 1;
        cfg.xcd_remap = 1;
        LBM_CHECK_EXIT(lbm_create(&cfg, &ctx), "lbm_create");

        if (dump_map || dump_f || dump_data)
            writer.reset(new lbm_out::Writer(ctx, dim, !std::is_same<T, float>::value, lbm_out::io_threads()));
    }

    // Enqueue the whole simulation (lbmcl.hpp:490-521).  Dumps are snapshots
    // written in the background; this blocks only when every snapshot slot is
    // still waiting for its file to be written.
    void performSimulation() {
        LBM_CHECK_EXIT(lbm_initialize(ctx), INITIALIZE_KERNEL_NAME);
        if (dump_map) storeMap();
        if (dump_data) storeData(0);
        if (dump_f) storeF(0, 0);
        for (size_t it = 1; it <= iterations; ++it) {
            const int store = (dump_data && (it % every == 0)) ? 1 : 0;
            LBM_CHECK_EXIT(lbm_compute(ctx, int64_t(it), store, 0), COMPUTE_KERNEL_NAME);
            if (store) storeData(it);
            if (dump_f) storeF((it % 2 == 0) ? 1 : 0, it);
        }
    }

    // The device is idle and every file is written.
    void waitCompletion() {
        LBM_CHECK_EXIT(lbm_finish(ctx), "lbm_finish");
        if (writer) {
            std::string what;
            const int rc = writer->drain(&what);
            if (rc) {
                std::cerr << __FILE__ << ":" << __LINE__ << " " << what << "(" << rc << ")" << std::endl;
                std::exit(rc);
            }
        }
    }

    void performSimulationAndWait() {
        performSimulation();
        waitCompletion();
    }

    // First event start to last event end (lbmcl.hpp:548-556).
    double totalTimeMS() {
        waitCompletion();
        double ms = 0.0;
        LBM_CHECK_EXIT(lbm_total_time_ms(ctx, &ms), "lbm_total_time_ms");
        return ms;
    }

    // Sum of the compute kernel durations (lbmcl.hpp:562-574).
    double kernelsTimeMS() {
        waitCompletion();
        double ms = 0.0;
        LBM_CHECK_EXIT(lbm_kernels_time_ms(ctx, &ms), "lbm_kernels_time_ms");
        return ms;
    }

    std::vector<std::pair<std::string, double>> kernelsTimingsMS() {
        waitCompletion();
        static const char* names[] = {INITIALIZE_KERNEL_NAME, COMPUTE_KERNEL_NAME, READ_MAP_NAME, READ_F_NAME,
                                      READ_RHO_NAME,          READ_U_NAME,         "diagnostics"};
        int64_t n = 0;
        LBM_CHECK_EXIT(lbm_event_count(ctx, &n), "lbm_event_count");
        std::vector<std::pair<std::string, double>> timings;
        timings.reserve(size_t(n));
        for (int64_t i = 0; i < n; ++i) {
            int32_t kind = 0;
            double ms = 0.0;
            LBM_CHECK_EXIT(lbm_event_get(ctx, i, &kind, &ms), "lbm_event_get");
            timings.emplace_back(names[kind], ms);
        }
        return timings;
    }

    // (dim-2)^3 * iterations / time (lbmcl.hpp:604-620).
    double MLUPS() { return double(wet_dim() * iterations) / (totalTimeMS() * 1000); }
    double kernelsMLUPS() { return double(wet_dim() * iterations) / (kernelsTimeMS() * 1000); }

    void printConfiguration() {
        const std::string prec = (std::is_same<T, float>::value ? "single" : "double");
        std::cout << std::boolalpha << "kernel options   = " << kernelOptionsStr() << "\n"
                  << "device           = " << device_name << "\n"
                  << "dim              = " << dim << "\n"
                  << "viscosity        = " << viscosity << "\n"
                  << "velocity         = " << velocity << "\n"
                  << "Device Mem. (B)  = " << device_memory_size_b() << "\n"
                  << "Device Mem. (KB) = " << device_memory_size_k() << "\n"
                  << "Device Mem. (MB) = " << device_memory_size_m() << "\n"
                  << "iterations       = " << iterations << "\n"
                  << "work_group_size  = (" << lws[0] << ", " << lws[1] << ", " << lws[2] << ")\n"
                  << "stride           = " << stride << "\n"
                  << "precision        = " << prec << "\n"
                  << "optimize         = " << optimize << "\n"
                  << "every            = " << every << "\n"
                  << "VTK PATH         = " << vtk_path << "\n"
                  << "DUMP F           = " << dump_f << "\n"
                  << "DUMP MAP         = " << dump_map << "\n";
    }

    // The ';'-separated line benchmark.sh parses (lbmcl.hpp:648-669).
    std::string statistics(char separator) {
        const std::string prec = (std::is_same<T, float>::value ? "single" : "double");
        std::stringstream stat;
        stat << device_name.c_str() << separator << prec << separator << dim << separator << iterations << separator
             << every << separator << std::setw(3) << std::setfill('0') << lws[0] << "," << std::setw(3)
             << std::setfill('0') << lws[1] << "," << std::setw(3) << std::setfill('0') << lws[2] << separator
             << stride << separator << optimize << separator << totalTimeMS() << separator << kernelsTimeMS()
             << separator << MLUPS() << separator << kernelsMLUPS() << "\n";
        return stat.str();
    }

    ~LBMCL() {
        writer.reset();  // finishes the pending files; uses ctx
        if (ctx) lbm_destroy(ctx);
    }
};
