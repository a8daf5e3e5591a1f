// sodecl::sodeclmgr -- the reference's host manager class
// (include/sodeclmgr.hpp:27-369 of avramidis/sodecl) with the same public
// surface, implemented over the C ABI of include/sodecl.h on one MI355X.
//
// OpenCL types become plain ones (cl_uint -> uint32_t, cl_double -> double);
// choose_device takes the HIP device ordinal as device_num and platform 0.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sodecl.h"

namespace sodecl {

// include/solver_Type.hpp:14-20
enum class solver_Type { StochasticEuler = 0, Euler = 1, RungeKutta = 2, ImplicitEuler = 3, ImplicitMidpoint = 4 };
// include/output_Type.hpp:14-18
enum class output_Type { None = 0, Array = 1, File = 2 };
// include/device_Type.hpp:14-19
enum class device_Type : uint64_t {
    ALL = SODECL_DEVICE_TYPE_ALL,
    CPU = SODECL_DEVICE_TYPE_CPU,
    GPU = SODECL_DEVICE_TYPE_GPU,
    ACCELERATOR = SODECL_DEVICE_TYPE_ACCELERATOR
};

class sodeclmgr {
public:
    std::vector<double> m_output;   // integration output, [snap][eq][orbit] per run
    int m_output_size = 0;

    sodeclmgr(std::string kernel_path_str, char* sode_system_str, solver_Type solver, double dt, double int_time,
              int kernel_steps, int num_equat, int num_params, int num_noi, int list_size, output_Type output_type);
    ~sodeclmgr();
    sodeclmgr(const sodeclmgr&) = delete;
    sodeclmgr& operator=(const sodeclmgr&) = delete;

    int choose_device(uint32_t platform_num, device_Type device_type, uint32_t device_num);
    void set_local_group_size(int local_group_size);
    void set_t0(double* t0);
    void set_y0(double* y0);
    void set_params(double* params);
    int read_binary_data_from_file(char* filename, double* data);
    int read_data_from_file(char* filename, double* data);
    int run_sode_solver();
    int setup_sode_solver();

    // no reference counterpart
    int set_rng_counters(const double* counters);
    int set_options(const sodecl_options& options);
    sodecl_ctx* context() { return m_ctx; }
    const char* last_error() const;

private:
    sodecl_ctx* m_ctx = nullptr;
    int m_created = 0;
    int m_num_equat, m_list_size;
};

}  // namespace sodecl
