// sodecl::sodeclmgr over the C ABI (reference: src/sodeclmgr.cpp).
#include "sodeclmgr.hpp"

#include <iostream>
#include <algorithm>
#include <limits>

namespace sodecl {

sodeclmgr::sodeclmgr(std::string kernel_path_str, char* sode_system_str, solver_Type solver, double dt,
                     double int_time, int kernel_steps, int num_equat, int num_params, int num_noi, int list_size,
                     output_Type output_type)
    : m_num_equat(num_equat), m_list_size(list_size)
{
    m_created = sodecl_create(&m_ctx, kernel_path_str.c_str(), sode_system_str, (int)solver, dt, int_time,
                              kernel_steps, num_equat, num_params, num_noi, list_size, (int)output_type);
    if (!m_created) std::cerr << sodecl_last_error(m_ctx) << std::endl;
}

sodeclmgr::~sodeclmgr() { sodecl_destroy(m_ctx); }

int sodeclmgr::choose_device(uint32_t platform_num, device_Type device_type, uint32_t device_num)
{
    return sodecl_choose_device(m_ctx, platform_num, (uint64_t)device_type, device_num);
}

void sodeclmgr::set_local_group_size(int local_group_size) { sodecl_set_local_group_size(m_ctx, local_group_size); }
void sodeclmgr::set_t0(double* t0) { sodecl_set_t0(m_ctx, t0); }
void sodeclmgr::set_y0(double* y0) { sodecl_set_y0(m_ctx, y0); }
void sodeclmgr::set_params(double* params) { sodecl_set_params(m_ctx, params); }

int sodeclmgr::read_binary_data_from_file(char* filename, double* data)
{
    size_t n = 0;
    return sodecl_read_binary_data_from_file(filename, data, std::numeric_limits<size_t>::max() / 8, &n);
}

int sodeclmgr::read_data_from_file(char* filename, double* data)
{
    size_t n = 0;
    return sodecl_read_data_from_file(filename, data, std::numeric_limits<size_t>::max() / 8, &n);
}

int sodeclmgr::setup_sode_solver()
{
    if (!m_created) return 0;
    int ok = sodecl_setup(m_ctx);
    if (!ok) std::cout << sodecl_last_error(m_ctx) << std::endl;
    return ok;
}

int sodeclmgr::run_sode_solver()
{
    // The context accumulates every run's output, as the reference's
    // push_back loop appends to m_output (sodeclmgr.cpp:979-1015).
    int ok = sodecl_run(m_ctx);
    if (!ok) {
        std::cout << sodecl_last_error(m_ctx) << std::endl;
        return 0;
    }
    const double* out = sodecl_output(m_ctx);
    size_t n = sodecl_output_size(m_ctx);
    m_output.assign(out, out + n);
    m_output_size = (int)std::min<size_t>(n, (size_t)std::numeric_limits<int>::max());
    return 1;
}

int sodeclmgr::set_rng_counters(const double* counters) { return sodecl_set_rng_counters(m_ctx, counters); }
int sodeclmgr::set_options(const sodecl_options& options) { return sodecl_set_options(m_ctx, &options); }
const char* sodeclmgr::last_error() const { return sodecl_last_error(m_ctx); }

}  // namespace sodecl
