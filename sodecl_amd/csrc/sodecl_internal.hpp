// Internal declarations shared by the translation units of libsodecl_amd.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sodecl.h"

namespace sodecl_amd {

// Optional plugin entry points found in a system source.
struct PluginFeatures {
    bool sode_system = false;
    bool sode_system_stoch = false;
    bool calc_jacobian = false;
    bool component = false;         // double sode_system_i(int, t, y[], p[])
    bool stoch_component = false;   // double sode_system_stoch_i(int, t, y[], p[], noise[])
    bool jacobian_band = false;     // calc_jacobian_band(y, lower, diag, upper, p)
    bool jacobian_band_i = false;   // calc_jacobian_band_i(i, y, p, *lower, *diag, *upper)
    bool vjp_group = false;         // sode_system_vjp_group(...) (adjoint, group schedule)
    bool pairs = false;             // sode_pair + sode_system_pairsum (antisymmetric pairwise drift)
    bool team = false;              // sode_system_team(...) (team schedule)
    bool jacobian_params = false;   // calc_jacobian_params(y, jp[neq][npar], p) (implicit adjoint)
    bool jacobian_exact = false;    // calc_jacobian_exact(y, jac[neq][neq], p) (implicit adjoint)
    bool calls_exp = false;         // the source calls exp( (the own exponential's table goes to LDS)
};

PluginFeatures detect_features(const std::string& source);
std::string strip_comments(const std::string& source);

// Everything that selects one compiled program.
struct ProgramKey {
    int solver = 2;
    int neq = 1, npar = 0, nnoi = 0, orbits = 1, ksteps = 1;
    double dt = 0.0;
    int schedule = 1;    // 1 lane, 2 group, 4 team
    int group = 64;      // lanes per orbit (group and team schedules)
    int block = 128;
    int newton_max_iter = 1000;
    double newton_tol = 0.0;     // > 0: _epsilon1_ in place of the reference's 1e-07
    int fp_contract = 0;
    int use_pairs = 0;           // pairwise drift under the group schedule: 0 off, 1 full matrix, 2 half storage
    int one_wave_per_simd = 0;   // the launch has at most one wavefront per SIMD (occupancy hint)
    int library_sin = 0;         // sin() in the system source: 0 sodecl_sin, 1 the device library's
    int library_exp = 0;         // exp() in the system source: 0 sodecl_exp, 1 the device library's
    std::string extra_defines;   // additional "#define" lines (adjoint, batch)
};

// printf("%g") of dt, as the reference's double2string (sodeclmgr.cpp:356-367).
std::string dt_literal(double dt);

// Program text handed to hipRTC: macro prelude, OpenCL-C compatibility,
// the user system source, the integrator templates.
std::string assemble_program(const std::string& system_source, const ProgramKey& key, const PluginFeatures& f);
// Program of the host share: the same system source and lane integrator as
// host C++ (sodecl_host_lane over an orbit range, OpenMP across orbits).
std::string assemble_host_program(const std::string& system_source, const ProgramKey& key, const PluginFeatures& f);
// Program of the batched implicit integrator: the plugin and the per-class
// solver once per system size in class_neq, then the persistent kernel.
std::string assemble_batch_program(const std::string& system_source, const ProgramKey& key, const PluginFeatures& f,
                                   const std::vector<int>& class_neq);

// A compiled and loaded program.
struct Program {
    std::vector<char> code;
    hipModule_t module = nullptr;
    int device = -1;
    std::map<std::string, hipFunction_t> functions;
    hipFunction_t get(const std::string& name);   // null when absent
};

// Compiles (through an in-memory and on-disk cache keyed by the program
// text, the options and the architecture) and returns the code object.
// ilp_schedule: schedule instructions for instruction-level parallelism
// (programs whose launches have at most one wavefront per SIMD).
bool compile_program(const std::string& text, const std::string& arch, int fp_contract, std::vector<char>& code,
                     std::string& log, bool ilp_schedule = false);
std::string device_arch(int device);
// Lanes of a wavefront per orbit in the batched integrator for a size class
// of neq equations: 16 (four orbits per wavefront) up to 32 equations, 32
// (two orbits) up to 128, else 64.  The 128-equation class ran 2.5-3% faster
// on 32 lanes (four rows per lane, 110 registers, four wavefronts per SIMD)
// than on 16 (eight rows per lane, two wavefronts per SIMD by registers).
inline int batch_lanes_per_orbit(int neq) { return neq <= 32 ? 16 : neq <= 128 ? 32 : 64; }
// Occupancy hint of a size class's kernel (the wavefronts per SIMD its
// register budget is sized for): up to two rows per lane, or four rows on
// 32 lanes, need few registers (four); a whole wavefront per orbit with up
// to 8 rows per lane fits 168 registers with no spill in its Newton loop
// (three); otherwise two.
inline int batch_waves_per_simd(int neq)
{
    const int lpo = batch_lanes_per_orbit(neq), rows = (neq + lpo - 1) / lpo;
    if (rows <= 2) return 4;
    if (lpo == 32 && rows <= 4) return 4;
    return (lpo == 64 && rows <= 8) ? 3 : 2;
}

long num_launches(double tspan, double dt, int ksteps);

}  // namespace sodecl_amd

// sodecl_host_lane of a host-share library (integrators.hip, SODECL_HOST_BUILD)
typedef void (*host_lane_fn)(double* t_state, double* y_state, const double* params, double* counters, double* traj,
                             long long* newton, int orbits, int first, int last, int s0, int nseg, int final_snap,
                             long long gstep0, int nthreads);

struct sodecl_ctx {
    // configuration (sodeclmgr ctor arguments, sodeclmgr.cpp:26-83)
    std::string kernel_path;
    std::string system_path;
    std::string system_source;
    int solver = 2;
    double dt = 0.0, tspan = 0.0;
    int ksteps = 1, neq = 1, npar = 0, nnoi = 0, orbits = 1;
    int output_type = SODECL_OUTPUT_ARRAY;
    int local_group_size = 0;
    int device = 0;
    bool device_chosen = false;
    sodecl_options opts{};
    sodecl_amd::PluginFeatures features;

    // stashed host inputs (read at setup)
    const double* h_t0 = nullptr;
    const double* h_y0 = nullptr;
    const double* h_params = nullptr;
    std::vector<double> counters;   // per-orbit RNG counters (SDE)

    // device state
    bool ready = false;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    double* d_t = nullptr;
    double* d_y = nullptr;
    double* d_params = nullptr;
    double* d_counters = nullptr;
    long long* d_newton = nullptr;
    // trajectory buffers: one for a single launch, kTrajPipe for the chunked
    // copy-out (sodecl_capi.cpp: run_fast), with pinned staging alike
    static constexpr int kTrajPipe = 3;
    double* d_traj[kTrajPipe] = {};
    size_t traj_capacity = 0;       // doubles per trajectory buffer
    double* h_stage[kTrajPipe] = {};   // pinned staging for host output
    size_t stage_capacity = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;

    // programs: one per (schedule, group, block)
    std::map<std::string, std::shared_ptr<sodecl_amd::Program>> programs;
    int used_schedule = 0, used_group = 0, used_block = 0;
    // autotuner result (schedule, group, block); tuned_schedule 0 = none
    int tuned_schedule = 0, tuned_group = 0, tuned_block = 0;
    std::string tune_report;

    // output (sodeclmgr::m_output)
    std::vector<double> output;
    double* ext_output = nullptr;
    size_t ext_capacity = 0;
    size_t ext_used = 0;

    double last_runtime = 0.0;
    double last_kernel_ms = 0.0;
    std::string last_error;
    std::string last_program;

    // hipGraph of the launch sequence of a run
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    std::string graph_key;

    // host share (sodecl_options.host_share): the host build of the lane
    // integrator, and host copies of the state of the orbits it runs
    void* host_handle = nullptr;
    host_lane_fn host_fn = nullptr;
    std::vector<double> hs_t, hs_y, hs_p, hs_cnt;
    std::vector<long long> hs_newton;
    double last_host_ms = 0.0;
    int last_host_orbits = 0;
};

namespace sodecl_amd {
int fail(sodecl_ctx* ctx, const std::string& msg);
ProgramKey make_key(const sodecl_ctx* ctx, int schedule, int group, int block, const std::string& extra);
bool hip_ok(sodecl_ctx* ctx, hipError_t e, const char* what);
// Program for the given schedule on the context's device (compiled on first use).
std::shared_ptr<Program> get_program(sodecl_ctx* ctx, int schedule, int group, int block,
                                     const std::string& extra_defines = std::string());
// The program if it is already compiled and loaded on the context, else null
// (no compilation, no module load).
std::shared_ptr<Program> find_program(sodecl_ctx* ctx, int schedule, int group, int block,
                                      const std::string& extra_defines = std::string());
// Backward sweep of sodecl_adjoint_device into a given workspace of
// work_doubles doubles, with the adjoint program already loaded (a
// checkpointed forward run loads it); no allocation.
int adjoint_into_workspace(sodecl_ctx* c, const double* ckpt_t, const double* ckpt_y, const double* params,
                           const double* cot, double* grad_y0, double* grad_params, double* work,
                           size_t work_doubles, int ckpt_every, void* stream);
// Device address of the library's static scratch array on device (the
// device becomes current); resolved once per device.  Returns null and sets
// err on failure.
double* scratch_address(int device, std::string& err);
// Doubles in the static scratch array.
size_t scratch_doubles();
// Host share: builds (or finds cached) the host library of the context's
// program; loads it into the context; unloads it.
bool build_host_library(const sodecl_ctx* c, std::string& path, std::string& err);
bool load_host_share(sodecl_ctx* c, std::string& err);
void unload_host_share(sodecl_ctx* c);
// stage_ws: when not null, the group kernel also writes every step's stage
// vectors there (the adjoint's whole-run workspace; group schedule only).
int run_device_impl(sodecl_ctx* c, double* t0, double* y0, const double* params, double* counters, double* traj,
                    double* ckpt_t, double* ckpt_y, int ckpt_every, void* stream, double* stage_ws = nullptr);
// The schedule, group size and block size runs of this context use.
int launch_config(const sodecl_ctx* c, int* schedule, int* group, int* block);
}  // namespace sodecl_amd
