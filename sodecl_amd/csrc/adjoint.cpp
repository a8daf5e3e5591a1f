// C ABI of the discrete adjoint (include/sodecl.h, no reference counterpart).
//
// Forward: sodecl_run_device_checkpointed stores the state every ckpt_every
// steps while it integrates.  Backward: sodecl_adjoint_device recomputes each
// checkpoint segment and sweeps it backwards (device/adjoint.hip), giving
// dL/dy0 and dL/dparams of every orbit from dL/d(trajectory).
#include <cmath>
#include <string>

#include "sodecl_internal.hpp"

using namespace sodecl_amd;

namespace {

const char* kAdjointDefines = "#define SODECL_ADJOINT 1\n";

int nstage(const sodecl_ctx* c) { return c->solver == SODECL_RUNGE_KUTTA ? 4 : 1; }

long total_steps(const sodecl_ctx* c) { return num_launches(c->tspan, c->dt, c->ksteps) * (long)c->ksteps; }

bool implicit(const sodecl_ctx* c)
{
    return c->solver == SODECL_IMPLICIT_EULER || c->solver == SODECL_IMPLICIT_MIDPOINT;
}

// Euler/RK4: group schedule with the plugin's group VJP.  IE/IM: lane
// schedule with the plugin's dense Jacobians (calc_jacobian,
// calc_jacobian_params).
int check_adjoint(sodecl_ctx* c, int& sched, int& group, int& block)
{
    launch_config(c, &sched, &group, &block);
    if (implicit(c)) {
        if (!c->features.calc_jacobian || (c->npar > 0 && !c->features.jacobian_params))
            return fail(c, "the implicit adjoint needs calc_jacobian and calc_jacobian_params in the system source");
        if (sched != SODECL_SCHEDULE_LANE) return fail(c, "the implicit adjoint runs under the lane schedule");
        return 1;
    }
    if (c->solver != SODECL_EULER && c->solver != SODECL_RUNGE_KUTTA)
        return fail(c, "the adjoint supports the Euler, RK4, implicit Euler and implicit midpoint solvers");
    if (!c->features.component || !c->features.vjp_group)
        return fail(c, "the adjoint needs sode_system_i and sode_system_vjp_group in the system source");
    if (sched != SODECL_SCHEDULE_GROUP)
        return fail(c, "the adjoint runs under the group schedule (set schedule to group)");
    return 1;
}

}  // namespace

extern "C" {

int sodecl_default_checkpoint_interval(const sodecl_ctx* c)
{
    if (!c) return 0;
    long s = total_steps(c);
    if (s <= 0) return 1;
    return (int)std::max(1.0, std::ceil(std::sqrt((double)s)));
}

long sodecl_checkpoint_count(const sodecl_ctx* c, int ckpt_every)
{
    if (!c || ckpt_every <= 0) return 0;
    long s = total_steps(c);
    return (s + ckpt_every - 1) / ckpt_every;
}

size_t sodecl_adjoint_workspace_size(const sodecl_ctx* c, int ckpt_every)
{
    if (!c || ckpt_every <= 0) return 0;
    int sched = 0, group = 0, block = 0;
    launch_config(c, &sched, &group, &block);
    if (implicit(c))   // rows 0..ckpt_every of [neq + 1][orbits]
        return sched == SODECL_SCHEDULE_LANE ? (size_t)c->orbits * (size_t)(ckpt_every + 1) * (size_t)(c->neq + 1) : 0;
    if (sched != SODECL_SCHEDULE_GROUP) return 0;
    const int opb = block / group;
    const size_t padded = (size_t)((c->orbits + opb - 1) / opb) * opb;
    return padded * (size_t)ckpt_every * (size_t)(nstage(c) * c->neq + 1);
}

int sodecl_run_device_checkpointed(sodecl_ctx* c, double* t0, double* y0, const double* params, double* counters,
                                   double* traj, double* ckpt_t, double* ckpt_y, int ckpt_every, void* stream)
{
    if (!c) return 0;
    if (!ckpt_t || !ckpt_y || ckpt_every <= 0) return fail(c, "checkpoint buffers and interval required");
    if (!run_device_impl(c, t0, y0, params, counters, traj, ckpt_t, ckpt_y, ckpt_every, stream)) return 0;
    // Checkpoints are taken for a backward sweep: load its program now, while
    // the forward kernel runs, so that the sweep itself compiles and loads
    // nothing.  A configuration without an adjoint keeps its old error state.
    const std::string before = c->last_error;
    int sched = 0, group = 0, block = 0;
    std::shared_ptr<Program> prog;
    if (check_adjoint(c, sched, group, block)) prog = get_program(c, sched, group, block, kAdjointDefines);
    if (prog)
        prog->get(implicit(c) ? "sodecl_lane_adjoint" : "sodecl_group_adjoint");
    else
        c->last_error = before;
    return 1;
}

}  // extern "C"

namespace {

// Launch of the adjoint kernel; stored: the workspace already holds every
// step's rows (written by sodecl_run_device_stored), no checkpoints.
int launch_adjoint(sodecl_ctx* c, const double* ckpt_t, const double* ckpt_y, const double* params,
                   const double* cot, double* grad_y0, double* grad_params, double* work, int ckpt_every, int stored,
                   void* stream, bool loaded_only = false)
{
    int sched = 0, group = 0, block = 0;
    if (!check_adjoint(c, sched, group, block)) return 0;
    if (stored && implicit(c)) return fail(c, "stored stage vectors are available for Euler and RK4 only");
    if (!hip_ok(c, hipSetDevice(c->device), "hipSetDevice")) return 0;
    auto prog = loaded_only ? find_program(c, sched, group, block, kAdjointDefines)
                            : get_program(c, sched, group, block, kAdjointDefines);
    if (!prog) {
        if (loaded_only)
            return fail(c, "the adjoint program is not loaded on this context: a checkpointed forward run "
                           "(sodecl_run_device_checkpointed) loads it");
        return 0;
    }
    const char* kname = implicit(c) ? "sodecl_lane_adjoint" : "sodecl_group_adjoint";
    hipFunction_t fn = prog->get(kname);
    if (!fn) return fail(c, std::string("kernel ") + kname + " missing from the program");
    const long steps = total_steps(c);
    if (steps > 2147483647L) return fail(c, "too many steps");
    int orbits = c->orbits, nsteps = (int)steps, every = ckpt_every;
    const int opb = implicit(c) ? block : block / group;
    const unsigned grid = (unsigned)((orbits + opb - 1) / opb);
    void* args[] = {(void*)&ckpt_t, (void*)&ckpt_y, (void*)&params, (void*)&cot, &grad_y0, &grad_params,
                    &work, &orbits, &nsteps, &every, &stored};
    hipStream_t s = (hipStream_t)stream;   // null: HIP's default stream
    if (!hip_ok(c, hipModuleLaunchKernel(fn, grid, 1, 1, (unsigned)block, 1, 1, 0, s, args, nullptr),
                "hipModuleLaunchKernel(adjoint)"))
        return 0;
    return 1;
}

}  // namespace

extern "C" {

int sodecl_adjoint_device(sodecl_ctx* c, const double* ckpt_t, const double* ckpt_y, const double* params,
                          const double* cot, double* grad_y0, double* grad_params, double* work, int ckpt_every,
                          void* stream)
{
    if (!c) return 0;
    if (!c->ready) return fail(c, "sodecl_setup has not succeeded");
    if (!ckpt_t || !ckpt_y || !cot || !grad_y0 || (c->npar > 0 && (!params || !grad_params)) || !work)
        return fail(c, "null device pointer");
    if (ckpt_every <= 0) return fail(c, "ckpt_every must be positive");
    return launch_adjoint(c, ckpt_t, ckpt_y, params, cot, grad_y0, grad_params, work, ckpt_every, 0, stream);
}

}  // extern "C"

namespace sodecl_amd {

int adjoint_into_workspace(sodecl_ctx* c, const double* ckpt_t, const double* ckpt_y, const double* params,
                           const double* cot, double* grad_y0, double* grad_params, double* work,
                           size_t work_doubles, int ckpt_every, void* stream)
{
    if (!c) return 0;
    if (!c->ready) return fail(c, "sodecl_setup has not succeeded");
    if (!ckpt_t || !ckpt_y || !cot || !grad_y0 || (c->npar > 0 && (!params || !grad_params)) || !work)
        return fail(c, "null device pointer");
    if (ckpt_every <= 0) return fail(c, "ckpt_every must be positive");
    const size_t need = sodecl_adjoint_workspace_size(c, ckpt_every);
    if (need == 0) return fail(c, "this configuration has no adjoint");
    if (need > work_doubles)
        return fail(c, "adjoint workspace of " + std::to_string(need) + " doubles exceeds the " +
                           std::to_string(work_doubles) +
                           " doubles of the library's scratch array (fewer orbits or a shorter checkpoint "
                           "interval per call)");
    return launch_adjoint(c, ckpt_t, ckpt_y, params, cot, grad_y0, grad_params, work, ckpt_every, 0, stream, true);
}

}  // namespace sodecl_amd

extern "C" {

size_t sodecl_stored_workspace_size(const sodecl_ctx* c)
{
    if (!c || implicit(c)) return 0;
    const long s = total_steps(c);
    if (s <= 0 || s > 2147483647L) return 0;
    return sodecl_adjoint_workspace_size(c, (int)s);
}

int sodecl_run_device_stored(sodecl_ctx* c, double* t0, double* y0, const double* params, double* traj, double* work,
                             void* stream)
{
    if (!c) return 0;
    if (!work) return fail(c, "null workspace");
    int sched = 0, group = 0, block = 0;
    if (!check_adjoint(c, sched, group, block)) return 0;
    if (implicit(c)) return fail(c, "stored stage vectors are available for Euler and RK4 only");
    return run_device_impl(c, t0, y0, params, nullptr, traj, nullptr, nullptr, 0, stream, work);
}

int sodecl_adjoint_device_stored(sodecl_ctx* c, const double* params, const double* cot, double* grad_y0,
                                 double* grad_params, double* work, void* stream)
{
    if (!c) return 0;
    if (!c->ready) return fail(c, "sodecl_setup has not succeeded");
    if (!cot || !grad_y0 || (c->npar > 0 && (!params || !grad_params)) || !work)
        return fail(c, "null device pointer");
    const long s = total_steps(c);
    if (s <= 0 || s > 2147483647L) return fail(c, "no steps to differentiate");
    return launch_adjoint(c, nullptr, nullptr, params, cot, grad_y0, grad_params, work, (int)s, 1, stream);
}

}  // extern "C"
