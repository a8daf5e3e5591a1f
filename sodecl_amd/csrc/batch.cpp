// C ABI of the batched integrator (include/sodecl.h, no reference
// counterpart): orbits of several system sizes and individual step counts,
// one kernel per size class with a workgroup per work item, the classes'
// kernels launched side by side on streams of their own
// (device/batch_class.hip, device/batch_kernel.hip).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "sodecl_internal.hpp"

using namespace sodecl_amd;

struct sodecl_batch {
    std::string source;
    PluginFeatures features;
    int solver = 3;
    double dt = 0.0;
    int ksteps = 1, npar = 0;
    std::vector<int> class_neq;
    sodecl_options opts{};
    int device = 0;
    // per-orbit layout
    int orbits = 0;
    std::vector<int> cls, steps, order;
    std::vector<long long> yoff, ooff;
    size_t state_size = 0, output_size = 0;
    // per class: its work items (four orbit slots each, -1: none, longest
    // first; one workgroup each), its kernel, LDS size and stream; classes
    // are launched in order of their total cost, the most expensive first
    struct Class {
        std::vector<int> items;
        int nitems = 0;
        double cost = 0.0;
        hipFunction_t fn = nullptr;
        int* d_items = nullptr;
        size_t lds_bytes = 0;
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<Class> classes;
    std::vector<int> launch_order;
    hipEvent_t fork = nullptr;
    // device
    std::vector<char> code;
    hipModule_t module = nullptr;
    int* d_steps = nullptr;
    long long *d_yoff = nullptr, *d_ooff = nullptr;
    bool ready = false;
    std::string last_error, program;
};

namespace {

int bfail(sodecl_batch* b, const std::string& m)
{
    if (b) b->last_error = m;
    return 0;
}

bool bhip(sodecl_batch* b, hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    bfail(b, std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

void free_batch(sodecl_batch* b)
{
    // work queued on the classes' streams ends before their memory goes
    for (auto& c : b->classes) {
        if (c.stream) {
            hipStreamSynchronize(c.stream);
            hipStreamDestroy(c.stream);
        }
        if (c.done) hipEventDestroy(c.done);
        hipFree(c.d_items);
    }
    b->classes.clear();
    b->launch_order.clear();
    if (b->fork) hipEventDestroy(b->fork);
    b->fork = nullptr;
    if (b->module) hipModuleUnload(b->module);
    hipFree(b->d_steps);
    hipFree(b->d_yoff);
    hipFree(b->d_ooff);
    b->module = nullptr;
    b->d_steps = nullptr;
    b->d_yoff = b->d_ooff = nullptr;
    b->ready = false;
}

ProgramKey batch_key(const sodecl_batch* b)
{
    ProgramKey k;
    k.solver = b->solver;
    k.npar = b->npar;
    k.orbits = b->orbits;
    k.ksteps = b->ksteps;
    k.dt = b->dt;
    k.newton_max_iter = b->opts.newton_max_iter > 0 ? b->opts.newton_max_iter : 1000;
    k.newton_tol = b->opts.newton_tol;
    k.fp_contract = b->opts.fp_contract;
    k.library_sin = b->opts.library_sin;
    k.library_exp = b->opts.library_exp;
    return k;
}

}  // namespace

extern "C" {

int sodecl_batch_create(sodecl_batch** out, const char* system_source, int solver, double dt, int ksteps, int npar,
                        int nclasses, const int* class_neq)
{
    if (!out) return 0;
    auto* b = new sodecl_batch();
    *out = b;
    sodecl_default_options(&b->opts);
    b->source = system_source ? system_source : "";
    b->features = detect_features(b->source);
    b->solver = solver;
    b->dt = dt;
    b->ksteps = ksteps;
    b->npar = npar;
    if (solver != SODECL_IMPLICIT_EULER && solver != SODECL_IMPLICIT_MIDPOINT)
        return bfail(b, "the batched integrator runs the implicit Euler and implicit midpoint solvers");
    if (!(dt > 0) || ksteps <= 0 || npar < 0 || nclasses <= 0 || !class_neq)
        return bfail(b, "invalid dt, ksteps, npar or size classes");
    for (int c = 0; c < nclasses; ++c) {
        if (class_neq[c] <= 0) return bfail(b, "size classes must be positive");
        b->class_neq.push_back(class_neq[c]);
    }
    if (!b->features.component || !b->features.jacobian_band_i)
        return bfail(b, "the batched integrator needs sode_system_i and calc_jacobian_band_i in the system source");
    return 1;
}

void sodecl_batch_destroy(sodecl_batch* b)
{
    if (!b) return;
    free_batch(b);
    delete b;
}

int sodecl_batch_set_options(sodecl_batch* b, const sodecl_options* o)
{
    if (!b || !o) return 0;
    b->opts = *o;
    return 1;
}

int sodecl_batch_choose_device(sodecl_batch* b, uint32_t device)
{
    if (!b) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || (int)device >= n) return bfail(b, "device number out of bounds");
    b->device = (int)device;
    return 1;
}

size_t sodecl_batch_compile_only(sodecl_batch* b, const char* arch)
{
    if (!b) return 0;
    b->program = assemble_batch_program(b->source, batch_key(b), b->features, b->class_neq);
    std::vector<char> code;
    std::string log;
    if (!compile_program(b->program, arch ? arch : "gfx950", b->opts.fp_contract, code, log)) {
        bfail(b, "Failed to build program executable:\n" + log);
        return 0;
    }
    return code.size();
}

int sodecl_batch_setup(sodecl_batch* b, int orbits, const int* orbit_class, const int* orbit_steps)
{
    if (!b) return 0;
    if (orbits <= 0 || !orbit_class || !orbit_steps) return bfail(b, "invalid orbit description");
    free_batch(b);
    b->orbits = orbits;
    b->cls.assign(orbit_class, orbit_class + orbits);
    b->steps.assign(orbit_steps, orbit_steps + orbits);
    b->yoff.resize(orbits);
    b->ooff.resize(orbits);
    size_t ys = 0, os = 0;
    std::vector<double> cost(orbits);
    for (int o = 0; o < orbits; ++o) {
        const int c = b->cls[o];
        if (c < 0 || c >= (int)b->class_neq.size()) return bfail(b, "orbit size class out of range");
        if (b->steps[o] < 0 || b->steps[o] % b->ksteps != 0)
            return bfail(b, "per-orbit step counts must be non-negative multiples of ksteps");
        const int n = b->class_neq[c];
        b->yoff[o] = (long long)ys;
        b->ooff[o] = (long long)os;
        ys += n;
        os += (size_t)(b->steps[o] / b->ksteps + 1) * n;
        cost[o] = (double)n * (double)b->steps[o];
    }
    b->state_size = ys;
    b->output_size = os;
    // longest first: the queue hands the most expensive orbits out first
    b->order.resize(orbits);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    // work items of each class: one orbit, or up to four consecutive orbits
    // (similar costs) where the class runs 32 or 16 lanes per orbit
    const int nclass = (int)b->class_neq.size();
    b->classes.assign(nclass, sodecl_batch::Class{});
    for (int o : b->order) {
        auto& c = b->classes[b->cls[o]];
        const int opw = 64 / batch_lanes_per_orbit(b->class_neq[b->cls[o]]);
        if (c.nitems == 0 || c.items[4 * (c.nitems - 1) + opw - 1] >= 0) {
            c.items.insert(c.items.end(), {-1, -1, -1, -1});
            ++c.nitems;
        }
        int* slot = &c.items[4 * (c.nitems - 1)];
        int k = 0;
        while (slot[k] >= 0) ++k;
        slot[k] = o;
        c.cost += cost[o];
    }
    b->launch_order.resize(nclass);
    std::iota(b->launch_order.begin(), b->launch_order.end(), 0);
    std::stable_sort(b->launch_order.begin(), b->launch_order.end(),
                     [&](int x, int y) { return b->classes[x].cost > b->classes[y].cost; });

    if (!bhip(b, hipSetDevice(b->device), "hipSetDevice")) return 0;
    b->program = assemble_batch_program(b->source, batch_key(b), b->features, b->class_neq);
    std::string log;
    if (!compile_program(b->program, device_arch(b->device), b->opts.fp_contract, b->code, log))
        return bfail(b, "Failed to build program executable:\n" + log);
    if (!bhip(b, hipModuleLoadData(&b->module, b->code.data()), "hipModuleLoadData")) return 0;
    auto up = [&](void** dst, const void* src, size_t bytes, const char* what) {
        return bhip(b, hipMalloc(dst, bytes), what) &&
               bhip(b, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    for (int ci = 0; ci < nclass; ++ci) {
        auto& c = b->classes[ci];
        const std::string name = "sodecl_batch_" + std::to_string(ci);
        if (!bhip(b, hipModuleGetFunction(&c.fn, b->module, name.c_str()), "class kernel")) return 0;
        // LDS per wavefront: per orbit four rows of the class's size (the
        // evaluation point shared with the right-hand side, three band
        // rows) and (more than 8) parameters, times the orbits per
        // wavefront (batch_class.hip)
        const long n = b->class_neq[ci];
        const long lpo = batch_lanes_per_orbit((int)n);
        const long br = (n + lpo - 1) / lpo;
        const long tlen = br >= 2 ? (n + 15) / 16 * 16 : n;
        const long per_orbit = 4 * tlen + (b->npar <= 8 ? 0 : b->npar);
        c.lds_bytes = (size_t)(per_orbit * (64 / lpo)) * sizeof(double);
        if (c.nitems > 0 &&
            !up((void**)&c.d_items, c.items.data(), c.items.size() * sizeof(int), "work items"))
            return 0;
        if (!bhip(b, hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate") ||
            !bhip(b, hipEventCreateWithFlags(&c.done, hipEventDisableTiming), "hipEventCreate"))
            return 0;
    }
    if (!bhip(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate")) return 0;
    if (!up((void**)&b->d_steps, b->steps.data(), orbits * sizeof(int), "steps")) return 0;
    if (!up((void**)&b->d_yoff, b->yoff.data(), orbits * sizeof(long long), "state offsets")) return 0;
    if (!up((void**)&b->d_ooff, b->ooff.data(), orbits * sizeof(long long), "output offsets")) return 0;
    b->ready = true;
    return 1;
}

size_t sodecl_batch_state_size(const sodecl_batch* b) { return b ? b->state_size : 0; }
size_t sodecl_batch_output_size(const sodecl_batch* b) { return b ? b->output_size : 0; }

int sodecl_batch_offsets(const sodecl_batch* b, long long* yoff, long long* ooff)
{
    if (!b || b->yoff.empty()) return 0;
    if (yoff) std::memcpy(yoff, b->yoff.data(), b->yoff.size() * sizeof(long long));
    if (ooff) std::memcpy(ooff, b->ooff.data(), b->ooff.size() * sizeof(long long));
    return 1;
}

int sodecl_batch_run_device(sodecl_batch* b, const double* y0, const double* params, double* out, long long* newton,
                            void* stream)
{
    if (!b) return 0;
    if (!b->ready) return bfail(b, "sodecl_batch_setup has not succeeded");
    if (!y0 || !out || !newton || (b->npar > 0 && !params)) return bfail(b, "null device pointer");
    if (!bhip(b, hipSetDevice(b->device), "hipSetDevice")) return 0;
    hipStream_t s = (hipStream_t)stream;
    // every class's kernel on its own stream after the work queued on the
    // caller's stream, then the caller's stream after all of them (its waits
    // are queued last: a class stream that shares a hardware queue with the
    // caller's would otherwise wait behind them)
    if (!bhip(b, hipEventRecord(b->fork, s), "hipEventRecord")) return 0;
    for (int ci : b->launch_order) {
        auto& c = b->classes[ci];
        if (c.nitems == 0) continue;
        void* args[] = {&c.d_items, &b->d_steps, &b->d_yoff, &b->d_ooff, (void*)&y0, (void*)&params, &out, &newton};
        if (!bhip(b, hipStreamWaitEvent(c.stream, b->fork, 0), "hipStreamWaitEvent") ||
            !bhip(b, hipModuleLaunchKernel(c.fn, (unsigned)c.nitems, 1, 1, 64, 1, 1, (unsigned)c.lds_bytes, c.stream,
                                           args, nullptr),
                  "hipModuleLaunchKernel(batch)") ||
            !bhip(b, hipEventRecord(c.done, c.stream), "hipEventRecord"))
            return 0;
    }
    for (int ci : b->launch_order)
        if (b->classes[ci].nitems > 0 && !bhip(b, hipStreamWaitEvent(s, b->classes[ci].done, 0), "hipStreamWaitEvent"))
            return 0;
    return 1;
}

const char* sodecl_batch_last_error(const sodecl_batch* b) { return b ? b->last_error.c_str() : "null batch"; }
const char* sodecl_batch_program_source(const sodecl_batch* b) { return b ? b->program.c_str() : ""; }

}  // extern "C"
