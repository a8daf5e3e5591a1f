// C ABI of the batched integrator (include/sodecl.h, no reference
// counterpart): orbits of several system sizes and individual step counts in
// one persistent launch (device/batch_class.hip, device/batch_kernel.hip).
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
    std::vector<int> cls, steps, order, items;   // items: four orbit slots each (-1: none), longest first
    int nitems = 0;
    std::vector<long long> yoff, ooff;
    size_t state_size = 0, output_size = 0;
    // device
    std::vector<char> code;
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr, fn_reset = nullptr;
    int *d_items = nullptr, *d_cls = nullptr, *d_steps = nullptr, *d_head = nullptr;
    long long *d_yoff = nullptr, *d_ooff = nullptr;
    unsigned grid = 0;
    size_t lds_bytes = 0;
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
    if (b->module) hipModuleUnload(b->module);
    hipFree(b->d_items);
    hipFree(b->d_cls);
    hipFree(b->d_steps);
    hipFree(b->d_head);
    hipFree(b->d_yoff);
    hipFree(b->d_ooff);
    b->module = nullptr;
    b->fn = b->fn_reset = nullptr;
    b->d_items = b->d_cls = b->d_steps = b->d_head = nullptr;
    b->d_yoff = b->d_ooff = nullptr;
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
    // work items: one orbit, or up to four consecutive orbits of the same
    // class (similar costs) where that class runs 32 or 16 lanes per orbit
    const int nclass = (int)b->class_neq.size();
    std::vector<std::vector<int>> per_class(nclass);
    for (int o : b->order) per_class[b->cls[o]].push_back(o);
    struct Item {
        double cost;
        int o[4];
    };
    std::vector<Item> items;
    for (int c = 0; c < nclass; ++c) {
        const int opw = 64 / batch_lanes_per_orbit(b->class_neq[c]);
        const auto& v = per_class[c];
        for (size_t i = 0; i < v.size(); i += opw) {
            Item it{cost[v[i]], {-1, -1, -1, -1}};
            for (int k = 0; k < opw && i + k < v.size(); ++k) it.o[k] = v[i + k];
            items.push_back(it);
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) { return a.cost > c.cost; });
    b->nitems = (int)items.size();
    b->items.resize(4 * items.size());
    for (size_t i = 0; i < items.size(); ++i)
        for (int k = 0; k < 4; ++k) b->items[4 * i + k] = items[i].o[k];

    if (!bhip(b, hipSetDevice(b->device), "hipSetDevice")) return 0;
    b->program = assemble_batch_program(b->source, batch_key(b), b->features, b->class_neq);
    std::string log;
    if (!compile_program(b->program, device_arch(b->device), b->opts.fp_contract, b->code, log))
        return bfail(b, "Failed to build program executable:\n" + log);
    if (!bhip(b, hipModuleLoadData(&b->module, b->code.data()), "hipModuleLoadData")) return 0;
    if (!bhip(b, hipModuleGetFunction(&b->fn, b->module, "sodecl_batch"), "sodecl_batch")) return 0;
    if (!bhip(b, hipModuleGetFunction(&b->fn_reset, b->module, "sodecl_batch_reset"), "sodecl_batch_reset"))
        return 0;
    // LDS per wavefront: per orbit the evaluation point, four padded solve
    // rows and (more than 8) parameters, times the orbits per wavefront
    // (batch_class.hip); the largest class sets the launch's size
    size_t lds_doubles = 1;
    for (int c = 0; c < nclass; ++c) {
        const long n = b->class_neq[c];
        const long lpo = batch_lanes_per_orbit((int)n);
        const long br = (n + lpo - 1) / lpo;
        const long tlen = br >= 2 ? (n + 15) / 16 * 16 : n;
        const long per_orbit = n + 4 * tlen + (b->npar <= 8 ? 0 : b->npar);
        lds_doubles = std::max(lds_doubles, (size_t)(per_orbit * (64 / lpo)));
    }
    b->lds_bytes = lds_doubles * sizeof(double);
    int per_cu = 0;
    if (!bhip(b, hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn, 64, b->lds_bytes),
              "occupancy"))
        return 0;
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
    long g = (long)std::max(per_cu, 1) * std::max(cus, 1);
    b->grid = (unsigned)std::min<long>(g, b->nitems);
    auto up = [&](void** dst, const void* src, size_t bytes, const char* what) {
        return bhip(b, hipMalloc(dst, bytes), what) &&
               bhip(b, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    if (!up((void**)&b->d_items, b->items.data(), b->items.size() * sizeof(int), "work items")) return 0;
    if (!up((void**)&b->d_cls, b->cls.data(), orbits * sizeof(int), "classes")) return 0;
    if (!up((void**)&b->d_steps, b->steps.data(), orbits * sizeof(int), "steps")) return 0;
    if (!up((void**)&b->d_yoff, b->yoff.data(), orbits * sizeof(long long), "state offsets")) return 0;
    if (!up((void**)&b->d_ooff, b->ooff.data(), orbits * sizeof(long long), "output offsets")) return 0;
    if (!bhip(b, hipMalloc(&b->d_head, sizeof(int)), "queue head")) return 0;
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
    if (!b->fn) return bfail(b, "sodecl_batch_setup has not succeeded");
    if (!y0 || !out || !newton || (b->npar > 0 && !params)) return bfail(b, "null device pointer");
    if (!bhip(b, hipSetDevice(b->device), "hipSetDevice")) return 0;
    hipStream_t s = (hipStream_t)stream;
    void* rargs[] = {&b->d_head};
    if (!bhip(b, hipModuleLaunchKernel(b->fn_reset, 1, 1, 1, 1, 1, 1, 0, s, rargs, nullptr), "reset")) return 0;
    int nitems = b->nitems;
    void* args[] = {&b->d_items, &b->d_cls, &b->d_steps, &b->d_yoff, &b->d_ooff, (void*)&y0, (void*)&params,
                    &out, &newton, &b->d_head, &nitems};
    return bhip(b, hipModuleLaunchKernel(b->fn, b->grid, 1, 1, 64, 1, 1, (unsigned)b->lds_bytes, s, args, nullptr),
                "hipModuleLaunchKernel(batch)")
               ? 1
               : 0;
}

const char* sodecl_batch_last_error(const sodecl_batch* b) { return b ? b->last_error.c_str() : "null batch"; }
const char* sodecl_batch_program_source(const sodecl_batch* b) { return b ? b->program.c_str() : ""; }

}  // extern "C"
