// C ABI of libsodecl_amd.so (include/sodecl.h): the MI355X counterpart of the
// reference's host manager sodecl::sodeclmgr (src/sodeclmgr.cpp).
//
// Differences in design from the reference's OpenCL host loop
// (sodeclmgr.cpp:921-1035), results unchanged:
//   * one kernel launch runs many ksteps segments and writes every snapshot
//     straight into a device trajectory buffer in the reference's output
//     order [snapshot][equation][orbit]; the host copies it once, instead of a
//     blocking read-back and a host transposition before every launch;
//   * when the trajectory exceeds the device buffer budget the run is cut
//     into chunks whose copies overlap the next chunk's kernel (two buffers,
//     a copy stream);
//   * the program is compiled once per configuration and cached.
// The reference's launch design is kept as SODECL_SCHEDULE_PLAIN, the
// baseline the fast schedules are measured against.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>

#include "host_copy.hpp"
#include "sodecl_internal.hpp"

using sodecl_amd::fail;
using sodecl_amd::hip_ok;

namespace sodecl_amd {

int fail(sodecl_ctx* ctx, const std::string& msg)
{
    if (ctx) ctx->last_error = msg;
    if (ctx && ctx->opts.verbose) std::cerr << "sodecl: " << msg << std::endl;
    return 0;
}

bool hip_ok(sodecl_ctx* ctx, hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    fail(ctx, std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

// LDS bytes per orbit of the pairwise drift's two layouts: the full matrix
// N x (N+1), or the evaluated values only, N x ceil((N-1)/2 + [N even]).
long pairs_full_bytes(const sodecl_ctx* c) { return (long)c->neq * (c->neq + 1) * 8; }
long pairs_half_bytes(const sodecl_ctx* c)
{
    return (long)c->neq * ((c->neq - 1) / 2 + (c->neq % 2 == 0 ? 1 : 0)) * 8;
}

// Layout of the group schedule's pairwise drift for group size G: 0 none
// (per-component entry point), 1 full LDS matrix, 2 half LDS storage, 3 lane
// exchange without storage (G >= N).  Options: -1 never, 1 always (the
// automatic layout, else half storage), 2 half, 3 full, 4 lane exchange
// (when G >= N, else none).  Automatic choice:
//   G >= 2N and the full matrix at most 16 KiB per orbit: the full matrix,
//     whose pair evaluations spread over all G lanes (small launches, where
//     the group size is raised to use the spare lanes);
//   G >= N: the lane exchange;
//   else the full matrix while it is at most 16 KiB per orbit, else none.
// Measured (MI355X): the lane exchange runs N = 64 RK4 (16384 orbits) 1.77x
// faster than the per-component form and N = 32 EM 1.11x faster than the
// full matrix, while LDS matrices above 16 KiB per orbit limit the kernel to
// about one wavefront per SIMD and lose to the per-component form.  Layouts
// 1 and 2 sum each row in j order (the per-component form's bits), layout 3
// in offset order (the same terms, the sum rounded differently).
int pairs_layout(const sodecl_ctx* c, int group)
{
    if (!c->features.pairs || c->opts.pairwise < 0) return 0;
    const bool full_fits = pairs_full_bytes(c) <= 16 * 1024;
    const int automatic = (group >= 2 * c->neq && full_fits) ? 1 : group >= c->neq ? 3 : full_fits ? 1 : 0;
    switch (c->opts.pairwise) {
    case 1: return automatic ? automatic : 2;
    case 2: return 2;
    case 3: return 1;
    case 4: return group >= c->neq ? 3 : 0;
    default: return automatic;
    }
}


// Compute units of the context's device (256 on MI355X when it cannot be read).
int device_cus(const sodecl_ctx* c)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
        cus = 256;
    return cus;
}

ProgramKey make_key(const sodecl_ctx* ctx, int schedule, int group, int block, const std::string& extra)
{
    ProgramKey k;
    k.solver = ctx->solver;
    k.neq = ctx->neq;
    k.npar = ctx->npar;
    k.nnoi = ctx->nnoi;
    k.orbits = ctx->orbits;
    k.ksteps = ctx->ksteps;
    k.dt = ctx->dt;
    k.schedule = schedule;
    k.group = group;
    k.block = block;
    k.newton_max_iter = ctx->opts.newton_max_iter > 0 ? ctx->opts.newton_max_iter : 1000;
    k.newton_tol = ctx->opts.newton_tol;
    k.fp_contract = ctx->opts.fp_contract;
    k.library_sin = ctx->opts.library_sin;
    k.library_exp = ctx->opts.library_exp;
    k.use_pairs = schedule == 2 ? pairs_layout(ctx, group) : 0;
    // At most one wavefront per SIMD: tell the compiler so.  It may then use
    // the whole register file and schedule for instruction-level
    // parallelism, and the dispatcher cannot stack two of the waves on one
    // SIMD while another idles (measured: 2x on 65536-orbit RK4 at block
    // sizes of 128 and 512 lanes).
    const long lanes = (long)ctx->orbits * (schedule == 2 || schedule == 4 ? std::max(group, 1) : 1);
    k.one_wave_per_simd = (lanes + 63) / 64 <= 4L * device_cus(ctx) ? 1 : 0;
    k.extra_defines = extra;
    return k;
}

std::shared_ptr<Program> get_program(sodecl_ctx* ctx, int schedule, int group, int block, const std::string& extra)
{
    std::ostringstream name;
    name << schedule << ":" << group << ":" << block << ":" << extra;
    auto it = ctx->programs.find(name.str());
    if (it != ctx->programs.end()) return it->second;
    ProgramKey key = make_key(ctx, schedule, group, block, extra);
    std::string text = assemble_program(ctx->system_source, key, ctx->features);
    ctx->last_program = text;
    auto prog = std::make_shared<Program>();
    std::string log;
    if (!compile_program(text, device_arch(ctx->device), key.fp_contract, prog->code, log, key.one_wave_per_simd)) {
        // the reference prints the build log on failure (sodeclmgr.cpp:606-619)
        fail(ctx, "Failed to build program executable:\n" + log);
        return nullptr;
    }
    if (!hip_ok(ctx, hipSetDevice(ctx->device), "hipSetDevice")) return nullptr;
    if (!hip_ok(ctx, hipModuleLoadData(&prog->module, prog->code.data()), "hipModuleLoadData")) return nullptr;
    prog->device = ctx->device;
    ctx->programs[name.str()] = prog;
    return prog;
}

std::shared_ptr<Program> find_program(sodecl_ctx* ctx, int schedule, int group, int block, const std::string& extra)
{
    std::ostringstream name;
    name << schedule << ":" << group << ":" << block << ":" << extra;
    auto it = ctx->programs.find(name.str());
    return it != ctx->programs.end() ? it->second : nullptr;
}

}  // namespace sodecl_amd

using namespace sodecl_amd;

namespace {

bool read_text(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

int init_ctx(sodecl_ctx* c, int solver, double dt, double tspan, int ksteps, int neq, int npar, int nnoi,
             int orbits, int output_type)
{
    c->solver = solver;
    c->dt = dt;
    c->tspan = tspan;
    c->ksteps = ksteps;
    c->neq = neq;
    c->npar = npar;
    c->nnoi = nnoi;
    c->orbits = orbits;
    c->output_type = output_type;
    sodecl_default_options(&c->opts);
    if (solver < 0 || solver > 4) return fail(c, "Unknown SODE solver selection.");
    if (neq <= 0 || npar < 0 || nnoi < 0 || orbits <= 0 || ksteps <= 0)
        return fail(c, "invalid sizes (neq, npar, nnoi, orbits, ksteps)");
    if (!(dt > 0) || !(tspan >= 0)) return fail(c, "invalid dt or tspan");
    if (output_type < 0 || output_type > 2) return fail(c, "invalid output type");
    // Per-orbit RNG counters for SDE runs: random, like the reference's
    // m_rcounter (sodeclmgr.cpp:48-59), but non-negative whole numbers below
    // 1e11 so the counter word is well defined.
    if (nnoi > 0) {
        std::random_device rd;
        std::mt19937_64 gen(((uint64_t)rd() << 32) ^ rd());
        std::uniform_int_distribution<long long> unif(0, 100000000000LL);
        c->counters.resize(orbits);
        for (auto& v : c->counters) v = (double)unif(gen);
    }
    return 1;
}

void drop_graph(sodecl_ctx* c)
{
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_key.clear();
}

void free_device(sodecl_ctx* c)
{
    drop_graph(c);
    for (auto& kv : c->programs)
        if (kv.second->module) hipModuleUnload(kv.second->module);
    c->programs.clear();
    hipFree(c->d_t);
    hipFree(c->d_y);
    hipFree(c->d_params);
    hipFree(c->d_counters);
    hipFree(c->d_newton);
    for (int b = 0; b < sodecl_ctx::kTrajPipe; ++b) {
        hipFree(c->d_traj[b]);
        hipHostFree(c->h_stage[b]);
        c->d_traj[b] = nullptr;
        c->h_stage[b] = nullptr;
    }
    if (c->ev_start) hipEventDestroy(c->ev_start);
    if (c->ev_stop) hipEventDestroy(c->ev_stop);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->copy_stream) hipStreamDestroy(c->copy_stream);
    c->d_t = c->d_y = c->d_params = c->d_counters = nullptr;
    c->d_newton = nullptr;
    c->ev_start = c->ev_stop = nullptr;
    c->stream = c->copy_stream = nullptr;
    c->traj_capacity = c->stage_capacity = 0;
    c->ready = false;
}

// Launch configuration for this context.
struct LaunchCfg {
    int schedule;   // 1 lane, 2 group, 4 team
    int group;
    int block;
};

const char* schedule_name(int s) { return s == 2 ? "group" : (s == 4 ? "team" : "lane"); }

bool team_possible(const sodecl_ctx* c)
{
    return c->features.team && (c->solver == SODECL_EULER || c->solver == SODECL_RUNGE_KUTTA);
}


bool group_possible(const sodecl_ctx* c)
{
    const bool implicit = c->solver == SODECL_IMPLICIT_EULER || c->solver == SODECL_IMPLICIT_MIDPOINT;
    if (implicit) return false;
    if (!c->features.component) return false;
    if (c->solver == SODECL_STOCHASTIC_EULER && !c->features.stoch_component) return false;
    return true;
}

int default_group(int neq)
{
    int g = 1;
    while (g < neq && g < 64) g <<= 1;
    return std::max(g, 4);
}

// The pairwise drift keeps its values per orbit in LDS.  Largest block (a
// power-of-two number of orbits) whose storage fits in 48 KiB; a large bound
// when the group schedule does not use the pairwise form.
int pairs_block_cap(const sodecl_ctx* c, int group)
{
    const int layout = pairs_layout(c, group);
    if (layout != 1 && layout != 2) return 1 << 20;
    const long per_orbit = layout == 1 ? pairs_full_bytes(c) : pairs_half_bytes(c);
    const long opb = std::max(1L, (48L * 1024) / per_orbit);
    long p2 = 1;
    while (p2 * 2 <= opb) p2 *= 2;
    return (int)(p2 * group);
}

LaunchCfg choose_config(const sodecl_ctx* c)
{
    LaunchCfg cfg{1, 64, 128};
    // a timed choice made by the autotuner overrides the heuristic
    if (c->opts.schedule == SODECL_SCHEDULE_AUTO && c->tuned_schedule)
        return LaunchCfg{c->tuned_schedule, c->tuned_group, c->tuned_block};
    int sched = c->opts.schedule;
    if (sched == SODECL_SCHEDULE_AUTO) {
        sched = (group_possible(c) && c->neq >= 8) ? 2 : 1;
    }
    if (sched == SODECL_SCHEDULE_GROUP && !group_possible(c)) sched = 1;
    if (sched == SODECL_SCHEDULE_TEAM && !team_possible(c)) sched = 1;
    if (sched == SODECL_SCHEDULE_PLAIN) sched = 1;
    cfg.schedule = sched;
    if (sched == 4) {
        int g = c->opts.group_size > 0 ? c->opts.group_size : 2;
        g = std::min(std::max(g, 2), 64);
        int p2 = 1;
        while (p2 < g) p2 <<= 1;
        cfg.group = p2;
        int b = c->opts.block_size > 0 ? c->opts.block_size : 128;
        b = std::max(b, cfg.group);
        b = (b / cfg.group) * cfg.group;
        cfg.block = std::min(b, 1024);
    } else if (sched == 2) {
        int g = c->opts.group_size > 0 ? c->opts.group_size : default_group(c->neq);
        if (c->opts.group_size <= 0 && c->features.pairs && c->opts.pairwise >= 0) {
            // a launch of less than one wavefront per SIMD: more lanes per
            // orbit share its pair evaluations (up to a wavefront per orbit)
            const long one_per_simd = 64L * 4 * device_cus(c);
            while (g < 64 && (long)c->orbits * g * 2 <= one_per_simd) g *= 2;
        }
        g = std::min(std::max(g, 1), 64);
        int p2 = 1;
        while (p2 < g) p2 <<= 1;
        cfg.group = p2;
        int b = c->opts.block_size > 0 ? c->opts.block_size : 256;
        if (c->opts.block_size <= 0) {
            b = std::min(b, pairs_block_cap(c, cfg.group));
            // small launches: smaller blocks (down to one wavefront) so that
            // every compute unit gets one (1024 ten-equation orbits: 256
            // blocks of 64 lanes ran 10% faster than 64 blocks of 256)
            const long lanes = (long)c->orbits * cfg.group;
            while (b > std::max(cfg.group, 64) && (lanes + b - 1) / b < device_cus(c)) b /= 2;
        }
        b = std::max(b, cfg.group);
        b = (b / cfg.group) * cfg.group;
        cfg.block = std::min(b, 1024);
    } else {
        // one wavefront per SIMD or fewer: any block size spreads evenly under
        // the occupancy hint; otherwise blocks of four wavefronts
        const bool few = (long)c->orbits <= 64L * 4 * device_cus(c);
        int b = c->opts.block_size > 0 ? c->opts.block_size
                                       : (c->local_group_size > 0 ? c->local_group_size : (few ? 128 : 256));
        cfg.block = std::min(std::max(b, 1), 1024);
        cfg.group = 1;
    }
    return cfg;
}

// Enqueues one launch of the chosen schedule: segments [s0, s0 + nseg) of the
// trajectory buffer traj (snapshot rows), final snapshot when final_snap.
int enqueue(sodecl_ctx* c, const LaunchCfg& cfg, double* t, double* y, const double* p, double* cnt, double* traj,
            long long* newton, int s0, int nseg, int final_snap, hipStream_t stream, double* ckpt_t = nullptr,
            double* ckpt_y = nullptr, int ckpt_every = 0, long long gstep0 = 0, double* stage_ws = nullptr,
            long long ws_steps = 0, int norbits = -1)
{
    auto prog = get_program(c, cfg.schedule, cfg.group, cfg.block);
    if (!prog) return 0;
    const char* kname = cfg.schedule == 2 ? "sodecl_group" : (cfg.schedule == 4 ? "sodecl_team" : "sodecl_lane");
    hipFunction_t fn = prog->get(kname);
    if (!fn) return fail(c, std::string("kernel ") + kname + " missing from the program");
    // the first norbits orbits only (the rest are the host share's)
    int orbits = norbits >= 0 ? norbits : c->orbits;
    unsigned grid;
    if (cfg.schedule == 2 || cfg.schedule == 4) {
        const int opb = cfg.block / cfg.group;
        grid = (unsigned)((orbits + opb - 1) / opb);
    } else {
        grid = (unsigned)((orbits + cfg.block - 1) / cfg.block);
    }
    if (!ckpt_y || ckpt_every <= 0) {
        ckpt_y = nullptr;
        ckpt_t = nullptr;
        ckpt_every = 1;
    }
    void* args[] = {&t, &y, (void*)&p, &cnt, &traj, &newton, &orbits, &s0, &nseg, &final_snap,
                    &ckpt_t, &ckpt_y, &ckpt_every, &gstep0, &stage_ws, &ws_steps};
    if (!hip_ok(c, hipModuleLaunchKernel(fn, grid, 1, 1, (unsigned)cfg.block, 1, 1, 0, stream, args, nullptr),
                "hipModuleLaunchKernel"))
        return 0;
    c->used_schedule = cfg.schedule;
    c->used_group = cfg.group;
    c->used_block = cfg.block;
    return 1;
}

int ensure_traj(sodecl_ctx* c, size_t doubles, int nbuf)
{
    if (c->traj_capacity >= doubles && c->d_traj[nbuf - 1]) return 1;
    for (int b = 0; b < sodecl_ctx::kTrajPipe; ++b) {
        hipFree(c->d_traj[b]);
        c->d_traj[b] = nullptr;
    }
    c->traj_capacity = 0;
    for (int b = 0; b < nbuf; ++b)
        if (!hip_ok(c, hipMalloc(&c->d_traj[b], doubles * sizeof(double)), "hipMalloc(trajectory)")) return 0;
    c->traj_capacity = doubles;
    return 1;
}

int ensure_stage(sodecl_ctx* c, size_t doubles, int nbuf)
{
    if (c->stage_capacity >= doubles && c->h_stage[nbuf - 1]) return 1;
    for (int b = 0; b < sodecl_ctx::kTrajPipe; ++b) {
        hipHostFree(c->h_stage[b]);
        c->h_stage[b] = nullptr;
    }
    c->stage_capacity = 0;
    for (int b = 0; b < nbuf; ++b)
        if (!hip_ok(c, hipHostMalloc((void**)&c->h_stage[b], doubles * sizeof(double), hipHostMallocDefault),
                    "hipHostMalloc"))
            return 0;
    c->stage_capacity = doubles;
    return 1;
}

// Sink for snapshots: host array, file, or nothing.
struct Sink {
    sodecl_ctx* c;
    std::ofstream file;
    double* dst = nullptr;   // ARRAY destination
    bool open(size_t total)
    {
        if (c->output_type == SODECL_OUTPUT_FILE) {
            // sodeclmgr.cpp:926-927: appended raw doubles
            file.open("sodecloutput.bin", std::ios::binary | std::ios::app | std::ios::out);
            return (bool)file;
        }
        if (c->output_type == SODECL_OUTPUT_ARRAY) {
            if (c->ext_output) {
                if (c->ext_used + total > c->ext_capacity) return false;
                dst = c->ext_output + c->ext_used;
                c->ext_used += total;
            } else {
                size_t old = c->output.size();
                c->output.resize(old + total);
                dst = c->output.data() + old;
            }
        }
        return true;
    }
    void put(const double* host, size_t n)
    {
        if (c->output_type == SODECL_OUTPUT_FILE) {
            file.write((const char*)host, (std::streamsize)(n * sizeof(double)));
        } else if (dst) {
            if (host != dst) parallel_copy(dst, host, n * sizeof(double));
            dst += n;
        }
    }
};

// The reference's launch design (sodeclmgr.cpp:951-1015): blocking read-back
// of the AoS state before each one-segment launch, host transposition into
// [equation][orbit] order.
int run_plain(sodecl_ctx* c, long launches)
{
    LaunchCfg cfg{1, 1, c->local_group_size > 0 ? std::min(c->local_group_size, 1024) : 128};
    if (c->opts.block_size > 0) cfg.block = std::min(c->opts.block_size, 1024);
    const size_t slab = (size_t)c->neq * c->orbits;
    Sink sink{c};
    if (!sink.open(slab * (size_t)(launches + 1))) return fail(c, "output buffer too small");
    std::vector<double> aos(slab), snap(slab);
    auto take = [&]() -> bool {
        if (!hip_ok(c, hipMemcpyAsync(aos.data(), c->d_y, slab * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                    "hipMemcpy(state)"))
            return false;
        if (!hip_ok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return false;
        if (c->output_type == SODECL_OUTPUT_NONE) return true;
        for (int e = 0; e < c->neq; ++e)
            for (int o = 0; o < c->orbits; ++o) snap[(size_t)e * c->orbits + o] = aos[(size_t)o * c->neq + e];
        sink.put(snap.data(), slab);
        return true;
    };
    hipEventRecord(c->ev_start, c->stream);
    for (long j = 0; j < launches; ++j) {
        if (!take()) return 0;
        if (!enqueue(c, cfg, c->d_t, c->d_y, c->d_params, c->d_counters, nullptr, c->d_newton, 0, 1, 0, c->stream))
            return 0;
    }
    if (!take()) return 0;
    hipEventRecord(c->ev_stop, c->stream);
    hipEventSynchronize(c->ev_stop);
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    c->last_kernel_ms = ms;
    c->used_schedule = SODECL_SCHEDULE_PLAIN;
    return 1;
}

size_t traj_budget(const sodecl_ctx* c)
{
    return c->opts.traj_budget_bytes ? c->opts.traj_budget_bytes : ((size_t)8 << 30);
}

// Orbits of a run the host share integrates: the last floor(orbits *
// host_share), when the host library is loaded and at least one orbit is
// left to the GPU.
int host_orbits(const sodecl_ctx* c)
{
    if (!c->host_fn || !(c->opts.host_share > 0) || c->hs_p.size() != (size_t)c->orbits * c->npar) return 0;
    const int n = (int)((double)c->orbits * c->opts.host_share);
    return n < c->orbits ? n : c->orbits - 1;
}

// One launch on the GPU for orbits [0, n_gpu) while the host share
// integrates [n_gpu, orbits) into the same output rows; the host orbits'
// final state (t, y, counters, Newton counts) goes back to the device
// buffers so that a following run continues from it.
int run_shared(sodecl_ctx* c, long launches, const LaunchCfg& cfg, Sink& sink, bool to_array_direct)
{
    const int n_host = host_orbits(c), n_gpu = c->orbits - n_host, n = c->neq;
    const size_t slab = (size_t)n * c->orbits, nsnap = (size_t)launches + 1;
    const bool want = c->output_type != SODECL_OUTPUT_NONE;
    double* dst = want ? (to_array_direct ? sink.dst : c->h_stage[0]) : nullptr;
    const size_t first = (size_t)n_gpu;
    auto rows = [&](double* host, const double* dev, size_t width, hipMemcpyKind kind, hipStream_t st) {
        const size_t off = first * width, cnt = (size_t)n_host * width;
        return kind == hipMemcpyDeviceToHost
                   ? hipMemcpyAsync(host + off, dev + off, cnt * sizeof(double), kind, st)
                   : hipMemcpyAsync((double*)dev + off, host + off, cnt * sizeof(double), kind, st);
    };
    c->hs_t.resize(c->orbits);
    c->hs_y.resize(slab);
    c->hs_newton.assign(c->orbits, 0);
    if (c->d_counters) c->hs_cnt.resize(c->orbits);
    // the host orbits' current state
    if (!hip_ok(c, rows(c->hs_t.data(), c->d_t, 1, hipMemcpyDeviceToHost, c->stream), "hipMemcpy(host share)") ||
        !hip_ok(c, rows(c->hs_y.data(), c->d_y, (size_t)n, hipMemcpyDeviceToHost, c->stream),
                "hipMemcpy(host share)") ||
        (c->d_counters &&
         !hip_ok(c, rows(c->hs_cnt.data(), c->d_counters, 1, hipMemcpyDeviceToHost, c->stream),
                 "hipMemcpy(host share)")) ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return 0;
    hipEventRecord(c->ev_start, c->stream);
    if (!enqueue(c, cfg, c->d_t, c->d_y, c->d_params, c->d_counters, want ? c->d_traj[0] : nullptr, c->d_newton, 0,
                 (int)launches, 1, c->stream, nullptr, nullptr, 0, 0, nullptr, 0, n_gpu))
        return 0;
    hipEventRecord(c->ev_stop, c->stream);
    // the host's orbits while the kernel runs
    const auto h0 = std::chrono::steady_clock::now();
    const int threads = host_threads();
    c->host_fn(c->hs_t.data(), c->hs_y.data(), c->hs_p.data(), c->d_counters ? c->hs_cnt.data() : nullptr, dst,
               c->hs_newton.data(), c->orbits, n_gpu, c->orbits, 0, (int)launches, 1, 0, threads);
    c->last_host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    c->last_host_orbits = n_host;
    // the GPU's columns [0, n_gpu) of every snapshot row, then the host
    // orbits' state back to the device
    if (want && !hip_ok(c, hipMemcpy2DAsync(dst, slab / n * sizeof(double), c->d_traj[0], (size_t)n_gpu * sizeof(double),
                                            (size_t)n_gpu * sizeof(double), nsnap * n, hipMemcpyDeviceToHost,
                                            c->stream),
                         "hipMemcpy2D(trajectory)"))
        return 0;
    std::vector<long long> nu(c->orbits, 0);
    if (!hip_ok(c, hipMemcpyAsync(nu.data(), c->d_newton, c->orbits * sizeof(long long), hipMemcpyDeviceToHost,
                                  c->stream),
                "hipMemcpy(newton)") ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "kernel"))
        return 0;
    for (int o = n_gpu; o < c->orbits; ++o) nu[o] += c->hs_newton[o];
    if (!hip_ok(c, rows(c->hs_t.data(), c->d_t, 1, hipMemcpyHostToDevice, c->stream), "hipMemcpy(host share)") ||
        !hip_ok(c, rows(c->hs_y.data(), c->d_y, (size_t)n, hipMemcpyHostToDevice, c->stream),
                "hipMemcpy(host share)") ||
        (c->d_counters &&
         !hip_ok(c, rows(c->hs_cnt.data(), c->d_counters, 1, hipMemcpyHostToDevice, c->stream),
                 "hipMemcpy(host share)")) ||
        !hip_ok(c, hipMemcpyAsync(c->d_newton, nu.data(), c->orbits * sizeof(long long), hipMemcpyHostToDevice,
                                  c->stream),
                "hipMemcpy(newton)") ||
        !hip_ok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return 0;
    if (want) {
        if (to_array_direct)
            sink.dst += nsnap * slab;
        else
            sink.put(c->h_stage[0], nsnap * slab);
    }
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    c->last_kernel_ms = ms;
    return 1;
}

// Fast path: many segments per launch, snapshots written by the kernel.
int run_fast(sodecl_ctx* c, long launches, const LaunchCfg& cfg)
{
    const size_t slab = (size_t)c->neq * c->orbits;
    const size_t nsnap = (size_t)launches + 1;
    const bool want = c->output_type != SODECL_OUTPUT_NONE;
    Sink sink{c};
    if (!sink.open(want ? slab * nsnap : 0)) return fail(c, "output buffer too small");
    if (!want) {
        // no output: integrate in as few launches as the options allow
        long per = c->opts.segments_per_launch > 0 ? c->opts.segments_per_launch : launches;
        if (per >= launches && host_orbits(c) > 0) return run_shared(c, launches, cfg, sink, false);
        hipEventRecord(c->ev_start, c->stream);
        for (long s = 0; s < launches; s += per) {
            int n = (int)std::min(per, launches - s);
            if (!enqueue(c, cfg, c->d_t, c->d_y, c->d_params, c->d_counters, nullptr, c->d_newton, 0, n, 0,
                         c->stream))
                return 0;
        }
        hipEventRecord(c->ev_stop, c->stream);
        if (!hip_ok(c, hipEventSynchronize(c->ev_stop), "kernel")) return 0;
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
        c->last_kernel_ms = ms;
        return 1;
    }
    // segments per chunk within the buffer budget (kTrajPipe buffers when
    // chunked)
    const int pipe = sodecl_ctx::kTrajPipe;
    long per = c->opts.segments_per_launch > 0 ? c->opts.segments_per_launch : launches;
    size_t budget_rows = std::max<size_t>(2, traj_budget(c) / (slab * sizeof(double)));
    if ((size_t)per + 1 > budget_rows) per = (long)(budget_rows / pipe) - 1;
    // A trajectory of more than 64 MiB for host memory streams in chunks of
    // about 16 MiB: chunk k integrates while earlier chunks cross PCIe into
    // pinned memory and are copied into the caller's array by the host
    // threads (one copy from device into pageable memory measured about
    // 8.5 GB/s).  Three buffers let each stage run a chunk ahead of the
    // next one's jitter.  A host share needs the single launch.
    const size_t snap_bytes = slab * sizeof(double);
    if (host_orbits(c) == 0 && snap_bytes * (size_t)(launches + 1) > ((size_t)64 << 20))
        per = std::min<long>(per, std::max<long>(1, (long)(((size_t)16 << 20) / snap_bytes)));
    if (per < 1) per = 1;
    const bool single = per >= launches;
    const size_t rows = (size_t)std::min<long>(per, launches) + 1;
    if (!ensure_traj(c, rows * slab, single ? 1 : pipe)) return 0;
    const bool to_array_direct = c->output_type == SODECL_OUTPUT_ARRAY && single;
    if (!to_array_direct && !ensure_stage(c, rows * slab, single ? 1 : pipe)) return 0;

    if (single && host_orbits(c) > 0) return run_shared(c, launches, cfg, sink, to_array_direct);
    hipEventRecord(c->ev_start, c->stream);
    if (single) {
        if (!enqueue(c, cfg, c->d_t, c->d_y, c->d_params, c->d_counters, c->d_traj[0], c->d_newton, 0,
                     (int)launches, 1, c->stream))
            return 0;
        hipEventRecord(c->ev_stop, c->stream);
        double* host = to_array_direct ? sink.dst : c->h_stage[0];
        if (!hip_ok(c, hipMemcpyAsync(host, c->d_traj[0], nsnap * slab * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream),
                    "hipMemcpy(trajectory)"))
            return 0;
        if (!hip_ok(c, hipStreamSynchronize(c->stream), "kernel")) return 0;
        if (to_array_direct)
            sink.dst += nsnap * slab;
        else
            sink.put(host, nsnap * slab);
    } else {
        // Chunk k runs on the compute stream into buffer k % pipe while the
        // copy stream moves the chunks before it to pinned host memory and
        // the host hands chunk k - pipe to the sink.
        hipEvent_t kdone[sodecl_ctx::kTrajPipe] = {}, cdone[sodecl_ctx::kTrajPipe] = {};
        bool ok = true;
        for (int q = 0; ok && q < pipe; ++q)
            ok = hip_ok(c, hipEventCreateWithFlags(&kdone[q], hipEventDisableTiming), "hipEventCreate") &&
                 hip_ok(c, hipEventCreateWithFlags(&cdone[q], hipEventDisableTiming), "hipEventCreate");
        size_t rows_of[sodecl_ctx::kTrajPipe] = {};
        long k = 0;
        for (long s = 0; ok && s < launches; s += per, ++k) {
            const int b = (int)(k % pipe);
            const int n = (int)std::min(per, launches - s);
            const bool last = s + n >= launches;
            if (k >= pipe && !hip_ok(c, hipStreamWaitEvent(c->stream, cdone[b], 0), "hipStreamWaitEvent")) {
                ok = false;
                break;
            }
            ok = enqueue(c, cfg, c->d_t, c->d_y, c->d_params, c->d_counters, c->d_traj[b], c->d_newton, 0, n,
                         last ? 1 : 0, c->stream, nullptr, nullptr, 0, (long long)s * c->ksteps) &&
                 hip_ok(c, hipEventRecord(kdone[b], c->stream), "hipEventRecord");
            if (!ok) break;
            if (k >= pipe) {
                ok = hip_ok(c, hipEventSynchronize(cdone[b]), "copy");
                if (!ok) break;
                sink.put(c->h_stage[b], rows_of[b] * slab);
            }
            rows_of[b] = (size_t)n + (last ? 1 : 0);
            ok = hip_ok(c, hipStreamWaitEvent(c->copy_stream, kdone[b], 0), "hipStreamWaitEvent") &&
                 hip_ok(c, hipMemcpyAsync(c->h_stage[b], c->d_traj[b], rows_of[b] * slab * sizeof(double),
                                          hipMemcpyDeviceToHost, c->copy_stream),
                        "hipMemcpy(trajectory)") &&
                 hip_ok(c, hipEventRecord(cdone[b], c->copy_stream), "hipEventRecord");
            if (!ok) break;
        }
        hipEventRecord(c->ev_stop, c->stream);
        // the last chunks, oldest first
        for (long q = std::max(0L, k - pipe); ok && q < k; ++q) {
            const int b = (int)(q % pipe);
            ok = hip_ok(c, hipEventSynchronize(cdone[b]), "copy");
            if (ok) sink.put(c->h_stage[b], rows_of[b] * slab);
        }
        if (ok) ok = hip_ok(c, hipStreamSynchronize(c->stream), "kernel");
        if (!ok) {
            // the first error is kept; the copies already queued into the
            // staging buffers must end before they can be reused
            (void)hipStreamSynchronize(c->copy_stream);
            (void)hipStreamSynchronize(c->stream);
        }
        for (int q = 0; q < pipe; ++q) {
            if (kdone[q]) hipEventDestroy(kdone[q]);
            if (cdone[q]) hipEventDestroy(cdone[q]);
        }
        if (!ok) return 0;
    }
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    c->last_kernel_ms = ms;
    return 1;
}

// Candidate launch configurations for the autotuner: the lane schedule at
// three block sizes; when the plugin allows it, the group schedule at every
// power-of-two group size from the system size's own power of two down to a
// quarter of it, at two block sizes (capped by the pairwise drift's LDS);
// and when the plugin has a team entry point, teams of two at three block
// sizes.
std::vector<LaunchCfg> tune_candidates(const sodecl_ctx* c)
{
    std::vector<LaunchCfg> out;
    const bool lane_ok = c->neq <= 64;   // larger states spill to scratch memory
    if (lane_ok)
        for (int b : {64, 128, 256}) out.push_back(LaunchCfg{1, 1, b});
    if (group_possible(c)) {
        int g0 = default_group(c->neq);
        for (int g = g0; g >= std::max(2, g0 / 4); g >>= 1) {
            const int cap = pairs_block_cap(c, g);
            bool any = false;
            for (int b : {64, 256})
                if (b >= g && b <= cap) {
                    out.push_back(LaunchCfg{2, g, b});
                    any = true;
                }
            if (!any) out.push_back(LaunchCfg{2, g, std::max(g, std::min(cap, 256))});
        }
    }
    if (team_possible(c))
        for (int b : {64, 128, 256}) out.push_back(LaunchCfg{4, 2, b});
    if (out.empty()) out.push_back(LaunchCfg{1, 1, 128});
    return out;
}

std::string tune_cache_path(const sodecl_ctx* c)
{
    // keyed by everything that selects the programs (text of one candidate)
    ProgramKey k = sodecl_amd::make_key(c, 0, 0, 0, std::string());
    // the pairwise option selects the group candidates' LDS layouts, which
    // the schedule-0 program text does not show
    std::string text = assemble_program(c->system_source, k, c->features) + device_arch(c->device) +
                       ":pairwise=" + std::to_string(c->opts.pairwise);
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : text) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    char name[64];
    std::snprintf(name, sizeof name, "tune_%016llx.txt", (unsigned long long)h);
    const char* dir = std::getenv("SODECL_AMD_CACHE_DIR");
    std::string d = dir ? dir : std::string();
    if (d.empty()) {
        // next to the code-object cache
        Dl_info info;
        if (dladdr((void*)&tune_candidates, &info) && info.dli_fname) {
            std::string p(info.dli_fname);
            d = p.substr(0, p.rfind('/')) + "/_jit_cache";
        } else {
            d = "/tmp/sodecl_amd_jit_cache";
        }
    }
    return d + "/" + name;
}

// Times every candidate on one launch of one ksteps segment over all orbits
// (on scratch copies of the state, so the context's state is untouched) and
// keeps the fastest.  The choice is cached on disk per configuration.
int autotune(sodecl_ctx* c)
{
    const std::string path = tune_cache_path(c);
    {
        std::ifstream f(path);
        int s = 0, g = 0, b = 0;
        if (f >> s >> g >> b && (s == 1 || s == 2 || s == 4) && b > 0 && b <= 1024 && g >= 1 && g <= 64) {
            // a cached group block never exceeds what the pairwise layout's
            // LDS allows now
            if (s == 2) b = std::max(g, std::min(b, pairs_block_cap(c, g)) / g * g);
            c->tuned_schedule = s;
            c->tuned_group = g;
            c->tuned_block = b;
            c->tune_report = "cached";
            return 1;
        }
    }
    const size_t n = (size_t)c->orbits;
    double *t = nullptr, *y = nullptr, *cnt = nullptr;
    auto release = [&]() {
        hipFree(t);
        hipFree(y);
        hipFree(cnt);
    };
    if (!hip_ok(c, hipMalloc(&t, n * sizeof(double)), "hipMalloc(tune)") ||
        !hip_ok(c, hipMalloc(&y, n * c->neq * sizeof(double)), "hipMalloc(tune)") ||
        (c->d_counters && !hip_ok(c, hipMalloc(&cnt, n * sizeof(double)), "hipMalloc(tune)"))) {
        release();
        return 0;
    }
    std::ostringstream rep;
    float best = 1e30f;
    LaunchCfg best_cfg{0, 0, 0};
    for (const LaunchCfg& cfg : tune_candidates(c)) {
        if (!get_program(c, cfg.schedule, cfg.group, cfg.block)) {
            c->last_error.clear();
            continue;
        }
        float ms_best = 1e30f;
        for (int rep_i = 0; rep_i < 2; ++rep_i) {   // the first launch also loads the code object
            if (!hip_ok(c, hipMemcpyAsync(t, c->d_t, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                        "hipMemcpy(tune)") ||
                !hip_ok(c, hipMemcpyAsync(y, c->d_y, n * c->neq * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                        "hipMemcpy(tune)") ||
                (cnt && !hip_ok(c, hipMemcpyAsync(cnt, c->d_counters, n * sizeof(double), hipMemcpyDeviceToDevice,
                                                  c->stream),
                                "hipMemcpy(tune)"))) {
                release();
                return 0;
            }
            hipEventRecord(c->ev_start, c->stream);
            if (!enqueue(c, cfg, t, y, c->d_params, cnt, nullptr, nullptr, 0, 1, 0, c->stream)) break;
            hipEventRecord(c->ev_stop, c->stream);
            if (!hip_ok(c, hipEventSynchronize(c->ev_stop), "autotune")) break;
            float ms = 0;
            hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
            ms_best = std::min(ms_best, ms);
        }
        rep << schedule_name(cfg.schedule) << " G=" << cfg.group << " B=" << cfg.block << ": "
            << ms_best << " ms\n";
        if (ms_best < best) {
            best = ms_best;
            best_cfg = cfg;
        }
    }
    release();
    if (!best_cfg.schedule) return fail(c, "autotune: no candidate ran");
    c->tuned_schedule = best_cfg.schedule;
    c->tuned_group = best_cfg.group;
    c->tuned_block = best_cfg.block;
    c->tune_report = rep.str();
    std::ofstream f(path);
    if (f) f << best_cfg.schedule << " " << best_cfg.group << " " << best_cfg.block << "\n";
    return 1;
}

}  // namespace

namespace sodecl_amd {

int run_device_impl(sodecl_ctx* c, double* t0, double* y0, const double* params, double* counters, double* traj,
                    double* ckpt_t, double* ckpt_y, int ckpt_every, void* stream, double* stage_ws)
{
    if (!c) return 0;
    if (!c->ready) return fail(c, "sodecl_setup has not succeeded");
    if (!t0 || !y0 || (c->npar > 0 && !params)) return fail(c, "null device pointer");
    if (c->solver == SODECL_STOCHASTIC_EULER && !counters) return fail(c, "the stochastic solver needs counters");
    if (ckpt_y && (!ckpt_t || ckpt_every <= 0)) return fail(c, "checkpoints need ckpt_t and ckpt_every > 0");
    if (!hip_ok(c, hipSetDevice(c->device), "hipSetDevice")) return 0;
    const long launches = num_launches(c->tspan, c->dt, c->ksteps);
    if (launches < 0) return fail(c, "too many kernel launches");
    // a null stream is HIP's default stream, ordered with the caller's other
    // default-stream work (PyTorch's current stream, typically)
    hipStream_t s = (hipStream_t)stream;
    LaunchCfg cfg = choose_config(c);
    long per = c->opts.segments_per_launch > 0 ? c->opts.segments_per_launch : launches;
    if (stage_ws && cfg.schedule != SODECL_SCHEDULE_GROUP)
        return fail(c, "stored stage vectors need the group schedule");
    if (launches == 0) {
        if (traj) return enqueue(c, cfg, t0, y0, params, counters, traj, nullptr, 0, 0, 1, s);
        return 1;
    }
    const long long ws_steps = (long long)launches * c->ksteps;
    auto enqueue_all = [&](hipStream_t st) -> int {
        for (long s0 = 0; s0 < launches; s0 += per) {
            int n = (int)std::min(per, launches - s0);
            bool last = s0 + n >= launches;
            if (!enqueue(c, cfg, t0, y0, params, counters, traj, nullptr, (int)s0, n, last ? 1 : 0, st, ckpt_t,
                         ckpt_y, ckpt_every, (long long)s0 * c->ksteps, stage_ws, ws_steps))
                return 0;
        }
        return 1;
    };
    if (!c->opts.use_graph || launches <= per) return enqueue_all(s);
    // Several launches (segments_per_launch set): record the sequence once
    // as a hipGraph on the context's stream and replay it on the caller's
    // stream; the recording is reused while pointers and shape stay the same.
    std::ostringstream key;
    key << (void*)t0 << (void*)y0 << (const void*)params << (void*)counters << (void*)traj << (void*)ckpt_t
        << (void*)ckpt_y << (void*)stage_ws << ":" << ckpt_every << ":" << launches << ":" << per << ":" << cfg.schedule << ":"
        << cfg.group << ":" << cfg.block;
    if (c->graph_key != key.str() || !c->graph_exec) {
        drop_graph(c);
        // compile before recording: no JIT inside a capture
        if (!get_program(c, cfg.schedule, cfg.group, cfg.block)) return 0;
        if (!hip_ok(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
            return 0;
        int ok = enqueue_all(c->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!ok) {
            if (g) hipGraphDestroy(g);
            return 0;
        }
        if (!hip_ok(c, e, "hipStreamEndCapture")) return 0;
        c->graph = g;
        if (!hip_ok(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate"))
            return 0;
        c->graph_key = key.str();
    }
    return hip_ok(c, hipGraphLaunch(c->graph_exec, s), "hipGraphLaunch") ? 1 : 0;
}

int launch_config(const sodecl_ctx* c, int* schedule, int* group, int* block)
{
    LaunchCfg cfg = choose_config(c);
    *schedule = cfg.schedule;
    *group = cfg.group;
    *block = cfg.block;
    return 1;
}

}  // namespace sodecl_amd

namespace sodecl_amd {

// Loads the host-share library and keeps a host copy of the parameters of
// the orbits it will integrate (taken from the device buffer).
int setup_host_share(sodecl_ctx* c)
{
    std::string err;
    if (!load_host_share(c, err)) return fail(c, err);
    if (c->hs_p.size() != (size_t)c->orbits * c->npar) {
        c->hs_p.assign((size_t)c->orbits * c->npar, 0.0);
        if (c->npar > 0 &&
            !hip_ok(c, hipMemcpy(c->hs_p.data(), c->d_params, c->hs_p.size() * sizeof(double), hipMemcpyDeviceToHost),
                    "hipMemcpy(params)"))
            return 0;
    }
    return 1;
}

}  // namespace sodecl_amd

extern "C" {

void sodecl_default_options(sodecl_options* o)
{
    std::memset(o, 0, sizeof *o);
    o->schedule = SODECL_SCHEDULE_AUTO;
}

long sodecl_num_launches(double tspan, double dt, int ksteps) { return num_launches(tspan, dt, ksteps); }

int sodecl_create(sodecl_ctx** out, const char* kernel_path, const char* system_path, int solver, double dt,
                  double tspan, int ksteps, int neq, int npar, int nnoi, int orbits, int output_type)
{
    if (!out) return 0;
    auto* c = new sodecl_ctx();
    *out = c;
    c->kernel_path = kernel_path ? kernel_path : "";
    c->system_path = system_path ? system_path : "";
    if (!init_ctx(c, solver, dt, tspan, ksteps, neq, npar, nnoi, orbits, output_type)) return 0;
    if (!read_text(c->system_path, c->system_source))
        return fail(c, "Error opening file " + c->system_path);   // sodeclmgr.cpp:325-328
    c->features = detect_features(c->system_source);
    return 1;
}

int sodecl_create_from_source(sodecl_ctx** out, const char* system_source, int solver, double dt, double tspan,
                              int ksteps, int neq, int npar, int nnoi, int orbits, int output_type)
{
    if (!out) return 0;
    auto* c = new sodecl_ctx();
    *out = c;
    if (!init_ctx(c, solver, dt, tspan, ksteps, neq, npar, nnoi, orbits, output_type)) return 0;
    c->system_source = system_source ? system_source : "";
    c->features = detect_features(c->system_source);
    return 1;
}

void sodecl_destroy(sodecl_ctx* c)
{
    if (!c) return;
    if (c->ready || c->stream) {
        // run_device may have queued work on caller streams that still reads
        // this context's modules: wait for the whole device
        hipSetDevice(c->device);
        hipDeviceSynchronize();
    }
    free_device(c);
    unload_host_share(c);
    delete c;
}

int sodecl_choose_device(sodecl_ctx* c, uint32_t platform, uint64_t device_type, uint32_t device)
{
    if (!c) return 0;
    // sodeclmgr.cpp:164-212: out-of-range platform/device or wrong type -> 0
    if (platform != 0) return fail(c, "Selected platform number is out of bounds.");
    if (device_type != SODECL_DEVICE_TYPE_ALL && device_type != SODECL_DEVICE_TYPE_GPU)
        return fail(c, "Selected device is not of the type selected.");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || (int)device >= n)
        return fail(c, "Selected device number is out of bounds.");
    c->device = (int)device;
    c->device_chosen = true;
    return 1;
}

void sodecl_set_local_group_size(sodecl_ctx* c, int local_group_size)
{
    if (c) c->local_group_size = local_group_size;
}

void sodecl_set_t0(sodecl_ctx* c, const double* t0)
{
    if (c) c->h_t0 = t0;
}
void sodecl_set_y0(sodecl_ctx* c, const double* y0)
{
    if (c) c->h_y0 = y0;
}
void sodecl_set_params(sodecl_ctx* c, const double* params)
{
    if (c) c->h_params = params;
}

int sodecl_set_rng_counters(sodecl_ctx* c, const double* counters)
{
    if (!c || !counters) return 0;
    c->counters.assign(counters, counters + c->orbits);
    if (c->ready && c->d_counters)
        return hip_ok(c, hipMemcpy(c->d_counters, c->counters.data(), c->orbits * sizeof(double),
                                   hipMemcpyHostToDevice),
                      "hipMemcpy(counters)")
                   ? 1
                   : 0;
    return 1;
}

int sodecl_get_rng_counters(sodecl_ctx* c, double* counters)
{
    if (!c || !counters || c->nnoi <= 0) return 0;
    if (c->ready && c->d_counters) {
        if (!hip_ok(c, hipMemcpy(c->counters.data(), c->d_counters, c->orbits * sizeof(double),
                                 hipMemcpyDeviceToHost),
                    "hipMemcpy(counters)"))
            return 0;
    }
    std::memcpy(counters, c->counters.data(), c->orbits * sizeof(double));
    return 1;
}

int sodecl_set_options(sodecl_ctx* c, const sodecl_options* o)
{
    if (!c || !o) return 0;
    if (o->schedule < 0 || o->schedule > SODECL_SCHEDULE_TEAM) return fail(c, "invalid schedule");
    if (o->group_size < 0 || o->group_size > 64) return fail(c, "group_size must be in 0..64");
    if (o->block_size < 0 || o->block_size > 1024) return fail(c, "block_size must be in 0..1024");
    if (!(o->newton_tol >= 0)) return fail(c, "newton_tol must be non-negative");
    if (o->pairwise < -1 || o->pairwise > 4) return fail(c, "pairwise must be in -1..4");
    if (o->library_sin < 0 || o->library_sin > 1) return fail(c, "library_sin must be 0 or 1");
    if (o->library_exp < 0 || o->library_exp > 1) return fail(c, "library_exp must be 0 or 1");
    if (!(o->host_share >= 0 && o->host_share < 1)) return fail(c, "host_share must be in [0, 1)");
    // options that change the program text
    bool recompile = o->newton_max_iter != c->opts.newton_max_iter || o->fp_contract != c->opts.fp_contract ||
                     o->pairwise != c->opts.pairwise || o->newton_tol != c->opts.newton_tol ||
                     o->library_sin != c->opts.library_sin || o->library_exp != c->opts.library_exp;
    c->opts = *o;
    if (recompile && (c->ready || !c->programs.empty())) {
        // Work enqueued by run_device on caller streams may still use the old
        // modules, and a recorded graph holds their kernels: wait for the
        // device, drop the graph, then unload.
        if (!hip_ok(c, hipSetDevice(c->device), "hipSetDevice")) return 0;
        if (!hip_ok(c, hipDeviceSynchronize(), "hipDeviceSynchronize")) return 0;
        drop_graph(c);
        for (auto& kv : c->programs)
            if (kv.second->module) hipModuleUnload(kv.second->module);
        c->programs.clear();
        unload_host_share(c);   // its program text follows the same options
    }
    if (c->ready && c->opts.host_share > 0 && !setup_host_share(c)) return 0;
    return 1;
}

int sodecl_setup(sodecl_ctx* c)
{
    if (!c) return 0;
    if (!c->last_error.empty() && c->system_source.empty()) return 0;
    if (!c->h_y0 || !c->h_params) return fail(c, "initial state and parameters must be set before setup");
    if (c->solver == SODECL_STOCHASTIC_EULER && (c->nnoi <= 0 || !c->features.sode_system_stoch))
        return fail(c, "the stochastic solver needs _numnoi_ > 0 and sode_system_stoch in the system source");
    if (!c->features.sode_system) return fail(c, "the system source defines no sode_system");
    if ((c->solver == SODECL_IMPLICIT_EULER || c->solver == SODECL_IMPLICIT_MIDPOINT) && !c->features.calc_jacobian &&
        !c->features.jacobian_band)
        return fail(c, "implicit solvers need calc_jacobian (or calc_jacobian_band) in the system source");
    free_device(c);
    if (!hip_ok(c, hipSetDevice(c->device), "hipSetDevice")) return 0;
    if (!hip_ok(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) return 0;
    if (!hip_ok(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking), "hipStreamCreate")) return 0;
    if (!hip_ok(c, hipEventCreate(&c->ev_start), "hipEventCreate")) return 0;
    if (!hip_ok(c, hipEventCreate(&c->ev_stop), "hipEventCreate")) return 0;
    const size_t n = (size_t)c->orbits;
    if (!hip_ok(c, hipMalloc(&c->d_t, n * sizeof(double)), "hipMalloc(t0)")) return 0;
    if (!hip_ok(c, hipMalloc(&c->d_y, n * c->neq * sizeof(double)), "hipMalloc(y0)")) return 0;
    if (!hip_ok(c, hipMalloc(&c->d_params, std::max<size_t>(1, n * c->npar) * sizeof(double)), "hipMalloc(params)"))
        return 0;
    if (!hip_ok(c, hipMalloc(&c->d_newton, n * sizeof(long long)), "hipMalloc(newton)")) return 0;
    if (!hip_ok(c, hipMemsetAsync(c->d_newton, 0, n * sizeof(long long), c->stream), "hipMemset(newton)")) return 0;
    std::vector<double> zeros;
    const double* t0 = c->h_t0;
    if (!t0) {
        zeros.assign(n, 0.0);
        t0 = zeros.data();
    }
    if (!hip_ok(c, hipMemcpy(c->d_t, t0, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(t0)")) return 0;
    if (!hip_ok(c, hipMemcpy(c->d_y, c->h_y0, n * c->neq * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(y0)"))
        return 0;
    if (c->npar > 0 &&
        !hip_ok(c, hipMemcpy(c->d_params, c->h_params, n * c->npar * sizeof(double), hipMemcpyHostToDevice),
                "hipMemcpy(params)"))
        return 0;
    if (c->nnoi > 0) {
        if (!hip_ok(c, hipMalloc(&c->d_counters, n * sizeof(double)), "hipMalloc(counters)")) return 0;
        if (!hip_ok(c, hipMemcpy(c->d_counters, c->counters.data(), n * sizeof(double), hipMemcpyHostToDevice),
                    "hipMemcpy(counters)"))
            return 0;
    }
    // the static scratch array's address on this device, resolved once per
    // device here so that the entry point that uses it resolves nothing; a
    // failure surfaces there, not in the entry points that never use it
    {
        std::string scratch_err;
        scratch_address(c->device, scratch_err);
    }
    // the autotuner times the candidate schedules once per configuration
    if (c->opts.schedule == SODECL_SCHEDULE_AUTO && c->opts.autotune && !autotune(c)) return 0;
    // compile the program the first run will use (JIT outside the timed loop)
    LaunchCfg cfg = choose_config(c);
    if (!get_program(c, cfg.schedule, cfg.group, cfg.block)) return 0;
    if (!hip_ok(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return 0;
    c->hs_p.clear();
    if (c->opts.host_share > 0 && !setup_host_share(c)) return 0;
    c->ready = true;
    return 1;
}

int sodecl_run(sodecl_ctx* c)
{
    if (!c) return 0;
    if (!c->ready) return fail(c, "sodecl_setup has not succeeded");
    if (!hip_ok(c, hipSetDevice(c->device), "hipSetDevice")) return 0;
    const long launches = num_launches(c->tspan, c->dt, c->ksteps);
    if (launches < 0) return fail(c, "too many kernel launches");
    auto t0 = std::chrono::steady_clock::now();
    c->last_host_ms = 0.0;
    c->last_host_orbits = 0;
    int ok;
    if (c->opts.schedule == SODECL_SCHEDULE_PLAIN)
        ok = run_plain(c, launches);
    else
        ok = run_fast(c, launches, choose_config(c));
    auto t1 = std::chrono::steady_clock::now();
    c->last_runtime = std::chrono::duration<double>(t1 - t0).count();
    if (ok && c->opts.verbose)
        std::cout << "Compute results runtime: " << c->last_runtime << " sec.\n";   // sodeclmgr.cpp:1019-1020
    return ok;
}

int sodecl_run_device(sodecl_ctx* c, double* t0, double* y0, const double* params, double* counters, double* traj,
                      void* stream)
{
    return sodecl_amd::run_device_impl(c, t0, y0, params, counters, traj, nullptr, nullptr, 0, stream, nullptr);
}

size_t sodecl_output_size(const sodecl_ctx* c)
{
    if (!c) return 0;
    return c->ext_output ? c->ext_used : c->output.size();
}

const double* sodecl_output(const sodecl_ctx* c)
{
    if (!c) return nullptr;
    return c->ext_output ? c->ext_output : c->output.data();
}

int sodecl_set_output_buffer(sodecl_ctx* c, double* host, size_t capacity)
{
    if (!c) return 0;
    c->ext_output = host;
    c->ext_capacity = host ? capacity : 0;
    c->ext_used = 0;
    return 1;
}

double sodecl_last_runtime(const sodecl_ctx* c) { return c ? c->last_runtime : 0.0; }
double sodecl_last_kernel_ms(const sodecl_ctx* c) { return c ? c->last_kernel_ms : 0.0; }
double sodecl_last_host_ms(const sodecl_ctx* c, int* host_orbits)
{
    if (host_orbits) *host_orbits = c ? c->last_host_orbits : 0;
    return c ? c->last_host_ms : 0.0;
}

int sodecl_get_launch_config(const sodecl_ctx* c, int* schedule, int* group, int* block)
{
    if (!c) return 0;
    if (schedule) *schedule = c->used_schedule;
    if (group) *group = c->used_group;
    if (block) *block = c->used_block;
    return 1;
}

int sodecl_get_newton_updates(sodecl_ctx* c, long long* counts)
{
    if (!c || !counts || !c->ready) return 0;
    return hip_ok(c, hipMemcpy(counts, c->d_newton, c->orbits * sizeof(long long), hipMemcpyDeviceToHost),
                  "hipMemcpy(newton)")
               ? 1
               : 0;
}

size_t sodecl_compile_only(sodecl_ctx* c, const char* arch)
{
    if (!c) return 0;
    LaunchCfg cfg = choose_config(c);
    c->used_schedule = cfg.schedule;   // reported by sodecl_get_launch_config
    c->used_group = cfg.group;
    c->used_block = cfg.block;
    ProgramKey key = sodecl_amd::make_key(c, cfg.schedule, cfg.group, cfg.block, std::string());
    size_t total = 0;
    std::vector<std::string> variants = {std::string()};
    // the adjoint program too, where this configuration supports one
    if (cfg.schedule == 2 && c->features.vjp_group && (c->solver == SODECL_EULER || c->solver == SODECL_RUNGE_KUTTA))
        variants.push_back("#define SODECL_ADJOINT 1\n");
    if (cfg.schedule == 1 && c->features.calc_jacobian && (c->npar == 0 || c->features.jacobian_params) &&
        (c->solver == SODECL_IMPLICIT_EULER || c->solver == SODECL_IMPLICIT_MIDPOINT))
        variants.push_back("#define SODECL_ADJOINT 1\n");
    for (auto& extra : variants) {
        key.extra_defines = extra;
        c->last_program = assemble_program(c->system_source, key, c->features);
        std::vector<char> code;
        std::string log;
        if (!compile_program(c->last_program, arch ? arch : "gfx950", key.fp_contract, code, log,
                             key.one_wave_per_simd)) {
            fail(c, "Failed to build program executable:\n" + log);
            return 0;
        }
        total += code.size();
    }
    return total;
}

const char* sodecl_program_source(sodecl_ctx* c) { return c ? c->last_program.c_str() : ""; }

const char* sodecl_autotune_report(const sodecl_ctx* c) { return c ? c->tune_report.c_str() : ""; }

const char* sodecl_last_error(const sodecl_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int sodecl_read_binary_data_from_file(const char* filename, double* data, size_t capacity, size_t* count)
{
    // sodeclmgr.cpp:839-867: the whole file as raw doubles
    std::ifstream f(filename, std::ios::binary | std::ios::ate);
    if (!f) {
        std::cout << "ERROR: can't open " << filename << " file." << std::endl;
        return 0;
    }
    std::streamsize size = f.tellg();
    f.seekg(0, std::ios::beg);
    size_t n = (size_t)size / sizeof(double);
    if (n > capacity) n = capacity;
    f.read((char*)data, (std::streamsize)(n * sizeof(double)));
    if (count) *count = n;
    return 1;
}

int sodecl_read_data_from_file(const char* filename, double* data, size_t capacity, size_t* count)
{
    // sodeclmgr.cpp:877-914: whitespace-separated numbers, line by line
    std::ifstream f(filename);
    if (!f) {
        std::cout << "ERROR: can't open " << filename << " file." << std::endl;
        return 0;
    }
    size_t j = 0;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream iss(line);
        std::string tok;
        while (iss >> tok) {
            if (j < capacity) data[j] = std::atof(tok.c_str());
            ++j;
        }
    }
    if (count) *count = std::min(j, capacity);
    return 1;
}

}  // extern "C"
