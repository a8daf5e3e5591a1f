/*
 * sodecl.h -- C ABI of the MI355X-native SODECL integrator library
 * (libsodecl_amd.so).
 *
 * The boundary mirrors the reference's host manager, class
 * sodecl::sodeclmgr (include/sodeclmgr.hpp:27-369, src/sodeclmgr.cpp of
 * avramidis/sodecl), as plain C: one opaque context per manager object,
 * plain pointers and sizes, no HIP or torch types.  Every function that can
 * fail returns 1 on success and 0 on failure, as the reference's member
 * functions do; sodecl_last_error() then holds the message.
 *
 * Ownership: the context owns all device memory and the host output.  Input
 * pointers given to sodecl_set_t0/y0/params are stashed and read at
 * sodecl_setup(), the same lifetime rule as the reference
 * (sodeclmgr.cpp:803-829 stash, :735-770 upload).  One context per thread.
 *
 * Layouts (the reference's):
 *   t0[orbits]; y0[orbits][neq]; params[orbits][npar]; counters[orbits];
 *   output[nsnap][neq][orbits], nsnap = launches + 1, launches = number of
 *   integers j >= 0 with j < tspan/(dt*ksteps) (sodeclmgr.cpp:951-1015).
 */
#ifndef SODECL_AMD_SODECL_H
#define SODECL_AMD_SODECL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: sodecl_options gained library_sin (callers compiled against 1 pass a shorter struct)
 * 3: sodecl_options gained library_exp
 * 4: sodecl_options gained host_share */
#define SODECL_AMD_ABI_VERSION 4

/* include/solver_Type.hpp:14-20 */
enum sodecl_solver {
    SODECL_STOCHASTIC_EULER = 0,
    SODECL_EULER = 1,
    SODECL_RUNGE_KUTTA = 2,
    SODECL_IMPLICIT_EULER = 3,
    SODECL_IMPLICIT_MIDPOINT = 4
};

/* include/output_Type.hpp:14-18 */
enum sodecl_output { SODECL_OUTPUT_NONE = 0, SODECL_OUTPUT_ARRAY = 1, SODECL_OUTPUT_FILE = 2 };

/* include/device_Type.hpp:14-19 (values of CL_DEVICE_TYPE_*) */
#define SODECL_DEVICE_TYPE_CPU 0x2ull
#define SODECL_DEVICE_TYPE_GPU 0x4ull
#define SODECL_DEVICE_TYPE_ACCELERATOR 0x8ull
#define SODECL_DEVICE_TYPE_ALL 0xFFFFFFFFull

/* Kernel schedules (no reference counterpart: the reference always runs one
 * work-item per orbit and one launch per ksteps segment). */
enum sodecl_schedule {
    SODECL_SCHEDULE_AUTO = 0,  /* group when the plugin has per-component entry points
                                  and the system is large enough, else lane; timed
                                  choice (team included) when autotune is on */
    SODECL_SCHEDULE_LANE = 1,  /* one orbit per lane, state in registers */
    SODECL_SCHEDULE_GROUP = 2, /* group_size lanes per orbit, stage vectors in LDS */
    SODECL_SCHEDULE_PLAIN = 3, /* the reference's launch design: one launch per ksteps
                                  segment with a blocking state read-back before each */
    SODECL_SCHEDULE_TEAM = 4   /* group_size lanes per orbit (default 2), each holding the
                                  whole state; the plugin's sode_system_team shares the
                                  drift's costly terms among them (Euler and RK4) */
};

/* Extra knobs (no reference counterpart).  Zero-initialised fields mean the
 * default, which is the reference's behaviour. */
typedef struct sodecl_options {
    int schedule;            /* enum sodecl_schedule */
    int group_size;          /* lanes per orbit for the group schedule (power of two <= 64); 0 = auto */
    int block_size;          /* threads per block; 0 = local_group_size if set, else auto */
    int newton_max_iter;     /* 0 = 1000 (src/kernels/ie.cl:69, im.cl:65) */
    int segments_per_launch; /* ksteps segments per kernel launch; 0 = as many as the
                                trajectory buffer budget allows */
    int use_graph;           /* device runs of several launches (segments_per_launch set): record the
                                launch sequence once as a hipGraph and replay it */
    int autotune;            /* with schedule AUTO: time the candidate schedules, group sizes and block
                                sizes on one segment at setup and keep the fastest (cached on disk) */
    int fp_contract;         /* 0 = default arithmetic: fused multiply-add contraction of the system
                                source's and the solvers' operations, and their divisions as products
                                with reciprocals (each within about an ulp of the IEEE quotient; the
                                device library's functions are compiled as written), 1 = strict
                                arithmetic: neither */
    int verbose;             /* print "Compute results runtime" like the reference */
    size_t traj_budget_bytes;/* device trajectory buffer budget; 0 = 8 GiB */
    int pairwise;            /* group schedule with a plugin that defines the pairwise drift (sode_pair,
                                sode_system_pairsum): 0 = automatic (the full N x (N+1) LDS matrix when
                                the group has two lanes per component and the matrix is at most 16 KiB
                                per orbit, else lane exchange when the group has a lane per component,
                                else the full matrix while it fits, else off), 1 = always (lane exchange, else half
                                storage), 2 = half storage, 3 = full matrix, 4 = lane exchange only,
                                -1 = never.  The LDS layouts reproduce the per-component drift bit for
                                bit; the lane exchange sums the same terms in another order */
    double newton_tol;       /* Newton stop threshold of the implicit solvers (_epsilon1_); 0 = 1e-07
                                (src/sodeclmgr.cpp:424-430) */
    int library_sin;         /* sin() of a double in the system source: 0 = SODECL's own sine (default;
                                Cody-Waite reduction by pi and a polynomial, at most 0.88 ulp measured,
                                about two thirds of the device library's instructions; the library's sine
                                for |x| > 2^30 and within ~|x| 1.6e-16 of a nonzero multiple of pi),
                                1 = the device library's sin */
    int library_exp;         /* exp() of a double in the system source: 0 = SODECL's own exponential
                                (default; Tang's table-driven method with 256 points, at most 0.68 ulp
                                measured, fewer instructions than the device library's), 1 = the device
                                library's exp */
    double host_share;       /* fraction of the orbits sodecl_run integrates on the host's cores while the
                                GPU integrates the rest (0 = none, below 1): the last
                                floor(orbits * host_share) orbits, with the same lane integrator and system
                                source compiled for the host by its C++ compiler ("c++", or
                                SODECL_AMD_HOST_CXX) at setup.  Applies to sodecl_run when the trajectory
                                fits one device buffer; device runs integrate every orbit on the GPU.  The
                                reference names such a CPU+GPU split as future work (README.md:147) */
} sodecl_options;

typedef struct sodecl_ctx sodecl_ctx;

/* Replaces sodeclmgr::sodeclmgr (sodeclmgr.hpp:122-124, sodeclmgr.cpp:26-83).
 * kernel_path: directory of solver sources; the solvers are built into this
 * library, so it is accepted and ignored.  system_path: the user system
 * source (OpenCL-C style functions).  RNG counters of SDE runs are drawn as
 * in the reference (random, per orbit) unless sodecl_set_rng_counters() is
 * called. */
int sodecl_create(sodecl_ctx** ctx, const char* kernel_path, const char* system_path, int solver, double dt,
                  double tspan, int ksteps, int neq, int npar, int nnoi, int orbits, int output_type);

/* As sodecl_create, with the system source given as text instead of a path
 * (no reference counterpart). */
int sodecl_create_from_source(sodecl_ctx** ctx, const char* system_source, int solver, double dt, double tspan,
                              int ksteps, int neq, int npar, int nnoi, int orbits, int output_type);

/* Replaces sodeclmgr::~sodeclmgr (sodeclmgr.cpp:88-149). */
void sodecl_destroy(sodecl_ctx* ctx);

/* Replaces sodeclmgr::choose_device (sodeclmgr.hpp:144, sodeclmgr.cpp:164-212):
 * platform must be 0; device is the HIP device ordinal; device_type must be
 * ALL or GPU. */
int sodecl_choose_device(sodecl_ctx* ctx, uint32_t platform, uint64_t device_type, uint32_t device);

/* Replaces sodeclmgr::set_local_group_size (sodeclmgr.hpp:151). */
void sodecl_set_local_group_size(sodecl_ctx* ctx, int local_group_size);

/* Replace sodeclmgr::set_t0/set_y0/set_params (sodeclmgr.hpp:320-334): pointers
 * are stashed and read at sodecl_setup(). */
void sodecl_set_t0(sodecl_ctx* ctx, const double* t0);
void sodecl_set_y0(sodecl_ctx* ctx, const double* y0);
void sodecl_set_params(sodecl_ctx* ctx, const double* params);

/* Per-orbit SDE counters, doubles as the reference's m_rcounter
 * (sodeclmgr.cpp:48-59), copied immediately.  Values below 2^53 reproduce the
 * reference's noise stream exactly.  No reference counterpart (the reference
 * draws them from std::random_device and never exposes them). */
int sodecl_set_rng_counters(sodecl_ctx* ctx, const double* counters);
/* Counters as they stand now (after a run: advanced by the calls made). */
int sodecl_get_rng_counters(sodecl_ctx* ctx, double* counters);

int sodecl_set_options(sodecl_ctx* ctx, const sodecl_options* options);
void sodecl_default_options(sodecl_options* options);

/* Replaces sodeclmgr::setup_sode_solver (sodeclmgr.hpp:368,
 * sodeclmgr.cpp:1042-1097): assembles and JIT-compiles the program, allocates
 * device memory, uploads inputs. */
int sodecl_setup(sodecl_ctx* ctx);

/* Replaces sodeclmgr::run_sode_solver (sodeclmgr.hpp:361,
 * sodeclmgr.cpp:921-1035).  Integrates from the current device state (a
 * second call continues where the first ended, as the reference's device
 * buffers do) and appends to the output: host array (ARRAY), raw doubles
 * appended to sodecloutput.bin in the working directory (FILE), or nothing
 * (NONE). */
int sodecl_run(sodecl_ctx* ctx);

/* Output of the runs so far: nsnap*neq*orbits doubles per run, [snap][eq][orbit]
 * (the reference's public m_output, sodeclmgr.hpp:33). */
size_t sodecl_output_size(const sodecl_ctx* ctx);
const double* sodecl_output(const sodecl_ctx* ctx);
/* Let the next sodecl_run() write its ARRAY output into caller memory of
 * capacity doubles instead of the context's own buffer. */
int sodecl_set_output_buffer(sodecl_ctx* ctx, double* host, size_t capacity);

/* Number of kernel-segment launches and snapshots of one run. */
long sodecl_num_launches(double tspan, double dt, int ksteps);

/* Device-resident run (no reference counterpart): all pointers are device
 * memory of the context's device.  t0/y0/params/counters as above (updated
 * in place); traj receives nsnap snapshots [snap][neq][orbits] (may be NULL);
 * stream is a hipStream_t; NULL is HIP's default (null) stream, ordered with
 * other default-stream work.  Returns after the work is enqueued;
 * synchronise the stream before reading results. */
int sodecl_run_device(sodecl_ctx* ctx, double* t0, double* y0, const double* params, double* counters,
                      double* traj, void* stream);

/* ---- Discrete adjoint (no reference counterpart) ----------------------
 * Euler and RK4 under the group schedule, with a plugin that defines
 * sode_system_i and sode_system_vjp_group: the exact transpose of the
 * forward steps.  Implicit Euler and implicit midpoint under the lane
 * schedule, with a plugin that defines calc_jacobian (or calc_jacobian_exact,
 * read instead when present) and calc_jacobian_params: the derivative of
 * each step's implicit equation at
 * its converged Newton solution (the transposed Newton system solved by
 * tridiagonal sweeps).  Given dL/dtraj for a trajectory of the context's
 * configuration, returns dL/dy0[orbits][neq] and dL/dparams[orbits][npar].
 * Memory is bounded by checkpointing every ckpt_every steps
 * (sodecl_default_checkpoint_interval: ceil(sqrt(steps))). */
int sodecl_default_checkpoint_interval(const sodecl_ctx* ctx);
/* Number of checkpoints: ceil(steps / ckpt_every); ckpt_y holds that many
 * [neq][orbits] states, ckpt_t that many [orbits] times. */
long sodecl_checkpoint_count(const sodecl_ctx* ctx, int ckpt_every);
/* Doubles of device workspace sodecl_adjoint_device needs. */
size_t sodecl_adjoint_workspace_size(const sodecl_ctx* ctx, int ckpt_every);
/* sodecl_run_device that also writes the checkpoints (device memory). */
int sodecl_run_device_checkpointed(sodecl_ctx* ctx, double* t0, double* y0, const double* params, double* counters,
                                   double* traj, double* ckpt_t, double* ckpt_y, int ckpt_every, void* stream);
/* Backward sweep: cot is dL/dtraj [nsnap][neq][orbits]; all device memory. */
int sodecl_adjoint_device(sodecl_ctx* ctx, const double* ckpt_t, const double* ckpt_y, const double* params,
                          const double* cot, double* grad_y0, double* grad_params, double* work, int ckpt_every,
                          void* stream);
/* As sodecl_adjoint_device, with no work argument and no allocation: no
 * reference counterpart.  For callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid it).
 * The workspace is the library's static scratch array of 8 Mi doubles
 * (64 MiB) per GPU; configurations whose sodecl_adjoint_workspace_size
 * exceeds that are rejected (return 0, sodecl_last_error says so).  The
 * adjoint program must already be loaded on the context: a checkpointed
 * forward run (sodecl_run_device_checkpointed) loads it.  Every call on a
 * GPU shares that GPU's array (nothing in it needs to start at zero: each
 * segment's rows are written before they are read), so calls that overlap
 * on one GPU race. */
int sodecl_adjoint_device_noalloc(sodecl_ctx* ctx, const double* ckpt_t, const double* ckpt_y, const double* params,
                                  const double* cot, double* grad_y0, double* grad_params, int ckpt_every,
                                  void* stream);
/* Without recomputation (Euler and RK4, group schedule), where memory allows:
 * sodecl_run_device_stored integrates like sodecl_run_device and also writes
 * every step's stage vectors to work (sodecl_stored_workspace_size doubles,
 * orbits x steps x (stages x neq + 1): 67 GB for 16384 orbits of a 64-equation
 * RK4 over 2000 steps); sodecl_adjoint_device_stored then sweeps backwards
 * from that workspace.  Gradients are bit-identical to the checkpointed
 * adjoint's. */
size_t sodecl_stored_workspace_size(const sodecl_ctx* ctx);
int sodecl_run_device_stored(sodecl_ctx* ctx, double* t0, double* y0, const double* params, double* traj,
                             double* work, void* stream);
int sodecl_adjoint_device_stored(sodecl_ctx* ctx, const double* params, const double* cot, double* grad_y0,
                                 double* grad_params, double* work, void* stream);

/* ---- On-device sampling of parameters and initial states --------------
 * (no reference counterpart; replaces the host-side draws of the
 * reference's benchmark scripts, e.g. benchmarks/kuramoto_benchmark.py:49-68)
 * Fills out[rows][cols] (device memory, row-major: rows = orbits) with
 * independent draws, column c from dists[c] (host memory).  Entry (r, c)
 * comes from one Threefry-2x64-20 block with counter {r, c} and key
 * {seed, 0x5EEDD15C0FFEE5ED}; u0, u1 are its two words converted to
 * open-interval (0, 1) doubles as the SDE noise does.
 *   CONSTANT     a
 *   UNIFORM      a + (b - a) u0                      (one fused multiply-add)
 *   LOG_UNIFORM  exp(log a + (log b - log a) u0)     (a, b > 0)
 *   NORMAL       a + b sqrt(-2 log u0) cos(2 pi u1)  (mean a, deviation b)
 * Enqueued on stream (NULL: the default stream); returns 0 on bad input. */
enum sodecl_dist_kind {
    SODECL_DIST_CONSTANT = 0,
    SODECL_DIST_UNIFORM = 1,
    SODECL_DIST_LOG_UNIFORM = 2,
    SODECL_DIST_NORMAL = 3
};
typedef struct sodecl_dist {
    int kind;   /* enum sodecl_dist_kind */
    double a, b;
} sodecl_dist;
int sodecl_sample_device(const sodecl_dist* dists, int rows, int cols, unsigned long long seed, double* out,
                         void* stream);

/* ---- Batched integration (no reference counterpart) ---------------------
 * Orbits of several system sizes ("size classes", each a value of _numeq_)
 * and individual step counts in one persistent launch, implicit Euler or
 * implicit midpoint with the reference's Newton iteration.  The plugin must
 * define sode_system_i and calc_jacobian_band_i; it is compiled once per
 * size class.  Packed device layouts: orbit o's initial state at
 * y0[yoff[o]], its steps[o]/ksteps + 1 snapshots [snap][neq_o] at
 * out[ooff[o]], params[orbits][npar]; newton[orbits] receives the Newton
 * updates each orbit took. */
typedef struct sodecl_batch sodecl_batch;
int sodecl_batch_create(sodecl_batch** batch, const char* system_source, int solver, double dt, int ksteps, int npar,
                        int nclasses, const int* class_neq);
void sodecl_batch_destroy(sodecl_batch* batch);
int sodecl_batch_set_options(sodecl_batch* batch, const sodecl_options* options);
int sodecl_batch_choose_device(sodecl_batch* batch, uint32_t device);
/* orbit_class[orbits] indexes class_neq; orbit_steps[orbits] are multiples of ksteps. */
int sodecl_batch_setup(sodecl_batch* batch, int orbits, const int* orbit_class, const int* orbit_steps);
size_t sodecl_batch_state_size(const sodecl_batch* batch);
size_t sodecl_batch_output_size(const sodecl_batch* batch);
int sodecl_batch_offsets(const sodecl_batch* batch, long long* yoff, long long* ooff);
int sodecl_batch_run_device(sodecl_batch* batch, const double* y0, const double* params, double* out,
                            long long* newton, void* stream);
size_t sodecl_batch_compile_only(sodecl_batch* batch, const char* arch);
const char* sodecl_batch_program_source(const sodecl_batch* batch);
const char* sodecl_batch_last_error(const sodecl_batch* batch);

/* Wall time of the last sodecl_run segment loop in seconds, and the kernel
 * time of the last run measured with HIP events on the launch stream. */
double sodecl_last_runtime(const sodecl_ctx* ctx);
double sodecl_last_kernel_ms(const sodecl_ctx* ctx);
/* Host share of the last sodecl_run: wall milliseconds of the host orbits'
 * integration and their number (0 without a host share). */
double sodecl_last_host_ms(const sodecl_ctx* ctx, int* host_orbits);
/* Builds (or finds in the cache) the host-share library of this context's
 * configuration and writes its file name to path (capacity bytes); returns
 * the name's length, 0 on failure.  The library exports
 *   void sodecl_host_lane(double* t, double* y, const double* params,
 *                         double* counters, double* traj, long long* newton,
 *                         int orbits, int first, int last, int s0, int nseg,
 *                         int final_snap, long long gstep0, int nthreads)
 * integrating orbits [first, last) of host arrays in the layouts above for
 * nseg ksteps segments, snapshots into traj[s0 ..][neq][orbits]. */
size_t sodecl_host_library(sodecl_ctx* ctx, char* path, size_t capacity);

/* Schedule, group size and block size the last run (or sodecl_compile_only) used. */
int sodecl_get_launch_config(const sodecl_ctx* ctx, int* schedule, int* group_size, int* block_size);

/* What the autotuner measured (candidate: time of one ksteps segment), or
 * "cached" when the choice came from its on-disk cache; empty without it. */
const char* sodecl_autotune_report(const sodecl_ctx* ctx);

/* Newton updates per orbit accumulated over the runs (implicit solvers);
 * counts[orbits]. */
int sodecl_get_newton_updates(sodecl_ctx* ctx, long long* counts);

/* Assembles and compiles the program for gfx950 without a device (build
 * check); returns the code-object size in bytes or 0. */
size_t sodecl_compile_only(sodecl_ctx* ctx, const char* arch);
/* The program text the JIT compiles (the counterpart of the reference's
 * Generated_kernel.cl dump, sodeclmgr.cpp:497-500). */
const char* sodecl_program_source(sodecl_ctx* ctx);

const char* sodecl_last_error(const sodecl_ctx* ctx);

/* Streaming device-to-device copy of bytes (a multiple of 16) on stream (a
 * hipStream_t or NULL), by a kernel compiled ahead of time for gfx950.  No
 * reference counterpart: a kernel of known traffic used to calibrate HBM
 * byte counters when profiling the integrators. */
int sodecl_copy_device(void* dst, const void* src, size_t bytes, void* stream);

/* Diagnostic (no reference counterpart): n Threefry-2x64-20 blocks computed
 * on the device by the same functions the SDE integrators and the sampler
 * use, so the device stream can be checked against published known-answer
 * vectors and against the reference's SDE key (Random123/threefry.h:178-293,
 * u01.h:129-131, src/kernels/stochastic_solver_caller.cl:31-52).  Device
 * memory: in[n][4] = {ctr0, ctr1, key0, key1}; out[n][4] = {word0, word1,
 * bits of u01(word0), bits of u01(word1)}, u01 the open-interval (0, 1)
 * double of the SDE noise.  Enqueued on stream (NULL: the default stream). */
int sodecl_threefry_device(const uint64_t* in, uint64_t* out, size_t n, void* stream);

/* Replace sodeclmgr::read_binary_data_from_file / read_data_from_file
 * (sodeclmgr.hpp:344,354; sodeclmgr.cpp:839-914).  data must hold capacity
 * doubles; *count receives the number read.  Return 1/0. */
int sodecl_read_binary_data_from_file(const char* filename, double* data, size_t capacity, size_t* count);
int sodecl_read_data_from_file(const char* filename, double* data, size_t capacity, size_t* count);

#ifdef __cplusplus
}
#endif

#endif
