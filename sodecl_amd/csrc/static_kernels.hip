// Kernels compiled ahead of time for gfx950 (everything that depends on the
// user's system is compiled at run time by hipRTC instead).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "../../include/sodecl.h"
#include "device/random.hip"
#include "sodecl_internal.hpp"

// ---------------------------------------------------------------------------
// Static scratch of sodecl_adjoint_device_noalloc: the backward sweep's
// workspace for callers that cannot allocate device memory.  Every GPU the
// library runs on holds its own copy (64 MiB) until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): the address each device gives for it is
// resolved once per device with hipGetSymbolAddress and used instead.
// ---------------------------------------------------------------------------
#define SODECL_SCRATCH_DOUBLES ((size_t)8 << 20)
#define SODECL_SCRATCH_DEVICES 64

static __device__ __align__(256) double scratch[SODECL_SCRATCH_DOUBLES];

namespace {
std::mutex scratch_mutex;
double* scratch_ptr[SODECL_SCRATCH_DEVICES];
}  // namespace

namespace sodecl_amd {

size_t scratch_doubles() { return SODECL_SCRATCH_DOUBLES; }

double* scratch_address(int device, std::string& err)
{
    if (device < 0 || device >= SODECL_SCRATCH_DEVICES) {
        err = "device ordinal out of range for the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return nullptr;
    }
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        scratch_ptr[device] = (double*)p;
    }
    return scratch_ptr[device];
}

}  // namespace sodecl_amd

extern "C" int sodecl_adjoint_device_noalloc(sodecl_ctx* ctx, const double* ckpt_t, const double* ckpt_y,
                                             const double* params, const double* cot, double* grad_y0,
                                             double* grad_params, int ckpt_every, void* stream)
{
    if (!ctx) return 0;
    std::string err;
    double* work = sodecl_amd::scratch_address(ctx->device, err);
    if (!work) return sodecl_amd::fail(ctx, err);
    return sodecl_amd::adjoint_into_workspace(ctx, ckpt_t, ckpt_y, params, cot, grad_y0, grad_params, work,
                                              SODECL_SCRATCH_DOUBLES, ckpt_every, stream);
}

// Plain streaming copy, 16 bytes per lane per iteration, grid-stride: the
// reference point for reading HBM traffic counters (bytes fetched and written
// by a kernel of known size).
__global__ void __launch_bounds__(256) sodecl_calibration_copy(const double2* __restrict__ src,
                                                                double2* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int sodecl_copy_device(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!dst || !src || bytes % 16 != 0) return 0;
    const size_t n = bytes / 16;
    if (n == 0) return 1;
    int device = 0;
    (void)hipGetDevice(&device);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    size_t blocks = (n + 255) / 256;
    const size_t cap = (size_t)cus * 8;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(sodecl_calibration_copy, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)src, (double2*)dst, n);
    return hipGetLastError() == hipSuccess ? 1 : 0;
}

// Diagnostic: Threefry-2x64-20 blocks and their open-interval doubles,
// computed by the same device functions the integrators and the sampler
// compile.  in[i] = {ctr0, ctr1, key0, key1}; out[i] = {word0, word1, bits of
// u01(word0), bits of u01(word1)}.
__global__ void __launch_bounds__(256) sodecl_threefry_blocks(const sodecl_u64* __restrict__ in,
                                                               sodecl_u64* __restrict__ out, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        sodecl_u64 w0, w1;
        sodecl_threefry2x64_20(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3], &w0, &w1);
        out[4 * i] = w0;
        out[4 * i + 1] = w1;
        out[4 * i + 2] = (sodecl_u64)__double_as_longlong(sodecl_u01(w0));
        out[4 * i + 3] = (sodecl_u64)__double_as_longlong(sodecl_u01(w1));
    }
}

extern "C" int sodecl_threefry_device(const uint64_t* in, uint64_t* out, size_t n, void* stream)
{
    if (!in || !out) return 0;
    if (n == 0) return 1;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(sodecl_threefry_blocks, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const sodecl_u64*)in, (sodecl_u64*)out, n);
    return hipGetLastError() == hipSuccess ? 1 : 0;
}

// Draws of sodecl_sample_device: entry (row, col) of out[rows][cols] comes
// from one Threefry-2x64-20 block with counter {row, col} and key
// {seed, SODECL_SAMPLE_KEY1}, so every entry is reproducible on its own
// whatever the launch shape.  u0, u1 = open-open (0, 1) doubles of the two
// output words.  The distributions of up to 64 columns travel by value in
// the kernel arguments (no device allocation; wider tables take several
// launches).
#define SODECL_SAMPLE_KEY1 0x5EEDD15C0FFEE5EDull
#define SODECL_SAMPLE_CHUNK 64

struct sodecl_dist_chunk {
    sodecl_dist d[SODECL_SAMPLE_CHUNK];
    int c0, nc;   // first column, columns in this chunk
};

__global__ void __launch_bounds__(256) sodecl_sample(const sodecl_dist_chunk chunk, int rows, int cols,
                                                     unsigned long long seed, double* __restrict__ out)
{
    const long long n = (long long)rows * chunk.nc;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const int row = (int)(e / chunk.nc);
        const int k = (int)(e - (long long)row * chunk.nc);
        const int col = chunk.c0 + k;
        const sodecl_dist d = chunk.d[k];
        double v = d.a;
        if (d.kind != SODECL_DIST_CONSTANT) {
            sodecl_u64 w0, w1;
            sodecl_threefry2x64_20((sodecl_u64)row, (sodecl_u64)col, (sodecl_u64)seed, SODECL_SAMPLE_KEY1, &w0, &w1);
            const double u0 = sodecl_u01(w0);
            if (d.kind == SODECL_DIST_UNIFORM) {
                v = fma(d.b - d.a, u0, d.a);
            } else if (d.kind == SODECL_DIST_LOG_UNIFORM) {
                const double la = log(d.a);
                v = exp(fma(log(d.b) - la, u0, la));
            } else {   // normal: Box-Muller, cosine branch
                const double u1 = sodecl_u01(w1);
                v = fma(d.b, sqrt(-2.0 * log(u0)) * cos(2.0 * M_PI * u1), d.a);
            }
        }
        out[(size_t)row * cols + col] = v;
    }
}

extern "C" int sodecl_sample_device(const sodecl_dist* dists, int rows, int cols, unsigned long long seed,
                                    double* out, void* stream)
{
    if (!dists || !out || rows < 0 || cols <= 0) return 0;
    for (int c = 0; c < cols; ++c) {
        const sodecl_dist& d = dists[c];
        if (d.kind < SODECL_DIST_CONSTANT || d.kind > SODECL_DIST_NORMAL) return 0;
        if (d.kind == SODECL_DIST_LOG_UNIFORM && !(d.a > 0 && d.b > 0)) return 0;
    }
    if (rows == 0) return 1;
    int device = 0;
    (void)hipGetDevice(&device);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    for (int c0 = 0; c0 < cols; c0 += SODECL_SAMPLE_CHUNK) {
        sodecl_dist_chunk chunk{};
        chunk.c0 = c0;
        chunk.nc = cols - c0 < SODECL_SAMPLE_CHUNK ? cols - c0 : SODECL_SAMPLE_CHUNK;
        for (int k = 0; k < chunk.nc; ++k) chunk.d[k] = dists[c0 + k];
        const long long n = (long long)rows * chunk.nc;
        long long blocks = (n + 255) / 256;
        const long long cap = (long long)cus * 16;
        if (blocks > cap) blocks = cap;
        hipLaunchKernelGGL(sodecl_sample, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, chunk, rows,
                           cols, seed, out);
        if (hipGetLastError() != hipSuccess) return 0;
    }
    return 1;
}
