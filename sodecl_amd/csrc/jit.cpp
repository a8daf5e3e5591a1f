// Program assembly and hipRTC compilation.
//
// The reference builds one OpenCL program per call by concatenating a macro
// prelude, the user system source, the solver source and the caller kernel
// (src/sodeclmgr.cpp:374-503) and compiles it with the OpenCL runtime
// (:559-622).  Here the same macro contract is emitted, the user's OpenCL-C
// style functions are compiled as host+device functions, and the MI355X
// integrator templates (device/integrators.hip, embedded in this library)
// follow; hipRTC compiles the text for gfx950.  Code objects are cached in
// memory and on disk, keyed by a hash of the text, the options and the
// target, so repeated calls skip the compiler.
#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <mutex>
#include <regex>

#include "sodecl_internal.hpp"

namespace {
#include "_embedded_device.inc"   // kPreludeSource, kIntegratorsSource (device/*.hip)
}

namespace sodecl_amd {

std::string strip_comments(const std::string& s)
{
    std::string out;
    out.reserve(s.size());
    size_t i = 0, n = s.size();
    while (i < n) {
        if (s[i] == '/' && i + 1 < n && s[i + 1] == '/') {
            while (i < n && s[i] != '\n') ++i;
        } else if (s[i] == '/' && i + 1 < n && s[i + 1] == '*') {
            i += 2;
            while (i + 1 < n && !(s[i] == '*' && s[i + 1] == '/')) ++i;
            i += 2;
            out.push_back(' ');
        } else if (s[i] == '"') {
            out.push_back(s[i++]);
            while (i < n && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < n) out.push_back(s[i++]);
                out.push_back(s[i++]);
            }
            if (i < n) out.push_back(s[i++]);
        } else {
            out.push_back(s[i++]);
        }
    }
    return out;
}

PluginFeatures detect_features(const std::string& source)
{
    const std::string s = strip_comments(source);
    auto defines = [&](const char* name) {
        // a function definition or declaration: name '(' ... at file scope
        std::regex re(std::string("(^|[^A-Za-z0-9_])") + name + "\\s*\\(");
        return std::regex_search(s, re);
    };
    PluginFeatures f;
    f.sode_system = defines("sode_system");
    f.sode_system_stoch = defines("sode_system_stoch");
    f.calc_jacobian = defines("calc_jacobian");
    f.component = defines("sode_system_i");
    f.stoch_component = defines("sode_system_stoch_i");
    f.jacobian_band = defines("calc_jacobian_band");
    f.jacobian_band_i = defines("calc_jacobian_band_i");
    f.vjp_group = defines("sode_system_vjp_group");
    f.pairs = defines("sode_pair") && defines("sode_system_pairsum");
    f.team = defines("sode_system_team");
    f.jacobian_params = defines("calc_jacobian_params");
    f.jacobian_exact = defines("calc_jacobian_exact");
    f.calls_exp = defines("exp");
    return f;
}

std::string dt_literal(double dt)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", dt);
    std::string s(buf);
    // a literal such as "1" must still be a double in the kernel
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

long num_launches(double tspan, double dt, int ksteps)
{
    // sodeclmgr.cpp:951: for (int j = 0; j < tspan/(dt*ksteps); j++)
    const double x = tspan / (dt * ksteps);
    if (!(x > 0)) return 0;
    if (x > 2147483647.0) return -1;
    long n = (long)std::ceil(x);
    while (n > 0 && !((double)(n - 1) < x)) --n;
    while ((double)n < x) ++n;
    return n;
}

// "#define _epsilon1_ 1e-07" as the reference prints it (sodeclmgr.cpp:424-430),
// or the caller's Newton threshold written exactly.
std::string epsilon_define(const ProgramKey& k)
{
    if (!(k.newton_tol > 0)) return "#define _epsilon1_ 1e-07\n";
    char buf[64];
    std::snprintf(buf, sizeof buf, "#define _epsilon1_ %.17g\n", k.newton_tol);
    return buf;
}

// sin() and exp() of the system source resolve to sodecl_sin (device/sin.hip)
// and sodecl_exp (device/exp.hip) unless the caller asked for the device
// library's; the solver templates after the system source keep the library's
// functions.
// Text around the user's system source: the own sine and exponential in
// place of the library's, and in the default arithmetic (fp_contract = 0)
// divisions as products with reciprocals (a division by a value fixed over
// the integration then costs one multiply per evaluation, its reciprocal
// hoisted out of the step loop; each quotient within about an ulp of the
// correctly rounded one, the same size of rounding change as contraction).
std::string system_begin(const ProgramKey& k)
{
    return std::string(k.library_sin ? "" : "#define sin sodecl_sin\n") +
           (k.library_exp ? "" : "#define exp sodecl_exp\n") +
           (k.fp_contract ? "" : "#pragma clang fp reciprocal(on)\n");
}
std::string system_end(const ProgramKey& k)
{
    return std::string(k.fp_contract ? "" : "#pragma clang fp reciprocal(off)\n") +
           (k.library_sin ? "" : "#undef sin\n") + (k.library_exp ? "" : "#undef exp\n");
}

std::string assemble_program(const std::string& system_source, const ProgramKey& k, const PluginFeatures& f)
{
    std::ostringstream o;
    o << "// SODECL program assembled for hipRTC\n";
    // The reference's macro contract (sodeclmgr.cpp:389-430)
    o << "#define _numnoi_ " << k.nnoi << "\n";
    o << "#define _listsize_ " << k.orbits << "\n";
    o << "#define _numsteps_ " << k.ksteps << "\n";
    o << "#define _m_dt_ " << dt_literal(k.dt) << "\n";
    o << "#define _numeq_ " << k.neq << "\n";
    o << "#define _numpar_ " << k.npar << "\n";
    if (k.solver == SODECL_IMPLICIT_EULER || k.solver == SODECL_IMPLICIT_MIDPOINT) o << epsilon_define(k);
    // MI355X build configuration
    o << "#define SODECL_HIP 1\n";
    o << "#define SODECL_STRICT " << (k.fp_contract ? 1 : 0) << "\n";
    o << "#define SODECL_EXP_IN_LDS " << (!k.library_exp && f.calls_exp ? 1 : 0) << "\n";
    o << "#define SODECL_SOLVER " << k.solver << "\n";
    o << "#define SODECL_SCHEDULE " << k.schedule << "\n";
    o << "#define SODECL_GROUP " << k.group << "\n";
    o << "#define SODECL_BLOCK " << k.block << "\n";
    o << "#define SODECL_NEWTON_MAXIT " << k.newton_max_iter << "\n";
    o << "#define SODECL_HAS_COMPONENT " << (f.component ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_STOCH_COMPONENT " << (f.stoch_component ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_BAND " << (f.jacobian_band ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_BAND_I " << (f.jacobian_band_i ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN " << (f.calc_jacobian ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_VJP_GROUP " << (f.vjp_group ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_PARAMS " << (f.jacobian_params ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_EXACT " << (f.jacobian_exact ? 1 : 0) << "\n";
    o << "#define SODECL_USE_PAIRS " << (k.schedule == 2 && f.pairs ? k.use_pairs : 0) << "\n";
    if (k.one_wave_per_simd) o << "#define SODECL_KERNEL_ATTR __attribute__((amdgpu_waves_per_eu(1, 1)))\n";
    o << k.extra_defines;
    // OpenCL C spellings a system source may use
    for (const char* kw : {"__kernel", "__global", "__private"})
        o << "#ifndef " << kw << "\n#define " << kw << "\n#endif\n";
    o << "#ifndef __local\n#define __local\n#endif\n#ifndef __constant\n#define __constant const\n#endif\n";
    o << "typedef unsigned long ulong;\ntypedef unsigned int uint;\ntypedef unsigned short ushort;\n"
         "typedef unsigned char uchar;\n";
    o << "#ifndef M_PI\n#define M_PI 3.14159265358979323846\n#endif\n";
    o << "#ifndef M_PI_F\n#define M_PI_F 3.14159265358979323846f\n#endif\n";
    o << kPreludeSource << "\n";
    // The plugin's functions become host+device functions.
    o << "#pragma clang force_cuda_host_device begin\n";
    o << system_begin(k);
    o << "#line 1 \"system.cl\"\n";
    o << system_source << "\n";
    o << system_end(k);
    o << "#pragma clang force_cuda_host_device end\n";
    o << "#line 1 \"integrators.hip\"\n";
    o << kIntegratorsSource << "\n";
    return o.str();
}

std::string assemble_host_program(const std::string& system_source, const ProgramKey& k, const PluginFeatures& f)
{
    std::ostringstream o;
    o << "// SODECL host-share program (the lane integrator for the host CPU)\n";
    o << "#include <math.h>\n#include <stddef.h>\n#include <string.h>\n";
    o << "#define _numnoi_ " << k.nnoi << "\n";
    o << "#define _listsize_ " << k.orbits << "\n";
    o << "#define _numsteps_ " << k.ksteps << "\n";
    o << "#define _m_dt_ " << dt_literal(k.dt) << "\n";
    o << "#define _numeq_ " << k.neq << "\n";
    o << "#define _numpar_ " << k.npar << "\n";
    if (k.solver == SODECL_IMPLICIT_EULER || k.solver == SODECL_IMPLICIT_MIDPOINT) o << epsilon_define(k);
    o << "#define SODECL_HOST_BUILD 1\n";
    o << "#define SODECL_STRICT " << (k.fp_contract ? 1 : 0) << "\n";
    o << "#define SODECL_EXP_IN_LDS 0\n";
    o << "#define SODECL_SOLVER " << k.solver << "\n";
    o << "#define SODECL_SCHEDULE 1\n#define SODECL_GROUP 1\n#define SODECL_BLOCK 1\n";
    o << "#define SODECL_NEWTON_MAXIT " << k.newton_max_iter << "\n";
    o << "#define SODECL_HAS_COMPONENT " << (f.component ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_STOCH_COMPONENT " << (f.stoch_component ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_BAND " << (f.jacobian_band ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN_BAND_I " << (f.jacobian_band_i ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_JACOBIAN " << (f.calc_jacobian ? 1 : 0) << "\n";
    o << "#define SODECL_HAS_VJP_GROUP 0\n#define SODECL_HAS_JACOBIAN_PARAMS 0\n#define SODECL_HAS_JACOBIAN_EXACT 0\n";
    o << "#define SODECL_USE_PAIRS 0\n";
    o << "#define SODECL_INLINE static inline\n#define SODECL_MATH_FN static inline\n";
    for (const char* kw : {"__kernel", "__global", "__private"})
        o << "#ifndef " << kw << "\n#define " << kw << "\n#endif\n";
    o << "#ifndef __local\n#define __local\n#endif\n#ifndef __constant\n#define __constant const\n#endif\n";
    o << "typedef unsigned long ulong;\ntypedef unsigned int uint;\ntypedef unsigned short ushort;\n"
         "typedef unsigned char uchar;\n";
    o << "#ifndef M_PI\n#define M_PI 3.14159265358979323846\n#endif\n";
    o << kHostMathSource << "\n";
    o << system_begin(k);
    o << "#line 1 \"system.cl\"\n";
    o << system_source << "\n";
    o << system_end(k);
    o << "#line 1 \"integrators.hip\"\n";
    o << kHostLaneSource << "\n";
    return o.str();
}

std::string assemble_batch_program(const std::string& system_source, const ProgramKey& k, const PluginFeatures& f,
                                   const std::vector<int>& class_neq)
{
    std::ostringstream o;
    o << "// SODECL batched program assembled for hipRTC\n";
    o << "#define _numnoi_ 0\n";
    o << "#define _listsize_ " << k.orbits << "\n";
    o << "#define _numsteps_ " << k.ksteps << "\n";
    o << "#define _m_dt_ " << dt_literal(k.dt) << "\n";
    o << "#define _numpar_ " << k.npar << "\n";
    o << epsilon_define(k);
    o << "#define SODECL_HIP 1\n";
    o << "#define SODECL_STRICT " << (k.fp_contract ? 1 : 0) << "\n";
    o << "#define SODECL_EXP_IN_LDS " << (!k.library_exp && f.calls_exp ? 1 : 0) << "\n";
    o << "#define SODECL_SOLVER " << k.solver << "\n";
    o << "#define SODECL_SCHEDULE 2\n";
    o << "#define SODECL_GROUP 64\n";
    o << "#define SODECL_BG 64\n";
    o << "#define SODECL_BLOCK 64\n";
    o << "#define SODECL_NEWTON_MAXIT " << k.newton_max_iter << "\n";
    o << "#define SODECL_INLINE __device__ __attribute__((always_inline)) inline\n";
    for (const char* kw : {"__kernel", "__global", "__private"})
        o << "#ifndef " << kw << "\n#define " << kw << "\n#endif\n";
    o << "#ifndef __local\n#define __local\n#endif\n#ifndef __constant\n#define __constant const\n#endif\n";
    o << "typedef unsigned long ulong;\ntypedef unsigned int uint;\ntypedef unsigned short ushort;\n"
         "typedef unsigned char uchar;\n";
    o << "#ifndef M_PI\n#define M_PI 3.14159265358979323846\n#endif\n";
    o << kPreludeSource << "\n";
    // one namespace per size class: the plugin and the per-class solver
    for (size_t c = 0; c < class_neq.size(); ++c) {
        o << "namespace sodecl_cls" << c << " {\n";
        o << "#define _numeq_ " << class_neq[c] << "\n";
        o << "#define SODECL_LPO " << batch_lanes_per_orbit(class_neq[c]) << "\n";
        o << "#pragma clang force_cuda_host_device begin\n";
        o << system_begin(k) << "#line 1 \"system.cl\"\n" << system_source << "\n" << system_end(k);
        o << "#pragma clang force_cuda_host_device end\n";
        o << kBatchClassSource << "\n";
        o << "#undef _numeq_\n#undef SODECL_LPO\n";
        o << "}  // namespace sodecl_cls" << c << "\n";
        // the class's own kernel
        o << "#define SODECL_BATCH_KERNEL sodecl_batch_" << c << "\n";
        o << "#define SODECL_BATCH_ITEM sodecl_cls" << c << "::sodecl_batch_item\n";
        o << "#define SODECL_BATCH_WAVES " << batch_waves_per_simd(class_neq[c]) << "\n";
        o << kBatchKernelSource << "\n";
        o << "#undef SODECL_BATCH_KERNEL\n#undef SODECL_BATCH_ITEM\n#undef SODECL_BATCH_WAVES\n";
    }

    return o.str();
}

hipFunction_t Program::get(const std::string& name)
{
    auto it = functions.find(name);
    if (it != functions.end()) return it->second;
    hipFunction_t fn = nullptr;
    if (module && hipModuleGetFunction(&fn, module, name.c_str()) != hipSuccess) fn = nullptr;
    functions[name] = fn;
    return fn;
}

namespace {

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull)
{
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string cache_dir()
{
    if (const char* e = std::getenv("SODECL_AMD_CACHE_DIR")) return e;
    Dl_info info;
    if (dladdr((void*)&fnv1a, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        auto pos = p.rfind('/');
        if (pos != std::string::npos) return p.substr(0, pos) + "/_jit_cache";
    }
    return "/tmp/sodecl_amd_jit_cache";
}

std::mutex g_cache_mu;
std::map<uint64_t, std::vector<char>> g_mem_cache;

bool read_file(const std::string& path, std::vector<char>& data)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    data.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !data.empty();
}

void write_file_atomic(const std::string& dir, const std::string& name, const std::vector<char>& data)
{
    mkdir(dir.c_str(), 0755);
    std::string tmp = dir + "/." + name + "." + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), (std::streamsize)data.size());
    }
    std::rename(tmp.c_str(), (dir + "/" + name).c_str());
}

}  // namespace

std::string device_arch(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "gfx950";
    std::string a(prop.gcnArchName);
    auto pos = a.find(':');
    if (pos != std::string::npos) a = a.substr(0, pos);
    return a.empty() ? "gfx950" : a;
}

bool compile_program(const std::string& text, const std::string& arch, int fp_contract, std::vector<char>& code,
                     std::string& log, bool ilp_schedule)
{
    int vmaj = 0, vmin = 0;
    hiprtcVersion(&vmaj, &vmin);
    std::vector<std::string> opts = {"--gpu-architecture=" + arch, "-O3", "-std=c++17",
                                     // contraction marked on the program's own operations only:
                                     // "fast" would also fuse products and sums inside the device
                                     // library's functions, whose double-double steps need them
                                     // separate (its sine of 1e22 then comes out 0.02 off)
                                     fp_contract ? "-ffp-contract=off" : "-ffp-contract=fast-honor-pragmas",
                                     // no packed FP32 instructions in any generated code
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    // One wavefront per SIMD leaves no other wavefront to hide latency
    // behind: the ILP scheduling strategy interleaves independent chains
    // (measured 3-4% on 65536-orbit oculomotor RK4, 2-3% on 1024-orbit
    // Kuramoto Euler).
    // Extra compiler options, whitespace-separated (part of the cache key);
    // a scheduling strategy among them replaces this one.
    const char* extra = std::getenv("SODECL_AMD_JIT_OPTIONS");
    if (ilp_schedule && !(extra && std::strstr(extra, "-amdgpu-sched-strategy="))) {
        opts.push_back("-mllvm");
        opts.push_back("-amdgpu-sched-strategy=max-ilp");
    }
    if (extra) {
        std::istringstream in(extra);
        for (std::string w; in >> w;) opts.push_back(w);
    }
    std::string okey;
    for (auto& s : opts) okey += s + " ";
    okey += "hiprtc" + std::to_string(vmaj) + "." + std::to_string(vmin);
    const uint64_t h = fnv1a(okey, fnv1a(text));
    char hname[40];
    std::snprintf(hname, sizeof hname, "%016llx.hsaco", (unsigned long long)h);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_mem_cache.find(h);
        if (it != g_mem_cache.end()) {
            code = it->second;
            return true;
        }
    }
    const std::string dir = cache_dir();
    if (!std::getenv("SODECL_AMD_NO_DISK_CACHE") && read_file(dir + "/" + hname, code)) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_mem_cache[h] = code;
        return true;
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, text.c_str(), "sodecl_program.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    std::vector<const char*> copts;
    for (auto& s : opts) copts.push_back(s.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t lsz = 0;
    hiprtcGetProgramLogSize(prog, &lsz);
    log.assign(lsz, '\0');
    if (lsz) hiprtcGetProgramLog(prog, &log[0]);
    while (!log.empty() && log.back() == '\0') log.pop_back();
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t csz = 0;
    hiprtcGetCodeSize(prog, &csz);
    code.resize(csz);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_mem_cache[h] = code;
    }
    if (!std::getenv("SODECL_AMD_NO_DISK_CACHE")) write_file_atomic(dir, hname, code);
    return true;
}

}  // namespace sodecl_amd
