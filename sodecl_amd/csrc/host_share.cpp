// Host share of a run (sodecl_options.host_share; no reference counterpart:
// the reference names a CPU+GPU split of the orbits as future work,
// README.md:147).  Orbits [n_gpu, orbits) of a sodecl_run are integrated on
// the host cores while the GPU integrates [0, n_gpu), with the product's own
// lane integrator and the user's system source compiled for the host: the
// program text the JIT builds for the GPU, restricted to the lane schedule's
// per-orbit function (assemble_host_program), built by the host's C++
// compiler into a shared library at set-up, cached next to the code-object
// cache and loaded with dlopen.
#include <dlfcn.h>
#include <fcntl.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "sodecl_internal.hpp"

extern char** environ;

namespace sodecl_amd {

namespace {

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull)
{
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string host_cache_dir()
{
    if (const char* e = std::getenv("SODECL_AMD_CACHE_DIR")) return e;
    Dl_info info;
    if (dladdr((void*)&host_cache_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        auto pos = p.rfind('/');
        if (pos != std::string::npos) return p.substr(0, pos) + "/_jit_cache";
    }
    return "/tmp/sodecl_amd_jit_cache";
}

bool file_exists(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && st.st_size > 0;
}

// Runs argv as a child process (no shell); stdout and stderr go to log_path.
int run_child(const std::vector<std::string>& argv, const std::string& log_path)
{
    std::vector<char*> av;
    for (auto& a : argv) av.push_back(const_cast<char*>(a.c_str()));
    av.push_back(nullptr);
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_addopen(&fa, 1, log_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    posix_spawn_file_actions_adddup2(&fa, 1, 2);
    pid_t pid = 0;
    const int e = posix_spawnp(&pid, av[0], &fa, nullptr, av.data(), environ);
    posix_spawn_file_actions_destroy(&fa);
    if (e != 0) return -1;
    int status = 0;
    if (waitpid(pid, &status, 0) < 0) return -1;
    return WIFEXITED(status) ? WEXITSTATUS(status) : -1;
}

}  // namespace

// Builds (or finds in the cache) the host-share library of this context's
// program; path receives its file name.
bool build_host_library(const sodecl_ctx* c, std::string& path, std::string& err)
{
    ProgramKey key = make_key(c, 1, 1, 1, std::string());
    const std::string text = assemble_host_program(c->system_source, key, c->features);
    const char* cxx_env = std::getenv("SODECL_AMD_HOST_CXX");
    const std::string cxx = cxx_env && *cxx_env ? cxx_env : "c++";
    // Same arithmetic as the GPU's: contraction in the default mode, none in
    // strict mode (fp_contract = 1).  x86-64-v3 (AVX2, FMA): a cached library
    // stays valid on every host that shares the cache directory.
    std::vector<std::string> flags = {"-O3", "-march=x86-64-v3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-w",
                                      c->opts.fp_contract ? "-ffp-contract=off" : "-ffp-contract=fast"};
    std::string fkey = cxx;
    for (auto& f : flags) fkey += " " + f;
    const uint64_t h = fnv1a(fkey, fnv1a(text));
    char name[48];
    std::snprintf(name, sizeof name, "host_%016llx", (unsigned long long)h);
    const std::string dir = host_cache_dir();
    mkdir(dir.c_str(), 0755);
    path = dir + "/" + name + ".so";
    if (file_exists(path)) return true;
    const std::string tag = "." + std::string(name) + "." + std::to_string(getpid());
    const std::string src = dir + "/" + tag + ".cpp", tmp = dir + "/" + tag + ".so", log = dir + "/" + tag + ".log";
    {
        std::ofstream f(src);
        if (!f) {
            err = "cannot write " + src;
            return false;
        }
        f << text;
    }
    std::vector<std::string> argv = {cxx};
    argv.insert(argv.end(), flags.begin(), flags.end());
    argv.insert(argv.end(), {"-x", "c++", src, "-o", tmp, "-lm"});
    const int rc = run_child(argv, log);
    if (rc != 0 || !file_exists(tmp)) {
        std::ifstream lf(log);
        std::stringstream ss;
        ss << lf.rdbuf();
        err = "host-share build failed (" + cxx + ", exit " + std::to_string(rc) + "):\n" + ss.str();
        std::remove(tmp.c_str());
        return false;
    }
    std::rename(tmp.c_str(), path.c_str());
    std::remove(src.c_str());
    std::remove(log.c_str());
    return true;
}

bool load_host_share(sodecl_ctx* c, std::string& err)
{
    if (c->host_fn) return true;
    std::string path;
    if (!build_host_library(c, path, err)) return false;
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        err = std::string("dlopen(host share): ") + dlerror();
        return false;
    }
    void* fn = dlsym(h, "sodecl_host_lane");
    if (!fn) {
        err = "sodecl_host_lane missing from the host-share library";
        dlclose(h);
        return false;
    }
    c->host_handle = h;
    c->host_fn = (host_lane_fn)fn;
    return true;
}

void unload_host_share(sodecl_ctx* c)
{
    if (c->host_handle) dlclose(c->host_handle);
    c->host_handle = nullptr;
    c->host_fn = nullptr;
}

}  // namespace sodecl_amd

extern "C" size_t sodecl_host_library(sodecl_ctx* c, char* path, size_t capacity)
{
    if (!c) return 0;
    std::string p, err;
    if (!sodecl_amd::build_host_library(c, p, err)) {
        sodecl_amd::fail(c, err);
        return 0;
    }
    if (path && capacity > p.size()) std::snprintf(path, capacity, "%s", p.c_str());
    return p.size();
}
