// Python module sodecl_interface: the reference's pybind11 entry point
// (interfaces/python/sodecl.cpp:16-101 of avramidis/sodecl) over the C ABI,
// plus a persistent Session (program compiled once, reused across calls) and
// a device-pointer entry for PyTorch tensors.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/sodecl.h"

namespace py = pybind11;
using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

void apply_options(sodecl_options& o, const py::dict& d)
{
    for (auto kv : d) {
        std::string k = py::str(kv.first);
        if (k == "schedule") {
            if (py::isinstance<py::str>(kv.second)) {
                std::string v = py::str(kv.second);
                o.schedule = v == "lane" ? SODECL_SCHEDULE_LANE
                             : v == "group" ? SODECL_SCHEDULE_GROUP
                             : v == "plain" ? SODECL_SCHEDULE_PLAIN
                             : v == "team"  ? SODECL_SCHEDULE_TEAM
                                            : SODECL_SCHEDULE_AUTO;
            } else {
                o.schedule = kv.second.cast<int>();
            }
        } else if (k == "group_size") o.group_size = kv.second.cast<int>();
        else if (k == "block_size") o.block_size = kv.second.cast<int>();
        else if (k == "newton_max_iter") o.newton_max_iter = kv.second.cast<int>();
        else if (k == "segments_per_launch") o.segments_per_launch = kv.second.cast<int>();
        else if (k == "use_graph") o.use_graph = kv.second.cast<int>();
        else if (k == "autotune") o.autotune = kv.second.cast<int>();
        else if (k == "fp_contract") o.fp_contract = kv.second.cast<int>();
        else if (k == "verbose") o.verbose = kv.second.cast<int>();
        else if (k == "traj_budget_bytes") o.traj_budget_bytes = kv.second.cast<size_t>();
        else if (k == "pairwise") o.pairwise = kv.second.cast<int>();
        else if (k == "newton_tol") o.newton_tol = kv.second.cast<double>();
        else if (k == "library_sin") o.library_sin = kv.second.cast<int>();
        else if (k == "library_exp") o.library_exp = kv.second.cast<int>();
        else if (k == "host_share") o.host_share = kv.second.cast<double>();
        else throw std::invalid_argument("unknown option " + k);
    }
}

// sodecl.cpp:16-95
py::array_t<double> sodeclcall(darr a_t0, darr a_y0, darr a_params, int a_platform, int a_device, std::string a_system,
                               int a_solver_type, int a_orbits, int a_equats, int a_nparams, int a_nnoi, double a_dt,
                               double a_tspan, int a_ksteps, int a_local_group_size)
{
    if (a_solver_type < 0 || a_solver_type > 4) {
        std::cout << "Unknown SODE solver selection." << std::endl;
        return py::array_t<double>(0);
    }
    if ((size_t)a_y0.size() < (size_t)a_orbits * a_equats || (size_t)a_params.size() < (size_t)a_orbits * a_nparams ||
        (size_t)a_t0.size() < (size_t)a_orbits)
        throw std::invalid_argument("t0, y0 or params is smaller than orbits x size");
    sodecl_ctx* ctx = nullptr;
    if (!sodecl_create(&ctx, "kernels", a_system.c_str(), a_solver_type, a_dt, a_tspan, a_ksteps, a_equats, a_nparams,
                       a_nnoi, a_orbits, SODECL_OUTPUT_ARRAY)) {
        std::cout << sodecl_last_error(ctx) << std::endl;
        sodecl_destroy(ctx);
        return py::array_t<double>(0);
    }
    if (!sodecl_choose_device(ctx, (uint32_t)a_platform, SODECL_DEVICE_TYPE_ALL, (uint32_t)a_device))
        std::cerr << "Error selecting OpenCL device!" << std::endl;
    sodecl_set_t0(ctx, a_t0.data());
    sodecl_set_y0(ctx, a_y0.data());
    sodecl_set_params(ctx, a_params.data());
    sodecl_set_local_group_size(ctx, a_local_group_size);
    const long launches = sodecl_num_launches(a_tspan, a_dt, a_ksteps);
    const size_t total = (size_t)(launches + 1) * (size_t)a_equats * (size_t)a_orbits;
    py::array_t<double> out(total);
    int ok;
    {
        py::gil_scoped_release nogil;
        ok = sodecl_setup(ctx);
        if (ok) {
            sodecl_set_output_buffer(ctx, out.mutable_data(), total);
            ok = sodecl_run(ctx);
        }
    }
    if (!ok) {
        std::cout << sodecl_last_error(ctx) << std::endl;
        sodecl_destroy(ctx);
        return py::array_t<double>(0);
    }
    sodecl_destroy(ctx);
    return out;
}

// A manager object kept alive across calls: the program is compiled once.
class Session {
public:
    Session(const std::string& system, int solver, int orbits, int neq, int npar, int nnoi, double dt, double tspan,
            int ksteps, int device, py::dict options, bool system_is_source, int output_type)
        : orbits_(orbits), neq_(neq), npar_(npar), nnoi_(nnoi), dt_(dt), tspan_(tspan), ksteps_(ksteps),
          output_type_(output_type)
    {
        int ok = system_is_source
                     ? sodecl_create_from_source(&ctx_, system.c_str(), solver, dt, tspan, ksteps, neq, npar, nnoi,
                                                 orbits, output_type)
                     : sodecl_create(&ctx_, "kernels", system.c_str(), solver, dt, tspan, ksteps, neq, npar, nnoi,
                                     orbits, output_type);
        if (!ok) throw std::runtime_error(sodecl_last_error(ctx_));
        sodecl_default_options(&opts_);
        apply_options(opts_, options);
        if (!sodecl_set_options(ctx_, &opts_)) throw std::runtime_error(sodecl_last_error(ctx_));
        device_ = device;
        compiled_only_ = false;
    }
    ~Session() { sodecl_destroy(ctx_); }

    long launches() const { return sodecl_num_launches(tspan_, dt_, ksteps_); }
    long nsnap() const { return launches() + 1; }

    void setup(darr y0, darr params, py::object t0, py::object counters)
    {
        check(y0.size() >= (py::ssize_t)orbits_ * neq_ && params.size() >= (py::ssize_t)orbits_ * npar_,
              "y0/params too small");
        if (!sodecl_choose_device(ctx_, 0, SODECL_DEVICE_TYPE_ALL, (uint32_t)device_)) fail();
        y0_ = y0;
        params_ = params;
        sodecl_set_y0(ctx_, y0_.data());
        sodecl_set_params(ctx_, params_.data());
        if (!t0.is_none()) {
            t0_ = darr::ensure(t0);
            check(t0_.size() >= orbits_, "t0 too small");
            sodecl_set_t0(ctx_, t0_.data());
        } else {
            sodecl_set_t0(ctx_, nullptr);
        }
        if (!counters.is_none()) {
            darr c = darr::ensure(counters);
            check(c.size() >= orbits_, "counters too small");
            if (!sodecl_set_rng_counters(ctx_, c.data())) fail();
        }
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_setup(ctx_);
        }
        if (!ok) fail();
    }

    // Host run: returns the flat [snap][eq][orbit] output of this run.
    // The trajectory [nsnap][neq][orbits] (flat) with ARRAY output; an empty
    // array with NONE or FILE output.
    py::array_t<double> run()
    {
        if (output_type_ != SODECL_OUTPUT_ARRAY) {
            int ok;
            {
                py::gil_scoped_release nogil;
                ok = sodecl_run(ctx_);
            }
            if (!ok) fail();
            return py::array_t<double>(0);
        }
        const size_t total = (size_t)nsnap() * neq_ * orbits_;
        py::array_t<double> out(total);
        int ok;
        {
            py::gil_scoped_release nogil;
            sodecl_set_output_buffer(ctx_, out.mutable_data(), total);
            ok = sodecl_run(ctx_);
            sodecl_set_output_buffer(ctx_, nullptr, 0);
        }
        if (!ok) fail();
        return out;
    }

    // Device run on caller memory (integer device pointers, e.g. tensor.data_ptr()).
    void run_device(uintptr_t t0, uintptr_t y0, uintptr_t params, uintptr_t counters, uintptr_t traj, uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_run_device(ctx_, (double*)t0, (double*)y0, (const double*)params, (double*)counters,
                                   (double*)traj, (void*)stream);
        }
        if (!ok) fail();
    }

    void run_device_checkpointed(uintptr_t t0, uintptr_t y0, uintptr_t params, uintptr_t counters, uintptr_t traj,
                                 uintptr_t ckpt_t, uintptr_t ckpt_y, int ckpt_every, uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_run_device_checkpointed(ctx_, (double*)t0, (double*)y0, (const double*)params,
                                                (double*)counters, (double*)traj, (double*)ckpt_t, (double*)ckpt_y,
                                                ckpt_every, (void*)stream);
        }
        if (!ok) fail();
    }
    void run_device_stored(uintptr_t t0, uintptr_t y0, uintptr_t params, uintptr_t traj, uintptr_t work,
                           uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_run_device_stored(ctx_, (double*)t0, (double*)y0, (const double*)params, (double*)traj,
                                          (double*)work, (void*)stream);
        }
        if (!ok) fail();
    }
    void adjoint_device_stored(uintptr_t params, uintptr_t cot, uintptr_t grad_y0, uintptr_t grad_p, uintptr_t work,
                               uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_adjoint_device_stored(ctx_, (const double*)params, (const double*)cot, (double*)grad_y0,
                                              (double*)grad_p, (double*)work, (void*)stream);
        }
        if (!ok) fail();
    }
    size_t stored_workspace_size() const { return sodecl_stored_workspace_size(ctx_); }
    void adjoint_device(uintptr_t ckpt_t, uintptr_t ckpt_y, uintptr_t params, uintptr_t cot, uintptr_t grad_y0,
                        uintptr_t grad_p, uintptr_t work, int ckpt_every, uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_adjoint_device(ctx_, (const double*)ckpt_t, (const double*)ckpt_y, (const double*)params,
                                       (const double*)cot, (double*)grad_y0, (double*)grad_p, (double*)work,
                                       ckpt_every, (void*)stream);
        }
        if (!ok) fail();
    }
    int default_checkpoint_interval() const { return sodecl_default_checkpoint_interval(ctx_); }
    long checkpoint_count(int every) const { return sodecl_checkpoint_count(ctx_, every); }
    size_t adjoint_workspace_size(int every) const { return sodecl_adjoint_workspace_size(ctx_, every); }

    py::array_t<double> get_counters()
    {
        py::array_t<double> c(orbits_);
        if (!sodecl_get_rng_counters(ctx_, c.mutable_data())) fail();
        return c;
    }
    py::array_t<long long> newton_updates()
    {
        py::array_t<long long> c(orbits_);
        if (!sodecl_get_newton_updates(ctx_, c.mutable_data())) fail();
        return c;
    }
    py::tuple launch_config()
    {
        int s = 0, g = 0, b = 0;
        sodecl_get_launch_config(ctx_, &s, &g, &b);
        return py::make_tuple(s, g, b);
    }
    double last_kernel_ms() const { return sodecl_last_kernel_ms(ctx_); }
    py::tuple host_stats() const
    {
        int n = 0;
        const double ms = sodecl_last_host_ms(ctx_, &n);
        return py::make_tuple(ms, n);
    }
    double last_runtime() const { return sodecl_last_runtime(ctx_); }
    size_t compile_only(const std::string& arch)
    {
        size_t n = sodecl_compile_only(ctx_, arch.c_str());
        if (!n) fail();
        return n;
    }
    std::string program_source() { return sodecl_program_source(ctx_); }
    std::string autotune_report() const { return sodecl_autotune_report(ctx_); }
    void set_options(py::dict options)
    {
        apply_options(opts_, options);
        if (!sodecl_set_options(ctx_, &opts_)) fail();
    }
    uintptr_t handle() const { return (uintptr_t)ctx_; }

private:
    void check(bool cond, const char* msg)
    {
        if (!cond) throw std::invalid_argument(msg);
    }
    [[noreturn]] void fail() { throw std::runtime_error(sodecl_last_error(ctx_)); }

    sodecl_ctx* ctx_ = nullptr;
    sodecl_options opts_{};
    int orbits_, neq_, npar_, nnoi_;
    double dt_, tspan_;
    int ksteps_;
    int output_type_ = SODECL_OUTPUT_ARRAY;
    int device_ = 0;
    bool compiled_only_;
    darr y0_, params_, t0_;
};

// Batched implicit integration of orbits of several sizes (sodecl_batch_*).
class Batch {
public:
    Batch(const std::string& system, int solver, double dt, int ksteps, int npar, std::vector<int> class_neq,
          int device, py::dict options, bool system_is_source)
    {
        std::string src = system;
        if (!system_is_source) {
            std::ifstream f(system);
            if (!f) throw std::runtime_error("Error opening file " + system);
            std::ostringstream ss;
            ss << f.rdbuf();
            src = ss.str();
        }
        if (!sodecl_batch_create(&b_, src.c_str(), solver, dt, ksteps, npar, (int)class_neq.size(),
                                 class_neq.data()))
            fail();
        sodecl_options o;
        sodecl_default_options(&o);
        apply_options(o, options);
        sodecl_batch_set_options(b_, &o);
        device_ = device;
    }
    ~Batch() { sodecl_batch_destroy(b_); }
    void setup(py::array_t<int, py::array::c_style | py::array::forcecast> cls,
               py::array_t<int, py::array::c_style | py::array::forcecast> steps)
    {
        if (cls.size() != steps.size()) throw std::invalid_argument("classes and steps differ in length");
        if (!sodecl_batch_choose_device(b_, (uint32_t)device_)) fail();
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_batch_setup(b_, (int)cls.size(), cls.data(), steps.data());
        }
        if (!ok) fail();
        orbits_ = (int)cls.size();
    }
    py::tuple offsets()
    {
        py::array_t<long long> y(orbits_), o(orbits_);
        if (!sodecl_batch_offsets(b_, y.mutable_data(), o.mutable_data())) fail();
        return py::make_tuple(y, o);
    }
    size_t state_size() const { return sodecl_batch_state_size(b_); }
    size_t output_size() const { return sodecl_batch_output_size(b_); }
    void run_device(uintptr_t y0, uintptr_t params, uintptr_t out, uintptr_t newton, uintptr_t stream)
    {
        int ok;
        {
            py::gil_scoped_release nogil;
            ok = sodecl_batch_run_device(b_, (const double*)y0, (const double*)params, (double*)out,
                                         (long long*)newton, (void*)stream);
        }
        if (!ok) fail();
    }
    size_t compile_only(const std::string& arch)
    {
        size_t n = sodecl_batch_compile_only(b_, arch.c_str());
        if (!n) fail();
        return n;
    }
    std::string program_source() { return sodecl_batch_program_source(b_); }

private:
    [[noreturn]] void fail() { throw std::runtime_error(sodecl_batch_last_error(b_)); }
    sodecl_batch* b_ = nullptr;
    int device_ = 0;
    int orbits_ = 0;
};

}  // namespace

PYBIND11_MODULE(sodecl_interface, m)
{
    m.doc() = "sodecl plugin (MI355X)";
    m.def("sodeclcall", &sodeclcall, "A function that integrates a SDE/ODE system.", py::arg("t0"), py::arg("y0"),
          py::arg("params"), py::arg("platform"), py::arg("device"), py::arg("system"), py::arg("solver_type"),
          py::arg("orbits"), py::arg("equats"), py::arg("nparams"), py::arg("nnoi"), py::arg("dt"), py::arg("tspan"),
          py::arg("ksteps"), py::arg("local_group_size"));
    m.def("num_launches", &sodecl_num_launches);
    // dists: sequence of (kind, a, b); out: device pointer to rows x len(dists) doubles
    m.def(
        "sample_device",
        [](py::sequence dists, int rows, unsigned long long seed, uintptr_t out, uintptr_t stream) {
            std::vector<sodecl_dist> d;
            for (auto item : dists) {
                auto t = item.cast<py::tuple>();
                d.push_back(sodecl_dist{t[0].cast<int>(), t[1].cast<double>(), t[2].cast<double>()});
            }
            int ok;
            {
                py::gil_scoped_release nogil;
                ok = sodecl_sample_device(d.data(), rows, (int)d.size(), seed, (double*)out, (void*)stream);
            }
            if (!ok) throw std::runtime_error("sodecl_sample_device: invalid distribution table or launch failure");
        },
        py::arg("dists"), py::arg("rows"), py::arg("seed"), py::arg("out"), py::arg("stream") = 0);
    m.attr("ABI_VERSION") = SODECL_AMD_ABI_VERSION;
    py::class_<Session>(m, "Session")
        .def(py::init<const std::string&, int, int, int, int, int, double, double, int, int, py::dict, bool, int>(),
             py::arg("system"), py::arg("solver"), py::arg("orbits"), py::arg("neq"), py::arg("npar"),
             py::arg("nnoi"), py::arg("dt"), py::arg("tspan"), py::arg("ksteps"), py::arg("device") = 0,
             py::arg("options") = py::dict(), py::arg("system_is_source") = false,
             py::arg("output_type") = (int)SODECL_OUTPUT_ARRAY)
        .def("setup", &Session::setup, py::arg("y0"), py::arg("params"), py::arg("t0") = py::none(),
             py::arg("counters") = py::none())
        .def("run", &Session::run)
        .def("run_device", &Session::run_device, py::arg("t0"), py::arg("y0"), py::arg("params"),
             py::arg("counters"), py::arg("traj"), py::arg("stream") = 0)
        .def("run_device_checkpointed", &Session::run_device_checkpointed, py::arg("t0"), py::arg("y0"),
             py::arg("params"), py::arg("counters"), py::arg("traj"), py::arg("ckpt_t"), py::arg("ckpt_y"),
             py::arg("ckpt_every"), py::arg("stream") = 0)
        .def("adjoint_device", &Session::adjoint_device, py::arg("ckpt_t"), py::arg("ckpt_y"), py::arg("params"),
             py::arg("cot"), py::arg("grad_y0"), py::arg("grad_p"), py::arg("work"), py::arg("ckpt_every"),
             py::arg("stream") = 0)
        .def("default_checkpoint_interval", &Session::default_checkpoint_interval)
        .def("checkpoint_count", &Session::checkpoint_count)
        .def("adjoint_workspace_size", &Session::adjoint_workspace_size)
        .def("run_device_stored", &Session::run_device_stored, py::arg("t0"), py::arg("y0"), py::arg("params"),
             py::arg("traj"), py::arg("work"), py::arg("stream") = 0)
        .def("adjoint_device_stored", &Session::adjoint_device_stored, py::arg("params"), py::arg("cot"),
             py::arg("grad_y0"), py::arg("grad_p"), py::arg("work"), py::arg("stream") = 0)
        .def("stored_workspace_size", &Session::stored_workspace_size)
        .def("counters", &Session::get_counters)
        .def("newton_updates", &Session::newton_updates)
        .def("launch_config", &Session::launch_config)
        .def("last_kernel_ms", &Session::last_kernel_ms)
        .def("host_stats", &Session::host_stats)
        .def("last_runtime", &Session::last_runtime)
        .def("compile_only", &Session::compile_only, py::arg("arch") = "gfx950")
        .def("program_source", &Session::program_source)
        .def("autotune_report", &Session::autotune_report)
        .def("set_options", &Session::set_options)
        .def("handle", &Session::handle)
        .def_property_readonly("nsnap", &Session::nsnap)
        .def_property_readonly("launches", &Session::launches);
    py::class_<Batch>(m, "Batch")
        .def(py::init<const std::string&, int, double, int, int, std::vector<int>, int, py::dict, bool>(),
             py::arg("system"), py::arg("solver"), py::arg("dt"), py::arg("ksteps"), py::arg("npar"),
             py::arg("class_neq"), py::arg("device") = 0, py::arg("options") = py::dict(),
             py::arg("system_is_source") = false)
        .def("setup", &Batch::setup, py::arg("classes"), py::arg("steps"))
        .def("offsets", &Batch::offsets)
        .def("state_size", &Batch::state_size)
        .def("output_size", &Batch::output_size)
        .def("run_device", &Batch::run_device, py::arg("y0"), py::arg("params"), py::arg("out"), py::arg("newton"),
             py::arg("stream") = 0)
        .def("compile_only", &Batch::compile_only, py::arg("arch") = "gfx950")
        .def("program_source", &Batch::program_source)
        ;
}
