// abi.cpp -- extern "C" entry points of libradiant_hip.so (declared in include/radiant_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "engine.hpp"

namespace {
thread_local std::string g_last_error;

}  // namespace

void rt::set_last_error(const std::string& m) { g_last_error = m; }

namespace {
int fail(int code, const std::string& m) {
    g_last_error = m;
    return code;
}

template <class F>
int guarded(F f) {
    try {
        f();
        return RT_OK;
    } catch (const rt::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(RT_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(RT_EDEVICE, e.what());
    }
}
}  // namespace

struct rt_problem {
    std::unique_ptr<rt::Problem> p;
};

extern "C" {

const char* rt_last_error(void) { return g_last_error.c_str(); }

const char* rt_version(void) {
    static std::string v;
    if (v.empty()) {
        v = "radiant_hip 0.1";
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            v += std::string(" ") + prop.gcnArchName + " " + prop.name;
        } else {
            (void)hipGetLastError();
            v += " (no device)";
        }
    }
    return v.c_str();
}

int rt_problem_device_bytes(const rt_problem_desc* p, int64_t* bytes) {
    if (!p || !bytes) return fail(RT_EINVAL, "null argument");
    return guarded([&] { *bytes = rt::Problem::device_bytes(*p); });
}

int rt_problem_create(const rt_problem_desc* p, rt_problem** out) {
    if (!p || !out) return fail(RT_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto h = std::make_unique<rt_problem>();
        h->p = std::make_unique<rt::Problem>(*p);
        *out = h.release();
    });
}

int rt_problem_solve(rt_problem* h, rt_group_stats* stats_out, rt_solve_info* info) {
    if (!h || !h->p) return fail(RT_EINVAL, "null handle");
    return guarded([&] { h->p->solve(stats_out, info); });
}

int rt_problem_get_flux(rt_problem* h, double* phi_out) {
    if (!h || !h->p || !phi_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] { h->p->get_flux(phi_out); });
}

int rt_problem_get_flux_cutoff(rt_problem* h, double* cutoff_out) {
    if (!h || !h->p || !cutoff_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] { h->p->get_flux_cutoff(cutoff_out); });
}

int rt_problem_destroy(rt_problem* h) {
    if (!h) return RT_OK;
    return guarded([&] { delete h; });
}

int rt_problem_pass(rt_problem* h, int32_t g, int32_t mode, const double* qlout, const double* z_in, double* z_out) {
    if (!h || !h->p || !qlout || !z_in || !z_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] { h->p->host_pass(g, mode, qlout, z_in, z_out); });
}

int rt_problem_pass_device(rt_problem* h, int32_t g, int32_t mode, const double* qlout, const double* z_in,
                           double* z_out) {
    if (!h || !h->p || !qlout || !z_in || !z_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] {
        size_t cap = 0;
        double* scr = rt::noalloc_scratch(&cap);
        h->p->device_pass(g, mode, qlout, z_in, z_out, scr, cap);
    });
}

int rt_problem_state_length(rt_problem* h, int64_t* S) {
    if (!h || !h->p || !S) return fail(RT_EINVAL, "null argument");
    *S = h->p->faithful_state_length();
    return RT_OK;
}

int rt_test_contraction(int32_t which, int32_t prec, int32_t M, int32_t K, int32_t N, const double* A, const double* B,
                        double* C) {
    if (!A || !B || !C || M < 1 || K < 1 || N < 1 || which < 0 || which > 1 || prec < 0 || prec > 2)
        return fail(RT_EINVAL, "invalid contraction test arguments");
    return guarded([&] { rt::test_contraction(which, prec, M, K, N, A, B, C); });
}

int rt_autotune(int32_t mode) {
    if (mode < 0 || mode > 2) return fail(RT_EINVAL, "autotune mode must be 0, 1 or 2");
    rt::set_autotune_mode(mode);
    return RT_OK;
}

int64_t rt_autotune_report(char* buf, int64_t cap) {
    const std::string r = rt::autotune_report();
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>(r.size(), (size_t)cap - 1);
        std::memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return (int64_t)r.size();
}

int rt_problem_kernel_times(rt_problem* h, double* ms, int64_t* launches) {
    if (!h || !h->p || !ms || !launches) return fail(RT_EINVAL, "null argument");
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 4; ++k) {
            ms[p * 4 + k] = h->p->ktimes.ms_p[p][k];
            launches[p * 4 + k] = h->p->ktimes.n_p[p][k];
        }
    return RT_OK;
}

int rt_solve(const rt_problem_desc* p, double* phi_out, rt_group_stats* stats_out) {
    if (!p || !phi_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] {
        rt::Problem prob(*p);
        prob.solve(stats_out, nullptr);
        prob.get_flux(phi_out);
    });
}

int rt_solve_cutoff(const rt_problem_desc* p, double* phi_out, double* cutoff_out, rt_group_stats* stats_out) {
    if (!p || !phi_out) return fail(RT_EINVAL, "null argument");
    return guarded([&] {
        rt::Problem prob(*p);
        prob.solve(stats_out, nullptr);
        prob.get_flux(phi_out);
        if (cutoff_out) prob.get_flux_cutoff(cutoff_out);
    });
}

}  // extern "C"
