// adjoint_host.cpp -- the discrete adjoint's own code (dugksfoam_amd/csrc/adjoint_core.h and
// adjoint_driver.h) instantiated with a host backend: every phase runs as plain loops over points
// and velocities.  Test infrastructure only: it lets the CPU test suite check the adjoint's
// arithmetic, indexing and checkpointed sweep against central differences of the oracle without a
// GPU.  The product library never contains or loads it.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../dugksfoam_amd/csrc/adjoint_driver.h"

namespace {

struct CX {
    template <class F>
    void for_k(int nv, F fn)
    {
        for (int k = 0; k < nv; ++k) fn(k);
    }
    template <int N>
    void sum(double *)
    {
    }
    bool leader() const { return true; }
};

struct HostBackend {
    bool bad = false;
    std::string msg;
    std::vector<void *> owned;
    std::chrono::steady_clock::time_point t[4];

    ~HostBackend()
    {
        for (void *p : owned) std::free(p);
    }
    template <class T>
    T *alloc(size_t n)
    {
        void *p = std::calloc(n ? n : 1, sizeof(T));
        if (!p) {
            bad = true;
            msg = "calloc";
            return nullptr;
        }
        owned.push_back(p);
        return (T *)p;
    }
    void release(void *p)
    {
        for (size_t i = 0; i < owned.size(); ++i)
            if (owned[i] == p) {
                std::free(p);
                owned.erase(owned.begin() + i);
                return;
            }
    }
    void upload(void *d, const void *h, size_t b) { std::memcpy(d, h, b); }
    void download(void *h, const void *d, size_t b) { std::memcpy(h, d, b); }
    void copy(void *d, const void *s, size_t b) { std::memmove(d, s, b); }
    void zero(void *d, size_t b) { std::memset(d, 0, b); }
    template <class Tag, class Fn>
    void launch(int n, Fn fn)
    {
        for (int i = 0; i < n; ++i) {
            CX x;
            fn(x, i);
        }
    }
    double free_bytes() { return 8.0e9; }
    void tic(int i) { t[i] = std::chrono::steady_clock::now(); }
    double ms(int a, int b) { return std::chrono::duration<double, std::milli>(t[b] - t[a]).count(); }
    void sync() {}
    void clear_error()
    {
        bad = false;
        msg.clear();
    }
};

}  // namespace

struct adjoint_host {
    adj::Engine<HostBackend> e;
};

extern "C" {

int adjoint_host_create(const dugks_case *cs, adjoint_host **out)
{
    adjoint_host *a = new (std::nothrow) adjoint_host();
    *out = a;
    if (!a) return DUGKS_ERR_MEMORY;
    return a->e.create(cs);
}

int adjoint_host_run(adjoint_host *a, const double *rho, const double *U, const double *T, double dt, int n_steps,
                     int first, int every, const double *coef, double *J, double *dJdT)
{
    return a->e.run(rho, U, T, dt, n_steps, first, every, coef, J, dJdT);
}

int adjoint_host_macro(adjoint_host *a, double *rho, double *U, double *T, double *q)
{
    return a->e.macro(rho, U, T, q);
}

const char *adjoint_host_error(adjoint_host *a) { return a->e.err.c_str(); }

void adjoint_host_destroy(adjoint_host *a) { delete a; }

}  // extern "C"
