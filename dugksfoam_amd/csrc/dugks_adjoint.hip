// dugks_adjoint.hip -- discrete adjoint of the DUGKS step on MI355X (gfx950), behind the
// dugks_adjoint_* entry points of include/dugks.h.
//
// The fused step engine (dugks_engine.hip) never materialises gbar+, its gradients or the face
// values of a step, and the reverse sweep needs all of them, so the adjoint runs its own FP64 step
// in a plain materialised form: one kernel per phase, one workgroup per cell or face, the lanes
// strided over the velocities and every velocity sum a deterministic workgroup reduction.  The
// per-point work lives in adjoint_core.h and the sweep in adjoint_driver.h; this file supplies the
// HIP backend: device memory, one stream (default priority), HIP events, and the kernel that runs a
// phase over its points.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "adjoint_driver.h"

namespace {

#ifndef DUGKS_ADJ_NT
#define DUGKS_ADJ_NT 128
#endif
constexpr int NT = DUGKS_ADJ_NT;     // lanes per workgroup (a multiple of the 64-lane wavefront)
constexpr int SUM_MAX = 32;          // widest workgroup reduction of adjoint_core.h

// Executor of adjoint_core.h inside a workgroup.
struct GX {
    double *lds;

    template <class F>
    __host__ __device__ void for_k(int nv, F fn)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        for (int k = threadIdx.x; k < nv; k += NT) fn(k);
#else
        (void)nv;
        (void)fn;
#endif
    }

    // Deterministic workgroup sum of v[N], result in every lane: butterfly within each wavefront,
    // then the wavefront totals added in order.
    template <int N>
    __host__ __device__ void sum(double *v)
    {
        static_assert(N <= SUM_MAX, "reduction width");
#if defined(__HIP_DEVICE_COMPILE__)
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double x = v[i];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            v[i] = x;
        }
        __syncthreads();
        if (lane == 0)
#pragma unroll
            for (int i = 0; i < N; ++i) lds[wave * N + i] = v[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) s += lds[w * N + i];
            v[i] = s;
        }
#else
        (void)v;
#endif
    }

    __host__ __device__ bool leader() const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return threadIdx.x == 0;
#else
        return true;
#endif
    }
};

template <class Tag, class Fn>
__global__ __launch_bounds__(NT) void k_points(Fn fn)
{
    __shared__ double lds[(NT / 64) * SUM_MAX];
    GX x{lds};
    fn(x, (int)blockIdx.x);
}

struct HipBackend {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool bad = false;
    std::string msg;
    std::vector<void *> owned;

    bool check(hipError_t e, const char *what)
    {
        if (e != hipSuccess && !bad) {
            bad = true;
            msg = std::string(what) + ": " + hipGetErrorString(e);
        }
        return e == hipSuccess;
    }

    int init(int dev)
    {
        device = dev;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DUGKS_ERR_DEVICE;
        if (dev < 0 || dev >= n) return DUGKS_ERR_INVALID;
        if (!check(hipSetDevice(dev), "hipSetDevice")) return DUGKS_ERR_DEVICE;
        if (!check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate")) return DUGKS_ERR_DEVICE;
        for (auto &e : ev)
            if (!check(hipEventCreate(&e), "hipEventCreate")) return DUGKS_ERR_DEVICE;
        return DUGKS_OK;
    }

    ~HipBackend()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : owned) (void)hipFree(p);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    T *alloc(size_t n)
    {
        void *p = nullptr;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (!check(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
        owned.push_back(p);
        check(hipMemsetAsync(p, 0, bytes, stream), "hipMemsetAsync");
        return (T *)p;
    }

    void release(void *p)
    {
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (size_t i = 0; i < owned.size(); ++i)
            if (owned[i] == p) {
                check(hipFree(p), "hipFree");
                owned.erase(owned.begin() + i);
                return;
            }
    }

    void upload(void *d, const void *h, size_t bytes)
    {
        if (bytes) check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream), "upload");
        check(hipStreamSynchronize(stream), "upload");
    }
    void download(void *h, const void *d, size_t bytes)
    {
        if (bytes) check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream), "download");
        check(hipStreamSynchronize(stream), "download");
    }
    void copy(void *d, const void *s, size_t bytes)
    {
        if (bytes) check(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, stream), "copy");
    }
    void zero(void *d, size_t bytes)
    {
        if (bytes) check(hipMemsetAsync(d, 0, bytes, stream), "memset");
    }

    template <class Tag, class Fn>
    void launch(int n, Fn fn)
    {
        if (n <= 0 || bad) return;
        hipLaunchKernelGGL((k_points<Tag, Fn>), dim3(n), dim3(NT), 0, stream, fn);
        check(hipGetLastError(), "kernel launch");
    }

    double free_bytes()
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0.0;
        return (double)fr;
    }

    void tic(int i) { check(hipEventRecord(ev[i], stream), "hipEventRecord"); }
    double ms(int a, int b)
    {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) return 0.0;
        return t;
    }
    void sync() { check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
    void clear_error()
    {
        (void)hipGetLastError();
        bad = false;
        msg.clear();
    }
};

}  // namespace

struct dugks_adjoint {
    adj::Engine<HipBackend> e;
};

extern "C" {

int dugks_adjoint_create(const dugks_case *cs, dugks_adjoint **out)
{
    if (!out) return DUGKS_ERR_INVALID;
    *out = nullptr;
    if (!cs) return DUGKS_ERR_INVALID;
    dugks_adjoint *a = new (std::nothrow) dugks_adjoint();
    if (!a) return DUGKS_ERR_MEMORY;
    int rc = a->e.be.init(cs->device);
    if (rc != DUGKS_OK) {
        a->e.err = "no usable HIP device: " + a->e.be.msg;
    } else {
        rc = a->e.create(cs);
    }
    *out = a;   // returned also on failure so that dugks_adjoint_last_error can explain it
    return rc;
}

int dugks_adjoint_run(dugks_adjoint *a, const double *rho, const double *U, const double *T, double dt,
                      int32_t n_steps, int32_t first_step, int32_t checkpoint_every, const double *face_coef,
                      double *J, double *dJ_dTw)
{
    if (!a) return DUGKS_ERR_INVALID;
    if (a->e.be.bad) return a->e.fail(DUGKS_ERR_DEVICE, "device error: " + a->e.be.msg);
    if (!a->e.p.xi) return a->e.fail(DUGKS_ERR_STATE, "adjoint was not created");
    if (hipSetDevice(a->e.be.device) != hipSuccess) return a->e.fail(DUGKS_ERR_DEVICE, "hipSetDevice failed");
    return a->e.run(rho, U, T, dt, n_steps, first_step, checkpoint_every, face_coef, J, dJ_dTw);
}

int dugks_adjoint_get_macro(dugks_adjoint *a, double *rho, double *U, double *T, double *q)
{
    if (!a) return DUGKS_ERR_INVALID;
    return a->e.macro(rho, U, T, q);
}

int dugks_adjoint_last_ms(dugks_adjoint *a, double *forward_ms, double *reverse_ms)
{
    if (!a) return DUGKS_ERR_INVALID;
    if (!a->e.have_run) return a->e.fail(DUGKS_ERR_STATE, "no run yet");
    if (forward_ms) *forward_ms = a->e.fwd_ms;
    if (reverse_ms) *reverse_ms = a->e.rev_ms;
    return DUGKS_OK;
}

void dugks_adjoint_destroy(dugks_adjoint *a) { delete a; }

const char *dugks_adjoint_last_error(dugks_adjoint *a)
{
    if (!a) return "null adjoint handle";
    return a->e.err.c_str();
}

}  // extern "C"
