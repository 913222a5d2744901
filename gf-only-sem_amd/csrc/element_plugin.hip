// element_plugin.hip -- HIP backend of the reference's element plugin
// gf::compute_element_residual<Backend>(n_elem, dxi_dx, jacobian, lambda_,
// mu_, D, weights, NGLL, u, r) (forward/share/include/gf/element.hpp:39-42).
//
// Host pointers in the reference's element-major AoS layout; r accumulates
// r -= K u (forward/elastic/src/element_cpu.cpp:116).  One workgroup per
// element, one lane per GLL point; displacement and fluxes are staged in LDS,
// so every point's gradient and divergence are contractions over LDS rows
// instead of the 45 atomics per point of the CUDA kernel
// (forward/elastic/src/element_cuda.cu:131-183).
#include <cstring>
#include <mutex>

#include "common.hpp"

#define GFX_HIP(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            throw ::gfx::Error(GFX_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#include <hip/hip_runtime.h>

namespace gfx {
namespace {

template <typename T, int NGL>
__global__ void __launch_bounds__(((NGL * NGL * NGL + 63) / 64) * 64)
    k_elem_plugin(int n_elem, const T* __restrict__ dxi_dx, const T* __restrict__ jac,
                  const T* __restrict__ lam_, const T* __restrict__ mu_, const T* __restrict__ Dm,
                  const T* __restrict__ wq, const T* __restrict__ u, T* __restrict__ r) {
    constexpr int NN = NGL * NGL * NGL;
    __shared__ T sD[NGL * NGL];
    __shared__ T sw[NGL];
    __shared__ T su[3][NN];
    __shared__ T sf[9][NN];
    const int p = threadIdx.x;
    const int64_t e = blockIdx.x;
    if (p < NGL * NGL) sD[p] = Dm[p];
    if (p < NGL) sw[p] = wq[p];
    const bool valid = p < NN;
    const int i = p / (NGL * NGL), j = (p / NGL) % NGL, k = p % NGL;
    const int64_t pt = e * NN + (valid ? p : 0);
    if (valid)
        for (int c = 0; c < 3; ++c) su[c][p] = u[pt * 3 + c];
    __syncthreads();
    if (valid) {
        const T* dd = dxi_dx + pt * 9;
        const T lam = lam_[pt], mu = mu_[pt];
        T F[3][3];
        if (mu <= T(0)) {  // element_cpu.cpp:42-43
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) F[c][a] = T(0);
        } else {
            T du[3][3];
            for (int c = 0; c < 3; ++c) {
                T a0 = 0, a1 = 0, a2 = 0;
                for (int s = 0; s < NGL; ++s) {
                    a0 += sD[i * NGL + s] * su[c][(s * NGL + j) * NGL + k];
                    a1 += sD[j * NGL + s] * su[c][(i * NGL + s) * NGL + k];
                    a2 += sD[k * NGL + s] * su[c][(i * NGL + j) * NGL + s];
                }
                du[c][0] = a0;
                du[c][1] = a1;
                du[c][2] = a2;
            }
            T g[3][3];
            for (int c = 0; c < 3; ++c)
                for (int m = 0; m < 3; ++m)
                    g[c][m] = du[c][0] * dd[m * 3 + 0] + du[c][1] * dd[m * 3 + 1] +
                              du[c][2] * dd[m * 3 + 2];
            T sig[3][3];
            const T tr = g[0][0] + g[1][1] + g[2][2];
            for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) sig[l][m] = T(2) * mu * (T(0.5) * (g[l][m] + g[m][l]));
            for (int l = 0; l < 3; ++l) sig[l][l] += lam * tr;
            const T fac = jac[pt] * sw[i] * sw[j] * sw[k];
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a)
                    F[c][a] = (sig[c][0] * dd[0 + a] + sig[c][1] * dd[3 + a] + sig[c][2] * dd[6 + a]) *
                              fac;
        }
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) sf[c * 3 + a][p] = F[c][a];
    }
    __syncthreads();
    if (valid) {
        for (int c = 0; c < 3; ++c) {
            T acc = 0;
            for (int s = 0; s < NGL; ++s) {
                acc += sD[s * NGL + i] * sf[c * 3 + 0][(s * NGL + j) * NGL + k];
                acc += sD[s * NGL + j] * sf[c * 3 + 1][(i * NGL + s) * NGL + k];
                acc += sD[s * NGL + k] * sf[c * 3 + 2][(i * NGL + j) * NGL + s];
            }
            r[pt * 3 + c] -= acc;
        }
    }
}

template <typename T>
struct Dev {
    T* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    void put(const double* h, size_t n) {
        GFX_HIP(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)));
        if (n == 0) return;
        if (std::is_same<T, double>::value) {
            GFX_HIP(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
        } else {
            std::vector<T> tmp(h, h + n);
            GFX_HIP(hipMemcpy(p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
        }
    }
    void get(double* h, size_t n) const {
        if (n == 0) return;
        if (std::is_same<T, double>::value) {
            GFX_HIP(hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost));
        } else {
            std::vector<T> tmp(n);
            GFX_HIP(hipMemcpy(tmp.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
            for (size_t q = 0; q < n; ++q) h[q] = double(tmp[q]);
        }
    }
};

template <typename T, int NGL>
void launch_plugin(int n_elem, const T* dd, const T* jac, const T* lam, const T* mu, const T* D,
                   const T* w, const T* u, T* r) {
    constexpr int NT = ((NGL * NGL * NGL + 63) / 64) * 64;
    k_elem_plugin<T, NGL><<<n_elem, NT>>>(n_elem, dd, jac, lam, mu, D, w, u, r);
}

template <typename T>
void element_residual(int n_elem, const double* dxi_dx, const double* jacobian,
                      const double* lambda_, const double* mu_, const double* D,
                      const double* weights, int ngll, const double* u, double* r) {
    require(n_elem >= 0, "n_elem must be >= 0");
    require(ngll >= 2 && ngll <= 8, "NGLL must be in [2, 8]");
    if (n_elem == 0) return;
    require(dxi_dx && jacobian && lambda_ && mu_ && D && weights && u && r, "null array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        throw Error(GFX_ERR_DEVICE, "no HIP device available");
    }
    const size_t nn = size_t(ngll) * ngll * ngll, P = size_t(n_elem) * nn;
    Dev<T> ddd, dj, dl, dm, dD, dw, du, dr;
    ddd.put(dxi_dx, P * 9);
    dj.put(jacobian, P);
    dl.put(lambda_, P);
    dm.put(mu_, P);
    dD.put(D, size_t(ngll) * ngll);
    dw.put(weights, ngll);
    du.put(u, P * 3);
    dr.put(r, P * 3);  // accumulate into the caller's r
    switch (ngll) {
        case 2: launch_plugin<T, 2>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        case 3: launch_plugin<T, 3>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        case 4: launch_plugin<T, 4>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        case 5: launch_plugin<T, 5>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        case 6: launch_plugin<T, 6>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        case 7: launch_plugin<T, 7>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
        default: launch_plugin<T, 8>(n_elem, ddd.p, dj.p, dl.p, dm.p, dD.p, dw.p, du.p, dr.p); break;
    }
    GFX_HIP(hipGetLastError());
    GFX_HIP(hipDeviceSynchronize());
    dr.get(r, P * 3);
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the device copies of the inputs and of r live in
// one file-scope device array instead of hipMalloc'd buffers.  Each GPU holds
// its own copy of the array (at its own address) from the first use until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: it passes the device address hipGetSymbolAddress returns for the
// current device, fetched once per device and kept in scratch_ptr.
constexpr size_t SCRATCH_BYTES = size_t(GFX_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;              // one call at a time in this process (all devices)
void* scratch_ptr[MAX_DEVICES] = {}; // device address of `scratch` per device

void* scratch_address(int dev) {
    require(dev >= 0 && dev < MAX_DEVICES, "device index out of range");
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(GFX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[dev] = p;
    }
    return scratch_ptr[dev];
}

inline size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// Scratch bytes the allocation-free plugin needs for n_elem elements of
// ngll^3 points (FP64: dxi_dx, J, lambda, mu, u, r per point; D and w).
size_t noalloc_bytes(int64_t n_elem, int ngll) {
    const size_t P = size_t(n_elem) * ngll * ngll * ngll, d = sizeof(double);
    return round256(P * 9 * d) + 3 * round256(P * d) + round256(size_t(ngll) * ngll * d) +
           round256(size_t(ngll) * d) + 2 * round256(P * 3 * d);
}

void element_residual_noalloc(int n_elem, const double* dxi_dx, const double* jacobian,
                              const double* lambda_, const double* mu_, const double* D,
                              const double* weights, int ngll, const double* u, double* r) {
    require(n_elem >= 0, "n_elem must be >= 0");
    require(ngll >= 2 && ngll <= 8, "NGLL must be in [2, 8]");
    if (n_elem == 0) return;
    require(dxi_dx && jacobian && lambda_ && mu_ && D && weights && u && r, "null array");
    const size_t need = noalloc_bytes(n_elem, ngll);
    if (need > SCRATCH_BYTES) {
        const size_t per = noalloc_bytes(1, ngll);
        throw Error(GFX_ERR_INVALID, "gfx_compute_element_residual_noalloc: " + std::to_string(n_elem) +
                                         " elements of NGLL " + std::to_string(ngll) + " need " +
                                         std::to_string(need) + " bytes of scratch; the maximum is " +
                                         std::to_string(SCRATCH_BYTES) + " bytes (about " +
                                         std::to_string(SCRATCH_BYTES / per) + " elements)");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        throw Error(GFX_ERR_DEVICE, "no HIP device available");
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    int dev = 0;
    GFX_HIP(hipGetDevice(&dev));
    unsigned char* base = static_cast<unsigned char*>(scratch_address(dev));
    const size_t nn = size_t(ngll) * ngll * ngll, P = size_t(n_elem) * nn, d = sizeof(double);
    size_t off = 0;
    auto put = [&](const double* h, size_t n) {
        double* p = reinterpret_cast<double*>(base + off);
        GFX_HIP(hipMemcpy(p, h, n * d, hipMemcpyHostToDevice));
        off += round256(n * d);
        return p;
    };
    double* ddd = put(dxi_dx, P * 9);
    double* dj = put(jacobian, P);
    double* dl = put(lambda_, P);
    double* dm = put(mu_, P);
    double* dD = put(D, size_t(ngll) * ngll);
    double* dw = put(weights, ngll);
    double* du = put(u, P * 3);
    double* dr = put(r, P * 3);  // every byte the kernel reads is written by this call
    switch (ngll) {
        case 2: launch_plugin<double, 2>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        case 3: launch_plugin<double, 3>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        case 4: launch_plugin<double, 4>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        case 5: launch_plugin<double, 5>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        case 6: launch_plugin<double, 6>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        case 7: launch_plugin<double, 7>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
        default: launch_plugin<double, 8>(n_elem, ddd, dj, dl, dm, dD, dw, du, dr); break;
    }
    GFX_HIP(hipGetLastError());
    GFX_HIP(hipDeviceSynchronize());
    GFX_HIP(hipMemcpy(r, dr, P * 3 * d, hipMemcpyDeviceToHost));
}

}  // namespace
}  // namespace gfx

extern "C" gfx_status gfx_compute_element_residual_noalloc(int n_elem, const double* dxi_dx,
                                                           const double* jacobian,
                                                           const double* lambda_, const double* mu_,
                                                           const double* D, const double* weights,
                                                           int ngll, const double* u, double* r) {
    return gfx::guarded([&] {
        gfx::element_residual_noalloc(n_elem, dxi_dx, jacobian, lambda_, mu_, D, weights, ngll, u, r);
    });
}

extern "C" gfx_status gfx_compute_element_residual(int n_elem, const double* dxi_dx,
                                                   const double* jacobian, const double* lambda_,
                                                   const double* mu_, const double* D,
                                                   const double* weights, int ngll,
                                                   const double* u, double* r) {
    return gfx::guarded([&] {
        gfx::element_residual<double>(n_elem, dxi_dx, jacobian, lambda_, mu_, D, weights, ngll, u,
                                      r);
    });
}

extern "C" gfx_status gfx_compute_element_residual_f32(int n_elem, const double* dxi_dx,
                                                       const double* jacobian,
                                                       const double* lambda_, const double* mu_,
                                                       const double* D, const double* weights,
                                                       int ngll, const double* u, double* r) {
    return gfx::guarded([&] {
        gfx::require(ngll == 5, "the FP32 element path supports NGLL = 5 only");
        gfx::element_residual<float>(n_elem, dxi_dx, jacobian, lambda_, mu_, D, weights, ngll, u,
                                     r);
    });
}
