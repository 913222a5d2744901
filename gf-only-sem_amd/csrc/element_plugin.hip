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

}  // namespace
}  // namespace gfx

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
