// kernels.hpp -- CDNA4 (gfx950) device kernels of the elastic CG-SEM step.
//
// Per step (FP32 throughout):
//   element kernel, once per element colour (or once with atomics):
//     gather predicted displacement of the element's 125 nodes into LDS,
//     reference-space gradients (GLL derivative contractions over LDS),
//     scaled physical gradient -> strain -> isotropic stress -> fluxes (LDS),
//     weighted divergence, plus the C-PML acceleration term and the point
//     source, then scatter-add into the assembled residual;
//   node kernel: diagonal-mass Newmark corrector fused with the next
//     predictor (beta = 0, gamma = 1/2), residual reset;
//   record kernel (snapshot steps): strain of the recording cells.
//
// Numerics follow the reference CPU path:
//   element force   forward/elastic/src/element_cpu.cpp:19-162
//   C-PML           forward/share/src/pml.cpp:55-216 (fused: the memory
//                   update only needs u_n, v_n, a_n and the previous "new",
//                   all known when the element kernel runs)
//   source          forward/share/src/solver.cpp:576-605
//   Newmark         forward/share/src/solver.cpp:40-69
//   strain          forward/share/src/solver.cpp:282-335
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfx {
namespace dev {

constexpr int NG = 5;    // GLL points per axis (polynomial order 4)
constexpr int NP = 125;  // points per element

struct ElemArgs {
    // geometry, SoA over element points in the solver's element order
    const float* __restrict__ G;      // [9][P]: dxi_dx[m*3+a] * sqrt(J w_i w_j w_k)
    const float* __restrict__ lam;    // [P]   (0 where mu <= 0)
    const float* __restrict__ mu;     // [P]   (0 where mu <= 0)
    const int32_t* __restrict__ ibool;  // [P]  solver node numbering
    int64_t P;
    // nodal fields, field-major [nf][fstride], xyz interleaved per node
    int nf;
    int64_t fstride;
    const float* __restrict__ ut;     // predicted displacement (u_{n+1})
    const float* __restrict__ u;      // u_n  (C-PML only)
    const float* __restrict__ v;      // v_n  (C-PML only)
    const float* __restrict__ a;      // a_n  (C-PML only)
    float* __restrict__ r;            // assembled residual
    // C-PML, compact over PML elements
    const int32_t* __restrict__ pml_slot;  // [E] slot or -1
    int64_t Ppml;
    const float* __restrict__ palpha;  // [9][Ppml] coef_alpha (dir*3 + {0,1,2})
    const float* __restrict__ pabar;   // [5][Ppml] w J / rho * A_k
    float* __restrict__ pmem;          // [nf][9][Ppml] rmemory (comp*3 + dir)
    float* __restrict__ pnew;          // [nf][3][Ppml] pml_displ_new
    float c1, c2;
    // point sources
    const int32_t* __restrict__ src_off;  // [E+1]
    const int4* __restrict__ src_ent;     // (field, dir, weight offset, 0)
    const float* __restrict__ src_w;      // weights
    const float* __restrict__ stf;        // [nf][nstf]
    int nstf, step;
    float D[25];
};

struct NodeArgs {
    int64_t G;
    int nf;
    int64_t fstride;
    const float* __restrict__ minv;  // [G] 1/m (0 where m <= 0)
    float* __restrict__ ucur;        // in: u_n,     out: predicted u_{n+2}
    float* __restrict__ unext;       // in: u_{n+1} (predicted); out: u_{n+1}
    float* __restrict__ v;
    float* __restrict__ a;
    float* __restrict__ r;           // in: residual; out: 0
    const float* __restrict__ damp;  // [G] legacy velocity damping or nullptr
    float dt;
};

struct RecArgs {
    int n_rec;
    const int32_t* __restrict__ ibool;  // [n_rec*125] solver node ids of recording cells
    const float* __restrict__ dd;       // [9][n_rec*125] unscaled dxi_dx
    int nf;
    int64_t fstride;
    const float* __restrict__ u;
    const float* __restrict__ v;
    const float* __restrict__ a;
    float* __restrict__ strain;  // [nf][n_rec*125][6]
    float* __restrict__ uva;     // [nf][n_rec*125][9] or nullptr
    float D[25];
};

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// works on a contiguous range of elements (neighbouring elements share nodes,
// and each XCD has its own L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, idx = b >> 3;
    return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

template <int EPB, bool ATOMIC>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64)
    k_element(ElemArgs A, int e0, int ne) {
    __shared__ float sD[25];
    __shared__ float su[EPB][3][NP];
    __shared__ float sf[EPB][9][NP];
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int le = tid / NP;
    const int p = tid - le * NP;
    const int eloc = blk * EPB + le;
    const bool valid = (le < EPB) && (eloc < ne);
    const int e = e0 + (valid ? eloc : 0);
    const int i = p / 25, j = (p / 5) % 5, k = p % 5;
    const int lel = valid ? le : 0;

    float g[9];
    float lam = 0.f, mu = 0.f;
    int node = 0;
    const int64_t pt = int64_t(e) * NP + p;
    if (valid) {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = __builtin_nontemporal_load(&A.G[c * A.P + pt]);
        lam = __builtin_nontemporal_load(&A.lam[pt]);
        mu = __builtin_nontemporal_load(&A.mu[pt]);
        node = __builtin_nontemporal_load(&A.ibool[pt]);
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = 0.f;
    }
    const int pslot = valid ? A.pml_slot[e] : -1;
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    __syncthreads();
    float Di[5], Dj[5], Dk[5], Dti[5], Dtj[5], Dtk[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        Di[s] = sD[i * 5 + s];
        Dj[s] = sD[j * 5 + s];
        Dk[s] = sD[k * 5 + s];
        Dti[s] = sD[s * 5 + i];
        Dtj[s] = sD[s * 5 + j];
        Dtk[s] = sD[s * 5 + k];
    }

    for (int f = 0; f < A.nf; ++f) {
        const int64_t fo = int64_t(f) * A.fstride;
        if (valid) {
            const float* ut = A.ut + fo + int64_t(node) * 3;
            su[le][0][p] = ut[0];
            su[le][1][p] = ut[1];
            su[le][2][p] = ut[2];
        }
        __syncthreads();
        if (valid) {
            float du[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    a0 += Di[s] * su[lel][c][s * 25 + j * 5 + k];
                    a1 += Dj[s] * su[lel][c][i * 25 + s * 5 + k];
                    a2 += Dk[s] * su[lel][c][i * 25 + j * 5 + s];
                }
                du[c][0] = a0;
                du[c][1] = a1;
                du[c][2] = a2;
            }
            // scaled physical gradient: gr[c][m] = sum_a du[c][a] * dxi_a/dx_m * sqrt(Jw)
            float gr[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    gr[c][m] = du[c][0] * g[m * 3 + 0] + du[c][1] * g[m * 3 + 1] +
                               du[c][2] * g[m * 3 + 2];
            const float exy = 0.5f * (gr[0][1] + gr[1][0]);
            const float exz = 0.5f * (gr[0][2] + gr[2][0]);
            const float eyz = 0.5f * (gr[1][2] + gr[2][1]);
            const float tr = gr[0][0] + gr[1][1] + gr[2][2];
            const float mu2 = 2.f * mu;
            float sig[3][3];
            sig[0][0] = mu2 * gr[0][0] + lam * tr;
            sig[1][1] = mu2 * gr[1][1] + lam * tr;
            sig[2][2] = mu2 * gr[2][2] + lam * tr;
            sig[0][1] = sig[1][0] = mu2 * exy;
            sig[0][2] = sig[2][0] = mu2 * exz;
            sig[1][2] = sig[2][1] = mu2 * eyz;
            // fluxes F[c][a] = sum_m sig[c][m] * dxi_a/dx_m * J w
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int aa = 0; aa < 3; ++aa)
                    sf[le][c * 3 + aa][p] = sig[c][0] * g[0 * 3 + aa] + sig[c][1] * g[1 * 3 + aa] +
                                            sig[c][2] * g[2 * 3 + aa];
        }
        __syncthreads();
        if (valid) {
            float rv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float acc = 0.f;
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    acc += Dti[s] * sf[lel][c * 3 + 0][s * 25 + j * 5 + k];
                    acc += Dtj[s] * sf[lel][c * 3 + 1][i * 25 + s * 5 + k];
                    acc += Dtk[s] * sf[lel][c * 3 + 2][i * 25 + j * 5 + s];
                }
                rv[c] = -acc;
            }
            if (pslot >= 0) {
                const int64_t q = int64_t(pslot) * NP + p;
                const int64_t gn = fo + int64_t(node) * 3;
                float un[3], vn[3], an[3], m[9], pn[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    un[c] = A.u[gn + c];
                    vn[c] = A.v[gn + c];
                    an[c] = A.a[gn + c];
                }
                float* pmem = A.pmem + int64_t(f) * 9 * A.Ppml;
                float* pnew = A.pnew + int64_t(f) * 3 * A.Ppml;
#pragma unroll
                for (int c = 0; c < 9; ++c) m[c] = pmem[c * A.Ppml + q];
#pragma unroll
                for (int c = 0; c < 3; ++c) pn[c] = pnew[c * A.Ppml + q];
                float al[9], S[5];
#pragma unroll
                for (int c = 0; c < 9; ++c) al[c] = A.palpha[c * A.Ppml + q];
#pragma unroll
                for (int c = 0; c < 5; ++c) S[c] = A.pabar[c * A.Ppml + q];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float oldv = pn[c] + A.c2 * an[c];   // pml.cpp:91-109
                    const float newv = un[c] + A.c1 * vn[c];   // pml.cpp:72-89
                    rv[c] += S[0] * vn[c] + S[1] * un[c] + S[2] * m[c * 3 + 0] +
                             S[3] * m[c * 3 + 1] + S[4] * m[c * 3 + 2];  // pml.cpp:209-212
#pragma unroll
                    for (int d = 0; d < 3; ++d)  // pml.cpp:133-146
                        pmem[(c * 3 + d) * A.Ppml + q] =
                            al[d * 3 + 0] * m[c * 3 + d] + al[d * 3 + 1] * newv + al[d * 3 + 2] * oldv;
                    pnew[c * A.Ppml + q] = newv;
                }
            }
            for (int s = sb; s < se; ++s) {
                const int4 en = A.src_ent[s];
                if (en.x == f && A.step < A.nstf) {
                    const float sv = A.stf[int64_t(f) * A.nstf + A.step];
                    rv[en.y] += sv * A.src_w[en.z + p];
                }
            }
            float* r = A.r + fo + int64_t(node) * 3;
            if (ATOMIC) {
                unsafeAtomicAdd(r + 0, rv[0]);
                unsafeAtomicAdd(r + 1, rv[1]);
                unsafeAtomicAdd(r + 2, rv[2]);
            } else {
                r[0] += rv[0];
                r[1] += rv[1];
                r[2] += rv[2];
            }
        }
        if (f + 1 < A.nf) __syncthreads();
    }
}

// Corrector of step n fused with the predictor of step n+1
// (forward/share/src/solver.cpp:40-69 with beta = 0, gamma = 1/2).
__global__ void __launch_bounds__(256) k_node(NodeArgs A) {
    const int64_t gidx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gidx >= A.G) return;
    const float mi = A.minv[gidx];
    const float dt = A.dt, hdt = 0.5f * dt, hdt2 = 0.5f * dt * dt;
    const float dmp = A.damp ? A.damp[gidx] : 0.f;
    for (int f = 0; f < A.nf; ++f) {
        const int64_t o = int64_t(f) * A.fstride + gidx * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float rr = A.r[o + c];
            const float v0 = A.v[o + c];
            const float ao = A.a[o + c];
            // legacy ramp v -= d v before the corrector (solver.cpp:565-573)
            float vv = (dmp > 0.f) ? v0 - dmp * v0 : v0;
            if (mi > 0.f) {
                const float an = rr * mi;
                // u_{n+1} = u_n + dt v + dt^2/2 a_n = predicted (beta = 0), corrected
                // for the damped velocity when the legacy ramp is active
                float un = A.unext[o + c];
                if (dmp > 0.f) {
                    un += dt * (vv - v0);
                    A.unext[o + c] = un;
                }
                vv += hdt * (ao + an);
                A.v[o + c] = vv;
                A.a[o + c] = an;
                A.ucur[o + c] = un + dt * vv + hdt2 * an;
            } else {  // m <= 0: a = 0, u and v kept (solver.cpp:57-61)
                const float uo = A.ucur[o + c];
                A.v[o + c] = vv;
                A.a[o + c] = 0.f;
                A.unext[o + c] = uo;
                A.ucur[o + c] = uo + dt * vv;
            }
            A.r[o + c] = 0.f;
        }
    }
}

// Strain of the recording cells (forward/share/src/solver.cpp:282-335) and
// copies of u, v, a at their points (solver.cpp:656-674).
template <int EPB>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64) k_record(RecArgs A) {
    __shared__ float sD[25];
    __shared__ float su[EPB][3][NP];
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int le = tid / NP;
    const int p = tid - le * NP;
    const int cell = blockIdx.x * EPB + le;
    const bool valid = (le < EPB) && (cell < A.n_rec);
    const int lel = valid ? le : 0;
    const int i = p / 25, j = (p / 5) % 5, k = p % 5;
    const int64_t pt = int64_t(valid ? cell : 0) * NP + p;
    const int64_t nrp = int64_t(A.n_rec) * NP;
    const int node = valid ? A.ibool[pt] : 0;
    float dd[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) dd[c] = valid ? A.dd[c * nrp + pt] : 0.f;
    __syncthreads();
    for (int f = 0; f < A.nf; ++f) {
        const int64_t go = int64_t(f) * A.fstride + int64_t(node) * 3;
        if (valid) {
            su[le][0][p] = A.u[go + 0];
            su[le][1][p] = A.u[go + 1];
            su[le][2][p] = A.u[go + 2];
        }
        __syncthreads();
        if (valid) {
            float du[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    a0 += sD[i * 5 + s] * su[lel][c][s * 25 + j * 5 + k];
                    a1 += sD[j * 5 + s] * su[lel][c][i * 25 + s * 5 + k];
                    a2 += sD[k * 5 + s] * su[lel][c][i * 25 + j * 5 + s];
                }
                du[c][0] = a0;
                du[c][1] = a1;
                du[c][2] = a2;
            }
            float gr[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    gr[c][m] = du[c][0] * dd[m * 3 + 0] + du[c][1] * dd[m * 3 + 1] +
                               du[c][2] * dd[m * 3 + 2];
            float* s6 = A.strain + (int64_t(f) * nrp + pt) * 6;
            s6[0] = gr[0][0];
            s6[1] = gr[1][1];
            s6[2] = gr[2][2];
            s6[3] = 0.5f * (gr[0][1] + gr[1][0]);
            s6[4] = 0.5f * (gr[0][2] + gr[2][0]);
            s6[5] = 0.5f * (gr[1][2] + gr[2][1]);
            if (A.uva) {
                float* o9 = A.uva + (int64_t(f) * nrp + pt) * 9;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    o9[c] = su[lel][c][p];
                    o9[3 + c] = A.v[go + c];
                    o9[6 + c] = A.a[go + c];
                }
            }
        }
        if (f + 1 < A.nf) __syncthreads();
    }
}

}  // namespace dev
}  // namespace gfx
