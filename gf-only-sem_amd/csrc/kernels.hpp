// kernels.hpp -- CDNA4 (gfx950) device kernels of the elastic CG-SEM step.
//
// Nodal state (FP32), one array per field ([field][node][xyz], field stride
// fstride floats): U = predicted displacement u~_n (= u_{n+1}, beta = 0),
// W = v_n + dt/2 a_n, R = assembled residual, and a_n only at "special" nodes
// (C-PML, legacy-damped, recording and massless nodes, numbered first:
// ids [0, Ga)).  The reference's Newmark pair (forward/share/src/solver.cpp:40-69,
// beta = 0, gamma = 1/2) is carried algebraically unchanged:
//   a_{n+1} = r / m,  w_{n+1} = w_n + dt a_{n+1},  u~_{n+1} = u~_n + dt w_{n+1},
//   u_n = u~_n - dt w_n,  v_n = w_n - dt/2 a_n.
//
// Per step:
//   element kernel: gather u~ into LDS, GLL contractions, scaled gradient ->
//     strain -> isotropic stress -> fluxes (LDS) -> divergence, point sources;
//     then, by assembly mode,
//       COLORED: one launch per element colour, store (first touch) or add into R;
//       ATOMIC:  float atomics into R (zeroed first);
//       GATHER:  one launch, element-local residual Rl[field][point][xyz];
//   special-node kernel: [GATHER: R = fixed-order sum of Rl] C-PML
//     (forward/share/src/pml.cpp:55-216, assembled per node), legacy damping,
//     massless rule, corrector/predictor;
//   regular-node kernel: [GATHER sum] streaming corrector/predictor;
//   record kernel (snapshot steps): strain of the recording cells
//     (forward/share/src/solver.cpp:282-335).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfx {
namespace dev {

constexpr int NG = 5;    // GLL points per axis (polynomial order 4)
constexpr int NP = 125;  // points per element
constexpr uint32_t FIRST_TOUCH = 0x80000000u;
constexpr int CPML_NC = 28;  // complete C-PML coefficients per point
constexpr int CPML_NM = 45;  // complete C-PML memories per point and field

enum { ASM_COLORED = 0, ASM_ATOMIC = 1, ASM_GATHER = 2 };

struct ElemArgs {
    // geometry, SoA over element points in the solver's element order
    const float* __restrict__ G;        // [9][P]: dxi_dx[m*3+a] * sqrt(J w_i w_j w_k)
    const float* __restrict__ lam;      // [P]   (0 where mu <= 0)
    const float* __restrict__ mu;       // [P]   (0 where mu <= 0)
    const uint32_t* __restrict__ ibool; // [P]  node id | FIRST_TOUCH
    int64_t P;
    int nf;
    int64_t fstride;                    // floats per field of the nodal arrays
    const float* __restrict__ U;        // [nf][fstride] predicted displacement
    float* __restrict__ R;              // [nf][fstride] assembled residual (COLORED, ATOMIC)
    float* __restrict__ Rl;             // [nf][PR][3] element-point residuals (GATHER): entry rdst[pt]
    const int32_t* __restrict__ rdst;   // [P] a node's first point: the node id; others: G4 + rank
    int64_t PR;
    // element-local C-PML (only when the nodal form is not applicable)
    const float* __restrict__ W;
    const float* __restrict__ Aacc;     // [nf][astride] a_n at special nodes
    int64_t astride;
    const int32_t* __restrict__ pml_slot;  // [E] slot or -1
    int64_t Ppml;
    const float* __restrict__ palpha;  // [9][Ppml]
    const float* __restrict__ pabar;   // [5][Ppml] w J / rho * A_k
    float* __restrict__ pmem;          // [nf][9][Ppml]
    float* __restrict__ pnew;          // [nf][3][Ppml]
    float c1, c2, dt;
    // complete C-PML (PML elements; slot from pml_slot): coefficients
    // [CPML_NC][Ppml] (axis q at 9q: K, d, alpha, a b c of alpha, a b c of
    // beta; 27: rho J w) and memories pmem [nf][CPML_NM][Ppml]
    const float* __restrict__ pcoef;
    // point sources
    const int32_t* __restrict__ src_off;  // [E+1]
    const int4* __restrict__ src_ent;     // (field, dir, weight offset, 0)
    const float* __restrict__ src_w;
    const float* __restrict__ stf;        // [nf][nstf]
    int nstf;
    const int* __restrict__ ctr;          // device counters: [0] step index, [1] snapshots taken
    // shear attenuation (n_sls > 0): lam / mu hold the unrelaxed moduli; per
    // mechanism l and point: att_a = exp(-dt/tau_sigma), att_b = Y b, att_c = Y c
    // ([n_sls][P]); memory zeta [nf][n_sls*5][P] and previous deviatoric
    // strain devp [nf][5][P], both in the scaled strain units of the kernels
    int n_sls;
    const float* __restrict__ att_a;
    const float* __restrict__ att_b;
    const float* __restrict__ att_c;
    float* __restrict__ zeta;
    float* __restrict__ devp;
    float D[25];
};

struct NodeArgs {
    int64_t G, Ga, Gpml, G4, P;
    int nf;
    int64_t fstride, astride;
    const float* __restrict__ minv;   // [G4] 1/m (0 where m <= 0 or padding)
    const float* __restrict__ damp;   // [Ga] legacy ramp or nullptr
    float* __restrict__ U;
    float* __restrict__ W;
    float* __restrict__ R;
    float* __restrict__ Aacc;         // [nf][astride]
    // GATHER assembly: node g's first element point (lowest solver order) has
    // its residual at Rl[f][g]; its other points are the slots [off[g],
    // off[g+1]), slot q holding the entry pts[q] (increasing in solver order)
    int gather;
    const float* __restrict__ Rl;     // [nf][PR][3]
    int64_t PR;
    const int32_t* __restrict__ off;  // [G4+1]
    const int32_t* __restrict__ pts;  // [P]
    // nodal C-PML
    int pml_nodal;
    const float* __restrict__ pc;      // [5][Gpml] sum over PML elements of w J / rho * A_k
    const float* __restrict__ palpha;  // [9][Gpml] coef_alpha (dir*3 + {0,1,2})
    const uint8_t* __restrict__ pmask; // [Gpml] memory directions in use (unused ones stay 0)
    float* __restrict__ pmem;          // [nf][9][Gpml] rmemory (comp*3 + dir)
    float* __restrict__ pnew;          // [nf][3][Gpml] pml_displ_new
    float c1, c2, dt;
};

struct RecArgs {
    const int* __restrict__ ctr;  // device counters: [0] step index, [1] snapshots taken
    int stride, cap;              // snapshot stride; ring capacity (snapshots)
    int n_rec;
    const int32_t* __restrict__ ibool;  // [n_rec*125] special node ids of recording cells
    const float* __restrict__ dd;       // [9][n_rec*125] unscaled dxi_dx
    int nf;
    int64_t fstride, astride;
    const float* __restrict__ U;
    const float* __restrict__ W;
    const float* __restrict__ Aacc;
    float dt;
    float* __restrict__ strain;  // ring [cap][nf][n_rec*125][6]
    float* __restrict__ uva;     // ring [cap][nf][n_rec*125][9] or nullptr
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

// Wave-level LDS visibility: the lanes of one wavefront exchange data through
// LDS with no workgroup barrier (same sequence as HIP's __syncwarp).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The flux of one point from its reference-space gradient du[c][a] =
// d u_c / d xi_a: chain rule with the scaled geometry g, small strain,
// isotropic stress, then F[c][a] = sum_m sigma[c][m] g[m][a]
// (element_cpu.cpp:51-108).
template <bool DIAG>
__device__ __forceinline__ void flux_from_du(const float (&du)[3][3], const float* g, float lam, float mu,
                                             float* F) {
    float gr[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            gr[c][m] = DIAG ? du[c][m] * g[m * 4]
                            : du[c][0] * g[m * 3 + 0] + du[c][1] * g[m * 3 + 1] + du[c][2] * g[m * 3 + 2];
    const float tr = gr[0][0] + gr[1][1] + gr[2][2];
    const float mu2 = 2.f * mu;
    float sig[3][3];
    sig[0][0] = mu2 * gr[0][0] + lam * tr;
    sig[1][1] = mu2 * gr[1][1] + lam * tr;
    sig[2][2] = mu2 * gr[2][2] + lam * tr;
    sig[0][1] = sig[1][0] = mu * (gr[0][1] + gr[1][0]);
    sig[0][2] = sig[2][0] = mu * (gr[0][2] + gr[2][0]);
    sig[1][2] = sig[2][1] = mu * (gr[1][2] + gr[2][1]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            F[c * 3 + a] = DIAG ? sig[c][a] * g[a * 4]
                                : sig[c][0] * g[0 * 3 + a] + sig[c][1] * g[1 * 3 + a] + sig[c][2] * g[2 * 3 + a];
}

// Element force of one GLL point (i,j,k) from its three rows of displacement
// neighbours ui[c][s] = u_c(s,j,k), uj[c][s] = u_c(i,s,k), uk[c][s] = u_c(i,j,s)
// and the scaled geometry g (dxi_a/dx_m sqrt(J w)): writes the 9 fluxes
// F[c][a] = sum_m sigma[c][m] dxi_a/dx_m J w  (element_cpu.cpp:51-108).  Every
// element kernel form goes through this one sequence of operations.  DIAG:
// the element is axis-aligned (dxi_a/dx_m = 0 for a != m, up to FP64
// round-off of the input), only g[0], g[4], g[8] are used and the products
// with the zero entries are left out.
template <bool DIAG>
__device__ __forceinline__ void flux_core(const float (&ui)[3][5], const float (&uj)[3][5], const float (&uk)[3][5],
                                          const float* Di, const float* Dj, const float* Dk, const float* g,
                                          float lam, float mu, float* F) {
    float du[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            a0 += Di[s] * ui[c][s];
            a1 += Dj[s] * uj[c][s];
            a2 += Dk[s] * uk[c][s];
        }
        du[c][0] = a0;
        du[c][1] = a1;
        du[c][2] = a2;
    }
    flux_from_du<DIAG>(du, g, lam, mu, F);
}

// Geometry of one element point: all 9 scaled entries, or (DIAG) the three
// diagonal ones with the others 0.
template <bool DIAG>
__device__ __forceinline__ void load_geometry(const ElemArgs& A, int64_t pt, float* g) {
#pragma unroll
    for (int c = 0; c < 9; ++c)
        g[c] = (DIAG && (c & 3) != 0) ? 0.f : __builtin_nontemporal_load(&A.G[c * A.P + pt]);
}

// point_flux with the neighbours read from a staged element (su_c[p], p = (i*5+j)*5+k).
template <bool DIAG>
__device__ __forceinline__ void point_flux(const float* __restrict__ su0, const float* __restrict__ su1,
                                           const float* __restrict__ su2, const float* Di,
                                           const float* Dj, const float* Dk, int i, int j, int k,
                                           const float* g, float lam, float mu, float* F) {
    const float* su[3] = {su0, su1, su2};
    float ui[3][5], uj[3][5], uk[3][5];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            ui[c][s] = su[c][s * 25 + j * 5 + k];
            uj[c][s] = su[c][i * 25 + s * 5 + k];
            uk[c][s] = su[c][i * 25 + j * 5 + s];
        }
    flux_core<DIAG>(ui, uj, uk, Di, Dj, Dk, g, lam, mu, F);
}

// One component of the weighted divergence from the point's flux rows
// a[s] = F_c0(s,j,k), b[s] = F_c1(i,s,k), e[s] = F_c2(i,j,s):
// -sum_s (D[s][i] a[s] + D[s][j] b[s] + D[s][k] e[s]), one fixed order.
__device__ __forceinline__ float div_core(const float (&a)[5], const float (&b)[5], const float (&e)[5],
                                         const float* Dti, const float* Dtj, const float* Dtk) {
    float acc = 0.f;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        acc += Dti[s] * a[s];
        acc += Dtj[s] * b[s];
        acc += Dtk[s] * e[s];
    }
    return -acc;
}

// Weighted divergence of the staged fluxes at point (i,j,k):
// r_c = -sum_s (D[s][i] F_c0(s,j,k) + D[s][j] F_c1(i,s,k) + D[s][k] F_c2(i,j,s))
// (element_cpu.cpp:110-157 rewritten as a gather over the point's neighbours).
__device__ __forceinline__ void point_div(const float* __restrict__ sf, int stride, const float* Dti,
                                          const float* Dtj, const float* Dtk, int i, int j, int k,
                                          float* rv) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float a[5], b[5], e[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            a[s] = sf[(c * 3 + 0) * stride + s * 25 + j * 5 + k];
            b[s] = sf[(c * 3 + 1) * stride + i * 25 + s * 5 + k];
            e[s] = sf[(c * 3 + 2) * stride + i * 25 + j * 5 + s];
        }
        rv[c] = div_core(a, b, e, Dti, Dtj, Dtk);
    }
}

template <int MODE>
__device__ __forceinline__ void emit_residual(const ElemArgs& A, int f, int64_t node, int64_t slot,
                                              const float* rold, const float* rv) {
    if (MODE == ASM_GATHER) {  // slot: the point's entry rdst[pt]
        float* o = A.Rl + (int64_t(f) * A.PR + slot) * 3;
        o[0] = rv[0];
        o[1] = rv[1];
        o[2] = rv[2];
    } else if (MODE == ASM_ATOMIC) {
        float* r = A.R + int64_t(f) * A.fstride + node * 3;
        unsafeAtomicAdd(r + 0, rv[0]);
        unsafeAtomicAdd(r + 1, rv[1]);
        unsafeAtomicAdd(r + 2, rv[2]);
    } else {  // the first element (in colour order) touching a node stores (rold = 0)
        float* r = A.R + int64_t(f) * A.fstride + node * 3;
        r[0] = rold[0] + rv[0];
        r[1] = rold[1] + rv[1];
        r[2] = rold[2] + rv[2];
    }
}

// Residual entry of element point pt: rdst[pt] for GATHER assembly (a node's
// first point at the node id, later points after the nodes), unused otherwise.
template <int MODE>
__device__ __forceinline__ int64_t residual_entry(const ElemArgs& A, int64_t pt, bool valid) {
    return (MODE == ASM_GATHER && valid) ? int64_t(A.rdst[pt]) : pt;
}

__device__ __forceinline__ void add_sources(const ElemArgs& A, int step, int f, int sb, int se, int p,
                                            float* rv) {
    for (int s = sb; s < se; ++s) {  // solver.cpp:576-605
        const int4 en = A.src_ent[s];
        if (en.x == f && step < A.nstf) rv[en.y] += A.stf[int64_t(f) * A.nstf + step] * A.src_w[en.z + p];
    }
}

// Closes a step: the step counter advances, and the snapshot counter when this
// step was a snapshot step (step % stride == 0), as the host mirrors it.
__global__ void k_advance(int* ctr, int stride) {
    if (threadIdx.x == 0) {
        const int st = ctr[0];
        if (stride > 0 && st % stride == 0) ctr[1] = ctr[1] + 1;
        ctr[0] = st + 1;
    }
}

// k_advance for several solvers in one launch (lane i: solver i).
__global__ void k_advance_grp(int* const* __restrict__ ctrs, const int* __restrict__ strides, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int* ctr = ctrs[i];
    const int st = ctr[0];
    if (strides[i] > 0 && st % strides[i] == 0) ctr[1] = ctr[1] + 1;
    ctr[0] = st + 1;
}

// One lane per GLL point, EPB elements per workgroup, workgroup barriers
// between the stages.  Supports the element-local C-PML fallback.
template <int EPB, int MODE, bool PML_ELEM, bool DIAG>
__device__ __forceinline__ void element_body(const ElemArgs& A, int e0, int ne, int blk) {
    __shared__ float sD[25];
    __shared__ float su[EPB][3][NP];
    __shared__ float sf[EPB][9][NP];
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int le = tid / NP;
    const int p = tid - le * NP;
    const int eloc = blk * EPB + le;
    const bool valid = (le < EPB) && (eloc < ne);
    const int e = e0 + (valid ? eloc : 0);
    const int i = p / 25, j = (p / 5) % 5, k = p % 5;
    const int lel = valid ? le : 0;

    float g[9];
    float lam = 0.f, mu = 0.f;
    uint32_t nodef = 0;
    const int64_t pt = int64_t(e) * NP + p;
    if (valid) {
        load_geometry<DIAG>(A, pt, g);
        lam = __builtin_nontemporal_load(&A.lam[pt]);
        mu = __builtin_nontemporal_load(&A.mu[pt]);
        nodef = __builtin_nontemporal_load(&A.ibool[pt]);
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = 0.f;
    }
    const int64_t node = int64_t(nodef & ~FIRST_TOUCH);
    const bool first = (nodef & FIRST_TOUCH) != 0;
    const int64_t rslot = residual_entry<MODE>(A, pt, valid);
    const int pslot = (PML_ELEM && valid) ? A.pml_slot[e] : -1;
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    const int step = A.ctr[0];
    float un[3] = {0.f, 0.f, 0.f};  // field 0; later fields are prefetched one iteration ahead
    if (valid) {
        const float* ut = A.U + node * 3;
        un[0] = ut[0];
        un[1] = ut[1];
        un[2] = ut[2];
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

    // Two barriers per field: (1) after the displacement is staged, (2) after
    // the fluxes are staged.  Staging the next field's displacement happens after
    // (2), when every lane has finished reading su; writing the next fluxes
    // happens after the next (1), when every lane has finished reading sf.
    for (int f = 0; f < A.nf; ++f) {
        const int64_t fo = int64_t(f) * A.fstride;
        if (valid) {
            su[le][0][p] = un[0];
            su[le][1][p] = un[1];
            su[le][2][p] = un[2];
        }
        if (valid && f + 1 < A.nf) {
            const float* ut = A.U + fo + A.fstride + node * 3;
            un[0] = ut[0];
            un[1] = ut[1];
            un[2] = ut[2];
        }
        float rold[3] = {0.f, 0.f, 0.f};
        if (MODE == ASM_COLORED && valid && !first) {
            const float* r = A.R + fo + node * 3;
            rold[0] = r[0];
            rold[1] = r[1];
            rold[2] = r[2];
        }
        __syncthreads();
        if (valid) {
            float F[9];
            point_flux<DIAG>(su[lel][0], su[lel][1], su[lel][2], Di, Dj, Dk, i, j, k, g, lam, mu, F);
#pragma unroll
            for (int q = 0; q < 9; ++q) sf[le][q][p] = F[q];
        }
        __syncthreads();
        if (valid) {
            float rv[3];
            point_div(&sf[lel][0][0], NP, Dti, Dtj, Dtk, i, j, k, rv);
            if (PML_ELEM && pslot >= 0) {
                // element-local C-PML (pml.cpp:55-216); PML nodes are special, so a_n is stored
                const int64_t q = int64_t(pslot) * NP + p;
                const int64_t gn = fo + node * 3;
                const float* an_p = A.Aacc + int64_t(f) * A.astride + node * 3;
                float upn[3], vpn[3], apn[3], m[9], pn[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    apn[c] = an_p[c];
                    const float w = A.W[gn + c];
                    upn[c] = A.U[gn + c] - A.dt * w;
                    vpn[c] = w - 0.5f * A.dt * apn[c];
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
                    const float oldv = pn[c] + A.c2 * apn[c];   // pml.cpp:91-109
                    const float newv = upn[c] + A.c1 * vpn[c];  // pml.cpp:72-89
                    rv[c] += S[0] * vpn[c] + S[1] * upn[c] + S[2] * m[c * 3 + 0] +
                             S[3] * m[c * 3 + 1] + S[4] * m[c * 3 + 2];  // pml.cpp:209-212
#pragma unroll
                    for (int d = 0; d < 3; ++d)  // pml.cpp:133-146
                        pmem[(c * 3 + d) * A.Ppml + q] =
                            al[d * 3 + 0] * m[c * 3 + d] + al[d * 3 + 1] * newv + al[d * 3 + 2] * oldv;
                    pnew[c * A.Ppml + q] = newv;
                }
            }
            add_sources(A, step, f, sb, se, p, rv);
            emit_residual<MODE>(A, f, node, rslot, rold, rv);
        }
    }
}

template <int EPB, int MODE, bool PML_ELEM, bool DIAG>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64)
    k_element(ElemArgs A, int e0, int ne) {
    element_body<EPB, MODE, PML_ELEM, DIAG>(A, e0, ne, xcd_block(blockIdx.x, gridDim.x));
}

// Grouped launches cover several solvers: workgroup b of such a launch reads
// its work from map[b] = (solver, first element of its range, elements in the
// range, workgroup within the range), or (solver, -, -, workgroup within the
// solver's node launch) for the node kernels, and the solver's arguments from
// tab[solver].  One table read per workgroup, no search.

// Row i of an argument table, read through the constant address space: the
// table does not change during a launch, and pointers read from constant
// memory are known to address global memory, so the kernel bodies issue the
// same global (not flat) loads and stores as with by-value arguments.
template <class T>
__device__ __forceinline__ const T& const_row(const T* tab, int i) {
    using CT = const __attribute__((address_space(4))) T;
    return *(const T*)((CT*)tab + i);
}

// One launch for the element ranges of several solvers (nodal C-PML only).
template <int EPB, int MODE, bool DIAG>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64)
    k_element_grp(const ElemArgs* __restrict__ tab, const int4* __restrict__ map) {
    const int4 E = map[xcd_block(blockIdx.x, gridDim.x)];
    element_body<EPB, MODE, false, DIAG>(const_row(tab, E.x), E.y, E.z, E.w);
}

// Stress with shear attenuation at one point of field f (general geometry):
// strain from the gradient, the memory update of every mechanism
// z <- a z + (Y b) dev_old + (Y c) dev_new, sigma = lam_U tr I + 2 mu_U eps -
// sum z, and the fluxes F = sigma . g.  L > 0: L solids whose coefficients
// the caller holds in ca, cb, cc; L = 0: A.n_sls solids read from memory.
template <int L>
__device__ __forceinline__ void flux_att(const float (&ui)[3][5], const float (&uj)[3][5], const float (&uk)[3][5],
                                         const float* Di, const float* Dj, const float* Dk, const float* g,
                                         float lam, float mu, const ElemArgs& A, int f, int64_t pt, const float* ca,
                                         const float* cb, const float* cc, float* F) {
    float du[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            a0 += Di[s] * ui[c][s];
            a1 += Dj[s] * uj[c][s];
            a2 += Dk[s] * uk[c][s];
        }
        du[c][0] = a0;
        du[c][1] = a1;
        du[c][2] = a2;
    }
    float gr[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int m = 0; m < 3; ++m) gr[c][m] = du[c][0] * g[m * 3 + 0] + du[c][1] * g[m * 3 + 1] + du[c][2] * g[m * 3 + 2];
    const float tr = gr[0][0] + gr[1][1] + gr[2][2];
    const float dev[5] = {gr[0][0] - tr * (1.f / 3.f), gr[1][1] - tr * (1.f / 3.f), 0.5f * (gr[0][1] + gr[1][0]),
                          0.5f * (gr[0][2] + gr[2][0]), 0.5f * (gr[1][2] + gr[2][1])};
    float* dp = A.devp + int64_t(f) * 5 * A.P + pt;
    float old[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) old[q] = dp[int64_t(q) * A.P];
    float zs[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    const int nl = L > 0 ? L : A.n_sls;
    float* zf = A.zeta + int64_t(f) * nl * 5 * A.P + pt;
    if constexpr (L > 0) {
        float zo[L * 5];
#pragma unroll
        for (int q = 0; q < L * 5; ++q) zo[q] = zf[int64_t(q) * A.P];  // all loads first
#pragma unroll
        for (int l = 0; l < L; ++l)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const float z = ca[l] * zo[l * 5 + q] + cb[l] * old[q] + cc[l] * dev[q];
                zf[int64_t(l * 5 + q) * A.P] = z;
                zs[q] += z;
            }
    } else {
        for (int l = 0; l < nl; ++l) {
            const float a = A.att_a[int64_t(l) * A.P + pt], b = A.att_b[int64_t(l) * A.P + pt],
                        c = A.att_c[int64_t(l) * A.P + pt];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                float* zq = zf + int64_t(l * 5 + q) * A.P;
                const float z = a * zq[0] + b * old[q] + c * dev[q];
                zq[0] = z;
                zs[q] += z;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) dp[int64_t(q) * A.P] = dev[q];
    const float mu2 = 2.f * mu;
    float sig[3][3];
    sig[0][0] = mu2 * gr[0][0] + lam * tr - zs[0];
    sig[1][1] = mu2 * gr[1][1] + lam * tr - zs[1];
    sig[2][2] = mu2 * gr[2][2] + lam * tr + zs[0] + zs[1];
    sig[0][1] = sig[1][0] = mu * (gr[0][1] + gr[1][0]) - zs[2];
    sig[0][2] = sig[2][0] = mu * (gr[0][2] + gr[2][0]) - zs[3];
    sig[1][2] = sig[2][1] = mu * (gr[1][2] + gr[2][1]) - zs[4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            F[c * 3 + a] = sig[c][0] * g[0 * 3 + a] + sig[c][1] * g[1 * 3 + a] + sig[c][2] * g[2 * 3 + a];
}

// Block kernel with shear attenuation: one element per workgroup, one lane per
// GLL point, general geometry, nodal C-PML only.  L > 0 solids keep their step
// coefficients in registers for all fields; L = 0 reads A.n_sls solids' from memory.
template <int MODE, int L>
__global__ void __launch_bounds__(128) k_element_att(ElemArgs A, int e0, int ne) {
    __shared__ float sD[25];
    __shared__ float su[3][NP];
    __shared__ float sf[9][NP];
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int p = tid;
    const bool valid = p < NP && blk < ne;
    const int e = e0 + (valid ? blk : 0);
    const int pc = valid ? p : 0;
    const int i = pc / 25, j = (pc / 5) % 5, k = pc % 5;
    float g[9];
    float lam = 0.f, mu = 0.f;
    uint32_t nodef = 0;
    const int64_t pt = int64_t(e) * NP + pc;
    if (valid) {
        load_geometry<false>(A, pt, g);
        lam = A.lam[pt];
        mu = A.mu[pt];
        nodef = A.ibool[pt];
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = 0.f;
    }
    const int64_t node = int64_t(nodef & ~FIRST_TOUCH);
    const bool first = (nodef & FIRST_TOUCH) != 0;
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    const int step = A.ctr[0];
    constexpr int LC = L > 0 ? L : 1;
    float ca[LC], cb[LC], cc[LC];
#pragma unroll
    for (int l = 0; l < LC; ++l) {
        const bool have = L > 0 && valid;
        ca[l] = have ? A.att_a[int64_t(l) * A.P + pt] : 0.f;
        cb[l] = have ? A.att_b[int64_t(l) * A.P + pt] : 0.f;
        cc[l] = have ? A.att_c[int64_t(l) * A.P + pt] : 0.f;
    }
    const int64_t rslot = residual_entry<MODE>(A, pt, valid);
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
            const float* ut = A.U + fo + node * 3;
            su[0][p] = ut[0];
            su[1][p] = ut[1];
            su[2][p] = ut[2];
        }
        float rold[3] = {0.f, 0.f, 0.f};
        if (MODE == ASM_COLORED && valid && !first) {
            const float* r = A.R + fo + node * 3;
            rold[0] = r[0];
            rold[1] = r[1];
            rold[2] = r[2];
        }
        __syncthreads();
        if (valid) {
            float ui[3][5], uj[3][5], uk[3][5];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    ui[c][s] = su[c][s * 25 + j * 5 + k];
                    uj[c][s] = su[c][i * 25 + s * 5 + k];
                    uk[c][s] = su[c][i * 25 + j * 5 + s];
                }
            float F[9];
            flux_att<L>(ui, uj, uk, Di, Dj, Dk, g, lam, mu, A, f, pt, ca, cb, cc, F);
#pragma unroll
            for (int q = 0; q < 9; ++q) sf[q][p] = F[q];
        }
        __syncthreads();
        if (valid) {
            float rv[3];
            point_div(&sf[0][0], NP, Dti, Dtj, Dtk, i, j, k, rv);
            add_sources(A, step, f, sb, se, p, rv);
            emit_residual<MODE>(A, f, node, rslot, rold, rv);
        }
        __syncthreads();  // su and sf are rewritten by the next field
    }
}

// Whether complete C-PML memory t of a point is in use, from its damped axes
// on[] (d != 0): gradient memory (k*3+l)*4 + r uses axis m, n, both, or l for
// r = 0..3 (m < n the axes other than l); inertia memory 36 + c*3 + r uses axis r.
__device__ __forceinline__ bool cpml_mem_on(const bool (&on)[3], int t) {
    if (t >= 36) return on[(t - 36) % 3];
    const int l = (t / 4) % 3, r = t % 4;
    const int m = l == 0 ? 1 : 0, n = l == 2 ? 1 : 2;
    return r == 0 ? on[m] : r == 1 ? on[n] : r == 2 ? (on[m] && on[n]) : on[l];
}

// The complete C-PML coefficients of one point, held in LDS (column p of a
// [CPML_NC][128] block; c[i] is coefficient o + i).  Read from LDS where they
// are used instead of kept in registers over the field loop, they leave the
// kernel room for three waves per SIMD.
struct CpmlCoef {
    const float (*b)[128];
    int p, o;
    __device__ __forceinline__ float operator[](int i) const { return b[o + i][p]; }
};

// One step of conv(g, p) (exact for g linear in time over the step): the
// carry q holds a M + b g of the previous level; returns M = q + c g and
// leaves a M + b g for the next step.
__device__ __forceinline__ float conv_step(const float* abc, float& q, float g) {
    const float m = q + abc[2] * g;
    q = abc[0] * m + abc[1] * g;
    return m;
}

// Complete C-PML element kernel (PML elements only): one element per
// workgroup, one lane per GLL point, general geometry.  Stress column j of a
// point sees the gradient G^(j)_kl = F^-1[s_x s_y s_z / (s_j s_l)] * du_k/dx_l,
// built from chains of first-order filters of the stretches
// s_q = K_q + d_q / (alpha_q + i w):
//   diagonal j = l: s_m s_n / s_l (m < n the other axes), memories
//     conv(g, alpha_m), conv(g, alpha_n), conv(conv(g, alpha_m), alpha_n),
//     conv(s_m s_n g, beta_l);
//   j != l: the stretch of the third axis, K g + d conv(g, alpha).
// S_ij = lambda delta_ij sum_k G^(j)_kk + mu (G^(j)_ij + G^(j)_ji) replaces the
// stress, and each point subtracts rho J w L, the inertia term of
// (i w)^2 s_x s_y s_z beyond K_x K_y K_z u'', from its residual, with u the
// predicted displacement and v = w_n = v_n + dt/2 a_n; the rest of the
// velocity terms at v_{n+1} (dt/2 a_{n+1} times their coefficient) is in the
// lumped mass, as is K_x K_y K_z.
// AX: the axes (bit 0 x, 1 y, 2 z) whose d is nonzero at some point of the
// launch's elements.  The memories of the other axes are known to stay zero
// at compile time, so they hold no registers (a single-axis layer keeps 12 of
// the 45).  The arithmetic is the same for every AX: the terms of an axis
// with d = 0 are exact zeros either way, so the residuals do not depend on
// how elements are grouped.
template <int MODE, int AX>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_element_cpml(ElemArgs A, int e0, int ne) {
    __shared__ float sD[25];
    __shared__ float su[3][NP];
    __shared__ float sf[9][NP];
    __shared__ float scf[CPML_NC][128];  // each lane's own coefficients
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int p = tid;
    const bool valid = p < NP && blk < ne;
    const int e = e0 + (valid ? blk : 0);
    const int pc = valid ? p : 0;
    const int i = pc / 25, j = (pc / 5) % 5, k = pc % 5;
    float g[9];
    const CpmlCoef cf{scf, p, 0};
    float lam = 0.f, mu = 0.f;
    uint32_t nodef = 0;
    const int64_t pt = int64_t(e) * NP + pc;
    int64_t q = 0;
    if (valid) {
        load_geometry<false>(A, pt, g);
        lam = A.lam[pt];
        mu = A.mu[pt];
        nodef = A.ibool[pt];
        q = int64_t(A.pml_slot[e]) * NP + p;
#pragma unroll
        for (int c = 0; c < CPML_NC; ++c) scf[c][p] = A.pcoef[c * A.Ppml + q];
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = 0.f;
#pragma unroll
        for (int c = 0; c < CPML_NC; ++c) scf[c][p] = 0.f;
    }
    const int64_t node = int64_t(nodef & ~FIRST_TOUCH);
    const bool first = (nodef & FIRST_TOUCH) != 0;
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    const int step = A.ctr[0];
    const int64_t rslot = residual_entry<MODE>(A, pt, valid);
    // a memory whose axis has d = 0 has zero recursion coefficients and stays
    // zero: it is neither read nor written (single-axis layers use 12 of 45)
    const bool on[3] = {(AX & 1) && cf[1] != 0.f, (AX & 2) && cf[10] != 0.f, (AX & 4) && cf[19] != 0.f};
    __syncthreads();
    // rows of D read from LDS where they are used (registers go to the memories)
    const float* Di = &sD[i * 5];
    const float* Dj = &sD[j * 5];
    const float* Dk = &sD[k * 5];
    for (int f = 0; f < A.nf; ++f) {
        const int64_t fo = int64_t(f) * A.fstride;
        if (valid) {
            const float* ut = A.U + fo + node * 3;
            su[0][p] = ut[0];
            su[1][p] = ut[1];
            su[2][p] = ut[2];
        }
        float rold[3] = {0.f, 0.f, 0.f};
        if (MODE == ASM_COLORED && valid && !first) {
            const float* r = A.R + fo + node * 3;
            rold[0] = r[0];
            rold[1] = r[1];
            rold[2] = r[2];
        }
        // this field's memories in use, loaded together before the barrier
        float* mem = A.pmem + int64_t(f) * CPML_NM * A.Ppml + q;
        float Qm[CPML_NM];
#pragma unroll
        for (int t = 0; t < CPML_NM; ++t) Qm[t] = (valid && cpml_mem_on(on, t)) ? mem[int64_t(t) * A.Ppml] : 0.f;
        __syncthreads();
        if (valid) {
            float du[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    a0 += Di[s] * su[c][s * 25 + j * 5 + k];
                    a1 += Dj[s] * su[c][i * 25 + s * 5 + k];
                    a2 += Dk[s] * su[c][i * 25 + j * 5 + s];
                }
                du[c][0] = a0;
                du[c][1] = a1;
                du[c][2] = a2;
            }
            float G[3][9];  // G[j][k*3+l]: the (scaled) gradient stress column j sees
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    const int m = l == 0 ? 1 : 0, n = l == 2 ? 1 : 2;
                    const CpmlCoef cm{scf, p, m * 9}, cn{scf, p, n * 9}, cl{scf, p, l * 9};
                    const float gv = du[kk][0] * g[l * 3 + 0] + du[kk][1] * g[l * 3 + 1] + du[kk][2] * g[l * 3 + 2];
                    float* Qk = Qm + (kk * 3 + l) * 4;
                    const float am[3] = {cm[3], cm[4], cm[5]}, an[3] = {cn[3], cn[4], cn[5]},
                                bl[3] = {cl[6], cl[7], cl[8]};
                    const float Mm = conv_step(am, Qk[0], gv);
                    const float Mn = conv_step(an, Qk[1], gv);
                    const float Mmn = conv_step(an, Qk[2], Mm);
                    const float y2 = cm[0] * cn[0] * gv + cn[0] * cm[1] * Mm + cn[1] * (cm[0] * Mn + cm[1] * Mmn);
                    const float Mb = conv_step(bl, Qk[3], y2);
                    G[l][kk * 3 + l] = (y2 - cl[1] / cl[0] * Mb) / cl[0];
                    G[n][kk * 3 + l] = cm[0] * gv + cm[1] * Mm;
                    G[m][kk * 3 + l] = cn[0] * gv + cn[1] * Mn;
                }
            float S[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    S[a][b] = mu * (G[b][a * 3 + b] + G[b][b * 3 + a]);
                    if (a == b) S[a][b] += lam * (G[b][0] + G[b][4] + G[b][8]);
                }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    sf[c * 3 + a][p] = S[c][0] * g[0 * 3 + a] + S[c][1] * g[1 * 3 + a] + S[c][2] * g[2 * 3 + a];
        }
        __syncthreads();
        if (valid) {
            float rv[3];
            {
                float Dti[5], Dtj[5], Dtk[5];
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    Dti[s] = sD[s * 5 + i];
                    Dtj[s] = sD[s * 5 + j];
                    Dtk[s] = sD[s * 5 + k];
                }
                point_div(&sf[0][0], NP, Dti, Dtj, Dtk, i, j, k, rv);
            }
            const CpmlCoef cx{scf, p, 0}, cy{scf, p, 9}, cz{scf, p, 18};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float u = su[c][p], v = A.W[fo + node * 3 + c];
                float* Qc = Qm + 36 + c * 3;
                const float ax[3] = {cx[3], cx[4], cx[5]}, ay[3] = {cy[3], cy[4], cy[5]},
                            az[3] = {cz[3], cz[4], cz[5]};
                const float C1 = conv_step(ax, Qc[0], u);
                const float z1 = cx[0] * u + cx[1] * C1;
                const float C2 = conv_step(ay, Qc[1], z1);
                const float z2 = cy[0] * z1 + cy[1] * C2;
                const float C3 = conv_step(az, Qc[2], z2);
                const float L1 = cx[1] * (v - cx[2] * u + cx[2] * cx[2] * C1);
                const float Dz1 = cx[0] * v + cx[1] * (u - cx[2] * C1);
                const float L2 = cy[0] * L1 + cy[1] * (Dz1 - cy[2] * z1 + cy[2] * cy[2] * C2);
                const float Dz2 = cy[0] * Dz1 + cy[1] * (z1 - cy[2] * C2);
                const float L3 = cz[0] * L2 + cz[1] * (Dz2 - cz[2] * z2 + cz[2] * cz[2] * C3);
                rv[c] -= cf[27] * L3;
            }
#pragma unroll
            for (int t = 0; t < CPML_NM; ++t)
                if (cpml_mem_on(on, t)) mem[int64_t(t) * A.Ppml] = Qm[t];
            add_sources(A, step, f, sb, se, p, rv);
            emit_residual<MODE>(A, f, node, rslot, rold, rv);
        }
        __syncthreads();  // su and sf are rewritten by the next field
    }
}

// Block kernel, two fields per pass: the same per-point work as k_element
// (nodal C-PML only), with fields f and f+1 staged and computed together, so
// each pass has two independent dependency chains per lane and one pair of
// workgroup barriers serves two fields.
template <int EPB, int MODE, bool DIAG>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64)
    k_element_pair(ElemArgs A, int e0, int ne) {
    __shared__ float sD[25];
    __shared__ float su[2][EPB][3][NP];
    __shared__ float sf[2][EPB][9][NP];
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
    uint32_t nodef = 0;
    const int64_t pt = int64_t(e) * NP + p;
    if (valid) {
        load_geometry<DIAG>(A, pt, g);
        lam = __builtin_nontemporal_load(&A.lam[pt]);
        mu = __builtin_nontemporal_load(&A.mu[pt]);
        nodef = __builtin_nontemporal_load(&A.ibool[pt]);
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) g[c] = 0.f;
    }
    const int64_t node = int64_t(nodef & ~FIRST_TOUCH);
    const bool first = (nodef & FIRST_TOUCH) != 0;
    const int64_t rslot = residual_entry<MODE>(A, pt, valid);
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    const int step = A.ctr[0];
    const int nf = A.nf;
    float un[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};  // fields f, f+1 (prefetched a pass ahead)
    if (valid) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (h < nf) {
                const float* ut = A.U + int64_t(h) * A.fstride + node * 3;
                un[h][0] = ut[0];
                un[h][1] = ut[1];
                un[h][2] = ut[2];
            }
    }
    __syncthreads();
    float Di[5], Dj[5], Dk[5], Dti[5], Dtj[5], Dtk[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        Di[q] = sD[i * 5 + q];
        Dj[q] = sD[j * 5 + q];
        Dk[q] = sD[k * 5 + q];
        Dti[q] = sD[q * 5 + i];
        Dtj[q] = sD[q * 5 + j];
        Dtk[q] = sD[q * 5 + k];
    }

    for (int f = 0; f < nf; f += 2) {
        const bool two = f + 1 < nf;
        if (valid) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                su[h][le][0][p] = un[h][0];
                su[h][le][1][p] = un[h][1];
                su[h][le][2][p] = un[h][2];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (f + 2 + h < nf) {
                    const float* ut = A.U + int64_t(f + 2 + h) * A.fstride + node * 3;
                    un[h][0] = ut[0];
                    un[h][1] = ut[1];
                    un[h][2] = ut[2];
                }
        }
        float rold[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        if (MODE == ASM_COLORED && valid && !first) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (h == 0 || two) {
                    const float* r = A.R + int64_t(f + h) * A.fstride + node * 3;
                    rold[h][0] = r[0];
                    rold[h][1] = r[1];
                    rold[h][2] = r[2];
                }
        }
        __syncthreads();
        if (valid) {
            float F0[9], F1[9];
            point_flux<DIAG>(su[0][lel][0], su[0][lel][1], su[0][lel][2], Di, Dj, Dk, i, j, k, g, lam, mu, F0);
            point_flux<DIAG>(su[1][lel][0], su[1][lel][1], su[1][lel][2], Di, Dj, Dk, i, j, k, g, lam, mu, F1);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                sf[0][le][q][p] = F0[q];
                sf[1][le][q][p] = F1[q];
            }
        }
        __syncthreads();
        if (valid) {
            float rv0[3], rv1[3];
            point_div(&sf[0][lel][0][0], NP, Dti, Dtj, Dtk, i, j, k, rv0);
            point_div(&sf[1][lel][0][0], NP, Dti, Dtj, Dtk, i, j, k, rv1);
            add_sources(A, step, f, sb, se, p, rv0);
            emit_residual<MODE>(A, f, node, rslot, rold[0], rv0);
            if (two) {
                add_sources(A, step, f + 1, sb, se, p, rv1);
                emit_residual<MODE>(A, f + 1, node, rslot, rold[1], rv1);
            }
        }
    }
}

// Pencil kernel: ten elements per 256-lane workgroup (250 lanes busy); lane
// (element, i, j) owns the five points (i, j, 0..4) of its element, a pencil
// along k.  The k-direction contractions of the gradient and of the
// divergence use the lane's own registers; the i and j directions read the
// other pencils of the element through LDS.  Per point and field this stages
// 9 floats in LDS and reads 60 (the per-point forms stage 12 and read 90), and
// geometry, material and node ids stay in registers for all fields.  The
// arithmetic per point is flux_core / div_core, so the residuals equal the
// other forms' bit for bit.  Nodal C-PML only.
constexpr int PENCIL_EPB = 10;
template <int MODE, bool DIAG>
__global__ void __launch_bounds__(256) k_element_pencil(ElemArgs A, int e0, int ne) {
    constexpr int EPB = PENCIL_EPB;
    __shared__ float sD[25];
    __shared__ float su[EPB][3][NP];  // displacement of the element's points
    __shared__ float s0[EPB][3][NP];  // F_c0 (xi-direction flux) of the points
    __shared__ float s1[EPB][3][NP];  // F_c1 (eta-direction flux)
    const int tid = threadIdx.x;
    if (tid < 25) sD[tid] = A.D[tid];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int le = tid / 25;
    const int q = tid - le * 25;  // pencil (i, j): q = i * 5 + j
    const int eloc = blk * EPB + le;
    const bool valid = (le < EPB) && (eloc < ne);
    const int lel = valid ? le : 0;
    const int e = e0 + (valid ? eloc : 0);
    const int i = q / 5, j = q - (q / 5) * 5;
    const int p0 = q * 5;  // first point of the pencil
    const int64_t pt0 = int64_t(e) * NP + p0;

    float g[5][9], lam[5], mu[5];
    uint32_t nodef[5];
    int64_t rs[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (valid) {
            load_geometry<DIAG>(A, pt0 + k, g[k]);
            lam[k] = __builtin_nontemporal_load(&A.lam[pt0 + k]);
            mu[k] = __builtin_nontemporal_load(&A.mu[pt0 + k]);
            nodef[k] = __builtin_nontemporal_load(&A.ibool[pt0 + k]);
            rs[k] = residual_entry<MODE>(A, pt0 + k, true);
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) g[k][c] = 0.f;
            lam[k] = mu[k] = 0.f;
            nodef[k] = 0;
        }
    }
    int sb = 0, se = 0;
    if (valid) {
        sb = A.src_off[e];
        se = A.src_off[e + 1];
    }
    const int step = A.ctr[0];
    const int nf = A.nf;
    float un[5][3];  // displacement of the pencil, loaded a field ahead
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) un[k][c] = 0.f;
    if (valid) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const float* ut = A.U + int64_t(nodef[k] & ~FIRST_TOUCH) * 3;
            un[k][0] = ut[0];
            un[k][1] = ut[1];
            un[k][2] = ut[2];
        }
    }
    __syncthreads();
    float Di[5], Dj[5], Dti[5], Dtj[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        Di[s] = sD[i * 5 + s];
        Dj[s] = sD[j * 5 + s];
        Dti[s] = sD[s * 5 + i];
        Dtj[s] = sD[s * 5 + j];
    }

    // Two barriers per field: (1) after the displacement is staged, (2) after
    // the fluxes are staged.  The next field's displacement is staged after (2),
    // when every lane has finished reading su; the next fluxes are written after
    // the next (1), when every lane has finished reading s0 and s1.
    for (int f = 0; f < nf; ++f) {
        const int64_t fo = int64_t(f) * A.fstride;
        if (valid) {
#pragma unroll
            for (int k = 0; k < 5; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c) su[lel][c][p0 + k] = un[k][c];
        }
        __syncthreads();
        float F2[5][3];  // zeta-direction fluxes stay in registers
        if (valid) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                float ui[3][5], uj[3][5], uk[3][5], Dk[5];
#pragma unroll
                for (int s = 0; s < 5; ++s) {
                    Dk[s] = A.D[k * 5 + s];  // uniform: scalar operands
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        ui[c][s] = su[lel][c][s * 25 + j * 5 + k];
                        uj[c][s] = su[lel][c][i * 25 + s * 5 + k];
                        uk[c][s] = un[s][c];
                    }
                }
                float F[9];
                flux_core<DIAG>(ui, uj, uk, Di, Dj, Dk, g[k], lam[k], mu[k], F);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    s0[lel][c][p0 + k] = F[c * 3 + 0];
                    s1[lel][c][p0 + k] = F[c * 3 + 1];
                    F2[k][c] = F[c * 3 + 2];
                }
            }
            // the next field's displacement, loaded while this one finishes
            if (f + 1 < nf) {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const float* ut = A.U + fo + A.fstride + int64_t(nodef[k] & ~FIRST_TOUCH) * 3;
                    un[k][0] = ut[0];
                    un[k][1] = ut[1];
                    un[k][2] = ut[2];
                }
            }
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                float Dtk[5];
#pragma unroll
                for (int s = 0; s < 5; ++s) Dtk[s] = A.D[s * 5 + k];
                float rv[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float a[5], b[5], ez[5];
#pragma unroll
                    for (int s = 0; s < 5; ++s) {
                        a[s] = s0[lel][c][s * 25 + j * 5 + k];
                        b[s] = s1[lel][c][i * 25 + s * 5 + k];
                        ez[s] = F2[s][c];
                    }
                    rv[c] = div_core(a, b, ez, Dti, Dtj, Dtk);
                }
                add_sources(A, step, f, sb, se, p0 + k, rv);
                const int64_t node = int64_t(nodef[k] & ~FIRST_TOUCH);
                float rold[3] = {0.f, 0.f, 0.f};
                if (MODE == ASM_COLORED && !(nodef[k] & FIRST_TOUCH)) {
                    const float* r = A.R + fo + node * 3;
                    rold[0] = r[0];
                    rold[1] = r[1];
                    rold[2] = r[2];
                }
                emit_residual<MODE>(A, f, node, rs[k], rold, rv);
            }
        }
    }
}

// One wavefront per element, lane l owns points l and l + 64 (l < 61).  All
// exchange between the element's points stays inside one wavefront, so the
// field loop has no workgroup barrier; WPB elements (wavefronts) per block.
template <int WPB, int MODE, bool DIAG>
__device__ __forceinline__ void wave_body(const ElemArgs& A, int e0, int ne, int blk) {
    __shared__ __align__(16) float sDr[5][8];  // rows of D, padded
    __shared__ __align__(16) float sDc[5][8];  // columns of D, padded
    __shared__ float su[WPB][3][128];
    __shared__ float sf[WPB][9][128];
    const int tid = threadIdx.x;
    if (tid < 25) {
        sDr[tid / 5][tid % 5] = A.D[tid];
        sDc[tid % 5][tid / 5] = A.D[tid];
    }
    __syncthreads();  // the only workgroup barrier
    const int lane = tid & 63, wv = tid >> 6;
    const int eloc = blk * WPB + wv;
    if (eloc >= ne) return;  // wave-uniform
    const int e = e0 + eloc;
    const int pq[2] = {lane, lane + 64};
    const bool vq[2] = {true, lane + 64 < NP};
    int iq[2], jq[2], kq[2];
    float g[2][9], lam[2], mu[2];
    int64_t node[2], ptq[2];
    bool first[2];
    const int sb = A.src_off[e], se = A.src_off[e + 1];
    const int step = A.ctr[0];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int p = vq[q] ? pq[q] : 0;
        iq[q] = p / 25;
        jq[q] = (p / 5) % 5;
        kq[q] = p % 5;
        const int64_t pt = int64_t(e) * NP + p;
#pragma unroll
        for (int c = 0; c < 9; ++c) g[q][c] = 0.f;
        if (vq[q]) load_geometry<DIAG>(A, pt, g[q]);
        lam[q] = vq[q] ? __builtin_nontemporal_load(&A.lam[pt]) : 0.f;
        mu[q] = vq[q] ? __builtin_nontemporal_load(&A.mu[pt]) : 0.f;
        const uint32_t nf_ = vq[q] ? __builtin_nontemporal_load(&A.ibool[pt]) : 0u;
        node[q] = int64_t(nf_ & ~FIRST_TOUCH);
        first[q] = (nf_ & FIRST_TOUCH) != 0;
        ptq[q] = residual_entry<MODE>(A, pt, vq[q]);
    }
    float un[2][3];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) un[q][c] = vq[q] ? A.U[node[q] * 3 + c] : 0.f;

    for (int f = 0; f < A.nf; ++f) {
        const int64_t fo = int64_t(f) * A.fstride;
#pragma unroll
        for (int q = 0; q < 2; ++q)
            if (vq[q])
#pragma unroll
                for (int c = 0; c < 3; ++c) su[wv][c][pq[q]] = un[q][c];
        float rold[2][3];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bool nx = vq[q] && f + 1 < A.nf;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                un[q][c] = nx ? A.U[fo + A.fstride + node[q] * 3 + c] : 0.f;
                rold[q][c] = (MODE == ASM_COLORED && vq[q] && !first[q]) ? A.R[fo + node[q] * 3 + c] : 0.f;
            }
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!vq[q]) continue;
            const int i = iq[q], j = jq[q], k = kq[q];
            const float4 di = *reinterpret_cast<const float4*>(&sDr[i][0]);
            const float4 dj = *reinterpret_cast<const float4*>(&sDr[j][0]);
            const float4 dk = *reinterpret_cast<const float4*>(&sDr[k][0]);
            const float Di[5] = {di.x, di.y, di.z, di.w, sDr[i][4]};
            const float Dj[5] = {dj.x, dj.y, dj.z, dj.w, sDr[j][4]};
            const float Dk[5] = {dk.x, dk.y, dk.z, dk.w, sDr[k][4]};
            float F[9];
            point_flux<DIAG>(su[wv][0], su[wv][1], su[wv][2], Di, Dj, Dk, i, j, k, g[q], lam[q], mu[q], F);
#pragma unroll
            for (int c = 0; c < 9; ++c) sf[wv][c][pq[q]] = F[c];
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!vq[q]) continue;
            const int i = iq[q], j = jq[q], k = kq[q];
            const float4 ci = *reinterpret_cast<const float4*>(&sDc[i][0]);
            const float4 cj = *reinterpret_cast<const float4*>(&sDc[j][0]);
            const float4 ck = *reinterpret_cast<const float4*>(&sDc[k][0]);
            const float Dti[5] = {ci.x, ci.y, ci.z, ci.w, sDc[i][4]};
            const float Dtj[5] = {cj.x, cj.y, cj.z, cj.w, sDc[j][4]};
            const float Dtk[5] = {ck.x, ck.y, ck.z, ck.w, sDc[k][4]};
            float rv[3];
            point_div(&sf[wv][0][0], 128, Dti, Dtj, Dtk, i, j, k, rv);
            add_sources(A, step, f, sb, se, pq[q], rv);
            emit_residual<MODE>(A, f, node[q], ptq[q], rold[q], rv);
        }
    }
}

template <int WPB, int MODE, bool DIAG>
__global__ void __launch_bounds__(64 * WPB) k_element_wave(ElemArgs A, int e0, int ne) {
    wave_body<WPB, MODE, DIAG>(A, e0, ne, xcd_block(blockIdx.x, gridDim.x));
}

// One launch for the element ranges of several solvers, wavefront form.
template <int WPB, int MODE, bool DIAG>
__global__ void __launch_bounds__(64 * WPB)
    k_element_wave_grp(const ElemArgs* __restrict__ tab, const int4* __restrict__ map) {
    const int4 E = map[xcd_block(blockIdx.x, gridDim.x)];
    wave_body<WPB, MODE, DIAG>(const_row(tab, E.x), E.y, E.z, E.w);
}

// Fixed-order sum of a node's element contributions (GATHER assembly): the
// node's first point (its entry at the node id) added to 0, then the other
// points' entries (slots [kb, ke) of the node) in increasing solver order.
__device__ __forceinline__ void gather_r(const NodeArgs& A, int f, int64_t n, int kb, int ke, float* r) {
    const float* Rl = A.Rl + int64_t(f) * A.PR * 3;
    r[0] = 0.f + Rl[n * 3 + 0];
    r[1] = 0.f + Rl[n * 3 + 1];
    r[2] = 0.f + Rl[n * 3 + 2];
    for (int q = kb; q < ke; ++q) {
        const float* c = Rl + int64_t(A.pts[q]) * 3;
        r[0] += c[0];
        r[1] += c[1];
        r[2] += c[2];
    }
}

// LDS-staged fixed-order slot sums for the nodes of one workgroup: the
// residuals of the workgroup's slots [s0, s1) of field f come into LDS in
// chunks (one 12-byte load per slot, spread over the workgroup), then each
// lane adds its nodes' slots in increasing slot order -- the order gather_r
// uses, so the sums are bit-identical.  Every lane of the workgroup must call
// it (it holds workgroup barriers); kb[0..NN] are the lane's node slot bounds.
constexpr int SLOT_CHUNK = 2048;

// Point ids of the workgroup's slots [s0, s1) into LDS when they fit one
// chunk and serve more than one field; staged_sums then reads them there for
// every field.  Returns whether they were staged (uniform over the
// workgroup); the ids become visible at the first barrier inside staged_sums.
__device__ __forceinline__ bool stage_points(const NodeArgs& A, int64_t s0, int64_t s1, int32_t* __restrict__ sp) {
    if (A.nf < 2 || s1 - s0 > SLOT_CHUNK) return false;
    for (int q = threadIdx.x; q < int(s1 - s0); q += blockDim.x) sp[q] = A.pts[s0 + q];
    return true;
}

// main[3 * NN]: the lane's nodes' first-point entries (read by the caller,
// coalesced: they sit at the node ids); the slots hold the other points.
template <int NN>
__device__ __forceinline__ void staged_sums(const NodeArgs& A, int f, int64_t s0, int64_t s1, const int* kb,
                                            bool active, float* __restrict__ sl, const int32_t* __restrict__ sp,
                                            bool ids_staged, const float* main, float* r) {
#pragma unroll
    for (int q = 0; q < 3 * NN; ++q) r[q] = 0.f + main[q];
    const float* Rl = A.Rl + int64_t(f) * A.PR * 3;
    for (int64_t c0 = s0; c0 < s1; c0 += SLOT_CHUNK) {
        const int64_t c1 = min(s1, c0 + SLOT_CHUNK);
        const int cnt = int(c1 - c0);
        __syncthreads();  // the previous chunk has been consumed
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            const int32_t pid = ids_staged ? sp[q] : A.pts[c0 + q];
            const float* src = Rl + int64_t(pid) * 3;
            sl[3 * q + 0] = src[0];
            sl[3 * q + 1] = src[1];
            sl[3 * q + 2] = src[2];
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int n = 0; n < NN; ++n) {
                const int64_t qb = max(int64_t(kb[n]), c0), qe = min(int64_t(kb[n + 1]), c1);
                for (int64_t q = qb; q < qe; ++q) {
                    const float* c = sl + (q - c0) * 3;
                    r[3 * n + 0] += c[0];
                    r[3 * n + 1] += c[1];
                    r[3 * n + 2] += c[2];
                }
            }
        }
    }
}

// Three floats moved as one 12-byte word (one dwordx3 access per lane).
struct Vec3 {
    float v[3];
};

// Special nodes [0, Ga): C-PML terms, legacy damping, massless rule, a_n kept.
// FG field groups per workgroup: the 256 lanes cover 256/FG nodes, and lane
// group fy handles fields fy, fy + FG, ... (each wavefront stays on one field
// group, so its loads remain contiguous over nodes).
template <int FG>
__device__ __forceinline__ void node_special_body(const NodeArgs& A, int blk) {
    constexpr int NPB = 256 / FG;
    // one field group: the workgroup's slot sums are staged through LDS
    constexpr bool STAGE = FG == 1;
    __shared__ float sl[STAGE ? SLOT_CHUNK * 3 : 1];
    __shared__ int32_t sp[STAGE ? SLOT_CHUNK : 1];
    const int fy = threadIdx.x / NPB;
    const int64_t g0 = int64_t(blk) * NPB + (threadIdx.x - fy * NPB);
    const bool active = g0 < A.Ga;
    const bool staged = STAGE && A.gather;  // uniform over the workgroup
    if (!active && !staged) return;
    const int64_t g = active ? g0 : A.Ga - 1;
    const float mi = A.minv[g];
    const float dt = A.dt, hdt = 0.5f * dt, hdt2 = 0.5f * dt * dt;
    const float dmp = A.damp ? A.damp[g] : 0.f;
    const bool pml = A.pml_nodal && g < A.Gpml;
    float C[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, al[9];
    bool use[3] = {false, false, false};
    if (pml) {
#pragma unroll
        for (int q = 0; q < 5; ++q) C[q] = A.pc[q * A.Gpml + g];
#pragma unroll
        for (int q = 0; q < 9; ++q) al[q] = A.palpha[q * A.Gpml + g];
        const uint32_t mk = A.pmask[g];
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) use[dd] = (mk >> dd) & 1u;
    }
    int kb[2] = {0, 0};
    int64_t s0 = 0, s1 = 0;
    bool ids = false;
    if (A.gather) {
        kb[0] = A.off[g];
        kb[1] = A.off[g + 1];
        if (staged) {
            const int64_t gb = int64_t(blk) * NPB;
            s0 = A.off[gb];
            s1 = A.off[min(A.Ga, gb + NPB)];
            ids = stage_points(A, s0, s1, sp);
        }
    }
    for (int f = fy; f < A.nf; f += FG) {
        const int64_t o = int64_t(f) * A.fstride + g * 3;
        const int64_t oa = int64_t(f) * A.astride + g * 3;
        float* pmem = A.pmem + int64_t(f) * 9 * A.Gpml;
        float* pnew = A.pnew + int64_t(f) * 3 * A.Gpml;
        float U[3], W[3], R[3], an[3], un[3], vn[3];
        // this field's node state first, so that its loads overlap the slot sums;
        // C-PML memory of unused directions is zero and stays zero (alpha = (1, 0, 0),
        // C = 0): neither read nor written
        float m[9], pn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            U[c] = A.U[o + c];
            W[c] = A.W[o + c];
            an[c] = A.Aacc[oa + c];
            pn[c] = pml ? pnew[c * A.Gpml + g] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = (pml && use[q % 3]) ? pmem[q * A.Gpml + g] : 0.f;
        if (staged) {
            const Vec3 rm = *reinterpret_cast<const Vec3*>(A.Rl + (int64_t(f) * A.PR + g) * 3);
            const float main[3] = {rm.v[0], rm.v[1], rm.v[2]};
            staged_sums<1>(A, f, s0, s1, kb, active, sl, sp, ids, main, R);
            if (!active) continue;
        } else if (A.gather) {
            gather_r(A, f, g, kb[0], kb[1], R);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) R[c] = A.R[o + c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            un[c] = U[c] - dt * W[c];
            vn[c] = W[c] - hdt * an[c];
        }
        if (pml) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float oldv = pn[c] + A.c2 * an[c];  // pml.cpp:91-109
                const float newv = un[c] + A.c1 * vn[c];                 // pml.cpp:72-89
                R[c] += C[0] * vn[c] + C[1] * un[c] + C[2] * m[c * 3 + 0] + C[3] * m[c * 3 + 1] +
                        C[4] * m[c * 3 + 2];  // pml.cpp:209-212, assembled over PML elements
#pragma unroll
                for (int d = 0; d < 3; ++d)  // pml.cpp:133-146
                    if (use[d])
                        pmem[(c * 3 + d) * A.Gpml + g] =
                            al[d * 3 + 0] * m[c * 3 + d] + al[d * 3 + 1] * newv + al[d * 3 + 2] * oldv;
                pnew[c * A.Gpml + g] = newv;
            }
        }
        Vec3 nu, nw, na;  // stored as 12-byte words
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // legacy ramp v -= d v before the corrector (solver.cpp:565-573)
            const float vp = (dmp > 0.f) ? vn[c] - dmp * vn[c] : vn[c];
            float u1, a1, v1;
            if (mi > 0.f) {
                u1 = (dmp > 0.f) ? un[c] + dt * vp + hdt2 * an[c] : U[c];
                a1 = R[c] * mi;
                v1 = vp + hdt * (an[c] + a1);
            } else {  // m <= 0: a = 0, u and v kept (solver.cpp:57-61)
                u1 = un[c];
                a1 = 0.f;
                v1 = vp;
            }
            const float w1 = v1 + hdt * a1;
            nu.v[c] = u1 + dt * w1;
            nw.v[c] = w1;
            na.v[c] = a1;
        }
        *reinterpret_cast<Vec3*>(A.U + o) = nu;
        *reinterpret_cast<Vec3*>(A.W + o) = nw;
        *reinterpret_cast<Vec3*>(A.Aacc + oa) = na;
    }
}

template <int FG>
__global__ void __launch_bounds__(256) k_node_special(NodeArgs A) { node_special_body<FG>(A, blockIdx.x); }

template <int FG>
__global__ void __launch_bounds__(256)
    k_node_special_grp(const NodeArgs* __restrict__ tab, const int4* __restrict__ map) {
    const int4 E = map[blockIdx.x];
    node_special_body<FG>(const_row(tab, E.x), E.w);
}

// Regular nodes [Ga, G4): a = r/m, w += dt a, u~ += dt w.  4 nodes per lane,
// 16-byte accesses (Ga and the field stride are multiples of 4 nodes).
template <bool GATHER, int FG>
__device__ __forceinline__ void node_regular_body(const NodeArgs& A, int blk) {
    constexpr int NPB = 256 / FG;
    constexpr bool STAGE = GATHER && FG == 1;  // slot sums staged through LDS
    __shared__ float sl[STAGE ? SLOT_CHUNK * 3 : 1];
    __shared__ int32_t sp[STAGE ? SLOT_CHUNK : 1];
    const int fy = threadIdx.x / NPB;
    const int64_t t = int64_t(blk) * NPB + (threadIdx.x - fy * NPB);
    const int64_t g0 = A.Ga + 4 * t;
    const bool active = g0 < A.G4;
    if (!active && !STAGE) return;
    const int64_t g = active ? g0 : A.G4 - 4;
    const float dt = A.dt;
    const float4 mi4 = *reinterpret_cast<const float4*>(A.minv + g);
    const float mi[4] = {mi4.x, mi4.y, mi4.z, mi4.w};
    int kb[5] = {0, 0, 0, 0, 0};
    int64_t s0 = 0, s1 = 0;
    bool ids = false;
    if (GATHER) {
#pragma unroll
        for (int q = 0; q < 5; ++q) kb[q] = A.off[g + q];
        if (STAGE) {
            const int64_t gb = A.Ga + int64_t(blk) * NPB * 4;
            s0 = A.off[gb];
            s1 = A.off[min(A.G4, gb + int64_t(NPB) * 4)];
            ids = stage_points(A, s0, s1, sp);
        }
    }
    for (int f = fy; f < A.nf; f += FG) {
        const int64_t o = int64_t(f) * A.fstride + g * 3;
        float4* U4 = reinterpret_cast<float4*>(A.U + o);
        float4* W4 = reinterpret_cast<float4*>(A.W + o);
        float u[12], w[12], r[12];
        // this field's U and W first, so that their loads overlap the slot sums
        float4 ua[3], wa[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            ua[q] = U4[q];
            wa[q] = W4[q];
        }
        if (STAGE) {
            // the 4 nodes' first-point entries: 12 floats at the node ids, 16-byte aligned
            const float4* m4 = reinterpret_cast<const float4*>(A.Rl + (int64_t(f) * A.PR + g) * 3);
            float main[12];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float4 c = m4[q];
                main[4 * q + 0] = c.x; main[4 * q + 1] = c.y; main[4 * q + 2] = c.z; main[4 * q + 3] = c.w;
            }
            staged_sums<4>(A, f, s0, s1, kb, active, sl, sp, ids, main, r);
            if (!active) continue;
        } else if (GATHER) {
#pragma unroll
            for (int n = 0; n < 4; ++n) gather_r(A, f, g + n, kb[n], kb[n + 1], &r[3 * n]);
        } else {
            const float4* R4 = reinterpret_cast<const float4*>(A.R + o);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float4 c = R4[q];
                r[4 * q + 0] = c.x; r[4 * q + 1] = c.y; r[4 * q + 2] = c.z; r[4 * q + 3] = c.w;
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float4 a = ua[q], b = wa[q];
            u[4 * q + 0] = a.x; u[4 * q + 1] = a.y; u[4 * q + 2] = a.z; u[4 * q + 3] = a.w;
            w[4 * q + 0] = b.x; w[4 * q + 1] = b.y; w[4 * q + 2] = b.z; w[4 * q + 3] = b.w;
        }
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const float a1 = r[q] * mi[q / 3];
            w[q] = w[q] + dt * a1;
            u[q] = u[q] + dt * w[q];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            U4[q] = make_float4(u[4 * q + 0], u[4 * q + 1], u[4 * q + 2], u[4 * q + 3]);
            W4[q] = make_float4(w[4 * q + 0], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
        }
    }
}

template <bool GATHER, int FG>
__global__ void __launch_bounds__(256) k_node_regular(NodeArgs A) {
    node_regular_body<GATHER, FG>(A, blockIdx.x);
}

template <bool GATHER, int FG>
__global__ void __launch_bounds__(256)
    k_node_regular_grp(const NodeArgs* __restrict__ tab, const int4* __restrict__ map) {
    const int4 E = map[blockIdx.x];
    node_regular_body<GATHER, FG>(const_row(tab, E.x), E.w);
}

// Fill kernel (used before atomic assembly; no memset nodes in a step).
__global__ void __launch_bounds__(256) k_zero(float4* __restrict__ p, int64_t n4) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n4) p[t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Strain of the recording cells (forward/share/src/solver.cpp:282-335) from
// u_{n+1} = U - dt W, and copies of u, v, a at their points (solver.cpp:656-674).
template <int EPB>
__global__ void __launch_bounds__(((EPB * NP + 63) / 64) * 64) k_record(RecArgs A) {
    __shared__ float sD[25];
    __shared__ float su[EPB][3][NP];
    // snapshot steps only (step % stride == 0), into ring slot n_snap % cap;
    // the condition is uniform over the grid
    const int st = A.ctr[0];
    if (A.stride <= 0 || A.cap <= 0 || st % A.stride != 0) return;
    const int64_t nrp_all = int64_t(A.n_rec) * NP;
    const int64_t slot = A.ctr[1] % A.cap;
    float* const strain_slot = A.strain + slot * A.nf * nrp_all * 6;
    float* const uva_slot = A.uva ? A.uva + slot * A.nf * nrp_all * 9 : nullptr;
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
    const int64_t node = valid ? A.ibool[pt] : 0;
    float dd[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) dd[c] = valid ? A.dd[c * nrp + pt] : 0.f;
    __syncthreads();
    for (int f = 0; f < A.nf; ++f) {
        const int64_t go = int64_t(f) * A.fstride + node * 3;
        const int64_t ga = int64_t(f) * A.astride + node * 3;
        float w[3] = {0.f, 0.f, 0.f};
        if (valid) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                w[c] = A.W[go + c];
                su[le][c][p] = A.U[go + c] - A.dt * w[c];
            }
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
            float* s6 = strain_slot + (int64_t(f) * nrp + pt) * 6;
            s6[0] = gr[0][0];
            s6[1] = gr[1][1];
            s6[2] = gr[2][2];
            s6[3] = 0.5f * (gr[0][1] + gr[1][0]);
            s6[4] = 0.5f * (gr[0][2] + gr[2][0]);
            s6[5] = 0.5f * (gr[1][2] + gr[2][1]);
            if (uva_slot) {
                float* o9 = uva_slot + (int64_t(f) * nrp + pt) * 9;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float a = A.Aacc[ga + c];
                    o9[c] = su[lel][c][p];
                    o9[3 + c] = w[c] - 0.5f * A.dt * a;
                    o9[6 + c] = a;
                }
            }
        }
        if (f + 1 < A.nf) __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stage kernels of the per-stage entry points (the reference's device branch,
// forward/share/src/solver.cpp:390-444, driven stage by stage).

// Assembly of the element-local residuals into R (GATHER): R[f][n] = the
// node's slots summed in increasing slot order, starting from 0 -- the sum the
// node kernels form in registers, so a corrector reading R gives the same
// bits.  One lane per node of [0, G4); nodes without slots get 0.
__global__ void __launch_bounds__(256) k_scatter_slots(NodeArgs A) {
    const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (n >= A.G4) return;
    float r[3];
    gather_r(A, f, n, A.off[n], A.off[n + 1], r);
    float* o = A.R + int64_t(f) * A.fstride + n * 3;
    o[0] = r[0];
    o[1] = r[1];
    o[2] = r[2];
}

// Point source of one field added to the element stage's residuals:
// r[cell point p][dir] += stf * w[c][p] (forward/share/src/cuda_step.cu:44-60).
// One workgroup; lane p takes point p of each cell in turn, with a barrier
// between cells, so points of different cells at one node never race (COLORED
// and ATOMIC hold assembled R) and the additions follow the cell order, as
// the element kernels' own source loop does.  GATHER: into Rl at the point.
__global__ void __launch_bounds__(128)
    k_inject(float* __restrict__ Rl, const int32_t* __restrict__ rdst, float* __restrict__ R,
             const uint32_t* __restrict__ ibool, const int32_t* __restrict__ cells, const float* __restrict__ w,
             int n_cell, int dir, float stf, int64_t PR, int64_t fstride, int f, int gather) {
    const int p = threadIdx.x;
    for (int c = 0; c < n_cell; ++c) {
        if (p < NP) {
            const int64_t pt = int64_t(cells[c]) * NP + p;
            float* o = gather ? Rl + (int64_t(f) * PR + rdst[pt]) * 3 + dir
                              : R + int64_t(f) * fstride + int64_t(ibool[pt] & ~FIRST_TOUCH) * 3 + dir;
            *o += stf * w[int64_t(c) * NP + p];
        }
        __syncthreads();
    }
}

}  // namespace dev
}  // namespace gfx
