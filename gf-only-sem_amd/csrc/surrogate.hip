// surrogate.hip -- strain Green's tensor tiles on the device, the GLL
// interpolation judge, and a Fourier-feature coordinate MLP trained on the
// tiles (BF16 MFMA, FP32 accumulation, hand-written forward / backward / Adam).
//
// Judge: greenfun/gll_interpolator.py:117-227 restated: exact-node shortcut
// (<= 1e-6 m), cell lookup by box with 1e-6 m slack (nearest cell centre
// among the boxes that contain the point), reference coordinates from the
// cell's corner box clipped to [-1, 1], tensor-product GLL Lagrange weights
// renormalised to sum 1, weighted sum of the cell's 125 node values.
//
// MLP: (x, y, z, t) scaled to [-1, 1] -> [sin, cos](2 pi B x) with B ~ N(0,
// sigma^2) (n_freq x 4) -> depth x (Linear(width) + GELU) -> Linear(n_out).
// Every product runs on v_mfma_f32_16x16x32_bf16 (LDS-tiled): operands
// rounded to BF16 (round to nearest even), FP32 accumulation; master weights
// and Adam moments stay FP32.
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "device.hpp"
#include "gfx.h"
#include "green.hpp"
#include "gfx_surrogate.h"

namespace gfx {
namespace sur {

constexpr int NP = 125;
constexpr double EXACT_TOL = 1e-6;  // gll_interpolator.py:19

typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
    return uint16_t(u >> 16);
}

// ------------------------------ judge ------------------------------------

__device__ __host__ inline void lagrange5(double xi, double* l) {
    const double x[5] = {-1.0, -0.6546536707079771, 0.0, 0.6546536707079771, 1.0};  // sqrt(3/7)
    for (int i = 0; i < 5; ++i) {
        double v = 1.0;
        for (int j = 0; j < 5; ++j)
            if (j != i) v *= (xi - x[j]) / (x[i] - x[j]);
        l[i] = v;
    }
}

struct JudgeArgs {
    int n_node, n_cell, n_time;
    const float* val;
    const int32_t* cell_idx;
    const double* xyz;
    const double* box;
    int nb0, nb1, nb2;
    double b00, b01, b02, bh0, bh1, bh2;
    const int32_t* bin_off;
    const int32_t* bin_cell;
};

// Containing cell of p (box with slack), nearest centre first; -1 when outside.
__device__ inline int find_cell(const JudgeArgs& A, const double p[3]) {
    int bi[3];
    const int nbv[3] = {A.nb0, A.nb1, A.nb2};
    const double b0[3] = {A.b00, A.b01, A.b02}, bh[3] = {A.bh0, A.bh1, A.bh2};
    for (int d = 0; d < 3; ++d) {
        const double q = (p[d] - b0[d]) / bh[d];
        if (q < -1e-9 || q > nbv[d] + 1e-9) return -1;
        bi[d] = min(max(int(floor(q)), 0), nbv[d] - 1);
    }
    const int b = (bi[2] * A.nb1 + bi[1]) * A.nb0 + bi[0];
    int best = -1;
    double bd = 1e300;
    for (int q = A.bin_off[b]; q < A.bin_off[b + 1]; ++q) {
        const int c = A.bin_cell[q];
        const double* bx = A.box + int64_t(c) * 6;
        if (p[0] < bx[0] - EXACT_TOL || p[0] > bx[1] + EXACT_TOL || p[1] < bx[2] - EXACT_TOL ||
            p[1] > bx[3] + EXACT_TOL || p[2] < bx[4] - EXACT_TOL || p[2] > bx[5] + EXACT_TOL)
            continue;
        const double cx = 0.5 * (bx[0] + bx[1]) - p[0], cy = 0.5 * (bx[2] + bx[3]) - p[1],
                     cz = 0.5 * (bx[4] + bx[5]) - p[2];
        const double dd = cx * cx + cy * cy + cz * cz;
        if (dd < bd || (dd == bd && c < best)) {
            bd = dd;
            best = c;
        }
    }
    return best;
}

// Region (cell, time window) of tile time index t: window t * n_window / n_time.
__host__ __device__ inline int64_t region_of(int cell, int t, int n_time, int n_window) {
    const int w = (t >= 0 && t < n_time) ? int(int64_t(t) * n_window / n_time) : 0;
    return int64_t(cell) * n_window + w;
}

// only (may be null): evaluate just the points whose region is marked there
__global__ void k_judge(JudgeArgs A, int n, const double* __restrict__ q, const int32_t* __restrict__ tix,
                        const uint8_t* __restrict__ only, int n_window, float* __restrict__ out,
                        int32_t* __restrict__ cell_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    const int c = find_cell(A, p);
    if (cell_out) cell_out[i] = c;
    const int t = tix[i];
    if (only && (c < 0 || !only[region_of(c, t, A.n_time, n_window)])) return;
    float* o = out + int64_t(i) * 18;
    if (c < 0 || t < 0 || t >= A.n_time) {
        for (int k = 0; k < 18; ++k) o[k] = __builtin_nanf("");
        return;
    }
    const int32_t* ci = A.cell_idx + int64_t(c) * NP;
    const float* vt = A.val + int64_t(t) * A.n_node * 18;
    // exact-node shortcut (gll_interpolator.py:172-177)
    for (int k = 0; k < NP; ++k) {
        const double* x = A.xyz + int64_t(ci[k]) * 3;
        const double dx = x[0] - p[0], dy = x[1] - p[1], dz = x[2] - p[2];
        if (dx * dx + dy * dy + dz * dz < EXACT_TOL * EXACT_TOL) {
            const float* v = vt + int64_t(ci[k]) * 18;
            for (int m = 0; m < 18; ++m) o[m] = v[m];
            return;
        }
    }
    const double* bx = A.box + int64_t(c) * 6;
    double xi[3];
    for (int d = 0; d < 3; ++d) {
        const double lo = bx[2 * d], hi = bx[2 * d + 1];
        xi[d] = fmin(1.0, fmax(-1.0, 2.0 * (p[d] - lo) / (hi - lo) - 1.0));
    }
    double lx[5], ly[5], lz[5];
    lagrange5(xi[0], lx);
    lagrange5(xi[1], ly);
    lagrange5(xi[2], lz);
    double sw = 0.0;
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b)
            for (int k = 0; k < 5; ++k) sw += lx[a] * ly[b] * lz[k];
    double acc[18];
    for (int m = 0; m < 18; ++m) acc[m] = 0.0;
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b)
            for (int k = 0; k < 5; ++k) {
                const double w = lx[a] * ly[b] * lz[k] / sw;
                const float* v = vt + int64_t(ci[(a * 5 + b) * 5 + k]) * 18;
                for (int m = 0; m < 18; ++m) acc[m] += w * double(v[m]);
            }
    for (int m = 0; m < 18; ++m) o[m] = float(acc[m]);
}

JudgeArgs judge_args(const gfx_green& G) {
    JudgeArgs a{};
    a.n_node = G.n_node;
    a.n_cell = G.n_cell;
    a.n_time = G.n_time;
    a.val = G.val.p;
    a.cell_idx = G.cell_idx.p;
    a.xyz = G.xyz.p;
    a.box = G.box.p;
    a.nb0 = G.nb[0];
    a.nb1 = G.nb[1];
    a.nb2 = G.nb[2];
    a.b00 = G.b0[0];
    a.b01 = G.b0[1];
    a.b02 = G.b0[2];
    a.bh0 = G.bh[0];
    a.bh1 = G.bh[1];
    a.bh2 = G.bh[2];
    a.bin_off = G.bin_off.p;
    a.bin_cell = G.bin_cell.p;
    return a;
}

// -------------------------------- MLP -------------------------------------

// C = op(A) op(B) with A(m,k) = TA ? A[k*lda+m] : A[m*lda+k] and
// B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n], on gfx950's
// v_mfma_f32_16x16x32_bf16, tiled through LDS: a workgroup of four wavefronts computes a 128 x 128 tile of C (each wavefront
// 64 x 64 = 4 x 4 MFMA tiles, 16 accumulators of 4 floats), stepping K by 32.
// Each step stages the A (128 x 32) and B (32 x 128) slices in LDS as BF16
// (round to nearest even), both k-contiguous per row / column, so every MFMA
// operand is one 16-byte LDS read; rows are padded to 40 BF16 (80 bytes), so
// the 16 lanes of a read phase hit distinct banks.  Lane layout of the 16 x 16
// x 32 BF16 MFMA: lane l feeds A row l % 16 and B column l % 16 with k =
// 8 (l / 16) .. + 7, and receives D rows 4 (l / 16) .. + 3 of column l % 16.
// The K range is split over gridDim.z (kc a multiple of 32); split z writes
// C + z*M*ldc, summed later in a fixed order.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int GT = 128, GK = 32, GKP = 40;

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_lds(const float* __restrict__ A, int lda, const float* __restrict__ B,
                                                  int ldb, float* __restrict__ C, int ldc, int M, int N, int K,
                                                  int kc) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[GT * GKP];
    __shared__ __attribute__((aligned(16))) uint16_t sB[GT * GKP];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int mb = blockIdx.y * GT, nb = blockIdx.x * GT;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    float* Cz = C + int64_t(blockIdx.z) * M * ldc;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        __syncthreads();  // the previous slices have been read
        // stage: 16 values of A and 16 of B per lane
        if (!TA) {  // A(m, k) = A[m * lda + k]: lane -> row t / 2, k half (t & 1)
            const int m = t >> 1, kk = (t & 1) * 16;
            const int gm = mb + m;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int k = k0 + kk + q;
                const float v = (gm < M && k < kend) ? A[int64_t(gm) * lda + k] : 0.f;
                sA[m * GKP + kk + q] = f2bf(v);
            }
        } else {  // A(m, k) = A[k * lda + m]: lane -> k = t / 8, 16 rows
            const int kk = t >> 3, m0 = (t & 7) * 16;
            const int k = k0 + kk;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int gm = mb + m0 + q;
                const float v = (gm < M && k < kend) ? A[int64_t(k) * lda + gm] : 0.f;
                sA[(m0 + q) * GKP + kk] = f2bf(v);
            }
        }
        if (TB) {  // B(k, n) = B[n * ldb + k]: lane -> column t / 2, k half (t & 1)
            const int n = t >> 1, kk = (t & 1) * 16;
            const int gn = nb + n;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int k = k0 + kk + q;
                const float v = (gn < N && k < kend) ? B[int64_t(gn) * ldb + k] : 0.f;
                sB[n * GKP + kk + q] = f2bf(v);
            }
        } else {  // B(k, n) = B[k * ldb + n]: lane -> k = t / 8, 16 columns
            const int kk = t >> 3, n0 = (t & 7) * 16;
            const int k = k0 + kk;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int gn = nb + n0 + q;
                const float v = (gn < N && k < kend) ? B[int64_t(k) * ldb + gn] : 0.f;
                sB[(n0 + q) * GKP + kk] = f2bf(v);
            }
        }
        __syncthreads();
        bf16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&sA[(wm + i * 16 + r16) * GKP + kq]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[(wn + j * 16 + r16) * GKP + kq]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = mb + wm + i * 16 + (lane >> 4) * 4 + q, n = nb + wn + j * 16 + r16;
                if (m < M && n < N) Cz[int64_t(m) * ldc + n] = acc[i][j][q];
            }
}

// dst[i] = sum_z src[z*n + i], z in order
__global__ void k_sum_splits(const float* __restrict__ src, int splits, int64_t n, float* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += src[int64_t(z) * n + i];
    dst[i] = s;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Z[b][n] += bias[n]; H = gelu(Z) (H may be null for the output layer)
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, float* __restrict__ H, int rows,
                           int cols) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(rows) * cols) return;
    const float z = Z[i] + bias[i % cols];
    Z[i] = z;
    if (H) H[i] = gelu(z);
}

// dZ = dH * gelu'(Z) in place on dH
__global__ void k_gelu_back(float* __restrict__ dH, const float* __restrict__ Z, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dH[i] *= gelu_d(Z[i]);
}

// column sums of X[rows][cols] in a fixed order: partial sums over chunks of
// 256 rows, then summed in chunk order
constexpr int COLSUM_ROWS = 32;  // rows per partial column sum
__global__ void k_colsum_part(const float* __restrict__ X, int rows, int cols, float* __restrict__ part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (c >= cols) return;
    float s = 0.f;
    const int r1 = min(rows, (chunk + 1) * COLSUM_ROWS);
    for (int r = chunk * COLSUM_ROWS; r < r1; ++r) s += X[int64_t(r) * cols + c];
    part[int64_t(chunk) * cols + c] = s;
}

// features: in [rows][n_in] scaled to [-1, 1]; Bf [n_freq][n_in]
__global__ void k_fourier(const float* __restrict__ X, int rows, int n_in, const float* __restrict__ Bf, int n_freq,
                          float* __restrict__ H) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(rows) * n_freq) return;
    const int r = int(i / n_freq), j = int(i % n_freq);
    float v = 0.f;
    for (int d = 0; d < n_in; ++d) v += Bf[j * n_in + d] * X[int64_t(r) * n_in + d];
    v *= 6.283185307179586f;
    H[int64_t(r) * 2 * n_freq + j] = sinf(v);
    H[int64_t(r) * 2 * n_freq + n_freq + j] = cosf(v);
}

// mean squared error over rows x n_out of (out - y), gradient into out
__global__ void k_mse(float* __restrict__ out, const float* __restrict__ y, int64_t n, float inv_n,
                      float* __restrict__ part) {
    __shared__ float sh[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        const float d = out[i] - y[i];
        e = d * d;
        out[i] = 2.f * d * inv_n;
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Output envelope e(r), r = horizontal distance from (cx, cy): piecewise
// linear through env[i] at r_i = (i + 0.5) h, h = r_max / n, constant outside
// [r_0, r_{n-1}]; 1 when n == 0.
// The network's time input is the retarded time t - slowness * r (moveout).
struct Envelope {
    const float* env;  // device or host array of n knots
    int n;
    double cx, cy, r_max, slowness;
};

__host__ __device__ inline double retarded_time(const Envelope& E, double x, double y, double t) {
    if (E.slowness == 0.0) return t;
    return t - E.slowness * sqrt((x - E.cx) * (x - E.cx) + (y - E.cy) * (y - E.cy));
}

__host__ __device__ inline float envelope_at(const Envelope& E, double x, double y) {
    if (E.n <= 0) return 1.f;
    const double r = sqrt((x - E.cx) * (x - E.cx) + (y - E.cy) * (y - E.cy));
    const double u = r * E.n / E.r_max - 0.5;
    if (u <= 0.0) return E.env[0];
    if (u >= E.n - 1) return E.env[E.n - 1];
    const int i = int(u);
    const float f = float(u - i);
    return E.env[i] * (1.f - f) + E.env[i + 1] * f;
}

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// training batch from a tile: random (node, time) pairs; inputs scaled to
// [-1, 1], targets divided by the per-output scale and the envelope
__global__ void k_sample_tile(const float* __restrict__ val, const double* __restrict__ xyz, int n_node, int n_time,
                              double t0, double dt, const double* __restrict__ lo, const double* __restrict__ hi,
                              const float* __restrict__ inv_scale, Envelope E, uint64_t seed, int rows,
                              float* __restrict__ X, float* __restrict__ Y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint64_t h = mix64(seed ^ (uint64_t(r) * 0xD1B54A32D192ED03ULL));
    const int node = int(h % uint64_t(n_node));
    const int t = int((h >> 32) % uint64_t(n_time));
    const double px = xyz[int64_t(node) * 3], py = xyz[int64_t(node) * 3 + 1];
    const double in[4] = {px, py, xyz[int64_t(node) * 3 + 2], retarded_time(E, px, py, t0 + t * dt)};
    for (int d = 0; d < 4; ++d) X[int64_t(r) * 4 + d] = float(2.0 * (in[d] - lo[d]) / (hi[d] - lo[d]) - 1.0);
    const float ie = 1.f / envelope_at(E, px, py);
    const float* v = val + (int64_t(t) * n_node + node) * 18;
    for (int k = 0; k < 18; ++k) Y[int64_t(r) * 18 + k] = v[k] * inv_scale[k] * ie;
}

// raw (x, y, z, t) -> scaled inputs
__global__ void k_scale_in(const double* __restrict__ raw, int rows, const double* __restrict__ lo,
                           const double* __restrict__ hi, float* __restrict__ X) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int d = 0; d < 4; ++d)
        X[int64_t(r) * 4 + d] = float(2.0 * (raw[int64_t(r) * 4 + d] - lo[d]) / (hi[d] - lo[d]) - 1.0);
}

}  // namespace sur
}  // namespace gfx

struct gfx_mlp {
    gfx_mlp_config cfg{};
    int n_feat = 0;
    std::vector<int> lin_in, lin_out;        // per linear layer
    std::vector<int64_t> w_off, b_off;       // offsets in the flat parameter array
    int64_t n_params = 0;
    gfx::DevBuf<float> w, g, m, v, Bf;
    gfx::DevBuf<double> lo, hi;
    gfx::DevBuf<float> inv_scale;
    std::vector<double> h_scale;
    // output envelope (knots on the host and the device)
    std::vector<float> h_env;
    gfx::DevBuf<float> env;
    double env_cx = 0, env_cy = 0, env_rmax = 1, slowness = 0;
    gfx::sur::Envelope envelope(bool device) const {
        return gfx::sur::Envelope{device ? env.p : h_env.data(), int(h_env.size()), env_cx, env_cy, env_rmax,
                                  slowness};
    }
    // activations for up to cap rows
    int cap = 0;
    gfx::DevBuf<float> X, H0, Y, work, part, lossp;
    std::vector<gfx::DevBuf<float>> Z, H;    // per hidden layer
    gfx::DevBuf<float> out, dA, dB;
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;
    ~gfx_mlp() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gfx {
namespace sur {
namespace {

void ensure_rows(gfx_mlp& M, int rows) {
    if (rows <= M.cap) return;
    const int cap = ((rows + 255) / 256) * 256;
    M.X.alloc(size_t(cap) * 4);
    M.H0.alloc(size_t(cap) * M.n_feat);
    M.Y.alloc(size_t(cap) * M.cfg.n_out);
    const int d = M.cfg.depth, wdt = M.cfg.width;
    M.Z.clear();
    M.H.clear();
    M.Z.resize(d);
    M.H.resize(d);
    for (int l = 0; l < d; ++l) {
        M.Z[l].alloc(size_t(cap) * wdt);
        M.H[l].alloc(size_t(cap) * wdt);
    }
    M.out.alloc(size_t(cap) * M.cfg.n_out);
    M.dA.alloc(size_t(cap) * std::max(wdt, M.n_feat));
    M.dB.alloc(size_t(cap) * std::max(wdt, M.n_feat));
    M.part.alloc(size_t(cap / sur::COLSUM_ROWS + 1) * std::max(wdt, 64));
    M.lossp.alloc(size_t(cap) * M.cfg.n_out / 256 + 1);
    M.cap = cap;
}

// C[M][N] = op(A) op(B), split over K when the output grid is small
template <bool TA, bool TB>
void gemm(gfx_mlp& Mm, const float* A, int lda, const float* B, int ldb, float* C, int ldc, int M, int N, int K) {
    const int gx = (N + sur::GT - 1) / sur::GT, gy = (M + sur::GT - 1) / sur::GT;
    int splits = 1;
    while (gx * gy * splits < 256 && K / (splits * 2) >= 256) splits *= 2;
    if (splits == 1) {
        sur::k_gemm_lds<TA, TB><<<dim3(gx, gy, 1), 256, 0, Mm.stream>>>(A, lda, B, ldb, C, ldc, M, N, K, K);
        return;
    }
    const int kc = ((K + splits - 1) / splits + sur::GK - 1) / sur::GK * sur::GK;
    const size_t need = size_t(splits) * M * ldc;
    if (Mm.work.n < need) {
        GFX_HIP(hipStreamSynchronize(Mm.stream));  // the old buffer may still be read
        Mm.work.alloc(need);
    }
    sur::k_gemm_lds<TA, TB><<<dim3(gx, gy, splits), 256, 0, Mm.stream>>>(A, lda, B, ldb, Mm.work.p, ldc, M, N, K,
                                                                         kc);
    const int64_t n = int64_t(M) * ldc;
    sur::k_sum_splits<<<int((n + 255) / 256), 256, 0, Mm.stream>>>(Mm.work.p, splits, n, C);
}

void colsum(gfx_mlp& M, const float* X, int rows, int cols, float* dst) {
    const int chunks = (rows + sur::COLSUM_ROWS - 1) / sur::COLSUM_ROWS;
    sur::k_colsum_part<<<dim3((cols + 63) / 64, chunks), 64, 0, M.stream>>>(X, rows, cols, M.part.p);
    sur::k_sum_splits<<<(cols + 255) / 256, 256, 0, M.stream>>>(M.part.p, chunks, cols, dst);
}

// forward of rows scaled inputs in M.X; output (scaled) in M.out
void forward(gfx_mlp& M, int rows) {
    const int nf = M.cfg.n_freq;
    sur::k_fourier<<<int((int64_t(rows) * nf + 255) / 256), 256, 0, M.stream>>>(M.X.p, rows, 4, M.Bf.p, nf, M.H0.p);
    const float* in = M.H0.p;
    int in_w = M.n_feat;
    const int d = M.cfg.depth;
    for (int l = 0; l < d; ++l) {
        const int ow = M.lin_out[l];
        gemm<false, false>(M, in, in_w, M.w.p + M.w_off[l], ow, M.Z[l].p, ow, rows, ow, in_w);
        sur::k_bias_act<<<int((int64_t(rows) * ow + 255) / 256), 256, 0, M.stream>>>(
            M.Z[l].p, M.w.p + M.b_off[l], M.H[l].p, rows, ow);
        in = M.H[l].p;
        in_w = ow;
    }
    const int no = M.cfg.n_out;
    gemm<false, false>(M, in, in_w, M.w.p + M.w_off[d], no, M.out.p, no, rows, no, in_w);
    sur::k_bias_act<<<int((int64_t(rows) * no + 255) / 256), 256, 0, M.stream>>>(M.out.p, M.w.p + M.b_off[d],
                                                                                   nullptr, rows, no);
}

// loss (mean over rows x n_out of squared scaled error) and gradient into M.g
double backward(gfx_mlp& M, int rows) {
    const int no = M.cfg.n_out, d = M.cfg.depth;
    const int64_t n = int64_t(rows) * no;
    const int nb = int((n + 255) / 256);
    sur::k_mse<<<nb, 256, 0, M.stream>>>(M.out.p, M.Y.p, n, 1.f / float(n), M.lossp.p);
    std::vector<float> hp(nb);
    GFX_HIP(hipMemcpyAsync(hp.data(), M.lossp.p, nb * 4, hipMemcpyDeviceToHost, M.stream));
    // output layer
    const float* dOut = M.out.p;
    const float* Hl = M.H[d - 1].p;
    const int wl = M.lin_in[d];
    gemm<true, false>(M, Hl, wl, dOut, no, M.g.p + M.w_off[d], no, wl, no, rows);
    colsum(M, dOut, rows, no, M.g.p + M.b_off[d]);
    // dH_{d-1} = dOut W_d^T
    float* dH = M.dA.p;
    float* spare = M.dB.p;
    gemm<false, true>(M, dOut, no, M.w.p + M.w_off[d], no, dH, wl, rows, wl, no);
    for (int l = d - 1; l >= 0; --l) {
        const int ow = M.lin_out[l], iw = M.lin_in[l];
        sur::k_gelu_back<<<int((int64_t(rows) * ow + 255) / 256), 256, 0, M.stream>>>(dH, M.Z[l].p,
                                                                                        int64_t(rows) * ow);
        const float* Hin = (l == 0) ? M.H0.p : M.H[l - 1].p;
        gemm<true, false>(M, Hin, iw, dH, ow, M.g.p + M.w_off[l], ow, iw, ow, rows);
        colsum(M, dH, rows, ow, M.g.p + M.b_off[l]);
        if (l > 0) {
            gemm<false, true>(M, dH, ow, M.w.p + M.w_off[l], ow, spare, iw, rows, iw, ow);
            std::swap(dH, spare);
        }
    }
    GFX_HIP(hipStreamSynchronize(M.stream));
    double loss = 0.0;
    for (float v : hp) loss += v;
    return loss / double(n);
}

void adam(gfx_mlp& M, float lr) {
    ++M.adam_t;
    const float b1 = M.cfg.beta1, b2 = M.cfg.beta2;
    const float c1 = 1.f - std::pow(b1, float(M.adam_t)), c2 = 1.f - std::pow(b2, float(M.adam_t));
    sur::k_adam<<<int((M.n_params + 255) / 256), 256, 0, M.stream>>>(M.w.p, M.g.p, M.m.p, M.v.p, M.n_params,
                                                                      lr, b1, b2, M.cfg.eps, c1, c2);
}

}  // namespace
}  // namespace sur
}  // namespace gfx

using namespace gfx;

// ------------------------------ C ABI: tiles -------------------------------

extern "C" gfx_status gfx_green_create(int n_time, int n_node, const double* node_xyz, int n_cell,
                                       const int32_t* cell_node_index, gfx_green** out) {
    return guarded([&] {
        require(out && node_xyz && cell_node_index, "null argument");
        require(n_time >= 1 && n_node >= 1 && n_cell >= 1, "empty tile");
        *out = nullptr;
        auto G = std::make_unique<gfx_green>();
        G->n_time = n_time;
        G->n_node = n_node;
        G->n_cell = n_cell;
        G->h_xyz.assign(node_xyz, node_xyz + size_t(n_node) * 3);
        G->h_cell.assign(cell_node_index, cell_node_index + size_t(n_cell) * sur::NP);
        for (int32_t k : G->h_cell) require(k >= 0 && k < n_node, "cell node index out of range");
        // corner boxes (gll_interpolator.py:80-105)
        const int corner[8] = {0, 4, 20, 24, 100, 104, 120, 124};
        G->h_box.resize(size_t(n_cell) * 6);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        double ext[3] = {0, 0, 0};
        for (int c = 0; c < n_cell; ++c) {
            double* b = &G->h_box[size_t(c) * 6];
            for (int d = 0; d < 3; ++d) {
                b[2 * d] = 1e300;
                b[2 * d + 1] = -1e300;
            }
            for (int k : corner) {
                const double* x = &G->h_xyz[size_t(G->h_cell[size_t(c) * sur::NP + k]) * 3];
                for (int d = 0; d < 3; ++d) {
                    b[2 * d] = std::min(b[2 * d], x[d]);
                    b[2 * d + 1] = std::max(b[2 * d + 1], x[d]);
                }
            }
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], b[2 * d]);
                hi[d] = std::max(hi[d], b[2 * d + 1]);
                ext[d] = std::max(ext[d], b[2 * d + 1] - b[2 * d]);
            }
        }
        // bins about one cell wide
        for (int d = 0; d < 3; ++d) {
            const double span = std::max(hi[d] - lo[d], 1e-9);
            G->nb[d] = std::max(1, std::min(4096, int(span / std::max(ext[d], 1e-9))));
            G->b0[d] = lo[d] - 2 * sur::EXACT_TOL;
            G->bh[d] = (span + 4 * sur::EXACT_TOL) / G->nb[d];
        }
        const int64_t nbins = int64_t(G->nb[0]) * G->nb[1] * G->nb[2];
        std::vector<std::vector<int32_t>> bins(nbins);
        for (int c = 0; c < n_cell; ++c) {
            const double* b = &G->h_box[size_t(c) * 6];
            int r0[3], r1[3];
            for (int d = 0; d < 3; ++d) {
                r0[d] = std::max(0, int(std::floor((b[2 * d] - sur::EXACT_TOL - G->b0[d]) / G->bh[d])));
                r1[d] = std::min(G->nb[d] - 1, int(std::floor((b[2 * d + 1] + sur::EXACT_TOL - G->b0[d]) / G->bh[d])));
            }
            for (int z = r0[2]; z <= r1[2]; ++z)
                for (int y = r0[1]; y <= r1[1]; ++y)
                    for (int x = r0[0]; x <= r1[0]; ++x)
                        bins[(int64_t(z) * G->nb[1] + y) * G->nb[0] + x].push_back(c);
        }
        std::vector<int32_t> off(nbins + 1, 0), cells;
        for (int64_t b = 0; b < nbins; ++b) {
            off[b + 1] = off[b] + int32_t(bins[b].size());
            cells.insert(cells.end(), bins[b].begin(), bins[b].end());
        }
        GFX_HIP(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
        G->val.alloc(size_t(n_time) * n_node * 18);
        G->val.zero(G->stream);
        G->cell_idx.upload(G->h_cell);
        G->xyz.upload(G->h_xyz);
        G->box.upload(G->h_box);
        G->bin_off.upload(off);
        G->bin_cell.upload(cells);
        GFX_HIP(hipStreamSynchronize(G->stream));
        *out = G.release();
    });
}

extern "C" gfx_status gfx_green_destroy(gfx_green* g) {
    return guarded([&] { delete g; });
}

extern "C" gfx_status gfx_green_set_values(gfx_green* g, int t, const float* v) {
    return guarded([&] {
        require(g && v, "null argument");
        require(t >= 0 && t < g->n_time, "time index out of range");
        GFX_HIP(hipMemcpy(g->val.p + int64_t(t) * g->n_node * 18, v, size_t(g->n_node) * 18 * 4,
                          hipMemcpyHostToDevice));
    });
}

extern "C" gfx_status gfx_green_get_values(const gfx_green* g, int t, float* v) {
    return guarded([&] {
        require(g && v, "null argument");
        require(t >= 0 && t < g->n_time, "time index out of range");
        GFX_HIP(hipStreamSynchronize(g->stream));
        GFX_HIP(hipMemcpy(v, g->val.p + int64_t(t) * g->n_node * 18, size_t(g->n_node) * 18 * 4,
                          hipMemcpyDeviceToHost));
    });
}

extern "C" gfx_status gfx_green_interpolate(const gfx_green* g, int n, const double* xyz, const int32_t* tix,
                                            float* out, int32_t* cell) {
    return guarded([&] {
        require(g && xyz && tix && out && n >= 0, "null argument");
        if (n == 0) return;
        DevBuf<double> q;
        DevBuf<int32_t> t, c;
        DevBuf<float> o;
        q.alloc(size_t(n) * 3);
        t.alloc(n);
        o.alloc(size_t(n) * 18);
        if (cell) c.alloc(n);
        GFX_HIP(hipMemcpy(q.p, xyz, size_t(n) * 24, hipMemcpyHostToDevice));
        GFX_HIP(hipMemcpy(t.p, tix, size_t(n) * 4, hipMemcpyHostToDevice));
        sur::k_judge<<<(n + 127) / 128, 128, 0, g->stream>>>(sur::judge_args(*g), n, q.p, t.p, nullptr, 1, o.p,
                                                              cell ? c.p : nullptr);
        GFX_HIP(hipGetLastError());
        GFX_HIP(hipStreamSynchronize(g->stream));
        GFX_HIP(hipMemcpy(out, o.p, size_t(n) * 18 * 4, hipMemcpyDeviceToHost));
        if (cell) GFX_HIP(hipMemcpy(cell, c.p, size_t(n) * 4, hipMemcpyDeviceToHost));
    });
}

// Host FP64 restatement of GLLInterpolator.interpolate for one point over
// node values [n_node][nv]; returns GFX_ERR_INVALID when the point lies
// outside all cells (the reference raises ValueError).
extern "C" gfx_status gfx_gll_interpolate_host(int n_node, const double* node_xyz, int n_cell,
                                               const int32_t* cell_node_index, int nv, const double* values,
                                               const double* point, double* out) {
    return guarded([&] {
        require(node_xyz && cell_node_index && values && point && out && nv >= 1, "null argument");
        const double p[3] = {point[0], point[1], point[2]};
        // exact-node shortcut over all nodes
        int best_node = -1;
        double bd = 1e300;
        for (int k = 0; k < n_node; ++k) {
            const double* x = node_xyz + size_t(k) * 3;
            const double dd = (x[0] - p[0]) * (x[0] - p[0]) + (x[1] - p[1]) * (x[1] - p[1]) + (x[2] - p[2]) * (x[2] - p[2]);
            if (dd < bd) {
                bd = dd;
                best_node = k;
            }
        }
        if (best_node >= 0 && std::sqrt(bd) < sur::EXACT_TOL) {
            std::memcpy(out, values + size_t(best_node) * nv, size_t(nv) * 8);
            return;
        }
        const int corner[8] = {0, 4, 20, 24, 100, 104, 120, 124};
        int cell = -1;
        double cdist = 1e300, cb[6] = {0, 0, 0, 0, 0, 0};
        for (int c = 0; c < n_cell; ++c) {
            double b[6] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300};
            for (int k : corner) {
                const double* x = node_xyz + size_t(cell_node_index[size_t(c) * sur::NP + k]) * 3;
                for (int d = 0; d < 3; ++d) {
                    b[2 * d] = std::min(b[2 * d], x[d]);
                    b[2 * d + 1] = std::max(b[2 * d + 1], x[d]);
                }
            }
            bool in = true;
            for (int d = 0; d < 3; ++d)
                in = in && p[d] >= b[2 * d] - sur::EXACT_TOL && p[d] <= b[2 * d + 1] + sur::EXACT_TOL;
            if (!in) continue;
            double dd = 0;
            for (int d = 0; d < 3; ++d) dd += (0.5 * (b[2 * d] + b[2 * d + 1]) - p[d]) * (0.5 * (b[2 * d] + b[2 * d + 1]) - p[d]);
            if (dd < cdist) {
                cdist = dd;
                cell = c;
                std::memcpy(cb, b, sizeof(b));
            }
        }
        if (cell < 0) throw Error(GFX_ERR_INVALID, "query point is outside all recording cells");
        double xi[3], l[3][5];
        for (int d = 0; d < 3; ++d) {
            xi[d] = std::min(1.0, std::max(-1.0, 2.0 * (p[d] - cb[2 * d]) / (cb[2 * d + 1] - cb[2 * d]) - 1.0));
            sur::lagrange5(xi[d], l[d]);
        }
        double sw = 0;
        for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b)
                for (int k = 0; k < 5; ++k) sw += l[0][a] * l[1][b] * l[2][k];
        for (int m = 0; m < nv; ++m) out[m] = 0.0;
        for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b)
                for (int k = 0; k < 5; ++k) {
                    const double w = l[0][a] * l[1][b] * l[2][k] / sw;
                    const double* v = values + size_t(cell_node_index[size_t(cell) * sur::NP + (a * 5 + b) * 5 + k]) * nv;
                    for (int m = 0; m < nv; ++m) out[m] += w * v[m];
                }
    });
}

// ------------------------------ C ABI: MLP ---------------------------------

extern "C" gfx_status gfx_mlp_config_default(gfx_mlp_config* c) {
    return guarded([&] {
        require(c != nullptr, "null argument");
        *c = gfx_mlp_config{};
        c->n_freq = 64;
        c->width = 256;
        c->depth = 6;
        c->n_out = 18;
        c->sigma = 2.0f;
        c->seed = 5;
        c->lr = 1e-3f;
        c->beta1 = 0.9f;
        c->beta2 = 0.999f;
        c->eps = 1e-8f;
        c->batch = 8192;
        c->lr_final = 5e-5f;
    });
}

extern "C" gfx_status gfx_mlp_create(const gfx_mlp_config* cfg, const double* in_lo, const double* in_hi,
                                     const double* out_scale, gfx_mlp** out) {
    return guarded([&] {
        require(cfg && in_lo && in_hi && out_scale && out, "null argument");
        require(cfg->n_freq >= 1 && cfg->n_freq <= 1024 && cfg->width >= 16 && cfg->width <= 4096 &&
                    cfg->depth >= 1 && cfg->depth <= 32 && cfg->n_out == 18,
                "bad MLP shape (n_out must be 18: 6 strain components x 3 force directions)");
        for (int d = 0; d < 4; ++d) require(in_hi[d] > in_lo[d], "input range must be non-empty");
        require(cfg->lr > 0 && cfg->lr_final > 0 && cfg->lr_final <= cfg->lr, "need 0 < lr_final <= lr");
        *out = nullptr;
        auto M = std::make_unique<gfx_mlp>();
        M->cfg = *cfg;
        M->n_feat = 2 * cfg->n_freq;
        int iw = M->n_feat;
        for (int l = 0; l <= cfg->depth; ++l) {
            const int ow = (l == cfg->depth) ? cfg->n_out : cfg->width;
            M->lin_in.push_back(iw);
            M->lin_out.push_back(ow);
            M->w_off.push_back(M->n_params);
            M->n_params += int64_t(iw) * ow;
            M->b_off.push_back(M->n_params);
            M->n_params += ow;
            iw = ow;
        }
        // init: weights U(-sqrt(6/(in+out)), ...) (Glorot), biases 0, B ~ N(0, sigma^2)
        std::mt19937_64 rng(cfg->seed);
        std::vector<float> hw(M->n_params, 0.f);
        for (size_t l = 0; l < M->lin_in.size(); ++l) {
            const double a = std::sqrt(6.0 / (M->lin_in[l] + M->lin_out[l]));
            std::uniform_real_distribution<double> U(-a, a);
            for (int64_t i = 0; i < int64_t(M->lin_in[l]) * M->lin_out[l]; ++i) hw[M->w_off[l] + i] = float(U(rng));
        }
        std::normal_distribution<double> Nd(0.0, cfg->sigma);
        std::vector<float> hb(size_t(cfg->n_freq) * 4);
        for (auto& x : hb) x = float(Nd(rng));
        GFX_HIP(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
        M->w.upload(hw);
        M->g.alloc(M->n_params);
        M->m.alloc(M->n_params);
        M->v.alloc(M->n_params);
        M->m.zero(M->stream);
        M->v.zero(M->stream);
        M->Bf.upload(hb);
        M->lo.upload(std::vector<double>(in_lo, in_lo + 4));
        M->hi.upload(std::vector<double>(in_hi, in_hi + 4));
        M->h_scale.assign(out_scale, out_scale + 18);
        std::vector<float> inv(18);
        for (int k = 0; k < 18; ++k) {
            require(out_scale[k] > 0, "output scales must be > 0");
            inv[k] = float(1.0 / out_scale[k]);
        }
        M->inv_scale.upload(inv);
        GFX_HIP(hipStreamSynchronize(M->stream));
        *out = M.release();
    });
}

extern "C" gfx_status gfx_mlp_destroy(gfx_mlp* m) {
    return guarded([&] { delete m; });
}

extern "C" gfx_status gfx_mlp_n_params(const gfx_mlp* m, size_t* n) {
    return guarded([&] {
        require(m && n, "null argument");
        *n = size_t(m->n_params);
    });
}

extern "C" gfx_status gfx_mlp_get_params(const gfx_mlp* m, float* w, size_t n, float* fourier) {
    return guarded([&] {
        require(m && w && n == size_t(m->n_params), "parameter buffer must hold n_params floats");
        GFX_HIP(hipStreamSynchronize(m->stream));
        GFX_HIP(hipMemcpy(w, m->w.p, n * 4, hipMemcpyDeviceToHost));
        if (fourier) GFX_HIP(hipMemcpy(fourier, m->Bf.p, size_t(m->cfg.n_freq) * 16, hipMemcpyDeviceToHost));
    });
}

extern "C" gfx_status gfx_mlp_set_params(gfx_mlp* m, const float* w, size_t n) {
    return guarded([&] {
        require(m && w && n == size_t(m->n_params), "parameter buffer must hold n_params floats");
        GFX_HIP(hipStreamSynchronize(m->stream));
        GFX_HIP(hipMemcpy(m->w.p, w, n * 4, hipMemcpyHostToDevice));
    });
}

extern "C" gfx_status gfx_mlp_set_envelope(gfx_mlp* m, double cx, double cy, double slowness, double r_max,
                                           int n_bins, const float* env) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        require(n_bins >= 0 && n_bins <= 65536, "n_bins must be in [0, 65536]");
        require(std::isfinite(slowness) && slowness >= 0, "slowness must be >= 0");
        GFX_HIP(hipStreamSynchronize(m->stream));
        m->env_cx = cx;
        m->env_cy = cy;
        m->slowness = slowness;
        if (n_bins == 0) {
            m->h_env.clear();
            m->env.release();
            return;
        }
        require(env != nullptr && r_max > 0, "envelope needs knots and r_max > 0");
        for (int i = 0; i < n_bins; ++i) require(env[i] > 0.f && std::isfinite(env[i]), "envelope knots must be > 0");
        m->h_env.assign(env, env + n_bins);
        m->env.upload(m->h_env);
        m->env_rmax = r_max;
    });
}

namespace gfx {
namespace sur {
namespace {
// raw (x, y, z, t) rows with t replaced by the network's time input
std::vector<double> network_inputs(const gfx_mlp& m, const double* x, int rows) {
    std::vector<double> v(x, x + size_t(rows) * 4);
    const Envelope E = m.envelope(false);
    if (E.slowness != 0.0)
        for (int r = 0; r < rows; ++r) v[size_t(r) * 4 + 3] = retarded_time(E, v[size_t(r) * 4], v[size_t(r) * 4 + 1], v[size_t(r) * 4 + 3]);
    return v;
}
}  // namespace
}  // namespace sur
}  // namespace gfx

// inputs raw (x, y, z, t) [n][4] FP64; outputs unscaled [n][18]
extern "C" gfx_status gfx_mlp_forward(gfx_mlp* m, int n, const double* x, float* y) {
    return guarded([&] {
        require(m && x && y && n >= 0, "null argument");
        const int chunk = 65536;
        DevBuf<double> raw;
        raw.alloc(size_t(std::min(n, chunk)) * 4 + 4);
        std::vector<float> ho;
        const sur::Envelope E = m->envelope(false);
        for (int b = 0; b < n; b += chunk) {
            const int rows = std::min(chunk, n - b);
            sur::ensure_rows(*m, rows);
            const std::vector<double> xin = sur::network_inputs(*m, x + size_t(b) * 4, rows);  // lives past the sync
            GFX_HIP(hipMemcpyAsync(raw.p, xin.data(), size_t(rows) * 32, hipMemcpyHostToDevice, m->stream));
            sur::k_scale_in<<<(rows + 255) / 256, 256, 0, m->stream>>>(raw.p, rows, m->lo.p, m->hi.p, m->X.p);
            sur::forward(*m, rows);
            ho.resize(size_t(rows) * 18);
            GFX_HIP(hipMemcpyAsync(ho.data(), m->out.p, ho.size() * 4, hipMemcpyDeviceToHost, m->stream));
            GFX_HIP(hipStreamSynchronize(m->stream));
            for (int r = 0; r < rows; ++r) {
                const double* xr = x + (size_t(b) + r) * 4;
                const double e = sur::envelope_at(E, xr[0], xr[1]);
                for (int k = 0; k < 18; ++k)
                    y[(size_t(b) + r) * 18 + k] = float(ho[size_t(r) * 18 + k] * m->h_scale[k] * e);
            }
        }
    });
}

// Gradient of the scaled MSE for given samples (checking against a CPU model):
// x raw [n][4] FP64, y unscaled targets [n][18]; grad [n_params]
extern "C" gfx_status gfx_mlp_gradient(gfx_mlp* m, int n, const double* x, const float* y, float* grad,
                                       double* loss) {
    return guarded([&] {
        require(m && x && y && grad && loss && n >= 1, "null argument");
        sur::ensure_rows(*m, n);
        DevBuf<double> raw;
        raw.alloc(size_t(n) * 4);
        const std::vector<double> xin = sur::network_inputs(*m, x, n);
        GFX_HIP(hipMemcpy(raw.p, xin.data(), size_t(n) * 32, hipMemcpyHostToDevice));
        std::vector<float> ys(size_t(n) * 18);
        const sur::Envelope E = m->envelope(false);
        for (int r = 0; r < n; ++r) {
            const double e = sur::envelope_at(E, x[size_t(r) * 4], x[size_t(r) * 4 + 1]);
            for (int k = 0; k < 18; ++k) ys[size_t(r) * 18 + k] = float(y[size_t(r) * 18 + k] / (m->h_scale[k] * e));
        }
        GFX_HIP(hipMemcpy(m->Y.p, ys.data(), ys.size() * 4, hipMemcpyHostToDevice));
        sur::k_scale_in<<<(n + 255) / 256, 256, 0, m->stream>>>(raw.p, n, m->lo.p, m->hi.p, m->X.p);
        sur::forward(*m, n);
        *loss = sur::backward(*m, n);
        GFX_HIP(hipMemcpy(grad, m->g.p, size_t(m->n_params) * 4, hipMemcpyDeviceToHost));
    });
}

// Adam training on random (node, time) samples of a tile; the tile's time
// index k maps to t = t0 + k * dt.  *loss receives the mean scaled loss of the
// last 10 % of the iterations.
extern "C" gfx_status gfx_mlp_train(gfx_mlp* m, const gfx_green* g, double t0, double dt, int iters, uint64_t seed,
                                    double* loss) {
    return guarded([&] {
        require(m && g && loss && iters >= 0, "null argument");
        const int rows = m->cfg.batch;
        require(rows >= 64, "batch must be >= 64");
        sur::ensure_rows(*m, rows);
        GFX_HIP(hipStreamSynchronize(g->stream));
        double tail = 0.0;
        int nt = 0;
        const sur::Envelope E = m->envelope(true);
        const double lr0 = m->cfg.lr, lr1 = m->cfg.lr_final;
        for (int it = 0; it < iters; ++it) {
            sur::k_sample_tile<<<(rows + 255) / 256, 256, 0, m->stream>>>(
                g->val.p, g->xyz.p, g->n_node, g->n_time, t0, dt, m->lo.p, m->hi.p, m->inv_scale.p, E,
                seed * 0x9E3779B97F4A7C15ULL + uint64_t(it), rows, m->X.p, m->Y.p);
            sur::forward(*m, rows);
            const double l = sur::backward(*m, rows);
            // cosine decay from lr to lr_final over this call
            const double ph = iters > 1 ? double(it) / double(iters - 1) : 1.0;
            sur::adam(*m, float(lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(3.14159265358979323846 * ph))));
            if (it >= iters - std::max(1, iters / 10)) {
                tail += l;
                ++nt;
            }
        }
        GFX_HIP(hipStreamSynchronize(m->stream));
        *loss = nt ? tail / nt : 0.0;
    });
}

// Surrogate query: the MLP everywhere, the judge for points in (cell, time
// window) regions marked for fallback (or outside the tile).  *n_fallback
// counts judge answers.
extern "C" gfx_status gfx_surrogate_query(gfx_mlp* m, const gfx_green* g, const uint8_t* fallback, int n_window,
                                          double t0, double dt, int n, const double* xyz, const int32_t* tix,
                                          float* out, int* n_fallback) {
    return guarded([&] {
        require(m && g && fallback && xyz && tix && out && n_fallback && n >= 0, "null argument");
        require(n_window >= 1 && n_window <= g->n_time, "n_window must be in [1, n_time]");
        if (n == 0) {
            *n_fallback = 0;
            return;
        }
        std::vector<double> x4(size_t(n) * 4);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) x4[size_t(i) * 4 + d] = xyz[size_t(i) * 3 + d];
            x4[size_t(i) * 4 + 3] = t0 + tix[i] * dt;
        }
        std::vector<float> mo(size_t(n) * 18);
        {
            const gfx_status st = gfx_mlp_forward(m, n, x4.data(), mo.data());
            if (st != GFX_OK) throw Error(st, gfx_last_error());
        }
        std::vector<int32_t> cell(n);
        std::vector<float> jo(size_t(n) * 18);
        DevBuf<double> q;
        DevBuf<int32_t> t, c;
        DevBuf<float> o;
        DevBuf<uint8_t> fb;
        q.alloc(size_t(n) * 3);
        t.alloc(n);
        c.alloc(n);
        o.alloc(size_t(n) * 18);
        std::vector<uint8_t> judge_only(size_t(g->n_cell) * n_window);
        for (size_t r = 0; r < judge_only.size(); ++r) judge_only[r] = fallback[r] == 1 ? 1 : 0;
        fb.upload(judge_only);
        GFX_HIP(hipMemcpy(q.p, xyz, size_t(n) * 24, hipMemcpyHostToDevice));
        GFX_HIP(hipMemcpy(t.p, tix, size_t(n) * 4, hipMemcpyHostToDevice));
        GFX_HIP(hipMemsetAsync(o.p, 0, size_t(n) * 72, g->stream));
        sur::k_judge<<<(n + 127) / 128, 128, 0, g->stream>>>(sur::judge_args(*g), n, q.p, t.p, fb.p, n_window, o.p,
                                                              c.p);
        GFX_HIP(hipGetLastError());
        GFX_HIP(hipStreamSynchronize(g->stream));
        GFX_HIP(hipMemcpy(jo.data(), o.p, jo.size() * 4, hipMemcpyDeviceToHost));
        GFX_HIP(hipMemcpy(cell.data(), c.p, size_t(n) * 4, hipMemcpyDeviceToHost));
        int nf = 0;
        const float zero[18] = {};
        for (int i = 0; i < n; ++i) {
            const uint8_t how = cell[i] < 0 ? 1 : fallback[sur::region_of(cell[i], tix[i], g->n_time, n_window)];
            const bool use_judge = how == 1;
            const float* src = use_judge ? &jo[size_t(i) * 18] : how == 2 ? zero : &mo[size_t(i) * 18];
            if (use_judge) {
                ++nf;
                if (cell[i] < 0)
                    for (int k = 0; k < 18; ++k) jo[size_t(i) * 18 + k] = __builtin_nanf("");
            }
            std::memcpy(out + size_t(i) * 18, src, 72);
        }
        *n_fallback = nf;
    });
}
