// common.cpp -- error state and GLL quadrature setup (host, FP64).
#include "common.hpp"

namespace gfx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

namespace {
double legendre(int n, double x) {
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double p0 = 1.0, p1 = x, pn = 0.0;
    for (int k = 1; k < n; ++k) {
        pn = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = pn;
    }
    return pn;
}
double legendre_prime(int n, double x) {
    if (n == 0) return 0.0;
    if (std::abs(x - 1.0) < 1e-14) return 0.5 * n * (n + 1.0);
    if (std::abs(x + 1.0) < 1e-14) return ((n % 2 == 0) ? -1.0 : 1.0) * 0.5 * n * (n + 1.0);
    return n * (x * legendre(n, x) - legendre(n - 1, x)) / (x * x - 1.0);
}
}  // namespace

GLL::GLL(int order) : N(order), n(order + 1) {
    require(order >= 1 && order <= 15, "GLL order must be in [1, 15]");
    x[0] = -1.0;
    x[N] = 1.0;
    for (int i = 1; i < N; ++i) {
        double xi = -std::cos(M_PI * i / N);
        for (int it = 0; it < 100; ++it) {
            const double f = legendre_prime(N, xi);
            const double df = (2.0 * xi * f - N * (N + 1.0) * legendre(N, xi)) / (1.0 - xi * xi);
            const double dx = -f / df;
            xi += dx;
            if (std::abs(dx) < 1.0e-15 * (1.0 + std::abs(xi))) break;
        }
        x[i] = xi;
    }
    double PN[16];
    for (int i = 0; i < n; ++i) {
        PN[i] = legendre(N, x[i]);
        w[i] = 2.0 / (N * (N + 1.0) * PN[i] * PN[i]);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double d;
            if (i == j)
                d = (i == 0) ? -0.25 * N * (N + 1.0) : (i == N ? 0.25 * N * (N + 1.0) : 0.0);
            else
                d = PN[i] / (PN[j] * (x[i] - x[j]));
            D[i * n + j] = d;
        }
}

void GLL::lagrange(double xi, double* ell) const {
    for (int j = 0; j < n; ++j) {
        double v = 1.0;
        for (int k = 0; k < n; ++k)
            if (k != j) v *= (xi - x[k]) / (x[j] - x[k]);
        ell[j] = v;
    }
}

}  // namespace gfx

extern "C" const char* gfx_last_error(void) { return gfx::g_last_error.c_str(); }
extern "C" const char* gfx_version(void) { return "gfx-mi355x 0.1.0"; }
