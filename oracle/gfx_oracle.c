/*
 * gfx_oracle.c -- CPU FP64 restatement of the reference's elastic CG-SEM
 * forward step.  TEST INFRASTRUCTURE ONLY (parity checker); see gfx_oracle.h
 * for which checks pin each function.
 *
 * Every function cites the reference file:line whose behaviour it restates,
 * including the quirks the parity target depends on:
 *   - the source at loop index `step` uses f(step*dt), i.e. f(t_n)
 *     (forward/share/src/solver.cpp:579-582);
 *   - the C-PML acceleration term is ADDED to the residual using u_n, v_n
 *     (forward/share/src/pml.cpp:197-212);
 *   - pml_displ_old mixes step indices: previous "new" plus c2*a_n
 *     (forward/share/src/pml.cpp:66-109);
 *   - a GLL point with mu <= 0 contributes nothing
 *     (forward/elastic/src/element_cpu.cpp:42-43);
 *   - a node with mass <= 0 gets a = 0 and keeps u, v
 *     (forward/share/src/solver.cpp:57-61).
 */
#include "gfx_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* GLL quadrature: forward/share/include/gf/gll.hpp                    */
/* ------------------------------------------------------------------ */

/* Bonnet recurrence for P_n(x) (gll.hpp:14-26). */
static double legendre(int n, double x) {
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

/* P'_n(x) from n (x P_n - P_{n-1}) / (x^2 - 1), endpoint limits (gll.hpp:29-41). */
static double legendre_prime(int n, double x) {
    if (n == 0) return 0.0;
    if (fabs(x - 1.0) < 1e-14) return 0.5 * n * (n + 1.0);
    if (fabs(x + 1.0) < 1e-14) return ((n % 2 == 0) ? -1.0 : 1.0) * 0.5 * n * (n + 1.0);
    return n * (x * legendre(n, x) - legendre(n - 1, x)) / (x * x - 1.0);
}

/* Interior nodes: Newton on P'_N from -cos(pi i/N) (gll.hpp:48-81). */
int gfo_gll_nodes(int N, double* nodes) {
    if (N < 1) return -1;
    nodes[0] = -1.0;
    nodes[N] = 1.0;
    for (int i = 1; i < N; ++i) {
        double xi = -cos(M_PI * i / N);
        for (int it = 0; it < 100; ++it) {
            double f = legendre_prime(N, xi);
            double df = (2.0 * xi * f - N * (N + 1.0) * legendre(N, xi)) / (1.0 - xi * xi);
            if (fabs(df) < 1.0e-14) return -2;
            double dx = -f / df;
            xi += dx;
            if (fabs(dx) < 1.0e-15 * (1.0 + fabs(xi))) break;
        }
        nodes[i] = xi;
    }
    return 0;
}

/* w_i = 2 / (N (N+1) P_N(x_i)^2) (gll.hpp:87-97). */
int gfo_gll_weights(int N, const double* nodes, double* w) {
    if (N < 1) return -1;
    for (int i = 0; i <= N; ++i) {
        double pn = legendre(N, nodes[i]);
        w[i] = 2.0 / (N * (N + 1.0) * pn * pn);
    }
    return 0;
}

/* D[i][j] = l_j'(x_i), row-major (gll.hpp:103-131). */
int gfo_gll_derivative_matrix(int N, const double* nodes, double* D) {
    if (N < 1) return -1;
    const int n = N + 1;
    double PN[64];
    if (n > 64) return -1;
    for (int i = 0; i < n; ++i) PN[i] = legendre(N, nodes[i]);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (i == j) {
                if (i == 0)
                    D[i * n + j] = -0.25 * N * (N + 1.0);
                else if (i == N)
                    D[i * n + j] = 0.25 * N * (N + 1.0);
                else
                    D[i * n + j] = 0.0;
            } else {
                D[i * n + j] = PN[i] / (PN[j] * (nodes[i] - nodes[j]));
            }
        }
    }
    return 0;
}

/* Lagrange basis values at xi (gll.hpp:136-147). */
void gfo_lagrange_basis(double xi, int ngll, const double* nodes, double* ell) {
    for (int j = 0; j < ngll; ++j) {
        ell[j] = 1.0;
        for (int k = 0; k < ngll; ++k) {
            if (k == j) continue;
            ell[j] *= (xi - nodes[k]) / (nodes[j] - nodes[k]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Element internal force: forward/elastic/src/element_cpu.cpp:19-162  */
/* ------------------------------------------------------------------ */

static inline int nidx(int i, int j, int k, int n) { return (i * n + j) * n + k; }

void gfo_element_residual(int n_elem, const double* dxi_dx, const double* jacobian,
                          const double* lambda_, const double* mu_, const double* D,
                          const double* weights, int ngll, const double* u, double* r) {
    const int n_node = ngll * ngll * ngll;
    /* elements are independent (each writes its own block of r) */
#pragma omp parallel for schedule(static)
    for (long e = 0; e < n_elem; ++e) {
        const double* edd = dxi_dx + e * n_node * 9;
        const double* ejac = jacobian + e * n_node;
        const double* elam = lambda_ + e * n_node;
        const double* emu = mu_ + e * n_node;
        const double* eu = u + e * n_node * 3;
        double* er = r + e * n_node * 3;
        for (int i = 0; i < ngll; ++i)
            for (int j = 0; j < ngll; ++j)
                for (int k = 0; k < ngll; ++k) {
                    const int n = nidx(i, j, k, ngll);
                    const double lam = elam[n], mu = emu[n];
                    if (mu <= 0.0) continue; /* element_cpu.cpp:42-43 */
                    const double* dd = edd + 9 * n;
                    double gx[3] = {0, 0, 0}, ge[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
                    for (int s = 0; s < ngll; ++s) {
                        const double di = D[i * ngll + s], dj = D[j * ngll + s],
                                     dk = D[k * ngll + s];
                        const int a = nidx(s, j, k, ngll), b = nidx(i, s, k, ngll),
                                  c = nidx(i, j, s, ngll);
                        for (int d = 0; d < 3; ++d) {
                            gx[d] += di * eu[3 * a + d];
                            ge[d] += dj * eu[3 * b + d];
                            gz[d] += dk * eu[3 * c + d];
                        }
                    }
                    double g[3][3];
                    for (int c = 0; c < 3; ++c) {
                        g[c][0] = gx[c] * dd[0] + ge[c] * dd[1] + gz[c] * dd[2];
                        g[c][1] = gx[c] * dd[3] + ge[c] * dd[4] + gz[c] * dd[5];
                        g[c][2] = gx[c] * dd[6] + ge[c] * dd[7] + gz[c] * dd[8];
                    }
                    double eps[3][3], sig[3][3];
                    for (int l = 0; l < 3; ++l)
                        for (int m = 0; m < 3; ++m) eps[l][m] = 0.5 * (g[l][m] + g[m][l]);
                    const double tr = eps[0][0] + eps[1][1] + eps[2][2];
                    for (int l = 0; l < 3; ++l) {
                        for (int m = 0; m < 3; ++m) sig[l][m] = 2.0 * mu * eps[l][m];
                        sig[l][l] += lam * tr;
                    }
                    const double fac = ejac[n] * weights[i] * weights[j] * weights[k];
                    /* xi-, eta-, zeta-direction test-function derivatives */
                    for (int dir = 0; dir < 3; ++dir) {
                        for (int s = 0; s < ngll; ++s) {
                            double Dq;
                            int ns;
                            if (dir == 0) {
                                Dq = D[i * ngll + s];
                                ns = nidx(s, j, k, ngll);
                            } else if (dir == 1) {
                                Dq = D[j * ngll + s];
                                ns = nidx(i, s, k, ngll);
                            } else {
                                Dq = D[k * ngll + s];
                                ns = nidx(i, j, s, ngll);
                            }
                            const double gn[3] = {Dq * dd[dir], Dq * dd[3 + dir], Dq * dd[6 + dir]};
                            for (int c = 0; c < 3; ++c)
                                er[3 * ns + c] -=
                                    (sig[c][0] * gn[0] + sig[c][1] * gn[1] + sig[c][2] * gn[2]) *
                                    fac;
                        }
                    }
                }
    }
}

/* ------------------------------------------------------------------ */
/* Shear attenuation (standard linear solids); no reference counterpart */
/* ------------------------------------------------------------------ */

void gfo_sls_coefficients(double tau_sigma, double tau_epsilon, double mu_relaxed, double dt, double* abcY) {
    const double a = exp(-dt / tau_sigma);
    const double q = tau_sigma * (1.0 - a) / dt;
    abcY[0] = a;
    abcY[1] = q - a;
    abcY[2] = 1.0 - q;
    abcY[3] = 2.0 * mu_relaxed * (tau_epsilon / tau_sigma - 1.0);
}

/* z^{n+1} = a z^n + Y (b dev^n + c dev^{n+1}) for the 5 stored components of
 * every mechanism; returns sum_l z_l in zs[5]. */
static void sls_update(int n_sls, const double* tsg, const double* teps, double mu_r, double dt,
                       const double* dev_old, const double* dev_new, double* z, double* zs) {
    for (int q = 0; q < 5; ++q) zs[q] = 0.0;
    for (int l = 0; l < n_sls; ++l) {
        double c[4];
        gfo_sls_coefficients(tsg[l], teps[l], mu_r, dt, c);
        for (int q = 0; q < 5; ++q) {
            z[l * 5 + q] = c[0] * z[l * 5 + q] + c[3] * (c[1] * dev_old[q] + c[2] * dev_new[q]);
            zs[q] += z[l * 5 + q];
        }
    }
}

static double sls_unrelaxed(int n_sls, const double* tsg, const double* teps, double mu_r) {
    double s = 1.0;
    for (int l = 0; l < n_sls; ++l) s += teps[l] / tsg[l] - 1.0;
    return mu_r * s;
}

void gfo_sls_point_history(int n_sls, const double* tau_sigma, const double* tau_epsilon, double mu_relaxed,
                           double dt, int n, const double* dev, double* sig) {
    double z[8 * 5] = {0}, zs[5];
    const double mu_u = sls_unrelaxed(n_sls, tau_sigma, tau_epsilon, mu_relaxed);
    for (int t = 1; t < n; ++t) {
        sls_update(n_sls, tau_sigma, tau_epsilon, mu_relaxed, dt, dev + (t - 1) * 5, dev + t * 5, z, zs);
        for (int q = 0; q < 5; ++q) sig[(t - 1) * 5 + q] = 2.0 * mu_u * dev[t * 5 + q] - zs[q];
    }
}

void gfo_element_residual_att(int n_elem, const double* dxi_dx, const double* jacobian, const double* lambda_,
                              const double* mu_, int n_sls, const double* tau_sigma, const double* tau_epsilon,
                              double dt, const double* D, const double* weights, int ngll, const double* u,
                              double* zeta, double* dev_prev, double* r) {
    const int n_node = ngll * ngll * ngll;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < n_elem; ++e) {
        const double* edd = dxi_dx + e * n_node * 9;
        const double* ejac = jacobian + e * n_node;
        const double* eu = u + e * n_node * 3;
        double* er = r + e * n_node * 3;
        for (int i = 0; i < ngll; ++i)
            for (int j = 0; j < ngll; ++j)
                for (int k = 0; k < ngll; ++k) {
                    const int n = nidx(i, j, k, ngll);
                    const long pt = e * n_node + n;
                    const double lam_r = lambda_[pt], mu_r = mu_[pt];
                    if (mu_r <= 0.0) continue; /* element_cpu.cpp:42-43 */
                    const double* tsg = tau_sigma + pt * n_sls;
                    const double* teps = tau_epsilon + pt * n_sls;
                    const double* dd = edd + 9 * n;
                    double gx[3] = {0, 0, 0}, ge[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
                    for (int s = 0; s < ngll; ++s) {
                        const double di = D[i * ngll + s], dj = D[j * ngll + s], dk = D[k * ngll + s];
                        const int a = nidx(s, j, k, ngll), b = nidx(i, s, k, ngll), c = nidx(i, j, s, ngll);
                        for (int d = 0; d < 3; ++d) {
                            gx[d] += di * eu[3 * a + d];
                            ge[d] += dj * eu[3 * b + d];
                            gz[d] += dk * eu[3 * c + d];
                        }
                    }
                    double g[3][3];
                    for (int c = 0; c < 3; ++c) {
                        g[c][0] = gx[c] * dd[0] + ge[c] * dd[1] + gz[c] * dd[2];
                        g[c][1] = gx[c] * dd[3] + ge[c] * dd[4] + gz[c] * dd[5];
                        g[c][2] = gx[c] * dd[6] + ge[c] * dd[7] + gz[c] * dd[8];
                    }
                    double eps[3][3];
                    for (int l = 0; l < 3; ++l)
                        for (int m = 0; m < 3; ++m) eps[l][m] = 0.5 * (g[l][m] + g[m][l]);
                    const double tr = eps[0][0] + eps[1][1] + eps[2][2];
                    const double dev[5] = {eps[0][0] - tr / 3.0, eps[1][1] - tr / 3.0, eps[0][1], eps[0][2],
                                           eps[1][2]};
                    double zs[5];
                    sls_update(n_sls, tsg, teps, mu_r, dt, dev_prev + pt * 5, dev, zeta + pt * 5 * n_sls, zs);
                    for (int q = 0; q < 5; ++q) dev_prev[pt * 5 + q] = dev[q];
                    const double mu_u = sls_unrelaxed(n_sls, tsg, teps, mu_r);
                    const double lam_u = lam_r + 2.0 * (mu_r - mu_u) / 3.0;
                    double sig[3][3];
                    for (int l = 0; l < 3; ++l) {
                        for (int m = 0; m < 3; ++m) sig[l][m] = 2.0 * mu_u * eps[l][m];
                        sig[l][l] += lam_u * tr;
                    }
                    sig[0][0] -= zs[0];
                    sig[1][1] -= zs[1];
                    sig[2][2] += zs[0] + zs[1];
                    sig[0][1] -= zs[2];
                    sig[1][0] -= zs[2];
                    sig[0][2] -= zs[3];
                    sig[2][0] -= zs[3];
                    sig[1][2] -= zs[4];
                    sig[2][1] -= zs[4];
                    const double fac = ejac[n] * weights[i] * weights[j] * weights[k];
                    for (int dir = 0; dir < 3; ++dir) {
                        for (int s = 0; s < ngll; ++s) {
                            double Dq;
                            int ns;
                            if (dir == 0) {
                                Dq = D[i * ngll + s];
                                ns = nidx(s, j, k, ngll);
                            } else if (dir == 1) {
                                Dq = D[j * ngll + s];
                                ns = nidx(i, s, k, ngll);
                            } else {
                                Dq = D[k * ngll + s];
                                ns = nidx(i, j, s, ngll);
                            }
                            const double gn[3] = {Dq * dd[dir], Dq * dd[3 + dir], Dq * dd[6 + dir]};
                            for (int c = 0; c < 3; ++c)
                                er[3 * ns + c] -=
                                    (sig[c][0] * gn[0] + sig[c][1] * gn[1] + sig[c][2] * gn[2]) * fac;
                        }
                    }
                }
    }
}

/* ------------------------------------------------------------------ */
/* Complete C-PML (stretched weak form); no reference counterpart: the */
/* reference defers the strain part (docs/deferred.md:46-80).          */
/* ------------------------------------------------------------------ */

/* conv(g, p) over one step for g linear in time between the two levels. */
void gfo_cpml_recursion(double p, double dt, double* abc) {
    const double x = p * dt;
    double i0, c;
    if (fabs(x) < 1e-2) { /* I0/dt and I1/dt^2 as power series in x */
        i0 = 1.0 - x / 2.0 + x * x / 6.0 - x * x * x / 24.0 + x * x * x * x / 120.0 - pow(x, 5) / 720.0 +
             pow(x, 6) / 5040.0;
        c = 0.5 - x / 6.0 + x * x / 24.0 - x * x * x / 120.0 + x * x * x * x / 720.0 - pow(x, 5) / 5040.0 +
            pow(x, 6) / 40320.0;
    } else {
        const double e = exp(-x);
        i0 = (1.0 - e) / x;
        c = (1.0 - i0) / x;
    }
    abc[0] = exp(-x);
    abc[1] = dt * (i0 - c);
    abc[2] = dt * c;
}

void gfo_cpml_point_coefficients(const double* K, const double* d, const double* alpha, double dt, double* coef) {
    for (int q = 0; q < 3; ++q) {
        double* cq = coef + q * 9;
        cq[0] = K[q];
        cq[1] = d[q];
        cq[2] = alpha[q];
        for (int t = 3; t < 9; ++t) cq[t] = 0.0;
        if (d[q] != 0.0) { /* an axis without damping keeps its memories at zero */
            gfo_cpml_recursion(alpha[q], dt, cq + 3);
            gfo_cpml_recursion(alpha[q] + d[q] / K[q], dt, cq + 6);
        }
    }
}

/* M^{n+1} = Q + c g^{n+1}; the carry becomes a M^{n+1} + b g^{n+1}. */
static inline double conv_step(const double* abc, double* q, double g) {
    const double m = *q + abc[2] * g;
    *q = abc[0] * m + abc[1] * g;
    return m;
}

void gfo_cpml_gradient(const double* coef, const double* g9, double* carry, double* G27) {
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
            const int m = l == 0 ? 1 : 0, n = l == 2 ? 1 : 2; /* the other two axes, m < n */
            const double* cm = coef + m * 9;
            const double* cn = coef + n * 9;
            const double* cl = coef + l * 9;
            const double g = g9[k * 3 + l];
            double* Q = carry + (k * 3 + l) * 4;
            const double Mm = conv_step(cm + 3, Q + 0, g);  /* conv(g, alpha_m) */
            const double Mn = conv_step(cn + 3, Q + 1, g);  /* conv(g, alpha_n) */
            const double Mmn = conv_step(cn + 3, Q + 2, Mm); /* conv(conv(g, alpha_m), alpha_n) */
            const double y2 = cm[0] * cn[0] * g + cn[0] * cm[1] * Mm + cn[1] * (cm[0] * Mn + cm[1] * Mmn);
            const double Mb = conv_step(cl + 6, Q + 3, y2); /* conv(s_m s_n g, beta_l) */
            G27[l * 9 + k * 3 + l] = (y2 - cl[1] / cl[0] * Mb) / cl[0]; /* s_m s_n / s_l */
            G27[n * 9 + k * 3 + l] = cm[0] * g + cm[1] * Mm;              /* column n: s_m */
            G27[m * 9 + k * 3 + l] = cn[0] * g + cn[1] * Mn;              /* column m: s_n */
        }
}

void gfo_cpml_inertia(const double* coef, const double* u3, const double* v3, double* carry, double* L3) {
    const double *cx = coef, *cy = coef + 9, *cz = coef + 18;
    for (int c = 0; c < 3; ++c) {
        const double u = u3[c], v = v3[c];
        double* Q = carry + c * 3;
        const double C1 = conv_step(cx + 3, Q + 0, u);
        const double z1 = cx[0] * u + cx[1] * C1;
        const double C2 = conv_step(cy + 3, Q + 1, z1);
        const double z2 = cy[0] * z1 + cy[1] * C2;
        const double C3 = conv_step(cz + 3, Q + 2, z2);
        const double L1 = cx[1] * (v - cx[2] * u + cx[2] * cx[2] * C1);
        const double Dz1 = cx[0] * v + cx[1] * (u - cx[2] * C1);
        const double L2 = cy[0] * L1 + cy[1] * (Dz1 - cy[2] * z1 + cy[2] * cy[2] * C2);
        const double Dz2 = cy[0] * Dz1 + cy[1] * (z1 - cy[2] * C2);
        L3[c] = cz[0] * L2 + cz[1] * (Dz2 - cz[2] * z2 + cz[2] * cz[2] * C3);
    }
}

void gfo_element_residual_cpml(int n_elem, const double* dxi_dx, const double* jacobian, const double* lambda_,
                               const double* mu_, const int32_t* pml_region, const double* coef,
                               const double* D, const double* weights, int ngll, const double* u, double* carry,
                               double* r) {
    const int n_node = ngll * ngll * ngll;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < n_elem; ++e) {
        const int pml = pml_region && pml_region[e] != 0;
        const double* edd = dxi_dx + e * n_node * 9;
        const double* eu = u + e * n_node * 3;
        double* er = r + e * n_node * 3;
        for (int i = 0; i < ngll; ++i)
            for (int j = 0; j < ngll; ++j)
                for (int k = 0; k < ngll; ++k) {
                    const int n = nidx(i, j, k, ngll);
                    const long pt = e * n_node + n;
                    const double lam = lambda_[pt], mu = mu_[pt];
                    if (mu <= 0.0) continue; /* element_cpu.cpp:42-43 */
                    const double* dd = edd + 9 * n;
                    double gx[3] = {0, 0, 0}, ge[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
                    for (int s = 0; s < ngll; ++s) {
                        const double di = D[i * ngll + s], dj = D[j * ngll + s], dk = D[k * ngll + s];
                        const int a = nidx(s, j, k, ngll), b = nidx(i, s, k, ngll), c = nidx(i, j, s, ngll);
                        for (int q = 0; q < 3; ++q) {
                            gx[q] += di * eu[3 * a + q];
                            ge[q] += dj * eu[3 * b + q];
                            gz[q] += dk * eu[3 * c + q];
                        }
                    }
                    double g[9]; /* g[c*3+m] = d u_c / d x_m */
                    for (int c = 0; c < 3; ++c) {
                        g[c * 3 + 0] = gx[c] * dd[0] + ge[c] * dd[1] + gz[c] * dd[2];
                        g[c * 3 + 1] = gx[c] * dd[3] + ge[c] * dd[4] + gz[c] * dd[5];
                        g[c * 3 + 2] = gx[c] * dd[6] + ge[c] * dd[7] + gz[c] * dd[8];
                    }
                    double G[27]; /* G[j*9 + k*3 + l]: the gradient seen by stress column j */
                    if (pml) {
                        gfo_cpml_gradient(coef + pt * 27, g, carry + pt * 45, G);
                    } else {
                        for (int c = 0; c < 3; ++c)
                            for (int q = 0; q < 9; ++q) G[c * 9 + q] = g[q];
                    }
                    double S[3][3];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            const double* Gb = G + b * 9;
                            S[a][b] = mu * (Gb[a * 3 + b] + Gb[b * 3 + a]);
                            if (a == b) S[a][b] += lam * (Gb[0] + Gb[4] + Gb[8]);
                        }
                    const double fac = jacobian[pt] * weights[i] * weights[j] * weights[k];
                    for (int dir = 0; dir < 3; ++dir) {
                        for (int s = 0; s < ngll; ++s) {
                            double Dq;
                            int ns;
                            if (dir == 0) {
                                Dq = D[i * ngll + s];
                                ns = nidx(s, j, k, ngll);
                            } else if (dir == 1) {
                                Dq = D[j * ngll + s];
                                ns = nidx(i, s, k, ngll);
                            } else {
                                Dq = D[k * ngll + s];
                                ns = nidx(i, j, s, ngll);
                            }
                            const double gn[3] = {Dq * dd[dir], Dq * dd[3 + dir], Dq * dd[6 + dir]};
                            for (int c = 0; c < 3; ++c)
                                er[3 * ns + c] -= (S[c][0] * gn[0] + S[c][1] * gn[1] + S[c][2] * gn[2]) * fac;
                        }
                    }
                }
    }
}

void gfo_cpml_filter_history(const double* coef, int n, const double* g, const double* u, const double* v,
                             double* G27, double* L3) {
    double cs[45] = {0};
    for (int t = 0; t < n; ++t) {
        if (g) gfo_cpml_gradient(coef, g + t * 9, cs, G27 + t * 27);
        if (u) gfo_cpml_inertia(coef, u + t * 3, v + t * 3, cs + 36, L3 + t * 3);
    }
}

/* ------------------------------------------------------------------ */
/* Assembly: forward/share/src/assembly.cpp:10-41                      */
/* ------------------------------------------------------------------ */

void gfo_scatter_to_rank(const double* loc, const int32_t* ibool, int n_elem, int n_node,
                         int n_rank_node, double* glob) {
    memset(glob, 0, sizeof(double) * 3 * (size_t)n_rank_node); /* zero-then-add, :15 */
    for (long e = 0; e < n_elem; ++e)
        for (int n = 0; n < n_node; ++n) {
            const long g = ibool[e * n_node + n];
            for (int d = 0; d < 3; ++d) glob[g * 3 + d] += loc[(e * n_node + n) * 3 + d];
        }
}

void gfo_gather_from_rank(const double* glob, const int32_t* ibool, int n_elem, int n_node,
                          double* loc) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < n_elem; ++e)
        for (int n = 0; n < n_node; ++n) {
            const long g = ibool[e * n_node + n];
            for (int d = 0; d < 3; ++d) loc[(e * n_node + n) * 3 + d] = glob[g * 3 + d];
        }
}

/* ------------------------------------------------------------------ */
/* Newmark: forward/share/src/solver.cpp:40-69                         */
/* ------------------------------------------------------------------ */

void gfo_newmark_predict(int n_dof, double dt, double beta, const double* u, const double* v,
                         const double* a, double* ut) {
    const double c = 0.5 * dt * dt * (1.0 - 2.0 * beta);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n_dof; ++i) ut[i] = u[i] + dt * v[i] + c * a[i];
}

void gfo_newmark_correct(int n_dof, double dt, double beta, double gamma, const double* mass,
                         double* u, double* v, double* a, const double* r) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n_dof; ++i) {
        const double m = mass[i / 3];
        if (m <= 0.0) {
            a[i] = 0.0;
            continue;
        }
        const double a_old = a[i];
        const double a_new = r[i] / m;
        u[i] += dt * v[i] + dt * dt * ((0.5 - beta) * a_old + beta * a_new);
        v[i] += dt * ((1.0 - gamma) * a_old + gamma * a_new);
        a[i] = a_new;
    }
}

/* ------------------------------------------------------------------ */
/* C-PML: forward/share/src/pml.cpp                                    */
/* ------------------------------------------------------------------ */

/* pml.cpp:13-29 (apply_pml_damping): v[i] -= d[i/3] v[i] where d > 0; the
 * solver's own loop over rank nodes (solver.cpp:565-573) is the same rule. */
void gfo_apply_pml_damping(const double* damping_node, double* v, long n_dof) {
    for (long i = 0; i < n_dof; ++i) {
        const double d = damping_node[i / 3];
        if (d > 0.0) v[i] -= d * v[i];
    }
}

#define GFO_THETA (1.0 / 8.0) /* pml.cpp:35 */

static inline int region_of(const gfo_cpml* p, int e) {
    return p->pml_region ? p->pml_region[e] : 0;
}

/* pml.cpp:55-110: swap(old,new); new = u + c1 v; old += c2 a. */
void gfo_cpml_update_displ_fields(gfo_cpml* p, const double* u, const double* v, const double* a,
                                  double dt) {
    const int n_node = p->ngll * p->ngll * p->ngll;
    const double c1 = (1.0 - 2.0 * GFO_THETA) * 0.5 * dt;
    const double c2 = (1.0 - GFO_THETA) * 0.5 * dt * dt;
    double* t = p->displ_old;
    p->displ_old = p->displ_new;
    p->displ_new = t;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < p->n_elem; ++e) {
        if (region_of(p, (int)e) == 0) continue;
        for (int n = 0; n < n_node; ++n) {
            const long g = p->ibool[e * n_node + n];
            for (int d = 0; d < 3; ++d)
                p->displ_new[(e * n_node + n) * 3 + d] = u[g * 3 + d] + c1 * v[g * 3 + d];
        }
    }
#pragma omp parallel for schedule(static)
    for (long e = 0; e < p->n_elem; ++e) {
        if (region_of(p, (int)e) == 0) continue;
        for (int n = 0; n < n_node; ++n) {
            const long g = p->ibool[e * n_node + n];
            for (int d = 0; d < 3; ++d) p->displ_old[(e * n_node + n) * 3 + d] += c2 * a[g * 3 + d];
        }
    }
}

/* pml.cpp:112-149: m[c*3+d] = a0_d m + a1_d new_c + a2_d old_c. */
void gfo_cpml_update_displ_memory(gfo_cpml* p) {
    const int n_node = p->ngll * p->ngll * p->ngll;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < p->n_elem; ++e) {
        if (region_of(p, (int)e) == 0) continue;
        for (int n = 0; n < n_node; ++n) {
            const long pt = e * n_node + n;
            for (int c = 0; c < 3; ++c) {
                const double nv = p->displ_new[pt * 3 + c];
                const double ov = p->displ_old[pt * 3 + c];
                for (int d = 0; d < 3; ++d) {
                    const double* al = p->coef_alpha + pt * 9 + d * 3;
                    double* m = p->rmemory + pt * 9 + c * 3 + d;
                    *m = al[0] * (*m) + al[1] * nv + al[2] * ov;
                }
            }
        }
    }
}

/* pml.cpp:151-216: r_loc += w J/rho (A1 v + A2 u + A3 mx + A4 my + A5 mz). */
void gfo_cpml_accel_contribution(const gfo_cpml* p, const double* u, const double* v,
                                 const double* gll_w, double* r_loc) {
    const int ngll = p->ngll;
    const int n_node = ngll * ngll * ngll;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < p->n_elem; ++e) {
        if (region_of(p, (int)e) == 0) continue;
        for (int n = 0; n < n_node; ++n) {
            const int nk = n % ngll, nj = (n / ngll) % ngll, ni = n / (ngll * ngll);
            const double wg = gll_w[ni] * gll_w[nj] * gll_w[nk];
            const long pt = e * n_node + n;
            const long g = p->ibool[pt];
            const double rho = p->density[pt];
            const double rho_inv = (rho > 0.0) ? 1.0 / rho : 0.0;
            const double* A = p->coef_abar + pt * 5;
            const double scale = wg * rho_inv * p->jacobian[pt];
            for (int c = 0; c < 3; ++c) {
                const double* m = p->rmemory + pt * 9 + c * 3;
                const double acc = scale * (A[0] * v[g * 3 + c] + A[1] * u[g * 3 + c] +
                                            A[2] * m[0] + A[3] * m[1] + A[4] * m[2]);
                r_loc[pt * 3 + c] += acc;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Strain: forward/share/src/solver.cpp:282-335 (recording cells only) */
/* ------------------------------------------------------------------ */

void gfo_full_strain(int n_cell, const int32_t* cells, const double* dxi_dx, const double* D,
                     int ngll, const double* u_loc_all, double* strain) {
    const int n_node = ngll * ngll * ngll;
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < n_cell; ++ci) {
        const long e = cells ? cells[ci] : ci;
        const double* eu = u_loc_all + e * n_node * 3;
        for (int i = 0; i < ngll; ++i)
            for (int j = 0; j < ngll; ++j)
                for (int k = 0; k < ngll; ++k) {
                    const int n = nidx(i, j, k, ngll);
                    const double* dd = dxi_dx + (e * n_node + n) * 9;
                    double gx[3] = {0, 0, 0}, ge[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
                    for (int s = 0; s < ngll; ++s) {
                        const double di = D[i * ngll + s], dj = D[j * ngll + s],
                                     dk = D[k * ngll + s];
                        const int a = nidx(s, j, k, ngll), b = nidx(i, s, k, ngll),
                                  c = nidx(i, j, s, ngll);
                        for (int d = 0; d < 3; ++d) {
                            gx[d] += di * eu[3 * a + d];
                            ge[d] += dj * eu[3 * b + d];
                            gz[d] += dk * eu[3 * c + d];
                        }
                    }
                    double g[3][3];
                    for (int c = 0; c < 3; ++c) {
                        g[c][0] = gx[c] * dd[0] + ge[c] * dd[1] + gz[c] * dd[2];
                        g[c][1] = gx[c] * dd[3] + ge[c] * dd[4] + gz[c] * dd[5];
                        g[c][2] = gx[c] * dd[6] + ge[c] * dd[7] + gz[c] * dd[8];
                    }
                    double* s6 = strain + (ci * n_node + n) * 6;
                    s6[0] = g[0][0];
                    s6[1] = g[1][1];
                    s6[2] = g[2][2];
                    s6[3] = 0.5 * (g[0][1] + g[1][0]);
                    s6[4] = 0.5 * (g[0][2] + g[2][0]);
                    s6[5] = 0.5 * (g[1][2] + g[2][1]);
                }
    }
}

/* ------------------------------------------------------------------ */
/* Whole run: CPU branch of forward/share/src/solver.cpp:168-688       */
/* ------------------------------------------------------------------ */

typedef struct {
    double *u, *v, *a, *r, *ut, *uloc, *rloc, *mass_g, *damp_g;
    double *pold, *pnew, *rmem;
    double *zeta, *devp;
    double *ccoef, *ccarry; /* complete C-PML: [P][27] coefficients, [P][45] memories */
    double *mu_in, *mu_pml;  /* complete C-PML with attenuation: mu zeroed in the PML / interior elements */
    double gll_x[16], gll_w[16], D[256];
    gfo_cpml pml;
} run_ws;

static void ws_free(run_ws* w) {
    free(w->u); free(w->v); free(w->a); free(w->r); free(w->ut);
    free(w->uloc); free(w->rloc); free(w->mass_g); free(w->damp_g);
    free(w->pold); free(w->pnew); free(w->rmem);
    free(w->zeta); free(w->devp);
    free(w->ccoef); free(w->ccarry);
    free(w->mu_in); free(w->mu_pml);
}

static int ws_init(run_ws* w, const gfo_problem* pb) {
    memset(w, 0, sizeof(*w));
    const int ngll = pb->ngll, N = ngll - 1;
    if (ngll < 2 || ngll > 16) return -1;
    const long n_node = (long)ngll * ngll * ngll;
    const long P = (long)pb->n_elem * n_node;
    const long G = pb->n_rank_node;
    gfo_gll_nodes(N, w->gll_x);
    gfo_gll_weights(N, w->gll_x, w->gll_w);
    gfo_gll_derivative_matrix(N, w->gll_x, w->D);
    w->u = calloc(G * 3, sizeof(double));
    w->v = calloc(G * 3, sizeof(double));
    w->a = calloc(G * 3, sizeof(double));
    w->r = calloc(G * 3, sizeof(double));
    w->ut = calloc(G * 3, sizeof(double));
    w->uloc = calloc(P * 3, sizeof(double));
    w->rloc = calloc(P * 3, sizeof(double));
    w->mass_g = calloc(G, sizeof(double));
    w->damp_g = calloc(G, sizeof(double));
    if (!w->u || !w->v || !w->a || !w->r || !w->ut || !w->uloc || !w->rloc || !w->mass_g ||
        !w->damp_g)
        return -2;
    if (pb->cpml_complete) {
        if (!pb->pml_region || !pb->pml_K || !pb->pml_d || !pb->pml_alpha) return -1;
        w->ccoef = calloc(P * 27, sizeof(double));
        w->ccarry = calloc(P * 45, sizeof(double));
        if (!w->ccoef || !w->ccarry) return -2;
        for (long p = 0; p < P; ++p)
            if (pb->pml_region[p / n_node] != 0)
                gfo_cpml_point_coefficients(pb->pml_K + p * 3, pb->pml_d + p * 3, pb->pml_alpha + p * 3, pb->dt,
                                            w->ccoef + p * 27);
        if (pb->n_sls > 0) {
            /* attenuation in the interior elements, elastic (relaxed moduli) in the
             * PML elements: each residual runs with mu zeroed where the other applies
             * (a point with mu <= 0 contributes nothing) */
            w->mu_in = malloc(sizeof(double) * P);
            w->mu_pml = malloc(sizeof(double) * P);
            if (!w->mu_in || !w->mu_pml) return -2;
            for (long p = 0; p < P; ++p) {
                const int pml = pb->pml_region[p / n_node] != 0;
                w->mu_in[p] = pml ? 0.0 : pb->mu_[p];
                w->mu_pml[p] = pml ? pb->mu_[p] : 0.0;
            }
        }
    }
    /* global mass sums, damping "last writer wins" (solver.cpp:193-202); the
     * complete C-PML weights the mass of PML element points by
     * K_x K_y K_z + dt/2 c_v (the inertia's velocity terms at v_{n+1}) */
    for (long p = 0; p < P; ++p) {
        const long g = pb->ibool[p];
        double m = pb->mass[p];
        if (pb->cpml_complete && pb->pml_region[p / n_node] != 0) {
            const double* K = pb->pml_K + p * 3;
            const double* d = pb->pml_d + p * 3;
            const double cv = K[2] * (K[1] * d[0] + K[0] * d[1]) + K[0] * K[1] * d[2];
            m *= K[0] * K[1] * K[2] + 0.5 * pb->dt * cv;
        }
        w->mass_g[g] += m;
        if (pb->damping) w->damp_g[g] = pb->damping[p];
    }
    if (pb->n_sls > 0) {
        if (pb->n_sls > 8 || !pb->tau_sigma || !pb->tau_epsilon) return -1;
        w->zeta = calloc(P * 5 * pb->n_sls, sizeof(double));
        w->devp = calloc(P * 5, sizeof(double));
        if (!w->zeta || !w->devp) return -2;
    }
    if (pb->has_cpml && !pb->cpml_complete) { /* cpml_initialize, pml.cpp:41-53 */
        w->pold = calloc(P * 3, sizeof(double));
        w->pnew = calloc(P * 3, sizeof(double));
        w->rmem = calloc(P * 9, sizeof(double));
        if (!w->pold || !w->pnew || !w->rmem) return -2;
        w->pml.n_elem = pb->n_elem;
        w->pml.ngll = ngll;
        w->pml.pml_region = pb->pml_region;
        w->pml.coef_alpha = pb->coef_alpha;
        w->pml.coef_abar = pb->coef_abar;
        w->pml.density = pb->density;
        w->pml.jacobian = pb->jacobian;
        w->pml.ibool = pb->ibool;
        w->pml.displ_old = w->pold;
        w->pml.displ_new = w->pnew;
        w->pml.rmemory = w->rmem;
    }
    return 0;
}

/* One solver step (solver.cpp:510-622). */
static void run_one_step(run_ws* w, const gfo_problem* pb, int step) {
    const int ngll = pb->ngll;
    const int n_node = ngll * ngll * ngll;
    const long P = (long)pb->n_elem * n_node;
    const int n_dof = pb->n_rank_node * 3;
    const double dt = pb->dt, beta = 0.0, gamma = 0.5;
    /* 1. predictor */
    gfo_newmark_predict(n_dof, dt, beta, w->u, w->v, w->a, w->ut);
    /* 3. gather predicted displacement (the reference calls it twice; idempotent) */
    gfo_gather_from_rank(w->ut, pb->ibool, pb->n_elem, n_node, w->uloc);
    const int ref_cpml = pb->has_cpml && !pb->cpml_complete;
    /* 3a. C-PML displacement fields from u_n, v_n, a_n */
    if (ref_cpml) gfo_cpml_update_displ_fields(&w->pml, w->u, w->v, w->a, dt);
    /* 3b. zero + element kernel */
    memset(w->rloc, 0, sizeof(double) * 3 * P);
    if (pb->cpml_complete) {
        /* stretched stresses in PML elements, then the inertia terms with the
         * predicted u and v_{n+1/2} = v_n + dt/2 a_n */
        if (pb->n_sls > 0)
            gfo_element_residual_att(pb->n_elem, pb->dxi_dx, pb->jacobian, pb->lambda_, w->mu_in, pb->n_sls,
                                     pb->tau_sigma, pb->tau_epsilon, dt, w->D, w->gll_w, ngll, w->uloc, w->zeta,
                                     w->devp, w->rloc);
        gfo_element_residual_cpml(pb->n_elem, pb->dxi_dx, pb->jacobian, pb->lambda_,
                                  pb->n_sls > 0 ? w->mu_pml : pb->mu_, pb->pml_region, w->ccoef, w->D, w->gll_w,
                                  ngll, w->uloc, w->ccarry, w->rloc);
#pragma omp parallel for schedule(static)
        for (long e = 0; e < pb->n_elem; ++e) {
            if (pb->pml_region[e] == 0) continue;
            for (int n = 0; n < n_node; ++n) {
                const long pt = e * n_node + n, g = pb->ibool[pt];
                double vh[3], L[3];
                for (int c = 0; c < 3; ++c) vh[c] = w->v[g * 3 + c] + 0.5 * dt * w->a[g * 3 + c];
                gfo_cpml_inertia(w->ccoef + pt * 27, w->ut + g * 3, vh, w->ccarry + pt * 45 + 36, L);
                for (int c = 0; c < 3; ++c) w->rloc[pt * 3 + c] -= pb->mass[pt] * L[c];
            }
        }
    } else if (pb->n_sls > 0)
        gfo_element_residual_att(pb->n_elem, pb->dxi_dx, pb->jacobian, pb->lambda_, pb->mu_, pb->n_sls,
                                 pb->tau_sigma, pb->tau_epsilon, dt, w->D, w->gll_w, ngll, w->uloc, w->zeta,
                                 w->devp, w->rloc);
    else
        gfo_element_residual(pb->n_elem, pb->dxi_dx, pb->jacobian, pb->lambda_, pb->mu_, w->D,
                             w->gll_w, ngll, w->uloc, w->rloc);
    /* 4. C-PML acceleration contribution, else legacy damping on v (solver.cpp:557-574) */
    if (ref_cpml) {
        gfo_cpml_accel_contribution(&w->pml, w->u, w->v, w->gll_w, w->rloc);
    } else if (!pb->cpml_complete) {
        gfo_apply_pml_damping(w->damp_g, w->v, (long)pb->n_rank_node * 3);
    }
    /* 5. source: stf_values[step] = f(t_n) (solver.cpp:576-605) */
    {
        double sv = 0.0;
        if (step < pb->n_stf) sv = pb->stf[step];
        if (sv != 0.0) {
            for (int si = 0; si < pb->n_src_cell; ++si) {
                const int e = pb->src_elem[si];
                if (e < 0) continue;
                for (int k = 0; k < ngll; ++k)
                    for (int j = 0; j < ngll; ++j)
                        for (int i = 0; i < ngll; ++i) {
                            const int no = nidx(i, j, k, ngll);
                            const double wt = pb->src_weights[(long)si * n_node + no];
                            if (wt == 0.0) continue;
                            w->rloc[((long)e * n_node + no) * 3 + pb->direction] += sv * wt;
                        }
            }
        }
    }
    /* 6. scatter (zero + add) */
    gfo_scatter_to_rank(w->rloc, pb->ibool, pb->n_elem, n_node, pb->n_rank_node, w->r);
    /* 7. corrector */
    gfo_newmark_correct(n_dof, dt, beta, gamma, w->mass_g, w->u, w->v, w->a, w->r);
    /* 8. C-PML memory */
    if (ref_cpml) gfo_cpml_update_displ_memory(&w->pml);
}

/* Steps [start_step, nsteps) of solver.cpp:389-688 from (u0, v0, a0) (zero when
 * NULL) with the C-PML state at zero: start_step > 0 is the reference's resume
 * (solver.cpp:369-388: state of the restart file, start_step = its step + 1). */
static int run_core(const gfo_problem* pb, int start_step, const double* u0, const double* v0,
                    const double* a0, double* rec_strain, double* rec_disp, double* rec_vel,
                    double* rec_acc, double* u_out, double* v_out, double* a_out) {
    run_ws w;
    if (!pb || pb->direction < 0 || pb->direction > 2 || start_step < 0) return -1;
    if (ws_init(&w, pb) != 0) {
        ws_free(&w);
        return -1;
    }
    const long n_dof = (long)pb->n_rank_node * 3;
    if (u0) memcpy(w.u, u0, sizeof(double) * n_dof);
    if (v0) memcpy(w.v, v0, sizeof(double) * n_dof);
    if (a0) memcpy(w.a, a0, sizeof(double) * n_dof);
    const int ngll = pb->ngll;
    const int n_node = ngll * ngll * ngll;
    const long nrec = pb->n_rec_cell;
    int snap = 0;
    double* strain_tmp = NULL;
    if (rec_strain && nrec > 0) strain_tmp = malloc(sizeof(double) * nrec * n_node * 6);
    for (int step = start_step; step < pb->nsteps; ++step) {
        run_one_step(&w, pb, step);
        if (pb->snapshot_stride > 0 && step % pb->snapshot_stride == 0) {
            /* solver.cpp:640-675: strain from gathered u after the corrector */
            if (nrec > 0) {
                gfo_gather_from_rank(w.u, pb->ibool, pb->n_elem, n_node, w.uloc);
                if (rec_strain) {
                    gfo_full_strain((int)nrec, pb->rec_cell, pb->dxi_dx, w.D, ngll, w.uloc,
                                    strain_tmp);
                    memcpy(rec_strain + (long)snap * nrec * n_node * 6, strain_tmp,
                           sizeof(double) * nrec * n_node * 6);
                }
                for (long ci = 0; ci < nrec; ++ci) {
                    const long e = pb->rec_cell[ci];
                    for (int n = 0; n < n_node; ++n) {
                        const long g = pb->ibool[e * n_node + n];
                        const long o = (((long)snap * nrec + ci) * n_node + n) * 3;
                        for (int d = 0; d < 3; ++d) {
                            if (rec_disp) rec_disp[o + d] = w.u[g * 3 + d];
                            if (rec_vel) rec_vel[o + d] = w.v[g * 3 + d];
                            if (rec_acc) rec_acc[o + d] = w.a[g * 3 + d];
                        }
                    }
                }
            }
            ++snap;
        }
    }
    const long nd = (long)pb->n_rank_node * 3;
    if (u_out) memcpy(u_out, w.u, sizeof(double) * nd);
    if (v_out) memcpy(v_out, w.v, sizeof(double) * nd);
    if (a_out) memcpy(a_out, w.a, sizeof(double) * nd);
    free(strain_tmp);
    ws_free(&w);
    return snap;
}

int gfo_run(const gfo_problem* pb, double* rec_strain, double* rec_disp, double* rec_vel,
            double* rec_acc, double* u_out, double* v_out, double* a_out) {
    return run_core(pb, 0, NULL, NULL, NULL, rec_strain, rec_disp, rec_vel, rec_acc, u_out, v_out, a_out);
}

int gfo_run_from(const gfo_problem* pb, int start_step, double* u, double* v, double* a,
                 double* rec_strain, double* rec_disp, double* rec_vel, double* rec_acc) {
    return run_core(pb, start_step, u, v, a, rec_strain, rec_disp, rec_vel, rec_acc, u, v, a);
}

int gfo_run_state(const gfo_problem* pb, double* u, double* v, double* a) {
    run_ws w;
    if (!pb || pb->direction < 0 || pb->direction > 2) return -1;
    if (ws_init(&w, pb) != 0) {
        ws_free(&w);
        return -1;
    }
    const long nd = (long)pb->n_rank_node * 3;
    memcpy(w.u, u, sizeof(double) * nd);
    memcpy(w.v, v, sizeof(double) * nd);
    memcpy(w.a, a, sizeof(double) * nd);
    for (int step = 0; step < pb->nsteps; ++step) run_one_step(&w, pb, step);
    memcpy(u, w.u, sizeof(double) * nd);
    memcpy(v, w.v, sizeof(double) * nd);
    memcpy(a, w.a, sizeof(double) * nd);
    ws_free(&w);
    return 0;
}

/* tests/test_element_cuda.cu:27-32 */
void gfo_random_fill(double* v, int n, unsigned seed) {
    srand(seed);
    for (int i = 0; i < n; ++i) v[i] = 2.0 * (double)rand() / RAND_MAX - 1.0;
}

/* Threads used by the parallel loops (OpenMP; 1 when built without it). */
int gfo_set_threads(int n) {
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

