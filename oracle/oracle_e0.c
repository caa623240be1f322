/*
 * CPU oracle: the E0 operator checks of verifica_operadores.cpp:36-272, evaluated on the
 * oracle's own operator functions.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_e0() fills err[] / tol[] in the order of the checks printed in E0_resultado.txt
 * (4 P operators x {M0, M1, M2}, 4 nz = 1 collapse rows, 3 Guo rows, 3 xi x 3 F operators x
 * {plane, disc}) and m4[] with the 4 x 5 table of M4_nnnn / (A |g|).  It returns the number of
 * failed checks.  The P-class moment checks use their own normal generator (the reference uses
 * std::mt19937(7)); only their tolerance is meaningful, every other row is deterministic.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;

static uint64_t sm_state = 7;
static double uniform01(void)
{
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double normal01(void)
{
    const double u1 = uniform01(), u2 = uniform01();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
}

static double dmax(double a, double b) { return (a < b) ? b : a; }

int orc_e0(double *err, double *tol, double *m4)
{
    double c[ORC_Q * 3], w[ORC_Q];
    orc_lattice(c, w);
    const double A = 0.37, mg = 0.81;
    const int ops[4] = {ORC_OP_P1, ORC_OP_P3, ORC_OP_P3, ORC_OP_P3};
    const double chis[4] = {0.5, -0.4, 0.5, 2.0};
    const double lambdas[4] = {0.0, 1.0, 1.0, 1.0};
    int k = 0, fails = 0;
    sm_state = 7;

    /* 1) P class: conservation and second moment, verifica_operadores.cpp:58-98 */
    for (int s = 0; s < 4; s++) {
        double e0 = 0, e1 = 0, e2 = 0;
        for (int t = 0; t < 200; t++) {
            double n[3] = {normal01(), normal01(), normal01()};
            const double m = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int a = 0; a < 3; a++) n[a] /= m;
            const double g[3] = {mg * n[0], mg * n[1], mg * n[2]};
            double f[ORC_Q];
            memset(f, 0, sizeof f);
            orc_perturbation(f, A, g, ops[s], chis[s]);
            double M0 = 0, M1[3] = {0, 0, 0}, M2[3][3] = {{0}};
            for (int i = 0; i < ORC_Q; i++) {
                const double *ci = c + 3 * i;
                M0 += f[i];
                for (int a = 0; a < 3; a++) {
                    M1[a] += f[i] * ci[a];
                    for (int b = 0; b < 3; b++) M2[a][b] += f[i] * ci[a] * ci[b];
                }
            }
            e0 = dmax(e0, fabs(M0));
            for (int a = 0; a < 3; a++) {
                e1 = dmax(e1, fabs(M1[a]));
                for (int b = 0; b < 3; b++) {
                    const double alvo = 2.0 / 9.0 * A * mg * (n[a] * n[b] - lambdas[s] * (a == b));
                    e2 = dmax(e2, fabs(M2[a][b] - alvo));
                }
            }
        }
        err[k] = e0; tol[k++] = 1e-14;
        err[k] = e1; tol[k++] = 1e-14;
        err[k] = e2; tol[k++] = 1e-14;
    }

    /* 4th-moment table, verifica_operadores.cpp:104-127 */
    const double ths[5] = {0.0, 15.0, 22.5, 30.0, 45.0};
    for (int s = 0; s < 4; s++)
        for (int j = 0; j < 5; j++) {
            const double t = ths[j] * PI / 180.0;
            const double n[3] = {cos(t), sin(t), 0.0};
            const double g[3] = {mg * n[0], mg * n[1], 0.0};
            double f[ORC_Q];
            memset(f, 0, sizeof f);
            orc_perturbation(f, A, g, ops[s], chis[s]);
            double M4 = 0;
            for (int i = 0; i < ORC_Q; i++) {
                const double *ci = c + 3 * i;
                const double cn = ci[0] * n[0] + ci[1] * n[1];
                M4 += f[i] * cn * cn * cn * cn;
            }
            m4[5 * s + j] = M4 / (A * mg);
        }

    /* 2) nz = 1 collapse onto the D2Q9 family, verifica_operadores.cpp:133-171 */
    const double chis2[4] = {-0.4, 0.5, 2.0, 7.0};
    for (int s = 0; s < 4; s++) {
        const double chi = chis2[s];
        const double t = 0.3;
        const double g[3] = {mg * cos(t), mg * sin(t), 0.0};
        double f[ORC_Q];
        memset(f, 0, sizeof f);
        orc_perturbation(f, A, g, ORC_OP_P3, chi);
        double f9[3][3] = {{0}};
        for (int i = 0; i < ORC_Q; i++) f9[(int)c[3 * i] + 1][(int)c[3 * i + 1] + 1] += f[i];
        const double c2x = chi + 2.0;
        const double B9[3] = {-c2x / (3.0 * (c2x + 2.0)), c2x / (6.0 * (c2x + 2.0)), 1.0 / (6.0 * (c2x + 2.0))};
        double erro = 0;
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++) {
                const int sh = abs(a) + abs(b);
                const double w9 = (sh == 0) ? 4.0 / 9.0 : (sh == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
                const double cg = a * g[0] + b * g[1];
                const double om = A * mg * (w9 * cg * cg / (mg * mg) - B9[sh]);
                erro = dmax(erro, fabs(om - f9[a + 1][b + 1]));
            }
        err[k] = erro; tol[k++] = 1e-15;
    }

    /* 3) Guo source, verifica_operadores.cpp:175-187 */
    {
        const double tau = 0.8, rho = 1.03, F[3] = {1.3e-3, -2.1e-3, 0.0}, u[3] = {0.02, -0.01, 0.0};
        double S[ORC_Q];
        orc_source(F[0], F[1], F[2], u[0], u[1], u[2], rho, tau, S);
        double M0 = 0, M1[3] = {0, 0, 0};
        for (int i = 0; i < ORC_Q; i++) {
            M0 += S[i];
            for (int a = 0; a < 3; a++) M1[a] += S[i] * c[3 * i + a];
        }
        err[k] = M0; tol[k++] = 1e-16;
        err[k] = M1[0] - (1 - 0.5 / tau) * F[0]; tol[k++] = 1e-16;
        err[k] = M1[1] - (1 - 0.5 / tau) * F[1]; tol[k++] = 1e-16;
    }

    /* 4) F class on tanh profiles, verifica_operadores.cpp:191-267 */
    {
        const int nx = 256, ny = 256;
        const double sigma = 0.1, R = 60.0;
        double *fR = (double *)malloc(sizeof(double) * (size_t)nx * ny * ORC_Q);
        double *fB = (double *)malloc(sizeof(double) * (size_t)nx * ny * ORC_Q);
        double *F = (double *)malloc(sizeof(double) * 3 * (size_t)nx * ny);
        const double xis[3] = {1.569988, 3.0, 6.0};
        const int fops[3] = {ORC_OP_F1, ORC_OP_F2, ORC_OP_F3};
        for (int ix = 0; ix < 3; ix++)
            for (int io = 0; io < 3; io++) {
                const double xi = xis[ix];
                orc_case cs;
                memset(&cs, 0, sizeof cs);
                cs.nx = nx; cs.ny = ny; cs.op = fops[io]; cs.xi = xi; cs.A = sigma; cs.tau = 1.0;
                for (int disco = 0; disco < 2; disco++) {
                    for (int y = 0; y < ny; y++)
                        for (int x = 0; x < nx; x++) {
                            const double s = disco ? R - hypot(x - nx / 2.0, y - ny / 2.0)
                                                   : (fabs(x - nx / 2.0) < nx / 4.0 ? nx / 4.0 - fabs(x - nx / 2.0)
                                                                                    : -(fabs(x - nx / 2.0) - nx / 4.0));
                            const double ph = tanh(s / xi);
                            const int p = x + y * nx;
                            for (int i = 0; i < ORC_Q; i++) {
                                fR[(size_t)p * ORC_Q + i] = w[i] * 0.5 * (1 + ph);
                                fB[(size_t)p * ORC_Q + i] = w[i] * 0.5 * (1 - ph);
                            }
                        }
                    orc_force_field(&cs, fR, fB, F);
                    if (!disco) {
                        double fmax = 0;
                        for (int p = 0; p < nx * ny; p++) fmax = dmax(fmax, fabs(F[3 * p]) + fabs(F[3 * p + 1]));
                        err[k] = fmax / (sigma / xi);
                        tol[k++] = (fops[io] == ORC_OP_F3) ? 0.2 / (xi * xi) : 1e-3;
                    } else {
                        double soma = 0;
                        for (int y = 0; y < ny; y++)
                            for (int x = 0; x < nx; x++) {
                                const double dx = x - nx / 2.0, dy = y - ny / 2.0, r = hypot(dx, dy);
                                if (r < 1e-9) continue;
                                const int p = x + y * nx;
                                soma += -(F[3 * p] * dx + F[3 * p + 1] * dy) / r / (2.0 * PI * r);
                            }
                        err[k] = soma / (sigma / R) - 1.0;
                        tol[k++] = 1.5 * (xi / R) * (xi / R) + 5e-4 + (fops[io] == ORC_OP_F3 ? 0.25 / (xi * xi) : 0.0);
                    }
                }
            }
        free(fR); free(fB); free(F);
    }
    for (int j = 0; j < k; j++)
        if (!(fabs(err[j]) <= tol[j])) fails++;
    return fails;
}
