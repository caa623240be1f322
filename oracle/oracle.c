/*
 * CPU oracle: lattice, collision, tension operators, streaming, initial conditions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the reference lines it
 * restates; every double expression keeps the reference's order of evaluation.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double C[ORC_Q * 3];
static double W[ORC_Q];
static int lattice_ready = 0;

/* D3Q19 ordering and weights: LBM_functions.cpp:366-497 */
void orc_lattice(double c[ORC_Q * 3], double w[ORC_Q])
{
    static const int cc[ORC_Q][3] = {
        {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
        {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
        {-1, 0, 1}, {0, -1, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}};
    for (int i = 0; i < ORC_Q; i++) {
        for (int a = 0; a < 3; a++) c[3 * i + a] = (double)cc[i][a];
        w[i] = (i == 0) ? 1. / 3. : (i < 7 ? 1. / 18. : 1. / 36.);
    }
}

static void ensure_lattice(void)
{
    if (!lattice_ready) {
        orc_lattice(C, W);
        lattice_ready = 1;
    }
}

/* dot_product, LBM_functions.cpp:865-877 (accumulates from 0.0) */
static double dot3(const double *a, const double *b)
{
    double r = 0.0;
    for (int k = 0; k < 3; k++) r = r + a[k] * b[k];
    return r;
}

/* density, LBM_functions.cpp:1581-1590 */
static double density(const double *f)
{
    double rho = 0.0;
    for (int i = 0; i < ORC_Q; i++) rho = rho + f[i];
    return rho;
}

/* momentum, LBM_functions.cpp:1056-1087 (3D branch) */
static void momentum(const double *f, double *mx, double *my, double *mz)
{
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = 0; i < ORC_Q; i++) {
        x = x + C[3 * i + 0] * f[i];
        y = y + C[3 * i + 1] * f[i];
        z = z + C[3 * i + 2] * f[i];
    }
    *mx = x; *my = y; *mz = z;
}

/* calcula, LBM_functions.cpp:1099-1171 (3D branch) */
static void calcula(const double *f, double *vx, double *vy, double *vz, double *rho_out)
{
    double mx = 0.0, my = 0.0, mz = 0.0, rho = 0.0;
    for (int i = 0; i < ORC_Q; i++) {
        rho = rho + f[i];
        mx = mx + C[3 * i + 0] * f[i];
        my = my + C[3 * i + 1] * f[i];
        mz = mz + C[3 * i + 2] * f[i];
    }
    if (rho != 0.0) {
        const double inv = 1.0 / rho;
        *vx = mx * inv; *vy = my * inv; *vz = mz * inv;
    } else {
        *vx = 0.; *vy = 0.; *vz = 0.;
    }
    *rho_out = rho;
}

/* dist_eq, LBM_functions.cpp:889-945 (3D branch), one_over_c_s2 = 3 */
static void dist_eq(double *feq, double vx, double vy, double vz, double rho)
{
    const double v[3] = {vx, vy, vz};
    const double vquad = (vx * vx + vy * vy + vz * vz);
    const double o = 3.0;
    for (int i = 0; i < ORC_Q; i++) {
        const double cv = dot3(C + 3 * i, v);
        feq[i] = W[i] * rho * (1.0 + cv * o + 0.5 * cv * cv * o * o - 0.5 * vquad * o);
    }
}

/* source (Guo), LBM_functions.cpp:1602-1673 (3D branch) */
void orc_source(double Fx, double Fy, double Fz, double vx, double vy, double vz, double rho,
                double tau, double S[ORC_Q])
{
    ensure_lattice();
    const double o = 3.0;
    const double v[3] = {vx, vy, vz};
    const double F[3] = {Fx, Fy, Fz};
    const double factor = (1. - 1. / (2. * tau));
    const double vF = dot3(F, v);
    (void)rho;
    for (int i = 0; i < ORC_Q; i++) {
        const double cF = dot3(C + 3 * i, F);
        const double cv = dot3(C + 3 * i, v);
        S[i] = factor * W[i] * o * (cF + o * cF * cv - vF);
    }
}

/* coll_BGK, Collision_operators.cpp:11-49 (nvel = 19 -> dist_eq) */
static void coll_bgk(double *f, double ax, double ay, double az, double tau)
{
    double feq[ORC_Q], S[ORC_Q];
    double vx, vy, vz, rho;
    const double one_over_tau = 1.0 / tau;
    calcula(f, &vx, &vy, &vz, &rho);
    const double vxa = vx + 0.5 * ax, vya = vy + 0.5 * ay, vza = vz + 0.5 * az;
    const double Fx = ax * rho, Fy = ay * rho, Fz = az * rho;
    orc_source(Fx, Fy, Fz, vxa, vya, vza, rho, tau, S);
    dist_eq(feq, vxa, vya, vza, rho);
    for (int i = 0; i < ORC_Q; i++) {
        const double opc = (feq[i] - f[i]) * one_over_tau;
        f[i] = f[i] + opc + S[i];
    }
}

/* coef_B, Operadores_tensao.cpp:94-109 */
static double coef_b(int op, double chi, double c2, double w)
{
    if (op == ORC_OP_P1) return w / 3.0;
    const double den = 6.0 * chi + 24.0;
    if (c2 < 0.5) return -(2.0 + 2.0 * chi) / (3.0 * chi + 12.0);
    if (c2 < 1.5) return chi / den;
    return 1.0 / den;
}

/* perturbacao, Operadores_tensao.cpp:116-133 */
void orc_perturbation(double *f, double A, const double g[3], int op, double chi)
{
    ensure_lattice();
    const double mod2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (mod2 <= 1.0e-60 || A == 0.0) return;
    const double mod = sqrt(mod2);
    for (int i = 0; i < ORC_Q; i++) {
        const double *c = C + 3 * i;
        const double cg = c[0] * g[0] + c[1] * g[1] + c[2] * g[2];
        const double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        f[i] += A * mod * (W[i] * cg * cg / mod2 - coef_b(op, chi, c2, W[i]));
    }
}

/* recolloring, LBM_functions.cpp:1712-1746 */
static void recolour(const double *f, double *fR, double *fB, const double *grad, double mod,
                     double cR, double cB, double rho, double beta)
{
    const double fator = beta * cR * cB;
    fR[0] = cR * f[0];
    fB[0] = cB * f[0];
    for (int i = 1; i < ORC_Q; i++) {
        const double *c = C + 3 * i;
        const double c_2 = dot3(c, c);
        const double prod = dot3(c, grad);
        double cosp;
        if (mod != 0.0) cosp = prod / (mod * sqrt(c_2));
        else cosp = 0.0;
        fR[i] = cR * f[i] - fator * W[i] * rho * cosp;
        fB[i] = cB * f[i] + fator * W[i] * rho * cosp;
    }
}

/* colisao_comum, Operadores_tensao.cpp:340-373 */
void orc_collide_site(double *fR, double *fB, const double *fm, const double F[3], const orc_case *cs)
{
    ensure_lattice();
    const double rho_R = density(fR);
    const double rho_B = density(fB);
    const double rho = rho_R + rho_B;
    if (rho <= 1.0e-30) return;
    const double cR = rho_R / rho;
    const double cB = 1.0 - cR;
    double gm[3];
    momentum(fm, &gm[0], &gm[1], &gm[2]);
    const double mod = sqrt(gm[0] * gm[0] + gm[1] * gm[1] + gm[2] * gm[2]);
    double f[ORC_Q];
    for (int i = 0; i < ORC_Q; i++) f[i] = fR[i] + fB[i];
    if (cs->op >= ORC_OP_F1) {
        coll_bgk(f, F[0] / rho, F[1] / rho, F[2] / rho, cs->tau);
    } else {
        coll_bgk(f, 0.0, 0.0, 0.0, cs->tau);
        orc_perturbation(f, cs->A, gm, cs->op, cs->chi);
    }
    recolour(f, fR, fB, gm, mod, cR, cB, rho, cs->beta);
}

/* site index of (x + dx, y + dy), periodic; z always wraps onto the single plane */
static inline int nb(int x, int y, int dx, int dy, int nx, int ny)
{
    int xx = x + dx, yy = y + dy;
    if (xx < 0) xx += nx;
    if (xx >= nx) xx -= nx;
    if (yy < 0) yy += ny;
    if (yy >= ny) yy -= ny;
    return xx + nx * yy;
}

/* direction opposite to i (c_opp = -c_i), found as def_dir_prop does (LBM_functions.cpp:713-733) */
static int opposite(int i)
{
    for (int j = 0; j < ORC_Q; j++)
        if (C[3 * j] == -C[3 * i] && C[3 * j + 1] == -C[3 * i + 1] && C[3 * j + 2] == -C[3 * i + 2]) return j;
    return i;
}

/* calcula_forca, Operadores_tensao.cpp:202-332 (viz = x + c_i, aloca_campos_F :149-182: a solid
 * neighbour reflects onto the site itself).  solid (n sites, 1 = solid) may be NULL: all fluid.
 * Solid sites hold no fields; their F is set to 0. */
static void force_field_impl(const orc_case *cs, const unsigned char *solid, const double *fR, const double *fB,
                             double *F)
{
    ensure_lattice();
    const int nx = cs->nx, ny = cs->ny, n = nx * ny;
    const double sigma = cs->A;
    double *fase = (double *)calloc((size_t)n, sizeof(double));
    double *g = (double *)calloc((size_t)3 * n, sizeof(double));
    double *aux = (double *)calloc((size_t)6 * n, sizeof(double));
    int *viz = (int *)malloc(sizeof(int) * ORC_Q * n);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++)
            for (int i = 0; i < ORC_Q; i++) {
                const int v = nb(x, y, (int)C[3 * i], (int)C[3 * i + 1], nx, ny);
                viz[(x + nx * y) * ORC_Q + i] = (solid && solid[v]) ? x + nx * y : v;
            }
#define ORC_FLUID(p) (!solid || !solid[p])
    for (int p = 0; p < n; p++) {
        if (!ORC_FLUID(p)) continue;
        const double rR = density(fR + (size_t)p * ORC_Q);
        const double rB = density(fB + (size_t)p * ORC_Q);
        fase[p] = (rR - rB) / (rR + rB);
    }
    for (int p = 0; p < n; p++) {
        if (!ORC_FLUID(p)) continue;
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int i = 1; i < ORC_Q; i++) {
            const double ph = W[i] * fase[viz[p * ORC_Q + i]];
            gx += ph * C[3 * i + 0];
            gy += ph * C[3 * i + 1];
            gz += ph * C[3 * i + 2];
        }
        g[3 * p + 0] = 3.0 * gx;
        g[3 * p + 1] = 3.0 * gy;
        g[3 * p + 2] = 3.0 * gz;
    }
    const double xi = cs->xi;
    const double kf = 0.75 * sigma * xi;
    const double bf = 3.0 * sigma / (8.0 * xi);
    for (int p = 0; p < n; p++) {
        if (!ORC_FLUID(p)) continue;
        const double *gp = g + 3 * p;
        double *a = aux + 6 * p;
        const double m = sqrt(gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2]);
        if (cs->op == ORC_OP_F1) {
            const double inv = (m > 1.0e-12) ? 1.0 / m : 0.0;
            a[0] = gp[0] * inv; a[1] = gp[1] * inv; a[2] = gp[2] * inv;
        } else if (cs->op == ORC_OP_F2) {
            const double h = 0.5 * sigma;
            const double inv = (m > 1.0e-30) ? 1.0 / m : 0.0;
            a[0] = h * (m - gp[0] * gp[0] * inv);
            a[1] = h * (m - gp[1] * gp[1] * inv);
            a[2] = h * (m - gp[2] * gp[2] * inv);
            a[3] = -h * gp[0] * gp[1] * inv;
            a[4] = -h * gp[0] * gp[2] * inv;
            a[5] = -h * gp[1] * gp[2] * inv;
        } else {
            const double ph = fase[p];
            double lap = 0.0;
            for (int i = 1; i < ORC_Q; i++) lap += W[i] * (fase[viz[p * ORC_Q + i]] - ph);
            lap *= 6.0;
            a[0] = 4.0 * bf * ph * (ph * ph - 1.0) - kf * lap;
        }
    }
    for (int p = 0; p < n; p++) {
        const double *gp = g + 3 * p;
        double *Fp = F + 3 * p;
        if (!ORC_FLUID(p)) {
            Fp[0] = Fp[1] = Fp[2] = 0.0;
            continue;
        }
        if (cs->op == ORC_OP_F1) {
            double d = 0.0;
            for (int i = 1; i < ORC_Q; i++) {
                const double *ci = C + 3 * i;
                const double *vv = aux + (size_t)viz[p * ORC_Q + i] * 6;
                d += W[i] * (ci[0] * vv[0] + ci[1] * vv[1] + ci[2] * vv[2]);
            }
            const double kappa = -(3.0 * d);
            Fp[0] = 0.5 * sigma * kappa * gp[0];
            Fp[1] = 0.5 * sigma * kappa * gp[1];
            Fp[2] = 0.5 * sigma * kappa * gp[2];
        } else if (cs->op == ORC_OP_F2) {
            double Fx = 0.0, Fy = 0.0, Fz = 0.0;
            for (int i = 1; i < ORC_Q; i++) {
                const double *ci = C + 3 * i;
                const double *T = aux + (size_t)viz[p * ORC_Q + i] * 6;
                const double w = W[i];
                Fx += w * (T[0] * ci[0] + T[3] * ci[1] + T[4] * ci[2]);
                Fy += w * (T[3] * ci[0] + T[1] * ci[1] + T[5] * ci[2]);
                Fz += w * (T[4] * ci[0] + T[5] * ci[1] + T[2] * ci[2]);
            }
            Fp[0] = 3.0 * Fx; Fp[1] = 3.0 * Fy; Fp[2] = 3.0 * Fz;
        } else {
            const double mu = aux[6 * p];
            Fp[0] = mu * gp[0]; Fp[1] = mu * gp[1]; Fp[2] = mu * gp[2];
        }
    }
    free(fase); free(g); free(aux); free(viz);
}
#undef ORC_FLUID

void orc_force_field(const orc_case *cs, const double *fR, const double *fB, double *F)
{
    force_field_impl(cs, NULL, fR, fB, F);
}

/* One step: Imm_OPS.cpp:246-372 (mediators :252-267, emite_mediadores LBM_functions.cpp:1286-1292,
 * propag_site_med :1309-1337, force :271, collide + propag_site :275-287, LBM_functions.cpp:1183-1253).
 * With solid sites (solid != NULL, 1 = solid): only fluid sites are collided; a population pushed
 * towards a solid site lands in the opposite direction at its own site (def_dir_prop
 * LBM_functions.cpp:509-853, odd number of inversions), and the mediator that would come from a solid
 * site is the wall's, w_i wett 3 (propag_site_med :1331).  Solid sites hold no populations: fR_new /
 * fB_new are 0 there. */
static void step_impl(const orc_case *cs, const unsigned char *solid, double wett, double *fR, double *fB,
                      double *fR_new, double *fB_new, double *F, double *gm)
{
    ensure_lattice();
    const int nx = cs->nx, ny = cs->ny, n = nx * ny;
    double *phi = (double *)calloc((size_t)n, sizeof(double));
    for (int p = 0; p < n; p++) {
        if (solid && solid[p]) continue;
        const double rho_R = density(fR + (size_t)p * ORC_Q);
        const double rho_B = density(fB + (size_t)p * ORC_Q);
        phi[p] = (rho_R - rho_B) / (rho_R + rho_B);
    }
    const int forced = cs->op >= ORC_OP_F1;
    if (forced) force_field_impl(cs, solid, fR, fB, F);
    static const double zero[3] = {0.0, 0.0, 0.0};
    if (solid)
        for (int p = 0; p < n; p++)
            if (solid[p])
                for (int i = 0; i < ORC_Q; i++) {
                    fR_new[(size_t)p * ORC_Q + i] = 0.0;
                    fB_new[(size_t)p * ORC_Q + i] = 0.0;
                }
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int p = x + nx * y;
            if (solid && solid[p]) {
                if (gm) gm[3 * p] = gm[3 * p + 1] = gm[3 * p + 2] = 0.0;
                continue;
            }
            /* mediator populations after propagation: M_i(x) = w_i * (3 phi(x - c_i)), or the wall's
             * (w_i wett) 3 when x - c_i is solid (w of the emitted direction -c_i equals w_i) */
            double fm[ORC_Q];
            for (int i = 0; i < ORC_Q; i++) {
                const int src = nb(x, y, -(int)C[3 * i], -(int)C[3 * i + 1], nx, ny);
                if (solid && solid[src]) {
                    fm[i] = W[i] * wett * 3.0;
                } else {
                    const double fat = phi[src] * 3.0;
                    fm[i] = W[i] * fat;
                }
            }
            if (gm) momentum(fm, &gm[3 * p], &gm[3 * p + 1], &gm[3 * p + 2]);
            orc_collide_site(fR + (size_t)p * ORC_Q, fB + (size_t)p * ORC_Q, fm,
                             forced ? F + 3 * p : zero, cs);
            for (int i = 0; i < ORC_Q; i++) {
                const int dst = nb(x, y, (int)C[3 * i], (int)C[3 * i + 1], nx, ny);
                const size_t to = (solid && solid[dst]) ? (size_t)p * ORC_Q + opposite(i) : (size_t)dst * ORC_Q + i;
                fR_new[to] = fR[(size_t)p * ORC_Q + i];
                fB_new[to] = fB[(size_t)p * ORC_Q + i];
            }
        }
    free(phi);
}

void orc_step(const orc_case *cs, double *fR, double *fB, double *fR_new, double *fB_new,
              double *F, double *gm)
{
    step_impl(cs, NULL, 0.0, fR, fB, fR_new, fB_new, F, gm);
}

void orc_step_geo(const orc_case *cs, const unsigned char *solid, double wett, double *fR, double *fB,
                  double *fR_new, double *fB_new, double *F, double *gm)
{
    step_impl(cs, solid, wett, fR, fB, fR_new, fB_new, F, gm);
}

/* initial_conditions_drop, Initial_conditions.cpp:37-169, EQ_CLASSICO branch :141-145,
 * cylinder (dz = 0), rho_ini_R = rho_ini_B = 1, delta_rho = 0 (Imm_OPS.cpp:199-202) */
void orc_init_drop(const orc_case *cs, double R, double k_perfil, int x0, int y0, double v0,
                   double *fR, double *fB)
{
    ensure_lattice();
    const int nx = cs->nx, ny = cs->ny;
    const double rho_ini_R = 1.0, rho_ini_B = 1.0, delta_rho = 0.0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            const double dx = (double)(x - x0), dy = (double)(y - y0), dz = 0.0;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            double phi;
            if (k_perfil > 0.0) {
                double arg = k_perfil * (r - R);
                if (arg > 60.0) arg = 60.0;
                if (arg < -60.0) arg = -60.0;
                phi = 1.0 / (1.0 + exp(arg));
            } else {
                phi = (r < R) ? 1.0 : 0.0;
            }
            const double rho_tot = rho_ini_B + phi * (rho_ini_R - rho_ini_B) + phi * delta_rho;
            const int p = x + nx * y;
            const double rR = phi * rho_tot;
            const double rB = (1.0 - phi) * rho_tot;
            dist_eq(fR + (size_t)p * ORC_Q, v0, 0., 0., rR);
            dist_eq(fB + (size_t)p * ORC_Q, v0, 0., 0., rB);
        }
}

static long long gcdll(long long a, long long b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { long long t = a % b; a = b; b = t; }
    return a;
}

/* Imm_INCL.cpp:309-334: rest equilibrium rho = 1, red fraction logistic in the signed distance to
 * the nearest interface.  Generalised from N x N to nx x ny with s = (p x + q y) mod M,
 * M = gcd(p nx, q ny) / gcd(p, q), which is N for every normal on an N x N box (the reference's
 * "% N", Imm_INCL.cpp:314-315, also for a non-coprime normal such as (2, 2)).  width <= 0 gives a
 * sharp band (cR = 1 / 0 off the interface, 1/2 on it: the limit of the logistic). */
void orc_init_band(const orc_case *cs, int p, int q, double width, double *fR, double *fB)
{
    ensure_lattice();
    const int nx = cs->nx, ny = cs->ny;
    const long long M = gcdll((long long)p * nx, (long long)q * ny) / gcdll(p, q);
    const double m = sqrt((double)(p * p + q * q));
    const double k = (width > 0) ? 2.0 * log(9.0) / width : -1.0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            long long s = ((long long)p * x + (long long)q * y) % M;
            if (s < 0) s += M;
            const double a = (double)s, b = 0.5 * (double)M - (double)s;
            const double c1 = (double)s - 0.5 * (double)M, c2 = (double)(M - s);
            const double dred = (2 * s < M) ? (a < b ? a : b) / m : -(c1 < c2 ? c1 : c2) / m;
            double cR;
            if (k > 0) cR = 1.0 / (1.0 + exp(-k * dred));
            else cR = (dred > 0) ? 1.0 : (dred < 0 ? 0.0 : 0.5);
            const int pto = x + nx * y;
            for (int i = 0; i < ORC_Q; i++) {
                fR[(size_t)pto * ORC_Q + i] = W[i] * cR;
                fB[(size_t)pto * ORC_Q + i] = W[i] * (1.0 - cR);
            }
        }
}
