/*
 * CPU oracle: diagnostics on post-collision populations.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Sequential loops in the reference's site order
 * (the 1-thread order of its OpenMP reductions); expressions keep the reference's order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;
static const double CS2 = 1. / 3.;

static double Cd[ORC_Q * 3];
static double Wd[ORC_Q];
static int ready = 0;
static void ensure(void)
{
    if (!ready) { orc_lattice(Cd, Wd); ready = 1; }
}

static double density(const double *f)
{
    double rho = 0.0;
    for (int i = 0; i < ORC_Q; i++) rho = rho + f[i];
    return rho;
}

/* calcula, LBM_functions.cpp:1099-1171 */
static void calcula(const double *f, double *vx, double *vy, double *vz, double *rho_out)
{
    double mx = 0.0, my = 0.0, mz = 0.0, rho = 0.0;
    for (int i = 0; i < ORC_Q; i++) {
        rho = rho + f[i];
        mx = mx + Cd[3 * i + 0] * f[i];
        my = my + Cd[3 * i + 1] * f[i];
        mz = mz + Cd[3 * i + 2] * f[i];
    }
    if (rho != 0.0) {
        const double inv = 1.0 / rho;
        *vx = mx * inv; *vy = my * inv; *vz = mz * inv;
    } else {
        *vx = 0.; *vy = 0.; *vz = 0.;
    }
    *rho_out = rho;
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* minimum-image distance, Other_functions.cpp:230-232 */
static inline double image(double d, int n)
{
    if (d > 0.5 * n) d -= n;
    if (d < -0.5 * n) d += n;
    return d;
}

/* gi_image, Galilean_diagnostics.cpp:18-20 */
static inline double gi_image(double dx, double length) { return dx - length * floor(dx / length + 0.5); }

typedef struct {
    double x0, y0, raio, raio_grad, raio_linha, massa_R, massa_B, qx, qy, qz, u_max, u_rms, area;
    int n_fluid, n_ruim;
    double p_in, p_out, sd_in, sd_out, delta_p, sigma, p_centro, p_canto, sigma_linha;
    int n_in, n_out;
} drop_t;

/* calc_drop_radius, Other_functions.cpp:60-276 (cilindro = true, nz = 1) */
static void drop_radius(const orc_case *cs, const double *fR, const double *fB, const double *F, drop_t *d)
{
    const int nx = cs->nx, ny = cs->ny, nz = 1;
    const int nv = nx * ny * nz;
    const double tx = 2.0 * PI / (double)nx, ty = 2.0 * PI / (double)ny;
    double massa_R = 0.0, massa_B = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
    double area = 0.0, u2_sum = 0.0, u_max = 0.0;
    double cx_c = 0.0, cx_s = 0.0, cy_c = 0.0, cy_s = 0.0;
    int n_ruim = 0, n_fluid = 0;
    for (int pos = 0; pos < nv; pos++) {
        double vxR, vyR, vzR, rho_R, vxB, vyB, vzB, rho_B;
        calcula(fR + (size_t)pos * ORC_Q, &vxR, &vyR, &vzR, &rho_R);
        calcula(fB + (size_t)pos * ORC_Q, &vxB, &vyB, &vzB, &rho_B);
        n_fluid = n_fluid + 1;
        if (!(rho_R == rho_R) || !(rho_B == rho_B) || !(vxR == vxR) || !(vxB == vxB)) {
            n_ruim = n_ruim + 1;
            continue;
        }
        massa_R = massa_R + rho_R;
        massa_B = massa_B + rho_B;
        qx = qx + rho_R * vxR + rho_B * vxB;
        qy = qy + rho_R * vyR + rho_B * vyB;
        qz = qz + rho_R * vzR + rho_B * vzB;
        const double rho = rho_R + rho_B;
        if (rho <= 0.0) continue;
        double ux = (rho_R * vxR + rho_B * vxB) / rho;
        double uy = (rho_R * vyR + rho_B * vyB) / rho;
        double uz = (rho_R * vzR + rho_B * vzB) / rho;
        if (F) {
            const double *Fp = F + (size_t)pos * 3;
            ux = ux - 0.5 * Fp[0] / rho;
            uy = uy - 0.5 * Fp[1] / rho;
            uz = uz - 0.5 * Fp[2] / rho;
            qx = qx - 0.5 * Fp[0];
            qy = qy - 0.5 * Fp[1];
            qz = qz - 0.5 * Fp[2];
        }
        const double u2 = ux * ux + uy * uy + uz * uz;
        u2_sum = u2_sum + u2;
        const double u = sqrt(u2);
        if (u > u_max) u_max = u;
        const double phi = rho_R / rho;
        area = area + phi;
        const int y = pos / nx;
        const int x = pos - y * nx;
        cx_c = cx_c + phi * cos(tx * x);
        cx_s = cx_s + phi * sin(tx * x);
        cy_c = cy_c + phi * cos(ty * y);
        cy_s = cy_s + phi * sin(ty * y);
    }
    double x0 = atan2(cx_s, cx_c) / tx;
    if (x0 < 0.0) x0 += nx;
    double y0 = atan2(cy_s, cy_c) / ty;
    if (y0 < 0.0) y0 += ny;
    const double z0 = 0.0;
    const double raio = (area > 0.0) ? sqrt(area / (PI * (double)nz)) : 0.0;

    double som_w = 0.0, som_wr = 0.0, sx_e = 0.0, sw_e = 0.0, sx_d = 0.0, sw_d = 0.0;
    int y_lin = (int)(y0 + 0.5);
    if (y_lin >= ny) y_lin -= ny;
    for (int pos = 0; pos < nv; pos++) {
        const double rho_R = density(fR + (size_t)pos * ORC_Q);
        const double rho_B = density(fB + (size_t)pos * ORC_Q);
        if (!(rho_R == rho_R) || !(rho_B == rho_B)) continue;
        const int y = pos / nx;
        const int x = pos - y * nx;
        const double dx = image((double)x - x0, nx);
        const double dy = image((double)y - y0, ny);
        const double dz = 0.0;
        (void)z0;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double w = rho_R * rho_B;
        som_w = som_w + w;
        som_wr = som_wr + w * r;
        if (y == y_lin) {
            if (dx < 0.0) { sx_e = sx_e + w * (double)x; sw_e = sw_e + w; }
            else { sx_d = sx_d + w * (double)x; sw_d = sw_d + w; }
        }
    }
    d->x0 = x0; d->y0 = y0; d->area = area; d->raio = raio;
    d->raio_grad = (som_w > 0.0) ? som_wr / som_w : 0.0;
    d->raio_linha = (sw_e > 0.0 && sw_d > 0.0) ? 0.5 * (sx_d / sw_d - sx_e / sw_e) : 0.0;
    d->massa_R = massa_R; d->massa_B = massa_B;
    d->qx = qx; d->qy = qy; d->qz = qz;
    d->u_max = u_max;
    d->u_rms = (n_fluid > 0) ? sqrt(u2_sum / (double)n_fluid) : 0.0;
    d->n_fluid = n_fluid; d->n_ruim = n_ruim;
}

/* calc_laplace_tension, Other_functions.cpp:453-554 (frac_in 0.5, frac_out 2, espessura 3.45) */
static void laplace_tension(const orc_case *cs, const double *fR, const double *fB, const double *F, drop_t *d)
{
    drop_radius(cs, fR, fB, F, d);
    const int nx = cs->nx, ny = cs->ny;
    const int nv = nx * ny;
    const double x0 = d->x0, y0 = d->y0, raio = d->raio;
    const double frac_in = 0.5, frac_out = 2.0, espessura = 3.45;
    const double r_in_max = frac_in * raio;
    double r_out_min = frac_out * raio;
    if (r_out_min < raio + 4.0 * espessura) r_out_min = raio + 4.0 * espessura;
    const double r_caixa = 0.45 * (double)((nx < ny) ? nx : ny);
    if (r_out_min > r_caixa) r_out_min = r_caixa;
    double p_in_s = 0.0, p_out_s = 0.0, p_in_s2 = 0.0, p_out_s2 = 0.0, p_ctr = 0.0, p_cnt = 0.0;
    int n_in = 0, n_out = 0, n_ctr = 0;
    for (int pos = 0; pos < nv; pos++) {
        const double rho_R = density(fR + (size_t)pos * ORC_Q);
        const double rho_B = density(fB + (size_t)pos * ORC_Q);
        if (!(rho_R == rho_R) || !(rho_B == rho_B)) continue;
        const int y = pos / nx;
        const int x = pos - y * nx;
        const double dx = image((double)x - x0, nx);
        const double dy = image((double)y - y0, ny);
        const double dz = 0.0;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double p = CS2 * (rho_R + rho_B);
        if (r < r_in_max) { p_in_s = p_in_s + p; p_in_s2 = p_in_s2 + p * p; n_in = n_in + 1; }
        else if (r >= r_out_min) { p_out_s = p_out_s + p; p_out_s2 = p_out_s2 + p * p; n_out = n_out + 1; }
        if (r < 1.5) { p_ctr = p_ctr + p; n_ctr = n_ctr + 1; }
        if (pos == 0) p_cnt = p_cnt + p;
    }
    const double p_in = (n_in > 0) ? p_in_s / (double)n_in : 0.0;
    const double p_out = (n_out > 0) ? p_out_s / (double)n_out : 0.0;
    const double var_in = (n_in > 0) ? p_in_s2 / (double)n_in - p_in * p_in : 0.0;
    const double var_out = (n_out > 0) ? p_out_s2 / (double)n_out - p_out * p_out : 0.0;
    const double fator_forma = 1.0;
    d->p_in = p_in; d->sd_in = (var_in > 0.0) ? sqrt(var_in) : 0.0;
    d->p_out = p_out; d->sd_out = (var_out > 0.0) ? sqrt(var_out) : 0.0;
    d->n_in = n_in; d->n_out = n_out;
    d->delta_p = p_in - p_out;
    d->sigma = fator_forma * raio * d->delta_p;
    d->p_centro = (n_ctr > 0) ? p_ctr / (double)n_ctr : 0.0;
    d->p_canto = p_cnt;
    d->sigma_linha = fator_forma * d->raio_linha * (d->p_centro - d->p_canto);
}

/* calc_perfil_logistico, Other_functions.cpp:318-448 (plana = false, cilindro = true) */
static void perfil_logistico(const orc_case *cs, const double *fR, const double *fB, double x0, double y0,
                             orc_drop_diag_t *o)
{
    const double eps = 1.0e-30;
    const int nx = cs->nx, ny = cs->ny, nz = 1;
    const int nv = nx * ny;
    const double d_casca = 0.5;
    const int n_casca = (int)(sqrt((double)(nx * nx + ny * ny + nz * nz)) / d_casca) + 2;
    double *soma_phi = (double *)calloc(n_casca, sizeof(double));
    double *soma_n = (double *)calloc(n_casca, sizeof(double));
    for (int pos = 0; pos < nv; pos++) {
        const double rho_R = density(fR + (size_t)pos * ORC_Q);
        const double rho_B = density(fB + (size_t)pos * ORC_Q);
        const double rho = rho_R + rho_B;
        if (rho <= eps) continue;
        const int y = pos / nx;
        const int x = pos - y * nx;
        const double dx = image((double)x - x0, nx);
        const double dy = image((double)y - y0, ny);
        const double dz = 0.0;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const int k = (int)(r / d_casca);
        if (k < 0 || k >= n_casca) continue;
        soma_phi[k] += rho_R / rho;
        soma_n[k] += 1.0;
    }
    double S = 0.0, Sr = 0.0, Sl = 0.0, Srr = 0.0, Srl = 0.0;
    int n_usadas = 0;
    for (int k = 0; k < n_casca; k++) {
        if (soma_n[k] < 1.0) continue;
        const double phi = soma_phi[k] / soma_n[k];
        if (phi <= 0.02 || phi >= 0.98) continue;
        const double r = ((double)k + 0.5) * d_casca;
        const double l = log(phi / (1.0 - phi));
        S += 1.0; Sr += r; Sl += l; Srr += r * r; Srl += r * l;
        n_usadas++;
    }
    double K = 0.0, r_meio = 0.0, res_max = 0.0;
    if (n_usadas >= 3) {
        const double den = S * Srr - Sr * Sr;
        if (fabs(den) > eps) {
            const double a = (S * Srl - Sr * Sl) / den;
            const double b = (Sl - a * Sr) / S;
            K = -a;
            if (fabs(a) > eps) r_meio = -b / a;
            for (int k = 0; k < n_casca; k++) {
                if (soma_n[k] < 1.0) continue;
                const double phi = soma_phi[k] / soma_n[k];
                if (phi <= 0.02 || phi >= 0.98) continue;
                const double r = ((double)k + 0.5) * d_casca;
                const double dd = log(phi / (1.0 - phi)) - (a * r + b);
                if (fabs(dd) > res_max) res_max = fabs(dd);
            }
        }
    }
    free(soma_phi);
    free(soma_n);
    o->K_perfil = K;
    o->r_perfil = r_meio;
    o->res_perfil = res_max;
    o->esp_perfil = (K > eps) ? 2.0 * log(9.0) / K : 0.0;
    o->n_perfil = n_usadas;
}

/* fase_em of modos_interface, Operadores_tensao.cpp:387-409 */
static double fase_em(const orc_case *cs, const double *fR, const double *fB, double x, double y)
{
    const int nx = cs->nx, ny = cs->ny;
    x = x - nx * floor(x / nx);
    y = y - ny * floor(y / ny);
    const int i0 = (int)floor(x), j0 = (int)floor(y);
    const double tx = x - i0, ty = y - j0;
    double v = 0.0;
    for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
            const int ii = (i0 + di) % nx, jj = (j0 + dj) % ny;
            const int p = ii + jj * nx;
            const double rR = density(fR + (size_t)p * ORC_Q);
            const double rB = density(fB + (size_t)p * ORC_Q);
            const double wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty);
            v += wgt * (rR - rB) / (rR + rB);
        }
    return v;
}

/* modos_interface, Operadores_tensao.cpp:382-449 */
static void modos(const orc_case *cs, const double *fR, const double *fB, double x0, double y0, double R_est,
                  orc_drop_diag_t *o)
{
    const int N_ANG = 256;
    double c2 = 0, s2 = 0, c4 = 0, s4 = 0, c8 = 0, s8 = 0, soma = 0;
    double rmin = 1e30, rmax = 0;
    for (int k = 0; k < N_ANG; k++) {
        const double th = 2.0 * PI * k / N_ANG, ct = cos(th), st = sin(th);
        double a = 0.3 * R_est, b = 1.7 * R_est;
        const double fa = fase_em(cs, fR, fB, x0 + a * ct, y0 + a * st);
        for (int it = 0; it < 60; it++) {
            const double mid = 0.5 * (a + b);
            const double fm = fase_em(cs, fR, fB, x0 + mid * ct, y0 + mid * st);
            if ((fm > 0) == (fa > 0)) a = mid;
            else b = mid;
        }
        const double r = 0.5 * (a + b);
        soma += r;
        c2 += r * cos(2 * th); s2 += r * sin(2 * th);
        c4 += r * cos(4 * th); s4 += r * sin(4 * th);
        c8 += r * cos(8 * th); s8 += r * sin(8 * th);
        rmin = dmin(rmin, r);
        rmax = dmax(rmax, r);
    }
    o->R0 = soma / N_ANG;
    o->a2 = 2.0 * sqrt(c2 * c2 + s2 * s2) / N_ANG / o->R0;
    o->a4 = 2.0 * sqrt(c4 * c4 + s4 * s4) / N_ANG / o->R0;
    o->a8 = 2.0 * sqrt(c8 * c8 + s8 * s8) / N_ANG / o->R0;
    o->rmin = rmin;
    o->rmax = rmax;
}

/* calc_galilean, Galilean_diagnostics.cpp:24-105, then (if ok) calc_perfil_logistico and
 * modos_interface as Imm_OPS.cpp:295-300 */
void orc_drop_diag(const orc_case *cs, const double *fR, const double *fB, const double *F,
                   double v0, int step, double x_ini, double y_ini, orc_drop_diag_t *o)
{
    ensure();
    memset(o, 0, sizeof(*o));
    const int nx = cs->nx, ny = cs->ny, n = nx * ny;
    double *ux = (double *)calloc(n, sizeof(double));
    double *uy = (double *)calloc(n, sizeof(double));
    double *uz = (double *)calloc(n, sizeof(double));
    double *ph = (double *)calloc(n, sizeof(double));
    double mass = 0, mass_r = 0, qx = 0, qy = 0, qz = 0, qrx = 0, qry = 0;
    double min_f = INFINITY, min_rho = INFINITY, min_phi = 1, max_phi = 0;
    long long neg_t = 0, neg_c = 0;
    int bad = 0;
    for (int p = 0; p < n; ++p) {
        double rr = 0, rb = 0, mx = 0, my = 0, mz = 0;
        for (int i = 0; i < ORC_Q; ++i) {
            const double fr = fR[(size_t)p * ORC_Q + i], fb = fB[(size_t)p * ORC_Q + i];
            const double f = fr + fb;
            if (!isfinite(fr) || !isfinite(fb)) ++bad;
            min_f = dmin(min_f, f);
            if (f < -1.e-12) ++neg_t;
            if (fr < -1.e-12) ++neg_c;
            if (fb < -1.e-12) ++neg_c;
            rr += fr; rb += fb;
            mx += f * Cd[i * 3]; my += f * Cd[i * 3 + 1]; mz += f * Cd[i * 3 + 2];
        }
        const double rho = rr + rb;
        if (!isfinite(rho) || rho <= 0 || !isfinite(mx) || !isfinite(my) || !isfinite(mz)) { ++bad; continue; }
        if (F) {
            const double *Fp = F + 3 * (size_t)p;
            if (!isfinite(Fp[0]) || !isfinite(Fp[1]) || !isfinite(Fp[2])) { ++bad; continue; }
            mx -= 0.5 * Fp[0]; my -= 0.5 * Fp[1]; mz -= 0.5 * Fp[2];
        }
        min_rho = dmin(min_rho, rho);
        ux[p] = mx / rho; uy[p] = my / rho; uz[p] = mz / rho; ph[p] = rr / rho;
        min_phi = dmin(min_phi, ph[p]);
        max_phi = dmax(max_phi, ph[p]);
        mass += rho; mass_r += rr; qx += mx; qy += my; qz += mz;
        qrx += rr * ux[p]; qry += rr * uy[p];
    }
    o->min_f = min_f; o->min_rho = min_rho; o->min_phi = min_phi; o->max_phi = max_phi;
    o->negative_total = neg_t; o->negative_color = neg_c;
    o->axis_ratio = 1;
    o->n_ruim = bad;
    if (bad || !(mass > 0) || !(mass_r > 0)) {
        o->ok = 0;
        free(ux); free(uy); free(uz); free(ph);
        return;
    }
    o->mean_x = qx / mass; o->mean_y = qy / mass; o->mean_z = qz / mass;
    o->drop_x = qrx / mass_r; o->drop_y = qry / mass_r;

    drop_t d;
    memset(&d, 0, sizeof d);
    laplace_tension(cs, fR, fB, F, &d);

    double sref = 0, smean = 0, slab = 0, sxx = 0, syy = 0, sxy = 0, sx = 0, sy = 0, weight = 0;
    double left_w = 0, right_w = 0, left_x = 0, right_x = 0;
    double max_ref = 0, max_mean = 0, max_lab = 0;
    const int y_line = ((int)floor(d.y0 + 0.5)) % ny;
    for (int pos = 0; pos < nx * ny; ++pos) {
        const int p = pos;
        const double vx = ux[p], vy = uy[p], vz = uz[p];
        const double a = (vx - v0) * (vx - v0) + vy * vy + vz * vz;
        const double b = (vx - o->mean_x) * (vx - o->mean_x) + (vy - o->mean_y) * (vy - o->mean_y) +
                         (vz - o->mean_z) * (vz - o->mean_z);
        const double c = vx * vx + vy * vy + vz * vz;
        sref += a; smean += b; slab += c;
        max_ref = dmax(max_ref, sqrt(a));
        max_mean = dmax(max_mean, sqrt(b));
        max_lab = dmax(max_lab, sqrt(c));
        const int x = pos % nx, y = pos / nx;
        const double dx = gi_image(x - d.x0, nx), dy = gi_image(y - d.y0, ny), phi = ph[p];
        weight += phi; sx += phi * dx; sy += phi * dy;
        sxx += phi * dx * dx; syy += phi * dy * dy; sxy += phi * dx * dy;
        if (y == y_line) {
            const double rr = density(fR + (size_t)p * ORC_Q), rb = density(fB + (size_t)p * ORC_Q), w = rr * rb;
            if (dx < 0) { left_w += w; left_x += w * dx; }
            else { right_w += w; right_x += w * dx; }
        }
    }
    o->max_ref = max_ref; o->max_mean = max_mean; o->max_lab = max_lab;
    o->rms_ref = sqrt(sref / n); o->rms_mean = sqrt(smean / n); o->rms_lab = sqrt(slab / n);
    const double xx = sxx / weight - (sx / weight) * (sx / weight), yy = syy / weight - (sy / weight) * (sy / weight);
    const double xy = sxy / weight - (sx / weight) * (sy / weight), trace = xx + yy;
    const double disc = sqrt((xx - yy) * (xx - yy) + 4 * xy * xy);
    o->axis_ratio = (trace > disc) ? sqrt((trace + disc) / (trace - disc)) : INFINITY;
    o->center_dx = gi_image(d.x0 - (x_ini + v0 * step), nx);
    o->center_dy = gi_image(d.y0 - y_ini, ny);
    o->x_unwrapped = x_ini + v0 * step + o->center_dx;
    d.raio_linha = (left_w > 0 && right_w > 0) ? 0.5 * (right_x / right_w - left_x / left_w) : 0;
    const int corner_x = ((int)floor(d.x0 + 0.5) + nx / 2) % nx;
    const int corner_y = ((int)floor(d.y0 + 0.5) + ny / 2) % ny;
    const int pc = corner_x + nx * corner_y;
    d.p_canto = density(fR + (size_t)pc * ORC_Q) * CS2 + density(fB + (size_t)pc * ORC_Q) * CS2;
    d.sigma_linha = d.raio_linha * (d.p_centro - d.p_canto);
    d.u_max = o->max_ref;
    d.u_rms = o->rms_ref;
    o->ok = isfinite(d.sigma) && isfinite(o->max_ref) && isfinite(o->axis_ratio);

    o->raio = d.raio; o->raio_grad = d.raio_grad; o->raio_linha = d.raio_linha;
    o->p_in = d.p_in; o->p_out = d.p_out; o->delta_p = d.delta_p; o->sigma = d.sigma;
    o->u_max = d.u_max; o->u_rms = d.u_rms; o->massa_R = d.massa_R; o->massa_B = d.massa_B;
    o->qx = d.qx; o->qy = d.qy; o->qz = d.qz; o->sd_in = d.sd_in; o->sd_out = d.sd_out;
    o->sigma_linha = d.sigma_linha; o->n_in = d.n_in; o->n_out = d.n_out;
    o->x0 = d.x0; o->y0 = d.y0; o->area = d.area; o->p_centro = d.p_centro; o->p_canto = d.p_canto;
    if (o->ok) {
        perfil_logistico(cs, fR, fB, d.x0, d.y0, o);
        modos(cs, fR, fB, d.x0, d.y0, d.raio, o);
    }
    free(ux); free(uy); free(uz); free(ph);
}

static long long gcdll(long long a, long long b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { long long t = a % b; a = b; b = t; }
    return a;
}

/* dist_central, Imm_INCL.cpp:74-83, with the band period M in place of N */
static double dist_central(int x, int y, int p, int q, long long M)
{
    const double m = sqrt((double)(p * p + q * q));
    long long s = ((long long)p * x + (long long)q * y) % M;
    if (s < 0) s += M;
    double d = ((double)s - 0.5 * M);
    if (d >= 0.5 * M) d -= M;
    if (d < -0.5 * M) d += M;
    return -d / m;
}

/* mede_inclinada, Imm_INCL.cpp:85-189.  Box nx x ny; period M = gcd(p nx, q ny) / gcd(p, q)
 * (= N for the reference's square boxes); interface length 2 (nx ny / M) m. */
void orc_incl_diag(const orc_case *cs, int p, int q, const double *fR, const double *fB,
                   const double *F, const double *gm, orc_incl_diag_t *r)
{
    ensure();
    memset(r, 0, sizeof(*r));
    const int nx = cs->nx, ny = cs->ny;
    const long long M = gcdll((long long)p * nx, (long long)q * ny) / gcdll(p, q);
    const double m = sqrt((double)(p * p + q * q));
    const double nx_ = p / m, ny_ = q / m;
    const double tx_ = -ny_, ty_ = nx_;
    r->min_f = 1e300;
    double u2 = 0, Sxx = 0, Sx = 0, Sxy = 0, Sy = 0;
    int nfit = 0;
    double *dd = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    double *ll = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int pto = x + y * nx;
            const double *fr = fR + (size_t)pto * ORC_Q, *fb = fB + (size_t)pto * ORC_Q;
            double rR = 0, rB = 0, jx = 0, jy = 0, Pxx = 0, Pyy = 0, Pxy = 0, Pzz = 0;
            for (int i = 0; i < ORC_Q; i++) {
                const double *c = Cd + 3 * i;
                const double f = fr[i] + fb[i];
                if (!isfinite(fr[i]) || !isfinite(fb[i])) r->n_ruim++;
                if (f < -1e-12) r->neg_total++;
                if (fr[i] < -1e-12) r->neg_cor++;
                if (fb[i] < -1e-12) r->neg_cor++;
                r->min_f = dmin(r->min_f, f);
                rR += fr[i]; rB += fb[i];
                jx += f * c[0]; jy += f * c[1];
                Pxx += f * c[0] * c[0]; Pyy += f * c[1] * c[1]; Pxy += f * c[0] * c[1]; Pzz += f * c[2] * c[2];
            }
            const double rho = rR + rB;
            if (!(rho > 0)) { r->n_ruim++; continue; }
            if (F) {
                jx -= 0.5 * F[3 * pto]; jy -= 0.5 * F[3 * pto + 1];
                r->SFx += F[3 * pto]; r->SFy += F[3 * pto + 1];
            }
            const double ux = jx / rho, uy = jy / rho;
            r->massa_R += rR; r->massa_B += rB; r->qx += jx; r->qy += jy;
            const double u = hypot(ux, uy), un = fabs(ux * nx_ + uy * ny_), ut = fabs(ux * tx_ + uy * ty_);
            r->u_max = dmax(r->u_max, u); r->un_max = dmax(r->un_max, un); r->ut_max = dmax(r->ut_max, ut);
            u2 += u * u;
            const double Pnn = nx_ * nx_ * Pxx + 2 * nx_ * ny_ * Pxy + ny_ * ny_ * Pyy;
            const double Ptt = tx_ * tx_ * Pxx + 2 * tx_ * ty_ * Pxy + ty_ * ty_ * Pyy;
            r->sigma_pop += Pnn - 0.5 * (Ptt + Pzz);
            const double *g = gm + 3 * (size_t)pto;
            r->int_grad += sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            const double d = dist_central(x, y, p, q, M);
            const double cR = rR / rho;
            if (cR > 0.02 && cR < 0.98 && fabs(d) < 8.0) {
                const double l = log(cR / (1.0 - cR));
                dd[nfit] = d; ll[nfit] = l;
                Sx += d; Sy += l; Sxx += d * d; Sxy += d * l; nfit++;
            }
        }
    const double comprimento = 2.0 * ((double)nx * (double)ny / (double)M) * m;
    r->sigma_pop /= comprimento;
    r->int_grad /= comprimento;
    r->u_rms = sqrt(u2 / ((double)nx * (double)ny));
    r->n_ajuste = nfit;
    if (nfit > 3) {
        const double K = (nfit * Sxy - Sx * Sy) / (nfit * Sxx - Sx * Sx);
        const double a = (Sy - K * Sx) / nfit;
        r->K = K;
        r->d0 = -a / K;
        for (int k = 0; k < nfit; k++) r->res = dmax(r->res, fabs(ll[k] - (a + K * dd[k])));
    }
    r->ok = (r->n_ruim == 0) && isfinite(r->u_max);
    free(dd); free(ll);
}
