/*
 * CPU oracle for the colour-gradient LBM step (D3Q19 slab, nz = 1).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference algorithm used
 * as the checker of the GPU library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never does.
 *
 * Every double expression keeps the reference's evaluation order and the library is
 * compiled with -ffp-contract=off, so its populations are bit-comparable both with the
 * reference built the same way (oracle/_ref/harness_*_nofma) and with the GPU FP64 kernel.
 *
 * Layout is the reference's: populations are [site][19] with site = x + nx*y
 * (Imm_OPS.cpp:158-181, read_geo all-fluid numbering LBM_functions.cpp:280-354).
 *
 * Parity pins (what checks each function, see tests/test_oracle_*.py):
 *   orc_step / orc_collide_site / orc_force_field : bitwise vs reference nofma build,
 *       E0 closed forms (E0_resultado.txt), Lotes/ fixtures.
 *   orc_step_geo : bitwise vs the reference build on boxes with solid sites read by the reference's
 *       own read_geo (harness geo=), all operators.
 *   orc_init_drop / orc_init_band : bitwise vs reference build.
 *   orc_drop_diag / orc_incl_diag : reference build (%.17g) and Lotes/ galilean.dat.
 */
#ifndef ITOLBM_ORACLE_H
#define ITOLBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_Q 19

/* operator ids, Operadores_tensao.cpp:45-49 */
#define ORC_OP_P1 1
#define ORC_OP_P3 3
#define ORC_OP_F1 11
#define ORC_OP_F2 12
#define ORC_OP_F3 13

typedef struct {
    int nx, ny;
    int op;        /* ORC_OP_* */
    double chi;    /* P3 */
    double xi;     /* F3 */
    double tau;    /* tau_R = tau_B */
    double A;      /* fat_R_B: amplitude A (P class) or sigma (F class) */
    double beta;   /* recolouring */
} orc_case;

/* lattice constants, LBM_functions.cpp:366-497 */
void orc_lattice(double c[ORC_Q * 3], double w[ORC_Q]);

/* per-site kernels (E0 probes) */
void orc_perturbation(double *f, double A, const double g[3], int op, double chi);
void orc_source(double Fx, double Fy, double Fz, double vx, double vy, double vz, double rho,
                double tau, double S[ORC_Q]);
void orc_collide_site(double *fR, double *fB, const double *fm, const double F[3], const orc_case *cs);

/* whole-field force (class F), Operadores_tensao.cpp:202-332.  fR/fB [site][19]; F [site][3]. */
void orc_force_field(const orc_case *cs, const double *fR, const double *fB, double *F);

/*
 * One time step (Imm_OPS.cpp:246-372):  mediators, force (F class), collide in place
 * (fR/fB become post-collision, pre-stream), push-stream into fR_new/fB_new.
 * F (n*3, class F only) receives the force field; gm (n*3, optional) the mediator momentum.
 */
void orc_step(const orc_case *cs, double *fR, double *fB, double *fR_new, double *fB_new,
              double *F, double *gm);

/*
 * The same step on a box with solid sites (solid: n bytes, 1 = solid, x + nx*y): bounce-back of the
 * populations pushed towards a solid site (def_dir_prop LBM_functions.cpp:509-853, propag_site
 * :1183-1253), wall mediators of phase wett (propag_site_med :1309-1337), solid neighbours of the force
 * stencils reflected onto the site (aloca_campos_F Operadores_tensao.cpp:149-182).  Solid sites are
 * not collided; their populations in fR_new / fB_new are set to 0.
 */
void orc_step_geo(const orc_case *cs, const unsigned char *solid, double wett, double *fR, double *fB,
                  double *fR_new, double *fB_new, double *F, double *gm);

/* initial conditions */
void orc_init_drop(const orc_case *cs, double R, double k_perfil, int x0, int y0, double v0,
                   double *fR, double *fB);
void orc_init_band(const orc_case *cs, int p, int q, double width, double *fR, double *fB);

/* diagnostics ---------------------------------------------------------------------- */

typedef struct {
    /* DROP (Definitions.cpp:251-339) */
    double raio, raio_grad, raio_linha, p_in, p_out, delta_p, sigma, u_max, u_rms;
    double massa_R, massa_B, qx, qy, qz, sd_in, sd_out, sigma_linha;
    double K_perfil, r_perfil, esp_perfil, res_perfil;
    int n_perfil, n_in, n_out;
    double x0, y0, area, p_centro, p_canto;
    /* GI_DATA (Galilean_diagnostics.cpp:7-16) */
    double mean_x, mean_y, mean_z, drop_x, drop_y, max_ref, rms_ref, max_mean, rms_mean;
    double max_lab, rms_lab, min_f, min_rho, min_phi, max_phi, axis_ratio;
    double center_dx, center_dy, x_unwrapped;
    long long negative_total, negative_color;
    /* MODOS (Operadores_tensao.cpp:380) */
    double R0, a2, a4, a8, rmin, rmax;
    int ok;
    int n_ruim;
} orc_drop_diag_t;

/* calc_galilean + calc_perfil_logistico + modos_interface on post-collision populations. */
void orc_drop_diag(const orc_case *cs, const double *fR, const double *fB, const double *F,
                   double v0, int step, double x_ini, double y_ini, orc_drop_diag_t *out);

typedef struct {
    double sigma_pop, int_grad, K, res, d0, u_max, u_rms, un_max, ut_max;
    double massa_R, massa_B, qx, qy, SFx, SFy, min_f;
    long long neg_total, neg_cor;
    int n_ruim, n_ajuste, ok;
} orc_incl_diag_t;

/* mede_inclinada (Imm_INCL.cpp:85-189), generalised to nx x ny with period M = gcd(p nx, q ny). */
void orc_incl_diag(const orc_case *cs, int p, int q, const double *fR, const double *fB,
                   const double *F, const double *gm, orc_incl_diag_t *out);

/* E0 checks (verifica_operadores.cpp); returns the number of failures.  err/tol hold 37
 * values, m4 holds 20 (see oracle_e0.c). */
int orc_e0(double *err, double *tol, double *m4);

#ifdef __cplusplus
}
#endif
#endif
