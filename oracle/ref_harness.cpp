// Reference harness: drives the reference's own library functions (compiled from the
// sources under the reference tree, never copied) through the same step loop as
// Imm_OPS.cpp:246-372 / Imm_INCL.cpp:358-442, with run-time parameters instead of the
// drivers' compile-time constants, and dumps full-precision results.
//
// This file is test infrastructure for the oracle: it is built only where the reference
// tree exists (oracle/Makefile), its outputs go to oracle/_ref/, and nothing in the
// product links or runs it.
//
// Usage:  ref_harness key=value ...
//   mode=drop|incl  nx ny  R (drop; default nx/4)  width (initial 10-90 width; <=0 = sharp)
//   op="P3 0.5"  tau  A  beta  v0 (drop)  p q (incl)  steps  diag (interval)
//   dump=s1,s2,...  (post-collision populations at these steps -> out/pop_<step>.bin)
//   geo=<meio.vtk> wett=<wett_R>  (drop mode: solid sites from the file, read by read_geo; the
//   populations are then dumped for the fluid sites only, in site order; diag=0 skips diagnostics)
//   out=<dir>
//   mode=readgeo geo=<file> pts_in pts_out: only read_geo's numbering of the file's sites -> out/meio.bin
// Writes out/diag.txt: one line per diagnostic step, %.17g, fields named in the header.

#define nvel 19
#define dim 3

#ifdef HARNESS_INCL
#define main reference_incl_main
#include "Imm_INCL.cpp"
#undef main
#include "Galilean_diagnostics.cpp"
#else
#include "Definitions.cpp"
#include "LBM_functions.cpp"
#include "Collision_operators.cpp"
#include "Boundary_conditions.cpp"
#include "Initial_conditions.cpp"
#include "Other_functions.cpp"
#include "Galilean_diagnostics.cpp"
#include "Operadores_tensao.cpp"
#endif

#include <map>
#include <set>
#include <cstdio>

static std::map<std::string, std::string> args;

static double argd(const char *k, double def) {
    auto it = args.find(k);
    return it == args.end() ? def : atof(it->second.c_str());
}
static int argi(const char *k, int def) {
    auto it = args.find(k);
    return it == args.end() ? def : atoi(it->second.c_str());
}
static std::string args_s(const char *k, const char *def) {
    auto it = args.find(k);
    return it == args.end() ? std::string(def) : it->second;
}

static void dump_bin(const std::string &path, const double *a, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    fwrite(a, sizeof(double), n, f);
    fclose(f);
}

int main(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
        std::string s(argv[i]);
        size_t e = s.find('=');
        if (e == std::string::npos) { fprintf(stderr, "bad arg %s\n", argv[i]); return 2; }
        args[s.substr(0, e)] = s.substr(e + 1);
    }
    omp_set_num_threads(argi("threads", 1));
    const std::string mode = args_s("mode", "drop");
    if (mode == "readgeo") {
        // the reference's read_geo on geo=<file> with pts_in / pts_out added columns: the numbering it
        // gives every site (0 solid, 1.. fluid) -> out/meio.bin (int32), the fluid count -> out/n.txt
        const std::string geo = args_s("geo", "");
        int gx = 0, gy = 0, gz = 0;
        {
            std::ifstream f(geo);
            std::string line, dump;
            for (int i = 0; i < 4; i++) std::getline(f, dump);
            std::getline(f, line);
            std::stringstream d(line);
            d >> dump >> gx >> gy >> gz;
        }
        const int pin = argi("pts_in", 0), pout = argi("pts_out", 0);
        const size_t total = (size_t)(gx + pin + pout) * gy * gz;
        int *meio = new int[total];
        const int n = read_geo(geo, meio, pin, pout);
        const std::string out = args_s("out", ".");
        FILE *f = fopen((out + "/meio.bin").c_str(), "wb");
        fwrite(meio, sizeof(int), total, f);
        fclose(f);
        f = fopen((out + "/n.txt").c_str(), "w");
        fprintf(f, "%d %d %d %d\n", n, gx + pin + pout, gy, gz);
        fclose(f);
        return 0;
    }
    const int nx = argi("nx", 100), ny = argi("ny", nx);
    const double tau = argd("tau", 1.0), A_in = argd("A", 0.225), beta = argd("beta", 0.8);
    const double width = argd("width", 3.45);
    const int n_steps = argi("steps", 10), passos_diag = argi("diag", 500);
    const std::string out = args_s("out", ".");
    std::set<int> dumps;
    {
        std::string d = args_s("dump", "");
        size_t p = 0;
        while (p < d.size()) {
            size_t c = d.find(',', p);
            if (c == std::string::npos) c = d.size();
            if (c > p) dumps.insert(atoi(d.substr(p, c - p).c_str()));
            p = c + 1;
        }
    }
    {
        std::ofstream of(out + "/operador.txt");
        of << args_s("op", "P3 0.5") << "\n";
    }
    OPERADOR op;
    if (!le_operador(out + "/operador.txt", op)) { fprintf(stderr, "bad op\n"); return 2; }

    GEOMETRY geometry;
    geometry.nx = nx; geometry.ny = ny; geometry.nz = 1;
    const std::string geo = args_s("geo", "");
    if (geo.empty()) {
        geometry.fluid = nx * ny;
        geometry.ini = new int[nx * ny];
        for (int p = 0; p < nx * ny; p++) geometry.ini[p] = p + 1;
    } else {
        // a geometry file with solid sites, read by the reference's own read_geo (its box must be
        // nx x ny x 1, the size given on the command line)
        geometry.ini = new int[nx * ny];
        geometry.fluid = read_geo(geo, geometry.ini, 0, 0);
    }
    const int n_fluid = geometry.fluid;

    PARAMETERS parameters;
    parameters.tau = tau; parameters.tau_R = tau; parameters.tau_B = tau; parameters.tau_m = 1.0;
    parameters.rho_ini = 1.0; parameters.rho_ini_R = 1.0; parameters.rho_ini_B = 1.0;
    parameters.A_fact = A_in; parameters.A_fact_R = 0.0; parameters.A_fact_B = 0.0;
    parameters.recoll = beta; parameters.wett_R = argd("wett", 0.0); parameters.tau_nh = 1.0;
    parameters.n_steps = n_steps; parameters.n_files = 0; parameters.n_threads = 1;

    LATTICE lattice;
    const int size_to_alloc = n_fluid * nvel;
    lattice.inif_R = new double[size_to_alloc];
    lattice.inif_R_new = new double[size_to_alloc];
    lattice.inif_B = new double[size_to_alloc];
    lattice.inif_B_new = new double[size_to_alloc];
    lattice.inif_m = new double[size_to_alloc];
    lattice.inif_m_new = new double[size_to_alloc];
    lattice.ini_stream = new int[size_to_alloc];
    lattice.ini_solid = new bool[size_to_alloc];
    double *forca = nullptr;
    if (op.forca()) {
        forca = new double[3 * n_fluid]();
        lattice.ini_force = forca;
    }
    lattice.c_i = new double[nvel * dim];
    def_lattice_d3q19(lattice);
    lattice.Q_i = new double[nvel * dim * dim];
    calc_Q(lattice);
    def_dir_prop(geometry, lattice);
    CAMPOS_F campos;
    if (op.forca()) aloca_campos_F(campos, geometry, lattice);

    const bool incl = (mode == "incl");
    const int pp = argi("p", 2), qq = argi("q", 1);
    const double raio_ini = argd("R", 0.25 * (double)nx);
    const double v0 = argd("v0", 0.0);
    const int x0_ini = nx / 2, y0_ini = ny / 2;

    if (!incl) {
        const double k_perfil = (width > 0) ? 2.0 * log(9.0) / width : -1.0;
        initial_conditions_drop(raio_ini, k_perfil, 0.0, x0_ini, y0_ini, 0, v0, 0., 0., true,
                                geometry, lattice, parameters, true, EQ_CLASSICO);
    } else {
#ifdef HARNESS_INCL
        // Same construction as Imm_INCL.cpp:309-334 (square box N only).
        const int N = nx;
        const double m = sqrt((double)(pp * pp + qq * qq));
        const double k = (width > 0) ? 2.0 * log(9.0) / width : -1.0;
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                long long s = ((long long)pp * x + (long long)qq * y) % N;
                if (s < 0) s += N;
                const double dred = (2 * s < N) ? min((double)s, 0.5 * N - s) / m
                                                : -min(s - 0.5 * N, (double)(N - s)) / m;
                const double cR = (k > 0) ? 1.0 / (1.0 + exp(-k * dred)) : (dred > 0 ? 1.0 : (dred < 0 ? 0.0 : 0.5));
                const int pto = geometry.ini[x + y * N] - 1;
                for (int i = 0; i < nvel; i++) {
                    lattice.inif_R[(size_t)pto * nvel + i] = lattice.w[i] * cR;
                    lattice.inif_B[(size_t)pto * nvel + i] = lattice.w[i] * (1.0 - cR);
                    lattice.inif_m[(size_t)pto * nvel + i] = 0.0;
                }
            }
#else
        fprintf(stderr, "incl mode needs the incl harness\n");
        return 2;
#endif
    }

    FILE *fd = fopen((out + "/diag.txt").c_str(), "w");
    if (!incl) {
        fprintf(fd, "# step raio raio_grad raio_linha p_in p_out delta_p sigma u_max u_rms massa_R massa_B "
                    "qx qy qz sd_in sd_out sigma_linha K_perfil r_perfil esp_perfil res_perfil n_perfil "
                    "n_in n_out x0 y0 area p_centro p_canto "
                    "gi_mean_x gi_mean_y gi_mean_z gi_drop_x gi_drop_y gi_max_ref gi_rms_ref gi_max_mean "
                    "gi_rms_mean gi_max_lab gi_rms_lab gi_min_f gi_min_rho gi_min_phi gi_max_phi gi_axis_ratio "
                    "gi_center_dx gi_center_dy gi_x_unwrapped gi_negative_total gi_negative_color "
                    "R0 a2 a4 a8 rmin rmax ok\n");
    } else {
        fprintf(fd, "# step sigma_pop int_grad K res d0 u_max u_rms un_max ut_max massa_R massa_B qx qy "
                    "SFx SFy min_f neg_total neg_cor n_ruim n_ajuste ok\n");
    }

    DROP drop;
    drop.cilindro = true;
    drop.frac_in = 0.5;
    drop.frac_out = 2.0;
    drop.espessura = 3.45;
    GI_DATA gi;

    for (int step = 0; step < n_steps; step++) {
#pragma omp parallel for
        for (int pto = 0; pto < n_fluid; pto++) {
            const double rho_R = density(lattice.inif_R + pto * nvel);
            const double rho_B = density(lattice.inif_B + pto * nvel);
            emite_mediadores(lattice.inif_m + pto * nvel, (rho_R - rho_B) / (rho_R + rho_B), lattice);
            propag_site_med(lattice, parameters, pto);
        }
        { double *t = lattice.inif_m; lattice.inif_m = lattice.inif_m_new; lattice.inif_m_new = t; }
        if (op.forca()) calcula_forca(campos, op, A_in, lattice, forca);
#pragma omp parallel for
        for (int pto = 0; pto < n_fluid; pto++) {
            static const double zero[3] = {0.0, 0.0, 0.0};
            const double *F = op.forca() ? forca + 3 * pto : zero;
            colisao_comum(lattice.inif_R + pto * nvel, lattice.inif_B + pto * nvel,
                          lattice.inif_m + pto * nvel, F, op, lattice, parameters);
            double mx = 0., my = 0., mz = 0.;
            propag_site(lattice, pto, mx, my, mz);
        }
        if (dumps.count(step)) {
            char nm[64];
            snprintf(nm, sizeof nm, "/popR_%d.bin", step);
            dump_bin(out + nm, lattice.inif_R, size_to_alloc);
            snprintf(nm, sizeof nm, "/popB_%d.bin", step);
            dump_bin(out + nm, lattice.inif_B, size_to_alloc);
            if (forca) {
                snprintf(nm, sizeof nm, "/force_%d.bin", step);
                dump_bin(out + nm, forca, 3 * (size_t)n_fluid);
            }
        }
        const bool medir = passos_diag > 0 && ((step % passos_diag == 0) || (step == n_steps - 1));
        if (medir) {
            if (!incl) {
                const bool ok = calc_galilean(geometry, lattice, drop, gi, v0, step, x0_ini, y0_ini);
                MODOS md;
                if (ok) {
                    calc_perfil_logistico(geometry, lattice, drop);
                    md = modos_interface(geometry, lattice, drop.x0, drop.y0, drop.raio);
                }
                fprintf(fd, "%d", step);
                const double v1[] = {drop.raio, drop.raio_grad, drop.raio_linha, drop.p_in, drop.p_out,
                                     drop.delta_p, drop.sigma, drop.u_max, drop.u_rms, drop.massa_R,
                                     drop.massa_B, drop.qx, drop.qy, drop.qz, drop.sd_in, drop.sd_out,
                                     drop.sigma_linha, drop.K_perfil, drop.r_perfil, drop.esp_perfil,
                                     drop.res_perfil, (double)drop.n_perfil, (double)drop.n_in,
                                     (double)drop.n_out, drop.x0, drop.y0, drop.area, drop.p_centro,
                                     drop.p_canto, gi.mean_x, gi.mean_y, gi.mean_z, gi.drop_x, gi.drop_y,
                                     gi.max_ref, gi.rms_ref, gi.max_mean, gi.rms_mean, gi.max_lab,
                                     gi.rms_lab, gi.min_f, gi.min_rho, gi.min_phi, gi.max_phi,
                                     gi.axis_ratio, gi.center_dx, gi.center_dy, gi.x_unwrapped,
                                     (double)gi.negative_total, (double)gi.negative_color, md.R0, md.a2,
                                     md.a4, md.a8, md.rmin, md.rmax, ok ? 1.0 : 0.0};
                for (double v : v1) fprintf(fd, " %.17g", v);
                fprintf(fd, "\n");
            }
#ifdef HARNESS_INCL
            else {
                INCL r;
                const bool ok = mede_inclinada(geometry, lattice, pp, qq, r, forca, false);
                fprintf(fd, "%d", step);
                const double v2[] = {r.sigma_pop, r.int_grad, r.K, r.res, r.d0, r.u_max, r.u_rms,
                                     r.un_max, r.ut_max, r.massa_R, r.massa_B, r.qx, r.qy, r.SFx, r.SFy,
                                     r.min_f, (double)r.neg_total, (double)r.neg_cor, (double)r.n_ruim,
                                     (double)r.n_ajuste, ok ? 1.0 : 0.0};
                for (double v : v2) fprintf(fd, " %.17g", v);
                fprintf(fd, "\n");
            }
#endif
            fflush(fd);
        }
        { double *t = lattice.inif_R; lattice.inif_R = lattice.inif_R_new; lattice.inif_R_new = t;
          t = lattice.inif_B; lattice.inif_B = lattice.inif_B_new; lattice.inif_B_new = t; }
    }
    fclose(fd);
    return 0;
}
