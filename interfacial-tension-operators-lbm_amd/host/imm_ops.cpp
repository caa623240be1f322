// Imm_OPS: cylindrical drop (static, or translating with v0) in a periodic nx x ny x 1 box, with the
// interfacial-tension operator of operador.txt, advanced on the GPU through the itolbm C ABI.
//
// Drop-in for the reference program of the same name (Imm_OPS.cpp): run with no argument in a case
// directory, it reads data_in.txt, operador.txt, velocity.txt and meio.vtk there and writes
// dat_out.txt, laplace.dat, galilean.dat, iso.dat, resumo.dat and the final rho_R / rho_B / vel VTK
// fields, with the reference's formats, diagnostic schedule (every 500 steps and the last step,
// Imm_OPS.cpp:293-351), stability rule and exit codes.  Given several case directories as arguments it
// runs them together in one GPU context (batched launches), each writing its own files and console
// log (execucao.log); cases of different storage precision run in separate contexts, one after the
// other.  Optional <case>/itolbm.txt: see case_files.h.
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "batch.h"
#include "case_files.h"

using namespace itolbm_host;

namespace {

constexpr double U_LIMITE = 0.4;                           // above this: unstable (Imm_OPS.cpp:43)
constexpr double CS2 = 1.0 / 3.0;

struct DropCase : Member {
    std::string dir;
    std::ostream* con = nullptr;   // console of the case
    std::unique_ptr<std::ofstream> log;
    DataIn din;
    Operator op;
    Options opt;
    double v0 = 0.0;
    int nx = 0, ny = 0, x0 = 0, y0 = 0;
    long long n_fluid = 0;
    double raio_ini = 0, tau = 0, A_in = 0, sigma_nominal = 0, nu = 0;
    std::ofstream f_lap, f_iso, f_gi;
    double s_sigma = 0, s_umax = 0, s_urms = 0, s_a4 = 0, s_a8 = 0, s_K = 0;
    int n_med = 0;
    long long neg_total_max = 0, neg_cor_max = 0;
    double min_f_global = 1e30;
    std::string estado = "estavel";
    int passo_falha = -1;

    const std::string& name() const override { return dir; }

    // Reads and checks the inputs; prints the reference's messages on failure.
    bool load()
    {
        read_data_in(dir, din, *con);
        std::string err;
        if (!read_options(dir, opt, err)) {
            *con << err << "\n";
            return false;
        }
        if (!read_operator(join(dir, "operador.txt"), op)) {
            *con << "\noperador.txt ausente ou invalido.  Uma linha: P1 | P3 chi | F1 | F2 | F3 xi\n";
            return false;
        }
        if (!read_velocity(join(dir, "velocity.txt"), v0)) {
            *con << "velocity.txt invalido.\n";
            return false;
        }
        if (std::fabs(din.tau_R - din.tau_B) > 1e-14) {
            *con << "\nEste programa usa BGK de tempo UNICO: tau_R e tau_B tem de ser iguais.\n";
            return false;
        }
        int nz = 0;
        if (!read_box(join(dir, din.geo_file), nx, ny, nz, n_fluid)) {
            *con << "\nNao foi possivel abrir " << din.geo_file << "\n";
            return false;
        }
        if (nz != 1 || n_fluid != (long long)nx * ny) {
            *con << "\nEste programa requer caixa 2D periodica toda fluida ( nz = 1 ).\n";
            return false;
        }
        raio_ini = (opt.radius > 0.0) ? opt.radius : 0.25 * (double)nx;
        x0 = nx / 2;
        y0 = ny / 2;
        tau = din.tau_R;
        A_in = din.A_fact;
        sigma_nominal = op.force() ? A_in : 4.0 / 9.0 * A_in * tau;
        nu = CS2 * (tau - 0.5);
        std::ostream& o = *con;
        o << "\nOperadores de tensao interfacial -- GPU ( D3Q19, BGK, Latva-Kokko )\n";
        o << "   operador          = " << op.name;
        if (op.op == LBM_OP_P3) o << "   chi = " << op.chi;
        if (op.op == LBM_OP_F3) o << "   xi = " << op.xi;
        o << "\n   caixa             = " << nx << " x " << ny << " x " << nz << "\n";
        o << "   raio inicial      = " << raio_ini << "\n";
        o << "   tau               = " << tau << "   ( nu = " << (tau - 0.5) / 3.0 << " )\n";
        o << "   beta              = " << din.recoll << "\n";
        o << "   fat_R_B           = " << A_in << (op.force() ? "   ( = sigma )" : "   ( = A )") << "\n";
        o << "   sigma nominal     = " << sigma_nominal << "\n";
        o << "   v0                = " << v0 << "\n";
        o << "   precisao          = " << precision_name(opt.precision)
          << std::endl;
        f_lap.open(join(dir, "laplace.dat"));
        f_lap << std::setprecision(opt.digits);
        laplace_header(f_lap);
        f_iso.open(join(dir, "iso.dat"));
        f_iso << std::setprecision(opt.digits);
        iso_header(f_iso);
        f_gi.open(join(dir, "galilean.dat"));
        galilean_header(f_gi);
        return true;
    }

    lbm_case_desc desc() const override
    {
        lbm_case_desc d{};
        d.nx = nx;
        d.ny = ny;
        d.op = op.op;
        d.chi = op.chi;
        d.xi = op.xi;
        d.tau = tau;
        d.A = A_in;
        d.beta = din.recoll;
        return d;
    }

    int steps() const override { return din.n_steps; }

    bool init(lbm_ctx* ctx, int id) override
    {
        return lbm_init_drop(ctx, id, raio_ini, opt.width, x0, y0, v0) == LBM_OK;
    }

    bool measure(int t) const { return t % opt.diag_every == 0 || t == din.n_steps - 1; }
    bool fields(int t) const
    {
        return t == din.n_steps - 1 ||
               (din.n_files > 0 && t % std::max(1, din.n_steps / din.n_files) == 0 && t > 0);
    }

    int next_event(int s) const override
    {
        int t = s;
        while (t < din.n_steps - 1 && !measure(t) && !fields(t)) ++t;
        return t;
    }

    bool on_event(lbm_ctx* ctx, int id, int t) override
    {
        if (measure(t)) {
            lbm_drop_diag d{};
            if (lbm_diag_drop(ctx, id, v0, t, x0, y0, &d) != LBM_OK) {
                *con << "\ndiagnostico falhou: " << lbm_last_error(ctx) << "\n";
                d.ok = 0;
            }
            const bool ok = d.ok != 0;
            if (ok) {
                laplace_row(f_lap, t, d, nu);
                galilean_row(f_gi, t, v0, CS2, d);
                iso_row(f_iso, t, d);
                neg_total_max = std::max(neg_total_max, d.negative_total);
                neg_cor_max = std::max(neg_cor_max, d.negative_color);
                min_f_global = std::min(min_f_global, d.min_f);
                if (t >= din.n_steps / 2) {
                    s_sigma += d.sigma; s_umax += d.u_max; s_urms += d.u_rms;
                    s_a4 += d.a4; s_a8 += d.a8; s_K += d.K_perfil;
                    n_med++;
                }
                *con << "step = " << t << "   R = " << d.raio << "   sigma = " << d.sigma << "   |u|max = " << d.u_max
                     << "   a4 = " << d.a4 << "   neg = " << d.negative_total << std::endl;
                // keep the reduced-precision storage centred on the measured phase densities
                if (opt.precision != LBM_PREC_F64 && opt.recentre && d.p_in > 0.0 && d.p_out > 0.0 &&
                    std::isfinite(d.p_in) && std::isfinite(d.p_out))
                    lbm_set_reference(ctx, id, 3.0 * d.p_in, 3.0 * d.p_out);
            }
            if (!ok || d.n_ruim > 0 || d.u_max > U_LIMITE) {
                estado = "instavel";
                passo_falha = t;
                *con << "\nSimulacao instavel no passo " << t << std::endl;
                return false;
            }
        }
        if (fields(t)) write_fields(ctx, id, t);
        return true;
    }

    void write_fields(lbm_ctx* ctx, int id, int t)
    {
        const size_t n = (size_t)nx * ny;
        std::vector<double> fR(n * 19), fB(n * 19), F;
        if (lbm_download_state(ctx, id, LBM_STATE_POST_COLLISION, fR.data(), fB.data()) != LBM_OK) {
            *con << "\ncampos: " << lbm_last_error(ctx) << "\n";
            return;
        }
        if (op.force()) {
            F.resize(3 * n);
            if (lbm_download_aux(ctx, id, nullptr, F.data()) != LBM_OK) F.clear();
        }
        const std::string ts = std::to_string((unsigned)t);
        const int fl = opt.vtk_flags;
        const bool ok = lbm_write_vtk(join(dir, "rho_R" + ts + ".vtk").c_str(), LBM_VTK_DENSITY, nx, ny, fR.data(), nullptr,
                                      nullptr, nullptr, fl) == LBM_OK &&
                        lbm_write_vtk(join(dir, "rho_B" + ts + ".vtk").c_str(), LBM_VTK_DENSITY, nx, ny, fB.data(), nullptr,
                                      nullptr, nullptr, fl) == LBM_OK &&
                        lbm_write_vtk(join(dir, "vel_" + ts + ".vtk").c_str(), LBM_VTK_VELOCITY, nx, ny, fR.data(), fB.data(),
                                      F.empty() ? nullptr : F.data(), nullptr, fl) == LBM_OK;
        if (!ok) *con << "\ncampos: nao foi possivel gravar os arquivos vtk\n";
    }

    void finish(int step, double elapsed) override
    {
        const double mlups = (elapsed > 0) ? (double)n_fluid * step / (elapsed * 1e6) : 0.0;
        const double m_sigma = n_med ? s_sigma / n_med : NAN;
        const double m_umax = n_med ? s_umax / n_med : NAN;
        std::ofstream f(join(dir, "resumo.dat"));
        f << std::setprecision(opt.digits);
        f << "# operador chi xi tau beta fat_R_B v0 nx R_ini sigma_nominal estado passo_falha passos "
             "sigma_med umax_med urms_med umax/sigma Ca_s=nu*umax/sigma a4_med a8_med K_med "
             "neg_total_max neg_cor_max min_f MLUPS\n";
        f << op.name << ' ' << op.chi << ' ' << op.xi << ' ' << tau << ' ' << din.recoll << ' ' << A_in << ' ' << v0
          << ' ' << nx << ' ' << raio_ini << ' ' << sigma_nominal << ' ' << estado << ' ' << passo_falha << ' ' << step
          << ' ' << m_sigma << ' ' << m_umax << ' ' << (n_med ? s_urms / n_med : NAN) << ' ' << m_umax / m_sigma << ' '
          << nu * m_umax / m_sigma << ' ' << (n_med ? s_a4 / n_med : NAN) << ' ' << (n_med ? s_a8 / n_med : NAN) << ' '
          << (n_med ? s_K / n_med : NAN) << ' ' << neg_total_max << ' ' << neg_cor_max << ' ' << min_f_global << ' '
          << mlups << '\n';
        std::ostream& o = *con;
        o << "\n\n================ Resultado ================\n";
        o << "   estado            = " << estado << "\n";
        o << "   sigma ( media )   = " << m_sigma << "   ( nominal " << sigma_nominal << " )\n";
        o << "   |u|max ( media )  = " << m_umax << "   |u|max / sigma = " << m_umax / m_sigma << "\n";
        o << "   populacoes < 0    = " << neg_total_max << " ( total )  " << neg_cor_max << " ( cor )\n";
        o << "   MLUPS             = " << mlups << std::endl;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> dirs;
    for (int i = 1; i < argc; ++i) dirs.emplace_back(argv[i]);
    const bool batch = dirs.size() > 1;
    // a case directory given as an argument gets its console output in <dir>/execucao.log, as the
    // reference's batch script keeps it (rodar_lote.py:234-235); with no argument the case is the
    // working directory and the output goes to the console
    const bool to_log = !dirs.empty();
    if (dirs.empty()) dirs.emplace_back(".");
    std::vector<std::unique_ptr<DropCase>> cases;
    std::map<int, std::vector<Member*>> by_prec;
    int rc = 0;
    for (const auto& d : dirs) {
        auto c = std::make_unique<DropCase>();
        c->dir = d;
        if (to_log) {
            c->log = std::make_unique<std::ofstream>(join(d, "execucao.log"));
            c->con = c->log.get();
        } else {
            c->con = &std::cout;
        }
        if (!c->load()) {
            if (to_log) std::fprintf(stderr, "%s: invalid case, skipped\n", d.c_str());
            rc = 1;
            continue;
        }
        by_prec[c->opt.precision].push_back(c.get());
        cases.push_back(std::move(c));
    }
    for (auto& [prec, members] : by_prec) {
        if (batch) std::printf("batch: %zu drop cases, precision %d\n", members.size(), prec);
        std::fflush(stdout);
        if (run_batch(members, prec)) rc = 1;
    }
    return rc;
}
