// Imm_INCL: a red band bounded by two straight interfaces of integer normal (p, q) in a periodic box,
// with the interfacial-tension operator of operador.txt, advanced on the GPU through the itolbm C ABI.
//
// Drop-in for the reference program of the same name (Imm_INCL.cpp): run with no argument in a case
// directory, it reads data_in.txt, operador.txt, inclinacao.txt and meio.vtk there and writes
// dat_out.txt, inclinada.dat (every 500 steps and the last step), perfil.dat (last step) and
// resumo.dat with the reference's formats, stability rule and exit codes (Imm_INCL.cpp:195-473).
// The reference requires a square N x N box; here nx x ny boxes are accepted too, with the band
// period M = gcd(p nx, q ny) / gcd(p, q) along p x + q y (M = N for a square box).  Several case
// directories as arguments run together in one GPU context, as for Imm_OPS.
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch.h"
#include "case_files.h"

using namespace itolbm_host;

namespace {

constexpr double U_LIMITE = 0.4;   // above this: unstable (Imm_INCL.cpp:44)

struct BandCase : Member {
    std::string dir;
    std::ostream* con = nullptr;
    std::unique_ptr<std::ofstream> log;
    DataIn din;
    Operator op;
    Options opt;
    int p = 1, q = 0, nx = 0, ny = 0;
    long long n_fluid = 0;
    double tau = 0, A_in = 0, sigma_nominal = 0, theta = 0;
    std::ofstream f_inc;
    double s_sig = 0, s_um = 0, s_un = 0, s_ut = 0, s_K = 0, s_res = 0, s_ig = 0;
    int n_med = 0;
    long long neg_t = 0, neg_c = 0;
    double min_f = 1e300, d0_ini = NAN, d0_fim = NAN;
    std::string estado = "estavel";
    int passo_falha = -1;

    const std::string& name() const override { return dir; }

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
        if (!read_inclination(join(dir, "inclinacao.txt"), p, q)) {
            *con << "inclinacao.txt: \"p q\" inteiros, nao ambos nulos.\n";
            return false;
        }
        if (std::fabs(din.tau_R - din.tau_B) > 1e-14) {
            *con << "tau_R e tau_B tem de ser iguais.\n";
            return false;
        }
        int nz = 0;
        if (!read_box(join(dir, din.geo_file), nx, ny, nz, n_fluid)) {
            *con << "\nNao foi possivel abrir " << din.geo_file << "\n";
            return false;
        }
        if (nz != 1 || n_fluid != (long long)nx * ny) {
            *con << "\nEste programa requer caixa N x M x 1, toda fluida.\n";
            return false;
        }
        const double m = std::sqrt((double)(p * p + q * q));
        tau = din.tau_R;
        A_in = din.A_fact;
        theta = std::atan2((double)q, (double)p) * 180.0 / M_PI;
        sigma_nominal = op.force() ? A_in : 4.0 / 9.0 * A_in * tau;
        std::ostream& o = *con;
        o << "\nInterface plana inclinada -- GPU ( D3Q19, BGK, Latva-Kokko )\n";
        o << "   operador          = " << op.name << "\n";
        o << "   caixa             = " << nx << " x " << ny << "\n";
        o << "   ( p , q )         = ( " << p << " , " << q << " )   theta = " << theta << " graus\n";
        if (nx == ny) o << "   distancia entre interfaces = " << 0.5 * nx / m << "\n";
        o << "   sigma nominal     = " << sigma_nominal << std::endl;
        f_inc.open(join(dir, "inclinada.dat"));
        f_inc << std::setprecision(opt.digits);
        inclinada_header(f_inc);
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

    bool init(lbm_ctx* ctx, int id) override { return lbm_init_band(ctx, id, p, q, opt.width) == LBM_OK; }

    int next_event(int s) const override
    {
        int t = s;
        while (t < din.n_steps - 1 && t % opt.diag_every != 0) ++t;
        return t;
    }

    bool on_event(lbm_ctx* ctx, int id, int t) override
    {
        const bool last = (t == din.n_steps - 1);
        lbm_incl_diag r{};
        std::vector<double> prof(last ? 5 * (size_t)nx * ny : 0);
        if (lbm_diag_incl(ctx, id, p, q, &r, last ? prof.data() : nullptr) != LBM_OK) {
            *con << "\ndiagnostico falhou: " << lbm_last_error(ctx) << "\n";
            r.ok = 0;
        }
        if (last) {
            std::ofstream fp(join(dir, "perfil.dat"));
            perfil_header(fp);
            fp << std::setprecision(opt.digits);
            for (size_t s = 0; s < (size_t)nx * ny; ++s) {
                const double* row = &prof[5 * s];
                if (!(row[2] > 0) || !(std::fabs(row[0]) < 12.0)) continue;
                fp << row[0] << ' ' << row[1] << ' ' << row[2] << ' ' << row[3] << ' ' << row[4] << '\n';
            }
        }
        inclinada_row(f_inc, t, r);
        if (t == 0) d0_ini = r.d0;
        d0_fim = r.d0;
        neg_t = std::max(neg_t, r.neg_total);
        neg_c = std::max(neg_c, r.neg_cor);
        min_f = std::min(min_f, r.min_f);
        if (t >= din.n_steps / 2) {
            s_sig += r.sigma_pop; s_um += r.u_max; s_un += r.un_max; s_ut += r.ut_max;
            s_K += r.K; s_res += r.res; s_ig += r.int_grad;
            n_med++;
        }
        *con << "step = " << t << "   sigma_pop = " << r.sigma_pop << "   K = " << r.K << "   |u|max = " << r.u_max
             << "   neg = " << r.neg_total << std::endl;
        if (!r.ok || r.u_max > U_LIMITE) {
            estado = "instavel";
            passo_falha = t;
            *con << "\nSimulacao instavel no passo " << t << std::endl;
            return false;
        }
        return true;
    }

    void finish(int step, double elapsed) override
    {
        const double mlups = (elapsed > 0) ? (double)n_fluid * step / (elapsed * 1e6) : 0.0;
        auto med = [&](double s) { return n_med ? s / n_med : NAN; };
        std::ofstream f(join(dir, "resumo.dat"));
        f << std::setprecision(opt.digits);
        f << "# operador chi xi tau beta fat_R_B N p q theta sigma_nominal estado passo_falha passos "
             "sigma_pop_med int_grad_med K_med residuo_med umax_med un_max_med ut_max_med umax/sigma_nominal "
             "d0_ini d0_fim neg_total_max neg_cor_max min_f MLUPS\n";
        f << op.name << ' ' << op.chi << ' ' << op.xi << ' ' << tau << ' ' << din.recoll << ' ' << A_in << ' ' << nx
          << ' ' << p << ' ' << q << ' ' << theta << ' ' << sigma_nominal << ' ' << estado << ' ' << passo_falha << ' '
          << step << ' ' << med(s_sig) << ' ' << med(s_ig) << ' ' << med(s_K) << ' ' << med(s_res) << ' ' << med(s_um)
          << ' ' << med(s_un) << ' ' << med(s_ut) << ' ' << med(s_um) / sigma_nominal << ' ' << d0_ini << ' ' << d0_fim
          << ' ' << neg_t << ' ' << neg_c << ' ' << min_f << ' ' << mlups << '\n';
        std::ostream& o = *con;
        o << "\n\n================ Resultado ================\n";
        o << "   estado            = " << estado << "\n";
        o << "   theta             = " << theta << " graus\n";
        o << "   sigma_pop         = " << med(s_sig) << "   ( nominal " << sigma_nominal << ";  ~0 nas forcas )\n";
        o << "   int |grad|        = " << med(s_ig) << "   ( deve ser 2 )\n";
        o << "   K / espessura     = " << med(s_K) << " / " << 2.0 * std::log(9.0) / med(s_K) << "\n";
        o << "   |u|max            = " << med(s_um) << "   ( |u_n| " << med(s_un) << ", |u_t| " << med(s_ut) << " )\n";
        o << "   d0 inicial / final= " << d0_ini << " / " << d0_fim << "\n";
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
    std::vector<std::unique_ptr<BandCase>> cases;
    std::map<int, std::vector<Member*>> by_prec;
    int rc = 0;
    for (const auto& d : dirs) {
        auto c = std::make_unique<BandCase>();
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
        if (batch) std::printf("batch: %zu band cases, precision %d\n", members.size(), prec);
        std::fflush(stdout);
        if (run_batch(members, prec)) rc = 1;
    }
    return rc;
}
