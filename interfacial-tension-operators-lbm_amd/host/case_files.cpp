// Case-directory inputs and outputs (see case_files.h).
#include "case_files.h"

#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace itolbm_host {

std::string join(const std::string& dir, const std::string& name)
{
    if (dir.empty() || dir == ".") return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

namespace {

// The echo of data_in.txt goes to the console and to dat_out.txt, one "\n<text>\n" block per value
// (LBM_functions.cpp:17-252): the labels below are that file's text.
struct Echo {
    std::ostream& a;
    std::ostream& b;
    template <class T> Echo& operator<<(const T& v)
    {
        a << v;
        b << v;
        return *this;
    }
};

}  // namespace

bool read_data_in(const std::string& dir, DataIn& d, std::ostream& echo)
{
    std::ifstream in(join(dir, "data_in.txt"));
    std::ofstream fdat(join(dir, "dat_out.txt"));
    Echo e{echo, fdat};
    echo << "\nName of the output file: dat_out.txt\n";
    fdat << "Name of the output file: dat_out.txt\n";
    if (!in) return false;
    std::string label;
    // each value after the first is preceded by one label token
    auto num = [&](auto& v) { return static_cast<bool>(in >> label >> v); };
    if (!(in >> d.geo_file)) return false;
    e << "\nName of the geometry file: " << d.geo_file << "\n";
    if (!num(d.ftesc)) return false;
    e << "\nPixel dimension = " << d.ftesc << " m\n";
    if (!num(d.n_steps)) return false;
    e << "\nNumber of steps: " << d.n_steps << "\n";
    if (!num(d.n_files)) return false;
    e << "\nNumber of files: " << d.n_files << "\n";
    if (!num(d.n_threads)) return false;
    e << "\nNumber of threads: " << d.n_threads << "\n";
    if (!num(d.tau)) return false;
    const double visc = (1. / 3.) * (d.tau - 0.5);
    e << "\nRelaxation time: " << d.tau << "\n";
    e << "\nKinematic viscosity = " << visc << "\n";
    if (!num(d.tau_nh)) return false;
    e << "\nRelaxation time (non-hidrodynamic): " << d.tau_nh << "\n";
    if (!num(d.rho)) return false;
    e << "\nDensity: " << d.rho << "\n";
    if (!num(d.tau_R)) return false;
    if (d.tau_R != 0.0)
        e << "\nRelaxation time (Red): " << d.tau_R << " => viscosity R = " << (1.0 / 3.0) * (d.tau_R - 0.5) << "\n";
    if (!num(d.tau_B)) return false;
    if (d.tau_B != 0.0)
        e << "\nRelaxation time (Blue): " << d.tau_B << " => viscosity B = " << (1.0 / 3.0) * (d.tau_B - 0.5)
          << "\n";
    if (!num(d.tau_m)) return false;
    if (d.tau_R != 0.0) e << "\nRelaxation time (mixture): " << d.tau_m << "\n";
    if (!num(d.rho_R)) return false;
    if (d.tau_R != 0.0) e << "\nInitial density (Red): " << d.rho_R << "\n";
    if (!num(d.rho_B)) return false;
    if (d.tau_B != 0.0) e << "\nInitial density (Blue): " << d.rho_B << "\n";
    if (!num(d.A_fact)) return false;
    if (d.tau_R != 0.0) e << "\nInterfacial factor[0;0.4]: " << d.A_fact << "\n";
    if (!num(d.A_fact_R)) return false;
    if (d.tau_R != 0.0) e << "\nInteraction factor (Red) [0;0.4]: " << d.A_fact_R << "\n";
    if (!num(d.A_fact_B)) return false;
    if (d.tau_B != 0.0) e << "\nInteraction factor (Blue) [0;0.4]: " << d.A_fact_B << "\n";
    if (!num(d.recoll)) return false;
    if (d.tau_R != 0.0) e << "\nRecolloring factor[0;1]: " << d.recoll << "\n";
    if (!num(d.wett_R)) return false;
    if (d.tau_R != 0.0) e << "\nRed_wall interaction [0.0 ; 1.0] : " << d.wett_R << "\n";
    return true;
}

bool read_operator(const std::string& path, Operator& op)
{
    std::ifstream in(path);
    std::string s;
    if (!in || !(in >> s)) return false;
    op.name = s;
    double v;
    if (s == "P1") {
        op.op = LBM_OP_P1;
    } else if (s == "P3") {
        op.op = LBM_OP_P3;
        if (in >> v) op.chi = v;
    } else if (s == "F1") {
        op.op = LBM_OP_F1;
    } else if (s == "F2") {
        op.op = LBM_OP_F2;
    } else if (s == "F3") {
        op.op = LBM_OP_F3;
        if (in >> v) op.xi = v;
    } else {
        return false;
    }
    return true;
}

bool read_velocity(const std::string& path, double& v0)
{
    v0 = 0.0;
    std::ifstream in(path);
    if (in && !(in >> v0)) v0 = 0.0;
    return std::isfinite(v0) && v0 >= 0.0;
}

bool read_inclination(const std::string& path, int& p, int& q)
{
    std::ifstream in(path);
    return (in >> p >> q) && !(p == 0 && q == 0);
}

bool read_box(const std::string& path, int& nx, int& ny, int& nz, long long& n_fluid)
{
    std::ifstream in(path);
    if (!in) return false;
    std::string line, word;
    for (int i = 0; i < 4; ++i) std::getline(in, line);
    if (!std::getline(in, line)) return false;
    std::istringstream dims(line);
    if (!(dims >> word >> nx >> ny >> nz) || nx < 1 || ny < 1 || nz < 1) return false;
    for (int i = 0; i < 5; ++i) std::getline(in, line);
    n_fluid = 0;
    const long long n = (long long)nx * ny * nz;
    double v;
    for (long long k = 0; k < n; ++k) {
        if (!(in >> v)) return false;
        if (v != 0.0) ++n_fluid;
    }
    return true;
}

const char* precision_name(int precision)
{
    switch (precision) {
    case LBM_PREC_F64: return "f64";
    case LBM_PREC_F32: return "f32";
    case LBM_PREC_F16: return "f16";
    case LBM_PREC_F32_ZSYM: return "f32z";
    case LBM_PREC_F16_ZSYM: return "f16z";
    default: return "?";
    }
}

bool read_options(const std::string& dir, Options& o, std::string& err)
{
    std::ifstream in(join(dir, "itolbm.txt"));
    if (!in) return true;
    std::string line;
    while (std::getline(in, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream ls(line);
        std::string key, val;
        if (!(ls >> key)) continue;
        if (!(ls >> val)) {
            err = "itolbm.txt: no value for " + key;
            return false;
        }
        try {
            if (key == "precisao") {
                if (val == "f64") o.precision = LBM_PREC_F64;
                else if (val == "f32") o.precision = LBM_PREC_F32;
                else if (val == "f16") o.precision = LBM_PREC_F16;
                else if (val == "f32z") o.precision = LBM_PREC_F32_ZSYM;
                else if (val == "f16z") o.precision = LBM_PREC_F16_ZSYM;
                else throw std::invalid_argument(val);
            } else if (key == "raio") {
                o.radius = std::stod(val);
            } else if (key == "espessura_ini") {
                o.width = std::stod(val);
            } else if (key == "passos_diag") {
                o.diag_every = std::stoi(val);
                if (o.diag_every < 1) throw std::invalid_argument(val);
            } else if (key == "digitos") {
                o.digits = std::stoi(val);
                if (o.digits < 1 || o.digits > 17) throw std::invalid_argument(val);
            } else if (key == "referencia") {
                o.recentre = std::stoi(val) != 0;
            } else if (key == "vtk") {
                if (val == "ascii") o.vtk_flags = 0;
                else if (val == "binario") o.vtk_flags = LBM_VTK_BINARY;
                else if (val == "binario_double") o.vtk_flags = LBM_VTK_BINARY | LBM_VTK_DOUBLE;
                else throw std::invalid_argument(val);
            } else {
                err = "itolbm.txt: unknown key " + key;
                return false;
            }
        } catch (const std::exception&) {
            err = "itolbm.txt: bad value for " + key + ": " + val;
            return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------- writers (Imm_OPS.cpp:199-218, 313-326)

void laplace_header(std::ostream& out)
{
    out << "# 1:passo  2:raio_area  3:raio_grad  4:raio_linha  5:p_in  6:p_out  7:delta_p"
           "  8:sigma  9:|u-v0|max  10:|u-v0|rms  11:Ca  12:massa_R  13:massa_B  14:|q|"
           "  15:sd_p_in  16:sd_p_out  17:sigma_2pontos"
           "  18:K_perfil  19:r_perfil  20:espessura_10_90  21:residuo_ajuste  22:n_camadas\n";
}

void laplace_row(std::ostream& out, int step, const lbm_drop_diag& d, double nu)
{
    const double Ca = (d.sigma > 0.0) ? nu * d.u_max / d.sigma : 0.0;
    const double q = std::sqrt(d.qx * d.qx + d.qy * d.qy + d.qz * d.qz);
    out << step << ' ' << d.raio << ' ' << d.raio_grad << ' ' << d.raio_linha << ' ' << d.p_in << ' ' << d.p_out
        << ' ' << d.delta_p << ' ' << d.sigma << ' ' << d.u_max << ' ' << d.u_rms << ' ' << Ca << ' ' << d.massa_R
        << ' ' << d.massa_B << ' ' << q << ' ' << d.sd_in << ' ' << d.sd_out << ' ' << d.sigma_linha << ' '
        << d.K_perfil << ' ' << d.r_perfil << ' ' << d.esp_perfil << ' ' << d.res_perfil << ' ' << d.n_perfil
        << '\n';
    out.flush();
}

void iso_header(std::ostream& out) { out << "# 1:passo  2:R0  3:|a2|/R0  4:|a4|/R0  5:|a8|/R0  6:rmin  7:rmax\n"; }

void iso_row(std::ostream& out, int step, const lbm_drop_diag& d)
{
    out << step << ' ' << d.R0 << ' ' << d.a2 << ' ' << d.a4 << ' ' << d.a8 << ' ' << d.rmin << ' ' << d.rmax
        << '\n';
    out.flush();
}

// galilean.dat (gi_header / gi_write, Galilean_diagnostics.cpp:107-123): 36 columns at 17 digits
void galilean_header(std::ostream& out)
{
    out << "# step v0 Mach sigma R umax_ref urms_ref umax_mean urms_mean umax_lab mean_x mean_y mean_z "
           "drop_ux drop_uy x0 y0 x_unwrapped center_dx center_dy axis_ratio mass_R mass_B min_f "
           "negative_total negative_color min_rho min_phi max_phi qx qy qz p_in p_out delta_p urms_lab\n";
}

void galilean_row(std::ostream& out, int step, double v0, double cs2, const lbm_drop_diag& d)
{
    out << std::setprecision(17) << step << ' ' << v0 << ' ' << v0 / std::sqrt(cs2) << ' ' << d.sigma << ' '
        << d.raio << ' ' << d.max_ref << ' ' << d.rms_ref << ' ' << d.max_mean << ' ' << d.rms_mean << ' '
        << d.max_lab << ' ' << d.mean_x << ' ' << d.mean_y << ' ' << d.mean_z << ' ' << d.drop_x << ' '
        << d.drop_y << ' ' << d.x0 << ' ' << d.y0 << ' ' << d.x_unwrapped << ' ' << d.center_dx << ' '
        << d.center_dy << ' ' << d.axis_ratio << ' ' << d.massa_R << ' ' << d.massa_B << ' ' << d.min_f << ' '
        << d.negative_total << ' ' << d.negative_color << ' ' << d.min_rho << ' ' << d.min_phi << ' '
        << d.max_phi << ' ' << d.qx << ' ' << d.qy << ' ' << d.qz << ' ' << d.p_in << ' ' << d.p_out << ' '
        << d.delta_p << ' ' << d.rms_lab << '\n';
    out.flush();
}

// inclinada.dat (Imm_INCL.cpp:339-342, 406-411)
void inclinada_header(std::ostream& out)
{
    out << "# 1:passo  2:sigma_pop  3:int_grad  4:K  5:espessura_10_90  6:residuo  7:d0  8:|u|max  9:|u|rms"
           "  10:|u_n|max  11:|u_t|max  12:massa_R  13:massa_B  14:qx  15:qy  16:soma_Fx  17:soma_Fy"
           "  18:min_f  19:neg_total  20:neg_cor\n";
}

void inclinada_row(std::ostream& out, int step, const lbm_incl_diag& r)
{
    out << step << ' ' << r.sigma_pop << ' ' << r.int_grad << ' ' << r.K << ' '
        << (r.K > 0 ? 2.0 * std::log(9.0) / r.K : 0.0) << ' ' << r.res << ' ' << r.d0 << ' ' << r.u_max << ' '
        << r.u_rms << ' ' << r.un_max << ' ' << r.ut_max << ' ' << r.massa_R << ' ' << r.massa_B << ' ' << r.qx
        << ' ' << r.qy << ' ' << r.SFx << ' ' << r.SFy << ' ' << r.min_f << ' ' << r.neg_total << ' ' << r.neg_cor
        << '\n';
    out.flush();
}

void perfil_header(std::ostream& out) { out << "# d  c=rho_R/rho  rho  ux  uy\n"; }

}  // namespace itolbm_host
