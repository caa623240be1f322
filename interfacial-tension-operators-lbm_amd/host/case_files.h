// Case-directory file contract of the drop and inclined-interface programs: the inputs they read
// (data_in.txt, operador.txt, velocity.txt, inclinacao.txt, meio.vtk) and the outputs they write
// (dat_out.txt, laplace.dat, galilean.dat, iso.dat, inclinada.dat, perfil.dat, resumo.dat, VTK fields).
// Formats follow the reference programs so that its batch scripts and analysis read them unchanged.
#pragma once

#include <fstream>
#include <string>
#include <vector>

#include "itolbm.h"

namespace itolbm_host {

// data_in.txt: label/value pairs in a fixed order (read_data, LBM_functions.cpp:9-255).
struct DataIn {
    std::string geo_file;
    double ftesc = 0;
    int n_steps = 0, n_files = 0, n_threads = 0;
    double tau = 0, tau_nh = 0, rho = 0, tau_R = 0, tau_B = 0, tau_m = 0, rho_R = 0, rho_B = 0;
    double A_fact = 0, A_fact_R = 0, A_fact_B = 0, recoll = 0, wett_R = 0;
};

// Reads <dir>/data_in.txt and writes <dir>/dat_out.txt with the echo of the values; the same echo
// goes to `echo` (the program's console).  False if the file is missing or truncated.
bool read_data_in(const std::string& dir, DataIn& d, std::ostream& echo);

// operador.txt: "<P1|P3|F1|F2|F3> [chi (P3) | xi (F3)]" (le_operador, Operadores_tensao.cpp:64-86).
struct Operator {
    int op = LBM_OP_P3;
    double chi = 0.5, xi = 1.569988;
    std::string name = "P3";
    bool force() const { return op >= LBM_OP_F1; }
};
bool read_operator(const std::string& path, Operator& op);

// velocity.txt: v0 >= 0, absent file = 0 (Imm_OPS.cpp:84-89).  False on an invalid value.
bool read_velocity(const std::string& path, double& v0);

// inclinacao.txt: "p q", integers not both zero (Imm_INCL.cpp:215-219).
bool read_inclination(const std::string& path, int& p, int& q);

// meio.vtk (legacy structured points): box size from the DIMENSIONS line and the number of
// fluid (non-zero) points (read_geo, LBM_functions.cpp:266-336).
bool read_box(const std::string& path, int& nx, int& ny, int& nz, long long& n_fluid);

// Optional <dir>/itolbm.txt, one "key value" per line, for what the reference fixes at compile
// time.  Absent keys keep the reference's values.
//   precisao        f64 | f32 | f16 | f32z | f16z   storage form (default f64; z: z-symmetric storage)
//   raio            initial drop radius    (default 0: nx / 4)
//   espessura_ini   initial 10-90 width    (default 3.45; <= 0: sharp colour field)
//   passos_diag     diagnostic interval    (default 500)
//   digitos         digits of laplace.dat / iso.dat / inclinada.dat / perfil.dat / resumo.dat (default 10)
//   referencia      1 | 0: re-centre the reduced-precision storage on the measured phase
//                   densities after each diagnostic (default 1)
//   vtk             ascii | binario | binario_double: form of the final rho_R / rho_B / vel files
//                   (default ascii, as the reference writes them; binario: big-endian float)
struct Options {
    int precision = LBM_PREC_F64;
    double radius = 0.0;
    double width = 3.45;
    int diag_every = 500;
    int digits = 10;
    bool recentre = true;
    int vtk_flags = 0;   // LBM_VTK_BINARY | LBM_VTK_DOUBLE
};
bool read_options(const std::string& dir, Options& o, std::string& err);
// the itolbm.txt name of a storage form
const char* precision_name(int precision);

// ---------------------------------------------------------------- writers
void laplace_header(std::ostream& out);
void laplace_row(std::ostream& out, int step, const lbm_drop_diag& d, double nu);
void iso_header(std::ostream& out);
void iso_row(std::ostream& out, int step, const lbm_drop_diag& d);
void galilean_header(std::ostream& out);
void galilean_row(std::ostream& out, int step, double v0, double cs2, const lbm_drop_diag& d);
void inclinada_header(std::ostream& out);
void inclinada_row(std::ostream& out, int step, const lbm_incl_diag& r);
void perfil_header(std::ostream& out);

std::string join(const std::string& dir, const std::string& name);

}  // namespace itolbm_host
