// Argument blocks and partial-result layouts of the diagnostic kernels (diag.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace itolbm {

// drop pass 1: sums / maxima / minima written per block, in this order
enum {
    P1_NFLUID, P1_NRUIM, P1_MASSA_R, P1_MASSA_B, P1_QXD, P1_QYD, P1_QZD, P1_U2, P1_AREA,
    P1_CXC, P1_CXS, P1_CYC, P1_CYS,
    P1_BAD, P1_NEGT, P1_NEGC, P1_MASS, P1_MASSR, P1_QX, P1_QY, P1_QZ, P1_QRX, P1_QRY,
    P1_NS
};
constexpr int P1_NX = 2;   // u_max (drop radius), max_phi
constexpr int P1_NN = 3;   // min_f, min_rho, min_phi
constexpr int P1_NP = P1_NS + P1_NX + P1_NN;

enum {
    P2_SOMW, P2_SOMWR, P2_SXE, P2_SWE, P2_SXD, P2_SWD,
    P2_PIN, P2_PIN2, P2_NIN, P2_POUT, P2_POUT2, P2_NOUT, P2_PCTR, P2_NCTR, P2_PCNT,
    P2_SREF, P2_SMEAN, P2_SLAB, P2_WEIGHT, P2_SX, P2_SY, P2_SXX, P2_SYY, P2_SXY,
    P2_LW, P2_LX, P2_RW, P2_RX, P2_PCANTO,
    P2_NS
};
constexpr int P2_NX = 3;   // max_ref, max_mean, max_lab
constexpr int P2_NP = P2_NS + P2_NX;

enum {
    I1_NRUIM, I1_NEGT, I1_NEGC, I1_SFX, I1_SFY, I1_MR, I1_MB, I1_QX, I1_QY, I1_U2, I1_SIGMA, I1_IGRAD,
    I1_SX, I1_SY, I1_SXX, I1_SXY, I1_NFIT,
    I1_NS
};
constexpr int I1_NX = 4;   // u_max, un_max, ut_max, fit residual
constexpr int I1_NN = 1;   // min_f
constexpr int I1_NP = I1_NS + I1_NX + I1_NN;

struct DropPass1Args {
    const double* fR;
    const double* fB;
    long long S;
    int nx, ny;
    const double* F;       // [3][S] force (class F) or nullptr
    const double* cosx;    // cos / sin (2 pi x / nx), (2 pi y / ny), from the host's libm
    const double* sinx;
    const double* cosy;
    const double* siny;
    double* gal;           // out [4][S]: u_x, u_y, u_z, phi = rho_R / rho (calc_galilean's arrays)
    double* dsum;          // out [2][S]: rho_R - rho_B, rho_R + rho_B
    double* aphi;          // out [S]: calc_drop_radius's phi term of each site (0 where it adds none)
    double* partial;       // out [nblk][P1_NP]
};

// Area and circular-centroid sums of calc_drop_radius (Other_functions.cpp:89-183) accumulated in
// site order, as the reference's loop does: they fix the drop centre and radius, and with them every
// region and shell the later passes select, so they must carry the reference's exact bits.
struct OrderedArgs {
    const double* aphi;    // [n] phi term per site
    long long n;
    int nx;
    const double* cosx;
    const double* sinx;
    const double* cosy;
    const double* siny;
    double* out;           // [5]: area, sum phi cos(tx x), sum phi sin(tx x), ... cos(ty y), sin(ty y)
};

struct DropPass2Args {
    const double* fR;
    const double* fB;
    long long S;
    int nx, ny;
    const double* gal;
    const double* dsum;
    double x0, y0;             // centre (calc_drop_radius)
    double gx0, gy0;           // centre used by calc_galilean (same values)
    int y_lin, y_line;         // measurement lines of calc_drop_radius / calc_galilean
    double r_in_max, r_out_min;
    double v0, mean_x, mean_y, mean_z;
    long long pc;              // corner site of calc_galilean
    int n_casca;
    double* hist;              // [n_casca] sum of phi per half-site shell
    unsigned long long* histn; // [n_casca] sites per shell
    double* partial;           // out [nblk][P2_NP]
};

struct ModesArgs {
    const double* dsum;
    long long S;
    int nx, ny;
    double x0, y0, R_est;
    const double* ct;          // cos / sin of the 256 ray angles, from the host's libm
    const double* st;
    double* r;                 // out [256]
};

struct InclArgs {
    const double* fR;
    const double* fB;
    long long S;
    int nx, ny;
    const double* F;           // [3][S] or nullptr
    const double* gm;          // [3][S] mediator momentum
    int p, q;
    long long M;               // band period in units of p x + q y
    double m;                  // sqrt(p^2 + q^2)
    double nxn, nyn, txn, tyn; // normal and tangent
    int fit_pass;              // 1: also the fit residual with (fa, fK)
    double fa, fK;
    double* prof;              // nullptr or out [5][S]: d, c_R, rho, u_x, u_y
    double* partial;           // out [nblk][I1_NP]
};

hipError_t launch_drop_pass1(const DropPass1Args& a, int nblk, hipStream_t st);
hipError_t launch_drop_pass2(const DropPass2Args& a, int nblk, hipStream_t st);
hipError_t launch_drop_ordered(const OrderedArgs& a, hipStream_t st);
hipError_t launch_drop_modes(const ModesArgs& a, hipStream_t st);
hipError_t launch_incl_pass(const InclArgs& a, int nblk, hipStream_t st);

}  // namespace itolbm
