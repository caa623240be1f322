/*
 * itolbm: C ABI of the MI355X colour-gradient LBM step (D3Q19 slab, one node in z).
 *
 * The reference has no C ABI: its hot path is the step loop of its drivers
 * (Imm_OPS.cpp:246-372, Imm_INCL.cpp:358-442) calling, per site, colisao_comum
 * (Operadores_tensao.cpp:340-373), emite_mediadores / propag_site_med / propag_site
 * (LBM_functions.cpp:1183-1337) and, for the force operators, calcula_forca
 * (Operadores_tensao.cpp:202-332), followed every PASSOS_DIAG steps by the diagnostics
 * calc_galilean / calc_perfil_logistico / modos_interface (Imm_OPS.cpp:289-353) or
 * mede_inclinada (Imm_INCL.cpp:85-189).  The entry points below replace that loop; each cites
 * the reference code it stands in for.  A maintainer binds them from C++ (the reference's
 * language) with a plain #include, or from Python with ctypes.
 *
 * Conventions: every function returns 0 on success and a negative LBM_E* code on failure, with a
 * message in lbm_last_error(ctx); no exception crosses the ABI; one host thread per context.
 * Host population arrays use the reference layout [site][19], site = x + nx * y
 * (Imm_OPS.cpp:158-181, read_geo numbering of an all-fluid box LBM_functions.cpp:280-354).
 */
#ifndef ITOLBM_H
#define ITOLBM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_ctx lbm_ctx;

/* operator ids: Operadores_tensao.cpp:45-49 */
enum { LBM_OP_P1 = 1, LBM_OP_P3 = 3, LBM_OP_F1 = 11, LBM_OP_F2 = 12, LBM_OP_F3 = 13 };

/* arithmetic / storage forms */
enum {
    LBM_PREC_F64 = 0, /* FP64 compute and storage: the drop-in form, bit-equal to the reference's arithmetic */
    LBM_PREC_F32 = 1, /* FP32 compute, FP32 storage of s_i / w_i and e_i (304 B per lattice update) */
    LBM_PREC_F16 = 2, /* FP32 compute, FP16 storage of s_i / w_i and of the minority colour fraction (152 B per lattice update) */
    LBM_PREC_F32_ZSYM = 3, /* as LBM_PREC_F32 for states mirror-symmetric in z: 14 of the 19 populations stored (224 B) */
    LBM_PREC_F16_ZSYM = 4  /* as LBM_PREC_F16 for states mirror-symmetric in z: 14 of the 19 populations stored (112 B) */
};
/* z-symmetric storage (no counterpart in the reference, whose D3Q19 slab stores all 19): with one node
 * in z, velocity i and its z mirror pull from the same site, and every state the reference's programs
 * create is symmetric under z -> -z (f_6 = f_5, f_13 = f_11, f_14 = f_12, f_17 = f_15, f_18 = f_16), a
 * symmetry the FP32/FP16 collision keeps exactly.  Only the 14 distinct populations are stored; the
 * results equal those of LBM_PREC_F32 / LBM_PREC_F16 on the same symmetric state.  lbm_upload_state
 * rejects (LBM_E_ARG) a state whose mirror populations differ beyond 1e-12 relative. */
/* In the FP32/FP16 forms each population is stored as A_i = s_i / w_i, s_i = fR_i + fB_i - w_i (deviation
 * from the rest state rho0 = 1, scaled per direction by 1 / w_i) and e_i = (fR_i - fB_i) / (fR_i + fB_i)
 * (colour fraction, exactly +-1 in a pure phase).  Transfers and diagnostics convert to and from FP64. */

/* which population state to transfer */
enum {
    LBM_STATE_POST_STREAM = 0,    /* state at the start of a step (the reference's inif_R after the swap) */
    LBM_STATE_POST_COLLISION = 1  /* post-collision, pre-stream (what the reference's diagnostics read) */
};

/* lbm_step flags */
enum { LBM_STEP_RECORD_AUX = 1 /* the last step keeps the mediator momentum and force for diagnostics */ };

/* error codes */
enum {
    LBM_OK = 0,
    LBM_E_ARG = -1,    /* invalid argument */
    LBM_E_HIP = -2,    /* HIP runtime error */
    LBM_E_STATE = -3,  /* call not valid in the current state (e.g. diagnostics without recorded aux) */
    LBM_E_NOMEM = -4
};

/* One case: the parameters the reference reads from data_in.txt / operador.txt
 * (read_data LBM_functions.cpp:10-268, le_operador Operadores_tensao.cpp:66-86). */
typedef struct {
    int nx, ny;        /* periodic box nx x ny x 1 */
    int op;            /* LBM_OP_* */
    double chi;        /* P3 chi (default 0.5) */
    double xi;         /* F3 xi (default 1.569988) */
    double tau;        /* tau_R = tau_B */
    double A;          /* fat_R_B: amplitude A (P class) or sigma (F class) */
    double beta;       /* recoll */
} lbm_case_desc;

/* Drop measurements: the DROP fields (Definitions.cpp:251-339) set by calc_laplace_tension /
 * calc_drop_radius / calc_perfil_logistico (Other_functions.cpp:60-554), the GI_DATA fields of
 * calc_galilean (Galilean_diagnostics.cpp:7-105) and the MODOS of modos_interface
 * (Operadores_tensao.cpp:380-449). */
typedef struct {
    double raio, raio_grad, raio_linha, p_in, p_out, delta_p, sigma, u_max, u_rms;
    double massa_R, massa_B, qx, qy, qz, sd_in, sd_out, sigma_linha;
    double K_perfil, r_perfil, esp_perfil, res_perfil;
    int n_perfil, n_in, n_out;
    double x0, y0, area, p_centro, p_canto;
    double mean_x, mean_y, mean_z, drop_x, drop_y, max_ref, rms_ref, max_mean, rms_mean;
    double max_lab, rms_lab, min_f, min_rho, min_phi, max_phi, axis_ratio;
    double center_dx, center_dy, x_unwrapped;
    long long negative_total, negative_color;
    double R0, a2, a4, a8, rmin, rmax;
    int ok;       /* calc_galilean's return value */
    int n_ruim;   /* non-finite or non-positive sites */
} lbm_drop_diag;

/* Inclined-interface measurements: INCL of Imm_INCL.cpp:62-70, filled as mede_inclinada. */
typedef struct {
    double sigma_pop, int_grad, K, res, d0, u_max, u_rms, un_max, ut_max;
    double massa_R, massa_B, qx, qy, SFx, SFy, min_f;
    long long neg_total, neg_cor;
    int n_ruim, n_ajuste, ok;
} lbm_incl_diag;

/* Allocates device state for n_cases cases (Imm_OPS.cpp:158-181 allocations) in form `precision`. */
int lbm_create(const lbm_case_desc* cases, int n_cases, int precision, lbm_ctx** out);

/* Frees everything (the reference never frees). */
int lbm_destroy(lbm_ctx* ctx);

/* Last error message of ctx ("" if none); a static message if ctx is NULL. */
const char* lbm_last_error(const lbm_ctx* ctx);

/* Number of device bytes held by ctx. */
long long lbm_device_bytes(const lbm_ctx* ctx);

/* Sets the state at the start of the next step from post-stream populations, as the reference's
 * initial_conditions_drop / Imm_INCL init fill inif_R / inif_B (Initial_conditions.cpp:37-169,
 * Imm_INCL.cpp:309-334). */
int lbm_upload_state(lbm_ctx* ctx, int case_id, const double* fR, const double* fB);

/* Initial drop: initial_conditions_drop (Initial_conditions.cpp:37-169), EQ_CLASSICO branch, for a
 * cylinder of radius R centred at the integer site (x0, y0), rho_R = rho_B = 1, velocity (v0, 0, 0),
 * logistic profile of 10-90 % width `width` (k = 2 ln 9 / width, Imm_OPS.cpp:197) or a sharp step
 * if width <= 0.  Computed on the host with libm, uploaded as the state at the next step. */
int lbm_init_drop(lbm_ctx* ctx, int case_id, double R, double width, int x0, int y0, double v0);

/* Initial band of Imm_INCL.cpp:309-334: rest equilibrium rho = 1, red fraction logistic (or sharp if
 * width <= 0, 1/2 on the interface sites) in the signed distance to the nearest interface of the
 * band 0 <= (p x + q y) mod M < M/2, M = gcd(p nx, q ny) / gcd(p, q) (= N for an N x N box). */
int lbm_init_band(lbm_ctx* ctx, int case_id, int p, int q, double width);

/* Copies populations to the host: LBM_STATE_POST_COLLISION gives what the reference holds in
 * inif_R / inif_B between collide and swap (Imm_OPS.cpp:289-364), LBM_STATE_POST_STREAM what it
 * holds after the swap. */
int lbm_download_state(lbm_ctx* ctx, int case_id, int which, double* fR, double* fB);

/* Advances every case by n_steps time steps (Imm_OPS.cpp:246-287 + swap :368-371), stream-ordered
 * and asynchronous.  flags: LBM_STEP_RECORD_AUX keeps the last step's mediator momentum and
 * class-F force for the diagnostics below.  LBM_PREC_F64: the P1/P3 and F3 cases of a family launch
 * covering 196,608 sites or more are stepped in the compact form (see lbm_stored_bytes), the others
 * in the full form; a case in the other form is converted first, exactly, on the device. */
int lbm_step(lbm_ctx* ctx, int n_steps, int flags);

/* Advances one case only (the form chosen by the case's own size, as for lbm_step's launches). */
int lbm_step_case(lbm_ctx* ctx, int case_id, int n_steps, int flags);

/* As lbm_step_case, allocating nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  The case's
 * launch table and tile map go into a static device array of 4 MiB + 256 B that each GPU the library
 * runs on holds until the process ends, instead of memory lbm_step_case allocates; the kernels, the
 * tile shape (the one lbm_step_case tuned for the case, else the default) and the results are the
 * same.  No tuning, no allocation, no copy from host memory, no synchronisation.  Limits: a case of
 * more than 2^20 tiles (64-column strips x row segments) is refused with LBM_E_ARG;
 * LBM_STEP_RECORD_AUX needs the aux fields that a first recording lbm_step / lbm_step_case allocates
 * (LBM_E_STATE otherwise).  Every call on one GPU shares that GPU's array, so calls that overlap on
 * one GPU (other streams, other host threads) race; calls queued on one stream do not.  An FP64 case
 * whose storage form must change (see lbm_step) is first converted by one kernel launch, nothing
 * allocated; under a graph capture that is refused (LBM_E_STATE). */
int lbm_step_case_noalloc(lbm_ctx* ctx, int case_id, int n_steps, int flags);

/* Waits for all queued work. */
int lbm_sync(lbm_ctx* ctx);

/* Queues all later work of ctx on the caller's HIP stream (a hipStream_t, default priority);
 * NULL returns to the context's own stream.  Waits for work already queued first.  No
 * reference counterpart: it lets a caller order the steps with its own work and time them. */
int lbm_set_stream(lbm_ctx* ctx, void* hip_stream);

/* lbm_step queues whole runs of steps_per_graph steps (even; default 32) as one captured HIP graph,
 * replayed with a single launch, and the remaining steps launch by launch; 0 turns graphs off.  The
 * results are the same either way.  The graph is captured on first use and again after the set of
 * active cases, a tile shape or the parity of a family changes.  No reference counterpart (it cuts
 * the per-step launch cost of small cases). */
int lbm_set_graph(lbm_ctx* ctx, int steps_per_graph);

/* Tile shape lbm_step uses for the kernel family of case case_id: rows per wavefront segment, the
 * pipeline depth (FP32/FP16: rows in flight ahead, 0 = the kernel's default; FP64 F1 / F2: 2 when a
 * row's totals and colour sums are kept until its collision) and whether the shape was tuned (1) or is
 * the default (0).  No reference counterpart (the reference has no tiles); for reports and benchmarks. */
int lbm_tile_shape(const lbm_ctx* ctx, int case_id, int* rows, int* depth, int* tuned);

/* How lbm_step launches the cases of the strip-kernel families: 0 one launch per family (on
 * concurrent streams), 1 one launch for all of them (FP64: a wavefront per tile, each running its
 * family's code), 2 that launch with a grid of the resident capacity, each wavefront looping over
 * tiles (persistent); -1 not chosen yet.  lbm_step chooses by timing real steps once the tile shapes
 * are tuned (FP32/FP16 forms: always 0); ITOLBM_FUSED=0|1|2 fixes it.  The results are the same bit
 * for bit.  No reference counterpart (the reference runs one case per process). */
int lbm_step_mode(const lbm_ctx* ctx, int* mode);

/* Bytes a time step reads and writes per lattice update of the case in its stored form (every stored
 * value read once and written once): 384 for the FP64 compact form of the P1/P3 and F3 cases (19
 * totals and 5 recolouring scalars per site, from which the receiving site forms the 38 colour
 * populations bit for bit), 608 for the FP64 form storing the 38 populations, 304 / 224 / 152 / 112
 * for the FP32, FP32 z-symmetric, FP16 and FP16 z-symmetric forms.  No reference counterpart; for
 * reports and roofline figures. */
int lbm_stored_bytes(const lbm_ctx* ctx, int case_id, int* bytes_per_update);

/* Copies `bytes` bytes of device memory from src to dst (both 16-byte aligned) on the HIP stream
 * hip_stream (NULL: the null stream) with the library's plain streaming kernel (16-byte loads,
 * nontemporal stores).  No reference counterpart: it measures the bandwidth a read-once / write-once
 * kernel reaches on the GPU, the ceiling the step kernels' rates are compared with. */
int lbm_stream_copy(void* dst, const void* src, unsigned long long bytes, void* hip_stream);

/* Includes (active != 0, the initial state) or leaves out (0) a case from lbm_step, e.g. once a batch
 * member has finished or become unstable; lbm_step_case still advances it.  No reference
 * counterpart (the reference runs one case per process). */
int lbm_set_active(lbm_ctx* ctx, int case_id, int active);

/* Reduced-precision forms (LBM_PREC_F32 / LBM_PREC_F16) store each population as its deviation from
 * w_i (1 + r(e)), r(e) = (rho_red + rho_blue) / 2 - 1 + e (rho_red - rho_blue) / 2, where e is the
 * population's colour fraction (+1 pure red, -1 pure blue).  Setting rho_red / rho_blue close to the
 * bulk phase densities (for a drop: 3 p_in and 3 p_out from lbm_diag_drop) keeps the stored numbers
 * small, so FP16 keeps its relative digits where the pressure is measured.  The stored state is
 * re-expressed in place (stream-ordered); no effect on LBM_PREC_F64.  Initial reference: 1, 1.  No
 * reference counterpart (the reference stores FP64 populations). */
int lbm_set_reference(lbm_ctx* ctx, int case_id, double rho_red, double rho_blue);

/* The last recorded step's mediator momentum gm = -grad(phi) (momentum of inif_m) and force field
 * (ini_force), each [site][3]; either pointer may be NULL. */
int lbm_download_aux(lbm_ctx* ctx, int case_id, double* gm, double* F);

/* Drop diagnostics on the post-collision state of the last step (needs LBM_STEP_RECORD_AUX for the
 * force operators): calc_galilean + calc_perfil_logistico + modos_interface as Imm_OPS.cpp:295-300,
 * with v0, step and the initial centre (x_ini, y_ini) as the reference passes them.  FP64,
 * reduced on the device. */
int lbm_diag_drop(lbm_ctx* ctx, int case_id, double v0, int step, double x_ini, double y_ini,
                  lbm_drop_diag* out);

/* Geometry file (meio.vtk), as read_geo reads it (LBM_functions.cpp:280-354): the box size from the
 * DIMENSIONS line (the 5th), 5 more header lines, then nx * ny * nz values in site order
 * (x + nx * (y + ny * z)), 0 = solid; pts_in / pts_out columns are added before / after the file's
 * columns, fluid except the outermost column on each side.  nx / ny / nz (may be NULL) receive the box
 * size (with the added columns); meio (cap >= nx * ny * nz entries) receives read_geo's numbering: 0 for
 * a solid site, 1, 2, ... for the fluid sites in site order.  meio NULL only queries the size.  Returns
 * the number of fluid sites, or LBM_E_ARG (unreadable or malformed file, cap too small).  Host only. */
long long lbm_read_geometry(const char* path, int pts_in, int pts_out, int* nx, int* ny, int* nz, int* meio,
                            long long cap);

/* Solid sites of a case and the wall's wettability, for the FP64 form: meio holds nx * ny entries as
 * lbm_read_geometry gives them (0 = solid, > 0 = fluid); NULL (or no 0 entry) makes the box all fluid
 * again.  On the next steps, a population streaming into a solid site comes back in the opposite
 * direction at its own site (def_dir_prop LBM_functions.cpp:509-853, propag_site :1183-1253), a mediator
 * that would come from a solid site carries the wall's phase wett_R (propag_site_med :1309-1337), and
 * the force operators read a solid neighbour's field at the site itself (aloca_campos_F,
 * Operadores_tensao.cpp:149-182).  The reference's drop and band programs require an all-fluid box
 * (Imm_OPS.cpp:123-127); its library supports solids, and so does this step.  The post-stream state
 * is kept across the call; solid sites hold no populations (0 in downloads).  Such a case steps with
 * site-parallel kernels (lbm_step, lbm_step_case), not in its family's launch; lbm_diag_drop /
 * lbm_diag_incl refuse it (LBM_E_STATE), lbm_step_case_noalloc too (LBM_E_ARG), and other forms than
 * LBM_PREC_F64 (LBM_E_ARG). */
int lbm_set_geometry(lbm_ctx* ctx, int case_id, const int* meio, double wett_R);

/* VTK field files of the final state: rec_density / rec_velocity (LBM_functions.cpp:1347-1510) write
 * legacy structured-points files of one colour's density (field LBM_VTK_DENSITY, populations fR) or the
 * colour-weighted velocity with Guo's half-force correction (LBM_VTK_VELOCITY: fR, fB and the force
 * field [site][3], which may be NULL), 0 at solid sites (meio as above, may be NULL), ASCII with the
 * stream's default 6 digits.  flags LBM_VTK_BINARY writes the binary form the reference describes
 * (LBM_functions.cpp:1515-1570: same header with BINARY, raw big-endian values, float unless
 * LBM_VTK_DOUBLE).  Host arrays [site][19]; no context needed.  LBM_E_ARG for bad arguments or a file
 * that cannot be written. */
enum { LBM_VTK_DENSITY = 0, LBM_VTK_VELOCITY = 1 };
enum { LBM_VTK_BINARY = 1, LBM_VTK_DOUBLE = 2 };
int lbm_write_vtk(const char* path, int field, int nx, int ny, const double* fR, const double* fB,
                  const double* force, const int* meio, int flags);

/* Inclined-interface diagnostics (mede_inclinada) for normal (p, q); needs LBM_STEP_RECORD_AUX.
 * profile (optional, NULL to skip) receives nx*ny rows of (d, c_R, rho, u_x, u_y) in site order,
 * the values perfil.dat is written from (Imm_INCL.cpp:101,162). */
int lbm_diag_incl(lbm_ctx* ctx, int case_id, int p, int q, lbm_incl_diag* out, double* profile);

#ifdef __cplusplus
}
#endif
#endif
