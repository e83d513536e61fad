/*
 * erm2d.h - C ABI of the 2D elastic / viscoelastic part of liberm.so:
 * 2D propagation with standard-linear-solid attenuation, P/S decomposition,
 * excitation-amplitude RTM, and the Born / adjoint operators of LSRTM with a
 * host CG driver.
 *
 * Reference counterparts (GeophysicsLab/2D_and_3D_elastic_reverse_time_migration,
 * 2D_elastic_RTM/): the reference has kernels but no host driver here; each
 * entry point names the kernels it drives.  Conventions as in erm.h (return
 * codes, erm2d_last_error, caller-owned host arrays, dt in ms).
 *
 * Layout: 2D volumes are z-fastest, idx = ix*nnz + iz, as the reference
 * (viscoelastic_equation.cu:90-92); model/image arrays cover the model region
 * [nx][nz].  Records are [2][lt][nrec] (vx, vz).
 */
#ifndef ERM2D_H
#define ERM2D_H
#include <stddef.h>
#include "erm.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct erm2d_params {
    int   nx, nz;
    float dx, dz;
    float dt;          /* ms */
    float freq;        /* Ricker peak frequency (Hz); also the SLS reference frequency */
    int   lt;          /* time steps */
    int   lt_forward;  /* source-pass steps for RTM (<0: lt) */
    int   bu, bd, bl, br;
    float coe_attenuation;  /* damping (cubic profile, zzzzz:238-325) */
    int   fd_order;    /* 8 (radius 4) or 12 (radius 6, reference coefficients) */
    int   nsls;        /* 0: elastic; 1: reference SLS (viscoelastic_equation.cu:22-58); 3: three mechanisms */
    int   shot_x, shot_z;              /* source, model coordinates (cells) */
    int   nrec, rec_x0, rec_dx, rec_z; /* receiver line, model coordinates */
    int   qcomp;       /* 1: Q-compensated receiver pass of erm2d_rtm (memory-term sign flipped,
                          rfwd_/receiver_fwd_ kernels, viscoelastic_equation.cu:1438-1822) */
    int   ex_tp_shift; /* the tp image's excitation time is recorded as it - ex_tp_shift
                          (caculate_ex_tp_time_new; its commented launch passes it - wavelet_half,
                          elastic_2D_kernel_3.cu:51); default 0 */
    int   ex_vzp_avg;  /* 1: the _new excitation/imaging variants (caculate_ex_x_z_new,
                          imaging_inner_product_ex_2D_new, elastic_2D_kernel_3.cu:91-109,582-605):
                          vzp averaged over its four neighbours (i, k +- 1), (i +- 1, k) in the
                          excitation components and in the receiver P and S motion; 0: own cell */
    int   boundary;    /* ERM_BOUNDARY_DAMP (default: the reference's cubic damping) or ERM_BOUNDARY_CPML
                          (convolutional PML, no reference counterpart; modeling and RTM only -- the
                          LSRTM operators need the damping boundary) */
    float cpml_r;      /* CPML target reflection coefficient (0: 1e-3) */
    float cpml_vmax;   /* CPML reference velocity in m/s (0: max vp of the loaded models) */
    int   inversion_para; /* LSRTM model parameters, relative perturbations (inversion_para of
                          cuda_cal_dem_parameter_viscoelastic_media_new, viscoelastic_equation.cu:2431-2486):
                          1 = (dlambda/lambda, dmu/mu, drho/rho); 2 (or 0, default) = (dvp/vp, dvs/vs,
                          drho/rho); 3 = impedances (dIp/Ip, dIs/Is, drho/rho), Ip = rho vp, Is = rho vs */
    int   adjoint_mode;   /* LSRTM adjoint of erm2d_adjoint / erm2d_adjoint_shots / erm2d_lsrtm_cg:
                          0 (default) = exact transpose of the discrete Born operator;
                          1 = the reference's continuous-adjoint equations and gradient
                          (erm2d_adjoint_ref), mapped to the inversion_para parameters */
    int   adjoint_source; /* background wavefield of the exact-transpose adjoint (adjoint_mode 0):
                          ERM2D_SRC_HISTORY (0, default) = the strain increments E and forces F of
                          every step stored in HBM ([lt][5][nx*nz] floats per shot, as the
                          reference's save_and_set_wavefiled history, viscoelastic_equation.cu:3-21);
                          ERM2D_SRC_STRIPS (2) = only a ring of width R + 1 around the model region
                          stored per half step ([lt][8][ring] floats per shot) and the background
                          stepped backward, E and F recomputed on the way (elastic media, nsls = 0;
                          equal to the history gradient up to FP32 rounding of the reversal) */
    int   tile_rows;      /* x rows per workgroup of the stencil launches (64 z lanes each): 0 (default) =
                          8 when a launch of 8-row workgroups has at least 4096 of them (batched shots;
                          measured 12.7-13.0 against 13.4 s per CG iteration at 600x300, 12 shots per launch),
                          else 4 (a single small shot then still fills the GPU); or 4, 8.  Results do
                          not depend on it */
} erm2d_params;

/* fields for erm2d_get_field / erm2d_set_field (padded grid, z fastest) */
enum erm2d_field {
    ERM2D_VX = 0, ERM2D_VZ, ERM2D_TXX, ERM2D_TZZ, ERM2D_TXZ, ERM2D_TP, ERM2D_VXP, ERM2D_VZP,
    ERM2D_DIV, ERM2D_CURL,              /* decom (elastic_2D_kernel_2.cu:563-669), read only */
    ERM2D_EX_TIME, ERM2D_EX_AMP, ERM2D_IMG_TP, ERM2D_IMG_PP, ERM2D_IMG_PS,
    ERM2D_EX_TP, ERM2D_EX_VXP, ERM2D_EX_VZP,  /* excitation state, read only */
    ERM2D_R_VXP, ERM2D_R_VZP,                /* receiver vector P at each cell's imaging step, read only */
    ERM2D_EX_TP_TIME, ERM2D_EX_TP_AMP,       /* tp excitation: step of max |tp| and the signed tp there
                                                (caculate_ex_tp_time_new, elastic_2D_kernel_3.cu:50-69) */
    ERM2D_PDIV, ERM2D_SCURL,                 /* decom_new (elastic_2D_kernel_2.cu:671-791): vp * div at the
                                                stress node, vs4 * curl at the sxz node, vs4 the mean of vs
                                                at (i,k), (i,k+1), (i+1,k), (i+1,k+1); read only */
    ERM2D_FIELD_COUNT
};
/* passes for erm2d_step */
#define ERM2D_PASS_MODEL    0
#define ERM2D_PASS_SOURCE   1
#define ERM2D_PASS_RECEIVER 2

typedef struct erm2d_ctx erm2d_ctx;

const char* erm2d_last_error(const erm2d_ctx* ctx);
size_t erm2d_params_size(void);
int   erm2d_params_default(erm2d_params* p);

/* create ≙ the reference's device allocations; which: 0 = true, 1 = background (smoothed) model.
 * cuda_cal_viscoelastic (viscoelastic_equation.cu:22-58) runs inside erm2d_load_model.
 * qp/qs may be NULL when nsls == 0. */
int   erm2d_create(const erm2d_params* p, int device, erm2d_ctx** out);
int   erm2d_destroy(erm2d_ctx* ctx);
int   erm2d_load_model(erm2d_ctx* ctx, int which, const float* vp, const float* vs, const float* rho,
                       const float* qp, const float* qs);
int   erm2d_set_shot(erm2d_ctx* ctx, int shot_x, int shot_z);
int   erm2d_zero(erm2d_ctx* ctx);

/* Kernel-level stepping (parity tests) and field access. */
int   erm2d_step(erm2d_ctx* ctx, int pass, int it_begin, int nsteps);
int   erm2d_get_field(erm2d_ctx* ctx, int field, float* dst);
int   erm2d_set_field(erm2d_ctx* ctx, int field, const float* src);
int   erm2d_get_dims(const erm2d_ctx* ctx, int* nnx, int* nnz);

/* Modeling in the true model (fwd_txxzzxzpp_viscoelastic_and_memory_3parameterization,
 * viscoelastic_equation.cu:184-325, + the velocity update): rec [2][lt][nrec]. */
int   erm2d_forward(erm2d_ctx* ctx, float* rec);
/* Excitation-amplitude RTM in the background model: source pass
 * (caculate_ex_amp_time_new, elastic_2D_kernel_3.cu:30-48, and caculate_ex_tp_time_new,
 * :50-69) then receiver pass with imaging_correlation_ex_2D / imaging_inner_product_ex_2D
 * (:542-580): I_PP = ex_vP.vP_r/|ex_vP|^2 and I_PS = ex_vP.vS_r/|ex_vP|^2 at the step of
 * max |vP|; I_tp = tp_r / ex_tp_amp at the step of max |tp| (its own excitation time and
 * the signed tp there).  A zero denominator images 0.
 * Images over the model region [nx][nz]; any pointer may be NULL. */
int   erm2d_rtm(erm2d_ctx* ctx, const float* rec, float* img_tp, float* img_pp, float* img_ps);

/* Cross-correlation RTM in the background model with a choice of
 * source-wavefield strategy (the reference's 2D options: full history,
 * save_all_wavefield / save_and_set_wavefiled, elastic_2D_kernel_2.cu:793-827,
 * viscoelastic_equation.cu:3-21; reversed propagation,
 * rfwd_txxzzxzpp_viscoelastic_and_memory_old, :1298-1435) and the reference's
 * PS polarity correction (set_sign_basedon_polarization_ps, filter_sign_new_share,
 * compare_sign, elastic_2D_kernel_2.cu:875-982):
 *   strategy ERM2D_SRC_HISTORY  the source pass stores vxp, vzp, div v of the model
 *                               region every step (lt_forward * 3 * nx * nz floats)
 *            ERM2D_SRC_REVERSE  the final source state is stepped backward (derivative
 *                               and memory terms negated, memory updated forward-like);
 *                               approximate: the absorbing boundary keeps absorbing
 *            ERM2D_SRC_STRIPS   as REVERSE with the fields of a ring of width R + 1 around
 *                               the model region (shrunk by one cell) stored every step and
 *                               restored before each reversed half step: exact for
 *                               elastic media (nsls = 0 only; no reference counterpart)
 * The receiver pass images after its step it with the source wavefield after forward
 * step it, for it < lt_forward; the source fields enter over the model region only.
 * images [5][nx][nz] (may be NULL): PP = sum vP_s.vP_r, PS = sum vP_s.(v_r - vP_r),
 * PPd = sum div_s div_r, PSc = sum div_s curl_r (decom stencils), and PSc with the
 * polarity correction: each step every interior cell votes +1 / -1 / 0 by the sign of
 * -sgn(vzp_s) sgn(rvxs) (vP_s . rvs), the 9x9 window majority sets the cell's sign
 * (+1 wins ties with either class; a -1 / 0 tie keeps the previous sign), and
 * PScp += sign div_s curl_r. */
#define ERM2D_SRC_HISTORY 0
#define ERM2D_SRC_REVERSE 1
#define ERM2D_SRC_STRIPS  2
int   erm2d_rtm_xcorr(erm2d_ctx* ctx, const float* rec, int strategy, float* images);

/* LSRTM operators in the background model, model perturbation dm = [3][nx][nz]
 * of the inversion_para parameters (default 2: dvp/vp, dvs/vs, drho/rho;
 * viscoelastic_equation.cu:2431-2489):
 *   erm2d_born    d = L dm   (demig_fwd_... :328-477 with cuda_cal_dem_parameter_..._new :2398-2505)
 *   erm2d_adjoint g = L^T r  (adjoint_fwd_vx/vz/memory :1827-2198, cal_gradient_..._new :2248-2295),
 * where L^T is the exact transpose of the discrete L (dot-product exact). */
int   erm2d_born(erm2d_ctx* ctx, const float* dm, float* d);
int   erm2d_adjoint(erm2d_ctx* ctx, const float* r, float* g);
/* The reference's own adjoint-state operator and gradient (the continuous
 * adjoint; adjoint_fwd_vx_viscoelastic / adjoint_fwd_vz_viscoelastic,
 * viscoelastic_equation.cu:1827-2122, adjoint_fwd_memory_new :2158-2198,
 * cal_gradient_in_viscoelastic_media_new :2248-2295) in the background model.
 * Adjoint state: velocities, strain-like stresses (updated by the strain
 * increments of the adjoint velocities: the reference does not ship that
 * kernel) and per-mechanism memory variables; the residual r [2][lt][nrec] is
 * added to the adjoint velocities at the receivers before each reversed step,
 * and the gradient of every step is accumulated after it from the stored
 * background strain rates and accelerations.
 *   g [3][nx][nz] = (g_lambda, g_mu, g_rho), the raw sums of the reference's
 *   gradient kernel for absolute perturbations.
 * Measured against the Born operator (tests/test_oracle2d.py): for
 * perturbations of lambda or mu alone <L dm, L dm> = (rho_r / dt) <dm, g>
 * to 0.3%, rho_r the density at the receivers; for rho alone the sign is
 * reversed (the reference adds rv . v_t where its Born source -drho v_t needs
 * the opposite sign).  adjoint_mode = 1 turns this into an approximate
 * transpose for erm2d_adjoint / erm2d_adjoint_shots / erm2d_lsrtm_cg: the
 * residual is scaled by dt / rho at each receiver, the density term negated,
 * and (g_lambda, g_mu, g_rho) mapped to the inversion_para parameters by the
 * transpose of the reference's parameter map (:2431-2486). */
int   erm2d_adjoint_ref(erm2d_ctx* ctx, const float* r, float* g);

/* Batched shots (no reference counterpart; the reference runs one shot per
 * launch): erm2d_set_batch(ctx, nb) gives the context per-shot states and
 * E/F histories for nb shots (nb * (lt*5*nx*nz + ~30*padded cells) floats of
 * device memory; 1 frees them), after which erm2d_born_shots /
 * erm2d_adjoint_shots run the shots in groups of nb, each time step of a
 * group one launch per kernel with the shots along the grid's z dimension,
 * and erm2d_lsrtm_cg uses them in place of the workers.  Results equal
 * erm2d_born / erm2d_adjoint shot by shot (records bit-identical; the
 * gradient is the sum over shots, in double on the host across groups).
 *   shots  [nshots][2] source (x, z) cells of the model region
 *   d      [nshots][2][lt][nrec] Born records (output)
 *   r      [nshots][2][lt][nrec] residual records; g [3][nx][nz] output */
int   erm2d_set_batch(erm2d_ctx* ctx, int nb);
int   erm2d_born_shots(erm2d_ctx* ctx, int nshots, const int* shots, const float* dm, float* d);
int   erm2d_adjoint_shots(erm2d_ctx* ctx, int nshots, const int* shots, const float* r, float* g);
/* Least-squares RTM by CGLS (conjugate gradients on the normal equations,
 * Paige & Saunders 1982) over nshots shots: minimise sum_s |L_s m - d_s|^2.
 * With a preconditioner the flexible form is used: each preconditioned
 * gradient is made L^T L-orthogonal to the last 4 search directions and the
 * step is the exact line minimiser, so the misfit never increases even for a
 * nonlinear (learned) preconditioner; with the identity it equals CGLS.
 * The reference ships the operators but no CG loop.
 *   shots   [nshots][2]  source (x, z), model coordinates
 *   dobs    [nshots][2][lt][nrec] observed data
 *   pmask   bit mask of inverted parameters (1: dvp/vp, 2: dvs/vs, 4: drho/rho)
 *   m       [3][nx][nz] initial model in, solution out
 *   misfit  [iters+1]   |L m - d| / |d| before iteration 1 and after each iteration
 *   seconds [iters+1]   wall time at the same points (may be NULL)
 *   precond optional z = P(g) over the model vector (NULL: identity).
 * Returns the number of iterations run (stops early when misfit <= tol), or
 * a negative error code. */
typedef int (*erm2d_precond_fn)(const float* g, float* z, void* user);
int   erm2d_lsrtm_cg(erm2d_ctx* ctx, int nshots, const int* shots, const float* dobs, int iters, unsigned pmask,
                     float* m, float* misfit, double* seconds, erm2d_precond_fn precond, void* user, float tol);

/* Angle-domain common-image gathers and PS polarity correction of the last
 * erm2d_rtm (excitation imaging).  The reference has the building blocks
 * (Poynting vectors, least-squares angles, opening angles, angle bins:
 * elastic_2D_kernel_2.cu:8-150,1236-1819, elastic_2D_kernel_3.cu:111-364) but
 * no driver calls them; here, per model cell with an excitation time:
 * source incidence ts = atan(sum Sx Sz / sum Sz^2) over a 5x5 window of the
 * excitation-time P Poynting vector S = -tp vP (degrees from the vertical,
 * +ve toward +x), receiver direction tr = atan(sum rvxp rvzp / sum rvzp^2) of
 * the receiver P motion at the imaging step over the same window, reflection
 * angle g = (ts - tr) / 2 (reliable below the critical angle).  Gathers
 * pp, ps [nang][nx][nz]: bin floor((g + nang dang / 2) / dang) receives the
 * cell's PP / PS image value.  pp_comp, ps_comp [nx][nz]: the reference's
 * angle-dependent compensation with the opening angle 2g
 * (elastic_2D_kernel_3.cu:698-770): PP * s(cos 2g) / (|cos 2g| + 0.03),
 * s = +1 for a positive argument else -1, and PS / (|sin 2g| + 0.1) (the
 * reference also negates both; here the images keep their sign).  The PS
 * image of the vector inner-product condition does not reverse polarity
 * across the source, so it needs no sign correction.  gamma [nx][nz] = g.
 * Any output may be NULL; nang * dang <= 180. */
int   erm2d_angle_gathers(erm2d_ctx* ctx, int nang, float dang, float* pp, float* ps, float* pp_comp,
                          float* ps_comp, float* gamma);
/* erm2d_angle_gathers without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as erm2d_angle_gathers; the device staging of the
 * outputs is a fixed per-GPU array of 2^23 floats (32 MiB, held until the
 * process ends), so 2 nang nx nz + 3 nx nz must not exceed 8388608
 * (ERM_E_SIZE otherwise).  The array is shared by every call on that GPU and
 * keeps the last call's contents (each call clears what it uses): calls that
 * overlap on one GPU, from different contexts or threads, race. */
int   erm2d_angle_gathers_noalloc(erm2d_ctx* ctx, int nang, float dang, float* pp, float* ps, float* pp_comp,
                                  float* ps_comp, float* gamma);
/* Source illumination of the background model: for each shot, the sum over
 * the time steps it = 0 .. lt-1 of vx^2 + vz^2 of the source wavefield
 * (after the step's velocity update, each component at its own staggered
 * point), summed over the shots; out [nx][nz] over the model region (z
 * fastest).  The reference only stacks per-shot illumination images
 * (cuda_sum_new_acqusition_illum, viscoelastic_equation.cu:2736-2754); here it
 * is the diagonal, physics-side part of the LSRTM preconditioner. */
int   erm2d_illumination(erm2d_ctx* ctx, int nshots, const int* shots, float* out);
/* Diagonal preconditioner for erm2d_lsrtm_cg (precond = erm2d_precond_diag,
 * user = erm2d_diag_precond*): z = w * g on each of the three parameter
 * planes, w [nx][nz] (e.g. 1 / (illumination + eps max illumination)). */
typedef struct erm2d_diag_precond {
    int nx, nz;
    const float* w;
} erm2d_diag_precond;
int   erm2d_precond_diag(const float* g, float* z, void* user);

/* Shot concurrency for erm2d_lsrtm_cg: n contexts in total (this one plus
 * n-1 workers created here with the same parameters and background model),
 * each with its own HIP stream and host thread, take the shots round-robin so
 * that shots too small to fill the GPU run side by side.  n = 1 (default)
 * runs the shots one after another.  Not part of the reference, which runs
 * one shot per GPU. */
int   erm2d_set_workers(erm2d_ctx* ctx, int n);
/* Optional per-iteration callback of erm2d_lsrtm_cg, called with the current
 * model m [3][nx][nz], the (masked) gradient g = L^T (d - L m) at that model
 * and the misfit, before iteration 1 (it = 0) and after each iteration that
 * is followed by another; a non-zero return aborts the solve.  Used to record
 * (gradient, remaining update) pairs for training the learned
 * preconditioner. */
typedef int (*erm2d_cg_monitor_fn)(int it, const float* m, const float* g, float misfit, void* user);
int   erm2d_set_cg_monitor(erm2d_ctx* ctx, erm2d_cg_monitor_fn fn, void* user);

/* Windowed multi-shot RTM over a full model, the reference's per-shot
 * aperture (cuda_get_partly_mode[_boundary][_z1_z2] and
 * cuda_sum_new_acqusition[_illum], viscoelastic_equation.cu:2606-2754): for
 * each shot the window [x0, x0 + width) x full depth is cut from the full
 * models (padded by edge replication like erm2d_load_model), the shot is
 * modelled in the true window model (or its record taken from `records`),
 * migrated in the background window model (erm2d_rtm: excitation imaging),
 * and the window's tp / PP / PS images and source illumination (as
 * erm2d_illumination) are added into the full-model outputs at column x0.
 * Receivers of a window: every p->rec_dx cells from its left edge at depth
 * p->rec_z, nrec = width / rec_dx (p->nx, p->nrec, p->rec_x0 and p->shot_* are
 * replaced per window; p->nx is the full model's width).  z2 >= 0 replaces the
 * background window model below depth z2 by its value at z2 (the
 * _boundary_z1_z2 cut); z2 < 0 keeps it.
 *   vp, vs, rho, qp, qs     true model [nx][nz] (qp, qs only when nsls > 0)
 *   bvp, bvs, brho, bqp, bqs background model [nx][nz]
 *   records  NULL (model every shot) or the shots' records one after another,
 *            shot s: [2][lt][width_s / rec_dx]
 *   img_tp, img_pp, img_ps, illum  [nx][nz] outputs (any may be NULL), overwritten */
typedef struct erm2d_window {
    int shot_x, shot_z;  /* source, full-model cell coordinates */
    int x0, width;       /* window origin (receiver_x_cord) and width (receiver_interval x receiver_num) */
} erm2d_window;
int   erm2d_rtm_windows(const erm2d_params* p, int device, const float* vp, const float* vs, const float* rho,
                        const float* qp, const float* qs, const float* bvp, const float* bvs, const float* brho,
                        const float* bqp, const float* bqs, int nshots, const erm2d_window* windows,
                        const float* records, int z2, float* img_tp, float* img_pp, float* img_ps, float* illum);

/* Timing of the last operator call (HIP events): milliseconds. */
int   erm2d_last_ms(const erm2d_ctx* ctx, double* ms);

#ifdef __cplusplus
}
#endif
#endif
