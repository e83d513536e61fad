/*
 * erm.h - C ABI of the MI355X elastic RTM library (liberm.so).
 *
 * This is the drop-in boundary for the reference's hot path
 * (GeophysicsLab/2D_and_3D_elastic_reverse_time_migration, 3D_elastic_RTM/).
 * The reference has no FFI layer: its host drivers are `void f(GPUdevice*)`
 * functions reading ~40 mutable globals, and its kernels are positional
 * `__global__` launches.  Each entry point below names the reference
 * function or struct it replaces (file:line).
 *
 * Conventions
 *  - Plain C types only; the context is opaque.
 *  - Every call returns 0 on success or a negative ERM_E* code; the message
 *    of the last failure is returned by erm_last_error(ctx) (or
 *    erm_last_error(NULL) for failures before a context exists).
 *    Nothing in the library calls exit() (the reference exits from err()
 *    and checkCudaErrors).
 *  - Ownership: the caller owns every host array it passes; the context owns
 *    every device buffer.  A context is bound to one HIP stream and is not
 *    re-entrant; distinct contexts may be used from distinct threads.
 *  - Units follow the reference: dt in milliseconds, dx/dy/dz in metres,
 *    velocities in m/s, density in kg/m^3.
 *  - Memory layout of 3D volumes on the device and in erm_get_field():
 *    x fastest, idx = (iz*nny + iy)*nnx + ix over the padded grid
 *    (3D_elastic_RTM/3D_zzzzz:21).  Model and image FILES use the reference
 *    disk order [nx][ny][nz], z fastest (3D_elastic_RTM/3D_zzzzz:15-24).
 */
#ifndef ERM_H
#define ERM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ERM_PATH_MAX 256

/* error codes */
#define ERM_OK            0
#define ERM_E_PARAM      -1   /* missing/invalid parameter (reference: err("... is not set!")) */
#define ERM_E_IO         -2   /* file could not be opened/read/written */
#define ERM_E_DEVICE     -3   /* HIP runtime error */
#define ERM_E_STATE      -4   /* call out of order (e.g. backward before a record exists) */
#define ERM_E_SIZE       -5   /* buffer or grid size out of range */
#define ERM_E_NOMEM      -6   /* allocation failed */

/* boundary kinds */
#define ERM_BOUNDARY_DAMP 0   /* reference Crank-Nicolson damping, 1-cos profile (3D_zzzzz:261-369) */
#define ERM_BOUNDARY_CPML 1   /* convolutional PML (no reference counterpart) */

/*
 * Parameters.  The first block carries the reference's getpar keys
 * (3D_elastic_RTM/3D_elastic_modeling_parameter_sh:6-162) under the same
 * names; boundary_* keys are stored as bu/bd/bl/br/bf/bb like the reference
 * globals (3D_elastic_modeling_parameter:38).  The second block holds the
 * extensions this library adds.
 */
typedef struct erm_params {
    /* --- reference keys --- */
    int   RTM, GPU_switch, GPU_N, GPU_start, exchange_device_bool;
    int   lt;
    float dt;       /* ms */
    float freq;     /* Hz */
    int   join_vs, join_den, join_shot, join_wavefield, wavefield_interval;
    char  infile_shot_name[ERM_PATH_MAX];
    int   nx, ny, nz;
    float dx, dy, dz;
    int   shot_num_x, shot_num_y, shot_num_z;
    int   shot_start_x, shot_interval_x;
    int   shot_start_y, shot_interval_y;
    int   shot_start_z, shot_interval_z;
    int   receiver_num_x, receiver_num_y, receiver_num_z;
    int   receiver_start_x, receiver_interval_x;
    int   receiver_start_y, receiver_interval_y;
    int   receiver_start_z, receiver_interval_z;
    int   receiver_offset, receiver_mark;
    int   bu, bd, bl, br, bf, bb;
    float coe_attenuation;
    float velocity_multiple, smooth_time_vp, smooth_time_vs, smooth_time_density;
    char  velocity[ERM_PATH_MAX], velocity1[ERM_PATH_MAX], density[ERM_PATH_MAX];
    char  s_velocity[ERM_PATH_MAX], s_velocity1[ERM_PATH_MAX], s_density[ERM_PATH_MAX];
    int   add_receiver_bool, cut_direct_wave, laplace, laplace_compensate;
    int   vsp, vsp_precon, precon, precon_z1, precon_z2;
    /* --- extensions --- */
    int   fd_order;     /* 12: reference radius-6 optimized coefficients; 8: radius-4 Taylor */
    int   boundary;     /* ERM_BOUNDARY_DAMP | ERM_BOUNDARY_CPML */
    int   lt_forward;   /* steps of the source-side forward pass; reference: 2*lt/3
                           (3D_forward_or_back_together.cu:173) */
    float cpml_r;       /* CPML target reflection coefficient (default 1e-3) */
    float cpml_vmax;    /* CPML reference velocity in m/s (0: max vp of the model) */
    int   tile_x, tile_y, zchunk;   /* kernel launch shape: tile_x 32 (default at radius 4; tile_y 8 or 16)
                                       or 64 (tile_y 4 or 8; radius 6 always); the tiles cover the updated
                                       columns from x = R, line-aligned.  The kernels use
                                       32-bit offsets: (zchunk + 2R) x-y planes must span < 4 GiB, else
                                       erm_create fails (ERM_E_PARAM; ERM_E_SIZE if one plane is too big) */
    int   pitch_align;  /* row pitch multiple in floats (0: 32 = 128 B; 1: dense rows) */
    int   xcd_map;      /* workgroup -> tile order (results do not depend on it): 0 plain (x, y, z chunk);
                           1 each XCD a contiguous chunk of tiles; 2 (default) each XCD a band of grid
                           rows of the region-split tile lists, its tiles in (z chunk, y, x) order --
                           measured fastest with C-PML (19.8 vs 19.4 Gcell/s at 296x296x168, L2 misses
                           -40%); without the region split 2 acts as 0 */
    int   cpml_split;   /* 0 (default): C-PML steps launch the plain kernel on tiles without C-PML cells and
                           the C-PML kernel on the rest; 1: the C-PML kernel on every tile (same results) */
    int   region_streams; /* 1: the two region launches of a step run concurrently on two streams (they
                             write disjoint cells and neither writes what the other reads; same results);
                             0 (default, measured faster: 18.8 vs 18.3 Gcell/s): one stream */
    int   poynting;     /* 1: the receiver pass (erm_backward_smoothed, erm_rtm_step) also keeps, at each
                           cell's imaging step, the receiver P and S Poynting vectors (ERM_F_POYN_R*): the
                           cuda_cal_receiver_poyn_3D call of 3D_forward_or_back_together.cu:432, commented
                           out in the reference; one-shot contexts only (0 in batched windows) */
    int   reserved[3];
} erm_params;

/* field ids for erm_get_field / erm_set_field */
enum erm_field {
    ERM_F_VX = 0, ERM_F_VY, ERM_F_VZ,
    ERM_F_TXX, ERM_F_TYY, ERM_F_TZZ, ERM_F_TXY, ERM_F_TXZ, ERM_F_TYZ,
    ERM_F_TP, ERM_F_VXP, ERM_F_VYP, ERM_F_VZP,
    ERM_F_VXS, ERM_F_VYS, ERM_F_VZS,            /* derived: v - vp (read only) */
    ERM_F_EX_TIME, ERM_F_EX_AMP, ERM_F_EX_TP,    /* ex_time returned as float; -1 = never excited */
    ERM_F_EX_VXP, ERM_F_EX_VYP, ERM_F_EX_VZP,
    ERM_F_IMG_TP, ERM_F_IMG_PP, ERM_F_IMG_PS,
    /* P/S decomposition by divergence and curl of v (read only; the 3D form of
     * decom, 2D_elastic_RTM/elastic_2D_kernel_2.cu:563-669): div at the stress
     * node (i,j,k) = D+x vx/dx + D+y vy/dy + D-z vz/dz; curl_x = D-y vz/dy - D+z vy/dz
     * at the syz node, curl_y = D+z vx/dz - D-x vz/dx at the sxz node,
     * curl_z = D-x vy/dx - D-y vx/dy at the sxy node; 0 on the outer R cells. */
    ERM_F_DIV, ERM_F_CURLX, ERM_F_CURLY, ERM_F_CURLZ,
    /* Poynting vectors (read only; the bodies of cuda_cal_source_poyn_3D and
     * cuda_cal_receiver_poyn_3D, whose calls are commented out at
     * 3D_forward_or_back_together.cu:283,432, are absent; this is the 3D form of
     * sum_poynting, 2D_elastic_RTM/elastic_2D_kernel_2.cu:8-51):
     * source P at the excitation step, -ex_tp ex_vP (derived from the excitation
     * state); receiver P and S at the imaging step (poynting = 1), -tp vP and
     * -(sigma - tp I) vS with vS = v - vP, each product at the cell's own index;
     * 0 where the step was not reached. */
    ERM_F_POYN_SX, ERM_F_POYN_SY, ERM_F_POYN_SZ,
    ERM_F_POYN_RPX, ERM_F_POYN_RPY, ERM_F_POYN_RPZ,
    ERM_F_POYN_RSX, ERM_F_POYN_RSY, ERM_F_POYN_RSZ,
    ERM_F_COUNT
};

/* passes for erm_rtm_step */
#define ERM_PASS_MODEL    0   /* true model, source on, receivers recorded   (3D_forward_or_back_together.cu:9-155)   */
#define ERM_PASS_SOURCE   1   /* smoothed model, source on, excitation kept  (3D_forward_or_back_together.cu:161-290) */
#define ERM_PASS_RECEIVER 2   /* smoothed model, record injected, imaging    (3D_forward_or_back_together.cu:296-453) */

typedef struct erm_ctx erm_ctx;

/* Library-level */
const char* erm_last_error(const erm_ctx* ctx);
int   erm_version(void);
size_t erm_params_size(void);           /* sizeof(erm_params), for binding checks */
int   erm_device_count(int* n);

/* Parameters.
 * erm_params_default fills the extension block and zeroes the rest.
 * erm_parse_par ≙ elastic_modeling_parameter_input()
 *   (3D_elastic_modeling_parameter_sh:2-164): `key=value` arguments and
 *   `par=file` indirection, same keys, same "is not set" failures (returned,
 *   not exited).  erm_derive checks the grid and fills nothing else. */
int   erm_params_default(erm_params* p);
int   erm_parse_par(int argc, const char* const* argv, erm_params* p);

/* Context ≙ GPUdevice + allocate_typedef_struct (3D_elastic_modeling_typedef_struct:1-210). */
int   erm_create(const erm_params* p, int device, erm_ctx** out);
int   erm_destroy(erm_ctx* ctx);
int   erm_get_dims(const erm_ctx* ctx, int* nnx, int* nny, int* nnz, int* wavelet_length);

/* Models.  which: 0 = true model (velocity/velocity1/density),
 * 1 = smoothed model (s_velocity/s_velocity1/s_density).
 * Arrays hold the MODEL region in reference disk order [nx][ny][nz]
 * (z fastest); the library pads them by edge replication
 * (add_pml_layers_v_h, 3D_zzzzz:141-259).  vs == NULL derives vs = vp/sqrt(3)
 * (vp_set_vs, zzzzz:1261-1270); rho == NULL derives rho = 230*vp^0.25
 * (vp_set_density, zzzzz:1272-1287). */
int   erm_load_model(erm_ctx* ctx, int which, const float* vp, const float* vs, const float* rho);
/* Same, reading the reference's raw-float32 files named in the params. */
int   erm_load_model_files(erm_ctx* ctx);

/* Shot selection: sx_real = shot_start_x + isx*shot_interval_x, etc. */
int   erm_set_shot(erm_ctx* ctx, int isx, int isy, int isz);
/* set_zero_typedef_struct_wavefield / _excitation (typedef_struct:212-293) */
int   erm_zero_wavefield(erm_ctx* ctx);
int   erm_zero_excitation(erm_ctx* ctx);

/* The three time loops.
 * rec layout: [3][lt][receiver_num_y][receiver_num_x] (vx, vy, vz).
 * erm_forward_real  ≙ forward_together_using_real_model     (3D_forward_or_back_together.cu:9-155);
 *                     rec may be NULL (the record stays on the device).  With
 *                     cut_direct_wave > 0 and vsp == 0 the direct wave is muted
 *                     (cut_direct_shot_3D_surface_three_new, :133-136, body not in
 *                     the reference): at each receiver the samples
 *                     it < ceil(offset / (vp_source dt)) + wavelet length +
 *                     cut_direct_wave are zeroed, offset = source-receiver distance
 *                     in metres, vp_source = the true model's vp at the source cell.
 * erm_forward_smoothed ≙ forward_together_using_smoothed_model (:161-290).
 * erm_backward_smoothed ≙ backward_together_using_smoothed_model (:296-453);
 *                     rec == NULL uses the device-resident record. */
int   erm_forward_real(erm_ctx* ctx, float* rec);
int   erm_forward_smoothed(erm_ctx* ctx);
int   erm_backward_smoothed(erm_ctx* ctx, const float* rec);
/* Upload / download the receiver record (join_shot ingest, parameter_sh:30,33). */
int   erm_set_record(erm_ctx* ctx, const float* rec);
int   erm_get_record(erm_ctx* ctx, float* rec);

/* Kernel-level stepping for parity tests: run `nsteps` iterations of `pass`
 * starting at time index it_begin (the receiver pass counts down). */
int   erm_rtm_step(erm_ctx* ctx, int pass, int it_begin, int nsteps);
/* Padded-grid field access (x fastest, nnx*nny*nnz floats). */
int   erm_get_field(erm_ctx* ctx, int field, float* dst);
int   erm_set_field(erm_ctx* ctx, int field, const float* src);

/* Images over the model region, reference disk order [nx][ny][nz]. */
int   erm_get_images(erm_ctx* ctx, float* tp, float* pp, float* ps);
/* output_3d_result (3D_output_file.cu:309-428): writes
 * dir/result-{tp,pp,ps}-{sy_real}-{sx_real} (vsp: -{sz_real}). */
int   erm_write_results(erm_ctx* ctx, const char* dir);
/* output_3d_wavefiled_excitation_amp_time (3D_output_file.cu:155-250):
 * dir/cut-ex-{time,amp,tp,vxp,vyp,vzp}-{sy}-{sx}. */
int   erm_write_excitation(erm_ctx* ctx, const char* dir);

/* Timing of the last pass: wall time (HIP events on the context stream) and,
 * when profiling is on, the summed durations and launch counts of the
 * stress and velocity kernels measured with events around each launch. */
int   erm_set_profiling(erm_ctx* ctx, int on);
int   erm_get_timing(erm_ctx* ctx, double* pass_ms, double* stress_ms, long* stress_launches,
                     double* velocity_ms, long* velocity_launches);
/* Algorithmic HBM bytes per cell-update of the stress and velocity kernels
 * for the given pass, as the kernels are built (see the kernel sources). */
int   erm_algorithmic_bytes(const erm_ctx* ctx, int pass, double* stress_bytes, double* velocity_bytes);
int   erm_synchronize(erm_ctx* ctx);
/* Source-wavefield reconstruction from boundary strips (no reference
 * counterpart; the reference images with the excitation amplitude only, and
 * its 2D code stores full histories, elastic_2D_kernel_2.cu:793-827).
 * erm_forward_smoothed_strips runs the source pass like erm_forward_smoothed
 * (excitation kept) and also saves, every step, sigma+tp and v+vP on the
 * R-cell face slabs around the model region shrunk by one cell (whose
 * update is free of damping and C-PML) plus the final wavefield, in HBM
 * (erm_strip_bytes tells how much: 52 B x slab cells x (lt+1)).
 * erm_backward_smoothed_xcorr then runs the receiver pass (excitation imaging
 * as erm_backward_smoothed) while reconstructing the source wavefield
 * backward in time in the model region (the step inverted with dt negated,
 * the slabs restored from the strips) and accumulates the zero-lag
 * cross-correlation images I_tp = sum_t tp_s tp_r, I_PP = sum_t vP_s.vP_r,
 * I_PS = sum_t vP_s.(v_r - vP_r), returned by erm_get_xcorr_images over the
 * model region in disk order.  Requires lt_forward == lt and every boundary
 * width >= the stencil radius - 1.  erm_reconstruct / erm_get_reconstructed_field
 * expose the reconstruction itself (step the saved final state back to the
 * start of step it_end; fields of the padded grid as erm_get_field). */
int   erm_strip_bytes(const erm_ctx* ctx, double* bytes);
int   erm_forward_smoothed_strips(erm_ctx* ctx);
int   erm_backward_smoothed_xcorr(erm_ctx* ctx, const float* rec);
int   erm_get_xcorr_images(erm_ctx* ctx, float* tp, float* pp, float* ps);
/* Cross-correlation imaging with a decimated source history (the reference's
 * full-history option, save_all_wavefield / set_all_wavefield,
 * 2D_elastic_RTM/elastic_2D_kernel_2.cu:793-827, in 3D and thinned in time):
 * erm_forward_smoothed_history runs the source pass (excitation kept) and
 * stores tp, vxp, vyp, vzp of the model region after every step it with
 * (it + 1) % decim == 0, in HBM (erm_history_bytes: 16 B x nx*ny*nz x
 * floor(lt / decim)); erm_backward_smoothed_xcorr_history runs the receiver
 * pass (excitation imaging as erm_backward_smoothed) and adds decim times the
 * zero-lag products with the stored steps to the images of erm_get_xcorr_images
 * (decim = 1: every step, the images of erm_backward_smoothed_xcorr).  No
 * reconstruction, so no boundary-width condition; lt_forward == lt. */
int   erm_history_bytes(const erm_ctx* ctx, int decim, double* bytes);
int   erm_forward_smoothed_history(erm_ctx* ctx, int decim);
int   erm_backward_smoothed_xcorr_history(erm_ctx* ctx, const float* rec);
int   erm_reconstruct(erm_ctx* ctx, int it_end);
int   erm_get_reconstructed_field(erm_ctx* ctx, int field, float* dst);
/* Launch-shape autotuner (no reference counterpart; the reference launches
 * one block per z plane): times `steps` source-pass steps (smoothed model)
 * for tiles 64 x {4, 8} (and 32 x {8, 16} at radius 4) x z sweep depths
 * {4, 5, 6, 7, 8, 12, 16, 32} and keeps the fastest for later passes (depths
 * whose sweep spans 4 GiB or more are skipped).  Results do not depend on
 * the shape.  Zeroes the wavefield and the excitation state afterwards.
 * erm_autotune reports the chosen tile height and depth only. */
int   erm_autotune_shape(erm_ctx* ctx, int steps, int* tile_x, int* tile_y, int* zchunk, double* ms_per_step);
int   erm_autotune(erm_ctx* ctx, int steps, int* tile_y, int* zchunk, double* ms_per_step);
/* Measurement utility (no reference counterpart): time `reps` launches of a
 * plain float4 grid-stride copy of `bytes` bytes on `device` for each grid of
 * 1024, 2048, .. 16384 workgroups and return the best time per copy; it is
 * the calibration kernel for HBM byte counters (every launch moves the same
 * bytes) and the practical bandwidth ceiling (2*bytes of traffic per copy). */
int   erm_copy_bandwidth(int device, size_t bytes, int reps, double* ms_per_copy);

/* Synthetic models (seeded; std::mt19937_64).  Arrays are model-region,
 * disk order [nx][ny][nz].  kind 0: flat layers + Gaussian anomaly,
 * kind 1: dipping layers + random Gaussian anomalies (see erm_model.cpp). */
int   erm_make_model3d(int kind, int nx, int ny, int nz, float dx, float dy, float dz,
                       uint64_t seed, float* vp, float* vs, float* rho);
/* Reference raw-float32 I/O in disk order (3D_zzzzz:3-100). */
int   erm_write_raw(const char* path, const float* data, size_t n);
int   erm_read_raw(const char* path, float* data, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* ERM_H */
