/*
 * erm_batch.h - batched multi-shot 3D elastic RTM in liberm.so.
 *
 * The reference migrates one shot per GPU pass, each over the whole grid or
 * over a per-shot window (cuda_get_partly_mode[_boundary], window extraction
 * at the receiver aperture, and cuda_sum_new_acqusition, the image stack:
 * 2D_elastic_RTM/viscoelastic_equation.cu:2606-2754).  Here many shots, each
 * with its own lateral window of any size (full depth), advance together:
 * every time step is ONE launch of the stress kernel and ONE of the velocity
 * kernel over the tiles of all windows (a block table maps workgroup ->
 * (shot, tile); the grid covers the table once or is persistent and strides
 * over it), plus one receiver launch; the step is captured in a hipGraph and
 * replayed (the step index lives in device memory).  Tile height and z sweep
 * depth are picked by a measured autotuner.  Images of the windows are
 * stacked into the full-model image.
 *
 * Each window is a full erm_ctx (include/erm.h) of size wx x wy x nz with the
 * boundary widths of the batch parameters, its model cut from the full model
 * and padded like erm_load_model, its source at (sx - x0, sy - y0, sz) and
 * receivers on the window surface every receiver_interval_x/_y cells from
 * the window origin at depth receiver_start_z.  A batched run produces, per
 * window, exactly the images erm_forward_real / erm_forward_smoothed /
 * erm_backward_smoothed produce on that context alone.
 *
 * Conventions as in erm.h; errors via erm_batch_last_error.
 */
#ifndef ERM_BATCH_H
#define ERM_BATCH_H
#include "erm.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct erm_batch erm_batch;

typedef struct erm_shot_window {
    int sx, sy, sz;   /* source, full-model cell coordinates */
    int x0, y0;       /* window origin in the full model */
    int wx, wy;       /* window size (cells); the window spans the full depth */
} erm_shot_window;

const char* erm_batch_last_error(const erm_batch* b);
/* p: the full model (nx, ny, nz, spacing, dt, lt, freq, fd_order, boundary,
 * boundary widths, receiver_interval_x/_y, receiver_start_z, ...). */
int   erm_batch_create(const erm_params* p, int device, erm_batch** out);
int   erm_batch_destroy(erm_batch* b);
/* Full-model arrays in reference disk order [nx][ny][nz]; which: 0 true, 1 smoothed. */
int   erm_batch_load_model(erm_batch* b, int which, const float* vp, const float* vs, const float* rho);
/* (Re)define the shots; windows must lie inside the model and contain their source. */
int   erm_batch_set_shots(erm_batch* b, const erm_shot_window* shots, int nshots);
/* Launch shape: tiles 64 x {4, 8} or, at radius 4, 32 x {8, 16} (the default
 * 32 x 8, as erm_create), zchunk >= 1, grid_blocks 0 = one workgroup per
 * table entry, > 0 = persistent grid of that many workgroups.  A zchunk whose
 * sweep (zchunk + 2R planes of a window) spans 4 GiB or more makes the next
 * pass fail with ERM_E_PARAM (32-bit kernel offsets).  erm_batch_set_launch
 * sets a 64-wide tile. */
int   erm_batch_set_launch_shape(erm_batch* b, int tile_x, int tile_y, int zchunk, int grid_blocks);
int   erm_batch_set_launch(erm_batch* b, int tile_y, int zchunk, int grid_blocks);
/* Measure candidate launch shapes on `steps` source-pass steps of the current
 * shots and keep the fastest; returns it (any pointer may be NULL;
 * erm_batch_autotune does not report the tile width). */
int   erm_batch_autotune_shape(erm_batch* b, int steps, int* tile_x, int* tile_y, int* zchunk, int* grid_blocks,
                               double* ms_per_step);
int   erm_batch_autotune(erm_batch* b, int steps, int* tile_y, int* zchunk, int* grid_blocks, double* ms_per_step);
/* 1 (default): replay a captured hipGraph per step; 0: plain launches. */
int   erm_batch_use_graph(erm_batch* b, int on);
/* Observed data of shot s ([3][lt][nr_s], see erm_batch_shot_context for nr).
 * When no shot holds a record, erm_batch_run models them all in the true
 * model first and keeps them for later runs until the true model or the
 * shots change; records must be set for all shots or for none. */
int   erm_batch_set_record(erm_batch* b, int shot, const float* rec);
/* Migrate every shot: model pass (true model) when there are no records,
 * source pass with excitation, receiver pass with imaging; then stack.
 * Stacked images over the model region in disk order [nx][ny][nz]; any may be NULL. */
int   erm_batch_run(erm_batch* b, float* stacked_tp, float* stacked_pp, float* stacked_ps);
/* Optional progress report (no reference counterpart): fn(pass, ms, user) is
 * called from erm_batch_run after each completed pass (ERM_PASS_MODEL /
 * _SOURCE / _RECEIVER) with that pass's device time; NULL removes it. */
typedef void (*erm_batch_progress_fn)(int pass, double ms, void* user);
int   erm_batch_set_progress(erm_batch* b, erm_batch_progress_fn fn, void* user);
/* The context of shot s (owned by the batch): dims, fields, records, images. */
int   erm_batch_shot_context(erm_batch* b, int shot, erm_ctx** ctx);
/* Device time of the last run per pass (ms) and the cell-updates per step of
 * all shots together (updated cells of every window). */
int   erm_batch_timing(const erm_batch* b, double* ms_model, double* ms_source, double* ms_receiver,
                       double* cells_per_step, int* steps_model, int* steps_source, int* steps_receiver);

#ifdef __cplusplus
}
#endif
#endif
