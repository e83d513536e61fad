/*
 * oracle.h - C interface of the CPU oracle (liboracle.so).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg; never by the product library.
 */
#ifndef ERM_ORACLE_H
#define ERM_ORACLE_H
#include "../include/erm.h"
#ifdef __cplusplus
extern "C" {
#endif
int   oracle_set_threads(int n);
int   oracle_coefficients(int radius, double* c);
int   oracle_ricker(double freq, double dt_ms, double* w, int cap);
int   oracle_damping_axis(const erm_params* p, int axis, double* out);
/* prec: 64 (oracle) or 32 (CPU baseline) */
void* oracle3d_create(const erm_params* p, int prec);
void  oracle3d_destroy(void* h);
int   oracle3d_load_model(void* h, int which, const float* vp, const float* vs, const float* rho);
int   oracle3d_set_shot(void* h, int isx, int isy, int isz);
int   oracle3d_zero(void* h, int what); /* bit 1 wavefield, 2 excitation, 4 images */
int   oracle3d_step(void* h, int pass, int it_begin, int nsteps);
int   oracle3d_get_field(void* h, int fid, double* dst);
int   oracle3d_set_field(void* h, int fid, const double* src);
int   oracle3d_get_record(void* h, double* dst);
int   oracle3d_set_record(void* h, const double* src);
int   oracle3d_rtm_shot(void* h, int use_record);
#ifdef __cplusplus
}
#endif
#endif
