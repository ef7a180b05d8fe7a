/* psst_oracle.h -- TEST INFRASTRUCTURE ONLY: C API of the CPU restatement of
 * the reference solver (psst_oracle.cpp).  Loaded by tests/, smoke() and
 * bench.py's CPU-baseline leg; never by the product library. */
#ifndef PSST_ORACLE_H
#define PSST_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
const char *oracle_last_error(void);
void *oracle_open(const char *mesh_path, const char *case_prefix);
void *oracle_open_raw(int dim, long nv, const double *xyz, long nf, const signed char *face_nv, const long *face_v,
                      const long *face_c, const long *face_zone, long nc, const long *cell_zone, long nz,
                      const long *zone_id, const long *zone_range, const long *zone_bc, const char *zone_names,
                      const char *case_prefix);
void oracle_close(void *h);
int oracle_info(void *h, int *info);          /* dim, N, G, F, rk, iter, max_iter, iter_step */
int oracle_constants(void *h, double *c);     /* gamma, rm, cp, cv, ref_vel, cfl, min_vol, uniform_dt */
int oracle_run(void *h, int n_iter, double *res_hist);
int oracle_get_state(void *h, double *prim);
int oracle_set_state(void *h, const double *prim);
int oracle_get_w(void *h, double *w);
int oracle_gradients(void *h, double *grads);
int oracle_residual(void *h, double *R);
int oracle_write_sav(void *h, const char *path, long iter_value);
double oracle_limiter(int kind, double num, double den, double eps2);
int oracle_face_flux(int flux, int dim, double gamma, double rm, double cp, double cv, const double *n,
                     const double *q0, const double *q1, double *F);
#ifdef __cplusplus
}
#endif
#endif
