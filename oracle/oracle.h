/*
 * oracle.h — CPU restatement of the reference heliostat-field ray tracer (JachyLikeCoding/
 * SolarEnergy_Chier), used ONLY as test infrastructure: by tests/, by __graft_entry__.smoke()
 * as the checker, and by bench.py's cpu_baseline leg. Nothing in the product links or calls it.
 *
 * It restates, ray for ray, the reference algorithm from the same scene descriptors and the same
 * pre-generated sample arrays the GPU consumes; see oracle.cpp for the file:line each function
 * follows and for which functions are pinned by which checks.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H

#include <stdint.h>

#include "../include/qmcrt.h"  /* descriptor structs only */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Build + process a scene (SceneProcessor geometry). Returns NULL on error (message in err). */
orc_scene *orc_scene_build(const qmc_config *cfg, const qmc_receiver_desc *R, int nr, const qmc_grid_desc *G,
                           int ng, const qmc_heliostat_desc *H, int nh, char *err, int errlen);
void orc_scene_free(orc_scene *s);

int orc_receiver_info(const orc_scene *s, int r, qmc_receiver_info *out);
int orc_heliostat_info(const orc_scene *s, int h, qmc_heliostat_info *out);
int orc_grid_info(const orc_scene *s, int g, qmc_grid_info *out);
int orc_grid_cells(const orc_scene *s, int g, int32_t *cell_start, int32_t *match);
/* 3 vertices (9 floats) per sub-heliostat of grid g; returns the vertex count. */
int orc_subheliostat_vertices(const orc_scene *s, int g, float *out);
float orc_factor(const orc_scene *s, int h);
/* Micro-facet origins / normals of heliostat h (F*3 floats each); returns F. */
int orc_facets(const orc_scene *s, int h, float *origins, float *normals);
int orc_micro_count(const orc_scene *s, int h);

/* Buie table: cdf_xy 2*(n+1), slopes n, abcig = {A, B, C, I, gamma}. */
void orc_sunshape_table(const qmc_config *cfg, float *cdf_xy, float *slopes, float *abcig);
/* Pools from uniforms: ut, up in (0,1], gx, gz slope errors; dirs/perts get P*3 floats. */
void orc_pool_from_uniforms(const qmc_config *cfg, const float *ut, const float *up, const float *gx,
                            const float *gz, int P, float *dirs, float *perts);
/* Seeded pool generation (scrambled Sobol + counter-hash Box-Muller). */
void orc_pool_generate(const qmc_config *cfg, uint64_t seed_sunshape, uint64_t seed_perturbation, float *dirs,
                       float *perts, float *ut, float *up, float *gx, float *gz);
/* Physical Buie pools (QMC_SUNSHAPE_BUIE): the same seeded uniforms as orc_pool_generate, theta drawn
 * from a tabulated sky density (bin edges theta[n+1] in mrad, cdf[n+1], end densities lo[n], hi[n]);
 * the table is the product's host table, checked separately against independent quadrature. */
void orc_pool_generate_buie(const qmc_config *cfg, uint64_t seed_sunshape, uint64_t seed_perturbation,
                            const double *theta, const double *cdf, const double *lo, const double *hi, int n,
                            float *dirs, float *perts);
/* The physical sampler's inverse CDF alone: out[i] = theta [mrad] at cumulative probability u[i]. */
void orc_table_theta(const double *theta, const double *cdf, const double *lo, const double *hi, int n,
                     const double *u, int count, double *out);
/* Seeded group offsets of heliostat h: F+1 entries. */
void orc_offsets_generate(const orc_scene *s, uint64_t seed, int P, int h, int32_t *out);

/* Trace one heliostat (reference map_raytracing) into the receiver's flux (double accumulation,
 * res_x*res_y entries, accumulated). rec (F*L entries) and stats may be NULL. */
int orc_trace_heliostat(const orc_scene *s, int h, const float *dirs, const float *perts, int P, int L,
                        const int32_t *offsets, double *flux, qmc_ray_record *rec, qmc_trace_stats *stats);
/* Trace a list of heliostats of one receiver with nthreads host threads. offsets_all holds F+1
 * offsets per heliostat for ALL heliostats of the scene, concatenated in index order.
 * ray_stride > 1 traces only every ray_stride-th ray of each heliostat (bounded CPU samples). */
int orc_trace_many(const orc_scene *s, const int32_t *ids, int n, const float *dirs, const float *perts, int P,
                   int L, const int32_t *offsets_all, double *flux, int nthreads, int64_t ray_stride,
                   qmc_trace_stats *stats);

/* FP64 host adjoint of the differentiable trace of grid `grid` (flat heliostats, rectangle receiver):
 * bilinear flux (double), loss, d(loss)/d(aim offsets) (2 per heliostat) and d(loss)/d(sun_dir).
 * delta (5 doubles, may be NULL) shifts heliostat fd_helio's aim offset (across, up) and the sun
 * direction while visibility stays that of the unshifted scene (finite differences of the model). */
int orc_adjoint_fp64(const orc_scene *s, int grid, const float *dirs, const float *perts, int P, int L,
                     const int32_t *offsets_all, float target, float peak, float lambda, int fd_helio,
                     const double *delta, double *flux_out, double *loss_out, double *grad_aim, double *grad_sun,
                     int64_t *n_visible);

/* Trimmed-mean smoother, in place (reference two-heap algorithm). */
void orc_smooth(float *img, int w, int h, int radius, float ratio);

/* Primitives (for closed-form and brute-force checks). */
int orc_ray_parallelogram(const float o[3], const float d[3], const float A[3], const float B[3],
                          const float C[3], float tuv[3]);
int orc_cylinder_intersect(const float pos[3], const float size[3], const float o[3], const float d[3],
                           float tuv[3]);
void orc_cylinder_focus(const float pos[3], const float size[3], const float hpos[3], float out[3]);
int orc_cylinder_resolution(const float size[3], int ppm, int res[2]);
/* DDA collision over grid g from o along d, skipping sub ids [own_lo, own_hi); *hit gets the id. */
int orc_dda(const orc_scene *s, int g, const float o[3], const float d[3], int own_lo, int own_hi, int *hit);
/* Brute force: does the ray hit any non-own quad of grid g with t >= eps (any order)? */
int orc_brute_any_hit(const orc_scene *s, int g, const float o[3], const float d[3], int own_lo, int own_hi);
/* The cells (x-major addresses) the traversal of orc_dda visits with no quad tested, in order, up to cap of
 * them into cells; returns how many it visits. */
int orc_dda_cells(const orc_scene *s, int g, const float o[3], const float d[3], int32_t *cells, int cap);

#ifdef __cplusplus
}
#endif
#endif
