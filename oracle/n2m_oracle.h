/*
 * n2m_oracle.h -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the reference's voxel-to-mesh path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Every function names the reference file:line it follows.
 */
#ifndef N2M_ORACLE_H
#define N2M_ORACLE_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include "../include/nii2mesh_amd.h"

#define ORC_CUBES_CLASSIC 0
#define ORC_CUBES_LEWINER 1

int orc_quick_smooth(float *img, int nx, int ny, int nz);
int orc_bwlabel(float *img, int conn, size_t dim[3], bool onlyLargest, bool fillBubbles);
void orc_dilate(float *img, const size_t dim[3]);
int orc_mc_classic(const float *img, const short dim[3], const int lo[3], const int hi[3],
                   float isolevel, vec3d **vs, vec3i **ts, int *nv, int *nt);
int orc_mc_lewiner(const float *img, const short dim[3], const int lo[3], const int hi[3],
                   int originalMC, float isolevel, vec3d **vs, vec3i **ts, int *nv, int *nt);
int orc_lewiner_select(const float cube[8], int lut, int originalMC, int *mc33case);
int orc_unify_vertices(vec3d **pts, vec3i *tris, int npt, int ntri);
int orc_remove_degenerate_triangles(vec3d *pts, vec3i **tris, int ntri);
int orc_meshify(float *img, short dim[3], int cubes, int originalMC, float isolevel, vec3i **t,
                vec3d **p, int *nt, int *np, bool preSmooth, bool onlyLargest, bool fillBubbles);
void orc_laplacian_smooth(vec3d *verts, const vec3i *tris, int nvert, int ntri);
void orc_laplacian_smoothHC(vec3d *verts, const vec3i *tris, int nvert, int ntri, double alpha,
                            double beta, int iter, bool lockEdges);
void orc_border_vertices(const vec3i *tris, int nvert, int ntri, uint8_t *border);
int orc_robust_range(const float *img, int nvox, float *pct2, float *pct98);
float orc_set_threshold(const float *img, int nvox, int darkMediumBright123);
int orc_otsu(const int *H, int nBin, int mode, int *dark, int *mid, int *bright);
void orc_apply_sform(vec3i *t, vec3d *p, int nt, int np, const float srow_x[4],
                     const float srow_y[4], const float srow_z[4]);
/* counts of 6/18/26-connected components and their voxel totals (tests) */
/* differentiable path: forward (exact != 0: FP64 positions) and adjoint */
void orc_laplacian_smooth_T(vec3d *g, const vec3i *tris, int nvert, int ntri);
int orc_diff_forward(const float *img, const short dim[3], float isolevel, int originalMC, int lapIter, int exact,
                     vec3i **t, vec3d **p, int *nt, int *np);
int orc_diff_gradient(const float *img, const short dim[3], float isolevel, int originalMC, int lapIter,
                      const double *dLdp, double *dLdvol);
int orc_decode(const void *raw, int datatype, int nvox, float slope, float inter, float *out);
int orc_label_components(const uint8_t *bw, const size_t dim[3], int conn, uint32_t *labels);

#endif
