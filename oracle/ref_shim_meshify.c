/*
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
 * reference's own src/meshify.c (included from where it lies, never copied)
 * to expose its file-static stages to the parity tests:
 * quick_smooth (src/meshify.c:170), unify_vertices (:45),
 * remove_degenerate_triangles (:113) and dilate (:218).
 */
#include "meshify.c"

int ref_quick_smooth(float *img, int nx, int ny, int nz) { return quick_smooth(img, nx, ny, nz); }
int ref_unify_vertices(vec3d **p, vec3i *t, int npt, int ntri) { return unify_vertices(p, t, npt, ntri, false); }
int ref_remove_degenerate_triangles(vec3d *p, vec3i **t, int ntri) { return remove_degenerate_triangles(p, t, ntri, false); }
void ref_dilate(float *img, size_t dim[3]) { dilate(img, dim, true); }
void ref_free(void *p) { free(p); }
