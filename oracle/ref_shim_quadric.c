/*
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
 * reference's own src/quadric.c (included from where it lies, never copied)
 * to expose the file-static laplacian_smooth (src/quadric.c:316).
 */
#include "quadric.c"

void ref_laplacian_smooth(vec3d *v, vec3i *t, int nv, int nt) { laplacian_smooth(v, t, nv, nt); }
