/*
 * nii2mesh_amd.h -- C ABI of the MI355X (gfx950) voxel-to-mesh library.
 *
 * The first block declares the reference's own entry points with the same
 * prototypes and semantics, so the nii2mesh CLI, obj2mesh and the mesh
 * writers link against this library unchanged (see the per-symbol citations).
 * Everything runs on the GPU; a missing or failing device is reported as
 * EXIT_FAILURE (or, for functions without a status, a message on stderr),
 * never silently computed on the CPU.
 *
 * The second block (n2m_*) is new: stage entry points used by parity tests,
 * a device-resident context API (inputs already in HBM, no PCIe in the timed
 * path), the batched atlas entry and the adjoint.
 */
#ifndef NII2MESH_AMD_H
#define NII2MESH_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

/* Same layout and include guard as the reference's src/meshtypes.h:1-11, so a
 * translation unit may include both headers. */
#ifndef MESHIFY_TYPES_H
#define MESHIFY_TYPES_H
typedef struct {
	double x, y, z;
} vec3d;
typedef struct {
	int x, y, z;
} vec3i;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Drop-in entry points (reference prototypes)
 * ------------------------------------------------------------------------- */

/* Replaces src/meshify.h:9 meshify() (implementation src/meshify.c:286-389):
 * pre-smooth, range check, largest cluster / bubble fill, dilate, edge darken,
 * marching cubes, vertex welding and degenerate-triangle removal.  img is
 * mutated in place like the reference.  *t and *p are malloc'd; the caller
 * frees them.  originalMC selects the Bourke table inside the Lewiner
 * algorithm (src/MarchingCubes.c:471-477); which marching-cubes algorithm
 * runs (classic src/oldcubes.c or Lewiner src/MarchingCubes.c, a compile-time
 * choice in the reference, src/meshify.c:25-29) is chosen by
 * n2m_set_cubes() or the N2M_CUBES environment variable ("classic" default,
 * "lewiner"). */
int meshify(float *img, short dim[3], int originalMC, float isolevel, vec3i **t, vec3d **p,
            int *nt, int *np, bool preSmooth, bool onlyLargest, bool fillBubbles, bool verbose);

/* Replaces src/quadric.h:7 laplacian_smoothHC() (src/quadric.c:343-394):
 * HC Laplacian smoothing in FP64, border vertices reset when lockEdges. */
void laplacian_smoothHC(vec3d *verts, vec3i *tris, int nvert, int ntri, double alpha, double beta,
                        int iter, bool lockEdges);

/* Replaces src/quadric.h:6 quadric_simplify_mesh() (src/quadric.c:396-518):
 * the reference's edge-collapse simplification, result bit-identical.  The
 * collapse loop is sequential (each collapse reads what the earlier ones
 * changed) and runs on the host; the first update pass (reference lists,
 * border flags, plane quadrics, edge errors: src/quadric.c:144-215) runs on
 * the GPU.  *vs and *ts are freed and replaced by malloc'd arrays, as in the
 * reference.  When the GPU pass fails the mesh is left unchanged and the
 * reason goes to stderr. */
void quadric_simplify_mesh(vec3d **vs, vec3i **ts, int *nvert, int *ntri, int target_count,
                           double agressiveness, bool verbose, bool finishLossless);

/* Replaces src/bwlabel.h:10 bwlabel() (src/bwlabel.c:478-543): connected
 * components (conn 6/18/26) numbered by first voxel in x-fastest order,
 * optional largest-cluster selection and bubble fill.  Returns the number of
 * labels (1 after onlyLargest), 0 on error. */
int bwlabel(float *img, int conn, size_t dim[3], bool onlyLargest, bool fillBubbles);

/* Replaces src/oldcubes.h:9 / src/MarchingCubes.h:19 marchingCubes(): the
 * flavour follows n2m_set_cubes().  Classic emits 3 private vertices per
 * triangle (src/oldcubes.c:465-522); Lewiner shares edge vertices
 * (src/MarchingCubes.c:1086-1143). */
int marchingCubes(float *img, short dim[3], int lo[3], int hi[3], int originalMC, float isolevel,
                  vec3d **vs, vec3i **ts, int *nv, int *nt);

/* Replaces src/isolevel.h:10 setThreshold() (src/isolevel.c:245-277):
 * histograms on the GPU, multi-Otsu search on the host. */
float setThreshold(float *img, int nvox, int darkMediumBright123);

/* ---------------------------------------------------------------------------
 * Configuration and diagnostics
 * ------------------------------------------------------------------------- */
#define N2M_CUBES_CLASSIC 0 /* src/oldcubes.c */
#define N2M_CUBES_LEWINER 1 /* src/MarchingCubes.c */

/* Select the marching-cubes algorithm used by meshify()/marchingCubes() in
 * this process.  Returns the previous setting. */
int n2m_set_cubes(int flavour);
int n2m_get_cubes(void);
/* Text of the last error on the calling thread ("" if none). */
const char *n2m_last_error(void);
/* 1 if a gfx9 device is usable, else 0 (and n2m_last_error says why). */
int n2m_device_ok(void);

/* ---------------------------------------------------------------------------
 * Stage entry points on host buffers (one H2D, one D2H each); each names the
 * reference function whose result it reproduces.
 * ------------------------------------------------------------------------- */

/* src/meshify.c:170-216 quick_smooth(): returns EXIT_FAILURE (image
 * untouched) if any dimension is < 5. */
int n2m_quick_smooth(float *img, int nx, int ny, int nz);

/* src/meshify.c:45-106 unify_vertices(): welds *pts in place (realloc'd with
 * malloc), remaps tris, returns the new vertex count. */
int n2m_unify_vertices(vec3d **pts, vec3i *tris, int npt, int ntri);

/* src/meshify.c:113-168 remove_degenerate_triangles(): compacts *tris
 * (re-malloc'd when triangles are removed), returns the new count. */
int n2m_remove_degenerate_triangles(vec3d *pts, vec3i **tris, int ntri);

/* src/quadric.c:316-341 laplacian_smooth(), applied iter times. */
void n2m_laplacian_smooth(vec3d *verts, vec3i *tris, int nvert, int ntri, int iter);

/* save_mesh() (src/meshify.c:984-1019), every format the reference writes,
 * byte-identical files (FreeSurfer: but its time stamp).  The binary formats
 * (.mz3, .gii, .ply, .stl, .vtk, .pial / .inflated) are encoded on the GPU:
 * float32 vertices, PLY face records, STL facets, big-endian words, GIfTI
 * base64; zlib / gzip (isGz) runs on the host.  The text formats (.obj,
 * .off, .json) are formatted by the host C library with the reference's
 * format strings, in parallel chunks. */
int n2m_save_mesh(const char *fnm, vec3i *tris, vec3d *pts, int ntri, int npt, bool isGz);

/* src/nii2mesh.c:119,155-172 (the image loader): raw voxels to the FP32
 * working volume, out = (raw * scl_slope) + scl_inter in float, slope 0
 * read as 1.  datatype: 2 uint8, 4 int16, 512 uint16, 16 float32 (the
 * types the CLI reads).  Ships 1/4 (uint8) or 1/2 (16-bit) of the float
 * bytes over PCIe; n2m_ctx_decode does the same on device buffers. */
int n2m_decode(const void *raw, int datatype, int64_t nvox, float scl_slope, float scl_inter, float *out);

/* src/isolevel.c:141-243 nii_otsu() (file-static in the reference): returns
 * the threshold bin and fills dark/mid/bright; modes 1 and 5 (the
 * three-threshold search) run on the GPU.  nBin must be 256 for them. */
int n2m_otsu(const int *H, int nBin, int mode, int *dark, int *mid, int *bright);

/* src/meshify.c:1025-1045 apply_sform() on the GPU (FP64). */
void n2m_apply_sform(vec3i *t, vec3d *p, int nt, int np, float srow_x[4], float srow_y[4],
                     float srow_z[4]);

/* ---------------------------------------------------------------------------
 * Device-resident context: one HIP stream, reusable device buffers.
 * All d_* pointers are device memory on the context's device.
 *
 * Streams: every n2m_ctx_* call runs on the context's own stream, created
 * with hipStreamNonBlocking, and returns after that stream is idle.  It does
 * not wait for other streams: work the caller queued on its own stream (or
 * the legacy default stream) that writes an input must be complete before
 * the call (hipStreamSynchronize / torch.cuda.synchronize()).  Destroy a
 * context from a thread that may make HIP calls (not from a finaliser that
 * can run during process teardown).
 * ------------------------------------------------------------------------- */
typedef struct n2m_ctx n2m_ctx;

typedef struct {
	int cubes;        /* N2M_CUBES_CLASSIC or N2M_CUBES_LEWINER */
	int originalMC;   /* Lewiner only: Bourke table */
	float isolevel;
	int preSmooth;    /* bool */
	int onlyLargest;  /* bool */
	int fillBubbles;  /* bool */
	int verbose;      /* bool: print the reference's stage lines */
} n2m_params;

n2m_ctx *n2m_ctx_create(void);
void n2m_ctx_destroy(n2m_ctx *ctx);
/* hipStream_t of the context (as void*). */
void *n2m_ctx_stream(n2m_ctx *ctx);

/* meshify() on a device volume (mutated in place, dims x-fastest).  The mesh
 * stays on the device in the context.  Returns EXIT_SUCCESS/EXIT_FAILURE
 * exactly where meshify() would. */
int n2m_ctx_meshify(n2m_ctx *ctx, float *d_img, const int dim[3], const n2m_params *prm);

/* Multi-isolevel pipeline: pre-smooth once (if preSmooth), then for each of
 * niso isolevels run the rest of meshify() on a private copy of the smoothed
 * volume followed by HC smoothing (alpha, beta, hcIter, lockEdges=true).
 * Meshes are kept in the context (mesh index 0..niso-1).  d_img is not
 * modified.  Equal to calling meshify()+laplacian_smoothHC() per isolevel on
 * copies of the input volume. */
int n2m_ctx_pipeline(n2m_ctx *ctx, const float *d_img, const int dim[3], const n2m_params *prm,
                     int niso, const float *isolevels, double alpha, double beta, int hcIter);

/* GPU histograms + host multi-Otsu on a device volume (setThreshold()). */
float n2m_ctx_set_threshold(n2m_ctx *ctx, const float *d_img, int64_t nvox, int darkMediumBright123);
/* Several isolevels (modes 1/2/3 = dark/mid/bright) from one set of
 * histograms; out[i] for modes[i]. */
int n2m_ctx_set_thresholds(n2m_ctx *ctx, const float *d_img, int64_t nvox, int n, const int *modes,
                           float *out);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): n2m_ctx_set_thresholds without a context and without any
 * allocation.  Runs on `stream` (a hipStream_t; NULL for the default
 * stream) and returns when it is done.  Its histograms, range words and Otsu
 * table live in one fixed 1 MiB device array per device, held until the
 * process ends; every call on a device shares that array, so calls that
 * overlap on one device race.  Voxel counts are 32-bit in the histograms:
 * nvox must be 1..N2M_NOALLOC_MAX_VOXELS, n 1..N2M_NOALLOC_MAX_MODES and each
 * mode 1, 2 or 3; anything else returns EXIT_FAILURE with the reason in
 * n2m_last_error().  out[i] equals n2m_ctx_set_thresholds' out[i]. */
#define N2M_NOALLOC_MAX_VOXELS 4294967295ull
#define N2M_NOALLOC_MAX_MODES 16
int n2m_set_thresholds_noalloc(const float *d_img, int64_t nvox, int n, const int *modes, float *out,
                               void *stream);

/* HC smoothing of mesh `mesh` held by the context. */
int n2m_ctx_smooth_hc(n2m_ctx *ctx, int mesh, double alpha, double beta, int iter, int lockEdges);

int n2m_ctx_mesh_count(n2m_ctx *ctx);
int n2m_ctx_mesh_size(n2m_ctx *ctx, int mesh, int *nt, int *np);
/* Copy mesh `mesh` to caller-allocated host arrays (nt / np entries). */
int n2m_ctx_download(n2m_ctx *ctx, int mesh, vec3i *t, vec3d *p);
/* Device pointers of mesh `mesh` (valid until the next run on the context). */
int n2m_ctx_device_mesh(n2m_ctx *ctx, int mesh, const vec3i **d_t, const vec3d **d_p);

/* ---------------------------------------------------------------------------
 * Atlas (src/nii2mesh.c:492-583, the -a path): every integer label i of an
 * indexed volume is binarised as (i-0.5 < v < i+0.5) and meshed at isolevel
 * prm->isolevel (the CLI passes 0.5) with onlyLargest forced off, then, like
 * nii2() (src/nii2mesh.c:321-333), mapped through the sform (srow = 12
 * floats: srow_x, srow_y, srow_z, or NULL for none) and HC-smoothed
 * postSmooth times.  Labels run 1..trunc(max); a label with no voxels is
 * skipped.  Each label's mesh equals the reference's for that label.
 * ------------------------------------------------------------------------- */
#define N2M_ATLAS_OK 0     /* meshed */
#define N2M_ATLAS_EMPTY 1  /* no voxel carries the label (skipped) */
#define N2M_ATLAS_FAILED 2 /* meshify() returned EXIT_FAILURE for it */

typedef struct {
	int label;
	int status;      /* N2M_ATLAS_* */
	int nt, np;      /* triangle / vertex counts */
	int tri_offset;  /* first triangle in the atlas union mesh */
	int vert_offset; /* first vertex in the atlas union mesh */
} n2m_atlas_entry;

/* Batched atlas: one pass over the label volume finds every label's voxel
 * count and box; all labels are then smoothed, classified, meshed, welded
 * and HC-smoothed together on per-label crops (box + 4 voxels).  With
 * fillBubbles the per-label loop below is used.  Returns EXIT_SUCCESS if at
 * least one label was meshed (the CLI's nOK > 0). */
int n2m_ctx_atlas(n2m_ctx *ctx, const float *d_labels, const int dim[3], const n2m_params *prm,
                  int postSmooth, const float *srow);
/* The reference's loop run on the GPU: per label, binarise the whole volume
 * and run meshify() + sform + HC on it.  Same results as n2m_ctx_atlas. */
int n2m_ctx_atlas_loop(n2m_ctx *ctx, const float *d_labels, const int dim[3], const n2m_params *prm,
                       int postSmooth, const float *srow);
/* Number of label entries of the last atlas run (labels 1..n). */
int n2m_ctx_atlas_count(n2m_ctx *ctx);
int n2m_ctx_atlas_entry(n2m_ctx *ctx, int idx, n2m_atlas_entry *e);
/* Copy one label's mesh (vertex ids local to the label) to host arrays. */
int n2m_ctx_atlas_download(n2m_ctx *ctx, int idx, vec3i *t, vec3d *p);
/* Device pointers of the union mesh (vertex ids global). */
int n2m_ctx_atlas_device_mesh(n2m_ctx *ctx, const vec3i **d_t, const vec3d **d_p, int *nt, int *np);

/* Host-memory form of the atlas loop for the CLI: meshes[k] for label k+1,
 * arrays malloc'd (free with n2m_atlas_free).  Returns as n2m_ctx_atlas. */
typedef struct {
	int label;
	int status;
	int nt, np;
	vec3i *t;
	vec3d *p;
} n2m_atlas_mesh;
int n2m_atlas(const float *img, short dim[3], int originalMC, bool preSmooth, bool fillBubbles, int postSmooth,
              const float *srow, bool verbose, n2m_atlas_mesh **meshes, int *nmesh);
void n2m_atlas_free(n2m_atlas_mesh *meshes, int nmesh);

/* ---------------------------------------------------------------------------
 * Differentiable path.  Forward: Lewiner marchingCubes over the whole volume
 * (lo = 0, hi = dim-1; src/MarchingCubes.c:1086-1143) at `isolevel`,
 * unify_vertices + remove_degenerate_triangles (src/meshify.c:45-168) and
 * lapIter laplacian_smooth iterations (src/quadric.c:316-341); the mesh is
 * mesh 0 of the context.  exact = 0 keeps the reference's FP32 vertex
 * positions; exact = 1 evaluates them in FP64 (the smooth map the adjoint
 * differentiates).  d_vol must stay unchanged until the backward.
 * Backward: dL/dvolume (nx*ny*nz doubles, x fastest) from dL/dp (np x 3
 * doubles) of the final vertices: adjoint of the edge interpolation
 * u = c0/(c0-c1), of the centre-vertex mean, of the weld and of the
 * transposed Laplacian.  No counterpart in the reference.
 * ------------------------------------------------------------------------- */
/* n2m_save_mesh for mesh `mesh` held by the context (no upload). */
int n2m_ctx_save_mesh(n2m_ctx *ctx, int mesh, const char *fnm, int isGz);

/* Device form of n2m_decode (d_raw and d_out on the context's device). */
int n2m_ctx_decode(n2m_ctx *ctx, const void *d_raw, int datatype, int64_t nvox, float scl_slope, float scl_inter,
                   float *d_out);

int n2m_ctx_diff_forward(n2m_ctx *ctx, const float *d_vol, const int dim[3], float isolevel, int originalMC,
                         int lapIter, int exact);
int n2m_ctx_diff_backward(n2m_ctx *ctx, const double *d_dLdp, double *d_dLdvol);
/* Transpose of n2m_laplacian_smooth (iter times) on host arrays. */
void n2m_laplacian_smooth_T(vec3d *g, vec3i *tris, int nvert, int ntri, int iter);

/* Per-kernel timing with HIP events on the context stream. */
void n2m_ctx_enable_timing(n2m_ctx *ctx, int on);
/* Time only the stage called `name` (NULL: all stages). */
void n2m_ctx_timing_only(n2m_ctx *ctx, const char *name);
void n2m_ctx_reset_timing(n2m_ctx *ctx);
/* Returns the number of timed kernels; fills entry idx if idx < count. */
int n2m_ctx_kernel_stat(n2m_ctx *ctx, int idx, const char **name, double *ms_total, int *launches,
                        double *bytes_total);

/* Launch shape the batched atlas's autotuner chose for stage `stage` on
   `device` (no reference counterpart: it reports the tuner of
   n2m_atlas / n2m_ctx_atlas).  Returns the number of stages; for a valid
   stage fills its name, its crop tile (-1 for stages without one) and its
   blocks per CU (0: one workgroup per work item), both -1 until that stage
   has been tuned on the device. */
int n2m_atlas_tuned(int device, int stage, const char **name, int *tile, int *blocks_per_cu);

#ifdef __cplusplus
}
#endif
#endif /* NII2MESH_AMD_H */
