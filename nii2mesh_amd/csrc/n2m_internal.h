// n2m_internal.h -- host-side declarations shared by the kernel files and the
// C-ABI layer (api.cpp).  Every launch_* enqueues work on stream s and
// returns immediately.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/nii2mesh_amd.h"

namespace n2m {

// inverse of the order-preserving float -> uint32 map used by the device
// min/max reductions
inline float ord2f_host(uint32_t o)
{
	uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
	float f;
	__builtin_memcpy(&f, &u, 4);
	return f;
}

// The isolevel meshify() takes for an intensity range [mn, mx] when asked
// for `req` (src/meshify.c:312-319) and its edge value (:347); false for no
// variability.  One definition for the host and the device, so both derive
// the same values.
__host__ __device__ inline bool atlas_iso(float req, float mn, float mx, float *iso, float *edgeMax, bool *fallback)
{
	if (mn == mx)
		return false;
	float isolevel = req;
	*fallback = (isolevel <= mn) || (isolevel > mx);
	if (*fallback)
		isolevel = 0.5 * (mn + mx);
	*iso = isolevel;
	*edgeMax = 0.75 * (mn + isolevel);
	return true;
}

// One label of the batched atlas (src/nii2mesh.c:551-573): its padded crop
// of the volume and the per-label values meshify() derives.
struct AtlasTask {
	int label;
	float lo, hi;       // voxel v belongs to the label when lo < v < hi
	int ox, oy, oz;     // crop origin in the volume
	int nx, ny, nz;     // crop size
	long long off;      // crop offset in the crop buffer (floats)
	float iso, mn, edgeMax;
	int lo3[3], hi3[3]; // marching-cubes box in crop coordinates
	float mnv;          // Lewiner padding value (min - iso)
	long long cube_off; // offset of the task's per-cube / per-point arrays
	int first_block;    // first workgroup of the task in the current launch
	int nseg;           // Lewiner segments per row
	int pad;
};

// sub-box geometry of the differentiable path's marching cubes
struct AdjGeo {
	int nx;
	long long nxy, nxyz;
	int lo0, lo1, lo2;
	int sx, sy;
	float iso, mnv;
};

struct CclTimer {
	virtual void begin(const char *name, double bytes) = 0;
	virtual void end() = 0;
	virtual ~CclTimer() {}
};

// k_volume.hip
// pre-smooth with the range scan; when bits != null it also writes the
// threshold masks (out >= isos[k]) of the first min(nmask,
// N2M_SMOOTH_MASKS) isolevels to bits + k * ceil(nx/32) * ny * nz
constexpr int N2M_SMOOTH_MASKS = 4;
// threshold masks (img >= isos[k]) of n isolevels to bits + k * ceil(nx/32) *
// ny * nz, one read of the volume when n <= N2M_SMOOTH_MASKS
void launch_mask_bits_multi(const float *img, int nx, int ny, int nz, const float *isos, int n, uint32_t *bits,
                            hipStream_t s);
void launch_smooth(const float *in, float *out, int nx, int ny, int nz, uint32_t *range, const float *isos,
                   int nmask, uint32_t *bits, hipStream_t s);
void launch_minmax(const float *img, size_t n, uint32_t *range, hipStream_t s);
void launch_mask_bits(const float *img, int nx, int ny, int nz, float iso, int nonzero, uint32_t *bits,
                      hipStream_t s);
void launch_dilate_bits(const uint32_t *in, uint32_t *out, int nx, int ny, int nz, hipStream_t s);
// img = apply(src) (src null: in place); every voxel of img is written
// sgp / sgn (when given and nx % 32 == 0 with 16-byte aligned volumes):
// the Lewiner sign bits of the result, lw_clamp(v - iso) > 0 / < 0, ceil(nx
// / 32) words per row; returns whether they were written
bool launch_apply(const float *src, float *img, int nx, int ny, int nz, float iso, float mn, float edgeMax,
                  const uint32_t *keep, const uint32_t *dil, int fill, int *bbox, uint32_t *vmin, hipStream_t s,
                  uint32_t *sgp = nullptr, uint32_t *sgn = nullptr);
void launch_hist_range(const float *img, size_t n, uint32_t *stats, hipStream_t s);
// mesh output encoding (k_encode.hip)
void launch_f64_to_f32(const double *in, size_t n, float *out, hipStream_t s);
void launch_ply_faces(const int *tris, size_t nt, uint8_t *out, hipStream_t s);
void launch_base64(const uint8_t *src, size_t len, uint8_t *out, hipStream_t s);
// binary STL facets (50 bytes each, written as 25 16-bit halves)
void launch_stl_facets(const double *pts, const int *tris, size_t nt, uint16_t *out, hipStream_t s);
// byte-swapped 32-bit words; VTK polygon records (3, v0, v1, v2), big-endian
void launch_bswap32(const uint32_t *in, size_t n, uint32_t *out, hipStream_t s);
void launch_vtk_polys(const int *tris, size_t nt, uint32_t *out, hipStream_t s);
// raw NIfTI voxels -> float (datatype 2 uint8, 4 int16, 512 uint16, 16 float32);
// returns -1 for another datatype
int launch_decode(const void *raw, int datatype, size_t n, float slope, float inter, float *out, hipStream_t s);
// mode-5 Otsu search, one workgroup per first threshold l (nBin <= 258)
void launch_otsu3(const double *P, int nBin, double *bv, int *bm, int *bh, hipStream_t s);
void launch_hist_bins(const float *img, size_t n, float mn, float scl, int nbins, uint32_t *hist, hipStream_t s);

// atlas volume stages (k_volume.hip)
void launch_atlas_stats(const float *labels, int nx, int ny, int nz, int nlabel, uint32_t *stats, hipStream_t s);
// The atlas stages are persistent: `grid` workgroups (at most nblocks) walk
// the nblocks items of the block map.
void launch_atlas_smooth(const float *labels, int gnx, int gny, int gnz, const AtlasTask *tasks,
                         const uint32_t *blockTask, unsigned nblocks, unsigned grid, int smooth, int tile,
                         float *crops, uint32_t *ranges, hipStream_t s);
int atlas_tile_count();
void atlas_tile_shape(int tile, int shape[3]);
void launch_atlas_apply(float *crops, int gnx, int gny, int gnz, const AtlasTask *tasks, const uint32_t *blockTask,
                        unsigned nblocks, unsigned grid, int *bbox, uint32_t *vmin, hipStream_t s);
// iso, mn, edgeMax of every task from its range (ranges[2t], ord min/max),
// as atlas_iso; a task without variability gets a NaN iso
void launch_atlas_iso(AtlasTask *tasks, int ntasks, const uint32_t *ranges, float req, hipStream_t s);
unsigned atlas_smooth_blocks(int nx, int ny, int nz, int smooth, int tile);
unsigned atlas_apply_blocks(int nx, int ny, int nz);
void launch_binarize(const float *labels, size_t n, float lo, float hi, float *out, uint32_t *count, hipStream_t s);

// k_ccl.hip.  Runs of set bits are named by dense ids in voxel order:
// launch_ccl fills rbase[w] (nwords + 1 entries) with the number of runs
// before word w, and the selection kernels take the same rbase.  L and AUX
// are indexed by run id (nvox / 2 + nwords entries bound the run count).
void launch_ccl(const uint32_t *bits, int nx, int ny, int nz, int conn, int invert, uint32_t *L, uint32_t *AUX,
                int mode, uint32_t *nroots, uint32_t *rootlist, uint32_t *rbase, uint32_t *scan_scratch,
                hipStream_t s, CclTimer *timer);
void launch_ccl_argmax(const uint32_t *rootlist, const uint32_t *nroots, const uint32_t *AUX,
                       unsigned long long *best, hipStream_t s);
void launch_select_root(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L,
                        const unsigned long long *best, uint32_t *out, const uint32_t *rbase, hipStream_t s);
// ninv: device word holding the background component count (no fill at <= 1)
void launch_fill_bubbles(uint32_t *bits, int nx, int ny, int nz, const uint32_t *L, const uint32_t *AUX,
                         const uint32_t *rbase, const uint32_t *ninv, hipStream_t s);
void launch_write_labels(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L, const uint32_t *rank,
                         float *img, const uint32_t *rbase, hipStream_t s);
void launch_scatter_rank(const uint32_t *sorted, uint32_t n, uint32_t *rank, hipStream_t s);
void launch_bits_to_float(const uint32_t *bits, int nx, int ny, int nz, float *img, hipStream_t s);

// k_mc.hip
size_t classic_blocks(int cx, int cy, int cz);
void launch_classic_count(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                          uint8_t *cfgs, uint32_t *blockSums, hipStream_t s);
void launch_classic_emit(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                         const uint8_t *cfgs, const uint32_t *blockOffs, double *pts, const int origin[3],
                         hipStream_t s);
size_t lw_nseg_total(int nx, int ny, int nz, const int lo[3], const int hi[3]);
// pass A of the whole-volume Lewiner cubes: point flags to pflags (one byte
// each), or, when fmask is given, to per-axis bit masks (lw_fmask_words
// words); tiling rows to crow, segment counts to segcnt
// launch_lw_classify (fmask form) from the sign bits the apply stage wrote
// (P: lw_clamp(v - iso) > 0, N: < 0; per volume row, ceil(nx / 32) words)
void launch_lw_classify_bits(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                             float mnv, int originalMC, const uint32_t *P, const uint32_t *N, uint16_t *crow,
                             uint32_t *segcnt, unsigned long long *fmask, hipStream_t s);
void launch_lw_classify(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                        float mnv, int originalMC, uint8_t *pflags, uint16_t *crow, uint32_t *segcnt,
                        hipStream_t s, unsigned long long *fmask = nullptr);
size_t lw_fmask_words(int nx, int ny, int nz, const int lo[3], const int hi[3]);
// pass B from the flag masks
void launch_lw_emit(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                    float mnv, int originalMC, const unsigned long long *fmask, const uint16_t *crow,
                    const uint32_t *segcnt, const uint32_t *bases, uint32_t nedge, double *pts, int *tris,
                    const int origin[3], hipStream_t s);

void launch_lw_emit_rec(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                        float mnv, int originalMC, const uint8_t *pflags, const uint16_t *crow,
                        const uint32_t *segcnt, const uint32_t *bases, uint32_t nedge, double *pts, int *tris,
                        int exact, unsigned long long *vrec, uint32_t *crec, hipStream_t s);

// adjoint (k_adjoint.hip)
void launch_adj_laplace_T(const uint32_t *start, const int2 *pairs, double *g, double *h, uint32_t n,
                          hipStream_t s);
void launch_adj_unweld(const int *cls, const int *rep, const double *g, uint32_t nmc, double *gm, hipStream_t s);
void launch_adj_centre(const unsigned long long *vrec, const uint32_t *crec, uint32_t nmc, double *gm,
                       hipStream_t s);
void launch_adj_edge(const unsigned long long *vrec, uint32_t nmc, const double *gm, const float *img,
                     const AdjGeo &g, double *dvol, hipStream_t s);

// atlas marching cubes (k_mc.hip): persistent, `grid` workgroups walk the
// nblocks items of the block map
void launch_atlas_classic_count(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                                unsigned nblocks, unsigned grid, uint8_t *cfgs, uint32_t *blockSums, hipStream_t s);
void launch_atlas_classic_emit(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                               unsigned nblocks, unsigned grid, const uint8_t *cfgs, const uint32_t *blockOffs,
                               double *pts, hipStream_t s);
void launch_atlas_lw_classify(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                              unsigned nblocks, unsigned grid, int originalMC, uint8_t *pflags, uint16_t *crow,
                              uint32_t *segcnt, size_t stride, hipStream_t s);
void launch_atlas_lw_emit(const float *crops, const AtlasTask *tasks, int ntasks, const uint32_t *blockTask,
                          unsigned nblocks, unsigned grid, int originalMC, const uint8_t *pflags, const uint16_t *crow,
                          const uint32_t *segcnt, const uint32_t *bases, size_t stride, uint32_t totalE,
                          double *pts, int *tris, hipStream_t s);
int lw_seg_len();

// k_mesh.hip
size_t scan_scratch_words(size_t n);
size_t scan_scratch_zero_words(); // leading words kept zero between scans
void scan_excl_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, uint32_t *scratch,
                   hipStream_t s);
size_t radix_counts_words(size_t n);
bool radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int bits,
                    uint32_t *counts, uint32_t *scan_scratch, hipStream_t s);
// one tile of a segmented radix sort: keys [start, start + len), digit d's
// count at counts[cbase + d * cstride]
struct RsTile {
	uint32_t start, len, cbase, cstride;
};
// tiles that sort every segment [sb[k], sb[k + 1]) on its own (the counts
// take 256 words per tile)
std::vector<RsTile> radix_seg_tiles(const std::vector<uint32_t> &sb);
bool radix_sort_seg_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int bits,
                        const RsTile *desc, size_t ntiles, uint32_t *counts, uint32_t *scan_scratch,
                        hipStream_t s);
void launch_weld_keys(const double *pts, uint32_t n, uint32_t *keys, uint32_t *vals, hipStream_t s);
void launch_weld_segflags(const uint32_t *keys, uint32_t n, uint32_t *flags, hipStream_t s);
void launch_weld_segstarts(const uint32_t *flags, const uint32_t *segid, uint32_t n, uint32_t *starts,
                           hipStream_t s);
// d_nseg / d_nnew / d_n: counts left on the device (no host round trip);
// the launches then cover the upper bound n
// lrep: n words written by the greedy pass (each seed's representative), from
// which the assign pass sets rep[] with plain stores (rep needs no fill)
void launch_weld_greedy(const uint32_t *starts, uint32_t nseg, uint32_t n, const uint32_t *keys,
                        const uint32_t *vals, const double *pts, int *lcls, uint32_t *seeds, uint32_t *lrep,
                        hipStream_t s, const uint32_t *d_nseg = nullptr);
void launch_weld_assign(const uint32_t *vals, const int *lcls, const uint32_t *flags, const uint32_t *segid,
                        const uint32_t *seedbase, const uint32_t *lrep, uint32_t n, int *cls, int *rep,
                        hipStream_t s, const uint32_t *d_nnew = nullptr);
void launch_gather_pts(const double *pts, const int *rep, uint32_t n, double *out, hipStream_t s,
                       const uint32_t *d_n = nullptr);
// wpos[v] = pts[rep[cls[v]]]: the welded position of every pre-weld vertex v
void launch_weld_wpos(const double *pts, const int *cls, const int *rep, uint32_t npt, double *wpos, hipStream_t s);
// remap (tris_in null: implicit 3t+k) and degenerate flags in one pass
void launch_remap_degen(const int *tris_in, const int *cls, uint32_t npt, const double *wpos, uint32_t nt,
                        int *tris_out, uint32_t *keep, hipStream_t s);
void launch_iota(int *out, uint32_t n, hipStream_t s);
void launch_remap_tris(const int *tris_in, const int *cls, uint32_t nt, uint32_t npt, int *tris_out,
                       hipStream_t s);
void launch_degen_flags(const double *pts, const int *tris, uint32_t nt, uint32_t np, uint32_t *keep,
                        hipStream_t s);
void launch_compact_tris(const int *tris, const uint32_t *keep, const uint32_t *pos, uint32_t nt, int *out,
                         hipStream_t s);
// CSR of the references: count (with each corner's arrival rank), scan,
// fill at start + rank, then sort each vertex's list and expand to pairs
// inv given: counts the references of inv[tris[i]] and writes those ids to
// tout (the relabelled triangles)
void launch_csr_count(const int *tris, uint32_t nt, uint32_t *cnt, uint32_t *rank, hipStream_t s,
                      const uint32_t *inv = nullptr, int *tout = nullptr);
void launch_csr_fill(const int *tris, uint32_t nt, const uint32_t *start, const uint32_t *rank, uint32_t *ref,
                     hipStream_t s);
void launch_csr_pairs(const int *tris, uint32_t nv, const uint32_t *start, uint32_t *ref, int2 *pairs,
                      hipStream_t s,
                      bool sorted_ref = false);
void launch_border(const int *tris, uint32_t nv, const uint32_t *start, const uint32_t *ref, uint8_t *border,
                   hipStream_t s);
void launch_hc_a(const uint32_t *start, const int2 *pairs, const double *o, const double *p, double *pn,
                 double *bq, uint32_t nv, double alpha, hipStream_t s);
void launch_hc_b(const uint32_t *start, const int2 *pairs, const double *pn, const double *bq, double *pout,
                 uint32_t nv, double beta, hipStream_t s);
void launch_laplace(const uint32_t *start, const int2 *pairs, const double *p, double *pout, uint32_t nv,
                    hipStream_t s);
// halo form of the smoothing gathers: per 1024-vertex block, the sorted
// distinct neighbours outside the block (halo_words(nv) words, stride 1024)
// and their count; the pairs as two 16-bit LDS slots (3*nt words)
uint32_t halo_blocks(uint32_t nv);
size_t halo_words(uint32_t nv);
void launch_halo_build(const uint32_t *start, const int2 *pairs, uint32_t nv, uint32_t *halo, uint32_t *hcount,
                       uint32_t *slots, uint32_t *overflow, hipStream_t s);
void launch_hc_a_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo, const uint32_t *hcount,
                      const double *o, const double *p, double *pn, double *bq, uint32_t nv, double alpha,
                      hipStream_t s);
void launch_hc_b_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo, const uint32_t *hcount,
                      const double *pn, const double *bq, double *pout, uint32_t nv, double beta, hipStream_t s);
void launch_laplace_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo,
                         const uint32_t *hcount, const double *p, double *pout, uint32_t nv, hipStream_t s);
// two-ring form of an HC iteration: per block of rb vertices, ring 1 and ring 2
// (ring_words, stride 2304 - rb; counts as uint2 in rcnt, 2 words per block), the
// block's pairs as 16-bit LDS row slots (slots, 3*nt words) and ring 1's pairs
// (roff: ring_off_words 16-bit offsets, rslots: ring_pair_words words), and
// ring 1's original positions (o1)
int ring_rb(); // vertices per block (N2M_HC_RB: 512, 640 or 768)
uint32_t ring_blocks(uint32_t nv, int rb);
size_t ring_words(uint32_t nv, int rb);
size_t ring_pair_words(uint32_t nv, int rb);
size_t ring_off_words(uint32_t nv, int rb);
size_t ring_o_rows(uint32_t nv, int rb); // ring 1's original positions, stride 1024 rows
void launch_ring_build(const uint32_t *start, const int2 *pairs, uint32_t nv, const double *o, uint32_t *ring,
                       uint32_t *rcnt, uint32_t *slots, uint16_t *roff, uint32_t *rslots, double *o1,
                       uint32_t *overflow, int rb, hipStream_t s);
void launch_hc_ring(const uint32_t *start, const uint32_t *slots, const uint16_t *roff, const uint32_t *rslots,
                    const uint32_t *ring, const uint32_t *rcnt, const double *o, const double *o1, const double *p,
                    double *pout, uint32_t nv, double alpha, double beta, int rb, hipStream_t s);
// smoothing state: 3 doubles per vertex, optionally gathered / scattered
// through a vertex permutation
void launch_hc_load(const double *src, const uint32_t *perm, uint32_t n, double *dst, hipStream_t s);
void launch_hc_store(const double *p, const double *o, const uint8_t *border, const uint32_t *perm, uint32_t n,
                  double *dst, hipStream_t s);
void launch_pts_bbox(const double *pts, uint32_t n, uint32_t *bb, hipStream_t s);
void launch_morton(const double *pts, uint32_t n, const uint32_t *bb, uint32_t *keys, uint32_t *vals,
                   hipStream_t s);
void launch_invert_perm(const uint32_t *perm, uint32_t n, uint32_t *inv, hipStream_t s);
void launch_tris_repeat(const int *tris, uint32_t nt, uint32_t *flag, hipStream_t s);
void launch_border_pairs(const uint32_t *start, const int2 *pairs, uint32_t nv, uint8_t *border, hipStream_t s);
// high-valence vertices (more than 16 references)
// the lists longer than 16 references, one wave each (the longest one
// workgroup each): sorted in place with their pairs expanded; their border
// flags (literal: the once-counted ids).  huge: 1 + (3 nt / 257 + 1) words
// of scratch.
void launch_long_sort(const int *tris, uint32_t nv, const uint32_t *start, uint32_t *ref, int2 *pairs,
                      uint32_t *huge, hipStream_t s);
void launch_long_border(const int *tris, uint32_t nv, const uint32_t *start, const uint32_t *ref, int literal,
                        uint8_t *border, uint32_t *huge, hipStream_t s);
size_t long_huge_words(size_t nt);
// batched atlas weld (k_mesh.hip)
void launch_atlas_weld_keys(const double *pts, uint32_t n, const uint32_t *vbase, int ntasks, uint32_t *keys,
                            uint32_t *vals, uint32_t *vtask, hipStream_t s);
void launch_atlas_weld_segflags(const uint32_t *keys, const uint32_t *tk, uint32_t n, uint32_t *flags,
                                hipStream_t s);
void launch_atlas_class_base(const uint32_t *tk, const uint32_t *segid, const uint32_t *seedbase, uint32_t n,
                             uint32_t *cbase, hipStream_t s);
void launch_atlas_weld_tasks(uint32_t *cbase, const uint32_t *vbase, const uint32_t *nextne, int ntasks,
                             const uint32_t *d_nnew, uint32_t *ncls, uint8_t *ident, hipStream_t s);
void launch_atlas_weld_identity(uint32_t n, const uint32_t *vtask, const uint32_t *vbase, const uint32_t *cbase,
                                const uint8_t *ident, int *cls, int *rep, hipStream_t s);
void launch_gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t n, uint32_t *dst, hipStream_t s);
void launch_add_i32(const int *src, size_t n, int add, int *dst, hipStream_t s);
void launch_sform(double *p, uint32_t np, const float *sx, const float *sy, const float *sz, hipStream_t s);
void launch_swap_winding(int *t, uint32_t nt, hipStream_t s);
// setThreshold() on the device (k_volume.hip): range, percentile walk, the
// two histograms, the Liao table and the Otsu searches run as one chain of
// launches, whose results land in one record for a single copy to the host.
struct ThrRecord {
	uint32_t stats[4]; // ordered min, ordered max, NaN count, 0
	float par[4];      // mn, scl of the 1001-bin pass; mn, scl of the 256-bin pass
	int32_t do1001;    // the 1001-bin pass ran
	int32_t pad[3];
	uint32_t h256[256];
	double probe[2 * 256 + 4]; // P[i], P[i+1][255], the four mode-5 start terms
	double bv[256];            // mode-5 search: best value, m and h per first threshold l
	int32_t bm[256], bh[256];
};
// rec: device record; h1001: 1024 words; P: 256 x 256 doubles.  need5: run
// the three-threshold search (modes 1 and 3).  Uses no host round trip.
void launch_threshold_chain(const float *img, size_t nvox, ThrRecord *rec, uint32_t *h1001, double *P, bool need5,
                            hipStream_t s);
// the isolevels of setThreshold(img, nvox, modes[i]) from the chain's record
// (src/isolevel.c:141-277): mode 1 dark, 3 bright, anything else medium
void thresholds_from_record(const ThrRecord &r, int n, const int *modes, float *out);

// quadric simplification (k_quadric.hip, quadric.cpp): the first update
// pass of quadric_simplify_mesh on the GPU
void launch_q_planes(const double *p, const int *tris, uint32_t nt, double *nd, hipStream_t s);
void launch_q_vertex(const uint32_t *start, const uint32_t *ref, uint32_t nv, const double *nd, double *q,
                     hipStream_t s);
void launch_q_errors(const double *p, const int *tris, uint32_t nt, const double *q, const uint8_t *border,
                     double *err, hipStream_t s);
// update_mesh(iteration 0) of src/quadric.c:144-215 for a host mesh: the
// vertex lists (start, nv + 1; ref, corners t*3+k in ascending order), the
// border flags, 10 quadric terms per vertex, normal + plane offset and the
// four edge errors per triangle.  Returns 0, or EXIT_FAILURE with the reason
// in n2m_last_error().
struct QuadricPrep {
	std::vector<uint32_t> start, ref;
	std::vector<uint8_t> border;
	std::vector<double> q, nd, err;
};
int quadric_prep(const vec3d *p, const vec3i *t, int nv, int nt, QuadricPrep &g);

// the 2 % / 98 % bins of the robust range (src/isolevel.c:102-134): lo is
// the bin where the count from below reaches n2pct, hi likewise from above;
// when they meet, both move outwards to the next non-empty bins.  Shared by
// the host and the device chain.
__host__ __device__ inline void robust_walk(const uint32_t *h, int kBins, size_t n2pct, int *plo, int *phi)
{
	size_t n = 0, lo = 0;
	while (n < n2pct)
		n += h[lo++];
	lo--;
	n = 0;
	int hi = kBins;
	while (n < n2pct)
		n += h[--hi];
	if ((int)lo == hi) { // src/isolevel.c:118-134
		int ok = -1;
		while (ok != 0) {
			if (lo > 0) {
				lo--;
				if (h[lo] > 0)
					ok = 0;
			}
			if (ok != 0 && hi < kBins - 1) {
				hi++;
				if (h[hi] > 0)
					ok = 0;
			}
			if (lo == 0 && hi == kBins - 1)
				ok = 0;
		}
	}
	*plo = (int)lo;
	*phi = hi;
}

// host helpers of api.cpp shared with noalloc.hip
void set_error(const char *msg); // n2m_last_error() text of this thread
void clear_error();

} // namespace n2m
