/*
 * n2m_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C, single-threaded restatement of neurolabusc/nii2mesh's voxel-to-
 * mesh path, written from the reference's behaviour (each function cites the
 * file:line it follows).  The GPU library is checked against it; the oracle
 * itself is pinned against the reference compiled from its own sources
 * (oracle/_ref, built by oracle/Makefile) and against closed-form checks in
 * tests/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * Floating point: every expression keeps the reference's operand order and
 * precision (float storage, double where the reference promotes); build
 * with -ffp-contract=off so that no multiply-add is fused.
 */
#include "n2m_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../nii2mesh_amd/csrc/mc_tables.h"

#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))
#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* pre-smooth: src/meshify.c:170-216                                         */
/* ------------------------------------------------------------------------ */

/* One 5-tap pass along an axis of length len with element stride `stride`;
 * voxels closer than 2 to either end of the axis are copied. */
static void smooth_axis(const float *src, float *dst, int nx, int ny, int nz, int axis)
{
	const double w0 = 0.45, w1 = 0.225, w2 = 0.05; /* src/meshify.c:176-178 */
	size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
	size_t s = axis == 0 ? sx : (axis == 1 ? sy : sz);
	for (int z = 0; z < nz; z++)
		for (int y = 0; y < ny; y++)
			for (int x = 0; x < nx; x++) {
				size_t v = x * sx + y * sy + z * sz;
				int c = axis == 0 ? x : (axis == 1 ? y : z);
				int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
				if (c < 2 || c >= len - 2) {
					dst[v] = src[v];
					continue;
				}
				double acc = src[v - 2 * s] * w2;
				acc = acc + src[v - s] * w1;
				acc = acc + src[v] * w0;
				acc = acc + src[v + s] * w1;
				acc = acc + src[v + 2 * s] * w2;
				dst[v] = (float)acc;
			}
}

int orc_quick_smooth(float *img, int nx, int ny, int nz)
{
	if (nx < 5 || ny < 5 || nz < 5) /* src/meshify.c:171-172 */
		return EXIT_FAILURE;
	size_t n = (size_t)nx * ny * nz;
	float *tmp = (float *)malloc(n * sizeof(float));
	smooth_axis(img, tmp, nx, ny, nz, 0); /* x: :184-192 */
	smooth_axis(tmp, img, nx, ny, nz, 1); /* y: :195-203 */
	memcpy(tmp, img, n * sizeof(float));
	smooth_axis(tmp, img, nx, ny, nz, 2); /* z: :205-213 */
	free(tmp);
	return EXIT_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* connected components: src/bwlabel.c:54-277, 429-543                       */
/* ------------------------------------------------------------------------ */

static uint32_t uf_find(uint32_t *par, uint32_t a)
{
	while (par[a] != a) {
		par[a] = par[par[a]];
		a = par[a];
	}
	return a;
}

/* Components of bw under conn-connectivity; labels[] gets 1..n numbered by
 * the first voxel of each component in x-fastest order, 0 for background
 * (the numbering translate_labels() produces, src/bwlabel.c:249-277). */
int orc_label_components(const uint8_t *bw, const size_t dim[3], int conn, uint32_t *labels)
{
	size_t nx = dim[0], ny = dim[1], nz = dim[2], n = nx * ny * nz;
	uint32_t *par = (uint32_t *)malloc(n * sizeof(uint32_t));
	/* backward neighbours (earlier in raster order), src/bwlabel.c:94-131,185-210 */
	static const int off6[3][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
	static const int off18[6][3] = {{-1, -1, 0}, {1, -1, 0}, {-1, 0, -1},
	                                {1, 0, -1},  {0, -1, -1}, {0, 1, -1}};
	static const int off26[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}};
	int nofs = 0;
	int ofs[13][3];
	for (int i = 0; i < 3; i++)
		memcpy(ofs[nofs++], off6[i], sizeof(ofs[0]));
	if (conn >= 18)
		for (int i = 0; i < 6; i++)
			memcpy(ofs[nofs++], off18[i], sizeof(ofs[0]));
	if (conn == 26)
		for (int i = 0; i < 4; i++)
			memcpy(ofs[nofs++], off26[i], sizeof(ofs[0]));
	for (size_t v = 0; v < n; v++)
		par[v] = (uint32_t)v;
	for (size_t z = 0; z < nz; z++)
		for (size_t y = 0; y < ny; y++)
			for (size_t x = 0; x < nx; x++) {
				size_t v = x + nx * (y + ny * z);
				if (!bw[v])
					continue;
				for (int k = 0; k < nofs; k++) {
					long xx = (long)x + ofs[k][0], yy = (long)y + ofs[k][1], zz = (long)z + ofs[k][2];
					if (xx < 0 || yy < 0 || zz < 0 || xx >= (long)nx || yy >= (long)ny)
						continue;
					size_t u = (size_t)xx + nx * ((size_t)yy + ny * (size_t)zz);
					if (!bw[u])
						continue;
					uint32_t ra = uf_find(par, (uint32_t)v), rb = uf_find(par, (uint32_t)u);
					if (ra < rb)
						par[rb] = ra;
					else if (rb < ra)
						par[ra] = rb;
				}
			}
	/* root = smallest index of the component; number roots in index order */
	uint32_t next = 0;
	for (size_t v = 0; v < n; v++) {
		if (!bw[v]) {
			labels[v] = 0;
			continue;
		}
		uint32_t r = uf_find(par, (uint32_t)v);
		if (r == v)
			labels[v] = ++next;
		else
			labels[v] = labels[r]; /* r < v, already numbered */
	}
	free(par);
	return (int)next;
}

/* bwlabelCore(): src/bwlabel.c:429-476 */
static int bwlabel_core(float *img, int conn, const size_t dim[3], bool onlyLargest)
{
	if (conn != 6 && conn != 18 && conn != 26) {
		fprintf(stderr, "bwlabel: conn must be 6, 18 or 26.\n");
		return 0;
	}
	if (dim[0] < 2 || dim[1] < 2 || dim[2] < 1) {
		fprintf(stderr, "bwlabel: img must be 2 or 3-dimensional\n");
		return 0;
	}
	size_t n = dim[0] * dim[1] * dim[2];
	uint8_t *bw = (uint8_t *)malloc(n);
	for (size_t i = 0; i < n; i++)
		bw[i] = img[i] != 0.0f;
	uint32_t *l = (uint32_t *)malloc(n * sizeof(uint32_t));
	int nl = orc_label_components(bw, dim, conn, l);
	free(bw);
	if (nl > 0 && onlyLargest) {
		uint32_t *cnt = (uint32_t *)calloc((size_t)nl + 1, sizeof(uint32_t));
		for (size_t i = 0; i < n; i++)
			cnt[l[i]]++;
		uint32_t best = 0, bestN = 0;
		for (int j = 1; j <= nl; j++)
			if (cnt[j] > bestN) { /* strict: ties keep the lowest label, :463 */
				bestN = cnt[j];
				best = (uint32_t)j;
			}
		for (size_t i = 0; i < n; i++)
			l[i] = (l[i] == best);
		nl = 1;
		free(cnt);
	}
	for (size_t i = 0; i < n; i++)
		img[i] = (float)l[i];
	free(l);
	return nl;
}

/* bwlabel(): src/bwlabel.c:478-543 */
int orc_bwlabel(float *img, int conn, size_t dim[3], bool onlyLargest, bool fillBubbles)
{
	if (!fillBubbles)
		return bwlabel_core(img, conn, dim, onlyLargest);
	size_t nx = dim[0], ny = dim[1], nz = dim[2], n = nx * ny * nz;
	float *inv = (float *)malloc(n * sizeof(float));
	for (size_t i = 0; i < n; i++)
		inv[i] = img[i] == 0.0f ? 1.0f : 0.0f;
	int nInv = bwlabel_core(inv, 6, dim, false);
	if (nInv <= 1) {
		free(inv);
		return bwlabel_core(img, conn, dim, onlyLargest);
	}
	/* background labels present on any face are "outside" */
	uint8_t *outside = (uint8_t *)calloc((size_t)nInv + 1, 1);
	for (size_t z = 0; z < nz; z++)
		for (size_t y = 0; y < ny; y++)
			for (size_t x = 0; x < nx; x++) {
				if (!(x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1))
					continue;
				outside[(uint32_t)lroundf(inv[x + nx * (y + ny * z)])] = 1;
			}
	for (size_t i = 0; i < n; i++)
		if (!outside[(uint32_t)lroundf(inv[i])])
			img[i] = 1.0f;
	free(outside);
	free(inv);
	return bwlabel_core(img, conn, dim, onlyLargest);
}

/* dilate(img, dim, true): src/meshify.c:218-264.  The neighbour list starts
 * at its second entry, so offset (-1,-1,-1) is never tested (25 neighbours);
 * only voxels at least one away from every face can be switched on. */
void orc_dilate(float *img, const size_t dim[3])
{
	size_t nx = dim[0], ny = dim[1], nz = dim[2], n = nx * ny * nz;
	uint8_t *on = (uint8_t *)calloc(n, 1);
	for (size_t z = 1; z + 1 < nz; z++)
		for (size_t y = 1; y + 1 < ny; y++)
			for (size_t x = 1; x + 1 < nx; x++) {
				size_t v = x + nx * (y + ny * z);
				for (int dz = -1; dz <= 1; dz++)
					for (int dy = -1; dy <= 1; dy++)
						for (int dx = -1; dx <= 1; dx++) {
							if ((dx == 0 && dy == 0 && dz == 0) || (dx == -1 && dy == -1 && dz == -1))
								continue;
							if (img[v + dx + (long)nx * (dy + (long)ny * dz)] > 0)
								on[v] = 1;
						}
			}
	for (size_t v = 1; v < n; v++)
		if (on[v])
			img[v] = 1;
	free(on);
}

/* ------------------------------------------------------------------------ */
/* classic marching cubes: src/oldcubes.c:22-40, 50-463, 465-522             */
/* ------------------------------------------------------------------------ */

/* corner p of a cube at (x,y,z): x+dx[p], y+dy[p], z+dz[p] (oldcubes.c:419-426) */
static const int kCornerX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int kCornerY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int kCornerZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
/* edge e joins corners (a,b), interpolated from a toward b (oldcubes.c:428-451) */
static const int kEdgeA[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
static const int kEdgeB[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

int orc_mc_classic(const float *img, const short dim[3], const int lo[3], const int hi[3],
                   float isolevel, vec3d **vs, vec3i **ts, int *nv, int *nt)
{
	const int NX = dim[0], NY = dim[1];
	const size_t NXY = (size_t)NX * NY;
	const double iso = isolevel;
	size_t cap = 1 << 16, npt = 0;
	vec3d *pts = (vec3d *)malloc(cap * sizeof(vec3d));
	for (int z = lo[2]; z < hi[2] - 1; z++)
		for (int y = lo[1]; y < hi[1] - 1; y++)
			for (int x = lo[0]; x < hi[0] - 1; x++) {
				double val[8];
				int cfg = 0;
				for (int p = 0; p < 8; p++) {
					val[p] = img[(x + kCornerX[p]) + (size_t)(y + kCornerY[p]) * NX +
					             (size_t)(z + kCornerZ[p]) * NXY];
					if (val[p] < iso)
						cfg |= 1 << p;
				}
				uint64_t word = n2m_classic_tri[cfg];
				if (npt + 16 > cap) {
					cap *= 2;
					pts = (vec3d *)realloc(pts, cap * sizeof(vec3d));
				}
				for (int k = 0; k < 15; k++) {
					int e = (int)((word >> (4 * k)) & 0xF);
					if (e == 0xF)
						break;
					int a = kEdgeA[e], b = kEdgeB[e];
					double mu = (iso - val[a]) / (val[b] - val[a]);
					double ax = x + kCornerX[a], ay = y + kCornerY[a], az = z + kCornerZ[a];
					double bx = x + kCornerX[b], by = y + kCornerY[b], bz = z + kCornerZ[b];
					vec3d q;
					q.x = ax + mu * (bx - ax);
					q.y = ay + mu * (by - ay);
					q.z = az + mu * (bz - az);
					pts[npt++] = q;
				}
			}
	if (npt < 3) {
		free(pts);
		printf("marching cubes failed to identify triangles with an isolevel of %g\n", isolevel);
		return EXIT_FAILURE;
	}
	*vs = (vec3d *)realloc(pts, npt * sizeof(vec3d));
	int ntri = (int)(npt / 3);
	*ts = (vec3i *)malloc((size_t)ntri * sizeof(vec3i));
	for (int i = 0; i < ntri; i++) {
		(*ts)[i].x = 3 * i;
		(*ts)[i].y = 3 * i + 1;
		(*ts)[i].z = 3 * i + 2;
	}
	*nv = (int)npt;
	*nt = ntri;
	return EXIT_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Lewiner MC33: src/MarchingCubes.c:116-155, 235-453, 458-859, 935-1143     */
/* ------------------------------------------------------------------------ */

/* face test, src/MarchingCubes.c:276-295; corner quadruples (A,B,C,D) */
static const int kFace[6][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                                {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};

static int lw_face(const float *c, int face)
{
	int f = face < 0 ? -face : face;
	float A = 0, B = 0, C = 0, D = 0;
	if (f >= 1 && f <= 6) {
		A = c[kFace[f - 1][0]];
		B = c[kFace[f - 1][1]];
		C = c[kFace[f - 1][2]];
		D = c[kFace[f - 1][3]];
	}
	if (fabs(A * C - B * D) < FLT_EPSILON)
		return face >= 0;
	return (float)face * A * (A * C - B * D) >= 0;
}

/* interior test along a reference edge, src/MarchingCubes.c:337-420:
 * t = c[p]/(c[p]-c[q]); B/C/D interpolated on the three parallel edges. */
static const int kInteriorEdge[12][8] = {
    {0, 1, 3, 2, 7, 6, 4, 5}, {1, 2, 0, 3, 4, 7, 5, 6}, {2, 3, 1, 0, 5, 4, 6, 7},
    {3, 0, 2, 1, 6, 5, 7, 4}, {4, 5, 7, 6, 3, 2, 0, 1}, {5, 6, 4, 7, 0, 3, 1, 2},
    {6, 7, 5, 4, 1, 0, 2, 3}, {7, 4, 6, 5, 2, 1, 3, 0}, {0, 4, 3, 7, 2, 6, 1, 5},
    {1, 5, 0, 4, 3, 7, 2, 6}, {2, 6, 1, 5, 0, 4, 3, 7}, {3, 7, 2, 6, 1, 5, 0, 4}};

static int lw_interior(const float *c, int mc33case, int config, int subconfig, int s)
{
	float t, At = 0, Bt = 0, Ct = 0, Dt = 0, a, b;
	int edge = -1;
	switch (mc33case) {
	case 4:
	case 10: /* :310-322 */
		a = (c[4] - c[0]) * (c[6] - c[2]) - (c[7] - c[3]) * (c[5] - c[1]);
		b = c[2] * (c[4] - c[0]) + c[0] * (c[6] - c[2]) - c[1] * (c[7] - c[3]) - c[3] * (c[5] - c[1]);
		t = -b / (2 * a);
		if (t < 0 || t > 1)
			return s > 0;
		At = c[0] + (c[4] - c[0]) * t;
		Bt = c[3] + (c[7] - c[3]) * t;
		Ct = c[2] + (c[6] - c[2]) * t;
		Dt = c[1] + (c[5] - c[1]) * t;
		break;
	case 6:
	case 7:
	case 12:
	case 13: /* :324-423 */
		if (mc33case == 6)
			edge = n2m_lw_test6[config * 3 + 2];
		else if (mc33case == 7)
			edge = n2m_lw_test7[config * 5 + 4];
		else if (mc33case == 12)
			edge = n2m_lw_test12[config * 4 + 3];
		else
			edge = n2m_lw_edges[n2m_lw_row_off[N2M_LW_T13_5_1 + config * 4 + subconfig]];
		if (edge >= 0 && edge < 12) {
			const int *e = kInteriorEdge[edge];
			t = c[e[0]] / (c[e[0]] - c[e[1]]);
			At = 0;
			Bt = c[e[2]] + (c[e[3]] - c[e[2]]) * t;
			Ct = c[e[4]] + (c[e[5]] - c[e[4]]) * t;
			Dt = c[e[6]] + (c[e[7]] - c[e[6]]) * t;
		} else {
			printf("Invalid edge %d\n", edge);
		}
		break;
	default:
		printf("Invalid ambiguous case %d\n", mc33case);
		break;
	}
	int test = (At >= 0) + 2 * (Bt >= 0) + 4 * (Ct >= 0) + 8 * (Dt >= 0);
	switch (test) { /* :432-452 */
	case 5:
		if (At * Ct - Bt * Dt < FLT_EPSILON)
			return s > 0;
		return s < 0;
	case 10:
		if (At * Ct - Bt * Dt >= FLT_EPSILON)
			return s > 0;
		return s < 0;
	case 7:
	case 11:
	case 13:
	case 14:
	case 15:
		return s < 0;
	default:
		return s > 0;
	}
}

/* process_cube() decision tree (src/MarchingCubes.c:458-795): returns the
 * tiling row of mc_tables.h, or -1 for no triangles. */
int orc_lewiner_select(const float c[8], int lut, int originalMC, int *mc33case)
{
	if (originalMC) {
		if (mc33case)
			*mc33case = -1;
		return N2M_LW_CLASSIC + lut;
	}
	int cs = n2m_lw_case[lut], cf = n2m_lw_config[lut], sub = 0;
	if (mc33case)
		*mc33case = cs;
	switch (cs) {
	case 0:
		return -1;
	case 1:
		return N2M_LW_T1 + cf;
	case 2:
		return N2M_LW_T2 + cf;
	case 3:
		return lw_face(c, n2m_lw_test3[cf]) ? N2M_LW_T3_2 + cf : N2M_LW_T3_1 + cf;
	case 4:
		return lw_interior(c, 4, cf, 0, n2m_lw_test4[cf]) ? N2M_LW_T4_1 + cf : N2M_LW_T4_2 + cf;
	case 5:
		return N2M_LW_T5 + cf;
	case 6:
		if (lw_face(c, n2m_lw_test6[cf * 3 + 0]))
			return N2M_LW_T6_2 + cf;
		return lw_interior(c, 6, cf, 0, n2m_lw_test6[cf * 3 + 1]) ? N2M_LW_T6_1_1 + cf
		                                                           : N2M_LW_T6_1_2 + cf;
	case 7:
		for (int i = 0; i < 3; i++)
			if (lw_face(c, n2m_lw_test7[cf * 5 + i]))
				sub += 1 << i;
		switch (sub) {
		case 0: return N2M_LW_T7_1 + cf;
		case 1: return N2M_LW_T7_2 + cf * 3 + 0;
		case 2: return N2M_LW_T7_2 + cf * 3 + 1;
		case 3: return N2M_LW_T7_3 + cf * 3 + 0;
		case 4: return N2M_LW_T7_2 + cf * 3 + 2;
		case 5: return N2M_LW_T7_3 + cf * 3 + 1;
		case 6: return N2M_LW_T7_3 + cf * 3 + 2;
		default:
			return lw_interior(c, 7, cf, sub, n2m_lw_test7[cf * 5 + 3]) ? N2M_LW_T7_4_2 + cf
			                                                            : N2M_LW_T7_4_1 + cf;
		}
	case 8:
		return N2M_LW_T8 + cf;
	case 9:
		return N2M_LW_T9 + cf;
	case 10:
	case 12: {
		const int8_t *tt = cs == 10 ? &n2m_lw_test10[cf * 3] : &n2m_lw_test12[cf * 4];
		int f0 = lw_face(c, tt[0]), f1 = lw_face(c, tt[1]);
		int b11 = cs == 10 ? N2M_LW_T10_1_1 : N2M_LW_T12_1_1;
		int b11_ = cs == 10 ? N2M_LW_T10_1_1_ : N2M_LW_T12_1_1_;
		int b12 = cs == 10 ? N2M_LW_T10_1_2 : N2M_LW_T12_1_2;
		int b2 = cs == 10 ? N2M_LW_T10_2 : N2M_LW_T12_2;
		int b2_ = cs == 10 ? N2M_LW_T10_2_ : N2M_LW_T12_2_;
		if (f0)
			return (f1 ? b11_ : b2) + cf;
		if (f1)
			return b2_ + cf;
		return (lw_interior(c, cs, cf, 0, tt[2]) ? b11 : b12) + cf;
	}
	case 11:
		return N2M_LW_T11 + cf;
	case 13: {
		for (int i = 0; i < 6; i++)
			if (lw_face(c, n2m_lw_test13[cf * 7 + i]))
				sub += 1 << i;
		int s = n2m_lw_subconfig13[sub];
		if (s == 0)
			return N2M_LW_T13_1 + cf;
		if (s >= 1 && s <= 6)
			return N2M_LW_T13_2 + cf * 6 + (s - 1);
		if (s >= 7 && s <= 18)
			return N2M_LW_T13_3 + cf * 12 + (s - 7);
		if (s >= 19 && s <= 22)
			return N2M_LW_T13_4 + cf * 4 + (s - 19);
		if (s >= 23 && s <= 26) {
			int k = s - 23;
			return lw_interior(c, 13, cf, k, n2m_lw_test13[cf * 7 + 6]) ? N2M_LW_T13_5_1 + cf * 4 + k
			                                                            : N2M_LW_T13_5_2 + cf * 4 + k;
		}
		if (s >= 27 && s <= 38)
			return N2M_LW_T13_3_ + cf * 12 + (s - 27);
		if (s >= 39 && s <= 44)
			return N2M_LW_T13_2_ + cf * 6 + (s - 39);
		if (s == 45)
			return N2M_LW_T13_1_ + cf;
		printf("Marching Cubes: Impossible case 13?\n");
		return -1;
	}
	case 14:
		return N2M_LW_T14 + cf;
	}
	return -1;
}

static float lw_clamp(float v)
{
	return fabs(v) < FLT_EPSILON ? FLT_EPSILON : v; /* src/MarchingCubes.c:132,252-255 */
}

/* Provenance of a Lewiner vertex (differentiable path): an edge vertex
 * records its sub-box point and axis as point*4+axis; a centre vertex
 * records -(1 + k) with its cube's 12 edge-vertex ids (-1 = none) in
 * cedges[12k..12k+11].  exact != 0 computes positions in double (x + u, u =
 * c0/(c0-c1) from the float corner values) instead of the reference's float
 * sums. */
static int lewiner_core(const float *img, const short dim[3], const int lo[3], const int hi[3], int originalMC,
                        float isolevel, int exact, vec3d **vs, vec3i **ts, int *nv, int *nt, long long **prov,
                        int **cedges, int *ncentre)
{
	/* sub-box copy as img-iso, pre-filled with min(img)-iso; the index only
	 * guards the end of the volume, so an extra column/row wraps (:1086-1115) */
	const int SX = hi[0] - lo[0] + 1, SY = hi[1] - lo[1] + 1, SZ = hi[2] - lo[2] + 1;
	const int inX = dim[0], inXY = dim[0] * dim[1], inXYZ = dim[0] * dim[1] * dim[2];
	const size_t ns = (size_t)SX * SY * SZ;
	float mn = img[0];
	for (int i = 0; i < inXYZ; i++)
		mn = (float)fmin(mn, img[i]);
	mn = mn - isolevel;
	float *d = (float *)malloc(ns * sizeof(float));
	size_t i = 0;
	for (int z = 0; z < SZ; z++)
		for (int y = 0; y < SY; y++) {
			int zy = (y + lo[1]) * inX + (z + lo[2]) * inXY;
			for (int x = 0; x < SX; x++, i++) {
				int j = lo[0] + x + zy;
				d[i] = j < inXYZ ? lw_clamp(img[j] - isolevel) : lw_clamp(mn);
			}
		}
#define D(x, y, z) d[(size_t)(x) + (size_t)SX * ((size_t)(y) + (size_t)SY * (size_t)(z))]
	/* edge vertices in point raster order, x then y then z edge (:235-270) */
	int *vid = (int *)malloc(3 * ns * sizeof(int));
	size_t vcap = 1 << 16, nverts = 0, ncen = 0;
	float *vp = (float *)malloc(vcap * 3 * sizeof(float));
	double *vx = (double *)malloc(vcap * 3 * sizeof(double)); /* exact positions */
	long long *pv = (long long *)malloc(vcap * sizeof(long long));
	size_t ccap = 1 << 12;
	int *ce = (int *)malloc(ccap * 12 * sizeof(int));
	i = 0;
	for (int z = 0; z < SZ; z++)
		for (int y = 0; y < SY; y++)
			for (int x = 0; x < SX; x++, i++) {
				float c0 = D(x, y, z);
				float cn[3];
				cn[0] = x < SX - 1 ? D(x + 1, y, z) : c0;
				cn[1] = y < SY - 1 ? D(x, y + 1, z) : c0;
				cn[2] = z < SZ - 1 ? D(x, y, z + 1) : c0;
				for (int a = 0; a < 3; a++) {
					int cross = c0 < 0 ? cn[a] > 0 : cn[a] < 0;
					vid[3 * i + a] = -1;
					if (!cross)
						continue;
					float denom = c0 - cn[a];
					float u = denom != 0.0f ? c0 / denom : 0.5f; /* :944-945 */
					if (nverts + 1 > vcap) {
						vcap *= 2;
						vp = (float *)realloc(vp, vcap * 3 * sizeof(float));
						vx = (double *)realloc(vx, vcap * 3 * sizeof(double));
						pv = (long long *)realloc(pv, vcap * sizeof(long long));
					}
					vp[3 * nverts + 0] = a == 0 ? (float)x + u : (float)x;
					vp[3 * nverts + 1] = a == 1 ? (float)y + u : (float)y;
					vp[3 * nverts + 2] = a == 2 ? (float)z + u : (float)z;
					{
						double dd = (double)c0 - (double)cn[a];
						double ux = dd != 0.0 ? (double)c0 / dd : 0.5;
						vx[3 * nverts + 0] = a == 0 ? (double)x + ux : (double)x;
						vx[3 * nverts + 1] = a == 1 ? (double)y + ux : (double)y;
						vx[3 * nverts + 2] = a == 2 ? (double)z + ux : (double)z;
					}
					pv[nverts] = (long long)i * 4 + a;
					vid[3 * i + a] = (int)nverts++;
				}
			}
	/* cube edge code -> (point offset, axis), src/MarchingCubes.c:811-824 */
	static const int kCodePt[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1},
	                                   {0, 0, 1, 0}, {1, 0, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1},
	                                   {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
	size_t tcap = 1 << 16, ntr = 0;
	int *tr = (int *)malloc(tcap * 3 * sizeof(int));
	for (int z = 0; z < SZ - 1; z++)
		for (int y = 0; y < SY - 1; y++)
			for (int x = 0; x < SX - 1; x++) {
				float c[8];
				int lut = 0;
				for (int p = 0; p < 8; p++) {
					c[p] = D(x + kCornerX[p], y + kCornerY[p], z + kCornerZ[p]);
					if (c[p] > 0)
						lut += 1 << p;
				}
				int row = orc_lewiner_select(c, lut, originalMC, NULL);
				if (row < 0)
					continue;
				int cid = -1;
				if (!originalMC && n2m_lw_row_centre[row]) { /* add_c_vertex :1029-1082 */
					float sx = 0, sy = 0, sz = 0, cnt = 0;
					double ex = 0, ey = 0, ez = 0;
					if (ncen + 1 > ccap) {
						ccap *= 2;
						ce = (int *)realloc(ce, ccap * 12 * sizeof(int));
					}
					for (int e = 0; e < 12; e++) {
						const int *q = kCodePt[e];
						size_t pi = (size_t)(x + q[0]) + (size_t)SX * ((size_t)(y + q[1]) + (size_t)SY * (size_t)(z + q[2]));
						int id = vid[3 * pi + q[3]];
						ce[12 * ncen + e] = id;
						if (id == -1)
							continue;
						++cnt;
						sx += vp[3 * id];
						sy += vp[3 * id + 1];
						sz += vp[3 * id + 2];
						ex += vx[3 * id];
						ey += vx[3 * id + 1];
						ez += vx[3 * id + 2];
					}
					if (cnt > 0) {
						sx /= cnt;
						sy /= cnt;
						sz /= cnt;
						ex /= cnt;
						ey /= cnt;
						ez /= cnt;
					}
					if (nverts + 1 > vcap) {
						vcap *= 2;
						vp = (float *)realloc(vp, vcap * 3 * sizeof(float));
						vx = (double *)realloc(vx, vcap * 3 * sizeof(double));
						pv = (long long *)realloc(pv, vcap * sizeof(long long));
					}
					vp[3 * nverts] = sx;
					vp[3 * nverts + 1] = sy;
					vp[3 * nverts + 2] = sz;
					vx[3 * nverts] = ex;
					vx[3 * nverts + 1] = ey;
					vx[3 * nverts + 2] = ez;
					pv[nverts] = -(long long)(1 + ncen);
					ncen++;
					cid = (int)nverts++;
				}
				int ntri = n2m_lw_row_ntri[row];
				const uint8_t *codes = &n2m_lw_edges[n2m_lw_row_off[row]];
				if (ntr + (size_t)ntri > tcap) {
					while (ntr + (size_t)ntri > tcap)
						tcap *= 2;
					tr = (int *)realloc(tr, tcap * 3 * sizeof(int));
				}
				for (int k = 0; k < 3 * ntri; k++) {
					int e = codes[k], id;
					if (e == 12) {
						id = cid;
					} else {
						const int *q = kCodePt[e];
						size_t pi = (size_t)(x + q[0]) + (size_t)SX * ((size_t)(y + q[1]) + (size_t)SY * (size_t)(z + q[2]));
						id = vid[3 * pi + q[3]];
					}
					if (id == -1)
						printf("Marching Cubes: invalid triangle %zu\n", ntr + 1 + k / 3);
					tr[3 * ntr + k] = id;
				}
				ntr += (size_t)ntri;
			}
#undef D
	free(d);
	free(vid);
	if (nverts < 3 || ntr < 1) {
		free(vp);
		free(vx);
		free(pv);
		free(ce);
		free(tr);
		return EXIT_FAILURE;
	}
	*vs = (vec3d *)malloc(nverts * sizeof(vec3d));
	for (size_t k = 0; k < nverts; k++) {
		if (exact) {
			(*vs)[k].x = vx[3 * k] + (double)lo[0];
			(*vs)[k].y = vx[3 * k + 1] + (double)lo[1];
			(*vs)[k].z = vx[3 * k + 2] + (double)lo[2];
		} else { /* float + (float)int, :1126-1130 */
			(*vs)[k].x = vp[3 * k] + (float)lo[0];
			(*vs)[k].y = vp[3 * k + 1] + (float)lo[1];
			(*vs)[k].z = vp[3 * k + 2] + (float)lo[2];
		}
	}
	if (prov) {
		*prov = pv;
		*cedges = ce;
		*ncentre = (int)ncen;
	} else {
		free(pv);
		free(ce);
	}
	free(vx);
	*ts = (vec3i *)malloc(ntr * sizeof(vec3i));
	for (size_t k = 0; k < ntr; k++) { /* winding v3,v2,v1, :1133-1137 */
		(*ts)[k].x = tr[3 * k + 2];
		(*ts)[k].y = tr[3 * k + 1];
		(*ts)[k].z = tr[3 * k];
	}
	free(vp);
	free(tr);
	*nv = (int)nverts;
	*nt = (int)ntr;
	return EXIT_SUCCESS;
}

int orc_mc_lewiner(const float *img, const short dim[3], const int lo[3], const int hi[3],
                   int originalMC, float isolevel, vec3d **vs, vec3i **ts, int *nv, int *nt)
{
	return lewiner_core(img, dim, lo, hi, originalMC, isolevel, 0, vs, ts, nv, nt, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* vertex welding: src/meshify.c:45-106 (+ radix11sort_f32, radixsort.c)    */
/* ------------------------------------------------------------------------ */

static double dist3(vec3d a, vec3d b)
{
	double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
	return sqrt(dx * dx + dy * dy + dz * dz);
}

typedef struct {
	float key;
	uint32_t idx;
} keyidx;

/* The reference radix-sorts the float keys by their "flipped" bit pattern
 * (src/radixsort.c:31-35, the usual sign-magnitude to unsigned map), a total
 * order that also places NaN keys.  A NaN key is ordered without its sign
 * (after every number, by payload): the sign a NaN takes through a
 * difference is not portable, and positive NaN voxels give the reference
 * positive NaN keys. */
static uint32_t key_order(float f)
{
	uint32_t u;
	memcpy(&u, &f, 4);
	if (f != f)
		return u | 0x80000000u;
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static int keyidx_cmp(const void *pa, const void *pb)
{
	const keyidx *a = (const keyidx *)pa, *b = (const keyidx *)pb;
	const uint32_t ka = key_order(a->key), kb = key_order(b->key);
	if (ka != kb)
		return ka < kb ? -1 : 1;
	return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* Weld; cls_out (npt entries, may be NULL) receives each vertex's class
 * (identity when nothing merged). */
static int weld_core(vec3d **inpt, vec3i *tris, int npt, int ntri, int *cls_out)
{
	vec3d *pts = *inpt;
	/* stable sort by FP32 distance to vertex 0: ties keep index order */
	keyidx *s = (keyidx *)malloc((size_t)npt * sizeof(keyidx));
	for (int i = 0; i < npt; i++) {
		s[i].key = (float)dist3(pts[0], pts[i]);
		s[i].idx = (uint32_t)i;
	}
	qsort(s, (size_t)npt, sizeof(keyidx), keyidx_cmp);
	int *cls = (int *)malloc((size_t)npt * sizeof(int));
	for (int i = 0; i < npt; i++)
		cls[i] = -1;
	const float tol = 0.00001f; /* :65 */
	int nnew = 0;
	for (int i = 0; i < npt; i++) { /* greedy window, :66-79 */
		if (cls[s[i].idx] >= 0)
			continue;
		vec3d seed = pts[s[i].idx];
		/* The seed always joins its own class.  For a finite position this
		 * is what :72-74 do (distance 0 < tol); a NaN position matches
		 * nothing, itself included, and the reference then leaves
		 * old2new at -1 and stores to pts[-1] at :99-100. */
		cls[s[i].idx] = nnew;
		for (int j = i + 1; j < npt && (s[j].key - s[i].key) < tol; j++)
			if (dist3(seed, pts[s[j].idx]) < tol)
				cls[s[j].idx] = nnew;
		nnew++;
	}
	free(s);
	if (nnew == npt) { /* nothing merged: unchanged, :82-87 */
		if (cls_out)
			for (int i = 0; i < npt; i++)
				cls_out[i] = i;
		free(cls);
		return npt;
	}
	/* A corner without a vertex (-1: the Lewiner tables found no vertex on
	 * the edge, src/MarchingCubes.c:829-833) stays -1; the reference reads
	 * old2new[-1] there (src/meshify.c:88-92). */
	for (int i = 0; i < ntri; i++) {
		tris[i].x = (tris[i].x >= 0 && tris[i].x < npt) ? cls[tris[i].x] : -1;
		tris[i].y = (tris[i].y >= 0 && tris[i].y < npt) ? cls[tris[i].y] : -1;
		tris[i].z = (tris[i].z >= 0 && tris[i].z < npt) ? cls[tris[i].z] : -1;
	}
	vec3d *out = (vec3d *)malloc((size_t)nnew * sizeof(vec3d));
	for (int i = 0; i < npt; i++) /* last member in index order wins, :99-100 */
		out[cls[i]] = pts[i];
	free(pts);
	*inpt = out;
	if (cls_out)
		memcpy(cls_out, cls, (size_t)npt * sizeof(int));
	free(cls);
	return nnew;
}

int orc_unify_vertices(vec3d **inpt, vec3i *tris, int npt, int ntri)
{
	return weld_core(inpt, tris, npt, ntri, NULL);
}

/* remove_degenerate_triangles(): src/meshify.c:113-168 */
int orc_remove_degenerate_triangles(vec3d *pts, vec3i **intris, int ntri)
{
	vec3i *tris = *intris;
	int keep = 0;
	uint8_t *bad = (uint8_t *)calloc((size_t)ntri + 1, 1);
	for (int i = 0; i < ntri; i++) {
		if (tris[i].x < 0 || tris[i].y < 0 || tris[i].z < 0) {
			bad[i] = 1; /* a corner without a vertex (the reference reads pts[-1]) */
			continue;
		}
		double l = dist3(pts[tris[i].x], pts[tris[i].y]);
		double m = dist3(pts[tris[i].x], pts[tris[i].z]);
		double n = dist3(pts[tris[i].y], pts[tris[i].z]);
		double c = fmin(fmin(l, m), n);
		double a = fmax(fmax(l, m), n);
		double b = l + m + n - a - c;
		if ((c - (a - b)) <= 0.0) {
			bad[i] = 1;
			continue;
		}
		double area4 = 0.25 * sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)));
		if (area4 < FLT_EPSILON)
			bad[i] = 1;
	}
	for (int i = 0; i < ntri; i++)
		keep += !bad[i];
	if (keep == ntri) {
		free(bad);
		return ntri;
	}
	vec3i *out = (vec3i *)malloc((size_t)(keep > 0 ? keep : 1) * sizeof(vec3i));
	int j = 0;
	for (int i = 0; i < ntri; i++)
		if (!bad[i])
			out[j++] = tris[i];
	free(tris);
	*intris = out;
	free(bad);
	return keep;
}

/* ------------------------------------------------------------------------ */
/* meshify(): src/meshify.c:286-389                                          */
/* ------------------------------------------------------------------------ */
int orc_meshify(float *img, short dim[3], int cubes, int originalMC, float isolevel, vec3i **t,
                vec3d **p, int *nt, int *np, bool preSmooth, bool onlyLargest, bool fillBubbles)
{
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const int nvox = NX * NY * NZ;
	if (preSmooth)
		orc_quick_smooth(img, NX, NY, NZ);
	float mx = img[0], mn = mx;
	for (int i = 0; i < nvox; i++) {
		mx = fmaxf(mx, img[i]);
		mn = fminf(mn, img[i]);
	}
	if (mn == mx) {
		printf("Error: No variability in image intensity.\n");
		return EXIT_FAILURE;
	}
	if ((isolevel <= mn) || (isolevel > mx)) { /* :316-319 */
		isolevel = 0.5 * (mn + mx);
		printf("Suggested isolevel out of range. Intensity range %g..%g, setting isolevel to %g\n",
		       mn, mx, isolevel);
	}
	if (onlyLargest || fillBubbles) {
		size_t sdim[3] = {(size_t)NX, (size_t)NY, (size_t)NZ};
		float *mask = (float *)malloc((size_t)nvox * sizeof(float));
		for (int i = 0; i < nvox; i++)
			mask[i] = img[i] >= isolevel ? 1.0f : 0.0f;
		orc_bwlabel(mask, 18, sdim, onlyLargest, fillBubbles);
		if (fillBubbles)
			for (int i = 0; i < nvox; i++)
				if (mask[i] != 0)
					img[i] = (float)fmax(img[i], isolevel);
		if (onlyLargest) {
			orc_dilate(mask, sdim);
			for (int i = 0; i < nvox; i++)
				if (mask[i] == 0)
					img[i] = mn;
		}
		free(mask);
	}
	float edgeMax = 0.75 * (mn + isolevel); /* :348 */
	int lo[3] = {NX, NY, NZ}, hi[3] = {0, 0, 0};
	int v = 0;
	for (int z = 0; z < NZ; z++)
		for (int y = 0; y < NY; y++)
			for (int x = 0; x < NX; x++, v++) {
				if (img[v] >= isolevel) {
					lo[0] = ORC_MIN(x, lo[0]);
					lo[1] = ORC_MIN(y, lo[1]);
					lo[2] = ORC_MIN(z, lo[2]);
					hi[0] = ORC_MAX(x, hi[0]);
					hi[1] = ORC_MAX(y, hi[1]);
					hi[2] = ORC_MAX(z, hi[2]);
				}
				if (x == 0 || y == 0 || z == 0 || x == NX - 1 || y == NY - 1 || z == NZ - 1)
					img[v] = fminf(edgeMax, img[v]);
			}
	for (int i = 0; i < 3; i++) {
		lo[i] = ORC_MAX(lo[i] - 1, 0);
		hi[i] = ORC_MIN(hi[i] + 2, dim[i]);
	}
	vec3d *pts = NULL;
	vec3i *tris = NULL;
	int ntri = 0, npt = 0;
	int rc = cubes == ORC_CUBES_LEWINER
	             ? orc_mc_lewiner(img, dim, lo, hi, originalMC, isolevel, &pts, &tris, &npt, &ntri)
	             : orc_mc_classic(img, dim, lo, hi, isolevel, &pts, &tris, &npt, &ntri);
	if (rc != EXIT_SUCCESS)
		return EXIT_FAILURE;
	npt = orc_unify_vertices(&pts, tris, npt, ntri);
	if (npt < 3)
		return EXIT_FAILURE;
	ntri = orc_remove_degenerate_triangles(pts, &tris, ntri);
	*t = tris;
	*p = pts;
	*nt = ntri;
	*np = npt;
	return EXIT_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Laplacian / HC smoothing: src/quadric.c:144-215, 316-394                  */
/* ------------------------------------------------------------------------ */

/* laplacian_smooth(): src/quadric.c:316-341 */
void orc_laplacian_smooth(vec3d *verts, const vec3i *tris, int nvert, int ntri)
{
	vec3d *sum = (vec3d *)calloc((size_t)nvert, sizeof(vec3d));
	int *num = (int *)calloc((size_t)nvert, sizeof(int));
	for (int i = 0; i < ntri; i++) {
		int c[3] = {tris[i].x, tris[i].y, tris[i].z};
		for (int k = 0; k < 3; k++) {
			/* the other two corners, lower corner first */
			int a = c[k == 0 ? 1 : 0], b = c[k == 2 ? 1 : 2];
			num[c[k]] += 2;
			sum[c[k]].x = sum[c[k]].x + (verts[a].x + verts[b].x);
			sum[c[k]].y = sum[c[k]].y + (verts[a].y + verts[b].y);
			sum[c[k]].z = sum[c[k]].z + (verts[a].z + verts[b].z);
		}
	}
	for (int i = 0; i < nvert; i++) {
		if (num[i] <= 0)
			continue;
		verts[i].x = sum[i].x / num[i];
		verts[i].y = sum[i].y / num[i];
		verts[i].z = sum[i].z / num[i];
	}
	free(sum);
	free(num);
}

/* Border vertices as update_mesh(iteration 0) marks them (quadric.c:188-215):
 * for each vertex i, count how often every vertex id occurs in the triangles
 * that reference i (once per reference); ids seen exactly once are border. */
void orc_border_vertices(const vec3i *tris, int nvert, int ntri, uint8_t *border)
{
	int *cnt = (int *)calloc((size_t)nvert + 1, sizeof(int));
	int *start = (int *)calloc((size_t)nvert + 1, sizeof(int));
	int *ref = (int *)malloc((size_t)(3 * ntri + 1) * sizeof(int));
	for (int i = 0; i < ntri; i++) {
		cnt[tris[i].x]++;
		cnt[tris[i].y]++;
		cnt[tris[i].z]++;
	}
	for (int i = 0; i < nvert; i++)
		start[i + 1] = start[i] + cnt[i];
	memset(cnt, 0, (size_t)nvert * sizeof(int));
	for (int i = 0; i < ntri; i++) {
		int c[3] = {tris[i].x, tris[i].y, tris[i].z};
		for (int k = 0; k < 3; k++)
			ref[start[c[k]] + cnt[c[k]]++] = i;
	}
	memset(border, 0, (size_t)nvert);
	int *ids = (int *)malloc((size_t)(9 * ntri + 1) * sizeof(int));
	int *num = (int *)malloc((size_t)(9 * ntri + 1) * sizeof(int));
	for (int v = 0; v < nvert; v++) {
		int n = 0;
		for (int r = start[v]; r < start[v + 1]; r++) {
			int c[3] = {tris[ref[r]].x, tris[ref[r]].y, tris[ref[r]].z};
			for (int k = 0; k < 3; k++) {
				int f = 0;
				while (f < n && ids[f] != c[k])
					f++;
				if (f == n) {
					ids[n] = c[k];
					num[n++] = 1;
				} else {
					num[f]++;
				}
			}
		}
		for (int f = 0; f < n; f++)
			if (num[f] == 1)
				border[ids[f]] = 1;
	}
	free(ids);
	free(num);
	free(ref);
	free(start);
	free(cnt);
}

/* laplacian_smoothHC(): src/quadric.c:343-394 */
void orc_laplacian_smoothHC(vec3d *verts, const vec3i *tris, int nvert, int ntri, double alpha,
                            double beta, int iter, bool lockEdges)
{
	double alpha1 = 1.0 - alpha, beta1 = 1.0 - beta;
	size_t bytes = (size_t)nvert * sizeof(vec3d);
	vec3d *p = (vec3d *)malloc(bytes), *q = (vec3d *)malloc(bytes), *b = (vec3d *)malloc(bytes);
	memcpy(p, verts, bytes);
	for (int it = 0; it < iter; it++) {
		memcpy(q, p, bytes);
		orc_laplacian_smooth(p, tris, nvert, ntri);
		for (int i = 0; i < nvert; i++) {
			b[i].x = p[i].x - (verts[i].x * alpha + q[i].x * alpha1);
			b[i].y = p[i].y - (verts[i].y * alpha + q[i].y * alpha1);
			b[i].z = p[i].z - (verts[i].z * alpha + q[i].z * alpha1);
		}
		memcpy(q, b, bytes);
		orc_laplacian_smooth(q, tris, nvert, ntri);
		for (int i = 0; i < nvert; i++) {
			p[i].x = p[i].x - (b[i].x * beta + q[i].x * beta1);
			p[i].y = p[i].y - (b[i].y * beta + q[i].y * beta1);
			p[i].z = p[i].z - (b[i].z * beta + q[i].z * beta1);
		}
	}
	free(q);
	free(b);
	if (!lockEdges) {
		memcpy(verts, p, bytes);
		free(p);
		return;
	}
	uint8_t *border = (uint8_t *)malloc((size_t)nvert + 1);
	orc_border_vertices(tris, nvert, ntri, border);
	for (int i = 0; i < nvert; i++)
		if (!border[i])
			verts[i] = p[i];
	free(border);
	free(p);
}

/* ------------------------------------------------------------------------ */
/* isolevel: src/isolevel.c:35-277                                           */
/* ------------------------------------------------------------------------ */

/* nifti_robust_range(..., ignoreZeroVoxels=0): src/isolevel.c:35-139 */
int orc_robust_range(const float *img, int nvox, float *pct2, float *pct98)
{
	*pct2 = 0.0f;
	*pct98 = 1.0f;
	if (nvox < 1)
		return 1;
	float mn = INFINITY, mx = -INFINITY;
	size_t nZero = 0, nNan = 0;
	for (int i = 0; i < nvox; i++) {
		if (isnan(img[i])) {
			nNan++;
			continue;
		}
		if (img[i] == 0.0f)
			nZero++;
		mn = (float)fmin(img[i], mn);
		mx = (float)fmax(img[i], mx);
	}
	if (nZero > 0 && mn > 0.0f)
		mn = 0.0f;
	if (mn > mx)
		return 0;
	if (mn == mx) {
		*pct2 = mn;
		*pct98 = mx;
		return 0;
	}
	size_t valid = (size_t)nvox - nNan;
	size_t n2pct = (size_t)round(valid * 0.02);
	if (n2pct < 1 || valid < 100) {
		*pct2 = mn;
		*pct98 = mx;
		return 0;
	}
	enum { kBins = 1001 };
	float scl = (kBins - 1) / (mx - mn);
	int hist[kBins];
	memset(hist, 0, sizeof(hist));
	for (int i = 0; i < nvox; i++) {
		if (isnan(img[i]))
			continue;
		hist[(int)round((img[i] - mn) * scl)]++;
	}
	size_t n = 0, lo = 0;
	while (n < n2pct)
		n += hist[lo++];
	lo--;
	n = 0;
	int hi = kBins;
	while (n < n2pct)
		n += hist[--hi];
	if ((int)lo == hi) { /* :118-134 */
		int ok = -1;
		while (ok != 0) {
			if (lo > 0) {
				lo--;
				if (hist[lo] > 0)
					ok = 0;
			}
			if (ok != 0 && hi < kBins - 1) {
				hi++;
				if (hist[hi] > 0)
					ok = 0;
			}
			if (lo == 0 && hi == kBins - 1)
				ok = 0;
		}
	}
	*pct2 = (lo) / scl + mn;
	*pct98 = (hi) / scl + mn;
	return 0;
}

/* nii_otsu(): src/isolevel.c:141-243 (P[0] holds the raw count H[0]) */
int orc_otsu(const int *H, int nBin, int mode, int *dark, int *mid, int *bright)
{
	*dark = *mid = *bright = 0;
	double Sum = 0.0;
	for (int v = 0; v < nBin; v++)
		Sum = Sum + H[v];
	if (Sum <= 0)
		return 0;
	double *P = (double *)malloc((size_t)nBin * nBin * sizeof(double));
	double *S = (double *)malloc((size_t)nBin * nBin * sizeof(double));
	P[0] = H[0];
	S[0] = H[0];
	for (int v = 1; v < nBin; v++) {
		double prob = H[v] / Sum;
		P[v] = P[v - 1] + prob;
		S[v] = S[v - 1] + (v + 1) * prob;
	}
	for (int u = 1; u < nBin; u++)
		for (int v = u; v < nBin; v++) {
			P[u * nBin + v] = P[v] - P[u - 1];
			S[u * nBin + v] = S[v] - S[u - 1];
		}
	for (int u = 0; u < nBin; u++)
		for (int v = u; v < nBin; v++)
			if (P[u * nBin + v] != 0)
				P[u * nBin + v] = (S[u * nBin + v] * S[u * nBin + v]) / P[u * nBin + v];
	int thresh = 0;
	if (mode == 1 || mode == 5) {
		int lo = (int)(0.25 * nBin), mi = (int)(0.50 * nBin), hi = (int)(0.75 * nBin);
		double best = P[lo] + P[(lo + 1) * nBin + mi] + P[(mi + 1) * nBin + hi] + P[(hi + 1) * nBin + 255];
		for (int l = 0; l < nBin - 3; l++)
			for (int m = l + 1; m < nBin - 2; m++)
				for (int h = m + 1; h < nBin - 1; h++) {
					double v = P[l] + P[(l + 1) * nBin + m] + P[(m + 1) * nBin + h] + P[(h + 1) * nBin + 255];
					if (v > best) {
						lo = l;
						mi = m;
						hi = h;
						best = v;
					}
				}
		thresh = mode == 1 ? hi : lo;
		*dark = lo;
		*mid = mi;
		*bright = hi;
	} else if (mode == 2 || mode == 4) {
		int lo = (int)(0.33 * nBin), hi = (int)(0.67 * nBin);
		double best = P[lo] + P[(lo + 1) * nBin + hi] + P[(hi + 1) * nBin + nBin - 1];
		for (int l = 0; l < nBin - 2; l++)
			for (int h = l + 1; h < nBin - 1; h++) {
				double v = P[l] + P[(l + 1) * nBin + h] + P[(h + 1) * nBin + nBin - 1];
				if (v > best) {
					lo = l;
					hi = h;
					best = v;
				}
			}
		thresh = lo; /* mode is never 1 here, :218-221 */
		*dark = lo;
		*mid = thresh;
		*bright = hi;
	} else {
		thresh = (int)(0.25 * nBin);
		double best = P[thresh] + P[(thresh + 1) * nBin + nBin - 1];
		for (int i = 0; i < nBin - 1; i++) {
			double v = P[i] + P[(i + 1) * nBin + nBin - 1];
			if (v > best) {
				thresh = i;
				best = v;
			}
		}
		*dark = *mid = *bright = thresh;
	}
	free(P);
	free(S);
	return thresh;
}

/* setThreshold(): src/isolevel.c:245-277 */
float orc_set_threshold(const float *img, int nvox, int darkMediumBright123)
{
	float mn, mx;
	if (orc_robust_range(img, nvox, &mn, &mx) != 0)
		return 1;
	enum { kOtsuBins = 256 };
	float scl = (kOtsuBins - 1) / (mx - mn);
	int hist[kOtsuBins];
	memset(hist, 0, sizeof(hist));
	for (int i = 0; i < nvox; i++) {
		if (isnan(img[i]))
			continue;
		int idx = (int)round((img[i] - mn) * scl);
		idx = ORC_MIN(idx, kOtsuBins - 1);
		idx = ORC_MAX(idx, 0);
		hist[idx]++;
	}
	int dark, mid, bright;
	if (darkMediumBright123 == 1 || darkMediumBright123 == 3) {
		orc_otsu(hist, kOtsuBins, 5, &dark, &mid, &bright);
		if (darkMediumBright123 == 1)
			return (dark / scl) + mn;
		return (bright / scl) + mn;
	}
	orc_otsu(hist, kOtsuBins, 3, &dark, &mid, &bright);
	return (mid / scl) + mn;
}

/* apply_sform(): src/meshify.c:1021-1045 */
void orc_apply_sform(vec3i *t, vec3d *p, int nt, int np, const float srow_x[4],
                     const float srow_y[4], const float srow_z[4])
{
	for (int i = 0; i < np; i++) {
		vec3d v = p[i];
		p[i].x = (v.x * srow_x[0]) + (v.y * srow_x[1]) + (v.z * srow_x[2]) + srow_x[3];
		p[i].y = (v.x * srow_y[0]) + (v.y * srow_y[1]) + (v.z * srow_y[2]) + srow_y[3];
		p[i].z = (v.x * srow_z[0]) + (v.y * srow_z[1]) + (v.z * srow_z[2]) + srow_z[3];
	}
	double px = srow_x[0] + srow_x[1] + srow_x[2];
	double py = srow_y[0] + srow_y[1] + srow_y[2];
	double pz = srow_z[0] + srow_z[1] + srow_z[2];
	float det = px * py * pz;
	if (det >= 0.0)
		return;
	for (int i = 0; i < nt; i++) {
		int a = t[i].x;
		t[i].x = t[i].y;
		t[i].y = a;
	}
}

/* ------------------------------------------------------------------------ */
/* Differentiable path (forward: Lewiner MC over the whole volume at the    */
/* isolevel, weld, degenerate removal, lapIter Laplacian iterations) and    */
/* its adjoint in FP64.  The forward is the reference's functions chained   */
/* as src/meshify.c:377-385 and src/quadric.c:316-341 chain them; the      */
/* adjoint is written here from the derivative of each step.               */
/* ------------------------------------------------------------------------ */

/* Transposed laplacian_smooth() (src/quadric.c:316-341): the forward maps
 * p_i -> sum_i / num_i (num_i > 0), with sum_i gathering the other two
 * corners of every triangle slot of i; so g -> A^T g scatters g_i / num_i
 * to those corners.  A vertex with num = 0 keeps its value. */
void orc_laplacian_smooth_T(vec3d *g, const vec3i *tris, int nvert, int ntri)
{
	vec3d *out = (vec3d *)calloc((size_t)nvert, sizeof(vec3d));
	int *num = (int *)calloc((size_t)nvert, sizeof(int));
	for (int i = 0; i < ntri; i++) {
		num[tris[i].x] += 2;
		num[tris[i].y] += 2;
		num[tris[i].z] += 2;
	}
	for (int i = 0; i < ntri; i++) {
		int c[3] = {tris[i].x, tris[i].y, tris[i].z};
		for (int k = 0; k < 3; k++) {
			int a = c[k == 0 ? 1 : 0], b = c[k == 2 ? 1 : 2];
			double w = 1.0 / num[c[k]];
			out[a].x += g[c[k]].x * w;
			out[a].y += g[c[k]].y * w;
			out[a].z += g[c[k]].z * w;
			out[b].x += g[c[k]].x * w;
			out[b].y += g[c[k]].y * w;
			out[b].z += g[c[k]].z * w;
		}
	}
	for (int i = 0; i < nvert; i++)
		g[i] = num[i] > 0 ? out[i] : g[i];
	free(out);
	free(num);
}

typedef struct {
	vec3d *p;
	vec3i *t;
	int np, nt;
	int nmc;         /* marching-cubes vertices before the weld */
	int *cls;        /* class of each marching-cubes vertex */
	long long *prov; /* provenance of each marching-cubes vertex */
	int *cedges;
	int ncentre;
} diff_state;

static void diff_free(diff_state *d)
{
	free(d->p);
	free(d->t);
	free(d->cls);
	free(d->prov);
	free(d->cedges);
}

static int diff_run(const float *img, const short dim[3], float isolevel, int originalMC, int lapIter, int exact,
                    diff_state *d)
{
	memset(d, 0, sizeof *d);
	const int lo[3] = {0, 0, 0}, hi[3] = {dim[0] - 1, dim[1] - 1, dim[2] - 1};
	if (lewiner_core(img, dim, lo, hi, originalMC, isolevel, exact, &d->p, &d->t, &d->nmc, &d->nt, &d->prov,
	                 &d->cedges, &d->ncentre) != EXIT_SUCCESS)
		return EXIT_FAILURE;
	d->cls = (int *)malloc((size_t)d->nmc * sizeof(int));
	d->np = weld_core(&d->p, d->t, d->nmc, d->nt, d->cls);
	if (d->np < 3) {
		diff_free(d);
		return EXIT_FAILURE;
	}
	d->nt = orc_remove_degenerate_triangles(d->p, &d->t, d->nt);
	for (int k = 0; k < lapIter; k++)
		orc_laplacian_smooth(d->p, d->t, d->np, d->nt);
	return EXIT_SUCCESS;
}

int orc_diff_forward(const float *img, const short dim[3], float isolevel, int originalMC, int lapIter, int exact,
                     vec3i **t, vec3d **p, int *nt, int *np)
{
	diff_state d;
	if (diff_run(img, dim, isolevel, originalMC, lapIter, exact, &d) != EXIT_SUCCESS)
		return EXIT_FAILURE;
	*t = d.t;
	*p = d.p;
	*nt = d.nt;
	*np = d.np;
	d.t = NULL;
	d.p = NULL;
	diff_free(&d);
	return EXIT_SUCCESS;
}

static float diff_clamp(float v) { return fabs(v) < FLT_EPSILON ? FLT_EPSILON : v; }

/* dL/dvolume (nx*ny*nz doubles, x fastest) from dL/dp of the forward's
 * final vertices (np x 3 doubles). */
int orc_diff_gradient(const float *img, const short dim[3], float isolevel, int originalMC, int lapIter,
                      const double *dLdp, double *dLdvol)
{
	diff_state d;
	if (diff_run(img, dim, isolevel, originalMC, lapIter, 1, &d) != EXIT_SUCCESS)
		return EXIT_FAILURE;
	const size_t nx = (size_t)dim[0], nxy = nx * (size_t)dim[1], nvox = nxy * (size_t)dim[2];
	vec3d *g = (vec3d *)malloc((size_t)d.np * sizeof(vec3d));
	memcpy(g, dLdp, (size_t)d.np * sizeof(vec3d));
	for (int k = 0; k < lapIter; k++)
		orc_laplacian_smooth_T(g, d.t, d.np, d.nt);
	/* weld: the class position is its last member's (src/meshify.c:99-100) */
	int *rep = (int *)malloc((size_t)d.np * sizeof(int));
	for (int c = 0; c < d.np; c++)
		rep[c] = -1;
	for (int m = 0; m < d.nmc; m++)
		rep[d.cls[m]] = m;
	vec3d *gm = (vec3d *)calloc((size_t)d.nmc, sizeof(vec3d));
	for (int m = 0; m < d.nmc; m++)
		if (rep[d.cls[m]] == m)
			gm[m] = g[d.cls[m]];
	/* centre vertex = mean of its cube's edge vertices */
	for (int m = 0; m < d.nmc; m++) {
		if (d.prov[m] >= 0)
			continue;
		const int *ce = &d.cedges[12 * (size_t)(-d.prov[m] - 1)];
		int cnt = 0;
		for (int e = 0; e < 12; e++)
			cnt += ce[e] >= 0;
		if (!cnt)
			continue;
		for (int e = 0; e < 12; e++) {
			if (ce[e] < 0)
				continue;
			gm[ce[e]].x += gm[m].x / cnt;
			gm[ce[e]].y += gm[m].y / cnt;
			gm[ce[e]].z += gm[m].z / cnt;
		}
	}
	/* edge vertex x + u, u = c0 / (c0 - c1), c = img - iso (clamped) */
	memset(dLdvol, 0, nvox * sizeof(double));
	for (int m = 0; m < d.nmc; m++) {
		if (d.prov[m] < 0)
			continue;
		const size_t ia = (size_t)(d.prov[m] >> 2); /* the sub-box is the volume */
		const int a = (int)(d.prov[m] & 3);
		const size_t ib = ia + (a == 0 ? 1 : (a == 1 ? nx : nxy));
		const float r0 = img[ia] - isolevel, r1 = img[ib] - isolevel;
		const float c0 = diff_clamp(r0), c1 = diff_clamp(r1);
		const double dd = (double)c0 - (double)c1;
		if (dd == 0.0 || (float)(c0 - c1) == 0.0f)
			continue;
		const double gu = a == 0 ? gm[m].x : (a == 1 ? gm[m].y : gm[m].z);
		if (c0 == r0) /* not clamped */
			dLdvol[ia] += gu * (-(double)c1 / (dd * dd));
		if (c1 == r1)
			dLdvol[ib] += gu * ((double)c0 / (dd * dd));
	}
	free(g);
	free(rep);
	free(gm);
	diff_free(&d);
	return EXIT_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Image loader conversion, src/nii2mesh.c:119,155-172: img32 = raw * slope   */
/* + inter in float for uint8 (2), int16 (4), uint16 (512), float32 (16).   */
/* ------------------------------------------------------------------------ */
int orc_decode(const void *raw, int datatype, int nvox, float slope, float inter, float *out)
{
	if (slope == 0.0f)
		slope = 1.0f;
	for (int i = 0; i < nvox; i++) {
		float v;
		switch (datatype) {
		case 2: v = ((const uint8_t *)raw)[i]; break;
		case 4: v = ((const int16_t *)raw)[i]; break;
		case 512: v = ((const uint16_t *)raw)[i]; break;
		case 16: v = ((const float *)raw)[i]; break;
		default: return EXIT_FAILURE;
		}
		float m = v * slope;
		out[i] = m + inter;
	}
	return EXIT_SUCCESS;
}
