// k_mc.hip -- marching cubes on gfx950: classify -> scan -> emit.
//
// Classic (src/oldcubes.c:465-522): configuration bit p = corner p below the
// isolevel; every triangle corner is its own vertex, interpolated in FP64
// exactly as VertexInterp (src/oldcubes.c:22-40).  Triangles come out in cube
// raster order, table order inside a cube, like the reference loop.
//
// Lewiner MC33 (src/MarchingCubes.c:116-1143): the volume is read through
// the reference's sub-box view (hi-lo+1 points per axis, img-iso, clamped
// away from 0 by FLT_EPSILON, wrap of the extra column/row, min-iso padding),
// edge vertices are numbered in point raster order (x, y, z edge of each
// point), centre vertices after them in cube order, triangles in cube order
// with the winding reversed (v3,v2,v1).
//
// Row segments: the Lewiner kernels work on segments of up to SEG points of
// one x-row; per-segment counts are scanned on the device, and inside a
// segment a workgroup-level scan gives each point/cube its offset.
#include <float.h>

#include "n2m_device.cuh"
#include "n2m_internal.h"

#define N2M_TABLE_QUAL static __constant__ const
#include "mc_tables.h"

namespace n2m {

// corner p -> (dx,dy,dz), src/oldcubes.c:419-426 / MarchingCubes.c:131
__constant__ int8_t c_cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int8_t c_cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__constant__ int8_t c_cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
// classic edge e interpolates from corner A[e] toward B[e]
__constant__ int8_t c_ea[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
__constant__ int8_t c_eb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

__device__ __forceinline__ int classic_ntri(uint64_t w)
{
	int n = 0;
	while (n < 5 && ((w >> (12 * n)) & 0xF) != 0xF)
		n++;
	return n;
}

// ---------------------------------------------------------------------------
// classic
// ---------------------------------------------------------------------------
struct ClsGeo {
	int nx, ny, nz;
	int lo0, lo1, lo2;
	int cx, cy, cz; // cube counts
	float iso;
	int ox, oy, oz; // origin of this volume in output coordinates
};

constexpr int MC_BS = 256;

// count the triangles of cube block `blk` (MC_BS cubes); returns the block
// total (valid in every thread)
__device__ __forceinline__ uint32_t classic_count_block(const float *__restrict__ img, const ClsGeo &g,
                                                        uint8_t *__restrict__ cfgs, size_t blk)
{
	__shared__ uint32_t scratch[MC_BS / N2M_WAVE + 1];
	size_t ncube = (size_t)g.cx * g.cy * g.cz;
	size_t c = blk * MC_BS + threadIdx.x;
	uint32_t nt = 0;
	if (c < ncube) {
		int x = g.lo0 + (int)(c % g.cx);
		size_t r = c / g.cx;
		int y = g.lo1 + (int)(r % g.cy), z = g.lo2 + (int)(r / g.cy);
		const size_t nxy = (size_t)g.nx * g.ny;
		const float *b = img + (size_t)x + (size_t)y * g.nx + (size_t)z * nxy;
		int cfg = 0;
#pragma unroll
		for (int p = 0; p < 8; p++) {
			float v = b[c_cx[p] + (size_t)c_cy[p] * g.nx + (size_t)c_cz[p] * nxy];
			if (v < g.iso)
				cfg |= 1 << p;
		}
		cfgs[c] = (uint8_t)cfg;
		nt = classic_ntri(n2m_classic_tri[cfg]);
	}
	return block_sum_u32<MC_BS>(nt, scratch);
}

// emit the triangle corners of cube block `blk`, first triangle at `off`.
// The block's triangles, not its cubes, are spread over the threads: the
// scan of the per-cube counts lists each triangle's cube and rank in LDS,
// then thread j builds triangles j, j + MC_BS, ... (the surface cubes are a
// few of the block's, so one thread per cube left most lanes idle while a
// few built up to 5 triangles each).  Triangle q is written at off + q,
// its corners in the cube's table order, as before.
__device__ __forceinline__ void classic_emit_block(const float *__restrict__ img, const ClsGeo &g,
                                                   const uint8_t *__restrict__ cfgs, size_t blk, uint32_t off0,
                                                   double *__restrict__ pts)
{
	__shared__ uint32_t scratch[MC_BS / N2M_WAVE + 1];
	__shared__ uint8_t s_cube[5 * MC_BS]; // the cube (block-local) of each triangle
	__shared__ uint8_t s_rank[5 * MC_BS]; // its rank within the cube
	const size_t ncube = (size_t)g.cx * g.cy * g.cz;
	const size_t c0 = blk * MC_BS;
	uint32_t tot;
	{
		const size_t c = c0 + threadIdx.x;
		const int cfg = c < ncube ? cfgs[c] : 0;
		const uint32_t nt = classic_ntri(n2m_classic_tri[cfg]);
		const uint32_t off = block_excl_scan<MC_BS>(nt, scratch, &tot);
		for (uint32_t k = 0; k < nt; k++) {
			s_cube[off + k] = (uint8_t)threadIdx.x;
			s_rank[off + k] = (uint8_t)k;
		}
		__syncthreads();
	}
	const size_t nxy = (size_t)g.nx * g.ny;
	const double iso = g.iso;
	for (uint32_t q = threadIdx.x; q < tot; q += MC_BS) {
		const size_t c = c0 + s_cube[q];
		const int kk = s_rank[q];
		const uint64_t w = n2m_classic_tri[cfgs[c]];
		const int x = g.lo0 + (int)(c % g.cx);
		const size_t r = c / g.cx;
		const int y = g.lo1 + (int)(r % g.cy), z = g.lo2 + (int)(r / g.cy);
		const float *b = img + (size_t)x + (size_t)y * g.nx + (size_t)z * nxy;
		double *o = pts + ((size_t)off0 + q) * 9;
#pragma unroll
		for (int j = 0; j < 3; j++) {
			const int e = (int)((w >> (4 * (3 * kk + j))) & 0xF);
			const int a = c_ea[e], bb = c_eb[e];
			const double va = b[c_cx[a] + (size_t)c_cy[a] * g.nx + (size_t)c_cz[a] * nxy];
			const double vb = b[c_cx[bb] + (size_t)c_cy[bb] * g.nx + (size_t)c_cz[bb] * nxy];
			const double mu = (iso - va) / (vb - va);
			const double ax = x + g.ox + c_cx[a], ay = y + g.oy + c_cy[a], az = z + g.oz + c_cz[a];
			const double bx = x + g.ox + c_cx[bb], by = y + g.oy + c_cy[bb], bz = z + g.oz + c_cz[bb];
			o[3 * j + 0] = ax + mu * (bx - ax);
			o[3 * j + 1] = ay + mu * (by - ay);
			o[3 * j + 2] = az + mu * (bz - az);
		}
	}
}

__global__ void __launch_bounds__(MC_BS)
k_classic_count(const float *__restrict__ img, ClsGeo g, uint8_t *__restrict__ cfgs,
                uint32_t *__restrict__ blockSums)
{
	uint32_t s = classic_count_block(img, g, cfgs, blockIdx.x);
	if (threadIdx.x == 0)
		blockSums[blockIdx.x] = s;
}

__global__ void __launch_bounds__(MC_BS)
k_classic_emit(const float *__restrict__ img, ClsGeo g, const uint8_t *__restrict__ cfgs,
               const uint32_t *__restrict__ blockOffs, double *__restrict__ pts)
{
	classic_emit_block(img, g, cfgs, blockIdx.x, blockOffs[blockIdx.x], pts);
}

size_t classic_blocks(int cx, int cy, int cz) { return ((size_t)cx * cy * cz + MC_BS - 1) / MC_BS; }

void launch_classic_count(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3],
                          float iso, uint8_t *cfgs, uint32_t *blockSums, hipStream_t s)
{
	ClsGeo g{nx, ny, nz, lo[0], lo[1], lo[2], hi[0] - 1 - lo[0], hi[1] - 1 - lo[1], hi[2] - 1 - lo[2], iso, 0, 0, 0};
	size_t nb = classic_blocks(g.cx, g.cy, g.cz);
	hipLaunchKernelGGL(k_classic_count, dim3((unsigned)nb), dim3(MC_BS), 0, s, img, g, cfgs, blockSums);
}

void launch_classic_emit(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3],
                         float iso, const uint8_t *cfgs, const uint32_t *blockOffs, double *pts,
                         const int origin[3], hipStream_t s)
{
	ClsGeo g{nx, ny, nz, lo[0], lo[1], lo[2], hi[0] - 1 - lo[0], hi[1] - 1 - lo[1], hi[2] - 1 - lo[2], iso,
	         origin[0], origin[1], origin[2]};
	size_t nb = classic_blocks(g.cx, g.cy, g.cz);
	hipLaunchKernelGGL(k_classic_emit, dim3((unsigned)nb), dim3(MC_BS), 0, s, img, g, cfgs, blockOffs, pts);
}

// ---------------------------------------------------------------------------
// Lewiner MC33
// ---------------------------------------------------------------------------
struct LwGeo {
	int nx;
	long long nxy, nxyz;
	int lo0, lo1, lo2;
	int sx, sy, sz; // sub-box points
	float iso, mnv; // mnv = min(img) - iso
	int originalMC;
	int nseg;       // segments per row
	int out0, out1, out2; // lo + origin: added (as float) to sub-box coordinates
};

constexpr int LW_SEG = 512, LW_BS = 256;

__device__ __forceinline__ float lw_clamp(float v) { return fabsf(v) < FLT_EPSILON ? FLT_EPSILON : v; }

// value of the sub-box copy at point (x,y,z), src/MarchingCubes.c:1097-1115
__device__ __forceinline__ float lw_data(const float *__restrict__ img, const LwGeo &g, int x, int y, int z)
{
	long long j = (long long)g.lo0 + x + (long long)(y + g.lo1) * g.nx + (long long)(z + g.lo2) * g.nxy;
	float v = j < g.nxyz ? img[j] - g.iso : g.mnv;
	return lw_clamp(v);
}

// test_face(), src/MarchingCubes.c:276-295
__constant__ int8_t c_face[6][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                                   {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};
__device__ __forceinline__ bool lw_face(const float *c, int face)
{
	int f = face < 0 ? -face : face;
	float A = 0.f, B = 0.f, C = 0.f, D = 0.f;
	if (f >= 1 && f <= 6) {
		A = c[c_face[f - 1][0]];
		B = c[c_face[f - 1][1]];
		C = c[c_face[f - 1][2]];
		D = c[c_face[f - 1][3]];
	}
	float det = A * C - B * D;
	if (fabsf(det) < FLT_EPSILON)
		return face >= 0;
	float lhs = (float)face * A;
	lhs = lhs * det;
	return lhs >= 0.f;
}

// test_interior(), src/MarchingCubes.c:301-453; the reference edge row gives
// (p, q) for t = c[p]/(c[p]-c[q]) and the (from, to) corners of B, C, D.
__constant__ int8_t c_iedge[12][8] = {
    {0, 1, 3, 2, 7, 6, 4, 5}, {1, 2, 0, 3, 4, 7, 5, 6}, {2, 3, 1, 0, 5, 4, 6, 7},
    {3, 0, 2, 1, 6, 5, 7, 4}, {4, 5, 7, 6, 3, 2, 0, 1}, {5, 6, 4, 7, 0, 3, 1, 2},
    {6, 7, 5, 4, 1, 0, 2, 3}, {7, 4, 6, 5, 2, 1, 3, 0}, {0, 4, 3, 7, 2, 6, 1, 5},
    {1, 5, 0, 4, 3, 7, 2, 6}, {2, 6, 1, 5, 0, 4, 3, 7}, {3, 7, 2, 6, 1, 5, 0, 4}};

__device__ bool lw_interior(const float *c, int cs, int cf, int sub, int s)
{
	float t, At = 0.f, Bt = 0.f, Ct = 0.f, Dt = 0.f;
	if (cs == 4 || cs == 10) {
		float a = (c[4] - c[0]) * (c[6] - c[2]) - (c[7] - c[3]) * (c[5] - c[1]);
		float b = c[2] * (c[4] - c[0]) + c[0] * (c[6] - c[2]) - c[1] * (c[7] - c[3]) - c[3] * (c[5] - c[1]);
		t = -b / (2.f * a);
		if (t < 0.f || t > 1.f)
			return s > 0;
		At = c[0] + (c[4] - c[0]) * t;
		Bt = c[3] + (c[7] - c[3]) * t;
		Ct = c[2] + (c[6] - c[2]) * t;
		Dt = c[1] + (c[5] - c[1]) * t;
	} else {
		int edge;
		if (cs == 6)
			edge = n2m_lw_test6[cf * 3 + 2];
		else if (cs == 7)
			edge = n2m_lw_test7[cf * 5 + 4];
		else if (cs == 12)
			edge = n2m_lw_test12[cf * 4 + 3];
		else
			edge = n2m_lw_edges[n2m_lw_row_off[N2M_LW_T13_5_1 + cf * 4 + sub]];
		if (edge >= 0 && edge < 12) {
			const int8_t *e = c_iedge[edge];
			t = c[e[0]] / (c[e[0]] - c[e[1]]);
			At = 0.f;
			Bt = c[e[2]] + (c[e[3]] - c[e[2]]) * t;
			Ct = c[e[4]] + (c[e[5]] - c[e[4]]) * t;
			Dt = c[e[6]] + (c[e[7]] - c[e[6]]) * t;
		}
	}
	int test = (At >= 0.f) + 2 * (Bt >= 0.f) + 4 * (Ct >= 0.f) + 8 * (Dt >= 0.f);
	switch (test) {
	case 5:
		return (At * Ct - Bt * Dt < FLT_EPSILON) ? s > 0 : s < 0;
	case 10:
		return (At * Ct - Bt * Dt >= FLT_EPSILON) ? s > 0 : s < 0;
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

// process_cube() decision tree, src/MarchingCubes.c:458-795 -> tiling row or -1.
// cs_tab / cf_tab: the case and configuration tables (or a copy in LDS).
__device__ int lw_select(const float c[8], int lut, int originalMC, const uint8_t *cs_tab = n2m_lw_case,
                         const uint8_t *cf_tab = n2m_lw_config)
{
	if (originalMC)
		return N2M_LW_CLASSIC + lut;
	const int cs = cs_tab[lut], cf = cf_tab[lut];
	int sub = 0;
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
		return lw_interior(c, 6, cf, 0, n2m_lw_test6[cf * 3 + 1]) ? N2M_LW_T6_1_1 + cf : N2M_LW_T6_1_2 + cf;
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
			return lw_interior(c, 7, cf, sub, n2m_lw_test7[cf * 5 + 3]) ? N2M_LW_T7_4_2 + cf : N2M_LW_T7_4_1 + cf;
		}
	case 8:
		return N2M_LW_T8 + cf;
	case 9:
		return N2M_LW_T9 + cf;
	case 10: {
		bool f0 = lw_face(c, n2m_lw_test10[cf * 3 + 0]), f1 = lw_face(c, n2m_lw_test10[cf * 3 + 1]);
		if (f0)
			return (f1 ? N2M_LW_T10_1_1_ : N2M_LW_T10_2) + cf;
		if (f1)
			return N2M_LW_T10_2_ + cf;
		return (lw_interior(c, 10, cf, 0, n2m_lw_test10[cf * 3 + 2]) ? N2M_LW_T10_1_1 : N2M_LW_T10_1_2) + cf;
	}
	case 11:
		return N2M_LW_T11 + cf;
	case 12: {
		bool f0 = lw_face(c, n2m_lw_test12[cf * 4 + 0]), f1 = lw_face(c, n2m_lw_test12[cf * 4 + 1]);
		if (f0)
			return (f1 ? N2M_LW_T12_1_1_ : N2M_LW_T12_2) + cf;
		if (f1)
			return N2M_LW_T12_2_ + cf;
		return (lw_interior(c, 12, cf, 0, n2m_lw_test12[cf * 4 + 2]) ? N2M_LW_T12_1_1 : N2M_LW_T12_1_2) + cf;
	}
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
		return -1;
	}
	case 14:
		return N2M_LW_T14 + cf;
	}
	return -1;
}

// Pass A for one row segment (seg, y, z): per point flags, per cube tiling
// row (a point's bit a is set when its a-edge carries a vertex,
// src/MarchingCubes.c:238-269); the segment's counts of edge vertices,
// triangles and centre vertices go to cnt[0], cnt[cstride], cnt[2*cstride].
// The four point rows the segment reads -- r = dy + 2*dz for (y+dy, z+dz) --
// are staged in LDS (sub-box values, clamped as lw_data returns them); row r
// sits in slot (r & 1) + 2 * (((r >> 1) + par) & 1), so that a z walk can
// keep the shared plane where it is.
// The small tables a cube's classification reads: the case and
// configuration of each corner pattern, the triangles and centre-vertex flag
// of each tiling row (global, or copies in LDS).
constexpr int LW_FMW = 4; // mask words per 64-point chunk: x, y, z edges, cubes with triangles

struct LwTabs {
	const uint8_t *cs, *cf, *ntri, *centre;
};
constexpr LwTabs LW_GLOBAL_TABS = {n2m_lw_case, n2m_lw_config, n2m_lw_row_ntri, n2m_lw_row_centre};

// The point flags go either to pflags (one byte per point) or, when fmask is
// given, to per-axis bit masks: fmask[(row * nchunk + chunk) * 4 + a] holds
// bit b for the point x = 64 * chunk + b of sub-box row (y + sy * z), and
// word 3 the cubes with a tiling row; crow is then written only for those.
__device__ __forceinline__ void lw_classify_core(const float (*s_row)[LW_SEG + 1], int par, const LwGeo &g, int seg,
                                                 int y, int z, uint8_t *__restrict__ pflags,
                                                 uint16_t *__restrict__ crow, uint32_t *__restrict__ cnt,
                                                 size_t cstride, uint32_t *scratch, const LwTabs &tab = LW_GLOBAL_TABS,
                                                 unsigned long long *__restrict__ fmask = nullptr)
{
	const int xs = seg * LW_SEG, xe = min(xs + LW_SEG, g.sx);
	const bool cuberow = y < g.sy - 1 && z < g.sz - 1;
	const int sl0 = 2 * (par & 1), sl1 = 2 * ((1 + par) & 1); // slot of dz = 0, 1
	const float *r00 = s_row[sl0], *r10 = s_row[sl0 + 1], *r01 = s_row[sl1];
	uint32_t nv = 0, nt = 0, nc = 0;
	const size_t rowbase = (size_t)g.sx * ((size_t)y + (size_t)g.sy * z);
	for (int x = xs + threadIdx.x; x < xe; x += LW_BS) {
		const int i = x - xs;
		const float c0 = r00[i];
		float cn[3];
		cn[0] = x < g.sx - 1 ? r00[i + 1] : c0;
		cn[1] = y < g.sy - 1 ? r10[i] : c0;
		cn[2] = z < g.sz - 1 ? r01[i] : c0;
		uint32_t f = 0;
#pragma unroll
		for (int a = 0; a < 3; a++) {
			bool cross = c0 < 0.f ? cn[a] > 0.f : cn[a] < 0.f;
			f |= (uint32_t)cross << a;
		}
		if (!fmask)
			pflags[rowbase + x] = (uint8_t)f;
		nv += __popc(f);
		uint16_t cr = 0;
		if (cuberow && x < g.sx - 1) {
			float c[8];
			int lut = 0;
#pragma unroll
			for (int p = 0; p < 8; p++) {
				c[p] = s_row[c_cy[p] + (c_cz[p] ? sl1 : sl0)][i + c_cx[p]];
				if (c[p] > 0.f)
					lut += 1 << p;
			}
			// all corners on one side: no triangle (case 0, or a classic row
			// without triangles for -o 1), the table walk is skipped
			const int row = (lut == 0 || lut == 255) ? -1 : lw_select(c, lut, g.originalMC, tab.cs, tab.cf);
			if (row >= 0) {
				cr = (uint16_t)(row + 1);
				nt += tab.ntri[row];
				nc += (!g.originalMC && tab.centre[row]) ? 1u : 0u;
			}
		}
		if (fmask) {
			// one wave holds one 64-point chunk (xs is a multiple of 64 and
			// a lane past the row end is inactive); its first lane writes
			const unsigned long long mx = __ballot(f & 1u), my = __ballot(f & 2u), mz = __ballot(f & 4u),
			                         mc = __ballot(cr != 0);
			if (lane_id() == 0) {
				const int nchunk = (g.sx + N2M_WAVE - 1) / N2M_WAVE;
				unsigned long long *o =
				    fmask + ((size_t)(y + (size_t)g.sy * z) * nchunk + (size_t)(x / N2M_WAVE)) * LW_FMW;
				o[0] = mx;
				o[1] = my;
				o[2] = mz;
				o[3] = mc;
			}
			if (cr)
				crow[rowbase + x] = cr;
		} else {
			crow[rowbase + x] = cr;
		}
	}
	// the three segment counts: wave sums, then one barrier (scratch holds
	// 3 words per wave)
	constexpr int NW = LW_BS / N2M_WAVE;
	nv = wave_sum(nv);
	nt = wave_sum(nt);
	nc = wave_sum(nc);
	if (lane_id() == 0) {
		scratch[wave_id()] = nv;
		scratch[NW + wave_id()] = nt;
		scratch[2 * NW + wave_id()] = nc;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t a = 0, b = 0, c = 0;
		for (int w = 0; w < NW; w++) {
			a += scratch[w];
			b += scratch[NW + w];
			c += scratch[2 * NW + w];
		}
		cnt[0] = a;
		cnt[cstride] = b;
		cnt[2 * cstride] = c;
	}
}

// plane zz of rows y and y+1 of segment seg into slots slot, slot + 1
__device__ __forceinline__ void lw_stage_plane(float (*s_row)[LW_SEG + 1], int slot, const float *__restrict__ img,
                                               const LwGeo &g, int seg, int y, int zz)
{
	const int xs = seg * LW_SEG, xe = min(xs + LW_SEG, g.sx);
	const int xl = min(xe + 1, g.sx) - xs; // staged points per row: [xs, xs + xl)
	if (zz >= g.sz)
		return;
#pragma unroll
	for (int dy = 0; dy < 2; dy++) {
		const int yy = y + dy;
		if (yy >= g.sy)
			continue;
		for (int i = threadIdx.x; i < xl; i += LW_BS)
			s_row[slot + dy][i] = lw_data(img, g, xs + i, yy, zz);
	}
}

// one row segment, its four rows staged here (the atlas path)
__device__ __forceinline__ void lw_classify_row(const float *__restrict__ img, const LwGeo &g, int seg, int y, int z,
                                                uint8_t *__restrict__ pflags, uint16_t *__restrict__ crow,
                                                uint32_t *__restrict__ cnt, size_t cstride)
{
	__shared__ uint32_t scratch[3 * (LW_BS / N2M_WAVE)];
	__shared__ float s_row[4][LW_SEG + 1];
	lw_stage_plane(s_row, 0, img, g, seg, y, z);
	lw_stage_plane(s_row, 2, img, g, seg, y, z + 1);
	__syncthreads();
	lw_classify_core(s_row, 0, g, seg, y, z, pflags, crow, cnt, cstride, scratch);
}

// The whole-volume pass: a workgroup walks LW_ZC planes of one (segment, y)
// column, loading one new plane per step (each point row is read about
// once, instead of once per neighbouring row segment).
constexpr int LW_ZC = 8;
// 7 waves per SIMD: the register cap (72) leaves no spills; the default
// allocation (95 registers, 5 waves) was 30 % slower.  Loading the plane two
// steps ahead into registers (80 registers, 6 waves) was 13 % slower.
__global__ void __launch_bounds__(LW_BS, 7)
k_lw_classify(const float *__restrict__ img, LwGeo g, uint8_t *__restrict__ pflags,
              uint16_t *__restrict__ crow, uint32_t *__restrict__ segcnt /* 3 arrays of nseg*sy*sz */,
              unsigned long long *__restrict__ fmask)
{
	__shared__ uint32_t scratch[3 * (LW_BS / N2M_WAVE)];
	__shared__ float s_row[4][LW_SEG + 1];
	// the tables a surface cube reads, in LDS (global reads of them, in the
	// loop over points, were a dependent memory round trip per cube)
	__shared__ uint8_t s_cs[256], s_cf[256], s_ntri[N2M_LW_NROWS], s_centre[N2M_LW_NROWS];
	s_cs[threadIdx.x] = n2m_lw_case[threadIdx.x];
	s_cf[threadIdx.x] = n2m_lw_config[threadIdx.x];
	for (int i = threadIdx.x; i < N2M_LW_NROWS; i += LW_BS) {
		s_ntri[i] = n2m_lw_row_ntri[i];
		s_centre[i] = n2m_lw_row_centre[i];
	}
	const LwTabs tab = {s_cs, s_cf, s_ntri, s_centre};
	const int seg = blockIdx.x, y = blockIdx.y, z0 = blockIdx.z * LW_ZC, z1 = min(z0 + LW_ZC, g.sz);
	const size_t ns = (size_t)g.nseg * g.sy * g.sz;
	lw_stage_plane(s_row, 0, img, g, seg, y, z0);
	for (int z = z0; z < z1; z++) {
		const int par = (z - z0) & 1;
		lw_stage_plane(s_row, 2 * ((1 + par) & 1), img, g, seg, y, z + 1);
		__syncthreads();
		const size_t si = (size_t)seg + (size_t)g.nseg * ((size_t)y + (size_t)g.sy * z);
		lw_classify_core(s_row, par, g, seg, y, z, pflags, crow, segcnt + si, ns, scratch, tab, fmask);
		__syncthreads(); // the slots of plane z are overwritten next
	}
}

// The whole-volume pass from sign bits.  P / N hold the bits
// lw_clamp(v - iso) > 0 and < 0 of every voxel in linear order (the apply
// stage writes them; rows of whole 32-voxel words).  A wave takes a chunk of
// 64 points of a row as 64-bit words: the three crossing masks
// (c0 < 0 ? c1 > 0 : c1 < 0, NaN-exact with the two bits) and the mask of
// cubes with corners on both sides come from a few word operations; only
// the lanes of those cubes build a case, and only the cases whose tiling
// depends on face or interior tests read the eight corner values.
struct SignBits {
	const uint32_t *P, *N;
	size_t nwords;         // words of each array (nx * ny * nz / 32)
	uint32_t padP, padN;   // a padding word (the signs of mnv, lw_clamp applied)
	int nx;
	long long nxy;
	// 65 bits of `a` from the linear index of sub-box point (x, y, z), as
	// lw_data indexes (bits past the volume take the padding)
	__device__ __forceinline__ unsigned long long win(const uint32_t *__restrict__ a, uint32_t pad, const LwGeo &g,
	                                                  int x, int y, int z, uint32_t &b64) const
	{
		const long long j = (long long)g.lo0 + x + (long long)(y + g.lo1) * nx + (long long)(z + g.lo2) * nxy;
		const size_t w = (size_t)(j >> 5);
		const int sh = (int)(j & 31);
		const uint32_t w0 = w < nwords ? a[w] : pad, w1 = w + 1 < nwords ? a[w + 1] : pad,
		               w2 = w + 2 < nwords ? a[w + 2] : pad;
		const unsigned long long lo = (unsigned long long)w0 | ((unsigned long long)w1 << 32);
		b64 = (w2 >> sh) & 1u;
		return sh ? (lo >> sh) | ((unsigned long long)w2 << (64 - sh)) : lo;
	}
};

// the tilings whose choice reads corner values (face / interior tests)
__device__ __forceinline__ bool lw_case_tests(int cs)
{
	return cs == 3 || cs == 4 || cs == 6 || cs == 7 || cs == 10 || cs == 12 || cs == 13;
}

// Each lane takes one 64-point chunk: the 8 lanes of an aligned group are
// the (up to) 8 chunks of one row segment (LW_SEG = 8 x 64), so a wave
// covers 8 row segments and the segment counts are a reduction over the
// group.  The surface cubes of a chunk (a few) are then visited one by one.
static_assert(LW_SEG == 8 * N2M_WAVE, "8 chunks per row segment");

__global__ void __launch_bounds__(LW_BS)
k_lw_classify_bits(const float *__restrict__ img, LwGeo g, SignBits sb, uint16_t *__restrict__ crow,
                   uint32_t *__restrict__ segcnt /* 3 arrays of nseg*sy*sz */, unsigned long long *__restrict__ fmask)
{
	__shared__ uint8_t s_cs[256], s_cf[256], s_ntri[N2M_LW_NROWS], s_centre[N2M_LW_NROWS];
	s_cs[threadIdx.x] = n2m_lw_case[threadIdx.x];
	s_cf[threadIdx.x] = n2m_lw_config[threadIdx.x];
	for (int i = threadIdx.x; i < N2M_LW_NROWS; i += LW_BS) {
		s_ntri[i] = n2m_lw_row_ntri[i];
		s_centre[i] = n2m_lw_row_centre[i];
	}
	__syncthreads();
	const size_t ns = (size_t)g.nseg * g.sy * g.sz;
	const size_t sr = ((size_t)blockIdx.x * LW_BS + threadIdx.x) / 8; // row segment
	const int k = threadIdx.x & 7;                                     // chunk within it
	if (sr >= ns)
		return; // whole 8-lane groups
	const int seg = (int)(sr % g.nseg);
	const size_t yz = sr / g.nseg;
	const int y = (int)(yz % g.sy), z = (int)(yz / g.sy);
	const int x0 = seg * LW_SEG + k * N2M_WAVE;
	const bool hy = y < g.sy - 1, hz = z < g.sz - 1, cuberow = hy && hz;
	uint32_t nv = 0, nt = 0, nc = 0;
	if (x0 < g.sx) {
		// bit i <-> point x0 + i of rows A (y, z), B (y + 1, z), C (y, z + 1),
		// D (y + 1, z + 1); the "x" words hold the x + 1 neighbours
		uint32_t t0, t1;
		const unsigned long long pA = sb.win(sb.P, sb.padP, g, x0, y, z, t0),
		                         nA = sb.win(sb.N, sb.padN, g, x0, y, z, t1);
		const unsigned long long pAx = (pA >> 1) | ((unsigned long long)t0 << 63),
		                         nAx = (nA >> 1) | ((unsigned long long)t1 << 63);
		unsigned long long pB = pA, nB = nA, pC = pA, nC = nA, pD = pA, pBx = pAx, pCx = pAx, pDx = pAx;
		if (hy) {
			pB = sb.win(sb.P, sb.padP, g, x0, y + 1, z, t0);
			pBx = (pB >> 1) | ((unsigned long long)t0 << 63);
			nB = sb.win(sb.N, sb.padN, g, x0, y + 1, z, t1);
		}
		if (hz) {
			pC = sb.win(sb.P, sb.padP, g, x0, y, z + 1, t0);
			pCx = (pC >> 1) | ((unsigned long long)t0 << 63);
			nC = sb.win(sb.N, sb.padN, g, x0, y, z + 1, t1);
		}
		if (cuberow) {
			pD = sb.win(sb.P, sb.padP, g, x0, y + 1, z + 1, t0);
			pDx = (pD >> 1) | ((unsigned long long)t0 << 63);
		}
		const int npt = min(N2M_WAVE, g.sx - x0), nxn = min(N2M_WAVE, g.sx - 1 - x0); // points; with an x + 1
		const unsigned long long valid = npt >= N2M_WAVE ? ~0ull : ((1ull << npt) - 1ull);
		const unsigned long long validx = nxn <= 0 ? 0ull : (nxn >= N2M_WAVE ? ~0ull : ((1ull << nxn) - 1ull));
		const unsigned long long fx = ((nA & pAx) | (~nA & nAx)) & validx;
		const unsigned long long fy = hy ? ((nA & pB) | (~nA & nB)) & valid : 0ull;
		const unsigned long long fz = hz ? ((nA & pC) | (~nA & nC)) & valid : 0ull;
		unsigned long long surf = 0, mc = 0;
		if (cuberow) {
			const unsigned long long all1 = pA & pAx & pB & pBx & pC & pCx & pD & pDx;
			const unsigned long long any1 = pA | pAx | pB | pBx | pC | pCx | pD | pDx;
			surf = any1 & ~all1 & validx;
		}
		const size_t rowbase = (size_t)g.sx * ((size_t)y + (size_t)g.sy * z);
		for (unsigned long long m = surf; m; m &= m - 1) {
			const int i = __ffsll((long long)m) - 1, x = x0 + i;
			const int lut = (int)((pA >> i) & 1ull) | ((int)((pAx >> i) & 1ull) << 1) |
			                ((int)((pBx >> i) & 1ull) << 2) | ((int)((pB >> i) & 1ull) << 3) |
			                ((int)((pC >> i) & 1ull) << 4) | ((int)((pCx >> i) & 1ull) << 5) |
			                ((int)((pDx >> i) & 1ull) << 6) | ((int)((pD >> i) & 1ull) << 7);
			float c[8];
			if (!g.originalMC && lw_case_tests(s_cs[lut])) {
#pragma unroll
				for (int p = 0; p < 8; p++)
					c[p] = lw_data(img, g, x + c_cx[p], y + c_cy[p], z + c_cz[p]);
			}
			const int row = lw_select(c, lut, g.originalMC, s_cs, s_cf);
			if (row >= 0) {
				nt += s_ntri[row];
				nc += (!g.originalMC && s_centre[row]) ? 1u : 0u;
				crow[rowbase + x] = (uint16_t)(row + 1);
				mc |= 1ull << i;
			}
		}
		const int nchunk = (g.sx + N2M_WAVE - 1) / N2M_WAVE;
		unsigned long long *o = fmask + ((size_t)(y + (size_t)g.sy * z) * nchunk + (size_t)(x0 / N2M_WAVE)) * LW_FMW;
		o[0] = fx;
		o[1] = fy;
		o[2] = fz;
		o[3] = mc;
		nv = (uint32_t)(__popcll(fx) + __popcll(fy) + __popcll(fz));
	}
#pragma unroll
	for (int d = 1; d < 8; d <<= 1) {
		nv += __shfl_xor(nv, d);
		nt += __shfl_xor(nt, d);
		nc += __shfl_xor(nc, d);
	}
	if (k == 0) {
		segcnt[sr] = nv;
		segcnt[ns + sr] = nt;
		segcnt[2 * ns + sr] = nc;
	}
}

// edge vertex position in sub-box coordinates, src/MarchingCubes.c:935-1026
__device__ __forceinline__ void lw_edge_pos(const float *__restrict__ img, const LwGeo &g, int x, int y, int z,
                                            int a, float *p)
{
	float c0 = lw_data(img, g, x, y, z);
	float c1 = lw_data(img, g, x + (a == 0), y + (a == 1), z + (a == 2));
	float denom = c0 - c1;
	float u = denom != 0.0f ? c0 / denom : 0.5f;
	p[0] = (float)x;
	p[1] = (float)y;
	p[2] = (float)z;
	p[a] = p[a] + u;
}

// the same position in double (differentiable path, exact mode): x + u with
// u = c0 / (c0 - c1) evaluated in double from the float corner values
__device__ __forceinline__ void lw_edge_pos_exact(const float *__restrict__ img, const LwGeo &g, int x, int y, int z,
                                                  int a, double *p)
{
	float c0 = lw_data(img, g, x, y, z);
	float c1 = lw_data(img, g, x + (a == 0), y + (a == 1), z + (a == 2));
	double dd = (double)c0 - (double)c1;
	double u = dd != 0.0 ? (double)c0 / dd : 0.5;
	p[0] = (double)x;
	p[1] = (double)y;
	p[2] = (double)z;
	p[a] = p[a] + u;
}

// Emit modes: positions as the reference computes them, or (differentiable
// path) with each vertex's provenance recorded -- an edge vertex stores
// point*4 + axis, a centre vertex 2^63 + its centre index with its cube's
// 12 edge-vertex ids in crec -- and optionally FP64 positions.
enum { LW_PLAIN = 0, LW_RECORD = 1, LW_RECORD_EXACT = 2 };

// cube edge code -> (point dx, dy, dz, axis), src/MarchingCubes.c:811-824
__constant__ int8_t c_code[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1},
                                    {0, 0, 1, 0}, {1, 0, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1},
                                    {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};

// Pass B for one row segment: write the edge vertices of the segment's
// points, and the triangles and centre vertices of its cubes.  bases holds
// three arrays (stride bstride) of exclusive segment offsets for the volume
// (segment index seg + nseg*(y + sy*z)); edge_add / centre_add are added to
// edge-vertex / centre-vertex numbers.
template <int MODE = LW_PLAIN>
__device__ __forceinline__ void lw_emit_row(const float *__restrict__ img, const LwGeo &g, int seg, int y, int z,
                                            const uint8_t *__restrict__ pflags, const uint16_t *__restrict__ crow,
                                            const uint32_t *__restrict__ bases, size_t bstride, uint32_t edge_add,
                                            uint32_t centre_add, double *__restrict__ pts, int *__restrict__ tris,
                                            const uint32_t *__restrict__ cnt, size_t cstride,
                                            unsigned long long *__restrict__ vrec = nullptr,
                                            uint32_t *__restrict__ crec = nullptr)
{
	static_assert(LW_BS == 4 * N2M_WAVE, "one wave per point row");
	__shared__ uint32_t s_pre[4][LW_SEG + 1]; // popcount prefix for the 4 point rows
	__shared__ uint8_t s_flag[4][LW_SEG + 1]; // the 4 rows' point flags, [0, len]
	__shared__ uint16_t s_crow[LW_SEG];
	__shared__ uint32_t s_base[4];
	__shared__ uint32_t scratch[LW_BS / N2M_WAVE + 1];
	// a segment without edge vertices, triangles or centre vertices of its
	// own writes nothing (its counts from pass A, the same for every lane)
	if (cnt[0] == 0 && cnt[cstride] == 0 && cnt[2 * cstride] == 0)
		return;
	const int xs = seg * LW_SEG, xe = min(xs + LW_SEG, g.sx);
	const int len = xe - xs; // points owned; prefix covers [xs, xe]
	const bool cuberow = y < g.sy - 1 && z < g.sz - 1;
	const int nrows = cuberow ? 4 : 1;
	const int ncube = min(xe, g.sx - 1) - xs;
	// the segment's tiling rows, staged while the prefixes are computed
	if (cuberow) {
		const size_t rb0 = (size_t)g.sx * ((size_t)y + (size_t)g.sy * z);
		for (int i = threadIdx.x; i < ncube; i += LW_BS)
			s_crow[i] = crow[rb0 + xs + i];
	}
	// row r = wave r: (y + (r&1), z + (r>>1)); exclusive popcount prefix.
	// All the row's flags are loaded first; s_pre is written only for
	// 64-point chunks with a flag (it is read only at flagged points) and at
	// len.
	{
		const int r = wave_id(), lane = lane_id();
		if (r < nrows) {
			const int yy = y + (r & 1), zz = z + (r >> 1);
			if (lane == 0)
				s_base[r] = edge_add + bases[(size_t)seg + (size_t)g.nseg * ((size_t)yy + (size_t)g.sy * zz)];
			const size_t rb = (size_t)g.sx * ((size_t)yy + (size_t)g.sy * zz);
			constexpr int NIT = LW_SEG / N2M_WAVE + 1; // chunks of [0, len]
			uint32_t f[NIT];
#pragma unroll
			for (int k = 0; k < NIT; k++) {
				const int i = k * N2M_WAVE + lane;
				// point len (the next segment's first point) closes the
				// segment's last cube; it exists unless the row ends at len
				const uint32_t fl = i <= len && xs + i < g.sx ? pflags[rb + xs + i] : 0u;
				if (i <= len)
					s_flag[r][i] = (uint8_t)fl;
				f[k] = i < len ? __popc(fl) : 0u;
			}
			uint32_t carry = 0;
#pragma unroll
			for (int k = 0; k < NIT; k++) {
				if (k * N2M_WAVE > len)
					break;
				if (!__ballot(f[k] != 0u))
					continue; // no flag in this chunk: carry unchanged
				const int i = k * N2M_WAVE + lane;
				const uint32_t incl = wave_incl_scan(f[k]);
				if (i <= len)
					s_pre[r][i] = carry + incl - f[k];
				carry += __shfl(incl, N2M_WAVE - 1, N2M_WAVE);
			}
			// the row total, at point len (the next segment's first point,
			// whose edges close this segment's last cube)
			if (lane == 0)
				s_pre[r][len] = carry;
		}
	}
	__syncthreads();
	// edge vertices of the owned points
	{
		const size_t rb = (size_t)g.sx * ((size_t)y + (size_t)g.sy * z);
		for (int i = threadIdx.x; i < len; i += LW_BS) {
			int x = xs + i;
			uint32_t f = s_flag[0][i];
			uint32_t id = s_base[0] + s_pre[0][i];
			for (int a = 0; a < 3; a++) {
				if (!((f >> a) & 1u))
					continue;
				double *o = pts + (size_t)id * 3;
				if (MODE == LW_RECORD_EXACT) {
					double q[3];
					lw_edge_pos_exact(img, g, x, y, z, a, q);
					o[0] = q[0] + (double)g.out0;
					o[1] = q[1] + (double)g.out1;
					o[2] = q[2] + (double)g.out2;
				} else {
					float p[3];
					lw_edge_pos(img, g, x, y, z, a, p);
					o[0] = (double)(p[0] + (float)g.out0);
					o[1] = (double)(p[1] + (float)g.out1);
					o[2] = (double)(p[2] + (float)g.out2);
				}
				if (MODE != LW_PLAIN)
					vrec[id] = ((unsigned long long)(rb + x) << 2) | (unsigned)a;
				id++;
			}
		}
	}
	if (!cuberow)
		return;
	const size_t si = (size_t)seg + (size_t)g.nseg * ((size_t)y + (size_t)g.sy * z);
	const uint32_t tbase = bases[bstride + si];
	const uint32_t cbase = centre_add + bases[2 * bstride + si];
	uint32_t tcarry = 0, ccarry = 0;
	for (int i0 = 0; i0 < ncube; i0 += LW_BS) {
		int i = i0 + threadIdx.x;
		int row = -1;
		if (i < ncube)
			row = (int)s_crow[i] - 1;
		if (!__syncthreads_or(row >= 0))
			continue; // no cube of this chunk has triangles
		uint32_t nt = row >= 0 ? n2m_lw_row_ntri[row] : 0u;
		uint32_t hc = (row >= 0 && !g.originalMC && n2m_lw_row_centre[row]) ? 1u : 0u;
		// one scan of (triangles | centres << 16): a chunk holds at most
		// LW_BS * 12 triangles and LW_BS centres
		uint32_t ptot;
		const uint32_t poff = block_excl_scan<LW_BS>(nt | (hc << 16), scratch, &ptot);
		const uint32_t toff = poff & 0xffffu, coff = poff >> 16;
		const uint32_t ttot = ptot & 0xffffu, ctot = ptot >> 16;
		if (row >= 0) {
			const int x = xs + i;
			// vertex id of cube edge code e
			uint32_t ids[12];
#pragma unroll
			for (int e = 0; e < 12; e++) {
				int dx = c_code[e][0], dy = c_code[e][1], dz = c_code[e][2], a = c_code[e][3];
				int r = dy + 2 * dz;
				uint32_t f = s_flag[r][i + dx];
				ids[e] = ((f >> a) & 1u) ? s_base[r] + s_pre[r][i + dx] + __popc(f & ((1u << a) - 1u))
				                         : 0xffffffffu;
			}
			uint32_t cid = 0xffffffffu;
			if (hc) { // add_c_vertex(): mean of the cube's edge vertices, in float
				float sxv = 0.f, syv = 0.f, szv = 0.f, cnt = 0.f;
				for (int e = 0; e < 12; e++) {
					if (ids[e] == 0xffffffffu)
						continue;
					float p[3];
					lw_edge_pos(img, g, x + c_code[e][0], y + c_code[e][1], z + c_code[e][2], c_code[e][3], p);
					++cnt;
					sxv += p[0];
					syv += p[1];
					szv += p[2];
				}
				if (cnt > 0.f) {
					sxv /= cnt;
					syv /= cnt;
					szv /= cnt;
				}
				cid = cbase + ccarry + coff;
				double *o = pts + (size_t)cid * 3;
				o[0] = (double)(sxv + (float)g.out0);
				o[1] = (double)(syv + (float)g.out1);
				o[2] = (double)(szv + (float)g.out2);
				if (MODE != LW_PLAIN) {
					const uint32_t ci = cid - centre_add;
					vrec[cid] = (1ull << 63) | ci;
					for (int e = 0; e < 12; e++)
						crec[(size_t)ci * 12 + e] = ids[e];
				}
				if (MODE == LW_RECORD_EXACT) { // mean of the FP64 edge positions
					double ex = 0.0, ey = 0.0, ez = 0.0, dc = 0.0;
					for (int e = 0; e < 12; e++) {
						if (ids[e] == 0xffffffffu)
							continue;
						double q[3];
						lw_edge_pos_exact(img, g, x + c_code[e][0], y + c_code[e][1], z + c_code[e][2],
						                  c_code[e][3], q);
						dc += 1.0;
						ex += q[0];
						ey += q[1];
						ez += q[2];
					}
					if (dc > 0.0) {
						ex /= dc;
						ey /= dc;
						ez /= dc;
					}
					o[0] = ex + (double)g.out0;
					o[1] = ey + (double)g.out1;
					o[2] = ez + (double)g.out2;
				}
			}
			const uint8_t *codes = &n2m_lw_edges[n2m_lw_row_off[row]];
			int *t = tris + (size_t)(tbase + tcarry + toff) * 3;
			for (uint32_t k = 0; k < nt; k++) {
				uint32_t v[3];
				for (int j = 0; j < 3; j++) {
					int e = codes[3 * k + j];
					v[j] = e == 12 ? cid : ids[e];
				}
				t[3 * k + 0] = (int)v[2];
				t[3 * k + 1] = (int)v[1];
				t[3 * k + 2] = (int)v[0];
			}
		}
		tcarry += ttot;
		ccarry += ctot;
	}
}

// The whole-volume pass B with one wave per row segment, from the per-axis
// flag masks of pass A: the output of lw_emit_row (the same vertex numbers,
// the same triangles in the same order).  A point's vertex number is its
// row's base plus the edge vertices before it: those of the chunks before
// (a running sum of mask popcounts) and those of its chunk below it
// (popcounts of the masks under its bit), so no per-point prefix pass is
// needed; the four point rows' masks (27 words each) are read into LDS with
// two loads per lane.
static_assert(LW_SEG % N2M_WAVE == 0, "segments are whole chunks");
constexpr int LW_SCH = LW_SEG / N2M_WAVE; // chunks per segment

struct LwMaskRows {
	unsigned long long m[4][LW_SCH + 1][3]; // rows (dy + 2 dz) x chunks (the last: the next segment's first) x axes
	unsigned long long cubes[LW_SCH];       // the owned row's cubes with a tiling row
	uint16_t cb[4][LW_SCH + 1];             // edge vertices of the row segment before each chunk
};

// number (from the row's base) of the a-edge vertex of point i of row r,
// or 0xffffffff when that edge has none
__device__ __forceinline__ uint32_t lw_mask_vid(const LwMaskRows &M, int r, int i, int a)
{
	const int c = i >> 6, b = i & 63;
	const unsigned long long m0 = M.m[r][c][0], m1 = M.m[r][c][1], m2 = M.m[r][c][2];
	const unsigned long long ma = a == 0 ? m0 : (a == 1 ? m1 : m2);
	if (!((ma >> b) & 1ull))
		return 0xffffffffu;
	const unsigned long long lo = (1ull << b) - 1ull;
	uint32_t id = M.cb[r][c] + (uint32_t)(__popcll(m0 & lo) + __popcll(m1 & lo) + __popcll(m2 & lo));
	if (a > 0)
		id += (uint32_t)((m0 >> b) & 1ull);
	if (a > 1)
		id += (uint32_t)((m1 >> b) & 1ull);
	return id;
}

__global__ void __launch_bounds__(N2M_WAVE)
k_lw_emit_m(const float *__restrict__ img, LwGeo g, const unsigned long long *__restrict__ fmask,
            const uint16_t *__restrict__ crow, const uint32_t *__restrict__ segcnt,
            const uint32_t *__restrict__ bases, uint32_t nedge, double *__restrict__ pts, int *__restrict__ tris)
{
	__shared__ LwMaskRows M;
	__shared__ uint16_t s_vl[3 * LW_SEG]; // the owned row's edge vertices in number order: point << 2 | axis
	const int seg = blockIdx.x, y = blockIdx.y, z = blockIdx.z, lane = (int)threadIdx.x;
	const size_t ns = (size_t)g.nseg * g.sy * g.sz;
	const size_t si = (size_t)seg + (size_t)g.nseg * ((size_t)y + (size_t)g.sy * z);
	if (segcnt[si] == 0 && segcnt[ns + si] == 0 && segcnt[2 * ns + si] == 0)
		return;
	const int xs = seg * LW_SEG;
	const bool cuberow = y < g.sy - 1 && z < g.sz - 1;
	const int nrows = cuberow ? 4 : 1;
	const int nchunk = (g.sx + N2M_WAVE - 1) / N2M_WAVE;
	// the masks of the four point rows; chunks past the row end are empty
	constexpr int NW = 4 * (LW_SCH + 1) * 3;
	unsigned long long *mflat = &M.m[0][0][0];
	for (int e = lane; e < NW; e += N2M_WAVE) {
		const int r = e / ((LW_SCH + 1) * 3), c = (e / 3) % (LW_SCH + 1), a = e % 3;
		const int ch = seg * LW_SCH + c;
		unsigned long long v = 0ull;
		if (r < nrows && ch < nchunk) {
			const int yy = y + (r & 1), zz = z + (r >> 1);
			v = fmask[((size_t)(yy + (size_t)g.sy * zz) * nchunk + ch) * LW_FMW + a];
		}
		mflat[e] = v;
	}
	if (lane < LW_SCH) {
		const int ch = seg * LW_SCH + lane;
		M.cubes[lane] =
		    cuberow && ch < nchunk ? fmask[((size_t)(y + (size_t)g.sy * z) * nchunk + ch) * LW_FMW + 3] : 0ull;
	}
	uint32_t base[4] = {0u, 0u, 0u, 0u};
#pragma unroll
	for (int r = 0; r < 4; r++)
		if (r < nrows)
			base[r] = bases[(size_t)seg + (size_t)g.nseg * ((size_t)(y + (r & 1)) + (size_t)g.sy * (z + (r >> 1)))];
	__syncthreads(); // one wave: orders the LDS writes before other lanes' reads
	// per row, the edge vertices before each chunk (the segment owns its
	// points: bits past the row end are zero, and the chunk after the
	// segment, the next one's, is not counted)
	if (lane < 4) {
		uint32_t run = 0;
		for (int c = 0; c <= LW_SCH; c++) {
			M.cb[lane][c] = (uint16_t)run;
			if (c < LW_SCH)
				run += (uint32_t)(__popcll(M.m[lane][c][0]) + __popcll(M.m[lane][c][1]) + __popcll(M.m[lane][c][2]));
		}
	}
	__syncthreads();
	// the owned row's edge vertices, listed in number order
	const uint32_t nv0 = M.cb[0][LW_SCH];
	for (int c = 0; c < LW_SCH; c++) {
		const unsigned long long m0 = M.m[0][c][0], m1 = M.m[0][c][1], m2 = M.m[0][c][2];
		if (!(m0 | m1 | m2))
			continue;
		const uint32_t f = (uint32_t)((m0 >> lane) & 1ull) | (uint32_t)((m1 >> lane) & 1ull) << 1 |
		                   (uint32_t)((m2 >> lane) & 1ull) << 2;
		if (f) {
			const unsigned long long lo = (1ull << lane) - 1ull;
			uint32_t at = M.cb[0][c] + (uint32_t)(__popcll(m0 & lo) + __popcll(m1 & lo) + __popcll(m2 & lo));
			const int i = c * N2M_WAVE + lane;
			for (int a = 0; a < 3; a++)
				if ((f >> a) & 1u)
					s_vl[at++] = (uint16_t)((i << 2) | a);
		}
	}
	// the tiling rows of the cubes that have one (crow holds only those),
	// all loaded before the first is used
	const size_t rb0 = (size_t)g.sx * ((size_t)y + (size_t)g.sy * z);
	int rows[LW_SCH];
#pragma unroll
	for (int c = 0; c < LW_SCH; c++) {
		const int i = c * N2M_WAVE + lane;
		rows[c] = ((M.cubes[c] >> lane) & 1ull) ? (int)crow[rb0 + xs + i] - 1 : -1;
	}
	__syncthreads();
	// edge vertices of the owned points (row 0), one vertex per lane
	for (uint32_t v = lane; v < nv0; v += N2M_WAVE) {
		const uint32_t e = s_vl[v];
		float p[3];
		lw_edge_pos(img, g, xs + (int)(e >> 2), y, z, (int)(e & 3u), p);
		double *o = pts + (size_t)(base[0] + v) * 3;
		o[0] = (double)(p[0] + (float)g.out0);
		o[1] = (double)(p[1] + (float)g.out1);
		o[2] = (double)(p[2] + (float)g.out2);
	}
	if (!cuberow)
		return;
	const uint32_t tbase = bases[ns + si];
	const uint32_t cbase = nedge + bases[2 * ns + si];
	uint32_t tcarry = 0, ccarry = 0;
#pragma unroll
	for (int c = 0; c < LW_SCH; c++) {
		const int i = c * N2M_WAVE + lane;
		const int row = rows[c];
		if (!__ballot(row >= 0))
			continue; // no cube of this chunk has triangles
		const uint32_t nt = row >= 0 ? n2m_lw_row_ntri[row] : 0u;
		const uint32_t hc = (row >= 0 && !g.originalMC && n2m_lw_row_centre[row]) ? 1u : 0u;
		// a chunk holds at most 64 * 12 triangles and 64 centres
		const uint32_t packed = nt | (hc << 16);
		const uint32_t incl = wave_incl_scan(packed);
		const uint32_t excl = incl - packed, tot = __shfl(incl, N2M_WAVE - 1, N2M_WAVE);
		if (row >= 0) {
			const int x = xs + i;
			uint32_t ids[12];
#pragma unroll
			for (int e = 0; e < 12; e++) {
				const int dx = c_code[e][0], dy = c_code[e][1], dz = c_code[e][2], a = c_code[e][3];
				const int r = dy + 2 * dz;
				const uint32_t k = lw_mask_vid(M, r, i + dx, a);
				ids[e] = k == 0xffffffffu ? k : base[r] + k;
			}
			uint32_t cid = 0xffffffffu;
			if (hc) { // add_c_vertex(): mean of the cube's edge vertices, in float
				float sxv = 0.f, syv = 0.f, szv = 0.f, cn = 0.f;
				for (int e = 0; e < 12; e++) {
					if (ids[e] == 0xffffffffu)
						continue;
					float p[3];
					lw_edge_pos(img, g, x + c_code[e][0], y + c_code[e][1], z + c_code[e][2], c_code[e][3], p);
					++cn;
					sxv += p[0];
					syv += p[1];
					szv += p[2];
				}
				if (cn > 0.f) {
					sxv /= cn;
					syv /= cn;
					szv /= cn;
				}
				cid = cbase + ccarry + (excl >> 16);
				double *o = pts + (size_t)cid * 3;
				o[0] = (double)(sxv + (float)g.out0);
				o[1] = (double)(syv + (float)g.out1);
				o[2] = (double)(szv + (float)g.out2);
			}
			const uint8_t *codes = &n2m_lw_edges[n2m_lw_row_off[row]];
			int *t = tris + (size_t)(tbase + tcarry + (excl & 0xffffu)) * 3;
			for (uint32_t k = 0; k < nt; k++) {
				uint32_t v[3];
				for (int j = 0; j < 3; j++) {
					const int e = codes[3 * k + j];
					v[j] = e == 12 ? cid : ids[e];
				}
				t[3 * k + 0] = (int)v[2];
				t[3 * k + 1] = (int)v[1];
				t[3 * k + 2] = (int)v[0];
			}
		}
		tcarry += tot & 0xffffu;
		ccarry += tot >> 16;
	}
}

template <int MODE>
__global__ void __launch_bounds__(LW_BS)
k_lw_emit_rec(const float *__restrict__ img, LwGeo g, const uint8_t *__restrict__ pflags,
              const uint16_t *__restrict__ crow, const uint32_t *__restrict__ segcnt,
              const uint32_t *__restrict__ bases, uint32_t nedge, double *__restrict__ pts, int *__restrict__ tris,
              unsigned long long *__restrict__ vrec, uint32_t *__restrict__ crec)
{
	const size_t ns = (size_t)g.nseg * g.sy * g.sz;
	const size_t si = (size_t)blockIdx.x + (size_t)g.nseg * ((size_t)blockIdx.y + (size_t)g.sy * blockIdx.z);
	lw_emit_row<MODE>(img, g, blockIdx.x, blockIdx.y, blockIdx.z, pflags, crow, bases, ns, 0u, nedge, pts, tris,
	                  segcnt + si, ns, vrec, crec);
}

void lw_geo(LwGeo &g, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso, float mnv,
            int originalMC)
{
	g.nx = nx;
	g.nxy = (long long)nx * ny;
	g.nxyz = (long long)nx * ny * nz;
	g.lo0 = lo[0];
	g.lo1 = lo[1];
	g.lo2 = lo[2];
	g.sx = hi[0] - lo[0] + 1;
	g.sy = hi[1] - lo[1] + 1;
	g.sz = hi[2] - lo[2] + 1;
	g.iso = iso;
	g.mnv = mnv;
	g.originalMC = originalMC;
	g.nseg = (g.sx + LW_SEG - 1) / LW_SEG;
	g.out0 = lo[0];
	g.out1 = lo[1];
	g.out2 = lo[2];
}

size_t lw_nseg_total(int nx, int ny, int nz, const int lo[3], const int hi[3])
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, 0.f, 0.f, 0);
	return (size_t)g.nseg * g.sy * g.sz;
}

void launch_lw_classify(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                        float mnv, int originalMC, uint8_t *pflags, uint16_t *crow, uint32_t *segcnt,
                        hipStream_t s, unsigned long long *fmask)
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, iso, mnv, originalMC);
	hipLaunchKernelGGL(k_lw_classify, dim3(g.nseg, g.sy, (g.sz + LW_ZC - 1) / LW_ZC), dim3(LW_BS), 0, s, img, g, pflags,
	                   crow, segcnt, fmask);
}

void launch_lw_classify_bits(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                             float mnv, int originalMC, const uint32_t *P, const uint32_t *N, uint16_t *crow,
                             uint32_t *segcnt, unsigned long long *fmask, hipStream_t s)
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, iso, mnv, originalMC);
	SignBits sb;
	sb.P = P;
	sb.N = N;
	sb.nwords = (size_t)nx * ny * nz / 32; // nx % 32 == 0
	sb.nx = nx;
	sb.nxy = (long long)nx * ny;
	const float m = fabsf(mnv) < FLT_EPSILON ? FLT_EPSILON : mnv; // lw_clamp
	sb.padP = m > 0.f ? 0xffffffffu : 0u;
	sb.padN = m < 0.f ? 0xffffffffu : 0u;
	const size_t lanes = 8 * (size_t)g.nseg * g.sy * g.sz;
	hipLaunchKernelGGL(k_lw_classify_bits, dim3((unsigned)((lanes + LW_BS - 1) / LW_BS)), dim3(LW_BS), 0, s, img, g, sb,
	                   crow, segcnt, fmask);
}

size_t lw_fmask_words(int nx, int ny, int nz, const int lo[3], const int hi[3])
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, 0.f, 0.f, 0);
	return (size_t)g.sy * g.sz * ((g.sx + N2M_WAVE - 1) / N2M_WAVE) * LW_FMW;
}

void launch_lw_emit(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                    float mnv, int originalMC, const unsigned long long *fmask, const uint16_t *crow,
                    const uint32_t *segcnt, const uint32_t *bases, uint32_t nedge, double *pts, int *tris,
                    const int origin[3], hipStream_t s)
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, iso, mnv, originalMC);
	g.out0 = lo[0] + origin[0];
	g.out1 = lo[1] + origin[1];
	g.out2 = lo[2] + origin[2];
	hipLaunchKernelGGL(k_lw_emit_m, dim3(g.nseg, g.sy, g.sz), dim3(N2M_WAVE), 0, s, img, g, fmask, crow, segcnt,
	                   bases, nedge, pts, tris);
}

void launch_lw_emit_rec(const float *img, int nx, int ny, int nz, const int lo[3], const int hi[3], float iso,
                        float mnv, int originalMC, const uint8_t *pflags, const uint16_t *crow,
                        const uint32_t *segcnt, const uint32_t *bases, uint32_t nedge, double *pts, int *tris,
                        int exact,
                        unsigned long long *vrec, uint32_t *crec, hipStream_t s)
{
	LwGeo g;
	lw_geo(g, nx, ny, nz, lo, hi, iso, mnv, originalMC);
	g.out0 = lo[0];
	g.out1 = lo[1];
	g.out2 = lo[2];
	if (exact)
		hipLaunchKernelGGL(k_lw_emit_rec<LW_RECORD_EXACT>, dim3(g.nseg, g.sy, g.sz), dim3(LW_BS), 0, s, img, g,
		                   pflags, crow, segcnt, bases, nedge, pts, tris, vrec, crec);
	else
		hipLaunchKernelGGL(k_lw_emit_rec<LW_RECORD>, dim3(g.nseg, g.sy, g.sz), dim3(LW_BS), 0, s, img, g, pflags,
		                   crow, segcnt, bases, nedge, pts, tris, vrec, crec);
}

// ---------------------------------------------------------------------------
// Batched atlas: every workgroup serves one task (label crop).
// ---------------------------------------------------------------------------
__device__ __forceinline__ ClsGeo cls_geo_task(const AtlasTask &t)
{
	return ClsGeo{t.nx, t.ny, t.nz, t.lo3[0], t.lo3[1], t.lo3[2], t.hi3[0] - 1 - t.lo3[0], t.hi3[1] - 1 - t.lo3[1],
	              t.hi3[2] - 1 - t.lo3[2], t.iso, t.ox, t.oy, t.oz};
}

// The atlas kernels: workgroup b takes item b of the block map, or with
// LOOP (persistent) the grid walks the `nwork` items in strides of its own
// size, with a barrier between items (the per-item helpers reuse their LDS).
#define N2M_ATLAS_ITEMS(ITEM)                                                                                      \
	if (!LOOP) {                                                                                                  \
		ITEM(blockIdx.x);                                                                                          \
		return;                                                                                                    \
	}                                                                                                             \
	for (unsigned b = blockIdx.x; b < nwork; b += gridDim.x) {                                                    \
		ITEM(b);                                                                                                   \
		__syncthreads();                                                                                           \
	}

template <bool LOOP>
__global__ void __launch_bounds__(MC_BS)
k_atlas_classic_count(const float *__restrict__ crops, const AtlasTask *__restrict__ tasks,
                      const uint32_t *__restrict__ blockTask, unsigned nwork, uint8_t *__restrict__ cfgs,
                      uint32_t *__restrict__ blockSums)
{
	auto item = [&](unsigned b) {
		const AtlasTask t = tasks[blockTask[b]];
		const ClsGeo g = cls_geo_task(t);
		uint32_t s = classic_count_block(crops + t.off, g, cfgs + t.cube_off, b - t.first_block);
		if (threadIdx.x == 0)
			blockSums[b] = s;
	};
	N2M_ATLAS_ITEMS(item)
}

template <bool LOOP>
__global__ void __launch_bounds__(MC_BS)
k_atlas_classic_emit(const float *__restrict__ crops, const AtlasTask *__restrict__ tasks,
                     const uint32_t *__restrict__ blockTask, unsigned nwork, const uint8_t *__restrict__ cfgs,
                     const uint32_t *__restrict__ blockOffs, double *__restrict__ pts)
{
	auto item = [&](unsigned b) {
		const AtlasTask t = tasks[blockTask[b]];
		const ClsGeo g = cls_geo_task(t);
		classic_emit_block(crops + t.off, g, cfgs + t.cube_off, b - t.first_block, blockOffs[b], pts);
	};
	N2M_ATLAS_ITEMS(item)
}

__device__ __forceinline__ LwGeo lw_geo_task(const AtlasTask &t, int originalMC)
{
	LwGeo g;
	g.nx = t.nx;
	g.nxy = (long long)t.nx * t.ny;
	g.nxyz = g.nxy * t.nz;
	g.lo0 = t.lo3[0];
	g.lo1 = t.lo3[1];
	g.lo2 = t.lo3[2];
	g.sx = t.hi3[0] - t.lo3[0] + 1;
	g.sy = t.hi3[1] - t.lo3[1] + 1;
	g.sz = t.hi3[2] - t.lo3[2] + 1;
	g.iso = t.iso;
	g.mnv = t.mnv;
	g.originalMC = originalMC;
	g.nseg = t.nseg;
	g.out0 = t.lo3[0] + t.ox;
	g.out1 = t.lo3[1] + t.oy;
	g.out2 = t.lo3[2] + t.oz;
	return g;
}

// segcnt: 3 arrays `stride` entries apart, one entry per item (= segment)
template <bool LOOP>
__global__ void __launch_bounds__(LW_BS)
k_atlas_lw_classify(const float *__restrict__ crops, const AtlasTask *__restrict__ tasks,
                    const uint32_t *__restrict__ blockTask, unsigned nwork, int originalMC,
                    uint8_t *__restrict__ pflags, uint16_t *__restrict__ crow, uint32_t *__restrict__ segcnt,
                    size_t stride)
{
	auto item = [&](unsigned b) {
		const AtlasTask t = tasks[blockTask[b]];
		const LwGeo g = lw_geo_task(t, originalMC);
		const int local = (int)b - t.first_block;
		const int seg = local % g.nseg, y = (local / g.nseg) % g.sy, z = local / (g.nseg * g.sy);
		lw_classify_row(crops + t.off, g, seg, y, z, pflags + t.cube_off, crow + t.cube_off, segcnt + b, stride);
	};
	N2M_ATLAS_ITEMS(item)
}

// Vertex numbers follow the reference within each task: the task's edge
// vertices, then its centre vertices.  With E and C the exclusive scans of
// the per-segment edge / centre counts over all tasks, task t (first segment
// f, next task's first segment f') starts at E(f) + C(f); its edge vertex
// of segment s is E(s) + C(f) and its centre vertex is E(f') + C(s).
template <bool LOOP>
__global__ void __launch_bounds__(LW_BS)
k_atlas_lw_emit(const float *__restrict__ crops, const AtlasTask *__restrict__ tasks, int ntasks,
                const uint32_t *__restrict__ blockTask, unsigned nwork, int originalMC,
                const uint8_t *__restrict__ pflags, const uint16_t *__restrict__ crow,
                const uint32_t *__restrict__ segcnt, const uint32_t *__restrict__ bases, size_t stride,
                uint32_t totalE, double *__restrict__ pts, int *__restrict__ tris)
{
	auto item = [&](unsigned b) {
		const int ti = (int)blockTask[b];
		const AtlasTask t = tasks[ti];
		const LwGeo g = lw_geo_task(t, originalMC);
		const int local = (int)b - t.first_block;
		const int seg = local % g.nseg, y = (local / g.nseg) % g.sy, z = local / (g.nseg * g.sy);
		const uint32_t cbase = bases[2 * stride + t.first_block];
		const uint32_t enext = ti + 1 < ntasks ? bases[tasks[ti + 1].first_block] : totalE;
		lw_emit_row(crops + t.off, g, seg, y, z, pflags + t.cube_off, crow + t.cube_off, bases + t.first_block,
		            stride, cbase, enext, pts, tris, segcnt + b, stride);
	};
	N2M_ATLAS_ITEMS(item)
}
#undef N2M_ATLAS_ITEMS

// the persistent form only when the grid is smaller than the work
#define N2M_ATLAS_LAUNCH(K, BS, ...)                                                                               \
	if (!nblocks)                                                                                                 \
		return;                                                                                                    \
	if (grid < nblocks)                                                                                           \
		hipLaunchKernelGGL(K<true>, dim3(grid), dim3(BS), 0, s, __VA_ARGS__);                                      \
	else                                                                                                          \
		hipLaunchKernelGGL(K<false>, dim3(nblocks), dim3(BS), 0, s, __VA_ARGS__);

void launch_atlas_classic_count(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                                unsigned nblocks, unsigned grid, uint8_t *cfgs, uint32_t *blockSums, hipStream_t s)
{
	N2M_ATLAS_LAUNCH(k_atlas_classic_count, MC_BS, crops, tasks, blockTask, nblocks, cfgs, blockSums)
}
void launch_atlas_classic_emit(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                               unsigned nblocks, unsigned grid, const uint8_t *cfgs, const uint32_t *blockOffs,
                               double *pts, hipStream_t s)
{
	N2M_ATLAS_LAUNCH(k_atlas_classic_emit, MC_BS, crops, tasks, blockTask, nblocks, cfgs, blockOffs, pts)
}
void launch_atlas_lw_classify(const float *crops, const AtlasTask *tasks, const uint32_t *blockTask,
                              unsigned nblocks, unsigned grid, int originalMC, uint8_t *pflags, uint16_t *crow,
                              uint32_t *segcnt, size_t stride, hipStream_t s)
{
	N2M_ATLAS_LAUNCH(k_atlas_lw_classify, LW_BS, crops, tasks, blockTask, nblocks, originalMC, pflags, crow, segcnt,
	                 stride)
}
void launch_atlas_lw_emit(const float *crops, const AtlasTask *tasks, int ntasks, const uint32_t *blockTask,
                          unsigned nblocks, unsigned grid, int originalMC, const uint8_t *pflags, const uint16_t *crow,
                          const uint32_t *segcnt, const uint32_t *bases, size_t stride, uint32_t totalE,
                          double *pts, int *tris, hipStream_t s)
{
	N2M_ATLAS_LAUNCH(k_atlas_lw_emit, LW_BS, crops, tasks, ntasks, blockTask, nblocks, originalMC, pflags, crow,
	                 segcnt, bases, stride, totalE, pts, tris)
}
#undef N2M_ATLAS_LAUNCH
int lw_seg_len() { return LW_SEG; }

} // namespace n2m
