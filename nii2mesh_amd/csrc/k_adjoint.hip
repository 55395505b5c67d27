// k_adjoint.hip -- adjoint of the differentiable path on gfx950.
//
// Forward (api.cpp): Lewiner marching cubes over the volume at the
// isolevel (src/MarchingCubes.c:1086-1143), vertex weld (src/meshify.c:45-106),
// degenerate removal (:113-168) and lapIter plain Laplacian iterations
// (src/quadric.c:316-341).  Backward, from dL/dp of the final vertices:
//
//  1. the Laplacian is linear, p' = A p with A_ij = mult(i,j) / num_i
//     (num_i = 2 x triangle slots of i; mult symmetric), so A^T g at j is
//     the sum, over the neighbour pairs of j, of g_i / num_i; a vertex with
//     no triangle keeps its value.  No forward iterate is needed, so nothing
//     is checkpointed;
//  2. the weld copies the position of each class's last member: only that
//     marching-cubes vertex receives the class gradient;
//  3. a centre vertex is the mean of its cube's edge vertices: its gradient
//     is shared equally among them;
//  4. an edge vertex sits at x + u along its axis with u = c0 / (c0 - c1),
//     c = value - iso (FLT_EPSILON-clamped, src/MarchingCubes.c:132):
//     du/dc0 = -c1 / (c0 - c1)^2, du/dc1 = c0 / (c0 - c1)^2; a clamped
//     value, or a padding value outside the volume, gets no gradient.
//
// Arithmetic is FP64; the scatters of steps 3 and 4 use atomics, so the last
// bits of the result depend on their order.
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

static unsigned g1(size_t n) { return (unsigned)((n + 255) / 256); }

// h_i = g_i / num_i (3 components)
__global__ void k_adj_lap_scale(const uint32_t *__restrict__ start, const double *__restrict__ g, uint32_t n,
                                double *__restrict__ h)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t num = 2 * (start[i + 1] - start[i]);
	for (int k = 0; k < 3; k++)
		h[3 * (size_t)i + k] = num ? g[3 * (size_t)i + k] / (double)num : 0.0;
}

// g_j <- sum over the pairs (a, b) of j of h_a + h_b  (in place on g)
__global__ void k_adj_lap_gather(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs,
                                 const double *__restrict__ h, uint32_t n, double *__restrict__ g)
{
	uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n)
		return;
	uint32_t b = start[j], e = start[j + 1];
	if (b == e)
		return; // no triangle: the forward left the vertex unchanged
	double sx = 0.0, sy = 0.0, sz = 0.0;
	for (uint32_t k = b; k < e; k++) {
		int2 q = pairs[k];
		sx += h[3 * (size_t)q.x] + h[3 * (size_t)q.y];
		sy += h[3 * (size_t)q.x + 1] + h[3 * (size_t)q.y + 1];
		sz += h[3 * (size_t)q.x + 2] + h[3 * (size_t)q.y + 2];
	}
	g[3 * (size_t)j] = sx;
	g[3 * (size_t)j + 1] = sy;
	g[3 * (size_t)j + 2] = sz;
}

// gradient of each marching-cubes vertex from its class (cls null: identity)
__global__ void k_adj_unweld(const int *__restrict__ cls, const int *__restrict__ rep, const double *__restrict__ g,
                             uint32_t nmc, double *__restrict__ gm)
{
	uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= nmc)
		return;
	int c = cls ? cls[m] : (int)m;
	bool own = !cls || rep[c] == (int)m;
	for (int k = 0; k < 3; k++)
		gm[3 * (size_t)m + k] = own ? g[3 * (size_t)c + k] : 0.0;
}

__global__ void k_adj_centre(const unsigned long long *__restrict__ vrec, const uint32_t *__restrict__ crec,
                             uint32_t nmc, double *__restrict__ gm)
{
	uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= nmc)
		return;
	unsigned long long r = vrec[m];
	if (!(r >> 63))
		return;
	const uint32_t *ids = crec + (size_t)(r & 0xffffffffull) * 12;
	int cnt = 0;
	for (int e = 0; e < 12; e++)
		cnt += ids[e] != 0xffffffffu;
	if (!cnt)
		return;
	const double gx = gm[3 * (size_t)m] / cnt, gy = gm[3 * (size_t)m + 1] / cnt, gz = gm[3 * (size_t)m + 2] / cnt;
	for (int e = 0; e < 12; e++) {
		uint32_t id = ids[e];
		if (id == 0xffffffffu || id >= nmc)
			continue;
		atomicAdd(&gm[3 * (size_t)id], gx);
		atomicAdd(&gm[3 * (size_t)id + 1], gy);
		atomicAdd(&gm[3 * (size_t)id + 2], gz);
	}
}

__device__ __forceinline__ float adj_clamp(float v) { return fabsf(v) < FLT_EPSILON ? FLT_EPSILON : v; }

__global__ void k_adj_edge(const unsigned long long *__restrict__ vrec, uint32_t nmc, const double *__restrict__ gm,
                           const float *__restrict__ img, AdjGeo g, double *__restrict__ dvol)
{
	uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= nmc)
		return;
	unsigned long long r = vrec[m];
	if (r >> 63)
		return;
	const int a = (int)(r & 3ull);
	const long long pt = (long long)(r >> 2);
	const int x = (int)(pt % g.sx), y = (int)((pt / g.sx) % g.sy), z = (int)(pt / ((long long)g.sx * g.sy));
	// voxel of a sub-box point (the sub-box copy indexes the volume linearly)
	auto vox = [&](int px, int py, int pz) {
		return (long long)g.lo0 + px + (long long)(py + g.lo1) * g.nx + (long long)(pz + g.lo2) * g.nxy;
	};
	const long long ja = vox(x, y, z), jb = vox(x + (a == 0), y + (a == 1), z + (a == 2));
	const float r0 = ja < g.nxyz ? img[ja] - g.iso : g.mnv;
	const float r1 = jb < g.nxyz ? img[jb] - g.iso : g.mnv;
	const float c0 = adj_clamp(r0), c1 = adj_clamp(r1);
	const double dd = (double)c0 - (double)c1;
	if (dd == 0.0)
		return;
	const double gu = gm[3 * (size_t)m + a];
	if (ja < g.nxyz && c0 == r0)
		atomicAdd(&dvol[ja], gu * (-(double)c1 / (dd * dd)));
	if (jb < g.nxyz && c1 == r1)
		atomicAdd(&dvol[jb], gu * ((double)c0 / (dd * dd)));
}

void launch_adj_laplace_T(const uint32_t *start, const int2 *pairs, double *g, double *h, uint32_t n,
                          hipStream_t s)
{
	if (!n)
		return;
	hipLaunchKernelGGL(k_adj_lap_scale, dim3(g1(n)), dim3(256), 0, s, start, g, n, h);
	hipLaunchKernelGGL(k_adj_lap_gather, dim3(g1(n)), dim3(256), 0, s, start, pairs, h, n, g);
}

void launch_adj_unweld(const int *cls, const int *rep, const double *g, uint32_t nmc, double *gm, hipStream_t s)
{
	if (nmc)
		hipLaunchKernelGGL(k_adj_unweld, dim3(g1(nmc)), dim3(256), 0, s, cls, rep, g, nmc, gm);
}

void launch_adj_centre(const unsigned long long *vrec, const uint32_t *crec, uint32_t nmc, double *gm,
                       hipStream_t s)
{
	if (nmc)
		hipLaunchKernelGGL(k_adj_centre, dim3(g1(nmc)), dim3(256), 0, s, vrec, crec, nmc, gm);
}

void launch_adj_edge(const unsigned long long *vrec, uint32_t nmc, const double *gm, const float *img,
                     const AdjGeo &g, double *dvol, hipStream_t s)
{
	if (nmc)
		hipLaunchKernelGGL(k_adj_edge, dim3(g1(nmc)), dim3(256), 0, s, vrec, nmc, gm, img, g, dvol);
}

} // namespace n2m
