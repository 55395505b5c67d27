// k_quadric.hip -- the order-free first pass of the quadric simplification on
// gfx950: update_mesh(iteration 0) of src/quadric.c:144-215.
//
// For every triangle its unit normal and plane offset; for every vertex the
// sum of the plane quadrics of its triangles, added in triangle order (the
// order of the reference's loop over triangles, src/quadric.c:197-207); for
// every triangle the collapse error of its three edges (src/quadric.c:208-214,
// calculate_error :108-133).  The reference list and the border flags come
// from the smoothing CSR (api.cpp), whose lists are in ascending corner order.
// FP64 throughout, each expression in the reference's operation order (the
// library is built with -ffp-contract=off), so the values are the reference's
// bit for bit.
#include <hip/hip_runtime.h>

#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

static unsigned gq(size_t n) { return (unsigned)((n + 255) / 256); }

// symMat4 (src/quadric.c:46-51): the plane quadric of (a, b, c, d)
__device__ __forceinline__ void plane_quadric(double a, double b, double c, double d, double *q)
{
	q[0] = a * a;
	q[1] = a * b;
	q[2] = a * c;
	q[3] = a * d;
	q[4] = b * b;
	q[5] = b * c;
	q[6] = b * d;
	q[7] = c * c;
	q[8] = c * d;
	q[9] = d * d;
}

// symMatDet (src/quadric.c:65-68), terms and signs in the reference's order
__device__ __forceinline__ double q_det(const double *m, int a11, int a12, int a13, int a21, int a22, int a23,
                                        int a31, int a32, int a33)
{
	return m[a11] * m[a22] * m[a33] + m[a13] * m[a21] * m[a32] + m[a12] * m[a23] * m[a31] -
	       m[a13] * m[a22] * m[a31] - m[a11] * m[a23] * m[a32] - m[a12] * m[a21] * m[a33];
}

// vertex_error (src/quadric.c:103-106)
__device__ __forceinline__ double q_vertex_error(const double *q, double x, double y, double z)
{
	return q[0] * x * x + 2 * q[1] * x * y + 2 * q[2] * x * z + 2 * q[3] * x + q[4] * y * y + 2 * q[5] * y * z +
	       2 * q[6] * y + q[7] * z * z + 2 * q[8] * z + q[9];
}

// per triangle: unit normal (vCross of the two edges from corner 0,
// vNormalize with its 0.001 floor, src/quadric.c:87-93,199-201) and the
// plane offset -n.p0
__global__ void k_q_planes(const double *__restrict__ p, const int *__restrict__ tris, uint32_t nt,
                           double *__restrict__ nd)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	const int a = tris[3 * (size_t)i], b = tris[3 * (size_t)i + 1], c = tris[3 * (size_t)i + 2];
	const double p0x = p[3 * (size_t)a], p0y = p[3 * (size_t)a + 1], p0z = p[3 * (size_t)a + 2];
	const double ux = p[3 * (size_t)b] - p0x, uy = p[3 * (size_t)b + 1] - p0y, uz = p[3 * (size_t)b + 2] - p0z;
	const double vx = p[3 * (size_t)c] - p0x, vy = p[3 * (size_t)c + 1] - p0y, vz = p[3 * (size_t)c + 2] - p0z;
	double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
	double len = sqrt((nx * nx) + (ny * ny) + (nz * nz));
	if (len <= 0)
		len = 0.001;
	nx = nx / len;
	ny = ny / len;
	nz = nz / len;
	nd[4 * (size_t)i] = nx;
	nd[4 * (size_t)i + 1] = ny;
	nd[4 * (size_t)i + 2] = nz;
	nd[4 * (size_t)i + 3] = -(nx * p0x + ny * p0y + nz * p0z);
}

// per vertex: q = sum of the plane quadrics of its references, in ascending
// corner order (ref holds corners t*3+k, sorted per vertex)
__global__ void k_q_vertex(const uint32_t *__restrict__ start, const uint32_t *__restrict__ ref, uint32_t nv,
                           const double *__restrict__ nd, double *__restrict__ q)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	double s[10];
#pragma unroll
	for (int k = 0; k < 10; k++)
		s[k] = 0.0;
	for (uint32_t i = start[v]; i < start[v + 1]; i++) {
		const uint32_t t = ref[i] / 3;
		double m[10];
		plane_quadric(nd[4 * (size_t)t], nd[4 * (size_t)t + 1], nd[4 * (size_t)t + 2], nd[4 * (size_t)t + 3], m);
#pragma unroll
		for (int k = 0; k < 10; k++)
			s[k] = s[k] + m[k];
	}
#pragma unroll
	for (int k = 0; k < 10; k++)
		q[10 * (size_t)v + k] = s[k];
}

// calculate_error (src/quadric.c:108-133) of the edge (i1, i2); the vertex it
// would collapse to is not needed here
__device__ double q_edge_error(const double *__restrict__ q, const uint8_t *__restrict__ border,
                               const double *__restrict__ p, int i1, int i2)
{
	double m[10];
#pragma unroll
	for (int k = 0; k < 10; k++)
		m[k] = q[10 * (size_t)i1 + k] + q[10 * (size_t)i2 + k];
	const int bd = (border ? border[i1] : 0) + (border ? border[i2] : 0);
	const double det = q_det(m, 0, 1, 2, 1, 4, 5, 2, 5, 7);
	if ((det != 0.0) && (bd == 0)) {
		const double x = -1.0 / det * (q_det(m, 1, 2, 3, 4, 5, 6, 5, 7, 8));
		const double y = 1.0 / det * (q_det(m, 0, 2, 3, 1, 5, 6, 2, 7, 8));
		const double z = -1.0 / det * (q_det(m, 0, 1, 3, 1, 4, 6, 2, 5, 8));
		return q_vertex_error(m, x, y, z);
	}
	const double ax = p[3 * (size_t)i1], ay = p[3 * (size_t)i1 + 1], az = p[3 * (size_t)i1 + 2];
	const double bx = p[3 * (size_t)i2], by = p[3 * (size_t)i2 + 1], bz = p[3 * (size_t)i2 + 2];
	const double cx = (ax + bx) * 0.5, cy = (ay + by) * 0.5, cz = (az + bz) * 0.5;
	const double e1 = q_vertex_error(m, ax, ay, az);
	const double e2 = q_vertex_error(m, bx, by, bz);
	const double e3 = q_vertex_error(m, cx, cy, cz);
	return fmin(e1, fmin(e2, e3));
}

// per triangle: err[0..2] of its edges (j, j+1) and err[3] their minimum
__global__ void k_q_errors(const double *__restrict__ p, const int *__restrict__ tris, uint32_t nt,
                           const double *__restrict__ q, const uint8_t *__restrict__ border, double *__restrict__ err)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	const int v[3] = {tris[3 * (size_t)i], tris[3 * (size_t)i + 1], tris[3 * (size_t)i + 2]};
	double e[3];
	for (int j = 0; j < 3; j++)
		e[j] = q_edge_error(q, border, p, v[j], v[(j + 1) % 3]);
	err[4 * (size_t)i] = e[0];
	err[4 * (size_t)i + 1] = e[1];
	err[4 * (size_t)i + 2] = e[2];
	err[4 * (size_t)i + 3] = fmin(e[0], fmin(e[1], e[2]));
}

void launch_q_planes(const double *p, const int *tris, uint32_t nt, double *nd, hipStream_t s)
{
	if (nt)
		hipLaunchKernelGGL(k_q_planes, dim3(gq(nt)), dim3(256), 0, s, p, tris, nt, nd);
}
void launch_q_vertex(const uint32_t *start, const uint32_t *ref, uint32_t nv, const double *nd, double *q,
                     hipStream_t s)
{
	if (nv)
		hipLaunchKernelGGL(k_q_vertex, dim3(gq(nv)), dim3(256), 0, s, start, ref, nv, nd, q);
}
void launch_q_errors(const double *p, const int *tris, uint32_t nt, const double *q, const uint8_t *border,
                     double *err, hipStream_t s)
{
	if (nt)
		hipLaunchKernelGGL(k_q_errors, dim3(gq(nt)), dim3(256), 0, s, p, tris, nt, q, border, err);
}

} // namespace n2m
