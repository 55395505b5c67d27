// k_encode.hip -- mesh output encoding on gfx950 (the byte streams of the
// reference's writers, src/meshify.c:602-930):
//   * float64 -> float32 vertices, as vec3d2vec4s (src/meshify.c:414-416);
//   * PLY face records: uint8 3 then three little-endian int32
//     (src/meshify.c:806-813);
//   * base64 without line breaks, '=' padded (src/base64.c, RFC 1341), for
//     the GIfTI data arrays;
//   * binary STL facets: a zero normal, the three corners as float32 and a
//     zero 16-bit spacer, 50 bytes (src/meshify.c:725-760);
//   * big-endian 32-bit words (FreeSurfer and VTK, whose writers swap every
//     word: src/meshify.c:418-456, 932-968) and VTK's polygon records
//     (3, v0, v1, v2).
#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

static unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535); }

__global__ void k_f64_to_f32(const double *__restrict__ in, size_t n, float *__restrict__ out)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] = (float)in[i];
}

__global__ void k_ply_faces(const int *__restrict__ tris, size_t nt, uint8_t *__restrict__ out)
{
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nt; t += (size_t)gridDim.x * blockDim.x) {
		uint8_t *o = out + 13 * t;
		o[0] = 3;
		for (int k = 0; k < 3; k++) {
			const uint32_t v = (uint32_t)tris[3 * t + k];
			o[1 + 4 * k] = (uint8_t)v;
			o[2 + 4 * k] = (uint8_t)(v >> 8);
			o[3 + 4 * k] = (uint8_t)(v >> 16);
			o[4 + 4 * k] = (uint8_t)(v >> 24);
		}
	}
}

__device__ __forceinline__ uint8_t b64(uint32_t v)
{
	// 'A'-'Z', 'a'-'z', '0'-'9', '+', '/'
	return (uint8_t)(v < 26 ? 'A' + v : v < 52 ? 'a' + (v - 26) : v < 62 ? '0' + (v - 52) : v == 62 ? '+' : '/');
}

// one thread per 3-byte group; the last group may be short
__global__ void k_base64(const uint8_t *__restrict__ src, size_t len, uint8_t *__restrict__ out)
{
	const size_t ng = (len + 2) / 3;
	for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < ng; g += (size_t)gridDim.x * blockDim.x) {
		const size_t i = 3 * g, left = len - i;
		const uint32_t a = src[i], b = left > 1 ? src[i + 1] : 0u, c = left > 2 ? src[i + 2] : 0u;
		uint8_t *o = out + 4 * g;
		o[0] = b64(a >> 2);
		o[1] = b64(((a & 3u) << 4) | (b >> 4));
		o[2] = left > 1 ? b64(((b & 15u) << 2) | (c >> 6)) : (uint8_t)'=';
		o[3] = left > 2 ? b64(c & 63u) : (uint8_t)'=';
	}
}

// one thread per facet; records are 2-byte aligned, so they are written as
// 25 16-bit halves
__global__ void k_stl_facets(const double *__restrict__ pts, const int *__restrict__ tris, size_t nt,
                             uint16_t *__restrict__ out)
{
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nt; t += (size_t)gridDim.x * blockDim.x) {
		uint16_t *o = out + 25 * t;
		for (int k = 0; k < 6; k++)
			o[k] = 0; // the normal, not written by the reference
		for (int c = 0; c < 3; c++) {
			const double *p = pts + 3 * (size_t)tris[3 * t + c];
			for (int k = 0; k < 3; k++) {
				const uint32_t u = __float_as_uint((float)p[k]);
				o[6 + 6 * c + 2 * k] = (uint16_t)u;
				o[7 + 6 * c + 2 * k] = (uint16_t)(u >> 16);
			}
		}
		o[24] = 0; // spacer
	}
}

__device__ __forceinline__ uint32_t bswap32(uint32_t v)
{
	return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

__global__ void k_bswap32(const uint32_t *__restrict__ in, size_t n, uint32_t *__restrict__ out)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] = bswap32(in[i]);
}

__global__ void k_vtk_polys(const int *__restrict__ tris, size_t nt, uint32_t *__restrict__ out)
{
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nt; t += (size_t)gridDim.x * blockDim.x) {
		out[4 * t] = bswap32(3u);
		for (int k = 0; k < 3; k++)
			out[4 * t + 1 + k] = bswap32((uint32_t)tris[3 * t + k]);
	}
}

void launch_stl_facets(const double *pts, const int *tris, size_t nt, uint16_t *out, hipStream_t s)
{
	if (nt)
		hipLaunchKernelGGL(k_stl_facets, dim3(blocks_for(nt)), dim3(256), 0, s, pts, tris, nt, out);
}
void launch_bswap32(const uint32_t *in, size_t n, uint32_t *out, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_bswap32, dim3(blocks_for(n)), dim3(256), 0, s, in, n, out);
}
void launch_vtk_polys(const int *tris, size_t nt, uint32_t *out, hipStream_t s)
{
	if (nt)
		hipLaunchKernelGGL(k_vtk_polys, dim3(blocks_for(nt)), dim3(256), 0, s, tris, nt, out);
}

void launch_f64_to_f32(const double *in, size_t n, float *out, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n)), dim3(256), 0, s, in, n, out);
}
void launch_ply_faces(const int *tris, size_t nt, uint8_t *out, hipStream_t s)
{
	if (nt)
		hipLaunchKernelGGL(k_ply_faces, dim3(blocks_for(nt)), dim3(256), 0, s, tris, nt, out);
}
void launch_base64(const uint8_t *src, size_t len, uint8_t *out, hipStream_t s)
{
	if (len)
		hipLaunchKernelGGL(k_base64, dim3(blocks_for((len + 2) / 3)), dim3(256), 0, s, src, len, out);
}

} // namespace n2m
