// k_volume.hip -- volume stages of meshify() on gfx950.
//
//  * pre-smooth (src/meshify.c:170-216) fused with the intensity range scan
//    (src/meshify.c:306-311) and a speculative threshold mask;
//  * threshold mask as packed bits (src/meshify.c:325-330);
//  * one-voxel dilation of a bit mask (src/meshify.c:218-264);
//  * fill/outside/edge-darken apply with bounding box (src/meshify.c:332-371);
//  * histograms of setThreshold() (src/isolevel.c:35-139, 245-262).
//
// Bit volumes: one uint32 word per 32 voxels of an x-row, W = ceil(nx/32)
// words per row, word index wx + W*(y + ny*z), bit b <-> x = 32*wx + b.
#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

// ---------------------------------------------------------------------------
// Pre-smooth.  2.5-D streaming: a workgroup owns a 64x8 column tile and a
// z-chunk; per input plane it stages the tile plus a 2-voxel halo in LDS,
// applies the x pass then the y pass (each rounded to float, as the
// reference stores between passes), and keeps the last five x/y-smoothed
// planes per column in registers for the z pass.  Every pass computes
// float*double products summed left to right in double (meshify.c:189).
// ---------------------------------------------------------------------------
constexpr int SM_TX = 64, SM_TY = 8, SM_ZC = 32, SM_BS = 256;
constexpr int SM_CPT = SM_TX * SM_TY / SM_BS; // columns per thread (2)

__device__ __forceinline__ float tap5(float a, float b, float c, float d, float e)
{
	const double w0 = 0.45, w1 = 0.225, w2 = 0.05;
	double acc = a * w2;
	acc = acc + b * w1;
	acc = acc + c * w0;
	acc = acc + d * w1;
	acc = acc + e * w2;
	return (float)acc;
}

__global__ void __launch_bounds__(SM_BS)
k_smooth(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz,
         uint32_t *__restrict__ range /* [0]=ord min, [1]=ord max */, float iso,
         uint32_t *__restrict__ bits, int W)
{
	__shared__ float s_in[SM_TY + 4][SM_TX + 4];
	__shared__ float s_x[SM_TY + 4][SM_TX];
	__shared__ uint32_t s_red[2][SM_BS / N2M_WAVE];
	const int x0 = blockIdx.x * SM_TX, y0 = blockIdx.y * SM_TY;
	const int z0 = blockIdx.z * SM_ZC, z1 = min(z0 + SM_ZC, nz);
	const int tx = threadIdx.x % SM_TX, ty = threadIdx.x / SM_TX; // ty 0..3
	const size_t nxy = (size_t)nx * ny;
	float ring[SM_CPT][5];
#pragma unroll
	for (int c = 0; c < SM_CPT; c++)
#pragma unroll
		for (int k = 0; k < 5; k++)
			ring[c][k] = 0.f;
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	const int x = x0 + tx;
	for (int zz = z0 - 2; zz <= z1 + 1; zz++) {
		const bool have = zz >= 0 && zz < nz;
		if (have) {
			const float *plane = in + (size_t)zz * nxy;
			for (int i = threadIdx.x; i < (SM_TY + 4) * (SM_TX + 4); i += SM_BS) {
				int ly = i / (SM_TX + 4), lx = i % (SM_TX + 4);
				int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
				float v = 0.f;
				if (gx >= 0 && gx < nx && gy >= 0 && gy < ny)
					v = plane[(size_t)gy * nx + gx];
				s_in[ly][lx] = v;
			}
			__syncthreads();
			for (int i = threadIdx.x; i < (SM_TY + 4) * SM_TX; i += SM_BS) {
				int ly = i / SM_TX, lx = i % SM_TX;
				int gx = x0 + lx;
				float v;
				if (gx >= 2 && gx < nx - 2)
					v = tap5(s_in[ly][lx], s_in[ly][lx + 1], s_in[ly][lx + 2], s_in[ly][lx + 3],
					         s_in[ly][lx + 4]);
				else
					v = s_in[ly][lx + 2];
				s_x[ly][lx] = v;
			}
			__syncthreads();
		}
#pragma unroll
		for (int c = 0; c < SM_CPT; c++) {
			const int ly = ty + c * (SM_BS / SM_TX);
			const int y = y0 + ly;
			float v = 0.f;
			if (have) {
				if (y >= 2 && y < ny - 2)
					v = tap5(s_x[ly][tx], s_x[ly + 1][tx], s_x[ly + 2][tx], s_x[ly + 3][tx],
					         s_x[ly + 4][tx]);
				else
					v = s_x[ly + 2][tx];
			}
			ring[c][0] = ring[c][1];
			ring[c][1] = ring[c][2];
			ring[c][2] = ring[c][3];
			ring[c][3] = ring[c][4];
			ring[c][4] = v;
		}
		const int zo = zz - 2;
		if (zo >= z0 && zo < z1) {
#pragma unroll
			for (int c = 0; c < SM_CPT; c++) {
				const int y = y0 + ty + c * (SM_BS / SM_TX);
				float o;
				if (zo >= 2 && zo < nz - 2)
					o = tap5(ring[c][0], ring[c][1], ring[c][2], ring[c][3], ring[c][4]);
				else
					o = ring[c][2];
				const bool valid = x < nx && y < ny;
				if (valid) {
					out[(size_t)zo * nxy + (size_t)y * nx + x] = o;
					lmin = fminf(lmin, o);
					lmax = fmaxf(lmax, o);
				}
				if (bits) { // speculative threshold mask (the common iso)
					unsigned long long m = __ballot(valid && o >= iso);
					if (y < ny) {
						const int wx = x0 / 32;
						size_t wi = (size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * zo);
						if (tx == 0)
							bits[wi] = (uint32_t)m;
						else if (tx == 32 && wx + 1 < W)
							bits[wi + 1] = (uint32_t)(m >> 32);
					}
				}
			}
		}
		if (have)
			__syncthreads();
	}
	uint32_t omin = wave_min_u32(f2ord(lmin)), omax = wave_max_u32(f2ord(lmax));
	if (lane_id() == 0) {
		s_red[0][wave_id()] = omin;
		s_red[1][wave_id()] = omax;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < SM_BS / N2M_WAVE; w++) {
			omin = min(omin, s_red[0][w]);
			omax = max(omax, s_red[1][w]);
		}
		atomicMin(&range[0], omin);
		atomicMax(&range[1], omax);
	}
}

void launch_smooth(const float *in, float *out, int nx, int ny, int nz, uint32_t *range, float iso,
                   uint32_t *bits, hipStream_t s)
{
	dim3 grid((nx + SM_TX - 1) / SM_TX, (ny + SM_TY - 1) / SM_TY, (nz + SM_ZC - 1) / SM_ZC);
	int W = (nx + 31) / 32;
	hipLaunchKernelGGL(k_smooth, grid, dim3(SM_BS), 0, s, in, out, nx, ny, nz, range, iso, bits, W);
}

// ---------------------------------------------------------------------------
// Intensity range without smoothing (src/meshify.c:306-311).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_minmax(const float *__restrict__ img, size_t n, int vec4, uint32_t *range)
{
	__shared__ uint32_t s_red[2][4];
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	size_t n4 = vec4 ? n / 4 : 0;
	const float4 *v4 = reinterpret_cast<const float4 *>(img);
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
		float4 v = v4[i];
		lmin = fminf(fminf(lmin, v.x), fminf(v.y, fminf(v.z, v.w)));
		lmax = fmaxf(fmaxf(lmax, v.x), fmaxf(v.y, fmaxf(v.z, v.w)));
	}
	for (size_t i = n4 * 4 + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * blockDim.x) {
		lmin = fminf(lmin, img[i]);
		lmax = fmaxf(lmax, img[i]);
	}
	uint32_t omin = wave_min_u32(f2ord(lmin)), omax = wave_max_u32(f2ord(lmax));
	if (lane_id() == 0) {
		s_red[0][wave_id()] = omin;
		s_red[1][wave_id()] = omax;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < 4; w++) {
			omin = min(omin, s_red[0][w]);
			omax = max(omax, s_red[1][w]);
		}
		atomicMin(&range[0], omin);
		atomicMax(&range[1], omax);
	}
}

void launch_minmax(const float *img, size_t n, uint32_t *range, hipStream_t s)
{
	size_t blocks = (n / 4 + 255) / 256;
	blocks = blocks < 1 ? 1 : (blocks > 8192 ? 8192 : blocks);
	int vec4 = ((uintptr_t)img & 15) == 0; // float4 loads need 16-byte alignment
	hipLaunchKernelGGL(k_minmax, dim3((unsigned)blocks), dim3(256), 0, s, img, n, vec4, range);
}

// ---------------------------------------------------------------------------
// Threshold mask bits: bit = (img >= iso) (or img != 0 when `nonzero`).
// One wave covers 64 consecutive x of one row; rows are padded to 64.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_mask_bits(const float *__restrict__ img, int nx, int ny, int nz, int W, int xpad, float iso,
            int nonzero, uint32_t *__restrict__ bits)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t rows = (size_t)ny * nz;
	size_t row = t / xpad;
	int x = (int)(t % xpad);
	if (row >= rows)
		return; // whole waves exit together (xpad is a multiple of 64)
	bool p = false;
	if (x < nx) {
		float v = img[row * nx + x];
		p = nonzero ? (v != 0.0f) : (v >= iso);
	}
	unsigned long long m = __ballot(p);
	int wx = (x & ~63) / 32;
	size_t wi = (size_t)wx + (size_t)W * row;
	if ((x & 63) == 0)
		bits[wi] = (uint32_t)m;
	else if ((x & 63) == 32 && wx + 1 < W)
		bits[wi + 1] = (uint32_t)(m >> 32);
}

void launch_mask_bits(const float *img, int nx, int ny, int nz, float iso, int nonzero, uint32_t *bits,
                      hipStream_t s)
{
	int W = (nx + 31) / 32, xpad = (nx + 63) / 64 * 64;
	size_t total = (size_t)xpad * ny * nz;
	hipLaunchKernelGGL(k_mask_bits, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, img, nx, ny,
	                   nz, W, xpad, iso, nonzero, bits);
}

// ---------------------------------------------------------------------------
// dilate(mask, dim, true), src/meshify.c:218-264: a voxel at least one away
// from every face is switched on if any of its 26 neighbours except
// (-1,-1,-1) is on.  One thread per word.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t word_at(const uint32_t *bits, int W, int ny, int wx, int y, int z)
{
	return (wx >= 0 && wx < W) ? bits[(size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * z)] : 0u;
}

__global__ void __launch_bounds__(256)
k_dilate_bits(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int nx, int ny, int nz, int W)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t nw = (size_t)W * ny * nz;
	if (t >= nw)
		return;
	const int wx = (int)(t % W);
	const size_t row = t / W;
	const int y = (int)(row % ny), z = (int)(row / ny);
	uint32_t self = in[t];
	uint32_t acc = 0;
	if (y >= 1 && y < ny - 1 && z >= 1 && z < nz - 1) {
		for (int dz = -1; dz <= 1; dz++)
			for (int dy = -1; dy <= 1; dy++) {
				uint32_t c = word_at(in, W, ny, wx, y + dy, z + dz);
				uint32_t p = word_at(in, W, ny, wx - 1, y + dy, z + dz);
				uint32_t n = word_at(in, W, ny, wx + 1, y + dy, z + dz);
				uint32_t xm = (c << 1) | (p >> 31); // neighbour at x-1
				uint32_t xp = (c >> 1) | (n << 31); // neighbour at x+1
				if (dz == -1 && dy == -1)
					acc |= c | xp; // (-1,-1,-1) is not tested
				else if (dz == 0 && dy == 0)
					acc |= xm | xp;
				else
					acc |= xm | c | xp;
			}
		// interior in x: 1 <= x <= nx-2
		uint32_t xmask = 0xffffffffu;
		int xb = wx * 32;
		for (int b = 0; b < 32; b++) {
			int xx = xb + b;
			if (xx < 1 || xx > nx - 2)
				xmask &= ~(1u << b);
		}
		acc &= xmask;
	}
	out[t] = self | acc;
}

void launch_dilate_bits(const uint32_t *in, uint32_t *out, int nx, int ny, int nz, hipStream_t s)
{
	int W = (nx + 31) / 32;
	size_t nw = (size_t)W * ny * nz;
	hipLaunchKernelGGL(k_dilate_bits, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, in, out, nx, ny,
	                   nz, W);
}

// ---------------------------------------------------------------------------
// Apply (src/meshify.c:332-371): raise filled voxels to iso, set voxels
// outside the dilated largest cluster to mn, record the bounding box of
// voxels >= iso, then clamp the outer faces to edgeMax.
// keep: mask after bwlabel (fill target); dil: dilated mask (or null).
// bbox: int[6] = {minx,miny,minz,maxx,maxy,maxz}.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_apply(float *__restrict__ img, int nx, int ny, int nz, int W, int xpad, float iso, float mn,
        float edgeMax, const uint32_t *__restrict__ keep, const uint32_t *__restrict__ dil, int fill,
        int *__restrict__ bbox, uint32_t *__restrict__ vmin)
{
	__shared__ int s_red[6][4];
	__shared__ uint32_t s_min[4];
	float lmin = __builtin_inff();
	int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, bmax[3] = {-1, -1, -1};
	const size_t rows = (size_t)ny * nz;
	const size_t total = rows * xpad;
	for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
		size_t row = t / xpad;
		const int x = (int)(t % xpad);
		if (x >= nx)
			continue;
		const int y = (int)(row % ny), z = (int)(row / ny);
		size_t v = row * nx + x;
		size_t wi = (size_t)(x >> 5) + (size_t)W * row;
		float val = img[v];
		if (fill && ((keep[wi] >> (x & 31)) & 1u))
			val = (float)fmax((double)val, (double)iso);
		if (dil && !((dil[wi] >> (x & 31)) & 1u))
			val = mn;
		if (val >= iso) {
			bmin[0] = min(bmin[0], x);
			bmax[0] = max(bmax[0], x);
			bmin[1] = min(bmin[1], y);
			bmax[1] = max(bmax[1], y);
			bmin[2] = min(bmin[2], z);
			bmax[2] = max(bmax[2], z);
		}
		if (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1)
			val = fminf(edgeMax, val);
		img[v] = val;
		lmin = fminf(lmin, val);
	}
	for (int k = 0; k < 3; k++) {
		bmin[k] = wave_min_i32(bmin[k]);
		bmax[k] = wave_max_i32(bmax[k]);
	}
	uint32_t om = wave_min_u32(f2ord(lmin));
	if (lane_id() == 0) {
		for (int k = 0; k < 3; k++) {
			s_red[k][wave_id()] = bmin[k];
			s_red[3 + k][wave_id()] = bmax[k];
		}
		s_min[wave_id()] = om;
	}
	__syncthreads();
	if (threadIdx.x < 6) {
		int k = threadIdx.x;
		int r = s_red[k][0];
		for (int w = 1; w < 4; w++)
			r = k < 3 ? min(r, s_red[k][w]) : max(r, s_red[k][w]);
		// a plain read first: most workgroups cannot improve the box
		const int cur = (int)ld_agent((const uint32_t *)&bbox[k]);
		if (k < 3) {
			if (r != 0x7fffffff && r < cur)
				atomicMin(&bbox[k], r);
		} else if (r >= 0 && r > cur) {
			atomicMax(&bbox[k], r);
		}
	}
	if (threadIdx.x == 6) {
		uint32_t m = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
		if (m < ld_agent(vmin))
			atomicMin(vmin, m);
	}
}

void launch_apply(float *img, int nx, int ny, int nz, float iso, float mn, float edgeMax,
                  const uint32_t *keep, const uint32_t *dil, int fill, int *bbox, uint32_t *vmin, hipStream_t s)
{
	int W = (nx + 31) / 32, xpad = (nx + 63) / 64 * 64;
	size_t total = (size_t)xpad * ny * nz;
	size_t blocks = (total + 255) / 256;
	blocks = blocks > 4096 ? 4096 : (blocks < 1 ? 1 : blocks);
	hipLaunchKernelGGL(k_apply, dim3((unsigned)blocks), dim3(256), 0, s, img, nx, ny, nz, W, xpad, iso, mn,
	                   edgeMax, keep, dil, fill, bbox, vmin);
}

// ---------------------------------------------------------------------------
// Histograms for setThreshold() (src/isolevel.c).
// stats: [0]=ord min, [1]=ord max over non-NaN, [2]=NaN count.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_hist_range(const float *__restrict__ img, size_t n, uint32_t *stats)
{
	__shared__ uint32_t s_red[3][4];
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	uint32_t nan = 0;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		float v = img[i];
		if (v != v) {
			nan++;
			continue;
		}
		lmin = fminf(lmin, v);
		lmax = fmaxf(lmax, v);
	}
	uint32_t omin = wave_min_u32(f2ord(lmin)), omax = wave_max_u32(f2ord(lmax));
	nan = wave_sum(nan);
	if (lane_id() == 0) {
		s_red[0][wave_id()] = omin;
		s_red[1][wave_id()] = omax;
		s_red[2][wave_id()] = nan;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < 4; w++) {
			omin = min(omin, s_red[0][w]);
			omax = max(omax, s_red[1][w]);
			nan += s_red[2][w];
		}
		atomicMin(&stats[0], omin);
		atomicMax(&stats[1], omax);
		if (nan)
			atomicAdd(&stats[2], nan);
	}
}

// bin = round((v - mn) * scl) in float, clamped to [0, nbins-1] (the
// 1001-bin pass never leaves that range for finite data).
__global__ void __launch_bounds__(256)
k_hist_bins(const float *__restrict__ img, size_t n, float mn, float scl, int nbins, uint32_t *hist)
{
	extern __shared__ uint32_t s_h[];
	for (int i = threadIdx.x; i < nbins; i += blockDim.x)
		s_h[i] = 0;
	__syncthreads();
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		float v = img[i];
		if (v != v)
			continue;
		float f = (v - mn) * scl;
		int b = (int)roundf(f);
		b = b < 0 ? 0 : (b > nbins - 1 ? nbins - 1 : b);
		atomicAdd(&s_h[b], 1u);
	}
	__syncthreads();
	for (int i = threadIdx.x; i < nbins; i += blockDim.x)
		if (s_h[i])
			atomicAdd(&hist[i], s_h[i]);
}

static unsigned hist_blocks(size_t n)
{
	size_t b = (n + 255) / 256;
	return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

void launch_hist_range(const float *img, size_t n, uint32_t *stats, hipStream_t s)
{
	hipLaunchKernelGGL(k_hist_range, dim3(hist_blocks(n)), dim3(256), 0, s, img, n, stats);
}

void launch_hist_bins(const float *img, size_t n, float mn, float scl, int nbins, uint32_t *hist,
                      hipStream_t s)
{
	hipLaunchKernelGGL(k_hist_bins, dim3(hist_blocks(n)), dim3(256), nbins * sizeof(uint32_t), s, img, n,
	                   mn, scl, nbins, hist);
}

} // namespace n2m
