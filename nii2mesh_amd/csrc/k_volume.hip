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
#include <float.h>

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
constexpr int SM_TX = 64, SM_TY = 8, SM_ZC = 32, SM_BS = 256; // 2 columns per thread

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

// Isolevels whose threshold masks the smoothing writes as it goes: mask k
// (out >= v[k]) goes to bits + k * nwords.
struct IsoMasks {
	float v[N2M_SMOOTH_MASKS];
	int n;
	size_t nwords;
};

// Smooth the TX x TY tile (x0, y0, [z0, z1)) of an nx*ny*nz volume read
// through `src(x, y, z)` (only called inside the volume) into out; returns
// the tile's min/max through lmin/lmax (per thread) and, for TX = 64 (one
// wave per tile row), writes speculative mask bits for the isolevels of
// `iso` when bits != null.
template <int TX, int TY, class Src>
__device__ __forceinline__ void smooth_tile(const Src &src, float *__restrict__ out, int nx, int ny, int nz,
                                            int x0, int y0, int z0, int z1, const IsoMasks &iso,
                                            uint32_t *__restrict__ bits, int W, float &lmin, float &lmax)
{
	constexpr int CPT = TX * TY / SM_BS, RPP = SM_BS / TX; // columns per thread, rows per pass
	static_assert(CPT * SM_BS == TX * TY && RPP * TX == SM_BS, "tile must tile the workgroup");
	__shared__ float s_in[TY + 4][TX + 4];
	__shared__ float s_x[TY + 4][TX];
	const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
	const size_t nxy = (size_t)nx * ny;
	float ring[CPT][5];
#pragma unroll
	for (int c = 0; c < CPT; c++)
#pragma unroll
		for (int k = 0; k < 5; k++)
			ring[c][k] = 0.f;
	const int x = x0 + tx;
	// the plane's halo tile is loaded into registers one plane ahead, so its
	// memory latency overlaps the previous plane's passes
	constexpr int NIN = (TY + 4) * (TX + 4), NL = (NIN + SM_BS - 1) / SM_BS;
	float pre[NL];
	auto fetch = [&](int zz) {
#pragma unroll
		for (int q = 0; q < NL; q++) {
			const int i = q * SM_BS + threadIdx.x;
			const int ly = i / (TX + 4), lx = i % (TX + 4);
			const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
			float v = 0.f;
			if (i < NIN && zz >= 0 && zz < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
				v = src(gx, gy, zz);
			pre[q] = v;
		}
	};
	fetch(z0 - 2);
	for (int zz = z0 - 2; zz <= z1 + 1; zz++) {
		const bool have = zz >= 0 && zz < nz;
		if (have) {
#pragma unroll
			for (int q = 0; q < NL; q++) {
				const int i = q * SM_BS + threadIdx.x;
				if (i < NIN)
					s_in[i / (TX + 4)][i % (TX + 4)] = pre[q];
			}
		}
		if (zz + 1 <= z1 + 1)
			fetch(zz + 1);
		if (have) {
			__syncthreads();
			for (int i = threadIdx.x; i < (TY + 4) * TX; i += SM_BS) {
				int ly = i / TX, lx = i % TX;
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
		for (int c = 0; c < CPT; c++) {
			const int ly = ty + c * RPP;
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
			for (int c = 0; c < CPT; c++) {
				const int y = y0 + ty + c * RPP;
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
				if (TX == 64 && bits) { // speculative threshold masks
					const int wx = x0 / 32;
					const size_t wi = (size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * zo);
					for (int k = 0; k < iso.n; k++) {
						const unsigned long long m = __ballot(valid && o >= iso.v[k]);
						uint32_t *bk = bits + k * iso.nwords;
						if (y < ny) {
							if (tx == 0)
								bk[wi] = (uint32_t)m;
							else if (tx == 32 && wx + 1 < W)
								bk[wi + 1] = (uint32_t)(m >> 32);
						}
					}
				}
			}
		}
		if (have)
			__syncthreads();
	}
}

// block-wide min/max into range[0] (ord min) and range[1] (ord max)
__device__ __forceinline__ void reduce_range(float lmin, float lmax, uint32_t *range)
{
	__shared__ uint32_t s_red[2][SM_BS / N2M_WAVE];
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
		if (omin < ld_agent(&range[0]))
			atomicMin(&range[0], omin);
		if (omax > ld_agent(&range[1]))
			atomicMax(&range[1], omax);
	}
}

struct ImgSrc {
	const float *__restrict__ img;
	int nx;
	size_t nxy;
	__device__ float operator()(int x, int y, int z) const { return img[(size_t)z * nxy + (size_t)y * nx + x]; }
};

__global__ void __launch_bounds__(SM_BS)
k_smooth(const float *__restrict__ in, float *__restrict__ out, int nx, int ny, int nz,
         uint32_t *__restrict__ range /* [0]=ord min, [1]=ord max */, IsoMasks iso,
         uint32_t *__restrict__ bits, int W)
{
	const int x0 = blockIdx.x * SM_TX, y0 = blockIdx.y * SM_TY;
	const int z0 = blockIdx.z * SM_ZC, z1 = min(z0 + SM_ZC, nz);
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	ImgSrc src{in, nx, (size_t)nx * ny};
	smooth_tile<SM_TX, SM_TY>(src, out, nx, ny, nz, x0, y0, z0, z1, iso, bits, W, lmin, lmax);
	reduce_range(lmin, lmax, range);
}

// ---------------------------------------------------------------------------
// Batched atlas (src/nii2mesh.c:492-583).  Each label is processed on its
// own crop (label box grown by 4 voxels, clipped to the volume): beyond 2
// voxels from the label its binarised and smoothed volume is 0, so the crop
// reproduces the full-volume smoothing, range, darkening and bounding box.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int atlas_label_of(float v, int nlabel)
{
	float r = rintf(v); // the only label whose open interval (i-0.5, i+0.5) can hold v
	if (!(r >= 1.0f && r <= (float)nlabel))
		return 0;
	int i = (int)r;
	float lo = (float)((double)i - 0.5), hi = (float)((double)i + 0.5); // src/nii2mesh.c:555-556
	return (v > lo && v < hi) ? i : 0;
}

// per label: voxel count and box (stats[7*i] = count, min xyz, max xyz);
// PRIV: accumulate in LDS (7 * (nlabel + 1) words) and flush once per block.
// Each block reads one contiguous run of voxels, AS_UNROLL loads per thread
// in flight, so it meets few labels and flushes few.  Within a wave the
// voxels are consecutive along x: a run of lanes with one label in one row
// is merged by its first lane (count = run length, x from its two ends, y
// and z shared), one set of atomics per run.  I32: the volume has fewer
// than 2^32 voxels, so coordinates take 32-bit division.
constexpr int AS_UNROLL = 4;

template <bool PRIV, bool I32>
__device__ __forceinline__ void atlas_stats_voxel(int i, size_t v, int nx, int ny, uint32_t *__restrict__ s_st,
                                                  uint32_t *__restrict__ stats)
{
	if (!__ballot(i != 0))
		return;
	uint32_t x, y, z;
	if (I32) {
		const uint32_t w = (uint32_t)v, r = w / (uint32_t)nx;
		x = w - r * (uint32_t)nx;
		y = r % (uint32_t)ny;
		z = r / (uint32_t)ny;
	} else {
		x = (uint32_t)(v % nx);
		y = (uint32_t)((v / nx) % ny);
		z = (uint32_t)(v / ((size_t)nx * ny));
	}
	const int lane = lane_id();
	const int iprev = __shfl_up(i, 1);
	const bool head = lane == 0 || i != iprev || x == 0; // a run starts
	const unsigned long long heads = __ballot(head);
	if (!head || !i)
		return;
	const unsigned long long after = lane == N2M_WAVE - 1 ? 0ull : heads & (~0ull << (lane + 1));
	const uint32_t len = (uint32_t)((after ? __ffsll((long long)after) - 1 : N2M_WAVE) - lane);
	uint32_t *st = PRIV ? s_st + 7 * i : stats + 7 * i;
	atomicAdd(&st[0], len);
	atomicMin(&st[1], x);
	atomicMin(&st[2], y);
	atomicMin(&st[3], z);
	atomicMax(&st[4], x + len - 1);
	atomicMax(&st[5], y);
	atomicMax(&st[6], z);
}

template <bool PRIV, bool I32>
__global__ void __launch_bounds__(256)
k_atlas_stats(const float *__restrict__ labels, int nx, int ny, int nz, int nlabel, uint32_t *__restrict__ stats)
{
	extern __shared__ uint32_t s_st[];
	if (PRIV) {
		for (int i = threadIdx.x; i <= nlabel; i += blockDim.x) {
			s_st[7 * i] = 0;
			s_st[7 * i + 1] = s_st[7 * i + 2] = s_st[7 * i + 3] = 0xffffffffu;
			s_st[7 * i + 4] = s_st[7 * i + 5] = s_st[7 * i + 6] = 0u;
		}
		__syncthreads();
	}
	const size_t n = (size_t)nx * ny * nz;
	constexpr size_t STEP = 256 * AS_UNROLL;
	const size_t per = ((n + gridDim.x - 1) / gridDim.x + STEP - 1) / STEP * STEP;
	const size_t v0 = blockIdx.x * per, v1 = v0 + per < n ? v0 + per : n;
	for (size_t vb = v0; vb < v1; vb += STEP) { // uniform trip count: whole waves take part
		float val[AS_UNROLL];
#pragma unroll
		for (int u = 0; u < AS_UNROLL; u++) {
			const size_t v = vb + u * 256 + threadIdx.x;
			val[u] = v < v1 ? labels[v] : 0.0f;
		}
#pragma unroll
		for (int u = 0; u < AS_UNROLL; u++)
			atlas_stats_voxel<PRIV, I32>(atlas_label_of(val[u], nlabel), vb + u * 256 + threadIdx.x, nx, ny, s_st,
			                             stats);
	}
	if (!PRIV)
		return;
	__syncthreads();
	for (int i = threadIdx.x; i <= nlabel; i += blockDim.x) {
		if (!s_st[7 * i])
			continue;
		atomicAdd(&stats[7 * i], s_st[7 * i]);
		for (int k = 1; k <= 3; k++)
			atomicMin(&stats[7 * i + k], s_st[7 * i + k]);
		for (int k = 4; k <= 6; k++)
			atomicMax(&stats[7 * i + k], s_st[7 * i + k]);
	}
}

void launch_atlas_stats(const float *labels, int nx, int ny, int nz, int nlabel, uint32_t *stats, hipStream_t s)
{
	size_t n = (size_t)nx * ny * nz;
	size_t b = (n + 1023) / 1024;
	b = b > 2048 ? 2048 : (b < 1 ? 1 : b);
	const size_t lds = (size_t)7 * (nlabel + 1) * 4;
	const bool i32 = n <= 0xffffffffull;
	if (nlabel <= 1023 && i32)
		hipLaunchKernelGGL((k_atlas_stats<true, true>), dim3((unsigned)b), dim3(256), lds, s, labels, nx, ny, nz, nlabel,
		                   stats);
	else if (nlabel <= 1023)
		hipLaunchKernelGGL((k_atlas_stats<true, false>), dim3((unsigned)b), dim3(256), lds, s, labels, nx, ny, nz,
		                   nlabel, stats);
	else if (i32)
		hipLaunchKernelGGL((k_atlas_stats<false, true>), dim3((unsigned)b), dim3(256), 0, s, labels, nx, ny, nz, nlabel,
		                   stats);
	else
		hipLaunchKernelGGL((k_atlas_stats<false, false>), dim3((unsigned)b), dim3(256), 0, s, labels, nx, ny, nz,
		                   nlabel, stats);
}

// the binarised label (src/nii2mesh.c:557-563) read inside a task's crop
struct LabelSrc {
	const float *__restrict__ labels;
	int gnx;
	size_t gnxy;
	int ox, oy, oz;
	float lo, hi;
	__device__ float operator()(int x, int y, int z) const
	{
		float v = labels[(size_t)(oz + z) * gnxy + (size_t)(oy + y) * gnx + (ox + x)];
		return (v > lo && v < hi) ? 1.0f : 0.0f;
	}
};

// Crop tile shapes the atlas autotuner chooses from (x, y, z-chunk).
#define N2M_ATLAS_TILES(X) X(0, 64, 8, 32) X(1, 32, 8, 16) X(2, 16, 16, 16) X(3, 32, 8, 32)

// smooth (or just binarise) every task's crop; ranges[2*t] = ord min/max.
// LOOP: persistent, the grid walks the `nwork` items of the block map in
// strides of its own size; otherwise workgroup b takes item b.
template <int TX, int TY, int ZC>
__device__ __forceinline__ void atlas_smooth_item(const float *__restrict__ labels, int gnx, int gny,
                                                  const AtlasTask *__restrict__ tasks,
                                                  const uint32_t *__restrict__ blockTask, unsigned b, int smooth,
                                                  float *__restrict__ crops, uint32_t *__restrict__ ranges)
{
	const int ti = (int)blockTask[b];
	const AtlasTask t = tasks[ti];
	const int local = (int)b - t.first_block;
	LabelSrc src{labels, gnx, (size_t)gnx * gny, t.ox, t.oy, t.oz, t.lo, t.hi};
	float *out = crops + t.off;
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	if (smooth) {
		const int tx = (t.nx + TX - 1) / TX, ty = (t.ny + TY - 1) / TY;
		const int bx = local % tx, by = (local / tx) % ty, bz = local / (tx * ty);
		const int z0 = bz * ZC, z1 = min(z0 + ZC, t.nz);
		smooth_tile<TX, TY>(src, out, t.nx, t.ny, t.nz, bx * TX, by * TY, z0, z1, IsoMasks{}, nullptr, 0, lmin, lmax);
	} else {
		const size_t n = (size_t)t.nx * t.ny * t.nz;
		const size_t v = (size_t)local * SM_BS + threadIdx.x;
		if (v < n) {
			int x = (int)(v % t.nx), y = (int)((v / t.nx) % t.ny), z = (int)(v / ((size_t)t.nx * t.ny));
			float o = src(x, y, z);
			out[v] = o;
			lmin = o;
			lmax = o;
		}
	}
	reduce_range(lmin, lmax, &ranges[2 * ti]);
}

template <int TX, int TY, int ZC, bool LOOP>
__global__ void __launch_bounds__(SM_BS)
k_atlas_smooth(const float *__restrict__ labels, int gnx, int gny, const AtlasTask *__restrict__ tasks,
               const uint32_t *__restrict__ blockTask, unsigned nwork, int smooth, float *__restrict__ crops,
               uint32_t *__restrict__ ranges)
{
	if (!LOOP) {
		atlas_smooth_item<TX, TY, ZC>(labels, gnx, gny, tasks, blockTask, blockIdx.x, smooth, crops, ranges);
		return;
	}
	for (unsigned b = blockIdx.x; b < nwork; b += gridDim.x) {
		atlas_smooth_item<TX, TY, ZC>(labels, gnx, gny, tasks, blockTask, b, smooth, crops, ranges);
		__syncthreads(); // the next item reuses the LDS
	}
}

int atlas_tile_count()
{
#define N2M_COUNT_TILE(i, tx, ty, zc) +1
	return 0 N2M_ATLAS_TILES(N2M_COUNT_TILE);
#undef N2M_COUNT_TILE
}

void atlas_tile_shape(int tile, int shape[3])
{
#define N2M_TILE_SHAPE(i, tx, ty, zc)                                                                              \
	if (tile == i) {                                                                                              \
		shape[0] = tx;                                                                                             \
		shape[1] = ty;                                                                                             \
		shape[2] = zc;                                                                                             \
		return;                                                                                                    \
	}
	N2M_ATLAS_TILES(N2M_TILE_SHAPE)
#undef N2M_TILE_SHAPE
	shape[0] = SM_TX;
	shape[1] = SM_TY;
	shape[2] = SM_ZC;
}

void launch_atlas_smooth(const float *labels, int gnx, int gny, int gnz, const AtlasTask *tasks,
                         const uint32_t *blockTask, unsigned nblocks, unsigned grid, int smooth, int tile,
                         float *crops, uint32_t *ranges, hipStream_t s)
{
	if (!nblocks)
		return;
#define N2M_LAUNCH_TILE(i, tx, ty, zc)                                                                             \
	if (tile == i) {                                                                                              \
		if (grid < nblocks)                                                                                        \
			hipLaunchKernelGGL((k_atlas_smooth<tx, ty, zc, true>), dim3(grid), dim3(SM_BS), 0, s, labels, gnx, gny, \
			                   tasks, blockTask, nblocks, smooth, crops, ranges);                                   \
		else                                                                                                       \
			hipLaunchKernelGGL((k_atlas_smooth<tx, ty, zc, false>), dim3(nblocks), dim3(SM_BS), 0, s, labels, gnx,  \
			                   gny, tasks, blockTask, nblocks, smooth, crops, ranges);                              \
		return;                                                                                                    \
	}
	N2M_ATLAS_TILES(N2M_LAUNCH_TILE)
#undef N2M_LAUNCH_TILE
}

// darken the voxels on the volume's faces, bounding box of >= iso and the
// minimum of the result, per task (src/meshify.c:347-371)
__device__ __forceinline__ void atlas_apply_item(float *__restrict__ crops, int gnx, int gny, int gnz,
                                                 const AtlasTask *__restrict__ tasks,
                                                 const uint32_t *__restrict__ blockTask, unsigned b,
                                                 int *__restrict__ bbox, uint32_t *__restrict__ vmin)
{
	__shared__ int s_red[6][4];
	__shared__ uint32_t s_min[4];
	const int ti = (int)blockTask[b];
	const AtlasTask t = tasks[ti];
	const int local = (int)b - t.first_block;
	const size_t n = (size_t)t.nx * t.ny * t.nz;
	float lmin = __builtin_inff();
	int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, bmax[3] = {-1, -1, -1};
	float *img = crops + t.off;
	for (int k = 0; k < 4; k++) {
		size_t v = ((size_t)local * 4 + k) * 256 + threadIdx.x;
		if (v >= n)
			break;
		const int x = (int)(v % t.nx), y = (int)((v / t.nx) % t.ny), z = (int)(v / ((size_t)t.nx * t.ny));
		float val = img[v];
		if (val >= t.iso) {
			bmin[0] = min(bmin[0], x);
			bmax[0] = max(bmax[0], x);
			bmin[1] = min(bmin[1], y);
			bmax[1] = max(bmax[1], y);
			bmin[2] = min(bmin[2], z);
			bmax[2] = max(bmax[2], z);
		}
		const int gx = t.ox + x, gy = t.oy + y, gz = t.oz + z;
		if (gx == 0 || gy == 0 || gz == 0 || gx == gnx - 1 || gy == gny - 1 || gz == gnz - 1) {
			val = fminf(t.edgeMax, val);
			img[v] = val;
		}
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
	int *bb = bbox + 6 * ti;
	if (threadIdx.x < 6) {
		int k = threadIdx.x;
		int r = s_red[k][0];
		for (int w = 1; w < 4; w++)
			r = k < 3 ? min(r, s_red[k][w]) : max(r, s_red[k][w]);
		const int cur = (int)ld_agent((const uint32_t *)&bb[k]);
		if (k < 3) {
			if (r != 0x7fffffff && r < cur)
				atomicMin(&bb[k], r);
		} else if (r >= 0 && r > cur) {
			atomicMax(&bb[k], r);
		}
	}
	if (threadIdx.x == 6) {
		uint32_t m = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
		if (m < ld_agent(&vmin[ti]))
			atomicMin(&vmin[ti], m);
	}
}

template <bool LOOP>
__global__ void __launch_bounds__(256)
k_atlas_apply(float *__restrict__ crops, int gnx, int gny, int gnz, const AtlasTask *__restrict__ tasks,
              const uint32_t *__restrict__ blockTask, unsigned nwork, int *__restrict__ bbox,
              uint32_t *__restrict__ vmin)
{
	if (!LOOP) {
		atlas_apply_item(crops, gnx, gny, gnz, tasks, blockTask, blockIdx.x, bbox, vmin);
		return;
	}
	for (unsigned b = blockIdx.x; b < nwork; b += gridDim.x) { // persistent, as k_atlas_smooth
		atlas_apply_item(crops, gnx, gny, gnz, tasks, blockTask, b, bbox, vmin);
		__syncthreads(); // the next item rewrites the LDS partials
	}
}

__global__ void k_atlas_iso(AtlasTask *__restrict__ tasks, int ntasks, const uint32_t *__restrict__ ranges, float req)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= ntasks)
		return;
	const float mn = ord2f(ranges[2 * t]), mx = ord2f(ranges[2 * t + 1]);
	float iso = __builtin_nanf(""), edge = __builtin_nanf("");
	bool fb;
	atlas_iso(req, mn, mx, &iso, &edge, &fb);
	tasks[t].iso = iso;
	tasks[t].mn = mn;
	tasks[t].edgeMax = edge;
}

void launch_atlas_iso(AtlasTask *tasks, int ntasks, const uint32_t *ranges, float req, hipStream_t s)
{
	if (ntasks)
		hipLaunchKernelGGL(k_atlas_iso, dim3((ntasks + 255) / 256), dim3(256), 0, s, tasks, ntasks, ranges, req);
}

void launch_atlas_apply(float *crops, int gnx, int gny, int gnz, const AtlasTask *tasks, const uint32_t *blockTask,
                        unsigned nblocks, unsigned grid, int *bbox, uint32_t *vmin, hipStream_t s)
{
	if (nblocks && grid < nblocks)
		hipLaunchKernelGGL(k_atlas_apply<true>, dim3(grid), dim3(256), 0, s, crops, gnx, gny, gnz, tasks, blockTask,
		                   nblocks, bbox, vmin);
	else if (nblocks)
		hipLaunchKernelGGL(k_atlas_apply<false>, dim3(nblocks), dim3(256), 0, s, crops, gnx, gny, gnz, tasks,
		                   blockTask, nblocks, bbox, vmin);
}

void launch_smooth(const float *in, float *out, int nx, int ny, int nz, uint32_t *range, const float *isos,
                   int nmask, uint32_t *bits, hipStream_t s)
{
	dim3 grid((nx + SM_TX - 1) / SM_TX, (ny + SM_TY - 1) / SM_TY, (nz + SM_ZC - 1) / SM_ZC);
	int W = (nx + 31) / 32;
	IsoMasks m{};
	m.n = bits ? (nmask < N2M_SMOOTH_MASKS ? nmask : N2M_SMOOTH_MASKS) : 0;
	for (int k = 0; k < m.n; k++)
		m.v[k] = isos[k];
	m.nwords = (size_t)W * ny * nz;
	hipLaunchKernelGGL(k_smooth, grid, dim3(SM_BS), 0, s, in, out, nx, ny, nz, range, m, bits, W);
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
		// same-address atomics serialise across the chip: only a block that
		// improves on what it reads issues one
		if (omin < ld_agent(&range[0]))
			atomicMin(&range[0], omin);
		if (omax > ld_agent(&range[1]))
			atomicMax(&range[1], omax);
	}
}

void launch_minmax(const float *img, size_t n, uint32_t *range, hipStream_t s)
{
	size_t blocks = (n / 4 + 255) / 256;
	blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
	int vec4 = ((uintptr_t)img & 15) == 0; // float4 loads need 16-byte alignment
	hipLaunchKernelGGL(k_minmax, dim3((unsigned)blocks), dim3(256), 0, s, img, n, vec4, range);
}

// ---------------------------------------------------------------------------
// Threshold mask bits: bit = (img >= iso) (or img != 0 when `nonzero`).
// A workgroup walks whole rows (the row index is uniform, so no per-voxel
// division); one wave covers 64 consecutive x of a row, padded to 64.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_mask_bits(const float *__restrict__ img, int nx, int ny, int nz, int W, int xpad, float iso,
            int nonzero, uint32_t *__restrict__ bits)
{
	const size_t rows = (size_t)ny * nz;
	for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
		const float *src = img + row * nx;
		uint32_t *dst = bits + (size_t)W * row;
		for (int x0 = 0; x0 < xpad; x0 += blockDim.x) {
			const int x = x0 + threadIdx.x;
			if (x - (int)lane_id() >= xpad)
				break; // whole waves past the padded row (xpad is a multiple of 64)
			bool p = false;
			if (x < nx) {
				const float v = src[x];
				p = nonzero ? (v != 0.0f) : (v >= iso);
			}
			const unsigned long long m = __ballot(p);
			const int wx = (x & ~63) / 32;
			if ((x & 63) == 0)
				dst[wx] = (uint32_t)m;
			else if ((x & 63) == 32 && wx + 1 < W)
				dst[wx + 1] = (uint32_t)(m >> 32);
		}
	}
}

// The same with 16-byte loads, for rows of a multiple of 32 voxels (nx = 32W)
// whose 16-byte chunks (lpr = nx/4 per row) either fit a workgroup (several
// rows per pass) or fill whole passes of it.  A lane tests 4 voxels; 8
// neighbouring lanes OR their nibbles into a word.
__global__ void __launch_bounds__(256)
k_mask_bits4(const float *__restrict__ img, int nx, int ny, int nz, int W, float iso, int nonzero,
             uint32_t *__restrict__ bits)
{
	const size_t rows = (size_t)ny * nz;
	const int lpr = nx >> 2, span = lpr < (int)blockDim.x ? lpr : (int)blockDim.x;
	const int rpi = (int)blockDim.x / span, sub = threadIdx.x / span, q0 = threadIdx.x % span;
	for (size_t row0 = (size_t)blockIdx.x * rpi; row0 < rows; row0 += (size_t)gridDim.x * rpi) {
		const size_t row = row0 + sub;
		const bool live = sub < rpi && row < rows;
		for (int q = q0; q < lpr; q += span) { // the same trip count on every lane
			uint32_t nib = 0;
			if (live) {
				const float4 v = reinterpret_cast<const float4 *>(img + row * nx)[q];
				if (nonzero)
					nib = (v.x != 0.0f) | (v.y != 0.0f) << 1 | (v.z != 0.0f) << 2 | (v.w != 0.0f) << 3;
				else
					nib = (v.x >= iso) | (v.y >= iso) << 1 | (v.z >= iso) << 2 | (v.w >= iso) << 3;
			}
			uint32_t wv = nib << (4 * (q & 7));
			wv |= __shfl_xor(wv, 1, N2M_WAVE);
			wv |= __shfl_xor(wv, 2, N2M_WAVE);
			wv |= __shfl_xor(wv, 4, N2M_WAVE);
			if (live && (q & 7) == 0)
				bits[(size_t)W * row + (q >> 3)] = wv;
		}
	}
}

// The masks of several isolevels in one read of the volume (same layout
// and conditions as k_mask_bits4): mask k to bits + k * iso.nwords.
__global__ void __launch_bounds__(256)
k_mask_bits4_multi(const float *__restrict__ img, int nx, int ny, int nz, int W, IsoMasks iso,
                   uint32_t *__restrict__ bits)
{
	const size_t rows = (size_t)ny * nz;
	const int lpr = nx >> 2, span = lpr < (int)blockDim.x ? lpr : (int)blockDim.x;
	const int rpi = (int)blockDim.x / span, sub = threadIdx.x / span, q0 = threadIdx.x % span;
	for (size_t row0 = (size_t)blockIdx.x * rpi; row0 < rows; row0 += (size_t)gridDim.x * rpi) {
		const size_t row = row0 + sub;
		const bool live = sub < rpi && row < rows;
		for (int q = q0; q < lpr; q += span) { // the same trip count on every lane
			float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
			if (live)
				v = reinterpret_cast<const float4 *>(img + row * nx)[q];
			for (int k = 0; k < iso.n; k++) {
				const float t = iso.v[k];
				uint32_t nib = 0;
				if (live)
					nib = (v.x >= t) | (v.y >= t) << 1 | (v.z >= t) << 2 | (v.w >= t) << 3;
				uint32_t wv = nib << (4 * (q & 7));
				wv |= __shfl_xor(wv, 1, N2M_WAVE);
				wv |= __shfl_xor(wv, 2, N2M_WAVE);
				wv |= __shfl_xor(wv, 4, N2M_WAVE);
				if (live && (q & 7) == 0)
					bits[k * iso.nwords + (size_t)W * row + (q >> 3)] = wv;
			}
		}
	}
}

static unsigned row_blocks(size_t rows) { return (unsigned)(rows < 8192 ? (rows ? rows : 1) : 8192); }

void launch_mask_bits_multi(const float *img, int nx, int ny, int nz, const float *isos, int n, uint32_t *bits,
                            hipStream_t s)
{
	const int W = (nx + 31) / 32, lpr = nx / 4;
	const size_t nwords = (size_t)W * ny * nz;
	if (n <= N2M_SMOOTH_MASKS && nx % 32 == 0 && (reinterpret_cast<uintptr_t>(img) & 15u) == 0 &&
	    (lpr <= 256 || lpr % 256 == 0)) {
		IsoMasks m{};
		m.n = n;
		for (int k = 0; k < n; k++)
			m.v[k] = isos[k];
		m.nwords = nwords;
		const size_t rows = (size_t)ny * nz, rpi = lpr < 256 ? 256 / lpr : 1;
		hipLaunchKernelGGL(k_mask_bits4_multi, dim3(row_blocks((rows + rpi - 1) / rpi)), dim3(256), 0, s, img, nx,
		                   ny, nz, W, m, bits);
		return;
	}
	for (int k = 0; k < n; k++)
		launch_mask_bits(img, nx, ny, nz, isos[k], 0, bits + k * nwords, s);
}

void launch_mask_bits(const float *img, int nx, int ny, int nz, float iso, int nonzero, uint32_t *bits,
                      hipStream_t s)
{
	int W = (nx + 31) / 32, xpad = (nx + 63) / 64 * 64;
	const int lpr = nx / 4;
	if (nx % 32 == 0 && (reinterpret_cast<uintptr_t>(img) & 15u) == 0 && (lpr <= 256 || lpr % 256 == 0)) {
		const size_t rows = (size_t)ny * nz, rpi = lpr < 256 ? 256 / lpr : 1;
		hipLaunchKernelGGL(k_mask_bits4, dim3(row_blocks((rows + rpi - 1) / rpi)), dim3(256), 0, s, img, nx, ny,
		                   nz, W, iso, nonzero, bits);
		return;
	}
	hipLaunchKernelGGL(k_mask_bits, dim3(row_blocks((size_t)ny * nz)), dim3(256), 0, s, img, nx, ny, nz, W, xpad,
	                   iso, nonzero, bits);
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
k_apply(const float *src, float *img, int nx, int ny, int nz, int W, int vec4, float iso, float mn,
        float edgeMax, const uint32_t *__restrict__ keep, const uint32_t *__restrict__ dil, int fill,
        int *__restrict__ bbox, uint32_t *__restrict__ vmin, uint32_t *__restrict__ sgp, uint32_t *__restrict__ sgn)
{
	__shared__ int s_red[6][4];
	__shared__ uint32_t s_min[4];
	float lmin = __builtin_inff();
	int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, bmax[3] = {-1, -1, -1};
	const size_t rows = (size_t)ny * nz;
	// one voxel: kw / dw are the keep / dilated-mask words holding bit x & 31
	auto voxel = [&](int x, int y, int z, bool face_row, float val, uint32_t kw, uint32_t dw, bool k_on,
	                 bool d_on) -> float {
		if (k_on && ((kw >> (x & 31)) & 1u))
			val = (float)fmax((double)val, (double)iso);
		if (d_on && !((dw >> (x & 31)) & 1u))
			val = mn;
		if (val >= iso) {
			bmin[0] = min(bmin[0], x);
			bmax[0] = max(bmax[0], x);
			bmin[1] = min(bmin[1], y);
			bmax[1] = max(bmax[1], y);
			bmin[2] = min(bmin[2], z);
			bmax[2] = max(bmax[2], z);
		}
		if (face_row || x == 0 || x == nx - 1)
			val = fminf(edgeMax, val);
		lmin = fminf(lmin, val);
		return val;
	};
	const bool k_on = fill && keep, d_on = dil != nullptr;
	if (vec4) {
		// 16-byte loads and stores; rows of nx = 4 * lpr voxels, several rows
		// per pass when a row has fewer 16-byte chunks than the workgroup
		const int lpr = nx >> 2, span = lpr < (int)blockDim.x ? lpr : (int)blockDim.x;
		const int rpi = (int)blockDim.x / span, sub = threadIdx.x / span, q0 = threadIdx.x % span;
		for (size_t row0 = (size_t)blockIdx.x * rpi; row0 < rows; row0 += (size_t)gridDim.x * rpi) {
			const size_t row = row0 + sub;
			if (sub >= rpi || row >= rows)
				continue;
			const int y = (int)(row % ny), z = (int)(row / ny);
			const bool face_row = y == 0 || z == 0 || y == ny - 1 || z == nz - 1;
			const float4 *in = reinterpret_cast<const float4 *>(src + row * nx);
			float4 *line = reinterpret_cast<float4 *>(img + row * nx);
			for (int q = q0; q < lpr; q += span) {
				const int x = 4 * q;
				const uint32_t kw = k_on ? keep[(size_t)W * row + (x >> 5)] : 0u;
				const uint32_t dw = d_on ? dil[(size_t)W * row + (x >> 5)] : 0u;
				float4 v = in[q];
				v.x = voxel(x, y, z, face_row, v.x, kw, dw, k_on, d_on);
				v.y = voxel(x + 1, y, z, face_row, v.y, kw, dw, k_on, d_on);
				v.z = voxel(x + 2, y, z, face_row, v.z, kw, dw, k_on, d_on);
				v.w = voxel(x + 3, y, z, face_row, v.w, kw, dw, k_on, d_on);
				line[q] = v;
				if (sgp) {
					// sign bits of the Lewiner data lw_clamp(v - iso): the 8
					// lanes of an aligned group hold one 32-voxel word of the
					// row (nx % 32 == 0 and the span a multiple of 8, so the
					// group shares its row and its trip count)
					const float vv[4] = {v.x, v.y, v.z, v.w};
					uint32_t bp = 0, bn = 0;
#pragma unroll
					for (int u = 0; u < 4; u++) {
						float d = vv[u] - iso;
						d = fabsf(d) < FLT_EPSILON ? FLT_EPSILON : d;
						bp |= (uint32_t)(d > 0.f) << u;
						bn |= (uint32_t)(d < 0.f) << u;
					}
					bp <<= 4 * (q & 7);
					bn <<= 4 * (q & 7);
#pragma unroll
					for (int m = 1; m < 8; m <<= 1) {
						bp |= __shfl_xor(bp, m);
						bn |= __shfl_xor(bn, m);
					}
					if ((q & 7) == 0) {
						sgp[(size_t)W * row + (x >> 5)] = bp;
						sgn[(size_t)W * row + (x >> 5)] = bn;
					}
				}
			}
		}
	} else {
		// whole rows per workgroup: y and z are uniform, no per-voxel division
		for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
			const int y = (int)(row % ny), z = (int)(row / ny);
			const bool face_row = y == 0 || z == 0 || y == ny - 1 || z == nz - 1;
			const float *in = src + row * nx;
			float *line = img + row * nx;
			for (int x = threadIdx.x; x < nx; x += blockDim.x) {
				const uint32_t kw = k_on ? keep[(size_t)W * row + (x >> 5)] : 0u;
				const uint32_t dw = d_on ? dil[(size_t)W * row + (x >> 5)] : 0u;
				line[x] = voxel(x, y, z, face_row, in[x], kw, dw, k_on, d_on);
			}
		}
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

bool apply_sign_bits_ok(const float *src, const float *img, int nx)
{
	const float *in = src ? src : img;
	return nx % 32 == 0 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(img)) & 15u) == 0;
}

bool launch_apply(const float *src, float *img, int nx, int ny, int nz, float iso, float mn, float edgeMax,
                  const uint32_t *keep, const uint32_t *dil, int fill, int *bbox, uint32_t *vmin, hipStream_t s,
                  uint32_t *sgp, uint32_t *sgn)
{
	const int W = (nx + 31) / 32;
	const float *in = src ? src : img;
	const bool vec4 = nx % 4 == 0 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(img)) & 15u) == 0;
	if (!sgp || !sgn || !apply_sign_bits_ok(src, img, nx))
		sgp = sgn = nullptr;
	const size_t rows = (size_t)ny * nz;
	const size_t rpi = vec4 && nx / 4 < 256 ? 256 / (nx / 4) : 1;
	hipLaunchKernelGGL(k_apply, dim3(row_blocks((rows + rpi - 1) / rpi)), dim3(256), 0, s, in, img, nx, ny, nz, W,
	                   vec4 ? 1 : 0, iso, mn, edgeMax, keep, dil, fill, bbox, vmin, sgp, sgn);
	return sgp != nullptr;
}

// ---------------------------------------------------------------------------
// Histograms for setThreshold() (src/isolevel.c).
// stats: [0]=ord min, [1]=ord max over non-NaN, [2]=NaN count.
// ---------------------------------------------------------------------------
// f(v) for every voxel of img[0, n), grid-stride; 16-byte loads when img is
// 16-byte aligned (the visiting order does not matter to the callers: min,
// max and counts)
template <typename F> __device__ __forceinline__ void each_voxel(const float *__restrict__ img, size_t n, F f)
{
	const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
	size_t done = 0;
	if ((reinterpret_cast<uintptr_t>(img) & 15u) == 0) {
		const float4 *v4 = reinterpret_cast<const float4 *>(img);
		const size_t n4 = n / 4;
#pragma unroll 2
		for (size_t i = tid; i < n4; i += stride) {
			const float4 q = v4[i];
			f(q.x);
			f(q.y);
			f(q.z);
			f(q.w);
		}
		done = 4 * n4;
	}
	for (size_t i = done + tid; i < n; i += stride)
		f(img[i]);
}

__global__ void __launch_bounds__(256) k_hist_range(const float *__restrict__ img, size_t n, uint32_t *stats)
{
	__shared__ uint32_t s_red[3][4];
	float lmin = __builtin_inff(), lmax = -__builtin_inff();
	uint32_t nan = 0;
	each_voxel(img, n, [&](float v) {
		if (v != v) {
			nan++;
			return;
		}
		lmin = fminf(lmin, v);
		lmax = fmaxf(lmax, v);
	});
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
__device__ __forceinline__ void hist_bins_body(const float *__restrict__ img, size_t n, float mn, float scl, int nbins,
                                               uint32_t *hist)
{
	extern __shared__ uint32_t s_h[];
	for (int i = threadIdx.x; i < nbins; i += blockDim.x)
		s_h[i] = 0;
	__syncthreads();
	// a run of voxels in the same bin (most of a background) is added with
	// one LDS atomic
	int cb = 0;
	uint32_t cc = 0;
	each_voxel(img, n, [&](float v) {
		if (v != v)
			return;
		float f = (v - mn) * scl;
		int b = (int)roundf(f);
		b = b < 0 ? 0 : (b > nbins - 1 ? nbins - 1 : b);
		if (b == cb) {
			cc++;
			return;
		}
		if (cc)
			atomicAdd(&s_h[cb], cc);
		cb = b;
		cc = 1;
	});
	if (cc)
		atomicAdd(&s_h[cb], cc);
	__syncthreads();
	for (int i = threadIdx.x; i < nbins; i += blockDim.x)
		if (s_h[i])
			atomicAdd(&hist[i], s_h[i]);
}

__global__ void __launch_bounds__(256)
k_hist_bins(const float *__restrict__ img, size_t n, float mn, float scl, int nbins, uint32_t *hist)
{
	hist_bins_body(img, n, mn, scl, nbins, hist);
}

// the same with mn = par[0], scl = par[1] from the device; nothing when
// *enable is 0
__global__ void __launch_bounds__(256)
k_hist_bins_dev(const float *__restrict__ img, size_t n, const float *__restrict__ par,
                const int32_t *__restrict__ enable, int nbins, uint32_t *hist)
{
	if (enable && *enable == 0)
		return;
	hist_bins_body(img, n, par[0], par[1], nbins, hist);
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

unsigned atlas_smooth_blocks(int nx, int ny, int nz, int smooth, int tile)
{
	if (smooth) {
		int sh[3];
		atlas_tile_shape(tile, sh);
		return (unsigned)(((nx + sh[0] - 1) / sh[0]) * ((ny + sh[1] - 1) / sh[1]) * ((nz + sh[2] - 1) / sh[2]));
	}
	return (unsigned)(((size_t)nx * ny * nz + SM_BS - 1) / SM_BS);
}
unsigned atlas_apply_blocks(int nx, int ny, int nz) { return (unsigned)(((size_t)nx * ny * nz + 1023) / 1024); }

// one label of the per-label loop: out = (lo < v < hi), count of ones
// (src/nii2mesh.c:553-563)
__global__ void __launch_bounds__(256)
k_binarize(const float *__restrict__ labels, size_t n, float lo, float hi, float *__restrict__ out,
           uint32_t *__restrict__ count)
{
	uint32_t c = 0;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		float v = labels[i];
		bool in = v > lo && v < hi;
		out[i] = in ? 1.0f : 0.0f;
		c += in;
	}
	__shared__ uint32_t s_c[4];
	c = wave_sum(c);
	if (lane_id() == 0)
		s_c[wave_id()] = c;
	__syncthreads();
	if (threadIdx.x == 0) {
		c = s_c[0] + s_c[1] + s_c[2] + s_c[3];
		if (c)
			atomicAdd(count, c);
	}
}

void launch_binarize(const float *labels, size_t n, float lo, float hi, float *out, uint32_t *count, hipStream_t s)
{
	size_t b = (n + 255) / 256;
	b = b > 4096 ? 4096 : (b < 1 ? 1 : b);
	hipLaunchKernelGGL(k_binarize, dim3((unsigned)b), dim3(256), 0, s, labels, n, lo, hi, out, count);
}

// ---------------------------------------------------------------------------
// Three-threshold Otsu search of nii_otsu() mode 5 (src/isolevel.c:176-197)
// over the Liao table P (nBin x nBin doubles, built on the host exactly as
// the reference builds it).  Workgroup l scans every (m, h) for that l;
// each lane keeps the first strict maximum in h order, the workgroup the
// first in m order, and the host the first in l order -- the reference's
// first occurrence of the maximum in (l, m, h) order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_otsu3(const double *__restrict__ P, int nBin, double *__restrict__ bv, int *__restrict__ bm, int *__restrict__ bh)
{
	__shared__ double s_v[256];
	__shared__ int s_h[256];
	const int l = blockIdx.x, m = l + 1 + threadIdx.x;
	double best = 0.0;
	int bestH = -1;
	if (m < nBin - 2) {
		const double a = P[l] + P[(size_t)(l + 1) * nBin + m];
		const double *row = P + (size_t)(m + 1) * nBin;
		for (int h = m + 1; h < nBin - 1; h++) {
			const double v = a + row[h] + P[(size_t)(h + 1) * nBin + 255];
			if (bestH < 0 || v > best) {
				best = v;
				bestH = h;
			}
		}
	}
	s_v[threadIdx.x] = best;
	s_h[threadIdx.x] = bestH;
	__syncthreads();
	if (threadIdx.x == 0) {
		double b = 0.0;
		int bt = -1;
		for (int t = 0; t < 256; t++) {
			if (s_h[t] < 0)
				continue;
			if (bt < 0 || s_v[t] > b) {
				b = s_v[t];
				bt = t;
			}
		}
		bv[l] = b;
		bm[l] = bt < 0 ? -1 : l + 1 + bt;
		bh[l] = bt < 0 ? -1 : s_h[bt];
	}
}

void launch_otsu3(const double *P, int nBin, double *bv, int *bm, int *bh, hipStream_t s)
{
	if (nBin > 3)
		hipLaunchKernelGGL(k_otsu3, dim3(nBin - 3), dim3(256), 0, s, P, nBin, bv, bm, bh);
}

// The Liao table of nii_otsu (src/isolevel.c:147-175) for a 256-bin
// histogram, built on the device, one workgroup per row u.  Each workgroup
// forms the running sums in the reference's order (the bin probabilities in
// parallel, the sums by one thread; P[0] is the raw count H[0], a quirk the
// thresholds depend on), then thread v writes P[u][v] = (S*S)/P of the bin
// range [u, v] (left as is when P is 0); entries with v < u are never read
// and are set to 0.
__global__ void __launch_bounds__(256) k_otsu_table(const uint32_t *__restrict__ H, double *__restrict__ P)
{
	constexpr int n = 256;
	__shared__ double s_p[n], s_s[n], s_prob[n];
	__shared__ double s_sum;
	if (threadIdx.x == 0) {
		double Sum = 0.0;
		for (int v = 0; v < n; v++)
			Sum = Sum + (double)(int)H[v];
		s_sum = Sum;
	}
	__syncthreads();
	s_prob[threadIdx.x] = (double)(int)H[threadIdx.x] / s_sum;
	__syncthreads();
	if (threadIdx.x == 0) {
		s_p[0] = (double)(int)H[0];
		s_s[0] = (double)(int)H[0];
		for (int v = 1; v < n; v++) {
			const double prob = s_prob[v];
			s_p[v] = s_p[v - 1] + prob;
			s_s[v] = s_s[v - 1] + (v + 1) * prob;
		}
	}
	__syncthreads();
	const int u = blockIdx.x, v = threadIdx.x;
	double val = 0.0;
	if (v >= u) {
		const double p = u == 0 ? s_p[v] : s_p[v] - s_p[u - 1];
		const double q = u == 0 ? s_s[v] : s_s[v] - s_s[u - 1];
		val = p != 0 ? (q * q) / p : p;
	}
	P[(size_t)u * n + v] = val;
}

// the table entries the host-side searches read: row 0 (P[i]), column 255
// shifted by one row (P[i+1][255]) and the four terms of the mode-5 start
// value (lo, mi, hi = 64, 128, 192)
__global__ void __launch_bounds__(256) k_otsu_probe(const double *__restrict__ P, double *__restrict__ probe)
{
	constexpr int n = 256;
	const int i = threadIdx.x;
	probe[i] = P[i];
	probe[n + i] = i < n - 1 ? P[(size_t)(i + 1) * n + (n - 1)] : 0.0;
	if (i == 0) {
		probe[2 * n + 0] = P[64];
		probe[2 * n + 1] = P[(size_t)65 * n + 128];
		probe[2 * n + 2] = P[(size_t)129 * n + 192];
		probe[2 * n + 3] = P[(size_t)193 * n + 255];
	}
}

// ---------------------------------------------------------------------------
// setThreshold() as one chain of launches (src/isolevel.c:35-277), results
// in a ThrRecord.
// ---------------------------------------------------------------------------
// range statistics {ordered min, ordered max, NaN count}, the parameters and
// both histograms cleared in one launch (no memset nodes)
__global__ void __launch_bounds__(256) k_thr_init(ThrRecord *__restrict__ r, uint32_t *__restrict__ h1001)
{
	for (int i = threadIdx.x; i < 1024; i += blockDim.x)
		h1001[i] = 0u;
	r->h256[threadIdx.x] = 0u;
	if (threadIdx.x < 4) {
		r->stats[threadIdx.x] = threadIdx.x == 0 ? 0xffffffffu : 0u;
		r->par[threadIdx.x] = 0.0f;
	}
	if (threadIdx.x == 0)
		r->do1001 = 0;
}

// One thread: the cases of the robust range (src/isolevel.c:59-139).  First
// call (final = 0): whether the 1001-bin pass runs, with its mn and scl.
// Second call: the percentiles (the walk over that histogram when it ran)
// and the 256-bin pass's mn and scl (src/isolevel.c:245-250).  All float
// and double arithmetic in the host code's order.
__global__ void k_thr_range(ThrRecord *__restrict__ r, const uint32_t *__restrict__ h1001, size_t nvox, int final)
{
	const size_t nNan = r->stats[2];
	const float mn = ord2f(r->stats[0]), mx = ord2f(r->stats[1]);
	float pct2 = 0.0f, pct98 = 1.0f;
	if (nvox == nNan) {
		// all NaN: the range keeps its 0..1 defaults
	} else if (mn == mx) {
		pct2 = mn;
		pct98 = mx;
	} else {
		const size_t valid = nvox - nNan;
		const size_t n2pct = (size_t)round(valid * 0.02);
		if (n2pct < 1 || valid < 100) {
			pct2 = mn;
			pct98 = mx;
		} else {
			const int kBins = 1001;
			const float scl = (kBins - 1) / (mx - mn);
			if (!final) {
				r->do1001 = 1;
				r->par[0] = mn;
				r->par[1] = scl;
				return;
			}
			int lo, hi;
			robust_walk(h1001, kBins, n2pct, &lo, &hi);
			pct2 = (lo) / scl + mn;
			pct98 = (hi) / scl + mn;
		}
	}
	if (final) {
		const int nBin = 256;
		r->par[2] = pct2;
		r->par[3] = (nBin - 1) / (pct98 - pct2);
	}
}

void launch_threshold_chain(const float *img, size_t nvox, ThrRecord *rec, uint32_t *h1001, double *P, bool need5,
                            hipStream_t s)
{
	const unsigned hb = hist_blocks(nvox);
	hipLaunchKernelGGL(k_thr_init, dim3(1), dim3(256), 0, s, rec, h1001);
	hipLaunchKernelGGL(k_hist_range, dim3(hb), dim3(256), 0, s, img, nvox, rec->stats);
	hipLaunchKernelGGL(k_thr_range, dim3(1), dim3(1), 0, s, rec, (const uint32_t *)h1001, nvox, 0);
	hipLaunchKernelGGL(k_hist_bins_dev, dim3(hb), dim3(256), 1001 * sizeof(uint32_t), s, img, nvox,
	                   (const float *)rec->par, (const int32_t *)&rec->do1001, 1001, h1001);
	hipLaunchKernelGGL(k_thr_range, dim3(1), dim3(1), 0, s, rec, (const uint32_t *)h1001, nvox, 1);
	hipLaunchKernelGGL(k_hist_bins_dev, dim3(hb), dim3(256), 256 * sizeof(uint32_t), s, img, nvox,
	                   (const float *)(rec->par + 2), (const int32_t *)nullptr, 256, rec->h256);
	hipLaunchKernelGGL(k_otsu_table, dim3(256), dim3(256), 0, s, (const uint32_t *)rec->h256, P);
	hipLaunchKernelGGL(k_otsu_probe, dim3(1), dim3(256), 0, s, (const double *)P, rec->probe);
	if (need5)
		launch_otsu3(P, 256, rec->bv, rec->bm, rec->bh, s);
}


// ---------------------------------------------------------------------------
// Raw NIfTI voxels to the FP32 working volume, src/nii2mesh.c:155-172:
// out = (raw * slope) + inter in float (two roundings), for uint8, int16,
// uint16 and float32 data.  Four voxels per thread.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_decode(const T *__restrict__ raw, size_t n, float slope, float inter, float *__restrict__ out)
{
	const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
	for (size_t i = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) * 4; i < n; i += stride) {
#pragma unroll
		for (int k = 0; k < 4; k++) {
			if (i + k < n) {
				const float v = (float)raw[i + k] * slope;
				out[i + k] = v + inter;
			}
		}
	}
}

int launch_decode(const void *raw, int datatype, size_t n, float slope, float inter, float *out, hipStream_t s)
{
	size_t b = (n + 1023) / 1024;
	b = b > 8192 ? 8192 : (b < 1 ? 1 : b);
	switch (datatype) {
	case 2: // DT_UINT8
		hipLaunchKernelGGL(k_decode<uint8_t>, dim3((unsigned)b), dim3(256), 0, s, (const uint8_t *)raw, n, slope,
		                   inter, out);
		return 0;
	case 4: // DT_INT16
		hipLaunchKernelGGL(k_decode<int16_t>, dim3((unsigned)b), dim3(256), 0, s, (const int16_t *)raw, n, slope,
		                   inter, out);
		return 0;
	case 512: // DT_UINT16
		hipLaunchKernelGGL(k_decode<uint16_t>, dim3((unsigned)b), dim3(256), 0, s, (const uint16_t *)raw, n, slope,
		                   inter, out);
		return 0;
	case 16: // DT_FLOAT32
		hipLaunchKernelGGL(k_decode<float>, dim3((unsigned)b), dim3(256), 0, s, (const float *)raw, n, slope, inter,
		                   out);
		return 0;
	}
	return -1;
}

} // namespace n2m
