// k_ccl.hip -- connected components for bwlabel() (src/bwlabel.c:429-543).
//
// The reference labels with a two-pass Thurfjell scan and numbers the final
// labels by first occurrence in x-fastest order (translate_labels,
// src/bwlabel.c:249-277).  Here: union-find whose root is always the
// smallest voxel index of its component, so "first occurrence" order is root
// order and the largest-cluster tie rule (lowest label wins,
// src/bwlabel.c:463) is "smallest root wins".
//
//  1. k_ccl_local: a workgroup labels a 32x8x8 tile in LDS (one bit word per
//     tile row) and writes each voxel's tile-local root to L; it also clears
//     AUX at every tile-local root (every global root is one of them).
//  2. k_ccl_merge: tile-boundary voxels union with their backward neighbours
//     in other tiles through global atomics.
//  3. k_ccl_flatten: L[v] = root(v); per-root voxel counts (wave-aggregated
//     atomics) or "touches a face" flags, root list / root count.
//  4. selection kernels turn roots into bit masks.
//
// Predicate: bit of `bits` (packed rows, see k_volume.hip), or its complement
// inside the volume when `invert` (background labelling for bubble fill).
#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

constexpr int CC_TX = 32, CC_TY = 8, CC_TZ = 8, CC_BS = 256;
constexpr int CC_N = CC_TX * CC_TY * CC_TZ;
constexpr uint32_t NONE = 0xffffffffu;

// backward neighbour offsets (dx,dy,dz): 6-, 18- and 26-connectivity, the
// voxels earlier in raster order that src/bwlabel.c:94-131,185-210 inspect.
__constant__ int8_t c_off[13][3] = {
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},                                          // 6
    {-1, -1, 0}, {1, -1, 0}, {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1},   // 18
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}};                         // 26

__host__ __device__ inline int n_offsets(int conn) { return conn == 6 ? 3 : (conn == 18 ? 9 : 13); }

struct Geo {
	int nx, ny, nz, W;
	__device__ uint32_t word(const uint32_t *bits, int wx, int y, int z, int invert) const
	{
		uint32_t w = bits[(size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * z)];
		if (invert) {
			int rem = nx - wx * 32;
			uint32_t valid = rem >= 32 ? 0xffffffffu : ((1u << rem) - 1u);
			w = ~w & valid;
		}
		return w;
	}
	__device__ bool fg(const uint32_t *bits, int x, int y, int z, int invert) const
	{
		if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz)
			return false;
		return (word(bits, x >> 5, y, z, invert) >> (x & 31)) & 1u;
	}
	__device__ uint32_t gidx(int x, int y, int z) const
	{
		return (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
	}
};

__device__ __forceinline__ uint32_t lds_find(const uint32_t *lab, uint32_t x)
{
	uint32_t p = lab[x];
	while (p != x) {
		x = p;
		p = lab[x];
	}
	return x;
}

__device__ __forceinline__ void lds_union(uint32_t *lab, uint32_t a, uint32_t b)
{
	for (;;) {
		a = lds_find(lab, a);
		b = lds_find(lab, b);
		if (a == b)
			return;
		if (a < b) {
			uint32_t t = a;
			a = b;
			b = t;
		}
		uint32_t old = atomicMin(&lab[a], b);
		if (old == a)
			return;
		a = old;
	}
}

__global__ void __launch_bounds__(CC_BS)
k_ccl_local(const uint32_t *__restrict__ bits, Geo g, int conn, int invert, uint32_t *__restrict__ L,
            uint32_t *__restrict__ AUX)
{
	__shared__ uint32_t lab[CC_N];
	__shared__ uint32_t rowbits[CC_TY * CC_TZ];
	const int wx = blockIdx.x, ty0 = blockIdx.y * CC_TY, tz0 = blockIdx.z * CC_TZ;
	const int lx = threadIdx.x % CC_TX, ly = threadIdx.x / CC_TX; // ly 0..7
	if (threadIdx.x < CC_TY * CC_TZ) {
		int ry = threadIdx.x % CC_TY, rz = threadIdx.x / CC_TY;
		int y = ty0 + ry, z = tz0 + rz;
		rowbits[threadIdx.x] = (y < g.ny && z < g.nz) ? g.word(bits, wx, y, z, invert) : 0u;
	}
	__syncthreads();
	for (int lz = 0; lz < CC_TZ; lz++) {
		uint32_t li = lx + CC_TX * (ly + CC_TY * lz);
		lab[li] = ((rowbits[ly + CC_TY * lz] >> lx) & 1u) ? li : NONE;
	}
	__syncthreads();
	const int no = n_offsets(conn);
	for (int lz = 0; lz < CC_TZ; lz++) {
		uint32_t li = lx + CC_TX * (ly + CC_TY * lz);
		if (lab[li] == NONE)
			continue;
		for (int k = 0; k < no; k++) {
			int nxl = lx + c_off[k][0], nyl = ly + c_off[k][1], nzl = lz + c_off[k][2];
			if (nxl < 0 || nxl >= CC_TX || nyl < 0 || nyl >= CC_TY || nzl < 0 || nzl >= CC_TZ)
				continue;
			if (!((rowbits[nyl + CC_TY * nzl] >> nxl) & 1u))
				continue;
			lds_union(lab, li, (uint32_t)(nxl + CC_TX * (nyl + CC_TY * nzl)));
		}
	}
	__syncthreads();
	const int x = wx * 32 + lx;
	for (int lz = 0; lz < CC_TZ; lz++) {
		uint32_t li = lx + CC_TX * (ly + CC_TY * lz);
		if (lab[li] == NONE)
			continue;
		uint32_t r = lds_find(lab, li);
		int rx = wx * 32 + (int)(r % CC_TX), ry = ty0 + (int)((r / CC_TX) % CC_TY), rz = tz0 + (int)(r / (CC_TX * CC_TY));
		uint32_t gi = g.gidx(x, ty0 + ly, tz0 + lz);
		uint32_t gr = g.gidx(rx, ry, rz);
		L[gi] = gr;
		if (gr == gi)
			AUX[gi] = 0;
	}
}

__global__ void __launch_bounds__(CC_BS)
k_ccl_merge(const uint32_t *__restrict__ bits, Geo g, int conn, int invert, uint32_t *__restrict__ L)
{
	const int wx = blockIdx.x, ty0 = blockIdx.y * CC_TY, tz0 = blockIdx.z * CC_TZ;
	const int lx = threadIdx.x % CC_TX, ly = threadIdx.x / CC_TX;
	const int x = wx * 32 + lx, y = ty0 + ly;
	const bool edge_xy = lx == 0 || lx == CC_TX - 1 || ly == 0 || ly == CC_TY - 1;
	const int no = n_offsets(conn);
	for (int lz = 0; lz < CC_TZ; lz++) {
		if (!edge_xy && lz != 0)
			continue;
		const int z = tz0 + lz;
		if (!g.fg(bits, x, y, z, invert))
			continue;
		uint32_t gi = g.gidx(x, y, z);
		for (int k = 0; k < no; k++) {
			int nxl = lx + c_off[k][0], nyl = ly + c_off[k][1], nzl = lz + c_off[k][2];
			if (nxl >= 0 && nxl < CC_TX && nyl >= 0 && nyl < CC_TY && nzl >= 0 && nzl < CC_TZ)
				continue; // same tile: done locally
			int xx = x + c_off[k][0], yy = y + c_off[k][1], zz = z + c_off[k][2];
			if (!g.fg(bits, xx, yy, zz, invert))
				continue;
			uf_union(L, gi, g.gidx(xx, yy, zz));
		}
	}
}

// mode 0: count voxels per root into AUX[root], append roots to rootlist.
// mode 1: AUX[root] = 1 for roots with a voxel on a volume face; count roots.
__global__ void __launch_bounds__(CC_BS)
k_ccl_flatten(const uint32_t *__restrict__ bits, Geo g, int invert, uint32_t *__restrict__ L,
              uint32_t *__restrict__ AUX, int mode, uint32_t *__restrict__ nroots,
              uint32_t *__restrict__ rootlist)
{
	const int wx = blockIdx.x, ty0 = blockIdx.y * CC_TY, tz0 = blockIdx.z * CC_TZ;
	const int lx = threadIdx.x % CC_TX, ly = threadIdx.x / CC_TX;
	const int x = wx * 32 + lx, y = ty0 + ly;
	for (int lz = 0; lz < CC_TZ; lz++) {
		const int z = tz0 + lz;
		const bool on = g.fg(bits, x, y, z, invert);
		uint32_t gi = 0, r = NONE;
		if (on) {
			gi = g.gidx(x, y, z);
			r = uf_find(L, gi);
			L[gi] = r;
		}
		bool isroot = on && r == gi;
		unsigned long long rm = __ballot(isroot);
		if (mode == 0) {
			// wave-aggregated per-root counts
			unsigned long long act = __ballot(on);
			while (act) {
				int leader = __ffsll((long long)act) - 1;
				uint32_t lr = __shfl(r, leader, N2M_WAVE);
				unsigned long long same = __ballot(on && r == lr);
				if (lane_id() == leader)
					atomicAdd(&AUX[lr], (uint32_t)__popcll(same));
				act &= ~same;
			}
			if (rm) {
				uint32_t base = 0;
				int leader = __ffsll((long long)rm) - 1;
				if (lane_id() == leader)
					base = atomicAdd(nroots, (uint32_t)__popcll(rm));
				base = __shfl(base, leader, N2M_WAVE);
				if (isroot)
					rootlist[base + __popcll(rm & ((1ull << lane_id()) - 1ull))] = gi;
			}
		} else {
			if (on && (x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1))
				AUX[r] = 1;
			if (rm && lane_id() == __ffsll((long long)rm) - 1)
				atomicAdd(nroots, (uint32_t)__popcll(rm));
		}
	}
}

static dim3 cc_grid(const Geo &g)
{
	return dim3(g.W, (g.ny + CC_TY - 1) / CC_TY, (g.nz + CC_TZ - 1) / CC_TZ);
}

void launch_ccl(const uint32_t *bits, int nx, int ny, int nz, int conn, int invert, uint32_t *L,
                uint32_t *AUX, int mode, uint32_t *nroots, uint32_t *rootlist, hipStream_t s,
                CclTimer *timer)
{
	Geo g{nx, ny, nz, (nx + 31) / 32};
	dim3 grid = cc_grid(g);
	const double nv = (double)nx * ny * nz;
	// algorithmic bytes: mask bits in, labels out (local); mask bits
	// (merge); labels in and out (flatten)
	if (timer) timer->begin("ccl_local", nv / 8 + 4 * nv);
	hipLaunchKernelGGL(k_ccl_local, grid, dim3(CC_BS), 0, s, bits, g, conn, invert, L, AUX);
	if (timer) timer->end();
	if (timer) timer->begin("ccl_merge", nv / 8);
	hipLaunchKernelGGL(k_ccl_merge, grid, dim3(CC_BS), 0, s, bits, g, conn, invert, L);
	if (timer) timer->end();
	if (timer) timer->begin("ccl_flatten", 8 * nv);
	hipLaunchKernelGGL(k_ccl_flatten, grid, dim3(CC_BS), 0, s, bits, g, invert, L, AUX, mode, nroots,
	                   rootlist);
	if (timer) timer->end();
}

// largest root: max count, ties -> smallest root.  best = (count<<32)|~root
__global__ void k_ccl_argmax(const uint32_t *__restrict__ rootlist, const uint32_t *__restrict__ nroots,
                             const uint32_t *__restrict__ AUX, unsigned long long *best)
{
	uint32_t n = *nroots;
	unsigned long long b = 0;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		uint32_t r = rootlist[i];
		unsigned long long k = ((unsigned long long)AUX[r] << 32) | (unsigned long long)(0xffffffffu - r);
		b = b > k ? b : k;
	}
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1) {
		unsigned long long o = __shfl_xor(b, d, N2M_WAVE);
		b = b > o ? b : o;
	}
	if (lane_id() == 0 && b)
		atomicMax(best, b);
}

void launch_ccl_argmax(const uint32_t *rootlist, const uint32_t *nroots, const uint32_t *AUX,
                       unsigned long long *best, hipStream_t s)
{
	hipLaunchKernelGGL(k_ccl_argmax, dim3(256), dim3(256), 0, s, rootlist, nroots, AUX, best);
}

// out bit = voxel in `bits` whose root is `best` (largest cluster mask).
__global__ void __launch_bounds__(256)
k_select_root(const uint32_t *__restrict__ bits, Geo g, int xpad, const uint32_t *__restrict__ L,
              const unsigned long long *__restrict__ bestp, uint32_t *__restrict__ out)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t rows = (size_t)g.ny * g.nz;
	size_t row = t / xpad;
	int x = (int)(t % xpad);
	if (row >= rows)
		return;
	unsigned long long bk = *bestp;
	uint32_t best = bk ? 0xffffffffu - (uint32_t)(bk & 0xffffffffu) : NONE;
	int y = (int)(row % g.ny), z = (int)(row / g.ny);
	bool p = x < g.nx && g.fg(bits, x, y, z, 0) && L[row * g.nx + x] == best;
	unsigned long long m = __ballot(p);
	int wx = (x & ~63) / 32;
	size_t wi = (size_t)wx + (size_t)g.W * row;
	if ((x & 63) == 0)
		out[wi] = (uint32_t)m;
	else if ((x & 63) == 32 && wx + 1 < g.W)
		out[wi + 1] = (uint32_t)(m >> 32);
}

void launch_select_root(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L,
                        const unsigned long long *best, uint32_t *out, hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32};
	int xpad = (nx + 63) / 64 * 64;
	size_t total = (size_t)xpad * ny * nz;
	hipLaunchKernelGGL(k_select_root, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bits, g, xpad,
	                   L, best, out);
}

// bubble fill (src/bwlabel.c:530-538): background voxels whose background
// component touches no face are set.  `bits` (foreground) is updated in place.
__global__ void __launch_bounds__(256)
k_fill_bubbles(uint32_t *__restrict__ bits, Geo g, int xpad, const uint32_t *__restrict__ L,
               const uint32_t *__restrict__ AUX)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t rows = (size_t)g.ny * g.nz;
	size_t row = t / xpad;
	int x = (int)(t % xpad);
	if (row >= rows)
		return;
	int y = (int)(row % g.ny), z = (int)(row / g.ny);
	int wx = (x & ~63) / 32;
	size_t wi = (size_t)wx + (size_t)g.W * row;
	bool set = false;
	if (x < g.nx && !g.fg(bits, x, y, z, 0))
		set = AUX[L[row * g.nx + x]] == 0;
	unsigned long long m = __ballot(set);
	if ((x & 63) == 0 && (uint32_t)m)
		bits[wi] |= (uint32_t)m;
	else if ((x & 63) == 32 && wx + 1 < g.W && (uint32_t)(m >> 32))
		bits[wi + 1] |= (uint32_t)(m >> 32);
}

void launch_fill_bubbles(uint32_t *bits, int nx, int ny, int nz, const uint32_t *L, const uint32_t *AUX,
                         hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32};
	int xpad = (nx + 63) / 64 * 64;
	size_t total = (size_t)xpad * ny * nz;
	hipLaunchKernelGGL(k_fill_bubbles, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bits, g,
	                   xpad, L, AUX);
}

// Full labelling output for bwlabel() without onlyLargest: label = rank of
// the root among all roots (1-based, ascending).  `rank` holds, at each root
// index, its rank (filled by the host from the sorted root list).
__global__ void __launch_bounds__(256)
k_write_labels(const uint32_t *__restrict__ bits, Geo g, const uint32_t *__restrict__ L,
               const uint32_t *__restrict__ rank, float *__restrict__ img)
{
	size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t n = (size_t)g.nx * g.ny * g.nz;
	if (v >= n)
		return;
	int x = (int)(v % g.nx);
	size_t row = v / g.nx;
	int y = (int)(row % g.ny), z = (int)(row / g.ny);
	img[v] = g.fg(bits, x, y, z, 0) ? (float)rank[L[v]] : 0.0f;
}

void launch_write_labels(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L,
                         const uint32_t *rank, float *img, hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32};
	size_t n = (size_t)nx * ny * nz;
	hipLaunchKernelGGL(k_write_labels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bits, g, L, rank,
	                   img);
}

// rank[root_sorted[i]] = i + 1
__global__ void k_scatter_rank(const uint32_t *__restrict__ sorted, uint32_t n, uint32_t *__restrict__ rank)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		rank[sorted[i]] = i + 1;
}

void launch_scatter_rank(const uint32_t *sorted, uint32_t n, uint32_t *rank, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_scatter_rank, dim3((n + 255) / 256), dim3(256), 0, s, sorted, n, rank);
}

// bits -> float image (0/1), for bwlabel(onlyLargest) output
__global__ void __launch_bounds__(256)
k_bits_to_float(const uint32_t *__restrict__ bits, int nx, int ny, int nz, int W, float *__restrict__ img)
{
	size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t n = (size_t)nx * ny * nz;
	if (v >= n)
		return;
	int x = (int)(v % nx);
	size_t row = v / nx;
	img[v] = ((bits[(size_t)(x >> 5) + (size_t)W * row] >> (x & 31)) & 1u) ? 1.0f : 0.0f;
}

void launch_bits_to_float(const uint32_t *bits, int nx, int ny, int nz, float *img, hipStream_t s)
{
	size_t n = (size_t)nx * ny * nz;
	hipLaunchKernelGGL(k_bits_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bits, nx, ny, nz,
	                   (nx + 31) / 32, img);
}

} // namespace n2m
