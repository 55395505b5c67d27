// k_mesh.hip -- surface stages on gfx950: scan, stable radix sort, vertex
// welding (src/meshify.c:45-106), degenerate-triangle removal
// (src/meshify.c:113-168), Laplacian / HC smoothing (src/quadric.c:316-394),
// border detection (src/quadric.c:188-215) and apply_sform
// (src/meshify.c:1021-1045).
#include <float.h>
#include <stdlib.h>

#include "n2m_device.cuh"
#include "n2m_internal.h"

namespace n2m {

// ---------------------------------------------------------------------------
// Exclusive scan of uint32: one workgroup for one tile, one pass with
// decoupled look-back up to SP_MAX_TILES tiles, else three levels.
// ---------------------------------------------------------------------------
constexpr int SC_BS = 256, SC_IPT = 8, SC_TILE = SC_BS * SC_IPT;

__global__ void __launch_bounds__(SC_BS) k_scan_reduce(const uint32_t *__restrict__ in, size_t n, uint32_t *sums)
{
	__shared__ uint32_t scratch[SC_BS / N2M_WAVE + 1];
	size_t base = blockIdx.x * (size_t)SC_TILE;
	uint32_t s = 0;
	for (int k = 0; k < SC_IPT; k++) {
		size_t i = base + k * SC_BS + threadIdx.x;
		if (i < n)
			s += in[i];
	}
	s = block_sum_u32<SC_BS>(s, scratch);
	if (threadIdx.x == 0)
		sums[blockIdx.x] = s;
}

__global__ void __launch_bounds__(SC_BS)
k_scan_apply(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, size_t n, const uint32_t *offs,
             uint32_t *total)
{
	__shared__ uint32_t scratch[SC_BS / N2M_WAVE + 1];
	size_t base = blockIdx.x * (size_t)SC_TILE + (size_t)threadIdx.x * SC_IPT;
	uint32_t v[SC_IPT], s = 0;
#pragma unroll
	for (int k = 0; k < SC_IPT; k++) {
		size_t i = base + k;
		v[k] = i < n ? in[i] : 0u;
		s += v[k];
	}
	uint32_t tot;
	uint32_t run = block_excl_scan<SC_BS>(s, scratch, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
	for (int k = 0; k < SC_IPT; k++) {
		size_t i = base + k;
		if (i < n)
			out[i] = run;
		run += v[k];
	}
	if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SC_BS - 1)
		*total = run;
}

// Scans of up to SP_MAX_TILES tiles run in one pass (k_scan_1p); longer ones
// reduce, scan the partial sums and apply (three launches: there the serial
// chain of look-backs costs more than the two extra launches).  The first
// SP_WORDS words of the scratch belong to the single-pass scan: two counters
// (tickets handed out, workgroups finished), then one 64-bit word per tile
// (status << 32 | sum).  They are zero between scans: the caller zeroes them
// once when the scratch is allocated, and the last workgroup of every
// single-pass scan clears what that scan used.  The partial sums of the
// multi-level form follow them.
constexpr int SP_MAX_TILES = 128, SP_WORDS = 320; // SP_WORDS >= 2 + 2 * SP_MAX_TILES
static_assert(SP_WORDS >= 2 + 2 * SP_MAX_TILES, "single-pass scan scratch");

size_t scan_scratch_words(size_t n)
{
	size_t w = 0;
	while (n > SC_TILE) {
		n = (n + SC_TILE - 1) / SC_TILE;
		w += n;
	}
	return SP_WORDS + w + 1;
}
size_t scan_scratch_zero_words() { return SP_WORDS; }

constexpr uint32_t SP_AGG = 1u, SP_INC = 2u; // tile status: own sum / inclusive prefix
constexpr uint32_t SP_SPIN_MAX = 1u << 20;   // bound on the polls of one predecessor

// Exclusive scan in one pass (decoupled look-back): each workgroup takes the
// next tile number from a counter, so every tile below it belongs to a
// workgroup that has already started; it publishes its tile's sum, adds the
// sums (or the inclusive prefix, where published) of the tiles before it,
// and publishes its inclusive prefix.  A predecessor's sum is published
// before that workgroup waits for anything, so every wait ends.
__global__ void __launch_bounds__(SC_BS)
k_scan_1p(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, uint32_t *scr, uint32_t ntiles)
{
	__shared__ uint32_t scratch[SC_BS / N2M_WAVE + 1];
	__shared__ uint32_t s_tile, s_prefix, s_last;
	uint32_t *ctr = scr;
	unsigned long long *flags = reinterpret_cast<unsigned long long *>(scr + 2);
	if (threadIdx.x == 0)
		s_tile = atomicAdd(&ctr[0], 1u);
	__syncthreads();
	const uint32_t tile = s_tile;
	const size_t base = tile * (size_t)SC_TILE + (size_t)threadIdx.x * SC_IPT;
	uint32_t v[SC_IPT], sum = 0;
#pragma unroll
	for (int k = 0; k < SC_IPT; k++) {
		const size_t i = base + k;
		v[k] = i < n ? in[i] : 0u;
		sum += v[k];
	}
	uint32_t agg;
	uint32_t run = block_excl_scan<SC_BS>(sum, scratch, &agg);
	// the look-back, by the first wave: lane l reads the word of tile
	// j0 - l, so one round trip covers 64 predecessors; it stops at the
	// nearest tile with an inclusive prefix (tile 0 always has one)
	if (threadIdx.x < N2M_WAVE) {
		const int lane = (int)threadIdx.x;
		uint32_t prefix = 0;
		if (tile > 0) {
			if (lane == 0)
				__hip_atomic_store(&flags[tile], ((unsigned long long)SP_AGG << 32) | agg, __ATOMIC_RELAXED,
				                   __HIP_MEMORY_SCOPE_AGENT);
			for (long long j0 = (long long)tile - 1;; j0 -= N2M_WAVE) {
				const long long j = j0 - lane;
				unsigned long long f;
				uint32_t spins = 0;
				bool all_ready;
				do {
					f = j >= 0 ? __hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
					           : ((unsigned long long)SP_INC << 32);
					all_ready = __all((f >> 32) != 0);
				} while (!all_ready && ++spins < SP_SPIN_MAX);
				const unsigned long long inc = __ballot((f >> 32) == SP_INC);
				const int stop = inc ? __ffsll((long long)inc) - 1 : N2M_WAVE - 1;
				prefix += wave_sum<uint32_t>(lane <= stop ? (uint32_t)f : 0u);
				if (inc || !all_ready)
					break;
			}
		}
		if (lane == 0) {
			__hip_atomic_store(&flags[tile], ((unsigned long long)SP_INC << 32) | (prefix + agg), __ATOMIC_RELAXED,
			                   __HIP_MEMORY_SCOPE_AGENT);
			s_prefix = prefix;
		}
	}
	__syncthreads();
	run += s_prefix;
#pragma unroll
	for (int k = 0; k < SC_IPT; k++) {
		const size_t i = base + k;
		if (i < n)
			out[i] = run;
		run += v[k];
	}
	if (total && tile == ntiles - 1 && threadIdx.x == SC_BS - 1)
		*total = run;
	// the last workgroup to finish clears the counters and the tile words:
	// every other one has finished its look-back by then
	if (threadIdx.x == 0)
		s_last = atomicAdd(&ctr[1], 1u) == ntiles - 1;
	__syncthreads();
	if (s_last) {
		for (uint32_t t = threadIdx.x; t < ntiles; t += SC_BS)
			flags[t] = 0ull;
		if (threadIdx.x == 0) {
			ctr[0] = 0u;
			ctr[1] = 0u;
		}
	}
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// In-place exclusive scan of up to SC_SMALL words by one workgroup: the
// partial sums level of a scan of up to SC_TILE * SC_SMALL words, in one
// launch instead of two or three.
constexpr int SC_SBS = 1024, SC_SIPT = 16, SC_SMALL = SC_SBS * SC_SIPT;

__global__ void __launch_bounds__(SC_SBS) k_scan_sums(uint32_t *__restrict__ a, uint32_t n)
{
	__shared__ uint32_t scratch[SC_SBS / N2M_WAVE + 1];
	const uint32_t base = threadIdx.x * SC_SIPT;
	uint32_t v[SC_SIPT], s = 0;
#pragma unroll
	for (int k = 0; k < SC_SIPT; k++) {
		v[k] = base + k < n ? a[base + k] : 0u;
		s += v[k];
	}
	uint32_t tot;
	uint32_t run = block_excl_scan<SC_SBS>(s, scratch, &tot);
#pragma unroll
	for (int k = 0; k < SC_SIPT; k++) {
		if (base + k < n)
			a[base + k] = run;
		run += v[k];
	}
}

// the multi-level form; sums: scan_scratch_words(n) - SP_WORDS words
static void scan_levels(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, uint32_t *sums, hipStream_t s)
{
	const size_t nb = (n + SC_TILE - 1) / SC_TILE;
	if (nb == 1) {
		hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SC_BS), 0, s, in, out, n, (const uint32_t *)nullptr, total);
		return;
	}
	hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SC_BS), 0, s, in, n, sums);
	if (nb <= (size_t)SC_SMALL)
		hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_SBS), 0, s, sums, (uint32_t)nb);
	else
		scan_levels(sums, sums, nb, nullptr, sums + nb, s);
	hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SC_BS), 0, s, in, out, n, (const uint32_t *)sums,
	                   total);
}

void scan_excl_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total, uint32_t *scratch,
                   hipStream_t s)
{
	if (n == 0) {
		if (total) // a kernel, not a memset: launch errors reach the caller's stream check
			hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, total);
		return;
	}
	const size_t nb = (n + SC_TILE - 1) / SC_TILE;
	if (nb > 1 && nb <= (size_t)SP_MAX_TILES)
		hipLaunchKernelGGL(k_scan_1p, dim3((unsigned)nb), dim3(SC_BS), 0, s, in, out, n, total, scratch,
		                   (uint32_t)nb);
	else
		scan_levels(in, out, n, total, scratch + SP_WORDS, s);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint32 key, uint32 value), 8-bit digits.
// Each workgroup ranks a tile of RS_TILE keys in RS_IPT rounds of RS_BS keys
// (tile order is preserved: round-major, thread order inside a round),
// places them in LDS in digit order, then writes each digit's keys as one
// contiguous run at its global offset (coalesced, instead of one scattered
// word per key).
// ---------------------------------------------------------------------------
constexpr int RS_BS = 256, RS_IPT = 16, RS_TILE = RS_BS * RS_IPT, RS_NW = RS_BS / N2M_WAVE;

// The tile a workgroup sorts: keys [start, start + len), digit d's count at
// counts[cbase + d * cstride].  Without descriptors tile b is keys
// [b * RS_TILE, ...) and the counts are digit-major over all tiles.
__device__ __forceinline__ RsTile rs_tile(const RsTile *__restrict__ desc, size_t n, size_t nblocks)
{
	if (desc)
		return desc[blockIdx.x];
	const size_t base = blockIdx.x * (size_t)RS_TILE;
	RsTile t;
	t.start = (uint32_t)base;
	t.len = (uint32_t)(n - base < (size_t)RS_TILE ? n - base : (size_t)RS_TILE);
	t.cbase = blockIdx.x;
	t.cstride = (uint32_t)nblocks;
	return t;
}

__global__ void __launch_bounds__(RS_BS)
k_radix_hist(const uint32_t *__restrict__ keys, size_t n, int shift, uint32_t *__restrict__ counts, size_t nblocks,
             const RsTile *__restrict__ desc)
{
	__shared__ uint32_t h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	const RsTile t = rs_tile(desc, n, nblocks);
	for (int k = 0; k < RS_IPT; k++) {
		const uint32_t i = k * RS_BS + threadIdx.x;
		if (i < t.len)
			atomicAdd(&h[(keys[t.start + i] >> shift) & 0xFF], 1u);
	}
	__syncthreads();
	counts[t.cbase + (size_t)threadIdx.x * t.cstride] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_BS)
k_radix_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
                uint32_t *__restrict__ vout, size_t n, int shift, const uint32_t *__restrict__ offs, size_t nblocks,
                const RsTile *__restrict__ desc)
{
	__shared__ uint32_t s_k[RS_TILE], s_v[RS_TILE];
	__shared__ uint32_t run[256];  // keys of each digit placed so far
	__shared__ uint32_t lofs[256]; // tile-local start of each digit
	__shared__ uint32_t gbase[256];
	__shared__ uint32_t wcnt[RS_NW][256];
	__shared__ uint32_t scratch[RS_BS / N2M_WAVE + 1];
	const RsTile t = rs_tile(desc, n, nblocks);
	const size_t base = t.start;
	const int tn = (int)t.len;
	run[threadIdx.x] = 0;
	gbase[threadIdx.x] = offs[t.cbase + (size_t)threadIdx.x * t.cstride];
	for (int w = 0; w < RS_NW; w++)
		wcnt[w][threadIdx.x] = 0;
	__syncthreads();
	for (int k = 0; k < RS_IPT; k++) {
		const int i = k * RS_BS + threadIdx.x;
		if (i < tn)
			atomicAdd(&run[(kin[base + i] >> shift) & 0xFF], 1u);
	}
	__syncthreads();
	uint32_t tot;
	lofs[threadIdx.x] = block_excl_scan<RS_BS>(run[threadIdx.x], scratch, &tot); // thread t <-> digit t
	run[threadIdx.x] = 0;
	__syncthreads();
	const int lane = lane_id(), wv = wave_id();
	const unsigned long long lt = (1ull << lane) - 1ull;
	for (int k = 0; k < RS_IPT; k++) {
		const int i = k * RS_BS + threadIdx.x;
		const bool act = i < tn;
		const uint32_t key = act ? kin[base + i] : 0u, val = act ? vin[base + i] : 0u;
		const uint32_t d = (key >> shift) & 0xFF;
		unsigned long long m = __ballot(act);
#pragma unroll
		for (int b = 0; b < 8; b++) {
			const bool bit = (d >> b) & 1u;
			const unsigned long long bb = __ballot(bit);
			m &= bit ? bb : ~bb;
		}
		const uint32_t rk = __popcll(m & lt);
		if (act && rk == 0)
			wcnt[wv][d] = __popcll(m);
		__syncthreads();
		if (act) {
			uint32_t pos = run[d] + rk;
			for (int w = 0; w < wv; w++)
				pos += wcnt[w][d];
			s_k[lofs[d] + pos] = key;
			s_v[lofs[d] + pos] = val;
		}
		__syncthreads();
		uint32_t add = 0;
		for (int w = 0; w < RS_NW; w++) {
			add += wcnt[w][threadIdx.x];
			wcnt[w][threadIdx.x] = 0;
		}
		run[threadIdx.x] += add;
		__syncthreads();
	}
	for (int i = threadIdx.x; i < tn; i += RS_BS) {
		const uint32_t key = s_k[i], d = (key >> shift) & 0xFF;
		const uint32_t o = gbase[d] + (uint32_t)(i - (int)lofs[d]);
		kout[o] = key;
		vout[o] = s_v[i];
	}
}

// Sorts (k0,v0) using (k1,v1) as the ping-pong buffers; returns true when
// the result is in (k1,v1).  counts/offs hold 256*nblocks words each.  With
// tile descriptors (radix_seg_tiles) every segment is sorted on its own and
// stays where it is: a segment's tiles count their digits in a region of
// their own, laid out digit-major after the regions of the earlier segments,
// so the one scan of all counts places each key inside its segment.
static bool radix_passes(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int bits,
                         uint32_t *counts, uint32_t *scan_scratch, size_t nb, const RsTile *desc, hipStream_t s)
{
	bool flip = false;
	for (int shift = 0; shift < bits; shift += 8) {
		const uint32_t *ki = flip ? k1 : k0, *vi = flip ? v1 : v0;
		uint32_t *ko = flip ? k0 : k1, *vo = flip ? v0 : v1;
		hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(RS_BS), 0, s, ki, n, shift, counts, nb, desc);
		scan_excl_u32(counts, counts, 256 * nb, nullptr, scan_scratch, s);
		hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(RS_BS), 0, s, ki, vi, ko, vo, n, shift,
		                   counts, nb, desc);
		flip = !flip;
	}
	return flip;
}

bool radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int bits,
                    uint32_t *counts, uint32_t *scan_scratch, hipStream_t s)
{
	if (n <= 1)
		return false;
	return radix_passes(k0, v0, k1, v1, n, bits, counts, scan_scratch, (n + RS_TILE - 1) / RS_TILE, nullptr, s);
}

std::vector<RsTile> radix_seg_tiles(const std::vector<uint32_t> &sb)
{
	std::vector<RsTile> d;
	for (size_t k = 0; k + 1 < sb.size(); k++) {
		const uint32_t len = sb[k + 1] - sb[k], nt = (len + RS_TILE - 1) / RS_TILE;
		const uint32_t first = (uint32_t)d.size();
		for (uint32_t j = 0; j < nt; j++) {
			RsTile t;
			t.start = sb[k] + j * RS_TILE;
			t.len = len - j * RS_TILE < (uint32_t)RS_TILE ? len - j * RS_TILE : (uint32_t)RS_TILE;
			t.cbase = 256 * first + j;
			t.cstride = nt;
			d.push_back(t);
		}
	}
	return d;
}

bool radix_sort_seg_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, size_t n, int bits,
                        const RsTile *desc, size_t ntiles, uint32_t *counts, uint32_t *scan_scratch, hipStream_t s)
{
	if (n <= 1 || !ntiles)
		return false;
	return radix_passes(k0, v0, k1, v1, n, bits, counts, scan_scratch, ntiles, desc, s);
}

size_t radix_counts_words(size_t n) { return 256 * ((n + RS_TILE - 1) / RS_TILE) + 1; }

// ---------------------------------------------------------------------------
// Vertex welding, src/meshify.c:45-106.  The reference sorts FP32 distances
// to vertex 0 (stable radix sort), then scans the sorted list: an unassigned
// vertex seeds a new class and every vertex in its window (key within 1e-5,
// FP32) that lies within 1e-5 in 3-D (FP64) joins that class, even one a
// previous seed claimed.  A window can never cross a gap where
// fl(key[k]-key[k-1]) >= tol, so the sorted list splits at such gaps into
// independent segments: one thread runs the reference's loop per segment;
// class numbers = number of seeds before, in sorted order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dist3(const double *a, const double *b)
{
	double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
	return sqrt(dx * dx + dy * dy + dz * dz);
}

// Sort key of a weld distance: the radix order of src/radixsort.c:31-35.
// A NaN distance is keyed without its sign (after every number, by
// payload): the sign a NaN picks up in a difference is not portable (x86
// returns the NaN operand, the GPU's subtract negates it), and a positive
// NaN is what the reference gets from positive NaN voxels.
__device__ __forceinline__ uint32_t weld_key(float d)
{
	return d != d ? (__float_as_uint(d) | 0x80000000u) : f2ord(d);
}

__global__ void k_weld_keys(const double *__restrict__ pts, uint32_t n, uint32_t *keys, uint32_t *vals)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float d = (float)dist3(pts, pts + (size_t)i * 3);
	keys[i] = weld_key(d);
	vals[i] = i;
}

__global__ void k_weld_segflags(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *flags)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float tol = 0.00001f;
	uint32_t f = 1;
	if (i > 0)
		f = (ord2f(keys[i]) - ord2f(keys[i - 1])) >= tol;
	flags[i] = f;
}

// segment starts from the exclusive scan of the flags
__global__ void k_weld_segstarts(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ segid,
                                 uint32_t n, uint32_t *starts)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flags[i])
		starts[segid[i]] = i;
}

// d_nseg (when given) holds the segment count on the device: the launch
// covers nseg = n threads, and those past the count write a zero seed count
// (so that the scan of seeds can run over n words without a host round trip)
__device__ __forceinline__ double dist3v(double ax, double ay, double az, double bx, double by, double bz)
{
	double dx = ax - bx, dy = ay - by, dz = az - bz;
	return sqrt(dx * dx + dy * dy + dz * dz);
}

// Segments of up to WG_REG vertices (most of them: a welded vertex's copies
// share one distance key) run the loop below from registers: their keys,
// ids and positions are loaded up front, all loads in flight together,
// instead of one dependent load chain per comparison.
constexpr int WG_REG = 8;

// lrep (when given): at each seed's sorted position the largest vertex id
// whose final class is the seed's (the member whose position the reference
// keeps, src/meshify.c:100-101), WG_NOSEED elsewhere, so that the class
// representatives need no atomics.  A seed can claim a vertex an earlier
// seed claimed; a segment where that happened takes the maxima again over
// the final classes.
constexpr uint32_t WG_NOSEED = 0xffffffffu;
// one segment of at most WG_REG vertices, by one lane from registers
__device__ __forceinline__ void weld_segment_regs(uint32_t s, uint32_t b, uint32_t len,
                                                  const uint32_t *__restrict__ keys,
                                                  const uint32_t *__restrict__ vals, const double *__restrict__ pts,
                                                  int *__restrict__ lcls, uint32_t *__restrict__ seeds,
                                                  uint32_t *__restrict__ lrep)
{
	const float tol = 0.00001f;
	float k[WG_REG];
	uint32_t v[WG_REG];
	double px[WG_REG], py[WG_REG], pz[WG_REG];
	int cl[WG_REG];
#pragma unroll
	for (int q = 0; q < WG_REG; q++) {
		cl[q] = -1;
		if ((uint32_t)q < len) {
			k[q] = ord2f(keys[b + q]);
			v[q] = vals[b + q];
		}
	}
#pragma unroll
	for (int q = 0; q < WG_REG; q++)
		if ((uint32_t)q < len) {
			const double *p = pts + (size_t)v[q] * 3;
			px[q] = p[0];
			py[q] = p[1];
			pz[q] = p[2];
		}
	int ns = 0;
	unsigned seedmask = 0;
#pragma unroll
	for (int i = 0; i < WG_REG; i++) {
		if ((uint32_t)i >= len || cl[i] >= 0)
			continue;
		cl[i] = ns++;
		seedmask |= 1u << i;
		// the reference's window stops at the first key >= k[i] + tol;
		// keys ascend and fl(k[j] - k[i]) ascends with them
#pragma unroll
		for (int j = i + 1; j < WG_REG; j++)
			if ((uint32_t)j < len && (k[j] - k[i]) < tol &&
			    dist3v(px[i], py[i], pz[i], px[j], py[j], pz[j]) < (double)tol)
				cl[j] = cl[i];
	}
#pragma unroll
	for (int q = 0; q < WG_REG; q++)
		if ((uint32_t)q < len) {
			lcls[b + q] = cl[q];
			if (lrep) {
				uint32_t hi = WG_NOSEED;
				if ((seedmask >> q) & 1u) { // members follow their seed
					hi = v[q];
#pragma unroll
					for (int j = q + 1; j < WG_REG; j++)
						if ((uint32_t)j < len && cl[j] == cl[q])
							hi = max(hi, v[j]);
				}
				lrep[b + q] = hi;
			}
		}
	seeds[s] = (uint32_t)ns;
}

// one segment of any length, by one lane (the reference's loop as written)
__device__ __forceinline__ void weld_segment_serial(uint32_t s, uint32_t b, uint32_t e,
                                                    const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals,
                                                    const double *__restrict__ pts, int *__restrict__ lcls,
                                                    uint32_t *__restrict__ seeds, uint32_t *__restrict__ lrep)
{
	const float tol = 0.00001f;
	for (uint32_t i = b; i < e; i++)
	lcls[i] = -1;
	int ns = 0;
	bool stolen = false;
	for (uint32_t i = b; i < e; i++) {
	if (lcls[i] >= 0) {
		if (lrep)
			lrep[i] = WG_NOSEED;
		continue;
	}
	const float k0 = ord2f(keys[i]);
	const double *seed = pts + (size_t)vals[i] * 3;
	// the seed is always in its own class: for a finite position its
	// distance to itself is 0 < tol anyway; a NaN position matches
	// nothing, and the reference would then leave it unassigned
	// (old2new -1, src/meshify.c:66-79)
	lcls[i] = ns;
	uint32_t hi = vals[i];
	for (uint32_t j = i + 1; j < e && (ord2f(keys[j]) - k0) < tol; j++)
		if (dist3(seed, pts + (size_t)vals[j] * 3) < (double)tol) {
			stolen |= lcls[j] >= 0;
			lcls[j] = ns;
			hi = max(hi, vals[j]);
		}
	if (lrep)
		lrep[i] = hi;
	ns++;
	}
	if (lrep && stolen) // members follow their seed, within its key window
	for (uint32_t i = b; i < e; i++) {
		if (lrep[i] == WG_NOSEED)
			continue;
		const int c = lcls[i];
		const float k0 = ord2f(keys[i]);
		uint32_t hi = vals[i];
		for (uint32_t j = i + 1; j < e && (ord2f(keys[j]) - k0) < tol; j++)
			if (lcls[j] == c)
				hi = max(hi, vals[j]);
		lrep[i] = hi;
	}
	seeds[s] = (uint32_t)ns;
}

// One segment of WG_REG < len <= 64 vertices by the whole wave: lane j holds
// vertex j; the seeds are taken in order, and for each the lanes after it
// test themselves against it at once.  The same comparisons as the serial
// loop (a later seed overwrites an earlier claim), with the loads of the
// segment in flight together instead of one dependent chain per test.
__device__ __forceinline__ void weld_segment_wave(uint32_t s, uint32_t b, uint32_t len,
                                                  const uint32_t *__restrict__ keys,
                                                  const uint32_t *__restrict__ vals, const double *__restrict__ pts,
                                                  int *__restrict__ lcls, uint32_t *__restrict__ seeds,
                                                  uint32_t *__restrict__ lrep)
{
	const float tol = 0.00001f;
	const int j = lane_id();
	const bool in = (uint32_t)j < len;
	float kj = 0.f;
	uint32_t vj = 0;
	double px = 0., py = 0., pz = 0.;
	if (in) {
		kj = ord2f(keys[b + j]);
		vj = vals[b + j];
		const double *p = pts + (size_t)vj * 3;
		px = p[0];
		py = p[1];
		pz = p[2];
	}
	int cl = -1, ns = 0;
	unsigned long long seedm = 0;
	for (int i = 0; i < (int)len; i++) {
		if (__shfl(cl, i) >= 0)
			continue;
		seedm |= 1ull << i;
		const float ki = __shfl(kj, i);
		const double ax = __shfl(px, i), ay = __shfl(py, i), az = __shfl(pz, i);
		if (j == i || (in && j > i && (kj - ki) < tol && dist3v(ax, ay, az, px, py, pz) < (double)tol))
			cl = ns;
		ns++;
	}
	// representative of each seed's final class: its largest member id
	uint32_t hi = WG_NOSEED;
	for (unsigned long long m = seedm; m; m &= m - 1) {
		const int i = __ffsll((long long)m) - 1;
		const int c = __shfl(cl, i);
		const uint32_t mx = wave_max_u32(in && cl == c ? vj : 0u);
		if (j == i)
			hi = mx;
	}
	if (in) {
		lcls[b + j] = cl;
		if (lrep)
			lrep[b + j] = hi;
	}
	if (j == 0)
		seeds[s] = (uint32_t)ns;
}

// lrep (when given): at each seed's sorted position the largest vertex id
// whose final class is the seed's (the member whose position the reference
// keeps, src/meshify.c:100-101), WG_NOSEED elsewhere, so that the class
// representatives need no atomics.  A seed can claim a vertex an earlier
// seed claimed; a segment where that happened takes the maxima again over
// the final classes.  One lane per segment; segments longer than WG_REG
// (and at most a wave) are taken by the lane's whole wave in turn.
__global__ void k_weld_greedy(const uint32_t *__restrict__ starts, uint32_t nseg, uint32_t n,
                              const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                              const double *__restrict__ pts, int *__restrict__ lcls, uint32_t *__restrict__ seeds,
                              const uint32_t *__restrict__ d_nseg, uint32_t *__restrict__ lrep)
{
	const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (d_nseg && s < nseg) {
		const uint32_t ns_dev = *d_nseg;
		if (s >= ns_dev)
			seeds[s] = 0u;
		nseg = ns_dev;
	}
	const bool mine = s < nseg;
	uint32_t b = 0, len = 0;
	if (mine) {
		b = starts[s];
		len = (s + 1 < nseg ? starts[s + 1] : n) - b;
		if (len == 1) { // a vertex alone in its key window is its own class
			lcls[b] = 0;
			if (lrep)
				lrep[b] = vals[b];
			seeds[s] = 1u;
		} else if (len <= (uint32_t)WG_REG)
			weld_segment_regs(s, b, len, keys, vals, pts, lcls, seeds, lrep);
		else if (len > (uint32_t)N2M_WAVE)
			weld_segment_serial(s, b, b + len, keys, vals, pts, lcls, seeds, lrep);
	}
	for (unsigned long long m = __ballot(mine && len > (uint32_t)WG_REG && len <= (uint32_t)N2M_WAVE); m;
	     m &= m - 1) {
		const int o = __ffsll((long long)m) - 1;
		weld_segment_wave(__shfl(s, o), __shfl(b, o), __shfl(len, o), keys, vals, pts, lcls, seeds, lrep);
	}
}

// d_nnew (when given): the class count on the device; when it equals n no
// vertex was merged and, as the reference then keeps its input unchanged
// (src/meshify.c:82-87), every vertex keeps its own number
// The representative of each class (the last member in index order provides
// the position): from lrep at the seeds when given, else by atomics on a rep
// array the caller filled with -1.
__global__ void k_weld_assign(const uint32_t *__restrict__ vals, const int *__restrict__ lcls,
                              const uint32_t *__restrict__ flags, const uint32_t *__restrict__ segid,
                              const uint32_t *__restrict__ seedbase, uint32_t n, int *__restrict__ cls,
                              int *__restrict__ rep, const uint32_t *__restrict__ d_nnew,
                              const uint32_t *__restrict__ lrep)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t sg = segid[i] + flags[i] - 1; // inclusive segment index
	uint32_t v = vals[i];
	const bool ident = d_nnew && *d_nnew == n;
	const int c = ident ? (int)v : (int)seedbase[sg] + lcls[i];
	cls[v] = c;
	if (!lrep)
		atomicMax(&rep[c], (int)v);
	else if (ident)
		rep[c] = (int)v;
	else if (lrep[i] != WG_NOSEED)
		rep[c] = (int)lrep[i];
}

__global__ void k_gather_pts(const double *__restrict__ pts, const int *__restrict__ rep, uint32_t n,
                             double *__restrict__ out, const uint32_t *__restrict__ d_n)
{
	uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= n || (d_n && c >= *d_n)) // d_n: the class count on the device
		return;
	const double *p = pts + (size_t)rep[c] * 3;
	out[(size_t)c * 3 + 0] = p[0];
	out[(size_t)c * 3 + 1] = p[1];
	out[(size_t)c * 3 + 2] = p[2];
}

// tris[t][k] = cls[tris[t][k]]; implicit (classic 3t+k) when tris_in == null.
// An id outside [0, npt) (a corner the tables left without a vertex) stays -1.
__global__ void k_remap_tris(const int *__restrict__ tris_in, const int *__restrict__ cls, uint32_t nt,
                             uint32_t npt, int *__restrict__ tris_out)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= 3 * nt)
		return;
	int v = tris_in ? tris_in[i] : (int)i;
	tris_out[i] = (v >= 0 && (uint32_t)v < npt) ? cls[v] : -1;
}

// workgroups of 256 for n items; at least one, since an empty mesh (no
// triangle, no vertex) still launches and every kernel here returns for
// items at or beyond its count
static unsigned g1(size_t n) { return n ? (unsigned)((n + 255) / 256) : 1u; }

__global__ void k_iota(int *__restrict__ out, uint32_t n)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = (int)i;
}
void launch_iota(int *out, uint32_t n, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_iota, dim3(g1(n)), dim3(256), 0, s, out, n);
}

void launch_weld_keys(const double *pts, uint32_t n, uint32_t *keys, uint32_t *vals, hipStream_t s)
{
	hipLaunchKernelGGL(k_weld_keys, dim3(g1(n)), dim3(256), 0, s, pts, n, keys, vals);
}
void launch_weld_segflags(const uint32_t *keys, uint32_t n, uint32_t *flags, hipStream_t s)
{
	hipLaunchKernelGGL(k_weld_segflags, dim3(g1(n)), dim3(256), 0, s, keys, n, flags);
}
void launch_weld_segstarts(const uint32_t *flags, const uint32_t *segid, uint32_t n, uint32_t *starts,
                           hipStream_t s)
{
	hipLaunchKernelGGL(k_weld_segstarts, dim3(g1(n)), dim3(256), 0, s, flags, segid, n, starts);
}
void launch_weld_greedy(const uint32_t *starts, uint32_t nseg, uint32_t n, const uint32_t *keys,
                        const uint32_t *vals, const double *pts, int *lcls, uint32_t *seeds, uint32_t *lrep,
                        hipStream_t s, const uint32_t *d_nseg)
{
	hipLaunchKernelGGL(k_weld_greedy, dim3(g1(nseg)), dim3(256), 0, s, starts, nseg, n, keys, vals, pts, lcls,
	                   seeds, d_nseg, lrep);
}
void launch_weld_assign(const uint32_t *vals, const int *lcls, const uint32_t *flags, const uint32_t *segid,
                        const uint32_t *seedbase, const uint32_t *lrep, uint32_t n, int *cls, int *rep,
                        hipStream_t s, const uint32_t *d_nnew)
{
	hipLaunchKernelGGL(k_weld_assign, dim3(g1(n)), dim3(256), 0, s, vals, lcls, flags, segid, seedbase, n, cls,
	                   rep, d_nnew, lrep);
}
void launch_gather_pts(const double *pts, const int *rep, uint32_t n, double *out, hipStream_t s,
                       const uint32_t *d_n)
{
	hipLaunchKernelGGL(k_gather_pts, dim3(g1(n)), dim3(256), 0, s, pts, rep, n, out, d_n);
}
void launch_remap_tris(const int *tris_in, const int *cls, uint32_t nt, uint32_t npt, int *tris_out,
                       hipStream_t s)
{
	hipLaunchKernelGGL(k_remap_tris, dim3(g1((size_t)3 * nt)), dim3(256), 0, s, tris_in, cls, nt, npt, tris_out);
}

// ---------------------------------------------------------------------------
// Degenerate triangles, src/meshify.c:113-168 (FP64, same operation order).
// ---------------------------------------------------------------------------
__global__ void k_degen_flags(const double *__restrict__ pts, const int *__restrict__ tris, uint32_t nt,
                              uint32_t np, uint32_t *__restrict__ keep)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	if ((uint32_t)tris[3 * i] >= np || (uint32_t)tris[3 * i + 1] >= np || (uint32_t)tris[3 * i + 2] >= np) {
		keep[i] = 0; // invalid vertex id: never read out of bounds
		return;
	}
	const double *A = pts + (size_t)tris[3 * i] * 3, *B = pts + (size_t)tris[3 * i + 1] * 3,
	             *C = pts + (size_t)tris[3 * i + 2] * 3;
	double l = dist3(A, B), m = dist3(A, C), n = dist3(B, C);
	double c = fmin(fmin(l, m), n);
	double a = fmax(fmax(l, m), n);
	double b = l + m + n - a - c;
	uint32_t k = 1;
	if ((c - (a - b)) <= 0.0) {
		k = 0;
	} else {
		double area4 = 0.25 * sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)));
		if (area4 < FLT_EPSILON)
			k = 0;
	}
	keep[i] = k;
}

// The welded position of every vertex in its own (pre-weld) order, wpos[v] =
// pts[rep[cls[v]]] (the weld's output position of v's class), so that
// triangles can read positions in emit order.  rep[c] is mostly v itself.
__global__ void k_weld_wpos(const double *__restrict__ pts, const int *__restrict__ cls,
                            const int *__restrict__ rep, uint32_t npt, double *__restrict__ wpos)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= npt)
		return;
	const size_t r = (size_t)(uint32_t)rep[cls[v]];
	wpos[(size_t)v * 3] = pts[r * 3];
	wpos[(size_t)v * 3 + 1] = pts[r * 3 + 1];
	wpos[(size_t)v * 3 + 2] = pts[r * 3 + 2];
}

// k_remap_tris and k_degen_flags in one pass: the welded ids of a
// triangle's corners (an id outside [0, npt) stays -1 and marks the
// triangle degenerate) and the degenerate test on the welded positions,
// read as wpos[pre-weld id] -- the same doubles as out[welded id].
__global__ void k_remap_degen(const int *__restrict__ tris_in, const int *__restrict__ cls, uint32_t npt,
                              const double *__restrict__ wpos, uint32_t nt, int *__restrict__ tris_out,
                              uint32_t *__restrict__ keep)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	int v[3];
	bool ok = true;
#pragma unroll
	for (int k = 0; k < 3; k++) {
		v[k] = tris_in ? tris_in[3 * (size_t)i + k] : (int)(3 * i + k);
		const bool valid = v[k] >= 0 && (uint32_t)v[k] < npt;
		tris_out[3 * (size_t)i + k] = valid ? cls[v[k]] : -1;
		ok = ok && valid;
	}
	if (!ok) {
		keep[i] = 0;
		return;
	}
	const double *A = wpos + (size_t)v[0] * 3, *B = wpos + (size_t)v[1] * 3, *C = wpos + (size_t)v[2] * 3;
	double l = dist3(A, B), m = dist3(A, C), n = dist3(B, C);
	double c = fmin(fmin(l, m), n);
	double a = fmax(fmax(l, m), n);
	double b = l + m + n - a - c;
	uint32_t k = 1;
	if ((c - (a - b)) <= 0.0) {
		k = 0;
	} else {
		double area4 = 0.25 * sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)));
		if (area4 < FLT_EPSILON)
			k = 0;
	}
	keep[i] = k;
}

void launch_weld_wpos(const double *pts, const int *cls, const int *rep, uint32_t npt, double *wpos, hipStream_t s)
{
	hipLaunchKernelGGL(k_weld_wpos, dim3(g1(npt)), dim3(256), 0, s, pts, cls, rep, npt, wpos);
}

void launch_remap_degen(const int *tris_in, const int *cls, uint32_t npt, const double *wpos, uint32_t nt,
                        int *tris_out, uint32_t *keep, hipStream_t s)
{
	hipLaunchKernelGGL(k_remap_degen, dim3(g1(nt)), dim3(256), 0, s, tris_in, cls, npt, wpos, nt, tris_out, keep);
}

__global__ void k_compact_tris(const int *__restrict__ tris, const uint32_t *__restrict__ keep,
                               const uint32_t *__restrict__ pos, uint32_t nt, int *__restrict__ out)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt || !keep[i])
		return;
	uint32_t o = pos[i];
	out[3 * o] = tris[3 * i];
	out[3 * o + 1] = tris[3 * i + 1];
	out[3 * o + 2] = tris[3 * i + 2];
}

void launch_degen_flags(const double *pts, const int *tris, uint32_t nt, uint32_t np, uint32_t *keep,
                        hipStream_t s)
{
	hipLaunchKernelGGL(k_degen_flags, dim3(g1(nt)), dim3(256), 0, s, pts, tris, nt, np, keep);
}
void launch_compact_tris(const int *tris, const uint32_t *keep, const uint32_t *pos, uint32_t nt, int *out,
                         hipStream_t s)
{
	hipLaunchKernelGGL(k_compact_tris, dim3(g1(nt)), dim3(256), 0, s, tris, keep, pos, nt, out);
}

// ---------------------------------------------------------------------------
// Smoothing.  CSR over triangle corners: for vertex v, its references
// (t*3+k) in ascending order -- the order src/quadric.c:321-332 adds them --
// stored as the pair of the other two corners (lower corner first).
// ---------------------------------------------------------------------------
// references per vertex; each corner keeps its arrival rank among its
// vertex's references (any order: k_csr_pairs sorts each vertex's list).
// A wave sorts its 64 corners by vertex (bitonic, with shuffles), and each
// run of equal vertices takes its ranks with one atomic: consecutive
// triangles share vertices, so this cuts the atomics about 2.5 times.
__global__ void __launch_bounds__(256)
k_csr_count(const int *__restrict__ tris, uint32_t nt, uint32_t *__restrict__ cnt, uint32_t *__restrict__ rank,
            const uint32_t *__restrict__ inv, int *__restrict__ tout)
{
	const uint32_t n = 3 * nt, i = blockIdx.x * blockDim.x + threadIdx.x, lane = lane_id();
	uint32_t key = 0xffffffffu, src = lane; // vertex ids are < 2^32 - 1
	if (i < n) {
		key = (uint32_t)tris[i];
		if (inv) { // the relabelling of the triangles, written on the way
			key = inv[key];
			tout[i] = (int)key;
		}
	}
	for (uint32_t k = 2; k <= N2M_WAVE; k <<= 1)
		for (uint32_t j = k >> 1; j > 0; j >>= 1) {
			const uint32_t ok = __shfl_xor(key, j, N2M_WAVE), os = __shfl_xor(src, j, N2M_WAVE);
			const bool other_less = ok < key || (ok == key && os < src);
			// the lower lane of the pair keeps the smaller entry in an
			// ascending block, the larger in a descending one
			const bool keep_less = ((lane & j) == 0) == ((lane & k) == 0);
			if (other_less == keep_less) {
				key = ok;
				src = os;
			}
		}
	// lane l now holds the l-th smallest corner; runs of equal vertices
	const uint32_t prev = __shfl(key, lane ? lane - 1 : 0, N2M_WAVE);
	const unsigned long long heads = __ballot(lane == 0 || prev != key);
	const unsigned long long upto = lane == N2M_WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
	const int head = 63 - __clzll(heads & upto);
	const unsigned long long later = heads & ~upto;
	const int next = later ? __ffsll((long long)later) - 1 : N2M_WAVE;
	uint32_t base = 0;
	if (lane == (uint32_t)head && key != 0xffffffffu)
		base = atomicAdd(&cnt[key], (uint32_t)(next - head));
	base = __shfl(base, head, N2M_WAVE);
	if (key != 0xffffffffu)
		rank[i - lane + src] = base + (lane - head);
}

__global__ void k_csr_fill(const int *__restrict__ tris, uint32_t nt, const uint32_t *__restrict__ start,
                           const uint32_t *__restrict__ rank, uint32_t *__restrict__ ref)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= 3 * nt)
		return;
	ref[start[tris[i]] + rank[i]] = i;
}

// sort each vertex's references, then expand to neighbour pairs.  Lists of
// up to CP_MAX references are sorted in LDS (element k of thread t at
// k * CP_BS + t); longer ones in place in global memory.  ref itself is left
// unsorted: its only other reader (k_border) does not depend on the order.
constexpr int CP_BS = 256, CP_MAX = 16;

// the other two corners of corner r (lower corner first)
__device__ __forceinline__ int2 corner_pair(const int *__restrict__ tris, uint32_t r)
{
	const uint32_t t = r / 3, k = r % 3;
	const int c0 = tris[3 * t], c1 = tris[3 * t + 1], c2 = tris[3 * t + 2];
	int2 p;
	p.x = k == 0 ? c1 : c0;
	p.y = k == 2 ? c1 : c2;
	return p;
}

// lists longer than CP_MAX are left to the high-valence pass below
// (sorted_ref: the sorted lists are also written back to ref)
__global__ void __launch_bounds__(CP_BS)
k_csr_pairs(const int *__restrict__ tris, uint32_t nv, const uint32_t *__restrict__ start,
            uint32_t *__restrict__ ref, int2 *__restrict__ pairs, int sorted_ref)
{
	__shared__ uint32_t s_ref[CP_MAX * CP_BS];
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	const uint32_t b = start[v], e = start[v + 1], n = e - b;
	auto expand = [&](uint32_t i, uint32_t r) { pairs[i] = corner_pair(tris, r); };
	if (n <= (uint32_t)CP_MAX) {
		uint32_t *L = s_ref + threadIdx.x;
		for (uint32_t i = 0; i < n; i++) { // insertion sort (valence is small)
			const uint32_t x = ref[b + i];
			uint32_t j = i;
			while (j > 0 && L[(j - 1) * CP_BS] > x) {
				L[j * CP_BS] = L[(j - 1) * CP_BS];
				j--;
			}
			L[j * CP_BS] = x;
		}
		for (uint32_t i = 0; i < n; i++)
			expand(b + i, L[i * CP_BS]);
		if (sorted_ref)
			for (uint32_t i = 0; i < n; i++)
				ref[b + i] = L[i * CP_BS];
	}
}

// Border vertices as update_mesh(iteration 0) marks them: for vertex v,
// every id in the triangles referencing v (once per reference) is counted;
// ids counted exactly once are border.
__global__ void k_border(const int *__restrict__ tris, uint32_t nv, const uint32_t *__restrict__ start,
                         const uint32_t *__restrict__ ref, uint8_t *__restrict__ border)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	uint32_t b = start[v], e = start[v + 1];
	if (e - b > (uint32_t)CP_MAX)
		return; // the high-valence pass counts this list
	for (uint32_t i = b; i < e; i++) {
		uint32_t t = ref[i] / 3;
		for (int k = 0; k < 3; k++) {
			int id = tris[3 * t + k];
			bool first = true;
			for (uint32_t i2 = b; i2 <= i && first; i2++) {
				uint32_t t2 = ref[i2] / 3;
				int kend = i2 == i ? k : 3;
				for (int k2 = 0; k2 < kend; k2++)
					if (tris[3 * t2 + k2] == id) {
						first = false;
						break;
					}
			}
			if (!first)
				continue;
			uint32_t count = 0;
			for (uint32_t i2 = b; i2 < e; i2++) {
				uint32_t t2 = ref[i2] / 3;
				for (int k2 = 0; k2 < 3; k2++)
					count += tris[3 * t2 + k2] == id;
			}
			if (count == 1)
				border[id] = 1;
		}
	}
}

struct d3 {
	double x, y, z;
};

// Neighbour gathers of the smoothers.  Vertices are in Morton order (see
// the relabelling below), so a workgroup of B consecutive vertices finds
// nearly all of its neighbours among the B + 2D vertices around it: that
// window is staged in LDS with coalesced loads, the rest come from global
// memory.  (B, D) is chosen at run time from a few compiled shapes.
// LDS pointer type: keeps ds_read instructions for window hits (a generic
// pointer would let the compiler merge the LDS and global paths into one
// flat load).
typedef __attribute__((address_space(3))) double lds_double;

template <int B, int D> struct Win {
	static constexpr int W = B + 2 * D;
	const lds_double *lds; // W x 3 doubles
	const d3 *__restrict__ g;
	int w0;
	__device__ __forceinline__ d3 get(int i) const
	{
		unsigned r = (unsigned)(i - w0);
		if (r < (unsigned)W) {
			d3 v;
			v.x = lds[3 * r];
			v.y = lds[3 * r + 1];
			v.z = lds[3 * r + 2];
			return v;
		}
		return g[i];
	}
};

// stage p[w0, w0+W) (clamped to [0, nv)) into LDS for vertex block blk;
// has a barrier
template <int B, int D>
__device__ __forceinline__ Win<B, D> stage_window(double *lds, const d3 *__restrict__ p, uint32_t nv, uint32_t blk)
{
	const int w0 = (int)(blk * B) - D;
	const double *src = reinterpret_cast<const double *>(p);
	const long long lo = 3ll * w0, hi = 3ll * nv;
	for (int i = threadIdx.x; i < 3 * Win<B, D>::W; i += B) {
		long long e = lo + i;
		lds[i] = (e >= 0 && e < hi) ? src[e] : 0.0;
	}
	__syncthreads();
	return Win<B, D>{(const lds_double *)lds, p, w0};
}

// Laplacian term of one vertex, src/quadric.c:321-338: the sum over its
// references (ascending triangle order) of (p[a] + p[b]), divided by twice
// the number of references; unchanged when it has none.  The first 8 pairs
// are loaded before any is summed, so their global loads overlap; the sum
// keeps the reference's order.
// Laplacian term of one vertex, src/quadric.c:321-338: the sum over its
// references (ascending triangle order) of (p[a] + p[b]), divided by twice
// the number of references; unchanged when it has none.  The first K pairs
// are loaded before any is summed, so their global loads overlap; the sum
// keeps the reference's order.
template <int K, class WIN>
__device__ __forceinline__ d3 laplace_at(uint32_t b, uint32_t e, const int2 *__restrict__ pairs, const WIN &w,
                                         const d3 self)
{
	if (e <= b)
		return self;
	double sx = 0.0, sy = 0.0, sz = 0.0;
	const uint32_t n = e - b;
	int2 q[K > 0 ? K : 1];
#pragma unroll
	for (int k = 0; k < K; k++)
		q[k] = (uint32_t)k < n ? pairs[b + k] : make_int2(0, 0);
#pragma unroll
	for (int k = 0; k < K; k++) {
		if ((uint32_t)k < n) {
			d3 P = w.get(q[k].x), Q = w.get(q[k].y);
			sx = sx + (P.x + Q.x);
			sy = sy + (P.y + Q.y);
			sz = sz + (P.z + Q.z);
		}
	}
	for (uint32_t i = b + K; i < e; i++) {
		int2 qq = pairs[i];
		d3 P = w.get(qq.x), Q = w.get(qq.y);
		sx = sx + (P.x + Q.x);
		sy = sy + (P.y + Q.y);
		sz = sz + (P.z + Q.z);
	}
	const double num = (double)(2 * n);
	d3 L;
	L.x = sx / num;
	L.y = sy / num;
	L.z = sz / num;
	return L;
}

// p' = L(p); b = p' - (o*alpha + p*alpha1)
template <int B, int D, int K>
__global__ void __launch_bounds__(B)
k_hc_a(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, const d3 *__restrict__ o,
       const d3 *__restrict__ p, d3 *__restrict__ pn, d3 *__restrict__ bq, uint32_t nv, double alpha, double alpha1)
{
	__shared__ double lds[3 * Win<B, D>::W];
	const uint32_t blk = blockIdx.x;
	const Win<B, D> w = stage_window<B, D>(lds, p, nv, blk);
	const uint32_t v = blk * B + threadIdx.x;
	if (v >= nv)
		return;
	const d3 pv = w.get(v), ov = o[v];
	const d3 L = laplace_at<K>(start[v], start[v + 1], pairs, w, pv);
	d3 bb;
	bb.x = L.x - (ov.x * alpha + pv.x * alpha1);
	bb.y = L.y - (ov.y * alpha + pv.y * alpha1);
	bb.z = L.z - (ov.z * alpha + pv.z * alpha1);
	pn[v] = L;
	bq[v] = bb;
}

// p = p' - (b*beta + L(b)*beta1)
template <int B, int D, int K>
__global__ void __launch_bounds__(B)
k_hc_b(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, const d3 *__restrict__ pn,
       const d3 *__restrict__ bq, d3 *__restrict__ pout, uint32_t nv, double beta, double beta1)
{
	__shared__ double lds[3 * Win<B, D>::W];
	const uint32_t blk = blockIdx.x;
	const Win<B, D> w = stage_window<B, D>(lds, bq, nv, blk);
	const uint32_t v = blk * B + threadIdx.x;
	if (v >= nv)
		return;
	const d3 bv = w.get(v), pv = pn[v];
	const d3 L = laplace_at<K>(start[v], start[v + 1], pairs, w, bv);
	d3 r;
	r.x = pv.x - (bv.x * beta + L.x * beta1);
	r.y = pv.y - (bv.y * beta + L.y * beta1);
	r.z = pv.z - (bv.z * beta + L.z * beta1);
	pout[v] = r;
}

// plain Laplacian step (src/quadric.c:316-341): pout = L(p)
template <int B, int D, int K>
__global__ void __launch_bounds__(B)
k_laplace(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, const d3 *__restrict__ p,
          d3 *__restrict__ pout, uint32_t nv)
{
	__shared__ double lds[3 * Win<B, D>::W];
	const uint32_t blk = blockIdx.x;
	const Win<B, D> w = stage_window<B, D>(lds, p, nv, blk);
	const uint32_t v = blk * B + threadIdx.x;
	if (v >= nv)
		return;
	pout[v] = laplace_at<K>(start[v], start[v + 1], pairs, w, w.get(v));
}

// ---------------------------------------------------------------------------
// Halo form of the same gathers.  A workgroup owns HB consecutive vertices
// (Morton order) and stages their positions plus its halo -- the sorted,
// deduplicated neighbours outside the block -- in LDS.  Every gather is then
// an LDS read, and a neighbour outside the block is fetched once per
// workgroup instead of once per reference.  A pair is two 16-bit LDS slots
// (slot < HB: vertex blk*HB + slot; otherwise halo entry slot - HB), 4 bytes
// instead of 8.  Built once per mesh by k_halo_build; a block with more than
// HRAW outside references or HCAP distinct outside neighbours raises
// *overflow, and the caller keeps the window kernels above.
// ---------------------------------------------------------------------------
constexpr int HB = 1024, HCAP = 1024, HRAW = 4096;

// slot of vertex x for block base `base` whose halo is the sorted h[0, H)
__device__ __forceinline__ uint32_t halo_slot(const uint32_t *h, uint32_t H, uint32_t base, uint32_t x)
{
	if (x - base < (uint32_t)HB)
		return x - base;
	uint32_t lo = 0, hi = H; // first entry >= x (x is present)
	while (lo < hi) {
		uint32_t mid = (lo + hi) >> 1;
		if (h[mid] < x)
			lo = mid + 1;
		else
			hi = mid;
	}
	return HB + lo;
}

__global__ void __launch_bounds__(HB)
k_halo_build(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, uint32_t nv,
             uint32_t *__restrict__ halo, uint32_t *__restrict__ hcount, uint32_t *__restrict__ slots,
             uint32_t *__restrict__ overflow)
{
	__shared__ uint32_t raw[HRAW];
	__shared__ uint32_t scratch[HB / N2M_WAVE + 1];
	__shared__ uint32_t s_n;
	const uint32_t blk = blockIdx.x, base = blk * HB, v = base + threadIdx.x;
	if (threadIdx.x == 0)
		s_n = 0;
	__syncthreads();
	uint32_t b = 0, e = 0;
	if (v < nv) {
		b = start[v];
		e = start[v + 1];
	}
	// 1. every reference to a vertex outside the block
	for (uint32_t k = b; k < e; k++) {
		const int2 q = pairs[k];
		const uint32_t x0 = (uint32_t)q.x, x1 = (uint32_t)q.y;
		if (x0 - base >= (uint32_t)HB) {
			uint32_t i = atomicAdd(&s_n, 1u);
			if (i < HRAW)
				raw[i] = x0;
		}
		if (x1 - base >= (uint32_t)HB) {
			uint32_t i = atomicAdd(&s_n, 1u);
			if (i < HRAW)
				raw[i] = x1;
		}
	}
	__syncthreads();
	const uint32_t n = s_n;
	if (n > HRAW) { // uniform: every thread read the same count
		if (threadIdx.x == 0)
			atomicOr(overflow, 1u);
		return;
	}
	// 2. bitonic sort of raw[0, P), padded with ~0
	uint32_t P = 1;
	while (P < n)
		P <<= 1;
	for (uint32_t i = n + threadIdx.x; i < P; i += HB)
		raw[i] = 0xffffffffu;
	__syncthreads();
	for (uint32_t size = 2; size <= P; size <<= 1)
		for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
			for (uint32_t i = threadIdx.x; i < P; i += HB) {
				const uint32_t j = i ^ stride;
				if (j > i) {
					const uint32_t a = raw[i], c = raw[j];
					if ((a > c) == ((i & size) == 0)) {
						raw[i] = c;
						raw[j] = a;
					}
				}
			}
			__syncthreads();
		}
	// 3. keep the first of each run of equal ids (4 entries per thread;
	// HRAW = 4 * HB); compacted in place after the scan's barriers
	uint32_t val[4], cnt = 0;
	bool keep[4];
#pragma unroll
	for (int q = 0; q < 4; q++) {
		const uint32_t i = 4 * threadIdx.x + q;
		val[q] = i < n ? raw[i] : 0u;
		keep[q] = i < n && (i == 0 || raw[i] != raw[i - 1]);
		cnt += keep[q];
	}
	uint32_t H;
	uint32_t off = block_excl_scan<HB>(cnt, scratch, &H);
#pragma unroll
	for (int q = 0; q < 4; q++)
		if (keep[q])
			raw[off++] = val[q];
	__syncthreads();
	if (H > HCAP) { // uniform
		if (threadIdx.x == 0)
			atomicOr(overflow, 1u);
		return;
	}
	for (uint32_t i = threadIdx.x; i < H; i += HB)
		halo[(size_t)blk * HCAP + i] = raw[i];
	if (threadIdx.x == 0)
		hcount[blk] = H;
	// 4. the pairs as LDS slots
	for (uint32_t k = b; k < e; k++) {
		const int2 q = pairs[k];
		slots[k] = halo_slot(raw, H, base, (uint32_t)q.x) | (halo_slot(raw, H, base, (uint32_t)q.y) << 16);
	}
}

// the first K slots of a vertex, loaded before the LDS staging so that their
// latency overlaps it
template <int K> struct SlotRefs {
	uint32_t b, e, q[K > 0 ? K : 1];
	__device__ __forceinline__ void load(const uint32_t *__restrict__ start, const uint32_t *__restrict__ slots,
	                                     uint32_t v, bool live)
	{
		b = live ? start[v] : 0u;
		e = live ? start[v + 1] : 0u;
#pragma unroll
		for (int k = 0; k < K; k++)
			q[k] = b + k < e ? slots[b + k] : 0u;
	}
};

__device__ __forceinline__ d3 lds_row(const lds_double *L, uint32_t s)
{
	d3 r;
	r.x = L[3 * s];
	r.y = L[3 * s + 1];
	r.z = L[3 * s + 2];
	return r;
}

// block rows and halo rows of p into LDS (has a barrier)
__device__ __forceinline__ void stage_halo(double *lds, const d3 *__restrict__ p, uint32_t nv, uint32_t blk,
                                           const uint32_t *__restrict__ halo, uint32_t H)
{
	const double *src = reinterpret_cast<const double *>(p);
	const size_t lo = 3 * (size_t)blk * HB, hi = 3 * (size_t)nv;
	for (int i = threadIdx.x; i < 3 * HB; i += HB) {
		const size_t e = lo + i;
		lds[i] = e < hi ? src[e] : 0.0;
	}
	const uint32_t *h = halo + (size_t)blk * HCAP;
	for (uint32_t i = threadIdx.x; i < H; i += HB) {
		const d3 q = p[h[i]];
		lds[3 * (HB + i)] = q.x;
		lds[3 * (HB + i) + 1] = q.y;
		lds[3 * (HB + i) + 2] = q.z;
	}
	__syncthreads();
}

// row s of an LDS array of 3-double rows (W = 3), or, for W = 2, of the
// split form of NR rows: NR 16-byte (x, y) pairs, then the z values at a
// 16-byte stride (so that both reads share one address).  Random rows of
// the split form are read about a third faster than 32-byte (x, y, z, pad)
// rows: its 16-byte reads fall on every bank group, those of 32-byte rows
// on half of them.
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) dbl2 lds_dbl2;
template <int W, int NR = 0> __device__ __forceinline__ d3 lds_row_w(const lds_double *L, uint32_t s)
{
	if constexpr (W == 2) {
		const dbl2 xy = ((const lds_dbl2 *)L)[s];
		d3 r;
		r.x = xy.x;
		r.y = xy.y;
		r.z = L[2 * NR + 2 * s];
		return r;
	} else {
		return lds_row(L, s);
	}
}

// writes row s of the split form of NR rows
template <int NR> __device__ __forceinline__ void lds_put_split(double *L, uint32_t s, const d3 v)
{
	dbl2 xy;
	xy.x = v.x;
	xy.y = v.y;
	((dbl2 *)L)[s] = xy;
	L[2 * NR + 2 * s] = v.z;
}

// laplace_at() over LDS slots: same terms, same order
template <int K, int W = 3, int NR = 0>
__device__ __forceinline__ d3 laplace_slots(const SlotRefs<K> &r, const uint32_t *__restrict__ slots,
                                            const lds_double *L, const d3 self)
{
	if (r.e <= r.b)
		return self;
	double sx = 0.0, sy = 0.0, sz = 0.0;
	const uint32_t n = r.e - r.b;
#pragma unroll
	for (int k = 0; k < K; k++) {
		if ((uint32_t)k < n) {
			const d3 P = lds_row_w<W, NR>(L, r.q[k] & 0xffffu), Q = lds_row_w<W, NR>(L, r.q[k] >> 16);
			sx = sx + (P.x + Q.x);
			sy = sy + (P.y + Q.y);
			sz = sz + (P.z + Q.z);
		}
	}
	for (uint32_t i = r.b + K; i < r.e; i++) {
		const uint32_t s = slots[i];
		const d3 P = lds_row_w<W, NR>(L, s & 0xffffu), Q = lds_row_w<W, NR>(L, s >> 16);
		sx = sx + (P.x + Q.x);
		sy = sy + (P.y + Q.y);
		sz = sz + (P.z + Q.z);
	}
	const double num = (double)(2 * n);
	d3 out;
	out.x = sx / num;
	out.y = sy / num;
	out.z = sz / num;
	return out;
}

constexpr int HK = 8; // slots prefetched per vertex

// k_hc_a in halo form
__global__ void __launch_bounds__(HB)
k_hc_a_h(const uint32_t *__restrict__ start, const uint32_t *__restrict__ slots, const uint32_t *__restrict__ halo,
         const uint32_t *__restrict__ hcount, const d3 *__restrict__ o, const d3 *__restrict__ p,
         d3 *__restrict__ pn, d3 *__restrict__ bq, uint32_t nv, double alpha, double alpha1)
{
	__shared__ double lds[3 * (HB + HCAP)];
	const uint32_t blk = blockIdx.x, v = blk * HB + threadIdx.x;
	const bool live = v < nv;
	SlotRefs<HK> r;
	r.load(start, slots, v, live);
	d3 ov = {0.0, 0.0, 0.0};
	if (live)
		ov = o[v];
	stage_halo(lds, p, nv, blk, halo, hcount[blk]);
	if (!live)
		return;
	const lds_double *L = (const lds_double *)lds;
	const d3 pv = lds_row(L, threadIdx.x);
	const d3 Lp = laplace_slots<HK>(r, slots, L, pv);
	d3 bb;
	bb.x = Lp.x - (ov.x * alpha + pv.x * alpha1);
	bb.y = Lp.y - (ov.y * alpha + pv.y * alpha1);
	bb.z = Lp.z - (ov.z * alpha + pv.z * alpha1);
	pn[v] = Lp;
	bq[v] = bb;
}

// k_hc_b in halo form
__global__ void __launch_bounds__(HB)
k_hc_b_h(const uint32_t *__restrict__ start, const uint32_t *__restrict__ slots, const uint32_t *__restrict__ halo,
         const uint32_t *__restrict__ hcount, const d3 *__restrict__ pn, const d3 *__restrict__ bq,
         d3 *__restrict__ pout, uint32_t nv, double beta, double beta1)
{
	__shared__ double lds[3 * (HB + HCAP)];
	const uint32_t blk = blockIdx.x, v = blk * HB + threadIdx.x;
	const bool live = v < nv;
	SlotRefs<HK> r;
	r.load(start, slots, v, live);
	d3 pv = {0.0, 0.0, 0.0};
	if (live)
		pv = pn[v];
	stage_halo(lds, bq, nv, blk, halo, hcount[blk]);
	if (!live)
		return;
	const lds_double *L = (const lds_double *)lds;
	const d3 bv = lds_row(L, threadIdx.x);
	const d3 Lb = laplace_slots<HK>(r, slots, L, bv);
	d3 out;
	out.x = pv.x - (bv.x * beta + Lb.x * beta1);
	out.y = pv.y - (bv.y * beta + Lb.y * beta1);
	out.z = pv.z - (bv.z * beta + Lb.z * beta1);
	pout[v] = out;
}

// k_laplace in halo form
__global__ void __launch_bounds__(HB)
k_laplace_h(const uint32_t *__restrict__ start, const uint32_t *__restrict__ slots, const uint32_t *__restrict__ halo,
            const uint32_t *__restrict__ hcount, const d3 *__restrict__ p, d3 *__restrict__ pout, uint32_t nv)
{
	__shared__ double lds[3 * (HB + HCAP)];
	const uint32_t blk = blockIdx.x, v = blk * HB + threadIdx.x;
	const bool live = v < nv;
	SlotRefs<HK> r;
	r.load(start, slots, v, live);
	stage_halo(lds, p, nv, blk, halo, hcount[blk]);
	if (!live)
		return;
	const lds_double *L = (const lds_double *)lds;
	pout[v] = laplace_slots<HK>(r, slots, L, lds_row(L, threadIdx.x));
}

// ---------------------------------------------------------------------------
// Two-ring form: one launch per HC iteration.  A workgroup of RT threads owns
// RB consecutive vertices.  Ring 1 is the sorted distinct neighbours outside
// the block, ring 2 the neighbours of ring 1 outside block and ring 1.  The
// kernel stages the positions of block, ring 1 and ring 2 in LDS (row slots:
// block 0..RB-1, ring 1 from RB, ring 2 after it), computes the Laplacian and
// b for block and ring-1 vertices (the hc_a step), replaces the rows with b,
// and finishes the block's vertices (the hc_b step).  Only the new positions
// go to memory: no pn / b arrays.  Ring-1 vertices are computed by every block
// that borders them, with the same terms in the same order, so the values are
// those of the two-launch form.  Ring-1 pair lists are kept per block as LDS
// slots (roff: 16-bit offsets, rslots: slot pairs).
// ---------------------------------------------------------------------------
// RB (vertices per block) is a template parameter: 512, 640 or 768.
constexpr int RT = 1024;
constexpr int R1CAP = 1024;      // ring-1 vertices per block (RB + R1CAP <= 2 * RT)
constexpr int RROWS = 2304;      // LDS rows: RB + ring 1 + ring 2 (ring words per block: RROWS - RB)
constexpr int RPCAP = 4096;      // ring-1 pair words per block
constexpr int ROFFW = R1CAP + 2; // 16-bit ring-1 offsets per block
constexpr int RK = 6;            // pairs per vertex held in registers by k_ring_build

// LDS hash set of vertex ids (open addressing, RHS slots) used by
// k_ring_build: key RH_EMPTY marks a free slot, val the id's LDS row.
constexpr int RHS = 4096;
constexpr uint32_t RH_EMPTY = 0xffffffffu;
__device__ __forceinline__ uint32_t rh_hash(uint32_t x) { return (x * 2654435761u) >> 20; } // 12 bits

// insert x; a new id is appended (as its slot) to list[*count] when *count <
// cap.  Returns false when the table is full or the list overflows.
__device__ __forceinline__ bool rh_insert(uint32_t *key, uint32_t x, uint32_t *list, uint32_t *count, uint32_t cap)
{
	uint32_t h = rh_hash(x);
	for (int p = 0; p < RHS; p++, h = (h + 1) & (RHS - 1)) {
		const uint32_t prev = atomicCAS(&key[h], RH_EMPTY, x);
		if (prev == RH_EMPTY) {
			const uint32_t i = atomicAdd(count, 1u);
			if (i >= cap)
				return false;
			list[i] = h;
			return true;
		}
		if (prev == x)
			return true;
	}
	return false;
}

// slot of x, which is present
__device__ __forceinline__ uint32_t rh_find(const uint32_t *key, uint32_t x)
{
	uint32_t h = rh_hash(x);
	for (int p = 0; p < RHS && key[h] != x; p++)
		h = (h + 1) & (RHS - 1);
	return h;
}

// The n ids at key[list[0, n)] in ascending order into out[0, n), and each
// id's row (row0 + its rank) into val.  Rank = number of smaller ids (the
// ids are distinct); O(n^2 / RT) LDS reads, broadcast across the wave.
__device__ __forceinline__ void rh_rank(const uint32_t *key, const uint32_t *list, uint32_t n, uint32_t *ids,
                                        uint32_t *out, uint32_t *val, uint32_t row0)
{
	for (uint32_t i = threadIdx.x; i < n; i += RT)
		ids[i] = key[list[i]];
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < n; i += RT) {
		const uint32_t x = ids[i];
		uint32_t r = 0;
		for (uint32_t j = 0; j < n; j++)
			r += ids[j] < x;
		out[r] = x;
		val[list[i]] = row0 + r;
	}
	__syncthreads();
}

// Per block: ring 1, ring 2, the block's pairs as row slots (slots, indexed
// like pairs) and ring 1's pairs (roff / rslots).  The rings are collected in
// an LDS hash set (an id already present -- block, ring 1 or ring 2 -- is
// skipped) and then ordered by id.  A block whose rings exceed the caps
// raises *overflow and the caller keeps the two-launch form.
template <int RB>
__global__ void __launch_bounds__(RT, 8)
k_ring_build(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, uint32_t nv,
             const d3 *__restrict__ o, uint32_t *__restrict__ ring, uint2 *__restrict__ rcnt,
             uint32_t *__restrict__ slots, uint16_t *__restrict__ roff, uint32_t *__restrict__ rslots,
             d3 *__restrict__ o1, uint32_t *__restrict__ overflow)
{
	constexpr int RCAP = RROWS - RB;
	__shared__ uint32_t hkey[RHS], hval[RHS];
	__shared__ uint32_t lst[RCAP], ids[RCAP], srt[RCAP]; // hash slots, ids, sorted ids (ring 1, then ring 2)
	__shared__ uint32_t scratch[RT / N2M_WAVE + 1];
	__shared__ uint32_t s_n1, s_n2, s_bad;
	const uint32_t blk = blockIdx.x, base = blk * RB, tid = threadIdx.x;
	for (uint32_t i = tid; i < RHS; i += RT)
		hkey[i] = RH_EMPTY;
	if (tid == 0) {
		s_n1 = 0;
		s_n2 = 0;
		s_bad = 0;
	}
	uint32_t b = 0, e = 0;
	if (tid < RB && base + tid < nv) {
		b = start[base + tid];
		e = start[base + tid + 1];
	}
	// the first RK pairs of a list are loaded at once (in-block filler beyond
	// the list); each pass loads them again rather than holding them
	auto load_pairs = [&](int2 *q, uint32_t lb, uint32_t le) {
#pragma unroll
		for (int k = 0; k < RK; k++)
			q[k] = lb + k < le ? pairs[lb + k] : make_int2((int)base, (int)base);
	};
	// ring 1: the block's references to vertices outside it
	auto add1 = [&](const int2 q) {
#pragma unroll
		for (int c = 0; c < 2; c++) {
			const uint32_t x = (uint32_t)(c ? q.y : q.x);
			if (x - base >= (uint32_t)RB && !rh_insert(hkey, x, lst, &s_n1, R1CAP))
				s_bad = 1;
		}
	};
	{
		int2 qa[RK];
		load_pairs(qa, b, e);
		__syncthreads();
#pragma unroll
		for (int k = 0; k < RK; k++)
			add1(qa[k]);
	}
	for (uint32_t k = b + RK; k < e; k++)
		add1(pairs[k]);
	__syncthreads();
	if (s_bad) { // uniform
		if (tid == 0)
			atomicOr(overflow, 1u);
		return;
	}
	const uint32_t n1 = s_n1;
	rh_rank(hkey, lst, n1, ids, srt, hval, RB);
	// ring 2: ring 1's references to vertices outside block and ring 1
	// (ring-1 vertex i = tid, n1 <= RT)
	uint32_t hb = 0, he = 0;
	d3 ov1 = {0.0, 0.0, 0.0}; // its original position, loaded with its offsets
	if (tid < n1) {
		const uint32_t h = srt[tid];
		hb = start[h];
		he = start[h + 1];
		ov1 = o[h];
	}
	auto add2 = [&](const int2 q) {
#pragma unroll
		for (int c = 0; c < 2; c++) {
			const uint32_t x = (uint32_t)(c ? q.y : q.x);
			if (x - base >= (uint32_t)RB && !rh_insert(hkey, x, lst + n1, &s_n2, RCAP - n1))
				s_bad = 1;
		}
	};
	{
		int2 qr[RK];
		load_pairs(qr, hb, he);
#pragma unroll
		for (int k = 0; k < RK; k++)
			add2(qr[k]);
	}
	for (uint32_t k = hb + RK; k < he; k++)
		add2(pairs[k]);
	__syncthreads();
	if (s_bad) { // uniform
		if (tid == 0)
			atomicOr(overflow, 1u);
		return;
	}
	const uint32_t n2 = s_n2;
	rh_rank(hkey, lst + n1, n2, ids + n1, srt + n1, hval, RB + n1);
	// ring-1 pair offsets (one ring-1 vertex per thread)
	uint32_t ptot;
	const uint32_t poff = block_excl_scan<RT>(he - hb, scratch, &ptot);
	if (ptot > (uint32_t)RPCAP) { // uniform
		if (tid == 0)
			atomicOr(overflow, 1u);
		return;
	}
	uint32_t *rg = ring + (size_t)blk * RCAP;
	for (uint32_t i = tid; i < n1 + n2; i += RT)
		rg[i] = srt[i];
	if (tid < n1)
		o1[(size_t)blk * R1CAP + tid] = ov1;
	if (tid == 0)
		rcnt[blk] = make_uint2(n1, n2);
	uint16_t *ro = roff + (size_t)blk * ROFFW;
	if (tid < n1)
		ro[tid] = (uint16_t)poff;
	if (tid == 0)
		ro[n1] = (uint16_t)ptot;
	// row slot of a vertex of block, ring 1 or ring 2
	auto row = [&](uint32_t x) -> uint32_t { return x - base < (uint32_t)RB ? x - base : hval[rh_find(hkey, x)]; };
	auto slot = [&](const int2 q) { return row((uint32_t)q.x) | (row((uint32_t)q.y) << 16); };
	{
		int2 qa[RK];
		load_pairs(qa, b, e);
#pragma unroll
		for (int k = 0; k < RK; k++)
			if (b + k < e)
				slots[b + k] = slot(qa[k]);
	}
	for (uint32_t k = b + RK; k < e; k++)
		slots[k] = slot(pairs[k]);
	uint32_t *rs = rslots + (size_t)blk * RPCAP + poff;
	{
		int2 qr[RK];
		load_pairs(qr, hb, he);
#pragma unroll
		for (int k = 0; k < RK; k++)
			if (hb + k < he)
				rs[k] = slot(qr[k]);
	}
	for (uint32_t k = hb + RK; k < he; k++)
		rs[k - hb] = slot(pairs[k]);
}

// one HC iteration (hc_a then hc_b) in two-ring form; pout != p.  LDS rows
// in split form (an (x, y) array and a z array, read as 16 + 8 bytes).
template <int RB>
__global__ void __launch_bounds__(RT)
k_hc_ring(const uint32_t *__restrict__ start, const uint32_t *__restrict__ slots, const uint16_t *__restrict__ roff,
          const uint32_t *__restrict__ rslots, const uint32_t *__restrict__ ring, const uint2 *__restrict__ rcnt,
          const d3 *__restrict__ o, const d3 *__restrict__ o1, const d3 *__restrict__ p, d3 *__restrict__ pout,
          uint32_t nv, double alpha, double alpha1, double beta, double beta1)
{
	constexpr int W = 2;
	__shared__ double lds[4 * RROWS];
	constexpr int RCAP = RROWS - RB;
	const uint32_t blk = blockIdx.x, base = blk * RB, tid = threadIdx.x;
	const uint2 nn = rcnt[blk];
	const uint32_t n1 = nn.x, n2 = nn.y;
	const uint32_t *rg = ring + (size_t)blk * RCAP;
	const uint16_t *ro = roff + (size_t)blk * ROFFW;
	const uint32_t *rs = rslots + (size_t)blk * RPCAP;
	const d3 *ob = o1 + (size_t)blk * R1CAP; // original positions of ring 1
	// row tid: a block vertex (tid < RB) or ring-1 vertex tid - RB
	const bool own = tid < RB && base + tid < nv;
	const bool r1a = tid >= RB && tid - RB < n1;
	SlotRefs<HK> r;
	const uint32_t *sl = own ? slots : rs;
	d3 ov = {0.0, 0.0, 0.0}, pv = {0.0, 0.0, 0.0};
	{
		uint32_t gid = 0, b = 0, e = 0;
		if (own) {
			gid = base + tid;
			b = start[gid];
			e = start[gid + 1];
		} else if (r1a) {
			gid = rg[tid - RB];
			b = ro[tid - RB];
			e = ro[tid - RB + 1];
		}
		r.b = b;
		r.e = e;
#pragma unroll
		for (int k = 0; k < HK; k++)
			r.q[k] = b + k < e ? sl[b + k] : 0u;
		if (own)
			ov = o[gid];
		else if (r1a)
			ov = ob[tid - RB];
		if (own)
			pv = p[gid];
	}
	// stage block (the own rows, already in registers), ring-1 and ring-2
	// positions
	if (tid < RB)
		lds_put_split<RROWS>(lds, tid, pv);
	for (uint32_t i = tid; i < n1 + n2; i += RT)
		lds_put_split<RROWS>(lds, RB + i, p[rg[i]]);
	__syncthreads();
	const lds_double *L = (const lds_double *)lds;
	// hc_a on row tid and, for rings longer than RT - RB, on row tid + RT
	d3 Lp = {0.0, 0.0, 0.0}, bb = {0.0, 0.0, 0.0}, bb2 = {0.0, 0.0, 0.0};
	if (own || r1a) {
		if (!own)
			pv = lds_row_w<W, RROWS>(L, tid);
		Lp = laplace_slots<HK, W, RROWS>(r, sl, L, pv);
		bb.x = Lp.x - (ov.x * alpha + pv.x * alpha1);
		bb.y = Lp.y - (ov.y * alpha + pv.y * alpha1);
		bb.z = Lp.z - (ov.z * alpha + pv.z * alpha1);
	}
	const uint32_t u2 = tid + RT; // always a ring-1 row (RT >= RB)
	const bool r2a = u2 - RB < n1;
	if (r2a) { // its first HK slots are loaded at once
		SlotRefs<HK> r2;
		r2.b = ro[u2 - RB];
		r2.e = ro[u2 - RB + 1];
#pragma unroll
		for (int k = 0; k < HK; k++)
			r2.q[k] = r2.b + k < r2.e ? rs[r2.b + k] : 0u;
		const d3 o2 = ob[u2 - RB];
		const d3 pv2 = lds_row_w<W, RROWS>(L, u2);
		const d3 L2 = laplace_slots<HK, W, RROWS>(r2, rs, L, pv2);
		bb2.x = L2.x - (o2.x * alpha + pv2.x * alpha1);
		bb2.y = L2.y - (o2.y * alpha + pv2.y * alpha1);
		bb2.z = L2.z - (o2.z * alpha + pv2.z * alpha1);
	}
	__syncthreads(); // every position row has been read
	if (own || r1a)
		lds_put_split<RROWS>(lds, tid, bb);
	if (r2a)
		lds_put_split<RROWS>(lds, u2, bb2);
	__syncthreads();
	if (!own)
		return;
	// hc_b on the block's vertices: their pairs reach block and ring 1 only
	const d3 Lb = laplace_slots<HK, W, RROWS>(r, sl, L, bb);
	d3 out;
	out.x = Lp.x - (bb.x * beta + Lb.x * beta1);
	out.y = Lp.y - (bb.y * beta + Lb.y * beta1);
	out.z = Lp.z - (bb.z * beta + Lb.z * beta1);
	pout[base + tid] = out;
}

// gather through a permutation: dst[i] = src[perm ? perm[i] : i]
__global__ void k_hc_load(const d3 *__restrict__ src, const uint32_t *__restrict__ perm, uint32_t n, d3 *__restrict__ dst)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = src[perm ? perm[i] : i];
}

// final HC positions (border vertices keep the original), scattered back
// through the permutation: dst[perm ? perm[i] : i] = ...
__global__ void k_hc_store(const d3 *__restrict__ p, const d3 *__restrict__ o, const uint8_t *__restrict__ border,
                        const uint32_t *__restrict__ perm, uint32_t n, d3 *__restrict__ dst)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[perm ? perm[i] : i] = (border && border[i]) ? o[i] : p[i];
}

// ---------------------------------------------------------------------------
// High-valence vertices: more than CP_MAX references (a fan apex, say).  One
// lane per list would take O(d^2) steps; instead all their references are
// handled at once.  The corners of such vertices are compacted in corner
// order, stably sorted by vertex, and placed (lists in ascending corner
// order, the order of src/quadric.c:321-332); their border flags come from
// a sort of (vertex, id) entries, in which an entry that occurs once is an
// id counted once for that vertex (src/quadric.c:188-215).
// ---------------------------------------------------------------------------
// lcnt[v] = the list length of a long vertex, else 0 (v <= nv: a 0 sentinel)
// ---------------------------------------------------------------------------
// High-valence lists (more than CP_MAX references), one wave per list: the
// lanes of a thread-per-vertex launch that hold such a vertex hand it to
// their whole wave in turn.  A list of up to LW_K * 64 elements is held in
// registers (element i in slot i / 64 of lane i % 64) and every element is
// compared with all the others through shuffles; a longer one is compared
// from global memory (the same counts, more loads).
// ---------------------------------------------------------------------------
constexpr int LW_K = 4;

// counts, for each element a of lane slots r[], the elements below it (less)
// and equal to it (eq) among the E elements of the wave
__device__ __forceinline__ void wave_rank_eq(const uint32_t (&r)[LW_K], uint32_t E, uint32_t (&less)[LW_K],
                                             uint32_t (&eq)[LW_K])
{
#pragma unroll
	for (int k = 0; k < LW_K; k++)
		less[k] = eq[k] = 0;
#pragma unroll
	for (int si = 0; si < LW_K; si++) {
		if ((uint32_t)si * N2M_WAVE >= E)
			break;
		const uint32_t nl = E - si * N2M_WAVE < (uint32_t)N2M_WAVE ? E - si * N2M_WAVE : (uint32_t)N2M_WAVE;
		for (uint32_t li = 0; li < nl; li++) {
			const uint32_t x = __shfl(r[si], (int)li);
#pragma unroll
			for (int k = 0; k < LW_K; k++) {
				less[k] += x < r[k];
				eq[k] += x == r[k];
			}
		}
	}
}

// Sorts each long list of ref ascending in place and expands its pairs
// (what k_csr_pairs does for the short lists).  A list longer than the
// register path is appended to `huge` (huge[0] = count, the vertices
// after it) for k_huge_sort.
__global__ void k_long_sort(const int *__restrict__ tris, uint32_t nv, const uint32_t *__restrict__ start,
                            uint32_t *__restrict__ ref, int2 *__restrict__ pairs, uint32_t *__restrict__ huge)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t deg = v < nv ? start[v + 1] - start[v] : 0u;
	const int lane = lane_id();
	for (unsigned long long m = __ballot(deg > (uint32_t)CP_MAX); m; m &= m - 1) {
		const int o = __ffsll((long long)m) - 1;
		const uint32_t u = __shfl(v, o), d = __shfl(deg, o), b = start[u];
		if (d > (uint32_t)(LW_K * N2M_WAVE)) {
			if (lane == 0)
				huge[1 + atomicAdd(&huge[0], 1u)] = u;
			continue;
		}
		uint32_t r[LW_K], less[LW_K], eq[LW_K];
#pragma unroll
		for (int k = 0; k < LW_K; k++) {
			const uint32_t i = k * N2M_WAVE + lane;
			r[k] = i < d ? ref[b + i] : 0xffffffffu;
		}
		wave_rank_eq(r, d, less, eq);
#pragma unroll
		for (int k = 0; k < LW_K; k++) {
			const uint32_t i = k * N2M_WAVE + lane;
			if (i < d) { // the references are distinct corners: ranks are a permutation
				ref[b + less[k]] = r[k];
				pairs[b + less[k]] = corner_pair(tris, r[k]);
			}
		}
	}
}

// Border flags of the long lists (k_border and k_border_pairs leave them):
// the ids of a list's entries -- the two pair ids of each reference, or
// (literal) the three ids of its triangle -- are counted within the list;
// one that occurs once marks the vertex (pair form) or the id (literal).
// Lists with more entries than the register path go to `huge`.
__device__ __forceinline__ uint32_t long_entry(const int *__restrict__ tris, const uint32_t *__restrict__ ref,
                                               uint32_t b, uint32_t e, int per)
{
	const uint32_t r = ref[b + e / per], m = e % per;
	if (per == 3)
		return (uint32_t)tris[3 * (r / 3) + m];
	const int2 q = corner_pair(tris, r);
	return (uint32_t)(m ? q.y : q.x);
}

__global__ void k_long_border(const int *__restrict__ tris, uint32_t nv, const uint32_t *__restrict__ start,
                              const uint32_t *__restrict__ ref, int literal, uint8_t *__restrict__ border,
                              uint32_t *__restrict__ huge)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t deg = v < nv ? start[v + 1] - start[v] : 0u;
	const int lane = lane_id(), per = literal ? 3 : 2;
	for (unsigned long long m = __ballot(deg > (uint32_t)CP_MAX); m; m &= m - 1) {
		const int o = __ffsll((long long)m) - 1;
		const uint32_t u = __shfl(v, o), d = __shfl(deg, o), b = start[u], E = d * per;
		if (E > (uint32_t)(LW_K * N2M_WAVE)) {
			if (lane == 0)
				huge[1 + atomicAdd(&huge[0], 1u)] = u;
			continue;
		}
		uint32_t r[LW_K], less[LW_K], eq[LW_K];
#pragma unroll
		for (int k = 0; k < LW_K; k++) {
			const uint32_t i = k * N2M_WAVE + lane;
			r[k] = i < E ? long_entry(tris, ref, b, i, per) : 0xffffffffu;
		}
		wave_rank_eq(r, E, less, eq);
		bool once = false; // this lane holds an entry that occurs once
#pragma unroll
		for (int k = 0; k < LW_K; k++)
			if (k * N2M_WAVE + lane < E && eq[k] == 1) {
				once = true;
				if (literal)
					border[r[k]] = 1;
			}
		if (!literal && __ballot(once) && lane == 0)
			border[u] = 1;
	}
}

// The lists left in `huge`, one workgroup each (a grid-stride loop over the
// count the list kernels left): up to HG_CAP values are sorted in LDS
// (bitonic); a longer list compares every pair from global memory.
constexpr int HG_BS = 1024, HG_CAP = 16384;

__device__ void lds_bitonic(uint32_t *s, uint32_t n, uint32_t npad)
{
	for (uint32_t i = n + threadIdx.x; i < npad; i += HG_BS)
		s[i] = 0xffffffffu;
	__syncthreads();
	for (uint32_t k = 2; k <= npad; k <<= 1)
		for (uint32_t j = k >> 1; j > 0; j >>= 1) {
			for (uint32_t i = threadIdx.x; i < npad; i += HG_BS) {
				const uint32_t x = i ^ j;
				if (x > i) {
					const uint32_t a = s[i], c = s[x];
					if ((a > c) == ((i & k) == 0)) {
						s[i] = c;
						s[x] = a;
					}
				}
			}
			__syncthreads();
		}
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n)
{
	uint32_t p = 1;
	while (p < n)
		p <<= 1;
	return p;
}

__global__ void __launch_bounds__(HG_BS)
k_huge_sort(const int *__restrict__ tris, const uint32_t *__restrict__ start, uint32_t *__restrict__ ref,
            int2 *__restrict__ pairs, const uint32_t *__restrict__ huge)
{
	__shared__ uint32_t s[HG_CAP];
	const uint32_t nh = huge[0];
	for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
		const uint32_t u = huge[1 + h], b = start[u], d = start[u + 1] - b;
		if (d <= (uint32_t)HG_CAP) {
			for (uint32_t i = threadIdx.x; i < d; i += HG_BS)
				s[i] = ref[b + i];
			lds_bitonic(s, d, pow2_at_least(d));
			for (uint32_t i = threadIdx.x; i < d; i += HG_BS) {
				ref[b + i] = s[i];
				pairs[b + i] = corner_pair(tris, s[i]);
			}
		} else { // ranks over a copy of the list kept in pairs[]
			for (uint32_t i = threadIdx.x; i < d; i += HG_BS)
				pairs[b + i] = make_int2((int)ref[b + i], 0);
			__threadfence_block();
			__syncthreads();
			for (uint32_t i = threadIdx.x; i < d; i += HG_BS) {
				const uint32_t x = (uint32_t)pairs[b + i].x;
				uint32_t less = 0;
				for (uint32_t j = 0; j < d; j++)
					less += (uint32_t)pairs[b + j].x < x;
				ref[b + less] = x;
			}
			__threadfence_block();
			__syncthreads();
			for (uint32_t i = threadIdx.x; i < d; i += HG_BS)
				pairs[b + i] = corner_pair(tris, ref[b + i]);
		}
		__syncthreads();
	}
}

__global__ void __launch_bounds__(HG_BS)
k_huge_border(const int *__restrict__ tris, const uint32_t *__restrict__ start, const uint32_t *__restrict__ ref,
              int literal, uint8_t *__restrict__ border, const uint32_t *__restrict__ huge)
{
	__shared__ uint32_t s[HG_CAP];
	__shared__ uint32_t s_once;
	const int per = literal ? 3 : 2;
	const uint32_t nh = huge[0];
	for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
		const uint32_t u = huge[1 + h], b = start[u], E = (start[u + 1] - b) * per;
		if (threadIdx.x == 0)
			s_once = 0;
		__syncthreads();
		if (E <= (uint32_t)HG_CAP) {
			for (uint32_t i = threadIdx.x; i < E; i += HG_BS)
				s[i] = long_entry(tris, ref, b, i, per);
			lds_bitonic(s, E, pow2_at_least(E));
			for (uint32_t i = threadIdx.x; i < E; i += HG_BS) {
				const uint32_t x = s[i];
				if ((i == 0 || s[i - 1] != x) && (i + 1 == E || s[i + 1] != x)) {
					if (literal)
						border[x] = 1;
					else
						s_once = 1;
				}
			}
		} else {
			for (uint32_t i = threadIdx.x; i < E; i += HG_BS) {
				const uint32_t x = long_entry(tris, ref, b, i, per);
				uint32_t n = 0;
				for (uint32_t j = 0; j < E && n < 2; j++)
					n += long_entry(tris, ref, b, j, per) == x;
				if (n == 1) {
					if (literal)
						border[x] = 1;
					else
						s_once = 1;
				}
			}
		}
		__syncthreads();
		if (threadIdx.x == 0 && !literal && s_once)
			border[u] = 1;
		__syncthreads();
	}
}

// vertices with more than 256 references, or more than 256 border entries
// (85 references of 3), plus the count
size_t long_huge_words(size_t nt) { return 2 + 3 * nt / 85; }

void launch_long_sort(const int *tris, uint32_t nv, const uint32_t *start, uint32_t *ref, int2 *pairs,
                      uint32_t *huge, hipStream_t s)
{
	if (!nv)
		return;
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, huge);
	hipLaunchKernelGGL(k_long_sort, dim3(g1(nv)), dim3(256), 0, s, tris, nv, start, ref, pairs, huge);
	hipLaunchKernelGGL(k_huge_sort, dim3(64), dim3(HG_BS), 0, s, tris, start, ref, pairs, (const uint32_t *)huge);
}
void launch_long_border(const int *tris, uint32_t nv, const uint32_t *start, const uint32_t *ref, int literal,
                        uint8_t *border, uint32_t *huge, hipStream_t s)
{
	if (!nv)
		return;
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, huge);
	hipLaunchKernelGGL(k_long_border, dim3(g1(nv)), dim3(256), 0, s, tris, nv, start, ref, literal, border, huge);
	hipLaunchKernelGGL(k_huge_border, dim3(64), dim3(HG_BS), 0, s, tris, start, ref, literal, border,
	                   (const uint32_t *)huge);
}
void launch_csr_count(const int *tris, uint32_t nt, uint32_t *cnt, uint32_t *rank, hipStream_t s,
                      const uint32_t *inv, int *tout)
{
	hipLaunchKernelGGL(k_csr_count, dim3(g1((size_t)3 * nt)), dim3(256), 0, s, tris, nt, cnt, rank, inv, tout);
}
void launch_csr_fill(const int *tris, uint32_t nt, const uint32_t *start, const uint32_t *rank, uint32_t *ref,
                     hipStream_t s)
{
	hipLaunchKernelGGL(k_csr_fill, dim3(g1((size_t)3 * nt)), dim3(256), 0, s, tris, nt, start, rank, ref);
}
void launch_csr_pairs(const int *tris, uint32_t nv, const uint32_t *start, uint32_t *ref, int2 *pairs,
                      hipStream_t s, bool sorted_ref)
{
	hipLaunchKernelGGL(k_csr_pairs, dim3(g1(nv)), dim3(CP_BS), 0, s, tris, nv, start, ref, pairs,
	                   sorted_ref ? 1 : 0);
}
void launch_border(const int *tris, uint32_t nv, const uint32_t *start, const uint32_t *ref, uint8_t *border,
                   hipStream_t s)
{
	hipLaunchKernelGGL(k_border, dim3(g1(nv)), dim3(256), 0, s, tris, nv, start, ref, border);
}
// Window shapes (B vertices per workgroup, D on each side, K pairs
// prefetched per vertex); N2M_HC_SHAPE picks one (default 0).
#define N2M_HC_SHAPES(X) X(0, 1024, 256, 8) X(1, 512, 256, 8) X(2, 1024, 128, 8) X(3, 512, 128, 8)

static int hc_shape()
{
	static const int v = [] { // read once, thread-safe initialisation
		const char *e = getenv("N2M_HC_SHAPE");
		const int k = e ? atoi(e) : 0;
		return k < 0 || k > 3 ? 0 : k;
	}();
	return v;
}

void launch_hc_a(const uint32_t *start, const int2 *pairs, const double *o, const double *p, double *pn,
                 double *bq, uint32_t nv, double alpha, hipStream_t s)
{
	switch (hc_shape()) {
#define X(k, B, D, K)                                                                                              \
	case k:                                                                                                      \
		hipLaunchKernelGGL((k_hc_a<B, D, K>), dim3((nv + B - 1) / B), dim3(B), 0, s, start, pairs, (const d3 *)o,   \
		                   (const d3 *)p, (d3 *)pn, (d3 *)bq, nv, alpha, 1.0 - alpha);                             \
		break;
		N2M_HC_SHAPES(X)
#undef X
	}
}
void launch_hc_b(const uint32_t *start, const int2 *pairs, const double *pn, const double *bq, double *pout,
                 uint32_t nv, double beta, hipStream_t s)
{
	switch (hc_shape()) {
#define X(k, B, D, K)                                                                                              \
	case k:                                                                                                      \
		hipLaunchKernelGGL((k_hc_b<B, D, K>), dim3((nv + B - 1) / B), dim3(B), 0, s, start, pairs, (const d3 *)pn,  \
		                   (const d3 *)bq, (d3 *)pout, nv, beta, 1.0 - beta);                                      \
		break;
		N2M_HC_SHAPES(X)
#undef X
	}
}
void launch_laplace(const uint32_t *start, const int2 *pairs, const double *p, double *pout, uint32_t nv,
                    hipStream_t s)
{
	switch (hc_shape()) {
#define X(k, B, D, K)                                                                                              \
	case k:                                                                                                      \
		hipLaunchKernelGGL((k_laplace<B, D, K>), dim3((nv + B - 1) / B), dim3(B), 0, s, start, pairs,              \
		                   (const d3 *)p, (d3 *)pout, nv);                                                         \
		break;
		N2M_HC_SHAPES(X)
#undef X
	}
}

uint32_t halo_blocks(uint32_t nv) { return (nv + HB - 1) / HB; }
size_t halo_words(uint32_t nv) { return (size_t)halo_blocks(nv) * HCAP; }

void launch_halo_build(const uint32_t *start, const int2 *pairs, uint32_t nv, uint32_t *halo, uint32_t *hcount,
                       uint32_t *slots, uint32_t *overflow, hipStream_t s)
{
	if (nv)
		hipLaunchKernelGGL(k_halo_build, dim3(halo_blocks(nv)), dim3(HB), 0, s, start, pairs, nv, halo, hcount, slots,
		                   overflow);
}
void launch_hc_a_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo, const uint32_t *hcount,
                      const double *o, const double *p, double *pn, double *bq, uint32_t nv, double alpha,
                      hipStream_t s)
{
	if (nv)
		hipLaunchKernelGGL(k_hc_a_h, dim3(halo_blocks(nv)), dim3(HB), 0, s, start, slots, halo, hcount,
		                   (const d3 *)o, (const d3 *)p, (d3 *)pn, (d3 *)bq, nv, alpha, 1.0 - alpha);
}
void launch_hc_b_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo, const uint32_t *hcount,
                      const double *pn, const double *bq, double *pout, uint32_t nv, double beta, hipStream_t s)
{
	if (nv)
		hipLaunchKernelGGL(k_hc_b_h, dim3(halo_blocks(nv)), dim3(HB), 0, s, start, slots, halo, hcount,
		                   (const d3 *)pn, (const d3 *)bq, (d3 *)pout, nv, beta, 1.0 - beta);
}
void launch_laplace_halo(const uint32_t *start, const uint32_t *slots, const uint32_t *halo,
                         const uint32_t *hcount, const double *p, double *pout, uint32_t nv, hipStream_t s)
{
	if (nv)
		hipLaunchKernelGGL(k_laplace_h, dim3(halo_blocks(nv)), dim3(HB), 0, s, start, slots, halo, hcount,
		                   (const d3 *)p, (d3 *)pout, nv);
}
// vertices per two-ring block: N2M_HC_RB = 512, 640 (default) or 768, read
// by each smoothing call
int ring_rb()
{
	const char *e = getenv("N2M_HC_RB");
	const int v = e ? atoi(e) : 640;
	return v == 512 || v == 768 ? v : 640;
}
uint32_t ring_blocks(uint32_t nv, int rb) { return (nv + rb - 1) / rb; }
size_t ring_words(uint32_t nv, int rb) { return (size_t)ring_blocks(nv, rb) * (RROWS - rb); }
size_t ring_pair_words(uint32_t nv, int rb) { return (size_t)ring_blocks(nv, rb) * RPCAP; }
size_t ring_off_words(uint32_t nv, int rb) { return (size_t)ring_blocks(nv, rb) * ROFFW; }
size_t ring_o_rows(uint32_t nv, int rb) { return (size_t)ring_blocks(nv, rb) * R1CAP; }

#define N2M_RING_RBS(X) X(512) X(640) X(768)

void launch_ring_build(const uint32_t *start, const int2 *pairs, uint32_t nv, const double *o, uint32_t *ring,
                       uint32_t *rcnt, uint32_t *slots, uint16_t *roff, uint32_t *rslots, double *o1,
                       uint32_t *overflow, int rbv, hipStream_t s)
{
	if (!nv)
		return;
	switch (rbv) {
#define X(rb)                                                                                                      \
	case rb:                                                                                                     \
		hipLaunchKernelGGL(k_ring_build<rb>, dim3(ring_blocks(nv, rb)), dim3(RT), 0, s, start, pairs, nv,           \
		                   (const d3 *)o, ring, (uint2 *)rcnt, slots, roff, rslots, (d3 *)o1, overflow);           \
		break;
		N2M_RING_RBS(X)
#undef X
	}
}
void launch_hc_ring(const uint32_t *start, const uint32_t *slots, const uint16_t *roff, const uint32_t *rslots,
                    const uint32_t *ring, const uint32_t *rcnt, const double *o, const double *o1, const double *p,
                    double *pout, uint32_t nv, double alpha, double beta, int rbv, hipStream_t s)
{
	if (!nv)
		return;
	switch (rbv) {
#define X(rb)                                                                                                      \
	case rb:                                                                                                     \
		hipLaunchKernelGGL(k_hc_ring<rb>, dim3(ring_blocks(nv, rb)), dim3(RT), 0, s, start, slots, roff, rslots, \
		                   ring, (const uint2 *)rcnt, (const d3 *)o, (const d3 *)o1, (const d3 *)p, (d3 *)pout, nv, \
		                   alpha, 1.0 - alpha, beta, 1.0 - beta);                                                  \
		break;
		N2M_RING_RBS(X)
#undef X
	}
}

void launch_hc_load(const double *src, const uint32_t *perm, uint32_t n, double *dst, hipStream_t s)
{
	hipLaunchKernelGGL(k_hc_load, dim3(g1(n)), dim3(256), 0, s, (const d3 *)src, perm, n, (d3 *)dst);
}
void launch_hc_store(const double *p, const double *o, const uint8_t *border, const uint32_t *perm, uint32_t n,
                  double *dst, hipStream_t s)
{
	hipLaunchKernelGGL(k_hc_store, dim3(g1(n)), dim3(256), 0, s, (const d3 *)p, (const d3 *)o, border, perm, n,
	                   (d3 *)dst);
}

// ---------------------------------------------------------------------------
// Vertex relabelling for the smoothing gathers: vertices are renumbered in
// Morton (Z-curve) order of their position so that a vertex's neighbours sit
// at nearby addresses.  Triangle order is untouched, so every vertex still
// sums its references in ascending triangle order: the result is unchanged.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_pts_bbox(const double *__restrict__ pts, uint32_t n, uint32_t *bb)
{
	__shared__ uint32_t s[6][4];
	float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
	float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		for (int k = 0; k < 3; k++) {
			float v = (float)pts[3 * (size_t)i + k];
			lo[k] = fminf(lo[k], v);
			hi[k] = fmaxf(hi[k], v);
		}
	for (int k = 0; k < 3; k++) {
		uint32_t a = wave_min_u32(f2ord(lo[k])), b = wave_max_u32(f2ord(hi[k]));
		if (lane_id() == 0) {
			s[k][wave_id()] = a;
			s[3 + k][wave_id()] = b;
		}
	}
	__syncthreads();
	if (threadIdx.x < 6) {
		int k = threadIdx.x;
		uint32_t r = s[k][0];
		for (int w = 1; w < 4; w++)
			r = k < 3 ? min(r, s[k][w]) : max(r, s[k][w]);
		if (k < 3)
			atomicMin(&bb[k], r);
		else
			atomicMax(&bb[k], r);
	}
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
	v &= 0x3ffu;
	v = (v | (v << 16)) & 0x030000ffu;
	v = (v | (v << 8)) & 0x0300f00fu;
	v = (v | (v << 4)) & 0x030c30c3u;
	v = (v | (v << 2)) & 0x09249249u;
	return v;
}

__global__ void k_morton(const double *__restrict__ pts, uint32_t n, const uint32_t *__restrict__ bb,
                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t key = 0;
	for (int k = 0; k < 3; k++) {
		float lo = ord2f(bb[k]), hi = ord2f(bb[3 + k]);
		float ext = hi - lo;
		float f = ext > 0.f ? ((float)pts[3 * (size_t)i + k] - lo) * (1023.0f / ext) : 0.f;
		uint32_t q = f > 0.f ? (f < 1023.f ? (uint32_t)f : 1023u) : 0u; // NaN -> 0
		key |= spread10(q) << k;
	}
	keys[i] = key >> 6; // 8 bits per axis: 24-bit keys, three radix passes
	vals[i] = i;
}

__global__ void k_invert_perm(const uint32_t *__restrict__ perm, uint32_t n, uint32_t *__restrict__ inv)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		inv[perm[i]] = i;
}

// 1 if any triangle repeats a vertex (then the general border rule applies)
__global__ void k_tris_repeat(const int *__restrict__ tris, uint32_t nt, uint32_t *flag)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	int a = tris[3 * i], b = tris[3 * i + 1], c = tris[3 * i + 2];
	if (a == b || b == c || a == c)
		*flag = 1;
}

// Border for triangles with three distinct vertices: v is border when some
// edge (v,a) lies in exactly one of v's triangles -- equivalent to the
// update_mesh rule (a vertex in a single triangle has two such edges).
__global__ void k_border_pairs(const uint32_t *__restrict__ start, const int2 *__restrict__ pairs, uint32_t nv,
                               uint8_t *__restrict__ border)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	uint32_t b = start[v], e = start[v + 1];
	uint8_t r = 0;
	constexpr int RB = 12;
	if (e - b <= RB) { // the common case: the pair ids held in registers
		int id[2 * RB];
#pragma unroll
		for (int k = 0; k < RB; k++) {
			const int2 q = b + k < e ? pairs[b + k] : make_int2(-1, -1); // ids are >= 0
			id[2 * k] = q.x;
			id[2 * k + 1] = q.y;
		}
#pragma unroll
		for (int i = 0; i < 2 * RB; i++) {
			uint32_t cnt = 0;
#pragma unroll
			for (int k = 0; k < 2 * RB; k++)
				cnt += id[k] == id[i];
			r |= (id[i] >= 0 && cnt == 1);
		}
		border[v] = r;
		return;
	}
	if (e - b > (uint32_t)CP_MAX) {
		border[v] = 0; // the high-valence pass marks it
		return;
	}
	for (uint32_t i = b; i < e && !r; i++) {
		int2 q = pairs[i];
		for (int s = 0; s < 2 && !r; s++) {
			int a = s ? q.y : q.x;
			uint32_t cnt = 0;
			for (uint32_t j = b; j < e; j++) {
				int2 w = pairs[j];
				cnt += (w.x == a) + (w.y == a);
			}
			r = cnt == 1;
		}
	}
	border[v] = r;
}

void launch_pts_bbox(const double *pts, uint32_t n, uint32_t *bb, hipStream_t s)
{
	unsigned b = g1(n);
	b = b > 1024 ? 1024 : (b < 1 ? 1 : b);
	hipLaunchKernelGGL(k_pts_bbox, dim3(b), dim3(256), 0, s, pts, n, bb);
}
void launch_morton(const double *pts, uint32_t n, const uint32_t *bb, uint32_t *keys, uint32_t *vals, hipStream_t s)
{
	hipLaunchKernelGGL(k_morton, dim3(g1(n)), dim3(256), 0, s, pts, n, bb, keys, vals);
}
void launch_invert_perm(const uint32_t *perm, uint32_t n, uint32_t *inv, hipStream_t s)
{
	hipLaunchKernelGGL(k_invert_perm, dim3(g1(n)), dim3(256), 0, s, perm, n, inv);
}
void launch_tris_repeat(const int *tris, uint32_t nt, uint32_t *flag, hipStream_t s)
{
	hipLaunchKernelGGL(k_tris_repeat, dim3(g1(nt)), dim3(256), 0, s, tris, nt, flag);
}
void launch_border_pairs(const uint32_t *start, const int2 *pairs, uint32_t nv, uint8_t *border, hipStream_t s)
{
	hipLaunchKernelGGL(k_border_pairs, dim3(g1(nv)), dim3(256), 0, s, start, pairs, nv, border);
}

// ---------------------------------------------------------------------------
// apply_sform, src/meshify.c:1021-1045 (positions only; the winding flip is
// decided on the host from the same float sums).
// ---------------------------------------------------------------------------
__global__ void k_sform(double *__restrict__ p, uint32_t np, float x0, float x1, float x2, float x3, float y0,
                        float y1, float y2, float y3, float z0, float z1, float z2, float z3)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= np)
		return;
	double vx = p[3 * (size_t)i], vy = p[3 * (size_t)i + 1], vz = p[3 * (size_t)i + 2];
	p[3 * (size_t)i + 0] = (vx * x0) + (vy * x1) + (vz * x2) + x3;
	p[3 * (size_t)i + 1] = (vx * y0) + (vy * y1) + (vz * y2) + y3;
	p[3 * (size_t)i + 2] = (vx * z0) + (vy * z1) + (vz * z2) + z3;
}

__global__ void k_swap_winding(int *__restrict__ t, uint32_t nt)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nt)
		return;
	int a = t[3 * i];
	t[3 * i] = t[3 * i + 1];
	t[3 * i + 1] = a;
}

void launch_sform(double *p, uint32_t np, const float *sx, const float *sy, const float *sz, hipStream_t s)
{
	hipLaunchKernelGGL(k_sform, dim3(g1(np)), dim3(256), 0, s, p, np, sx[0], sx[1], sx[2], sx[3], sy[0], sy[1],
	                   sy[2], sy[3], sz[0], sz[1], sz[2], sz[3]);
}
void launch_swap_winding(int *t, uint32_t nt, hipStream_t s)
{
	hipLaunchKernelGGL(k_swap_winding, dim3(g1(nt)), dim3(256), 0, s, t, nt);
}

// ---------------------------------------------------------------------------
// Batched atlas weld: every label mesh is welded on its own, as
// unify_vertices() does per label (src/meshify.c:45-106), in one segmented
// pass over the union of the label meshes.  Vertices of task t occupy
// [vbase[t], vbase[t+1]); keys are distances to the task's own vertex 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int task_of(const uint32_t *__restrict__ vbase, int ntasks, uint32_t i)
{
	int lo = 0, hi = ntasks - 1; // largest t with vbase[t] <= i
	while (lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		if (vbase[mid] <= i)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__global__ void k_atlas_weld_keys(const double *__restrict__ pts, uint32_t n, const uint32_t *__restrict__ vbase,
                                  int ntasks, uint32_t *keys, uint32_t *vals, uint32_t *vtask)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	int t = task_of(vbase, ntasks, i);
	float d = (float)dist3(pts + (size_t)vbase[t] * 3, pts + (size_t)i * 3);
	keys[i] = weld_key(d);
	vals[i] = i;
	vtask[i] = (uint32_t)t;
}

__global__ void k_atlas_weld_segflags(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ tk, uint32_t n,
                                      uint32_t *flags)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float tol = 0.00001f;
	uint32_t f = 1;
	if (i > 0 && tk[i] == tk[i - 1])
		f = (ord2f(keys[i]) - ord2f(keys[i - 1])) >= tol;
	flags[i] = f;
}

// first class of each task (the seed at its first sorted position)
__global__ void k_atlas_class_base(const uint32_t *__restrict__ tk, const uint32_t *__restrict__ segid,
                                   const uint32_t *__restrict__ seedbase, uint32_t n, uint32_t *cbase)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (i == 0 || tk[i] != tk[i - 1])
		cbase[tk[i]] = seedbase[segid[i]];
}

// Per task, from the class bases the assign pass left: its class count
// (next non-empty task's base, or the total, minus its own), and whether
// its weld merged nothing (class count = vertex count).  nextne[t]: the
// next task with vertices, ntasks for none.  An empty task takes the next
// base and count 0.
__global__ void k_atlas_weld_tasks(uint32_t *__restrict__ cbase, const uint32_t *__restrict__ vbase,
                                   const uint32_t *__restrict__ nextne, int ntasks,
                                   const uint32_t *__restrict__ d_nnew, uint32_t *__restrict__ ncls,
                                   uint8_t *__restrict__ ident)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= ntasks)
		return;
	const uint32_t nx = nextne[t];
	const uint32_t next = nx < (uint32_t)ntasks ? cbase[nx] : *d_nnew;
	const uint32_t nv = vbase[t + 1] - vbase[t];
	if (!nv) {
		cbase[t] = next; // no thread reads an empty task's base
		ncls[t] = 0;
		ident[t] = 0;
		return;
	}
	const uint32_t c = next - cbase[t];
	ncls[t] = c;
	ident[t] = c == nv;
}

// a task whose weld merged nothing keeps its vertex order (src/meshify.c:82-87)
__global__ void k_atlas_weld_identity(uint32_t n, const uint32_t *__restrict__ vtask,
                                      const uint32_t *__restrict__ vbase, const uint32_t *__restrict__ cbase,
                                      const uint8_t *__restrict__ ident, int *__restrict__ cls, int *__restrict__ rep)
{
	uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= n)
		return;
	uint32_t t = vtask[v];
	if (!ident[t])
		return;
	int c = (int)(cbase[t] + (v - vbase[t]));
	cls[v] = c;
	rep[c] = (int)v;
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t n,
                             uint32_t *__restrict__ dst)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = src[idx[i]];
}

__global__ void k_add_i32(const int *__restrict__ src, size_t n, int add, int *__restrict__ dst)
{
	size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = src[i] + add;
}

void launch_atlas_weld_keys(const double *pts, uint32_t n, const uint32_t *vbase, int ntasks, uint32_t *keys,
                            uint32_t *vals, uint32_t *vtask, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_atlas_weld_keys, dim3(g1(n)), dim3(256), 0, s, pts, n, vbase, ntasks, keys, vals,
		                   vtask);
}
void launch_atlas_weld_segflags(const uint32_t *keys, const uint32_t *tk, uint32_t n, uint32_t *flags,
                                hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_atlas_weld_segflags, dim3(g1(n)), dim3(256), 0, s, keys, tk, n, flags);
}
void launch_atlas_class_base(const uint32_t *tk, const uint32_t *segid, const uint32_t *seedbase, uint32_t n,
                             uint32_t *cbase, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_atlas_class_base, dim3(g1(n)), dim3(256), 0, s, tk, segid, seedbase, n, cbase);
}
void launch_atlas_weld_tasks(uint32_t *cbase, const uint32_t *vbase, const uint32_t *nextne, int ntasks,
                             const uint32_t *d_nnew, uint32_t *ncls, uint8_t *ident, hipStream_t s)
{
	if (ntasks)
		hipLaunchKernelGGL(k_atlas_weld_tasks, dim3((ntasks + 255) / 256), dim3(256), 0, s, cbase, vbase, nextne,
		                   ntasks, d_nnew, ncls, ident);
}
void launch_atlas_weld_identity(uint32_t n, const uint32_t *vtask, const uint32_t *vbase, const uint32_t *cbase,
                                const uint8_t *ident, int *cls, int *rep, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_atlas_weld_identity, dim3(g1(n)), dim3(256), 0, s, n, vtask, vbase, cbase, ident, cls,
		                   rep);
}
void launch_gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t n, uint32_t *dst, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_gather_u32, dim3(g1(n)), dim3(256), 0, s, src, idx, n, dst);
}
void launch_add_i32(const int *src, size_t n, int add, int *dst, hipStream_t s)
{
	if (n)
		hipLaunchKernelGGL(k_add_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, add, dst);
}

} // namespace n2m
