// n2m_device.cuh -- device-side helpers shared by the gfx950 kernels.
//
// Wave64 everywhere: shuffles and ballots assume 64 lanes (CDNA).  All code
// is compiled with -ffp-contract=off so that float/double expressions keep
// the reference's separate roundings (no fused multiply-add).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define N2M_WAVE 64

namespace n2m {

__device__ __forceinline__ int lane_id() { return threadIdx.x & (N2M_WAVE - 1); }
// set bits of the wave mask m at lanes below this one
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m)
{
	return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ int wave_id() { return threadIdx.x / N2M_WAVE; }

// Order-preserving map float <-> uint32 for atomic min/max on floats.
__device__ __forceinline__ uint32_t f2ord(float f)
{
	uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t o)
{
	uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
	float f;
	__builtin_memcpy(&f, &u, 4);
	return f;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
	const int lane = lane_id();
#pragma unroll
	for (int d = 1; d < N2M_WAVE; d <<= 1) {
		uint32_t t = __shfl_up(v, d, N2M_WAVE);
		if (lane >= d)
			v += t;
	}
	return v;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1)
		v += __shfl_xor(v, d, N2M_WAVE);
	return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1)
		v = max(v, (uint32_t)__shfl_xor(v, d, N2M_WAVE));
	return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1)
		v = min(v, (uint32_t)__shfl_xor(v, d, N2M_WAVE));
	return v;
}
__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1)
		v = max(v, __shfl_xor(v, d, N2M_WAVE));
	return v;
}
__device__ __forceinline__ int wave_min_i32(int v)
{
#pragma unroll
	for (int d = N2M_WAVE / 2; d > 0; d >>= 1)
		v = min(v, __shfl_xor(v, d, N2M_WAVE));
	return v;
}

// Exclusive block scan of one value per thread; BS threads (multiple of 64).
// `scratch` must hold BS/64 + 1 uint32 in shared memory.  Returns the
// exclusive prefix; *total receives the block sum.  Contains barriers.
template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *scratch, uint32_t *total)
{
	constexpr int NW = BS / N2M_WAVE;
	uint32_t incl = wave_incl_scan(v);
	if (lane_id() == N2M_WAVE - 1)
		scratch[wave_id()] = incl;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t run = 0;
		for (int w = 0; w < NW; w++) {
			uint32_t t = scratch[w];
			scratch[w] = run;
			run += t;
		}
		scratch[NW] = run;
	}
	__syncthreads();
	uint32_t r = scratch[wave_id()] + incl - v;
	*total = scratch[NW];
	__syncthreads();
	return r;
}

template <int BS> __device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *scratch)
{
	constexpr int NW = BS / N2M_WAVE;
	v = wave_sum(v);
	if (lane_id() == 0)
		scratch[wave_id()] = v;
	__syncthreads();
	uint32_t r = 0;
	for (int w = 0; w < NW; w++)
		r += scratch[w];
	__syncthreads();
	return r;
}

// Global union-find on voxel indices: L[x] <= x always, roots satisfy
// L[x] == x, and union hooks the larger root under the smaller, so every
// component's root is its smallest voxel index.  While unions are in flight
// (other workgroups writing L) the parent reads go to L2 (agent-scope atomic
// loads) so a stale L1 line cannot hide a hook.
__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Find with path halving: every visited non-root is re-pointed to its
// grandparent.  That is always an ancestor in the same tree with a smaller
// index, and only roots are ever hooked, so the stores cannot lose a union.
__device__ __forceinline__ uint32_t uf_find_live(uint32_t *L, uint32_t x)
{
	uint32_t p = ld_agent(&L[x]);
	while (p != x) {
		const uint32_t gp = ld_agent(&L[p]);
		if (gp != p)
			__hip_atomic_store(&L[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		x = p;
		p = gp;
	}
	return x;
}
__device__ __forceinline__ void uf_union(uint32_t *L, uint32_t a, uint32_t b)
{
	for (;;) {
		a = uf_find_live(L, a);
		b = uf_find_live(L, b);
		if (a == b)
			return;
		if (a < b) {
			uint32_t t = a;
			a = b;
			b = t;
		}
		uint32_t old = atomicMin(&L[a], b);
		if (old == a)
			return;
		a = old;
	}
}

} // namespace n2m
