// k_ccl.hip -- connected components for bwlabel() (src/bwlabel.c:429-543).
//
// The reference labels with a two-pass Thurfjell scan and numbers the final
// labels by first occurrence in x-fastest order (translate_labels,
// src/bwlabel.c:249-277).  Here: union-find whose root is always the
// smallest voxel index of its component, so "first occurrence" order is root
// order and the largest-cluster tie rule (lowest label wins,
// src/bwlabel.c:463) is "smallest root wins".
//
//  1. k_ccl_local: a workgroup labels the runs of a 512x4x4 tile in LDS
//     (16 bit words per tile row) and writes each run's tile-local root to
//     L; it also clears AUX at every tile-local root (every global root is
//     one of them).
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

constexpr int CC_KW = 16, CC_TY = 4, CC_TZ = 4, CC_BS = CC_KW * CC_TY * CC_TZ; // 256
constexpr uint32_t NONE = 0xffffffffu;

// Run-based labelling.  A row word is split into maximal runs of set bits
// (runs are also cut at word boundaries); a run is named by its first bit.
// Two runs of neighbouring rows are connected when their bits overlap after
// the x shift of a backward neighbour offset; every contiguous segment of
// the overlap needs one union.  Backward neighbours of a row (y,z), with the
// x shifts they allow (src/bwlabel.c:94-131,185-210):
//   6:  (y-1,z) {0}, (y,z-1) {0}, and x-1 in the same row
//   18: (y-1,z) {-1,0,1}, (y,z-1) {-1,0,1}, (y-1,z-1) {0}, (y+1,z-1) {0}
//   26: 18 plus (y-1,z-1) {-1,1}, (y+1,z-1) {-1,1}
struct NbrRow {
	int8_t dy, dz, dxlo, dxhi;
};
__constant__ NbrRow c_nbr6[2] = {{-1, 0, 0, 0}, {0, -1, 0, 0}};
__constant__ NbrRow c_nbr18[4] = {{-1, 0, -1, 1}, {0, -1, -1, 1}, {-1, -1, 0, 0}, {1, -1, 0, 0}};
__constant__ NbrRow c_nbr26[4] = {{-1, 0, -1, 1}, {0, -1, -1, 1}, {-1, -1, -1, 1}, {1, -1, -1, 1}};

__device__ __forceinline__ const NbrRow *nbr_rows(int conn, int *n)
{
	if (conn == 6) {
		*n = 2;
		return c_nbr6;
	}
	*n = 4;
	return conn == 18 ? c_nbr18 : c_nbr26;
}

__device__ __forceinline__ uint32_t run_starts(uint32_t w) { return w & ~(w << 1); }

// first bit of the run of w (starts = run_starts(w)) that contains bit b
__device__ __forceinline__ int run_of(uint32_t starts, int b)
{
	uint32_t m = b == 31 ? starts : (starts & ((2u << b) - 1u));
	return 31 - __clz(m);
}

// length of the run of w starting at bit s
__device__ __forceinline__ int run_len(uint32_t w, int s)
{
	uint32_t t = ~(w >> s);
	return t ? __ffs(t) - 1 : 32 - s;
}

__device__ __forceinline__ uint32_t run_mask(uint32_t w, int s)
{
	int l = run_len(w, s);
	return (l == 32 ? 0xffffffffu : ((1u << l) - 1u)) << s;
}

// index, among the runs of a word, of the run that contains bit `bit`
__device__ __forceinline__ uint32_t run_index(uint32_t starts, int bit)
{
	return __popc(bit == 31 ? starts : (starts & ((2u << bit) - 1u))) - 1u;
}

// Runs are named by dense ids in voxel order: rb[word] = number of runs in
// the words before it (exclusive scan of the per-word counts), and a run's id
// is rb[word] + its index in the word.  The smallest id of a component is its
// first run in x-fastest order.
struct Geo {
	int nx, ny, nz, W;
	const uint32_t *rb;
	__device__ size_t widx(int wx, int y, int z) const
	{
		return (size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * z);
	}
	__device__ uint32_t word(const uint32_t *bits, int wx, int y, int z, int invert) const
	{
		if (wx < 0 || wx >= W || y < 0 || y >= ny || z < 0 || z >= nz)
			return 0u;
		uint32_t w = bits[(size_t)wx + (size_t)W * ((size_t)y + (size_t)ny * z)];
		if (invert) {
			int rem = nx - wx * 32;
			uint32_t valid = rem >= 32 ? 0xffffffffu : ((1u << rem) - 1u);
			w = ~w & valid;
		}
		return w;
	}
};

// LDS labels: parent node in the low LAB_BITS bits (nodes < CC_RUNS); once
// the unions of a tile are done, the voxel count of each tile component is
// accumulated above them at its root
constexpr uint32_t LAB_BITS = 12, LAB_MASK = (1u << LAB_BITS) - 1u;

__device__ __forceinline__ uint32_t lds_find(const uint32_t *lab, uint32_t x)
{
	uint32_t p = lab[x] & LAB_MASK;
	while (p != x) {
		x = p;
		p = lab[x] & LAB_MASK;
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

// Visit every (run of w, run of u) pair connected by the x shift dx
// (u is the neighbour row word at the same word index; up / un are the
// neighbour row words at word index -1 / +1, used for the bit that crosses
// the word boundary).  f(bit_in_w, word_offset_of_u, bit_in_u).
template <typename F>
__device__ __forceinline__ void for_links(uint32_t w, uint32_t u, uint32_t up, uint32_t un, int dx, F f)
{
	uint32_t m;
	if (dx == 0)
		m = w & u;
	else if (dx < 0)
		m = w & (u << 1); // bit b of w touches bit b-1 of u
	else
		m = w & (u >> 1); // bit b of w touches bit b+1 of u
	for (uint32_t sg = m & ~(m << 1); sg; sg &= sg - 1) {
		int b = __ffs(sg) - 1;
		f(b, 0, b + dx);
	}
	if (dx < 0 && (w & 1u) && (up >> 31))
		f(0, -1, 31);
	if (dx > 0 && (w >> 31) && (un & 1u))
		f(31, 1, 0);
}

// per word: the number of runs (starts) of the predicate; cnt[nw] = 0, so
// that the exclusive scan of cnt[0, nw] ends with the total
__global__ void __launch_bounds__(256)
k_run_count(const uint32_t *__restrict__ bits, Geo g, int invert, uint32_t *__restrict__ cnt)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t nw = (uint32_t)g.W * g.ny * g.nz;
	if (t > nw)
		return;
	uint32_t c = 0;
	if (t < nw) {
		const uint32_t row = t / g.W;
		c = __popc(run_starts(g.word(bits, (int)(t % g.W), (int)(row % g.ny), (int)(row / g.ny), invert)));
	}
	cnt[t] = c;
}

// Pass 1: label runs inside a tile of 16 words x 4 rows x 4 slices in LDS.
// A run's tile-local node is its rank among the tile's runs in (slice, row,
// word, bit) order -- the order of their global ids -- so the smallest node
// of a tile component is its smallest run id.  A tile holds at most
// 256 words x 16 runs.  L[run] = the tile-local root run.  AUX: with counts
// (mode 0), AUX[run] = TILE_ROOT | the tile component's voxel count at each
// tile-local root and 0 at every other run; without (mode 1), AUX[root] = 0
// at each tile-local root.
constexpr int CC_RUNS = CC_BS * 16;
constexpr uint32_t TILE_ROOT = 0x80000000u;
static_assert((uint32_t)CC_RUNS <= (1u << LAB_BITS), "node ids fit the label bits");
static_assert((uint64_t)CC_KW * 32 * CC_TY * CC_TZ < (1ull << (31 - LAB_BITS)), "tile counts fit above them");

// word of the tile (flat index r * CC_KW + k) holding tile-local run `node`:
// the last word whose first run is at or before it
__device__ __forceinline__ int tile_word_of(const uint32_t *pre, uint32_t node)
{
	int lo = 0, hi = CC_BS - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (pre[mid] <= node)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__global__ void __launch_bounds__(CC_BS)
k_ccl_local(const uint32_t *__restrict__ bits, Geo g, int conn, int invert, uint32_t *__restrict__ L,
            uint32_t *__restrict__ AUX, int mode)
{
	__shared__ uint32_t lab[CC_RUNS];
	__shared__ uint32_t s_w[CC_TY * CC_TZ][CC_KW];
	__shared__ uint32_t s_rb[CC_TY * CC_TZ][CC_KW];
	__shared__ uint32_t s_pre[CC_TY * CC_TZ][CC_KW]; // first tile-local node of each word
	__shared__ uint32_t scratch[CC_BS / N2M_WAVE + 1];
	const int k = threadIdx.x % CC_KW, r = threadIdx.x / CC_KW; // word in tile, row in tile
	const int ry = r % CC_TY, rz = r / CC_TY;
	const int wx = blockIdx.x * CC_KW + k, y = blockIdx.y * CC_TY + ry, z = blockIdx.z * CC_TZ + rz;
	const uint32_t w = g.word(bits, wx, y, z, invert);
	s_w[r][k] = w;
	s_rb[r][k] = w ? g.rb[g.widx(wx, y, z)] : 0u;
	const uint32_t st = run_starts(w);
	uint32_t tot;
	const uint32_t pre = block_excl_scan<CC_BS>(__popc(st), scratch, &tot); // thread order = (r, k)
	s_pre[r][k] = pre;
	for (uint32_t j = 0, n = __popc(st); j < n; j++)
		lab[pre + j] = pre + j;
	__syncthreads();
	// x-1 across the word boundary inside the tile
	if (k > 0 && (w & 1u) && (s_w[r][k - 1] >> 31))
		lds_union(lab, pre, s_pre[r][k - 1] + run_index(run_starts(s_w[r][k - 1]), 31));
	int nn;
	const NbrRow *nb = nbr_rows(conn, &nn);
	for (int i = 0; i < nn; i++) {
		const int ny2 = ry + nb[i].dy, nz2 = rz + nb[i].dz;
		if (ny2 < 0 || ny2 >= CC_TY || nz2 < 0 || nz2 >= CC_TZ)
			continue; // another tile: merge pass
		const int r2 = ny2 + CC_TY * nz2;
		const uint32_t u = s_w[r2][k];
		const uint32_t up = k > 0 ? s_w[r2][k - 1] : 0u, un = k + 1 < CC_KW ? s_w[r2][k + 1] : 0u;
		for (int dx = nb[i].dxlo; dx <= nb[i].dxhi; dx++)
			for_links(w, u, up, un, dx, [&](int bw, int wo, int bu) {
				uint32_t uw = wo == 0 ? u : (wo < 0 ? up : un);
				lds_union(lab, pre + run_index(st, bw), s_pre[r2][k + wo] + run_index(run_starts(uw), bu));
			});
	}
	__syncthreads();
	if (mode == 0) { // voxel count of each tile component, at its root
		uint32_t j = 0;
		for (uint32_t t = st; t; t &= t - 1, j++)
			atomicAdd(&lab[lds_find(lab, pre + j)], (uint32_t)run_len(w, __ffs(t) - 1) << LAB_BITS);
		__syncthreads();
	}
	uint32_t gi = s_rb[r][k];
	for (uint32_t j = 0, n = __popc(st); j < n; j++, gi++) {
		const uint32_t root = lds_find(lab, pre + j);
		const int rw = tile_word_of(&s_pre[0][0], root);
		const uint32_t gr = (&s_rb[0][0])[rw] + (root - (&s_pre[0][0])[rw]);
		L[gi] = gr;
		if (mode == 0)
			AUX[gi] = gr == gi ? TILE_ROOT | (lab[root] >> LAB_BITS) : 0u;
		else if (gr == gi)
			AUX[gi] = 0;
	}
}

// Pass 2 links runs of different tiles through the global union-find.
// Along a tile face most links join the same two tile components, so each
// link first takes the runs' current parents (ancestors: linking them links
// the runs) and a workgroup keeps the pairs it has linked in a small LDS
// hash set; a pair already there is skipped.
constexpr int CC_SEEN = 512;

__device__ __forceinline__ void ccl_link(uint32_t *L, unsigned long long *seen, uint32_t a, uint32_t b)
{
	uint32_t pa = ld_agent(&L[a]), pb = ld_agent(&L[b]);
	if (pa == pb)
		return;
	if (pa > pb) {
		const uint32_t t = pa;
		pa = pb;
		pb = t;
	}
	const unsigned long long key = ((unsigned long long)pa << 32) | pb; // never ~0: ids < 2^32 - 1
	uint32_t h = (pa * 0x9E3779B1u ^ pb * 0x85EBCA77u) >> 23;         // 9 bits: CC_SEEN slots
	for (int probe = 0; probe < 8; probe++) {
		const unsigned long long old = atomicCAS(&seen[h], ~0ull, key);
		if (old == key)
			return; // another thread of the workgroup links this pair
		if (old == ~0ull)
			break; // inserted: this thread links it
		h = (h + 1) & (CC_SEEN - 1);
	}
	uf_union(L, pa, pb);
}

__global__ void __launch_bounds__(CC_BS)
k_ccl_merge(const uint32_t *__restrict__ bits, Geo g, int conn, int invert, uint32_t *__restrict__ L)
{
	__shared__ unsigned long long s_seen[CC_SEEN];
	for (int i = threadIdx.x; i < CC_SEEN; i += CC_BS)
		s_seen[i] = ~0ull;
	__syncthreads();
	const int k = threadIdx.x % CC_KW, r = threadIdx.x / CC_KW;
	const int ry = r % CC_TY, rz = r / CC_TY;
	const int wx = blockIdx.x * CC_KW + k, y = blockIdx.y * CC_TY + ry, z = blockIdx.z * CC_TZ + rz;
	const uint32_t w = g.word(bits, wx, y, z, invert);
	if (!w)
		return;
	const uint32_t st = run_starts(w);
	const size_t wi = g.widx(wx, y, z);
	const uint32_t rb0 = g.rb[wi];
	// x-1 across a tile boundary
	if (k == 0 && wx > 0 && (w & 1u)) {
		uint32_t pw = g.word(bits, wx - 1, y, z, invert);
		if (pw >> 31)
			ccl_link(L, s_seen, rb0, g.rb[wi - 1] + run_index(run_starts(pw), 31));
	}
	int nn;
	const NbrRow *nb = nbr_rows(conn, &nn);
	for (int i = 0; i < nn; i++) {
		const int y2 = y + nb[i].dy, z2 = z + nb[i].dz;
		const int ny2 = ry + nb[i].dy, nz2 = rz + nb[i].dz;
		const bool other_tile_row = ny2 < 0 || ny2 >= CC_TY || nz2 < 0 || nz2 >= CC_TZ;
		if (y2 < 0 || y2 >= g.ny || z2 < 0 || z2 >= g.nz)
			continue;
		// a row inside the tile links across the tile only through the
		// neighbouring words of the tile's first and last word
		if (!other_tile_row && k != 0 && k != CC_KW - 1)
			continue;
		const uint32_t u = g.word(bits, wx, y2, z2, invert);
		const uint32_t up = g.word(bits, wx - 1, y2, z2, invert), un = g.word(bits, wx + 1, y2, z2, invert);
		for (int dx = nb[i].dxlo; dx <= nb[i].dxhi; dx++)
			for_links(w, u, up, un, dx, [&](int bw, int wo, int bu) {
				// same tile unless the row or the word lies outside it
				bool other = other_tile_row || (wo < 0 && k == 0) || (wo > 0 && k == CC_KW - 1);
				if (!other)
					return;
				uint32_t uw = wo == 0 ? u : (wo < 0 ? up : un);
				ccl_link(L, s_seen, rb0 + run_index(st, bw),
				         g.rb[g.widx(wx + wo, y2, z2)] + run_index(run_starts(uw), bu));
			});
	}
}

// Pass 3, per run: L[run] = root.  mode 0: voxel count per root into
// AUX[root] and the root list; mode 1: AUX[root] = 1 when the run touches a
// volume face, and the number of roots.  The counts: each tile component's
// size is already at its tile-local root (pass 1), which is the global root
// for a component inside one tile; a tile root that the merge hooked adds
// its size to its global root.  So a mask of millions of small components
// costs no atomics here (one per run had cost about 0.8 ms at 512^3).  The
// additions are summed per thread while consecutive ones share a root, then
// per wave, then per workgroup.
//
// The runs of a tile point at a few parents (their tile-local roots, or
// another tile's run where the merge hooked a tile root), so the parents
// are collected in an LDS table first and each distinct parent is walked to
// its root by one thread; the runs then take their root from the table.
// Without it every run of a giant component walks the same long chain of
// hooked tile roots at once.
constexpr int CC_MEMO = CC_BS, CC_PROBE = 16;

__device__ __forceinline__ uint32_t memo_hash(uint32_t p) { return (p * 0x9E3779B1u) >> 24; } // 8 bits

// The flatten pass reads and writes L with plain (cached) accesses.  It
// starts after the merge pass has ended, so every value a read can return --
// the post-merge parent, or a root written by this pass, possibly seen late
// through another XCD's L2 -- is an ancestor of the node, and a node that
// reads as its own parent was a root when the merge ended, which makes it a
// root for good.  Device-scope accesses to L cost about 0.8 ms on a 512^3
// mask with one giant component; plain ones stay in the L2.

// root of x0 (all unions done); the walked path is then pointed at the
// root.  Only the root is ever written here: every run's final value must be
// its root, and a walker that pointed a node at an older ancestor (path
// splitting) could overwrite a finished node's root.
__device__ __forceinline__ uint32_t ccl_root(uint32_t *L, uint32_t x0)
{
	uint32_t root = x0;
	for (uint32_t p = L[root]; p != root; p = L[root])
		root = p;
	for (uint32_t x = x0; x != root;) {
		const uint32_t nx = L[x];
		if (nx != root)
			L[x] = root;
		x = nx;
	}
	return root;
}

// Between merge and flatten: the distinct parents of each tile's runs (their
// tile roots, hooked by the merge into chains that can cross the volume) are
// walked to their roots with path splitting, every node on the way pointed
// at its grandparent.  Walkers that start on one long chain at the same time
// shorten it for each other, so the chain is halved round after round
// instead of walked in full by each of them; the flatten pass then finds
// every root in a few steps.  No value here needs to be final, which is what
// makes splitting safe (the flatten pass writes only roots).
__global__ void __launch_bounds__(CC_BS)
k_ccl_jump(const uint32_t *__restrict__ bits, Geo g, int invert, uint32_t *__restrict__ L)
{
	__shared__ uint32_t s_key[CC_MEMO];
	const int k = threadIdx.x % CC_KW, r = threadIdx.x / CC_KW;
	const int wx = blockIdx.x * CC_KW + k, y = blockIdx.y * CC_TY + r % CC_TY, z = blockIdx.z * CC_TZ + r / CC_TY;
	const uint32_t w = g.word(bits, wx, y, z, invert);
	s_key[threadIdx.x] = NONE;
	__syncthreads();
	uint32_t gi = w ? g.rb[g.widx(wx, y, z)] : 0u;
	for (uint32_t t = run_starts(w); t; t &= t - 1, gi++) {
		const uint32_t p = ld_agent(&L[gi]);
		if (p == gi)
			continue;
		uint32_t h = memo_hash(p);
		for (int probe = 0; probe < CC_PROBE; probe++) {
			const uint32_t old = atomicCAS(&s_key[h], NONE, p);
			if (old == NONE || old == p)
				break;
			h = (h + 1) & (CC_MEMO - 1);
		}
	}
	__syncthreads();
	if (s_key[threadIdx.x] != NONE)
		(void)uf_find_live(L, s_key[threadIdx.x]);
}

// A workgroup flattens several tiles (grid-stride over the tile index) and
// carries the count of its current root from tile to tile, so the root of a
// component that spans the volume receives one device-scope atomic per
// workgroup rather than one per tile: atomics on one address from
// thousands of tiles serialise (they took about 0.8 ms on a 512^3 mask).
constexpr unsigned CC_FLAT_WG = 1024;

__global__ void __launch_bounds__(CC_BS)
k_ccl_flatten(const uint32_t *__restrict__ bits, Geo g, int invert, uint32_t *__restrict__ L,
              uint32_t *__restrict__ AUX, int mode, uint32_t *__restrict__ nroots,
              uint32_t *__restrict__ rootlist, int gx, int gy, uint32_t ntiles)
{
	__shared__ uint32_t s_root[CC_BS / N2M_WAVE], s_cnt[CC_BS / N2M_WAVE];
	__shared__ uint32_t s_key[CC_MEMO], s_val[CC_MEMO];
	const int k = threadIdx.x % CC_KW, r = threadIdx.x / CC_KW;
	uint32_t carry_root = NONE, carry_cnt = 0; // thread 0: the workgroup's pending count
	uint32_t my_roots = 0;                      // roots among this thread's runs, all tiles
	for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const int bx = (int)(tile % (uint32_t)gx), by = (int)((tile / (uint32_t)gx) % (uint32_t)gy),
		          bz = (int)(tile / ((uint32_t)gx * (uint32_t)gy));
		const int wx = bx * CC_KW + k, y = by * CC_TY + r % CC_TY, z = bz * CC_TZ + r / CC_TY;
		const uint32_t w = g.word(bits, wx, y, z, invert);
		s_key[threadIdx.x] = NONE;
		__syncthreads();
		const uint32_t gi0 = w ? g.rb[g.widx(wx, y, z)] : 0u;
		// the distinct parents of the tile's runs (a run that finds no free
		// slot walks on its own below)
		{
			uint32_t gi = gi0;
			for (uint32_t t = run_starts(w); t; t &= t - 1, gi++) {
				const uint32_t p = L[gi];
				if (p == gi)
					continue;
				uint32_t h = memo_hash(p);
				for (int probe = 0; probe < CC_PROBE; probe++) {
					const uint32_t old = atomicCAS(&s_key[h], NONE, p);
					if (old == NONE || old == p)
						break;
					h = (h + 1) & (CC_MEMO - 1);
				}
			}
		}
		__syncthreads();
		if (s_key[threadIdx.x] != NONE)
			s_val[threadIdx.x] = ccl_root(L, s_key[threadIdx.x]);
		__syncthreads();
		uint32_t nroot_local = 0, cur = NONE, cnt = 0;
		const bool face_row = y == 0 || z == 0 || y == g.ny - 1 || z == g.nz - 1;
		uint32_t gi = gi0;
		for (uint32_t t = run_starts(w); t; t &= t - 1, gi++) {
			int b = __ffs(t) - 1;
			const uint32_t p = L[gi];
			uint32_t root = p;
			if (p != gi) {
				uint32_t h = memo_hash(p);
				int probe = 0;
				for (; probe < CC_PROBE; probe++) {
					const uint32_t key = s_key[h];
					if (key == p || key == NONE)
						break;
					h = (h + 1) & (CC_MEMO - 1);
				}
				root = (probe < CC_PROBE && s_key[h] == p) ? s_val[h] : ccl_root(L, p);
				if (root != p)
					L[gi] = root;
			}
			if (mode == 0) {
				const uint32_t a = AUX[gi]; // stable: adds go to global roots only
				if ((a & TILE_ROOT) && root != gi) {
					if (root != cur) {
						if (cnt)
							atomicAdd(&AUX[cur], cnt);
						cur = root;
						cnt = 0;
					}
					cnt += a & ~TILE_ROOT;
				}
			} else {
				int x0 = wx * 32 + b, x1 = x0 + run_len(w, b) - 1;
				if ((face_row || x0 == 0 || x1 == g.nx - 1) && AUX[root] == 0)
					AUX[root] = 1;
			}
			nroot_local += root == gi;
		}
		my_roots += nroot_local;
		if (mode == 1) {
			__syncthreads(); // the table is cleared for the next tile
			continue;
		}
		// wave: lanes holding the same root add their counts
		bool pend = cnt != 0;
		bool first = true;
		unsigned long long act = __ballot(pend);
		while (act) {
			int leader = __ffsll((long long)act) - 1;
			uint32_t lr = __shfl(cur, leader, N2M_WAVE);
			bool mine = pend && cur == lr;
			uint32_t sum = wave_sum(mine ? cnt : 0u);
			if (first) { // the wave's first root goes through LDS
				if (lane_id() == 0) {
					s_root[wave_id()] = lr;
					s_cnt[wave_id()] = sum;
				}
				first = false;
			} else if (lane_id() == leader) {
				atomicAdd(&AUX[lr], sum);
			}
			if (mine)
				pend = false;
			act = __ballot(pend);
		}
		if (first && lane_id() == 0) {
			s_root[wave_id()] = NONE;
			s_cnt[wave_id()] = 0;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			for (int i = 0; i < CC_BS / N2M_WAVE; i++) {
				if (s_root[i] == NONE)
					continue;
				uint32_t sum = s_cnt[i];
				for (int j = i + 1; j < CC_BS / N2M_WAVE; j++)
					if (s_root[j] == s_root[i]) {
						sum += s_cnt[j];
						s_root[j] = NONE;
					}
				if (s_root[i] == carry_root) {
					carry_cnt += sum;
				} else {
					if (carry_cnt)
						atomicAdd(&AUX[carry_root], carry_cnt);
					carry_root = s_root[i];
					carry_cnt = sum;
				}
			}
		}
		__syncthreads(); // s_root, s_cnt and the table are reused by the next tile
	}
	if (threadIdx.x == 0 && carry_cnt)
		atomicAdd(&AUX[carry_root], carry_cnt);
	// The roots: one atomic on the count per workgroup.  Returning atomics
	// on one address serialise (one per wave, 65 thousand on a noisy 512^3
	// mask, took 0.8 ms).  A root never changes (L[root] == root from the
	// merge on), so the second walk over the tiles finds the same runs.
	__shared__ uint32_t s_scan[CC_BS / N2M_WAVE + 1], s_base;
	uint32_t total;
	const uint32_t off = block_excl_scan<CC_BS>(my_roots, s_scan, &total);
	if (threadIdx.x == 0)
		s_base = total ? atomicAdd(nroots, total) : 0u;
	__syncthreads();
	if (mode == 1 || total == 0)
		return;
	uint32_t at = s_base + off;
	for (uint32_t tile = blockIdx.x; tile < ntiles && my_roots; tile += gridDim.x) {
		const int bx = (int)(tile % (uint32_t)gx), by = (int)((tile / (uint32_t)gx) % (uint32_t)gy),
		          bz = (int)(tile / ((uint32_t)gx * (uint32_t)gy));
		const int wx = bx * CC_KW + k, y = by * CC_TY + r % CC_TY, z = bz * CC_TZ + r / CC_TY;
		const uint32_t w = g.word(bits, wx, y, z, invert);
		if (!w)
			continue;
		uint32_t gi = g.rb[g.widx(wx, y, z)];
		for (uint32_t t = run_starts(w); t && my_roots; t &= t - 1, gi++)
			if (L[gi] == gi) {
				rootlist[at++] = gi;
				my_roots--;
			}
	}
}

static dim3 cc_grid(const Geo &g)
{
	return dim3((g.W + CC_KW - 1) / CC_KW, (g.ny + CC_TY - 1) / CC_TY, (g.nz + CC_TZ - 1) / CC_TZ);
}

void launch_ccl(const uint32_t *bits, int nx, int ny, int nz, int conn, int invert, uint32_t *L,
                uint32_t *AUX, int mode, uint32_t *nroots, uint32_t *rootlist, uint32_t *rbase,
                uint32_t *scan_scratch, hipStream_t s, CclTimer *timer)
{
	Geo g{nx, ny, nz, (nx + 31) / 32, rbase};
	dim3 grid = cc_grid(g);
	const double nb = (double)nx * ny * nz / 8;
	const size_t nw = (size_t)g.W * ny * nz;
	// algorithmic bytes: mask bits in per pass; run labels are not counted
	if (timer) timer->begin("ccl_runs", nb + 8.0 * nw);
	hipLaunchKernelGGL(k_run_count, dim3((unsigned)((nw + 1 + 255) / 256)), dim3(256), 0, s, bits, g, invert, rbase);
	scan_excl_u32(rbase, rbase, nw + 1, nullptr, scan_scratch, s);
	if (timer) timer->end();
	if (timer) timer->begin("ccl_local", nb);
	hipLaunchKernelGGL(k_ccl_local, grid, dim3(CC_BS), 0, s, bits, g, conn, invert, L, AUX, mode);
	if (timer) timer->end();
	if (timer) timer->begin("ccl_merge", nb);
	hipLaunchKernelGGL(k_ccl_merge, grid, dim3(CC_BS), 0, s, bits, g, conn, invert, L);
	if (timer) timer->end();
	if (timer) timer->begin("ccl_flatten", nb);
	hipLaunchKernelGGL(k_ccl_jump, grid, dim3(CC_BS), 0, s, bits, g, invert, L);
	const uint32_t ntiles = grid.x * grid.y * grid.z;
	hipLaunchKernelGGL(k_ccl_flatten, dim3(ntiles < CC_FLAT_WG ? ntiles : CC_FLAT_WG), dim3(CC_BS), 0, s, bits, g,
	                   invert, L, AUX, mode, nroots, rootlist, (int)grid.x, (int)grid.y, ntiles);
	if (timer) timer->end();
}

// largest root: max count, ties -> smallest root.  best = (count<<32)|~root;
// a root's AUX holds TILE_ROOT | its count (mode 0 of the labelling)
__global__ void k_ccl_argmax(const uint32_t *__restrict__ rootlist, const uint32_t *__restrict__ nroots,
                             const uint32_t *__restrict__ AUX, unsigned long long *best)
{
	uint32_t n = *nroots;
	unsigned long long b = 0;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		uint32_t r = rootlist[i];
		unsigned long long key = ((unsigned long long)(AUX[r] & ~TILE_ROOT) << 32) | (unsigned long long)(0xffffffffu - r);
		b = b > key ? b : key;
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

// out = runs of `bits` whose root is `best` (the largest cluster), per word
__global__ void __launch_bounds__(256)
k_select_root(const uint32_t *__restrict__ bits, Geo g, const uint32_t *__restrict__ L,
              const unsigned long long *__restrict__ bestp, uint32_t *__restrict__ out)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t nw = (size_t)g.W * g.ny * g.nz;
	if (t >= nw)
		return;
	unsigned long long bk = *bestp;
	uint32_t best = bk ? 0xffffffffu - (uint32_t)(bk & 0xffffffffu) : NONE;
	const uint32_t w = bits[t];
	uint32_t o = 0;
	uint32_t id = w ? g.rb[t] : 0u;
	for (uint32_t s = run_starts(w); s; s &= s - 1, id++) {
		int b = __ffs(s) - 1;
		if (L[id] == best)
			o |= run_mask(w, b);
	}
	out[t] = o;
}

void launch_select_root(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L,
                        const unsigned long long *best, uint32_t *out, const uint32_t *rbase, hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32, rbase};
	size_t nw = (size_t)g.W * ny * nz;
	hipLaunchKernelGGL(k_select_root, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, bits, g, L, best, out);
}

// bubble fill (src/bwlabel.c:530-538): background runs whose component
// touches no face are set in the foreground `bits` (in place, per word).
// Nothing when the background has at most one component (*ninv, the
// reference's "no bubbles" case, src/bwlabel.c:488-491).
__global__ void __launch_bounds__(256)
k_fill_bubbles(uint32_t *__restrict__ bits, Geo g, const uint32_t *__restrict__ L, const uint32_t *__restrict__ AUX,
               const uint32_t *__restrict__ ninv)
{
	size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t nw = (size_t)g.W * g.ny * g.nz;
	if (t >= nw || *ninv <= 1u)
		return;
	const int wx = (int)(t % g.W);
	const uint32_t fg = bits[t];
	const int rem = g.nx - wx * 32;
	const uint32_t bg = ~fg & (rem >= 32 ? 0xffffffffu : ((1u << rem) - 1u));
	uint32_t add = 0;
	uint32_t id = bg ? g.rb[t] : 0u; // run ids of the background labelling
	for (uint32_t s = run_starts(bg); s; s &= s - 1, id++) {
		int b = __ffs(s) - 1;
		if (AUX[L[id]] == 0)
			add |= run_mask(bg, b);
	}
	if (add)
		bits[t] = fg | add;
}

void launch_fill_bubbles(uint32_t *bits, int nx, int ny, int nz, const uint32_t *L, const uint32_t *AUX,
                         const uint32_t *rbase, const uint32_t *ninv, hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32, rbase};
	size_t nw = (size_t)g.W * ny * nz;
	hipLaunchKernelGGL(k_fill_bubbles, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, bits, g, L, AUX, ninv);
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
	const size_t wi = (size_t)(x >> 5) + (size_t)g.W * row;
	uint32_t w = bits[wi];
	int b = x & 31;
	float out = 0.0f;
	if ((w >> b) & 1u)
		out = (float)rank[L[g.rb[wi] + run_index(run_starts(w), b)]];
	img[v] = out;
}

void launch_write_labels(const uint32_t *bits, int nx, int ny, int nz, const uint32_t *L,
                         const uint32_t *rank, float *img, const uint32_t *rbase, hipStream_t s)
{
	Geo g{nx, ny, nz, (nx + 31) / 32, rbase};
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
