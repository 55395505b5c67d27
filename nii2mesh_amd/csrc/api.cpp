// api.cpp -- C ABI of the MI355X voxel-to-mesh library (include/nii2mesh_amd.h).
//
// Host orchestration follows the reference's control flow stage by stage
// (each function cites the reference lines it replaces); all arithmetic runs
// in the kernels of k_*.hip on the context's HIP stream.  Errors inside are
// C++ exceptions, turned into EXIT_FAILURE plus a message at the ABI edge.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nii2mesh_amd.h"
#include "n2m_internal.h"

using namespace n2m;

namespace {

thread_local std::string tl_err;

struct Error : std::runtime_error {
	using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	throw Error(buf);
}

#define HIPX(call)                                                                                       \
	do {                                                                                                 \
		hipError_t e_ = (call);                                                                          \
		if (e_ != hipSuccess)                                                                            \
			fail("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);             \
	} while (0)

double now_ms()
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec * 1000.0 + t.tv_nsec / 1.0e6;
}

int g_cubes = -1; // process-wide marching-cubes flavour

int cubes_flavour()
{
	if (g_cubes < 0) {
		const char *e = getenv("N2M_CUBES");
		g_cubes = (e && (e[0] == 'l' || e[0] == 'L' || e[0] == '1')) ? N2M_CUBES_LEWINER : N2M_CUBES_CLASSIC;
	}
	return g_cubes;
}

// buffer slots of a context
enum Slot {
	S_VOL,
	S_VOL2,
	S_BITS,
	S_BITS2,
	S_BITS3,
	S_L,
	S_AUX,
	S_ROOTS,
	S_CFG,
	S_CROW,
	S_PFLAGS,
	S_SEGCNT,
	S_SEGBASE,
	S_BLOCK,
	S_BLOCKOFF,
	S_SCAN,
	S_PTS0,
	S_TRIS0,
	S_K0,
	S_V0,
	S_K1,
	S_V1,
	S_RCOUNT,
	S_FLAGS,
	S_SEGID,
	S_STARTS,
	S_LCLS,
	S_SEEDS,
	S_SEEDBASE,
	S_CLS,
	S_REP,
	S_TRIS1,
	S_PTS1,
	S_KEEP,
	S_KEEPPOS,
	S_HC_START,
	S_HC_CNT,
	S_HC_REF,
	S_HC_PAIRS,
	S_HC_BORDER,
	S_HC_O,
	S_HC_P,
	S_HC_PN,
	S_HC_B,
	S_HC_Q,
	S_HIST,
	S_HC_TRIS,
	S_HC_INV,
	S_HC_OUT,
	S_COUNT
};

struct DevBuf {
	void *p = nullptr;
	size_t cap = 0;
};

struct Mesh {
	DevBuf tris, pts;
	int nt = 0, np = 0;
};

// small per-context scalars (device + pinned host mirror)
enum Small {
	SM_RANGE = 0,  // 2 words: ord min, ord max
	SM_NROOTS = 2, // 1 word
	SM_BEST = 4,   // 2 words (u64)
	SM_BBOX = 6,   // 6 words
	SM_VMIN = 12,  // 1 word
	SM_TOTAL = 16, // 4 words of totals
	SM_HSTAT = 24, // 3 words
	SM_PBOX = 32,  // 6 words: float box of vertex positions
	SM_FLAG = 40,  // 1 word
	SM_WORDS = 64
};

struct Stat {
	std::string name;
	double ms = 0, bytes = 0;
	int n = 0;
};

struct Pending {
	int stat;
	hipEvent_t a, b;
	double bytes;
};

} // namespace

struct n2m_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	DevBuf bufs[S_COUNT];
	std::vector<Mesh> meshes;
	uint32_t *d_small = nullptr;
	uint32_t *h_small = nullptr;
	bool timing = false;
	std::vector<Stat> stats;
	std::vector<Pending> pending;
	std::vector<hipEvent_t> evpool;
	int cur_stat = -1;
	hipEvent_t cur_a = nullptr;
	double cur_bytes = 0;

	void *get(Slot s, size_t bytes)
	{
		DevBuf &b = bufs[s];
		if (bytes == 0)
			bytes = 1;
		if (b.cap < bytes) {
			if (b.p)
				HIPX(hipFree(b.p));
			b.p = nullptr;
			size_t want = bytes + bytes / 8 + 256;
			HIPX(hipMalloc(&b.p, want));
			b.cap = want;
		}
		return b.p;
	}
	template <typename T> T *buf(Slot s, size_t count) { return (T *)get(s, count * sizeof(T)); }

	static void *grow(DevBuf &b, size_t bytes)
	{
		if (bytes == 0)
			bytes = 1;
		if (b.cap < bytes) {
			if (b.p)
				HIPX(hipFree(b.p));
			b.p = nullptr;
			HIPX(hipMalloc(&b.p, bytes));
			b.cap = bytes;
		}
		return b.p;
	}

	void sync() { HIPX(hipStreamSynchronize(stream)); }
	void check() { HIPX(hipGetLastError()); }

	// read `n` small words at offset `off` back to the host
	uint32_t *fetch(int off, int n)
	{
		HIPX(hipMemcpyAsync(h_small + off, d_small + off, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
		sync();
		return h_small + off;
	}

	// ---- timing ----
	hipEvent_t ev()
	{
		if (!evpool.empty()) {
			hipEvent_t e = evpool.back();
			evpool.pop_back();
			return e;
		}
		hipEvent_t e;
		HIPX(hipEventCreate(&e));
		return e;
	}
	void tbegin(const char *name, double bytes = 0)
	{
		check();
		if (!timing)
			return;
		int k = -1;
		for (size_t i = 0; i < stats.size(); i++)
			if (stats[i].name == name)
				k = (int)i;
		if (k < 0) {
			stats.push_back(Stat());
			stats.back().name = name;
			k = (int)stats.size() - 1;
		}
		cur_stat = k;
		cur_a = ev();
		cur_bytes = bytes;
		HIPX(hipEventRecord(cur_a, stream));
	}
	void tend()
	{
		check();
		if (!timing || cur_stat < 0)
			return;
		hipEvent_t b = ev();
		HIPX(hipEventRecord(b, stream));
		pending.push_back({cur_stat, cur_a, b, cur_bytes});
		cur_stat = -1;
	}
	void resolve()
	{
		if (pending.empty())
			return;
		sync();
		for (auto &p : pending) {
			float ms = 0;
			HIPX(hipEventElapsedTime(&ms, p.a, p.b));
			stats[p.stat].ms += ms;
			stats[p.stat].n += 1;
			stats[p.stat].bytes += p.bytes;
			evpool.push_back(p.a);
			evpool.push_back(p.b);
		}
		pending.clear();
	}
};

namespace {

struct Timer : CclTimer {
	n2m_ctx *c;
	explicit Timer(n2m_ctx *c_) : c(c_) {}
	void begin(const char *n, double bytes) override { c->tbegin(n, bytes); }
	void end() override { c->tend(); }
};

n2m_ctx *ctx_new()
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
		fail("no HIP device available");
	n2m_ctx *c = new n2m_ctx();
	HIPX(hipGetDevice(&c->device));
	HIPX(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
	HIPX(hipMalloc(&c->d_small, SM_WORDS * sizeof(uint32_t)));
	HIPX(hipHostMalloc((void **)&c->h_small, SM_WORDS * sizeof(uint32_t), hipHostMallocDefault));
	return c;
}

void ctx_free(n2m_ctx *c)
{
	if (!c)
		return;
	(void)hipStreamSynchronize(c->stream);
	for (auto &b : c->bufs)
		if (b.p)
			(void)hipFree(b.p);
	for (auto &m : c->meshes) {
		if (m.tris.p)
			(void)hipFree(m.tris.p);
		if (m.pts.p)
			(void)hipFree(m.pts.p);
	}
	for (auto &p : c->pending) {
		(void)hipEventDestroy(p.a);
		(void)hipEventDestroy(p.b);
	}
	for (auto e : c->evpool)
		(void)hipEventDestroy(e);
	if (c->d_small)
		(void)hipFree(c->d_small);
	if (c->h_small)
		(void)hipHostFree(c->h_small);
	(void)hipStreamDestroy(c->stream);
	delete c;
}

// one context per host thread for the drop-in entry points (meshify is
// called concurrently from OpenMP threads in atlas mode, src/nii2mesh.c:546)
struct TlsCtx {
	n2m_ctx *c = nullptr;
	~TlsCtx() { ctx_free(c); }
	n2m_ctx *get()
	{
		if (!c)
			c = ctx_new();
		return c;
	}
};
thread_local TlsCtx tl_ctx;

Mesh &mesh_slot(n2m_ctx *c, int i)
{
	if ((int)c->meshes.size() <= i)
		c->meshes.resize(i + 1);
	return c->meshes[i];
}

// scratch of the device scans: sized for 2^31 elements up front, so a later
// scan never reallocates a buffer that queued kernels still read
uint32_t *scan_scratch(n2m_ctx *C, size_t n)
{
	size_t w = scan_scratch_words(n) + 16;
	return C->buf<uint32_t>(S_SCAN, w > ((size_t)1 << 21) ? w : ((size_t)1 << 21));
}

uint32_t scan(n2m_ctx *C, const uint32_t *in, uint32_t *out, size_t n, bool want_total)
{
	uint32_t *scr = scan_scratch(C, n);
	uint32_t *tot = C->d_small + SM_TOTAL;
	scan_excl_u32(in, out, n, want_total ? tot : nullptr, scr, C->stream);
	C->check();
	if (!want_total)
		return 0;
	return C->fetch(SM_TOTAL, 1)[0];
}

// ------------------------------------------------------------------------
// Volume stages of meshify(), src/meshify.c:294-371.
// Returns false where meshify() returns EXIT_FAILURE.  vol is the working
// volume (already smoothed when smoothed_range is given).
// ------------------------------------------------------------------------
struct VolOut {
	float iso;
	int lo[3], hi[3];
	float vmin; // min of the final volume (Lewiner padding)
};

bool volume_stages(n2m_ctx *C, float *vol, const int dim[3], const n2m_params &P, bool have_range,
                   bool spec_bits_valid_for_iso, VolOut &out)
{
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const size_t nvox = (size_t)NX * NY * NZ;
	hipStream_t s = C->stream;
	if (!have_range) {
		uint32_t init[2] = {0xffffffffu, 0u};
		HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
		C->tbegin("minmax", 4.0 * nvox);
		launch_minmax(vol, nvox, C->d_small + SM_RANGE, s);
		C->tend();
	}
	uint32_t *r = C->fetch(SM_RANGE, 2);
	float mn = ord2f_host(r[0]), mx = ord2f_host(r[1]);
	if (mn == mx) {
		printf("Error: No variability in image intensity.\n");
		return false;
	}
	float isolevel = P.isolevel;
	bool iso_changed = false;
	if ((isolevel <= mn) || (isolevel > mx)) { // src/meshify.c:316-319
		isolevel = 0.5 * (mn + mx);
		iso_changed = true;
		printf("Suggested isolevel out of range. Intensity range %g..%g, setting isolevel to %g\n", mn, mx,
		       isolevel);
	}
	if (P.verbose)
		printf("intensity range %g..%g, isolevel %g\n", mn, mx, isolevel);
	const int W = (NX + 31) / 32;
	const size_t nwords = (size_t)W * NY * NZ;
	uint32_t *keep = nullptr, *dil = nullptr;
	double t0 = now_ms();
	// bwlabelCore() rejects dim[0]<2 || dim[1]<2 || dim[2]<1 and leaves the
	// mask untouched (src/bwlabel.c:434-437)
	const bool label_ok = NX >= 2 && NY >= 2 && NZ >= 1;
	if (P.onlyLargest || P.fillBubbles) {
		uint32_t *bits = C->buf<uint32_t>(S_BITS, nwords);
		if (!spec_bits_valid_for_iso || iso_changed) {
			C->tbegin("mask_bits", 4.0 * nvox + nvox / 8.0);
			launch_mask_bits(vol, NX, NY, NZ, isolevel, 0, bits, s);
			C->tend();
		}
		keep = bits;
		if (label_ok) {
			uint32_t *L = C->buf<uint32_t>(S_L, nvox);
			uint32_t *AUX = C->buf<uint32_t>(S_AUX, nvox);
			Timer tm(C);
			if (P.fillBubbles) { // src/bwlabel.c:478-541
				HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
				launch_ccl(bits, NX, NY, NZ, 6, 1, L, AUX, 1, C->d_small + SM_NROOTS, nullptr, s, &tm);
				uint32_t ninv = C->fetch(SM_NROOTS, 1)[0];
				if (ninv > 1) {
					C->tbegin("fill_bubbles", 4.0 * nvox);
					launch_fill_bubbles(bits, NX, NY, NZ, L, AUX, s);
					C->tend();
				}
			}
			if (P.onlyLargest) { // src/bwlabel.c:453-471
				uint32_t *roots = C->buf<uint32_t>(S_ROOTS, nvox / 2 + 1024);
				HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
				HIPX(hipMemsetAsync(C->d_small + SM_BEST, 0, 8, s));
				launch_ccl(bits, NX, NY, NZ, 18, 0, L, AUX, 0, C->d_small + SM_NROOTS, roots, s, &tm);
				C->tbegin("ccl_argmax");
				launch_ccl_argmax(roots, C->d_small + SM_NROOTS, AUX, (unsigned long long *)(C->d_small + SM_BEST), s);
				C->tend();
				uint32_t *sel = C->buf<uint32_t>(S_BITS2, nwords);
				C->tbegin("select_root", 4.0 * nvox);
				launch_select_root(bits, NX, NY, NZ, L, (unsigned long long *)(C->d_small + SM_BEST), sel, s);
				C->tend();
				keep = sel;
			}
		}
		if (P.onlyLargest) { // dilate(mask, dim, true), src/meshify.c:337-342
			dil = C->buf<uint32_t>(S_BITS3, nwords);
			C->tbegin("dilate_bits", 2.0 * nwords * 4);
			launch_dilate_bits(keep, dil, NX, NY, NZ, s);
			C->tend();
		}
	}
	// apply + edge darken + bounding box, src/meshify.c:332-371
	float edgeMax = 0.75 * (mn + isolevel);
	int init[7] = {NX, NY, NZ, 0, 0, 0, 0};
	uint32_t minit = 0xffffffffu;
	HIPX(hipMemcpyAsync(C->d_small + SM_BBOX, init, 6 * sizeof(int), hipMemcpyHostToDevice, s));
	HIPX(hipMemcpyAsync(C->d_small + SM_VMIN, &minit, 4, hipMemcpyHostToDevice, s));
	C->tbegin("apply", 8.0 * nvox);
	launch_apply(vol, NX, NY, NZ, isolevel, mn, edgeMax, P.fillBubbles ? keep : nullptr, dil, P.fillBubbles ? 1 : 0,
	             (int *)(C->d_small + SM_BBOX), C->d_small + SM_VMIN, s);
	C->tend();
	uint32_t *bb = C->fetch(SM_BBOX, 7);
	if ((P.onlyLargest || P.fillBubbles) && P.verbose)
		printf("voxel clustering (largest cluster, bubbles): %ld ms\n", (long)round(now_ms() - t0));
	int lo[3] = {(int)bb[0], (int)bb[1], (int)bb[2]}, hi[3] = {(int)bb[3], (int)bb[4], (int)bb[5]};
	for (int i = 0; i < 3; i++) {
		lo[i] = lo[i] - 1 > 0 ? lo[i] - 1 : 0;
		hi[i] = hi[i] + 2 < dim[i] ? hi[i] + 2 : dim[i];
		out.lo[i] = lo[i];
		out.hi[i] = hi[i];
	}
	out.iso = isolevel;
	out.vmin = ord2f_host(bb[6]);
	return true;
}

// ------------------------------------------------------------------------
// marching cubes into (S_PTS0, S_TRIS0).  classic: tris implicit.
// ------------------------------------------------------------------------
struct McOut {
	uint32_t npt = 0, ntri = 0;
	bool implicit = false;
};

bool marching_cubes(n2m_ctx *C, const float *vol, const int dim[3], const int lo[3], const int hi[3], float iso,
                    int cubes, int originalMC, float vmin, McOut &o)
{
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	hipStream_t s = C->stream;
	if (cubes == N2M_CUBES_CLASSIC) { // src/oldcubes.c:465-522
		int cx = hi[0] - 1 - lo[0], cy = hi[1] - 1 - lo[1], cz = hi[2] - 1 - lo[2];
		uint32_t T = 0;
		size_t ncube = (cx > 0 && cy > 0 && cz > 0) ? (size_t)cx * cy * cz : 0;
		size_t nb = 0;
		if (ncube) {
			nb = classic_blocks(cx, cy, cz);
			uint8_t *cfg = C->buf<uint8_t>(S_CFG, ncube);
			uint32_t *bs = C->buf<uint32_t>(S_BLOCK, nb);
			uint32_t *bo = C->buf<uint32_t>(S_BLOCKOFF, nb);
			C->tbegin("mc_classic_count", 4.0 * ncube + ncube);
			launch_classic_count(vol, NX, NY, NZ, lo, hi, iso, cfg, bs, s);
			C->tend();
			T = scan(C, bs, bo, nb, true);
		}
		if (3 * (size_t)T < 3) {
			printf("marching cubes failed to identify triangles with an isolevel of %g\n", iso);
			return false;
		}
		double *pts = C->buf<double>(S_PTS0, (size_t)9 * T);
		C->tbegin("mc_classic_emit", 1.0 * ncube + 72.0 * T);
		launch_classic_emit(vol, NX, NY, NZ, lo, hi, iso, C->buf<uint8_t>(S_CFG, ncube),
		                    C->buf<uint32_t>(S_BLOCKOFF, nb), pts, s);
		C->tend();
		o.npt = 3 * T;
		o.ntri = T;
		o.implicit = true;
		return true;
	}
	// Lewiner, src/MarchingCubes.c:1086-1143
	size_t sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
	size_t npts = sx * sy * sz;
	size_t nseg = lw_nseg_total(NX, NY, NZ, lo, hi);
	uint8_t *pf = C->buf<uint8_t>(S_PFLAGS, npts);
	uint16_t *cr = C->buf<uint16_t>(S_CROW, npts);
	uint32_t *cnt = C->buf<uint32_t>(S_SEGCNT, 3 * nseg);
	uint32_t *base = C->buf<uint32_t>(S_SEGBASE, 3 * nseg);
	float mnv = vmin - iso;
	C->tbegin("mc_lewiner_classify", 4.0 * npts + 3.0 * npts);
	launch_lw_classify(vol, NX, NY, NZ, lo, hi, iso, mnv, originalMC, pf, cr, cnt, s);
	C->tend();
	uint32_t nE = scan(C, cnt, base, nseg, true);
	uint32_t nT = scan(C, cnt + nseg, base + nseg, nseg, true);
	uint32_t nC = scan(C, cnt + 2 * nseg, base + 2 * nseg, nseg, true);
	uint32_t V = nE + nC;
	if (V < 3 || nT < 1)
		return false;
	double *pts = C->buf<double>(S_PTS0, (size_t)3 * V);
	int *tris = C->buf<int>(S_TRIS0, (size_t)3 * nT);
	C->tbegin("mc_lewiner_emit", 3.0 * npts + 24.0 * V + 12.0 * nT);
	launch_lw_emit(vol, NX, NY, NZ, lo, hi, iso, mnv, originalMC, pf, cr, base, nE, pts, tris, s);
	C->tend();
	o.npt = V;
	o.ntri = nT;
	o.implicit = false;
	return true;
}

// ------------------------------------------------------------------------
// weld (src/meshify.c:45-106) from (pts, tris) -> (ptsOut, trisOut).
// Returns the new vertex count; trisOut always receives explicit triangles.
// ------------------------------------------------------------------------
uint32_t weld(n2m_ctx *C, const double *pts, uint32_t npt, const int *tris_in /* null: implicit */, uint32_t nt,
              double *&pts_out, int *&tris_out, bool verbose)
{
	hipStream_t s = C->stream;
	double t0 = now_ms();
	uint32_t *k0 = C->buf<uint32_t>(S_K0, npt), *v0 = C->buf<uint32_t>(S_V0, npt);
	uint32_t *k1 = C->buf<uint32_t>(S_K1, npt), *v1 = C->buf<uint32_t>(S_V1, npt);
	C->tbegin("weld_keys", 24.0 * npt + 8.0 * npt);
	launch_weld_keys(pts, npt, k0, v0, s);
	C->tend();
	uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, radix_counts_words(npt));
	uint32_t *scr = scan_scratch(C, radix_counts_words(npt));
	C->tbegin("radix_sort", 4 * 16.0 * npt);
	bool in1 = radix_sort_u32(k0, v0, k1, v1, npt, 32, rc, scr, s);
	C->tend();
	uint32_t *keys = in1 ? k1 : k0, *vals = in1 ? v1 : v0;
	uint32_t *flags = C->buf<uint32_t>(S_FLAGS, npt), *segid = C->buf<uint32_t>(S_SEGID, npt);
	C->tbegin("weld_segments", 8.0 * npt);
	launch_weld_segflags(keys, npt, flags, s);
	C->tend();
	uint32_t nseg = scan(C, flags, segid, npt, true);
	uint32_t *starts = C->buf<uint32_t>(S_STARTS, nseg);
	int *lcls = C->buf<int>(S_LCLS, npt);
	uint32_t *seeds = C->buf<uint32_t>(S_SEEDS, nseg), *seedbase = C->buf<uint32_t>(S_SEEDBASE, nseg);
	C->tbegin("weld_greedy", 4.0 * npt + 4.0 * npt + 24.0 * npt + 4.0 * npt);
	launch_weld_segstarts(flags, segid, npt, starts, s);
	launch_weld_greedy(starts, nseg, npt, keys, vals, pts, lcls, seeds, s);
	C->tend();
	uint32_t nnew = scan(C, seeds, seedbase, nseg, true);
	tris_out = C->buf<int>(S_TRIS1, (size_t)3 * nt);
	if (nnew == npt) { // src/meshify.c:82-87: unchanged
		if (verbose)
			printf("Unify vertices found no shared vertices\n");
		pts_out = const_cast<double *>(pts);
		if (tris_in)
			HIPX(hipMemcpyAsync(tris_out, tris_in, (size_t)12 * nt, hipMemcpyDeviceToDevice, s));
		else // implicit classic triangles 3t, 3t+1, 3t+2
			launch_iota(tris_out, 3 * nt, s);
		C->check();
		return npt;
	}
	int *cls = C->buf<int>(S_CLS, npt), *rep = C->buf<int>(S_REP, nnew);
	HIPX(hipMemsetAsync(rep, 0xff, (size_t)4 * nnew, s));
	C->tbegin("weld_assign", 4.0 * npt * 3 + 8.0 * npt);
	launch_weld_assign(vals, lcls, flags, segid, seedbase, npt, cls, rep, s);
	C->tend();
	pts_out = C->buf<double>(S_PTS1, (size_t)3 * nnew);
	C->tbegin("weld_gather", 28.0 * nnew + 24.0 * nnew);
	launch_gather_pts(pts, rep, nnew, pts_out, s);
	launch_remap_tris(tris_in, cls, nt, npt, tris_out, s);
	C->tend();
	if (verbose) {
		C->sync();
		printf("vertex welding %d -> %d: %ld ms\n", (int)npt, (int)nnew, (long)round(now_ms() - t0));
	}
	return nnew;
}

} // namespace

namespace {

// remove_degenerate_triangles (src/meshify.c:113-168): tris -> S_TRIS0 (or in
// place when nothing is removed).  Returns the kept count.
uint32_t remove_degenerate(n2m_ctx *C, const double *pts, uint32_t np, int *&tris, uint32_t nt, bool verbose)
{
	hipStream_t s = C->stream;
	double t0 = now_ms();
	uint32_t *keep = C->buf<uint32_t>(S_KEEP, nt), *pos = C->buf<uint32_t>(S_KEEPPOS, nt);
	C->tbegin("degenerate", 12.0 * nt + 72.0 * nt + 4.0 * nt);
	launch_degen_flags(pts, tris, nt, np, keep, s);
	C->tend();
	uint32_t k = scan(C, keep, pos, nt, true);
	if (k == nt)
		return nt;
	int *out = C->buf<int>(S_TRIS0, (size_t)3 * (k ? k : 1));
	C->tbegin("compact_tris", 12.0 * nt + 12.0 * k);
	launch_compact_tris(tris, keep, pos, nt, out, s);
	C->tend();
	tris = out;
	if (verbose) {
		C->sync();
		printf("remove degenerate triangles %d -> %d: %ld ms\n", (int)nt, (int)k, (long)round(now_ms() - t0));
	}
	return k;
}

// meshify() after the volume stages: MC, weld, de-degenerate; result into mesh m
bool surface_stages(n2m_ctx *C, const float *vol, const int dim[3], const VolOut &vo, const n2m_params &P, Mesh &m)
{
	double t0 = now_ms();
	McOut mc;
	if (!marching_cubes(C, vol, dim, vo.lo, vo.hi, vo.iso, P.cubes, P.originalMC, vo.vmin, mc))
		return false;
	if (P.verbose) {
		C->sync();
		printf("marching cubes (%dx%dx%d): %ld ms\n", dim[0], dim[1], dim[2], (long)round(now_ms() - t0));
	}
	const double *pts0 = C->buf<double>(S_PTS0, (size_t)3 * mc.npt);
	const int *tris0 = mc.implicit ? nullptr : C->buf<int>(S_TRIS0, (size_t)3 * mc.ntri);
	double *pts = nullptr;
	int *tris = nullptr;
	uint32_t np = weld(C, pts0, mc.npt, tris0, mc.ntri, pts, tris, P.verbose != 0);
	if (np < 3)
		return false;
	uint32_t nt = remove_degenerate(C, pts, np, tris, mc.ntri, P.verbose != 0);
	// copy out of the scratch slots into the mesh's own buffers
	int *mt = (int *)n2m_ctx::grow(m.tris, (size_t)12 * (nt ? nt : 1));
	double *mp = (double *)n2m_ctx::grow(m.pts, (size_t)24 * np);
	HIPX(hipMemcpyAsync(mt, tris, (size_t)12 * nt, hipMemcpyDeviceToDevice, C->stream));
	HIPX(hipMemcpyAsync(mp, pts, (size_t)24 * np, hipMemcpyDeviceToDevice, C->stream));
	m.nt = (int)nt;
	m.np = (int)np;
	return true;
}

// full meshify() on a device volume: src (input) -> vol (working copy, may
// equal src).  src/meshify.c:286-389
bool meshify_device(n2m_ctx *C, const float *src, float *vol, const int dim[3], const n2m_params &P, Mesh &m)
{
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const size_t nvox = (size_t)NX * NY * NZ;
	hipStream_t s = C->stream;
	bool have_range = false, spec = false;
	if (P.preSmooth && NX >= 5 && NY >= 5 && NZ >= 5) { // src/meshify.c:299-304
		double t0 = now_ms();
		uint32_t init[2] = {0xffffffffu, 0u};
		HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
		float *out = vol;
		if (src == vol) // the stencil cannot run in place
			out = C->buf<float>(S_VOL2, nvox);
		uint32_t *bits = nullptr;
		if (P.onlyLargest || P.fillBubbles) {
			bits = C->buf<uint32_t>(S_BITS, (size_t)((NX + 31) / 32) * NY * NZ);
			spec = true;
		}
		C->tbegin("smooth", 8.0 * nvox);
		launch_smooth(src, out, NX, NY, NZ, C->d_small + SM_RANGE, P.isolevel, bits, s);
		C->tend();
		if (out != vol)
			HIPX(hipMemcpyAsync(vol, out, 4 * nvox, hipMemcpyDeviceToDevice, s));
		have_range = true;
		if (P.verbose) {
			C->sync();
			printf("pre-smooth: %ld ms\n", (long)round(now_ms() - t0));
		}
	} else if (src != vol) {
		HIPX(hipMemcpyAsync(vol, src, 4 * nvox, hipMemcpyDeviceToDevice, s));
	}
	VolOut vo;
	if (!volume_stages(C, vol, dim, P, have_range, spec, vo))
		return false;
	return surface_stages(C, vol, dim, vo, P, m);
}

// CSR of a device mesh (tris int[3*nt]) for smoothing
struct Csr {
	uint32_t *start;
	int2 *pairs;
	uint32_t *ref;
};

Csr build_csr(n2m_ctx *C, const int *tris, uint32_t nv, uint32_t nt)
{
	hipStream_t s = C->stream;
	Csr c;
	uint32_t *cnt = C->buf<uint32_t>(S_HC_CNT, nv + 1);
	c.start = C->buf<uint32_t>(S_HC_START, nv + 1);
	c.ref = C->buf<uint32_t>(S_HC_REF, (size_t)3 * nt + 1);
	c.pairs = C->buf<int2>(S_HC_PAIRS, (size_t)3 * nt + 1);
	HIPX(hipMemsetAsync(cnt, 0, 4 * ((size_t)nv + 1), s));
	C->tbegin("csr_build", 12.0 * nt * 4 + 8.0 * nv);
	launch_csr_count(tris, nt, cnt, s);
	scan_excl_u32(cnt, c.start, nv + 1, nullptr, scan_scratch(C, nv + 1), s);
	HIPX(hipMemsetAsync(cnt, 0, 4 * ((size_t)nv + 1), s));
	launch_csr_fill(tris, nt, c.start, cnt, c.ref, s);
	launch_csr_pairs(tris, nv, c.start, c.ref, c.pairs, s);
	C->tend();
	return c;
}

void check_tris(const int *tris, int ntri, int nvert)
{
	for (int i = 0; i < 3 * ntri; i++)
		if (tris[i] < 0 || tris[i] >= nvert)
			fail("triangle %d references vertex %d (mesh has %d vertices)", i / 3, tris[i], nvert);
}

// Vertex renumbering used by the smoothers: Morton order of the positions.
// perm[new] = old.  Returns false (identity) for small meshes.
struct Relabel {
	bool on = false;
	const uint32_t *perm = nullptr;
	const int *tris = nullptr;
};

Relabel morton_relabel(n2m_ctx *C, const double *pts, const int *tris, uint32_t nv, uint32_t nt)
{
	Relabel r;
	if (nv < 32768)
		return r;
	hipStream_t s = C->stream;
	C->tbegin("hc_relabel", 24.0 * nv * 2 + 16.0 * 4 * nv + 24.0 * nt);
	uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
	HIPX(hipMemcpyAsync(C->d_small + SM_PBOX, init, sizeof init, hipMemcpyHostToDevice, s));
	launch_pts_bbox(pts, nv, C->d_small + SM_PBOX, s);
	uint32_t *k0 = C->buf<uint32_t>(S_K0, nv), *v0 = C->buf<uint32_t>(S_V0, nv);
	uint32_t *k1 = C->buf<uint32_t>(S_K1, nv), *v1 = C->buf<uint32_t>(S_V1, nv);
	launch_morton(pts, nv, C->d_small + SM_PBOX, k0, v0, s);
	uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, radix_counts_words(nv));
	bool in1 = radix_sort_u32(k0, v0, k1, v1, nv, 32, rc, scan_scratch(C, radix_counts_words(nv)), s);
	r.perm = in1 ? v1 : v0;
	uint32_t *inv = C->buf<uint32_t>(S_HC_INV, nv);
	launch_invert_perm(r.perm, nv, inv, s);
	int *t2 = C->buf<int>(S_HC_TRIS, (size_t)3 * (nt ? nt : 1));
	launch_relabel_tris(tris, inv, nt, t2, s);
	C->tend();
	r.tris = t2;
	r.on = true;
	return r;
}

// border vertices of update_mesh(iteration 0), src/quadric.c:188-215
uint8_t *border_flags(n2m_ctx *C, const Csr &c, const int *tris, uint32_t nv, uint32_t nt)
{
	hipStream_t s = C->stream;
	uint8_t *border = C->buf<uint8_t>(S_HC_BORDER, nv);
	HIPX(hipMemsetAsync(C->d_small + SM_FLAG, 0, 4, s));
	launch_tris_repeat(tris, nt, C->d_small + SM_FLAG, s);
	bool repeats = C->fetch(SM_FLAG, 1)[0] != 0;
	C->tbegin("border");
	if (repeats) { // a triangle lists a vertex twice: count references literally
		HIPX(hipMemsetAsync(border, 0, nv, s));
		launch_border(tris, nv, c.start, c.ref, border, s);
	} else {
		launch_border_pairs(c.start, c.pairs, nv, border, s);
	}
	C->tend();
	return border;
}

// laplacian_smoothHC on a device mesh in place, src/quadric.c:343-394
void smooth_hc_device(n2m_ctx *C, double *pts, const int *tris, uint32_t nv, uint32_t nt, double alpha,
                      double beta, int iter, bool lockEdges)
{
	if (nv == 0)
		return;
	hipStream_t s = C->stream;
	Relabel R = morton_relabel(C, pts, tris, nv, nt);
	const int *T = R.on ? R.tris : tris;
	Csr c = build_csr(C, T, nv, nt);
	const size_t n3 = 3 * (size_t)nv;
	double *o = C->buf<double>(S_HC_O, n3), *p = C->buf<double>(S_HC_P, n3);
	double *pn = C->buf<double>(S_HC_PN, n3), *bq = C->buf<double>(S_HC_B, n3);
	launch_to4(pts, R.on ? R.perm : nullptr, nv, o, s);
	HIPX(hipMemcpyAsync(p, o, 8 * n3, hipMemcpyDeviceToDevice, s));
	const double pair_bytes = 8.0 * 3 * nt;
	for (int it = 0; it < iter; it++) {
		C->tbegin("hc_a", 4.0 * nv + pair_bytes + 24.0 * nv * 2 + 48.0 * nv);
		launch_hc_a(c.start, c.pairs, o, p, pn, bq, nv, alpha, s);
		C->tend();
		C->tbegin("hc_b", 4.0 * nv + pair_bytes + 24.0 * nv * 2 + 24.0 * nv);
		launch_hc_b(c.start, c.pairs, pn, bq, p, nv, beta, s);
		C->tend();
	}
	uint8_t *border = lockEdges ? border_flags(C, c, T, nv, nt) : nullptr;
	launch_from4(p, o, border, R.on ? R.perm : nullptr, nv, pts, s);
	C->check();
}

void laplace_device(n2m_ctx *C, double *pts, const int *tris, uint32_t nv, uint32_t nt, int iter)
{
	if (nv == 0 || iter <= 0)
		return;
	hipStream_t s = C->stream;
	Relabel R = morton_relabel(C, pts, tris, nv, nt);
	Csr c = build_csr(C, R.on ? R.tris : tris, nv, nt);
	const size_t n3 = 3 * (size_t)nv;
	double *p = C->buf<double>(S_HC_P, n3), *q = C->buf<double>(S_HC_PN, n3);
	launch_to4(pts, R.on ? R.perm : nullptr, nv, q, s);
	double *cur = q, *nxt = p;
	for (int it = 0; it < iter; it++) {
		C->tbegin("laplace", 4.0 * nv + 8.0 * 3 * nt + 48.0 * nv);
		launch_laplace(c.start, c.pairs, cur, nxt, nv, s);
		C->tend();
		double *t = cur;
		cur = nxt;
		nxt = t;
	}
	launch_from4(cur, cur, nullptr, R.on ? R.perm : nullptr, nv, pts, s);
	C->check();
}

template <typename F> int guarded(F f, int on_error = EXIT_FAILURE)
{
	try {
		tl_err.clear();
		return f();
	} catch (const std::exception &e) {
		tl_err = e.what();
		fprintf(stderr, "nii2mesh_amd: %s\n", e.what());
		return on_error;
	}
}

template <typename T> T *host_copy_out(n2m_ctx *C, const void *d, size_t count)
{
	T *h = (T *)malloc(count ? count * sizeof(T) : 1);
	if (!h)
		fail("out of host memory");
	if (count)
		HIPX(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, C->stream));
	return h;
}

int set_threshold_device(n2m_ctx *C, const float *d_img, int64_t nvox, int mode, float *result);
struct ThrHist {
	float mn, scl;
	std::vector<int> h;
};
bool threshold_hist(n2m_ctx *C, const float *d_img, int64_t nvox, ThrHist &out);
void thresholds_from(const ThrHist &H, int n, const int *modes, float *out);

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int n2m_set_cubes(int flavour)
{
	int prev = cubes_flavour();
	g_cubes = flavour == N2M_CUBES_LEWINER ? N2M_CUBES_LEWINER : N2M_CUBES_CLASSIC;
	return prev;
}

int n2m_get_cubes(void) { return cubes_flavour(); }

const char *n2m_last_error(void) { return tl_err.c_str(); }

int n2m_device_ok(void)
{
	return guarded([] {
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
			fail("no HIP device");
		hipDeviceProp_t prop;
		HIPX(hipGetDeviceProperties(&prop, 0));
		if (strncmp(prop.gcnArchName, "gfx9", 4) != 0)
			fail("device %s is not a gfx9 (CDNA) GPU", prop.gcnArchName);
		return 1;
	}, 0);
}

n2m_ctx *n2m_ctx_create(void)
{
	n2m_ctx *c = nullptr;
	guarded([&] {
		c = ctx_new();
		return 0;
	});
	return c;
}

void n2m_ctx_destroy(n2m_ctx *ctx) { ctx_free(ctx); }

void *n2m_ctx_stream(n2m_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int n2m_ctx_meshify(n2m_ctx *C, float *d_img, const int dim[3], const n2m_params *prm)
{
	return guarded([&] {
		Mesh &m = mesh_slot(C, 0);
		C->meshes.resize(1);
		if (!meshify_device(C, d_img, d_img, dim, *prm, m))
			return EXIT_FAILURE;
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_pipeline(n2m_ctx *C, const float *d_img, const int dim[3], const n2m_params *prm, int niso,
                     const float *isolevels, double alpha, double beta, int hcIter)
{
	return guarded([&] {
		const int NX = dim[0], NY = dim[1], NZ = dim[2];
		const size_t nvox = (size_t)NX * NY * NZ;
		hipStream_t s = C->stream;
		C->meshes.resize(niso);
		float *sm = C->buf<float>(S_VOL2, nvox); // smoothed volume, shared by all isolevels
		float *vol = C->buf<float>(S_VOL, nvox);
		bool smoothed = prm->preSmooth && NX >= 5 && NY >= 5 && NZ >= 5;
		uint32_t range[2];
		if (smoothed) {
			uint32_t init[2] = {0xffffffffu, 0u};
			HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
			C->tbegin("smooth", 8.0 * nvox);
			launch_smooth(d_img, sm, NX, NY, NZ, C->d_small + SM_RANGE, 0.f, nullptr, s);
			C->tend();
			uint32_t *r = C->fetch(SM_RANGE, 2);
			range[0] = r[0];
			range[1] = r[1];
		}
		int ok = 0;
		for (int k = 0; k < niso; k++) {
			n2m_params P = *prm;
			P.isolevel = isolevels[k];
			HIPX(hipMemcpyAsync(vol, smoothed ? sm : d_img, 4 * nvox, hipMemcpyDeviceToDevice, s));
			if (smoothed)
				HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, range, sizeof range, hipMemcpyHostToDevice, s));
			Mesh &m = C->meshes[k];
			m.nt = m.np = 0;
			VolOut vo;
			if (!volume_stages(C, vol, dim, P, smoothed, false, vo) || !surface_stages(C, vol, dim, vo, P, m))
				continue;
			if (hcIter > 0)
				smooth_hc_device(C, (double *)m.pts.p, (const int *)m.tris.p, m.np, m.nt, alpha, beta, hcIter,
				                 true);
			ok++;
		}
		C->sync();
		return ok == niso ? EXIT_SUCCESS : EXIT_FAILURE;
	});
}

float n2m_ctx_set_threshold(n2m_ctx *C, const float *d_img, int64_t nvox, int mode)
{
	float r = 1.0f;
	guarded([&] { return set_threshold_device(C, d_img, nvox, mode, &r); });
	return r;
}

int n2m_ctx_set_thresholds(n2m_ctx *C, const float *d_img, int64_t nvox, int n, const int *modes, float *out)
{
	return guarded([&] {
		ThrHist H;
		if (!threshold_hist(C, d_img, nvox, H)) {
			for (int i = 0; i < n; i++)
				out[i] = 1.0f;
			return EXIT_SUCCESS;
		}
		thresholds_from(H, n, modes, out);
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_smooth_hc(n2m_ctx *C, int mesh, double alpha, double beta, int iter, int lockEdges)
{
	return guarded([&] {
		if (mesh < 0 || mesh >= (int)C->meshes.size())
			fail("no mesh %d", mesh);
		Mesh &m = C->meshes[mesh];
		smooth_hc_device(C, (double *)m.pts.p, (const int *)m.tris.p, m.np, m.nt, alpha, beta, iter,
		                 lockEdges != 0);
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_mesh_count(n2m_ctx *C) { return (int)C->meshes.size(); }

int n2m_ctx_mesh_size(n2m_ctx *C, int mesh, int *nt, int *np)
{
	if (mesh < 0 || mesh >= (int)C->meshes.size())
		return EXIT_FAILURE;
	*nt = C->meshes[mesh].nt;
	*np = C->meshes[mesh].np;
	return EXIT_SUCCESS;
}

int n2m_ctx_download(n2m_ctx *C, int mesh, vec3i *t, vec3d *p)
{
	return guarded([&] {
		if (mesh < 0 || mesh >= (int)C->meshes.size())
			fail("no mesh %d", mesh);
		Mesh &m = C->meshes[mesh];
		if (m.nt)
			HIPX(hipMemcpyAsync(t, m.tris.p, (size_t)12 * m.nt, hipMemcpyDeviceToHost, C->stream));
		if (m.np)
			HIPX(hipMemcpyAsync(p, m.pts.p, (size_t)24 * m.np, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_device_mesh(n2m_ctx *C, int mesh, const vec3i **d_t, const vec3d **d_p)
{
	if (mesh < 0 || mesh >= (int)C->meshes.size())
		return EXIT_FAILURE;
	*d_t = (const vec3i *)C->meshes[mesh].tris.p;
	*d_p = (const vec3d *)C->meshes[mesh].pts.p;
	return EXIT_SUCCESS;
}

void n2m_ctx_enable_timing(n2m_ctx *C, int on) { C->timing = on != 0; }

void n2m_ctx_reset_timing(n2m_ctx *C)
{
	guarded([&] {
		C->resolve();
		C->stats.clear();
		return 0;
	});
}

int n2m_ctx_kernel_stat(n2m_ctx *C, int idx, const char **name, double *ms_total, int *launches,
                        double *bytes_total)
{
	guarded([&] {
		C->resolve();
		return 0;
	});
	if (idx >= 0 && idx < (int)C->stats.size()) {
		*name = C->stats[idx].name.c_str();
		*ms_total = C->stats[idx].ms;
		*launches = C->stats[idx].n;
		*bytes_total = C->stats[idx].bytes;
	}
	return (int)C->stats.size();
}

// ---------------------------------------------------------------------------
// drop-in entry points
// ---------------------------------------------------------------------------
int meshify(float *img, short dim[3], int originalMC, float isolevel, vec3i **t, vec3d **p, int *nt, int *np,
            bool preSmooth, bool onlyLargest, bool fillBubbles, bool verbose)
{
	return guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		int d[3] = {dim[0], dim[1], dim[2]};
		const size_t nvox = (size_t)d[0] * d[1] * d[2];
		float *src = C->buf<float>(S_VOL, nvox);
		HIPX(hipMemcpyAsync(src, img, 4 * nvox, hipMemcpyHostToDevice, C->stream));
		n2m_params P{cubes_flavour(), originalMC, isolevel, preSmooth, onlyLargest, fillBubbles, verbose};
		Mesh m;
		struct Free {
			Mesh &m;
			~Free()
			{
				if (m.tris.p)
					(void)hipFree(m.tris.p);
				if (m.pts.p)
					(void)hipFree(m.pts.p);
			}
		} fr{m};
		bool ok = meshify_device(C, src, src, d, P, m);
		// the reference leaves img smoothed, filled and darkened
		HIPX(hipMemcpyAsync(img, src, 4 * nvox, hipMemcpyDeviceToHost, C->stream));
		if (!ok) {
			C->sync();
			return EXIT_FAILURE;
		}
		*t = host_copy_out<vec3i>(C, m.tris.p, m.nt);
		*p = host_copy_out<vec3d>(C, m.pts.p, m.np);
		C->sync();
		*nt = m.nt;
		*np = m.np;
		return EXIT_SUCCESS;
	});
}

void laplacian_smoothHC(vec3d *verts, vec3i *tris, int nvert, int ntri, double alpha, double beta, int iter,
                        bool lockEdges)
{
	guarded([&] {
		if (nvert <= 0)
			return 0;
		check_tris((const int *)tris, ntri, nvert);
		n2m_ctx *C = tl_ctx.get();
		double *dp = C->buf<double>(S_PTS1, 3 * (size_t)nvert);
		int *dt = C->buf<int>(S_TRIS1, 3 * (size_t)(ntri > 0 ? ntri : 1));
		HIPX(hipMemcpyAsync(dp, verts, (size_t)24 * nvert, hipMemcpyHostToDevice, C->stream));
		if (ntri > 0)
			HIPX(hipMemcpyAsync(dt, tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		smooth_hc_device(C, dp, dt, nvert, ntri, alpha, beta, iter, lockEdges);
		HIPX(hipMemcpyAsync(verts, dp, (size_t)24 * nvert, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return 0;
	});
}

void n2m_laplacian_smooth(vec3d *verts, vec3i *tris, int nvert, int ntri, int iter)
{
	guarded([&] {
		if (nvert <= 0)
			return 0;
		check_tris((const int *)tris, ntri, nvert);
		n2m_ctx *C = tl_ctx.get();
		double *dp = C->buf<double>(S_PTS1, 3 * (size_t)nvert);
		int *dt = C->buf<int>(S_TRIS1, 3 * (size_t)(ntri > 0 ? ntri : 1));
		HIPX(hipMemcpyAsync(dp, verts, (size_t)24 * nvert, hipMemcpyHostToDevice, C->stream));
		if (ntri > 0)
			HIPX(hipMemcpyAsync(dt, tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		laplace_device(C, dp, dt, nvert, ntri, iter);
		HIPX(hipMemcpyAsync(verts, dp, (size_t)24 * nvert, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return 0;
	});
}

int n2m_quick_smooth(float *img, int nx, int ny, int nz)
{
	return guarded([&] {
		if (nx < 5 || ny < 5 || nz < 5)
			return EXIT_FAILURE;
		n2m_ctx *C = tl_ctx.get();
		size_t n = (size_t)nx * ny * nz;
		float *a = C->buf<float>(S_VOL, n), *b = C->buf<float>(S_VOL2, n);
		HIPX(hipMemcpyAsync(a, img, 4 * n, hipMemcpyHostToDevice, C->stream));
		HIPX(hipMemsetAsync(C->d_small + SM_RANGE, 0, 8, C->stream));
		launch_smooth(a, b, nx, ny, nz, C->d_small + SM_RANGE, 0.f, nullptr, C->stream);
		C->check();
		HIPX(hipMemcpyAsync(img, b, 4 * n, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_unify_vertices(vec3d **pts, vec3i *tris, int npt, int ntri)
{
	return guarded([&] {
		if (npt <= 0)
			return npt;
		check_tris((const int *)tris, ntri, npt);
		n2m_ctx *C = tl_ctx.get();
		double *dp = C->buf<double>(S_PTS0, 3 * (size_t)npt);
		int *dt = C->buf<int>(S_TRIS0, 3 * (size_t)(ntri > 0 ? ntri : 1));
		HIPX(hipMemcpyAsync(dp, *pts, (size_t)24 * npt, hipMemcpyHostToDevice, C->stream));
		if (ntri > 0)
			HIPX(hipMemcpyAsync(dt, tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		double *po = nullptr;
		int *to = nullptr;
		uint32_t nn = weld(C, dp, npt, dt, ntri, po, to, false);
		if ((int)nn == npt) {
			C->sync();
			return npt;
		}
		vec3d *h = host_copy_out<vec3d>(C, po, nn);
		if (ntri > 0)
			HIPX(hipMemcpyAsync(tris, to, (size_t)12 * ntri, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		free(*pts);
		*pts = h;
		return (int)nn;
	});
}

int n2m_remove_degenerate_triangles(vec3d *pts, vec3i **tris, int ntri)
{
	return guarded([&] {
		if (ntri <= 0)
			return ntri;
		n2m_ctx *C = tl_ctx.get();
		int maxv = 0;
		const int *ti = (const int *)*tris;
		for (int i = 0; i < 3 * ntri; i++)
			maxv = ti[i] > maxv ? ti[i] : maxv;
		double *dp = C->buf<double>(S_PTS0, 3 * (size_t)(maxv + 1));
		int *dt = C->buf<int>(S_TRIS1, 3 * (size_t)ntri);
		HIPX(hipMemcpyAsync(dp, pts, (size_t)24 * (maxv + 1), hipMemcpyHostToDevice, C->stream));
		HIPX(hipMemcpyAsync(dt, *tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		int *to = dt;
		uint32_t k = remove_degenerate(C, dp, (uint32_t)(maxv + 1), to, ntri, false);
		if ((int)k == ntri) {
			C->sync();
			return ntri;
		}
		vec3i *h = host_copy_out<vec3i>(C, to, k);
		C->sync();
		free(*tris);
		*tris = h;
		return (int)k;
	});
}

void n2m_apply_sform(vec3i *t, vec3d *p, int nt, int np, float srow_x[4], float srow_y[4], float srow_z[4])
{
	guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		if (np > 0) {
			double *dp = C->buf<double>(S_PTS1, 3 * (size_t)np);
			HIPX(hipMemcpyAsync(dp, p, (size_t)24 * np, hipMemcpyHostToDevice, C->stream));
			launch_sform(dp, np, srow_x, srow_y, srow_z, C->stream);
			C->check();
			HIPX(hipMemcpyAsync(p, dp, (size_t)24 * np, hipMemcpyDeviceToHost, C->stream));
			C->sync();
		}
		// winding: sign of the product of the row sums, src/meshify.c:1033-1044
		double px = srow_x[0] + srow_x[1] + srow_x[2];
		double py = srow_y[0] + srow_y[1] + srow_y[2];
		double pz = srow_z[0] + srow_z[1] + srow_z[2];
		float det = px * py * pz;
		if (det < 0.0)
			for (int i = 0; i < nt; i++) {
				int a = t[i].x;
				t[i].x = t[i].y;
				t[i].y = a;
			}
		return 0;
	});
}

int marchingCubes(float *img, short dim[3], int lo[3], int hi[3], int originalMC, float isolevel, vec3d **vs,
                  vec3i **ts, int *nv, int *nt)
{
	return guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		int d[3] = {dim[0], dim[1], dim[2]};
		const size_t nvox = (size_t)d[0] * d[1] * d[2];
		float *vol = C->buf<float>(S_VOL, nvox);
		HIPX(hipMemcpyAsync(vol, img, 4 * nvox, hipMemcpyHostToDevice, C->stream));
		float vmin = 0.f;
		int cubes = cubes_flavour();
		if (cubes == N2M_CUBES_LEWINER) {
			uint32_t init[2] = {0xffffffffu, 0u};
			HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, C->stream));
			launch_minmax(vol, nvox, C->d_small + SM_RANGE, C->stream);
			vmin = ord2f_host(C->fetch(SM_RANGE, 1)[0]);
		}
		McOut mc;
		if (!marching_cubes(C, vol, d, lo, hi, isolevel, cubes, originalMC, vmin, mc))
			return EXIT_FAILURE;
		*vs = host_copy_out<vec3d>(C, C->buf<double>(S_PTS0, 3 * (size_t)mc.npt), mc.npt);
		if (mc.implicit) {
			vec3i *h = (vec3i *)malloc((size_t)12 * mc.ntri);
			for (uint32_t i = 0; i < mc.ntri; i++) {
				h[i].x = 3 * i;
				h[i].y = 3 * i + 1;
				h[i].z = 3 * i + 2;
			}
			*ts = h;
		} else {
			*ts = host_copy_out<vec3i>(C, C->buf<int>(S_TRIS0, 3 * (size_t)mc.ntri), mc.ntri);
		}
		C->sync();
		*nv = (int)mc.npt;
		*nt = (int)mc.ntri;
		return EXIT_SUCCESS;
	});
}

int bwlabel(float *img, int conn, size_t dim[3], bool onlyLargest, bool fillBubbles)
{
	return guarded([&]() -> int {
		if (conn != 6 && conn != 18 && conn != 26) {
			fprintf(stderr, "bwlabel: conn must be 6, 18 or 26.\n");
			return 0;
		}
		if (dim[0] < 2 || dim[1] < 2 || dim[2] < 1) {
			fprintf(stderr, "bwlabel: img must be 2 or 3-dimensional\n");
			return 0;
		}
		n2m_ctx *C = tl_ctx.get();
		hipStream_t s = C->stream;
		const int NX = (int)dim[0], NY = (int)dim[1], NZ = (int)dim[2];
		const size_t nvox = (size_t)NX * NY * NZ, nwords = (size_t)((NX + 31) / 32) * NY * NZ;
		float *vol = C->buf<float>(S_VOL, nvox);
		HIPX(hipMemcpyAsync(vol, img, 4 * nvox, hipMemcpyHostToDevice, s));
		uint32_t *bits = C->buf<uint32_t>(S_BITS, nwords);
		launch_mask_bits(vol, NX, NY, NZ, 0.f, 1, bits, s);
		uint32_t *L = C->buf<uint32_t>(S_L, nvox), *AUX = C->buf<uint32_t>(S_AUX, nvox);
		uint32_t *roots = C->buf<uint32_t>(S_ROOTS, nvox / 2 + 1024);
		if (fillBubbles) { // src/bwlabel.c:478-541
			HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
			launch_ccl(bits, NX, NY, NZ, 6, 1, L, AUX, 1, C->d_small + SM_NROOTS, nullptr, s, nullptr);
			if (C->fetch(SM_NROOTS, 1)[0] > 1)
				launch_fill_bubbles(bits, NX, NY, NZ, L, AUX, s);
		}
		HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
		launch_ccl(bits, NX, NY, NZ, conn, 0, L, AUX, 0, C->d_small + SM_NROOTS, roots, s, nullptr);
		uint32_t nl = C->fetch(SM_NROOTS, 1)[0];
		if (nl > 0 && onlyLargest) { // src/bwlabel.c:453-471
			HIPX(hipMemsetAsync(C->d_small + SM_BEST, 0, 8, s));
			launch_ccl_argmax(roots, C->d_small + SM_NROOTS, AUX, (unsigned long long *)(C->d_small + SM_BEST), s);
			uint32_t *sel = C->buf<uint32_t>(S_BITS2, nwords);
			launch_select_root(bits, NX, NY, NZ, L, (unsigned long long *)(C->d_small + SM_BEST), sel, s);
			launch_bits_to_float(sel, NX, NY, NZ, vol, s);
			nl = 1;
		} else {
			// labels numbered by root order = first occurrence (translate_labels)
			uint32_t *k1 = C->buf<uint32_t>(S_K1, nl + 1), *v0 = C->buf<uint32_t>(S_V0, nl + 1),
			         *v1 = C->buf<uint32_t>(S_V1, nl + 1);
			uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, radix_counts_words(nl + 1));
			uint32_t *scr = scan_scratch(C, radix_counts_words(nl + 1));
			bool in1 = radix_sort_u32(roots, v0, k1, v1, nl, 32, rc, scr, s);
			launch_scatter_rank(in1 ? k1 : roots, nl, AUX, s);
			launch_write_labels(bits, NX, NY, NZ, L, AUX, vol, s);
		}
		C->check();
		HIPX(hipMemcpyAsync(img, vol, 4 * nvox, hipMemcpyDeviceToHost, s));
		C->sync();
		return (int)nl;
	}, 0);
}

float setThreshold(float *img, int nvox, int darkMediumBright123)
{
	float r = 1.0f;
	guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		float *vol = C->buf<float>(S_VOL, nvox > 0 ? nvox : 1);
		if (nvox > 0)
			HIPX(hipMemcpyAsync(vol, img, 4 * (size_t)nvox, hipMemcpyHostToDevice, C->stream));
		return set_threshold_device(C, vol, nvox, darkMediumBright123, &r);
	});
	return r;
}

} // extern "C"

namespace {

// nii_otsu(), src/isolevel.c:141-243: stays on the host (256 bins).
int otsu(const int *H, int nBin, int mode, int *dark, int *mid, int *bright)
{
	*dark = *mid = *bright = 0;
	double Sum = 0.0;
	for (int v = 0; v < nBin; v++)
		Sum = Sum + H[v];
	if (Sum <= 0)
		return 0;
	std::vector<double> P((size_t)nBin * nBin), S((size_t)nBin * nBin);
	P[0] = H[0]; // a raw count, not a probability (src/isolevel.c:157)
	S[0] = H[0];
	for (int v = 1; v < nBin; v++) {
		double prob = H[v] / Sum;
		P[v] = P[v - 1] + prob;
		S[v] = S[v - 1] + (v + 1) * prob;
	}
	for (int u = 1; u < nBin; u++)
		for (int v = u; v < nBin; v++) {
			P[(size_t)u * nBin + v] = P[v] - P[u - 1];
			S[(size_t)u * nBin + v] = S[v] - S[u - 1];
		}
	for (int u = 0; u < nBin; u++)
		for (int v = u; v < nBin; v++) {
			size_t i = (size_t)u * nBin + v;
			if (P[i] != 0)
				P[i] = (S[i] * S[i]) / P[i];
		}
	int thresh = 0;
	if (mode == 1 || mode == 5) {
		int lo = (int)(0.25 * nBin), mi = (int)(0.50 * nBin), hi = (int)(0.75 * nBin);
		double best = P[lo] + P[(size_t)(lo + 1) * nBin + mi] + P[(size_t)(mi + 1) * nBin + hi] +
		              P[(size_t)(hi + 1) * nBin + 255];
		for (int l = 0; l < nBin - 3; l++)
			for (int m = l + 1; m < nBin - 2; m++) {
				const double a = P[l] + P[(size_t)(l + 1) * nBin + m];
				const double *row = &P[(size_t)(m + 1) * nBin];
				for (int h = m + 1; h < nBin - 1; h++) {
					double v = a + row[h] + P[(size_t)(h + 1) * nBin + 255];
					if (v > best) {
						lo = l;
						mi = m;
						hi = h;
						best = v;
					}
				}
			}
		thresh = mode == 1 ? hi : lo;
		*dark = lo;
		*mid = mi;
		*bright = hi;
	} else if (mode == 2 || mode == 4) {
		int lo = (int)(0.33 * nBin), hi = (int)(0.67 * nBin);
		double best = P[lo] + P[(size_t)(lo + 1) * nBin + hi] + P[(size_t)(hi + 1) * nBin + nBin - 1];
		for (int l = 0; l < nBin - 2; l++)
			for (int h = l + 1; h < nBin - 1; h++) {
				double v = P[l] + P[(size_t)(l + 1) * nBin + h] + P[(size_t)(h + 1) * nBin + nBin - 1];
				if (v > best) {
					lo = l;
					hi = h;
					best = v;
				}
			}
		thresh = lo;
		*dark = lo;
		*mid = thresh;
		*bright = hi;
	} else {
		thresh = (int)(0.25 * nBin);
		double best = P[thresh] + P[(size_t)(thresh + 1) * nBin + nBin - 1];
		for (int i = 0; i < nBin - 1; i++) {
			double v = P[i] + P[(size_t)(i + 1) * nBin + nBin - 1];
			if (v > best) {
				thresh = i;
				best = v;
			}
		}
		*dark = *mid = *bright = thresh;
	}
	return thresh;
}

// setThreshold(): GPU range + histograms, host percentiles and Otsu.
// src/isolevel.c:35-139 (nifti_robust_range, ignoreZeroVoxels = 0), 245-277

// range + the two histograms of setThreshold(); returns false where the
// reference returns 1 early (nvox < 1)
bool threshold_hist(n2m_ctx *C, const float *d_img, int64_t nvox, ThrHist &out)
{
	hipStream_t s = C->stream;
	float mn, mx;
	{
		float pct2 = 0.0f, pct98 = 1.0f;
		if (nvox < 1)
			return false;
		uint32_t init[3] = {0xffffffffu, 0u, 0u};
		HIPX(hipMemcpyAsync(C->d_small + SM_HSTAT, init, sizeof init, hipMemcpyHostToDevice, s));
		C->tbegin("hist_range", 4.0 * nvox);
		launch_hist_range(d_img, nvox, C->d_small + SM_HSTAT, s);
		C->tend();
		uint32_t *st = C->fetch(SM_HSTAT, 3);
		size_t nNan = st[2];
		mn = ord2f_host(st[0]);
		mx = ord2f_host(st[1]);
		// zero voxels are already part of the minimum (src/isolevel.c:65-66)
		if ((size_t)nvox == nNan) {
			// all NaN: the range keeps its 0..1 defaults (src/isolevel.c:67-68)
		} else if (mn == mx) {
			pct2 = mn;
			pct98 = mx;
		} else {
			size_t valid = (size_t)nvox - nNan;
			size_t n2pct = (size_t)round(valid * 0.02);
			if (n2pct < 1 || valid < 100) {
				pct2 = mn;
				pct98 = mx;
			} else {
				const int kBins = 1001;
				float scl = (kBins - 1) / (mx - mn);
				uint32_t *hist = C->buf<uint32_t>(S_HIST, 2048);
				HIPX(hipMemsetAsync(hist, 0, 4 * kBins, s));
				C->tbegin("hist_1001", 4.0 * nvox);
				launch_hist_bins(d_img, nvox, mn, scl, kBins, hist, s);
				C->tend();
				std::vector<uint32_t> h(kBins);
				HIPX(hipMemcpyAsync(h.data(), hist, 4 * kBins, hipMemcpyDeviceToHost, s));
				C->sync();
				size_t n = 0, lo = 0;
				while (n < n2pct)
					n += h[lo++];
				lo--;
				n = 0;
				int hi = kBins;
				while (n < n2pct)
					n += h[--hi];
				if ((int)lo == hi) { // src/isolevel.c:118-134
					int ok = -1;
					while (ok != 0) {
						if (lo > 0) {
							lo--;
							if (h[lo] > 0)
								ok = 0;
						}
						if (ok != 0 && hi < kBins - 1) {
							hi++;
							if (h[hi] > 0)
								ok = 0;
						}
						if (lo == 0 && hi == kBins - 1)
							ok = 0;
					}
				}
				pct2 = (lo) / scl + mn;
				pct98 = (hi) / scl + mn;
			}
		}
		mn = pct2;
		mx = pct98;
	}
	const int kOtsuBins = 256;
	float scl = (kOtsuBins - 1) / (mx - mn);
	uint32_t *hist = C->buf<uint32_t>(S_HIST, 2048);
	HIPX(hipMemsetAsync(hist, 0, 4 * kOtsuBins, s));
	C->tbegin("hist_256", 4.0 * nvox);
	launch_hist_bins(d_img, nvox, mn, scl, kOtsuBins, hist, s);
	C->tend();
	std::vector<uint32_t> hu(kOtsuBins);
	HIPX(hipMemcpyAsync(hu.data(), hist, 4 * kOtsuBins, hipMemcpyDeviceToHost, s));
	C->sync();
	out.mn = mn;
	out.scl = scl;
	out.h.assign(hu.begin(), hu.end());
	return true;
}

// isolevels for several darkMediumBright123 modes from one set of histograms
// (src/isolevel.c:264-276: modes 1 and 3 share the 3-threshold search)
void thresholds_from(const ThrHist &H, int n, const int *modes, float *out)
{
	int d5 = -1, b5 = -1, m3 = -1, dark, mid, bright;
	for (int i = 0; i < n; i++) {
		if (modes[i] == 1 || modes[i] == 3) {
			if (d5 < 0) {
				otsu(H.h.data(), 256, 5, &dark, &mid, &bright);
				d5 = dark;
				b5 = bright;
			}
			out[i] = modes[i] == 1 ? (d5 / H.scl) + H.mn : (b5 / H.scl) + H.mn;
		} else {
			if (m3 < 0) {
				otsu(H.h.data(), 256, 3, &dark, &mid, &bright);
				m3 = mid;
			}
			out[i] = (m3 / H.scl) + H.mn;
		}
	}
}

int set_threshold_device(n2m_ctx *C, const float *d_img, int64_t nvox, int mode, float *result)
{
	ThrHist H;
	*result = 1.0f;
	if (!threshold_hist(C, d_img, nvox, H))
		return 0;
	thresholds_from(H, 1, &mode, result);
	return 0;
}

} // namespace
