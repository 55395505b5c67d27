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

#include <atomic>
#include <algorithm>
#include <functional>
#include <limits>
#include <condition_variable>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <exception>
#include <utility>
#include <vector>

#include <zlib.h>

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

} // namespace

void n2m::thresholds_from_record(const ThrRecord &r, int n, const int *modes, float *out)
{
	const int nBin = 256;
	const float mn = r.par[2], scl = r.par[3];
	double Sum = 0.0;
	for (int v = 0; v < nBin; v++)
		Sum = Sum + (double)(int)r.h256[v];
	int dark5 = 0, bright5 = 0, mid3 = 0;
	bool need5 = false, need3 = false;
	for (int i = 0; i < n; i++)
		(modes[i] == 1 || modes[i] == 3 ? need5 : need3) = true;
	if (Sum > 0) { // an empty histogram gives 0 for every threshold (src/isolevel.c:152)
		const double *probe = r.probe;
		if (need3) { // nii_otsu mode 3: one threshold (src/isolevel.c:225-239)
			int t = (int)(0.25 * nBin);
			double best = probe[t] + probe[nBin + t];
			for (int i = 0; i < nBin - 1; i++) {
				const double v = probe[i] + probe[nBin + i];
				if (v > best) {
					t = i;
					best = v;
				}
			}
			mid3 = t;
		}
		if (need5) { // nii_otsu mode 5: three thresholds (src/isolevel.c:177-200)
			int lo = (int)(0.25 * nBin), hi = (int)(0.75 * nBin);
			double best = probe[2 * nBin] + probe[2 * nBin + 1] + probe[2 * nBin + 2] + probe[2 * nBin + 3];
			for (int l = 0; l < nBin - 3; l++) // first strict maximum in l order
				if (r.bm[l] >= 0 && r.bv[l] > best) {
					lo = l;
					hi = r.bh[l];
					best = r.bv[l];
				}
			dark5 = lo;
			bright5 = hi;
		}
	}
	for (int i = 0; i < n; i++) { // src/isolevel.c:264-276
		const int t = modes[i] == 1 ? dark5 : (modes[i] == 3 ? bright5 : mid3);
		out[i] = (t / scl) + mn;
	}
}

void n2m::set_error(const char *msg)
{
	tl_err = msg;
	fprintf(stderr, "nii2mesh_amd: %s\n", msg);
}

void n2m::clear_error() { tl_err.clear(); }

namespace {

double now_ms()
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec * 1000.0 + t.tv_nsec / 1.0e6;
}

std::atomic<int> g_cubes{-1}; // process-wide marching-cubes flavour (-1: not chosen yet)

int cubes_flavour()
{
	int v = g_cubes.load();
	if (v < 0) {
		const char *e = getenv("N2M_CUBES");
		int want = (e && (e[0] == 'l' || e[0] == 'L' || e[0] == '1')) ? N2M_CUBES_LEWINER : N2M_CUBES_CLASSIC;
		g_cubes.compare_exchange_strong(v, want); // keeps a value set meanwhile
		v = g_cubes.load();
	}
	return v;
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
	S_RBASE,
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
	S_HIST,
	S_HC_TRIS,
	S_HC_INV,
	S_HC_SLOTS,
	S_HC_RANK,
	S_HC_HALO,
	S_HC_HCNT,
	S_HC_RING,
	S_HC_RCNT,
	S_HC_ROFF,
	S_HC_RSLOTS,
	S_HC_RO,
	S_AT_STATS,
	S_AT_TASKS,
	S_AT_BMAP,
	S_AT_CROPS,
	S_AT_RED,
	S_AT_VBASE,
	S_AT_VTASK,
	S_AT_RSTILES,
	S_LREP,
	S_HC_HUGE,
	S_SIGNP,
	S_SIGNN,
	S_AT_NEXT,
	S_AT_CBASE,
	S_AT_IDENT,
	S_AT_IDX,
	S_AT_OUT,
	S_AT_LABELS,
	S_DF_VREC,
	S_DF_CREC,
	S_DF_CLS,
	S_DF_REP,
	S_DF_G,
	S_DF_H,
	S_DF_GM,
	S_OTSU,
	S_ENC_A,
	S_ENC_B,
	S_ENC_C,
	S_Q_ND,
	S_Q_Q,
	S_Q_ERR,
	S_PIPE_BITS,
	S_WPOS,
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
	SM_HALO = 44,  // 1 word: a smoothing block overflowed its halo
	SM_RING = 45,  // 1 word: a smoothing block overflowed its two rings
	SM_WNSEG = 48, // 1 word: weld segment count (left on the device)
	SM_WNNEW = 49, // 1 word: welded vertex count (left on the device)
	SM_DEGK = 50,  // 1 word: triangles kept by the degenerate test (follows SM_WNNEW)
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
	int launches; // kernel launches between the two events
};

} // namespace

struct n2m_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	DevBuf bufs[S_COUNT];
	std::vector<Mesh> meshes;
	Mesh atlas_mesh;                    // union of the label meshes of the last atlas run
	std::vector<n2m_atlas_entry> atlas; // one entry per label 1..n
	struct {                            // last differentiable forward
		bool valid = false;
		const float *vol = nullptr;
		int dim[3] = {0, 0, 0};
		float iso = 0.f;
		int lapIter = 0;
		uint32_t nmc = 0; // marching-cubes vertices before the weld
		bool welded = false;
	} diff;
	uint32_t *d_small = nullptr;
	uint32_t *h_small = nullptr;
	ThrRecord *h_thr = nullptr; // pinned copy of the threshold chain's record
	bool timing = false;
	std::string timing_only; // when set, only this stage is timed
	std::vector<Stat> stats;
	std::vector<Pending> pending;
	std::vector<hipEvent_t> evpool;
	int cur_stat = -1;
	hipEvent_t cur_a = nullptr;
	double cur_bytes = 0;
	std::vector<n2m_ctx *> workers; // per-isolevel contexts of the pipeline's threaded meshing
	// large copies between pageable host memory and the device (xfer_big):
	// one stream, two pinned chunk buffers and two events per host thread
	struct XferLane {
		hipStream_t st = nullptr;
		uint8_t *buf[2] = {nullptr, nullptr};
		hipEvent_t ev[2] = {nullptr, nullptr};
	};
	std::vector<XferLane> xfer;
	void *scan_zero_p = nullptr;    // the scan scratch allocation last zeroed
	size_t scan_zero_cap = 0;

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

	// Pinned staging of host-to-device uploads: the copy runs asynchronously
	// from here, so an upload costs no round trip.  Space is handed out in
	// stream order and reclaimed at the next stream synchronisation, when
	// every copy from it has completed.
	uint8_t *h_stage = nullptr;
	size_t stage_cap = 0, stage_used = 0;

	// Waits for the stream.  It polls for up to kSpinNs first: a blocking
	// wait sleeps until an interrupt and wakes some tens of microseconds
	// after the work ends, and most waits here are the host round trips of
	// a few small counts, which end sooner than that.
	static constexpr long long kSpinNs = 200000;
	void sync()
	{
		timespec t0, t;
		clock_gettime(CLOCK_MONOTONIC, &t0);
		for (;;) {
			const hipError_t e = hipStreamQuery(stream);
			if (e == hipSuccess)
				break;
			if (e != hipErrorNotReady)
				HIPX(e);
			clock_gettime(CLOCK_MONOTONIC, &t);
			if ((t.tv_sec - t0.tv_sec) * 1000000000LL + (t.tv_nsec - t0.tv_nsec) > kSpinNs) {
				HIPX(hipStreamSynchronize(stream));
				break;
			}
		}
		stage_used = 0;
	}
	void check() { HIPX(hipGetLastError()); }

	void *stage(const void *src, size_t bytes)
	{
		const size_t need = (bytes + 255) & ~(size_t)255;
		if (stage_used + need > stage_cap) {
			sync();
			if (need > stage_cap) {
				if (h_stage)
					HIPX(hipHostFree(h_stage));
				h_stage = nullptr;
				stage_cap = 0;
				size_t want = need > ((size_t)1 << 20) ? need : ((size_t)1 << 20);
				HIPX(hipHostMalloc((void **)&h_stage, want, hipHostMallocDefault));
				stage_cap = want;
			}
		}
		void *p = h_stage + stage_used;
		stage_used += need;
		memcpy(p, src, bytes);
		return p;
	}

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
		cur_stat = -1;
		if (!timing || (!timing_only.empty() && timing_only != name))
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
	// launches: how many launches of the stage the bracket holds (a loop of
	// back-to-back launches is timed with one event pair)
	void tend(int launches = 1)
	{
		check();
		if (!timing || cur_stat < 0)
			return;
		hipEvent_t b = ev();
		HIPX(hipEventRecord(b, stream));
		pending.push_back({cur_stat, cur_a, b, cur_bytes, launches});
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
			stats[p.stat].n += p.launches;
			stats[p.stat].bytes += p.bytes;
			evpool.push_back(p.a);
			evpool.push_back(p.b);
		}
		pending.clear();
	}
	// add the (resolved) stage statistics of another context into these
	void absorb(n2m_ctx &o)
	{
		o.resolve();
		for (auto &st : o.stats) {
			size_t k = 0;
			while (k < stats.size() && stats[k].name != st.name)
				k++;
			if (k == stats.size()) {
				stats.push_back(Stat());
				stats.back().name = st.name;
			}
			stats[k].ms += st.ms;
			stats[k].n += st.n;
			stats[k].bytes += st.bytes;
		}
		o.stats.clear();
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
	for (n2m_ctx *w : c->workers)
		ctx_free(w);
	(void)hipStreamSynchronize(c->stream);
	for (auto &b : c->bufs)
		if (b.p)
			(void)hipFree(b.p);
	c->meshes.push_back(c->atlas_mesh);
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
	if (c->h_thr)
		(void)hipHostFree(c->h_thr);
	if (c->h_stage)
		(void)hipHostFree(c->h_stage);
	for (auto &x : c->xfer) {
		if (x.st)
			(void)hipStreamSynchronize(x.st);
		for (int b = 0; b < 2; b++) {
			if (x.buf[b])
				(void)hipHostFree(x.buf[b]);
			if (x.ev[b])
				(void)hipEventDestroy(x.ev[b]);
		}
		if (x.st)
			(void)hipStreamDestroy(x.st);
	}
	(void)hipStreamDestroy(c->stream);
	delete c;
}

// ---------------------------------------------------------------------------
// Large copies between pageable host memory and the device.  A plain copy
// from or into pageable memory runs at 10-15 GB/s here: its host side is
// one thread, and the pages of freshly allocated memory fault in on first
// touch.  Pinned DMA runs at about 54 GB/s, and a host copy spread over
// eight threads keeps up with it.  So the range is split over kXferLanes host
// threads, each with its own stream and two pinned chunk buffers: the DMA of
// one chunk overlaps the host copy of the other.  Synchronous: the copy is
// complete on return.  Work queued on the context's stream before the call
// is ordered before it.
// ---------------------------------------------------------------------------
constexpr size_t kXferMin = (size_t)32 << 20, kXferChunk = (size_t)4 << 20;
constexpr int kXferLanes = 8, kMaxXferLanes = 32; // N2M_XFER_LANES overrides the default

void xfer_lanes(n2m_ctx *C, int n)
{
	while ((int)C->xfer.size() < n) {
		n2m_ctx::XferLane x;
		HIPX(hipStreamCreateWithFlags(&x.st, hipStreamNonBlocking));
		C->xfer.push_back(x); // owned by the context from here (freed by ctx_free)
		n2m_ctx::XferLane &y = C->xfer.back();
		for (int b = 0; b < 2; b++) {
			HIPX(hipHostMalloc((void **)&y.buf[b], kXferChunk, hipHostMallocDefault));
			HIPX(hipEventCreateWithFlags(&y.ev[b], hipEventDisableTiming));
		}
	}
}

// d2h: dst host, src device; else dst device, src host
void xfer_big(n2m_ctx *C, void *dst, const void *src, size_t bytes, bool d2h)
{
	const hipMemcpyKind kind = d2h ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
	if (bytes < kXferMin) {
		HIPX(hipMemcpyAsync(dst, src, bytes, kind, C->stream));
		C->sync();
		return;
	}
	unsigned hc = std::thread::hardware_concurrency();
	const char *el = getenv("N2M_XFER_LANES");
	const size_t want = el && atoi(el) > 0 ? (size_t)atoi(el) : (size_t)kXferLanes;
	const int nl = (int)std::min<size_t>({want, (size_t)kMaxXferLanes, hc ? (size_t)hc : 1, bytes / kXferChunk});
	xfer_lanes(C, nl);
	hipEvent_t ready = C->ev(); // the context's queued work comes first
	HIPX(hipEventRecord(ready, C->stream));
	// lane k copies [k * per, min((k + 1) * per, bytes)), per a multiple of the chunk
	const size_t per = ((bytes + nl - 1) / nl + kXferChunk - 1) / kXferChunk * kXferChunk;
	std::vector<std::exception_ptr> err(nl);
	auto lane = [&](int k) {
		try {
			if (k > 0) // a new host thread starts on device 0
				HIPX(hipSetDevice(C->device));
			n2m_ctx::XferLane &x = C->xfer[k];
			const size_t a = (size_t)k * per, b = std::min(bytes, a + per);
			if (a >= b)
				return;
			HIPX(hipStreamWaitEvent(x.st, ready, 0));
			const size_t nc = (b - a + kXferChunk - 1) / kXferChunk;
			auto len = [&](size_t c) { return std::min(kXferChunk, b - a - c * kXferChunk); };
			auto off = [&](size_t c) { return a + c * kXferChunk; };
			if (d2h) {
				auto issue = [&](size_t c) {
					HIPX(hipMemcpyAsync(x.buf[c & 1], (const uint8_t *)src + off(c), len(c), kind, x.st));
					HIPX(hipEventRecord(x.ev[c & 1], x.st));
				};
				issue(0);
				if (nc > 1)
					issue(1);
				for (size_t c = 0; c < nc; c++) {
					HIPX(hipEventSynchronize(x.ev[c & 1]));
					memcpy((uint8_t *)dst + off(c), x.buf[c & 1], len(c));
					if (c + 2 < nc)
						issue(c + 2);
				}
			} else {
				for (size_t c = 0; c < nc; c++) {
					if (c >= 2) // the buffer's previous chunk has left it
						HIPX(hipEventSynchronize(x.ev[c & 1]));
					memcpy(x.buf[c & 1], (const uint8_t *)src + off(c), len(c));
					HIPX(hipMemcpyAsync((uint8_t *)dst + off(c), x.buf[c & 1], len(c), kind, x.st));
					HIPX(hipEventRecord(x.ev[c & 1], x.st));
				}
				HIPX(hipStreamSynchronize(x.st));
			}
		} catch (...) {
			err[k] = std::current_exception();
		}
	};
	std::vector<std::thread> th;
	for (int k = 1; k < nl; k++)
		th.emplace_back(lane, k);
	lane(0);
	for (auto &t : th)
		t.join();
	C->evpool.push_back(ready);
	for (auto &e : err)
		if (e)
			std::rethrow_exception(e);
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

// keep n mesh slots; the device arrays of the slots dropped are freed
void set_mesh_count(n2m_ctx *c, int n)
{
	c->sync(); // no queued kernel still reads a dropped mesh
	for (size_t i = (size_t)n; i < c->meshes.size(); i++) {
		Mesh &m = c->meshes[i];
		if (m.tris.p)
			HIPX(hipFree(m.tris.p));
		if (m.pts.p)
			HIPX(hipFree(m.pts.p));
		m.tris = DevBuf();
		m.pts = DevBuf();
	}
	c->meshes.resize(n);
}

// scratch of the device scans: sized for 2^31 elements up front, so a later
// scan never reallocates a buffer that queued kernels still read
uint32_t *scan_scratch(n2m_ctx *C, size_t n)
{
	size_t w = scan_scratch_words(n) + 16;
	uint32_t *p = C->buf<uint32_t>(S_SCAN, w > ((size_t)1 << 21) ? w : ((size_t)1 << 21));
	// the single-pass scan keeps the scratch's leading words zero between
	// scans: zero them once per allocation, stream-ordered before the scans
	const DevBuf &b = C->bufs[S_SCAN];
	if (b.p != C->scan_zero_p || b.cap != C->scan_zero_cap) {
		HIPX(hipMemsetAsync(p, 0, 4 * scan_scratch_zero_words(), C->stream));
		C->scan_zero_p = b.p;
		C->scan_zero_cap = b.cap;
	}
	return p;
}

// three exclusive scans of n elements (in + k*n -> out + k*n) with one host
// round trip for the three totals
void scan3(n2m_ctx *C, const uint32_t *in, uint32_t *out, size_t n, uint32_t tot[3])
{
	uint32_t *scr = scan_scratch(C, n);
	for (int k = 0; k < 3; k++)
		scan_excl_u32(in + k * n, out + k * n, n, C->d_small + SM_TOTAL + k, scr, C->stream);
	C->check();
	const uint32_t *t = C->fetch(SM_TOTAL, 3);
	for (int k = 0; k < 3; k++)
		tot[k] = t[k];
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
	// the apply stage's Lewiner sign bits of the final volume (null when not
	// written): lw_clamp(v - iso) > 0 / < 0, ceil(nx / 32) words per row
	const uint32_t *sgp = nullptr, *sgn = nullptr;
};

// spec_bits (when given) holds the threshold mask of P.isolevel, written by
// the smoothing; it is recomputed there if the range check moves the isolevel.
// known_range (when given) holds the ordered min / max words of the input on
// the host, which saves the device words' round trip.
bool volume_stages(n2m_ctx *C, float *vol, const int dim[3], const n2m_params &P, bool have_range,
                   uint32_t *spec_bits, VolOut &out, const float *src = nullptr,
                   const uint32_t *known_range = nullptr)
{
	// src (when given) holds the voxels; the stages read it and the apply
	// stage writes every voxel of vol, so vol needs no copy of src
	const float *in = src ? src : vol;
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const size_t nvox = (size_t)NX * NY * NZ;
	hipStream_t s = C->stream;
	if (!have_range && !known_range) {
		uint32_t init[2] = {0xffffffffu, 0u};
		HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
		C->tbegin("minmax", 4.0 * nvox);
		launch_minmax(in, nvox, C->d_small + SM_RANGE, s);
		C->tend();
	}
	const uint32_t *r = known_range ? known_range : C->fetch(SM_RANGE, 2);
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
		uint32_t *bits = spec_bits ? spec_bits : C->buf<uint32_t>(S_BITS, nwords);
		if (!spec_bits || iso_changed) {
			C->tbegin("mask_bits", 4.0 * nvox + nvox / 8.0);
			launch_mask_bits(in, NX, NY, NZ, isolevel, 0, bits, s);
			C->tend();
		}
		keep = bits;
		if (label_ok) {
			uint32_t *L = C->buf<uint32_t>(S_L, nvox);
			uint32_t *AUX = C->buf<uint32_t>(S_AUX, nvox);
			uint32_t *rbase = C->buf<uint32_t>(S_RBASE, nwords + 1);
			uint32_t *scr = scan_scratch(C, nwords + 1);
			Timer tm(C);
			if (P.fillBubbles) { // src/bwlabel.c:478-541
				HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
				launch_ccl(bits, NX, NY, NZ, 6, 1, L, AUX, 1, C->d_small + SM_NROOTS, nullptr, rbase, scr, s, &tm);
				C->tbegin("fill_bubbles", 4.0 * nvox);
				launch_fill_bubbles(bits, NX, NY, NZ, L, AUX, rbase, C->d_small + SM_NROOTS, s);
				C->tend();
			}
			if (P.onlyLargest) { // src/bwlabel.c:453-471
				uint32_t *roots = C->buf<uint32_t>(S_ROOTS, nvox / 2 + 1024);
				HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
				HIPX(hipMemsetAsync(C->d_small + SM_BEST, 0, 8, s));
				launch_ccl(bits, NX, NY, NZ, 18, 0, L, AUX, 0, C->d_small + SM_NROOTS, roots, rbase, scr, s, &tm);
				C->tbegin("ccl_argmax");
				launch_ccl_argmax(roots, C->d_small + SM_NROOTS, AUX, (unsigned long long *)(C->d_small + SM_BEST), s);
				C->tend();
				uint32_t *sel = C->buf<uint32_t>(S_BITS2, nwords);
				C->tbegin("select_root", 4.0 * nvox);
				launch_select_root(bits, NX, NY, NZ, L, (unsigned long long *)(C->d_small + SM_BEST), sel, rbase, s);
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
	uint32_t *sgp = nullptr, *sgn = nullptr;
	if (P.cubes == N2M_CUBES_LEWINER) { // the sign bits the classify pass reads
		const size_t sw = (size_t)((NX + 31) / 32) * NY * NZ;
		sgp = C->buf<uint32_t>(S_SIGNP, sw);
		sgn = C->buf<uint32_t>(S_SIGNN, sw);
	}
	C->tbegin("apply", 8.0 * nvox);
	const bool signs = launch_apply(in, vol, NX, NY, NZ, isolevel, mn, edgeMax, P.fillBubbles ? keep : nullptr, dil,
	                                P.fillBubbles ? 1 : 0, (int *)(C->d_small + SM_BBOX), C->d_small + SM_VMIN, s, sgp,
	                                sgn);
	C->tend();
	out.sgp = signs ? sgp : nullptr;
	out.sgn = signs ? sgn : nullptr;
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
                    int cubes, int originalMC, float vmin, McOut &o, const int *origin = nullptr,
                    const uint32_t *sgp = nullptr, const uint32_t *sgn = nullptr)
{
	static const int zero[3] = {0, 0, 0};
	if (!origin)
		origin = zero;
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
		                    C->buf<uint32_t>(S_BLOCKOFF, nb), pts, origin, s);
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
	unsigned long long *fm = C->buf<unsigned long long>(S_PFLAGS, lw_fmask_words(NX, NY, NZ, lo, hi));
	uint16_t *cr = C->buf<uint16_t>(S_CROW, npts);
	uint32_t *cnt = C->buf<uint32_t>(S_SEGCNT, 3 * nseg);
	uint32_t *base = C->buf<uint32_t>(S_SEGBASE, 3 * nseg);
	float mnv = vmin - iso;
	if (sgp && sgn) {
		C->tbegin("mc_lewiner_classify", 2.0 * npts / 8 + 3.0 * npts / 8);
		launch_lw_classify_bits(vol, NX, NY, NZ, lo, hi, iso, mnv, originalMC, sgp, sgn, cr, cnt, fm, s);
	} else {
		C->tbegin("mc_lewiner_classify", 4.0 * npts + 2.0 * npts + 3.0 * npts / 8);
		launch_lw_classify(vol, NX, NY, NZ, lo, hi, iso, mnv, originalMC, nullptr, cr, cnt, s, fm);
	}
	C->tend();
	uint32_t tot[3];
	scan3(C, cnt, base, nseg, tot);
	const uint32_t nE = tot[0], nT = tot[1], nC = tot[2];
	uint32_t V = nE + nC;
	if (V < 3 || nT < 1)
		return false;
	double *pts = C->buf<double>(S_PTS0, (size_t)3 * V);
	int *tris = C->buf<int>(S_TRIS0, (size_t)3 * nT);
	C->tbegin("mc_lewiner_emit", 2.0 * npts + 4 * 3.0 * npts / 8 + 24.0 * V + 12.0 * nT);
	launch_lw_emit(vol, NX, NY, NZ, lo, hi, iso, mnv, originalMC, fm, cr, cnt, base, nE, pts, tris, origin, s);
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
// With wpos given, a weld that merges vertices leaves tris_out unfilled and
// returns in *wpos the welded position of every input vertex (in input
// order), for remove_degenerate_welded; *wpos is null when nothing merged.
uint32_t weld(n2m_ctx *C, const double *pts, uint32_t npt, const int *tris_in /* null: implicit */, uint32_t nt,
              double *&pts_out, int *&tris_out, bool verbose, double **wpos = nullptr)
{
	if (wpos)
		*wpos = nullptr;
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
	uint32_t *lrep = C->buf<uint32_t>(S_LREP, npt);
	C->tbegin("weld_greedy", 4.0 * npt + 4.0 * npt + 24.0 * npt + 8.0 * npt);
	launch_weld_segstarts(flags, segid, npt, starts, s);
	launch_weld_greedy(starts, nseg, npt, keys, vals, pts, lcls, seeds, lrep, s);
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
	C->tbegin("weld_assign", 4.0 * npt * 5 + 4.0 * npt + 4.0 * nnew);
	launch_weld_assign(vals, lcls, flags, segid, seedbase, lrep, npt, cls, rep, s);
	C->tend();
	pts_out = C->buf<double>(S_PTS1, (size_t)3 * nnew);
	if (wpos) {
		double *wp = C->buf<double>(S_WPOS, (size_t)3 * npt);
		C->tbegin("weld_gather", 28.0 * nnew + 24.0 * nnew + 8.0 * npt + 24.0 * npt + 24.0 * npt);
		launch_gather_pts(pts, rep, nnew, pts_out, s);
		launch_weld_wpos(pts, cls, rep, npt, wp, s);
		C->tend();
		*wpos = wp;
	} else {
		C->tbegin("weld_gather", 28.0 * nnew + 24.0 * nnew);
		launch_gather_pts(pts, rep, nnew, pts_out, s);
		launch_remap_tris(tris_in, cls, nt, npt, tris_out, s);
		C->tend();
	}
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

// remove_degenerate after a merging weld(..., &wpos): the remap of the
// input triangles (tris_in, null for implicit) through the class map and the
// degenerate flags in one pass over positions in input order, then the
// compaction.  tris: the weld's tris_out buffer in, the result out.
uint32_t remove_degenerate_welded(n2m_ctx *C, const int *tris_in, uint32_t npt, const double *wpos, int *&tris,
                                  uint32_t nt, bool verbose)
{
	hipStream_t s = C->stream;
	double t0 = now_ms();
	uint32_t *keep = C->buf<uint32_t>(S_KEEP, nt), *pos = C->buf<uint32_t>(S_KEEPPOS, nt);
	C->tbegin("degenerate", 12.0 * nt + 12.0 * nt + 72.0 * nt + 4.0 * nt);
	launch_remap_degen(tris_in, C->buf<int>(S_CLS, npt), npt, wpos, nt, tris, keep, s);
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

// weld + remove_degenerate with every count left on the device (no host
// round trip between the stages): the buffers are sized for the upper bounds
// (npt vertices, nt triangles), the launches cover them and read the counts
// from the device, and a weld that merges nothing numbers each vertex as
// itself (the reference's unchanged output).  pts_out / tris_out receive the
// result arrays; the vertex and triangle counts are left in SM_WNNEW and
// SM_DEGK, for one copy to the host.  The same kernels and order of
// operations as weld() and remove_degenerate_welded().
void weld_clean_deferred(n2m_ctx *C, const double *pts, uint32_t npt, const int *tris_in, uint32_t nt,
                         double *&pts_out, int *&tris_out)
{
	hipStream_t s = C->stream;
	uint32_t *k0 = C->buf<uint32_t>(S_K0, npt), *v0 = C->buf<uint32_t>(S_V0, npt);
	uint32_t *k1 = C->buf<uint32_t>(S_K1, npt), *v1 = C->buf<uint32_t>(S_V1, npt);
	C->tbegin("weld_keys", 24.0 * npt + 8.0 * npt);
	launch_weld_keys(pts, npt, k0, v0, s);
	C->tend();
	uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, radix_counts_words(npt));
	uint32_t *scr = scan_scratch(C, radix_counts_words(npt) > npt ? radix_counts_words(npt) : npt);
	C->tbegin("radix_sort", 4 * 16.0 * npt);
	bool in1 = radix_sort_u32(k0, v0, k1, v1, npt, 32, rc, scr, s);
	C->tend();
	uint32_t *keys = in1 ? k1 : k0, *vals = in1 ? v1 : v0;
	uint32_t *flags = C->buf<uint32_t>(S_FLAGS, npt), *segid = C->buf<uint32_t>(S_SEGID, npt);
	uint32_t *d_nseg = C->d_small + SM_WNSEG, *d_nnew = C->d_small + SM_WNNEW, *d_k = C->d_small + SM_DEGK;
	C->tbegin("weld_segments", 8.0 * npt);
	launch_weld_segflags(keys, npt, flags, s);
	scan_excl_u32(flags, segid, npt, d_nseg, scr, s);
	C->tend();
	uint32_t *starts = C->buf<uint32_t>(S_STARTS, npt);
	int *lcls = C->buf<int>(S_LCLS, npt);
	uint32_t *seeds = C->buf<uint32_t>(S_SEEDS, npt), *seedbase = C->buf<uint32_t>(S_SEEDBASE, npt);
	uint32_t *lrep = C->buf<uint32_t>(S_LREP, npt);
	C->tbegin("weld_greedy", 4.0 * npt + 4.0 * npt + 24.0 * npt + 8.0 * npt);
	launch_weld_segstarts(flags, segid, npt, starts, s);
	launch_weld_greedy(starts, npt, npt, keys, vals, pts, lcls, seeds, lrep, s, d_nseg);
	scan_excl_u32(seeds, seedbase, npt, d_nnew, scr, s);
	C->tend();
	int *cls = C->buf<int>(S_CLS, npt), *rep = C->buf<int>(S_REP, npt);
	C->tbegin("weld_assign", 4.0 * npt * 5 + 4.0 * npt + 4.0 * npt);
	launch_weld_assign(vals, lcls, flags, segid, seedbase, lrep, npt, cls, rep, s, d_nnew);
	C->tend();
	pts_out = C->buf<double>(S_PTS1, (size_t)3 * npt);
	double *wpos = C->buf<double>(S_WPOS, (size_t)3 * npt);
	C->tbegin("weld_gather", 28.0 * npt + 24.0 * npt + 8.0 * npt + 24.0 * npt + 24.0 * npt);
	launch_gather_pts(pts, rep, npt, pts_out, s, d_nnew);
	launch_weld_wpos(pts, cls, rep, npt, wpos, s);
	C->tend();
	int *t1 = C->buf<int>(S_TRIS1, (size_t)3 * (nt ? nt : 1));
	uint32_t *keep = C->buf<uint32_t>(S_KEEP, nt), *pos = C->buf<uint32_t>(S_KEEPPOS, nt);
	C->tbegin("degenerate", 12.0 * nt + 12.0 * nt + 72.0 * nt + 4.0 * nt);
	launch_remap_degen(tris_in, cls, npt, wpos, nt, t1, keep, s);
	scan_excl_u32(keep, pos, nt, d_k, scan_scratch(C, nt), s);
	C->tend();
	int *out = C->buf<int>(S_TRIS0, (size_t)3 * (nt ? nt : 1));
	C->tbegin("compact_tris", 12.0 * nt + 12.0 * nt);
	launch_compact_tris(t1, keep, pos, nt, out, s);
	C->tend();
	tris_out = out;
}

// The result array p (bytes long) into the mesh buffer dst: when p is a
// scratch slot of C, the two buffers trade places (no copy; the slot reuses
// the mesh's old buffer next time), otherwise p is copied.
void adopt(n2m_ctx *C, DevBuf &dst, const void *p, size_t bytes)
{
	for (DevBuf &b : C->bufs)
		if (b.p == p) {
			std::swap(b, dst);
			return;
		}
	void *q = n2m_ctx::grow(dst, bytes);
	HIPX(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, C->stream));
}

// meshify() after the volume stages: MC, weld, de-degenerate; result into mesh m
bool surface_stages(n2m_ctx *C, const float *vol, const int dim[3], const VolOut &vo, const n2m_params &P, Mesh &m)
{
	double t0 = now_ms();
	McOut mc;
	if (!marching_cubes(C, vol, dim, vo.lo, vo.hi, vo.iso, P.cubes, P.originalMC, vo.vmin, mc, nullptr, vo.sgp,
	                    vo.sgn))
		return false;
	if (P.verbose) {
		C->sync();
		printf("marching cubes (%dx%dx%d): %ld ms\n", dim[0], dim[1], dim[2], (long)round(now_ms() - t0));
	}
	const double *pts0 = C->buf<double>(S_PTS0, (size_t)3 * mc.npt);
	const int *tris0 = mc.implicit ? nullptr : C->buf<int>(S_TRIS0, (size_t)3 * mc.ntri);
	double *pts = nullptr;
	int *tris = nullptr;
	if (!P.verbose) { // the counts come back once, at the end
		weld_clean_deferred(C, pts0, mc.npt, tris0, mc.ntri, pts, tris);
		const uint32_t *cnt = C->fetch(SM_WNNEW, 2);
		const uint32_t np = cnt[0], nt = cnt[1];
		if (np < 3) // src/meshify.c:380
			return false;
		adopt(C, m.tris, tris, (size_t)12 * nt);
		adopt(C, m.pts, pts, (size_t)24 * np);
		m.nt = (int)nt;
		m.np = (int)np;
		return true;
	}
	double *wpos = nullptr;
	uint32_t np = weld(C, pts0, mc.npt, tris0, mc.ntri, pts, tris, P.verbose != 0, &wpos);
	if (np < 3)
		return false;
	uint32_t nt = wpos ? remove_degenerate_welded(C, tris0, mc.npt, wpos, tris, mc.ntri, P.verbose != 0)
	                   : remove_degenerate(C, pts, np, tris, mc.ntri, P.verbose != 0);
	adopt(C, m.tris, tris, (size_t)12 * nt);
	adopt(C, m.pts, pts, (size_t)24 * np);
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
	bool have_range = false;
	uint32_t *spec = nullptr;
	if (P.preSmooth && NX >= 5 && NY >= 5 && NZ >= 5) { // src/meshify.c:299-304
		double t0 = now_ms();
		uint32_t init[2] = {0xffffffffu, 0u};
		HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
		float *out = vol;
		if (src == vol) // the stencil cannot run in place
			out = C->buf<float>(S_VOL2, nvox);
		if (P.onlyLargest || P.fillBubbles)
			spec = C->buf<uint32_t>(S_BITS, (size_t)((NX + 31) / 32) * NY * NZ);
		C->tbegin("smooth", 8.0 * nvox);
		launch_smooth(src, out, NX, NY, NZ, C->d_small + SM_RANGE, &P.isolevel, 1, spec, s);
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

// Vertex renumbering used by the smoothers: Morton order of the positions.
// perm[new] = old.  Returns false (identity) for small meshes.  The
// relabelled triangles (tris) are written by build_csr's count pass, from
// src (the triangles as given) and inv (old -> new).
struct Relabel {
	bool on = false;
	const uint32_t *perm = nullptr, *inv = nullptr;
	const int *src = nullptr;
	int *tris = nullptr;
};

// CSR of a device mesh (tris int[3*nt]) for smoothing
struct Csr {
	uint32_t *start;
	int2 *pairs;
	uint32_t *ref;
};

// sorted_ref: ref holds each vertex's corners in ascending order (by
// default only the pairs are ordered)
// With R (on), tris is R->tris, which the count pass writes from R->src.
Csr build_csr(n2m_ctx *C, const int *tris, uint32_t nv, uint32_t nt, bool sorted_ref = false,
              const Relabel *R = nullptr)
{
	hipStream_t s = C->stream;
	Csr c;
	uint32_t *cnt = C->buf<uint32_t>(S_HC_CNT, nv + 1);
	c.start = C->buf<uint32_t>(S_HC_START, nv + 1);
	c.ref = C->buf<uint32_t>(S_HC_REF, (size_t)3 * nt + 1);
	c.pairs = C->buf<int2>(S_HC_PAIRS, (size_t)3 * nt + 1);
	uint32_t *rank = C->buf<uint32_t>(S_HC_RANK, (size_t)3 * nt + 1);
	HIPX(hipMemsetAsync(cnt, 0, 4 * ((size_t)nv + 1), s));
	const bool rl = R && R->on;
	C->tbegin("csr_build", 12.0 * nt * 4 + 8.0 * nv + (rl ? 12.0 * nt : 0.0));
	if (rl)
		launch_csr_count(R->src, nt, cnt, rank, s, R->inv, R->tris);
	else
		launch_csr_count(tris, nt, cnt, rank, s);
	scan_excl_u32(cnt, c.start, nv + 1, nullptr, scan_scratch(C, nv + 1), s);
	launch_csr_fill(tris, nt, c.start, rank, c.ref, s);
	launch_csr_pairs(tris, nv, c.start, c.ref, c.pairs, s, sorted_ref);
	C->tend();
	// lists longer than 16 (k_csr_pairs left them): one wave each
	launch_long_sort(tris, nv, c.start, c.ref, c.pairs, C->buf<uint32_t>(S_HC_HUGE, long_huge_words(nt)), s);
	C->check();
	return c;
}

// border flags of the high-valence vertices (k_border / k_border_pairs
// skip them), one wave per list
void border_long(n2m_ctx *C, const Csr &c, const int *tris, uint32_t nv, uint32_t nt, bool literal, uint8_t *border)
{
	launch_long_border(tris, nv, c.start, c.ref, literal ? 1 : 0, border,
	                   C->buf<uint32_t>(S_HC_HUGE, long_huge_words(nt)), C->stream);
	C->check();
}

void check_tris(const int *tris, int ntri, int nvert)
{
	for (int i = 0; i < 3 * ntri; i++)
		if (tris[i] < 0 || tris[i] >= nvert)
			fail("triangle %d references vertex %d (mesh has %d vertices)", i / 3, tris[i], nvert);
}

Relabel morton_relabel(n2m_ctx *C, const double *pts, const int *tris, uint32_t nv, uint32_t nt)
{
	Relabel r;
	if (nv < 32768)
		return r;
	hipStream_t s = C->stream;
	C->tbegin("hc_relabel", 24.0 * nv * 2 + 16.0 * 4 * nv);
	uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
	HIPX(hipMemcpyAsync(C->d_small + SM_PBOX, init, sizeof init, hipMemcpyHostToDevice, s));
	launch_pts_bbox(pts, nv, C->d_small + SM_PBOX, s);
	uint32_t *k0 = C->buf<uint32_t>(S_K0, nv), *v0 = C->buf<uint32_t>(S_V0, nv);
	uint32_t *k1 = C->buf<uint32_t>(S_K1, nv), *v1 = C->buf<uint32_t>(S_V1, nv);
	launch_morton(pts, nv, C->d_small + SM_PBOX, k0, v0, s);
	uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, radix_counts_words(nv));
	bool in1 = radix_sort_u32(k0, v0, k1, v1, nv, 24, rc, scan_scratch(C, radix_counts_words(nv)), s);
	r.perm = in1 ? v1 : v0;
	uint32_t *inv = C->buf<uint32_t>(S_HC_INV, nv);
	launch_invert_perm(r.perm, nv, inv, s);
	C->tend();
	r.inv = inv;
	r.src = tris;
	r.tris = C->buf<int>(S_HC_TRIS, (size_t)3 * (nt ? nt : 1)); // written by build_csr
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
	border_long(C, c, tris, nv, nt, repeats, border);
	C->tend();
	return border;
}

// Halo form of the smoothing CSR (k_halo_build).  on = false when some
// vertex block has too many outside neighbours for its LDS halo (or
// N2M_HC_HALO=0): the window kernels are used then.
struct Halo {
	bool on = false;
	uint32_t *slots = nullptr, *halo = nullptr, *hcount = nullptr;
};

Halo build_halo(n2m_ctx *C, const Csr &c, uint32_t nv, uint32_t nt)
{
	Halo h;
	const char *e = getenv("N2M_HC_HALO");
	if ((e && *e == '0') || nv == 0)
		return h;
	hipStream_t s = C->stream;
	h.slots = C->buf<uint32_t>(S_HC_SLOTS, (size_t)3 * nt + 1);
	h.halo = C->buf<uint32_t>(S_HC_HALO, halo_words(nv));
	h.hcount = C->buf<uint32_t>(S_HC_HCNT, halo_blocks(nv));
	HIPX(hipMemsetAsync(C->d_small + SM_HALO, 0, 4, s));
	C->tbegin("hc_halo", 8.0 * nv + 12.0 * 3 * nt);
	launch_halo_build(c.start, c.pairs, nv, h.halo, h.hcount, h.slots, C->d_small + SM_HALO, s);
	C->tend();
	h.on = C->fetch(SM_HALO, 1)[0] == 0;
	return h;
}

struct Ring {
	bool on = false;
	uint32_t *slots = nullptr, *ring = nullptr, *rcnt = nullptr, *rslots = nullptr;
	uint16_t *roff = nullptr;
	double *o1 = nullptr;
	int rb = 0; // vertices per block
};

// o: the original positions in the smoothing order
Ring build_ring(n2m_ctx *C, const Csr &c, uint32_t nv, uint32_t nt, const double *o)
{
	Ring r;
	const char *e = getenv("N2M_HC_RING");
	if ((e && *e == '0') || nv == 0)
		return r;
	hipStream_t s = C->stream;
	r.rb = ring_rb();
	r.slots = C->buf<uint32_t>(S_HC_SLOTS, (size_t)3 * nt + 1);
	r.ring = C->buf<uint32_t>(S_HC_RING, ring_words(nv, r.rb));
	r.rcnt = C->buf<uint32_t>(S_HC_RCNT, 2 * (size_t)ring_blocks(nv, r.rb));
	r.roff = C->buf<uint16_t>(S_HC_ROFF, ring_off_words(nv, r.rb));
	r.rslots = C->buf<uint32_t>(S_HC_RSLOTS, ring_pair_words(nv, r.rb));
	r.o1 = C->buf<double>(S_HC_RO, 3 * ring_o_rows(nv, r.rb));
	HIPX(hipMemsetAsync(C->d_small + SM_RING, 0, 4, s));
	C->tbegin("hc_rings", 8.0 * nv + 12.0 * 3 * nt);
	launch_ring_build(c.start, c.pairs, nv, o, r.ring, r.rcnt, r.slots, r.roff, r.rslots, r.o1,
	                  C->d_small + SM_RING, r.rb, s);
	C->tend();
	r.on = C->fetch(SM_RING, 1)[0] == 0;
	return r;
}

// laplacian_smoothHC on a device mesh in place, src/quadric.c:343-394
// before_iter (when given) runs on the host after the preparation (every host
// round trip done) and before the iterations are enqueued; stat names the
// timing bracket of the one-launch iterations
void smooth_hc_device(n2m_ctx *C, double *pts, const int *tris, uint32_t nv, uint32_t nt, double alpha,
                      double beta, int iter, bool lockEdges, const std::function<void()> *before_iter = nullptr,
                      const char *stat = "hc_ring")
{
	if (nv == 0)
		return;
	hipStream_t s = C->stream;
	Relabel R = morton_relabel(C, pts, tris, nv, nt);
	const int *T = R.on ? R.tris : tris;
	Csr c = build_csr(C, T, nv, nt, false, &R);
	// one launch per iteration (two-ring form) where every block's rings fit,
	// else the halo pair of launches, else the window pair
	const size_t n3 = 3 * (size_t)nv;
	double *o = C->buf<double>(S_HC_O, n3), *p = C->buf<double>(S_HC_P, n3);
	double *pn = C->buf<double>(S_HC_PN, n3);
	launch_hc_load(pts, R.on ? R.perm : nullptr, nv, o, s);
	const Ring G = iter > 0 ? build_ring(C, c, nv, nt, o) : Ring();
	const Halo H = iter > 0 && !G.on ? build_halo(C, c, nv, nt) : Halo();
	// the border flags depend on the topology only; computed before the
	// iterations, so that every host round trip comes first and the
	// iterations are the last work enqueued
	uint8_t *border = lockEdges ? border_flags(C, c, T, nv, nt) : nullptr;
	if (before_iter)
		(*before_iter)();
	const double *res = o;
	if (G.on) {
		// algorithmic bytes: CSR offsets, one 4-byte slot pair per reference,
		// the original and current positions in, the new positions out.  The
		// iterations are back-to-back launches, timed with one event pair.
		C->tbegin(stat, (4.0 * nv + 12.0 * nt + 72.0 * nv) * iter);
		for (int it = 0; it < iter; it++) {
			double *nxt = it % 2 == 0 ? p : pn;
			launch_hc_ring(c.start, G.slots, G.roff, G.rslots, G.ring, G.rcnt, o, G.o1, res, nxt, nv, alpha, beta,
			               G.rb, s);
			res = nxt;
		}
		C->tend(iter);
	} else {
		double *bq = C->buf<double>(S_HC_B, n3);
		// algorithmic bytes: CSR offsets, one pair word per reference (4
		// bytes as LDS slots, 8 as vertex ids), each position array once,
		// the outputs
		const double pair_bytes = (H.on ? 4.0 : 8.0) * 3 * nt;
		for (int it = 0; it < iter; it++) {
			const double *cur = it == 0 ? o : p; // p starts as the original
			C->tbegin("hc_a", 4.0 * nv + pair_bytes + 24.0 * nv * 2 + 48.0 * nv);
			if (H.on)
				launch_hc_a_halo(c.start, H.slots, H.halo, H.hcount, o, cur, pn, bq, nv, alpha, s);
			else
				launch_hc_a(c.start, c.pairs, o, cur, pn, bq, nv, alpha, s);
			C->tend();
			C->tbegin("hc_b", 4.0 * nv + pair_bytes + 24.0 * nv * 2 + 24.0 * nv);
			if (H.on)
				launch_hc_b_halo(c.start, H.slots, H.halo, H.hcount, pn, bq, p, nv, beta, s);
			else
				launch_hc_b(c.start, c.pairs, pn, bq, p, nv, beta, s);
			C->tend();
			res = p;
		}
	}
	launch_hc_store(res, o, border, R.on ? R.perm : nullptr, nv, pts, s);
	C->check();
}

void laplace_device(n2m_ctx *C, double *pts, const int *tris, uint32_t nv, uint32_t nt, int iter)
{
	if (nv == 0 || iter <= 0)
		return;
	hipStream_t s = C->stream;
	Relabel R = morton_relabel(C, pts, tris, nv, nt);
	Csr c = build_csr(C, R.on ? R.tris : tris, nv, nt, false, &R);
	const Halo H = build_halo(C, c, nv, nt);
	const size_t n3 = 3 * (size_t)nv;
	double *p = C->buf<double>(S_HC_P, n3), *q = C->buf<double>(S_HC_PN, n3);
	launch_hc_load(pts, R.on ? R.perm : nullptr, nv, q, s);
	double *cur = q, *nxt = p;
	for (int it = 0; it < iter; it++) {
		C->tbegin("laplace", 4.0 * nv + (H.on ? 4.0 : 8.0) * 3 * nt + 48.0 * nv);
		if (H.on)
			launch_laplace_halo(c.start, H.slots, H.halo, H.hcount, cur, nxt, nv, s);
		else
			launch_laplace(c.start, c.pairs, cur, nxt, nv, s);
		C->tend();
		double *t = cur;
		cur = nxt;
		nxt = t;
	}
	launch_hc_store(cur, cur, nullptr, R.on ? R.perm : nullptr, nv, pts, s);
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
	if (count * sizeof(T) >= kXferMin)
		xfer_big(C, h, d, count * sizeof(T), true);
	else if (count)
		HIPX(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, C->stream));
	return h;
}

int set_thresholds_device(n2m_ctx *C, const float *d_img, int64_t nvox, int n, const int *modes, float *out);
int otsu(n2m_ctx *C, const int *H, int nBin, int mode, int *dark, int *mid, int *bright);

} // namespace

// ===========================================================================
// Atlas (src/nii2mesh.c:492-583)
// ===========================================================================
namespace {

template <typename T> void upload(n2m_ctx *C, T *d, const std::vector<T> &h)
{
	if (h.empty())
		return;
	const size_t bytes = h.size() * sizeof(T);
	HIPX(hipMemcpyAsync(d, C->stage(h.data(), bytes), bytes, hipMemcpyHostToDevice, C->stream));
}

template <typename T> std::vector<T> download(n2m_ctx *C, const T *d, size_t n)
{
	std::vector<T> h(n);
	if (n) {
		HIPX(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, C->stream));
		C->sync();
	}
	return h;
}

// d[idx[i]] for each i
std::vector<uint32_t> values_at(n2m_ctx *C, const uint32_t *d, const std::vector<uint32_t> &idx)
{
	uint32_t *di = C->buf<uint32_t>(S_AT_IDX, idx.size());
	upload(C, di, idx);
	uint32_t *dout = C->buf<uint32_t>(S_AT_OUT, idx.size());
	launch_gather_u32(d, di, (uint32_t)idx.size(), dout, C->stream);
	C->check();
	return download(C, dout, idx.size());
}

// grow b to `bytes`, keeping its first `keep` bytes
void grow_keep(n2m_ctx *C, DevBuf &b, size_t bytes, size_t keep)
{
	if (b.cap >= bytes)
		return;
	size_t want = bytes > 2 * b.cap ? bytes : 2 * b.cap;
	void *p = nullptr;
	HIPX(hipMalloc(&p, want));
	if (keep)
		HIPX(hipMemcpyAsync(p, b.p, keep, hipMemcpyDeviceToDevice, C->stream));
	C->sync();
	if (b.p)
		HIPX(hipFree(b.p));
	b.p = p;
	b.cap = want;
}

// U += (pts, tris): vertex ids of tris shifted by U's vertex count
void append_mesh(n2m_ctx *C, Mesh &U, const double *pts, uint32_t np, const int *tris, uint32_t nt)
{
	grow_keep(C, U.pts, (size_t)24 * ((size_t)U.np + np), (size_t)24 * U.np);
	grow_keep(C, U.tris, (size_t)12 * ((size_t)U.nt + nt), (size_t)12 * U.nt);
	if (np)
		HIPX(hipMemcpyAsync((double *)U.pts.p + (size_t)3 * U.np, pts, (size_t)24 * np, hipMemcpyDeviceToDevice,
		                    C->stream));
	launch_add_i32(tris, (size_t)3 * nt, U.np, (int *)U.tris.p + (size_t)3 * U.nt, C->stream);
	C->check();
	if ((size_t)U.np + np > (size_t)INT32_MAX || (size_t)U.nt + nt > (size_t)INT32_MAX)
		fail("atlas mesh exceeds 2^31 vertices or triangles");
	U.np += (int)np;
	U.nt += (int)nt;
}

// label count as the CLI finds it: trunc(max), src/nii2mesh.c:494-501
int atlas_label_count(n2m_ctx *C, const float *labels, size_t nvox)
{
	uint32_t init[2] = {0xffffffffu, 0u};
	HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, C->stream));
	launch_minmax(labels, nvox, C->d_small + SM_RANGE, C->stream);
	C->check();
	uint32_t *r = C->fetch(SM_RANGE, 2);
	float mn = ord2f_host(r[0]), mx = ord2f_host(r[1]);
	if (!(mx >= 1.0f)) {
		printf("intensity range not consistent with an indexed atlas %g..%g\n", mn, mx);
		fail("no label >= 1 in the atlas volume");
	}
	if (mx >= 16777216.0f)
		fail("atlas label %g is beyond the supported 2^24", mx);
	return (int)truncf(mx);
}

void sform_and_smooth(n2m_ctx *C, Mesh &m, int postSmooth, const float *srow)
{
	if (srow && m.np) { // src/meshify.c:1025-1045
		launch_sform((double *)m.pts.p, m.np, srow, srow + 4, srow + 8, C->stream);
		double px = srow[0] + srow[1] + srow[2];
		double py = srow[4] + srow[5] + srow[6];
		double pz = srow[8] + srow[9] + srow[10];
		float det = px * py * pz;
		if (det < 0.0)
			launch_swap_winding((int *)m.tris.p, m.nt, C->stream);
		C->check();
	}
	if (postSmooth > 0 && m.np) // src/nii2mesh.c:327-331
		smooth_hc_device(C, (double *)m.pts.p, (const int *)m.tris.p, m.np, m.nt, 0.1, 0.5, postSmooth, true);
}

// Launch shapes of the batched atlas.  Every stage is a persistent launch of
// `blocks per CU` x CUs workgroups over its block map (0: one workgroup per
// item); the smoothing also picks its crop tile.  The first atlas run on a
// device times every candidate on its own crops and keeps the fastest for
// the process.  N2M_ATLAS_TILE / N2M_ATLAS_BPC force a tile / one blocks-per-CU
// value for every stage.
constexpr int kAtlasBpc[] = {0, 1, 2, 4, 8};
constexpr int kNBpc = (int)(sizeof kAtlasBpc / sizeof kAtlasBpc[0]);
enum AtlasStage { AST_SMOOTH, AST_APPLY, AST_CCOUNT, AST_CEMIT, AST_LWCLS, AST_LWEMIT, AST_N };
const char *const kAtlasStageName[AST_N] = {"smooth", "apply", "classic_count", "classic_emit", "lewiner_classify",
                                            "lewiner_emit"};
std::atomic<int> g_atlas_cfg[64][AST_N]; // per device: tuned configuration + 1, 0 until tuned

int env_choice(const char *name, int n)
{
	const char *e = getenv(name);
	if (!e || !*e)
		return -1;
	int v = atoi(e);
	if (v < 0 || v >= n)
		fail("%s=%s: expected 0..%d", name, e, n - 1);
	return v;
}

int device_cus(int dev)
{
	static std::atomic<int> cus[64];
	const int d = dev >= 0 && dev < 64 ? dev : 0;
	int n = cus[d].load(std::memory_order_relaxed);
	if (!n) {
		HIPX(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
		cus[d].store(n, std::memory_order_relaxed);
	}
	return n;
}

// workgroups of a persistent launch over nwork items at blocks-per-CU index k
unsigned atlas_grid(n2m_ctx *C, unsigned nwork, int k)
{
	if (!kAtlasBpc[k])
		return nwork;
	const size_t g = (size_t)kAtlasBpc[k] * device_cus(C->device);
	return g < nwork ? (unsigned)g : nwork;
}

// configuration of a stage: cfg = tile * kNBpc + blocks-per-CU index (tile 0
// for the stages without one).  run(cfg) must be safe to repeat.
template <typename Run> int atlas_tune(n2m_ctx *C, int stage, int ntiles, Run run)
{
	const int ft = ntiles > 1 ? env_choice("N2M_ATLAS_TILE", ntiles) : -1;
	const int fb = env_choice("N2M_ATLAS_BPC", kNBpc);
	std::vector<int> cand;
	for (int t = 0; t < ntiles; t++)
		for (int k = 0; k < kNBpc; k++)
			if ((ft < 0 || t == ft) && (fb < 0 || k == fb))
				cand.push_back(t * kNBpc + k);
	if (cand.size() == 1)
		return cand[0];
	const int dev = C->device >= 0 && C->device < 64 ? C->device : 0;
	const int known = g_atlas_cfg[dev][stage].load(std::memory_order_relaxed);
	if (known > 0 && std::find(cand.begin(), cand.end(), known - 1) != cand.end())
		return known - 1;
	hipEvent_t a = C->ev(), b = C->ev();
	float best = 0.f;
	int pick = cand[0];
	for (size_t i = 0; i < cand.size(); i++) {
		run(cand[i]); // warms the code object and the caches
		float ms = 0.f;
		for (int r = 0; r < 2; r++) {
			HIPX(hipEventRecord(a, C->stream));
			run(cand[i]);
			HIPX(hipEventRecord(b, C->stream));
			HIPX(hipEventSynchronize(b));
			float m = 0.f;
			HIPX(hipEventElapsedTime(&m, a, b));
			ms = r == 0 || m < ms ? m : ms;
		}
		if (i == 0 || ms < best) {
			best = ms;
			pick = cand[i];
		}
	}
	C->evpool.push_back(a);
	C->evpool.push_back(b);
	if (ft < 0 && fb < 0)
		g_atlas_cfg[dev][stage].store(pick + 1, std::memory_order_relaxed);
	return pick;
}

// Tasks (labels) whose crops share one pass.  entry[k] = index of task k's
// label in C->atlas.  Meshes of the tasks are appended to U in task order.
void atlas_chunk(n2m_ctx *C, const float *labels, const int dim[3], const n2m_params &P, std::vector<AtlasTask> T,
                 std::vector<int> entry, Mesh &U)
{
	hipStream_t s = C->stream;
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const bool smooth = P.preSmooth && NX >= 5 && NY >= 5 && NZ >= 5; // src/meshify.c:171
	auto failed = [&](int k) { C->atlas[entry[k]].status = N2M_ATLAS_FAILED; };
	size_t nfl = 0;
	for (auto &t : T) {
		t.off = (long long)nfl;
		nfl += (size_t)t.nx * t.ny * t.nz;
	}
	float *crops = C->buf<float>(S_AT_CROPS, nfl);
	std::vector<uint32_t> bmap;
	auto plan = [&](auto blocks_of) {
		bmap.clear();
		for (size_t k = 0; k < T.size(); k++) {
			T[k].first_block = (int)bmap.size();
			bmap.insert(bmap.end(), (size_t)blocks_of(T[k]), (uint32_t)k);
		}
		if (bmap.size() >= ((size_t)1 << 31))
			fail("atlas launch of %zu workgroups", bmap.size());
		upload(C, C->buf<AtlasTask>(S_AT_TASKS, T.size()), T);
		upload(C, C->buf<uint32_t>(S_AT_BMAP, bmap.size()), bmap);
		return (unsigned)bmap.size();
	};
	auto dT = [&] { return C->buf<AtlasTask>(S_AT_TASKS, T.size()); };
	auto dmap = [&] { return C->buf<uint32_t>(S_AT_BMAP, bmap.size()); };

	// 1. binarise + pre-smooth every crop, per-task range (src/meshify.c:299-311)
	uint32_t *red = C->buf<uint32_t>(S_AT_RED, (size_t)9 * T.size());
	std::vector<uint32_t> rinit(2 * T.size());
	for (size_t k = 0; k < T.size(); k++) {
		rinit[2 * k] = 0xffffffffu;
		rinit[2 * k + 1] = 0u;
	}
	int planned_tile = -1;
	unsigned nb = 0;
	auto smooth_with = [&](int cfg) {
		const int tile = cfg / kNBpc;
		if (tile != planned_tile) {
			nb = plan([&](const AtlasTask &t) { return atlas_smooth_blocks(t.nx, t.ny, t.nz, smooth, tile); });
			planned_tile = tile;
		}
		upload(C, red, rinit);
		launch_atlas_smooth(labels, NX, NY, NZ, dT(), dmap(), nb, atlas_grid(C, nb, cfg % kNBpc), smooth ? 1 : 0,
		                    tile, crops, red, s);
		C->check();
	};
	const int scfg = atlas_tune(C, AST_SMOOTH, smooth ? atlas_tile_count() : 1, smooth_with);
	C->tbegin("atlas_smooth", 8.0 * nfl);
	smooth_with(scfg);
	C->tend();

	// 2. edge darkening, box of >= iso, minimum (src/meshify.c:347-371).  The
	// isolevels come from the ranges on the device (atlas_iso), so the
	// ranges, boxes and minima come back in one copy; a task without
	// variability runs with a NaN isolevel (it darkens nothing and finds no
	// voxel >= iso) and is dropped afterwards.
	{
		const size_t n0 = T.size();
		nb = plan([&](const AtlasTask &t) { return atlas_apply_blocks(t.nx, t.ny, t.nz); });
		launch_atlas_iso(dT(), (int)n0, red, P.isolevel, s);
		std::vector<uint32_t> init(7 * n0);
		for (size_t k = 0; k < n0; k++) {
			for (int j = 0; j < 3; j++) {
				init[6 * k + j] = 0x7fffffffu;
				init[6 * k + 3 + j] = 0xffffffffu; // -1
			}
			init[6 * n0 + k] = 0xffffffffu;
		}
		upload(C, red + 2 * n0, init);
		int *bbox = (int *)red + 2 * n0;
		uint32_t *vmin = red + 8 * n0;
		// repeatable for the tuner: the darkening is a min, and a repeat sees
		// a subset of the voxels >= iso, so the reduced box and minimum stay
		auto apply_with = [&](int cfg) {
			launch_atlas_apply(crops, NX, NY, NZ, dT(), dmap(), nb, atlas_grid(C, nb, cfg), bbox, vmin, s);
			C->check();
		};
		const int acfg = atlas_tune(C, AST_APPLY, 1, apply_with);
		C->tbegin("atlas_apply", 8.0 * nfl);
		apply_with(acfg);
		C->tend();
		const std::vector<uint32_t> rb = download(C, red, 9 * n0);
		const uint32_t *rr = rb.data(), *bb = rb.data() + 2 * n0;
		std::vector<AtlasTask> L;
		std::vector<int> LE;
		std::vector<size_t> LK; // each kept task's index in T
		for (size_t k = 0; k < n0; k++) {
			AtlasTask t = T[k];
			const float mn = ord2f_host(rr[2 * k]), mx = ord2f_host(rr[2 * k + 1]);
			bool fallback = false;
			if (!atlas_iso(P.isolevel, mn, mx, &t.iso, &t.edgeMax, &fallback)) { // src/meshify.c:312-315
				printf("Error: No variability in image intensity.\n");
				failed((int)k);
				continue;
			}
			if (fallback) // src/meshify.c:316-319
				printf("Suggested isolevel out of range. Intensity range %g..%g, setting isolevel to %g\n", mn, mx,
				       t.iso);
			if (P.verbose)
				printf("intensity range %g..%g, isolevel %g\n", mn, mx, t.iso);
			t.mn = mn;
			L.push_back(t);
			LE.push_back(entry[k]);
			LK.push_back(k);
		}
		for (size_t j = 0; j < L.size(); j++) {
			const size_t k = LK[j];
			AtlasTask &t = L[j];
			const int org[3] = {t.ox, t.oy, t.oz}, csz[3] = {t.nx, t.ny, t.nz};
			t.mnv = ord2f_host(bb[6 * n0 + k]) - t.iso;
			if ((int)bb[6 * k + 3] < 0) {
				// no voxel >= iso: cannot happen (the maximum is >= iso); the
				// reference's marching cubes would find nothing
				printf("marching cubes failed to identify triangles with an isolevel of %g\n", t.iso);
				C->atlas[LE[j]].status = N2M_ATLAS_FAILED;
				for (int i = 0; i < 3; i++)
					t.lo3[i] = t.hi3[i] = 0;
			} else {
				for (int i = 0; i < 3; i++) {
					int glo = (int)bb[6 * k + i] + org[i], ghi = (int)bb[6 * k + 3 + i] + org[i];
					glo = glo - 1 > 0 ? glo - 1 : 0;
					ghi = ghi + 2 < dim[i] ? ghi + 2 : dim[i];
					t.lo3[i] = glo - org[i];
					t.hi3[i] = ghi - org[i];
					// the crop holds the label's box + 4 (clipped), the box of
					// >= iso lies within the label's box + 2, so lo3 >= 0 and
					// hi3 <= csz
					if (t.lo3[i] < 0 || t.hi3[i] > csz[i])
						fail("atlas label %d: box outside its crop", t.label);
				}
			}
		}
		T.swap(L);
		entry.swap(LE);
	}
	if (T.empty())
		return;
	const size_t nt_ = T.size();

	// 3. marching cubes
	uint32_t Vtot = 0, Ttot = 0;
	bool implicit = P.cubes == N2M_CUBES_CLASSIC;
	std::vector<uint32_t> vbase(nt_ + 1), tbase(nt_ + 1);
	if (implicit) { // src/oldcubes.c:465-522
		size_t cubes = 0;
		nb = plan([&](AtlasTask &t) {
			int cx = t.hi3[0] - 1 - t.lo3[0], cy = t.hi3[1] - 1 - t.lo3[1], cz = t.hi3[2] - 1 - t.lo3[2];
			size_t nc = (cx > 0 && cy > 0 && cz > 0) ? (size_t)cx * cy * cz : 0;
			t.cube_off = (long long)cubes;
			cubes += nc;
			return nc ? (unsigned)classic_blocks(cx, cy, cz) : 0u;
		});
		uint8_t *cfg = C->buf<uint8_t>(S_CFG, cubes);
		uint32_t *bs = C->buf<uint32_t>(S_BLOCK, (size_t)nb + 1), *bo = C->buf<uint32_t>(S_BLOCKOFF, (size_t)nb + 1);
		HIPX(hipMemsetAsync(bs + nb, 0, 4, s));
		auto count_with = [&](int k) {
			launch_atlas_classic_count(crops, dT(), dmap(), nb, atlas_grid(C, nb, k), cfg, bs, s);
			C->check();
		};
		const int ccfg = atlas_tune(C, AST_CCOUNT, 1, count_with);
		C->tbegin("atlas_mc_classic_count", 4.0 * cubes + cubes);
		count_with(ccfg);
		C->tend();
		scan_excl_u32(bs, bo, (size_t)nb + 1, nullptr, scan_scratch(C, (size_t)nb + 1), s);
		C->check();
		std::vector<uint32_t> idx(nt_ + 1);
		for (size_t k = 0; k < nt_; k++)
			idx[k] = (uint32_t)T[k].first_block;
		idx[nt_] = nb;
		tbase = values_at(C, bo, idx);
		Ttot = tbase[nt_];
		for (size_t k = 0; k <= nt_; k++)
			vbase[k] = 3 * tbase[k];
		Vtot = 3 * Ttot;
		for (size_t k = 0; k < nt_; k++)
			if (tbase[k + 1] == tbase[k]) {
				printf("marching cubes failed to identify triangles with an isolevel of %g\n", T[k].iso);
				failed((int)k);
			}
		if (Ttot) {
			double *pts = C->buf<double>(S_PTS0, (size_t)9 * Ttot);
			auto emit_with = [&](int k) {
				launch_atlas_classic_emit(crops, dT(), dmap(), nb, atlas_grid(C, nb, k), cfg, bo, pts, s);
				C->check();
			};
			const int ecfg = atlas_tune(C, AST_CEMIT, 1, emit_with);
			C->tbegin("atlas_mc_classic_emit", 1.0 * cubes + 72.0 * Ttot);
			emit_with(ecfg);
			C->tend();
		}
	} else { // src/MarchingCubes.c:1086-1143
		size_t npts = 0;
		const int seg = lw_seg_len();
		nb = plan([&](AtlasTask &t) {
			int sx = t.hi3[0] - t.lo3[0] + 1, sy = t.hi3[1] - t.lo3[1] + 1, sz = t.hi3[2] - t.lo3[2] + 1;
			t.nseg = (sx + seg - 1) / seg;
			t.cube_off = (long long)npts;
			npts += (size_t)sx * sy * sz;
			return (unsigned)((size_t)t.nseg * sy * sz);
		});
		const size_t stride = (size_t)nb + 1;
		uint8_t *pf = C->buf<uint8_t>(S_PFLAGS, npts);
		uint16_t *cr = C->buf<uint16_t>(S_CROW, npts);
		uint32_t *cnt = C->buf<uint32_t>(S_SEGCNT, 3 * stride), *base = C->buf<uint32_t>(S_SEGBASE, 3 * stride);
		for (int j = 0; j < 3; j++)
			HIPX(hipMemsetAsync(cnt + j * stride + nb, 0, 4, s));
		auto classify_with = [&](int k) {
			launch_atlas_lw_classify(crops, dT(), dmap(), nb, atlas_grid(C, nb, k), P.originalMC, pf, cr, cnt, stride,
			                         s);
			C->check();
		};
		const int lcfg = atlas_tune(C, AST_LWCLS, 1, classify_with);
		C->tbegin("atlas_mc_lewiner_classify", 7.0 * npts);
		classify_with(lcfg);
		C->tend();
		for (int j = 0; j < 3; j++)
			scan_excl_u32(cnt + j * stride, base + j * stride, stride, nullptr, scan_scratch(C, stride), s);
		C->check();
		std::vector<uint32_t> idx(3 * (nt_ + 1));
		for (int j = 0; j < 3; j++) {
			for (size_t k = 0; k < nt_; k++)
				idx[j * (nt_ + 1) + k] = (uint32_t)(j * stride + T[k].first_block);
			idx[j * (nt_ + 1) + nt_] = (uint32_t)(j * stride + nb);
		}
		std::vector<uint32_t> v = values_at(C, base, idx);
		const uint32_t *E = v.data(), *Tb = v.data() + (nt_ + 1), *Cb = v.data() + 2 * (nt_ + 1);
		for (size_t k = 0; k <= nt_; k++) {
			vbase[k] = E[k] + Cb[k];
			tbase[k] = Tb[k];
		}
		Vtot = vbase[nt_];
		Ttot = tbase[nt_];
		for (size_t k = 0; k < nt_; k++)
			if (vbase[k + 1] - vbase[k] < 3 || tbase[k + 1] == tbase[k])
				failed((int)k);
		if (Vtot) {
			double *pts = C->buf<double>(S_PTS0, (size_t)3 * Vtot);
			int *tris = C->buf<int>(S_TRIS0, (size_t)3 * (Ttot ? Ttot : 1));
			auto emit_with = [&](int k) {
				launch_atlas_lw_emit(crops, dT(), (int)nt_, dmap(), nb, atlas_grid(C, nb, k), P.originalMC, pf, cr, cnt,
				                     base, stride, E[nt_], pts, tris, s);
				C->check();
			};
			const int ecfg = atlas_tune(C, AST_LWEMIT, 1, emit_with);
			C->tbegin("atlas_mc_lewiner_emit", 3.0 * npts + 24.0 * Vtot + 12.0 * Ttot);
			emit_with(ecfg);
			C->tend();
		}
	}
	if (!Vtot)
		return;

	// 4. weld each task's mesh on its own (src/meshify.c:45-106)
	const double *pts0 = C->buf<double>(S_PTS0, (size_t)3 * Vtot);
	uint32_t *dvb = C->buf<uint32_t>(S_AT_VBASE, nt_ + 1);
	upload(C, dvb, vbase);
	uint32_t *k0 = C->buf<uint32_t>(S_K0, Vtot), *v0 = C->buf<uint32_t>(S_V0, Vtot);
	uint32_t *k1 = C->buf<uint32_t>(S_K1, Vtot), *v1 = C->buf<uint32_t>(S_V1, Vtot);
	uint32_t *vtask = C->buf<uint32_t>(S_AT_VTASK, Vtot);
	// every task's vertices sorted on their own: one segmented sort, the
	// tasks keep their places, so the task of sorted position i is vtask[i]
	const std::vector<RsTile> tiles = radix_seg_tiles(vbase);
	RsTile *dtiles = C->buf<RsTile>(S_AT_RSTILES, tiles.size());
	upload(C, dtiles, tiles);
	uint32_t *rc = C->buf<uint32_t>(S_RCOUNT, 256 * tiles.size() + 1);
	uint32_t *scr = scan_scratch(C, 256 * tiles.size() + 1);
	C->tbegin("atlas_weld_sort", 24.0 * Vtot + 12.0 * Vtot + 4 * 16.0 * Vtot);
	launch_atlas_weld_keys(pts0, Vtot, dvb, (int)nt_, k0, v0, vtask, s);
	bool in1 = radix_sort_seg_u32(k0, v0, k1, v1, Vtot, 32, dtiles, tiles.size(), rc, scr, s);
	uint32_t *keys = in1 ? k1 : k0, *vals = in1 ? v1 : v0;
	const uint32_t *TK = vtask;
	C->tend();
	uint32_t *flags = C->buf<uint32_t>(S_FLAGS, Vtot), *segid = C->buf<uint32_t>(S_SEGID, Vtot);
	C->tbegin("atlas_weld_segments", 12.0 * Vtot);
	launch_atlas_weld_segflags(keys, TK, Vtot, flags, s);
	C->tend();
	// The counts stay on the device until one copy at the end: the launches
	// cover their upper bound Vtot.
	uint32_t *d_nseg = C->d_small + SM_WNSEG, *d_nnew = C->d_small + SM_WNNEW;
	scan_excl_u32(flags, segid, Vtot, d_nseg, scr, s);
	uint32_t *starts = C->buf<uint32_t>(S_STARTS, Vtot);
	int *lcls = C->buf<int>(S_LCLS, Vtot);
	uint32_t *seeds = C->buf<uint32_t>(S_SEEDS, Vtot), *seedbase = C->buf<uint32_t>(S_SEEDBASE, Vtot);
	uint32_t *lrep = C->buf<uint32_t>(S_LREP, Vtot);
	C->tbegin("atlas_weld_greedy", 4.0 * Vtot + 4.0 * Vtot + 24.0 * Vtot + 8.0 * Vtot);
	launch_weld_segstarts(flags, segid, Vtot, starts, s);
	launch_weld_greedy(starts, Vtot, Vtot, keys, vals, pts0, lcls, seeds, lrep, s, d_nseg);
	C->tend();
	scan_excl_u32(seeds, seedbase, Vtot, d_nnew, scr, s);
	// per task: next task with vertices (ntasks for none)
	std::vector<uint32_t> nextne(nt_);
	for (size_t k = nt_, nx = nt_; k-- > 0;) {
		nextne[k] = (uint32_t)nx;
		if (vbase[k + 1] != vbase[k])
			nx = k;
	}
	uint32_t *dnext = C->buf<uint32_t>(S_AT_NEXT, nt_);
	upload(C, dnext, nextne);
	// result words: [0, nt_] kept-triangle bases, [nt_ + 1, 2 nt_ + 1)
	// class bases, [2 nt_ + 1, 3 nt_ + 1) class counts, then the total
	uint32_t *res = C->buf<uint32_t>(S_AT_OUT, 3 * nt_ + 2);
	uint32_t *dcb = res + nt_ + 1, *dncls = res + 2 * nt_ + 1;
	uint8_t *di = C->buf<uint8_t>(S_AT_IDENT, nt_);
	int *cls = C->buf<int>(S_CLS, Vtot), *rep = C->buf<int>(S_REP, Vtot);
	C->tbegin("atlas_weld_assign", 4.0 * Vtot * 6 + 4.0 * Vtot + 4.0 * Vtot);
	launch_atlas_class_base(TK, segid, seedbase, Vtot, dcb, s);
	launch_weld_assign(vals, lcls, flags, segid, seedbase, lrep, Vtot, cls, rep, s);
	launch_atlas_weld_tasks(dcb, dvb, dnext, (int)nt_, d_nnew, dncls, di, s);
	launch_atlas_weld_identity(Vtot, vtask, dvb, dcb, di, cls, rep, s);
	C->tend();
	double *pts1 = C->buf<double>(S_PTS1, (size_t)3 * Vtot);
	int *tris1 = C->buf<int>(S_TRIS1, (size_t)3 * (Ttot ? Ttot : 1));
	C->tbegin("atlas_weld_gather", 28.0 * Vtot + 24.0 * Vtot + 24.0 * Ttot);
	launch_gather_pts(pts0, rep, Vtot, pts1, s, d_nnew);
	launch_remap_tris(implicit ? nullptr : C->buf<int>(S_TRIS0, (size_t)3 * (Ttot ? Ttot : 1)), cls, Ttot, Vtot,
	                  tris1, s);
	C->tend();

	// 5. degenerate triangles (src/meshify.c:113-168); every class number is
	// below the class count, so Vtot bounds the vertex ids as well
	uint32_t *keep = C->buf<uint32_t>(S_KEEP, (size_t)Ttot + 1), *kpos = C->buf<uint32_t>(S_KEEPPOS, (size_t)Ttot + 1);
	HIPX(hipMemsetAsync(keep + Ttot, 0, 4, s));
	C->tbegin("atlas_degenerate", 12.0 * Ttot + 72.0 * Ttot + 4.0 * Ttot);
	if (Ttot)
		launch_degen_flags(pts1, tris1, Ttot, Vtot, keep, s);
	C->tend();
	scan_excl_u32(keep, kpos, (size_t)Ttot + 1, nullptr, scan_scratch(C, (size_t)Ttot + 1), s);
	uint32_t *dtb = C->buf<uint32_t>(S_AT_IDX, nt_ + 1);
	upload(C, dtb, tbase);
	launch_gather_u32(kpos, dtb, (uint32_t)(nt_ + 1), res, s);
	HIPX(hipMemcpyAsync(res + 3 * nt_ + 1, d_nnew, 4, hipMemcpyDeviceToDevice, s));
	C->check();
	const std::vector<uint32_t> r = download(C, res, 3 * nt_ + 2);
	const std::vector<uint32_t> kb(r.begin(), r.begin() + nt_ + 1);
	const uint32_t *cb = r.data() + nt_ + 1, *ncls = r.data() + 2 * nt_ + 1;
	const uint32_t nnew = r[3 * nt_ + 1];
	for (size_t k = nt_; k-- > 0;) {
		if (vbase[k + 1] == vbase[k])
			continue;
		if (ncls[k] == vbase[k + 1] - vbase[k] && P.verbose)
			printf("Unify vertices found no shared vertices\n");
		if (ncls[k] < 3) // src/meshify.c:384
			failed((int)k);
	}
	const uint32_t kept = kb[nt_];
	int *tris2 = C->buf<int>(S_TRIS0, (size_t)3 * (kept ? kept : 1));
	if (kept && Ttot)
		launch_compact_tris(tris1, keep, kpos, Ttot, tris2, s);
	C->check();

	// 6. append to the union mesh
	const uint32_t v0u = (uint32_t)U.np, t0u = (uint32_t)U.nt;
	append_mesh(C, U, pts1, nnew, tris2, kept);
	for (size_t k = 0; k < nt_; k++) {
		n2m_atlas_entry &e = C->atlas[entry[k]];
		e.tri_offset = (int)(t0u + kb[k]);
		e.nt = (int)(kb[k + 1] - kb[k]);
		e.vert_offset = (int)(v0u + cb[k]);
		e.np = (int)ncls[k];
		if (e.status != N2M_ATLAS_FAILED)
			e.status = N2M_ATLAS_OK;
	}
}

// chunk budget in crop voxels (env N2M_ATLAS_CHUNK overrides)
size_t atlas_chunk_voxels()
{
	const char *e = getenv("N2M_ATLAS_CHUNK");
	if (e && atoll(e) > 0)
		return (size_t)atoll(e);
	return (size_t)1 << 29; // 2 GiB of crops
}

int atlas_batched(n2m_ctx *C, const float *labels, const int dim[3], const n2m_params &prm, int postSmooth,
                  const float *srow)
{
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	const size_t nvox = (size_t)NX * NY * NZ;
	hipStream_t s = C->stream;
	n2m_params P = prm;
	P.onlyLargest = 0; // src/nii2mesh.c:493
	const int nL = atlas_label_count(C, labels, nvox);
	C->atlas.assign(nL, n2m_atlas_entry{0, N2M_ATLAS_EMPTY, 0, 0, 0, 0});
	Mesh &U = C->atlas_mesh;
	U.nt = U.np = 0;
	// one pass: voxel count and box of every label
	uint32_t *st = C->buf<uint32_t>(S_AT_STATS, (size_t)7 * (nL + 1));
	{
		std::vector<uint32_t> init((size_t)7 * (nL + 1));
		for (int i = 0; i <= nL; i++) {
			init[7 * i] = 0;
			init[7 * i + 1] = init[7 * i + 2] = init[7 * i + 3] = 0xffffffffu;
			init[7 * i + 4] = init[7 * i + 5] = init[7 * i + 6] = 0u;
		}
		upload(C, st, init);
	}
	C->tbegin("atlas_stats", 4.0 * nvox);
	launch_atlas_stats(labels, NX, NY, NZ, nL, st, s);
	C->tend();
	std::vector<uint32_t> hs = download(C, st, (size_t)7 * (nL + 1));
	std::vector<AtlasTask> tasks;
	std::vector<int> entry;
	for (int i = 1; i <= nL; i++) {
		n2m_atlas_entry &e = C->atlas[i - 1];
		e.label = i;
		const uint32_t *q = &hs[(size_t)7 * i];
		if (!q[0]) { // src/nii2mesh.c:564-567
			if (P.verbose)
				printf("Skipping %d: no voxels with this intensity\n", i);
			continue;
		}
		AtlasTask t;
		memset(&t, 0, sizeof t);
		t.label = i;
		t.lo = (float)((double)i - 0.5); // src/nii2mesh.c:555-556
		t.hi = (float)((double)i + 0.5);
		const int lo[3] = {(int)q[1], (int)q[2], (int)q[3]}, hi[3] = {(int)q[4], (int)q[5], (int)q[6]};
		int o[3], n[3];
		for (int a = 0; a < 3; a++) {
			o[a] = lo[a] - 4 > 0 ? lo[a] - 4 : 0;
			int e2 = hi[a] + 4 < dim[a] - 1 ? hi[a] + 4 : dim[a] - 1;
			n[a] = e2 - o[a] + 1;
		}
		t.ox = o[0];
		t.oy = o[1];
		t.oz = o[2];
		t.nx = n[0];
		t.ny = n[1];
		t.nz = n[2];
		tasks.push_back(t);
		entry.push_back(i - 1);
	}
	// chunks of tasks whose crops fit the budget
	const size_t budget = atlas_chunk_voxels();
	size_t b = 0;
	while (b < tasks.size()) {
		size_t e = b, vox = 0;
		while (e < tasks.size()) {
			size_t v = (size_t)tasks[e].nx * tasks[e].ny * tasks[e].nz;
			if (e > b && vox + v > budget)
				break;
			vox += v;
			e++;
		}
		atlas_chunk(C, labels, dim, P, std::vector<AtlasTask>(tasks.begin() + b, tasks.begin() + e),
		            std::vector<int>(entry.begin() + b, entry.begin() + e), U);
		b = e;
	}
	sform_and_smooth(C, U, postSmooth, srow);
	C->sync();
	int ok = 0;
	for (auto &e : C->atlas)
		ok += e.status == N2M_ATLAS_OK;
	return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

// the reference's per-label loop on the GPU (src/nii2mesh.c:551-573)
int atlas_loop(n2m_ctx *C, const float *labels, const int dim[3], const n2m_params &prm, int postSmooth,
               const float *srow)
{
	const size_t nvox = (size_t)dim[0] * dim[1] * dim[2];
	hipStream_t s = C->stream;
	n2m_params P = prm;
	P.onlyLargest = 0;
	const int nL = atlas_label_count(C, labels, nvox);
	C->atlas.assign(nL, n2m_atlas_entry{0, N2M_ATLAS_EMPTY, 0, 0, 0, 0});
	Mesh &U = C->atlas_mesh;
	U.nt = U.np = 0;
	float *bin = C->buf<float>(S_VOL, nvox);
	Mesh &m = mesh_slot(C, 0);
	int ok = 0;
	for (int i = 1; i <= nL; i++) {
		n2m_atlas_entry &e = C->atlas[i - 1];
		e.label = i;
		HIPX(hipMemsetAsync(C->d_small + SM_FLAG, 0, 4, s));
		C->tbegin("atlas_binarize", 8.0 * nvox);
		launch_binarize(labels, nvox, (float)((double)i - 0.5), (float)((double)i + 0.5), bin, C->d_small + SM_FLAG,
		                s);
		C->tend();
		if (C->fetch(SM_FLAG, 1)[0] == 0) {
			if (P.verbose)
				printf("Skipping %d: no voxels with this intensity\n", i);
			continue;
		}
		m.nt = m.np = 0;
		if (!meshify_device(C, bin, bin, dim, P, m)) {
			e.status = N2M_ATLAS_FAILED;
			continue;
		}
		sform_and_smooth(C, m, postSmooth, srow);
		e.status = N2M_ATLAS_OK;
		e.tri_offset = U.nt;
		e.vert_offset = U.np;
		e.nt = m.nt;
		e.np = m.np;
		append_mesh(C, U, (const double *)m.pts.p, m.np, (const int *)m.tris.p, m.nt);
		ok++;
	}
	C->sync();
	return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

int atlas_run(n2m_ctx *C, const float *labels, const int dim[3], const n2m_params &prm, int postSmooth,
              const float *srow, bool loop)
{
	if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1)
		fail("bad atlas dimensions %d x %d x %d", dim[0], dim[1], dim[2]);
	// bubble filling needs the whole-volume label pass of meshify()
	if (loop || prm.fillBubbles)
		return atlas_loop(C, labels, dim, prm, postSmooth, srow);
	return atlas_batched(C, labels, dim, prm, postSmooth, srow);
}

} // namespace


// ===========================================================================
// Differentiable path: forward and adjoint (kernels in k_adjoint.hip)
// ===========================================================================
namespace {

// Lewiner over the whole volume (lo = 0, hi = dim-1), weld, degenerate
// removal, lapIter Laplacian iterations; the mesh goes to mesh slot 0.
int diff_forward(n2m_ctx *C, const float *vol, const int dim[3], float iso, int originalMC, int lapIter, bool exact)
{
	hipStream_t s = C->stream;
	C->diff.valid = false;
	const int NX = dim[0], NY = dim[1], NZ = dim[2];
	if (NX < 2 || NY < 2 || NZ < 2)
		fail("differentiable path needs at least 2 voxels per axis");
	const int lo[3] = {0, 0, 0}, hi[3] = {NX - 1, NY - 1, NZ - 1};
	const size_t npts = (size_t)NX * NY * NZ;
	const size_t nseg = lw_nseg_total(NX, NY, NZ, lo, hi);
	uint8_t *pf = C->buf<uint8_t>(S_PFLAGS, npts);
	uint16_t *cr = C->buf<uint16_t>(S_CROW, npts);
	uint32_t *cnt = C->buf<uint32_t>(S_SEGCNT, 3 * nseg);
	uint32_t *base = C->buf<uint32_t>(S_SEGBASE, 3 * nseg);
	// the sub-box is the volume: no padding value is ever read
	C->tbegin("mc_lewiner_classify", 7.0 * npts);
	launch_lw_classify(vol, NX, NY, NZ, lo, hi, iso, 0.f, originalMC, pf, cr, cnt, s);
	C->tend();
	uint32_t tot[3];
	scan3(C, cnt, base, nseg, tot);
	const uint32_t nE = tot[0], nT = tot[1], nC = tot[2];
	uint32_t V = nE + nC;
	if (V < 3 || nT < 1)
		return EXIT_FAILURE;
	double *pts0 = C->buf<double>(S_PTS0, (size_t)3 * V);
	int *tris0 = C->buf<int>(S_TRIS0, (size_t)3 * nT);
	unsigned long long *vrec = C->buf<unsigned long long>(S_DF_VREC, V);
	uint32_t *crec = C->buf<uint32_t>(S_DF_CREC, (size_t)12 * (nC ? nC : 1));
	C->tbegin("mc_lewiner_emit", 3.0 * npts + 32.0 * V + 12.0 * nT + 48.0 * nC);
	launch_lw_emit_rec(vol, NX, NY, NZ, lo, hi, iso, 0.f, originalMC, pf, cr, cnt, base, nE, pts0, tris0,
	                   exact ? 1 : 0, vrec, crec, s);
	C->tend();
	double *pts = nullptr;
	int *tris = nullptr;
	uint32_t np = weld(C, pts0, V, tris0, nT, pts, tris, false);
	bool welded = np != V;
	if (welded) { // keep the class map for the backward
		HIPX(hipMemcpyAsync(C->buf<int>(S_DF_CLS, V), C->buf<int>(S_CLS, V), (size_t)4 * V, hipMemcpyDeviceToDevice,
		                    s));
		HIPX(hipMemcpyAsync(C->buf<int>(S_DF_REP, np), C->buf<int>(S_REP, np), (size_t)4 * np,
		                    hipMemcpyDeviceToDevice, s));
	}
	if (np < 3)
		return EXIT_FAILURE;
	uint32_t nt = remove_degenerate(C, pts, np, tris, nT, false);
	set_mesh_count(C, 1);
	Mesh &m = C->meshes[0];
	int *mt = (int *)n2m_ctx::grow(m.tris, (size_t)12 * (nt ? nt : 1));
	double *mp = (double *)n2m_ctx::grow(m.pts, (size_t)24 * np);
	HIPX(hipMemcpyAsync(mt, tris, (size_t)12 * nt, hipMemcpyDeviceToDevice, s));
	HIPX(hipMemcpyAsync(mp, pts, (size_t)24 * np, hipMemcpyDeviceToDevice, s));
	m.nt = (int)nt;
	m.np = (int)np;
	laplace_device(C, mp, mt, np, nt, lapIter);
	C->sync();
	C->diff.valid = true;
	C->diff.vol = vol;
	for (int i = 0; i < 3; i++)
		C->diff.dim[i] = dim[i];
	C->diff.iso = iso;
	C->diff.lapIter = lapIter;
	C->diff.nmc = V;
	C->diff.welded = welded;
	return EXIT_SUCCESS;
}

int diff_backward(n2m_ctx *C, const double *dLdp, double *dLdvol)
{
	if (!C->diff.valid || C->meshes.empty())
		fail("no differentiable forward on this context");
	hipStream_t s = C->stream;
	const Mesh &m = C->meshes[0];
	const uint32_t np = (uint32_t)m.np, nt = (uint32_t)m.nt, nmc = C->diff.nmc;
	const int *D = C->diff.dim;
	const size_t nvox = (size_t)D[0] * D[1] * D[2];
	double *g = C->buf<double>(S_DF_G, (size_t)3 * np), *h = C->buf<double>(S_DF_H, (size_t)3 * np);
	HIPX(hipMemcpyAsync(g, dLdp, (size_t)24 * np, hipMemcpyDeviceToDevice, s));
	if (C->diff.lapIter > 0) {
		Csr c = build_csr(C, (const int *)m.tris.p, np, nt);
		C->tbegin("adj_laplace", 4.0 * np + 8.0 * 3 * nt + 72.0 * np);
		for (int it = 0; it < C->diff.lapIter; it++)
			launch_adj_laplace_T(c.start, c.pairs, g, h, np, s);
		C->tend();
	}
	double *gm = C->buf<double>(S_DF_GM, (size_t)3 * nmc);
	AdjGeo ag;
	ag.nx = D[0];
	ag.nxy = (long long)D[0] * D[1];
	ag.nxyz = (long long)nvox;
	ag.lo0 = ag.lo1 = ag.lo2 = 0;
	ag.sx = D[0];
	ag.sy = D[1];
	ag.iso = C->diff.iso;
	ag.mnv = 0.f;
	HIPX(hipMemsetAsync(dLdvol, 0, 8 * nvox, s));
	C->tbegin("adj_mc", 24.0 * nmc * 2 + 8.0 * nmc + 16.0 * nmc);
	launch_adj_unweld(C->diff.welded ? C->buf<int>(S_DF_CLS, nmc) : nullptr,
	                  C->diff.welded ? C->buf<int>(S_DF_REP, np) : nullptr, g, nmc, gm, s);
	launch_adj_centre(C->buf<unsigned long long>(S_DF_VREC, nmc), C->buf<uint32_t>(S_DF_CREC, 12), nmc, gm, s);
	launch_adj_edge(C->buf<unsigned long long>(S_DF_VREC, nmc), nmc, gm, C->diff.vol, ag, dLdvol, s);
	C->tend();
	C->sync();
	return EXIT_SUCCESS;
}

} // namespace


// ===========================================================================
// Mesh writers for the binary formats (src/meshify.c:602-930): the byte
// streams are encoded on the GPU (FP32 vertices, PLY face records, base64),
// the headers and the optional zlib / gzip layer on the host.  The header
// text must equal the reference's byte for byte, so it is written with the
// same strings.
// ===========================================================================
namespace {

std::vector<uint8_t> dl_bytes(n2m_ctx *C, const void *d, size_t n)
{
	std::vector<uint8_t> h(n);
	if (n)
		HIPX(hipMemcpyAsync(h.data(), d, n, hipMemcpyDeviceToHost, C->stream));
	C->sync();
	return h;
}

// float32 copy of the vertices (vec3d2vec4s, src/meshify.c:414-416)
const float *pts_f32(n2m_ctx *C, const Mesh &m)
{
	float *d = C->buf<float>(S_ENC_A, (size_t)3 * (m.np ? m.np : 1));
	C->tbegin("encode_f32", 24.0 * m.np + 12.0 * m.np);
	launch_f64_to_f32((const double *)m.pts.p, (size_t)3 * m.np, d, C->stream);
	C->tend();
	return d;
}

// base64 of n host or device bytes (src/base64.c: no line breaks)
std::vector<uint8_t> b64_of(n2m_ctx *C, const void *src, size_t n, bool src_on_host)
{
	const uint8_t *d = (const uint8_t *)src;
	if (src_on_host) {
		uint8_t *dd = C->buf<uint8_t>(S_ENC_B, n ? n : 1);
		if (n)
			HIPX(hipMemcpyAsync(dd, src, n, hipMemcpyHostToDevice, C->stream));
		d = dd;
	}
	const size_t olen = 4 * ((n + 2) / 3);
	uint8_t *o = C->buf<uint8_t>(S_ENC_C, olen ? olen : 1);
	C->tbegin("encode_base64", 1.0 * n + 1.0 * olen);
	launch_base64(d, n, o, C->stream);
	C->tend();
	return dl_bytes(C, o, olen);
}

// zlib compress() as the reference's GIfTI writer calls it (src/meshify.c:862-870)
std::vector<uint8_t> zlib_of(const std::vector<uint8_t> &raw)
{
	uLongf dl = compressBound((uLong)raw.size());
	std::vector<uint8_t> out(dl);
	if (compress(out.data(), &dl, raw.data(), (uLong)raw.size()) != Z_OK)
		printf("Compression error\n");
	out.resize(dl);
	return out;
}

bool write_ply(n2m_ctx *C, const Mesh &m, const char *fnm) // src/meshify.c:764-816
{
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	fputs("ply\n", fp);
	fputs("format binary_little_endian 1.0\n", fp);
	fputs("comment niimath\n", fp);
	fprintf(fp, "element vertex %d\n", m.np);
	fputs("property float x\n", fp);
	fputs("property float y\n", fp);
	fputs("property float z\n", fp);
	fprintf(fp, "element face %d\n", m.nt);
	fputs("property list uchar int vertex_indices\n", fp);
	fputs("end_header\n", fp);
	const float *p32 = pts_f32(C, m);
	uint8_t *faces = C->buf<uint8_t>(S_ENC_B, (size_t)13 * (m.nt ? m.nt : 1));
	C->tbegin("encode_ply", 12.0 * m.nt + 13.0 * m.nt);
	launch_ply_faces((const int *)m.tris.p, (size_t)m.nt, faces, C->stream);
	C->tend();
	std::vector<uint8_t> a = dl_bytes(C, p32, (size_t)12 * m.np), b = dl_bytes(C, faces, (size_t)13 * m.nt);
	fwrite(a.data(), 1, a.size(), fp);
	fwrite(b.data(), 1, b.size(), fp);
	fclose(fp);
	return true;
}

void gii_array_head(FILE *fp, const char *type, int n, bool gz, const char *intent)
{
	fputs("   <DataArray  ArrayIndexingOrder=\"RowMajorOrder\"\n", fp);
	fprintf(fp, "               DataType=\"%s\"\n", type);
	fprintf(fp, "               Dim0=\"%d\"\n", n);
	fputs("               Dim1=\"3\"\n", fp);
	fputs("               Dimensionality=\"2\"\n", fp);
	fputs(gz ? "               Encoding=\"GZipBase64Binary\"\n" : "               Encoding=\"Base64Binary\"\n", fp);
	fputs("               Endian=\"LittleEndian\"\n", fp);
	fputs("               ExternalFileName=\"\"\n", fp);
	fputs("               ExternalFileOffset=\"\"\n", fp);
	fprintf(fp, "               Intent=\"%s\">\n", intent);
	fputs("      <MetaData>\n", fp);
	fputs("      </MetaData>\n", fp);
}

bool write_gii(n2m_ctx *C, const Mesh &m, const char *fnm, bool gz) // src/meshify.c:819-930
{
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	fputs("<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n", fp);
	fputs("<!DOCTYPE GIFTI SYSTEM \"http://www.nitrc.org/frs/download.php/115/gifti.dtd\">\n", fp);
	fputs("<GIFTI Version=\"1.0\"  NumberOfDataArrays=\"2\">\n", fp);
	fputs("   <MetaData>\n", fp);
	fputs("      <MD>\n", fp);
	fputs("          <Name><![CDATA[nii2mesh-version]]></Name>\n", fp);
	fputs("          <Value><![CDATA[nii2mesh, 1 Jan 2022]]></Value>\n", fp);
	fputs("      </MD>\n", fp);
	fputs("   </MetaData>\n", fp);
	fputs("   <LabelTable/>\n", fp);
	gii_array_head(fp, "NIFTI_TYPE_INT32", m.nt, gz, "NIFTI_INTENT_TRIANGLE");
	fputs("      <Data>", fp);
	std::vector<uint8_t> e;
	if (gz) {
		std::vector<uint8_t> z = zlib_of(dl_bytes(C, m.tris.p, (size_t)12 * m.nt));
		e = b64_of(C, z.data(), z.size(), true);
	} else {
		e = b64_of(C, m.tris.p, (size_t)12 * m.nt, false);
	}
	fwrite(e.data(), 1, e.size(), fp);
	fputs("</Data>\n", fp);
	fputs("   </DataArray>\n", fp);
	gii_array_head(fp, "NIFTI_TYPE_FLOAT32", m.np, gz, "NIFTI_INTENT_POINTSET");
	fputs("      <CoordinateSystemTransformMatrix>\n", fp);
	fputs("         <DataSpace><![CDATA[NIFTI_XFORM_UNKNOWN]]></DataSpace>\n", fp);
	fputs("         <TransformedSpace><![CDATA[NIFTI_XFORM_UNKNOWN]]></TransformedSpace>\n", fp);
	fputs("         <MatrixData>1.000000 0.000000 0.000000 0.000000 0.000000 1.000000 0.000000 0.000000 0.000000 "
	      "0.000000 1.000000 0.000000 0.000000 0.000000 0.000000 1.000000 </MatrixData>\n",
	      fp);
	fputs("      </CoordinateSystemTransformMatrix>\n", fp);
	fputs("      <Data>", fp);
	const float *p32 = pts_f32(C, m);
	if (gz) {
		std::vector<uint8_t> z = zlib_of(dl_bytes(C, p32, (size_t)12 * m.np));
		e = b64_of(C, z.data(), z.size(), true);
	} else {
		e = b64_of(C, p32, (size_t)12 * m.np, false);
	}
	fwrite(e.data(), 1, e.size(), fp);
	fputs("</Data>\n", fp);
	fputs("   </DataArray>\n", fp);
	fputs("</GIFTI>\n", fp);
	fclose(fp);
	return true;
}

bool write_mz3(n2m_ctx *C, const Mesh &m, const char *fnm, bool gz) // src/meshify.c:602-660
{
	uint8_t h[16];
	const uint16_t sig = 0x5A4D, attr = 3;
	const uint32_t nf = (uint32_t)m.nt, nv = (uint32_t)m.np, nskip = 0;
	memcpy(h, &sig, 2);
	memcpy(h + 2, &attr, 2);
	memcpy(h + 4, &nf, 4);
	memcpy(h + 8, &nv, 4);
	memcpy(h + 12, &nskip, 4);
	std::vector<uint8_t> t = dl_bytes(C, m.tris.p, (size_t)12 * m.nt);
	std::vector<uint8_t> p = dl_bytes(C, pts_f32(C, m), (size_t)12 * m.np);
	if (gz) {
		gzFile f = gzopen(fnm, "w");
		if (!f)
			return false;
		gzwrite(f, h, sizeof h);
		if (!t.empty())
			gzwrite(f, t.data(), (unsigned)t.size());
		if (!p.empty())
			gzwrite(f, p.data(), (unsigned)p.size());
		gzclose(f);
		return true;
	}
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	fwrite(h, 1, sizeof h, fp);
	fwrite(t.data(), 1, t.size(), fp);
	fwrite(p.data(), 1, p.size(), fp);
	fclose(fp);
	return true;
}

// save_mesh() dispatch (src/meshify.c:984-1019) for the binary formats
// FreeSurfer binary surface (src/meshify.c:418-456): magic, a time stamp
// line, then counts, float32 vertices and triangles, all big-endian (the
// reference swaps on every platform: its endianness test is a function
// address, never null)
bool write_freesurfer(n2m_ctx *C, const Mesh &m, const char *fnm)
{
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	const uint8_t magic[3] = {0xFF, 0xFF, 0xFE};
	fwrite(magic, 3, 1, fp);
	time_t t = time(nullptr);
	char stamp[128] = "";
	struct tm tmv;
	localtime_r(&t, &tmv);
	strftime(stamp, sizeof stamp, "created by niimath on %c\n\n", &tmv);
	fwrite(stamp, strlen(stamp), 1, fp);
	const uint32_t counts[2] = {__builtin_bswap32((uint32_t)m.np), __builtin_bswap32((uint32_t)m.nt)};
	fwrite(counts, 4, 2, fp);
	const float *p32 = pts_f32(C, m);
	uint32_t *a = C->buf<uint32_t>(S_ENC_B, (size_t)3 * (m.np ? m.np : 1));
	uint32_t *b = C->buf<uint32_t>(S_ENC_C, (size_t)3 * (m.nt ? m.nt : 1));
	C->tbegin("encode_swap", 24.0 * m.np + 24.0 * m.nt);
	launch_bswap32((const uint32_t *)p32, (size_t)3 * m.np, a, C->stream);
	launch_bswap32((const uint32_t *)m.tris.p, (size_t)3 * m.nt, b, C->stream);
	C->tend();
	std::vector<uint8_t> ha = dl_bytes(C, a, (size_t)12 * m.np), hb = dl_bytes(C, b, (size_t)12 * m.nt);
	fwrite(ha.data(), 1, ha.size(), fp);
	fwrite(hb.data(), 1, hb.size(), fp);
	fclose(fp);
	return true;
}

// binary STL (src/meshify.c:725-760): 80 zero bytes, the facet count, then
// 50-byte facets
bool write_stl(n2m_ctx *C, const Mesh &m, const char *fnm)
{
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	const uint8_t hdr[80] = {0};
	fwrite(hdr, 80, 1, fp);
	const int32_t nf = m.nt;
	fwrite(&nf, sizeof nf, 1, fp);
	uint16_t *f = C->buf<uint16_t>(S_ENC_B, (size_t)25 * (m.nt ? m.nt : 1));
	C->tbegin("encode_stl", 12.0 * m.nt + 72.0 * m.nt + 50.0 * m.nt);
	launch_stl_facets((const double *)m.pts.p, (const int *)m.tris.p, (size_t)m.nt, f, C->stream);
	C->tend();
	std::vector<uint8_t> h = dl_bytes(C, f, (size_t)50 * m.nt);
	fwrite(h.data(), 1, h.size(), fp);
	fclose(fp);
	return true;
}

// legacy VTK polydata, binary (src/meshify.c:932-968): big-endian float32
// points and (3, v0, v1, v2) polygons
bool write_vtk(n2m_ctx *C, const Mesh &m, const char *fnm)
{
	FILE *fp = fopen(fnm, "wb");
	if (!fp)
		return false;
	fputs("# vtk DataFile Version 3.0\n", fp);
	fputs("this file was written using niimath\n", fp);
	fputs("BINARY\n", fp);
	fputs("DATASET POLYDATA\n", fp);
	fprintf(fp, "POINTS %d float\n", m.np);
	const float *p32 = pts_f32(C, m);
	uint32_t *a = C->buf<uint32_t>(S_ENC_B, (size_t)3 * (m.np ? m.np : 1));
	uint32_t *b = C->buf<uint32_t>(S_ENC_C, (size_t)4 * (m.nt ? m.nt : 1));
	C->tbegin("encode_swap", 24.0 * m.np + 28.0 * m.nt);
	launch_bswap32((const uint32_t *)p32, (size_t)3 * m.np, a, C->stream);
	launch_vtk_polys((const int *)m.tris.p, (size_t)m.nt, b, C->stream);
	C->tend();
	std::vector<uint8_t> ha = dl_bytes(C, a, (size_t)12 * m.np), hb = dl_bytes(C, b, (size_t)16 * m.nt);
	fwrite(ha.data(), 1, ha.size(), fp);
	fprintf(fp, "POLYGONS %d %d\n", m.nt, m.nt * 4);
	fwrite(hb.data(), 1, hb.size(), fp);
	fclose(fp);
	return true;
}

// rows [0, n) formatted by fmt(i, buf) (returning the length), in parallel
// chunks on the host threads, written in order
template <typename F> void write_rows(FILE *fp, size_t n, F fmt)
{
	unsigned nth = std::thread::hardware_concurrency();
	nth = n < 65536 ? 1u : (nth < 1 ? 1u : (nth > 16 ? 16u : nth));
	std::vector<std::string> out(nth);
	auto run = [&](unsigned k) {
		const size_t b = n * k / nth, e = n * (k + 1) / nth;
		std::string &o = out[k];
		o.reserve((e - b) * 36);
		char buf[160];
		for (size_t i = b; i < e; i++)
			o.append(buf, (size_t)fmt(i, buf));
	};
	std::vector<std::thread> th;
	for (unsigned k = 1; k < nth; k++)
		th.emplace_back(run, k);
	run(0);
	for (auto &t : th)
		t.join();
	for (auto &o : out)
		fwrite(o.data(), 1, o.size(), fp);
}

// the text formats (src/meshify.c:683-723): numbers are formatted by the
// host C library with the reference's format strings (%g of each float64
// coordinate, %d of each 1-based index); the GPU only supplies the arrays
enum TextFmt { TXT_JSON, TXT_OFF, TXT_OBJ };
bool write_text(n2m_ctx *C, const Mesh &m, const char *fnm, TextFmt kind)
{
	std::vector<uint8_t> hp = dl_bytes(C, m.pts.p, (size_t)24 * m.np), ht = dl_bytes(C, m.tris.p, (size_t)12 * m.nt);
	const double *p = (const double *)hp.data();
	const int *t = (const int *)ht.data();
	FILE *fp = fopen(fnm, "w");
	if (!fp)
		return false;
	if (kind == TXT_JSON) {
		fprintf(fp, "{\n");
		fprintf(fp, "\t\"_DataInfo_\":{\n\t\t\"JMeshVersion\":\"0.5\",\n\t\t\"Comment\":\"Created by nii2mesh\"\n\t},\n");
		fprintf(fp, "\t\"MeshVertex3\":[\n");
		write_rows(fp, (size_t)m.np, [&](size_t i, char *b) {
			return snprintf(b, 160, "[%g,\t%g,\t%g],\n", p[3 * i], p[3 * i + 1], p[3 * i + 2]);
		});
		fprintf(fp, "\t],\n\t\"MeshTri3\":[\n");
		write_rows(fp, (size_t)m.nt, [&](size_t i, char *b) {
			return snprintf(b, 160, "[%d,\t%d,\t%d],\n", t[3 * i] + 1, t[3 * i + 1] + 1, t[3 * i + 2] + 1);
		});
		fprintf(fp, "\t]\n}\n");
	} else {
		if (kind == TXT_OFF)
			fprintf(fp, "OFF\n%d\t%d\t0\n", m.np, m.nt);
		const char *vf = kind == TXT_OFF ? "%g %g %g\n" : "v %g %g %g\n";
		const char *ff = kind == TXT_OFF ? "%d %d %d\n" : "f %d %d %d\n";
		write_rows(fp, (size_t)m.np,
		           [&](size_t i, char *b) { return snprintf(b, 160, vf, p[3 * i], p[3 * i + 1], p[3 * i + 2]); });
		write_rows(fp, (size_t)m.nt, [&](size_t i, char *b) {
			return snprintf(b, 160, ff, t[3 * i] + 1, t[3 * i + 1] + 1, t[3 * i + 2] + 1);
		});
	}
	fclose(fp);
	return true;
}

int save_mesh_device(n2m_ctx *C, const Mesh &m, const char *fnm, bool gz)
{
	// extension = what strip_ext() (src/meshify.c:973-982) removes: the text
	// from the last '.' after the last path separator, unless that '.' starts
	// the name
	std::string f(fnm), ext;
	size_t e = f.size();
	while (e > 0 && f[e] != '.' && f[e] != '\\' && f[e] != '/')
		--e;
	if (e > 0 && e < f.size() && f[e] == '.' && f[e - 1] != '\\' && f[e - 1] != '/')
		ext = f.substr(e);
	// the reference's order of tests (src/meshify.c:989-1018)
	auto has = [&](const char *x) { return ext.find(x) != std::string::npos; };
	bool ok;
	if (has(".mz3"))
		ok = write_mz3(C, m, fnm, gz);
	else if (has(".gii"))
		ok = write_gii(C, m, fnm, gz);
	else if (has(".inflated") || has(".pial"))
		ok = write_freesurfer(C, m, fnm);
	else if (has(".json"))
		ok = write_text(C, m, fnm, TXT_JSON);
	else if (has(".off"))
		ok = write_text(C, m, fnm, TXT_OFF);
	else if (has(".obj"))
		ok = write_text(C, m, fnm, TXT_OBJ);
	else if (has(".ply"))
		ok = write_ply(C, m, fnm);
	else if (has(".stl"))
		ok = write_stl(C, m, fnm);
	else if (has(".vtk"))
		ok = write_vtk(C, m, fnm);
	else
		ok = write_mz3(C, m, (f + ".mz3").c_str(), gz);
	return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
int n2m::quadric_prep(const vec3d *p, const vec3i *t, int nv, int nt, QuadricPrep &g)
{
	return guarded([&] {
		if (nv < 0 || nt < 0)
			fail("quadric_simplify_mesh: %d vertices, %d triangles", nv, nt);
		check_tris((const int *)t, nt, nv);
		n2m_ctx *C = tl_ctx.get();
		hipStream_t s = C->stream;
		const uint32_t NV = (uint32_t)nv, NT = (uint32_t)nt;
		double *dp = C->buf<double>(S_PTS1, (size_t)3 * NV);
		int *dt = C->buf<int>(S_TRIS1, (size_t)3 * NT);
		if (NV)
			HIPX(hipMemcpyAsync(dp, p, (size_t)24 * NV, hipMemcpyHostToDevice, s));
		if (NT)
			HIPX(hipMemcpyAsync(dt, t, (size_t)12 * NT, hipMemcpyHostToDevice, s));
		const Csr c = build_csr(C, dt, NV, NT, true);
		uint8_t *border = border_flags(C, c, dt, NV, NT);
		double *nd = C->buf<double>(S_Q_ND, (size_t)4 * NT);
		double *q = C->buf<double>(S_Q_Q, (size_t)10 * NV);
		double *err = C->buf<double>(S_Q_ERR, (size_t)4 * NT);
		C->tbegin("quadric_first_update", 36.0 * NT + 8.0 * NV);
		launch_q_planes(dp, dt, NT, nd, s);
		launch_q_vertex(c.start, c.ref, NV, nd, q, s);
		launch_q_errors(dp, dt, NT, q, border, err, s);
		C->tend();
		g.start = download(C, c.start, (size_t)NV + 1);
		g.ref = download(C, c.ref, (size_t)3 * NT);
		g.border = download(C, border, NV);
		g.q = download(C, q, (size_t)10 * NV);
		g.nd = download(C, nd, (size_t)4 * NT);
		g.err = download(C, err, (size_t)4 * NT);
		return EXIT_SUCCESS;
	});
}

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
		set_mesh_count(C, 1);
		Mesh &m = mesh_slot(C, 0);
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
		set_mesh_count(C, niso);
		float *sm = C->buf<float>(S_VOL2, nvox); // smoothed volume, shared by all isolevels
		bool smoothed = prm->preSmooth && NX >= 5 && NY >= 5 && NZ >= 5;
		uint32_t range[2];
		// the threshold masks of the first isolevels in one pass over the
		// smoothed volume (valid while the range check keeps the isolevel)
		const size_t nwords = (size_t)((NX + 31) / 32) * NY * NZ;
		const int nmask = smoothed && (prm->onlyLargest || prm->fillBubbles)
		                      ? (niso < N2M_SMOOTH_MASKS ? niso : N2M_SMOOTH_MASKS)
		                      : 0;
		uint32_t *masks = nmask ? C->buf<uint32_t>(S_PIPE_BITS, nmask * nwords) : nullptr;
		if (smoothed) {
			uint32_t init[2] = {0xffffffffu, 0u};
			HIPX(hipMemcpyAsync(C->d_small + SM_RANGE, init, sizeof init, hipMemcpyHostToDevice, s));
			C->tbegin("smooth", 8.0 * nvox);
			launch_smooth(d_img, sm, NX, NY, NZ, C->d_small + SM_RANGE, nullptr, 0, nullptr, s);
			C->tend();
			if (nmask) {
				C->tbegin("mask_bits", 4.0 * nvox + nmask * nvox / 8.0);
				launch_mask_bits_multi(sm, NX, NY, NZ, isolevels, nmask, masks, s);
				C->tend();
			}
			uint32_t *r = C->fetch(SM_RANGE, 2);
			range[0] = r[0];
			range[1] = r[1];
		}
		// The isolevels are meshed by one host thread and context (stream and
		// scratch) each: the meshing stages are mostly bound by latency, so
		// they overlap well.  This thread meshes the last isolevel on this
		// context and worker threads the others on C->workers (kept between
		// calls): with one stream fewer, every meshing stream gets a hardware
		// queue of its own (HIP maps streams onto 4 queues per process by
		// default, and two streams on one queue run one after the other).
		// N2M_PIPE_THREADS=0 meshes the isolevels one after another.
		//
		// HC with threads: a mesh whose meshing ends while another isolevel is
		// still being meshed is smoothed at once on its own stream (timing
		// bracket "hc_ring_side": it shares the GPU).  The mesh meshed last
		// (in practice the largest: its weld takes longest) prepares its
		// smoothing (relabelling, CSR, rings, border flags) while those run,
		// then waits for them, so that its iterations ("hc_ring") have the
		// GPU to themselves.  N2M_PIPE_HC=1: every mesh is smoothed right
		// after its meshing, the HC passes one at a time (a mutex);
		// N2M_PIPE_HC=2: the same, concurrently.
		const char *pt = getenv("N2M_PIPE_THREADS");
		const bool threaded = niso > 1 && !(pt && *pt == '0');
		const char *ph = getenv("N2M_PIPE_HC");
		const int hc_mode = ph && *ph == '1' ? 1 : (ph && *ph == '2' ? 2 : 0);
		std::mutex hc_mu, st_mu;
		std::condition_variable st_cv;
		int finished = 0, side_pending = 0;
		std::vector<char> done(niso, 0);
		auto mesh_one = [&](n2m_ctx *D, int k) {
			n2m_params P = *prm;
			P.isolevel = isolevels[k];
			float *v = D->buf<float>(S_VOL, nvox); // written whole by the apply stage
			Mesh &m = C->meshes[k];
			m.nt = m.np = 0;
			VolOut vo;
			if (volume_stages(D, v, dim, P, smoothed, k < nmask ? masks + k * nwords : nullptr, vo,
			                  smoothed ? sm : d_img, smoothed ? range : nullptr) &&
			    surface_stages(D, v, dim, vo, P, m))
				done[k] = 1;
			D->sync();
		};
		auto smooth_k = [&](n2m_ctx *D, int k, const char *stat, const std::function<void()> *before) {
			Mesh &m = C->meshes[k];
			smooth_hc_device(D, (double *)m.pts.p, (const int *)m.tris.p, m.np, m.nt, alpha, beta, hcIter, true,
			                 before, stat);
			D->sync();
		};
		// one isolevel on a thread of its own: mesh it, then smooth it
		auto run_iso = [&](n2m_ctx *D, int k) {
			try {
				mesh_one(D, k);
			} catch (...) {
				std::lock_guard<std::mutex> g(st_mu);
				finished++;
				st_cv.notify_all();
				throw;
			}
			bool last, side = false;
			{
				std::lock_guard<std::mutex> g(st_mu);
				last = ++finished == niso;
				if (!last && hc_mode == 0 && done[k] && hcIter > 0) {
					side = true;
					side_pending++;
				}
			}
			st_cv.notify_all();
			if (!done[k] || hcIter <= 0)
				return;
			if (hc_mode != 0) {
				std::unique_lock<std::mutex> lock(hc_mu, std::defer_lock);
				if (hc_mode == 1)
					lock.lock();
				smooth_k(D, k, "hc_ring", nullptr);
			} else if (side) {
				try {
					smooth_k(D, k, "hc_ring_side", nullptr);
				} catch (...) {
					std::lock_guard<std::mutex> g(st_mu);
					side_pending--;
					st_cv.notify_all();
					throw;
				}
				std::lock_guard<std::mutex> g(st_mu);
				side_pending--;
				st_cv.notify_all();
			} else { // the last one: its iterations wait for the side passes
				const std::function<void()> wait = [&] {
					std::unique_lock<std::mutex> l(st_mu);
					st_cv.wait(l, [&] { return side_pending == 0; });
				};
				smooth_k(D, k, "hc_ring", &wait);
			}
		};
		if (threaded) {
			C->sync(); // the shared smoothed volume is complete
			const int kmain = niso - 1;
			while ((int)C->workers.size() < niso - 1)
				C->workers.push_back(ctx_new());
			std::vector<std::string> err(niso);
			std::vector<std::thread> th;
			for (int k = 0; k < kmain; k++) {
				n2m_ctx *D = C->workers[k];
				D->timing = C->timing;
				D->timing_only = C->timing_only;
				th.emplace_back([&, D, k] {
					try {
						HIPX(hipSetDevice(C->device));
						run_iso(D, k);
					} catch (const std::exception &e) {
						err[k] = e.what();
					}
				});
			}
			try {
				run_iso(C, kmain);
			} catch (const std::exception &e) {
				err[kmain] = e.what();
			}
			for (auto &t : th)
				t.join();
			for (int k = 0; k < kmain; k++)
				C->absorb(*C->workers[k]);
			for (int k = 0; k < niso; k++)
				if (!err[k].empty())
					throw std::runtime_error(err[k]);
		} else {
			for (int k = 0; k < niso; k++)
				mesh_one(C, k);
			int big = -1;
			for (int k = 0; k < niso; k++)
				if (done[k] && (big < 0 || C->meshes[k].nt > C->meshes[big].nt))
					big = k;
			if (hcIter > 0)
				for (int k = 0; k < niso; k++)
					if (done[k])
						smooth_k(C, k, k == big ? "hc_ring" : "hc_ring_side", nullptr);
		}
		int ok = 0;
		for (int k = 0; k < niso; k++)
			ok += done[k];
		C->sync();
		return ok == niso ? EXIT_SUCCESS : EXIT_FAILURE;
	});
}

float n2m_ctx_set_threshold(n2m_ctx *C, const float *d_img, int64_t nvox, int mode)
{
	float r = 1.0f;
	guarded([&] { return set_thresholds_device(C, d_img, nvox, 1, &mode, &r); });
	return r;
}

int n2m_ctx_set_thresholds(n2m_ctx *C, const float *d_img, int64_t nvox, int n, const int *modes, float *out)
{
	return guarded([&] { return set_thresholds_device(C, d_img, nvox, n, modes, out); });
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
			xfer_big(C, t, m.tris.p, (size_t)12 * m.nt, true);
		if (m.np)
			xfer_big(C, p, m.pts.p, (size_t)24 * m.np, true);
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

int n2m_ctx_atlas(n2m_ctx *C, const float *d_labels, const int dim[3], const n2m_params *prm, int postSmooth,
                  const float *srow)
{
	return guarded([&] { return atlas_run(C, d_labels, dim, *prm, postSmooth, srow, false); });
}

int n2m_ctx_atlas_loop(n2m_ctx *C, const float *d_labels, const int dim[3], const n2m_params *prm, int postSmooth,
                       const float *srow)
{
	return guarded([&] { return atlas_run(C, d_labels, dim, *prm, postSmooth, srow, true); });
}

int n2m_ctx_atlas_count(n2m_ctx *C) { return (int)C->atlas.size(); }

int n2m_ctx_atlas_entry(n2m_ctx *C, int idx, n2m_atlas_entry *e)
{
	if (idx < 0 || idx >= (int)C->atlas.size())
		return EXIT_FAILURE;
	*e = C->atlas[idx];
	return EXIT_SUCCESS;
}

int n2m_ctx_atlas_download(n2m_ctx *C, int idx, vec3i *t, vec3d *p)
{
	return guarded([&] {
		if (idx < 0 || idx >= (int)C->atlas.size())
			fail("no atlas entry %d", idx);
		const n2m_atlas_entry &e = C->atlas[idx];
		if (e.status != N2M_ATLAS_OK)
			fail("atlas label %d has no mesh", e.label);
		const Mesh &U = C->atlas_mesh;
		if (e.nt)
			HIPX(hipMemcpyAsync(t, (const int *)U.tris.p + (size_t)3 * e.tri_offset, (size_t)12 * e.nt,
			                    hipMemcpyDeviceToHost, C->stream));
		if (e.np)
			HIPX(hipMemcpyAsync(p, (const double *)U.pts.p + (size_t)3 * e.vert_offset, (size_t)24 * e.np,
			                    hipMemcpyDeviceToHost, C->stream));
		C->sync();
		for (int i = 0; i < e.nt; i++) {
			t[i].x -= e.vert_offset;
			t[i].y -= e.vert_offset;
			t[i].z -= e.vert_offset;
		}
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_atlas_device_mesh(n2m_ctx *C, const vec3i **d_t, const vec3d **d_p, int *nt, int *np)
{
	*d_t = (const vec3i *)C->atlas_mesh.tris.p;
	*d_p = (const vec3d *)C->atlas_mesh.pts.p;
	*nt = C->atlas_mesh.nt;
	*np = C->atlas_mesh.np;
	return EXIT_SUCCESS;
}

int n2m_atlas(const float *img, short dim[3], int originalMC, bool preSmooth, bool fillBubbles, int postSmooth,
              const float *srow, bool verbose, n2m_atlas_mesh **meshes, int *nmesh)
{
	*meshes = nullptr;
	*nmesh = 0;
	return guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		const int d[3] = {dim[0], dim[1], dim[2]};
		const size_t nvox = (size_t)d[0] * d[1] * d[2];
		float *dl = C->buf<float>(S_AT_LABELS, nvox);
		xfer_big(C, dl, img, 4 * nvox, false);
		n2m_params P{cubes_flavour(), originalMC, 0.5f, preSmooth, 0, fillBubbles, verbose};
		int rc = atlas_run(C, dl, d, P, postSmooth, srow, false);
		const int n = (int)C->atlas.size();
		n2m_atlas_mesh *out = (n2m_atlas_mesh *)calloc(n ? n : 1, sizeof(n2m_atlas_mesh));
		if (!out)
			fail("out of host memory");
		for (int k = 0; k < n; k++) {
			const n2m_atlas_entry &e = C->atlas[k];
			out[k].label = e.label;
			out[k].status = e.status;
			if (e.status != N2M_ATLAS_OK)
				continue;
			out[k].nt = e.nt;
			out[k].np = e.np;
			out[k].t = (vec3i *)malloc((size_t)12 * (e.nt ? e.nt : 1));
			out[k].p = (vec3d *)malloc((size_t)24 * (e.np ? e.np : 1));
			if (!out[k].t || !out[k].p) {
				n2m_atlas_free(out, n);
				fail("out of host memory");
			}
			if (n2m_ctx_atlas_download(C, k, out[k].t, out[k].p) != EXIT_SUCCESS) {
				n2m_atlas_free(out, n);
				fail("%s", tl_err.c_str());
			}
		}
		*meshes = out;
		*nmesh = n;
		return rc;
	});
}

void n2m_atlas_free(n2m_atlas_mesh *meshes, int nmesh)
{
	if (!meshes)
		return;
	for (int k = 0; k < nmesh; k++) {
		free(meshes[k].t);
		free(meshes[k].p);
	}
	free(meshes);
}

int n2m_ctx_decode(n2m_ctx *C, const void *d_raw, int datatype, int64_t nvox, float scl_slope, float scl_inter,
                   float *d_out)
{
	return guarded([&] {
		if (nvox < 0)
			fail("negative voxel count");
		if (scl_slope == 0.0f) // src/nii2mesh.c:119
			scl_slope = 1.0f;
		C->tbegin("decode", 4.0 * nvox + 4.0 * nvox);
		if (launch_decode(d_raw, datatype, (size_t)nvox, scl_slope, scl_inter, d_out, C->stream) != 0)
			fail("unsupported NIfTI datatype %d (uint8 2, int16 4, uint16 512, float32 16)", datatype);
		C->tend();
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_decode(const void *raw, int datatype, int64_t nvox, float scl_slope, float scl_inter, float *out)
{
	return guarded([&] {
		n2m_ctx *C = tl_ctx.get();
		const size_t bpp = datatype == 2 ? 1 : (datatype == 16 ? 4 : 2);
		void *draw = C->buf<char>(S_AT_LABELS, (size_t)nvox * bpp);
		float *dout = C->buf<float>(S_VOL, (size_t)nvox);
		HIPX(hipMemcpyAsync(draw, raw, (size_t)nvox * bpp, hipMemcpyHostToDevice, C->stream));
		int rc = n2m_ctx_decode(C, draw, datatype, nvox, scl_slope, scl_inter, dout);
		if (rc != EXIT_SUCCESS)
			fail("%s", tl_err.c_str());
		HIPX(hipMemcpyAsync(out, dout, (size_t)nvox * 4, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return EXIT_SUCCESS;
	});
}

int n2m_ctx_save_mesh(n2m_ctx *C, int mesh, const char *fnm, int isGz)
{
	return guarded([&] {
		if (mesh < 0 || mesh >= (int)C->meshes.size())
			fail("no mesh %d", mesh);
		return save_mesh_device(C, C->meshes[mesh], fnm, isGz != 0);
	});
}

int n2m_save_mesh(const char *fnm, vec3i *tris, vec3d *pts, int ntri, int npt, bool isGz)
{
	return guarded([&] {
		if (ntri < 0 || npt < 0)
			fail("negative mesh size");
		n2m_ctx *C = tl_ctx.get();
		Mesh m;
		m.nt = ntri;
		m.np = npt;
		m.tris.p = C->buf<int>(S_TRIS1, (size_t)3 * (ntri ? ntri : 1));
		m.pts.p = C->buf<double>(S_PTS1, (size_t)3 * (npt ? npt : 1));
		if (ntri)
			HIPX(hipMemcpyAsync(m.tris.p, tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		if (npt)
			HIPX(hipMemcpyAsync(m.pts.p, pts, (size_t)24 * npt, hipMemcpyHostToDevice, C->stream));
		return save_mesh_device(C, m, fnm, isGz);
	});
}

int n2m_otsu(const int *H, int nBin, int mode, int *dark, int *mid, int *bright)
{
	int r = 0;
	guarded([&] {
		r = otsu(tl_ctx.get(), H, nBin, mode, dark, mid, bright);
		return 0;
	});
	return r;
}

int n2m_ctx_diff_forward(n2m_ctx *C, const float *d_vol, const int dim[3], float isolevel, int originalMC,
                         int lapIter, int exact)
{
	return guarded([&] { return diff_forward(C, d_vol, dim, isolevel, originalMC, lapIter, exact != 0); });
}

int n2m_ctx_diff_backward(n2m_ctx *C, const double *d_dLdp, double *d_dLdvol)
{
	return guarded([&] { return diff_backward(C, d_dLdp, d_dLdvol); });
}

void n2m_laplacian_smooth_T(vec3d *g, vec3i *tris, int nvert, int ntri, int iter)
{
	guarded([&] {
		if (nvert <= 0 || iter <= 0)
			return 0;
		check_tris((const int *)tris, ntri, nvert);
		n2m_ctx *C = tl_ctx.get();
		double *dg = C->buf<double>(S_DF_G, (size_t)3 * nvert), *dh = C->buf<double>(S_DF_H, (size_t)3 * nvert);
		int *dt = C->buf<int>(S_TRIS1, (size_t)3 * (ntri ? ntri : 1));
		HIPX(hipMemcpyAsync(dg, g, (size_t)24 * nvert, hipMemcpyHostToDevice, C->stream));
		HIPX(hipMemcpyAsync(dt, tris, (size_t)12 * ntri, hipMemcpyHostToDevice, C->stream));
		Csr c = build_csr(C, dt, nvert, ntri);
		for (int it = 0; it < iter; it++)
			launch_adj_laplace_T(c.start, c.pairs, dg, dh, nvert, C->stream);
		C->check();
		HIPX(hipMemcpyAsync(g, dg, (size_t)24 * nvert, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		return 0;
	});
}

void n2m_ctx_enable_timing(n2m_ctx *C, int on) { C->timing = on != 0; }

void n2m_ctx_timing_only(n2m_ctx *C, const char *name) { C->timing_only = name ? name : ""; }

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

int n2m_atlas_tuned(int device, int stage, const char **name, int *tile, int *blocks_per_cu)
{
	if (stage >= 0 && stage < AST_N) {
		const int k = device >= 0 && device < 64 ? g_atlas_cfg[device][stage].load(std::memory_order_relaxed) : 0;
		if (name)
			*name = kAtlasStageName[stage];
		if (tile)
			*tile = k ? (stage == AST_SMOOTH ? (k - 1) / kNBpc : -1) : -1;
		if (blocks_per_cu)
			*blocks_per_cu = k ? kAtlasBpc[(k - 1) % kNBpc] : -1;
	}
	return AST_N;
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
		xfer_big(C, src, img, 4 * nvox, false);
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
		xfer_big(C, img, src, 4 * nvox, true);
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
		xfer_big(C, dp, verts, (size_t)24 * nvert, false);
		if (ntri > 0)
			xfer_big(C, dt, tris, (size_t)12 * ntri, false);
		smooth_hc_device(C, dp, dt, nvert, ntri, alpha, beta, iter, lockEdges);
		xfer_big(C, verts, dp, (size_t)24 * nvert, true);
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
		launch_smooth(a, b, nx, ny, nz, C->d_small + SM_RANGE, nullptr, 0, nullptr, C->stream);
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
		if (det < 0.0 && nt > 0) {
			int *dt = C->buf<int>(S_TRIS1, 3 * (size_t)nt);
			HIPX(hipMemcpyAsync(dt, t, (size_t)12 * nt, hipMemcpyHostToDevice, C->stream));
			launch_swap_winding(dt, (uint32_t)nt, C->stream);
			C->check();
			HIPX(hipMemcpyAsync(t, dt, (size_t)12 * nt, hipMemcpyDeviceToHost, C->stream));
			C->sync();
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
		uint32_t *rbase = C->buf<uint32_t>(S_RBASE, nwords + 1);
		if (fillBubbles) { // src/bwlabel.c:478-541
			HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
			launch_ccl(bits, NX, NY, NZ, 6, 1, L, AUX, 1, C->d_small + SM_NROOTS, nullptr, rbase,
			           scan_scratch(C, nwords + 1), s, nullptr);
			launch_fill_bubbles(bits, NX, NY, NZ, L, AUX, rbase, C->d_small + SM_NROOTS, s);
		}
		HIPX(hipMemsetAsync(C->d_small + SM_NROOTS, 0, 4, s));
		launch_ccl(bits, NX, NY, NZ, conn, 0, L, AUX, 0, C->d_small + SM_NROOTS, roots, rbase,
		           scan_scratch(C, nwords + 1), s, nullptr);
		uint32_t nl = C->fetch(SM_NROOTS, 1)[0];
		if (nl > 0 && onlyLargest) { // src/bwlabel.c:453-471
			HIPX(hipMemsetAsync(C->d_small + SM_BEST, 0, 8, s));
			launch_ccl_argmax(roots, C->d_small + SM_NROOTS, AUX, (unsigned long long *)(C->d_small + SM_BEST), s);
			uint32_t *sel = C->buf<uint32_t>(S_BITS2, nwords);
			launch_select_root(bits, NX, NY, NZ, L, (unsigned long long *)(C->d_small + SM_BEST), sel, rbase, s);
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
			launch_write_labels(bits, NX, NY, NZ, L, AUX, vol, rbase, s);
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
		return set_thresholds_device(C, vol, nvox, 1, &darkMediumBright123, &r);
	});
	return r;
}

} // extern "C"

namespace {

// nii_otsu(), src/isolevel.c:141-243.  The Liao table is built on the host
// in the reference's order; the O(nBin^3) three-threshold search of modes
// 1 and 5 runs on the GPU (k_otsu3), the other modes on the host.
int otsu(n2m_ctx *C, const int *H, int nBin, int mode, int *dark, int *mid, int *bright)
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
		if (nBin != 256)
			fail("three-threshold Otsu search needs 256 bins");
		int lo = (int)(0.25 * nBin), mi = (int)(0.50 * nBin), hi = (int)(0.75 * nBin);
		double best = P[lo] + P[(size_t)(lo + 1) * nBin + mi] + P[(size_t)(mi + 1) * nBin + hi] +
		              P[(size_t)(hi + 1) * nBin + 255];
		const int nl = nBin - 3;
		char *buf = C->buf<char>(S_OTSU, P.size() * 8 + (size_t)nl * 16 + 64);
		double *dP = (double *)buf, *bv = dP + P.size();
		int *bm = (int *)(bv + nl), *bh = bm + nl;
		HIPX(hipMemcpyAsync(dP, P.data(), P.size() * 8, hipMemcpyHostToDevice, C->stream));
		C->tbegin("otsu_search", 8.0 * P.size());
		launch_otsu3(dP, nBin, bv, bm, bh, C->stream);
		C->tend();
		std::vector<double> hv(nl);
		std::vector<int> hm(nl), hh(nl);
		HIPX(hipMemcpyAsync(hv.data(), bv, (size_t)nl * 8, hipMemcpyDeviceToHost, C->stream));
		HIPX(hipMemcpyAsync(hm.data(), bm, (size_t)nl * 4, hipMemcpyDeviceToHost, C->stream));
		HIPX(hipMemcpyAsync(hh.data(), bh, (size_t)nl * 4, hipMemcpyDeviceToHost, C->stream));
		C->sync();
		for (int l = 0; l < nl; l++) // first strict maximum in l order
			if (hm[l] >= 0 && hv[l] > best) {
				lo = l;
				mi = hm[l];
				hi = hh[l];
				best = hv[l];
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

// setThreshold() for a device volume (src/isolevel.c:35-277): the device
// chain, then one copy of its record to pinned host memory.  nvox < 1
// returns 1 for every mode, as the reference does.
int set_thresholds_device(n2m_ctx *C, const float *d_img, int64_t nvox, int n, const int *modes, float *out)
{
	if (nvox < 1) {
		for (int i = 0; i < n; i++)
			out[i] = 1.0f;
		return 0;
	}
	bool need5 = false;
	for (int i = 0; i < n; i++)
		need5 = need5 || modes[i] == 1 || modes[i] == 3;
	constexpr size_t REC = (sizeof(ThrRecord) + 255) / 256 * 256, H1001 = 4096;
	char *b = C->buf<char>(S_OTSU, REC + H1001 + (size_t)256 * 256 * 8);
	ThrRecord *rec = (ThrRecord *)b;
	C->tbegin("set_threshold", 12.0 * nvox);
	launch_threshold_chain(d_img, (size_t)nvox, rec, (uint32_t *)(b + REC), (double *)(b + REC + H1001), need5,
	                       C->stream);
	C->tend();
	if (!C->h_thr)
		HIPX(hipHostMalloc((void **)&C->h_thr, sizeof(ThrRecord), hipHostMallocDefault));
	HIPX(hipMemcpyAsync(C->h_thr, rec, sizeof(ThrRecord), hipMemcpyDeviceToHost, C->stream));
	C->sync();
	thresholds_from_record(*C->h_thr, n, modes, out);
	return 0;
}

} // namespace
