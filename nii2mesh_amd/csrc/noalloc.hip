// noalloc.hip -- n2m_set_thresholds_noalloc: the isolevels of setThreshold()
// (src/isolevel.c:245-277) for a device volume, with no allocation at all.
//
// n2m_ctx_set_thresholds keeps its histograms, range words and Otsu table in
// the context's device buffers, which the library allocates.  This entry
// takes the caller's stream and uses one fixed device array instead
// (`scratch` below, 1 MiB per device, held until the process ends); host
// side it uses only fixed-size stack arrays.  Every call on a device shares
// that device's array: calls that overlap on one device race.  Whatever the
// previous call left there is cleared by a kernel before use.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <mutex>

#include "../../include/nii2mesh_amd.h"
#include "n2m_internal.h"

using namespace n2m;

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)1 << 20;
// byte offsets inside scratch
constexpr size_t OFF_STATS = 0;                         // 4 words: ordered min, ordered max, NaN count
constexpr size_t OFF_H1001 = 256;                       // 1024 words (1001 bins used)
constexpr size_t OFF_H256 = OFF_H1001 + 4096;           // 256 words
constexpr size_t OFF_BV = OFF_H256 + 1024;              // 256 doubles: k_otsu3 best value per l
constexpr size_t OFF_BM = OFF_BV + 2048;                // 256 ints: its m
constexpr size_t OFF_BH = OFF_BM + 1024;                // 256 ints: its h
constexpr size_t OFF_PROBE = OFF_BH + 1024;             // 516 doubles: k_otsu_probe
constexpr size_t OFF_P = 16384;                         // 256 x 256 doubles: the Otsu table
static_assert(OFF_PROBE + 516 * 8 <= OFF_P, "scratch layout");
static_assert(OFF_P + 256 * 256 * 8 <= SCRATCH_BYTES, "scratch size");

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
std::atomic<unsigned char *> g_ptr[MAX_DEVICES];

} // namespace

// The device array.  In host code `scratch` is only a handle for
// hipGetSymbolAddress: kernels and copies use the per-device address.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

struct Fail {
	char msg[256];
};

[[noreturn]] void raise(const char *what, hipError_t e)
{
	Fail f;
	snprintf(f.msg, sizeof f.msg, "%s failed: %s", what, hipGetErrorString(e));
	throw f;
}

#define NAX(call)                                                                                        \
	do {                                                                                                 \
		hipError_t e_ = (call);                                                                          \
		if (e_ != hipSuccess)                                                                            \
			raise(#call, e_);                                                                            \
	} while (0)

// address of this device's copy of scratch, looked up once per device with
// that device current
unsigned char *device_scratch()
{
	int dev = 0;
	NAX(hipGetDevice(&dev));
	if (dev < 0 || dev >= MAX_DEVICES) {
		Fail f;
		snprintf(f.msg, sizeof f.msg, "device %d is beyond the %d supported", dev, MAX_DEVICES);
		throw f;
	}
	unsigned char *p = g_ptr[dev].load(std::memory_order_acquire);
	if (p)
		return p;
	std::lock_guard<std::mutex> lock(g_mu);
	p = g_ptr[dev].load(std::memory_order_relaxed);
	if (!p) {
		void *q = nullptr;
		NAX(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
		p = (unsigned char *)q;
		g_ptr[dev].store(p, std::memory_order_release);
	}
	return p;
}

template <typename T> void d2h(T *dst, const unsigned char *src, size_t count, hipStream_t s)
{
	NAX(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
	NAX(hipStreamSynchronize(s));
}

void check_launch() { NAX(hipGetLastError()); }

int thresholds(const float *d_img, int64_t nvox, int n, const int *modes, float *out, hipStream_t s)
{
	unsigned char *base = device_scratch();
	uint32_t *stats = (uint32_t *)(base + OFF_STATS);
	uint32_t *h1001 = (uint32_t *)(base + OFF_H1001);
	uint32_t *h256 = (uint32_t *)(base + OFF_H256);
	launch_threshold_init(stats, h1001, h256, s);
	check_launch();
	// robust range (src/isolevel.c:35-139)
	float mn, mx, pct2 = 0.0f, pct98 = 1.0f;
	launch_hist_range(d_img, (size_t)nvox, stats, s);
	check_launch();
	uint32_t st[3];
	d2h(st, base + OFF_STATS, 3, s);
	const size_t nNan = st[2];
	mn = ord2f_host(st[0]);
	mx = ord2f_host(st[1]);
	if ((size_t)nvox == nNan) {
		// all NaN: 0..1 (src/isolevel.c:67-68)
	} else if (mn == mx) {
		pct2 = mn;
		pct98 = mx;
	} else {
		const size_t valid = (size_t)nvox - nNan;
		const size_t n2pct = (size_t)round(valid * 0.02);
		if (n2pct < 1 || valid < 100) {
			pct2 = mn;
			pct98 = mx;
		} else {
			const int kBins = 1001;
			const float scl = (kBins - 1) / (mx - mn);
			launch_hist_bins(d_img, (size_t)nvox, mn, scl, kBins, h1001, s);
			check_launch();
			uint32_t h[kBins];
			d2h(h, base + OFF_H1001, kBins, s);
			int lo, hi;
			robust_walk(h, kBins, n2pct, &lo, &hi);
			pct2 = (lo) / scl + mn;
			pct98 = (hi) / scl + mn;
		}
	}
	mn = pct2;
	mx = pct98;
	// 256-bin histogram and Otsu (src/isolevel.c:141-277)
	const int nBin = 256;
	const float scl = (nBin - 1) / (mx - mn);
	launch_hist_bins(d_img, (size_t)nvox, mn, scl, nBin, h256, s);
	check_launch();
	uint32_t H[nBin];
	d2h(H, base + OFF_H256, nBin, s);
	double Sum = 0.0;
	for (int v = 0; v < nBin; v++)
		Sum = Sum + (double)(int)H[v];
	int dark5 = 0, bright5 = 0, mid3 = 0;
	bool need5 = false, need3 = false;
	for (int i = 0; i < n; i++)
		(modes[i] == 2 ? need3 : need5) = true;
	if (Sum > 0) { // an empty histogram gives 0 for every threshold
		double *P = (double *)(base + OFF_P);
		launch_otsu_table((const uint32_t *)h256, P, s);
		launch_otsu_probe(P, (double *)(base + OFF_PROBE), s);
		check_launch();
		double probe[2 * nBin + 4];
		d2h(probe, base + OFF_PROBE, 2 * nBin + 4, s);
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
			int lo = (int)(0.25 * nBin), mi = (int)(0.50 * nBin), hi = (int)(0.75 * nBin);
			double best = probe[2 * nBin] + probe[2 * nBin + 1] + probe[2 * nBin + 2] + probe[2 * nBin + 3];
			const int nl = nBin - 3;
			launch_otsu3(P, nBin, (double *)(base + OFF_BV), (int *)(base + OFF_BM), (int *)(base + OFF_BH), s);
			check_launch();
			double bv[nBin];
			int bm[nBin], bh[nBin];
			NAX(hipMemcpyAsync(bv, base + OFF_BV, (size_t)nl * 8, hipMemcpyDeviceToHost, s));
			NAX(hipMemcpyAsync(bm, base + OFF_BM, (size_t)nl * 4, hipMemcpyDeviceToHost, s));
			NAX(hipMemcpyAsync(bh, base + OFF_BH, (size_t)nl * 4, hipMemcpyDeviceToHost, s));
			NAX(hipStreamSynchronize(s));
			for (int l = 0; l < nl; l++) // first strict maximum in l order
				if (bm[l] >= 0 && bv[l] > best) {
					lo = l;
					mi = bm[l];
					hi = bh[l];
					best = bv[l];
				}
			dark5 = lo;
			bright5 = hi;
		}
	}
	for (int i = 0; i < n; i++) { // src/isolevel.c:264-276
		const int t = modes[i] == 1 ? dark5 : (modes[i] == 3 ? bright5 : mid3);
		out[i] = (t / scl) + mn;
	}
	return EXIT_SUCCESS;
}

} // namespace

extern "C" int n2m_set_thresholds_noalloc(const float *d_img, int64_t nvox, int n, const int *modes, float *out,
                                          void *stream)
{
	clear_error();
	char msg[160];
	if (nvox < 1 || (uint64_t)nvox > (uint64_t)N2M_NOALLOC_MAX_VOXELS) {
		snprintf(msg, sizeof msg, "n2m_set_thresholds_noalloc: %lld voxels; 1 to %llu supported (32-bit bin counts)",
		         (long long)nvox, (unsigned long long)N2M_NOALLOC_MAX_VOXELS);
		set_error(msg);
		return EXIT_FAILURE;
	}
	if (n < 1 || n > N2M_NOALLOC_MAX_MODES || !modes || !out || !d_img) {
		snprintf(msg, sizeof msg, "n2m_set_thresholds_noalloc: %d modes; 1 to %d supported", n,
		         N2M_NOALLOC_MAX_MODES);
		set_error(msg);
		return EXIT_FAILURE;
	}
	for (int i = 0; i < n; i++)
		if (modes[i] < 1 || modes[i] > 3) {
			snprintf(msg, sizeof msg, "n2m_set_thresholds_noalloc: mode %d; expected 1, 2 or 3", modes[i]);
			set_error(msg);
			return EXIT_FAILURE;
		}
	try {
		return thresholds(d_img, nvox, n, modes, out, (hipStream_t)stream);
	} catch (const Fail &f) {
		set_error(f.msg);
		return EXIT_FAILURE;
	}
}
