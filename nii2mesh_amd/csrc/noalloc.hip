// noalloc.hip -- n2m_set_thresholds_noalloc: the isolevels of setThreshold()
// (src/isolevel.c:245-277) for a device volume, with no allocation at all.
//
// n2m_ctx_set_thresholds keeps its histograms, range words and Otsu table in
// the context's device buffers, which the library allocates.  This entry
// takes the caller's stream and uses one fixed device array instead
// (`scratch` below, 1 MiB per device, held until the process ends); host
// side it uses only a fixed-size stack record.  Every call on a device
// shares that device's array: calls that overlap on one device race.
// Whatever the previous call left there is cleared by a kernel before use.
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
constexpr size_t OFF_REC = 0;                                           // the chain's record
constexpr size_t OFF_H1001 = (sizeof(ThrRecord) + 255) / 256 * 256;    // 1024 words (1001 bins used)
constexpr size_t OFF_P = OFF_H1001 + 4096;                              // 256 x 256 doubles: the Otsu table
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

int thresholds(const float *d_img, int64_t nvox, int n, const int *modes, float *out, hipStream_t s)
{
	unsigned char *base = device_scratch();
	bool need5 = false;
	for (int i = 0; i < n; i++)
		need5 = need5 || modes[i] != 2;
	// the whole of setThreshold() on the device (src/isolevel.c:35-277), one
	// copy of the record back
	launch_threshold_chain(d_img, (size_t)nvox, (ThrRecord *)(base + OFF_REC), (uint32_t *)(base + OFF_H1001),
	                       (double *)(base + OFF_P), need5, s);
	NAX(hipGetLastError());
	ThrRecord rec;
	NAX(hipMemcpyAsync(&rec, base + OFF_REC, sizeof rec, hipMemcpyDeviceToHost, s));
	NAX(hipStreamSynchronize(s));
	thresholds_from_record(rec, n, modes, out);
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
