// Error reporting and device queries for the C-ABI.
#include "bm_internal.h"

#include <algorithm>
#include <vector>

namespace bm {

namespace { thread_local std::string g_err; }

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

int hip_fail(hipError_t e, const char* what)
{
	g_err = std::string(what) + ": " + hipGetErrorString(e);
	return e == hipErrorOutOfMemory ? BM_ERR_NOMEM : BM_ERR_HIP;
}

namespace {
struct Prof {
	bool on = false;
	int cap = 0, used = 0, open_kind = -1;
	std::vector<hipEvent_t> a, b;
	std::vector<int> kind;
	std::vector<hipStream_t> stream;
};
Prof g_prof;
}

void prof_mark_begin(int kind, hipStream_t st)
{
	if (!g_prof.on || g_prof.used >= g_prof.cap) return;
	(void)hipEventRecord(g_prof.a[g_prof.used], st);
	g_prof.open_kind = kind;
}

void prof_mark_end(int kind, hipStream_t st)
{
	if (!g_prof.on || g_prof.used >= g_prof.cap || g_prof.open_kind != kind) return;
	(void)hipEventRecord(g_prof.b[g_prof.used], st);
	g_prof.kind[g_prof.used] = kind;
	g_prof.stream[g_prof.used] = st;
	g_prof.used++;
	g_prof.open_kind = -1;
}

int device_cu_count()
{
	int dev = 0, n = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 256;
	if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
	return n;
}

} // namespace bm

extern "C" const char* bm_last_error(void) { return bm::g_err.c_str(); }

extern "C" int bm_prof_begin(int max_launches)
{
	using bm::g_prof;
	if (max_launches < 1) return bm::fail(BM_ERR_ARG, "bm_prof_begin: max_launches < 1");
	for (auto e : g_prof.a) (void)hipEventDestroy(e);
	for (auto e : g_prof.b) (void)hipEventDestroy(e);
	g_prof = bm::Prof();
	g_prof.a.resize(max_launches); g_prof.b.resize(max_launches);
	g_prof.kind.assign(max_launches, -1); g_prof.stream.assign(max_launches, nullptr);
	for (int i = 0; i < max_launches; i++) {
		if (hipEventCreate(&g_prof.a[i]) != hipSuccess || hipEventCreate(&g_prof.b[i]) != hipSuccess)
			return bm::fail(BM_ERR_HIP, "bm_prof_begin: hipEventCreate");
	}
	g_prof.cap = max_launches;
	g_prof.on = true;
	return BM_OK;
}

extern "C" int bm_prof_end(void)
{
	bm::g_prof.on = false;
	return BM_OK;
}

extern "C" int bm_prof_query(int kind, long long* count, double* total_ms, double* min_ms, double* max_ms)
{
	using bm::g_prof;
	long long c = 0; double tot = 0, mn = 1e300, mx = 0;
	for (int i = 0; i < g_prof.used; i++) {
		if (g_prof.kind[i] != kind) continue;
		if (hipEventSynchronize(g_prof.b[i]) != hipSuccess) return bm::fail(BM_ERR_HIP, "bm_prof_query: sync");
		float ms = 0;
		if (hipEventElapsedTime(&ms, g_prof.a[i], g_prof.b[i]) != hipSuccess) return bm::fail(BM_ERR_HIP, "bm_prof_query: elapsed");
		c++; tot += ms; mn = std::min(mn, (double)ms); mx = std::max(mx, (double)ms);
	}
	if (count) *count = c;
	if (total_ms) *total_ms = tot;
	if (min_ms) *min_ms = c ? mn : 0;
	if (max_ms) *max_ms = mx;
	return BM_OK;
}

extern "C" int bm_device_ok(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
	return std::string(prop.gcnArchName).rfind("gfx950", 0) == 0 ? 1 : 0;
}
