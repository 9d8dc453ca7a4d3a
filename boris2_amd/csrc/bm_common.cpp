// Error reporting and device queries for the C-ABI.
#include "bm_internal.h"

#include <mutex>

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

int device_cu_count()
{
	int dev = 0, n = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 256;
	if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
	return n;
}

} // namespace bm

extern "C" const char* bm_last_error(void) { return bm::g_err.c_str(); }

extern "C" int bm_device_ok(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
	return std::string(prop.gcnArchName).rfind("gfx950", 0) == 0 ? 1 : 0;
}
