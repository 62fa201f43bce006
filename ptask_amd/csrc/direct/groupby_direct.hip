// groupby_direct.hip -- libptk_groupby.so: the groupby graph's task kernels as direct library
// calls (include/ptk_groupby_api.h).
//
// The kernels are the graph's own (groupby.hip is compiled into this library unchanged), launched
// on the caller's stream in the order the graph dispatches them, with the geometry the groupby
// application gives its tasks.  Results are therefore bitwise the graph's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <string>

#include "groupby.hip"
#include "ptk_groupby_api.h"

namespace {

thread_local std::string g_lastError;

int Fail(int code, const std::string &what) {
    g_lastError = what;
    return code;
}

int HipFail(hipError_t e, const char *what) {
    return Fail(PTK_GB_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

long long Align256(long long b) { return (b + 255) / 256 * 256; }

// table (count | sum | cursor | errors), then the scratch of include/ptk_groupby.h
long long TableBytes(int numKeys) { return Align256(8 * (3 * ((long long)numKeys + 1) + 1)); }

bool ShapeOk(long long n, int numKeys) {
    return n >= 1 && n <= INT_MAX && numKeys >= 0 && (long long)numKeys + 1 <= PTK_GB_MAX_KEYS;
}

// The seven launches on `stream`; table and scratch lie in device memory the caller provides.
int Launch(const int *keys, const int *vals, int n, int numKeys, long long *table, void *scratch, long long *offsets,
           long long *groups, int *outKeys, int *outVals, hipStream_t stream) {
    const unsigned chunks = (unsigned)ptk_gb_chunks(n), nb = (unsigned)ptk_gb_buckets(numKeys);
    const unsigned tiles = (unsigned)ptk_gb_tiles_max(n, numKeys);
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const unsigned partGrid = std::min(chunks, (unsigned)(2 * std::max(cus, 1)));  // two blocks per CU
    hipLaunchKernelGGL(GroupbyHistogram, dim3(chunks), dim3(PTK_GB_HIST_THREADS), 0, stream, keys, n, numKeys, scratch);
    hipLaunchKernelGGL(GroupbyBucketScan, dim3(nb), dim3(PTK_GB_SCAN_THREADS), 0, stream, scratch, n, numKeys);
    hipLaunchKernelGGL(GroupbyPartition, dim3(partGrid), dim3(PTK_GB_PART_THREADS), 0, stream, keys, vals, n, numKeys,
                       scratch);
    hipLaunchKernelGGL(GroupbyKeyCount, dim3(tiles), dim3(PTK_GB_COUNT_THREADS), 0, stream, scratch, n, numKeys);
    hipLaunchKernelGGL(GroupbyKeyScan, dim3(2 * nb), dim3(PTK_GB_KEYSCAN_THREADS), 0, stream, scratch, n, numKeys, offsets);
    hipLaunchKernelGGL(GroupbyReshuffle, dim3(tiles), dim3(PTK_GB_TILE_THREADS), 0, stream, scratch, n, numKeys,
                       (const long long *)offsets, outKeys, outVals);
    hipLaunchKernelGGL(GroupbyTotals, dim3(4 * nb), dim3(PTK_GB_TOTALS_THREADS), 0, stream, scratch, n, numKeys, table,
                       groups);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PTK_GB_OK : HipFail(e, "groupby launch");
}

}  // namespace

extern "C" long long ptk_groupby_workspace_bytes(long long n, int numKeys) {
    if (!ShapeOk(n, numKeys)) return -1;
    return TableBytes(numKeys) + ptk_gb_scratch_bytes(n, numKeys);
}

extern "C" int ptk_groupby_run(const int *keys, const int *vals, long long n, int numKeys, void *workspace,
                               unsigned long long workspaceBytes, long long *offsets, long long *groups, int *outKeys,
                               int *outVals, void *stream) {
    if (!ShapeOk(n, numKeys)) return Fail(PTK_GB_EINVAL, "ptk_groupby_run: n must be in [1, 2^31 - 1] and numKeys in [0, 2^19 - 2]");
    if (!keys || !vals || !workspace || !offsets || !groups || !outKeys || !outVals)
        return Fail(PTK_GB_EINVAL, "ptk_groupby_run: null device pointer");
    if (((unsigned long long)workspace & 15) != 0) return Fail(PTK_GB_EINVAL, "ptk_groupby_run: workspace is not 16-byte aligned");
    const long long need = ptk_groupby_workspace_bytes(n, numKeys);
    if (workspaceBytes < (unsigned long long)need)
        return Fail(PTK_GB_ESPACE, "ptk_groupby_run: workspace of " + std::to_string(workspaceBytes) + " bytes, " +
                                       std::to_string(need) + " needed");
    long long *table = (long long *)workspace;
    void *scratch = (char *)workspace + TableBytes(numKeys);
    return Launch(keys, vals, (int)n, numKeys, table, scratch, offsets, groups, outKeys, outVals, (hipStream_t)stream);
}

// ---- the no-allocation entry point ------------------------------------------------------------
// Its workspace is this file-scope device array: each GPU the library runs on holds one copy for
// the life of the process.  Host code never passes the array's name to a kernel (on the host it
// is a stand-in, not a device address): the first call on a device, with that device current,
// asks the runtime for the device's own address of the array and keeps it per device.
static __device__ __align__(256) unsigned char scratch[PTK_GB_NOALLOC_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchPtr[kMaxDevices] = {};

int ScratchOfCurrentDevice(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return HipFail(e, "ptk_groupby_run_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return Fail(PTK_GB_EINVAL, "ptk_groupby_run_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return HipFail(e, "ptk_groupby_run_noalloc: hipGetSymbolAddress");
        if (!p) return Fail(PTK_GB_EHIP, "ptk_groupby_run_noalloc: no device address for the workspace");
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return PTK_GB_OK;
}

}  // namespace

extern "C" unsigned long long ptk_groupby_noalloc_capacity(void) { return PTK_GB_NOALLOC_BYTES; }

extern "C" int ptk_groupby_run_noalloc(const int *keys, const int *vals, long long n, int numKeys, long long *offsets,
                                       long long *groups, int *outKeys, int *outVals, void *stream) {
    if (!ShapeOk(n, numKeys))
        return Fail(PTK_GB_EINVAL, "ptk_groupby_run_noalloc: n must be in [1, 2^31 - 1] and numKeys in [0, 2^19 - 2]");
    if (!keys || !vals || !offsets || !groups || !outKeys || !outVals)
        return Fail(PTK_GB_EINVAL, "ptk_groupby_run_noalloc: null device pointer");
    const long long need = ptk_groupby_workspace_bytes(n, numKeys);
    if ((unsigned long long)need > PTK_GB_NOALLOC_BYTES)
        return Fail(PTK_GB_ESPACE, "ptk_groupby_run_noalloc: " + std::to_string(n) + " rows over " +
                                       std::to_string((long long)numKeys + 1) + " keys need " + std::to_string(need) +
                                       " bytes of workspace; the static workspace holds " +
                                       std::to_string((unsigned long long)PTK_GB_NOALLOC_BYTES));
    void *ws = nullptr;
    if (int r = ScratchOfCurrentDevice(&ws)) return r;
    long long *table = (long long *)ws;
    void *s = (char *)ws + TableBytes(numKeys);
    return Launch(keys, vals, (int)n, numKeys, table, s, offsets, groups, outKeys, outVals, (hipStream_t)stream);
}

extern "C" const char *ptk_groupby_last_error(void) { return g_lastError.c_str(); }
