// ptk_hip.cpp -- implementation of the thin HIP C ABI declared in include/ptk_hip.h.
//
// Handles are the HIP runtime's own handles cast to opaque pointer types, so no wrapper objects
// are allocated and nothing needs freeing beyond what HIP itself owns.  Every entry point maps
// the hipError_t it gets onto a PTK_HIP_* status and remembers the HIP error text per thread.

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

#include "ptk_hip.h"

namespace {

thread_local char t_last_error[512] = {0};

int record(hipError_t err, const char *what) {
    if (err == hipSuccess) return PTK_HIP_OK;
    if (err == hipErrorNotReady) return PTK_HIP_NOT_READY;
    std::snprintf(t_last_error, sizeof(t_last_error), "%s: %s (%d)", what, hipGetErrorString(err),
                  static_cast<int>(err));
    (void)hipGetLastError();  // clear the sticky per-thread error so later calls start clean
    switch (err) {
        case hipErrorInvalidValue:
        case hipErrorInvalidHandle:
        case hipErrorInvalidDevice:
        case hipErrorInvalidDevicePointer: return PTK_HIP_ERR_INVALID;
        case hipErrorNoDevice:
        case hipErrorInsufficientDriver: return PTK_HIP_ERR_NO_DEVICE;
        case hipErrorOutOfMemory: return PTK_HIP_ERR_OOM;
        case hipErrorFileNotFound:
        case hipErrorNotFound:
        case hipErrorSharedObjectSymbolNotFound:
        case hipErrorInvalidImage: return PTK_HIP_ERR_NOT_FOUND;
        case hipErrorLaunchFailure:
        case hipErrorInvalidConfiguration:
        case hipErrorLaunchOutOfResources: return PTK_HIP_ERR_LAUNCH;
        default: return PTK_HIP_ERR_RUNTIME;
    }
}

int invalid(const char *what) {
    std::snprintf(t_last_error, sizeof(t_last_error), "%s: invalid argument", what);
    return PTK_HIP_ERR_INVALID;
}

inline hipStream_t S(ptk_hip_stream s) { return reinterpret_cast<hipStream_t>(s); }
inline hipEvent_t E(ptk_hip_event e) { return reinterpret_cast<hipEvent_t>(e); }

}  // namespace

extern "C" {

int ptk_hip_get_device_count(int *count) {
    if (!count) return invalid("ptk_hip_get_device_count");
    *count = 0;
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err == hipErrorNoDevice || err == hipErrorInsufficientDriver) {
        (void)hipGetLastError();
        return PTK_HIP_OK;  // no device is a valid answer, not a failure
    }
    if (err != hipSuccess) return record(err, "hipGetDeviceCount");
    *count = n;
    return PTK_HIP_OK;
}

int ptk_hip_get_device_props(int device, ptk_hip_device_props_t *props) {
    if (!props) return invalid("ptk_hip_get_device_props");
    hipDeviceProp_t p;
    hipError_t err = hipGetDeviceProperties(&p, device);
    if (err != hipSuccess) return record(err, "hipGetDeviceProperties");
    std::memset(props, 0, sizeof(*props));
    std::snprintf(props->name, sizeof(props->name), "%s", p.name);
    std::snprintf(props->arch, sizeof(props->arch), "%s", p.gcnArchName);
    props->compute_units = p.multiProcessorCount;
    props->clock_khz = p.clockRate;
    props->mem_clock_khz = p.memoryClockRate;
    props->mem_bus_width_bits = p.memoryBusWidth;
    props->l2_bytes = p.l2CacheSize;
    props->max_threads_per_block = p.maxThreadsPerBlock;
    props->warp_size = p.warpSize;
    props->lds_bytes_per_block = static_cast<int>(p.sharedMemPerBlock);
    props->concurrent_kernels = p.concurrentKernels;
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    props->runtime_version = rt;
    props->total_global_mem = p.totalGlobalMem;
    return PTK_HIP_OK;
}

int ptk_hip_set_device(int device) { return record(hipSetDevice(device), "hipSetDevice"); }

int ptk_hip_get_device(int *device) {
    if (!device) return invalid("ptk_hip_get_device");
    return record(hipGetDevice(device), "hipGetDevice");
}

int ptk_hip_device_synchronize(void) {
    return record(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int ptk_hip_mem_get_info(size_t *free_bytes, size_t *total_bytes) {
    if (!free_bytes || !total_bytes) return invalid("ptk_hip_mem_get_info");
    return record(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int ptk_hip_malloc(void **dptr, size_t bytes) {
    if (!dptr) return invalid("ptk_hip_malloc");
    *dptr = nullptr;
    if (bytes == 0) return PTK_HIP_OK;
    return record(hipMalloc(dptr, bytes), "hipMalloc");
}

int ptk_hip_free(void *dptr) {
    if (!dptr) return PTK_HIP_OK;
    return record(hipFree(dptr), "hipFree");
}

int ptk_hip_host_malloc(void **hptr, size_t bytes) {
    if (!hptr) return invalid("ptk_hip_host_malloc");
    *hptr = nullptr;
    if (bytes == 0) return PTK_HIP_OK;
    return record(hipHostMalloc(hptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}

int ptk_hip_host_free(void *hptr) {
    if (!hptr) return PTK_HIP_OK;
    return record(hipHostFree(hptr), "hipHostFree");
}

int ptk_hip_memcpy_htod_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s) {
    if (bytes == 0) return PTK_HIP_OK;
    if (!dst || !src) return invalid("ptk_hip_memcpy_htod_async");
    return record(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)),
                  "hipMemcpyAsync(HtoD)");
}

int ptk_hip_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s) {
    if (bytes == 0) return PTK_HIP_OK;
    if (!dst || !src) return invalid("ptk_hip_memcpy_dtoh_async");
    return record(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)),
                  "hipMemcpyAsync(DtoH)");
}

int ptk_hip_memcpy_dtod_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s) {
    if (bytes == 0) return PTK_HIP_OK;
    if (!dst || !src) return invalid("ptk_hip_memcpy_dtod_async");
    return record(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)),
                  "hipMemcpyAsync(DtoD)");
}

int ptk_hip_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return PTK_HIP_OK;
    if (!dst || !src) return invalid("ptk_hip_memcpy_htod");
    return record(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(HtoD)");
}

int ptk_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return PTK_HIP_OK;
    if (!dst || !src) return invalid("ptk_hip_memcpy_dtoh");
    return record(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy(DtoH)");
}

int ptk_hip_stream_create(ptk_hip_stream *s) {
    if (!s) return invalid("ptk_hip_stream_create");
    hipStream_t h = nullptr;
    // Non-blocking with respect to the legacy null stream; priority stays the default.
    int rc = record(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    *s = reinterpret_cast<ptk_hip_stream>(h);
    return rc;
}

int ptk_hip_stream_destroy(ptk_hip_stream s) {
    if (!s) return PTK_HIP_OK;
    return record(hipStreamDestroy(S(s)), "hipStreamDestroy");
}

int ptk_hip_stream_synchronize(ptk_hip_stream s) {
    return record(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
}

int ptk_hip_stream_query(ptk_hip_stream s) {
    return record(hipStreamQuery(S(s)), "hipStreamQuery");
}

int ptk_hip_stream_wait_event(ptk_hip_stream s, ptk_hip_event e) {
    if (!e) return invalid("ptk_hip_stream_wait_event");
    return record(hipStreamWaitEvent(S(s), E(e), 0), "hipStreamWaitEvent");
}

int ptk_hip_event_create(ptk_hip_event *e, int timing) {
    if (!e) return invalid("ptk_hip_event_create");
    hipEvent_t h = nullptr;
    unsigned flags = timing ? hipEventDefault : hipEventDisableTiming;
    int rc = record(hipEventCreateWithFlags(&h, flags), "hipEventCreateWithFlags");
    *e = reinterpret_cast<ptk_hip_event>(h);
    return rc;
}

int ptk_hip_event_destroy(ptk_hip_event e) {
    if (!e) return PTK_HIP_OK;
    return record(hipEventDestroy(E(e)), "hipEventDestroy");
}

int ptk_hip_event_record(ptk_hip_event e, ptk_hip_stream s) {
    if (!e) return invalid("ptk_hip_event_record");
    return record(hipEventRecord(E(e), S(s)), "hipEventRecord");
}

int ptk_hip_event_query(ptk_hip_event e) {
    if (!e) return invalid("ptk_hip_event_query");
    return record(hipEventQuery(E(e)), "hipEventQuery");
}

int ptk_hip_event_synchronize(ptk_hip_event e) {
    if (!e) return invalid("ptk_hip_event_synchronize");
    return record(hipEventSynchronize(E(e)), "hipEventSynchronize");
}

int ptk_hip_event_elapsed_ms(float *ms, ptk_hip_event start, ptk_hip_event stop) {
    if (!ms || !start || !stop) return invalid("ptk_hip_event_elapsed_ms");
    return record(hipEventElapsedTime(ms, E(start), E(stop)), "hipEventElapsedTime");
}

int ptk_hip_module_load(ptk_hip_module *m, const char *path) {
    if (!m || !path) return invalid("ptk_hip_module_load");
    hipModule_t h = nullptr;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        std::snprintf(t_last_error, sizeof(t_last_error), "code object not found: %s", path);
        *m = nullptr;
        return PTK_HIP_ERR_NOT_FOUND;
    }
    std::fclose(fp);
    int rc = record(hipModuleLoad(&h, path), "hipModuleLoad");
    *m = reinterpret_cast<ptk_hip_module>(h);
    return rc;
}

int ptk_hip_module_load_data(ptk_hip_module *m, const void *image) {
    if (!m || !image) return invalid("ptk_hip_module_load_data");
    hipModule_t h = nullptr;
    int rc = record(hipModuleLoadData(&h, image), "hipModuleLoadData");
    *m = reinterpret_cast<ptk_hip_module>(h);
    return rc;
}

int ptk_hip_module_unload(ptk_hip_module m) {
    if (!m) return PTK_HIP_OK;
    return record(hipModuleUnload(reinterpret_cast<hipModule_t>(m)), "hipModuleUnload");
}

int ptk_hip_module_get_function(ptk_hip_function *f, ptk_hip_module m, const char *name) {
    if (!f || !m || !name) return invalid("ptk_hip_module_get_function");
    hipFunction_t h = nullptr;
    int rc = record(hipModuleGetFunction(&h, reinterpret_cast<hipModule_t>(m), name),
                    "hipModuleGetFunction");
    *f = reinterpret_cast<ptk_hip_function>(h);
    return rc;
}

int ptk_hip_function_get_attrs(ptk_hip_function f, int *num_vgprs, int *lds_bytes,
                               int *scratch_bytes, int *max_threads_per_block) {
    if (!f) return invalid("ptk_hip_function_get_attrs");
    hipFunction_t h = reinterpret_cast<hipFunction_t>(f);
    int v = 0;
    if (num_vgprs) {
        if (int rc = record(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, h),
                            "hipFuncGetAttribute(NUM_REGS)"))
            return rc;
        *num_vgprs = v;
    }
    if (lds_bytes) {
        if (int rc = record(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, h),
                            "hipFuncGetAttribute(SHARED_SIZE_BYTES)"))
            return rc;
        *lds_bytes = v;
    }
    if (scratch_bytes) {
        if (int rc = record(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, h),
                            "hipFuncGetAttribute(LOCAL_SIZE_BYTES)"))
            return rc;
        *scratch_bytes = v;
    }
    if (max_threads_per_block) {
        if (int rc = record(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, h),
                            "hipFuncGetAttribute(MAX_THREADS_PER_BLOCK)"))
            return rc;
        *max_threads_per_block = v;
    }
    return PTK_HIP_OK;
}

int ptk_hip_launch(ptk_hip_function f, unsigned grid_x, unsigned grid_y, unsigned grid_z,
                   unsigned block_x, unsigned block_y, unsigned block_z, unsigned lds_bytes,
                   ptk_hip_stream s, void *args, size_t args_bytes) {
    if (!f) return invalid("ptk_hip_launch(function)");
    if (grid_x == 0 || grid_y == 0 || grid_z == 0 || block_x == 0 || block_y == 0 ||
        block_z == 0 || static_cast<unsigned long long>(block_x) * block_y * block_z > 1024)
        return invalid("ptk_hip_launch(geometry)");
    if (args_bytes && !args) return invalid("ptk_hip_launch(args)");
    size_t size = args_bytes;
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                      HIP_LAUNCH_PARAM_END};
    hipError_t err = hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(f), grid_x, grid_y,
                                           grid_z, block_x, block_y, block_z, lds_bytes, S(s),
                                           nullptr, args_bytes ? config : nullptr);
    return record(err, "hipModuleLaunchKernel");
}

int ptk_hip_stream_begin_capture(ptk_hip_stream s) {
    if (!s) return invalid("ptk_hip_stream_begin_capture");
    return record(hipStreamBeginCapture(S(s), hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
}

int ptk_hip_stream_end_capture(ptk_hip_stream s, ptk_hip_graph *g) {
    if (!s || !g) return invalid("ptk_hip_stream_end_capture");
    *g = nullptr;
    hipGraph_t graph = nullptr;
    int rc = record(hipStreamEndCapture(S(s), &graph), "hipStreamEndCapture");
    if (rc != PTK_HIP_OK) return rc;
    hipGraphExec_t exec = nullptr;
    rc = record(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(graph);  // the executable graph keeps what it needs
    if (rc != PTK_HIP_OK) return rc;
    *g = reinterpret_cast<ptk_hip_graph>(exec);
    return PTK_HIP_OK;
}

int ptk_hip_graph_launch(ptk_hip_graph g, ptk_hip_stream s) {
    if (!g) return invalid("ptk_hip_graph_launch");
    return record(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(g), S(s)), "hipGraphLaunch");
}

int ptk_hip_graph_destroy(ptk_hip_graph g) {
    if (!g) return PTK_HIP_OK;
    return record(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g)), "hipGraphExecDestroy");
}

const char *ptk_hip_status_string(int status) {
    switch (status) {
        case PTK_HIP_OK: return "ok";
        case PTK_HIP_NOT_READY: return "not ready";
        case PTK_HIP_ERR_INVALID: return "invalid argument";
        case PTK_HIP_ERR_NO_DEVICE: return "no HIP device";
        case PTK_HIP_ERR_OOM: return "out of memory";
        case PTK_HIP_ERR_NOT_FOUND: return "code object or symbol not found";
        case PTK_HIP_ERR_LAUNCH: return "launch failed";
        case PTK_HIP_ERR_RUNTIME: return "HIP runtime error";
        default: return "unknown status";
    }
}

const char *ptk_hip_last_error_string(void) { return t_last_error; }

}  // extern "C"
