/*
 * ptk_hip.h -- the thin C ABI through which the PTask HIP backend reaches the device.
 *
 * Every HIP runtime call made by HIPAccelerator, HIPTask, PHIPBuffer and HIPAsyncContext goes
 * through one of these entry points (libptk_hip.so).  The library is plain C at its boundary:
 * opaque handles, raw pointers and byte counts, an int status (PTK_HIP_OK == 0) from every call,
 * and nothing ever throws across it.  Handles are valid on the device that was current when they
 * were created; the current device is per thread, as in HIP itself.
 *
 * Each declaration names the reference call site whose CUDA driver-API call it takes over.
 * Reference paths are relative to the PTask source root.
 */
#ifndef PTK_HIP_H
#define PTK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTK_HIP_OK              0
#define PTK_HIP_NOT_READY       1    /* query of a stream/event whose work is still pending */
#define PTK_HIP_ERR_INVALID     2    /* bad handle, null pointer or out-of-range argument */
#define PTK_HIP_ERR_NO_DEVICE   3    /* no usable HIP device in this process */
#define PTK_HIP_ERR_OOM         4    /* device or pinned-host allocation failed */
#define PTK_HIP_ERR_NOT_FOUND   5    /* code object file or kernel symbol not found */
#define PTK_HIP_ERR_LAUNCH      6    /* the runtime rejected a launch */
#define PTK_HIP_ERR_RUNTIME     7    /* any other HIP runtime failure; see ptk_hip_last_error_string */

typedef struct ptk_hip_module_s   *ptk_hip_module;    /* a loaded code object (.hsaco) */
typedef struct ptk_hip_function_s *ptk_hip_function;  /* a kernel symbol inside a module */
typedef struct ptk_hip_stream_s   *ptk_hip_stream;    /* an in-order queue of device work */
typedef struct ptk_hip_event_s    *ptk_hip_event;     /* a point in a stream's work */
typedef struct ptk_hip_graph_s    *ptk_hip_graph;     /* a captured, instantiated work graph */

typedef struct ptk_hip_device_props_t {
    char     name[256];             /* marketing name */
    char     arch[64];              /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int      compute_units;
    int      clock_khz;
    int      mem_clock_khz;
    int      mem_bus_width_bits;
    int      l2_bytes;
    int      max_threads_per_block;
    int      warp_size;             /* 64 on CDNA */
    int      lds_bytes_per_block;
    int      concurrent_kernels;
    int      runtime_version;
    size_t   total_global_mem;
} ptk_hip_device_props_t;

/* ---- devices ---------------------------------------------------------------------------- */

/* cuInit + cuDeviceGetCount (ptask/cuaccelerator.cpp:341,347). 0 devices is not an error. */
int ptk_hip_get_device_count(int *count);
/* cuDeviceGetName / cuDeviceComputeCapability / cuDeviceGetAttribute
 * (ptask/cuaccelerator.cpp:359-383). */
int ptk_hip_get_device_props(int device, ptk_hip_device_props_t *props);
/* cuCtxPushCurrent / cuCtxSetCurrent (ptask/cuaccelerator.cpp:2247, ptask/cutask.cpp:43-54). */
int ptk_hip_set_device(int device);
int ptk_hip_get_device(int *device);
/* cuCtxSynchronize (ptask/cuaccelerator.cpp:1971, ptask/cutask.cpp:691). */
int ptk_hip_device_synchronize(void);
/* cuMemGetInfo (ptask/DeviceMemoryStatus.cpp). */
int ptk_hip_mem_get_info(size_t *free_bytes, size_t *total_bytes);

/* ---- memory ----------------------------------------------------------------------------- */

/* cuMemAlloc (ptask/PCUBuffer.cpp:971). Device memory, 256-byte aligned. */
int ptk_hip_malloc(void **dptr, size_t bytes);
/* cuMemFree (ptask/PCUBuffer.cpp:43-51). */
int ptk_hip_free(void *dptr);
/* cuMemHostAlloc, page-locked (ptask/cuaccelerator.cpp AllocatePagelockedHostMemory). */
int ptk_hip_host_malloc(void **hptr, size_t bytes);
int ptk_hip_host_free(void *hptr);
/* cuMemcpyHtoDAsync (ptask/PCUBuffer.cpp:591,669). */
int ptk_hip_memcpy_htod_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s);
/* cuMemcpyDtoHAsync (ptask/PCUBuffer.cpp:272). */
int ptk_hip_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s);
/* cuMemcpyDtoDAsync (ptask/PCUBuffer.cpp:756). */
int ptk_hip_memcpy_dtod_async(void *dst, const void *src, size_t bytes, ptk_hip_stream s);
/* Synchronous forms used when no stream is bound (ptask/PCUBuffer.cpp:143). */
int ptk_hip_memcpy_htod(void *dst, const void *src, size_t bytes);
int ptk_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes);

/* ---- streams ---------------------------------------------------------------------------- */

/* cuStreamCreate (ptask/CUAsyncContext.cpp:142). Always the default priority. */
int ptk_hip_stream_create(ptk_hip_stream *s);
/* cuStreamDestroy (ptask/CUAsyncContext.cpp:75). */
int ptk_hip_stream_destroy(ptk_hip_stream s);
/* cuStreamSynchronize (ptask/CUAsyncContext.cpp:473, ptask/cutask.cpp:689). */
int ptk_hip_stream_synchronize(ptk_hip_stream s);
/* Returns PTK_HIP_OK when all work is done, PTK_HIP_NOT_READY otherwise. */
int ptk_hip_stream_query(ptk_hip_stream s);
/* cuStreamWaitEvent (ptask/CUAsyncContext.cpp:287). */
int ptk_hip_stream_wait_event(ptk_hip_stream s, ptk_hip_event e);

/* ---- events ----------------------------------------------------------------------------- */

/* cuEventCreate (ptask/CUAsyncContext.cpp:193-198; ptask/cutask.cpp:720 for timing events).
 * timing != 0 creates an event usable with ptk_hip_event_elapsed_ms. */
int ptk_hip_event_create(ptk_hip_event *e, int timing);
/* cuEventDestroy (ptask/CUAsyncContext.cpp:248). */
int ptk_hip_event_destroy(ptk_hip_event e);
/* cuEventRecord (ptask/CUAsyncContext.cpp:211). */
int ptk_hip_event_record(ptk_hip_event e, ptk_hip_stream s);
/* cuEventQuery (ptask/CUAsyncContext.cpp:353,405,444): PTK_HIP_OK when complete,
 * PTK_HIP_NOT_READY when pending. */
int ptk_hip_event_query(ptk_hip_event e);
/* cuEventSynchronize (ptask/CUAsyncContext.cpp:365). */
int ptk_hip_event_synchronize(ptk_hip_event e);
/* cuEventElapsedTime (ptask/cutask.cpp:773). */
int ptk_hip_event_elapsed_ms(float *ms, ptk_hip_event start, ptk_hip_event stop);

/* ---- code objects and launch ------------------------------------------------------------ */

/* cuModuleLoadDataEx (ptask/cuaccelerator.cpp:1379), from a gfx950 code object file. */
int ptk_hip_module_load(ptk_hip_module *m, const char *path);
/* Same, from an in-memory image. */
int ptk_hip_module_load_data(ptk_hip_module *m, const void *image);
int ptk_hip_module_unload(ptk_hip_module m);
/* cuModuleGetFunction (ptask/cuaccelerator.cpp:1420). */
int ptk_hip_module_get_function(ptk_hip_function *f, ptk_hip_module m, const char *name);
/* Static resource use of a kernel: registers, LDS, spill scratch, max block size. */
int ptk_hip_function_get_attrs(ptk_hip_function f, int *num_vgprs, int *lds_bytes,
                               int *scratch_bytes, int *max_threads_per_block);
/* cuParamSetSize + cuParamSetv/i/f + cuFuncSetBlockShape + cuLaunchGridAsync
 * (ptask/cutask.cpp:633-697). `args` is the packed kernel-argument buffer laid out as the
 * kernel's formal parameters (each aligned to its natural alignment, buffers as 8-byte device
 * pointers, by-value structs as raw bytes); `args_bytes` is its total size. */
int ptk_hip_launch(ptk_hip_function f,
                   unsigned grid_x, unsigned grid_y, unsigned grid_z,
                   unsigned block_x, unsigned block_y, unsigned block_z,
                   unsigned lds_bytes, ptk_hip_stream s,
                   void *args, size_t args_bytes);

/* ---- captured work graphs (no reference counterpart: the CUDA backend launches every
 * ---- dispatch individually, ptask/cutask.cpp:675) ------------------------------------------ */

/* Starts recording the work issued to `s` instead of running it. */
int ptk_hip_stream_begin_capture(ptk_hip_stream s);
/* Ends recording and instantiates the recorded work as a launchable graph. */
int ptk_hip_stream_end_capture(ptk_hip_stream s, ptk_hip_graph *g);
/* Replays an instantiated graph on `s`. */
int ptk_hip_graph_launch(ptk_hip_graph g, ptk_hip_stream s);
int ptk_hip_graph_destroy(ptk_hip_graph g);

/* ---- diagnostics ------------------------------------------------------------------------ */

/* Static text for a PTK_HIP_* status. */
const char *ptk_hip_status_string(int status);
/* Text of the most recent HIP failure seen on the calling thread ("" if none). */
const char *ptk_hip_last_error_string(void);

#ifdef __cplusplus
}
#endif

#endif /* PTK_HIP_H */
