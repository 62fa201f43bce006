/*
 * primitive_types.h -- result codes, enumerations and small value types of the PTask API.
 *
 * Names and numeric values match the reference (ptask/primitive_types.h:33-1233) because they
 * are part of the application-facing contract: existing graphs pass these constants to
 * CreatePort, SetPredicationType, SetMetaFunction and friends.  One value is added:
 * ACCELERATOR_CLASS_HIP, the class of the MI355X backend.
 */
#ifndef PTK_PRIMITIVE_TYPES_H
#define PTK_PRIMITIVE_TYPES_H

#include <stdio.h>

#include "ptk_win32.h"

#ifndef PTSRELEASE
#define PTSRELEASE(p) { if (p) { (p)->Release(); (p) = NULL; } }
#endif

namespace PTask {

typedef long PTRESULT;
static const long PTASK_OK = 0;
static const long PTASK_ERR = PTASK_OK + 1;
static const long PTASK_ERR_NOT_FOUND = PTASK_ERR + 1;
static const long PTASK_ERR_INVALID_PARAMETER = PTASK_ERR + 2;
static const long PTASK_ERR_EXISTS = PTASK_ERR + 3;
static const long PTASK_ERR_GRAPH_MALFORMED = PTASK_OK + 4;
static const long PTASK_ERR_UNINITIALIZED = PTASK_OK + 5;
static const long PTASK_ERR_INFLIGHT = PTASK_OK + 6;
static const long PTASK_ERR_TOO_MANY_ACCELERATORS = PTASK_OK + 7;
static const long PTASK_ERR_ALREADY_INITIALIZED = PTASK_OK + 8;
static const long PTASK_WARNING = PTASK_OK - 1;
static const long PTASK_WARNING_GRAPH_MALFORMED = PTASK_OK - 2;
#define PTSUCCESS(x) ((x) <= PTask::PTASK_OK)
#define PTFAILED(x) ((x) > PTask::PTASK_OK)

typedef unsigned int PORTINDEX;

typedef enum affinity_type_t {
    AFFINITYTYPE_NONE = 676,
    AFFINITYTYPE_WEAK = 677,
    AFFINITYTYPE_STRONG = 678,
    AFFINITYTYPE_MANDATORY = 679
} AFFINITYTYPE;

typedef enum acctype_t {
    ACCELERATOR_CLASS_DIRECT_X = 0,
    ACCELERATOR_CLASS_OPEN_CL = 1,
    ACCELERATOR_CLASS_CUDA = 2,
    ACCELERATOR_CLASS_REFERENCE = 3,
    ACCELERATOR_CLASS_HOST = 4,
    ACCELERATOR_CLASS_UNKNOWN = 5,
    ACCELERATOR_CLASS_SUPER = 6,
    ACCELERATOR_CLASS_HIP = 7
} ACCELERATOR_CLASS;

const char *AccClassString(ACCELERATOR_CLASS cls);

typedef struct ACCELERATOR_DESCRIPTOR_t {
    UINT uiAcceleratorId;
    ACCELERATOR_CLASS accClass;
    BOOL bEnabled;
    UINT nCoreCount;
    UINT nRuntimeVersion;
    UINT nMemorySize;
    UINT nClockRate;
    UINT nPlatformIndex;
    BOOL bSupportsConcurrentKernels;
    char szDescription[256];
    void *pAccelerator;
} ACCELERATOR_DESCRIPTOR, *PACCELERATOR_DESCRIPTOR;

typedef unsigned int BUFFERACCESSFLAGS;
static const unsigned int PT_ACCESS_DEFAULT = 0x0;
static const unsigned int PT_ACCESS_HOST_READ = 0x1;
static const unsigned int PT_ACCESS_HOST_WRITE = 0x2;
static const unsigned int PT_ACCESS_ACCELERATOR_READ = 0x4;
static const unsigned int PT_ACCESS_ACCELERATOR_WRITE = 0x8;
static const unsigned int PT_ACCESS_IMMUTABLE = 0x10;
static const unsigned int PT_ACCESS_BYTE_ADDRESSABLE = 0x20;
static const unsigned int PT_ACCESS_SHARED_HINT = 0x40;
static const unsigned int PT_ACCESS_POOLED_HINT = 0x80;

static const int MAX_DATABLOCK_DIMENSIONS = 3;
static const int XDIM = 0;
static const int YDIM = 1;
static const int ZDIM = 2;

typedef enum ptask_parm_t {
    PTPARM_INT = 0,
    PTPARM_FLOAT = 1,
    PTPARM_DOUBLE = 2,
    PTPARM_BYVALSTRUCT = 3,
    PTPARM_BYREFSTRUCT = 4,
    PTPARM_BUFFER = 5,
    PTPARM_NONE = 6
} PTASK_PARM_TYPE;

typedef unsigned long long CONTROLSIGNAL;
const CONTROLSIGNAL DBCTLC_NONE = 0x0;
const CONTROLSIGNAL DBCTLC_BOF = 0x1;
const CONTROLSIGNAL DBCTLC_EOF = 0x2;
const CONTROLSIGNAL DBCTLC_BEGINITERATION = 0x4;
const CONTROLSIGNAL DBCTLC_ENDITERATION = 0x8;
const CONTROLSIGNAL DBCTLC_ERR = 0x10;
const CONTROLSIGNAL DBCTLC_DEVNULL = 0x20;
#define TESTSIGNAL(x, y) (((x) & (y)) != 0)
#define HASSIGNAL(x) ((x) != PTask::DBCTLC_NONE)

typedef enum buffer_coh_state_t {
    BSTATE_NO_ENTRY = 0,
    BSTATE_INVALID = 1,
    BSTATE_SHARED = 2,
    BSTATE_EXCLUSIVE = 3
} BUFFER_COHERENCE_STATE;

const UINT NUM_DATABLOCK_CHANNELS = 3;
const UINT DBDATA_IDX = 0;
const UINT DBMETADATA_IDX = 1;
const UINT DBTEMPLATE_IDX = 2;

typedef enum porttyp_t {
    INPUT_PORT,
    OUTPUT_PORT,
    STICKY_PORT,
    META_PORT,
    INITIALIZER_PORT
} PORTTYPE;

typedef enum descfun_t {
    DF_SIZE,
    DF_EOF,
    DF_BOF,
    DF_METADATA_SPLITTER,
    DF_CONTROL_CODE,
    DF_BLOCK_UID,
    DF_DATA_SOURCE,
    DF_METADATA_SOURCE,
    DF_TEMPLATE_SOURCE,
    DF_DATA_DIMENSIONS
} DESCRIPTORFUNC;

typedef enum metafun_t {
    MF_NONE = 0,
    MF_ALLOCATION_SIZE = 1,
    MF_SIMPLE_ITERATOR = 2,
    MF_GENERAL_ITERATOR = 3,
    MF_PADDED_ALLOCATION_SIZE = 4,
    MF_USER_DEFINED = 5
} METAFUNCTION;

typedef struct __bufferDims_t {
    UINT uiXElements;
    UINT uiYElements;
    UINT uiZElements;
    UINT cbElementStride;
    UINT cbPitch;
    __bufferDims_t() { Initialize(0, 0, 0, 0, 0); }
    explicit __bufferDims_t(UINT uiBytes) { Initialize(uiBytes, 1, 1, 1, uiBytes); }
    __bufferDims_t(UINT x, UINT y, UINT z, UINT stride) { Initialize(x, y, z, stride, stride * x); }
    __bufferDims_t(UINT x, UINT y, UINT z, UINT stride, UINT pitch) { Initialize(x, y, z, stride, pitch); }
    void Initialize(UINT x, UINT y, UINT z, UINT stride, UINT pitch) {
        uiXElements = x; uiYElements = y; uiZElements = z; cbElementStride = stride; cbPitch = pitch;
    }
    UINT AllocationSizeBytes() const { return cbPitch * uiYElements * uiZElements; }
    UINT TotalElements() const { return uiXElements * uiYElements * uiZElements; }
    BOOL HasPaddedPitch() const { return uiXElements * cbElementStride != cbPitch; }
} BUFFERDIMENSIONS;

typedef enum __outstanding_op_type {
    OT_LAUNCH_TARGET_READABLE,
    OT_LAUNCH_TARGET_WRITABLE,
    OT_MEMCPY_TARGET,
    OT_MEMCPY_SOURCE
} ASYNCHRONOUS_OPTYPE;

typedef enum __async_ctxt_type_t {
    ASYNCCTXT_DEFAULT,
    ASYNCCTXT_TASK,
    ASYNCCTXT_XFERHTOD,
    ASYNCCTXT_XFERDTOH,
    ASYNCCTXT_XFERDTOD,
    ASYNCCTXT_UNDEFINED
} ASYNCCONTEXTTYPE;

typedef enum _graph_state_t {
    PTGS_INITIALIZING = 0,
    PTGS_RUNNABLE = 1,
    PTGS_RUNNING = 2,
    PTGS_QUIESCING = 3,
    PTGS_TEARINGDOWN = 4,
    PTGS_TEARDOWNCOMPLETE = 5
} GRAPHSTATE;

typedef enum viewmaterializationpolicy_t {
    VIEWMATERIALIZATIONPOLICY_ON_DEMAND = 0,
    VIEWMATERIALIZATIONPOLICY_EAGER = 1
} VIEWMATERIALIZATIONPOLICY;

typedef struct ptask_dim3_t {
    int x;
    int y;
    int z;
    ptask_dim3_t(int xx = 0, int yy = 0, int zz = 0) : x(xx), y(yy), z(zz) {}
} PTASKDIM3;

typedef struct ptask_usage_stats_t {
    double dLastWaitTime;
    double dAverageWaitTime;
    double dLastDispatchTime;
    double dAverageDispatchTime;
    double dAverageOSPrio;
} PTASKUSAGESTATS, *PPTASKUSAGESTATS;

typedef enum schedmode_t {
    SCHEDMODE_COOPERATIVE = 0,
    SCHEDMODE_PRIORITY = 1,
    SCHEDMODE_DATADRIVEN = 2,
    SCHEDMODE_FIFO = 3
} SCHEDULINGMODE;

typedef enum graphpartitionmode_t {
    GRAPHPARTITIONINGMODE_NONE = 0,
    GRAPHPARTITIONINGMODE_HINTED = 1,
    GRAPHPARTITIONINGMODE_HEURISTIC = 2,
    GRAPHPARTITIONINGMODE_OPTIMAL = 3
} GRAPHPARTITIONINGMODE;

typedef enum threadpoolpolicy_t {
    TPP_AUTOMATIC,
    TPP_EXPLICIT,
    TPP_THREADPERTASK
} THREADPOOLPOLICY;

typedef enum threadrole_t {
    PTTR_UNKNOWN = 0,
    PTTR_GRAPHRUNNER = 1,
    PTTR_SCHEDULER = 2,
    PTTR_GC = 3,
    PTTR_APPLICATION = 4
} PTTHREADROLE;

}  // namespace PTask

#endif /* PTK_PRIMITIVE_TYPES_H */
