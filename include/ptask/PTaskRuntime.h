/* PTaskRuntime.h -- runtime-wide entry points: initialisation, factories and configuration
 * (reference: ptask/PTaskRuntime.h).  Signatures follow the reference so application code
 * compiles unchanged; configuration setters take effect before Initialize() unless noted. */
#ifndef PTK_PTASKRUNTIME_H
#define PTK_PTASKRUNTIME_H

#include <string>

#include "primitive_types.h"

namespace PTask {

class Accelerator;
class Channel;
class CompiledKernel;
class Datablock;
class DatablockTemplate;
class Graph;
class Port;
class Task;

namespace Runtime {

PTRESULT Initialize();
BOOL IsInitialized();
PTRESULT Terminate();

/* ---- templates (reference PTaskRuntime.h:392-429) ---------------------------------------- */
DatablockTemplate *GetDatablockTemplate(char *lpszType, unsigned int uiStride, unsigned int x,
                                        unsigned int y, unsigned int z, bool bRecordStream = false,
                                        bool bRaw = false);
DatablockTemplate *GetDatablockTemplate(char *lpszTemplateName, unsigned int uiElementStride,
                                        unsigned int uiElementsX, unsigned int uiElementsY,
                                        unsigned int uiElementsZ, unsigned int uiPitch,
                                        bool bIsRecordStream, bool bIsByteAddressable);
DatablockTemplate *GetDatablockTemplate(char *lpszTemplateName, BUFFERDIMENSIONS *pBufferDims,
                                        unsigned int uiNumBufferDims, bool bIsRecordStream,
                                        bool bIsByteAddressable);
DatablockTemplate *GetDatablockTemplate(char *lpszType, unsigned int uiStride,
                                        PTASK_PARM_TYPE pttype);
DatablockTemplate *GetDatablockTemplate(char *lpszType);

/* ---- datablocks (reference PTaskRuntime.h:431-464) --------------------------------------- */
Datablock *AllocateDatablock(DatablockTemplate *pTemplate, void *pInitData, UINT cbInitData,
                             Channel *pDestChannel, BUFFERACCESSFLAGS flags = 0,
                             CONTROLSIGNAL uiBlockControlCode = 0);
Datablock *AllocateDatablock(DatablockTemplate *pTemplate, UINT uiDataSize, UINT uiMetaSize,
                             UINT uiTemplateSize, CONTROLSIGNAL uiBlockControlCode = 0);
Datablock *AllocateDatablock(CONTROLSIGNAL uiBlockControlCode);
void FreeDatablock(Datablock *pBlock);

/* ---- kernels and ports (reference PTaskRuntime.h:205-237) -------------------------------- */
/* The accelerator class follows the file extension: ".hsaco"/".co" -> HIP (a gfx950 code
 * object), ".so"/".dll" -> host (a shared library of host-task functions). */
CompiledKernel *GetCompiledKernel(char *lpszFile, char *lpszOperation,
                                  char *lpszCompilerOutput = NULL, int uiCompilerOutput = 0,
                                  int tgx = 1, int tgy = 1, int tgz = 1);
CompiledKernel *GetHostFunctionCompiledKernel(char *lpszOperation, FARPROC lpfn);
Port *CreatePort(PORTTYPE type, DatablockTemplate *pTemplate, UINT uiId,
                 char *lpszVariableBinding = NULL, UINT nKernelParameterIndex = (UINT)-1,
                 UINT nInOutParmOutputIndex = (UINT)-1);
PTRESULT SetTaskAffinity(Task *pTask, Accelerator *pAccelerator, AFFINITYTYPE eAffinityType);

/* ---- block pools (reference PTaskRuntime.h:111-112, GlobalPoolManager.h) ------------------
 * Runtime-wide pools of pre-allocated blocks (host storage, and device storage on the first HIP
 * accelerator): typed by template, or untyped by data size.  AllocateDatablock takes a free block
 * from the template's pool, else from the smallest untyped pool that holds the size; released
 * blocks go back.  Called before Initialize, the pools are allocated by Initialize; after, at
 * once.  nBlocks 0: the default block pool size. */
void RequireBlockPool(int nDataSize, int nMetaSize, int nTemplateSize, int nBlocks = 0);
void RequireBlockPool(DatablockTemplate *pTemplate, int nBlocks);

/* ---- memory statistics and failure injection (no counterpart in the reference) -----------
 * A host or device allocation that fails makes the runtime free the free blocks of every pool
 * and try again, then once more after every accelerator has finished its work.  These count
 * what happened; SetAllocationFailureInjection makes the next nFailures allocation attempts in a
 * memory space (0 host, 1 the HIP device) fail, for tests of that path. */
typedef struct _ptk_memory_statistics {
    unsigned long long ullAllocationFailures;      /* allocations whose first attempt failed */
    unsigned long long ullAllocationsRecovered;    /* of those, allocations a retry satisfied */
    unsigned long long ullAllocationsUnrecovered;  /* of those, allocations that failed for good */
    unsigned long long ullBytesReclaimed;          /* storage freed from pools for retries */
    unsigned long long ullPooledBlockHits;         /* pooled blocks handed out by AllocateDatablock / RequestPooledBlock */
    unsigned long long ullPooledBlockMisses;       /* requests a pool could not serve from its free blocks */
} PTK_MEMORY_STATISTICS;
void GetMemoryStatistics(PTK_MEMORY_STATISTICS *pStatistics);
void ResetMemoryStatistics();
void SetAllocationFailureInjection(UINT uiMemorySpaceId, int nFailures);
/* By accelerator id (ACCELERATOR_DESCRIPTOR::uiAcceleratorId) among the accelerators of the task's
 * class; PTASK_ERR_INVALID_PARAMETER for an id of no such accelerator (reference
 * PTaskRuntime.h:472-502, both forms). */
PTRESULT SetTaskAffinity(Task *pTask, UINT uiAcceleratorId, AFFINITYTYPE eAffinityType);
PTRESULT SetTaskAffinity(Task *pTask, UINT *pvAcceleratorIds, AFFINITYTYPE *pvAffinityTypes, UINT nAcceleratorIds);

/* ---- accelerators --------------------------------------------------------------------------- */
/* The uiEnumerationIndex-th enabled accelerator of a class (reference:
 * Runtime::EnumerateAccelerators); the descriptor is malloc'd and owned by the caller (free). */
PTRESULT EnumerateAccelerators(ACCELERATOR_CLASS eClass, UINT uiEnumerationIndex,
                               ACCELERATOR_DESCRIPTOR **ppDescriptor);
Accelerator *GetAccelerator(ACCELERATOR_CLASS eClass, UINT uiIndex);
UINT GetAcceleratorCount(ACCELERATOR_CLASS eClass);

/* ---- configuration ------------------------------------------------------------------------ */
void SetUseHost(BOOL b);
void SetUseHIP(BOOL b);
void SetUseCUDA(BOOL b);      // accepted for source compatibility; there is no CUDA backend
void SetUseOpenCL(BOOL b);    // accepted for source compatibility; there is no OpenCL backend
void SetUseDirectX(BOOL b);   // accepted for source compatibility; there is no DirectX backend
BOOL GetUseHost();
BOOL GetUseHIP();
void SetMaximumConcurrency(int nGPUs);
int GetMaximumConcurrency();
void SetMaximumHostConcurrency(int nHostAccelerators);
int GetMaximumHostConcurrency();
void SetDefaultChannelCapacity(int n);
int GetDefaultChannelCapacity();
void SetTaskThreadPoolSize(UINT uiThreadPoolSize);
UINT GetTaskThreadPoolSize();
void SetTaskThreadPoolPolicy(THREADPOOLPOLICY ePolicy);
void SetSchedulingMode(int mode);
int GetSchedulingMode();
/* Dispatch admission across all live graphs.  Each accelerator class (host, HIP) admits at most
 * a number of dispatches at once (host: SetMaximumHostConcurrency; HIP: the launches being
 * enqueued, 8 by default) and admits waiting ones best first: highest task priority
 * (Task::SetPriority) in SCHEDMODE_PRIORITY, arrival order in the other modes (reference: the
 * global ready queue of Scheduler.cpp:4551-4673).  GlobalQuiesce admits no more dispatches and
 * returns when the admitted ones have finished; EndGlobalQuiescence resumes (reference:
 * Scheduler::GlobalQuiesce / EndGlobalQuiescence, Scheduler.cpp:526-600, internal there).
 * GetWaitingDispatchCount: dispatches waiting for admission. */
void SetMaximumDeviceDispatchConcurrency(int n);
int GetMaximumDeviceDispatchConcurrency();
void GlobalQuiesce();
void EndGlobalQuiescence();
int GetWaitingDispatchCount();
void SetForceSynchronous(BOOL b);
BOOL GetForceSynchronous();
void SetProfilePSDispatch(BOOL b);
BOOL GetProfilePSDispatch();
void SetDebugMode(BOOL b);
/* The graph monitor watchdog (reference PTaskRuntime.h:1476,1599-1619; default threshold
 * 10000 ms as in PTaskRuntime.cpp:99): graphs started while it is on get a monitor thread that
 * reports a stall (Graph::ReportGraphState to stderr).  GetWatchdogReportCount counts the
 * reports of every graph (no reference counterpart). */
void SetUseGraphMonitorWatchdog(BOOL b);
BOOL GetUseGraphMonitorWatchdog();
void SetDispatchWatchdogThreshold(DWORD dwThreshold);
DWORD GetDispatchWatchdogThreshold();
unsigned long long GetWatchdogReportCount();
void SetDefaultViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY e);
void SetDefaultOutputViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY e);
VIEWMATERIALIZATIONPOLICY GetDefaultOutputViewMaterializationPolicy();
void SetLoggingLevel(UINT nLevel);
UINT GetLoggingLevel();
void SetExitOnRuntimeFailure(BOOL bExit);
BOOL GetExitOnRuntimeFailure();
/* Stream assignment of HIP tasks (no reference counterpart; the CUDA backend gives every task
 * its own context, cutask.cpp:667-675).  PER_TASK lets independent tasks overlap on the device;
 * SHARED puts every task on one in-order stream, which removes the cross-stream event waits of
 * chains of dependent tasks.  Takes effect for tasks that have not dispatched yet. */
#define PTK_STREAMS_PER_TASK 0
#define PTK_STREAMS_SHARED 1
void SetTaskStreamPolicy(int ePolicy);
int GetTaskStreamPolicy();
/* Blocks kept per output-port pool before recycled blocks are freed instead. */
void SetDefaultBlockPoolSize(UINT n);
UINT GetDefaultBlockPoolSize();

/* ---- graph helpers (reference PTaskRuntime.h:2441-2560): forwarders to the Graph, Port and
 * Channel calls, for applications written against the Runtime:: form ---------------------- */
void SetPortMetaFunction(Port *pPort, METAFUNCTION eCanonicalMetaFunction);
Task *AddTask(Graph *pGraph, CompiledKernel *pKernel, UINT uiInputPortCount, Port **pvInputPorts,
              UINT uiOutputPorts, Port **pvOutputPorts, char *lpszTaskName);
void BindDerivedPort(Graph *pGraph, Port *pDescribedPort, Port *pDescriberPort, DESCRIPTORFUNC func = DF_SIZE);
void RunGraph(Graph *pGraph);
void Push(Channel *pChannel, Datablock *pBlock);
Datablock *Pull(Channel *pChannel, DWORD dwTimeout = INFINITE);
PTRESULT CheckGraphSemantics(Graph *pGraph, BOOL bVerbose = TRUE, BOOL bFailOnWarning = FALSE);
/* Accelerators the runtime may use per class: the devices enumerated at Initialize unless set
 * (before Initialize) to fewer (reference: g_nPhysicalAccelerators). */
UINT GetPhysicalAcceleratorCount();
void SetPhysicalAcceleratorCount(UINT nAcceleratorCount);
/* TRUE when bCondition holds; otherwise a warning with the text and FALSE (the reference also
 * asserts in debug builds). */
BOOL CheckInvariantCondition(BOOL bCondition, char *lpszErrorText = NULL);
BOOL CheckInvariantCondition(BOOL bCondition, std::string strErrorText);

/* ---- diagnostics -------------------------------------------------------------------------- */
/* Prints the message; then exits the process unless SetExitOnRuntimeFailure(FALSE), in which
 * case the error is recorded (GetLastError) and FALSE returned (reference:
 * PTaskRuntime.cpp:3884-3899). */
BOOL HandleError(const char *szMessage, ...);
BOOL HandleError(const std::string &strError);
const char *GetLastErrorText();
void ClearLastError();
void Inform(const char *szMessage, ...);
void Warning(const char *szMessage, UINT nLogLevel = 0xFFFFFFFF);
BOOL MandatoryInform(const char *szMessage, ...);
void Trace(const char *szMessage, ...);

/* Host<->device traffic since Initialize (or the last reset): counts and bytes of block
 * populations in each direction.  Used to show that device-resident graphs move nothing. */
void GetTransferStatistics(unsigned long long *pHtoDCount, unsigned long long *pHtoDBytes,
                           unsigned long long *pDtoHCount, unsigned long long *pDtoHBytes);
void ResetTransferStatistics();

}  // namespace Runtime
}  // namespace PTask

#endif
