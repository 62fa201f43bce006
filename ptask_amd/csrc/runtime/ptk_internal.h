// ptk_internal.h -- runtime-private classes: the Host and HIP backends, block pools and the
// process-wide runtime state.  Not installed; application code sees include/ptask/*.h only.
#pragma once

#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "ptaskapi.h"
#include "ptk_hip.h"

namespace PTask {

// ---------------------------------------------------------------------------------------------
// Block pools (reference: BlockPool.cpp, OutputPort::GetPooledDestinationBuffer
// OutputPort.cpp:603).  A pool keeps blocks whose last reference was released so the next
// dispatch reuses their device storage: no allocation and no migration in steady state.
class BlockPool : public ReferenceCounted {
public:
    BlockPool(DatablockTemplate *pTemplate, UINT uiMaxFree);
    ~BlockPool() override;
    // A free block of exactly cb bytes (bAtLeast: the smallest free block of at least cb bytes,
    // its logical size set to cb), or a new one of cb bytes; returned with one reference.
    // bCreate FALSE: nullptr instead of a new block.
    Datablock *Request(UINT cb, BOOL bAtLeast = FALSE, BOOL bCreate = TRUE);
    // Called by Datablock::OnLastRelease for blocks owned by this pool.
    void Return(Datablock *pBlock);
    // No more recycling: free blocks are deleted now, returning blocks are deleted later.
    void Close();
    size_t FreeCount();
    // Adds n free blocks of cb data bytes (and cbMeta metadata bytes) with host storage and, when
    // pDevice is given, storage on that accelerator; returns how many were added.
    UINT Prefill(UINT n, UINT cb, UINT cbMeta, Accelerator *pDevice);
    // Deletes the free blocks (their host and device storage) for an allocation that failed;
    // returns the bytes freed.  The pool stays open.
    size_t Trim();

private:
    std::mutex m_lock;
    DatablockTemplate *m_pTemplate;
    std::vector<Datablock *> m_vFree;
    UINT m_uiMaxFree;
    BOOL m_bClosed;
};

// ---------------------------------------------------------------------------------------------
// Memory: allocation with recovery and runtime-wide pools (pools.cpp).  A failed allocation
// reclaims the free blocks of every live pool and tries again, then once more after every
// accelerator has finished its work (reference: the forced collection and retry of
// PCUBuffer.cpp:979-1015 with the collector of ptgc.cpp:326).
void *AllocHostStorage(size_t cb);                    // page-aligned, free() releases it
PBuffer *AllocDeviceStorage(Accelerator *pAcc, size_t cb);
size_t ReclaimPooledMemory();                         // bytes freed from every pool's free list
void RegisterPool(BlockPool *pPool);
void UnregisterPool(BlockPool *pPool);
// Runtime-wide pools of Runtime::RequireBlockPool (reference GlobalPoolManager.cpp:120-250):
// typed by template, or untyped by data size; a request takes the template's pool, else the
// smallest untyped pool whose blocks hold cb bytes.
Datablock *RequestGlobalPooledBlock(DatablockTemplate *pTemplate, UINT cb, UINT cbMeta);
void AllocateGlobalPools();
void DestroyGlobalPools();

// ---------------------------------------------------------------------------------------------
// Host backend: memory space 0, host functions called in place (reference: hostaccelerator.cpp,
// hosttask.cpp).  The oracle path of every parity check runs here.

class HostAsyncContext : public AsyncContext {
public:
    explicit HostAsyncContext(Accelerator *pAcc, ASYNCCONTEXTTYPE e) : AsyncContext(pAcc, e) {}
    void *GetPlatformContextObject() override { return nullptr; }
    SyncPoint *CreateSyncPoint() override;
    BOOL InsertFence(SyncPoint *) override { return TRUE; }
    BOOL SynchronizeContext() override { return TRUE; }
    void RecycleEvent(void *) override {}
};

class PHBuffer : public PBuffer {
public:
    PHBuffer(Accelerator *pAcc, size_t cb);
    ~PHBuffer() override;
};

class HostAccelerator : public Accelerator {
public:
    HostAccelerator(UINT uiId);
    ~HostAccelerator() override;
    BOOL Open() override;
    BOOL Compile(const char *lpszFile, const char *lpszOperation, CompiledKernel *pKernel,
                 char *lpszCompilerOutput, int cbCompilerOutput) override;
    BOOL Synchronize() override { return TRUE; }
    BOOL MakeDeviceContextCurrent() override { return TRUE; }
    BOOL SupportsByvalArguments() override { return FALSE; }
    AsyncContext *CreateAsyncContext(ASYNCCONTEXTTYPE e) override;
    AsyncContext *GetTransferContext() override { return m_pXfer; }
    PBuffer *NewPlatformSpecificBuffer(size_t cb) override;
    BOOL CopyHostToDevice(PBuffer *, const void *, size_t, AsyncContext *) override { return FALSE; }
    BOOL CopyDeviceToHost(void *, PBuffer *, size_t) override { return FALSE; }
    BOOL CopyDeviceToDevice(PBuffer *, PBuffer *, size_t, AsyncContext *) override { return FALSE; }

private:
    AsyncContext *m_pXfer;
    std::map<std::string, void *> m_mLibraries;
};

class HIPAccelerator;
class HIPAsyncContext;

class HostTask : public Task {
public:
    HostTask(Graph *pGraph, CompiledKernel *pKernel, const char *lpszName);
    ~HostTask() override;
    void ReleaseDeviceResources() override;

protected:
    BOOL PlatformSpecificDispatch(std::vector<Datablock *> &vIn, std::vector<Datablock *> &vOut,
                                  UINT uiIterations) override;
    // dispatch with a dependent accelerator (hosttask.h forms)
    BOOL DispatchWithDependences(FARPROC fn, std::vector<Datablock *> &vIn, std::vector<Datablock *> &vOut,
                                 UINT uiIterations);
    HIPAccelerator *m_pDepAccelerator = nullptr;
    HIPAsyncContext *m_pDepContext = nullptr;
};

// ---------------------------------------------------------------------------------------------
// HIP backend: one MI355X, memory space 1 (reference roles: cuaccelerator.cpp, cutask.cpp,
// PCUBuffer.cpp, CUAsyncContext.cpp).  Every device call goes through include/ptk_hip.h.
class HIPAsyncContext : public AsyncContext {
public:
    HIPAsyncContext(Accelerator *pAcc, ASYNCCONTEXTTYPE e);
    ~HIPAsyncContext() override;
    void *GetPlatformContextObject() override { return m_hStream; }
    ptk_hip_stream GetStream() { return m_hStream; }
    SyncPoint *CreateSyncPoint() override;
    BOOL InsertFence(SyncPoint *pSP) override;
    BOOL SynchronizeContext() override;
    void RecycleEvent(void *pEvent) override;

private:
    ptk_hip_stream m_hStream;
    std::mutex m_evLock;
    std::vector<ptk_hip_event> m_vFreeEvents;
};

class PHIPBuffer : public PBuffer {
public:
    PHIPBuffer(Accelerator *pAcc, size_t cb);
    ~PHIPBuffer() override;
};

class HIPAccelerator : public Accelerator {
public:
    HIPAccelerator(UINT uiId, int nDevice, UINT uiMemSpace);
    ~HIPAccelerator() override;
    BOOL Open() override;
    BOOL Compile(const char *lpszFile, const char *lpszOperation, CompiledKernel *pKernel,
                 char *lpszCompilerOutput, int cbCompilerOutput) override;
    BOOL Synchronize() override;
    BOOL MakeDeviceContextCurrent() override;
    BOOL SupportsByvalArguments() override { return TRUE; }
    AsyncContext *CreateAsyncContext(ASYNCCONTEXTTYPE e) override;
    AsyncContext *GetTransferContext() override { return m_pXfer; }
    PBuffer *NewPlatformSpecificBuffer(size_t cb) override;
    BOOL CopyHostToDevice(PBuffer *pDst, const void *pSrc, size_t cb, AsyncContext *pCtx) override;
    BOOL CopyDeviceToHost(void *pDst, PBuffer *pSrc, size_t cb) override;
    BOOL CopyDeviceToDevice(PBuffer *pDst, PBuffer *pSrc, size_t cb, AsyncContext *pCtx) override;
    size_t GetMemorySize() override { return m_props.total_global_mem; }
    UINT GetCoreCount() override { return (UINT)m_props.compute_units; }
    int GetDeviceOrdinal() { return m_nDevice; }
    // The one stream every task shares when Runtime::SetTaskStreamPolicy(PTK_STREAMS_SHARED),
    // else NULL (each task then gets its own stream).
    AsyncContext *GetSharedTaskContext();
    const ptk_hip_device_props_t &GetProps() { return m_props; }

private:
    int m_nDevice;
    ptk_hip_device_props_t m_props;
    AsyncContext *m_pXfer;
    std::mutex m_modLock;
    std::map<std::string, ptk_hip_module> m_mModules;
    AsyncContext *m_pShared = nullptr;
};

class HIPTask : public Task {
public:
    HIPTask(Graph *pGraph, CompiledKernel *pKernel, const char *lpszName);
    ~HIPTask() override;
    void ReleaseDeviceResources() override;
    // The stream this task launches on (created on first dispatch).
    HIPAsyncContext *GetContext() { return m_pContext; }
    Accelerator *GetHIPAccelerator() { return (Accelerator *)m_pAccelerator; }

protected:
    BOOL PlatformSpecificDispatch(std::vector<Datablock *> &vIn, std::vector<Datablock *> &vOut,
                                  UINT uiIterations) override;

private:
    HIPAccelerator *m_pAccelerator;
    HIPAsyncContext *m_pContext;
    std::vector<unsigned char> m_vArgs;
    // per-dispatch (start, end) events while dispatch profiling is on; resolved lazily so the
    // profiled dispatches run at full speed
    std::vector<std::pair<ptk_hip_event, ptk_hip_event>> m_vProfEvents;
    void ResolveProfile();

public:
    double GetTotalDispatchDeviceMsResolved();
    double GetTotalDispatchDeviceMs() override { return GetTotalDispatchDeviceMsResolved(); }
};

// ---------------------------------------------------------------------------------------------
// Launch capture of a graph (replay.cpp): a recorded run replayed as one hipGraph launch.
class GraphReplay {
public:
    enum State { RECORDING, ARMED, FAILED };
    explicit GraphReplay(Graph *pGraph);
    ~GraphReplay();
    State GetState() { return m_eState; }
    BOOL IsArmed() { return m_eState == ARMED; }
    UINT GetReplayCount() { return m_uiReplays; }
    UINT GetLaunchesPerReplay() { return m_uiLaunchesPerReplay; }
    // recording hooks: graph input and output channel enqueues (graph lock held), HIP launches
    // (any runner thread), completed dispatches (graph lock held)
    void OnInputPush(Channel *pChannel, Datablock *pBlock);
    void OnOutputEnqueue(Channel *pChannel, Datablock *pBlock);
    struct PtrArg {
        UINT uiOffset;      // byte offset of the pointer in the packed arguments
        Datablock *pBlock;
        BOOL bWrite;        // the launch writes the block
    };
    void OnLaunch(Task *pTask, Accelerator *pAcc, ptk_hip_function fn, const PTASKDIM3 &grid, const PTASKDIM3 &block,
                  const unsigned char *args, UINT cbArgs, const std::vector<PtrArg> &ptrs, UINT uiIterations);
    void OnDispatchDone(Task *pTask);
    // graph lock held: one replay when every data input holds a block; TRUE when it ran
    BOOL TryReplay();
    // the stream's work is finished and every resource returned
    void Release();

private:
    struct Launch {
        Task *pTask;
        ptk_hip_function fn;
        PTASKDIM3 grid, block;
        std::vector<unsigned char> args;
        std::vector<PtrArg> ptrs;
        UINT uiIterations;
    };
    struct Staging {
        void *dptr = nullptr;
        size_t bytes = 0, logical = 0;
        bool in = false, out = false;
    };
    void Fail(const char *why);
    void Seal();

public:
    // A runtime-side copy of a block for an in/out consumer (OutputPort::Push) is not a launch,
    // so a run that makes one cannot be replayed from its recorded launches.
    void OnBlockCopied() {
        if (m_eState == RECORDING) Fail("a run copied a block for an in/out consumer");
    }

private:
    bool CaptureLaunches();

    Graph *m_pGraph;
    State m_eState = RECORDING;
    std::mutex m_recLock;
    Accelerator *m_pAccelerator = nullptr;
    HIPAsyncContext *m_pContext = nullptr;
    std::vector<Launch> m_vLaunches;
    std::map<Channel *, Datablock *> m_mInputs, m_mOutputs;  // recorded blocks (keys once sealed)
    bool m_bKeysOnly = false;
    std::set<Task *> m_sDispatched;
    std::map<Datablock *, Staging> m_mStaging;
    std::vector<Datablock *> m_vPinned;
    ptk_hip_graph m_hGraph = nullptr;
    UINT m_uiReplays = 0, m_uiLaunchesPerReplay = 0;
};

// ---------------------------------------------------------------------------------------------
// Admission to one accelerator class's dispatch path, shared by every live graph (reference: the
// global ready queue and accelerator assignment of Scheduler.cpp:4551-4673, quiescence
// :526-700).  At most `slots` dispatches run at once; waiting dispatches are admitted best first:
// the highest task priority when the scheduling mode is SCHEDMODE_PRIORITY, else arrival order.
// While quiesced, none is admitted; Quiesce(TRUE) returns once the admitted ones have finished.
class DispatchArbiter {
public:
    explicit DispatchArbiter(int nSlots) : m_nSlots(nSlots) {}
    void Acquire(int nPriority);
    void Release();
    void SetSlots(int n);
    int GetSlots();
    void Quiesce(BOOL bQuiesce);
    int GetWaiting();
    unsigned long long GetAdmitted();

private:
    struct Ticket {
        int prio;
        unsigned long long seq;
    };
    std::mutex m_lock;
    std::condition_variable m_cv;
    int m_nSlots, m_nInUse = 0;
    BOOL m_bQuiesced = FALSE;
    unsigned long long m_uiSeq = 0, m_uiAdmitted = 0;
    std::vector<Ticket> m_vWaiting;
};

// ---------------------------------------------------------------------------------------------
// Process-wide state.
struct RuntimeState {
    std::recursive_mutex lock;
    bool initialized = false;
    BOOL useHost = TRUE, useHIP = TRUE;
    int maxConcurrency = 0;         // HIP devices to use (0 = all)
    UINT physicalAccelerators = 0;  // devices per class the runtime may open (0 = all found)
    int maxHostConcurrency = 4;     // host tasks running at once
    int channelCapacity = 4;
    UINT threadPoolSize = 0;        // graph runner threads (0 = automatic)
    int schedulingMode = SCHEDMODE_PRIORITY;
    BOOL forceSynchronous = FALSE;
    BOOL profilePSDispatch = FALSE;
    BOOL exitOnFailure = TRUE;
    UINT loggingLevel = 0;
    UINT blockPoolSize = 8;
    int streamPolicy = 0;           // PTK_STREAMS_PER_TASK or PTK_STREAMS_SHARED
    VIEWMATERIALIZATIONPOLICY outputViewPolicy = VIEWMATERIALIZATIONPOLICY_ON_DEMAND;
    std::vector<Accelerator *> hostAccelerators;
    std::vector<Accelerator *> hipAccelerators;
    std::map<std::string, DatablockTemplate *> templates;
    std::vector<CompiledKernel *> kernels;
    std::vector<Port *> ports;
    std::string lastError;
    std::atomic<unsigned long long> h2dCount{0}, h2dBytes{0}, d2hCount{0}, d2hBytes{0};
    std::atomic<UINT> nextDBUID{1};
    // dispatch admission per accelerator class: host tasks (maxHostConcurrency slots) and HIP
    // tasks (enqueueing a launch; the device runs launched work concurrently)
    DispatchArbiter hostArbiter{4}, hipArbiter{8};
    // graph monitor watchdog
    std::atomic<BOOL> useWatchdog{FALSE};
    std::atomic<DWORD> watchdogThresholdMs{10000};
    std::atomic<unsigned long long> watchdogReports{0};
    // allocation recovery and pools (pools.cpp)
    std::atomic<unsigned long long> allocFailures{0}, allocRecovered{0}, allocUnrecovered{0}, bytesReclaimed{0};
    std::atomic<unsigned long long> poolHits{0}, poolMisses{0};
    std::atomic<int> injectedFailures[PTK_MAX_MEMORY_SPACES] = {};
    // templates used for descriptor blocks
    DatablockTemplate *sizeDescriptorTemplate = nullptr;
    DatablockTemplate *controlTemplate = nullptr;
};

RuntimeState &RT();
Accelerator *HostAcceleratorInstance();
HIPAccelerator *DefaultHIPAccelerator();
Accelerator *AcceleratorForMemorySpace(UINT uiMemSpace);
// Fails loudly: HandleError with the HIP C ABI's last error text.
BOOL CheckHip(int status, const char *what);

namespace Runtime {
// Adds a new template to the runtime's table under its name and returns it, or deletes it and
// returns the template already registered under that name.
DatablockTemplate *RegisterTemplate(DatablockTemplate *t);
}  // namespace Runtime

}  // namespace PTask
