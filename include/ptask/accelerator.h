/* accelerator.h -- the backend interface: devices, memory, asynchronous contexts and buffers.
 *
 * The pure virtuals mirror the reference's backend contract (ptask/accelerator.h:293-715,
 * ptask/AsyncContext.h:204-604, ptask/PBuffer.h:765-1045) at the granularity this runtime
 * needs.  Two implementations exist: HostAccelerator (host functions, the oracle path) and
 * HIPAccelerator (gfx950 code objects through the C ABI of include/ptk_hip.h). */
#ifndef PTK_ACCELERATOR_H
#define PTK_ACCELERATOR_H

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class AsyncContext;
class CompiledKernel;
class Task;

/* A point in an asynchronous context's work (a recorded HIP event, or "already complete" on
 * the host).  Reference counted: the last holder returns the event to its context's pool. */
class SyncPoint : public ReferenceCounted {
public:
    SyncPoint(AsyncContext *pCtx, void *pEvent) : m_pContext(pCtx), m_pEvent(pEvent) {}
    AsyncContext *GetContext() { return m_pContext; }
    void *GetPlatformEvent() { return m_pEvent; }
    BOOL QueryOutstanding();  // TRUE while the work before it is still pending
    void Wait();              // blocks the calling thread until the work is done

protected:
    void OnLastRelease() override;

private:
    AsyncContext *m_pContext;
    void *m_pEvent;
};

/* An in-order queue of device work: one per task per accelerator, plus transfer contexts. */
class AsyncContext : public ReferenceCounted, public Lockable {
public:
    AsyncContext(class Accelerator *pAcc, ASYNCCONTEXTTYPE eType) : m_pAccelerator(pAcc), m_eType(eType) {}
    virtual ~AsyncContext() {}
    class Accelerator *GetDeviceContext() { return m_pAccelerator; }
    ASYNCCONTEXTTYPE GetAsyncContextType() { return m_eType; }
    virtual void *GetPlatformContextObject() = 0;
    /* Records a sync point after everything enqueued so far (returns it with one reference). */
    virtual SyncPoint *CreateSyncPoint() = 0;
    /* Makes later work in this context wait for `pSP` (no host blocking). */
    virtual BOOL InsertFence(SyncPoint *pSP) = 0;
    virtual BOOL SynchronizeContext() = 0;
    virtual void RecycleEvent(void *pEvent) = 0;

protected:
    class Accelerator *m_pAccelerator;
    ASYNCCONTEXTTYPE m_eType;
};

/* Storage for one datablock channel in one memory space, with the asynchronous-dependence
 * frontier of that storage: the last write and every read issued since it. */
class PBuffer {
public:
    PBuffer(class Accelerator *pAcc, size_t cb) : m_pAccelerator(pAcc), m_cb(cb), m_pBuffer(nullptr), m_pLastWrite(nullptr) {}
    virtual ~PBuffer();
    void *GetBuffer() { return m_pBuffer; }
    size_t GetSize() { return m_cb; }
    class Accelerator *GetAccelerator() { return m_pAccelerator; }
    /* Orders `pCtx` after the frontier: after the last write for a read, after the last write
     * and all reads for a write. */
    void WaitOutstanding(AsyncContext *pCtx, BOOL bForWrite);
    /* Records an access completed at `pSP`. */
    void AddOutstanding(SyncPoint *pSP, BOOL bWrite);
    /* Blocks the calling thread until every outstanding access is complete. */
    void WaitAllOnHost();

protected:
    class Accelerator *m_pAccelerator;
    size_t m_cb;
    void *m_pBuffer;
    std::mutex m_frontierLock;
    SyncPoint *m_pLastWrite;
    std::vector<SyncPoint *> m_vReads;
};

class Accelerator : public ReferenceCounted, public Lockable {
public:
    Accelerator(ACCELERATOR_CLASS eClass, UINT uiId, UINT uiMemSpace)
        : m_eClass(eClass), m_uiAcceleratorId(uiId), m_uiMemorySpaceId(uiMemSpace), m_bEnabled(TRUE) {}
    virtual ~Accelerator() {}

    ACCELERATOR_CLASS GetClass() { return m_eClass; }
    UINT GetAcceleratorId() { return m_uiAcceleratorId; }
    UINT GetMemorySpaceId() { return m_uiMemorySpaceId; }
    const char *GetDeviceName() { return m_strName.c_str(); }
    BOOL IsEnabled() { return m_bEnabled; }
    void SetEnabled(BOOL b) { m_bEnabled = b; }
    BOOL IsHost() { return m_eClass == ACCELERATOR_CLASS_HOST; }
    virtual void GetDescriptor(ACCELERATOR_DESCRIPTOR *pDesc);

    virtual BOOL Open() = 0;
    /* Resolves `lpszOperation` in `lpszFile` for this accelerator (stores it in pKernel). */
    virtual BOOL Compile(const char *lpszFile, const char *lpszOperation, CompiledKernel *pKernel,
                         char *lpszCompilerOutput, int cbCompilerOutput) = 0;
    virtual BOOL Synchronize() = 0;
    virtual BOOL MakeDeviceContextCurrent() = 0;
    virtual BOOL SupportsByvalArguments() = 0;
    virtual AsyncContext *CreateAsyncContext(ASYNCCONTEXTTYPE eType) = 0;
    /* Default context for transfers that belong to no task (host reads of device data). */
    virtual AsyncContext *GetTransferContext() = 0;
    virtual PBuffer *NewPlatformSpecificBuffer(size_t cb) = 0;
    virtual BOOL CopyHostToDevice(PBuffer *pDst, const void *pSrc, size_t cb, AsyncContext *pCtx) = 0;
    virtual BOOL CopyDeviceToHost(void *pDst, PBuffer *pSrc, size_t cb) = 0;
    virtual BOOL CopyDeviceToDevice(PBuffer *pDst, PBuffer *pSrc, size_t cb, AsyncContext *pCtx) = 0;
    virtual size_t GetMemorySize() { return 0; }
    virtual UINT GetCoreCount() { return 0; }  // compute units of a device (0: not known)

protected:
    ACCELERATOR_CLASS m_eClass;
    UINT m_uiAcceleratorId;
    UINT m_uiMemorySpaceId;
    BOOL m_bEnabled;
    std::string m_strName;
};

}  // namespace PTask

#endif
