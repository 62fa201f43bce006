// backends.cpp -- the Host backend (oracle path) and the HIP backend (MI355X path).
//
// Host (reference roles: hostaccelerator.cpp, hosttask.cpp, PHBuffer.cpp, HostAsyncContext.cpp):
// host functions with the void f(UINT, void**) ABI, loaded with dlopen/dlsym or registered as
// function pointers, called tgx*tgy*tgz times per dispatch on host views of the blocks.
//
// HIP (reference roles: cuaccelerator.cpp, cutask.cpp, PCUBuffer.cpp, CUAsyncContext.cpp):
// gfx950 code objects loaded with ptk_hip_module_load, one stream per task, device-resident
// buffers, one reusable event per dispatch ordering every buffer the dispatch touched.

#include <dlfcn.h>

#include <algorithm>
#include <cstring>

#include "ptk_internal.h"

namespace PTask {

void Accelerator::GetDescriptor(ACCELERATOR_DESCRIPTOR *d) {
    memset(d, 0, sizeof(*d));
    d->uiAcceleratorId = m_uiAcceleratorId;
    d->accClass = m_eClass;
    d->bEnabled = m_bEnabled;
    d->nMemorySize = (UINT)(GetMemorySize() >> 20);
    d->nCoreCount = GetCoreCount();
    snprintf(d->szDescription, sizeof(d->szDescription), "%s", m_strName.c_str());
    d->pAccelerator = this;
}

// =============================================================================================
// Host backend
// =============================================================================================

SyncPoint *HostAsyncContext::CreateSyncPoint() {
    AddRef();
    SyncPoint *sp = new SyncPoint(this, nullptr);
    sp->AddRef();
    return sp;
}

PHBuffer::PHBuffer(Accelerator *pAcc, size_t cb) : PBuffer(pAcc, cb) { m_pBuffer = calloc(1, cb ? cb : 1); }
PHBuffer::~PHBuffer() { free(m_pBuffer); }

HostAccelerator::HostAccelerator(UINT uiId)
    : Accelerator(ACCELERATOR_CLASS_HOST, uiId, HOST_MEMORY_SPACE_ID), m_pXfer(nullptr) {
    m_strName = "host";
}

HostAccelerator::~HostAccelerator() {
    if (m_pXfer) m_pXfer->Release();
    // code in loaded libraries may still be referenced by function pointers the application
    // holds, so libraries stay mapped for the life of the process
}

BOOL HostAccelerator::Open() {
    m_pXfer = CreateAsyncContext(ASYNCCTXT_DEFAULT);
    return TRUE;
}

AsyncContext *HostAccelerator::CreateAsyncContext(ASYNCCONTEXTTYPE e) {
    AsyncContext *c = new HostAsyncContext(this, e);
    c->AddRef();
    return c;
}

PBuffer *HostAccelerator::NewPlatformSpecificBuffer(size_t cb) { return new PHBuffer(this, cb); }

// reference: HostAccelerator::Compile, hostaccelerator.cpp:327-410 (LoadLibraryExA +
// GetProcAddress) -- here dlopen + dlsym.
BOOL HostAccelerator::Compile(const char *lpszFile, const char *lpszOperation,
                              CompiledKernel *pKernel, char *lpszOut, int cbOut) {
    Lock();
    void *h = nullptr;
    auto it = m_mLibraries.find(lpszFile);
    if (it != m_mLibraries.end()) {
        h = it->second;
    } else {
        h = dlopen(lpszFile, RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            const char *e = dlerror();
            if (lpszOut && cbOut > 0) snprintf(lpszOut, (size_t)cbOut, "%s", e ? e : "dlopen failed");
            Unlock();
            Runtime::Warning((std::string("host library not loadable: ") + (e ? e : lpszFile)).c_str());
            return FALSE;
        }
        m_mLibraries[lpszFile] = h;
    }
    void *fn = dlsym(h, lpszOperation);
    Unlock();
    if (!fn) {
        if (lpszOut && cbOut > 0) snprintf(lpszOut, (size_t)cbOut, "symbol %s not found", lpszOperation);
        return FALSE;
    }
    pKernel->SetPlatformSpecificBinary(this, fn, h);
    return TRUE;
}

HostTask::HostTask(Graph *pGraph, CompiledKernel *pKernel, const char *lpszName)
    : Task(pGraph, pKernel, ACCELERATOR_CLASS_HOST, lpszName) {}

HostTask::~HostTask() { ReleaseDeviceResources(); }

void HostTask::ReleaseDeviceResources() {
    if (m_pDepContext) {
        m_pDepAccelerator->MakeDeviceContextCurrent();
        m_pDepContext->SynchronizeContext();
        m_pDepContext->Release();
        m_pDepContext = nullptr;
    }
}

// reference: HostTask::PlatformSpecificDispatchNoDependences, hosttask.cpp:305-343 -- buffers
// bind as host pointers, INT/FLOAT values are bit-copied into the argument slot, by-value
// structs are passed by pointer, and the function runs once per point of the geometry.
BOOL HostTask::PlatformSpecificDispatch(std::vector<Datablock *> &vIn, std::vector<Datablock *> &vOut,
                                        UINT uiIterations) {
    Accelerator *host = HostAcceleratorInstance();
    void *binary = m_pKernel->GetPlatformSpecificBinary(host);
    if (!binary) return Runtime::HandleError("%s: host function not resolved", GetTaskName());
    if (HasDependentAcceleratorBindings()) return DispatchWithDependences((FARPROC)binary, vIn, vOut, uiIterations);
    LPFNHOSTTASK fn = (LPFNHOSTTASK)binary;
    std::vector<void *> args(m_vBindings.size(), nullptr);
    for (size_t a = 0; a < m_vBindings.size(); ++a) {
        Binding &b = m_vBindings[a];
        Datablock *blk = b.bInput ? vIn[b.uiSlot] : vOut[b.uiSlot];
        if (!blk) return Runtime::HandleError("%s: no block for argument %zu", GetTaskName(), a);
        void *p = blk->AcquireHostView(b.bWritable, b.bInput ? TRUE : FALSE);
        switch (b.eType) {
            case PTPARM_INT: {
                int v = 0;
                memcpy(&v, p, sizeof(int));
                args[a] = (void *)(intptr_t)v;
                break;
            }
            case PTPARM_FLOAT: {
                float f = 0.f;
                memcpy(&f, p, sizeof(float));
                uintptr_t bits = 0;
                memcpy(&bits, &f, sizeof(float));
                args[a] = (void *)bits;
                break;
            }
            default: args[a] = p; break;
        }
    }
    // the host arbiter has admitted this dispatch (MaximumHostConcurrency slots)
    const long long reps = (long long)std::max(1, m_tgx) * std::max(1, m_tgy) * std::max(1, m_tgz);
    for (UINT it = 0; it < uiIterations; ++it)
        for (long long r = 0; r < reps; ++r) fn((UINT)args.size(), args.data());
    return TRUE;
}

// reference: HostTask::PlatformSpecificDispatchWithDependences(Ex), hosttask.cpp:357-563.  Arguments
// of ports bound to the dependent accelerator are that device's pointers (the block is made valid
// there, in order after its producers on the task's dependent stream); the others bind as in a
// plain host task.  EX form (platform objects requested): the function gets the stream and the
// runtime orders later readers after the work it enqueued there.  Legacy form: the stream is
// drained before the call and the device after it, so the function may use the device freely.
BOOL HostTask::DispatchWithDependences(FARPROC binary, std::vector<Datablock *> &vIn,
                                       std::vector<Datablock *> &vOut, UINT uiIterations) {
    if (!m_pDepAccelerator) {
        m_pDepAccelerator = DefaultHIPAccelerator();
        if (!m_pDepAccelerator) return Runtime::HandleError("%s: no HIP accelerator for the dependent binding", GetTaskName());
    }
    m_pDepAccelerator->MakeDeviceContextCurrent();
    if (!m_pDepContext) {
        m_pDepContext = static_cast<HIPAsyncContext *>(m_pDepAccelerator->CreateAsyncContext(ASYNCCTXT_TASK));
        if (!m_pDepContext) return Runtime::HandleError("%s: dependent stream creation failed", GetTaskName());
    }
    const size_t nArgs = m_vBindings.size();
    if (nArgs > (size_t)MAXARGS) return Runtime::HandleError("%s: %zu arguments, at most %d", GetTaskName(), nArgs, MAXARGS);
    std::vector<void *> args(nArgs, nullptr), devBindings(nArgs, nullptr);
    std::vector<BOOL> isDep(nArgs, FALSE);
    std::vector<Datablock *> blocks(nArgs, nullptr);
    struct Touch { PBuffer *buf; BOOL write; };
    std::vector<Touch> touched;
    void *device = (void *)(intptr_t)m_pDepAccelerator->GetDeviceOrdinal();
    for (size_t a = 0; a < nArgs; ++a) {
        Binding &b = m_vBindings[a];
        Datablock *blk = b.bInput ? vIn[b.uiSlot] : vOut[b.uiSlot];
        if (!blk) return Runtime::HandleError("%s: no block for argument %zu", GetTaskName(), a);
        blocks[a] = blk;
        const bool buffer = b.eType != PTPARM_INT && b.eType != PTPARM_FLOAT && b.eType != PTPARM_BYVALSTRUCT;
        if (buffer && b.pPort->HasDependentAcceleratorBinding()) {
            void *dptr = blk->AcquireDeviceView(m_pDepAccelerator, m_pDepContext, b.bWritable, b.bInput ? TRUE : FALSE);
            if (!dptr) return Runtime::HandleError("%s: no device view for %s", GetTaskName(), b.pPort->GetVariableBinding());
            args[a] = dptr;
            isDep[a] = TRUE;
            devBindings[a] = device;
            touched.push_back({blk->GetPlatformBuffer(m_pDepAccelerator->GetMemorySpaceId()), b.bWritable});
            continue;
        }
        void *p = blk->AcquireHostView(b.bWritable, b.bInput ? TRUE : FALSE);
        switch (b.eType) {
            case PTPARM_INT: {
                int v = 0;
                memcpy(&v, p, sizeof(int));
                args[a] = (void *)(intptr_t)v;
                break;
            }
            case PTPARM_FLOAT: {
                float f = 0.f;
                memcpy(&f, p, sizeof(float));
                uintptr_t bits = 0;
                memcpy(&bits, &f, sizeof(float));
                args[a] = (void *)bits;
                break;
            }
            default: args[a] = p; break;
        }
    }
    for (Touch &t : touched) t.buf->WaitOutstanding(m_pDepContext, t.write);
    void *deps[1] = {device};
    void *streams[1] = {m_bRequestDependentPSObjects ? (void *)m_pDepContext->GetStream() : nullptr};
    const long long reps = (long long)std::max(1, m_tgx) * std::max(1, m_tgy) * std::max(1, m_tgz);
    if (m_bRequestDependentPSObjects) {
        DEPENDENTCONTEXT ctx;
        memset(&ctx, 0, sizeof(ctx));
        ctx.cbDependentContext = sizeof(ctx);
        ctx.nArguments = (UINT)nArgs;
        ctx.nDependentAccelerators = 1;
        ctx.ppArguments = args.data();
        ctx.ppDatablocks = blocks.data();
        ctx.pbIsDependentBinding = isDep.data();
        ctx.pvDeviceBindings = devBindings.data();
        ctx.pDependentDevices = deps;
        ctx.pStreams = streams;
        ctx.lpszTaskName = const_cast<char *>(GetTaskName());
        LPFNDEPHOSTTASKEX fn = (LPFNDEPHOSTTASKEX)binary;
        for (UINT it = 0; it < uiIterations; ++it)
            for (long long r = 0; r < reps; ++r) fn(&ctx);
        SyncPoint *sp = m_pDepContext->CreateSyncPoint();
        if (!sp) return FALSE;
        for (Touch &t : touched) t.buf->AddOutstanding(sp, t.write);
        sp->Release();
        if (RT().forceSynchronous) m_pDepContext->SynchronizeContext();
    } else {
        if (!m_pDepContext->SynchronizeContext()) return FALSE;
        LPFNDEPHOSTTASK fn = (LPFNDEPHOSTTASK)binary;
        for (UINT it = 0; it < uiIterations; ++it)
            for (long long r = 0; r < reps; ++r) fn((UINT)nArgs, args.data(), isDep.data(), devBindings.data(), 1, deps);
        if (!CheckHip(ptk_hip_device_synchronize(), "device synchronize after a dependent host task")) return FALSE;
    }
    return TRUE;
}

// =============================================================================================
// HIP backend
// =============================================================================================

HIPAsyncContext::HIPAsyncContext(Accelerator *pAcc, ASYNCCONTEXTTYPE e)
    : AsyncContext(pAcc, e), m_hStream(nullptr) {
    pAcc->MakeDeviceContextCurrent();
    CheckHip(ptk_hip_stream_create(&m_hStream), "stream create");
}

HIPAsyncContext::~HIPAsyncContext() {
    m_pAccelerator->MakeDeviceContextCurrent();
    if (m_hStream) {
        ptk_hip_stream_synchronize(m_hStream);
        ptk_hip_stream_destroy(m_hStream);
    }
    for (ptk_hip_event e : m_vFreeEvents) ptk_hip_event_destroy(e);
}

SyncPoint *HIPAsyncContext::CreateSyncPoint() {
    ptk_hip_event ev = nullptr;
    {
        std::lock_guard<std::mutex> g(m_evLock);
        if (!m_vFreeEvents.empty()) {
            ev = m_vFreeEvents.back();
            m_vFreeEvents.pop_back();
        }
    }
    if (!ev && !CheckHip(ptk_hip_event_create(&ev, 0), "event create")) return nullptr;
    CheckHip(ptk_hip_event_record(ev, m_hStream), "event record");
    AddRef();
    SyncPoint *sp = new SyncPoint(this, ev);
    sp->AddRef();
    return sp;
}

BOOL HIPAsyncContext::InsertFence(SyncPoint *pSP) {
    if (!pSP || !pSP->GetPlatformEvent()) return TRUE;
    if (pSP->GetContext() == this) return TRUE;  // same stream: already ordered
    return CheckHip(ptk_hip_stream_wait_event(m_hStream, (ptk_hip_event)pSP->GetPlatformEvent()),
                    "stream wait event");
}

BOOL HIPAsyncContext::SynchronizeContext() {
    return CheckHip(ptk_hip_stream_synchronize(m_hStream), "stream synchronize");
}

void HIPAsyncContext::RecycleEvent(void *pEvent) {
    // Re-recording an event later is safe: waits already enqueued captured its earlier state.
    std::lock_guard<std::mutex> g(m_evLock);
    m_vFreeEvents.push_back((ptk_hip_event)pEvent);
}

// A failed allocation leaves the buffer empty without reporting: the caller
// (AllocDeviceStorage) reclaims pooled memory, retries, and reports only the final failure.
PHIPBuffer::PHIPBuffer(Accelerator *pAcc, size_t cb) : PBuffer(pAcc, cb) {
    pAcc->MakeDeviceContextCurrent();
    if (ptk_hip_malloc(&m_pBuffer, cb) != PTK_HIP_OK) m_pBuffer = nullptr;
}

PHIPBuffer::~PHIPBuffer() {
    if (m_pBuffer) {
        m_pAccelerator->MakeDeviceContextCurrent();
        ptk_hip_free(m_pBuffer);
    }
}

HIPAccelerator::HIPAccelerator(UINT uiId, int nDevice, UINT uiMemSpace)
    : Accelerator(ACCELERATOR_CLASS_HIP, uiId, uiMemSpace), m_nDevice(nDevice), m_pXfer(nullptr) {
    memset(&m_props, 0, sizeof(m_props));
}

HIPAccelerator::~HIPAccelerator() {
    MakeDeviceContextCurrent();
    if (m_pShared) m_pShared->Release();
    if (m_pXfer) m_pXfer->Release();
    for (auto &kv : m_mModules) ptk_hip_module_unload(kv.second);
}

// reference: CUAccelerator::Open, cuaccelerator.cpp:731-950
BOOL HIPAccelerator::Open() {
    if (!CheckHip(ptk_hip_set_device(m_nDevice), "set device")) return FALSE;
    if (!CheckHip(ptk_hip_get_device_props(m_nDevice, &m_props), "device properties")) return FALSE;
    m_strName = m_props.name;
    m_pXfer = CreateAsyncContext(ASYNCCTXT_XFERDTOH);
    return m_pXfer != nullptr;
}

// reference: CUAccelerator::Compile, cuaccelerator.cpp:1229-1440 (cuModuleLoadDataEx +
// cuModuleGetFunction) -- here a prebuilt gfx950 code object and its symbol.
BOOL HIPAccelerator::Compile(const char *lpszFile, const char *lpszOperation, CompiledKernel *pKernel,
                             char *lpszOut, int cbOut) {
    MakeDeviceContextCurrent();
    std::lock_guard<std::mutex> g(m_modLock);
    ptk_hip_module m = nullptr;
    auto it = m_mModules.find(lpszFile);
    if (it != m_mModules.end()) {
        m = it->second;
    } else {
        int rc = ptk_hip_module_load(&m, lpszFile);
        if (rc != PTK_HIP_OK) {
            if (lpszOut && cbOut > 0) snprintf(lpszOut, (size_t)cbOut, "%s", ptk_hip_last_error_string());
            Runtime::Warning((std::string("code object not loadable: ") + ptk_hip_last_error_string()).c_str());
            return FALSE;
        }
        m_mModules[lpszFile] = m;
    }
    ptk_hip_function f = nullptr;
    int rc = ptk_hip_module_get_function(&f, m, lpszOperation);
    if (rc != PTK_HIP_OK) {
        if (lpszOut && cbOut > 0) snprintf(lpszOut, (size_t)cbOut, "%s", ptk_hip_last_error_string());
        return FALSE;
    }
    pKernel->SetPlatformSpecificBinary(this, (void *)f, (void *)m);
    return TRUE;
}

AsyncContext *HIPAccelerator::GetSharedTaskContext() {
    if (RT().streamPolicy != PTK_STREAMS_SHARED) return nullptr;
    std::lock_guard<std::mutex> g(m_modLock);
    if (!m_pShared) m_pShared = CreateAsyncContext(ASYNCCTXT_TASK);
    return m_pShared;
}

BOOL HIPAccelerator::Synchronize() {
    MakeDeviceContextCurrent();
    return CheckHip(ptk_hip_device_synchronize(), "device synchronize");
}

BOOL HIPAccelerator::MakeDeviceContextCurrent() {
    // the current device is per thread in HIP; setting it is cheap when already current
    thread_local int t_current = -1;
    if (t_current == m_nDevice) return TRUE;
    if (ptk_hip_set_device(m_nDevice) != PTK_HIP_OK) return FALSE;
    t_current = m_nDevice;
    return TRUE;
}

AsyncContext *HIPAccelerator::CreateAsyncContext(ASYNCCONTEXTTYPE e) {
    HIPAsyncContext *c = new HIPAsyncContext(this, e);
    if (!c->GetStream()) {
        delete c;
        return nullptr;
    }
    c->AddRef();
    return c;
}

PBuffer *HIPAccelerator::NewPlatformSpecificBuffer(size_t cb) {
    PHIPBuffer *b = new PHIPBuffer(this, cb);
    if (!b->GetBuffer()) {
        delete b;
        return nullptr;
    }
    return b;
}

BOOL HIPAccelerator::CopyHostToDevice(PBuffer *pDst, const void *pSrc, size_t cb, AsyncContext *pCtx) {
    MakeDeviceContextCurrent();
    ptk_hip_stream s = static_cast<HIPAsyncContext *>(pCtx ? pCtx : m_pXfer)->GetStream();
    return CheckHip(ptk_hip_memcpy_htod_async(pDst->GetBuffer(), pSrc, cb, s), "host->device copy");
}

BOOL HIPAccelerator::CopyDeviceToHost(void *pDst, PBuffer *pSrc, size_t cb) {
    MakeDeviceContextCurrent();
    // order after the last write of the source, then copy on the transfer stream
    pSrc->WaitOutstanding(m_pXfer, FALSE);
    HIPAsyncContext *x = static_cast<HIPAsyncContext *>(m_pXfer);
    if (!CheckHip(ptk_hip_memcpy_dtoh_async(pDst, pSrc->GetBuffer(), cb, x->GetStream()),
                  "device->host copy"))
        return FALSE;
    return x->SynchronizeContext();
}

BOOL HIPAccelerator::CopyDeviceToDevice(PBuffer *pDst, PBuffer *pSrc, size_t cb, AsyncContext *pCtx) {
    MakeDeviceContextCurrent();
    ptk_hip_stream s = static_cast<HIPAsyncContext *>(pCtx ? pCtx : m_pXfer)->GetStream();
    return CheckHip(ptk_hip_memcpy_dtod_async(pDst->GetBuffer(), pSrc->GetBuffer(), cb, s),
                    "device->device copy");
}

HIPTask::HIPTask(Graph *pGraph, CompiledKernel *pKernel, const char *lpszName)
    : Task(pGraph, pKernel, ACCELERATOR_CLASS_HIP, lpszName), m_pAccelerator(nullptr),
      m_pContext(nullptr) {}

HIPTask::~HIPTask() { ReleaseDeviceResources(); }

void HIPTask::ReleaseDeviceResources() {
    if (m_pAccelerator) m_pAccelerator->MakeDeviceContextCurrent();
    ResolveProfile();
    if (m_pContext) {
        m_pContext->SynchronizeContext();
        m_pContext->Release();
        m_pContext = nullptr;
    }
}

// reference: CUTask::PlatformSpecificDispatch, cutask.cpp:633-697, with the bindings of
// BindInputs/BindOutputs/BindConstants (task.cpp:4130-4381) and the async dependences of
// WaitIncomingBufferAsyncDependences / CreateBufferAsyncDependences (task.cpp:3756-4086).
BOOL HIPTask::PlatformSpecificDispatch(std::vector<Datablock *> &vIn, std::vector<Datablock *> &vOut,
                                       UINT uiIterations) {
    if (!m_pAccelerator) {
        m_pAccelerator = static_cast<HIPAccelerator *>(m_pMandatoryAccelerator ? m_pMandatoryAccelerator
                                                                               : DefaultHIPAccelerator());
        if (!m_pAccelerator) return Runtime::HandleError("%s: no HIP accelerator", GetTaskName());
    }
    m_pAccelerator->MakeDeviceContextCurrent();
    if (!m_pContext) {
        m_pContext = static_cast<HIPAsyncContext *>(m_pAccelerator->GetSharedTaskContext());
        if (m_pContext) m_pContext->AddRef();
        else m_pContext = static_cast<HIPAsyncContext *>(m_pAccelerator->CreateAsyncContext(ASYNCCTXT_TASK));
        if (!m_pContext) return Runtime::HandleError("%s: stream creation failed", GetTaskName());
    }
    ptk_hip_function fn = (ptk_hip_function)m_pKernel->GetPlatformSpecificBinary(m_pAccelerator);
    if (!fn) return Runtime::HandleError("%s: kernel %s not resolved on %s", GetTaskName(),
                                         m_pKernel->GetOperation(), m_pAccelerator->GetDeviceName());
    m_vArgs.assign(m_uiArgBytes ? m_uiArgBytes : 8, 0);
    // (1) views: populate what is read, make exclusive what is written
    struct Touch { PBuffer *buf; BOOL write; };
    std::vector<Touch> touched;
    touched.reserve(m_vBindings.size());
    GraphReplay *replay = m_pGraph->GetReplay();
    std::vector<GraphReplay::PtrArg> ptrArgs;  // for a launch capture being recorded
    for (Binding &b : m_vBindings) {
        Datablock *blk = b.bInput ? vIn[b.uiSlot] : vOut[b.uiSlot];
        if (!blk) return Runtime::HandleError("%s: argument %s has no block", GetTaskName(),
                                              b.pPort->GetVariableBinding());
        if (b.eType == PTPARM_INT || b.eType == PTPARM_FLOAT || b.eType == PTPARM_BYVALSTRUCT) {
            const void *src = blk->AcquireHostView(FALSE, TRUE);
            UINT n = std::min<UINT>(b.uiBytes, blk->GetDataBufferLogicalSizeBytes());
            memcpy(m_vArgs.data() + b.uiOffset, src, n);
            continue;
        }
        void *dptr = blk->AcquireDeviceView(m_pAccelerator, m_pContext, b.bWritable,
                                            b.bInput ? TRUE : FALSE);
        if (!dptr) return Runtime::HandleError("%s: no device view for %s", GetTaskName(),
                                               b.pPort->GetVariableBinding());
        memcpy(m_vArgs.data() + b.uiOffset, &dptr, sizeof(void *));
        touched.push_back({blk->GetPlatformBuffer(m_pAccelerator->GetMemorySpaceId()), b.bWritable});
        if (replay) ptrArgs.push_back({b.uiOffset, blk, b.bWritable});
    }
    // (2) order after every producer (and, for writes, every reader) on other streams
    for (Touch &t : touched) t.buf->WaitOutstanding(m_pContext, t.write);
    // (3) launch
    PTASKDIM3 grid, block;
    GeometryFor(vIn, &grid, &block);
    if (replay && replay->GetState() == GraphReplay::RECORDING)
        replay->OnLaunch(this, m_pAccelerator, fn, grid, block, m_vArgs.data(), m_uiArgBytes, ptrArgs, uiIterations);
    ptk_hip_stream s = m_pContext->GetStream();
    const BOOL prof = Runtime::GetProfilePSDispatch();
    ptk_hip_event evStart = nullptr, evEnd = nullptr;
    if (prof) {
        CheckHip(ptk_hip_event_create(&evStart, 1), "event create");
        CheckHip(ptk_hip_event_create(&evEnd, 1), "event create");
        ptk_hip_event_record(evStart, s);
    }
    for (UINT it = 0; it < uiIterations; ++it) {
        int rc = ptk_hip_launch(fn, (unsigned)grid.x, (unsigned)grid.y, (unsigned)grid.z,
                                (unsigned)block.x, (unsigned)block.y, (unsigned)block.z, 0, s,
                                m_vArgs.data(), m_uiArgBytes);
        if (rc != PTK_HIP_OK)
            return Runtime::HandleError("%s: launch of %s (grid %d,%d,%d block %d,%d,%d) failed: %s",
                                        GetTaskName(), m_pKernel->GetOperation(), grid.x, grid.y,
                                        grid.z, block.x, block.y, block.z, ptk_hip_last_error_string());
    }
    if (prof) {
        ptk_hip_event_record(evEnd, s);
        m_vProfEvents.push_back(std::make_pair(evStart, evEnd));
    }
    // (4) one sync point orders every later access to the buffers touched here
    SyncPoint *sp = m_pContext->CreateSyncPoint();
    if (!sp) return FALSE;
    for (Touch &t : touched) t.buf->AddOutstanding(sp, t.write);
    sp->Release();
    if (RT().forceSynchronous) m_pContext->SynchronizeContext();
    return TRUE;
}

// Device time of the profiled dispatches: waits for their end events, accumulates the
// start->end intervals (reference: the ProfilePSDispatch timing of cutask.cpp:709-776, without
// its per-dispatch synchronisation) and recycles the events.
void HIPTask::ResolveProfile() {
    for (auto &pr : m_vProfEvents) {
        float ms = 0.f;
        if (ptk_hip_event_synchronize(pr.second) == PTK_HIP_OK &&
            ptk_hip_event_elapsed_ms(&ms, pr.first, pr.second) == PTK_HIP_OK) {
            m_dLastDeviceMs = ms;
            m_dTotalDeviceMs += ms;
        }
        ptk_hip_event_destroy(pr.first);
        ptk_hip_event_destroy(pr.second);
    }
    m_vProfEvents.clear();
}

double HIPTask::GetTotalDispatchDeviceMsResolved() {
    if (m_pAccelerator) m_pAccelerator->MakeDeviceContextCurrent();
    ResolveProfile();
    return m_dTotalDeviceMs;
}

}  // namespace PTask
