// runtime.cpp -- PTask::Runtime: initialisation, accelerator enumeration, factories for
// templates, datablocks, kernels and ports, configuration and error handling
// (reference roles: ptask/PTaskRuntime.cpp, ptask/ptaskutils.cpp).

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>

#include "ptk_internal.h"

namespace PTask {

RuntimeState &RT() {
    static RuntimeState *s = new RuntimeState();  // never destroyed: runner threads may outlive main
    return *s;
}

// ---- DispatchArbiter ---------------------------------------------------------------------------

void DispatchArbiter::Acquire(int nPriority) {
    std::unique_lock<std::mutex> lk(m_lock);
    const Ticket me{nPriority, m_uiSeq++};
    m_vWaiting.push_back(me);
    // best waiting ticket: highest priority, then earliest arrival
    auto first = [&] {
        for (const Ticket &w : m_vWaiting)
            if (w.prio > me.prio || (w.prio == me.prio && w.seq < me.seq)) return false;
        return true;
    };
    m_cv.wait(lk, [&] { return !m_bQuiesced && m_nInUse < m_nSlots && first(); });
    for (size_t i = 0; i < m_vWaiting.size(); ++i)
        if (m_vWaiting[i].seq == me.seq) {
            m_vWaiting.erase(m_vWaiting.begin() + (long)i);
            break;
        }
    m_nInUse++;
    m_uiAdmitted++;
    lk.unlock();
    m_cv.notify_all();  // the next best may take a remaining slot
}

void DispatchArbiter::Release() {
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_nInUse--;
    }
    m_cv.notify_all();
}

void DispatchArbiter::SetSlots(int n) {
    {
        std::lock_guard<std::mutex> g(m_lock);
        m_nSlots = n < 1 ? 1 : n;
    }
    m_cv.notify_all();
}

int DispatchArbiter::GetSlots() {
    std::lock_guard<std::mutex> g(m_lock);
    return m_nSlots;
}

void DispatchArbiter::Quiesce(BOOL bQuiesce) {
    std::unique_lock<std::mutex> lk(m_lock);
    m_bQuiesced = bQuiesce;
    if (bQuiesce) {
        m_cv.wait(lk, [&] { return m_nInUse == 0; });
    } else {
        lk.unlock();
        m_cv.notify_all();
    }
}

int DispatchArbiter::GetWaiting() {
    std::lock_guard<std::mutex> g(m_lock);
    return (int)m_vWaiting.size();
}

unsigned long long DispatchArbiter::GetAdmitted() {
    std::lock_guard<std::mutex> g(m_lock);
    return m_uiAdmitted;
}

const char *AccClassString(ACCELERATOR_CLASS cls) {
    static const char *names[] = {"ACCELERATOR_CLASS_DIRECT_X", "ACCELERATOR_CLASS_OPEN_CL",
                                  "ACCELERATOR_CLASS_CUDA",     "ACCELERATOR_CLASS_REFERENCE",
                                  "ACCELERATOR_CLASS_HOST",     "ACCELERATOR_CLASS_UNKNOWN",
                                  "ACCELERATOR_CLASS_SUPER",    "ACCELERATOR_CLASS_HIP"};
    int i = (int)cls;
    return (i >= 0 && i <= 7) ? names[i] : "ACCELERATOR_CLASS_INVALID";
}

Accelerator *HostAcceleratorInstance() {
    auto &v = RT().hostAccelerators;
    return v.empty() ? nullptr : v[0];
}

HIPAccelerator *DefaultHIPAccelerator() {
    for (Accelerator *a : RT().hipAccelerators)
        if (a->IsEnabled()) return static_cast<HIPAccelerator *>(a);
    return nullptr;
}

Accelerator *AcceleratorForMemorySpace(UINT uiMemSpace) {
    if (uiMemSpace == HOST_MEMORY_SPACE_ID) return HostAcceleratorInstance();
    for (Accelerator *a : RT().hipAccelerators)
        if (a->GetMemorySpaceId() == uiMemSpace) return a;
    return nullptr;
}

BOOL CheckHip(int status, const char *what) {
    if (status == PTK_HIP_OK) return TRUE;
    return Runtime::HandleError("%s failed: %s (%s)", what, ptk_hip_status_string(status),
                                ptk_hip_last_error_string());
}

// ---- ptaskutils ------------------------------------------------------------------------------

ACCELERATOR_CLASS ptaskutils::SelectAcceleratorClass(const char *lpszFileName) {
    if (!lpszFileName) return ACCELERATOR_CLASS_UNKNOWN;
    if (!strcasecmp(lpszFileName, "HostFunction")) return ACCELERATOR_CLASS_HOST;
    const char *dot = strrchr(lpszFileName, '.');
    if (!dot) return ACCELERATOR_CLASS_UNKNOWN;
    if (!strcasecmp(dot, ".hsaco") || !strcasecmp(dot, ".co")) return ACCELERATOR_CLASS_HIP;
    if (!strcasecmp(dot, ".so") || !strcasecmp(dot, ".dll")) return ACCELERATOR_CLASS_HOST;
    if (!strcasecmp(dot, ".ptx") || !strcasecmp(dot, ".cu")) return ACCELERATOR_CLASS_CUDA;
    if (!strcasecmp(dot, ".hlsl")) return ACCELERATOR_CLASS_DIRECT_X;
    if (!strcasecmp(dot, ".cl")) return ACCELERATOR_CLASS_OPEN_CL;
    return ACCELERATOR_CLASS_UNKNOWN;
}

UINT ptaskutils::nextpow2(UINT x) {
    UINT p = 1;
    while (p < x) p <<= 1;
    return p;
}

// ---- CompiledKernel ----------------------------------------------------------------------------

CompiledKernel::CompiledKernel(const char *lpszSourceFile, const char *lpszOperation,
                               ACCELERATOR_CLASS eClass)
    : m_strFile(lpszSourceFile ? lpszSourceFile : ""), m_strOp(lpszOperation ? lpszOperation : ""),
      m_eClass(eClass), m_tgx(1), m_tgy(1), m_tgz(1) {}

CompiledKernel::~CompiledKernel() {}
const char *CompiledKernel::GetSourceFile() { return m_strFile.c_str(); }
const char *CompiledKernel::GetOperation() { return m_strOp.c_str(); }
ACCELERATOR_CLASS CompiledKernel::GetAcceleratorClass() { return m_eClass; }

void *CompiledKernel::GetPlatformSpecificBinary(Accelerator *pAcc) {
    auto it = m_binaries.find(pAcc);
    return it == m_binaries.end() ? nullptr : it->second.first;
}
void *CompiledKernel::GetPlatformSpecificModule(Accelerator *pAcc) {
    auto it = m_binaries.find(pAcc);
    return it == m_binaries.end() ? nullptr : it->second.second;
}
void CompiledKernel::SetPlatformSpecificBinary(Accelerator *pAcc, void *pBinary, void *pModule) {
    m_binaries[pAcc] = std::make_pair(pBinary, pModule);
}
BOOL CompiledKernel::HasBinaryFor(Accelerator *pAcc) { return GetPlatformSpecificBinary(pAcc) != nullptr; }

// ---- Runtime -------------------------------------------------------------------------------------

namespace Runtime {

static void vreport(const char *prefix, const char *fmt, va_list ap) {
    char buf[2048];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    fprintf(stderr, "%s%s%s", prefix, buf, (buf[0] && buf[strlen(buf) - 1] == '\n') ? "" : "\n");
}

BOOL HandleError(const char *szMessage, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, szMessage);
    vsnprintf(buf, sizeof(buf), szMessage, ap);
    va_end(ap);
    {
        std::lock_guard<std::recursive_mutex> g(RT().lock);
        RT().lastError = buf;
    }
    fprintf(stderr, "PTask runtime error: %s\n", buf);
    if (RT().exitOnFailure) {
        fflush(stderr);
        exit(1);
    }
    return FALSE;
}

BOOL HandleError(const std::string &strError) { return HandleError("%s", strError.c_str()); }

const char *GetLastErrorText() { return RT().lastError.c_str(); }
void ClearLastError() {
    std::lock_guard<std::recursive_mutex> g(RT().lock);
    RT().lastError.clear();
}

void Inform(const char *szMessage, ...) {
    if (RT().loggingLevel == 0) return;
    va_list ap;
    va_start(ap, szMessage);
    vreport("PTask: ", szMessage, ap);
    va_end(ap);
}

void Trace(const char *szMessage, ...) {
    if (RT().loggingLevel < 2) return;
    va_list ap;
    va_start(ap, szMessage);
    vreport("PTask trace: ", szMessage, ap);
    va_end(ap);
}

void Warning(const char *szMessage, UINT nLogLevel) {
    if (nLogLevel != 0xFFFFFFFF && RT().loggingLevel < nLogLevel) return;
    fprintf(stderr, "PTask warning: %s\n", szMessage);
}

BOOL MandatoryInform(const char *szMessage, ...) {
    va_list ap;
    va_start(ap, szMessage);
    vreport("", szMessage, ap);
    va_end(ap);
    return TRUE;
}

// Accelerator discovery (reference: Scheduler::CreateAccelerators, Scheduler.cpp:2129-2230):
// one host accelerator (memory space 0), then every HIP device up to the concurrency cap.
PTRESULT Initialize() {
    RuntimeState &S = RT();
    std::lock_guard<std::recursive_mutex> g(S.lock);
    if (S.initialized) return PTASK_ERR_ALREADY_INITIALIZED;
    if (S.useHost) {
        HostAccelerator *h = new HostAccelerator(0);
        h->AddRef();
        h->Open();
        S.hostAccelerators.push_back(h);
    }
    if (S.useHIP) {
        int n = 0;
        if (ptk_hip_get_device_count(&n) != PTK_HIP_OK) n = 0;
        if (S.maxConcurrency > 0 && n > S.maxConcurrency) n = S.maxConcurrency;
        if (S.physicalAccelerators > 0 && n > (int)S.physicalAccelerators) n = (int)S.physicalAccelerators;
        for (int d = 0; d < n; ++d) {
            HIPAccelerator *a = new HIPAccelerator((UINT)(1 + d), d, (UINT)(1 + d));
            a->AddRef();
            if (!a->Open()) {
                a->Release();
                continue;
            }
            S.hipAccelerators.push_back(a);
            Inform("HIP accelerator %d: %s", d, a->GetDeviceName());
        }
    }
    S.sizeDescriptorTemplate = new DatablockTemplate("__ptk_size_descriptor", sizeof(int), PTPARM_INT);
    S.sizeDescriptorTemplate->AddRef();
    S.controlTemplate = new DatablockTemplate("__ptk_control", sizeof(CONTROLSIGNAL), PTPARM_BYVALSTRUCT);
    S.controlTemplate->AddRef();
    S.initialized = true;
    AllocateGlobalPools();
    return PTASK_OK;
}

BOOL IsInitialized() { return RT().initialized; }

PTRESULT Terminate() {
    RuntimeState &S = RT();
    std::lock_guard<std::recursive_mutex> g(S.lock);
    if (!S.initialized) return PTASK_ERR_UNINITIALIZED;
    DestroyGlobalPools();  // while the accelerators that hold their device storage are open
    for (CompiledKernel *k : S.kernels) delete k;
    S.kernels.clear();
    for (auto &kv : S.templates) kv.second->Release();
    S.templates.clear();
    for (Port *p : S.ports) p->Release();
    S.ports.clear();
    for (Accelerator *a : S.hipAccelerators) {
        a->Synchronize();
        a->Release();
    }
    S.hipAccelerators.clear();
    for (Accelerator *a : S.hostAccelerators) a->Release();
    S.hostAccelerators.clear();
    PTSRELEASE(S.sizeDescriptorTemplate);
    PTSRELEASE(S.controlTemplate);
    S.initialized = false;
    return PTASK_OK;
}

DatablockTemplate *RegisterTemplate(DatablockTemplate *t) {
    RuntimeState &S = RT();
    std::lock_guard<std::recursive_mutex> g(S.lock);
    auto it = S.templates.find(t->GetTemplateName());
    if (it != S.templates.end()) {
        // same name: the first registration wins, as in the reference's template table
        delete t;
        return it->second;
    }
    t->AddRef();
    S.templates[t->GetTemplateName()] = t;
    return t;
}

DatablockTemplate *GetDatablockTemplate(char *lpszType, unsigned int uiStride, unsigned int x,
                                        unsigned int y, unsigned int z, bool bRecordStream,
                                        bool bRaw) {
    return RegisterTemplate(new DatablockTemplate(lpszType, uiStride, x, y, z, uiStride * x,
                                                  bRecordStream, bRaw));
}

DatablockTemplate *GetDatablockTemplate(char *lpszTemplateName, unsigned int uiElementStride,
                                        unsigned int uiElementsX, unsigned int uiElementsY,
                                        unsigned int uiElementsZ, unsigned int uiPitch,
                                        bool bIsRecordStream, bool bIsByteAddressable) {
    return RegisterTemplate(new DatablockTemplate(lpszTemplateName, uiElementStride, uiElementsX,
                                                  uiElementsY, uiElementsZ, uiPitch,
                                                  bIsRecordStream, bIsByteAddressable));
}

DatablockTemplate *GetDatablockTemplate(char *lpszTemplateName, BUFFERDIMENSIONS *pBufferDims,
                                        unsigned int uiNumBufferDims, bool bIsRecordStream,
                                        bool bIsByteAddressable) {
    return RegisterTemplate(new DatablockTemplate(lpszTemplateName, pBufferDims, uiNumBufferDims,
                                                  bIsRecordStream, bIsByteAddressable));
}

DatablockTemplate *GetDatablockTemplate(char *lpszType, unsigned int uiStride,
                                        PTASK_PARM_TYPE pttype) {
    return RegisterTemplate(new DatablockTemplate(lpszType, uiStride, pttype));
}

DatablockTemplate *GetDatablockTemplate(char *lpszType) {
    RuntimeState &S = RT();
    std::lock_guard<std::recursive_mutex> g(S.lock);
    auto it = S.templates.find(lpszType ? lpszType : "");
    return it == S.templates.end() ? nullptr : it->second;
}

// reference: PTaskRuntime.cpp:3279 -- the block starts valid on the host with the given bytes
Datablock *AllocateDatablock(DatablockTemplate *pTemplate, void *pInitData, UINT cbInitData,
                             Channel *pDestChannel, BUFFERACCESSFLAGS flags,
                             CONTROLSIGNAL uiBlockControlCode) {
    (void)pDestChannel;
    UINT cb = cbInitData;
    if (pTemplate && cb == 0) cb = pTemplate->GetDatablockByteCount(DBDATA_IDX);
    // a pooled block when a runtime-wide pool holds one (RequireBlockPool), else a new block
    Datablock *b = cb ? RequestGlobalPooledBlock(pTemplate, cb, 0) : nullptr;
    if (b) {
        if (uiBlockControlCode) b->SetControlSignal(uiBlockControlCode);
        b->SetAccessFlags(flags ? flags : PT_ACCESS_DEFAULT);
    } else {
        b = new Datablock(pTemplate, cb, 0, uiBlockControlCode);
        b->AddRef();
        if (flags) b->SetAccessFlags(flags);
    }
    if (pInitData && cbInitData) {
        b->Lock();
        // a recycled block's old contents are read back only when the new data covers part of it
        void *dst = b->AcquireHostView(TRUE, cbInitData < b->GetDataBufferLogicalSizeBytes());
        if (dst) memcpy(dst, pInitData, cbInitData);
        b->Unlock();
        if (!dst) {  // out of host memory even after reclaiming pooled blocks (reported)
            b->Release();
            return nullptr;
        }
    }
    if (pTemplate && pTemplate->DescribesRecordStream() == FALSE && pTemplate->GetStride() &&
        !pTemplate->IsScalarParameter())
        b->SetRecordCount(cb / pTemplate->GetStride());
    return b;
}

Datablock *AllocateDatablock(DatablockTemplate *pTemplate, UINT uiDataSize, UINT uiMetaSize,
                             UINT uiTemplateSize, CONTROLSIGNAL uiBlockControlCode) {
    (void)uiTemplateSize;
    Datablock *b = uiDataSize ? RequestGlobalPooledBlock(pTemplate, uiDataSize, uiMetaSize) : nullptr;
    if (b) {
        if (uiBlockControlCode) b->SetControlSignal(uiBlockControlCode);
        return b;
    }
    b = new Datablock(pTemplate, uiDataSize, uiMetaSize, uiBlockControlCode);
    b->AddRef();
    return b;
}

Datablock *AllocateDatablock(CONTROLSIGNAL uiBlockControlCode) {
    Datablock *b = new Datablock(nullptr, 0, 0, uiBlockControlCode);
    b->AddRef();
    return b;
}

void FreeDatablock(Datablock *pBlock) {
    if (pBlock) pBlock->Release();
}

// reference: GetCompiledKernelEx, PTaskRuntime.cpp:2904-2970 -- compile for every capable
// accelerator of the class the file extension selects; NULL (after HandleError) on failure.
CompiledKernel *GetCompiledKernel(char *lpszFile, char *lpszOperation, char *lpszCompilerOutput,
                                  int uiCompilerOutput, int tgx, int tgy, int tgz) {
    RuntimeState &S = RT();
    if (!S.initialized) {
        HandleError("GetCompiledKernel(%s, %s) before Runtime::Initialize", lpszFile, lpszOperation);
        return nullptr;
    }
    ACCELERATOR_CLASS cls = ptaskutils::SelectAcceleratorClass(lpszFile);
    std::vector<Accelerator *> *accs = nullptr;
    if (cls == ACCELERATOR_CLASS_HIP) accs = &S.hipAccelerators;
    else if (cls == ACCELERATOR_CLASS_HOST) accs = &S.hostAccelerators;
    if (!accs) {
        HandleError("GetCompiledKernel(%s): %s kernels are not supported by this runtime (HIP "
                    "code objects and host libraries are)", lpszFile, AccClassString(cls));
        return nullptr;
    }
    if (accs->empty()) {
        HandleError("GetCompiledKernel(%s, %s): no %s accelerator is available", lpszFile,
                    lpszOperation, AccClassString(cls));
        return nullptr;
    }
    CompiledKernel *k = new CompiledKernel(lpszFile, lpszOperation, cls);
    k->SetGroupSize(tgx, tgy, tgz);
    BOOL any = FALSE;
    for (Accelerator *a : *accs) {
        if (!a->IsEnabled()) continue;
        if (a->Compile(lpszFile, lpszOperation, k, lpszCompilerOutput, uiCompilerOutput)) any = TRUE;
    }
    if (!any) {
        delete k;
        HandleError("GetCompiledKernel(%s, %s): the kernel could not be loaded", lpszFile, lpszOperation);
        return nullptr;
    }
    std::lock_guard<std::recursive_mutex> g(S.lock);
    S.kernels.push_back(k);
    return k;
}

// reference: GetHostFunctionCompiledKernel, PTaskRuntime.cpp:2992-3019
CompiledKernel *GetHostFunctionCompiledKernel(char *lpszOperation, FARPROC lpfn) {
    RuntimeState &S = RT();
    if (!S.initialized || S.hostAccelerators.empty() || !lpfn) {
        HandleError("GetHostFunctionCompiledKernel(%s): no host accelerator or null function",
                    lpszOperation);
        return nullptr;
    }
    CompiledKernel *k = new CompiledKernel("HostFunction", lpszOperation, ACCELERATOR_CLASS_HOST);
    for (Accelerator *a : S.hostAccelerators) k->SetPlatformSpecificBinary(a, (void *)lpfn, nullptr);
    std::lock_guard<std::recursive_mutex> g(S.lock);
    S.kernels.push_back(k);
    return k;
}

Port *CreatePort(PORTTYPE type, DatablockTemplate *pTemplate, UINT uiId, char *lpszVariableBinding,
                 UINT nKernelParameterIndex, UINT nInOutParmOutputIndex) {
    Port *p = nullptr;
    switch (type) {
        case INPUT_PORT:
            p = new InputPort(pTemplate, uiId, lpszVariableBinding, nKernelParameterIndex,
                              nInOutParmOutputIndex);
            break;
        case OUTPUT_PORT:
            p = new OutputPort(pTemplate, uiId, lpszVariableBinding, nKernelParameterIndex,
                               nInOutParmOutputIndex);
            break;
        case STICKY_PORT:
            p = new StickyPort(pTemplate, uiId, lpszVariableBinding, nKernelParameterIndex);
            break;
        case META_PORT:
            p = new MetaPort(pTemplate, uiId, lpszVariableBinding, nKernelParameterIndex, nInOutParmOutputIndex);
            break;
        case INITIALIZER_PORT:
            p = new InitializerPort(pTemplate, uiId, lpszVariableBinding, nKernelParameterIndex,
                                    nInOutParmOutputIndex);
            break;
        default:
            HandleError("CreatePort: unknown port type %d", (int)type);
            return nullptr;
    }
    p->AddRef();  // the runtime's reference, dropped at Terminate
    std::lock_guard<std::recursive_mutex> g(RT().lock);
    RT().ports.push_back(p);
    return p;
}

PTRESULT SetTaskAffinity(Task *pTask, Accelerator *pAccelerator, AFFINITYTYPE eAffinityType) {
    if (!pTask) return PTASK_ERR_INVALID_PARAMETER;
    return pTask->SetAffinity(pAccelerator, eAffinityType) ? PTASK_OK : PTASK_ERR;
}

// the accelerator of the task's class with this id (ids as EnumerateAccelerators reports them)
static Accelerator *AcceleratorById(Task *pTask, UINT uiAcceleratorId) {
    const ACCELERATOR_CLASS c = pTask->GetAcceleratorClass();
    for (UINT i = 0; i < GetAcceleratorCount(c); ++i) {
        Accelerator *a = GetAccelerator(c, i);
        if (a && a->GetAcceleratorId() == uiAcceleratorId) return a;
    }
    return nullptr;
}

PTRESULT SetTaskAffinity(Task *pTask, UINT uiAcceleratorId, AFFINITYTYPE eAffinityType) {
    if (!pTask) return PTASK_ERR_INVALID_PARAMETER;
    Accelerator *a = AcceleratorById(pTask, uiAcceleratorId);
    if (!a) return PTASK_ERR_INVALID_PARAMETER;
    return SetTaskAffinity(pTask, a, eAffinityType);
}

PTRESULT SetTaskAffinity(Task *pTask, UINT *pvAcceleratorIds, AFFINITYTYPE *pvAffinityTypes, UINT nAcceleratorIds) {
    if (!pTask || (nAcceleratorIds && (!pvAcceleratorIds || !pvAffinityTypes))) return PTASK_ERR_INVALID_PARAMETER;
    for (UINT i = 0; i < nAcceleratorIds; ++i) {
        PTRESULT r = SetTaskAffinity(pTask, pvAcceleratorIds[i], pvAffinityTypes[i]);
        if (PTFAILED(r)) return r;
    }
    return PTASK_OK;
}

static std::vector<Accelerator *> *ListFor(ACCELERATOR_CLASS eClass) {
    if (eClass == ACCELERATOR_CLASS_HIP) return &RT().hipAccelerators;
    if (eClass == ACCELERATOR_CLASS_HOST) return &RT().hostAccelerators;
    return nullptr;
}

PTRESULT EnumerateAccelerators(ACCELERATOR_CLASS eClass, UINT uiEnumerationIndex,
                               ACCELERATOR_DESCRIPTOR **ppDescriptor) {
    if (!ppDescriptor) return PTASK_ERR_INVALID_PARAMETER;
    *ppDescriptor = nullptr;
    Accelerator *a = GetAccelerator(eClass, uiEnumerationIndex);
    if (!a) return PTASK_ERR_NOT_FOUND;
    // malloc'd, as the reference's callers free() it (PTaskRuntime.cpp:1140, graphgroupby.cpp:538)
    ACCELERATOR_DESCRIPTOR *d = (ACCELERATOR_DESCRIPTOR *)calloc(1, sizeof(ACCELERATOR_DESCRIPTOR));
    if (!d) return PTASK_ERR;
    a->GetDescriptor(d);
    *ppDescriptor = d;
    return PTASK_OK;
}

Accelerator *GetAccelerator(ACCELERATOR_CLASS eClass, UINT uiIndex) {
    auto *v = ListFor(eClass);
    if (!v || uiIndex >= v->size()) return nullptr;
    return (*v)[uiIndex];
}

UINT GetAcceleratorCount(ACCELERATOR_CLASS eClass) {
    auto *v = ListFor(eClass);
    return v ? (UINT)v->size() : 0;
}

// ---- graph helpers: the Runtime:: forms of Graph, Port and Channel calls -----------------------
// (reference PTaskRuntime.cpp:4690-4790, PTaskRuntime.h:2441-2560)

void SetPortMetaFunction(Port *pPort, METAFUNCTION eCanonicalMetaFunction) {
    if (pPort) pPort->SetMetaFunction(eCanonicalMetaFunction);
}

Task *AddTask(Graph *pGraph, CompiledKernel *pKernel, UINT uiInputPortCount, Port **pvInputPorts, UINT uiOutputPorts,
              Port **pvOutputPorts, char *lpszTaskName) {
    if (!pGraph) {
        HandleError("AddTask: no graph");
        return nullptr;
    }
    return pGraph->AddTask(pKernel, uiInputPortCount, pvInputPorts, uiOutputPorts, pvOutputPorts, lpszTaskName);
}

void BindDerivedPort(Graph *pGraph, Port *pDescribedPort, Port *pDescriberPort, DESCRIPTORFUNC func) {
    if (pGraph) pGraph->BindDescriptorPort(pDescribedPort, pDescriberPort, func);
}

void RunGraph(Graph *pGraph) {
    if (pGraph) pGraph->Run();
}

void Push(Channel *pChannel, Datablock *pBlock) {
    if (pChannel) pChannel->Push(pBlock);
}

Datablock *Pull(Channel *pChannel, DWORD dwTimeout) { return pChannel ? pChannel->Pull(dwTimeout) : nullptr; }

PTRESULT CheckGraphSemantics(Graph *pGraph, BOOL bVerbose, BOOL bFailOnWarning) {
    return pGraph ? pGraph->CheckGraphSemantics(bVerbose, bFailOnWarning) : PTASK_ERR;
}

UINT GetPhysicalAcceleratorCount() {
    RuntimeState &S = RT();
    if (S.physicalAccelerators) return S.physicalAccelerators;
    return (UINT)S.hipAccelerators.size();
}

void SetPhysicalAcceleratorCount(UINT nAcceleratorCount) { RT().physicalAccelerators = nAcceleratorCount; }

BOOL CheckInvariantCondition(BOOL bCondition, char *lpszErrorText) {
    if (bCondition) return TRUE;
    Warning(lpszErrorText ? lpszErrorText : "invariant violated");
    return FALSE;
}

BOOL CheckInvariantCondition(BOOL bCondition, std::string strErrorText) {
    if (bCondition) return TRUE;
    Warning(strErrorText.c_str());
    return FALSE;
}

void SetUseHost(BOOL b) { RT().useHost = b; }
void SetUseHIP(BOOL b) { RT().useHIP = b; }
void SetUseCUDA(BOOL) {}
void SetUseOpenCL(BOOL) {}
void SetUseDirectX(BOOL) {}
BOOL GetUseHost() { return RT().useHost; }
BOOL GetUseHIP() { return RT().useHIP; }
void SetMaximumConcurrency(int n) { RT().maxConcurrency = n; }
int GetMaximumConcurrency() { return RT().maxConcurrency; }
void SetMaximumHostConcurrency(int n) {
    RT().maxHostConcurrency = n < 1 ? 1 : n;
    RT().hostArbiter.SetSlots(RT().maxHostConcurrency);
}
int GetMaximumHostConcurrency() { return RT().maxHostConcurrency; }
void SetMaximumDeviceDispatchConcurrency(int n) { RT().hipArbiter.SetSlots(n < 1 ? 1 : n); }
int GetMaximumDeviceDispatchConcurrency() { return RT().hipArbiter.GetSlots(); }
void GlobalQuiesce() {
    RT().hostArbiter.Quiesce(TRUE);
    RT().hipArbiter.Quiesce(TRUE);
}
void EndGlobalQuiescence() {
    RT().hostArbiter.Quiesce(FALSE);
    RT().hipArbiter.Quiesce(FALSE);
}
int GetWaitingDispatchCount() { return RT().hostArbiter.GetWaiting() + RT().hipArbiter.GetWaiting(); }
void SetDefaultChannelCapacity(int n) { RT().channelCapacity = n < 1 ? 1 : n; }
int GetDefaultChannelCapacity() { return RT().channelCapacity; }
void SetTaskThreadPoolSize(UINT n) { RT().threadPoolSize = n; }
UINT GetTaskThreadPoolSize() { return RT().threadPoolSize; }
void SetTaskThreadPoolPolicy(THREADPOOLPOLICY) {}
void SetSchedulingMode(int mode) { RT().schedulingMode = mode; }
int GetSchedulingMode() { return RT().schedulingMode; }
void SetForceSynchronous(BOOL b) { RT().forceSynchronous = b; }
BOOL GetForceSynchronous() { return RT().forceSynchronous; }
void SetProfilePSDispatch(BOOL b) { RT().profilePSDispatch = b; }
BOOL GetProfilePSDispatch() { return RT().profilePSDispatch; }
void SetDebugMode(BOOL) {}
void SetUseGraphMonitorWatchdog(BOOL b) { RT().useWatchdog = b; }
BOOL GetUseGraphMonitorWatchdog() { return RT().useWatchdog; }
void SetDispatchWatchdogThreshold(DWORD dwThreshold) { RT().watchdogThresholdMs = dwThreshold ? dwThreshold : 1; }
DWORD GetDispatchWatchdogThreshold() { return RT().watchdogThresholdMs; }
unsigned long long GetWatchdogReportCount() { return RT().watchdogReports; }
void SetDefaultViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY) {}
void SetDefaultOutputViewMaterializationPolicy(VIEWMATERIALIZATIONPOLICY e) { RT().outputViewPolicy = e; }
VIEWMATERIALIZATIONPOLICY GetDefaultOutputViewMaterializationPolicy() { return RT().outputViewPolicy; }
void SetLoggingLevel(UINT n) { RT().loggingLevel = n; }
UINT GetLoggingLevel() { return RT().loggingLevel; }
void SetExitOnRuntimeFailure(BOOL b) { RT().exitOnFailure = b; }
BOOL GetExitOnRuntimeFailure() { return RT().exitOnFailure; }
void SetTaskStreamPolicy(int e) { RT().streamPolicy = e; }
int GetTaskStreamPolicy() { return RT().streamPolicy; }
void SetDefaultBlockPoolSize(UINT n) { RT().blockPoolSize = n; }
UINT GetDefaultBlockPoolSize() { return RT().blockPoolSize; }

void GetTransferStatistics(unsigned long long *pHtoDCount, unsigned long long *pHtoDBytes,
                           unsigned long long *pDtoHCount, unsigned long long *pDtoHBytes) {
    if (pHtoDCount) *pHtoDCount = RT().h2dCount.load();
    if (pHtoDBytes) *pHtoDBytes = RT().h2dBytes.load();
    if (pDtoHCount) *pDtoHCount = RT().d2hCount.load();
    if (pDtoHBytes) *pDtoHBytes = RT().d2hBytes.load();
}

void ResetTransferStatistics() {
    RT().h2dCount = 0;
    RT().h2dBytes = 0;
    RT().d2hCount = 0;
    RT().d2hBytes = 0;
}

}  // namespace Runtime
}  // namespace PTask
