/* datablock.h -- the unit of data that flows through channels (reference: ptask/datablock.h).
 *
 * A block owns one logical byte range in up to two memory spaces: the host (memory space 0)
 * and the HIP device (memory space 1).  Each space has a coherence state (BSTATE_*); a view is
 * brought up to date on demand and a write invalidates the other space, so a block that only
 * ever moves between HIP tasks never crosses PCIe.  Host access is bracketed by Lock/Unlock
 * (reference usage: ptaskUnitTest/graphfdtd.cpp:1118-1124). */
#ifndef PTK_DATABLOCK_H
#define PTK_DATABLOCK_H

#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class DatablockTemplate;
class Accelerator;
class AsyncContext;
class PBuffer;
class BlockPool;

static const UINT HOST_MEMORY_SPACE_ID = 0;
static const UINT PTK_MAX_MEMORY_SPACES = 2;

class Datablock : public ReferenceCounted, public Lockable {
public:
    Datablock(DatablockTemplate *pTemplate, UINT cbData, UINT cbMeta, CONTROLSIGNAL luiCode);
    virtual ~Datablock();

    /* Host view of the data channel.  bWriteable invalidates device copies; bSynchronize=TRUE
     * waits for outstanding device writes and copies back when the host copy is stale. */
    void *GetDataPointer(BOOL bWriteable, BOOL bSynchronize = TRUE);
    void *GetMetadataPointer(BOOL bWriteable, BOOL bSynchronize = TRUE);

    DatablockTemplate *GetTemplate();
    UINT GetDataBufferLogicalSizeBytes();
    UINT GetDataBufferAllocatedSizeBytes();
    UINT GetMetaBufferLogicalSizeBytes();
    void GrowDataBuffer(UINT uiNewSize);
    void Seal(UINT uiRecordCount, UINT uiDataSize, UINT uiMetaDataSize, UINT uiTemplateSize);
    BOOL IsSealed();
    UINT GetRecordCount();
    void SetRecordCount(UINT uiRecordCount);
    BOOL HasValidDataBuffer();
    CONTROLSIGNAL GetControlSignals();
    BOOL TestControlSignal(CONTROLSIGNAL luiCode);
    VOID SetControlSignal(CONTROLSIGNAL luiCode);
    VOID ClearControlSignal(CONTROLSIGNAL luiCode);
    VOID ClearAllControlSignals();
    BUFFERACCESSFLAGS GetAccessFlags();
    void SetAccessFlags(BUFFERACCESSFLAGS e);
    UINT GetDBUID();
    BUFFER_COHERENCE_STATE GetCoherenceState(UINT uiMemorySpaceId);

    /* ---- runtime internals (used by tasks, ports and backends) ---------------------------- */

    /* Device view on `pAcc` for a dispatch running on `pCtx`.  bPopulate copies host data in
     * (on pCtx's stream) when the device copy is stale; bWriteable makes the device exclusive. */
    void *AcquireDeviceView(Accelerator *pAcc, AsyncContext *pCtx, BOOL bWriteable,
                            BOOL bPopulate);
    /* Host view for a host task (same rules as GetDataPointer, without the public checks). */
    void *AcquireHostView(BOOL bWriteable, BOOL bPopulate);
    PBuffer *GetPlatformBuffer(UINT uiMemSpace);
    void SetPoolOwner(BlockPool *pPool);
    BlockPool *GetPoolOwner();
    /* Returns a pooled block to its initial state: no control signals, record count 0. */
    void ResetForReuse();
    /* Makes the block a copy-free alias target: called when an output block is recycled so a
     * later writer does not trigger a pointless host->device population. */
    void MarkDeviceOnly(UINT uiMemSpace);
    /* Allocates the block's storage on pAcc now, leaving every view's coherence state as it is
     * (pre-allocated pool blocks); FALSE when the allocation failed. */
    BOOL ReserveDeviceBuffer(Accelerator *pAcc);
    /* Bytes of host and device storage the block holds. */
    size_t GetFootprintBytes();
    /* A new block holding a copy of this one (contents, record count, metadata, control
     * signals), for a consumer that writes its input in place while other consumers still read
     * this block.  The copy is made where the current data lives, a device first: device to
     * device on that accelerator's transfer stream (ordered after the block's pending writes),
     * else on the host.
     * NULL when the storage could not be allocated. */
    Datablock *CloneForWriter();

protected:
    void OnLastRelease() override;

private:
    BOOL EnsureHostBuffer();
    void SyncHostFromDevice(UINT uiMemSpace);

    DatablockTemplate *m_pTemplate;
    UINT m_cbLogical;
    UINT m_cbAllocated;
    UINT m_uiRecordCount;
    CONTROLSIGNAL m_luiControl;
    BUFFERACCESSFLAGS m_eFlags;
    BOOL m_bSealed;
    UINT m_uiDBUID;
    unsigned char *m_pHost;
    std::vector<unsigned char> m_vMeta;
    PBuffer *m_pBuffers[PTK_MAX_MEMORY_SPACES];
    BUFFER_COHERENCE_STATE m_eState[PTK_MAX_MEMORY_SPACES];
    BlockPool *m_pPoolOwner;
};

}  // namespace PTask

#endif
