/* datablocktemplate.h -- the shape and parameter kind of the blocks a port carries
 * (reference: ptask/datablocktemplate.h; created by Runtime::GetDatablockTemplate,
 * ptask/PTaskRuntime.h:392-429). */
#ifndef PTK_DATABLOCKTEMPLATE_H
#define PTK_DATABLOCKTEMPLATE_H

#include <string>
#include <vector>

#include "Lockable.h"
#include "ReferenceCounted.h"
#include "primitive_types.h"

namespace PTask {

class DatablockTemplate : public ReferenceCounted, public Lockable {
public:
    DatablockTemplate(const char *lpszName, UINT uiStride, UINT x, UINT y, UINT z, UINT uiPitch,
                      bool bRecordStream, bool bByteAddressable);
    DatablockTemplate(const char *lpszName, const BUFFERDIMENSIONS *pDims, UINT nDims,
                      bool bRecordStream, bool bByteAddressable);
    DatablockTemplate(const char *lpszName, UINT uiStride, PTASK_PARM_TYPE eType);
    virtual ~DatablockTemplate();

    char *GetTemplateName();
    UINT GetStride(UINT uiChannel = DBDATA_IDX);
    UINT GetPitch(UINT uiChannel = DBDATA_IDX);
    UINT GetXElementCount(UINT uiChannel = DBDATA_IDX);
    UINT GetYElementCount(UINT uiChannel = DBDATA_IDX);
    UINT GetZElementCount(UINT uiChannel = DBDATA_IDX);
    UINT GetDatablockByteCount(UINT uiChannel = DBDATA_IDX);
    BUFFERDIMENSIONS GetBufferDimensions(UINT uiChannel = DBDATA_IDX);
    PTASK_PARM_TYPE GetParameterBaseType();
    BOOL IsByteAddressable();
    BOOL DescribesRecordStream();
    BOOL IsVariableDimensioned();
    BOOL IsScalarParameter();  // INT, FLOAT or BYVALSTRUCT: bound by value, never a buffer

    void SetInitialValue(void *lpvInitData, UINT cbData, UINT nRecordCount,
                         BOOL bExplicitlyEmpty = FALSE);
    const void *GetInitialValue();
    UINT GetInitialValueSizeBytes();
    UINT GetInitialValueElements();
    BOOL HasInitialValue();
    BOOL IsInitialValueEmpty();

private:
    std::string m_strName;
    BUFFERDIMENSIONS m_dims[NUM_DATABLOCK_CHANNELS];
    PTASK_PARM_TYPE m_eParamType;
    bool m_bRecordStream;
    bool m_bByteAddressable;
    std::vector<unsigned char> m_vInitialValue;
    UINT m_uiInitialElements;
    BOOL m_bHasInitialValue;
    BOOL m_bExplicitlyEmpty;
};

}  // namespace PTask

#endif
