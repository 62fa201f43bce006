/* ReferenceCounted.h -- AddRef/Release ownership (reference: ptask/ReferenceCounted.h:66-76).
 * Objects start at zero references; factories that hand an object to the caller return it with
 * one reference the caller owns. */
#ifndef PTK_REFERENCECOUNTED_H
#define PTK_REFERENCECOUNTED_H

#include <atomic>

#include "ptk_win32.h"

namespace PTask {

class ReferenceCounted {
public:
    ReferenceCounted() : m_uiRefCount(0) {}
    virtual ~ReferenceCounted() {}
    virtual LONG AddRef() { return ++m_uiRefCount; }
    virtual LONG Release() {
        LONG n = --m_uiRefCount;
        if (n == 0) OnLastRelease();
        return n;
    }
    LONG RefCount() const { return m_uiRefCount.load(); }

protected:
    /* Called once the count reaches zero; pooled objects override this to return home. */
    virtual void OnLastRelease() { delete this; }
    std::atomic<LONG> m_uiRefCount;
};

}  // namespace PTask

#endif
