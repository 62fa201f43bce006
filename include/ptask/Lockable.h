/* Lockable.h -- recursive locks that know their owner (reference: ptask/Lockable.h:8-40). */
#ifndef PTK_LOCKABLE_H
#define PTK_LOCKABLE_H

#include <atomic>
#include <mutex>
#include <thread>

#include "ptk_win32.h"

namespace PTask {

class Lockable {
public:
    Lockable() : m_nDepth(0) {}
    virtual ~Lockable() {}
    int Lock() {
        m_mutex.lock();
        m_owner.store(std::this_thread::get_id());
        return ++m_nDepth;
    }
    int Unlock() {
        int d = --m_nDepth;
        if (d == 0) m_owner.store(std::thread::id());
        m_mutex.unlock();
        return d;
    }
    BOOL LockIsHeld() const { return m_owner.load() == std::this_thread::get_id(); }
    int GetLockDepth() const { return LockIsHeld() ? m_nDepth : 0; }

private:
    std::recursive_mutex m_mutex;
    std::atomic<std::thread::id> m_owner;
    int m_nDepth;
};

}  // namespace PTask

#endif
