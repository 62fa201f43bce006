/* hrperft.h -- a high-resolution wall-clock timer (reference: ptask/hrperft.h). */
#ifndef PTK_HRPERFT_H
#define PTK_HRPERFT_H

#include <chrono>

typedef enum gran_t { gran_nanosec, gran_usec, gran_msec, gran_sec } hpf_granularity;

class CHighResolutionTimer {
public:
    explicit CHighResolutionTimer(hpf_granularity g = gran_msec) : m_gran(g) { reset(); }
    void reset() { m_start = std::chrono::steady_clock::now(); }
    /* Time since reset() in the timer's unit; bReset restarts the clock. */
    double elapsed(bool bReset = false) {
        auto now = std::chrono::steady_clock::now();
        double ns = (double)std::chrono::duration_cast<std::chrono::nanoseconds>(now - m_start).count();
        if (bReset) m_start = now;
        switch (m_gran) {
            case gran_nanosec: return ns;
            case gran_usec: return ns / 1e3;
            case gran_msec: return ns / 1e6;
            default: return ns / 1e9;
        }
    }

private:
    hpf_granularity m_gran;
    std::chrono::steady_clock::time_point m_start;
};

#endif
