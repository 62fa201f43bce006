// ptk_copy.cpp -- the reference point for HBM traffic counters and the measured copy bandwidth:
// PtkCopyWords (ptk_util.hsaco) copies --mib MiB from one device buffer to another --repeat
// times.  A copy of B bytes moves exactly B bytes in and B bytes out, so the FETCH_SIZE and
// WRITE_SIZE counters of these dispatches give the factor that turns the counters of any other
// kernel into bytes.  Prints one JSON object with the copy bandwidth (read + write bytes / time).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptk_hip.h"

#define CHK(call)                                                                          \
    do {                                                                                   \
        if ((call) != PTK_HIP_OK) {                                                        \
            fprintf(stderr, "%s failed: %s\n", #call, ptk_hip_last_error_string());        \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

int main(int argc, char **argv) {
    std::string kernels = "ptask_amd/lib/kernels";
    long long mib = 1024;
    int repeat = 20;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--kernels" && i + 1 < argc) kernels = argv[++i];
        else if (a == "--mib" && i + 1 < argc) mib = atoll(argv[++i]);
        else if (a == "--repeat" && i + 1 < argc) repeat = atoi(argv[++i]);
        else {
            fprintf(stderr, "usage: ptk_copy [--kernels DIR] [--mib M] [--repeat R]\n");
            return 2;
        }
    }
    if (mib < 1 || repeat < 1) return 2;
    const size_t bytes = (size_t)mib << 20;
    ptk_hip_module m;
    ptk_hip_function f, fill;
    CHK(ptk_hip_module_load(&m, (kernels + "/ptk_util.hsaco").c_str()));
    CHK(ptk_hip_module_get_function(&f, m, "PtkCopyWords"));
    CHK(ptk_hip_module_get_function(&fill, m, "PtkFillWords"));
    void *src = nullptr, *dst = nullptr;
    CHK(ptk_hip_malloc(&src, bytes));
    CHK(ptk_hip_malloc(&dst, bytes));
    ptk_hip_stream s;
    ptk_hip_event e0, e1;
    CHK(ptk_hip_stream_create(&s));
    CHK(ptk_hip_event_create(&e0, 1));
    CHK(ptk_hip_event_create(&e1, 1));
    struct {
        void *d;
        unsigned long long n;
        unsigned v;
    } fa = {src, bytes / 4, 0x3f800000u};
    unsigned char fbuf[24] = {0};
    memcpy(fbuf, &fa.d, 8);
    memcpy(fbuf + 8, &fa.n, 8);
    memcpy(fbuf + 16, &fa.v, 4);
    CHK(ptk_hip_launch(fill, 2048, 1, 1, 256, 1, 1, 0, s, fbuf, 20));
    unsigned char cbuf[24];
    const unsigned long long n16 = bytes / 16;
    memcpy(cbuf, &dst, 8);
    memcpy(cbuf + 8, &src, 8);
    memcpy(cbuf + 16, &n16, 8);
    CHK(ptk_hip_launch(f, 4096, 1, 1, 256, 1, 1, 0, s, cbuf, 24));  // warm
    CHK(ptk_hip_event_record(e0, s));
    for (int r = 0; r < repeat; ++r) CHK(ptk_hip_launch(f, 4096, 1, 1, 256, 1, 1, 0, s, cbuf, 24));
    CHK(ptk_hip_event_record(e1, s));
    CHK(ptk_hip_event_synchronize(e1));
    float ms = 0;
    CHK(ptk_hip_event_elapsed_ms(&ms, e0, e1));
    const double avg = ms / repeat;
    printf("{\"kernel\": \"PtkCopyWords\", \"bytes_copied\": %zu, \"launches_timed\": %d, \"avg_ms\": %.5f, "
           "\"read_plus_write_gbs\": %.1f}\n",
           bytes, repeat, avg, 2.0 * bytes / (avg / 1e3) / 1e9);
    ptk_hip_free(src);
    ptk_hip_free(dst);
    ptk_hip_module_unload(m);
    return 0;
}
