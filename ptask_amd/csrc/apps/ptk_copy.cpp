// ptk_copy.cpp -- the reference point for HBM traffic counters and the measured copy bandwidth:
// PtkCopyWords (ptk_util.hsaco) copies --mib MiB from one device buffer to another --repeat
// times.  A copy of B bytes moves exactly B bytes in and B bytes out, so the FETCH_SIZE and
// WRITE_SIZE counters of these dispatches give the factor that turns the counters of any other
// kernel into bytes.  Prints one JSON object with the copy bandwidth (read + write bytes / time).
// --mode mix times PtkStreamMix instead (six --mib arrays read, three written: the half-step FDTD
// access mix without its stencil) and --mode read PtkReadWords (reads only); --grid sets the
// launch's block count.

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

// PtkStreamMix / PtkReadWords over arrays of `bytes` each.
static int Probe(const std::string &kernels, const std::string &mode, size_t bytes, int repeat, int grid) {
    ptk_hip_module m;
    ptk_hip_function f, fill;
    CHK(ptk_hip_module_load(&m, (kernels + "/ptk_util.hsaco").c_str()));
    CHK(ptk_hip_module_get_function(&f, m, mode == "mix" ? "PtkStreamMix" : "PtkReadWords"));
    CHK(ptk_hip_module_get_function(&fill, m, "PtkFillWords"));
    const int narr = mode == "mix" ? 9 : 1;
    std::vector<void *> buf(narr + 1, nullptr);
    ptk_hip_stream s;
    ptk_hip_event e0, e1;
    CHK(ptk_hip_stream_create(&s));
    CHK(ptk_hip_event_create(&e0, 1));
    CHK(ptk_hip_event_create(&e1, 1));
    for (int a = 0; a < narr; ++a) {
        CHK(ptk_hip_malloc(&buf[a], bytes));
        unsigned char fbuf[24] = {0};
        const unsigned long long n = bytes / 4;
        const unsigned v = 0x3f800000u;
        memcpy(fbuf, &buf[a], 8);
        memcpy(fbuf + 8, &n, 8);
        memcpy(fbuf + 16, &v, 4);
        CHK(ptk_hip_launch(fill, 2048, 1, 1, 256, 1, 1, 0, s, fbuf, 20));
    }
    CHK(ptk_hip_malloc(&buf[narr], 256));  // PtkReadWords' sink
    const unsigned long long n16 = bytes / 16;
    std::vector<unsigned char> args;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        args.insert(args.end(), c, c + n);
    };
    if (mode == "mix") {
        for (int a = 0; a < 9; ++a) put(&buf[a], 8);
        put(&n16, 8);
    } else {
        put(&buf[0], 8);
        put(&n16, 8);
        put(&buf[1], 8);
    }
    CHK(ptk_hip_launch(f, grid, 1, 1, 256, 1, 1, 0, s, args.data(), args.size()));  // warm
    CHK(ptk_hip_event_record(e0, s));
    for (int r = 0; r < repeat; ++r) CHK(ptk_hip_launch(f, grid, 1, 1, 256, 1, 1, 0, s, args.data(), args.size()));
    CHK(ptk_hip_event_record(e1, s));
    CHK(ptk_hip_event_synchronize(e1));
    float ms = 0;
    CHK(ptk_hip_event_elapsed_ms(&ms, e0, e1));
    const double avg = ms / repeat, moved = (double)bytes * narr;
    printf("{\"kernel\": \"%s\", \"bytes_moved_per_launch\": %.0f, \"grid\": %d, \"launches_timed\": %d, "
           "\"avg_ms\": %.5f, \"gbs\": %.1f}\n",
           mode == "mix" ? "PtkStreamMix" : "PtkReadWords", moved, grid, repeat, avg, moved / (avg / 1e3) / 1e9);
    for (void *p : buf) ptk_hip_free(p);
    ptk_hip_module_unload(m);
    return 0;
}

int main(int argc, char **argv) {
    std::string kernels = "ptask_amd/lib/kernels";
    long long mib = 1024;
    int repeat = 20, grid = 4096;
    std::string mode = "copy";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--kernels" && i + 1 < argc) kernels = argv[++i];
        else if (a == "--mib" && i + 1 < argc) mib = atoll(argv[++i]);
        else if (a == "--repeat" && i + 1 < argc) repeat = atoi(argv[++i]);
        else if (a == "--mode" && i + 1 < argc) mode = argv[++i];
        else if (a == "--grid" && i + 1 < argc) grid = atoi(argv[++i]);
        else {
            fprintf(stderr, "usage: ptk_copy [--kernels DIR] [--mib M] [--repeat R] [--mode copy|mix|read] [--grid G]\n");
            return 2;
        }
    }
    if (mib < 1 || repeat < 1 || grid < 1 || (mode != "copy" && mode != "mix" && mode != "read")) return 2;
    if (mode != "copy") return Probe(kernels, mode, (size_t)mib << 20, repeat, grid);
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
    CHK(ptk_hip_launch(f, grid, 1, 1, 256, 1, 1, 0, s, cbuf, 24));  // warm
    CHK(ptk_hip_event_record(e0, s));
    for (int r = 0; r < repeat; ++r) CHK(ptk_hip_launch(f, grid, 1, 1, 256, 1, 1, 0, s, cbuf, 24));
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
