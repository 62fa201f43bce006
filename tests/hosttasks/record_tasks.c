/* Host tasks for the multi-graph scheduling tests (ptk_multigraph): the runtime calls
 * f(nArgs, args) with buffers as pointers and INT / FLOAT parameters as their bits in the
 * pointer slot.  Built by tests/test_runtime_multigraph.py with gcc -shared -fPIC. */
#include <pthread.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#define LOG_MAX 4096
static int g_log[LOG_MAX];
static int g_n;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

/* RecordTask(int* tag, int* tagOut): log the tag in dispatch order, hold the dispatch ~2 ms so
 * that waiting dispatches queue behind it, copy the tag out. */
void RecordTask(unsigned int n, void **a) {
    (void)n;
    const int tag = *(const int *)a[0];
    pthread_mutex_lock(&g_lock);
    if (g_n < LOG_MAX) g_log[g_n++] = tag;
    pthread_mutex_unlock(&g_lock);
    struct timespec ts = {0, 2000000};
    nanosleep(&ts, NULL);
    *(int *)a[1] = tag;
}

/* The log so far: copies at most max entries, returns how many there are. */
int RecordLog(int *out, int max) {
    pthread_mutex_lock(&g_lock);
    const int n = g_n < max ? g_n : max;
    memcpy(out, g_log, (size_t)n * sizeof(int));
    pthread_mutex_unlock(&g_lock);
    return n;
}

/* ScaleVector(float* A, float s, int n, float* out): out = s * A. */
void ScaleVector(unsigned int nArgs, void **a) {
    (void)nArgs;
    const float *A = (const float *)a[0];
    float s;
    const uintptr_t bits = (uintptr_t)a[1];
    memcpy(&s, &bits, sizeof(s));
    const int n = (int)(intptr_t)a[2];
    float *out = (float *)a[3];
    for (int i = 0; i < n; ++i) out[i] = s * A[i];
}

/* ScaleInPlace(float* A, float s, VECADD_PARAMS n): A[i] *= s for i < n, in place -- the vector
 * task of ptaskUnitTest/iteration.cpp run_simple_iteration_task (an in/out port, a sticky
 * scalar, the element count from a port derived from A). */
void ScaleInPlace(unsigned int nArgs, void **a) {
    (void)nArgs;
    float *A = (float *)a[0];
    float s;
    const uintptr_t bits = (uintptr_t)a[1];
    memcpy(&s, &bits, sizeof(s));
    const int n = (int)(intptr_t)a[2];
    for (int i = 0; i < n; ++i) A[i] = s * A[i];
}

/* AddVectors(float* A, float* B, int n, float* out): out = A + B. */
void AddVectors(unsigned int nArgs, void **a) {
    (void)nArgs;
    const float *A = (const float *)a[0];
    const float *B = (const float *)a[1];
    const int n = (int)(intptr_t)a[2];
    float *out = (float *)a[3];
    for (int i = 0; i < n; ++i) out[i] = A[i] + B[i];
}
