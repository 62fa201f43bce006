/*
 * ptk_batch.h -- the batched / served executor: many graph instances of uneven size run as one
 * hipGraph-captured iteration of persistent multi-instance task kernels, or, as the baseline,
 * by plain per-task dispatch (one kernel launch per task per instance).
 *
 * A graph instance is one FDTD graph (the six Yee tasks per step, ptaskUnitTest/graphfdtd.cpp),
 * one groupby graph (clear/count/offsets/reshuffle, ptaskUnitTest/graphgroupby.cpp:106-387) or
 * one EM-GMM graph (E+M per iteration, ptaskUnitTest/graphemgmm.cpp:124-211).  There is no
 * counterpart of the executor in the reference: it is the batching and serving layer of this
 * build, and its instances compute exactly what the single graphs compute (bitwise for FDTD,
 * for every groupby output, and for the GMM model).
 *
 * Lifecycle: submit (queued, inputs described by seeds) -> admitted at an iteration boundary
 * when its HBM footprint fits the arena (inputs generated on the device, datablocks resident
 * from then on) -> one step per iteration -> complete (checksums taken, memory returned unless
 * the instance was submitted with keep = 1).  All functions return 0 on success or a
 * PTK_BATCH_E* code; they are not thread-safe (one host thread drives an executor).
 */
#ifndef PTK_BATCH_H
#define PTK_BATCH_H

#include "ptk_fdtd.h"
#include "ptk_gmm.h"
#include "ptk_groupby.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PTK_BATCH_OK = 0,
    PTK_BATCH_EINVAL = 1,   /* bad argument or shape */
    PTK_BATCH_ENOMEM = 2,   /* cannot ever fit the arena / host allocation failed */
    PTK_BATCH_EHIP = 3,     /* a HIP call failed (ptk_batch_last_error has the text) */
    PTK_BATCH_ESTATE = 4,   /* wrong instance state for the call */
};

enum { PTK_BATCH_PLAIN = 0, PTK_BATCH_BATCHED = 1 };              /* execution modes */
enum { PTK_INST_FDTD = 0, PTK_INST_GROUPBY = 1, PTK_INST_GMM = 2 };
enum { PTK_STATE_QUEUED = 0, PTK_STATE_ACTIVE = 1, PTK_STATE_DONE = 2, PTK_STATE_RELEASED = 3 };

/* Device-side instance descriptors (the batched kernels read arrays of these). */
typedef struct _ptk_fdtd_inst {
    float *f[2][6];       /* field sets A and B: Hx, Hy, Hz, Ex, Ey, Ez */
    FDTD_PARAMS p;
    int steps, done;      /* target steps; launches completed (advanced on the device): each
                             launch, of one or three steps, writes the other set, so the current
                             set is done & 1 */
} PTK_FDTD_INST;

typedef struct _ptk_groupby_inst {
    const int *keys, *vals;
    long long n;          /* rows */
    long long *table, *offsets, *groups;
    int *outKeys, *outVals;
    void *scratch;        /* ptk_gb_scratch_bytes(n, numKeys) (include/ptk_groupby.h) */
    int numKeys, chunkRows;  /* chunkRows: rows per accumulate / reshuffle work unit, a
                                multiple of PTK_GB_CHUNK */
} PTK_GROUPBY_INST;

typedef struct _ptk_gmm_inst {
    const float *X;
    double *model, *partials;
    GMM_DIMS d;
} PTK_GMM_INST;

/* One checksum of a retiring instance's result array (BatchChecksumJobs): mode 0 the words of
 * a[0..n), mode 1 the (key, value) pairs (a[i], b[i]), mode 2 a groupby groups block (its
 * length read from its header); the sum is added to *out. */
typedef struct _ptk_csum_job {
    const void *a, *b;
    long long n;
    unsigned long long *out;
    int mode, pad;
} PTK_CSUM_JOB;

typedef struct ptk_batch ptk_batch;

/* kernel_dir holds batch.hsaco, fdtd.hsaco, groupby.hsaco, gmm.hsaco; arena_bytes of HBM are
 * reserved for instance datablocks. */
int ptk_batch_create(const char *kernel_dir, unsigned long long arena_bytes, int mode, ptk_batch **out);
int ptk_batch_destroy(ptk_batch *b);
const char *ptk_batch_last_error(void);

/* Queue an instance; *id receives its handle.  Inputs are generated on the device at admission:
 * FDTD fields U[-1,1) (the FdtdInitField generator), groupby keys Zipf(s) over num_keys+1 keys
 * (or uniform when zipf_s <= 0) and int32 values, GMM points from a seeded mixture (means
 * U[-10,10], sd U[0.5,2]) with one initial mean per generating component. */
int ptk_batch_submit_fdtd(ptk_batch *b, int nx, int ny, int nz, int steps, unsigned long long seed, int keep, int *id);
int ptk_batch_submit_groupby(ptk_batch *b, long long rows, int num_keys, double zipf_s, unsigned long long seed,
                             int keep, int *id);
int ptk_batch_submit_gmm(ptk_batch *b, int n, int d, int k, int steps, unsigned long long seed, int keep, int *id);
/* Bytes of arena an instance occupies while resident. */
unsigned long long ptk_batch_footprint(int kind, long long a, int b_, int c, int steps);

/* One iteration: admit queued instances in submission order while they fit (at most max_admit,
 * < 0 = no limit), run one step of every active instance, retire completed ones.  *active and
 * *completed (either may be NULL) receive the counts after the iteration. */
int ptk_batch_tick(ptk_batch *b, int max_admit, int *active, int *completed);
/* As ptk_batch_tick, but runs up to max_iters iterations back to back without waiting between
 * them: as many as every resident instance still needs before the first of them completes (one
 * when a groupby instance is resident, since it completes in one).  *iters receives the number
 * run.  Results are the same as with single iterations. */
int ptk_batch_run(ptk_batch *b, int max_admit, int max_iters, int *iters, int *active, int *completed);
/* Ticks until every submitted instance is complete. */
int ptk_batch_drain(ptk_batch *b);

/* state: PTK_STATE_*; times in ms since ptk_batch_create (submit, admit, done; -1 if not yet). */
int ptk_batch_status(ptk_batch *b, int id, int *state, double *t_submit, double *t_admit, double *t_done);
/* Checksums of a completed instance, taken on the device at completion:
 *   FDTD: one per field (bit patterns, position-weighted), [6] unused;
 *   groupby: [0] groups (header and the G triples), [1] reshuffled keys, [2] reshuffled values (both
 *            position-weighted), [3] multiset of (key, value) rows of the input, [4] multiset of (key, value) rows of
 *            the output (equal to [3] when every key is valid);
 *   GMM: [0] model block. */
int ptk_batch_checksums(ptk_batch *b, int id, unsigned long long out[8]);
/* Copy a result array of a kept instance to the host.  FDTD: which = field 0..5 (current set);
 * groupby: 0 keys, 1 vals, 2 groups, 3 offsets, 4 outKeys, 5 outVals; GMM: 0 X, 1 model.
 * *bytes receives the array size; host may be NULL to query it. */
int ptk_batch_read(ptk_batch *b, int id, int which, void *host, unsigned long long *bytes);
int ptk_batch_release(ptk_batch *b, int id);

/* Tile shapes of the batched kernels: FDTD columns per tile and k planes per tile, groupby
 * rows per work unit (a multiple of PTK_GB_CHUNK).  fdtd_cols == 0 (the default, with 8 planes)
 * selects the whole-step kernel on overlapped 62 x 6-column tiles instead of the two half-step
 * kernels; instances with (nx+1)(ny+1)(nz+1) >= 2^29 use the half steps.  ptk_batch_autotune
 * times candidates of both forms on the active instances (the timed launches are idempotent, so
 * results are unchanged) and keeps the fastest. */
int ptk_batch_set_tiles(ptk_batch *b, int fdtd_cols, int fdtd_depth, int gb_rows);
int ptk_batch_get_tiles(ptk_batch *b, int *fdtd_cols, int *fdtd_depth, int *gb_rows);
int ptk_batch_autotune(ptk_batch *b, double *best_ms);

/* The autotuner's choice as five integers {fdtd_cols, fdtd_depth, fdtd_grid, fdtd_steps_per_launch,
 * fdtd_blocks3} (tile shape, persistent grid of the one-step kernels, steps per launch, virtual
 * blocks of the three-step schedule), so a tuning run's decision can be applied to later executors
 * of the same size class without tuning inside their timed work.  set validates every field. */
int ptk_batch_get_tuning(ptk_batch *b, int out[5]);
int ptk_batch_set_tuning(ptk_batch *b, const int in[5]);

/* FDTD time steps per batched launch with the whole-step form (fdtd_cols == 0): 3 (the default)
 * advances every resident FDTD instance three steps per launch on overlapped 58 x 10-column tiles
 * (BatchFdtdSteps3) while each has at least three steps left, then one step per launch for the
 * rest; 1 keeps one step per launch.  Results are bitwise those of single steps.  No
 * counterpart in the reference, whose graphs run one task per step (graphfdtd.cpp:735-739).
 * The three-step launches follow a persistent schedule over *blocks virtual blocks (by default
 * one per compute unit): each takes whole-depth tiles round-robin, then an equal run of the
 * remaining tiles' planes.  ptk_batch_autotune also compares both, per step. */
int ptk_batch_set_fdtd_steps(ptk_batch *b, int steps_per_launch);
int ptk_batch_get_fdtd_steps(ptk_batch *b, int *steps_per_launch, int *blocks);

/* Launch form of batched iterations: 1 (default) one captured hipGraph per mix of instance kinds,
 * replayed every iteration; 0 the same kernels launched directly with grids cut to the current
 * work-list lengths.  Results are identical; the cheaper form depends on how much work an
 * iteration holds (a graph replay costs more than a few direct launches on a nearly idle device). */
int ptk_batch_set_launch(ptk_batch *b, int graphs);

/* Counters: [0] iterations run, [1] kernel launches issued, [2] graph launches, [3] arena bytes in
 * use, [4] peak, [5] arena size, [6] microseconds spent executing iterations (from the first
 * launch of an iteration's work to its completion; admission and retirement excluded). */
int ptk_batch_stats(ptk_batch *b, unsigned long long out[7]);

#ifdef __cplusplus
}
#endif

#endif /* PTK_BATCH_H */
