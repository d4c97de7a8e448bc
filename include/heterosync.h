/*
 * heterosync.h -- C ABI of the MI355X (gfx950) synchronization hot path.
 *
 * One shared library, libheterosync_amd.so, built by hipcc.  Plain C types
 * only: host pointers and sizes in, results out.  Every function returns
 * HS_OK (0) or a negative HS_E* code; hs_strerror() gives the text.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   hs_open / hs_close         hipLocksInit / hipLocksDestroy
 *                              (hip/syncPrims/uvm/hipLocksImpl.h:3-93, :95-108)
 *                              plus the device query of
 *                              hip/syncPrims/uvm/main.hip.cpp:1995-2011
 *   hs_prim_id / hs_prim_name  the string->id table main.hip.cpp:2036-2089
 *   hs_storage_words           numLocsMult/numStorageLocs main.hip.cpp:2109-2136
 *   hs_syncprim_expected       the host golden and its compare,
 *                              main.hip.cpp:2524-2702
 *   hs_syncprim_run            hipMutexCreate* / hipSemaphoreCreate* +
 *                              the ROI and invoke* launchers
 *                              main.hip.cpp:1423-1927, :2354-2511
 *   hs_seqlock_run             hip/relaxedAtomics/uvm/seqlocks/seqlocks.cpp:6-112
 *                              (+ seqlocks_hipAtomics/seqlocks.cpp)
 *   hs_flags_run               cuda/relaxedAtomics/uvm/flags/flags.cu:19-176
 *   hs_refcounter_run          cuda/relaxedAtomics/uvm/refCounter/refCounter.cu:5-155
 *   hs_splitcounter_run        cuda/relaxedAtomics/uvm/splitCounter/splitCounter.cu:5-93
 *   hs_uts_ref_run             cuda/uts/uvm/uts.cu:244-455 (callGPU/doWork)
 *   hs_uts_*_run_cpu           the same with host worker threads taking part
 *                              (uts.cu:457-539's pthreads, README.txt:30)
 *   hs_uts_geo_run             no reference implementation (SHA-1 UTS,
 *                              the public UTS algorithm; see hs_uts_geo_run)
 *   hs_batch_run, hs_uts_serve extensions with no reference counterpart:
 *                              batched persistent syncPrims and the UTS
 *                              serving loop.
 */
#ifndef HETEROSYNC_H
#define HETEROSYNC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HS_OK = 0,
  HS_EINVAL = -1,      /* bad argument / unsupported shape (usage error) */
  HS_EHIP = -2,        /* a HIP runtime call failed */
  HS_ENOMEM = -3,      /* device allocation failed */
  HS_EOCCUPANCY = -4,  /* workgroups cannot all be co-resident */
  HS_ETIMEOUT = -5,    /* device watchdog fired: a wait exceeded timeout_s */
  HS_ENODEV = -6,      /* no HIP device */
  HS_EUNSUPPORTED = -7, /* primitive id without a kernel in this library (every
                          id of the reference table has one; kept for ABI
                          stability) */
  HS_ESCRATCH = -8     /* hs_syncprim_run_noalloc: the run's state does not
                          fit the 4 MiB scratch */
};

const char *hs_strerror(int code);

typedef struct hs_ctx hs_ctx;

/* Opens device `device` (hipSetDevice) and queries it. */
int hs_open(int device, hs_ctx **out);
void hs_close(hs_ctx *ctx);

/* NUM_CU, maxThreadsPerBlock, MAX_WGS = (maxThreadsPerBlock/64)*NUM_CU
 * (main.hip.cpp:2004-2008), gcnArchName. */
int hs_device_info(hs_ctx *ctx, int *num_cu, int *max_threads_per_block,
                   int *max_wgs, char *arch, int arch_len);

/* ---- syncPrims -------------------------------------------------------- */

/* Primitive ids follow main.hip.cpp:2036-2089 exactly (0..42, 36 and
 * 43-46 reserved).  Returns -1 for unknown names. */
int hs_prim_id(const char *name);
const char *hs_prim_name(int id);
/* 1 if this library implements the id (0-35 and 37-42; 36 is reserved in the
 * reference table), else 0. */
int hs_prim_supported(int id);

/* Number of int32 storage words a run uses (main.hip.cpp:2109-2136 with
 * NUM_WIS_PER_WG = wg_size).  Negative on error. */
long hs_storage_words(int prim, int num_wgs, int num_cu, int wg_size,
                      int num_ldst);

typedef struct {
  int prim;          /* hs_prim_id() value */
  int num_ldst;      /* argv[2] of allSyncPrims-1kernel */
  int num_wgs;       /* argv[3] */
  int num_cs_iters;  /* argv[4] */
  int wg_size;       /* threads per WG; 0 = 64 (main.hip.cpp:9) */
  int reps;          /* launches per run; 0 = 10 (main.hip.cpp:21) */
  int mode;          /* 0 = copy chain (reference), 1 = in-place increment */
  int backoff_min;   /* EBO backoff start, s_sleep quanta; 0 = 1 */
  int backoff_max;   /* EBO cap mask (2^k-1); 0 = 1023 (MAX_BACKOFF-1) */
  int poll_sleep;    /* quanta between plain polls; <0 = 0, 0 = default
                        (1, or 0 for the XCD-aware tree barriers) */
  double timeout_s;  /* per-launch watchdog; 0 = 20 s */
  int sync_scope;    /* 0 = XCD-aware barriers where the device has several
                        XCDs (one cache writeback/invalidate per XCD and
                        episode), 1 = agent-scope release/acquire by every
                        WG (strict HIP memory model everywhere) */
} hs_syncprim_cfg;

typedef struct {
  double roi_ms;        /* region of interest: all launches + syncs */
  double kernel_ms;     /* sum of per-launch kernel times (HIP events) */
  int64_t cs_entries;   /* exclusive entries (mutex CS / semaphore writers) */
  int64_t rd_entries;   /* semaphore reader entries */
  int64_t episodes;     /* completed barrier episodes */
  int state_ok;         /* 1 if every lock/barrier word is back in its
                           idle state (see hs_syncprim_run) */
  uint32_t lock_words[8]; /* first lock block / barrier words, for tests */
  int sync_scope_used;  /* 0 = XCD-aware, 1 = strict */
} hs_syncprim_result;

/* Runs one microbenchmark exactly as the reference's ROI does: `reps`
 * launches of numWGs x wg_size, each followed by a device synchronize.
 * storage: host array of hs_storage_words() int32 words; copied to the
 * device before the ROI and back after it (not timed).  Lock and barrier
 * state is created per call, sized from the workload (no fixed worst-case
 * buffers).  Barriers and semaphores need every WG co-resident and return
 * HS_EOCCUPANCY otherwise. */
int hs_syncprim_run(hs_ctx *ctx, const hs_syncprim_cfg *cfg, int32_t *storage,
                    long n_words, hs_syncprim_result *res);

/* hs_syncprim_run without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_storage is the caller's DEVICE array of n_words =
 * hs_storage_words(...) words, read and written in place (initialise it to
 * storage[i] = i for the reference's workload).  Every other word of state
 * (lock blocks, the queue lock's ring, barrier counters and flags, control
 * words) is carved from one static 4 MiB device array per GPU, whose
 * address hs_open looks up; the call makes no hipMalloc, hipFree or event
 * and clears what it uses with a fill kernel.  Runs whose state exceeds
 * 4 MiB (long queue-lock rings, more than about 16 WGs per CU) return
 * HS_ESCRATCH.  All calls on one GPU share that array and keep the last
 * call's contents: calls that overlap on one GPU (other host threads or
 * contexts on the same device) race.  roi_ms and kernel_ms are host-clock
 * times of the launches and their synchronizes.  If the XCD-aware barriers'
 * placement check fails (they then run in the strict form in
 * hs_syncprim_run), this call cannot restart and returns HS_ETIMEOUT with
 * sync_scope_used = 2. */
int hs_syncprim_run_noalloc(hs_ctx *ctx, const hs_syncprim_cfg *cfg,
                            int32_t *d_storage, long n_words,
                            hs_syncprim_result *res);

/* The storage a correct run leaves: the reference binary's own self-check
 * (the host golden of main.hip.cpp:2524-2702), in closed form.  Fills
 * out[0..n_words) for a run of hs_syncprim_run with these arguments and
 * storage initialised to storage[i] = i (main.hip.cpp:2141).  Host code
 * only; needs no device. */
int hs_syncprim_expected(int prim, int num_ldst, int num_wgs, int num_cs_iters,
                         int reps, int wg_size, int num_cu, int mode,
                         int32_t *out, long n_words);

/* ---- relaxed atomics ---------------------------------------------------- */

typedef struct {
  double kernel_ms;
  uint64_t atomic_ops;  /* device-scope atomic RMWs/loads/stores issued on
                           sync words (algorithmic count) */
  int aborted;
} hs_relaxed_result;

/* seqlock: writer WGs are b % writer_period == 0 (reference: 16), lock
 * index b % group_size, `iters` passes per WG (reference: 10).  variant bit
 * 0: 0 = scoped acquire/release atomics (seqlocks_hipAtomics), 1 = relaxed
 * atomics plus fences with exchanges as loads (seqlocks, useTFs).  Variant
 * bit 1 (variants 2, 3) keeps the reference's fences and acquire/release
 * orderings; without it the same orderings come from drained agent-scope
 * accesses with no cache fences (the default form).  seq_out has
 * n_seq >= group_size words; out_arr has num_wgs*wg_size words;
 * torn_reads counts reader snapshots whose two data words differed
 * (variant 0 only; must be 0). */
int hs_seqlock_run(hs_ctx *ctx, int num_wgs, int wg_size, int group_size,
                   int writer_period, int iters, int variant,
                   uint32_t *seq_out, int n_seq, int32_t *out_arr,
                   uint64_t *torn_reads, hs_relaxed_result *res);

/* flags: the last TB is the main TB; workers loop num_iters times (and
 * until stop is seen), bumping `dirty` when rand[b + it*num_tbs] <= cutoff.
 * rand_arr has num_tbs*num_iters floats; out_arr num_tbs*tb_size words. */
int hs_flags_run(hs_ctx *ctx, int num_tbs, int tb_size, int num_accs_per_thr,
                 int num_repeats, float cutoff, int num_iters,
                 const float *rand_arr, uint32_t *out_arr, uint32_t *dirty,
                 hs_relaxed_result *res);

/* refCounter: num_repeats x {+1 on c0,c1; `spacing` LDS steps;
 * atomicDec(c, 1e9) on c0, c1; del = old <= 1}.  spacing < 0 = the
 * reference's 2*num_repeats.  strict = 1 keeps the fences around the
 * decrements (refCounter_kernel.cu:42-45); 0 orders them by drained
 * agent-scope atomics.  c0/c1 have num_tbs*tb_size/sharers words,
 * del0/del1 num_tbs*tb_size. */
int hs_refcounter_run(hs_ctx *ctx, int sharers, int num_tbs, int tb_size,
                      int num_repeats, int spacing, int strict, uint32_t *c0,
                      uint32_t *c1, uint32_t *del0, uint32_t *del1,
                      hs_relaxed_result *res);

/* splitCounter: even waves of a WG are writers (one atomicAdd(+1) per
 * iteration on their own counter); odd waves are readers.  The read is
 * WG-collective (splitCounter_kernel.cu:33-76): the first reader wave of
 * each WG makes it every `read_every` writer iterations (0 = min(64,
 * iters); capped at iters, so every WG reads at least once), summing the WG's counters into its per-WG partial and then every
 * partial into the approximate global total, which every reader thread
 * outputs.  iters / read_every reads per WG in all.  counters:
 * num_tbs*tb_size/2 words; out_arr num_tbs*tb_size words. */
int hs_splitcounter_run(hs_ctx *ctx, int num_tbs, int tb_size, int iters,
                        int read_every, uint32_t *counters, uint32_t *out_arr,
                        hs_relaxed_result *res);

/* ---- UTS ----------------------------------------------------------------- */

typedef struct {
  double kernel_ms;
  uint64_t nodes;       /* nodes processed (device count) */
  uint64_t leaves;
  uint64_t leaf_sum;    /* reference tree only */
  int max_depth;        /* geometric trees only */
  int aborted;
} hs_uts_result;

/* Reference binary tree (uts.cu:290-339 layout: value, child_a, child_b per
 * node; child_a == 0 means leaf).  Work-stealing over per-WG deques in
 * HBM; num_blocks WGs of 64 lanes (0 = 8 per CU, two waves per SIMD).
 * Each deque holds max(2^14, tree height + 4096) records. */
int hs_uts_ref_run(hs_ctx *ctx, int num_nodes, const uint32_t *value,
                   const int32_t *child_a, const int32_t *child_b,
                   int num_blocks, hs_uts_result *res);

/* One geometric (fixed shape) SHA-1 tree: root seed, expected branching
 * b0, depth limit gen_mx (UTS -t 1 -a 3 -b b0 -d gen_mx -r seed).
 * num_blocks 0 = 8 per CU, or every WG that can be resident (four waves
 * per SIMD) for trees expected above 2.5e8 nodes.  Deques hold 2^16
 * records; HS_EINVAL when gen_mx * (max_children - 1) exceeds that (the
 * deque bound, see hs_uts_serve: gen_mx <= 689 at b0 = 4, whose nodes
 * have at most 96 children).  The context keeps the engines (deques,
 * inboxes, state; about 1.5 MB per WG) of the two most recent grids
 * between calls; hs_close frees them. */
int hs_uts_geo_run(hs_ctx *ctx, int seed, double b0, int gen_mx,
                   int num_blocks, hs_uts_result *res);

/* Host worker threads taking part in one search: README.txt:30 promises
 * load balancing "across CPU and GPU threads"; the reference's pthreads only
 * meet at a barrier (cuda/uts/uvm/uts.cu:457-539).  Here each of
 * cpu_threads host threads has a mailbox in pinned, host-coherent memory
 * served by one GPU workgroup: when the thread asks, its WG hands it up to
 * 64 of its newest records (with their termination credits); the thread
 * expands their subtrees depth-first (the same SHA-1 children, threshold
 * counts or reference-tree arrays), then publishes its cumulative counts
 * and returned credits, which its WG folds into the tree's device counters.
 * The counts are the GPU-only run's exactly.  cpu_threads <= the grid's WGs.
 * *cpu (may be NULL) receives what the host workers did. */
typedef struct {
  uint64_t nodes;       /* nodes the host workers expanded */
  uint64_t deliveries;  /* record batches they received */
  double busy_s;        /* their summed expansion time */
  int threads;
} hs_uts_cpu_result;

int hs_uts_ref_run_cpu(hs_ctx *ctx, int num_nodes, const uint32_t *value,
                       const int32_t *child_a, const int32_t *child_b,
                       int num_blocks, int cpu_threads, hs_uts_result *res,
                       hs_uts_cpu_result *cpu);
int hs_uts_geo_run_cpu(hs_ctx *ctx, int seed, double b0, int gen_mx,
                       int num_blocks, int cpu_threads, hs_uts_result *res,
                       hs_uts_cpu_result *cpu);

/* ---- extensions ---------------------------------------------------------- */

typedef struct {
  int prim;       /* mutex or semaphore id (4-15) */
  int num_wgs;
  int num_cs_iters;
  int num_ldst;
} hs_batch_instance;

typedef struct {
  int wg_size;       /* lanes per persistent worker WG: 64, 128 or 256
                        (0 = 64).  Every virtual WG keeps the reference's
                        64 threads (main.hip.cpp:9); a wider worker splits
                        each critical section's lines over its waves. */
  int reps;          /* 0 = 10 */
  int mode;          /* 0 = copy chain (reference), 1 = in-place increment
                        (count-exact: a lost or doubled pass changes the
                        final storage) */
  int backoff_min;   /* EBO backoff start I_min, in skipped polls x 4;
                        0 = 1 (hipLocksMutexEBO.h:26) */
  int backoff_max;   /* EBO cap I_max (2^k-1); 0 = 1023 */
  int poll_sleep;    /* s_sleep quanta after a pass with no progress;
                        <0 = 0, 0 = 1 */
  int use_graph;     /* capture the launches in a hipGraph */
  int persistent_wgs; /* worker WGs of the persistent grid; 0 = 8 waves
                         per CU */
  double timeout_s;
  int probe;         /* attempt form: 0 = lane 0 tries each virtual WG in
                        turn (one round trip each), 1 = one load per lane
                        probes all of a worker's virtual WGs at once, then
                        lane 0 acts on the hopeful ones */
  int cohort;        /* spin and EBO mutexes and semaphore writers: most
                        critical sections per exclusive hold.  A worker that
                        holds an instance's lock (or a semaphore's whole
                        count) passes it to another of its virtual WGs of
                        that instance still waiting for it, without the
                        release and the next attempt (the same outcome as
                        that WG winning right after a release), up to
                        `cohort` critical sections, then releases.  Every
                        mutex instance's virtual WGs, and those of the
                        semaphores of count 1 and 2, are then placed in
                        blocks on workers instead of round robin (a ticket
                        or queue lock's next ticket is then usually on the
                        releasing worker, whose WG enters at once).  0 or
                        1 = off (every critical section releases). */
} hs_batch_cfg;

typedef struct {
  double roi_ms;
  double kernel_ms;
  int64_t acquires;   /* total acquire+release pairs (sum over instances) */
  int all_ok;         /* every instance's state/counters matched */
} hs_batch_result;

/* Runs n instances in one persistent launch per rep.  storages[i] is the
 * host array of instance i (hs_storage_words(prim,num_wgs,...) words);
 * entries[i] receives instance i's counted CS entries. */
int hs_batch_run(hs_ctx *ctx, const hs_batch_instance *inst, int n,
                 const hs_batch_cfg *cfg, int32_t *const *storages,
                 int64_t *entries, hs_batch_result *res);

/* Autotuner for hs_batch_run: times a calibration copy of the batch (CS
 * iterations capped at 20, one launch, in cfg->mode) over waves per CU
 * {2,4,8,16} x poll interval {0,1,4} x EBO cap I_max {63,255,1023} with
 * 64-lane workers, then, at the winning waves per CU, worker width
 * {64,128,256} x backoff start I_min {1,4,16} x attempt form (probe 0,
 * 1), and last the mutex cohort {off, 8, 64} at the winning knobs on the
 * batch itself (iterations capped at 2000), one launch each (a cohort's
 * holds grow with the sections each WG still owes, which a short copy does
 * not show).  Writes the
 * fastest wg_size / persistent_wgs / poll_sleep / backoff_min / backoff_max
 * / probe / cohort into *cfg (grids that cannot be co-resident are
 * skipped), and the chosen configuration's time for one launch of the
 * batch into *best_ms. */
int hs_batch_autotune(hs_ctx *ctx, const hs_batch_instance *inst, int n,
                      hs_batch_cfg *cfg, double *best_ms);

typedef struct {
  int seed;
  int gen_mx;
  double b0;
  double arrival_s;   /* offered arrival time relative to start */
} hs_uts_request;

typedef struct {
  uint64_t nodes, leaves;
  int max_depth;
  double admit_s, done_s;  /* device-observed admission / completion */
} hs_uts_response;

typedef struct {
  int slots;          /* concurrent trees on the device (the response
                         board); 0 = 8, at most 64 */
  int num_blocks;     /* persistent WGs including the dispatcher; 0 = every
                         WG that can be resident (three waves per SIMD at
                         the kernel's register use).  HS_EOCCUPANCY if they
                         cannot all be resident. */
  int deque_nodes;    /* records per WG deque (HBM); 0 = 65536 */
  double timeout_s;
} hs_serve_cfg;

typedef struct {
  double wall_s;
  uint64_t total_nodes;
  double p50_s, p99_s;
  uint32_t cut_batches;      /* batches the deque bound cut short */
  uint32_t deque_high_water; /* most records any WG's deque held */
  uint32_t deque_soft;       /* the bound: deque_nodes minus the reserve */
} hs_serve_summary;

/* UTS serving loop: one persistent launch admits the requests as they
 * arrive (arrival_s after the call starts) into free slots and reports
 * each tree when its last record is done.  Admission is by deque memory:
 * a tree is taken by a worker whose deque is empty, and every worker keeps
 * free the records a one-node-at-a-time depth-first walk of the deepest
 * tree of the session could add, max gen_mx * (max_children - 1), cutting
 * batches and holding back deliveries to stay under deque_nodes minus that
 * reserve.  No deque overflows and no tree is dropped.  HS_EINVAL if
 * deque_nodes cannot hold the reserve. */
int hs_uts_serve(hs_ctx *ctx, const hs_uts_request *req, int n,
                 const hs_serve_cfg *cfg, hs_uts_response *resp,
                 hs_serve_summary *sum);

#ifdef __cplusplus
}
#endif

#endif
