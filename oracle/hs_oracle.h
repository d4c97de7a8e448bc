/*
 * hs_oracle.h -- CPU oracle for the HeteroSync synchronization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library or CLIs links,
 * loads or calls this code: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker (or, for the
 * std::thread emulation, as the timed CPU baseline).
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference tree (hip/syncPrims/uvm/main.hip.cpp = HM, ...).
 *
 * Pinning (what each function is checked against, in tests/test_oracle.py):
 *   hso_golden_literal   restatement of HM:2524-2682; pinned by brute-force
 *                        simulation of the kernels' region walk at tiny sizes
 *                        and the hand-derived closed form (HM:60-72 shift).
 *   hso_golden_closed    closed form; pinned against hso_golden_literal and
 *                        the std::thread emulation (a real concurrent run).
 *   hso_sha1             FIPS 180-4 known-answer vectors ("abc", "", the
 *                        448-bit and million-'a' messages).
 *   hso_uts_geo_*        published UTS sample tree T1 (geometric, fixed
 *                        shape, b=4, d=10, r=19: 4130071 nodes, 3305118
 *                        leaves) and brute force on tiny depths.
 *   hso_uts_ref_tree     restatement of cuda/uts/uvm/uts.cu:290-382; pinned
 *                        by structural invariants (binary tree, node count,
 *                        leaves = (nodes+1)/2, golden leaf sum recomputed).
 *   seqlock/flags/split/refcounter expected values: closed forms checked
 *                        against brute-force literal loops (tests).
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- syncPrims (hip/syncPrims/uvm/main.hip.cpp) ------------------------ */

/* Primitive ids, identical to the string table at HM:2036-2089. */
int hso_prim_id(const char *name);

/* numLocsMult of HM:2109-2128 (how many WG-sized regions the storage has). */
int hso_num_locs_mult(int prim, int num_wgs, int num_cu);

/* numStorageLocs of HM:2130-2136 with NUM_WIS_PER_WG replaced by wg_size. */
long hso_num_storage_locs(int prim, int num_wgs, int num_cu, int wg_size,
                          int num_ldst);

/* Literal restatement of the host golden loop HM:2524-2682 (with
 * accessData_golden HM:33-51).  mode 0 = copy chain (reference), mode 1 =
 * "inc" (each visited non-first-line word is incremented by 1).  out must
 * hold hso_num_storage_locs() words; it is initialised to out[i] = i here.
 * Returns 0, or -1 for an unsupported primitive. */
int hso_golden_literal(int prim, int num_ldst, int num_wgs, int iters,
                       int reps, int wg_size, int num_cu, int mode,
                       int32_t *out, long n);

/* Closed form of the same final state, O(n): passes per region k_r come
 * from the kernels' region walk (HM:104-149, 315-373, 542-643, 645-857,
 * 861-930, 1005-1073) and each chain of L+1 lines ends as
 *   copy: line j = original line max(0, j-k)
 *   inc : line j = original + k (j >= 1), line 0 untouched.
 * Returns 0, or -1 for an unsupported primitive / shape. */
int hso_golden_closed(int prim, int num_ldst, int num_wgs, int iters,
                      int reps, int wg_size, int num_cu, int mode,
                      int32_t *out, long n);

/* Passes over each region (size hso_num_locs_mult()) implied by the
 * kernels' access pattern; the basis of hso_golden_closed. */
int hso_region_passes(int prim, int num_wgs, int iters, int reps, int num_cu,
                      int64_t *passes, int nregions);

/* Number of critical-section entries (acquires) in one invocation:
 * mutexes num_wgs*iters*reps; semaphores split into writer/reader entries;
 * barriers: episodes = iters*reps. */
int hso_expected_counts(int prim, int num_wgs, int iters, int reps,
                        int num_cu, int64_t *writer_entries,
                        int64_t *reader_entries, int64_t *episodes);

/* ---- relaxed atomics ----------------------------------------------------- */

/* seqlock (hip/relaxedAtomics/uvm/seqlocks/seqlocks_kernel.h:187-240):
 * writer WGs are b % writer_period == 0, lock index b % group_size; every
 * writer pass adds 2.  seq must hold n_locks >= group_size entries. */
void hso_seqlock_expected(int num_wgs, int group_size, int writer_period,
                          int iters, uint32_t *seq, int n_locks);

/* flags (cuda/relaxedAtomics/uvm/flags/flags.cu:92-101): glibc
 * srand(seed); r[i] = (float)rand() / (float)UINT32_MAX. */
void hso_flags_rand(float *out, long n, unsigned seed);

/* flags dirty count: sum over worker TBs b < num_tbs-1 and it < num_iters
 * of [rand[b + it*num_tbs] <= cutoff]  (flags_kernel.cu:118-135). */
uint64_t hso_flags_dirty(const float *rand_arr, int num_tbs, int num_iters,
                         float cutoff);

/* ---- SHA-1 and UTS ------------------------------------------------------- */

/* FIPS 180-4 SHA-1 of an arbitrary message. */
void hso_sha1(const uint8_t *msg, size_t len, uint8_t out[20]);

/* UTS RNG (public UTS distribution, rng/brg_sha1.c): root state =
 * SHA1(16 zero bytes || seed as 4 big-endian bytes); child i state =
 * SHA1(parent state || i as 4 big-endian bytes); rand = last 4 state bytes
 * big-endian & 0x7fffffff. */
void hso_uts_root(int seed, uint8_t state[20]);
void hso_uts_spawn(const uint8_t parent[20], int i, uint8_t child[20]);
int32_t hso_uts_rand(const uint8_t state[20]);

/* Geometric/fixed-shape child count, the classic double-precision formula
 * floor(log(1-u)/log(1-p)), p = 1/(1+b_i), u = rand/2^31, b_i = b0 for
 * depth < gen_mx else 0, capped at 100 (uts.c uts_numChildren_geo). */
int hso_uts_geo_children(const uint8_t state[20], int depth, double b0,
                         int gen_mx);

/* Integer thresholds: children = #{k >= 1 : rand >= thr[k-1]}.  Fills up
 * to max_k entries, returns how many are finite (the rest are 2^31). */
int hso_uts_geo_thresholds(double b0, uint32_t *thr, int max_k);

/* Exhaustive depth-first count of one geometric tree. */
void hso_uts_geo_count(int seed, double b0, int gen_mx, uint64_t *nodes,
                       uint64_t *leaves, int *max_depth);

/* Reference binary tree of cuda/uts/uvm/uts.cu:290-339 (glibc srand(200)).
 * Arrays have num_nodes entries.  Returns restart count.  Also returns the
 * golden leaf count and leaf sum (uts.cu:327-339). */
int hso_uts_ref_tree(int num_nodes, int fertility, uint32_t *value,
                     int32_t *child_a, int32_t *child_b,
                     uint32_t *leaves_golden, uint32_t *leaf_sum_golden);

/* ---- std::thread + std::atomic emulation (CPU baseline) ------------------ */

/* Runs the same primitive on host threads (one std::thread per emulated WG,
 * inner loop over its lanes) and writes final storage.  Returns elapsed
 * seconds (>= 0) or a negative value on an unsupported primitive.
 * threads_cap bounds concurrent host threads for the lock primitives (the
 * barrier needs all participants live, so it uses num_wgs threads). */
double hso_cpu_run(int prim, int num_ldst, int num_wgs, int iters, int reps,
                   int wg_size, int num_cu, int mode, int32_t *storage,
                   long n, int64_t *cs_count);

/* Barrier workload (prims 0-3) on `threads` host threads, each owning a
 * contiguous block of the num_wgs emulated WGs: per episode every thread
 * runs its WGs' critical-section walks (same region rotation as the
 * kernels), then all threads meet in one barrier.  Final storage equals
 * the closed form.  Returns seconds, or < 0 for an unsupported shape. */
double hso_cpu_barrier_run(int prim, int num_ldst, int num_wgs, int iters,
                           int reps, int wg_size, int num_cu, int mode,
                           int threads, int32_t *storage, long n);

#ifdef __cplusplus
}
#endif

#endif
