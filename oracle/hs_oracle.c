/*
 * hs_oracle.c -- CPU restatement of the reference's host-side goldens and of
 * the algorithms whose results the GPU path must reproduce bit-exactly.
 *
 * TEST INFRASTRUCTURE ONLY (see hs_oracle.h).  Citations are to the
 * reference tree:
 *   HM  = hip/syncPrims/uvm/main.hip.cpp
 *   HQ  = hip/relaxedAtomics/uvm/seqlocks/seqlocks_kernel.h
 *   CF  = cuda/relaxedAtomics/uvm/flags/flags.cu
 *   CFK = cuda/relaxedAtomics/uvm/flags/flags_kernel.cu
 *   CU  = cuda/uts/uvm/uts.cu
 */
#include "hs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define WORDS_PER_LINE 16 /* NUM_WORDS_PER_CACHELINE, HM:12 */

/* ------------------------------------------------------------------------ */
/* primitive table: HM:2036-2089                                            */
/* ------------------------------------------------------------------------ */
static const struct {
  const char *name;
  int id;
} k_prims[] = {
    {"atomicTreeBarrUniq", 0},
    {"atomicTreeBarrUniqLocalExch", 1},
    {"lfTreeBarrUniq", 2},
    {"lfTreeBarrUniqLocalExch", 3},
    {"spinMutex", 4},
    {"spinMutexEBO", 5},
    {"sleepMutex", 6},
    {"faMutex", 7},
    {"spinSem1", 8},
    {"spinSem2", 9},
    {"spinSem10", 10},
    {"spinSem120", 11},
    {"spinSemEBO1", 12},
    {"spinSemEBO2", 13},
    {"spinSemEBO10", 14},
    {"spinSemEBO120", 15},
    {"PRspinSem1", 16},
    {"PRspinSem2", 17},
    {"PRspinSem10", 18},
    {"PRspinSem120", 19},
    {"spinMutexUniq", 20},
    {"spinMutexEBOUniq", 21},
    {"sleepMutexUniq", 22},
    {"faMutexUniq", 23},
    {"spinSemUniq1", 24},
    {"spinSemUniq2", 25},
    {"spinSemUniq10", 26},
    {"spinSemUniq120", 27},
    {"spinSemEBOUniq1", 28},
    {"spinSemEBOUniq2", 29},
    {"spinSemEBOUniq10", 30},
    {"spinSemEBOUniq120", 31},
    {"PRspinSemEBO1", 32},
    {"PRspinSemEBO2", 33},
    {"PRspinSemEBO10", 34},
    {"PRspinSemEBO120", 35},
    {"atomicTreeBarrUniqSRB", 37},
    {"atomicTreeBarrUniqLocalExchSRB", 38},
    {"atomicTreeBarrUniqNaiveSRB", 39},
    {"atomicTreeBarrUniqLocalExchNaiveSRB", 40},
    {"atomicTreeBarrUniqNaiveAllSRB", 41},
    {"atomicTreeBarrUniqLocalExchNaiveAllSRB", 42},
};

int hso_prim_id(const char *name) {
  for (size_t i = 0; i < sizeof(k_prims) / sizeof(k_prims[0]); ++i)
    if (strcmp(name, k_prims[i].name) == 0) return k_prims[i].id;
  return -1;
}

static int is_barrier(int p) { return p < 4 || p >= 37; }
static int is_lx(int p) {
  return p == 1 || p == 3 || p == 38 || p == 40 || p == 42 || p == 44 ||
         p == 46;
}
static int is_mutex(int p) { return p >= 4 && p <= 7; }
static int is_sem(int p) { return p >= 8 && p <= 19; }
static int is_uniq_mutex(int p) { return p >= 20 && p <= 23; }
static int is_uniq_sem(int p) { return p >= 24 && p <= 35; }

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

/* numWGs_perCU = ceil(numWGs / NUM_CU), HM:2032 */
static int per_cu(int num_wgs, int num_cu) { return ceil_div(num_wgs, num_cu); }

/* HM:2109-2128 */
int hso_num_locs_mult(int p, int num_wgs, int num_cu) {
  if (is_barrier(p) || is_uniq_sem(p)) return num_wgs;
  if (is_mutex(p)) return 1;
  if (is_sem(p)) return per_cu(num_wgs, num_cu);
  if (is_uniq_mutex(p) || p == 36) return num_wgs < num_cu ? num_wgs : num_cu;
  return -1;
}

/* HM:2130-2136 */
long hso_num_storage_locs(int p, int num_wgs, int num_cu, int wg_size,
                          int num_ldst) {
  int m = hso_num_locs_mult(p, num_wgs, num_cu);
  if (m < 0) return -1;
  return (long)m * (long)wg_size * (long)(num_ldst + 1);
}

/* accessData_golden, HM:33-51 (mode 1 replaces the copy by +1). */
static void access_golden(int32_t *s, long cur, int num_ldst, int mode) {
  if (cur % ((long)WORDS_PER_LINE * (num_ldst + 1)) >= WORDS_PER_LINE) {
    if (mode == 0)
      s[cur] = s[cur - WORDS_PER_LINE];
    else
      s[cur] += 1;
  }
}

/* Literal restatement of HM:2524-2682. */
int hso_golden_literal(int p, int L, int num_wgs, int iters, int reps,
                       int wg_size, int num_cu, int mode, int32_t *s,
                       long n) {
  const long U = (long)wg_size * (L + 1); /* numUniqLocsAccPerWG */
  const long N = hso_num_storage_locs(p, num_wgs, num_cu, wg_size, L);
  const int perCU = per_cu(num_wgs, num_cu);
  if (N < 0 || N != n) return -1;
  if (!(is_barrier(p) || is_mutex(p) || is_sem(p) || is_uniq_mutex(p) ||
        is_uniq_sem(p)))
    return -1;
  for (long i = 0; i < n; ++i) s[i] = (int32_t)i;

  for (int rep = 0; rep < reps; ++rep) {
    for (int j = 0; j < iters; ++j) {
      if (is_barrier(p)) { /* HM:2541-2579 */
        const long acc = (long)num_wgs * U;
        for (long i = acc - 1; i >= 0; --i)
          access_golden(s, (i + (long)j * U) % acc, L, mode);
        if (is_lx(p) && perCU > 1) {
          for (long i = acc - 1; i >= 0; --i) {
            long cur = (i + (long)perCU * U) % acc;
            cur = (cur + (long)j * U) % acc;
            access_golden(s, cur, L, mode);
          }
        }
      } else if (is_mutex(p)) { /* HM:2585-2597 */
        for (int wg = 0; wg < num_wgs; ++wg)
          for (long i = U - 1; i >= 0; --i) access_golden(s, i, L, mode);
      } else if (is_sem(p)) { /* HM:2604-2637 */
        const int ncu = num_wgs < num_cu ? num_wgs : num_cu;
        for (int wg = 0; wg < num_wgs; ++wg) {
          const int cu = wg % ncu, pw = wg / ncu;
          if (pw != cu % perCU) continue;
          for (int k = 0; k < perCU; ++k)
            for (long i = U - 1; i >= 0; --i)
              access_golden(s, (i + (long)k * U) % N, L, mode);
        }
      } else if (is_uniq_mutex(p)) { /* HM:2642-2659 */
        const long acc = (long)num_wgs * U;
        for (long i = acc - 1; i >= 0; --i)
          access_golden(s, i % ((long)num_cu * U), L, mode);
      } else { /* uniq semaphores, HM:2667-2680 */
        const long acc = (long)num_wgs * U;
        for (long i = acc - 1; i >= 0; --i) access_golden(s, i, L, mode);
      }
    }
  }
  return 0;
}

/* Region walk of the kernels.  Barrier kernels visit region
 * (b + shift) mod grid once per episode (HM:135-148, 542-584) and, for the
 * local-exchange variants with more than one WG per group, a second region
 * (HM:351-361, 620-631); mutex kernels all use region 0 (HM:664-672) or the
 * per-CU region blockIdx % NUM_CU (HM:683-686); semaphore writers walk
 * their regions (HM:899-921, HM:1040-1058). */
int hso_region_passes(int p, int num_wgs, int iters, int reps, int num_cu,
                      int64_t *passes, int nregions) {
  const int m = hso_num_locs_mult(p, num_wgs, num_cu);
  const int64_t ir = (int64_t)iters * reps;
  const int nab = num_wgs < num_cu ? num_wgs : num_cu;
  const int perCU = ceil_div(num_wgs, nab);
  if (m < 0 || m != nregions) return -1;
  for (int r = 0; r < m; ++r) passes[r] = 0;
  if (is_barrier(p)) {
    if (num_wgs > num_cu && num_wgs % num_cu) return -1; /* RD:80 */
    for (int r = 0; r < m; ++r) passes[r] = ir * ((is_lx(p) && perCU > 1) ? 2 : 1);
  } else if (is_mutex(p)) {
    passes[0] = (int64_t)num_wgs * ir;
  } else if (is_sem(p) || (p >= 32 && p <= 35)) {
    /* global semaphores; the priority EBO ids 32-35 run the global kernel
     * too (HM:2455-2470 -> invokeEBOSemaphorePriority) over storage sized
     * like the uniq semaphores' (HM:2109-2128), so only the first perCU of
     * their num_wgs regions are written */
    int64_t writers = 0;
    for (int b = 0; b < num_wgs; ++b)
      if (b / nab == (b % nab) % perCU) ++writers;
    for (int r = 0; r < m && r < perCU; ++r) passes[r] = writers * ir;
  } else if (is_uniq_mutex(p)) {
    for (int b = 0; b < num_wgs; ++b) passes[b % num_cu] += ir;
  } else if (is_uniq_sem(p)) {
    if (num_wgs % nab) return -1; /* every group needs its writer */
    for (int b = 0; b < num_wgs; ++b) {
      if (b / nab != (b % nab) % perCU) continue;
      int cur = b;
      for (int j = 0; j < perCU; ++j) {
        passes[cur] += ir;
        cur = (cur + nab) % num_wgs;
      }
    }
  } else {
    return -1;
  }
  return 0;
}

int hso_golden_closed(int p, int L, int num_wgs, int iters, int reps,
                      int wg_size, int num_cu, int mode, int32_t *s, long n) {
  const long N = hso_num_storage_locs(p, num_wgs, num_cu, wg_size, L);
  const int m = hso_num_locs_mult(p, num_wgs, num_cu);
  if (N < 0 || N != n) return -1;
  int64_t *passes = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
  if (!passes) return -1;
  if (hso_region_passes(p, num_wgs, iters, reps, num_cu, passes, m)) {
    free(passes);
    return -1;
  }
  const long U = (long)wg_size * (L + 1);
  const long chain = (long)WORDS_PER_LINE * (L + 1);
  for (long i = 0; i < n; ++i) {
    const int64_t k = passes[i / U];
    const long c0 = (i / chain) * chain;       /* chain start */
    const long j = (i - c0) / WORDS_PER_LINE;  /* line in chain */
    const long w = (i - c0) % WORDS_PER_LINE;  /* word in line */
    if (mode == 0) {
      const long src = j - k > 0 ? j - k : 0;
      s[i] = (int32_t)(c0 + src * WORDS_PER_LINE + w);
    } else {
      s[i] = (int32_t)(i + (j >= 1 ? k : 0));
    }
  }
  free(passes);
  return 0;
}

int hso_expected_counts(int p, int num_wgs, int iters, int reps, int num_cu,
                        int64_t *wr, int64_t *rd, int64_t *ep) {
  const int64_t ir = (int64_t)iters * reps;
  const int nab = num_wgs < num_cu ? num_wgs : num_cu;
  const int perCU = ceil_div(num_wgs, nab);
  *wr = *rd = *ep = 0;
  if (is_barrier(p)) {
    *ep = ir;
  } else if (is_mutex(p) || is_uniq_mutex(p)) {
    *wr = (int64_t)num_wgs * ir;
  } else if (is_sem(p) || is_uniq_sem(p)) {
    int64_t w = 0;
    for (int b = 0; b < num_wgs; ++b)
      if (b / nab == (b % nab) % perCU) ++w;
    *wr = w * ir;
    *rd = (int64_t)(num_wgs - w) * ir;
  } else {
    return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* relaxed atomics                                                          */
/* ------------------------------------------------------------------------ */

/* HQ:187-240: writer iff blockIdx % 16 == 0 (writer_period generalises the
 * hard-coded 16), lock = blockIdx % groupSize; each writer pass moves the
 * sequence number from even s to s+2 (HQ:166-176). */
void hso_seqlock_expected(int num_wgs, int group_size, int writer_period,
                          int iters, uint32_t *seq, int n_locks) {
  for (int l = 0; l < n_locks; ++l) seq[l] = 0;
  for (int b = 0; b < num_wgs; ++b)
    if (b % writer_period == 0) seq[b % group_size] += 2u * (uint32_t)iters;
}

/* CF:92-101 (UINT32_MAX redefined as (unsigned)-1 at CF:7). */
void hso_flags_rand(float *out, long n, unsigned seed) {
  srand(seed);
  for (long i = 0; i < n; ++i)
    out[i] = ((float)rand() / (float)(unsigned int)-1);
}

uint64_t hso_flags_dirty(const float *r, int num_tbs, int num_iters,
                         float cutoff) {
  uint64_t d = 0;
  for (int b = 0; b < num_tbs - 1; ++b)
    for (int it = 0; it < num_iters; ++it)
      if (r[(long)b + (long)it * num_tbs] <= cutoff) ++d;
  return d;
}

/* ------------------------------------------------------------------------ */
/* SHA-1, FIPS 180-4 section 6.1                                            */
/* ------------------------------------------------------------------------ */
static uint32_t rol(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

static void sha1_block(uint32_t h[5], const uint8_t blk[64]) {
  uint32_t w[80];
  for (int t = 0; t < 16; ++t)
    w[t] = ((uint32_t)blk[4 * t] << 24) | ((uint32_t)blk[4 * t + 1] << 16) |
           ((uint32_t)blk[4 * t + 2] << 8) | (uint32_t)blk[4 * t + 3];
  for (int t = 16; t < 80; ++t)
    w[t] = rol(w[t - 3] ^ w[t - 8] ^ w[t - 14] ^ w[t - 16], 1);
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
  for (int t = 0; t < 80; ++t) {
    uint32_t f, k;
    if (t < 20) {
      f = (b & c) | (~b & d);
      k = 0x5a827999u;
    } else if (t < 40) {
      f = b ^ c ^ d;
      k = 0x6ed9eba1u;
    } else if (t < 60) {
      f = (b & c) | (b & d) | (c & d);
      k = 0x8f1bbcdcu;
    } else {
      f = b ^ c ^ d;
      k = 0xca62c1d6u;
    }
    uint32_t tmp = rol(a, 5) + f + e + k + w[t];
    e = d;
    d = c;
    c = rol(b, 30);
    b = a;
    a = tmp;
  }
  h[0] += a;
  h[1] += b;
  h[2] += c;
  h[3] += d;
  h[4] += e;
}

void hso_sha1(const uint8_t *msg, size_t len, uint8_t out[20]) {
  uint32_t h[5] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u,
                   0xc3d2e1f0u};
  size_t off = 0;
  for (; off + 64 <= len; off += 64) sha1_block(h, msg + off);
  uint8_t blk[128];
  size_t rem = len - off;
  memset(blk, 0, sizeof blk);
  memcpy(blk, msg + off, rem);
  blk[rem] = 0x80;
  size_t tot = (rem + 1 + 8 <= 64) ? 64 : 128;
  uint64_t bits = (uint64_t)len * 8u;
  for (int i = 0; i < 8; ++i) blk[tot - 1 - i] = (uint8_t)(bits >> (8 * i));
  sha1_block(h, blk);
  if (tot == 128) sha1_block(h, blk + 64);
  for (int i = 0; i < 5; ++i) {
    out[4 * i] = (uint8_t)(h[i] >> 24);
    out[4 * i + 1] = (uint8_t)(h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(h[i] >> 8);
    out[4 * i + 3] = (uint8_t)h[i];
  }
}

/* ------------------------------------------------------------------------ */
/* UTS geometric trees (public UTS algorithm; no reference implementation)  */
/* ------------------------------------------------------------------------ */
void hso_uts_root(int seed, uint8_t st[20]) {
  uint8_t m[20];
  memset(m, 0, sizeof m);
  m[16] = (uint8_t)(seed >> 24);
  m[17] = (uint8_t)(seed >> 16);
  m[18] = (uint8_t)(seed >> 8);
  m[19] = (uint8_t)seed;
  hso_sha1(m, 20, st);
}

void hso_uts_spawn(const uint8_t parent[20], int i, uint8_t child[20]) {
  uint8_t m[24];
  memcpy(m, parent, 20);
  m[20] = (uint8_t)(i >> 24);
  m[21] = (uint8_t)(i >> 16);
  m[22] = (uint8_t)(i >> 8);
  m[23] = (uint8_t)i;
  hso_sha1(m, 24, child);
}

int32_t hso_uts_rand(const uint8_t st[20]) {
  uint32_t b = ((uint32_t)st[16] << 24) | ((uint32_t)st[17] << 16) |
               ((uint32_t)st[18] << 8) | (uint32_t)st[19];
  return (int32_t)(b & 0x7fffffffu);
}

static int geo_children_h(int32_t h, double b_i) {
  double p = 1.0 / (1.0 + b_i);
  double u = (double)h / 2147483648.0;
  int nc = (int)floor(log(1.0 - u) / log(1.0 - p));
  return nc > 100 ? 100 : nc;
}

int hso_uts_geo_children(const uint8_t st[20], int depth, double b0,
                         int gen_mx) {
  double b_i = (depth < gen_mx) ? b0 : 0.0;
  if (depth == 0) b_i = b0;
  return geo_children_h(hso_uts_rand(st), b_i);
}

int hso_uts_geo_thresholds(double b0, uint32_t *thr, int max_k) {
  int finite = 0;
  for (int k = 1; k <= max_k; ++k) {
    /* smallest h in [0, 2^31) with children(h) >= k, else 2^31 */
    int64_t lo = 0, hi = 2147483648LL;
    while (lo < hi) {
      int64_t mid = (lo + hi) / 2;
      if (geo_children_h((int32_t)mid, b0) >= k)
        hi = mid;
      else
        lo = mid + 1;
    }
    thr[k - 1] = (uint32_t)lo;
    if (lo < 2147483648LL) finite = k;
  }
  return finite;
}

typedef struct {
  uint8_t st[20];
  int depth;
} geo_node;

void hso_uts_geo_count(int seed, double b0, int gen_mx, uint64_t *nodes,
                       uint64_t *leaves, int *max_depth) {
  size_t cap = 1 << 16, top = 0;
  geo_node *stk = (geo_node *)malloc(cap * sizeof(geo_node));
  uint64_t n = 0, l = 0;
  int md = 0;
  hso_uts_root(seed, stk[0].st);
  stk[0].depth = 0;
  top = 1;
  while (top) {
    geo_node cur = stk[--top];
    ++n;
    if (cur.depth > md) md = cur.depth;
    int nc = hso_uts_geo_children(cur.st, cur.depth, b0, gen_mx);
    if (nc <= 0) {
      ++l;
      continue;
    }
    if (top + (size_t)nc > cap) {
      cap *= 2;
      stk = (geo_node *)realloc(stk, cap * sizeof(geo_node));
    }
    for (int i = 0; i < nc; ++i) {
      hso_uts_spawn(cur.st, i, stk[top].st);
      stk[top].depth = cur.depth + 1;
      ++top;
    }
  }
  free(stk);
  *nodes = n;
  *leaves = l;
  *max_depth = md;
}

/* ------------------------------------------------------------------------ */
/* Reference UTS binary tree, CU:290-339                                    */
/* ------------------------------------------------------------------------ */
int hso_uts_ref_tree(int num_nodes, int fertility, uint32_t *value,
                     int32_t *child_a, int32_t *child_b,
                     uint32_t *leaves_golden, uint32_t *leaf_sum_golden) {
  /* undecided_nodes: FIFO; leaf_nodes: inserted at the front (CU:320), kept
   * here in push order and reversed when it becomes the undecided list. */
  int32_t *und = (int32_t *)malloc(sizeof(int32_t) * (size_t)(num_nodes + 1));
  int32_t *lf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(num_nodes + 1));
  long uh = 0, ut = 0, nl = 0;
  int restarts = 0;
  for (int i = 0; i < num_nodes; ++i) {
    value[i] = 0;
    child_a[i] = child_b[i] = 0;
  }
  srand(200);
  und[ut++] = 0;
  int next = 1;
  while (next != num_nodes) {
    if (ut == uh) { /* CU:296-301 */
      for (long i = 0; i < nl; ++i) und[i] = lf[nl - 1 - i];
      uh = 0;
      ut = nl;
      nl = 0;
      ++restarts;
    }
    int cur = und[uh++];
    value[cur] = (uint32_t)(rand() % 100);
    if ((rand() % 100) < fertility) {
      child_a[cur] = next;
      und[ut++] = next++;
      child_b[cur] = next;
      und[ut++] = next++;
    } else {
      child_a[cur] = child_b[cur] = 0;
      lf[nl++] = cur;
    }
  }
  uint32_t sum = 0;
  for (long i = uh; i < ut; ++i) { /* CU:327-334 */
    child_a[und[i]] = child_b[und[i]] = 0;
    value[und[i]] = 1;
    sum += 1;
  }
  for (long i = 0; i < nl; ++i) sum += value[lf[i]];
  *leaves_golden = (uint32_t)(nl + (ut - uh));
  *leaf_sum_golden = sum;
  free(und);
  free(lf);
  return restarts;
}
