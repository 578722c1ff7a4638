/*
 * eclcc_oracle.c -- CPU oracle for the race-free ECL-CC path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into the product library.
 * See oracle.h for what each function restates and how it is pinned.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------ */
/* union-find with min-linking (MST/ECL-MST_10-atomic.cu:85-109 pattern,
   link direction reversed so that roots are component minima)         */

static int32_t uf_find(int32_t* parent, int32_t x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]]; /* path halving */
    x = parent[x];
  }
  return x;
}

int oracle_cc_unionfind(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* label) {
  if (n < 1 || !nindex || !label) return -1;
  for (int32_t v = 0; v < n; v++) label[v] = v;
  for (int32_t v = 0; v < n; v++) {
    for (int32_t i = nindex[v]; i < nindex[v + 1]; i++) {
      const int32_t u = nlist[i];
      if (u < 0 || u >= n) return -2;
      int32_t a = uf_find(label, v), b = uf_find(label, u);
      if (a == b) continue;
      if (a < b) label[b] = a; /* larger root under the smaller */
      else label[a] = b;
    }
  }
  for (int32_t v = 0; v < n; v++) label[v] = uf_find(label, v);
  return 0;
}

/* ------------------------------------------------------------------ */
/* BFS labelling in increasing start order                             */

int oracle_cc_bfs(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* label) {
  if (n < 1 || !nindex || !label) return -1;
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!queue) return -3;
  for (int32_t v = 0; v < n; v++) label[v] = -1;
  for (int32_t s = 0; s < n; s++) {
    if (label[s] >= 0) continue;
    int32_t head = 0, tail = 0;
    label[s] = s;
    queue[tail++] = s;
    while (head < tail) {
      const int32_t x = queue[head++];
      for (int32_t i = nindex[x]; i < nindex[x + 1]; i++) {
        const int32_t y = nlist[i];
        if (y < 0 || y >= n) { free(queue); return -2; }
        if (label[y] < 0) {
          label[y] = s;
          queue[tail++] = y;
        }
      }
    }
  }
  free(queue);
  return 0;
}

/* ------------------------------------------------------------------ */
/* init (ECL-CC-atomic.cu:69-88)                                       */

void oracle_init(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat) {
  for (int32_t v = 0; v < n; v++) {
    int32_t m = v;
    for (int32_t i = nindex[v]; i < nindex[v + 1] && m == v; i++)
      if (nlist[i] < m) m = nlist[i];
    nstat[v] = m;
  }
}

/* ------------------------------------------------------------------ */
/* verifier (ECL-CC-atomic.cu:365-395, verify :327-336)                */

int64_t oracle_count_components(int32_t n, const int32_t* label) {
  unsigned char* seen = (unsigned char*)calloc((size_t)n, 1);
  if (!seen) return -1;
  int64_t c = 0;
  for (int32_t v = 0; v < n; v++) {
    const int32_t l = label[v];
    if (l >= 0 && l < n && !seen[l]) { seen[l] = 1; c++; }
  }
  free(seen);
  return c;
}

int oracle_verify(int32_t n, const int32_t* nindex, const int32_t* nlist, const int32_t* label,
                  int64_t* ncomp) {
  if (n < 1 || !nindex || !label) return ORACLE_VERIFY_ALLOC;
  /* labels must be vertex IDs before anything indexes by them */
  for (int32_t v = 0; v < n; v++) {
    if (label[v] < 0) return ORACLE_VERIFY_NEGATIVE;
    if (label[v] >= n) return ORACLE_VERIFY_RANGE;
  }
  const int64_t s1 = oracle_count_components(n, label);
  /* adjacent vertices share a label (:373-377) */
  for (int32_t v = 0; v < n; v++)
    for (int32_t i = nindex[v]; i < nindex[v + 1]; i++)
      if (label[nlist[i]] != label[v]) return ORACLE_VERIFY_ADJACENT;
  /* one DFS per component on a copy; each reaches only its own ID (:383-393) */
  int32_t* stat = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  unsigned char* seen = (unsigned char*)calloc((size_t)n, 1);
  if (!stat || !stack || !seen) { free(stat); free(stack); free(seen); return ORACLE_VERIFY_ALLOC; }
  memcpy(stat, label, sizeof(int32_t) * (size_t)n);
  int64_t count = 0, s2 = 0;
  int rc = ORACLE_VERIFY_OK, notmin = 0;
  for (int32_t v = 0; v < n && rc == ORACLE_VERIFY_OK; v++) {
    if (stat[v] < 0) continue;
    const int32_t id = stat[v];
    count++;
    if (!seen[id]) { seen[id] = 1; s2++; }
    /* the ID must be the component's minimum: v is the first vertex met in
       ID order, so it is the minimum, and the label must name it */
    if (id != v) notmin = 1;
    int32_t top = 0;
    stack[top++] = v;
    stat[v] = -1;
    while (top > 0 && rc == ORACLE_VERIFY_OK) {
      const int32_t x = stack[--top];
      for (int32_t i = nindex[x]; i < nindex[x + 1]; i++) {
        const int32_t y = nlist[i];
        if (stat[y] >= 0) {
          if (stat[y] != id) { rc = ORACLE_VERIFY_WRONG_ID; break; }
          stat[y] = -1;
          stack[top++] = y;
        }
      }
    }
  }
  free(stat);
  free(stack);
  free(seen);
  if (rc != ORACLE_VERIFY_OK) return rc;
  if (s1 != s2) return ORACLE_VERIFY_COUNT;
  if (s1 != count) return ORACLE_VERIFY_NOT_UNIQUE;
  if (notmin) return ORACLE_VERIFY_NOT_MIN;
  if (ncomp) *ncomp = s1;
  return ORACLE_VERIFY_OK;
}

/* ------------------------------------------------------------------ */
/* serial ECL-CC restatement (CPU-1)                                   */

enum { THREAD_MAX_DEG = 16, WARP_MAX_DEG = 352 }; /* ECL-CC-atomic.cu:120,122 */

/* representative with intermediate pointer jumping (:92-104) */
static int32_t rep_serial(int32_t idx, int32_t* nstat) {
  int32_t curr = nstat[idx];
  if (curr != idx) {
    int32_t next, prev = idx;
    while (curr > (next = nstat[curr])) {
      nstat[prev] = next;
      prev = curr;
      curr = next;
    }
  }
  return curr;
}

/* hook the edge from vertex v (root estimate vstat) to neighbour u (:134-151) */
static int32_t hook_serial(int32_t vstat, int32_t u, int32_t* nstat) {
  int32_t ostat = rep_serial(u, nstat);
  while (vstat != ostat) {
    if (vstat < ostat) {
      if (nstat[ostat] == ostat) { nstat[ostat] = vstat; break; }
      ostat = nstat[ostat];
    } else {
      if (nstat[vstat] == vstat) { nstat[vstat] = ostat; break; }
      vstat = nstat[vstat];
    }
  }
  return vstat;
}

static void hook_all_serial(int32_t v, const int32_t* nindex, const int32_t* nlist, int32_t* nstat) {
  int32_t vstat = rep_serial(v, nstat);
  for (int32_t i = nindex[v]; i < nindex[v + 1]; i++)
    if (v > nlist[i]) vstat = hook_serial(vstat, nlist[i], nstat);
}

int oracle_eclcc_serial(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat) {
  if (n < 1 || !nindex || !nstat) return -1;
  int32_t* wl = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!wl) return -3;
  oracle_init(n, nindex, nlist, nstat);
  int32_t topL = 0, topH = n - 1;
  for (int32_t v = 0; v < n; v++) { /* compute1 */
    if (nstat[v] == v) continue;
    const int32_t deg = nindex[v + 1] - nindex[v];
    if (deg > THREAD_MAX_DEG) {
      if (deg <= WARP_MAX_DEG) wl[topL++] = v;
      else wl[topH--] = v;
    } else {
      hook_all_serial(v, nindex, nlist, nstat);
    }
  }
  for (int32_t k = 0; k < topL; k++) hook_all_serial(wl[k], nindex, nlist, nstat); /* compute2 */
  for (int32_t k = n - 1; k > topH; k--) hook_all_serial(wl[k], nindex, nlist, nstat); /* compute3 */
  for (int32_t v = 0; v < n; v++) { /* flatten (:250-264) */
    int32_t s = nstat[v], next;
    while (s > (next = nstat[s])) s = next;
    nstat[v] = s;
  }
  free(wl);
  return 0;
}

/* ------------------------------------------------------------------ */
/* threaded ECL-CC restatement (CPU-N)                                 */

static inline int32_t ald(const int32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
static inline void ast(int32_t* p, int32_t v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
static inline int32_t acas(int32_t* p, int32_t expect, int32_t desired) {
  __atomic_compare_exchange_n(p, &expect, desired, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
  return expect; /* value seen */
}

static int32_t rep_atomic(int32_t idx, int32_t* nstat) {
  int32_t curr = ald(&nstat[idx]);
  if (curr != idx) {
    int32_t next, prev = idx;
    while (curr > (next = ald(&nstat[curr]))) {
      ast(&nstat[prev], next);
      prev = curr;
      curr = next;
    }
  }
  return curr;
}

static int32_t hook_atomic(int32_t vstat, int32_t u, int32_t* nstat) {
  int32_t ostat = rep_atomic(u, nstat);
  while (vstat != ostat) {
    if (vstat < ostat) {
      const int32_t seen = acas(&nstat[ostat], ostat, vstat);
      if (seen == ostat) break;
      ostat = seen;
    } else {
      const int32_t seen = acas(&nstat[vstat], vstat, ostat);
      if (seen == vstat) break;
      vstat = seen;
    }
  }
  return vstat;
}

typedef struct {
  int32_t n;
  const int32_t* nindex;
  const int32_t* nlist;
  int32_t* nstat;
  int32_t* wl;
  int32_t topL, topH;   /* worklist ends (fetch-add) */
  int64_t next;         /* dynamic chunk cursor */
  int phase;
  /* high-degree edge tasks */
  int64_t ntasks;
  int64_t* task_first;  /* first task index of hub k (prefix) */
  int32_t nhubs;
} par_ctx;

enum { CHUNK_V = 1024, HUB_CHUNK = 4096 };

static void* par_worker(void* arg) {
  par_ctx* c = (par_ctx*)arg;
  const int32_t n = c->n;
  const int32_t* ni = c->nindex;
  const int32_t* nl = c->nlist;
  int32_t* ns = c->nstat;
  if (c->phase == 0) { /* init */
    for (;;) {
      const int64_t lo = __atomic_fetch_add(&c->next, CHUNK_V, __ATOMIC_RELAXED);
      if (lo >= n) break;
      const int32_t hi = (int32_t)(lo + CHUNK_V < n ? lo + CHUNK_V : n);
      for (int32_t v = (int32_t)lo; v < hi; v++) {
        int32_t m = v;
        for (int32_t i = ni[v]; i < ni[v + 1] && m == v; i++)
          if (nl[i] < m) m = nl[i];
        ns[v] = m;
      }
    }
  } else if (c->phase == 1) { /* compute1: low degree, fill worklists */
    for (;;) {
      const int64_t lo = __atomic_fetch_add(&c->next, CHUNK_V, __ATOMIC_RELAXED);
      if (lo >= n) break;
      const int32_t hi = (int32_t)(lo + CHUNK_V < n ? lo + CHUNK_V : n);
      for (int32_t v = (int32_t)lo; v < hi; v++) {
        if (ald(&ns[v]) == v) continue;
        const int32_t deg = ni[v + 1] - ni[v];
        if (deg > THREAD_MAX_DEG) {
          if (deg <= WARP_MAX_DEG) c->wl[__atomic_fetch_add(&c->topL, 1, __ATOMIC_RELAXED)] = v;
          else c->wl[__atomic_fetch_sub(&c->topH, 1, __ATOMIC_RELAXED)] = v;
        } else {
          int32_t vstat = rep_atomic(v, ns);
          for (int32_t i = ni[v]; i < ni[v + 1]; i++)
            if (v > nl[i]) vstat = hook_atomic(vstat, nl[i], ns);
        }
      }
    }
  } else if (c->phase == 2) { /* compute2: medium degree, whole vertices */
    for (;;) {
      const int64_t k = __atomic_fetch_add(&c->next, 16, __ATOMIC_RELAXED);
      if (k >= c->topL) break;
      const int64_t kend = k + 16 < c->topL ? k + 16 : c->topL;
      for (int64_t j = k; j < kend; j++) {
        const int32_t v = c->wl[j];
        int32_t vstat = rep_atomic(v, ns);
        for (int32_t i = ni[v]; i < ni[v + 1]; i++)
          if (v > nl[i]) vstat = hook_atomic(vstat, nl[i], ns);
      }
    }
  } else if (c->phase == 3) { /* compute3: high degree, edge chunks */
    for (;;) {
      const int64_t t = __atomic_fetch_add(&c->next, 1, __ATOMIC_RELAXED);
      if (t >= c->ntasks) break;
      int32_t lo = 0, hi = c->nhubs - 1; /* last hub with task_first <= t */
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (c->task_first[mid] <= t) lo = mid; else hi = mid - 1;
      }
      const int32_t v = c->wl[n - 1 - lo];
      const int64_t first = (int64_t)ni[v] + (t - c->task_first[lo]) * HUB_CHUNK;
      const int64_t last = first + HUB_CHUNK < ni[v + 1] ? first + HUB_CHUNK : ni[v + 1];
      int32_t vstat = rep_atomic(v, ns);
      for (int64_t i = first; i < last; i++)
        if (v > nl[i]) vstat = hook_atomic(vstat, nl[i], ns);
    }
  } else { /* flatten */
    for (;;) {
      const int64_t lo = __atomic_fetch_add(&c->next, CHUNK_V, __ATOMIC_RELAXED);
      if (lo >= n) break;
      const int32_t hi = (int32_t)(lo + CHUNK_V < n ? lo + CHUNK_V : n);
      for (int32_t v = (int32_t)lo; v < hi; v++) {
        int32_t s = ald(&ns[v]), next;
        const int32_t old = s;
        while (s > (next = ald(&ns[s]))) s = next;
        if (s != old) ast(&ns[v], s);
      }
    }
  }
  return NULL;
}

static int run_phase(par_ctx* c, int phase, int nthreads, pthread_t* th) {
  c->phase = phase;
  c->next = 0;
  int started = 0;
  for (int t = 1; t < nthreads; t++) {
    if (pthread_create(&th[t], NULL, par_worker, c) != 0) break;
    started++;
  }
  par_worker(c);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
  return 0;
}

int oracle_eclcc_threads(int32_t n, const int32_t* nindex, const int32_t* nlist, int32_t* nstat,
                         int nthreads) {
  if (n < 1 || !nindex || !nstat) return -1;
  if (nthreads <= 0) {
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    nthreads = c > 0 ? (int)c : 1;
  }
  par_ctx c;
  memset(&c, 0, sizeof c);
  c.n = n;
  c.nindex = nindex;
  c.nlist = nlist;
  c.nstat = nstat;
  c.wl = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  if (!c.wl || !th) { free(c.wl); free(th); return -3; }
  c.topL = 0;
  c.topH = n - 1;
  run_phase(&c, 0, nthreads, th);
  run_phase(&c, 1, nthreads, th);
  run_phase(&c, 2, nthreads, th);
  c.nhubs = (n - 1) - c.topH;
  c.task_first = (int64_t*)malloc(sizeof(int64_t) * (size_t)(c.nhubs + 1));
  if (!c.task_first) { free(c.wl); free(th); return -3; }
  int64_t acc = 0;
  for (int32_t k = 0; k < c.nhubs; k++) {
    const int32_t v = c.wl[n - 1 - k];
    c.task_first[k] = acc;
    acc += ((int64_t)(nindex[v + 1] - nindex[v]) + HUB_CHUNK - 1) / HUB_CHUNK;
  }
  c.ntasks = acc;
  if (c.nhubs > 0) run_phase(&c, 3, nthreads, th);
  run_phase(&c, 4, nthreads, th);
  free(c.task_first);
  free(c.wl);
  free(th);
  return nthreads;
}
