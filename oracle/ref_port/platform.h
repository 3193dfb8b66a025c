/*
 * Linux implementation of the reference's OS interface (the plat_* names
 * listed in src/platform.h:14-38 of shwwwa/fastfib), written for this repo so
 * that the unmodified reference sources compile on Linux.
 *
 * The reference's porting procedure (src/platform.h:8-12) is "write
 * platform_<os>.h with the same names". Its platform.h has no Linux branch,
 * so oracle/Makefile puts this file in front of it with gcc's `-I-` (quoted
 * includes search oracle/ref_port first, and not the including file's own
 * directory); every other reference file is read where it lies.
 *
 * Test infrastructure only: this builds the CPU reference that checks the
 * GPU library and times the CPU baseline. Nothing in fastfib_amd links it.
 */
#pragma once

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <time.h>
#include <unistd.h>

/* Memory: reserve address space without backing, commit by mprotect. */
static inline void *plat_reserve(size_t bytes) {
    void *p = mmap(NULL, bytes, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    return p == MAP_FAILED ? NULL : p;
}

static inline int plat_commit(void *p, size_t bytes) {
    return mprotect(p, bytes, PROT_READ | PROT_WRITE) == 0;
}

/* Pages go back to the kernel; the range stays mapped read-write and reads
   back as zeros (contents lost, as the interface allows). */
static inline void plat_discard(void *p, size_t bytes) {
    madvise(p, bytes, MADV_DONTNEED);
}

static inline void *plat_aligned_alloc(size_t bytes, size_t align) {
    void *p = NULL;
    if (align < sizeof(void *)) align = sizeof(void *);
    return posix_memalign(&p, align, bytes ? bytes : 1) == 0 ? p : NULL;
}

static inline void plat_aligned_free(void *p) {
    free(p);
}

/* Locks and condition variables: pthreads. */
typedef pthread_mutex_t plat_lock;
#define PLAT_LOCK_INIT PTHREAD_MUTEX_INITIALIZER
static inline void plat_lock_acquire(plat_lock *l) { pthread_mutex_lock(l); }
static inline void plat_lock_release(plat_lock *l) { pthread_mutex_unlock(l); }

typedef pthread_cond_t plat_cond;
#define PLAT_COND_INIT PTHREAD_COND_INITIALIZER
static inline void plat_cond_wait(plat_cond *c, plat_lock *l) { pthread_cond_wait(c, l); }
static inline void plat_cond_signal(plat_cond *c) { pthread_cond_signal(c); }
static inline void plat_cond_broadcast(plat_cond *c) { pthread_cond_broadcast(c); }

typedef pthread_once_t plat_once;
#define PLAT_ONCE_INIT PTHREAD_ONCE_INIT
static inline void plat_once_run(plat_once *o, void (*fn)(void)) { pthread_once(o, fn); }

/* Detached threads: fn(arg) through pthread's one pointer. */
typedef struct {
    void (*fn)(void *);
    void *arg;
} PlatThreadStart;

static void *plat_thread_main(void *p) {
    PlatThreadStart s = *(PlatThreadStart *)p;
    free(p);
    s.fn(s.arg);
    return NULL;
}

static inline int plat_thread_start(void (*fn)(void *), void *arg) {
    PlatThreadStart *s = (PlatThreadStart *)malloc(sizeof *s);
    if (!s) return 0;
    s->fn = fn;
    s->arg = arg;
    pthread_t t;
    if (pthread_create(&t, NULL, plat_thread_main, s) != 0) {
        free(s);
        return 0;
    }
    pthread_detach(t);
    return 1;
}

static inline long plat_atomic_inc(volatile long *x) { return __atomic_add_fetch(x, 1, __ATOMIC_SEQ_CST); }
static inline long plat_atomic_dec(volatile long *x) { return __atomic_sub_fetch(x, 1, __ATOMIC_SEQ_CST); }
static inline long plat_atomic_load(volatile long *x) { return __atomic_load_n(x, __ATOMIC_ACQUIRE); }
static inline void plat_atomic_store(volatile long *x, long v) { __atomic_store_n(x, v, __ATOMIC_RELEASE); }

/* Logical CPUs this process may run on (its affinity mask). FF_REF_THREADS,
   when set to a positive number, caps it: a GPU host shares its cores, and
   the pool would otherwise start one spinning worker per visible CPU. */
static inline int plat_cpu_count(void) {
    int n = 1;
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n < 1) n = 1;
    const char *cap = getenv("FF_REF_THREADS");
    if (cap) {
        int c = atoi(cap);
        if (c > 0 && c < n) n = c;
    }
    return n;
}

static inline double plat_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
