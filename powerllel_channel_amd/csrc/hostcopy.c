/*
 * A small pool of host threads that split one memcpy between them.
 *
 * Host-pointer solves move the haloed pressure array between the caller's
 * (pageable) memory and pinned staging buffers; one core copies about
 * 10 GB/s, several keep up with the DMA engines.  The workers are created
 * once (with the solver's first host-pointer call) and wait on a condition
 * variable between copies, so a copy costs no thread start-up.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "pll_internal.h"
#include "powerllel.h"

#define POOL_MAX 64

struct pll_copy_pool {
    pthread_t th[POOL_MAX];
    int n;                 /* worker threads; the caller copies a share too */
    pthread_mutex_t mu;
    pthread_cond_t go, done;
    unsigned long gen;     /* bumped for every copy */
    int pending;           /* workers still copying the current one */
    int quit;
    char *dst;
    const char *src;
    size_t bytes, piece;
};

/* share `idx` of n + 1 (the caller takes share n) */
static void copy_share(struct pll_copy_pool *p, int idx)
{
    const size_t lo = (size_t)idx * p->piece;
    if (lo >= p->bytes) return;
    size_t len = p->bytes - lo;
    if (len > p->piece) len = p->piece;
    memcpy(p->dst + lo, p->src + lo, len);
}

typedef struct {
    struct pll_copy_pool *p;
    int idx;
} worker_arg;

static void *worker(void *arg)
{
    worker_arg *wa = (worker_arg *)arg;
    struct pll_copy_pool *p = wa->p;
    const int idx = wa->idx;
    free(wa);
    unsigned long seen = 0;
    pthread_mutex_lock(&p->mu);
    for (;;) {
        while (p->gen == seen && !p->quit) pthread_cond_wait(&p->go, &p->mu);
        if (p->quit) break;
        seen = p->gen;
        pthread_mutex_unlock(&p->mu);
        copy_share(p, idx);
        pthread_mutex_lock(&p->mu);
        if (--p->pending == 0) pthread_cond_signal(&p->done);
    }
    pthread_mutex_unlock(&p->mu);
    return NULL;
}

int pll_copy_pool_create(pll_copy_pool **out, int nthreads)
{
    *out = NULL;
    struct pll_copy_pool *p = (struct pll_copy_pool *)calloc(1, sizeof *p);
    if (!p) return PLL_ENOMEM;
    pthread_mutex_init(&p->mu, NULL);
    pthread_cond_init(&p->go, NULL);
    pthread_cond_init(&p->done, NULL);
    if (nthreads > POOL_MAX + 1) nthreads = POOL_MAX + 1;
    for (int i = 0; i < nthreads - 1; i++) {
        worker_arg *wa = (worker_arg *)malloc(sizeof *wa);
        if (!wa) break;
        wa->p = p;
        wa->idx = i;
        if (pthread_create(&p->th[i], NULL, worker, wa) != 0) {
            free(wa);
            break;
        }
        p->n++;
    }
    *out = p;
    return 0;
}

void pll_copy_pool_memcpy(pll_copy_pool *p, void *dst, const void *src, size_t n)
{
    if (!p || p->n == 0 || n < ((size_t)4 << 20)) {
        memcpy(dst, src, n);
        return;
    }
    pthread_mutex_lock(&p->mu);
    p->dst = (char *)dst;
    p->src = (const char *)src;
    p->bytes = n;
    /* shares of whole 4 KiB pages */
    p->piece = ((n + (size_t)p->n) / ((size_t)p->n + 1) + 4095) & ~(size_t)4095;
    p->pending = p->n;
    p->gen++;
    pthread_cond_broadcast(&p->go);
    pthread_mutex_unlock(&p->mu);
    copy_share(p, p->n);
    pthread_mutex_lock(&p->mu);
    while (p->pending > 0) pthread_cond_wait(&p->done, &p->mu);
    pthread_mutex_unlock(&p->mu);
}

void pll_copy_pool_destroy(pll_copy_pool *p)
{
    if (!p) return;
    pthread_mutex_lock(&p->mu);
    p->quit = 1;
    pthread_cond_broadcast(&p->go);
    pthread_mutex_unlock(&p->mu);
    for (int i = 0; i < p->n; i++) pthread_join(p->th[i], NULL);
    pthread_mutex_destroy(&p->mu);
    pthread_cond_destroy(&p->go);
    pthread_cond_destroy(&p->done);
    free(p);
}
