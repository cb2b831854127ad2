/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).
 * Fork-join over pthreads; see par.h.
 */
#include "par.h"
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

static int g_threads = 1;

int par_threads(void) { return g_threads; }

void par_set_threads(int t) {
    if (t <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        t = c > 0 ? (int)c : 1;
    }
    g_threads = t;
}

typedef struct {
    par_body_fn body;
    void *ctx;
    size_t begin, end;
    int thread;
} par_job;

static void *par_tramp(void *p) {
    par_job *j = (par_job *)p;
    j->body(j->ctx, j->begin, j->end, j->thread);
    return NULL;
}

void par_for(size_t count, size_t grain, par_body_fn body, void *ctx) {
    if (count == 0) return;
    if (grain == 0) grain = 1;
    size_t units = (count + grain - 1) / grain;
    size_t t = (size_t)g_threads;
    if (t > units) t = units;
    if (t <= 1) {
        body(ctx, 0, count, 0);
        return;
    }
    size_t per = (units + t - 1) / t;
    par_job *jobs = (par_job *)calloc(t, sizeof(par_job));
    pthread_t *th = (pthread_t *)calloc(t, sizeof(pthread_t));
    size_t nj = 0;
    for (size_t i = 0; i < t; i++) {
        size_t b = i * per * grain, e = (i + 1) * per * grain;
        if (b >= count) break;
        if (e > count) e = count;
        jobs[nj].body = body;
        jobs[nj].ctx = ctx;
        jobs[nj].begin = b;
        jobs[nj].end = e;
        jobs[nj].thread = (int)nj;
        nj++;
    }
    for (size_t i = 1; i < nj; i++) pthread_create(&th[i], NULL, par_tramp, &jobs[i]);
    par_tramp(&jobs[0]);
    for (size_t i = 1; i < nj; i++) pthread_join(th[i], NULL);
    free(jobs);
    free(th);
}
