/* SPDX-License-Identifier: 0BSD
 *
 * The reference CPU solver's time loop on a thread pool, for the CPU
 * baseline.  TEST INFRASTRUCTURE ONLY: the benchmark's CPU-baseline leg and
 * the tests call it; nothing in the product links or calls it.
 *
 * It follows the shape of src/main.cpp:316-348 on Hedgehog: the driver
 * thread builds every row-strip tile of the step (make_tile,
 * main.cpp:125-155: a fresh buffer holding the owned rows and a neighbour
 * or mirrored edge row above and below), pushes them to the task, whose
 * `nthreads` clones each take tiles, write the updated tile into a new
 * buffer and drop the old one (StepTileTask::execute,
 * include/step_task.hpp:54-206), and the driver stitches the owned rows of
 * each result back into the global field (main.cpp:327-343).  The cell
 * arithmetic is ora_step_tile / ora_step_tile_f32, so the field equals the
 * full-grid oracle step for any tile and thread count.
 */
#define _POSIX_C_SOURCE 200809L
#include "lpbf_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct Pool {
    int prec, phys, nx, ntiles;
    const int* y0;
    const int* y1;
    StepConst k;            /* the current step's constants (laser set) */
    void** in;              /* per tile: the built tile (freed by the worker) */
    void** out;             /* per tile: the result (freed by the driver) */
    int* done;              /* per tile: result ready (under mu) */
    atomic_int next;        /* next tile to take */
    int gen, stop, failed;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
} Pool;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void* worker(void* arg) {
    Pool* p = (Pool*)arg;
    int seen = 0;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        while (!p->stop && p->gen == seen) pthread_cond_wait(&p->cv_work, &p->mu);
        if (p->stop) {
            pthread_mutex_unlock(&p->mu);
            return NULL;
        }
        seen = p->gen;
        pthread_mutex_unlock(&p->mu);
        for (;;) {
            const int i = atomic_fetch_add(&p->next, 1);
            if (i >= p->ntiles) break;
            const int owned = p->y1[i] - p->y0[i] + 1;
            const size_t n = (size_t)(owned + 2) * (size_t)p->nx;
            void* o = malloc(n * (p->prec == 32 ? sizeof(float) : sizeof(double)));
            if (o) {
                if (p->prec == 32)
                    ora_step_tile_f32((const float*)p->in[i], (float*)o, owned, p->y0[i], &p->k);
                else
                    ora_step_tile((const double*)p->in[i], (double*)o, owned, p->y0[i], &p->k,
                                  p->phys);
            }
            free(p->in[i]);
            p->in[i] = NULL;
            pthread_mutex_lock(&p->mu);
            p->out[i] = o;
            if (!o) p->failed = 1;
            p->done[i] = 1;
            pthread_cond_broadcast(&p->cv_done);
            pthread_mutex_unlock(&p->mu);
        }
    }
}

int ora_steptask_run(int prec, const StepConst* k, int phys, const LaserTick* plan,
                     int n_steps, int ntiles, int nthreads, void* T, double* phase_s) {
    if (!k || !plan || !T || n_steps < 0 || ntiles < 1 || nthreads < 1 ||
        (prec != 32 && prec != 64) || (prec == 32 && phys != LPBF_PHYS_BASELINE) ||
        k->nx < 1 || k->ny < 1)
        return -1;
    const int nx = k->nx, ny = k->ny;
    if (ntiles > ny) ntiles = ny;   /* main.cpp:237 */
    const size_t es = prec == 32 ? sizeof(float) : sizeof(double);
    int* y0 = (int*)malloc(sizeof(int) * (size_t)ntiles);
    int* y1 = (int*)malloc(sizeof(int) * (size_t)ntiles);
    void** in = (void**)calloc((size_t)ntiles, sizeof(void*));
    void** out = (void**)calloc((size_t)ntiles, sizeof(void*));
    int* done = (int*)calloc((size_t)ntiles, sizeof(int));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    if (!y0 || !y1 || !in || !out || !done || !th) {
        free(y0); free(y1); free(in); free(out); free(done); free(th);
        return -1;
    }
    ora_tile_bounds(ny, ntiles, y0, y1);
    Pool p;
    memset(&p, 0, sizeof p);
    p.prec = prec; p.phys = phys; p.nx = nx; p.ntiles = ntiles;
    p.y0 = y0; p.y1 = y1; p.in = in; p.out = out; p.done = done;
    p.k = *k;
    atomic_init(&p.next, 0);
    pthread_mutex_init(&p.mu, NULL);
    pthread_cond_init(&p.cv_work, NULL);
    pthread_cond_init(&p.cv_done, NULL);
    int started = 0;
    for (; started < nthreads; ++started)
        if (pthread_create(&th[started], NULL, worker, &p) != 0) break;
    int rc = started == nthreads ? 0 : -1;
    double t_build = 0.0, t_wait = 0.0;
    const double t0 = now_s();
    char* G = (char*)T;
    const size_t row = (size_t)nx * es;
    for (int s = 0; s < n_steps && rc == 0; ++s) {
        const double tb = now_s();
        for (int i = 0; i < ntiles; ++i) {   /* make_tile, main.cpp:125-155 */
            const int owned = y1[i] - y0[i] + 1, rows = owned + 2;
            char* t = (char*)malloc((size_t)rows * row);
            if (!t) { rc = -1; break; }
            const int top = (y0[i] == 0) ? y0[i] : y0[i] - 1;
            const int bot = (y1[i] == ny - 1) ? y1[i] : y1[i] + 1;
            memcpy(t, G + (size_t)top * row, row);
            memcpy(t + row, G + (size_t)y0[i] * row, (size_t)owned * row);
            memcpy(t + (size_t)(rows - 1) * row, G + (size_t)bot * row, row);
            in[i] = t;
        }
        if (rc) {
            for (int i = 0; i < ntiles; ++i) { free(in[i]); in[i] = NULL; }
            break;
        }
        const double tw = now_s();
        t_build += tw - tb;
        pthread_mutex_lock(&p.mu);
        p.k.x_l = plan[s].x_l;
        p.k.y_l = plan[s].y_l;
        p.k.q_scale = plan[s].q_scale;
        for (int i = 0; i < ntiles; ++i) done[i] = 0;
        atomic_store(&p.next, 0);
        p.gen++;
        pthread_cond_broadcast(&p.cv_work);
        pthread_mutex_unlock(&p.mu);
        for (int i = 0; i < ntiles; ++i) {   /* collect and stitch, main.cpp:327-343 */
            pthread_mutex_lock(&p.mu);
            while (!done[i]) pthread_cond_wait(&p.cv_done, &p.mu);
            void* o = out[i];
            out[i] = NULL;
            pthread_mutex_unlock(&p.mu);
            if (o) {
                const int owned = y1[i] - y0[i] + 1;
                memcpy(G + (size_t)y0[i] * row, (char*)o + row, (size_t)owned * row);
                free(o);
            }
        }
        t_wait += now_s() - tw;
        if (p.failed) rc = -1;
    }
    const double total = now_s() - t0;
    pthread_mutex_lock(&p.mu);
    p.stop = 1;
    pthread_cond_broadcast(&p.cv_work);
    pthread_mutex_unlock(&p.mu);
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&p.mu);
    pthread_cond_destroy(&p.cv_work);
    pthread_cond_destroy(&p.cv_done);
    if (phase_s) { phase_s[0] = t_build; phase_s[1] = t_wait; phase_s[2] = total; }
    free(y0); free(y1); free(in); free(out); free(done); free(th);
    return rc;
}
