/*
 * oracle_hm.c -- multi-threaded FP64 restatement of NaiveDPDEngineHalfMergeTBB
 * (include/dpd/engines/naive/naive_dpd_engine_half_merge_tbb.hpp:151-262 on top
 * of naive_dpd_engine_half_merge.hpp:283-690). TEST INFRASTRUCTURE / CPU
 * BASELINE ONLY: bench.py times it on the host cores; tests pin it against the
 * FP64 naive restatement.
 *
 * Step (same order as the reference):
 *   1. t_hash = get_t_hash(t, seed)
 *   2. per cell, last bead first: half-step update_pos (dpd_maths_core_half_step.hpp:
 *      22-50), migrate to the destination cell's incoming list  [8 colour waves]
 *   3. append incoming lists
 *   4. per cell: 13 forward neighbours then intra-cell pairs, Newton-3 updates
 *      with dpd_maths_core_half_step::calc_force (:69-158)       [8 colour waves]
 *   5. bonds + angles per polymer (half_merge.hpp:638-690)
 *   6. v += dt/2 f (half_step.hpp:52-62); t += 1
 *
 * Waves: 2x2x2 super-cells on a 4-cell period give 8 colours whose members never
 * share a neighbour (naive_dpd_engine_half_merge_tbb.hpp:73-108). Boxes whose
 * sides are not multiples of 4 run the same step on one thread.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int32_t *v; int n, cap; } ivec;

static void ivec_push(ivec *a, int32_t x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 8;
        a->v = (int32_t *)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}

typedef struct {
    int pos[3];
    int is_edge;
    int nbr[13];
    ivec beads, incoming;
} hm_cell;

typedef struct hm_engine {
    orc_world *w;
    int dims[3];
    int ncells;
    hm_cell *cells;
    double sqrt_diss[256];
    double scale;            /* sqrt(24 kT / dt) */
    uint64_t t_hash;
    /* waves: 8 lists of super-cell base indices */
    int nwaves;
    int32_t *wave_items[8];
    int wave_n[8];
    int serial;
    /* thread pool */
    int nthreads;
    pthread_t *threads;
    pthread_barrier_t bar;
    atomic_int next;
    int phase;               /* 0 move, 1 forces, -1 exit */
    int cur_wave;
    int poly_lo, poly_hi;
} hm_engine;

static int cidx(const hm_engine *e, int x, int y, int z)
{
    return x * e->dims[1] * e->dims[2] + y * e->dims[2] + z;
}

static void move_cell(hm_engine *e, int ci)
{
    orc_world *w = e->w;
    hm_cell *c = &e->cells[ci];
    const double dt = w->dt;
    for (int bi = c->beads.n - 1; bi >= 0; bi--) {
        int b = c->beads.v[bi];
        if (w->stationary && w->stationary[w->bead_type[b]]) continue;
        double *x = &w->x[3 * b], *v = &w->v[3 * b], *f = &w->f[3 * b];
        int np[3];
        for (int d = 0; d < 3; d++) {
            double xd = (x[d] + v[d] * dt) + f[d] * (dt * dt * 0.5);
            if (xd < 0) {
                xd += w->box[d];
                if (xd == w->box[d]) xd = 0;
            } else if (xd >= w->box[d]) {
                xd -= w->box[d];
            }
            x[d] = xd;
            v[d] = v[d] + f[d] * (dt * 0.5);
            f[d] = 0;
            np[d] = (int)floor(xd);
        }
        int ni = cidx(e, np[0], np[1], np[2]);
        if (ni != ci) {
            ivec_push(&e->cells[ni].incoming, b);
            c->beads.v[bi] = c->beads.v[c->beads.n - 1];
            c->beads.n--;
        }
    }
}

static void transfer_incoming(hm_cell *c)
{
    for (int k = 0; k < c->incoming.n; k++) ivec_push(&c->beads, c->incoming.v[k]);
    c->incoming.n = 0;
}

static void pair(hm_engine *e, int hb, int ob, const double *dx_in, double dr2)
{
    orc_world *w = e->w;
    const int nt = w->n_types;
    double dr = sqrt(dr2);
    double inv_dr = 1.0 / dr;
    int ti = w->bead_type[hb] * nt + w->bead_type[ob];
    double con = w->conservative[ti];
    double sd = e->sqrt_diss[ti];
    double dv[3], dx[3];
    for (int d = 0; d < 3; d++) dv[d] = w->v[3 * hb + d] - w->v[3 * ob + d];
    double wr = 1.0 - dr;
    double conForce = con * wr;
    for (int d = 0; d < 3; d++) dx[d] = dx_in[d] * inv_dr;
    double rdotv = dx[0] * dv[0] + dx[1] * dv[1] + dx[2] * dv[2];
    double sg = sd * wr;
    double dissForce = -sg * sg * rdotv;
    double u = (double)orc_default_hash(e->t_hash, w->hash[hb], w->hash[ob]);
    double randForce = (sg * e->scale) * u;
    double scaled = conForce + dissForce + randForce + 0.0 * (0.0 - dr);
    for (int d = 0; d < 3; d++) {
        double fd = dx[d] * scaled;
        w->f[3 * hb + d] += fd;
        w->f[3 * ob + d] -= fd;
    }
}

static void forces_cell(hm_engine *e, int ci)
{
    orc_world *w = e->w;
    hm_cell *home = &e->cells[ci];
    /* forward neighbours, last first (naive_dpd_engine_half_merge_tbb.hpp:284-318) */
    for (int k = 12; k >= 0; k--) {
        hm_cell *oth = &e->cells[home->nbr[k]];
        double delta[3] = {0, 0, 0};
        if (home->is_edge) {
            for (int d = 0; d < 3; d++) {
                if (home->pos[d] == 0 && oth->pos[d] == e->dims[d] - 1) delta[d] = -e->dims[d];
                else if (home->pos[d] == e->dims[d] - 1 && oth->pos[d] == 0) delta[d] = e->dims[d];
            }
        }
        for (int oi = 0; oi < oth->beads.n; oi++) {
            int ob = oth->beads.v[oi];
            double ox[3];
            for (int d = 0; d < 3; d++) ox[d] = w->x[3 * ob + d] + delta[d];
            for (int hi = 0; hi < home->beads.n; hi++) {
                int hb = home->beads.v[hi];
                double dx[3];
                for (int d = 0; d < 3; d++) dx[d] = w->x[3 * hb + d] - ox[d];
                double dr2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
                if (dr2 >= 1 || dr2 < (double)(1e-9f * 1e-9f)) continue;
                pair(e, hb, ob, dx, dr2);
            }
        }
    }
    /* intra (naive_dpd_engine_half_merge.hpp:505-534) */
    for (int i = 0; i < home->beads.n - 1; i++) {
        int hb = home->beads.v[i];
        for (int j = i + 1; j < home->beads.n; j++) {
            int ob = home->beads.v[j];
            double dx[3];
            for (int d = 0; d < 3; d++) dx[d] = w->x[3 * hb + d] - w->x[3 * ob + d];
            double dr2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (dr2 >= 1 || dr2 < (double)(1e-9f * 1e-9f)) continue;
            pair(e, hb, ob, dx, dr2);
        }
    }
}

static void dist_from_to(const double *L, const double *base, const double *other, double *res)
{
    for (int i = 0; i < 3; i++) {
        double len = L[i];
        double dx = other[i] - base[i];
        int wrap_neg = dx > len * 0.5;
        int wrap_pos = dx < -len * 0.5;
        res[i] = dx + (wrap_pos - wrap_neg) * len;
    }
}

static void polymer_bonds(hm_engine *e, int p)
{
    orc_world *w = e->w;
    for (int bi = w->poly_bond_start[p]; bi < w->poly_bond_start[p + 1]; bi++) {
        int h = w->bond_head[bi], t = w->bond_tail[bi];
        double dx[3];
        dist_from_to(w->box, &w->x[3 * h], &w->x[3 * t], dx);
        double dr = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double hookeanForce = w->bond_kappa[bi] * (w->bond_r0[bi] - dr);
        for (int d = 0; d < 3; d++) {
            double fd = dx[d] * (hookeanForce / dr);
            w->f[3 * h + d] -= fd;
            w->f[3 * t + d] += fd;
        }
    }
    for (int ai = w->poly_angle_start[p]; ai < w->poly_angle_start[p + 1]; ai++) {
        int h = w->angle_head[ai], m = w->angle_mid[ai], t = w->angle_tail[ai];
        double a[3], b[3];
        dist_from_to(w->box, &w->x[3 * h], &w->x[3 * m], a);
        dist_from_to(w->box, &w->x[3 * m], &w->x[3 * t], b);
        double r01 = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double r12 = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        double hf[3] = {0, 0, 0}, tf[3] = {0, 0, 0}, mf[3] = {0, 0, 0};
        double mag = r01 * r12;
        if (mag > 0.0001) {
            double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            double o1 = dot / (r01 * r01), o2 = dot / (r12 * r12);
            double cos2 = o1 * o2;
            double fm = w->angle_kappa[ai] / mag;
            double s0 = sin(w->angle_theta0[ai]);
            if (fabs(dot) > 0.000001 && s0 > 0) {
                double ip = 1.0 / sqrt(1.0 / cos2 - 1.0);
                fm = fm * (cos(w->angle_theta0[ai]) - s0 * ip);
            }
            for (int d = 0; d < 3; d++) {
                hf[d] = ((a[d] * o1) - b[d]) * fm;
                tf[d] = (a[d] - (b[d] * o2)) * fm;
                mf[d] = -(hf[d] + tf[d]);
            }
        }
        for (int d = 0; d < 3; d++) {
            w->f[3 * h + d] += hf[d];
            w->f[3 * t + d] += tf[d];
            w->f[3 * m + d] += mf[d];
        }
    }
}

/* Run every super-cell of one wave: the 2x2x2 block based at item. */
static void run_super(hm_engine *e, int base, int phase)
{
    int bx = base / (e->dims[1] * e->dims[2]);
    int by = (base / e->dims[2]) % e->dims[1];
    int bz = base % e->dims[2];
    for (int lx = 0; lx < 2; lx++)
    for (int ly = 0; ly < 2; ly++)
    for (int lz = 0; lz < 2; lz++) {
        int ci = cidx(e, bx + lx, by + ly, bz + lz);
        if (phase == 0) move_cell(e, ci);
        else forces_cell(e, ci);
    }
}

static void *worker(void *arg)
{
    hm_engine *e = (hm_engine *)arg;
    for (;;) {
        pthread_barrier_wait(&e->bar);
        if (e->phase < 0) break;
        if (e->phase <= 1) {
            int wv = e->cur_wave;
            for (;;) {
                int k = atomic_fetch_add(&e->next, 1);
                if (k >= e->wave_n[wv]) break;
                run_super(e, e->wave_items[wv][k], e->phase);
            }
        } else if (e->phase == 2) {
            for (;;) {
                int k = atomic_fetch_add(&e->next, 64);
                if (k >= e->ncells) break;
                int hi = k + 64 < e->ncells ? k + 64 : e->ncells;
                for (int c = k; c < hi; c++) transfer_incoming(&e->cells[c]);
            }
        } else if (e->phase == 3) {
            for (;;) {
                int k = atomic_fetch_add(&e->next, 256);
                if (k >= e->w->n_polymers) break;
                int hi = k + 256 < e->w->n_polymers ? k + 256 : e->w->n_polymers;
                for (int p = k; p < hi; p++) polymer_bonds(e, p);
            }
        } else if (e->phase == 4) {
            orc_world *w = e->w;
            for (;;) {
                int k = atomic_fetch_add(&e->next, 1024);
                if (k >= w->n_beads) break;
                int hi = k + 1024 < w->n_beads ? k + 1024 : w->n_beads;
                for (int i = k; i < hi; i++) {
                    if (w->stationary && w->stationary[w->bead_type[i]]) {
                        for (int d = 0; d < 3; d++) { w->v[3 * i + d] = 0; w->f[3 * i + d] = 0; }
                    } else {
                        for (int d = 0; d < 3; d++) w->v[3 * i + d] += w->f[3 * i + d] * (w->dt * 0.5);
                    }
                }
            }
        }
        pthread_barrier_wait(&e->bar);
    }
    return NULL;
}

static void run_phase(hm_engine *e, int phase, int wave)
{
    e->phase = phase;
    e->cur_wave = wave;
    atomic_store(&e->next, 0);
    pthread_barrier_wait(&e->bar);   /* release workers */
    /* main thread participates */
    if (phase <= 1) {
        for (;;) {
            int k = atomic_fetch_add(&e->next, 1);
            if (k >= e->wave_n[wave]) break;
            run_super(e, e->wave_items[wave][k], phase);
        }
    } else if (phase == 2) {
        for (;;) {
            int k = atomic_fetch_add(&e->next, 64);
            if (k >= e->ncells) break;
            int hi = k + 64 < e->ncells ? k + 64 : e->ncells;
            for (int c = k; c < hi; c++) transfer_incoming(&e->cells[c]);
        }
    } else if (phase == 3) {
        for (;;) {
            int k = atomic_fetch_add(&e->next, 256);
            if (k >= e->w->n_polymers) break;
            int hi = k + 256 < e->w->n_polymers ? k + 256 : e->w->n_polymers;
            for (int p = k; p < hi; p++) polymer_bonds(e, p);
        }
    } else if (phase == 4) {
        orc_world *w = e->w;
        for (;;) {
            int k = atomic_fetch_add(&e->next, 1024);
            if (k >= w->n_beads) break;
            int hi = k + 1024 < w->n_beads ? k + 1024 : w->n_beads;
            for (int i = k; i < hi; i++) {
                if (w->stationary && w->stationary[w->bead_type[i]]) {
                    for (int d = 0; d < 3; d++) { w->v[3 * i + d] = 0; w->f[3 * i + d] = 0; }
                } else {
                    for (int d = 0; d < 3; d++) w->v[3 * i + d] += w->f[3 * i + d] * (w->dt * 0.5);
                }
            }
        }
    }
    pthread_barrier_wait(&e->bar);   /* wait for workers */
}

static void hm_step_serial(hm_engine *e)
{
    orc_world *w = e->w;
    for (int c = 0; c < e->ncells; c++) move_cell(e, c);
    for (int c = 0; c < e->ncells; c++) transfer_incoming(&e->cells[c]);
    for (int c = 0; c < e->ncells; c++) forces_cell(e, c);
    for (int p = 0; p < w->n_polymers; p++) polymer_bonds(e, p);
    for (int i = 0; i < w->n_beads; i++) {
        if (w->stationary && w->stationary[w->bead_type[i]]) {
            for (int d = 0; d < 3; d++) { w->v[3 * i + d] = 0; w->f[3 * i + d] = 0; }
        } else {
            for (int d = 0; d < 3; d++) w->v[3 * i + d] += w->f[3 * i + d] * (w->dt * 0.5);
        }
    }
}

int orc_hm_run(orc_world *w, int nsteps, int nthreads)
{
    hm_engine E;
    memset(&E, 0, sizeof(E));
    hm_engine *e = &E;
    e->w = w;
    for (int d = 0; d < 3; d++) {
        if (w->box[d] != floor(w->box[d]) || w->box[d] < 3) return -1;
        e->dims[d] = (int)w->box[d];
    }
    e->ncells = e->dims[0] * e->dims[1] * e->dims[2];
    e->cells = (hm_cell *)calloc((size_t)e->ncells, sizeof(hm_cell));
    for (int i = 0; i < w->n_types * w->n_types; i++) e->sqrt_diss[i] = sqrt(w->dissipative[i]);
    e->scale = sqrt(24 * 1.0 / w->dt);
    e->serial = (e->dims[0] % 4) || (e->dims[1] % 4) || (e->dims[2] % 4) || nthreads <= 1;

    /* half of the 26-neighbourhood: make_nhood.hpp:40-49 keeps the first of each
     * {d,-d} met in x,y,z order, i.e. the 13 offsets before the centre. */
    int fwd[13][3], nf = 0;
    for (int x = -1; x <= 1; x++)
    for (int y = -1; y <= 1; y++)
    for (int z = -1; z <= 1; z++) {
        int key = (x + 1) * 9 + (y + 1) * 3 + (z + 1);
        if (key < 13) { fwd[nf][0] = x; fwd[nf][1] = y; fwd[nf][2] = z; nf++; }
    }
    for (int c = 0; c < e->ncells; c++) {
        hm_cell *cc = &e->cells[c];
        cc->pos[0] = c / (e->dims[1] * e->dims[2]);
        cc->pos[1] = (c / e->dims[2]) % e->dims[1];
        cc->pos[2] = c % e->dims[2];
        cc->is_edge = 0;
        for (int d = 0; d < 3; d++) cc->is_edge |= cc->pos[d] == 0 || cc->pos[d] == e->dims[d] - 1;
        for (int k = 0; k < 13; k++) {
            int p[3];
            for (int d = 0; d < 3; d++) p[d] = (cc->pos[d] + fwd[k][d] + e->dims[d]) % e->dims[d];
            cc->nbr[k] = cidx(e, p[0], p[1], p[2]);
        }
    }
    for (int i = 0; i < w->n_beads; i++) {
        int p[3];
        for (int d = 0; d < 3; d++) p[d] = (int)floor(w->x[3 * i + d]);
        ivec_push(&e->cells[cidx(e, p[0], p[1], p[2])].beads, i);
    }

    if (!e->serial) {
        int g = 0;
        for (int gx = 0; gx < 4; gx += 2)
        for (int gy = 0; gy < 4; gy += 2)
        for (int gz = 0; gz < 4; gz += 2) {
            int cnt = (e->dims[0] / 4) * (e->dims[1] / 4) * (e->dims[2] / 4);
            e->wave_items[g] = (int32_t *)malloc(sizeof(int32_t) * (size_t)cnt);
            int n = 0;
            for (int ix = gx; ix < e->dims[0]; ix += 4)
            for (int iy = gy; iy < e->dims[1]; iy += 4)
            for (int iz = gz; iz < e->dims[2]; iz += 4) e->wave_items[g][n++] = cidx(e, ix, iy, iz);
            e->wave_n[g] = n;
            g++;
        }
        e->nwaves = 8;
        e->nthreads = nthreads;
        pthread_barrier_init(&e->bar, NULL, (unsigned)nthreads);
        e->threads = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int k = 1; k < nthreads; k++) pthread_create(&e->threads[k], NULL, worker, e);
    }

    for (int s = 0; s < nsteps; s++) {
        e->t_hash = orc_t_hash((uint32_t)w->t, w->seed);
        if (e->serial) {
            hm_step_serial(e);
        } else {
            for (int g = 0; g < 8; g++) run_phase(e, 0, g);
            run_phase(e, 2, 0);
            for (int g = 0; g < 8; g++) run_phase(e, 1, g);
            run_phase(e, 3, 0);
            run_phase(e, 4, 0);
        }
        w->t += 1;
    }

    if (!e->serial) {
        e->phase = -1;
        pthread_barrier_wait(&e->bar);
        for (int k = 1; k < e->nthreads; k++) pthread_join(e->threads[k], NULL);
        pthread_barrier_destroy(&e->bar);
        free(e->threads);
        for (int g = 0; g < 8; g++) free(e->wave_items[g]);
    }
    for (int c = 0; c < e->ncells; c++) {
        free(e->cells[c].beads.v);
        free(e->cells[c].incoming.v);
    }
    free(e->cells);
    return 0;
}
