/*
 * oracle.c -- CPU restatement of the poets-dpd step (test infrastructure only;
 * see oracle.h for what each part restates and what pins it).
 *
 * Build: oracle/Makefile, gcc -O2 -mfma -ffp-contract=off. Contraction is off so
 * that every fused multiply-add is an explicit fmaf()/fma() call: the FP32 part
 * then mirrors the HIP kernels operation for operation, and the FP64 part
 * mirrors the reference's unfused expression order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <pthread.h>

/* ------------------------------------------------------------------------ */
/* RNG: include/dpd/core/hash.hpp:20-32 (riscv_mix64_m2), 61-65 (get_t_hash), */
/* 81-88 (hash_rng_sym); include/dpd/maths/dpd_maths_core.hpp:55-79           */
/* (default_hash); naive_dpd_engine.hpp:413-435 (RandSym, FP64 scale).        */
/* ------------------------------------------------------------------------ */

uint64_t orc_mix64_m2(uint64_t x)
{
    const uint64_t C = 0xd392d2a7u;   /* 32-bit constant, zero-extended */
    x ^= x >> 32;
    x *= C;
    x ^= x >> 32;
    x *= C;
    x ^= x >> 32;
    return x;
}

uint64_t orc_t_hash(uint32_t t, uint64_t seed)
{
    return orc_mix64_m2(seed + orc_mix64_m2((uint64_t)t));
}

uint32_t orc_hash_rng_sym(uint64_t t_hash, uint32_t a, uint32_t b)
{
    if (a > b) { uint32_t tmp = a; a = b; b = tmp; }
    return (uint32_t)orc_mix64_m2(((uint64_t)a | ((uint64_t)b << 32)) ^ t_hash);
}

float orc_default_hash(uint64_t t_hash, uint32_t a, uint32_t b)
{
    uint32_t ru = orc_hash_rng_sym(t_hash, a, b);
    int32_t rs;
    memcpy(&rs, &ru, 4);                      /* two's complement reinterpretation */
    return (float)rs * 0.00000000023283064365386962890625f;   /* 2^-32 */
}

double orc_rand_sym_f64(uint64_t t_hash, uint32_t a, uint32_t b)
{
    uint32_t ru = orc_hash_rng_sym(t_hash, a, b);
    int32_t rs;
    memcpy(&rs, &ru, 4);
    return rs * ldexp(1.0, -32);
}

/* orc_default_hash for the pair-force loop (same arithmetic, inlined). */
static inline float default_hash_inl(uint64_t t_hash, uint32_t a, uint32_t b)
{
    if (a > b) { uint32_t tmp = a; a = b; b = tmp; }
    uint32_t ru = (uint32_t)orc_mix64_m2(((uint64_t)a | ((uint64_t)b << 32)) ^ t_hash);
    int32_t rs;
    memcpy(&rs, &ru, 4);
    return (float)rs * 0.00000000023283064365386962890625f;
}

/* include/dpd/core/dpd_state.hpp:18-60: bbbb 1ppp.. (monomer) / bbbb 0ooo oooo pppp.. */
uint32_t orc_bead_hash(uint32_t bead_type, int is_monomer, uint32_t polymer_id, uint32_t polymer_offset)
{
    uint32_t base = (polymer_offset << 20) | polymer_id;
    base |= (uint32_t)(is_monomer ? 1u : 0u) << 27;
    base |= bead_type << 28;
    return base;
}

void orc_t_hash_n(int n, const uint32_t *t, const uint64_t *seed, uint64_t *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_t_hash(t[i], seed[i]);
}

void orc_hash_rng_sym_n(int n, const uint64_t *t_hash, const uint32_t *a, const uint32_t *b,
                        uint32_t *out_u32, float *out_f32)
{
    for (int i = 0; i < n; i++) {
        if (out_u32) out_u32[i] = orc_hash_rng_sym(t_hash[i], a[i], b[i]);
        if (out_f32) out_f32[i] = orc_default_hash(t_hash[i], a[i], b[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 NaiveDPDEngine<false> (naive_dpd_engine.hpp)                          */
/* ------------------------------------------------------------------------ */

typedef struct { int32_t *start; int32_t *items; int ncells; } cell_lists;

/* Bucket beads by cell, preserving bead-index order inside a cell (the naive
 * engine pushes beads into per-cell vectors in bead order: :295-310). */
static void build_cells_f64(const orc_world *w, const int *dims, cell_lists *cl)
{
    int nc = dims[0] * dims[1] * dims[2];
    cl->ncells = nc;
    cl->start = (int32_t *)calloc((size_t)nc + 1, sizeof(int32_t));
    cl->items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w->n_beads > 0 ? w->n_beads : 1));
    int32_t *cell_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w->n_beads > 0 ? w->n_beads : 1));
    for (int i = 0; i < w->n_beads; i++) {
        int c[3];
        for (int d = 0; d < 3; d++) {
            c[d] = (int)floor(w->x[3 * i + d]);
            if (c[d] < 0) c[d] = 0;
            if (c[d] >= dims[d]) c[d] = dims[d] - 1;
        }
        /* index = x*Ly*Lz + y*Lz + z (naive_dpd_engine.hpp:210-213) */
        int idx = c[0] * dims[1] * dims[2] + c[1] * dims[2] + c[2];
        cell_of[i] = idx;
        cl->start[idx + 1]++;
    }
    for (int c = 0; c < nc; c++) cl->start[c + 1] += cl->start[c];
    int32_t *fill = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
    for (int i = 0; i < w->n_beads; i++) {
        int c = cell_of[i];
        cl->items[cl->start[c] + fill[c]++] = i;
    }
    free(fill);
    free(cell_of);
}

static void free_cells(cell_lists *cl)
{
    free(cl->start);
    free(cl->items);
}

static int is_stationary(const orc_world *w, int i)
{
    return w->stationary && w->stationary[w->bead_type[i]];
}

/* calc_distance_from_to (naive_dpd_engine.hpp:243-254) */
static void dist_from_to_f64(const double *L, const double *base, const double *other, double *res)
{
    for (int i = 0; i < 3; i++) {
        double len = L[i];
        double dx = other[i] - base[i];
        int wrap_neg = dx > len * 0.5;
        int wrap_pos = dx < -len * 0.5;
        res[i] = dx + (wrap_pos - wrap_neg) * len;
    }
}

static double l2_norm3(const double *x)
{
    return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
}

/* Bonds and angles of one polymer, naive form (:567-742). */
static void naive_polymer_bonds(const orc_world *w, int p, double *F)
{
    for (int bi = w->poly_bond_start[p]; bi < w->poly_bond_start[p + 1]; bi++) {
        int h = w->bond_head[bi], t = w->bond_tail[bi];
        double dx[3];
        dist_from_to_f64(w->box, &w->x[3 * h], &w->x[3 * t], dx);
        double r = l2_norm3(dx);
        double force = w->bond_kappa[bi] * (r - w->bond_r0[bi]);
        double s = force / r;
        for (int d = 0; d < 3; d++) {
            double fd = dx[d] * s;
            F[3 * h + d] += fd;
            F[3 * t + d] -= fd;
        }
    }
    for (int ai = w->poly_angle_start[p]; ai < w->poly_angle_start[p + 1]; ai++) {
        int h = w->angle_head[ai], m = w->angle_mid[ai], t = w->angle_tail[ai];
        double first[3], second[3];
        dist_from_to_f64(w->box, &w->x[3 * h], &w->x[3 * m], first);
        dist_from_to_f64(w->box, &w->x[3 * m], &w->x[3 * t], second);
        double FirstLength = l2_norm3(first);
        double SecondLength = l2_norm3(second);
        double hf[3] = {0, 0, 0}, mf[3] = {0, 0, 0}, tf[3] = {0, 0, 0};
        double magProduct = FirstLength * SecondLength;
        if (magProduct > 0.0001) {
            double b1MagSq = FirstLength * FirstLength;
            double b2MagSq = SecondLength * SecondLength;
            double b1Dotb2 = first[0] * second[0] + first[1] * second[1] + first[2] * second[2];
            double b1b2Overb1Sq = b1Dotb2 / b1MagSq;
            double b1b2Overb2Sq = b1Dotb2 / b2MagSq;
            double cosPhiSq = b1b2Overb1Sq * b1b2Overb2Sq;
            double Modulus = w->angle_kappa[ai];
            double theta0 = w->angle_theta0[ai];
            double forceMag;
            if (fabs(b1Dotb2) > 0.000001) {
                double Prefactor = sqrt(1.0 / cosPhiSq - 1.0);
                Prefactor = (Prefactor < 0.000001) ? 0.000001 : Prefactor;   /* std::max(a,b) */
                double CosPhi0 = cos(theta0);
                if (theta0 > 0.0) {
                    double SinPhi0 = sin(theta0);
                    forceMag = Modulus * (CosPhi0 - SinPhi0 / Prefactor) / magProduct;
                } else {
                    forceMag = Modulus / magProduct;
                }
            } else {
                forceMag = Modulus / magProduct;
            }
            for (int d = 0; d < 3; d++) {
                hf[d] = ((first[d] * b1b2Overb1Sq) - second[d]) * forceMag;
                tf[d] = (first[d] - (second[d] * b1b2Overb2Sq)) * forceMag;
                mf[d] = -(hf[d] + tf[d]);
            }
        }
        for (int d = 0; d < 3; d++) {
            F[3 * h + d] += hf[d];
            F[3 * t + d] += tf[d];
            F[3 * m + d] += mf[d];
        }
    }
}

static void naive_step(orc_world *w, double *F)
{
    int dims[3];
    for (int d = 0; d < 3; d++) dims[d] = (int)w->box[d];
    const int nt = w->n_types;
    const double dt = w->dt;
    const uint64_t t_hash = orc_t_hash((uint32_t)w->t, w->seed);
    const int n = w->n_beads;

    memset(F, 0, sizeof(double) * 3 * (size_t)n);

    /* update_bead_pos (:746-789) */
    for (int i = 0; i < n; i++) {
        if (is_stationary(w, i)) continue;
        double *x = &w->x[3 * i], *v = &w->v[3 * i], *f = &w->f[3 * i];
        for (int d = 0; d < 3; d++) {
            double xd = (x[d] + v[d] * dt) + f[d] * (dt * dt / 2);
            int add = xd < 0;
            int sub = xd >= w->box[d];
            double adj = (add - sub) * w->box[d];
            double xa = xd + adj;
            double top = nextafter(w->box[d], -DBL_MAX);
            xa = (top < xa) ? top : xa;      /* std::min(xa, nexttoward(box,-DBL_MAX)) */
            x[d] = xa;
        }
    }

    cell_lists cl;
    build_cells_f64(w, dims, &cl);

    const double lambda_dt = w->lambda * w->dt;
    const double scaled_inv_root_dt = sqrt(1.0 * 24 / dt);

    /* step_cell over x,y,z with the full 27-cell stencil (:348-392) */
    int pos[3];
    for (pos[0] = 0; pos[0] < dims[0]; pos[0]++)
    for (pos[1] = 0; pos[1] < dims[1]; pos[1]++)
    for (pos[2] = 0; pos[2] < dims[2]; pos[2]++) {
        int hc = pos[0] * dims[1] * dims[2] + pos[1] * dims[2] + pos[2];
        int dir[3];
        for (dir[0] = -1; dir[0] <= 1; dir[0]++)
        for (dir[1] = -1; dir[1] <= 1; dir[1]++)
        for (dir[2] = -1; dir[2] <= 1; dir[2]++) {
            int op[3];
            double delta[3];
            for (int d = 0; d < 3; d++) {     /* make_relative_cell_pos (:226-241) */
                int raw = pos[d] + dir[d];
                int left = raw < 0;
                int right = raw == dims[d];
                op[d] = raw + (left - right) * dims[d];
                delta[d] = (right - left) * w->box[d];
            }
            int oc = op[0] * dims[1] * dims[2] + op[1] * dims[2] + op[2];
            for (int hk = cl.start[hc]; hk < cl.start[hc + 1]; hk++) {
                int hb = cl.items[hk];
                for (int ok = cl.start[oc]; ok < cl.start[oc + 1]; ok++) {
                    int ob = cl.items[ok];
                    if (hb == ob) continue;
                    /* update_bead_forces (:437-565) */
                    const double *hx = &w->x[3 * hb], *ox = &w->x[3 * ob];
                    double dx[3];
                    for (int d = 0; d < 3; d++) dx[d] = (hx[d] - ox[d]) - delta[d];
                    double dr = l2_norm3(dx);
                    double inv_dr = 1.0 / dr;
                    if (dr >= 1 || dr < 0.000000001) continue;
                    double dv[3];
                    for (int d = 0; d < 3; d++) {
                        double hv = w->v[3 * hb + d] + w->f[3 * hb + d] * lambda_dt;
                        double ov = w->v[3 * ob + d] + w->f[3 * ob + d] * lambda_dt;
                        dv[d] = hv - ov;
                    }
                    double wr = (1.0 - dr);
                    double wr2 = wr * wr;
                    int ti = w->bead_type[hb] * nt + w->bead_type[ob];
                    double conForce = w->conservative[ti] * wr;
                    double rdotv = (dx[0] * dv[0] + dx[1] * dv[1] + dx[2] * dv[2]) * inv_dr;
                    double gammap = w->dissipative[ti] * wr2;
                    double dissForce = -gammap * rdotv;
                    double u = orc_rand_sym_f64(t_hash, w->hash[hb], w->hash[ob]);
                    double randScale = sqrt(gammap) * scaled_inv_root_dt;
                    double randForce = randScale * u;
                    double scaled_force = (conForce + dissForce + randForce) * inv_dr;
                    for (int d = 0; d < 3; d++) F[3 * hb + d] += dx[d] * scaled_force;
                }
            }
        }
    }
    free_cells(&cl);

    for (int p = 0; p < w->n_polymers; p++) naive_polymer_bonds(w, p, F);

    /* update_bead_mom (:791-824) */
    for (int i = 0; i < n; i++) {
        double *v = &w->v[3 * i], *f = &w->f[3 * i];
        if (is_stationary(w, i)) {
            for (int d = 0; d < 3; d++) { v[d] = 0; f[d] = 0; }
            continue;
        }
        for (int d = 0; d < 3; d++) {
            double fNow = F[3 * i + d];
            v[d] = v[d] + (f[d] + fNow) * (dt / 2);
            f[d] = fNow;
        }
    }
    w->t += 1;
}

int orc_naive_run(orc_world *w, int nsteps)
{
    for (int d = 0; d < 3; d++) {
        if (w->box[d] != floor(w->box[d]) || w->box[d] < 2) return -1;
    }
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)(w->n_beads > 0 ? w->n_beads : 1));
    for (int s = 0; s < nsteps; s++) naive_step(w, F);
    free(F);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 restatement of the HIP engine                                         */
/* ------------------------------------------------------------------------ */

#define ORC_MIN_R2 (1e-9f * 1e-9f)   /* MIN_DISTANCE_CUTOFF_SQR, dpd_state.hpp:138-139 */

/* Fused finish-kick + drift + half-kick + wrap, per component:
 *   v  = w + f*dt/2            (update_mom of the previous step, raw:114-124)
 *   v' = v + f*dt/2            (half-kick, raw:70)
 *   x' = x + dt*v'             (drift, raw:54-58, written as (f*dt/2 + v)*dt + x)
 *   wrap: x += (x<0?L:0); x -= (x>=L?L:0)   (raw:61-68)
 * w is replaced by v'. */
static int f32_stationary(const orc_f32_world *w, int i)
{
    return w->stationary && w->stationary[w->hash[i] >> 28];
}

void orc_f32_integrate(orc_f32_world *w)
{
    const float dt = w->dt;
    const float hdt = 0.5f * dt;
    for (int i = 0; i < w->n_beads; i++) {
        if (f32_stationary(w, i)) {
            /* x stays; the pair force sees v(t) = w + f dt/2 (naive_dpd_engine_half_merge.hpp:327-331) */
            for (int d = 0; d < 3; d++) w->w[3 * i + d] = fmaf(w->f[3 * i + d], hdt, w->w[3 * i + d]);
            continue;
        }
        for (int d = 0; d < 3; d++) {
            float Lf = (float)w->L[d];
            float fd = w->f[3 * i + d];
            float v = fmaf(fd, hdt, w->w[3 * i + d]);
            float vh = fmaf(fd, hdt, v);
            float x = fmaf(vh, dt, w->x[3 * i + d]);
            x += (x < 0.0f) ? Lf : 0.0f;
            x -= (x >= Lf) ? Lf : 0.0f;
            w->x[3 * i + d] = x;
            w->w[3 * i + d] = vh;
        }
    }
}

void orc_f32_import(int n, const double *x, const double *v, const double *f, const int32_t *L, float dt,
                    float *x32, float *w32, float *f32)
{
    const float hdt = 0.5f * dt;
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            const float Lf = (float)L[d];
            float xx = (float)x[3 * i + d];
            if (xx >= Lf) xx = nextafterf(Lf, 0.0f);
            float fv = (float)f[3 * i + d];
            x32[3 * i + d] = xx;
            f32[3 * i + d] = fv;
            w32[3 * i + d] = fmaf(fv, -hdt, (float)v[3 * i + d]);
        }
    }
}

static int cell_key_f32(const orc_f32_world *w, int i, int *c3)
{
    for (int d = 0; d < 3; d++) {
        float x = w->x[3 * i + d];
        if (!(x >= 0.0f && x < (float)w->L[d])) return -1;
        int c = (int)floorf(x);
        if (c >= w->L[d]) c = w->L[d] - 1;   /* unreachable for x < L; kept for symmetry with the kernel */
        c3[d] = c;
    }
    return (c3[0] * w->L[1] + c3[1]) * w->L[2] + c3[2];
}

int orc_f32_cells(const orc_f32_world *w, int32_t *key, int32_t *order, int32_t *cell_start)
{
    const int n = w->n_beads;
    const int nc = w->L[0] * w->L[1] * w->L[2];
    int32_t *k = key ? key : (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *cs = (int32_t *)calloc((size_t)nc + 1, sizeof(int32_t));
    int rc = 0;
    for (int i = 0; i < n; i++) {
        int c3[3];
        int kk = cell_key_f32(w, i, c3);
        if (kk < 0) { rc = -1; kk = 0; }
        k[i] = kk;
        cs[kk + 1]++;
    }
    for (int c = 0; c < nc; c++) cs[c + 1] += cs[c];
    if (order) {
        /* stable counting sort by key: bead-index order inside each cell */
        int32_t *fill = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
        for (int i = 0; i < n; i++) order[cs[k[i]] + fill[k[i]]++] = i;
        free(fill);
    }
    if (cell_start) memcpy(cell_start, cs, sizeof(int32_t) * ((size_t)nc + 1));
    free(cs);
    if (!key) free(k);
    return rc;
}

/* Worker state of the pair-force pass: the world's beads gathered into cell
 * order (positions, shifted velocities, hashes), so that each row of the
 * stencil reads contiguous memory. Every bead's force is still the same
 * sequence of FP32 operations in the same order as the single-threaded
 * definition below; threads only split the beads between them. */
typedef struct f32_pair_job {
    const orc_f32_world *w;
    uint64_t t_hash;
    const int32_t *order, *cs;
    const float *sx, *sw;          /* cell order, 3 per bead */
    const uint32_t *sh;            /* cell order */
    float *f_out, *fabs_sum;
    float scale;
    int begin, end;                /* beads [begin, end) of the cell order */
} f32_pair_job;

/* Pair force of bead i: one running sum per x-plane of the stencil over the
 * plane's rows (dy = -1, 0, +1) in order, each row's cells in z order (z-1, z,
 * z+1) and each cell in slot order; planes combined (p0 + p1) + p2.
 * dpd_maths_core_half_step_raw.hpp:129-216 with the terms grouped as
 * a wr + sgw (scale u - sgw rdotv). */
static void f32_pair_force_bead(const f32_pair_job *J, int k0)
{
    const orc_f32_world *w = J->w;
    const int nt = w->n_types;
    float *f_out = J->f_out;
    const int i = J->order[k0];
    if (f32_stationary(w, i)) {
        /* a stationary bead ends the step with f = 0 (naive_dpd_engine_half_merge.hpp:387-390) */
        f_out[3 * i] = f_out[3 * i + 1] = f_out[3 * i + 2] = 0.0f;
        if (J->fabs_sum) J->fabs_sum[i] = 0.0f;
        return;
    }
    int hc[3];
    cell_key_f32(w, i, hc);
    const float hx = J->sx[3 * (size_t)k0], hy = J->sx[3 * (size_t)k0 + 1], hz = J->sx[3 * (size_t)k0 + 2];
    const float hvx = J->sw[3 * (size_t)k0], hvy = J->sw[3 * (size_t)k0 + 1], hvz = J->sw[3 * (size_t)k0 + 2];
    const uint32_t hh = J->sh[k0];
    const int htype = (int)(hh >> 28);
    float px[3] = {0, 0, 0}, py[3] = {0, 0, 0}, pz[3] = {0, 0, 0}, fa = 0.0f;
    for (int ddx = -1; ddx <= 1; ddx++) {
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int ddy = -1; ddy <= 1; ddy++) {
    for (int ddz = -1; ddz <= 1; ddz++) {
        int oc3[3];
        float sh[3];
        int dd[3] = {ddx, ddy, ddz};
        for (int d = 0; d < 3; d++) {
            int o = hc[d] + dd[d];
            sh[d] = 0.0f;
            if (o < 0) { o += w->L[d]; sh[d] = -(float)w->L[d]; }
            else if (o >= w->L[d]) { o -= w->L[d]; sh[d] = (float)w->L[d]; }
            oc3[d] = o;
        }
        int oc = (oc3[0] * w->L[1] + oc3[1]) * w->L[2] + oc3[2];
        for (int k = J->cs[oc]; k < J->cs[oc + 1]; k++) {
            if (k == k0) continue;
            const float *xo = J->sx + 3 * (size_t)k, *wo = J->sw + 3 * (size_t)k;
            float ox = xo[0] + sh[0];
            float oy = xo[1] + sh[1];
            float oz = xo[2] + sh[2];
            float dx = hx - ox, dy = hy - oy, dz = hz - oz;
            float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (!(r2 < 1.0f && r2 >= ORC_MIN_R2)) continue;
            float r = sqrtf(r2);
            float inv_r = 1.0f / r;
            float wr = 1.0f - r;
            float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
            float dvx = hvx - wo[0], dvy = hvy - wo[1], dvz = hvz - wo[2];
            float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
            uint32_t oh = J->sh[k];
            int ti = htype * nt + (int)(oh >> 28);
            float a, sg;
            if (w->cell_par) {
                const int hcl = (hc[0] * w->L[1] + hc[1]) * w->L[2] + hc[2];
                const float *ph = w->cell_par + 2 * ((size_t)hcl * nt * nt + ti);
                const float *po = w->cell_par + 2 * ((size_t)oc * nt * nt + ti);
                a = (ph[0] + po[0]) * 0.5f;
                sg = sqrtf((ph[1] + po[1]) * 0.5f);
            } else {
                a = w->a[ti];
                sg = w->sg[ti];
            }
            float sgw = sg * wr;
            float u = default_hash_inl(J->t_hash, hh, oh);
            float F = fmaf(a, wr, sgw * fmaf(-sgw, rdotv, J->scale * u));
            fx = fmaf(ex, F, fx);
            fy = fmaf(ey, F, fy);
            fz = fmaf(ez, F, fz);
            fa += fabsf(F);
        }
    }
    }
    px[ddx + 1] = fx;
    py[ddx + 1] = fy;
    pz[ddx + 1] = fz;
    }
    f_out[3 * i] = (px[0] + px[1]) + px[2];
    f_out[3 * i + 1] = (py[0] + py[1]) + py[2];
    f_out[3 * i + 2] = (pz[0] + pz[1]) + pz[2];
    if (J->fabs_sum) J->fabs_sum[i] = fa;
}

static void *f32_pair_worker(void *arg)
{
    const f32_pair_job *J = (const f32_pair_job *)arg;
    for (int k = J->begin; k < J->end; k++) f32_pair_force_bead(J, k);
    return NULL;
}

static int g_orc_threads = 1;

/* Threads used by the FP32 pair-force pass (results do not depend on it). */
void orc_set_threads(int n) { g_orc_threads = n < 1 ? 1 : (n > 256 ? 256 : n); }

void orc_f32_pair_forces(const orc_f32_world *w, uint64_t t_hash, float *f_out, float *fabs_sum)
{
    const int n = w->n_beads;
    const int nc = w->L[0] * w->L[1] * w->L[2];
    const size_t nn = (size_t)(n > 0 ? n : 1);
    int32_t *key = (int32_t *)malloc(sizeof(int32_t) * nn);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * nn);
    int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nc + 1));
    float *sx = (float *)malloc(sizeof(float) * 3 * nn);
    float *sw = (float *)malloc(sizeof(float) * 3 * nn);
    uint32_t *shs = (uint32_t *)malloc(sizeof(uint32_t) * nn);
    orc_f32_cells(w, key, order, cs);
    for (int k = 0; k < n; k++) {
        const int j = order[k];
        memcpy(sx + 3 * (size_t)k, w->x + 3 * (size_t)j, 3 * sizeof(float));
        memcpy(sw + 3 * (size_t)k, w->w + 3 * (size_t)j, 3 * sizeof(float));
        shs[k] = w->hash[j];
    }
    f32_pair_job base;
    base.w = w;
    base.t_hash = t_hash;
    base.order = order;
    base.cs = cs;
    base.sx = sx;
    base.sw = sw;
    base.sh = shs;
    base.f_out = f_out;
    base.fabs_sum = fabs_sum;
    base.scale = (float)sqrt(24.0 / (double)w->dt);
    int nth = g_orc_threads;
    if (n < 4096 * nth) nth = n / 4096 > 0 ? n / 4096 : 1;
    if (nth <= 1) {
        base.begin = 0;
        base.end = n;
        f32_pair_worker(&base);
    } else {
        f32_pair_job *jobs = (f32_pair_job *)malloc(sizeof(f32_pair_job) * (size_t)nth);
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nth);
        for (int k = 0; k < nth; k++) {
            jobs[k] = base;
            jobs[k].begin = (int)((int64_t)n * k / nth);
            jobs[k].end = (int)((int64_t)n * (k + 1) / nth);
        }
        for (int k = 1; k < nth; k++) pthread_create(&th[k], NULL, f32_pair_worker, &jobs[k]);
        f32_pair_worker(&jobs[0]);
        for (int k = 1; k < nth; k++) pthread_join(th[k], NULL);
        free(jobs);
        free(th);
    }
    free(key);
    free(order);
    free(cs);
    free(sx);
    free(sw);
    free(shs);
}

static double min_image_f64(double d, double len)
{
    if (d > 0.5 * len) d -= len;
    else if (d < -0.5 * len) d += len;
    return d;
}

/* Difference b - a of FP32 coordinates, exact in double, wrapped to the minimum image. */
static void diff_f64(const int *L, const float *a, const float *b, double *res)
{
    for (int d = 0; d < 3; d++) res[d] = min_image_f64((double)b[d] - (double)a[d], (double)L[d]);
}

static double norm3_f64(const double *v)
{
    return sqrt(fma(v[2], v[2], fma(v[1], v[1], v[0] * v[0])));
}

/* Bonded terms in FP64 from the FP32 positions. Each bead's terms are summed in
 * a double accumulator starting from 0, in the order the per-polymer loop
 * applies them (the polymer's bonds, then its bond pairs); a bead with terms
 * gets f = (float)((double)f_pair + sum), rounded once. Hookean: naive_dpd_engine_half_merge.hpp:638-654; angle:
 * dpd_maths_core_half_step.hpp:178-231 via naive_dpd_engine_half_merge.hpp:656-690,
 * with the prefactor floor of naive_dpd_engine.hpp:681-682 (argument clamped at 0). */
void orc_f32_bonds(const orc_f32_world *w, float *F)
{
    int n = w->n_beads;
    double *acc = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    unsigned char *touched = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    for (int i = 0; i < 3 * n; i++) acc[i] = 0.0;
    for (int p = 0; p < w->n_polymers; p++) {
        for (int bi = w->poly_bond_start[p]; bi < w->poly_bond_start[p + 1]; bi++) {
            int h = w->bond_head[bi], t = w->bond_tail[bi];
            double dx[3];
            diff_f64(w->L, &w->x[3 * h], &w->x[3 * t], dx);
            double r = norm3_f64(dx);
            double hf = w->bond_kappa[bi] * (w->bond_r0[bi] - r);
            double sc = hf / r;
            for (int d = 0; d < 3; d++) {
                acc[3 * h + d] += -(dx[d] * sc);
                acc[3 * t + d] += dx[d] * sc;
            }
            touched[h] = touched[t] = 1;
        }
        for (int ai = w->poly_angle_start[p]; ai < w->poly_angle_start[p + 1]; ai++) {
            int h = w->angle_head[ai], m = w->angle_mid[ai], t = w->angle_tail[ai];
            double d01[3], d12[3];
            diff_f64(w->L, &w->x[3 * h], &w->x[3 * m], d01);
            diff_f64(w->L, &w->x[3 * m], &w->x[3 * t], d12);
            double r01 = norm3_f64(d01), r12 = norm3_f64(d12);
            double mag = r01 * r12;
            double hf[3] = {0, 0, 0}, tf[3] = {0, 0, 0};
            if (mag > 0.0001) {
                double b1 = r01 * r01, b2 = r12 * r12;
                double dot = fma(d01[2], d12[2], fma(d01[1], d12[1], d01[0] * d12[0]));
                double o1 = dot / b1, o2 = dot / b2;
                double cos2 = o1 * o2;
                double fm = w->angle_kappa[ai] / mag;
                if (fabs(dot) > 0.000001 && w->angle_sin0[ai] > 0.0) {
                    double pref = fmax(sqrt(fmax(1.0 / cos2 - 1.0, 0.0)), 0.000001);
                    fm = fm * (w->angle_cos0[ai] - w->angle_sin0[ai] * (1.0 / pref));
                }
                for (int d = 0; d < 3; d++) {
                    hf[d] = (d01[d] * o1 - d12[d]) * fm;
                    tf[d] = (d01[d] - d12[d] * o2) * fm;
                }
            }
            for (int d = 0; d < 3; d++) {
                acc[3 * h + d] += hf[d];
                acc[3 * t + d] += tf[d];
                acc[3 * m + d] += -(hf[d] + tf[d]);
            }
            touched[h] = touched[m] = touched[t] = 1;
        }
    }
    for (int i = 0; i < n; i++)
        if (touched[i] && !f32_stationary(w, i))
            for (int d = 0; d < 3; d++) F[3 * i + d] = (float)((double)F[3 * i + d] + acc[3 * i + d]);
    free(acc);
    free(touched);
}

int orc_f32_step(orc_f32_world *w, int32_t *order, int32_t *cell_start, int32_t *key, float *fabs_sum)
{
    orc_f32_integrate(w);
    int rc = orc_f32_cells(w, key, order, cell_start);
    if (rc) return rc;
    uint64_t t_hash = orc_t_hash(w->t, w->seed);
    orc_f32_pair_forces(w, t_hash, w->f, fabs_sum);
    orc_f32_bonds(w, w->f);
    /* stationary beads end the step with v = 0 (naive_dpd_engine_half_merge.hpp:387-390) */
    for (int i = 0; i < w->n_beads; i++)
        if (f32_stationary(w, i)) w->w[3 * i] = w->w[3 * i + 1] = w->w[3 * i + 2] = 0.0f;
    w->t += 1;
    return 0;
}
