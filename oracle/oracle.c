/*
 * CPU restatement of pRRTC's planning hot path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h.  Built with -ffp-contract=off: the only fused multiply-adds
 * are the explicit fmaf() calls, which stand where the reference's CUDA build
 * contracts a*b+c (nvcc's default) in the planner arithmetic.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* collision primitives: utils.cuh:19-208                                    */

static inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;                   /* utils.cuh:24-33 */
}
static inline float sql2(float ax, float ay, float az, float bx, float by, float bz) {
    float xs = ax - bx, ys = ay - by, zs = az - bz;       /* utils.cuh:35-48 */
    return dot3(xs, ys, zs, xs, ys, zs);
}
static inline float clampf(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); } /* :50-53 */

/* Rounding: by default the tests are evaluated exactly as written in the
 * reference (utils.cuh:19-208, separate multiplies and adds).  In "fused"
 * mode (or_set_fused(1)) they use the explicit fmaf association of the HIP
 * kernels (prrtc_amd/csrc/device_common.h), so that verdicts -- and planner
 * runs of the point robot -- can be compared bit for bit. */
static int g_fused = 0;
void or_set_fused(int on) { g_fused = on; }

static inline float sq3f(float dx, float dy, float dz) { return fmaf(dx, dx, fmaf(dy, dy, dz * dz)); }

/* each returns the squared-form value (< 0 = collision) and the metric clearance */
static inline float t_sphere(const float *s, float x, float y, float z, float r, double *clr) {
    float d = g_fused ? sq3f(s[0] - x, s[1] - y, s[2] - z)
                      : sql2(s[0], s[1], s[2], x, y, z);  /* utils.cuh:58-81 */
    float rs = s[3] + r;
    if (clr) *clr = sqrt((double)d) - (double)rs;
    return g_fused ? fmaf(-rs, rs, d) : d - rs * rs;
}
static inline float t_capsule(const float *c, float x, float y, float z, float r, double *clr) {
    float dot, cdf, sum;
    if (g_fused) {
        dot = fmaf(x - c[0], c[3], fmaf(y - c[1], c[4], (z - c[2]) * c[5]));
        cdf = clampf(dot * c[7], 0.f, 1.f);
        sum = sq3f(x - fmaf(c[3], cdf, c[0]), y - fmaf(c[4], cdf, c[1]), z - fmaf(c[5], cdf, c[2]));
    } else {
        dot = dot3(x - c[0], y - c[1], z - c[2], c[3], c[4], c[5]);   /* utils.cuh:113-126 */
        cdf = clampf(dot * c[7], 0.f, 1.f);
        sum = sql2(x, y, z, c[0] + c[3] * cdf, c[1] + c[4] * cdf, c[2] + c[5] * cdf);
    }
    float rs = r + c[6];
    if (clr) *clr = sqrt((double)sum) - (double)rs;
    return g_fused ? fmaf(-rs, rs, sum) : sum - rs * rs;
}
static inline float t_zcapsule(const float *c, float x, float y, float z, float r, double *clr) {
    float dot = (z - c[2]) * c[5];                         /* utils.cuh:135-148 */
    float cdf = clampf(dot * c[7], 0.f, 1.f);
    float sum = g_fused ? sq3f(x - c[0], y - c[1], z - fmaf(c[5], cdf, c[2]))
                        : sql2(x, y, z, c[0], c[1], c[2] + c[5] * cdf);
    float rs = r + c[6];
    if (clr) *clr = sqrt((double)sum) - (double)rs;
    return g_fused ? fmaf(-rs, rs, sum) : sum - rs * rs;
}
static inline float t_cuboid(const float *c, float x, float y, float z, float r, double *clr) {
    float xs = x - c[0], ys = y - c[1], zs = z - c[2];     /* utils.cuh:158-175 */
    float p1, p2, p3;
    if (g_fused) {
        p1 = fmaf(c[3], xs, fmaf(c[4], ys, c[5] * zs));
        p2 = fmaf(c[6], xs, fmaf(c[7], ys, c[8] * zs));
        p3 = fmaf(c[9], xs, fmaf(c[10], ys, c[11] * zs));
    } else {
        p1 = dot3(c[3], c[4], c[5], xs, ys, zs);
        p2 = dot3(c[6], c[7], c[8], xs, ys, zs);
        p3 = dot3(c[9], c[10], c[11], xs, ys, zs);
    }
    float a1 = fmaxf(0.f, fabsf(p1) - c[12]);
    float a2 = fmaxf(0.f, fabsf(p2) - c[13]);
    float a3 = fmaxf(0.f, fabsf(p3) - c[14]);
    float sum = g_fused ? sq3f(a1, a2, a3) : dot3(a1, a2, a3, a1, a2, a3);
    if (clr) *clr = sqrt((double)sum) - (double)r;
    return g_fused ? fmaf(-r, r, sum) : sum - r * r;
}
static inline float t_zcuboid(const float *c, float x, float y, float z, float r, double *clr) {
    float xs = x - c[0], ys = y - c[1], zs = z - c[2];     /* utils.cuh:184-201 */
    float p1 = g_fused ? fmaf(c[3], xs, c[4] * ys) : c[3] * xs + c[4] * ys;
    float p2 = g_fused ? fmaf(c[6], xs, c[7] * ys) : c[6] * xs + c[7] * ys;
    float a1 = fmaxf(0.f, fabsf(p1) - c[12]);
    float a2 = fmaxf(0.f, fabsf(p2) - c[13]);
    float a3 = fmaxf(0.f, fabsf(zs) - c[14]);
    float sum = g_fused ? sq3f(a1, a2, a3) : dot3(a1, a2, a3, a1, a2, a3);
    if (clr) *clr = sqrt((double)sum) - (double)r;
    return g_fused ? fmaf(-r, r, sum) : sum - r * r;
}

static inline void upd_margin(double *margin, double clr) {
    double a = fabs(clr);
    if (a < *margin) *margin = a;
}

/* sphere_environment_in_collision (utils.cuh:211-242): OR over the five
 * tables in this order, each loop stopping at its first hit.  counts gets the
 * tests executed; with margin != NULL every obstacle is also measured. */
static int env_collide(const or_scene *s, float x, float y, float z, float r, int32_t *counts,
                       double *margin) {
    int hit = 0;
    for (int i = 0; i < s->n_spheres && !hit; i++) {
        if (counts) counts[0]++;
        hit |= t_sphere(s->spheres + 4 * i, x, y, z, r, NULL) < 0;
    }
    for (int i = 0; i < s->n_capsules && !hit; i++) {
        if (counts) counts[1]++;
        hit |= t_capsule(s->capsules + 8 * i, x, y, z, r, NULL) < 0;
    }
    for (int i = 0; i < s->n_z_capsules && !hit; i++) {
        if (counts) counts[2]++;
        hit |= t_zcapsule(s->z_capsules + 8 * i, x, y, z, r, NULL) < 0;
    }
    for (int i = 0; i < s->n_cuboids && !hit; i++) {
        if (counts) counts[3]++;
        hit |= t_cuboid(s->cuboids + 15 * i, x, y, z, r, NULL) < 0;
    }
    for (int i = 0; i < s->n_z_cuboids && !hit; i++) {
        if (counts) counts[4]++;
        hit |= t_zcuboid(s->z_cuboids + 15 * i, x, y, z, r, NULL) < 0;
    }
    if (margin) {
        double c;
        for (int i = 0; i < s->n_spheres; i++) { t_sphere(s->spheres + 4 * i, x, y, z, r, &c); upd_margin(margin, c); }
        for (int i = 0; i < s->n_capsules; i++) { t_capsule(s->capsules + 8 * i, x, y, z, r, &c); upd_margin(margin, c); }
        for (int i = 0; i < s->n_z_capsules; i++) { t_zcapsule(s->z_capsules + 8 * i, x, y, z, r, &c); upd_margin(margin, c); }
        for (int i = 0; i < s->n_cuboids; i++) { t_cuboid(s->cuboids + 15 * i, x, y, z, r, &c); upd_margin(margin, c); }
        for (int i = 0; i < s->n_z_cuboids; i++) { t_zcuboid(s->z_cuboids + 15 * i, x, y, z, r, &c); upd_margin(margin, c); }
    }
    return hit;
}

/* sphere_sphere_self_collision (utils.cuh:83-86) */
static inline int self_collide(const float *a, float ra, const float *b, float rb, double *margin) {
    float d = g_fused ? sq3f(a[0] - b[0], a[1] - b[1], a[2] - b[2]) : sql2(a[0], a[1], a[2], b[0], b[1], b[2]);
    float rs = ra + rb;
    if (margin) upd_margin(margin, sqrt((double)d) - (double)rs);
    return (g_fused ? fmaf(-rs, rs, d) : d - rs * rs) < 0;
}

/* ------------------------------------------------------------------------ */
/* forward kinematics from the model tables                                  */

static void fk_one(const or_model *m, const float *q, float *cen) {
    if (m->kind == 0) {                      /* point robot: the sphere is q itself */
        cen[0] = q[0]; cen[1] = q[1]; cen[2] = q[2];
        return;
    }
    float R[OR_MAX_JOINTS][9], t[OR_MAX_JOINTS][3];
    for (int j = 0; j < m->n_joints; j++) {
        static const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        static const float Z[3] = {0, 0, 0};
        const float *Rp = m->parent[j] < 0 ? I : R[m->parent[j]];
        const float *tp = m->parent[j] < 0 ? Z : t[m->parent[j]];
        const float *F = m->rot[j], *Ft = m->trans[j];
        float Rn[9], tn[3];
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++)
                Rn[3 * r + c] = Rp[3 * r] * F[c] + Rp[3 * r + 1] * F[3 + c] + Rp[3 * r + 2] * F[6 + c];
            tn[r] = tp[r] + (Rp[3 * r] * Ft[0] + Rp[3 * r + 1] * Ft[1] + Rp[3 * r + 2] * Ft[2]);
        }
        float qj = q[j];
        if (m->prismatic[j]) {
            for (int r = 0; r < 3; r++) tn[r] += Rn[3 * r + 2] * qj;
        } else {
            float c = cosf(qj), s = sinf(qj);
            for (int r = 0; r < 3; r++) {
                float x = Rn[3 * r], y = Rn[3 * r + 1];
                Rn[3 * r] = c * x + s * y;
                Rn[3 * r + 1] = c * y - s * x;
            }
        }
        memcpy(R[j], Rn, sizeof Rn);
        memcpy(t[j], tn, sizeof tn);
    }
    for (int i = 0; i < m->n_spheres; i++) {
        const float *l = m->local + 3 * i;
        int f = m->frame[i];
        if (f < 0) {
            cen[3 * i] = l[0]; cen[3 * i + 1] = l[1]; cen[3 * i + 2] = l[2];
        } else {
            for (int r = 0; r < 3; r++)
                cen[3 * i + r] = t[f][r] + (R[f][3 * r] * l[0] + R[f][3 * r + 1] * l[1] + R[f][3 * r + 2] * l[2]);
        }
    }
}

void or_fk(const or_model *m, const float *q, int n, float *centres) {
    int S = m->kind == 0 ? 1 : m->n_spheres;
    for (int i = 0; i < n; i++) fk_one(m, q + (size_t)i * m->dof, centres + (size_t)i * S * 3);
}

/* fkcc<Robot> (utils.cuh:326-33131): the generated checks in their order,
 * returning at the first collision.  Returns 1 when the pose is free. */
static int check_one(const or_model *m, const or_scene *s, const float *q, int32_t *counts, float *margin_out) {
    float cen[3 * 256];
    fk_one(m, q, cen);
    if (m->kind == 0) {                       /* utils.cuh:330-334 */
        double mg = INFINITY;
        int hit = env_collide(s, cen[0], cen[1], cen[2], m->point_radius, counts, margin_out ? &mg : NULL);
        if (margin_out) *margin_out = (float)mg;
        return !hit;
    }
    int coll = 0;
    for (int o = 0; o < m->n_order && !coll; o++) {
        int k = m->order_idx[o];
        if (m->order_kind[o] == 0) {
            int g = m->env_gate[k];
            if (env_collide(s, cen[3 * g], cen[3 * g + 1], cen[3 * g + 2], m->radius[g], counts, NULL)) {
                if (m->env_direct[k]) { coll = 1; break; }
                for (int li = m->env_leaf_off[k]; li < m->env_leaf_off[k + 1]; li++) {
                    int l = m->env_leaf[li];
                    if (env_collide(s, cen[3 * l], cen[3 * l + 1], cen[3 * l + 2], m->radius[l], counts, NULL)) {
                        coll = 1;
                        break;
                    }
                }
            }
        } else {
            int a = m->self_ga[k], b = m->self_gb[k];
            if (counts) counts[5]++;
            if (self_collide(cen + 3 * a, m->radius[a], cen + 3 * b, m->radius[b], NULL)) {
                if (m->self_direct[k]) { coll = 1; break; }
                for (int pi = m->self_pair_off[k]; pi < m->self_pair_off[k + 1]; pi++) {
                    int pa = m->self_pa[pi], pb = m->self_pb[pi];
                    if (counts) counts[5]++;
                    if (self_collide(cen + 3 * pa, m->radius[pa], cen + 3 * pb, m->radius[pb], NULL)) {
                        coll = 1;
                        break;
                    }
                }
            }
        }
    }
    if (margin_out) {
        /* every test the checker could run, hierarchy ignored */
        double mg = INFINITY;
        for (int k = 0; k < m->n_env; k++) {
            int g = m->env_gate[k];
            env_collide(s, cen[3 * g], cen[3 * g + 1], cen[3 * g + 2], m->radius[g], NULL, &mg);
            for (int li = m->env_leaf_off[k]; li < m->env_leaf_off[k + 1]; li++) {
                int l = m->env_leaf[li];
                env_collide(s, cen[3 * l], cen[3 * l + 1], cen[3 * l + 2], m->radius[l], NULL, &mg);
            }
        }
        for (int k = 0; k < m->n_self; k++) {
            int a = m->self_ga[k], b = m->self_gb[k];
            self_collide(cen + 3 * a, m->radius[a], cen + 3 * b, m->radius[b], &mg);
            for (int pi = m->self_pair_off[k]; pi < m->self_pair_off[k + 1]; pi++) {
                int pa = m->self_pa[pi], pb = m->self_pb[pi];
                self_collide(cen + 3 * pa, m->radius[pa], cen + 3 * pb, m->radius[pb], &mg);
            }
        }
        *margin_out = (float)mg;
    }
    return !coll;
}

void or_check_poses(const or_model *m, const or_scene *s, const float *q, int n, uint8_t *free_out,
                    float *margin, int32_t *counts) {
    for (int i = 0; i < n; i++) {
        int32_t *c = counts ? counts + 6 * (size_t)i : NULL;
        if (c) memset(c, 0, 6 * sizeof(int32_t));
        free_out[i] = (uint8_t)check_one(m, s, q + (size_t)i * m->dof, c, margin ? margin + i : NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* edges: pRRTC.cu:430-445                                                   */

static int edge_free(const or_model *m, const or_scene *s, const float *a, const float *b, int g) {
    int D = m->dof;
    float delta[OR_MAX_DOF], pose[OR_MAX_DOF];
    for (int i = 0; i < D; i++) delta[i] = (b[i] - a[i]) / (float)g;
    for (int k = 0; k < g; k++) {
        for (int i = 0; i < D; i++) pose[i] = fmaf((float)(k + 1), delta[i], a[i]);
        if (!check_one(m, s, pose, NULL, NULL)) return 0;
    }
    return 1;
}

void or_validate_edges(const or_model *m, const or_scene *s, const float *edges, int n, int g, uint8_t *free_out) {
    int D = m->dof;
    for (int e = 0; e < n; e++)
        free_out[e] = (uint8_t)edge_free(m, s, edges + (size_t)e * 2 * D, edges + (size_t)e * 2 * D + D, g);
}

typedef struct {
    const or_model *m; const or_scene *s; const float *edges; int lo, hi, g; uint8_t *out;
} mt_job;

static void *mt_worker(void *p) {
    mt_job *j = (mt_job *)p;
    int D = j->m->dof;
    for (int e = j->lo; e < j->hi; e++)
        j->out[e] = (uint8_t)edge_free(j->m, j->s, j->edges + (size_t)e * 2 * D, j->edges + (size_t)e * 2 * D + D, j->g);
    return NULL;
}

void or_validate_edges_mt(const or_model *m, const or_scene *s, const float *edges, int n, int g,
                          uint8_t *free_out, int threads) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    mt_job *jobs = (mt_job *)calloc((size_t)threads, sizeof(mt_job));
    for (int t = 0; t < threads; t++) {
        jobs[t] = (mt_job){m, s, edges, (int)((int64_t)n * t / threads), (int)((int64_t)n * (t + 1) / threads), g, free_out};
        pthread_create(&th[t], NULL, mt_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* ------------------------------------------------------------------------ */
/* Halton sampler: pRRTC.cu:42-121                                           */

static uint64_t splitmix(uint64_t *st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_halton_init(or_halton *h, int dof, uint64_t seed, int slot) {
    float primes[16] = {3.f, 5.f, 7.f, 11.f, 13.f, 17.f, 19.f, 23.f,
                        29.f, 31.f, 37.f, 41.f, 43.f, 47.f, 53.f, 59.f};
    if (slot != 0) {                         /* Fisher-Yates, pRRTC.cu:49-56, 65 */
        uint64_t st = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)slot * 0xD1B54A32D192ED03ULL;
        for (int i = 15; i > 0; i--) {
            int j = (int)(splitmix(&st) % (uint64_t)(i + 1));
            float tmp = primes[i]; primes[i] = primes[j]; primes[j] = tmp;
        }
    }
    for (int i = 0; i < dof; i++) { h->b[i] = primes[i]; h->n[i] = 0.f; h->d[i] = 1.f; }
}

void or_halton_next(or_halton *h, int dof, float *out) {
    for (int i = 0; i < dof; i++) {            /* pRRTC.cu:81-105 */
        float xf = h->d[i] - h->n[i];
        if (xf == 1.0f) {
            h->d[i] = floorf(h->d[i] * h->b[i]);
            h->n[i] = 1.0f;
        } else {
            float y = floorf(h->d[i] / h->b[i]);
            while (xf <= y) y = floorf(y / h->b[i]);
            h->n[i] = floorf((h->b[i] + 1.0f) * y) - xf;
        }
        out[i] = h->n[i] / h->d[i];
    }
}

/* ------------------------------------------------------------------------ */
/* planner: pRRTC.cu:305-602 (device loop), 607-766 (host result)           */

typedef struct {
    int D, cap;
    float *nodes[2];
    int *parent[2];
    float *radii[2];
    int size[2];
} trees_t;

static inline float sq_dist(const float *a, const float *b, int D) {
    float acc = 0.f;                       /* utils.cuh:100-109, contracted */
    for (int i = 0; i < D; i++) { float d = a[i] - b[i]; acc = fmaf(d, d, acc); }
    return acc;
}

/* argmin over [0, n); ties keep the smallest index */
static int nearest(const trees_t *T, int t, int n, const float *q, float *best_sq) {
    int best = 0;
    float bd = FLT_MAX;
    for (int i = 0; i < n; i++) {
        float d = sq_dist(T->nodes[t] + (size_t)i * T->D, q, T->D);
        if (d < bd) { bd = d; best = i; }
    }
    *best_sq = bd;
    return best;
}

typedef struct {
    int t, o;            /* target / other tree */
    int skip;
    int nn;
    int ext_ok;          /* steered edge free */
    float cfg[OR_MAX_DOF];
    int nn_o;
    unsigned n_ext;
    int n_ok;            /* extension sub-edges that were free */
    float *ext;          /* [n_ok][D] */
    int connected;
} slot_out;

/* Planted mistakes (tests only): or_set_plant(k) makes the planner commit
 * mistake k, so the known-answer tests can show that each one is caught.
 * 0 (the default) is the restatement itself. */
enum {
    PLANT_NONE = 0,
    PLANT_ITER1_TREE,     /* first iteration extends the other tree */
    PLANT_B1_RATIO,       /* balance 1: ratio test inverted */
    PLANT_B1_SPLIT,       /* balance 1: near-equal sizes split the other way */
    PLANT_B2_RULE,        /* balance 2: swap when the ratio is NOT below tree_ratio */
    PLANT_DD_SIGN,        /* dynamic domain: grow and shrink factors exchanged */
    PLANT_DD_SKIP,        /* dynamic domain: skip when the radius is above the distance */
    PLANT_CHAIN_DROP,     /* connect: a partial chain keeps no nodes */
    PLANT_STEER,          /* steering by min(1, d / range) instead of range / d */
    PLANT_PATH_ORDER,     /* path returned start -> goal */
    PLANT_N_EXT,          /* floor instead of ceil for the connect step count */
    PLANT_COUNT
};
static int g_plant = PLANT_NONE;
int or_set_plant(int k) {
    if (k < 0 || k >= PLANT_COUNT) return -1;
    g_plant = k;
    return PLANT_COUNT;
}

static int choose_tree(const or_settings *st, int slot, int iter, int *t_prev, int n0, int n1) {
    int nnc = st->num_new_configs;
    int t = *t_prev;                                         /* pRRTC.cu:348-371 */
    if (st->balance == 0 || iter == 1) {
        t = (slot < (nnc / 2)) ? 0 : 1;
        if (g_plant == PLANT_ITER1_TREE && iter == 1) t = 1 - t;
    } else if (st->balance == 1 && abs(n0 - n1) < 1.5 * nnc) {
        float ratio = n0 / (float)(n0 + n1);
        float bf = 1 - ratio;
        t = (slot < (nnc * bf)) ? 0 : 1;
        if (g_plant == PLANT_B1_SPLIT) t = 1 - t;
    } else if (st->balance == 1) {
        float ratio = n0 / (float)(n0 + n1);
        t = (ratio < st->tree_ratio) ? 0 : 1;
        if (g_plant == PLANT_B1_RATIO) t = 1 - t;
    } else if (st->balance == 2) {
        int nt = t == 0 ? n0 : n1, no = t == 0 ? n1 : n0;
        float ratio = abs(nt - no) / (float)nt;
        if ((ratio < st->tree_ratio) != (g_plant == PLANT_B2_RULE)) t = 1 - t;
    }
    *t_prev = t;
    return t;
}

static void slot_step(const or_model *m, const or_scene *s, const or_settings *st, const trees_t *T,
                      const int *sizes, or_halton *h, int slot, int iter, int *t_prev, slot_out *out,
                      int64_t *poses, int64_t *edges) {
    int D = m->dof, g = st->granularity;
    int t = choose_tree(st, slot, iter, t_prev, sizes[0], sizes[1]);
    int o = 1 - t;
    out->t = t; out->o = o; out->skip = 0; out->ext_ok = 0; out->n_ok = 0; out->connected = 0;
    float sample[OR_MAX_DOF];
    or_halton_next(h, D, sample);
    for (int i = 0; i < D; i++) sample[i] = fmaf(sample[i], m->span[i], m->lower[i]);  /* Robots.hh scale_cfg */

    float bsq;
    int nn = nearest(T, t, sizes[t], sample, &bsq);              /* pRRTC.cu:384-412 */
    float d = sqrtf(bsq);
    float scale = fminf(1.0f, g_plant == PLANT_STEER ? d / st->range : st->range / d);   /* :415-418 */
    out->nn = nn;
    if (st->dynamic_domain && ((T->radii[t][nn] < d) != (g_plant == PLANT_DD_SKIP))) {    /* :420-427 */
        out->skip = 1;
        return;
    }
    const float *nnv = T->nodes[t] + (size_t)nn * D;
    float delta[OR_MAX_DOF];
    for (int i = 0; i < D; i++) {                                 /* :430-434 */
        out->cfg[i] = fmaf(sample[i] - nnv[i], scale, nnv[i]);
        delta[i] = (out->cfg[i] - nnv[i]) / (float)g;
    }
    float pose[OR_MAX_DOF];
    int good = 1;
    (*edges)++;
    for (int k = 0; k < g && good; k++) {                        /* :436-445 */
        for (int i = 0; i < D; i++) pose[i] = fmaf((float)(k + 1), delta[i], nnv[i]);
        (*poses)++;
        good = check_one(m, s, pose, NULL, NULL);
    }
    if (!good) return;
    out->ext_ok = 1;

    int nn_o = nearest(T, o, sizes[o], out->cfg, &bsq);          /* :478-505 */
    float d_o = sqrtf(bsq);
    unsigned n_ext = (unsigned)(g_plant == PLANT_N_EXT ? floorf(d_o / st->range) : ceilf(d_o / st->range)); /* :511 */
    out->nn_o = nn_o;
    out->n_ext = n_ext;
    const float *nno = T->nodes[o] + (size_t)nn_o * D;
    float vec[OR_MAX_DOF], step[OR_MAX_DOF], cur[OR_MAX_DOF];
    for (int i = 0; i < D; i++) {
        vec[i] = (nno[i] - out->cfg[i]) / (float)n_ext;           /* :516-518 */
        step[i] = vec[i] / (float)g;
        cur[i] = out->cfg[i];
    }
    unsigned i_ext = 0;
    while (i_ext < n_ext) {                                       /* :521-551 */
        int ok = 1;
        (*edges)++;
        for (int k = 0; k < g && ok; k++) {
            for (int i = 0; i < D; i++) pose[i] = fmaf((float)(k + 1), step[i], cur[i]);
            (*poses)++;
            ok = check_one(m, s, pose, NULL, NULL);
        }
        if (!ok) break;
        for (int i = 0; i < D; i++) {
            cur[i] = cur[i] + vec[i];
            out->ext[(size_t)i_ext * D + i] = cur[i];
        }
        i_ext++;
    }
    out->n_ok = (int)i_ext;
    out->connected = (i_ext == n_ext);
    if (g_plant == PLANT_CHAIN_DROP && !out->connected) out->n_ok = 0;
}

static void dd_update(const or_settings *st, float *r, int grow) {
    if (!st->dynamic_domain) return;
    const float up = g_plant == PLANT_DD_SIGN ? 1.f - st->dd_alpha : 1 + st->dd_alpha;
    const float down = g_plant == PLANT_DD_SIGN ? 1 + st->dd_alpha : 1.f - st->dd_alpha;
    if (grow) {                                                   /* :456-468 */
        if (*r != FLT_MAX) *r = *r * up;
    } else {                                                      /* :584-598 */
        *r = (*r == FLT_MAX) ? st->dd_radius : fmaxf(*r * down, st->dd_min_radius);
    }
}

static int trace(const trees_t *T, int t, int cur, float *buf, int cap_rows, int *rows, float *cost) {
    int D = T->D, n = 0;                                          /* :555-576 */
    while (T->parent[t][cur] != cur) {
        int p = T->parent[t][cur];
        *cost += sqrtf(sq_dist(T->nodes[t] + (size_t)cur * D, T->nodes[t] + (size_t)p * D, D));
        if (n >= cap_rows) return -1;
        memcpy(buf + (size_t)n * D, T->nodes[t] + (size_t)cur * D, sizeof(float) * D);
        n++;
        cur = p;
    }
    *rows = n;
    return cur;
}

/* commit one slot's staged work; returns 1 solved, -1 failed (capacity / path), 0 go on */
static int commit(const or_model *m, const or_settings *st, trees_t *T, slot_out *so, int iter,
                  const float *start, const float *goals, float *path_out, int path_cap, or_result *res) {
    int D = m->dof;
    if (so->skip) return 0;
    int t = so->t, o = so->o;
    if (!so->ext_ok) { dd_update(st, &T->radii[t][so->nn], 0); return 0; }
    int idx = T->size[t];
    if (idx >= st->max_samples || idx >= T->cap) return -1;                /* :451-452 */
    memcpy(T->nodes[t] + (size_t)idx * D, so->cfg, sizeof(float) * D);
    T->parent[t][idx] = so->nn;
    T->radii[t][idx] = FLT_MAX;
    dd_update(st, &T->radii[t][so->nn], 1);
    T->size[t]++;
    int last = idx;
    for (int e = 0; e < so->n_ok; e++) {                                  /* :534-546 */
        idx = T->size[t];
        if (idx >= st->max_samples || idx >= T->cap) return -1;
        memcpy(T->nodes[t] + (size_t)idx * D, so->ext + (size_t)e * D, sizeof(float) * D);
        T->parent[t][idx] = last;
        T->radii[t][idx] = FLT_MAX;
        T->size[t]++;
        last = idx;
    }
    if (!so->connected) return 0;
    /* solved: trace both trees to their roots (:552-582), then assemble the
     * path goal -> start as the host does (:731-743) */
    int half = path_cap > 2 ? (path_cap - 2) : 0;
    float *tbuf = (float *)malloc(sizeof(float) * (size_t)(half + 1) * D);
    float *obuf = (float *)malloc(sizeof(float) * (size_t)(half + 1) * D);
    float cost = 0.f;
    int tn = 0, on = 0;
    int root_t = trace(T, t, last, tbuf, half, &tn, &cost);
    int root_o = root_t < 0 ? -1 : trace(T, o, so->nn_o, obuf, half, &on, &cost);
    if (root_t < 0 || root_o < 0 || tn + on + 2 > path_cap) { free(tbuf); free(obuf); return -2; }
    int reached = (t == 1) ? root_t : root_o;
    int n1 = (t == 1) ? tn : on, n0 = (t == 1) ? on : tn;
    const float *p1 = (t == 1) ? tbuf : obuf, *p0 = (t == 1) ? obuf : tbuf;
    int row = 0;
    memcpy(path_out + (size_t)row++ * D, goals + (size_t)reached * D, sizeof(float) * D);
    for (int i = n1 - 1; i >= 0; i--) memcpy(path_out + (size_t)row++ * D, p1 + (size_t)i * D, sizeof(float) * D);
    for (int i = 0; i < n0; i++) memcpy(path_out + (size_t)row++ * D, p0 + (size_t)i * D, sizeof(float) * D);
    memcpy(path_out + (size_t)row++ * D, start, sizeof(float) * D);
    if (g_plant == PLANT_PATH_ORDER)
        for (int i = 0; i < row / 2; i++)
            for (int k = 0; k < D; k++) {
                const float tmp = path_out[(size_t)i * D + k];
                path_out[(size_t)i * D + k] = path_out[(size_t)(row - 1 - i) * D + k];
                path_out[(size_t)(row - 1 - i) * D + k] = tmp;
            }
    res->path_nodes = row;
    res->path_length = tn + on;
    res->cost = cost;
    res->iters = iter;
    free(tbuf);
    free(obuf);
    return 1;
}

int or_plan(const or_model *m, const or_scene *s, const float *start, const float *goals, int n_goals,
            const or_settings *st, int mode, float *path_out, int path_cap, or_result *res) {
    int D = m->dof, nnc = st->num_new_configs, g = st->granularity;
    memset(res, 0, sizeof *res);
    if (D > OR_MAX_DOF || n_goals < 1 || nnc < 1 || g < 1 || st->range <= 0.f) return -1;
    trees_t T;
    T.D = D;
    T.cap = st->max_samples > n_goals ? st->max_samples : n_goals;
    for (int t = 0; t < 2; t++) {
        T.nodes[t] = (float *)calloc((size_t)T.cap * D, sizeof(float));
        T.parent[t] = (int *)calloc((size_t)T.cap, sizeof(int));
        T.radii[t] = (float *)calloc((size_t)T.cap, sizeof(float));
    }
    memcpy(T.nodes[0], start, sizeof(float) * D);                 /* :686-693 */
    T.parent[0][0] = 0;
    T.radii[0][0] = FLT_MAX;
    memcpy(T.nodes[1], goals, sizeof(float) * D * n_goals);
    for (int i = 0; i < n_goals; i++) { T.parent[1][i] = i; T.radii[1][i] = FLT_MAX; }
    T.size[0] = 1;
    T.size[1] = n_goals;

    or_halton *h = (or_halton *)calloc((size_t)nnc, sizeof(or_halton));
    int *tprev = (int *)calloc((size_t)nnc, sizeof(int));
    slot_out *so = (slot_out *)calloc((size_t)nnc, sizeof(slot_out));
    /* longest connect: the joint-space diagonal / range */
    double diag = 0;
    for (int i = 0; i < D; i++) diag += (double)m->span[i] * m->span[i];
    int max_ext = (int)ceil(sqrt(diag) / st->range) + 2;
    for (int i = 0; i < nnc; i++) {
        or_halton_init(&h[i], D, st->seed, i);
        so[i].ext = (float *)malloc(sizeof(float) * (size_t)max_ext * D);
    }
    int status = 0;
    for (int iter = 1; iter <= st->max_iters && status == 0; iter++) {
        if (mode == 0) {
            int snap[2] = {T.size[0], T.size[1]};
            for (int sl = 0; sl < nnc; sl++)
                slot_step(m, s, st, &T, snap, &h[sl], sl, iter, &tprev[sl], &so[sl], &res->pose_checks, &res->edge_checks);
            for (int sl = 0; sl < nnc && status == 0; sl++)
                status = commit(m, st, &T, &so[sl], iter, start, goals, path_out, path_cap, res);
        } else {
            for (int sl = 0; sl < nnc && status == 0; sl++) {
                int cur[2] = {T.size[0], T.size[1]};
                slot_step(m, s, st, &T, cur, &h[sl], sl, iter, &tprev[sl], &so[sl], &res->pose_checks, &res->edge_checks);
                status = commit(m, st, &T, &so[sl], iter, start, goals, path_out, path_cap, res);
            }
        }
    }
    res->solved = status == 1;
    if (!res->solved) { res->iters = 0; res->cost = 0.f; res->path_length = 0; res->path_nodes = 0; }
    res->start_tree_size = T.size[0];
    res->goal_tree_size = T.size[1];
    for (int i = 0; i < nnc; i++) free(so[i].ext);
    free(so); free(tprev); free(h);
    for (int t = 0; t < 2; t++) { free(T.nodes[t]); free(T.parent[t]); free(T.radii[t]); }
    return status == -2 ? -3 : 0;
}

/* ------------------------------------------------------------------------ */
/* textbook RRT-Connect (Kuffner & LaValle 2000), the CPU baseline           */

static int rc_extend(const or_model *m, const or_scene *s, trees_t *T, int t, const float *target, int g,
                     float range, int *new_idx, or_result *res) {
    int D = m->dof;
    float bsq;
    int nn = nearest(T, t, T->size[t], target, &bsq);
    float d = sqrtf(bsq);
    const float *nv = T->nodes[t] + (size_t)nn * D;
    float q[OR_MAX_DOF], delta[OR_MAX_DOF], pose[OR_MAX_DOF];
    int reached = d <= range;
    float scale = reached ? 1.0f : range / d;
    for (int i = 0; i < D; i++) {
        q[i] = reached ? target[i] : fmaf(target[i] - nv[i], scale, nv[i]);
        delta[i] = (q[i] - nv[i]) / (float)g;
    }
    res->edge_checks++;
    for (int k = 0; k < g; k++) {
        for (int i = 0; i < D; i++) pose[i] = fmaf((float)(k + 1), delta[i], nv[i]);
        res->pose_checks++;
        if (!check_one(m, s, pose, NULL, NULL)) return 0;          /* trapped */
    }
    int idx = T->size[t];
    if (idx >= T->cap) return -1;
    memcpy(T->nodes[t] + (size_t)idx * D, q, sizeof(float) * D);
    T->parent[t][idx] = nn;
    T->size[t]++;
    *new_idx = idx;
    return reached ? 2 : 1;                                          /* reached / advanced */
}

int or_rrt_connect(const or_model *m, const or_scene *s, const float *start, const float *goal,
                   int granularity, float range, int max_samples, uint64_t seed,
                   float *path_out, int path_cap, or_result *res) {
    int D = m->dof;
    memset(res, 0, sizeof *res);
    trees_t T;
    T.D = D;
    T.cap = max_samples + 2;
    for (int t = 0; t < 2; t++) {
        T.nodes[t] = (float *)calloc((size_t)T.cap * D, sizeof(float));
        T.parent[t] = (int *)calloc((size_t)T.cap, sizeof(int));
        T.radii[t] = NULL;
    }
    memcpy(T.nodes[0], start, sizeof(float) * D);
    memcpy(T.nodes[1], goal, sizeof(float) * D);
    T.size[0] = T.size[1] = 1;
    uint64_t st = seed;
    int a = 0, status = 0, ja = -1, jb = -1;
    for (int it = 0; it < max_samples && status == 0; it++) {
        float sample[OR_MAX_DOF];
        for (int i = 0; i < D; i++)
            sample[i] = m->lower[i] + m->span[i] * (float)((splitmix(&st) >> 40) * (1.0 / 16777216.0));
        int idx;
        int r = rc_extend(m, s, &T, a, sample, granularity, range, &idx, res);
        if (r < 0) { status = -1; break; }
        if (r > 0) {
            const float *qn = T.nodes[a] + (size_t)idx * D;
            int b = 1 - a, bidx = -1, rr;
            do {                                                     /* CONNECT */
                rr = rc_extend(m, s, &T, b, qn, granularity, range, &bidx, res);
            } while (rr == 1);
            if (rr < 0) { status = -1; break; }
            if (rr == 2) {
                status = 1;
                ja = a == 0 ? idx : bidx;
                jb = a == 0 ? bidx : idx;
                res->iters = it + 1;
            }
        }
        a = 1 - a;
    }
    if (status == 1) {
        /* path goal -> start, same layout as pRRTC's result */
        int row = 0, cur;
        float cost = 0.f;
        int nb = 0;
        for (cur = jb; T.parent[1][cur] != cur && cur != 0; cur = T.parent[1][cur]) nb++;
        int *stackb = (int *)malloc(sizeof(int) * (size_t)(nb + 2));
        int k = 0;
        for (cur = jb;; cur = T.parent[1][cur]) { stackb[k++] = cur; if (cur == 0) break; }
        /* stackb: junction .. goal(0); emit reversed: goal .. junction */
        for (int i = k - 1; i >= 0 && row < path_cap; i--)
            memcpy(path_out + (size_t)row++ * D, T.nodes[1] + (size_t)stackb[i] * D, sizeof(float) * D);
        for (cur = ja;; cur = T.parent[0][cur]) {
            if (row >= path_cap) { status = -2; break; }
            memcpy(path_out + (size_t)row++ * D, T.nodes[0] + (size_t)cur * D, sizeof(float) * D);
            if (cur == 0) break;
        }
        free(stackb);
        for (int i = 0; i + 1 < row; i++)
            cost += sqrtf(sq_dist(path_out + (size_t)i * D, path_out + (size_t)(i + 1) * D, D));
        res->cost = cost;
        res->path_nodes = row;
        res->path_length = row - 2;
    }
    res->solved = status == 1;
    res->start_tree_size = T.size[0];
    res->goal_tree_size = T.size[1];
    for (int t = 0; t < 2; t++) { free(T.nodes[t]); free(T.parent[t]); }
    return status < 0 ? status : 0;
}

/* ------------------------------------------------------------------------ */

int or_revalidate_path(const or_model *m, const or_scene *s, const float *path, int n_nodes, int g) {
    int D = m->dof;
    float delta[OR_MAX_DOF], pose[OR_MAX_DOF];
    for (int e = 0; e + 1 < n_nodes; e++) {
        const float *a = path + (size_t)e * D, *b = path + (size_t)(e + 1) * D;
        for (int i = 0; i < D; i++) delta[i] = (b[i] - a[i]) / (float)g;
        for (int k = 0; k <= g; k++) {
            for (int i = 0; i < D; i++) pose[i] = fmaf((float)k, delta[i], a[i]);
            if (!check_one(m, s, pose, NULL, NULL)) return e;
        }
    }
    return -1;
}
