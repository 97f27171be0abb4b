/* hbref.c -- ORACLE (test infrastructure only; see hbref.h).
 *
 * Restates, single-threaded and in plain C, what Brian2's cython runtime does
 * in one clock-driven step, for the neuron models of the benchmark
 * configurations.  Each function cites the reference code it follows.
 *
 * How each part is pinned (tests/test_oracle.py):
 *  - queue (hbr_queue_*): against the reference's own CSpikeQueue compiled from
 *    brian2/synapses/spikequeue.h into oracle/_ref (random push/peek/advance
 *    sequences, heterogeneous and scalar delays, subgroup clipping), and the
 *    known answers of brian2/tests/test_spikequeue.py:36-61.
 *  - step loop, threshold, reset, propagation, refractoriness, monitors: the
 *    known answers of brian2/tests/test_synapses.py:1026-1218 and
 *    test_refractory.py:45-75, restated in the tests, and bit-exact agreement
 *    with the spike lists / traces that Brian's cython runtime wrote into
 *    tests/golden (CUBA 4000, COBAHH 4000, CUBA with delays and weights).
 *  - the per-model update code: statement-by-statement against the optimised
 *    statements Brian emitted (tests/golden/spec_*.json), and against the
 *    closed-form solution of the linear CUBA system.
 *  - generator: hand-checked values and statistical properties.
 *  - PoissonGroup threshold (HBR_POISSON): the spike rate of every neuron within
 *    binomial bounds of rates * dt, and independence across neurons and steps
 *    (the reference draws with numpy, so only statistics can be compared).
 *  - queues kept across runs with a dt change in between: the known answer of
 *    brian2/tests/test_synapses.py:1310-1337, and the re-binning against the
 *    reference's CSpikeQueue (oracle/_ref).
 *  - the OpenMP variant (threads > 1, the CPU baseline's multi-core leg): final
 *    state, spikes and event counts identical to one thread's.
 */
#include "hbref.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#else
static int omp_get_thread_num(void) { return 0; }
#endif

#include "../brian2_amd/csrc/hb_math.h"

static char g_err[512];
const char* hbr_last_error(void) { return g_err; }
static int fail(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

double hbr_exp(double x) { return hb_exp(x); }
double hbr_expm1(double x) { return hb_expm1(x); }
double hbr_log(double x) { return hb_log(x); }
double hbr_pow(double x, double y) { return hb_pow(x, y); }

/* ------------------------------------------------------------------------- */
/* Input generator.  SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): the ctr-th */
/* output of a SplitMix64 sequence whose state is keyed by (seed, stream).     */
/* ------------------------------------------------------------------------- */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t hbr_rand64(uint64_t seed, uint64_t stream, uint64_t ctr) {
    uint64_t key = mix64(seed ^ mix64(stream + 0x9e3779b97f4a7c15ULL));
    return mix64(key + (ctr + 1) * 0x9e3779b97f4a7c15ULL);
}

double hbr_uniform(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return (double)(hbr_rand64(seed, stream, ctr) >> 11) * 0x1p-53;
}

void hbr_uniform_fill(uint64_t seed, uint64_t stream, uint64_t ctr0, int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) out[i] = hbr_uniform(seed, stream, ctr0 + (uint64_t)i);
}

uint32_t hbr_uniform_int(uint64_t h, uint32_t n) {
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

/* (i, j) with u(seed, stream, i*n_post + j) < p, for i in [i0, i1), sorted by (i, j) */
int64_t hbr_bernoulli(uint64_t seed, uint64_t stream, int32_t i0, int32_t i1, int32_t n_post,
                      double p, int32_t* pre, int32_t* post, int64_t cap) {
    int64_t n = 0;
    for (int32_t i = i0; i < i1; i++)
        for (int32_t j = 0; j < n_post; j++) {
            uint64_t ctr = (uint64_t)i * (uint64_t)n_post + (uint64_t)j;
            if (hbr_uniform(seed, stream, ctr) < p) {
                if (pre && n < cap) {
                    pre[n] = i;
                    post[n] = j;
                }
                n++;
            }
        }
    return n;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* For each target j: K distinct sources drawn uniformly from [0, n_pre), ascending.
 * Round r draws the missing count m with ctr = (j*64 + r)*K + t, t < m; the union
 * is de-duplicated after each round.  The procedure commutes with relabelling
 * of the sources, so every K-subset is equally likely. */
int hbr_fixed_indegree(uint64_t seed, uint64_t stream, int32_t n_pre, int32_t n_post, int32_t K,
                       int32_t* pre_of_post) {
    if (K > n_pre) return fail("fixed in-degree larger than the source population");
    int failed = 0;
    /* targets are independent draws: any number of threads gives the same output */
#pragma omp parallel reduction(| : failed)
    {
        int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)K * 2);
#pragma omp for schedule(dynamic, 256)
        for (int32_t j = 0; j < n_post; j++) {
            int32_t have = 0;
            int r = 0;
            while (have < K) {
                if (r >= 64) {
                    failed = 1;
                    break;
                }
                int32_t m = K - have;
                for (int32_t t = 0; t < m; t++) {
                    uint64_t ctr = ((uint64_t)j * 64 + (uint64_t)r) * (uint64_t)K + (uint64_t)t;
                    buf[have + t] = (int32_t)hbr_uniform_int(hbr_rand64(seed, stream, ctr), (uint32_t)n_pre);
                }
                qsort(buf, (size_t)K, sizeof(int32_t), cmp_i32);
                int32_t u = 0;
                for (int32_t t = 0; t < K; t++)
                    if (t == 0 || buf[t] != buf[u - 1]) buf[u++] = buf[t];
                have = u;
                r++;
            }
            memcpy(pre_of_post + (size_t)j * K, buf, sizeof(int32_t) * (size_t)K);
        }
        free(buf);
    }
    return failed ? fail("fixed in-degree sampling did not converge") : 0;
}

/* The synapses of a fixed in-degree table (n_post rows of K ascending sources) in
 * (source, target) order, the order Synapses.connect(i=, j=) receives them in: a
 * counting sort by source of the target-major table, stable, so each source's
 * targets stay ascending. */
void hbr_order_by_pre(const int32_t* pre_of_post, int32_t n_pre, int32_t n_post, int32_t K, int32_t* pre,
                      int32_t* post) {
    int64_t* start = (int64_t*)calloc((size_t)n_pre + 1, sizeof(int64_t));
    const int64_t S = (int64_t)n_post * K;
    for (int64_t e = 0; e < S; e++) start[pre_of_post[e] + 1]++;
    for (int32_t i = 0; i < n_pre; i++) start[i + 1] += start[i];
    for (int64_t e = 0; e < S; e++) {
        int32_t i = pre_of_post[e];
        int64_t q = start[i]++;
        pre[q] = i;
        post[q] = (int32_t)(e / K);
    }
    free(start);
}

/* lo + uniform_int(h(pair), hi - lo + 1) per synapse, keyed by (pre, post) */
void hbr_pair_int(uint64_t seed, uint64_t stream, int32_t n_post, const int32_t* pre,
                  const int32_t* post, int64_t n, int32_t lo, int32_t hi, int32_t* out) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; s++) {
        uint64_t ctr = (uint64_t)pre[s] * (uint64_t)n_post + (uint64_t)post[s];
        out[s] = lo + (int32_t)hbr_uniform_int(hbr_rand64(seed, stream, ctr), (uint32_t)(hi - lo + 1));
    }
}

/* (float)(lo + u (hi - lo)) per synapse, keyed by (pre, post) */
void hbr_pair_uniform_f32(uint64_t seed, uint64_t stream, int32_t n_post, const int32_t* pre,
                          const int32_t* post, int64_t n, double lo, double hi, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; s++) {
        uint64_t ctr = (uint64_t)pre[s] * (uint64_t)n_post + (uint64_t)post[s];
        out[s] = (float)(lo + hbr_uniform(seed, stream, ctr) * (hi - lo));
    }
}

/* ------------------------------------------------------------------------- */
/* CSpikeQueue restatement (brian2/synapses/spikequeue.h:14-206)              */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t* v;
    int64_t n, cap;
} ivec;

static void ivec_push(ivec* a, int32_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}

struct hbr_queue {
    ivec* ring;          /* queue[(offset + d) % size] holds delay d (spikequeue.h:17) */
    int64_t size;
    int64_t offset;
    double dt;
    int scalar_delay;
    int32_t* delays;     /* integer delays per synapse (spikequeue.h:20) */
    int64_t n_delays;
    int32_t source_start, source_end;
    int64_t* syn_start;  /* per-source synapse lists in increasing synapse index, as CSR */
    int32_t* syn_ids;
};

hbr_queue* hbr_queue_create(int32_t source_start, int32_t source_end) {
    hbr_queue* q = (hbr_queue*)calloc(1, sizeof(hbr_queue));
    q->size = 1; /* spikequeue.h:29 */
    q->ring = (ivec*)calloc(1, sizeof(ivec));
    q->source_start = source_start;
    q->source_end = source_end;
    return q;
}

void hbr_queue_destroy(hbr_queue* q) {
    if (!q) return;
    for (int64_t i = 0; i < q->size; i++) free(q->ring[i].v);
    free(q->ring);
    free(q->delays);
    free(q->syn_start);
    free(q->syn_ids);
    free(q);
}

/* expand (spikequeue.h:131-141): rotate offset to 0, then grow */
static void queue_expand(hbr_queue* q, int64_t newsize) {
    if (newsize <= q->size) return;
    ivec* r = (ivec*)calloc((size_t)newsize, sizeof(ivec));
    for (int64_t i = 0; i < q->size; i++) r[i] = q->ring[(q->offset + i) % q->size];
    free(q->ring);
    q->ring = r;
    q->size = newsize;
    q->offset = 0;
}

/* prepare (spikequeue.h:48-105) */
int hbr_queue_prepare(hbr_queue* q, const double* delays, int64_t n_delays, const int32_t* sources,
                      int64_t n_synapses, double dt) {
    if (!(n_delays == 1 || n_delays == n_synapses)) return fail("n_delays must be 1 or n_synapses");
    if (q->dt != 0.0 && q->dt != dt) {
        /* dt changed: re-bin queued spikes (spikequeue.h:58-74) */
        double factor = q->dt / dt;
        int64_t oldsize = q->size;
        int64_t newsize = (int64_t)(int)((double)oldsize * factor) + 1;
        ivec* r = (ivec*)calloc((size_t)newsize, sizeof(ivec));
        for (int64_t i = 0; i < oldsize; i++) {
            ivec* src = &q->ring[(i + q->offset) % oldsize];
            int64_t tgt = (int64_t)(int)((double)i * factor + 0.5);
            /* with a factor below 0.5 the last slots can map to newsize itself: the reference
               then assigns past the vector's size (its content is never peeked and is
               overwritten when the queue grows), i.e. those spikes are lost */
            if (tgt >= newsize) continue;
            free(r[tgt].v); /* assignment replaces the slot */
            r[tgt].n = r[tgt].cap = src->n;
            r[tgt].v = src->n ? (int32_t*)malloc(sizeof(int32_t) * (size_t)src->n) : NULL;
            if (src->n) memcpy(r[tgt].v, src->v, sizeof(int32_t) * (size_t)src->n);
        }
        for (int64_t i = 0; i < oldsize; i++) free(q->ring[i].v);
        free(q->ring);
        q->ring = r;
        q->size = newsize;
        q->offset = 0;
    }
    free(q->delays);
    q->delays = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_delays > 0 ? n_delays : 1));
    q->n_delays = n_delays;
    int min_delay = n_delays > 0 ? (int)(delays[0] / dt + 0.5) : 0;
    int max_delay = min_delay;
    for (int64_t i = 0; i < n_delays; i++) {
        q->delays[i] = (int)(delays[i] / dt + 0.5); /* spikequeue.h:90 */
        if (q->delays[i] > max_delay)
            max_delay = q->delays[i];
        else if (q->delays[i] < min_delay)
            min_delay = q->delays[i];
    }
    /* per-source synapse lists, increasing synapse index (spikequeue.h:96-97) */
    int32_t nsrc = q->source_end - q->source_start;
    free(q->syn_start);
    free(q->syn_ids);
    q->syn_start = (int64_t*)calloc((size_t)nsrc + 1, sizeof(int64_t));
    q->syn_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_synapses > 0 ? n_synapses : 1));
    for (int64_t i = 0; i < n_synapses; i++) {
        int32_t s = sources[i] - q->source_start;
        if (s < 0 || s >= nsrc) return fail("synapse source outside the queue's source range");
        q->syn_start[s + 1]++;
    }
    for (int32_t s = 0; s < nsrc; s++) q->syn_start[s + 1] += q->syn_start[s];
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nsrc > 0 ? nsrc : 1));
    memcpy(cur, q->syn_start, sizeof(int64_t) * (size_t)nsrc);
    for (int64_t i = 0; i < n_synapses; i++) q->syn_ids[cur[sources[i] - q->source_start]++] = (int32_t)i;
    free(cur);
    q->dt = dt;
    if (max_delay >= q->size) queue_expand(q, max_delay + 1); /* ensure_delay 143-149 */
    q->scalar_delay = (min_delay == max_delay);
    return 0;
}

static int64_t lower_bound_i32(const int32_t* a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int64_t upper_bound_i32(const int32_t* a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* push (spikequeue.h:151-191) */
void hbr_queue_push(hbr_queue* q, const int32_t* spikes, int32_t n) {
    if (n == 0) return;
    int64_t start = lower_bound_i32(spikes, n, q->source_start);
    int64_t stop = upper_bound_i32(spikes, n, q->source_end - 1);
    for (int64_t k = start; k < stop; k++) {
        int32_t s = spikes[k] - q->source_start;
        for (int64_t p = q->syn_start[s]; p < q->syn_start[s + 1]; p++) {
            int32_t syn = q->syn_ids[p];
            int32_t d = q->scalar_delay ? q->delays[0] : q->delays[syn];
            ivec_push(&q->ring[(q->offset + d) % q->size], syn);
        }
    }
}

/* peek (spikequeue.h:193-196) */
int64_t hbr_queue_peek(hbr_queue* q, int32_t* out, int64_t cap) {
    ivec* v = &q->ring[q->offset];
    if (out) memcpy(out, v->v, sizeof(int32_t) * (size_t)(v->n < cap ? v->n : cap));
    return v->n;
}

/* advance (spikequeue.h:198-205) */
void hbr_queue_advance(hbr_queue* q) {
    q->ring[q->offset].n = 0;
    q->offset = (q->offset + 1) % q->size;
}

int64_t hbr_queue_size(hbr_queue* q) { return q->size; }
int32_t hbr_queue_offset(hbr_queue* q) { return (int32_t)q->offset; }

/* ------------------------------------------------------------------------- */
/* Models: the optimised statements Brian emits (tests/golden/spec_*.json),   */
/* restated by hand.  `timestep` is brian2/core/functions.py:771-802 (C form  */
/* cpp_generator.py:651-659).                                                 */
/* ------------------------------------------------------------------------- */
static int64_t timestep(double t, double dt) { return (int64_t)((t + 1e-3 * dt) / dt); }

/* CUBA, method='exact', v (unless refractory); params: taum taue taui El Vt Vr refractory.
 * State: v, ge, gi.  Statements: spec_cuba.json, stateupdate/threshold/reset. */
typedef struct {
    int64_t lio1;
    double lio2, lio3, lio4, lio5, lio6, lio7, lio8, lio9, lio10, lio11, lio12, lio13, lio14, lio15;
    double Vt, Vr;
} cuba_consts;

static void cuba_scalar(const double* P, double dt, cuba_consts* c) {
    double taum = P[0], taue = P[1], taui = P[2], El = P[3];
    c->Vt = P[4];
    c->Vr = P[5];
    c->lio1 = timestep(P[6], dt);
    c->lio2 = hb_exp((-dt) / taue);
    c->lio3 = hb_exp((-dt) / taui);
    c->lio4 = taue * hb_exp((-dt) / taue);
    c->lio5 = -hb_exp(dt / taue);
    c->lio6 = dt / taum;
    c->lio7 = (-dt) / taum;
    c->lio8 = 0.0 - taum;
    c->lio9 = taui * hb_exp((-dt) / taui);
    c->lio10 = -hb_exp(dt / taui);
    c->lio11 = El - El;
    c->lio12 = El - (El * hb_exp(c->lio7));
    c->lio13 = ((c->lio4 * (c->lio5 + hb_exp(c->lio6))) * hb_exp(c->lio7)) / (c->lio8 + taue);
    c->lio14 = ((c->lio9 * (c->lio10 + hb_exp(c->lio6))) * hb_exp(c->lio7)) / (c->lio8 + taui);
    c->lio15 = hb_exp(c->lio7);
}

static void cuba_update(const cuba_consts* c, hbr_group* g, double t, double dt, int threads) {
    int32_t N = g->N;
    double *v = g->state, *ge = g->state + N, *gi = g->state + 2 * (size_t)N;
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
    for (int32_t i = 0; i < N; i++) {
        int nr = timestep(t - g->lastspike[i], dt) >= c->lio1;
        g->not_refractory[i] = (uint8_t)nr;
        double _ge = c->lio2 * ge[i];
        double _gi = c->lio3 * gi[i];
        double _v;
        if (!nr)
            _v = c->lio11 + v[i];
        else
            _v = c->lio12 + (((c->lio13 * ge[i]) + (c->lio14 * gi[i])) + (c->lio15 * v[i]));
        ge[i] = _ge;
        gi[i] = _gi;
        if (nr) v[i] = _v;
    }
}

/* LIF with an exponential current, method='exact', no refractoriness
 * (dv/dt = (I + bias - v)/tau_m, dI/dt = -I/tau_s, threshold v > theta, reset v = 0).
 * params: tau_m tau_s theta bias.  State: v, I.  Statements: spec_lif.json. */
typedef struct {
    double lio1, lio2, lio3, lio4, theta;
} lif_consts;

static void lif_scalar(const double* P, double dt, lif_consts* c) {
    double tau_m = P[0], tau_s = P[1], bias = P[3];
    c->theta = P[2];
    c->lio1 = hb_exp((-dt) / tau_s);
    c->lio2 = bias - (bias * hb_exp((-dt) / tau_m));
    c->lio3 = (((tau_s * ((-hb_exp(dt / tau_m)) + hb_exp(dt / tau_s))) * hb_exp((-dt) / tau_m)) * hb_exp((-dt) / tau_s)) /
              (tau_m - tau_s);
    c->lio4 = hb_exp((-dt) / tau_m);
}

static void lif_update(const lif_consts* c, hbr_group* g, int threads) {
    int32_t N = g->N;
    double *v = g->state, *I = g->state + N;
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
    for (int32_t i = 0; i < N; i++) {
        double _I = c->lio1 * I[i];
        double _v = c->lio2 + ((c->lio3 * I[i]) + (c->lio4 * v[i]));
        I[i] = _I;
        v[i] = _v;
    }
}

/* COBAHH (examples/COBAHH.py), method='exponential_euler'.
 * params: Cm gl El EK ENa g_na g_kd VT taue taui Ee Ei mV ms refractory
 * State: v, m, n, h, ge, gi.  Statements: spec_cobahh.json. */
typedef struct {
    int64_t lio1;
    double l[28];
    double ms, thr;
} hh_consts;

static void hh_scalar(const double* P, double dt, hh_consts* c) {
    double Cm = P[0], gl = P[1], El = P[2], EK = P[3], ENa = P[4], g_na = P[5], g_kd = P[6];
    double VT = P[7], taue = P[8], taui = P[9], Ee = P[10], Ei = P[11], mV = P[12], ms = P[13];
    double* l = c->l;
    c->ms = ms;
    c->lio1 = timestep(P[14], dt);
    l[2] = hb_exp(((-1.0) * dt) / taue);
    l[3] = hb_exp(((-1.0) * dt) / taui);
    l[4] = (0.128 * (2.5713844347880297 * hb_exp((0.05555555555555555 * VT) / mV))) / ms;
    l[5] = 0.05555555555555555 / mV;
    l[6] = (ms * 2980.9579870417283) * hb_exp((0.2 * VT) / mV);
    l[7] = 0.2 / mV;
    l[8] = (0.128 * (2.5713844347880297 * hb_pow(hb_exp(VT / mV), 0.05555555555555555))) / ms;
    l[9] = 1.0 / mV;
    l[10] = 1.28 / ms;
    l[11] = (-1.28) / ms;
    l[12] = ((0.25 * VT) / mV) + 3.25;
    l[13] = 0.25 / mV;
    l[14] = 1.4 / ms;
    l[15] = (-8.0) + ((0.2 * (-VT)) / mV);
    l[16] = 0.16 / ms;
    l[17] = ((-0.6420127083438707) * hb_pow(hb_exp(VT / mV), 0.025)) / ms;
    l[18] = 3.0 + ((0.2 * VT) / mV);
    l[19] = (El * gl) / Cm;
    l[20] = (EK * g_kd) / Cm;
    l[21] = (ENa * g_na) / Cm;
    l[22] = Ee / Cm;
    l[23] = Ei / Cm;
    l[24] = 0.0 - (gl / Cm);
    l[25] = (-g_kd) / Cm;
    l[26] = g_na / Cm;
    l[27] = 1.0 / Cm;
    c->thr = (-20.0) * mV; /* threshold block: _lio_1 := (-20.0) * mV */
}

static void hh_update(const hh_consts* c, hbr_group* g, double t, double dt, int threads) {
    int32_t N = g->N;
    const double* l = c->l;
    double ms = c->ms;
    double *V = g->state, *M = V + N, *Nn = V + 2 * (size_t)N, *H = V + 3 * (size_t)N;
    double *GE = V + 4 * (size_t)N, *GI = V + 5 * (size_t)N;
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
    for (int32_t i = 0; i < N; i++) {
        double v = V[i], m = M[i], n = Nn[i], h = H[i], ge = GE[i], gi = GI[i];
        g->not_refractory[i] = (uint8_t)(timestep(t - g->lastspike[i], dt) >= c->lio1);
        double _ge = l[2] * ge;
        double _gi = l[3] * gi;
        double _BA_h = (l[4] * hb_exp(l[5] * (-v))) /
                       (((-4.0) / (ms + (l[6] * hb_exp(l[7] * (-v))))) -
                        (l[8] / hb_pow(hb_exp(l[9] * v), 0.05555555555555555)));
        double _h = (-_BA_h) + ((_BA_h + h) *
                                hb_exp(dt * (((-4.0) / (ms + (l[6] * hb_exp(l[7] * (-v))))) -
                                             (l[8] / hb_pow(hb_exp(l[9] * v), 0.05555555555555555)))));
        double _BA_m = l[10] / (((l[11] / hb_exprel(l[12] - (l[13] * v))) -
                                 (l[14] / hb_exprel(l[15] + (l[7] * v)))) *
                                hb_exprel(l[12] - (l[13] * v)));
        double _m = (-_BA_m) + ((_BA_m + m) * hb_exp(dt * ((l[11] / hb_exprel(l[12] - (l[13] * v))) -
                                                           (l[14] / hb_exprel(l[15] + (l[7] * v))))));
        double _BA_n = l[16] / (((l[17] / hb_pow(hb_exp(l[9] * v), 0.025)) -
                                 (l[16] / hb_exprel(l[18] - (l[7] * v)))) *
                                hb_exprel(l[18] - (l[7] * v)));
        double _n = (-_BA_n) + ((_BA_n + n) * hb_exp(dt * ((l[17] / hb_pow(hb_exp(l[9] * v), 0.025)) -
                                                           (l[16] / hb_exprel(l[18] - (l[7] * v))))));
        double n4 = hb_pow(n, 4.0), m3 = hb_pow(m, 3.0);
        double _BA_v = (l[19] + ((((l[20] * n4) + (l[21] * (h * m3))) + (l[22] * ge)) + (l[23] * gi))) /
                       ((l[24] + (l[25] * n4)) - (((l[26] * (h * m3)) + (l[27] * ge)) + (l[27] * gi)));
        double _v = (-_BA_v) + ((_BA_v + v) * hb_exp(dt * ((l[24] + (l[25] * n4)) -
                                                           (((l[26] * (h * m3)) + (l[27] * ge)) + (l[27] * gi)))));
        GE[i] = _ge;
        GI[i] = _gi;
        H[i] = _h;
        M[i] = _m;
        Nn[i] = _n;
        V[i] = _v;
    }
}

/* ------------------------------------------------------------------------- */
/* The step loop (brian2/core/network.py:1194-1250 with the default schedule  */
/* network.py:42-50): groups -> thresholds (+ monitor, order+1) -> synapses   */
/* (push, then propagate, per pathway) -> resets -> clock.advance.             */
/* ------------------------------------------------------------------------- */
/* The left fold of one slot split by target over T threads (the OpenMP leg of the CPU
 * baseline, not a reference restatement): thread t owns the targets [t*N/T, (t+1)*N/T).
 * The slot is cut into T chunks; each thread counts its chunk's events per owner, a
 * scan gives every (chunk, owner) pair its place, each thread copies its chunk's events
 * to the owners' lists in slot order, and each owner then folds its list in order.  A
 * target's events are thus applied in queue order, as in the sequential loop. */
typedef struct {
    int32_t *syn, *post;
    int64_t cap;
    int64_t* cnt;   /* [T][T] */
    int32_t T;
} fold_buf;

static void fold_parallel(const hbr_pathway* P, double* tv, int32_t N, const int32_t* slot, int64_t n, int T,
                          fold_buf* fb) {
    if (fb->cap < n) {
        free(fb->syn);
        free(fb->post);
        fb->cap = n * 2;
        fb->syn = (int32_t*)malloc(sizeof(int32_t) * (size_t)fb->cap);
        fb->post = (int32_t*)malloc(sizeof(int32_t) * (size_t)fb->cap);
    }
    if (fb->T != T) {
        free(fb->cnt);
        fb->cnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(T * T + 1));
        fb->T = T;
    }
    int64_t* cnt = fb->cnt;
#pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num();
        int64_t a = n * t / T, b = n * (t + 1) / T;
        int64_t* c = cnt + (size_t)t * T;
        for (int o = 0; o < T; o++) c[o] = 0;
        for (int64_t e = a; e < b; e++) c[(int64_t)P->post[slot[e]] * T / N]++;
#pragma omp barrier
#pragma omp single
        {
            /* owner-major positions: owner o's list holds chunk 0's events, then chunk 1's ... */
            int64_t acc = 0;
            for (int o = 0; o < T; o++)
                for (int k = 0; k < T; k++) {
                    int64_t x = cnt[(size_t)k * T + o];
                    cnt[(size_t)k * T + o] = acc;
                    acc += x;
                }
        }
        for (int64_t e = a; e < b; e++) {
            int32_t syn = slot[e], post = P->post[syn];
            int64_t q = c[(int64_t)post * T / N]++;
            fb->syn[q] = syn;
            fb->post[q] = post;
        }
#pragma omp barrier
        /* after the copy cnt[k][o] is the end of chunk k's part of owner o's list, so owner
           t's list ends at cnt[T-1][t] and starts where owner t-1's ends */
        const int64_t lo = t > 0 ? cnt[(size_t)(T - 1) * T + t - 1] : 0;
        const int64_t hi = cnt[(size_t)(T - 1) * T + t];
        for (int64_t q = lo; q < hi; q++) {
            int32_t syn = fb->syn[q], post = fb->post[q];
            double x = tv[post];
            x += P->w ? P->w[syn] : P->w_scalar;
            tv[post] = x;
            if (P->counts) P->counts[syn]++;
        }
    }
}

typedef struct {
    cuba_consts cuba;
    lif_consts lif;
    hh_consts hh;
    int32_t* spikes;   /* _spikespace[0:n] */
    int32_t nspk;
} group_rt;

int hbr_run(hbr_sim* S) {
    g_err[0] = 0;
    double dt = S->dt;
    group_rt* rt = (group_rt*)calloc((size_t)S->n_groups, sizeof(group_rt));
    for (int32_t gi = 0; gi < S->n_groups; gi++) {
        hbr_group* g = &S->groups[gi];
        rt[gi].spikes = (int32_t*)malloc(sizeof(int32_t) * (size_t)(g->N > 0 ? g->N : 1));
        if (g->model == HBR_CUBA_EXACT) cuba_scalar(g->params, dt, &rt[gi].cuba);
        else if (g->model == HBR_COBAHH_EE) hh_scalar(g->params, dt, &rt[gi].hh);
        else if (g->model == HBR_LIF_EXACT) lif_scalar(g->params, dt, &rt[gi].lif);
        else if (g->model != HBR_SPIKEGEN && g->model != HBR_PASSIVE && g->model != HBR_POISSON) {
            free(rt);
            return fail("unknown model");
        }
    }
    /* SynapticPathway.initialise_queue -> CSpikeQueue::prepare (synapses.py:421-458) */
    /* the queues: the caller's (kept across runs, prepared again with this run's dt,
       synapses.py:421-437) or this call's own */
#ifdef _OPENMP
    const int threads = S->threads > 1 ? S->threads : 1;
#else
    const int threads = 1;
#endif
    hbr_queue** qs = S->queues ? S->queues
                               : (hbr_queue**)calloc((size_t)(S->n_paths > 0 ? S->n_paths : 1), sizeof(hbr_queue*));
    for (int32_t p = 0; p < S->n_paths; p++) {
        hbr_pathway* P = &S->paths[p];
        if (!qs[p]) qs[p] = hbr_queue_create(P->source_start, P->source_end);
        int rc;
        if (P->delay)
            rc = hbr_queue_prepare(qs[p], P->delay, P->n_syn, P->pre, P->n_syn, dt);
        else
            rc = hbr_queue_prepare(qs[p], &P->delay_scalar, 1, P->pre, P->n_syn, dt);
        if (rc) return rc;
    }
    fold_buf fb = {0};
    int32_t* slot = NULL;
    int64_t slot_cap = 0;
    S->n_spk = 0;
    int rc = 0;
    for (int64_t k = 0; k < S->n_steps; k++) {
        int64_t step = S->start_step + k;
        double t = (double)step * dt; /* Clock.advance, clocks.py:315-325 */
        /* 'start': StateMonitor (statemonitor.py:171) */
        if (S->rec_out) {
            hbr_group* g = &S->groups[S->rec_group];
            for (int32_t a = 0; a < S->n_rec_vars; a++)
                for (int32_t r = 0; r < S->n_rec; r++)
                    S->rec_out[((size_t)k * S->n_rec_vars + a) * S->n_rec + r] =
                        g->state[(size_t)S->rec_vars[a] * g->N + S->rec_idx[r]];
        }
        /* 'groups': state updaters */
        for (int32_t gi = 0; gi < S->n_groups; gi++) {
            hbr_group* g = &S->groups[gi];
            if (g->model == HBR_CUBA_EXACT) cuba_update(&rt[gi].cuba, g, t, dt, threads);
            else if (g->model == HBR_COBAHH_EE) hh_update(&rt[gi].hh, g, t, dt, threads);
            else if (g->model == HBR_LIF_EXACT) lif_update(&rt[gi].lif, g, threads);
        }
        /* 'thresholds': threshold.pyx:18-32 / spikegenerator.pyx:6-26 */
        for (int32_t gi = 0; gi < S->n_groups; gi++) {
            hbr_group* g = &S->groups[gi];
            group_rt* R = &rt[gi];
            R->nspk = 0;
            if (g->model == HBR_PASSIVE) continue; /* "v : 1": no threshold */
            if (g->model == HBR_POISSON) {
                /* poissongroup.py:106: "rand() < rates * dt", the uniform drawn from the
                   counter-based generator keyed by (seed, step, neuron) */
                for (int32_t i = 0; i < g->N; i++)
                    if (hbr_uniform(g->poisson_seed, g->poisson_stream, (uint64_t)step * (uint64_t)g->N + (uint64_t)i) <
                        g->poisson_rdt[i])
                        R->spikes[R->nspk++] = i;
                continue;
            }
            if (g->model == HBR_SPIKEGEN) {
                int32_t timebin = (int32_t)step; /* t_in_timesteps */
                int64_t idx;
                for (idx = g->gen_lastindex; idx < g->n_gen; idx++) {
                    if (g->gen_timebin[idx] > timebin) break;
                    R->spikes[R->nspk++] = g->gen_index[idx];
                }
                g->gen_lastindex += R->nspk;
                continue;
            }
            double thr = g->model == HBR_CUBA_EXACT ? R->cuba.Vt : g->model == HBR_LIF_EXACT ? R->lif.theta : R->hh.thr;
            double* v = g->state;
            for (int32_t i = 0; i < g->N; i++) {
                /* without refractoriness the condition is v > thr alone (neurongroup.py:392-395) */
                int cond = (v[i] > thr) && (!g->not_refractory || g->not_refractory[i]);
                if (cond) {
                    R->spikes[R->nspk++] = i;
                    if (g->not_refractory) {
                        g->not_refractory[i] = 0;
                        g->lastspike[i] = t;
                    }
                }
            }
        }
        /* SpikeMonitor (order+1 in 'thresholds'; spikemonitor.pyx:9-58) */
        if (S->mon_group >= 0) {
            group_rt* R = &rt[S->mon_group];
            for (int32_t s = 0; s < R->nspk; s++) {
                if (S->n_spk < S->spk_cap) {
                    S->spk_i[S->n_spk] = R->spikes[s];
                    S->spk_step[S->n_spk] = step;
                }
                S->n_spk++;
            }
        }
        /* 'synapses': per pathway push_spikes then propagate (synapses.py:416-419) */
        for (int32_t p = 0; p < S->n_paths; p++) {
            hbr_pathway* P = &S->paths[p];
            group_rt* R = &rt[P->source_group];
            hbr_queue_push(qs[p], R->spikes, R->nspk);
            int64_t n = hbr_queue_peek(qs[p], NULL, 0);
            if (n > slot_cap) {
                slot_cap = n * 2;
                slot = (int32_t*)realloc(slot, sizeof(int32_t) * (size_t)slot_cap);
            }
            hbr_queue_peek(qs[p], slot, slot_cap);
            hbr_group* tg = &S->groups[P->target_group];
            double* tv = tg->state + (size_t)P->target_var * tg->N;
            if (threads > 1 && n >= 4096) {
                fold_parallel(P, tv, tg->N, slot, n, threads, &fb);
            } else {
                /* synapses.pyx:42-50: sequential left fold in queue order */
                for (int64_t e = 0; e < n; e++) {
                    int32_t syn = slot[e];
                    int32_t post = P->post[syn];
                    double x = tv[post];
                    x += P->w ? P->w[syn] : P->w_scalar;
                    tv[post] = x;
                    if (P->counts) P->counts[syn]++;
                }
            }
            hbr_queue_advance(qs[p]);
        }
        /* 'resets': reset.pyx:10-15 */
        for (int32_t gi = 0; gi < S->n_groups; gi++) {
            hbr_group* g = &S->groups[gi];
            if (g->model == HBR_CUBA_EXACT) {
                double* v = g->state;
                for (int32_t s = 0; s < rt[gi].nspk; s++) v[rt[gi].spikes[s]] = rt[gi].cuba.Vr;
            } else if (g->model == HBR_LIF_EXACT) {
                double* v = g->state;
                for (int32_t s = 0; s < rt[gi].nspk; s++) v[rt[gi].spikes[s]] = 0.0;
            }
        }
    }
    if (!S->queues) {
        for (int32_t p = 0; p < S->n_paths; p++) hbr_queue_destroy(qs[p]);
        free(qs);
    }
    for (int32_t gi = 0; gi < S->n_groups; gi++) free(rt[gi].spikes);
    free(fb.syn);
    free(fb.post);
    free(fb.cnt);
    free(rt);
    free(slot);
    return rc;
}
