/* hbref.h -- ORACLE (test infrastructure only).
 *
 * A plain-C, single-threaded restatement of Brian2's clock-driven step as the
 * cython runtime executes it.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load this library, and only as the checker.
 * The product path never links, imports or calls it.
 *
 * Pinning (what each part is checked against) is listed in hbref.c's header.
 */
#ifndef HBREF_H
#define HBREF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- input generator (the inputs of the configurations) ------------------ */
uint64_t hbr_rand64(uint64_t seed, uint64_t stream, uint64_t ctr);
double hbr_uniform(uint64_t seed, uint64_t stream, uint64_t ctr);
void hbr_uniform_fill(uint64_t seed, uint64_t stream, uint64_t ctr0, int64_t n, double* out);
uint32_t hbr_uniform_int(uint64_t h, uint32_t n);
int64_t hbr_bernoulli(uint64_t seed, uint64_t stream, int32_t i0, int32_t i1, int32_t n_post,
                      double p, int32_t* pre, int32_t* post, int64_t cap);
int hbr_fixed_indegree(uint64_t seed, uint64_t stream, int32_t n_pre, int32_t n_post, int32_t K,
                       int32_t* pre_of_post);
void hbr_order_by_pre(const int32_t* pre_of_post, int32_t n_pre, int32_t n_post, int32_t K, int32_t* pre,
                      int32_t* post);
void hbr_pair_int(uint64_t seed, uint64_t stream, int32_t n_post, const int32_t* pre,
                  const int32_t* post, int64_t n, int32_t lo, int32_t hi, int32_t* out);
void hbr_pair_uniform_f32(uint64_t seed, uint64_t stream, int32_t n_post, const int32_t* pre,
                          const int32_t* post, int64_t n, double lo, double hi, float* out);

/* ---- CSpikeQueue restatement ------------------------------------------------ */
typedef struct hbr_queue hbr_queue;
hbr_queue* hbr_queue_create(int32_t source_start, int32_t source_end);
void hbr_queue_destroy(hbr_queue* q);
int hbr_queue_prepare(hbr_queue* q, const double* delays, int64_t n_delays, const int32_t* sources,
                      int64_t n_synapses, double dt);
void hbr_queue_push(hbr_queue* q, const int32_t* spikes, int32_t n);
int64_t hbr_queue_peek(hbr_queue* q, int32_t* out, int64_t cap);
void hbr_queue_advance(hbr_queue* q);
int64_t hbr_queue_size(hbr_queue* q);
int32_t hbr_queue_offset(hbr_queue* q);

/* ---- network step restatement ---------------------------------------------- */
enum { HBR_CUBA_EXACT = 1, HBR_COBAHH_EE = 2, HBR_LIF_EXACT = 3, HBR_SPIKEGEN = 4, HBR_PASSIVE = 5, HBR_POISSON = 6 };

typedef struct {
    int32_t model;          /* HBR_* */
    int32_t N;
    double* state;          /* nvar * N, variable-major; in/out */
    double* lastspike;      /* N; in/out (NULL if no refractoriness) */
    uint8_t* not_refractory;/* N; in/out */
    const double* params;   /* model parameters, see hbref.c */
    /* HBR_SPIKEGEN: spikes sorted by (timebin, index) */
    int64_t n_gen;
    const int32_t* gen_index;
    const int32_t* gen_timebin;
    int64_t gen_lastindex;  /* in/out */
    /* HBR_POISSON: neuron i spikes at step s when hbr_uniform(seed, stream, s*N + i) < rdt[i] */
    uint64_t poisson_seed, poisson_stream;
    const double* poisson_rdt;
} hbr_group;

typedef struct {
    int32_t source_group, target_group;
    int32_t source_start, source_end; /* absolute index range of the source (subgroup clip) */
    int64_t n_syn;
    const int32_t* pre;     /* absolute source index per synapse, in synapse order */
    const int32_t* post;    /* target index per synapse */
    const double* w;        /* per-synapse increment, or NULL for w_scalar */
    double w_scalar;
    const double* delay;    /* per-synapse delay (s), or NULL for delay_scalar */
    double delay_scalar;
    int32_t target_var;     /* index of the incremented state variable of the target */
    int32_t* counts;        /* optional per-synapse event counter (in/out) */
} hbr_pathway;

typedef struct {
    double dt;
    int64_t start_step, n_steps;
    int32_t n_groups;
    hbr_group* groups;
    int32_t n_paths;
    hbr_pathway* paths;     /* in Brian's schedule order */
    /* spike monitor of group `mon_group` (-1: none) */
    int32_t mon_group;
    int32_t* spk_i;
    int64_t* spk_step;
    int64_t spk_cap;
    int64_t n_spk;          /* out */
    /* state monitor (when='start'): record `rec_vars` of `rec_group` at rec_idx */
    int32_t rec_group;
    int32_t n_rec, n_rec_vars;
    const int32_t* rec_idx;
    const int32_t* rec_vars;
    double* rec_out;        /* [n_steps][n_rec_vars][n_rec] */
    /* threads > 1: the neuron updates and each pathway's fold run on that many OpenMP
       threads (the fold split by target, each target's events still applied in queue
       order: the results are identical to one thread's) */
    int32_t threads;
    /* queues: NULL, or n_paths queue handles kept across calls the way a pathway keeps
       its SpikeQueue across runs (synapses.py:421-437): a NULL entry is created on the
       first call; every call prepares each queue with this call's dt, which re-bins the
       spikes in flight when dt changed (spikequeue.h:58-74).  The caller destroys them
       with hbr_queue_destroy. */
    hbr_queue** queues;
} hbr_sim;

int hbr_run(hbr_sim* sim);
const char* hbr_last_error(void);

/* elementary functions shared with the device (hb_math.h), for tests */
double hbr_exp(double x);
double hbr_expm1(double x);
double hbr_log(double x);
double hbr_pow(double x, double y);

#ifdef __cplusplus
}
#endif
#endif
