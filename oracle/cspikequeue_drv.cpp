// Test-only driver: exposes the reference's header-only CSpikeQueue
// (brian2/synapses/spikequeue.h:14-206) through a C ABI so that the oracle's
// own queue restatement can be checked against it.  Built by `make -C oracle ref`
// into oracle/_ref/; never part of the product.
#include "spikequeue.h"
#include <cstring>

extern "C" {

void* csq_create(int source_start, int source_end) {
    return new CSpikeQueue(source_start, source_end);
}

void csq_destroy(void* q) { delete static_cast<CSpikeQueue*>(q); }

// CSpikeQueue::prepare (spikequeue.h:48-105) with float64 delays in seconds.
void csq_prepare(void* q, const double* delays, int n_delays, const int32_t* sources,
                 int n_synapses, double dt) {
    static_cast<CSpikeQueue*>(q)->prepare<double>(const_cast<double*>(delays), n_delays,
                                                  const_cast<int32_t*>(sources), n_synapses, dt);
}

// CSpikeQueue::push (spikequeue.h:151-191); spikes must be ascending.
void csq_push(void* q, const int32_t* spikes, int n) {
    static_cast<CSpikeQueue*>(q)->push(const_cast<int32_t*>(spikes), n);
}

// CSpikeQueue::peek (spikequeue.h:193-196): copies up to cap ids, returns the slot size.
long csq_peek(void* q, int32_t* out, long cap) {
    std::vector<int32_t>* v = static_cast<CSpikeQueue*>(q)->peek();
    long n = (long)v->size();
    long m = n < cap ? n : cap;
    if (m > 0) std::memcpy(out, v->data(), sizeof(int32_t) * (size_t)m);
    return n;
}

// CSpikeQueue::advance (spikequeue.h:198-205).
void csq_advance(void* q) { static_cast<CSpikeQueue*>(q)->advance(); }

long csq_size(void* q) { return (long)static_cast<CSpikeQueue*>(q)->queue.size(); }
int csq_offset(void* q) { return static_cast<CSpikeQueue*>(q)->offset; }

}
