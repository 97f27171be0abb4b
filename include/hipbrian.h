/* hipbrian.h -- C ABI of libhipbrian.so, the MI355X (gfx950) back end of the
 * `hip` code generation target for Brian2's clock-driven step.
 *
 * All functions return 0 on success and a negative code on failure; the
 * message is then available from hb_last_error() (per thread).  Device
 * pointers are plain `void*` obtained from hb_malloc or from any other HIP
 * allocation (e.g. a torch tensor's data_ptr()).  Streams are `void*`
 * hipStream_t handles; NULL means the default stream.
 *
 * Reference counterparts (paths relative to the brian2 source tree):
 *  - The per-model kernels are Brian's code objects: hb_module_load replaces
 *    CythonCodeObject.compile_block (codegen/runtime/cython_rt/cython_rt.py:177-193)
 *    and hb_launch replaces CythonCodeObject.run_block (cython_rt.py:195-205).
 *  - hb_plan_run replaces the Network.run hot loop (core/network.py:1194-1250)
 *    together with Clock.advance (core/clocks.py:315-325): the ordered per-step
 *    code objects are replayed n_steps times on the device.
 *  - hb_csr_* replace CSpikeQueue::prepare (synapses/spikequeue.h:48-105): per-source
 *    synapse lists in synapse order, with the integer delay (int)(d/dt + 0.5)
 *    (spikequeue.h:90) made a segment of each list.
 *  - hb_poisson_function's kernel replaces PoissonGroup's Thresholder
 *    (input/poissongroup.py:106 with codegen/runtime/cython_rt/templates/threshold.pyx).
 *  - hb_spikegen_function's kernel replaces the spikegenerator template
 *    (codegen/runtime/cython_rt/templates/spikegenerator.pyx:6-26).
 */
#ifndef HIPBRIAN_H
#define HIPBRIAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hb_ctx hb_ctx;
typedef struct hb_module hb_module;
typedef struct hb_plan hb_plan;

/* ---- errors / version --------------------------------------------------- */
int hb_version(void);
/* Copies the calling thread's last error message; returns its full length. */
int hb_last_error(char* buf, size_t n);

/* ---- context ------------------------------------------------------------ */
/* One context per device: owns a default-priority stream used for graph
 * replay, and the modules/plans created from it.  (Device.activate,
 * devices/device.py:86-467, RuntimeDevice 470-615) */
int hb_ctx_create(int device, hb_ctx** out);
int hb_ctx_destroy(hb_ctx* ctx);
int hb_device_count(int* n);

/* ---- memory (RuntimeDevice.add_array / get_value / set_value,
 *      devices/device.py:500-555) ---------------------------------------- */
int hb_malloc(hb_ctx* ctx, size_t nbytes, void** dptr);
int hb_free(hb_ctx* ctx, void* dptr);
/* kind: 1 host->device, 2 device->host, 3 device->device (ordered on stream) */
int hb_memcpy(void* dst, const void* src, size_t nbytes, int kind, void* stream);
/* Fills nbytes (multiple of 4) with a 32-bit pattern using a kernel. */
int hb_fill_u32(void* dptr, uint32_t value, size_t nbytes, void* stream);
int hb_stream_sync(void* stream);

/* ---- per-model kernels (CodeObject.compile / __call__) ------------------ */
/* Compiles HIP source for gfx950 with hiprtc (-O3 -ffp-contract=off, packed
 * fp32 off) and loads it.  `log` (may be NULL) receives the compiler log. */
int hb_module_load(hb_ctx* ctx, const char* source, const char* extra_options, hb_module** out, char* log,
                   size_t loglen);
int hb_module_unload(hb_module* mod);
int hb_module_function(hb_module* mod, const char* name, void** fn);
/* Resources of a generated kernel: registers per thread, private (spill) memory
 * per thread, static LDS per workgroup.  No counterpart in Brian (the cython
 * runtime compiles once per code object); the runtime uses it to pick between a
 * two-workgroups-per-CU build of a step kernel and a one-workgroup build. */
int hb_function_info(void* fn, int* vgprs, int64_t* scratch_bytes, int64_t* lds_bytes);
/* Registers spilled to memory by a generated kernel's compilation (VGPRs and SGPRs,
 * the code object's .vgpr_spill_count / .sgpr_spill_count; -1 when its metadata does not
 * say).  A private segment can be reserved with nothing spilled, so this, not the
 * scratch size, decides whether a build for two workgroups per CU is kept.  No
 * counterpart in Brian. */
int hb_function_spills(void* fn, int* vgpr_spills, int* sgpr_spills);
/* The same counts read from a code object in memory (largest over its kernels). */
int hb_code_object_spills(const void* code, size_t nbytes, int* vgpr_spills, int* sgpr_spills);
/* Launches a module function with a packed argument buffer (the kernel takes
 * one struct by value whose bytes are `args`). */
int hb_launch(void* fn, unsigned grid, unsigned block, unsigned shmem, void* stream, const void* args,
              size_t argsize);

/* ---- step plans (Network.run hot loop + Clock.advance) ------------------ */
/* A plan is the ordered list of launches of one time step.  `clock` points to a
 * device int64 holding the current time step; each launch's argument buffer
 * may hold an int64 "step offset" at byte offset `step_arg_offset` (or -1):
 * replaying k steps patches it with 0..k-1 and then advances *clock by k. */
int hb_plan_create(hb_ctx* ctx, int64_t* clock, hb_plan** out);
int hb_plan_add(hb_plan* plan, void* fn, unsigned grid, unsigned block, unsigned shmem, const void* args,
                size_t argsize, long step_arg_offset);
/* Puts launch `index` in lane 0 (default: in plan order on the plan's stream) or
 * lane 1, 2 or 3 (further default-priority streams, forked at the start of each
 * step and joined at its end): the launches of different lanes of a step may run
 * concurrently, those of one lane run in plan order.  For pathways whose delays
 * are all at least one step, the propagation of step s reads only spikes of
 * earlier steps, so it can overlap the group update of step s and the other
 * pathways' propagation (Network.run's per-step order, core/network.py:
 * 1194-1250, is kept wherever results depend on it). */
int hb_plan_set_lane(hb_plan* plan, int index, int lane);
/* Replaces the argument buffer of launch `index` (drops a captured graph). */
int hb_plan_set_args(hb_plan* plan, int index, const void* args, size_t argsize);
/* Runs n_steps steps, replaying a hipGraph of `graph_steps` steps when
 * graph_steps > 1 (the remaining n_steps mod graph_steps steps are enqueued one
 * launch at a time), ordered after prior work on `stream`.  A plan keeps the
 * graphs of its two most recently used lengths.  Returns after enqueueing. */
int hb_plan_run(hb_plan* plan, void* stream, int64_t n_steps, int graph_steps);
/* Device time of hb_plan_run calls (Network.run's profile / get_profiling_info,
 * core/network.py:461-486, at the granularity of a run call): returns in
 * *ms the summed time between a HIP event recorded on the plan's stream before the
 * first launch of each hb_plan_run call since the previous hb_plan_timing call and one
 * after its last launch, and in *n_runs the number of such calls (waits for them);
 * then turns the recording on (on != 0) or off for the calls that follow. */
int hb_plan_timing(hb_plan* plan, int on, double* ms, int64_t* n_runs);
/* Runs n_steps steps as individual launches with a HIP event after each one, on
 * the plan's stream, and accumulates each launch's duration over the steps into
 * ms_per_launch[index] (synchronous). */
int hb_plan_profile(hb_plan* plan, void* stream, int64_t n_steps, float* ms_per_launch);
int hb_plan_destroy(hb_plan* plan);

/* ---- built-in kernels --------------------------------------------------- */
/* SpikeGeneratorGroup step (spikegenerator.pyx:6-26) for the step *clock + k:
 * emits neuron_index[lastindex .. first timebin > step) into the spike ring
 * slot (step mod ring_slots); `state` holds {lastindex}. period_bins > 0
 * repeats the pattern.  Launch from a plan with hb_spikegen_function. */
int hb_spikegen_function(void** fn);
/* PoissonGroup step (input/poissongroup.py:106: threshold "rand() < rates * dt",
 * run by codegen/runtime/cython_rt/templates/threshold.pyx:18-32) for the step
 * s = *clock + k: neuron i spikes when the uniform hb_rand64(seed, stream, s*N + i)
 * mapped to [0, 1) is below rates[i] * dt (computed once on the host, the same fp64
 * product), so a run is reproducible from (seed, step) alone and store/restore
 * needs no generator state.  Spikes go to the ring slot (step mod ring_slots);
 * optional SpikeMonitor (spikemonitor.pyx:9-58) and PopulationRateMonitor
 * (ratemonitor.pyx:7-48) counts.  Launch from a plan with hb_poisson_function
 * (argument block: struct hb_poisson_args in hipbrian.hip, 256 threads per
 * workgroup). */
int hb_poisson_function(void** fn);

/* ---- connectivity / CSR (CSpikeQueue::prepare) -------------------------- */
/* Counts synapses per (source, delay segment): seg_count[(pre-src0)*n_seg + d-dmin]
 * for pre in [src0, src1).  delay_steps may be NULL (scalar delay dmin). */
int hb_csr_count(const int32_t* pre, const int32_t* delay_steps, int64_t n_syn, int32_t src0, int32_t src1,
                 int32_t dmin, int32_t n_seg, uint32_t* seg_count, void* stream);
/* In-place exclusive scan of n uint32 counts into n+1 offsets (last = total). */
int hb_scan_u32(uint32_t* counts_to_offsets, int64_t n, void* stream);
/* hb_scan_u32 without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight loops, harnesses that forbid hipMalloc).  Its block sums live in a
 * fixed device array of the library (256 KiB per GPU, held until the process
 * ends), so n is at most 65536 * 1024 = 67108864 counts; larger n is rejected
 * with an error and nothing is launched.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (different streams or threads) race. */
int hb_scan_u32_noalloc(uint32_t* counts_to_offsets, int64_t n, void* stream);

/* Synthetic connectivity (deterministic, counter-based SplitMix64 streams):
 * fixed in-degree K per target, sources ascending per target. */
int hb_gen_fixed_indegree(uint64_t seed, uint64_t stream_id, int32_t n_pre, int32_t n_post, int32_t K,
                          int32_t* pre_of_post, int32_t* status, void* stream);
/* CSR by (source, delay, target) from pre_of_post (n_post x K): for sources in
 * [src0, src1) with pair delays dmin + U{0..n_seg-1} (stream delay_stream)
 * and fp32 weights w_base * U(w_lo, w_hi) (stream w_stream).  seg_ptr has
 * (src1-src0)*n_seg+1 entries; post/w have seg_ptr[last] entries. */
int hb_csr_from_indegree(const int32_t* pre_of_post, int32_t n_post, int32_t K, int32_t src0, int32_t src1,
                         uint64_t seed, uint64_t delay_stream, int32_t dmin, int32_t n_seg, uint64_t w_stream,
                         double w_base, double w_lo, double w_hi, uint32_t* seg_ptr, uint32_t* cursor,
                         int32_t* post, float* w, void* stream);
/* Bernoulli(p) connectivity keyed by (i*n_post + j), CSR by source for
 * i in [src0, src1), targets ascending: first call with post == NULL fills
 * row_count, then (after hb_scan_u32 into row_ptr) with post != NULL. */
int hb_gen_bernoulli_rows(uint64_t seed, uint64_t stream_id, int32_t src0, int32_t src1, int32_t n_post,
                          double p, uint32_t* row_count_or_ptr, int32_t* post, void* stream);

/* ---- host-side elementary functions (identical to the device ones) ------ */
double hb_host_exp(double x);
double hb_host_expm1(double x);
double hb_host_log(double x);
double hb_host_pow(double x, double y);
double hb_host_exprel(double x);

#ifdef __cplusplus
}
#endif
#endif
