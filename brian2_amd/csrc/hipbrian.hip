// libhipbrian.so: C ABI declared in include/hipbrian.h.
//
// Host runtime (contexts, hiprtc modules, step plans replayed as hipGraphs)
// plus the built-in kernels that are not generated per model: clock advance,
// 32-bit fill, SpikeGeneratorGroup replay, prefix scan, and the
// connectivity/CSR builders used by CSpikeQueue::prepare's counterpart.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hipbrian.h"
#include "hb_math.h"
#include "hb_rng.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int set_err(const std::string& m) {
    g_err = m;
    return -1;
}

#define HB_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return set_err(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int hb_scratch_for_current_device(uint32_t** out);

extern "C" int hb_version(void) { return 1; }

extern "C" int hb_last_error(char* buf, size_t n) {
    if (buf && n) {
        size_t m = g_err.size() < n - 1 ? g_err.size() : n - 1;
        memcpy(buf, g_err.data(), m);
        buf[m] = 0;
    }
    return (int)g_err.size();
}

// ---------------------------------------------------------------------------
// context, memory
// ---------------------------------------------------------------------------
#define HB_LANES 4  // lane 0: the plan's stream; lanes 1..HB_LANES-1: side streams

struct hb_ctx {
    int device;
    hipStream_t stream;                // default priority; graph replay happens here
    hipStream_t side[HB_LANES - 1];    // default priority; launches of lanes 1.. (concurrent within a step)
    hipEvent_t ev_in, ev_out, ev_fork;
    hipEvent_t ev_join[HB_LANES - 1];
};

extern "C" int hb_device_count(int* n) {
    HB_HIP(hipGetDeviceCount(n));
    return 0;
}

extern "C" int hb_ctx_create(int device, hb_ctx** out) {
    HB_HIP(hipSetDevice(device));
    std::unique_ptr<hb_ctx> c(new hb_ctx());
    c->device = device;
    HB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int l = 0; l < HB_LANES - 1; l++) {
        HB_HIP(hipStreamCreateWithFlags(&c->side[l], hipStreamNonBlocking));
        HB_HIP(hipEventCreateWithFlags(&c->ev_join[l], hipEventDisableTiming));
    }
    HB_HIP(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    HB_HIP(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
    HB_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    uint32_t* scr = nullptr;
    if (hb_scratch_for_current_device(&scr)) return -1;
    *out = c.release();
    return 0;
}

extern "C" int hb_ctx_destroy(hb_ctx* c) {
    if (!c) return 0;
    HB_HIP(hipStreamSynchronize(c->stream));
    for (int l = 0; l < HB_LANES - 1; l++) HB_HIP(hipStreamSynchronize(c->side[l]));
    HB_HIP(hipEventDestroy(c->ev_in));
    HB_HIP(hipEventDestroy(c->ev_out));
    HB_HIP(hipEventDestroy(c->ev_fork));
    for (int l = 0; l < HB_LANES - 1; l++) {
        HB_HIP(hipEventDestroy(c->ev_join[l]));
        HB_HIP(hipStreamDestroy(c->side[l]));
    }
    HB_HIP(hipStreamDestroy(c->stream));
    delete c;
    return 0;
}

extern "C" int hb_malloc(hb_ctx* c, size_t nbytes, void** dptr) {
    HB_HIP(hipSetDevice(c->device));
    HB_HIP(hipMalloc(dptr, nbytes ? nbytes : 4));
    return 0;
}

extern "C" int hb_free(hb_ctx* c, void* dptr) {
    (void)c;
    HB_HIP(hipFree(dptr));
    return 0;
}

extern "C" int hb_memcpy(void* dst, const void* src, size_t nbytes, int kind, void* stream) {
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                    : hipMemcpyDeviceToDevice;
    HB_HIP(hipMemcpyAsync(dst, src, nbytes, k, (hipStream_t)stream));
    if (kind == 2) HB_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

__global__ void hb_k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

extern "C" int hb_fill_u32(void* dptr, uint32_t value, size_t nbytes, void* stream) {
    if (nbytes % 4) return set_err("hb_fill_u32: nbytes must be a multiple of 4");
    size_t n = nbytes / 4;
    if (!n) return 0;
    unsigned grid = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(hb_k_fill_u32, dim3(grid), dim3(256), 0, (hipStream_t)stream, (uint32_t*)dptr, value, n);
    HB_HIP(hipGetLastError());
    return 0;
}

extern "C" int hb_stream_sync(void* stream) {
    HB_HIP(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

// ---------------------------------------------------------------------------
// functions: generated (module) or built in
// ---------------------------------------------------------------------------
struct hb_module;
struct hb_fn {
    hipFunction_t mfn;        // module function (hiprtc)
    const void* kernel;       // built-in __global__ taking one struct argument
    const hb_module* mod;     // module of mfn
};

static int launch_fn(const hb_fn* f, unsigned grid, unsigned block, unsigned shmem, hipStream_t s,
                     const void* args, size_t argsize) {
    if (f->mfn) {
        size_t sz = argsize;
        void* extra[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                         HIP_LAUNCH_PARAM_END};
        HB_HIP(hipModuleLaunchKernel(f->mfn, grid, 1, 1, block, 1, 1, shmem, s, nullptr, extra));
    } else {
        void* argv[] = {const_cast<void*>(args)};
        HB_HIP(hipLaunchKernel(f->kernel, dim3(grid), dim3(block), argv, shmem, s));
    }
    return 0;
}

struct hb_module {
    hb_ctx* ctx;
    hipModule_t mod;
    std::vector<std::unique_ptr<hb_fn>> fns;
    int vgpr_spills = -1, sgpr_spills = -1;  // from the code object's metadata (-1: not found)
};

// The largest value of a msgpack map key `key` (a fixstr) followed by an unsigned integer,
// over every occurrence in [p, p + n); -1 if there is none.  The kernels' metadata in an
// AMDGPU code object (the NT_AMDGPU_METADATA note) is a msgpack document with one map per
// kernel.
static long long msgpack_uint_after(const unsigned char* p, size_t n, const char* key) {
    const size_t k = strlen(key);
    long long best = -1;
    for (size_t i = 0; k < 32 && i + k + 2 <= n; i++) {
        if (p[i] != (unsigned char)(0xa0 | k) || memcmp(p + i + 1, key, k) != 0) continue;
        const unsigned char* v = p + i + 1 + k;
        const size_t left = n - (i + 1 + k);
        long long x = -1;
        if (v[0] < 0x80) x = v[0];
        else if (v[0] == 0xcc && left >= 2) x = v[1];
        else if (v[0] == 0xcd && left >= 3) x = ((long long)v[1] << 8) | v[2];
        else if (v[0] == 0xce && left >= 5)
            x = ((long long)v[1] << 24) | ((long long)v[2] << 16) | ((long long)v[3] << 8) | v[4];
        if (x > best) best = x;
    }
    return best;
}

extern "C" int hb_code_object_spills(const void* code, size_t nbytes, int* vgpr_spills, int* sgpr_spills) {
    if (!code || !vgpr_spills || !sgpr_spills) return set_err("hb_code_object_spills: null argument");
    const unsigned char* p = (const unsigned char*)code;
    *vgpr_spills = (int)msgpack_uint_after(p, nbytes, ".vgpr_spill_count");
    *sgpr_spills = (int)msgpack_uint_after(p, nbytes, ".sgpr_spill_count");
    return 0;
}

extern "C" int hb_module_load(hb_ctx* c, const char* source, const char* extra_options, hb_module** out,
                              char* log, size_t loglen) {
    HB_HIP(hipSetDevice(c->device));
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source, "hb_generated.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return set_err("hiprtcCreateProgram failed");
    std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3",     "-ffp-contract=off", "-std=c++17",
                                     "-Xclang",               "-target-feature", "-Xclang", "-packed-fp32-ops"};
    if (extra_options && *extra_options) {
        std::string e(extra_options);
        size_t p = 0;
        while (p < e.size()) {
            size_t q = e.find(' ', p);
            if (q == std::string::npos) q = e.size();
            if (q > p) opts.push_back(e.substr(p, q - p));
            p = q + 1;
        }
    }
    std::vector<const char*> o;
    for (auto& s : opts) o.push_back(s.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)o.size(), o.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string lg(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &lg[0]);
    if (log && loglen) {
        size_t m = lg.size() < loglen - 1 ? lg.size() : loglen - 1;
        memcpy(log, lg.data(), m);
        log[m] = 0;
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return set_err("hiprtc compilation failed:\n" + lg);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::unique_ptr<hb_module> m(new hb_module());
    m->ctx = c;
    hb_code_object_spills(code.data(), code.size(), &m->vgpr_spills, &m->sgpr_spills);
    HB_HIP(hipModuleLoadData(&m->mod, code.data()));
    *out = m.release();
    return 0;
}

extern "C" int hb_module_unload(hb_module* m) {
    if (!m) return 0;
    HB_HIP(hipModuleUnload(m->mod));
    delete m;
    return 0;
}

extern "C" int hb_module_function(hb_module* m, const char* name, void** fn) {
    std::unique_ptr<hb_fn> f(new hb_fn());
    f->kernel = nullptr;
    f->mod = m;
    HB_HIP(hipModuleGetFunction(&f->mfn, m->mod, name));
    *fn = f.get();
    m->fns.push_back(std::move(f));
    return 0;
}

extern "C" int hb_function_info(void* fn, int* vgprs, int64_t* scratch_bytes, int64_t* lds_bytes) {
    const hb_fn* f = (const hb_fn*)fn;
    if (!f || !f->mfn) return set_err("hb_function_info: not a generated kernel");
    int regs = 0, local = 0, shared = 0;
    HB_HIP(hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f->mfn));
    HB_HIP(hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f->mfn));
    HB_HIP(hipFuncGetAttribute(&shared, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f->mfn));
    *vgprs = regs;
    *scratch_bytes = local;
    *lds_bytes = shared;
    return 0;
}

extern "C" int hb_function_spills(void* fn, int* vgpr_spills, int* sgpr_spills) {
    const hb_fn* f = (const hb_fn*)fn;
    if (!f || !f->mfn || !f->mod) return set_err("hb_function_spills: not a generated kernel");
    *vgpr_spills = f->mod->vgpr_spills;
    *sgpr_spills = f->mod->sgpr_spills;
    return 0;
}

extern "C" int hb_launch(void* fn, unsigned grid, unsigned block, unsigned shmem, void* stream, const void* args,
                         size_t argsize) {
    if (grid == 0) return 0;
    return launch_fn((const hb_fn*)fn, grid, block, shmem, (hipStream_t)stream, args, argsize);
}

// ---------------------------------------------------------------------------
// step plans
// ---------------------------------------------------------------------------
__global__ void hb_k_clock_advance(int64_t* clock, int64_t n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) clock[0] += n;
}

struct hb_launch_rec {
    const hb_fn* fn;
    unsigned grid, block, shmem;
    std::vector<char> args;
    long step_off;
    int lane = 0;  // 0: in order on the plan's stream; l > 0: on side stream l, concurrently within the step
};

// a captured hipGraph of `steps` steps of a plan
struct hb_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int steps = 0;
};

#define HB_PLAN_GRAPHS 2  // graph lengths kept per plan (a run's whole graphs and its remainder)

struct hb_plan {
    hb_ctx* ctx;
    int64_t* clock;
    std::vector<hb_launch_rec> launches;
    hb_graph graphs[HB_PLAN_GRAPHS];
    int next_graph = 0;                                 // slot replaced by the next capture
    bool timing = false;                                // record device time of hb_plan_run calls
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timers;
};

static int graph_drop(hb_graph& g) {
    if (g.exec) HB_HIP(hipGraphExecDestroy(g.exec));
    if (g.graph) HB_HIP(hipGraphDestroy(g.graph));
    g.exec = nullptr;
    g.graph = nullptr;
    g.steps = 0;
    return 0;
}

static int plan_drop_graph(hb_plan* p) {
    for (auto& g : p->graphs) {
        int rc = graph_drop(g);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int hb_plan_create(hb_ctx* c, int64_t* clock, hb_plan** out) {
    hb_plan* p = new hb_plan();
    p->ctx = c;
    p->clock = clock;
    *out = p;
    return 0;
}

extern "C" int hb_plan_add(hb_plan* p, void* fn, unsigned grid, unsigned block, unsigned shmem, const void* args,
                           size_t argsize, long step_arg_offset) {
    if (step_arg_offset >= 0 && (size_t)step_arg_offset + 8 > argsize)
        return set_err("hb_plan_add: step offset outside the argument buffer");
    hb_launch_rec r;
    r.fn = (const hb_fn*)fn;
    r.grid = grid;
    r.block = block;
    r.shmem = shmem;
    r.args.assign((const char*)args, (const char*)args + argsize);
    r.step_off = step_arg_offset;
    p->launches.push_back(std::move(r));
    return plan_drop_graph(p);
}

extern "C" int hb_plan_set_lane(hb_plan* p, int index, int lane) {
    if (index < 0 || index >= (int)p->launches.size()) return set_err("hb_plan_set_lane: bad index");
    if (lane < 0 || lane >= HB_LANES) return set_err("hb_plan_set_lane: lane must be in [0, 4)");
    p->launches[index].lane = lane;
    return plan_drop_graph(p);
}

extern "C" int hb_plan_set_args(hb_plan* p, int index, const void* args, size_t argsize) {
    if (index < 0 || index >= (int)p->launches.size()) return set_err("hb_plan_set_args: bad index");
    auto& r = p->launches[index];
    if (argsize != r.args.size()) return set_err("hb_plan_set_args: argument size changed");
    memcpy(r.args.data(), args, argsize);
    return plan_drop_graph(p);
}

// Enqueues `k` steps (launch records with patched step offsets) and the clock advance.
// Within a step, the launches of lane l > 0 run in order on side stream l, forked after
// the previous step and joined before the next one; lane-0 launches run in order on `s`.
static int plan_enqueue(hb_plan* p, hipStream_t s, int64_t k) {
    bool used[HB_LANES] = {false};
    bool lanes = false;
    for (auto& r : p->launches)
        if (r.grid && r.lane > 0) { used[r.lane] = true; lanes = true; }
    hb_ctx* c = p->ctx;
    for (int64_t j = 0; j < k; j++) {
        if (lanes) {
            HB_HIP(hipEventRecord(c->ev_fork, s));
            for (int l = 1; l < HB_LANES; l++)
                if (used[l]) HB_HIP(hipStreamWaitEvent(c->side[l - 1], c->ev_fork, 0));
        }
        for (auto& r : p->launches) {
            if (r.grid == 0) continue;
            if (r.step_off >= 0) memcpy(r.args.data() + r.step_off, &j, 8);
            hipStream_t ls = r.lane > 0 ? c->side[r.lane - 1] : s;
            int rc = launch_fn(r.fn, r.grid, r.block, r.shmem, ls, r.args.data(), r.args.size());
            if (rc) return rc;
        }
        for (int l = 1; l < HB_LANES; l++) {
            if (!used[l]) continue;
            HB_HIP(hipEventRecord(c->ev_join[l - 1], c->side[l - 1]));
            HB_HIP(hipStreamWaitEvent(s, c->ev_join[l - 1], 0));
        }
    }
    hipLaunchKernelGGL(hb_k_clock_advance, dim3(1), dim3(64), 0, s, p->clock, k);
    HB_HIP(hipGetLastError());
    return 0;
}

static int plan_capture(hb_plan* p, hipStream_t s, int steps, hipGraphExec_t* out) {
    for (auto& g : p->graphs)
        if (g.exec && g.steps == steps) {
            *out = g.exec;
            return 0;
        }
    hb_graph& g = p->graphs[p->next_graph];
    p->next_graph = (p->next_graph + 1) % HB_PLAN_GRAPHS;
    int rc = graph_drop(g);
    if (rc) return rc;
    HB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = plan_enqueue(p, s, steps);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(s, &gr);
    if (rc) {
        if (gr) hipGraphDestroy(gr);
        return rc;
    }
    if (e != hipSuccess) return set_err(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    g.graph = gr;
    HB_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    g.steps = steps;
    *out = g.exec;
    return 0;
}

extern "C" int hb_plan_run(hb_plan* p, void* stream, int64_t n_steps, int graph_steps) {
    if (n_steps <= 0) return 0;
    hb_ctx* c = p->ctx;
    HB_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    // order after the caller's stream
    HB_HIP(hipEventRecord(c->ev_in, (hipStream_t)stream));
    HB_HIP(hipStreamWaitEvent(s, c->ev_in, 0));
    std::pair<hipEvent_t, hipEvent_t> tm(nullptr, nullptr);
    if (p->timing) {
        HB_HIP(hipEventCreate(&tm.first));
        HB_HIP(hipEventCreate(&tm.second));
        p->timers.push_back(tm);
        HB_HIP(hipEventRecord(tm.first, s));
    }
    int64_t done = 0;
    if (graph_steps > 1 && n_steps >= graph_steps) {
        hipGraphExec_t exec = nullptr;
        int rc = plan_capture(p, s, graph_steps, &exec);
        if (rc) return rc;
        int64_t reps = n_steps / graph_steps;
        for (int64_t r = 0; r < reps; r++) HB_HIP(hipGraphLaunch(exec, s));
        done = reps * graph_steps;
    }
    if (n_steps > done) {
        int rc = plan_enqueue(p, s, n_steps - done);
        if (rc) return rc;
    }
    if (p->timing) HB_HIP(hipEventRecord(tm.second, s));
    HB_HIP(hipEventRecord(c->ev_out, s));
    HB_HIP(hipStreamWaitEvent((hipStream_t)stream, c->ev_out, 0));
    return 0;
}

extern "C" int hb_plan_timing(hb_plan* p, int on, double* ms, int64_t* n_runs) {
    double total = 0.0;
    int64_t n = 0;
    int rc = 0;
    for (auto& tm : p->timers) {
        if (!rc) {
            hipError_t e = hipEventSynchronize(tm.second);
            float x = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&x, tm.first, tm.second);
            if (e != hipSuccess) rc = set_err(std::string("hb_plan_timing: ") + hipGetErrorString(e));
            total += x;
            n++;
        }
        hipEventDestroy(tm.first);
        hipEventDestroy(tm.second);
    }
    p->timers.clear();
    p->timing = on != 0;
    if (ms) *ms = total;
    if (n_runs) *n_runs = n;
    return rc;
}

extern "C" int hb_plan_profile(hb_plan* p, void* stream, int64_t n_steps, float* ms_per_launch) {
    hb_ctx* c = p->ctx;
    HB_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t nl = p->launches.size();
    // steps are enqueued in chunks with an event after every launch, so that the device never
    // waits for the host inside a chunk and each interval between events is one launch
    const int64_t chunk = 64;
    std::vector<hipEvent_t> ev((size_t)chunk * (nl + 1));
    // timing-only events: no system-scope fence (its cache writeback would be timed as part of
    // the kernel before it)
    for (auto& e : ev) HB_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (size_t k = 0; k < nl; k++) ms_per_launch[k] = 0.0f;
    HB_HIP(hipEventRecord(c->ev_in, (hipStream_t)stream));
    HB_HIP(hipStreamWaitEvent(s, c->ev_in, 0));
    int rc = 0;
    for (int64_t j0 = 0; j0 < n_steps && !rc; j0 += chunk) {
        const int64_t nj = n_steps - j0 < chunk ? n_steps - j0 : chunk;
        for (int64_t jj = 0; jj < nj && !rc; jj++) {
            const int64_t j = j0 + jj;
            hipEvent_t* e = ev.data() + (size_t)jj * (nl + 1);
            HB_HIP(hipEventRecord(e[0], s));
            for (size_t k = 0; k < nl && !rc; k++) {
                auto& r = p->launches[k];
                if (r.step_off >= 0) memcpy(r.args.data() + r.step_off, &j, 8);
                if (r.grid) rc = launch_fn(r.fn, r.grid, r.block, r.shmem, s, r.args.data(), r.args.size());
                HB_HIP(hipEventRecord(e[k + 1], s));
            }
        }
        if (rc) break;
        HB_HIP(hipEventSynchronize(ev[(size_t)(nj - 1) * (nl + 1) + nl]));
        for (int64_t jj = 0; jj < nj; jj++) {
            hipEvent_t* e = ev.data() + (size_t)jj * (nl + 1);
            for (size_t k = 0; k < nl; k++) {
                float ms = 0.0f;
                HB_HIP(hipEventElapsedTime(&ms, e[k], e[k + 1]));
                ms_per_launch[k] += ms;
            }
        }
    }
    if (!rc) {
        hipLaunchKernelGGL(hb_k_clock_advance, dim3(1), dim3(64), 0, s, p->clock, n_steps);
        HB_HIP(hipGetLastError());
    }
    HB_HIP(hipStreamSynchronize(s));
    for (auto& e : ev) hipEventDestroy(e);
    HB_HIP(hipEventRecord(c->ev_out, s));
    HB_HIP(hipStreamWaitEvent((hipStream_t)stream, c->ev_out, 0));
    HB_HIP(hipStreamSynchronize(s));
    return rc;
}

extern "C" int hb_plan_destroy(hb_plan* p) {
    if (!p) return 0;
    for (auto& tm : p->timers) {
        hipEventSynchronize(tm.second);
        hipEventDestroy(tm.first);
        hipEventDestroy(tm.second);
    }
    int rc = plan_drop_graph(p);
    delete p;
    return rc;
}

// ---------------------------------------------------------------------------
// SpikeGeneratorGroup replay (spikegenerator.pyx:6-26)
// ---------------------------------------------------------------------------
struct hb_spikegen_args {
    const int32_t* timebins;
    const int32_t* indices;
    int64_t n;
    int64_t* state;       // [0]: _lastindex
    int32_t* ring_spk;    // [ring_slots][ring_stride]
    int32_t* ring_cnt;    // [ring_slots]
    const int64_t* clock;
    int64_t step_offset;
    int32_t ring_slots;
    int32_t ring_stride;
    int32_t period_bins;
    int32_t pad;
};

__global__ void hb_k_spikegen(hb_spikegen_args a) {
    __shared__ int64_t s_lo, s_hi;
    const int64_t step = a.clock[0] + a.step_offset;
    if (threadIdx.x == 0) {
        int64_t timebin = step;
        int64_t last = a.state[0];
        if (a.period_bins > 0) {
            timebin %= a.period_bins;
            if (last > 0 && a.timebins[last - 1] >= timebin) last = 0;
        }
        // first index >= last whose timebin > current bin (upper bound)
        int64_t lo = last, hi = a.n;
        while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (a.timebins[mid] <= timebin) lo = mid + 1; else hi = mid;
        }
        s_lo = last;
        s_hi = lo;
    }
    __syncthreads();
    const int64_t lo = s_lo, hi = s_hi;
    const int slot = (int)(step % a.ring_slots);
    int32_t* out = a.ring_spk + (size_t)slot * a.ring_stride;
    for (int64_t k = lo + threadIdx.x; k < hi; k += blockDim.x) out[k - lo] = a.indices[k];
    if (threadIdx.x == 0) {
        a.ring_cnt[slot] = (int32_t)(hi - lo);
        a.ring_cnt[(step + 1) % a.ring_slots] = 0;
        a.state[0] = hi;
    }
}

static hb_fn g_spikegen_fn = {nullptr, (const void*)hb_k_spikegen};

extern "C" int hb_spikegen_function(void** fn) {
    *fn = &g_spikegen_fn;
    return 0;
}

// ---------------------------------------------------------------------------
// PoissonGroup threshold (poissongroup.py:106, "rand() < rates * dt", run by
// threshold.pyx:18-32) with a counter-based generator
// ---------------------------------------------------------------------------
struct hb_poisson_args {
    const double* rdt;       // rates * dt per neuron (the product Brian's threshold evaluates)
    const int64_t* clock;
    int64_t step_offset;
    uint64_t seed, stream;   // uniform of neuron i at step s: hb_rand64(seed, stream, s * N + i)
    int32_t* ring_spk;       // [ring_slots][ring_stride]
    int32_t* ring_cnt;       // [ring_slots]
    int32_t ring_slots, ring_stride, N, pad;
    // SpikeMonitor (mon_i == nullptr: none): (index, step) appended, per-neuron counts
    int32_t* mon_i;
    int32_t* mon_step;
    uint64_t* mon_n;
    int32_t* mon_count;
    int64_t mon_cap;
    int32_t mon_start, mon_stop;
    // PopulationRateMonitor (rate_cnt == nullptr: none): spikes in [rate_start, rate_stop)
    // of step s added to rate_cnt[s - rate_row0]
    int32_t* rate_cnt;
    int64_t rate_row0, rate_rows;
    int32_t rate_start, rate_stop;
    int32_t* status;         // bit 2: spike monitor full
};

__global__ void __launch_bounds__(256) hb_k_poisson(hb_poisson_args a) {
    __shared__ int s_cnt[3];           // spikes, monitored spikes, rate-monitored spikes of the block
    __shared__ int s_base;
    __shared__ unsigned long long s_mbase;
    const int64_t step = a.clock[0] + a.step_offset;
    const int slot = (int)(step % a.ring_slots);
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
    const bool sp = i < a.N &&
        hb_uniform01(hb_rand64(a.seed, a.stream, (uint64_t)step * (uint64_t)a.N + (uint64_t)i)) < a.rdt[i];
    const bool mon = a.mon_i && sp && i >= a.mon_start && i < a.mon_stop;
    const bool rate = a.rate_cnt && sp && i >= a.rate_start && i < a.rate_stop;
    // offsets in the block: ballot prefix within the wave, one LDS atomic per wave
    const unsigned long long b0 = __ballot(sp), b1 = __ballot(mon), b2 = __ballot(rate);
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    __syncthreads();
    int w0 = 0, w1 = 0;
    if (lane == 0) {
        if (b0) w0 = atomicAdd(&s_cnt[0], __popcll(b0));
        if (b1) w1 = atomicAdd(&s_cnt[1], __popcll(b1));
        if (b2) atomicAdd(&s_cnt[2], __popcll(b2));
    }
    w0 = __shfl(w0, 0, 64);
    w1 = __shfl(w1, 0, 64);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base = s_cnt[0] ? atomicAdd(&a.ring_cnt[slot], s_cnt[0]) : 0;
        s_mbase = s_cnt[1] ? atomicAdd(a.mon_n, (unsigned long long)s_cnt[1]) : 0ull;
        if (s_cnt[1] && s_mbase + (unsigned long long)s_cnt[1] > (unsigned long long)a.mon_cap) atomicOr(a.status, 2);
        if (s_cnt[2] && step - a.rate_row0 >= 0 && step - a.rate_row0 < a.rate_rows)
            atomicAdd(&a.rate_cnt[step - a.rate_row0], s_cnt[2]);
        if (blockIdx.x == 0) a.ring_cnt[(step + 1) % a.ring_slots] = 0;
    }
    __syncthreads();
    if (sp) a.ring_spk[(size_t)slot * a.ring_stride + s_base + w0 + __popcll(b0 & lt)] = i;
    if (mon) {
        const unsigned long long o = s_mbase + (unsigned long long)(w1 + __popcll(b1 & lt));
        if (o < (unsigned long long)a.mon_cap) { a.mon_i[o] = i; a.mon_step[o] = (int32_t)step; }
        a.mon_count[i - a.mon_start] += 1;
    }
}

static hb_fn g_poisson_fn = {nullptr, (const void*)hb_k_poisson};

extern "C" int hb_poisson_function(void** fn) {
    *fn = &g_poisson_fn;
    return 0;
}

// ---------------------------------------------------------------------------
// prefix scan (uint32 counts -> n+1 offsets), three phases
// ---------------------------------------------------------------------------
#define HB_SCAN_BLOCK 1024  // elements per workgroup (256 threads x 4)

__device__ inline uint32_t hb_block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // up to 1024 threads: wave64 scan, then a scan of the wave sums; sh holds 17 words
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            uint32_t t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[16] = acc;
    }
    __syncthreads();
    uint32_t r = x - v + sh[wid];
    *total = sh[16];
    __syncthreads();
    return r;
}

__global__ void hb_k_scan_reduce(const uint32_t* in, int64_t n, uint32_t* block_sums) {
    __shared__ uint32_t sh[17];
    int64_t base = (int64_t)blockIdx.x * HB_SCAN_BLOCK + threadIdx.x * 4;
    uint32_t s = 0;
    for (int k = 0; k < 4; k++)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    hb_block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void hb_k_scan_sums(uint32_t* sums, int64_t nb) {
    // single workgroup: exclusive scan of nb block sums, chunk by chunk
    __shared__ uint32_t sh[17];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < nb; c0 += blockDim.x) {
        int64_t i = c0 + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t tot;
        uint32_t ex = hb_block_exclusive_scan(v, sh, &tot);
        if (i < nb) sums[i] = ex + carry;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
}

__global__ void hb_k_scan_apply(uint32_t* data, int64_t n, const uint32_t* block_off, uint32_t* last) {
    __shared__ uint32_t sh[17];
    int64_t base = (int64_t)blockIdx.x * HB_SCAN_BLOCK + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? data[base + k] : 0;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = hb_block_exclusive_scan(s, sh, &tot) + block_off[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        if (base + k < n) data[base + k] = ex;
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) *last = ex;
}

extern "C" int hb_scan_u32(uint32_t* d, int64_t n, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (n <= 0) {
        hipLaunchKernelGGL(hb_k_fill_u32, dim3(1), dim3(64), 0, s, d, 0u, (size_t)1);
        HB_HIP(hipGetLastError());
        return 0;
    }
    int64_t nb = (n + HB_SCAN_BLOCK - 1) / HB_SCAN_BLOCK;
    uint32_t* sums = nullptr;
    HB_HIP(hipMallocAsync((void**)&sums, (size_t)nb * 4, s));
    hipLaunchKernelGGL(hb_k_scan_reduce, dim3((unsigned)nb), dim3(256), 0, s, d, n, sums);
    hipLaunchKernelGGL(hb_k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nb);
    hipLaunchKernelGGL(hb_k_scan_apply, dim3((unsigned)nb), dim3(256), 0, s, d, n, sums, d + n);
    HB_HIP(hipGetLastError());
    HB_HIP(hipFreeAsync(sums, s));
    return 0;
}

// The allocation-free scan keeps its block sums in a file-scope device array (one
// copy per GPU, held until the process ends).  Host code never passes `scratch`
// itself to a kernel: its host-side name is not a device address.  The address of
// each GPU's copy is looked up once per device, with that device current
// (hb_ctx_create, or the entry point's first call on a device), and used instead.
#define HB_SCAN_SCRATCH_SUMS 65536  // block sums: up to 65536 * 1024 elements per call
#define HB_MAX_DEVICES 64
static __device__ __align__(256) uint32_t scratch[HB_SCAN_SCRATCH_SUMS];
static uint32_t* g_scratch_ptr[HB_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int hb_scratch_for_current_device(uint32_t** out) {
    int dev = 0;
    HB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= HB_MAX_DEVICES) return set_err("hb_scan_u32_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err("hb_scan_u32_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int hb_scan_u32_noalloc(uint32_t* d, int64_t n, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int64_t nmax = (int64_t)HB_SCAN_SCRATCH_SUMS * HB_SCAN_BLOCK;
    if (n > nmax)
        return set_err("hb_scan_u32_noalloc: " + std::to_string(n) + " counts exceed the maximum of " +
                       std::to_string(nmax) + " (use hb_scan_u32)");
    uint32_t* sums = nullptr;
    if (hb_scratch_for_current_device(&sums)) return -1;
    if (n <= 0) {
        hipLaunchKernelGGL(hb_k_fill_u32, dim3(1), dim3(64), 0, s, d, 0u, (size_t)1);
        HB_HIP(hipGetLastError());
        return 0;
    }
    // every block sum the later kernels read is written by hb_k_scan_reduce first,
    // so the scratch needs no clearing between calls
    int64_t nb = (n + HB_SCAN_BLOCK - 1) / HB_SCAN_BLOCK;
    hipLaunchKernelGGL(hb_k_scan_reduce, dim3((unsigned)nb), dim3(256), 0, s, d, n, sums);
    hipLaunchKernelGGL(hb_k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nb);
    hipLaunchKernelGGL(hb_k_scan_apply, dim3((unsigned)nb), dim3(256), 0, s, d, n, sums, d + n);
    HB_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// connectivity: per-(source, delay) counts, synthetic generators
// ---------------------------------------------------------------------------
__global__ void hb_k_csr_count(const int32_t* pre, const int32_t* dsteps, int64_t n, int32_t src0, int32_t src1,
                               int32_t dmin, int32_t n_seg, uint32_t* cnt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t i = pre[e];
        if (i < src0 || i >= src1) continue;
        int32_t d = dsteps ? dsteps[e] - dmin : 0;
        atomicAdd(&cnt[(int64_t)(i - src0) * n_seg + d], 1u);
    }
}

extern "C" int hb_csr_count(const int32_t* pre, const int32_t* delay_steps, int64_t n_syn, int32_t src0, int32_t src1,
                            int32_t dmin, int32_t n_seg, uint32_t* seg_count, void* stream) {
    if (n_syn <= 0) return 0;
    unsigned grid = (unsigned)((n_syn + 255) / 256 < 65536 ? (n_syn + 255) / 256 : 65536);
    hipLaunchKernelGGL(hb_k_csr_count, dim3(grid), dim3(256), 0, (hipStream_t)stream, pre, delay_steps, n_syn, src0,
                       src1, dmin, n_seg, seg_count);
    HB_HIP(hipGetLastError());
    return 0;
}

// Fixed in-degree: one workgroup (256 threads) per target, K <= 4096.
#define HB_INDEG_MAX 4096

__device__ void hb_bitonic_sort_i32(int32_t* a, int n2) {
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                int l = i ^ j;
                if (l > i) {
                    int32_t x = a[i], y = a[l];
                    bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(256) hb_k_gen_indegree(uint64_t seed, uint64_t stream_id, int32_t n_pre,
                                                        int32_t n_post, int32_t K, int32_t* out, int32_t* status) {
    __shared__ int32_t buf[HB_INDEG_MAX];
    __shared__ int32_t cnt[256];
    __shared__ int32_t s_have;
    int n2 = 1;
    while (n2 < K) n2 <<= 1;
    for (int32_t j = blockIdx.x; j < n_post; j += gridDim.x) {
        if (threadIdx.x == 0) s_have = 0;
        __syncthreads();
        int r = 0;
        while (s_have < K) {
            int have = s_have;
            if (r >= 64) {
                if (threadIdx.x == 0) atomicExch(status, 1);
                return;
            }
            int m = K - have;
            for (int t = threadIdx.x; t < m; t += blockDim.x) {
                uint64_t ctr = ((uint64_t)j * 64 + (uint64_t)r) * (uint64_t)K + (uint64_t)t;
                buf[have + t] = (int32_t)hb_uniform_int(hb_rand64(seed, stream_id, ctr), (uint32_t)n_pre);
            }
            for (int t = K + threadIdx.x; t < n2; t += blockDim.x) buf[t] = 0x7fffffff;
            __syncthreads();
            hb_bitonic_sort_i32(buf, n2);
            // de-duplicate: each thread owns a contiguous chunk of the K sorted values
            int chunk = (K + blockDim.x - 1) / blockDim.x;
            int b0 = threadIdx.x * chunk, b1 = min(K, b0 + chunk);
            int c = 0;
            for (int t = b0; t < b1; t++) c += (t == 0 || buf[t] != buf[t - 1]);
            cnt[threadIdx.x] = c;
            __syncthreads();
            if (threadIdx.x == 0) {
                int acc = 0;
                for (int w = 0; w < (int)blockDim.x; w++) {
                    int x = cnt[w];
                    cnt[w] = acc;
                    acc += x;
                }
                s_have = acc;
            }
            __syncthreads();
            int32_t keep[16];
            int nk = 0;
            for (int t = b0; t < b1; t++)
                if (t == 0 || buf[t] != buf[t - 1]) keep[nk++] = buf[t];
            __syncthreads();
            for (int q = 0; q < nk; q++) buf[cnt[threadIdx.x] + q] = keep[q];
            __syncthreads();
            r++;
        }
        for (int t = threadIdx.x; t < K; t += blockDim.x) out[(size_t)j * K + t] = buf[t];
        __syncthreads();
    }
}

extern "C" int hb_gen_fixed_indegree(uint64_t seed, uint64_t stream_id, int32_t n_pre, int32_t n_post, int32_t K,
                                     int32_t* pre_of_post, int32_t* status, void* stream) {
    if (K > HB_INDEG_MAX) return set_err("hb_gen_fixed_indegree: K > 4096 is not supported");
    if (K > n_pre) return set_err("hb_gen_fixed_indegree: K larger than the source population");
    if ((K + 255) / 256 > 16) return set_err("hb_gen_fixed_indegree: K too large for the dedupe chunking");
    unsigned grid = (unsigned)(n_post < 8192 ? n_post : 8192);
    if (grid == 0) return 0;
    hipLaunchKernelGGL(hb_k_gen_indegree, dim3(grid), dim3(256), 0, (hipStream_t)stream, seed, stream_id, n_pre,
                       n_post, K, pre_of_post, status);
    HB_HIP(hipGetLastError());
    return 0;
}

// CSR by (source, delay, target) from the fixed in-degree lists.
__global__ void hb_k_indeg_count(const int32_t* pre_of_post, int32_t n_post, int32_t K, int32_t src0, int32_t src1,
                                 uint64_t seed, uint64_t dstream, int32_t n_seg, uint32_t* cnt) {
    int64_t n = (int64_t)n_post * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t i = pre_of_post[e];
        if (i < src0 || i >= src1) continue;
        int32_t j = (int32_t)(e / K);
        uint32_t d = hb_uniform_int(hb_rand64(seed, dstream, (uint64_t)i * (uint64_t)n_post + (uint64_t)j), (uint32_t)n_seg);
        atomicAdd(&cnt[(int64_t)(i - src0) * n_seg + d], 1u);
    }
}

__global__ void hb_k_indeg_fill(const int32_t* pre_of_post, int32_t n_post, int32_t K, int32_t src0, int32_t src1,
                                uint64_t seed, uint64_t dstream, int32_t n_seg, uint32_t* cursor, int32_t* post) {
    int64_t n = (int64_t)n_post * K;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t i = pre_of_post[e];
        if (i < src0 || i >= src1) continue;
        int32_t j = (int32_t)(e / K);
        uint32_t d = hb_uniform_int(hb_rand64(seed, dstream, (uint64_t)i * (uint64_t)n_post + (uint64_t)j), (uint32_t)n_seg);
        uint32_t pos = atomicAdd(&cursor[(int64_t)(i - src0) * n_seg + d], 1u);
        post[pos] = j;
    }
}

// one thread per (source, delay) segment: sort targets ascending, then weights
__global__ void hb_k_indeg_finish(const uint32_t* seg_ptr, int64_t n_segs, int32_t n_seg, int32_t src0,
                                  int32_t n_post, uint64_t seed, uint64_t wstream, double w_base, double w_lo,
                                  double w_hi, int32_t* post, float* w) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_segs; s += (int64_t)gridDim.x * blockDim.x) {
        uint32_t a = seg_ptr[s], b = seg_ptr[s + 1];
        for (uint32_t x = a + 1; x < b; x++) {
            int32_t v = post[x];
            uint32_t y = x;
            while (y > a && post[y - 1] > v) {
                post[y] = post[y - 1];
                y--;
            }
            post[y] = v;
        }
        if (w) {
            int32_t i = src0 + (int32_t)(s / n_seg);
            for (uint32_t x = a; x < b; x++) {
                uint64_t ctr = (uint64_t)i * (uint64_t)n_post + (uint64_t)post[x];
                float u = (float)(w_lo + hb_uniform01(hb_rand64(seed, wstream, ctr)) * (w_hi - w_lo));
                w[x] = (float)((double)u * w_base);
            }
        }
    }
}

extern "C" int hb_csr_from_indegree(const int32_t* pre_of_post, int32_t n_post, int32_t K, int32_t src0, int32_t src1,
                                    uint64_t seed, uint64_t delay_stream, int32_t dmin, int32_t n_seg, uint64_t w_stream,
                                    double w_base, double w_lo, double w_hi, uint32_t* seg_ptr, uint32_t* cursor,
                                    int32_t* post, float* w, void* stream) {
    (void)dmin;
    hipStream_t s = (hipStream_t)stream;
    int64_t n_segs = (int64_t)(src1 - src0) * n_seg;
    int64_t n = (int64_t)n_post * K;
    unsigned grid = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    int rc = hb_fill_u32(seg_ptr, 0u, (size_t)(n_segs + 1) * 4, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(hb_k_indeg_count, dim3(grid), dim3(256), 0, s, pre_of_post, n_post, K, src0, src1, seed,
                       delay_stream, n_seg, seg_ptr);
    HB_HIP(hipGetLastError());
    rc = hb_scan_u32(seg_ptr, n_segs, stream);
    if (rc) return rc;
    HB_HIP(hipMemcpyAsync(cursor, seg_ptr, (size_t)n_segs * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(hb_k_indeg_fill, dim3(grid), dim3(256), 0, s, pre_of_post, n_post, K, src0, src1, seed,
                       delay_stream, n_seg, cursor, post);
    unsigned g2 = (unsigned)((n_segs + 255) / 256 < 65536 ? (n_segs + 255) / 256 : 65536);
    hipLaunchKernelGGL(hb_k_indeg_finish, dim3(g2 ? g2 : 1), dim3(256), 0, s, seg_ptr, n_segs, n_seg, src0, n_post, seed,
                       w_stream, w_base, w_lo, w_hi, post, w);
    HB_HIP(hipGetLastError());
    return 0;
}

// Bernoulli rows: one workgroup per source row, targets scanned in chunks of 256.
__global__ void __launch_bounds__(256) hb_k_bernoulli_rows(uint64_t seed, uint64_t sid, int32_t src0, int32_t src1,
                                                          int32_t n_post, double p, uint32_t* rc, int32_t* post) {
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int32_t i = src0 + blockIdx.x; i < src1; i += gridDim.x) {
        uint32_t base = post ? rc[i - src0] : 0;
        uint32_t total = 0;
        for (int32_t j0 = 0; j0 < n_post; j0 += blockDim.x) {
            int32_t j = j0 + threadIdx.x;
            bool f = false;
            if (j < n_post) f = hb_uniform01(hb_rand64(seed, sid, (uint64_t)i * (uint64_t)n_post + (uint64_t)j)) < p;
            uint64_t b = __ballot(f);
            uint32_t rank = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
            if (lane == 0) wsum[wid] = (uint32_t)__popcll(b);
            __syncthreads();
            uint32_t off = 0, tot = 0;
            for (int w = 0; w < 4; w++) {
                if (w < wid) off += wsum[w];
                tot += wsum[w];
            }
            if (post && f) post[base + total + off + rank] = j;
            total += tot;
            __syncthreads();
        }
        if (!post && threadIdx.x == 0) rc[i - src0] = total;
    }
}

extern "C" int hb_gen_bernoulli_rows(uint64_t seed, uint64_t stream_id, int32_t src0, int32_t src1, int32_t n_post,
                                     double p, uint32_t* row_count_or_ptr, int32_t* post, void* stream) {
    int32_t rows = src1 - src0;
    if (rows <= 0) return 0;
    unsigned grid = (unsigned)(rows < 16384 ? rows : 16384);
    hipLaunchKernelGGL(hb_k_bernoulli_rows, dim3(grid), dim3(256), 0, (hipStream_t)stream, seed, stream_id, src0, src1,
                       n_post, p, row_count_or_ptr, post);
    HB_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// host-side elementary functions (the same code as on the device)
// ---------------------------------------------------------------------------
extern "C" double hb_host_exp(double x) { return hb_exp(x); }
extern "C" double hb_host_expm1(double x) { return hb_expm1(x); }
extern "C" double hb_host_log(double x) { return hb_log(x); }
extern "C" double hb_host_pow(double x, double y) { return hb_pow(x, y); }
extern "C" double hb_host_exprel(double x) { return hb_exprel(x); }
