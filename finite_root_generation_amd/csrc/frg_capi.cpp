// frg_capi.cpp -- the C ABI of include/frg.h over the HIP runtime.
//
// Owns: the device, one stream (default priority), device buffers grown on demand, and the
// launch-shape choice per (schedule, precision).  Replaces the host side of the reference's
// round (frg.cu:138-236 set-up, :356-404 launch and copies, :518-531 tear-down).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/frg.h"
#include "frg_dedupe_kernels.h"
#include "frg_kernels.h"
#include "frg_mlp_kernels.h"

using frg::cx;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

}  // namespace

struct frg_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    // ev0 -> ev_mid: the tracking kernel; ev_mid -> ev1: the refinement kernel
    hipEvent_t ev0 = nullptr, ev_mid = nullptr, ev1 = nullptr;
    // recorded after work that an entry point leaves running on a caller's stream
    // (frg_track_device, frg_sort_dedupe_merge_device); every entry point waits for it before it
    // touches the context's buffers
    hipEvent_t ev_busy = nullptr;
    frg_opts opts{};
    DevBuf inst_first, fpm, gamma, seed, pid0, spt, spm, fpt, err, stats, yfinal, queue, dbg_a, dbg_b, dbg_c,
        dbg_d, dbg_e, dbg_x;
    // step model: FP16 weights on the device, normalisation on the host
    DevBuf model16, steplog, steplog_count, dedupe_tmp;
    bool have_model = false;
    frg::StepModel model{};
    // training state
    DevBuf tw, tm, tv, tgrad, tsamples, tnorm, tloss, tidx, tw16;
    // cached graph of the launch sequence (queue reset + tracker) and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    struct GraphKey {
        frg::Job job;
        int kind, block, grid, lanes, waves;
    } graph_key;
    bool graph_valid = false;
    // this GPU's fixed scratch array of the allocation-free entry point (device address)
    void *noalloc_scratch = nullptr;
};

namespace {

// Largest path or record count that one call accepts (FRG_MAX_COUNT = 2^36): beyond any GPU's
// memory (2^36 roots are 23 TB), and small enough that no byte count computed from it overflows.
bool bad_count(int64_t n) { return n < 0 || n > FRG_MAX_COUNT; }

#define HIPCHK(x)                              \
    do {                                       \
        hipError_t e_ = (x);                   \
        if (e_ != hipSuccess) return FRG_EHIP; \
    } while (0)

// Makes the context's device current and waits until no earlier work of the context (on its own
// stream or on a caller's stream) can still read or write the context's buffers.
int begin(frg_ctx *c) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventSynchronize(c->ev_busy));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FRG_OK;
}

#define BEGIN(c)                       \
    do {                               \
        const int rc_ = begin(c);      \
        if (rc_ != FRG_OK) return rc_; \
    } while (0)

frg::TrackOpts to_track_opts(const frg_opts &o) {
    frg::TrackOpts t{};
    t.sinit = o.sinit;
    t.smax = o.smax;
    t.smin = o.smin;
    t.rk_tol = o.rk_tol;
    t.caution_steps = o.caution_steps;
    t.newt_its = o.newt_its;
    t.newt_tol = o.newt_tol;
    t.newt_attempts = o.newt_attempts;
    t.sharp_its = o.sharp_its;
    t.sharp_tol = o.sharp_tol;
    t.sharp_attempts = o.sharp_attempts;
    t.polish_its = o.polish_its;
    t.sharp_trigger = o.sharp_trigger;
    t.nudge_trigger = o.nudge_trigger;
    t.max_steps = o.max_steps;
    t.max_ministeps = o.max_ministeps;
    t.patch_lo = o.patch_lo;
    t.patch_hi = o.patch_hi;
    t.start_lo = o.start_lo;
    t.start_hi = o.start_hi;
    t.start_mode = o.start_mode;
    t.precision = o.precision;
    t.fp64_band = o.fp64_band;
    t.ml_safety = o.ml_safety;
    t.ml_growth = o.ml_growth;
    return t;
}

void drop_graph(frg_ctx *c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_valid = false;
}

int check_opts(const frg_opts &o) {
    if (o.newt_residual != 1) return FRG_EUNSUPPORTED;
    if (o.use_graph != 0 && o.use_graph != 1) return FRG_EINVAL;
    if (o.start_mode != FRG_START_UNIFORM && o.start_mode != FRG_START_GAUSSIAN) return FRG_EINVAL;
    if (o.precision != FRG_PREC_FP64 && o.precision != FRG_PREC_MIXED) return FRG_EINVAL;
    if (o.schedule != FRG_SCHED_STATIC && o.schedule != FRG_SCHED_PERSISTENT) return FRG_EINVAL;
    if (o.step_control != FRG_STEP_CLASSIC && o.step_control != FRG_STEP_LEARNED) return FRG_EINVAL;
    if (o.block_size != 0 && (o.block_size < 64 || o.block_size > 256 || o.block_size % 64 != 0)) return FRG_EINVAL;
    if (o.groups_per_cu < 0 || o.max_steps < 0 || o.max_ministeps < 0) return FRG_EINVAL;
    if (o.polish_its < 0 || o.polish_its > 8) return FRG_EINVAL;
    if (o.lanes_per_path != 0 && o.lanes_per_path != 8 && o.lanes_per_path != 16) return FRG_EINVAL;
    if (o.lanes_per_path == 16 && o.precision != FRG_PREC_FP64) return FRG_EUNSUPPORTED;  // FP64 solve only
    if (o.waves_per_simd < 0 || o.waves_per_simd > 2) return FRG_EINVAL;
    if (o.waves_per_simd == 2 && o.lanes_per_path != 16) return FRG_EUNSUPPORTED;  // 16-lane kernel only
    if (!(o.smin > 0.0) || !(o.rk_tol > 0.0)) return FRG_EINVAL;
    if (!(o.ml_safety > 0.0) || !(o.ml_growth >= 1.0) || !(o.ml_safety < 1e6) || !(o.ml_growth < 1e6)) return FRG_EINVAL;
    return FRG_OK;
}

// Default launch shapes: 64-thread blocks (8 paths per block); persistent grids fill every CU
// to the occupancy the kernel's register use allows.
int pick_block(const frg_ctx *c) { return c->opts.block_size ? c->opts.block_size : 64; }
int pick_lanes(const frg_ctx *c) { return c->opts.lanes_per_path ? c->opts.lanes_per_path : 8; }
int pick_waves(const frg_ctx *c) { return c->opts.waves_per_simd ? c->opts.waves_per_simd : 1; }

// allow_graph = false launches eagerly whatever opts.use_graph says (no graph is created); a job
// that brings its own work-queue counter (job.queue) uses it instead of the context's.
int run_job(frg_ctx *c, frg::Job &job, hipStream_t stream, frg_stats *summary, bool allow_graph = true) {
    job.o = to_track_opts(c->opts);
    job.learned = c->opts.step_control == FRG_STEP_LEARNED;
    if (job.learned && !c->have_model) return FRG_EINVAL;
    job.model = c->model;
    const int block = pick_block(c);
    const int lanes = pick_lanes(c);
    const int waves = pick_waves(c);
    const int gpb = block / lanes;
    int grid;
    int kind = c->opts.schedule == FRG_SCHED_PERSISTENT ? frg::LK_PERSISTENT : frg::LK_STATIC;
    if (kind == frg::LK_PERSISTENT) {
        int per_cu = c->opts.groups_per_cu;
        if (per_cu == 0) per_cu = frg::track_max_groups_per_cu(block, kind, c->opts.precision, lanes, waves);
        if (per_cu < gpb) per_cu = gpb;
        grid = std::max(1, (per_cu / gpb) * c->num_cu);
        const int64_t need = (job.n_total + gpb - 1) / gpb;
        if ((int64_t)grid > need) grid = (int)need;
        if (!job.queue) {
            HIPCHK(c->queue.ensure(sizeof(unsigned long long)));
            job.queue = c->queue.as<unsigned long long>();
        }
    } else {
        const int64_t g = (job.n_total + gpb - 1) / gpb;
        if (g > 0x7fffffff) return FRG_EINVAL;
        grid = (int)g;
    }
    if (job.n_total == 0) return FRG_OK;
    if (c->opts.use_graph && allow_graph) {
        // one graph per launch signature: the kernel arguments (the Job) and the launch shape
        frg_ctx::GraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.job = job;
        key.kind = kind;
        key.block = block;
        key.grid = grid;
        key.lanes = lanes;
        key.waves = waves;
        if (!c->graph_valid || std::memcmp(&key, &c->graph_key, sizeof(key)) != 0) {
            drop_graph(c);
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            if (kind == frg::LK_PERSISTENT) e = frg::launch_reset_counter(job.queue, stream);
            if (e == hipSuccess) e = frg::launch_track(job, kind, block, grid, lanes, waves, stream);
            if (e == hipSuccess) e = frg::launch_refine(job, stream);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return FRG_EHIP;
            }
            c->graph = g;
            HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            std::memcpy(&c->graph_key, &key, sizeof(key));
            c->graph_valid = true;
        }
        // a graph replays as one unit: ev_mid marks its end too, so the tracking time of a graph
        // launch includes the refinement
        HIPCHK(hipEventRecord(c->ev0, stream));
        HIPCHK(hipGraphLaunch(c->graph_exec, stream));
        HIPCHK(hipEventRecord(c->ev_mid, stream));
        HIPCHK(hipEventRecord(c->ev1, stream));
    } else {
        if (kind == frg::LK_PERSISTENT) HIPCHK(frg::launch_reset_counter(job.queue, stream));
        HIPCHK(hipEventRecord(c->ev0, stream));
        HIPCHK(frg::launch_track(job, kind, block, grid, lanes, waves, stream));
        HIPCHK(hipEventRecord(c->ev_mid, stream));
        HIPCHK(frg::launch_refine(job, stream));
        HIPCHK(hipEventRecord(c->ev1, stream));
    }
    if (summary) {
        summary->lanes_per_path = lanes;
        summary->waves_per_simd = waves;
        summary->block_size = block;
        summary->grid_size = grid;
        summary->schedule = c->opts.schedule;
        summary->precision = c->opts.precision;
    }
    return FRG_OK;
}

void fill_summary_counts(frg_stats *s, int64_t n, const int32_t *err, const int32_t *st) {
    s->n_paths = n;
    for (int k = 0; k < 8; ++k) s->code_count[k] = 0;
    s->steps = s->rk_attempts = s->newton_calls = s->newton_its = s->tangent_solves = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int e = err[i];
        if (e >= 0 && e < 8) s->code_count[e]++;
        if (st) {
            const int32_t *r = st + i * 8;
            s->steps += r[0];
            s->rk_attempts += r[1];
            s->newton_calls += r[2];
            s->newton_its += r[3];
            s->tangent_solves += r[6];
        }
    }
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Uploads the single-instance description.
int upload_instances(frg_ctx *c, int32_t n_inst, const frg_c64 *fpm, const frg_c64 *gamma, const uint64_t *seeds,
                     const uint64_t *pid0s, const int64_t *counts, frg::Job &job, hipStream_t stream) {
    std::vector<int64_t> first(n_inst + 1, 0);
    for (int i = 0; i < n_inst; ++i) {
        if (bad_count(counts[i]) || first[i] + counts[i] > FRG_MAX_COUNT) return FRG_EINVAL;
        first[i + 1] = first[i] + counts[i];
    }
    HIPCHK(c->inst_first.ensure(sizeof(int64_t) * (n_inst + 1)));
    HIPCHK(c->fpm.ensure(sizeof(frg_c64) * 30 * n_inst));
    HIPCHK(c->gamma.ensure(sizeof(frg_c64) * n_inst));
    HIPCHK(c->seed.ensure(sizeof(uint64_t) * n_inst));
    HIPCHK(c->pid0.ensure(sizeof(uint64_t) * n_inst));
    HIPCHK(hipMemcpyAsync(c->inst_first.p, first.data(), sizeof(int64_t) * (n_inst + 1), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->fpm.p, fpm, sizeof(frg_c64) * 30 * n_inst, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->gamma.p, gamma, sizeof(frg_c64) * n_inst, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->seed.p, seeds, sizeof(uint64_t) * n_inst, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->pid0.p, pid0s, sizeof(uint64_t) * n_inst, hipMemcpyHostToDevice, stream));
    // the host vector dies at return: make sure the copies have consumed it
    HIPCHK(hipStreamSynchronize(stream));
    job.n_total = first[n_inst];
    job.n_inst = n_inst;
    job.inst_first = c->inst_first.as<int64_t>();
    job.fpm = c->fpm.as<cx<double>>();
    job.gamma = c->gamma.as<cx<double>>();
    job.seed = c->seed.as<uint64_t>();
    job.pid0 = c->pid0.as<uint64_t>();
    return FRG_OK;
}

// Runs a job whose outputs live in the context's buffers and copies them to the host.
int run_host(frg_ctx *c, frg::Job &job, frg_c64 *fpt, int32_t *err, int32_t *path_stats, frg_c64 *yfinal,
             frg_stats *summary, std::chrono::steady_clock::time_point t0) {
    const int64_t n = job.n_total;
    const bool want_stats = path_stats || summary;
    HIPCHK(c->fpt.ensure(sizeof(frg_c64) * 21 * std::max<int64_t>(n, 1)));
    HIPCHK(c->err.ensure(sizeof(int32_t) * std::max<int64_t>(n, 1)));
    job.fpt = c->fpt.as<cx<double>>();
    job.err = c->err.as<int>();
    job.stats = nullptr;
    job.yfinal = nullptr;
    if (want_stats) {
        HIPCHK(c->stats.ensure(sizeof(int32_t) * 8 * std::max<int64_t>(n, 1)));
        job.stats = c->stats.as<int>();
    }
    if (yfinal) {
        HIPCHK(c->yfinal.ensure(sizeof(frg_c64) * 23 * std::max<int64_t>(n, 1)));
        job.yfinal = c->yfinal.as<cx<double>>();
    }
    int rc = run_job(c, job, c->stream, summary);
    if (rc != FRG_OK) return rc;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(fpt, c->fpt.p, sizeof(frg_c64) * 21 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(err, c->err.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
        if (yfinal)
            HIPCHK(hipMemcpyAsync(yfinal, c->yfinal.p, sizeof(frg_c64) * 23 * n, hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<int32_t> tmp;
    int32_t *st = path_stats;
    if (want_stats && n > 0) {
        if (!st) {
            tmp.resize((size_t)n * 8);
            st = tmp.data();
        }
        HIPCHK(hipMemcpyAsync(st, c->stats.p, sizeof(int32_t) * 8 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (summary) {
        fill_summary_counts(summary, n, err, st);
        float ms = 0.f, ms_ref = 0.f;
        if (n > 0) {
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev_mid));
            HIPCHK(hipEventElapsedTime(&ms_ref, c->ev_mid, c->ev1));
        }
        summary->kernel_ms = ms;
        summary->refine_ms = ms_ref;
        summary->total_ms = ms_since(t0);
    }
    return FRG_OK;
}

}  // namespace

extern "C" {

int frg_version(void) { return 10000; }

const char *frg_strerror(int code) {
    switch (code) {
        case FRG_OK: return "ok";
        case FRG_EINVAL: return "invalid argument";
        case FRG_EHIP: return "HIP runtime error";
        case FRG_ENOMEM: return "out of memory";
        case FRG_EUNSUPPORTED: return "unsupported option";
        case FRG_ELIMIT: return "size beyond the fixed limit of the entry point";
        default: return "unknown error";
    }
}

int frg_default_opts(frg_opts *o) {
    if (!o) return FRG_EINVAL;
    std::memset(o, 0, sizeof(*o));
    o->sinit = 1.;
    o->smax = 100.;
    o->smin = 0.000001;
    o->rk_tol = 0.001;
    o->caution_steps = 2;
    o->newt_its = 10;
    o->newt_tol = 0.00000001;
    o->newt_attempts = 3;
    o->sharp_its = 30;
    o->sharp_tol = 0.000000000001;
    o->sharp_attempts = 3;
    o->newt_residual = 1;
    o->sharp_trigger = 0.00001;
    o->nudge_trigger = 0.00001;
    o->max_steps = 2000;
    o->max_ministeps = 3000;
    o->patch_lo = -2.;
    o->patch_hi = 2.;
    o->start_lo = -5.;
    o->start_hi = 5.;
    o->start_mode = FRG_START_UNIFORM;
    o->precision = FRG_PREC_FP64;
    o->fp64_band = 1e-2;
    o->ml_safety = 0.6;
    o->ml_growth = 5.0;
    o->schedule = FRG_SCHED_STATIC;
    o->block_size = 0;
    o->groups_per_cu = 0;
    o->step_control = FRG_STEP_CLASSIC;
    o->use_graph = 0;
    o->lanes_per_path = 0;
    o->waves_per_simd = 0;
    o->polish_its = 2;
    return FRG_OK;
}

int frg_create(frg_ctx **out, int device, const frg_opts *opts) {
    if (!out) return FRG_EINVAL;
    *out = nullptr;
    frg_opts o;
    if (opts) o = *opts;
    else frg_default_opts(&o);
    int rc = check_opts(o);
    if (rc != FRG_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FRG_EINVAL;
    frg_ctx *c = new (std::nothrow) frg_ctx();
    if (!c) return FRG_ENOMEM;
    c->device = device;
    c->opts = o;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev_mid) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_busy, hipEventDisableTiming) != hipSuccess) {
        frg_destroy(c);
        return FRG_EHIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        frg_destroy(c);
        return FRG_EHIP;
    }
    c->num_cu = prop.multiProcessorCount;
    // the allocation-free entry point's scratch array: its address on this device, looked up once
    // with the device current
    if (frg::noalloc_scratch_address(&c->noalloc_scratch) != hipSuccess || !c->noalloc_scratch) {
        frg_destroy(c);
        return FRG_EHIP;
    }
    *out = c;
    return FRG_OK;
}

void frg_destroy(frg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->ev_busy) (void)hipEventSynchronize(c->ev_busy);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    DevBuf *bufs[] = {&c->inst_first, &c->fpm, &c->gamma, &c->seed, &c->pid0, &c->spt, &c->spm, &c->fpt, &c->err,
                      &c->stats, &c->yfinal, &c->queue, &c->dbg_a, &c->dbg_b, &c->dbg_c, &c->dbg_d, &c->dbg_e,
                      &c->dbg_x, &c->model16, &c->steplog, &c->steplog_count, &c->tw, &c->tm, &c->tv, &c->tgrad,
                      &c->tsamples, &c->tnorm, &c->tloss, &c->tidx, &c->tw16, &c->dedupe_tmp};
    for (DevBuf *b : bufs) b->release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev_mid) (void)hipEventDestroy(c->ev_mid);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_busy) (void)hipEventDestroy(c->ev_busy);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int frg_set_opts(frg_ctx *c, const frg_opts *o) {
    if (!c || !o) return FRG_EINVAL;
    int rc = check_opts(*o);
    if (rc != FRG_OK) return rc;
    BEGIN(c);
    c->opts = *o;
    drop_graph(c);
    return FRG_OK;
}

int frg_get_opts(const frg_ctx *c, frg_opts *o) {
    if (!c || !o) return FRG_EINVAL;
    *o = c->opts;
    return FRG_OK;
}

int frg_track(frg_ctx *c, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id, int64_t n_paths,
              frg_c64 *fpt, int32_t *err, int32_t *path_stats, frg_stats *summary) try {
    if (!c || !fpm || bad_count(n_paths) || (n_paths > 0 && (!fpt || !err))) return FRG_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    BEGIN(c);
    frg::Job job{};
    int rc = upload_instances(c, 1, fpm, &gamma, &seed, &first_path_id, &n_paths, job, c->stream);
    if (rc != FRG_OK) return rc;
    return run_host(c, job, fpt, err, path_stats, nullptr, summary, t0);
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_track_device(frg_ctx *c, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                     int64_t n_paths, frg_c64 *d_fpt, int32_t *d_err, int32_t *d_path_stats, void *stream) try {
    if (!c || !fpm || bad_count(n_paths) || (n_paths > 0 && (!d_fpt || !d_err))) return FRG_EINVAL;
    BEGIN(c);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    frg::Job job{};
    int rc = upload_instances(c, 1, fpm, &gamma, &seed, &first_path_id, &n_paths, job, s);
    if (rc != FRG_OK) return rc;
    job.fpt = reinterpret_cast<cx<double> *>(d_fpt);
    job.err = d_err;
    job.stats = d_path_stats;
    job.yfinal = nullptr;
    rc = run_job(c, job, s, nullptr);
    if (rc != FRG_OK) return rc;
    HIPCHK(hipEventRecord(c->ev_busy, s));
    return FRG_OK;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_track_explicit(frg_ctx *c, const frg_c64 *fpm, frg_c64 gamma, const frg_c64 *spt, const frg_c64 *spm,
                       uint64_t seed, uint64_t first_path_id, int64_t n_paths, frg_c64 *fpt, int32_t *err,
                       int32_t *path_stats, frg_c64 *yfinal, frg_stats *summary) try {
    if (!c || !fpm || bad_count(n_paths) || (n_paths > 0 && (!spt || !spm || !fpt || !err))) return FRG_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    BEGIN(c);
    frg::Job job{};
    int rc = upload_instances(c, 1, fpm, &gamma, &seed, &first_path_id, &n_paths, job, c->stream);
    if (rc != FRG_OK) return rc;
    if (n_paths > 0) {
        HIPCHK(c->spt.ensure(sizeof(frg_c64) * 21 * n_paths));
        HIPCHK(c->spm.ensure(sizeof(frg_c64) * 30 * n_paths));
        HIPCHK(hipMemcpyAsync(c->spt.p, spt, sizeof(frg_c64) * 21 * n_paths, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->spm.p, spm, sizeof(frg_c64) * 30 * n_paths, hipMemcpyHostToDevice, c->stream));
        job.spt = c->spt.as<cx<double>>();
        job.spm = c->spm.as<cx<double>>();
    }
    return run_host(c, job, fpt, err, path_stats, yfinal, summary, t0);
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_track_batch(frg_ctx *c, int32_t n_inst, const frg_c64 *fpm, const frg_c64 *gamma, const uint64_t *seeds,
                    const uint64_t *first_path_ids, const int64_t *paths_per_inst, frg_c64 *fpt, int32_t *err,
                    int32_t *path_stats, frg_stats *summary) try {
    if (!c || n_inst <= 0 || !fpm || !gamma || !seeds || !first_path_ids || !paths_per_inst) return FRG_EINVAL;
    auto t0 = std::chrono::steady_clock::now();
    BEGIN(c);
    frg::Job job{};
    int rc = upload_instances(c, n_inst, fpm, gamma, seeds, first_path_ids, paths_per_inst, job, c->stream);
    if (rc != FRG_OK) return rc;
    if (job.n_total > 0 && (!fpt || !err)) return FRG_EINVAL;
    return run_host(c, job, fpt, err, path_stats, nullptr, summary, t0);
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_track_batch_noalloc(frg_ctx *c, int32_t n_inst, const frg_c64 *d_fpm, const frg_c64 *gamma,
                            const uint64_t *seeds, const uint64_t *first_path_ids, const int64_t *paths_per_inst,
                            frg_c64 *d_fpt, int32_t *d_err, int32_t *d_path_stats, void *stream) {
    if (!c || n_inst < 1 || !d_fpm || !gamma || !seeds || !first_path_ids || !paths_per_inst) return FRG_EINVAL;
    if (n_inst > FRG_NOALLOC_MAX_INSTANCES) return FRG_ELIMIT;
    if (!c->noalloc_scratch) return FRG_EHIP;
    frg::NoAllocStage st;  // kernel arguments of the staging kernel (host stack, no allocation)
    st.n_inst = n_inst;
    st.inst_first[0] = 0;
    for (int i = 0; i < n_inst; ++i) {
        if (bad_count(paths_per_inst[i]) || st.inst_first[i] + paths_per_inst[i] > FRG_MAX_COUNT) return FRG_EINVAL;
        st.inst_first[i + 1] = st.inst_first[i] + paths_per_inst[i];
        st.gamma[i] = frg::mk(gamma[i].re, gamma[i].im);
        st.seed[i] = seeds[i];
        st.pid0[i] = first_path_ids[i];
    }
    const int64_t total = st.inst_first[n_inst];
    if (total > 0 && (!d_fpt || !d_err)) return FRG_EINVAL;
    BEGIN(c);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const frg::NoAllocLayout L = frg::noalloc_layout(static_cast<unsigned char *>(c->noalloc_scratch));
    frg::Job job{};
    job.n_total = total;
    job.n_inst = n_inst;
    job.inst_first = L.inst_first;
    job.fpm = reinterpret_cast<const cx<double> *>(d_fpm);
    job.gamma = L.gamma;
    job.seed = L.seed;
    job.pid0 = L.pid0;
    job.fpt = reinterpret_cast<cx<double> *>(d_fpt);
    job.err = d_err;
    job.stats = d_path_stats;
    job.yfinal = nullptr;
    job.queue = L.queue;
    HIPCHK(frg::launch_stage_instances(st, c->noalloc_scratch, s));
    const int rc = run_job(c, job, s, nullptr, /*allow_graph=*/false);
    if (rc != FRG_OK) return rc;
    HIPCHK(hipEventRecord(c->ev_busy, s));
    return FRG_OK;
}

int frg_autotune(frg_ctx *c, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                 int64_t n_sample, frg_tune_entry *table, int32_t table_cap, int32_t *n_entries) try {
    if (!c || !fpm || n_sample <= 0 || bad_count(n_sample) || table_cap < 0 || (table_cap > 0 && !table))
        return FRG_EINVAL;
    BEGIN(c);
    const frg_opts saved = c->opts;
    frg::Job job{};
    int rc = upload_instances(c, 1, fpm, &gamma, &seed, &first_path_id, &n_sample, job, c->stream);
    if (rc != FRG_OK) return rc;
    HIPCHK(c->fpt.ensure(sizeof(frg_c64) * 21 * n_sample));
    HIPCHK(c->err.ensure(sizeof(int32_t) * n_sample));
    job.fpt = c->fpt.as<cx<double>>();
    job.err = c->err.as<int>();
    job.stats = nullptr;
    job.yfinal = nullptr;
    struct Cand {
        int schedule, block, gpc, lanes, waves;
    };
    // lanes per path: 8 (8 paths per wave) for every precision, 16 (4 paths per wave) for FP64
    // with a register budget of 1 or 2 resident waves per SIMD
    std::vector<Cand> cands;
    const int variants[3][2] = {{8, 1}, {16, 1}, {16, 2}};
    for (const auto &v : variants) {
        const int lanes = v[0], waves = v[1];
        if (lanes == 16 && saved.precision != FRG_PREC_FP64) continue;
        for (int block : {64, 128, 256}) cands.push_back({FRG_SCHED_STATIC, block, 0, lanes, waves});
        for (int block : {64, 128, 256}) {
            const int gpb = block / lanes;
            const int max_gpc =
                frg::track_max_groups_per_cu(block, frg::LK_PERSISTENT, saved.precision, lanes, waves);
            for (int bpc = std::max(1, max_gpc / gpb); bpc >= 1; --bpc)
                cands.push_back({FRG_SCHED_PERSISTENT, block, bpc * gpb, lanes, waves});
        }
    }
    int best = -1;
    double best_ms = 0.0;
    int32_t k = 0;
    c->opts.use_graph = 0;
    for (size_t i = 0; i < cands.size(); ++i) {
        c->opts.schedule = cands[i].schedule;
        c->opts.block_size = cands[i].block;
        c->opts.groups_per_cu = cands[i].gpc;
        c->opts.lanes_per_path = cands[i].lanes;
        c->opts.waves_per_simd = cands[i].waves;
        rc = run_job(c, job, c->stream, nullptr);
        if (rc != FRG_OK) {
            c->opts = saved;
            return rc;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev_mid));
        if (k < table_cap) {
            table[k].schedule = cands[i].schedule;
            table[k].block_size = cands[i].block;
            table[k].groups_per_cu = cands[i].gpc;
            table[k].lanes_per_path = cands[i].lanes;
            table[k].waves_per_simd = cands[i].waves;
            table[k].pad0 = 0;
            table[k].kernel_ms = ms;
            table[k].paths_per_s = ms > 0.f ? n_sample / (ms * 1e-3) : 0.0;
            ++k;
        }
        if (best < 0 || ms < best_ms) {
            best = (int)i;
            best_ms = ms;
        }
    }
    c->opts = saved;
    c->opts.schedule = cands[best].schedule;
    c->opts.block_size = cands[best].block;
    c->opts.groups_per_cu = cands[best].gpc;
    c->opts.lanes_per_path = cands[best].lanes;
    c->opts.waves_per_simd = cands[best].waves;
    drop_graph(c);
    if (n_entries) *n_entries = k;
    return FRG_OK;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_sort_dedupe_merge_device(frg_ctx *c, const frg_c64 *d_coll, int64_t m, const frg_c64 *d_fpt,
                                 int32_t *d_err, int64_t n, double tol, frg_c64 *d_coll_out, int64_t *m_out,
                                 void *stream) {
    if (!c || bad_count(m) || bad_count(n) || (m > 0 && !d_coll) || (n > 0 && (!d_fpt || !d_err)) || !d_coll_out || !m_out ||
        !(tol >= 0.0) || m + n >= ((int64_t)1 << 31))
        return FRG_EINVAL;
    BEGIN(c);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const size_t need = frg::dedupe_temp_bytes(m, n);
    HIPCHK(c->dedupe_tmp.ensure(need));
    HIPCHK(frg::sort_dedupe_merge_device(reinterpret_cast<const double2 *>(d_coll), m,
                                         reinterpret_cast<const double2 *>(d_fpt), d_err, n, tol,
                                         reinterpret_cast<double2 *>(d_coll_out), m_out, c->dedupe_tmp.p,
                                         c->dedupe_tmp.bytes, s));
    HIPCHK(hipEventRecord(c->ev_busy, s));
    return FRG_OK;
}

int frg_last_kernel_ms(frg_ctx *c, double *ms) {
    if (!c || !ms) return FRG_EINVAL;
    HIPCHK(hipSetDevice(c->device));  // no BEGIN: reads only the events
    HIPCHK(hipEventSynchronize(c->ev1));
    float v = 0.f;
    HIPCHK(hipEventElapsedTime(&v, c->ev0, c->ev_mid));
    *ms = v;
    return FRG_OK;
}

int frg_last_refine_ms(frg_ctx *c, double *ms) {
    if (!c || !ms) return FRG_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventSynchronize(c->ev1));
    float v = 0.f;
    HIPCHK(hipEventElapsedTime(&v, c->ev_mid, c->ev1));
    *ms = v;
    return FRG_OK;
}

int frg_generate_starts(frg_ctx *c, uint64_t seed, uint64_t first_path_id, int64_t n, frg_c64 *spt, frg_c64 *spm) {
    if (!c || bad_count(n) || (n > 0 && (!spt || !spm))) return FRG_EINVAL;
    if (n == 0) return FRG_OK;
    BEGIN(c);
    HIPCHK(c->spt.ensure(sizeof(frg_c64) * 21 * n));
    HIPCHK(c->spm.ensure(sizeof(frg_c64) * 30 * n));
    frg::DebugStartArgs a{};
    a.n = n;
    a.seed = seed;
    a.pid0 = first_path_id;
    a.o = to_track_opts(c->opts);
    a.spt = c->spt.as<cx<double>>();
    a.spm = c->spm.as<cx<double>>();
    HIPCHK(frg::launch_debug_start(a, c->stream));
    HIPCHK(hipMemcpyAsync(spt, c->spt.p, sizeof(frg_c64) * 21 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(spm, c->spm.p, sizeof(frg_c64) * 30 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FRG_OK;
}

int frg_solve_step(frg_ctx *c, int32_t mode, int32_t precision, int64_t n, const frg_c64 *Y, const frg_c64 *u,
                   const frg_c64 *w, const frg_c64 *spm, const frg_c64 *fpm, frg_c64 gamma, const double *tgoal,
                   frg_c64 *x) {
    if (!c || bad_count(n) || mode < 0 || mode > 2) return FRG_EINVAL;
    if (precision != FRG_PREC_FP64 && precision != FRG_PREC_MIXED) return FRG_EINVAL;
    if (pick_lanes(c) == 16 && precision != FRG_PREC_FP64) return FRG_EUNSUPPORTED;
    if (n == 0) return FRG_OK;
    if (!Y || !u || !spm || !fpm || !x || (mode < 2 && !w) || (mode == 2 && !tgoal)) return FRG_EINVAL;
    BEGIN(c);
    const size_t v23 = sizeof(frg_c64) * 23 * n;
    HIPCHK(c->dbg_a.ensure(v23));
    HIPCHK(c->dbg_b.ensure(v23));
    HIPCHK(c->dbg_c.ensure(v23));
    HIPCHK(c->dbg_d.ensure(sizeof(frg_c64) * 30 * n + sizeof(frg_c64) * 30));
    HIPCHK(c->dbg_e.ensure(sizeof(double) * n));
    HIPCHK(c->dbg_x.ensure(v23));
    HIPCHK(hipMemcpyAsync(c->dbg_a.p, Y, v23, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dbg_b.p, u, v23, hipMemcpyHostToDevice, c->stream));
    if (w) HIPCHK(hipMemcpyAsync(c->dbg_c.p, w, v23, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemsetAsync(c->dbg_c.p, 0, v23, c->stream));
    HIPCHK(hipMemcpyAsync(c->dbg_d.p, spm, sizeof(frg_c64) * 30 * n, hipMemcpyHostToDevice, c->stream));
    frg_c64 *dfpm = c->dbg_d.as<frg_c64>() + 30 * n;
    HIPCHK(hipMemcpyAsync(dfpm, fpm, sizeof(frg_c64) * 30, hipMemcpyHostToDevice, c->stream));
    if (tgoal) HIPCHK(hipMemcpyAsync(c->dbg_e.p, tgoal, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemsetAsync(c->dbg_e.p, 0, sizeof(double) * n, c->stream));
    frg::DebugSolveArgs a{};
    a.n = n;
    a.mode = mode;
    a.Y = c->dbg_a.as<cx<double>>();
    a.u = c->dbg_b.as<cx<double>>();
    a.w = c->dbg_c.as<cx<double>>();
    a.spm = c->dbg_d.as<cx<double>>();
    a.fpm = reinterpret_cast<cx<double> *>(dfpm);
    a.tgoal = c->dbg_e.as<double>();
    a.gamma = frg::mk(gamma.re, gamma.im);
    a.x = c->dbg_x.as<cx<double>>();
    HIPCHK(frg::launch_debug_solve(a, precision, pick_lanes(c), c->stream));
    HIPCHK(hipMemcpyAsync(x, c->dbg_x.p, v23, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FRG_OK;
}

int frg_step_factor(frg_ctx *c, int64_t n, const double *err, double *scale) {
    if (!c || bad_count(n) || (n > 0 && (!err || !scale))) return FRG_EINVAL;
    if (n == 0) return FRG_OK;
    BEGIN(c);
    HIPCHK(c->dbg_e.ensure(sizeof(double) * n));
    HIPCHK(c->dbg_x.ensure(sizeof(double) * n));
    HIPCHK(hipMemcpyAsync(c->dbg_e.p, err, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(frg::launch_debug_step_factor(c->dbg_e.as<double>(), n, c->opts.rk_tol, c->dbg_x.as<double>(), c->stream));
    HIPCHK(hipMemcpyAsync(scale, c->dbg_x.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FRG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Learned step control
// ------------------------------------------------------------------------------------------
namespace {

void fill_model_norm(frg::StepModel &m, const float *model) {
    for (int i = 0; i < frg::kMlpIn; ++i) {
        m.xmu[i] = model[frg::kMlpXmu + i];
        m.xsd[i] = model[frg::kMlpXsd + i];
    }
    m.ymu = model[frg::kMlpYmu];
    m.ysd = model[frg::kMlpYsd];
}

// mean and deviation of every feature and of the target (deviation floored at 1e-3)
void record_stats(const float *rec, int64_t n, float *norm) {
    for (int q = 0; q <= frg::kMlpIn; ++q) {
        double s = 0, s2 = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double v = rec[i * frg::kStepRecord + q];
            s += v;
            s2 += v * v;
        }
        const double mu = s / (double)n;
        const double sd = std::sqrt(std::max(s2 / (double)n - mu * mu, 1e-6));
        const int mi = q < frg::kMlpIn ? q : 2 * frg::kMlpIn;
        const int si = q < frg::kMlpIn ? frg::kMlpIn + q : 2 * frg::kMlpIn + 1;
        norm[mi] = (float)mu;
        norm[si] = (float)std::max(sd, 1e-3);
    }
}

}  // namespace

extern "C" {

int frg_collect_step_log(frg_ctx *c, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                         int64_t n_paths, float *records, int64_t cap, int64_t *n_records) try {
    if (!c || !fpm || bad_count(n_paths) || cap < 0 || cap > INT32_MAX || (cap > 0 && !records) || !n_records)
        return FRG_EINVAL;
    BEGIN(c);
    frg::Job job{};
    int rc = upload_instances(c, 1, fpm, &gamma, &seed, &first_path_id, &n_paths, job, c->stream);
    if (rc != FRG_OK) return rc;
    HIPCHK(c->fpt.ensure(sizeof(frg_c64) * 21 * std::max<int64_t>(n_paths, 1)));
    HIPCHK(c->err.ensure(sizeof(int32_t) * std::max<int64_t>(n_paths, 1)));
    HIPCHK(c->steplog.ensure(sizeof(float) * frg::kStepRecord * std::max<int64_t>(cap, 1)));
    HIPCHK(c->steplog_count.ensure(sizeof(unsigned long long)));
    HIPCHK(frg::launch_reset_counter(c->steplog_count.as<unsigned long long>(), c->stream));
    job.fpt = c->fpt.as<cx<double>>();
    job.err = c->err.as<int>();
    job.steplog = c->steplog.as<float>();
    job.steplog_count = c->steplog_count.as<unsigned long long>();
    job.steplog_cap = cap;
    rc = run_job(c, job, c->stream, nullptr);
    if (rc != FRG_OK) return rc;
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, c->steplog_count.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t got = std::min<int64_t>((int64_t)cnt, cap);
    if (got > 0)
        HIPCHK(hipMemcpy(records, c->steplog.p, sizeof(float) * frg::kStepRecord * got, hipMemcpyDeviceToHost));
    // keep the proposals whose outcome was observed (a path's last proposal and proposals
    // followed by a landing step have none)
    int64_t kept = 0;
    for (int64_t i = 0; i < got; ++i) {
        const float *r = records + i * frg::kStepRecord;
        if (!std::isfinite(r[frg::kStepRecord - 1])) continue;
        if (kept != i) std::memmove(records + kept * frg::kStepRecord, r, sizeof(float) * frg::kStepRecord);
        ++kept;
    }
    *n_records = kept;
    return FRG_OK;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_set_step_model(frg_ctx *c, const float *model) {
    if (!c || !model) return FRG_EINVAL;
    for (int i = 0; i < FRG_STEP_MODEL_SIZE; ++i)
        if (!std::isfinite(model[i])) return FRG_EINVAL;
    for (int i = 0; i < frg::kMlpIn; ++i)
        if (!(model[frg::kMlpXsd + i] > 0.f)) return FRG_EINVAL;
    if (!(model[frg::kMlpYsd] > 0.f)) return FRG_EINVAL;
    BEGIN(c);
    HIPCHK(c->tw.ensure(sizeof(float) * frg::kMlpNW));
    HIPCHK(c->model16.ensure(sizeof(__half) * frg::kMlpNW));
    HIPCHK(hipMemcpyAsync(c->tw.p, model, sizeof(float) * frg::kMlpNW, hipMemcpyHostToDevice, c->stream));
    HIPCHK(frg::launch_mlp_to_half(c->tw.as<float>(), c->model16.as<__half>(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    fill_model_norm(c->model, model);
    c->model.w = c->model16.as<__half>();
    c->have_model = true;
    return FRG_OK;
}

int frg_step_model_eval(frg_ctx *c, const float *features, int64_t n, float *out) {
    if (!c || bad_count(n) || (n > 0 && (!features || !out)) || !c->have_model) return FRG_EINVAL;
    if (n == 0) return FRG_OK;
    BEGIN(c);
    HIPCHK(c->tsamples.ensure(sizeof(float) * frg::kMlpIn * n));
    HIPCHK(c->tloss.ensure(sizeof(float) * n));
    HIPCHK(hipMemcpyAsync(c->tsamples.p, features, sizeof(float) * frg::kMlpIn * n, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(frg::launch_mlp_eval(c->model, c->tsamples.as<float>(), n, c->tloss.as<float>(), c->stream));
    HIPCHK(hipMemcpyAsync(out, c->tloss.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FRG_OK;
}

int frg_step_model_grad(frg_ctx *c, const float *model, const float *records, int64_t n, float *grad,
                        float *loss) try {
    if (!c || !model || !records || n <= 0 || bad_count(n) || !grad) return FRG_EINVAL;
    BEGIN(c);
    HIPCHK(c->tw.ensure(sizeof(float) * frg::kMlpNW));
    HIPCHK(c->tgrad.ensure(sizeof(float) * frg::kMlpNW));
    HIPCHK(c->tsamples.ensure(sizeof(float) * frg::kStepRecord * n));
    HIPCHK(c->tnorm.ensure(sizeof(float) * 18));
    HIPCHK(c->tloss.ensure(sizeof(float)));
    HIPCHK(c->tidx.ensure(sizeof(int64_t) * n));
    std::vector<int64_t> idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    HIPCHK(hipMemcpyAsync(c->tw.p, model, sizeof(float) * frg::kMlpNW, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tnorm.p, model + frg::kMlpXmu, sizeof(float) * 18, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tsamples.p, records, sizeof(float) * frg::kStepRecord * n, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(c->tidx.p, idx.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->tgrad.p, 0, sizeof(float) * frg::kMlpNW, c->stream));
    HIPCHK(hipMemsetAsync(c->tloss.p, 0, sizeof(float), c->stream));
    frg::MlpGradArgs a{};
    a.w = c->tw.as<float>();
    a.samples = c->tsamples.as<float>();
    a.norm = c->tnorm.as<float>();
    a.indices = c->tidx.as<int64_t>();
    a.n_samples = n;
    a.batch = n;
    a.grad = c->tgrad.as<float>();
    a.loss = c->tloss.as<float>();
    HIPCHK(frg::launch_mlp_grad(a, c->stream));
    HIPCHK(hipMemcpyAsync(grad, c->tgrad.p, sizeof(float) * frg::kMlpNW, hipMemcpyDeviceToHost, c->stream));
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, c->tloss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return FRG_OK;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

int frg_train_step_model(frg_ctx *c, const float *records, int64_t n, int32_t steps, int32_t batch, float lr,
                         uint64_t seed, float *model, float *loss) try {
    if (!c || !records || n <= 0 || bad_count(n) || steps < 0 || batch <= 0 || !(lr > 0.f) || !model)
        return FRG_EINVAL;
    BEGIN(c);
    // normalisation from the data; Glorot-uniform initial weights from a host generator
    std::vector<float> h((size_t)FRG_STEP_MODEL_SIZE, 0.f);
    record_stats(records, n, h.data() + frg::kMlpXmu);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto urand = [&]() {  // xorshift64*
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return (double)((st * 0x2545F4914F6CDD1Dull) >> 11) * 0x1.0p-53;
    };
    const float a1 = std::sqrt(6.f / (frg::kMlpIn + frg::kMlpH)), a2 = std::sqrt(6.f / (2 * frg::kMlpH)),
                a3 = std::sqrt(6.f / (frg::kMlpH + 1));
    for (int i = 0; i < frg::kMlpH * frg::kMlpIn; ++i) h[frg::kMlpW1 + i] = (float)((2 * urand() - 1) * a1);
    for (int i = 0; i < frg::kMlpH * frg::kMlpH; ++i) h[frg::kMlpW2 + i] = (float)((2 * urand() - 1) * a2);
    for (int i = 0; i < frg::kMlpH; ++i) h[frg::kMlpW3 + i] = (float)((2 * urand() - 1) * a3);
    const size_t wbytes = sizeof(float) * frg::kMlpNW;
    HIPCHK(c->tw.ensure(wbytes));
    HIPCHK(c->tm.ensure(wbytes));
    HIPCHK(c->tv.ensure(wbytes));
    HIPCHK(c->tgrad.ensure(wbytes));
    HIPCHK(c->tw16.ensure(sizeof(__half) * frg::kMlpNW));
    HIPCHK(c->tsamples.ensure(sizeof(float) * frg::kStepRecord * n));
    HIPCHK(c->tnorm.ensure(sizeof(float) * 18));
    HIPCHK(c->tloss.ensure(sizeof(float) * std::max(steps, 1)));
    HIPCHK(hipMemcpyAsync(c->tw.p, h.data(), wbytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->tm.p, 0, wbytes, c->stream));
    HIPCHK(hipMemsetAsync(c->tv.p, 0, wbytes, c->stream));
    HIPCHK(hipMemsetAsync(c->tgrad.p, 0, wbytes, c->stream));
    HIPCHK(hipMemsetAsync(c->tloss.p, 0, sizeof(float) * std::max(steps, 1), c->stream));
    HIPCHK(hipMemcpyAsync(c->tnorm.p, h.data() + frg::kMlpXmu, sizeof(float) * 18, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tsamples.p, records, sizeof(float) * frg::kStepRecord * n, hipMemcpyHostToDevice,
                          c->stream));
    const float b1 = 0.9f, b2 = 0.999f;
    float b1t = 1.f, b2t = 1.f;
    for (int k = 0; k < steps; ++k) {
        frg::MlpGradArgs g{};
        g.w = c->tw.as<float>();
        g.samples = c->tsamples.as<float>();
        g.norm = c->tnorm.as<float>();
        g.indices = nullptr;
        g.n_samples = n;
        g.batch = batch;
        g.seed = seed;
        g.step = (uint64_t)k;
        g.grad = c->tgrad.as<float>();
        g.loss = c->tloss.as<float>() + k;
        HIPCHK(frg::launch_mlp_grad(g, c->stream));
        b1t *= b1;
        b2t *= b2;
        frg::MlpAdamArgs ad{};
        ad.w = c->tw.as<float>();
        ad.m = c->tm.as<float>();
        ad.v = c->tv.as<float>();
        ad.grad = c->tgrad.as<float>();
        ad.w16 = c->tw16.as<__half>();
        ad.lr = lr;
        ad.b1 = b1;
        ad.b2 = b2;
        ad.eps = 1e-8f;
        ad.b1t = b1t;
        ad.b2t = b2t;
        HIPCHK(frg::launch_mlp_adam(ad, c->stream));
    }
    HIPCHK(hipMemcpyAsync(h.data(), c->tw.p, wbytes, hipMemcpyDeviceToHost, c->stream));
    if (loss && steps > 0)
        HIPCHK(hipMemcpyAsync(loss, c->tloss.p, sizeof(float) * steps, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memcpy(model, h.data(), sizeof(float) * FRG_STEP_MODEL_SIZE);
    return FRG_OK;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

}  // extern "C"
