// gf_batch.hip — batched sweep: many independent magnet contexts of uneven
// grid sizes stepped together.
//
// Each instance keeps its own gf_ctx (state, materials, code array).  The
// batch holds, in device memory, the stage arguments of every instance for
// the four stages, so one launch per stage and magnet kind covers all
// instances; a per-instance clock kernel advances t and writes each stage's
// sinusoidal drive into those arguments.  One RK4 step (clock + 4 or 8 stage
// launches) is captured in a HIP graph and replayed.  The arithmetic per
// cell is the single-context stage kernels' (the same device functions), so
// a batch step equals gf_step of each context up to the device sin() of the
// drive against the host's (at most an ulp of the drive).
//
// The workgroup size and the grid shape of the batched launches are chosen
// by gf_batch_tune, which times stage 1 of every candidate on the live
// instances.  Stage 1 writes only the first stage store and the RK
// accumulator, which the next stage 1 overwrites before anything reads
// them, so tuning leaves the state as it was.
#include "gf_internal.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace gf;

struct gf_batch {
    std::vector<gf_ctx*> ctxs;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t order = nullptr;
    std::string err;

    StageArgs* d_fm = nullptr;       // [4*count], stage s of instance i at 4*i + s - 1
    AfmStageArgs* d_afm = nullptr;
    std::vector<StageArgs> h_fm;     // host tables of the last upload
    std::vector<AfmStageArgs> h_afm;
    BatchClock* d_clk = nullptr;
    std::vector<BatchClock> clk;

    bool any_fm = false, any_afm = false, fm_narrow = true, afm_narrow = true;
    int64_t fm_cells = 0, afm_cells = 0;
    uint32_t* d_fm_item = nullptr;   // workgroup slot -> instance
    uint32_t* d_fm_block0 = nullptr; // instance -> first slot
    uint32_t* d_afm_item = nullptr;
    uint32_t* d_afm_block0 = nullptr;
    uint32_t fm_total = 0, afm_total = 0;

    int threads = 256;               // workgroup size of the batched stage launches
    int cpt = 1;                     // cells per thread of a workgroup slot
    unsigned grid_cap = 0;           // 0: one workgroup per slot; else at most this many, striding
    hipGraphExec_t graph1 = nullptr; // one RK4 step
    hipGraphExec_t graphN = nullptr; // kGraphSteps RK4 steps
};

namespace {

constexpr int kGraphSteps = 16;
constexpr uint32_t kAllowedFlags = GF_FLAG_ORIENT_ROTATE | GF_FLAG_HEXT_NONUNIF;

thread_local std::string g_batch_create_error;

gf_status berr(gf_batch* b, gf_status code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

gf_status bhip(gf_batch* b, hipError_t e, const char* what) {
    if (e == hipSuccess) return GF_OK;
    return berr(b, GF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Why a context cannot join a batch, or "" when it can.
std::string unbatchable(const gf_ctx* c) {
    if (!(c->any_fm || c->any_afm)) return "has no magnetic material";
    if (c->any_fe || c->el_on || c->em_on) return "couples ferroelectric, elastic or electrodynamic physics";
    if (c->flags & ~kAllowedFlags)
        return "uses magnetostatics, the 1D model, prescribed drive tables or coupled physics "
               "(a batch steps magnets with fixed Hstat and sinusoidal drives)";
    return "";
}

bool at_step_boundary(const gf_ctx* c) {
    return (!c->any_fm || c->mag_next_stage == 1) && (!c->any_afm || c->afm_next_stage == 1);
}

template <class T>
gf_status dev_copy(gf_batch* b, T** dst, const std::vector<T>& src, const char* what) {
    if (*dst) { hipFree(*dst); *dst = nullptr; }
    if (src.empty()) return GF_OK;
    gf_status st = bhip(b, hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * src.size()), what);
    if (st) return st;
    return bhip(b, hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice), what);
}

// The stage arguments of every instance, from the contexts as they are now
// (fields uploaded after the batch was made, the static field's state).
// Stage s reads store {0,1,2,1}[s-1] and writes {1,2,1,0}[s-1], the
// sequence single-context stages follow from a step boundary; with
// `from_store`, stage 1 reads the store a context's stage 1 would read
// instead (a stage store uploaded or loaded at the boundary, selector 2).
void build_args(gf_batch* b, bool from_store) {
    const int in_sel[4] = {0, 1, 2, 1}, out_sel[4] = {1, 2, 1, 0};
    const size_t count = b->ctxs.size();
    b->h_fm.assign(4 * count, StageArgs{});
    b->h_afm.assign(4 * count, AfmStageArgs{});
    for (size_t i = 0; i < count; i++) {
        gf_ctx* c = b->ctxs[i];
        for (int s = 0; s < 4; s++) {
            const int fin = s == 0 && from_store ? c->mag_store_sel : in_sel[s];
            const int ain = s == 0 && from_store ? c->afm_store_sel : in_sel[s];
            if (c->any_fm) b->h_fm[4 * i + static_cast<size_t>(s)] = fm_stage_args(c, s + 1, fin, out_sel[s]);
            if (c->any_afm) b->h_afm[4 * i + static_cast<size_t>(s)] = afm_stage_args(c, s + 1, ain, out_sel[s]);
        }
    }
}

// Upload the host tables into the existing device tables (the graphs keep
// their addresses); the copies are ordered on the batch stream.
gf_status upload_args(gf_batch* b) {
    gf_status st = GF_OK;
    if (b->any_fm)
        st = bhip(b, hipMemcpyAsync(b->d_fm, b->h_fm.data(), sizeof(StageArgs) * b->h_fm.size(),
                                    hipMemcpyHostToDevice, b->stream), "batch stage arguments");
    if (!st && b->any_afm)
        st = bhip(b, hipMemcpyAsync(b->d_afm, b->h_afm.data(), sizeof(AfmStageArgs) * b->h_afm.size(),
                                    hipMemcpyHostToDevice, b->stream), "batch stage arguments");
    // the host tables are rewritten by the next build: wait for the copies
    if (!st) st = bhip(b, hipStreamSynchronize(b->stream), "batch stage arguments");
    return st;
}

// Workgroup slot tables of one magnet kind for a slot of `chunk` cells.
gf_status build_tables(gf_batch* b, bool fm, int64_t chunk) {
    std::vector<uint32_t> item, block0(b->ctxs.size(), 0);
    for (size_t i = 0; i < b->ctxs.size(); i++) {
        const gf_ctx* c = b->ctxs[i];
        if (fm ? !c->any_fm : !c->any_afm) continue;
        block0[i] = static_cast<uint32_t>(item.size());
        const int64_t blocks = (c->n + chunk - 1) / chunk;
        if (static_cast<int64_t>(item.size()) + blocks > int64_t{0xffffffff})
            return berr(b, GF_ERR_UNSUPPORTED, "batch has more than 2^32 workgroups");
        item.insert(item.end(), static_cast<size_t>(blocks), static_cast<uint32_t>(i));
    }
    gf_status st;
    if (fm) {
        if ((st = dev_copy(b, &b->d_fm_item, item, "batch tables"))) return st;
        if ((st = dev_copy(b, &b->d_fm_block0, block0, "batch tables"))) return st;
        b->fm_total = static_cast<uint32_t>(item.size());
    } else {
        if ((st = dev_copy(b, &b->d_afm_item, item, "batch tables"))) return st;
        if ((st = dev_copy(b, &b->d_afm_block0, block0, "batch tables"))) return st;
        b->afm_total = static_cast<uint32_t>(item.size());
    }
    return GF_OK;
}

unsigned grid_for(uint32_t total, unsigned cap) {
    if (cap == 0 || total < cap) return total;
    return cap;
}

hipError_t launch_stage(gf_batch* b, int stage) {
    hipError_t e = hipSuccess;
    if (b->any_fm && b->fm_total) {
        const BatchLaunch L{b->d_fm_item, b->d_fm_block0, b->fm_total, grid_for(b->fm_total, b->grid_cap),
                            static_cast<unsigned>(b->threads), b->cpt, b->fm_narrow};
        e = launch_mag_batch(b->d_fm, L, stage, b->stream);
    }
    if (e == hipSuccess && b->any_afm && b->afm_total) {
        const BatchLaunch L{b->d_afm_item, b->d_afm_block0, b->afm_total, grid_for(b->afm_total, b->grid_cap),
                            static_cast<unsigned>(b->threads), b->cpt, b->afm_narrow};
        e = launch_afm_batch(b->d_afm, L, stage, b->stream);
    }
    return e;
}

hipError_t launch_step(gf_batch* b) {
    hipError_t e = launch_batch_clock(b->d_clk, b->d_fm, b->d_afm, static_cast<int>(b->ctxs.size()), b->stream);
    for (int s = 1; s <= 4 && e == hipSuccess; s++) e = launch_stage(b, s);
    return e;
}

void drop_graphs(gf_batch* b) {
    if (b->graph1) hipGraphExecDestroy(b->graph1);
    if (b->graphN) hipGraphExecDestroy(b->graphN);
    b->graph1 = b->graphN = nullptr;
}

gf_status capture(gf_batch* b, int steps, hipGraphExec_t* out) {
    gf_status st = bhip(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    if (st) return st;
    hipError_t e = hipSuccess;
    for (int k = 0; k < steps && e == hipSuccess; k++) e = launch_step(b);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return bhip(b, e, "batched launch during capture");
    }
    if ((st = bhip(b, e2, "graph capture end"))) return st;
    st = bhip(b, hipGraphInstantiate(out, g, nullptr, nullptr, 0), "graph instantiate");
    hipGraphDestroy(g);
    return st;
}

gf_status rebuild(gf_batch* b) {
    drop_graphs(b);
    gf_status st;
    const int64_t chunk = int64_t{b->threads} * b->cpt;
    if (b->any_fm && (st = build_tables(b, true, chunk))) return st;
    if (b->any_afm && (st = build_tables(b, false, chunk))) return st;
    if ((st = capture(b, 1, &b->graph1))) return st;
    return capture(b, kGraphSteps, &b->graphN);
}

void release(gf_batch* b) {
    drop_graphs(b);
    hipFree(b->d_fm); hipFree(b->d_afm); hipFree(b->d_clk);
    hipFree(b->d_fm_item); hipFree(b->d_fm_block0); hipFree(b->d_afm_item); hipFree(b->d_afm_block0);
    if (b->order) hipEventDestroy(b->order);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

// The batch stream waits for every context's queued work, or every context
// stream waits for the batch.
gf_status join(gf_batch* b, bool batch_waits) {
    gf_status st = GF_OK;
    if (batch_waits) {
        for (gf_ctx* c : b->ctxs) {
            if ((st = bhip(b, hipEventRecord(b->order, c->stream), "order event"))) return st;
            if ((st = bhip(b, hipStreamWaitEvent(b->stream, b->order, 0), "order wait"))) return st;
        }
        return GF_OK;
    }
    if ((st = bhip(b, hipEventRecord(b->order, b->stream), "order event"))) return st;
    for (gf_ctx* c : b->ctxs)
        if ((st = bhip(b, hipStreamWaitEvent(c->stream, b->order, 0), "order wait"))) return st;
    return GF_OK;
}

}  // namespace

extern "C" {

gf_status gf_batch_create(gf_ctx* const* ctxs, int count, gf_batch** out) {
    g_batch_create_error.clear();
    if (!out) { g_batch_create_error = "out is null"; return GF_ERR_ARG; }
    *out = nullptr;
    if (!ctxs || count <= 0) { g_batch_create_error = "a batch needs at least one context"; return GF_ERR_ARG; }
    for (int i = 0; i < count; i++) {
        if (!ctxs[i]) { g_batch_create_error = "context " + std::to_string(i) + " is null"; return GF_ERR_ARG; }
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) {
                g_batch_create_error = "context " + std::to_string(i) + " appears twice";
                return GF_ERR_ARG;
            }
        const std::string why = unbatchable(ctxs[i]);
        if (!why.empty()) {
            g_batch_create_error = "context " + std::to_string(i) + " " + why;
            return GF_ERR_UNSUPPORTED;
        }
        if (ctxs[i]->device != ctxs[0]->device) {
            g_batch_create_error = "contexts of a batch must share one device";
            return GF_ERR_ARG;
        }
    }
    gf_batch* b = new gf_batch();
    b->ctxs.assign(ctxs, ctxs + count);
    b->device = ctxs[0]->device;
    gf_status st = bhip(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "batch stream");
    if (!st) st = bhip(b, hipEventCreateWithFlags(&b->order, hipEventDisableTiming), "batch event");
    build_args(b, false);
    for (int i = 0; i < count && !st; i++) {
        gf_ctx* c = ctxs[i];
        if (c->any_fm) {
            b->any_fm = true;
            b->fm_cells += c->n;
            if (!narrow(c->n)) b->fm_narrow = false;
        }
        if (c->any_afm) {
            b->any_afm = true;
            b->afm_cells += c->n;
            if (!narrow(c->n)) b->afm_narrow = false;
        }
    }
    b->clk.resize(static_cast<size_t>(count));
    if (!st && b->any_fm) st = dev_copy(b, &b->d_fm, b->h_fm, "batch stage arguments");
    if (!st && b->any_afm) st = dev_copy(b, &b->d_afm, b->h_afm, "batch stage arguments");
    if (!st) st = dev_copy(b, &b->d_clk, b->clk, "batch clocks");
    if (!st) st = rebuild(b);
    if (st) {
        g_batch_create_error = b->err;
        release(b);
        return st;
    }
    *out = b;
    return GF_OK;
}

gf_status gf_batch_destroy(gf_batch* b) {
    if (!b) return GF_ERR_ARG;
    hipStreamSynchronize(b->stream);
    release(b);
    return GF_OK;
}

const char* gf_batch_last_error(const gf_batch* b) {
    if (!b) return g_batch_create_error.c_str();
    return b->err.c_str();
}

gf_status gf_batch_step(gf_batch* b, int64_t nsteps) {
    if (!b) return GF_ERR_ARG;
    if (nsteps < 0) return berr(b, GF_ERR_ARG, "nsteps must be >= 0");
    for (size_t i = 0; i < b->ctxs.size(); i++)
        if (!at_step_boundary(b->ctxs[i]))
            return berr(b, GF_ERR_STATE, "context " + std::to_string(i) + " is in the middle of a step");
    if (nsteps == 0) return GF_OK;
    // the previous call's clock upload has finished before the host
    // buffer is rewritten
    gf_status st = bhip(b, hipStreamSynchronize(b->stream), "batch sync");
    if (st) return st;
    for (size_t i = 0; i < b->ctxs.size(); i++) {
        const gf_ctx* c = b->ctxs[i];
        BatchClock& k = b->clk[i];
        k.t = c->time; k.dt = c->dt;
        for (int a = 0; a < 3; a++) { k.stat[a] = c->hext_stat[a]; k.altr[a] = c->hext_altr[a]; k.freq[a] = c->h_altr_freq[a]; }
    }
    if ((st = join(b, true))) return st;
    st = bhip(b, hipMemcpyAsync(b->d_clk, b->clk.data(), sizeof(BatchClock) * b->clk.size(), hipMemcpyHostToDevice,
                                b->stream), "batch clock upload");
    // the stage arguments follow the contexts as they are now; a context
    // whose stage-1 store is not m (an uploaded store) runs its first step
    // from that store, as gf_step would
    int64_t left = nsteps;
    bool odd = false;
    for (const gf_ctx* c : b->ctxs)
        odd = odd || (c->any_fm && c->mag_store_sel != 0) || (c->any_afm && c->afm_store_sel != 0);
    if (!st && odd) {
        build_args(b, true);
        st = upload_args(b);
        if (!st) st = bhip(b, hipGraphLaunch(b->graph1, b->stream), "batch graph launch");
        left--;
    }
    if (!st) {
        build_args(b, false);
        st = upload_args(b);
    }
    for (int64_t k = 0; k + kGraphSteps <= left && !st; k += kGraphSteps)
        st = bhip(b, hipGraphLaunch(b->graphN, b->stream), "batch graph launch");
    for (int64_t k = 0; k < left % kGraphSteps && !st; k++)
        st = bhip(b, hipGraphLaunch(b->graph1, b->stream), "batch graph launch");
    if (st) return st;
    if ((st = join(b, false))) return st;
    // every context is now where a stage 4 leaves it: the state in m,
    // stage 4 having read store 1
    for (gf_ctx* c : b->ctxs) {
        if (c->any_fm) { c->mag_store_sel = 0; c->mag_in_sel = 1; }
        if (c->any_afm) { c->afm_store_sel = 0; c->afm_in_sel = 1; }
    }
    // host copies of the clocks follow the same additions as the device
    for (gf_ctx* c : b->ctxs) {
        for (int64_t k = 0; k < nsteps; k++) c->time = c->time + c->dt;
        for (int a = 0; a < 3; a++)
            c->hext[a] = c->hext_stat[a] + c->hext_altr[a] * std::sin(2. * kPi * c->h_altr_freq[a] * c->time);
    }
    return GF_OK;
}

gf_status gf_batch_sync(gf_batch* b) {
    if (!b) return GF_ERR_ARG;
    return bhip(b, hipStreamSynchronize(b->stream), "batch sync");
}

gf_status gf_batch_get_stream(gf_batch* b, void** stream) {
    if (!b || !stream) return GF_ERR_ARG;
    *stream = b->stream;
    return GF_OK;
}

gf_status gf_batch_config(const gf_batch* b, int* threads, int* cells_per_thread, int* grid_cap) {
    if (!b || !threads || !cells_per_thread || !grid_cap) return GF_ERR_ARG;
    *threads = b->threads;
    *cells_per_thread = b->cpt;
    *grid_cap = static_cast<int>(b->grid_cap);
    return GF_OK;
}

gf_status gf_batch_set_config(gf_batch* b, int threads, int cells_per_thread, int grid_cap) {
    if (!b) return GF_ERR_ARG;
    if (threads != 64 && threads != 128 && threads != 256)
        return berr(b, GF_ERR_ARG, "threads must be 64, 128 or 256");
    if (cells_per_thread < 1 || cells_per_thread > 16)
        return berr(b, GF_ERR_ARG, "cells_per_thread must be in 1..16");
    if (grid_cap < 0) return berr(b, GF_ERR_ARG, "grid_cap must be >= 0");
    gf_status st = bhip(b, hipStreamSynchronize(b->stream), "batch sync");
    if (st) return st;
    b->threads = threads;
    b->cpt = cells_per_thread;
    b->grid_cap = static_cast<unsigned>(grid_cap);
    return rebuild(b);
}

gf_status gf_batch_tune(gf_batch* b, int reps, double* best_ms) {
    if (!b) return GF_ERR_ARG;
    if (reps < 1) return berr(b, GF_ERR_ARG, "reps must be >= 1");
    for (size_t i = 0; i < b->ctxs.size(); i++)
        if (!at_step_boundary(b->ctxs[i]))
            return berr(b, GF_ERR_STATE, "context " + std::to_string(i) + " is in the middle of a step");
    // the slot tables are replaced below: no queued batch work may still use them
    gf_status st = bhip(b, hipStreamSynchronize(b->stream), "batch sync");
    if (!st) st = join(b, true);
    if (st) return st;
    int cu = 0;
    if ((st = bhip(b, hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, b->device), "CU count")))
        return st;
    const int threads[3] = {64, 128, 256};
    const int cpts[4] = {1, 2, 4, 8};
    const unsigned caps[2] = {0u, static_cast<unsigned>(16 * cu)};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((st = bhip(b, hipEventCreate(&e0), "tune event"))) return st;
    if ((st = bhip(b, hipEventCreate(&e1), "tune event"))) { hipEventDestroy(e0); return st; }
    double best = -1.;
    int best_t = b->threads, best_p = b->cpt;
    unsigned best_c = b->grid_cap;
    for (int k = 0; k < 12 && !st; k++) {
        b->threads = threads[k / 4];
        b->cpt = cpts[k % 4];
        const int64_t chunk = int64_t{b->threads} * b->cpt;
        if (b->any_fm && (st = build_tables(b, true, chunk))) break;
        if (b->any_afm && (st = build_tables(b, false, chunk))) break;
        for (int ci = 0; ci < 2 && !st; ci++) {
            b->grid_cap = caps[ci];
            st = bhip(b, launch_stage(b, 1), "tune launch");
            if (!st) st = bhip(b, hipEventRecord(e0, b->stream), "tune event");
            for (int r = 0; r < reps && !st; r++) st = bhip(b, launch_stage(b, 1), "tune launch");
            if (!st) st = bhip(b, hipEventRecord(e1, b->stream), "tune event");
            if (!st) st = bhip(b, hipEventSynchronize(e1), "tune sync");
            float ms = 0.f;
            if (!st) st = bhip(b, hipEventElapsedTime(&ms, e0, e1), "tune time");
            if (!st && (best < 0. || ms < best)) {
                best = ms; best_t = b->threads; best_p = b->cpt; best_c = b->grid_cap;
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->threads = best_t;
    b->cpt = best_p;
    b->grid_cap = best_c;
    gf_status st2 = rebuild(b);
    if (st) return st;
    if (st2) return st2;
    if ((st = join(b, false))) return st;
    if (best_ms) *best_ms = best / reps;
    return GF_OK;
}

}  // extern "C"
