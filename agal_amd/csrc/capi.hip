// capi.hip -- the extern "C" entry points of include/agal_capi.h.
//
// Host orchestration only: building the root grid (Mesh::M_Init, src/mesh/mesh_init.cu:10-567),
// the recursive sub-cycling order of Solver_LBM::S_Advance (solver_lbm_advance.cu:30-174), the
// solution-based rebuild driver (mesh_advance.cu:184-243) and data exchange.  All arithmetic on
// populations runs in the kernels of lbm_kernels.hip / amr_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

using namespace agal;

namespace agal {

int hip_fail(Ctx *c, hipError_t e, const char *what)
{
    if (c) c->last_error = std::string(what) + ": " + hipGetErrorString(e);
    return AGAL_ERR_HIP;
}

static hipError_t sched_event(Ctx *c, int &idx)
{
    if (c->ev_next >= (int)c->evpool.size()) {
        hipEvent_t e;
        hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (r != hipSuccess) return r;
        c->evpool.push_back(e);
    }
    idx = c->ev_next++;
    return hipSuccess;
}

static hipStream_t slot_stream(Ctx *c, int slot) { return slot == 0 ? c->stream : c->lstream[slot - 1]; }

hipStream_t sched_begin(Ctx *c, int slot, unsigned rmask, unsigned wmask)
{
    if (!c->sched_on) return c->stream;
    hipStream_t s = slot_stream(c, slot);
    if (!c->slot_joined[slot]) {
        (void)hipStreamWaitEvent(s, c->sched_start, 0);
        c->slot_joined[slot] = true;
    }
    auto wait = [&](int ev, int from) {
        if (from == slot || ev <= c->seen[slot][from]) return;
        (void)hipStreamWaitEvent(s, c->evpool[ev], 0);
        c->seen[slot][from] = ev;
    };
    for (int r = 0; r < Ctx::NRES; r++) {
        if (!((rmask | wmask) >> r & 1)) continue;
        const auto &d = c->deps[r];
        if (d.w_ev >= 0) wait(d.w_ev, d.w_slot);
        if (wmask >> r & 1)
            for (auto &rd : d.readers) wait(rd.first, rd.second);
    }
    return s;
}

int sched_done(Ctx *c, hipStream_t s, int slot, unsigned rmask, unsigned wmask)
{
    if (!c->sched_on) return AGAL_OK;
    int e;
    AGAL_HIP(c, sched_event(c, e));
    AGAL_HIP(c, hipEventRecord(c->evpool[e], s));
    for (int r = 0; r < Ctx::NRES; r++) {
        auto &d = c->deps[r];
        if (wmask >> r & 1) {
            d.w_ev = e;
            d.w_slot = slot;
            d.readers.clear();
        } else if (rmask >> r & 1) {
            d.readers.push_back({e, slot});
        }
    }
    return AGAL_OK;
}

// Element-wise gather (dir 0) or scatter (dir 1) between a block's current population buffer
// and a contiguous staging array of decoded values, Q rows of n_blocks*cells.
struct Weights {
    double w[27];
};

template <typename T, typename S>
__global__ void k_stage_pops(S *F0, S *F1, T *stage, long long ncm, long long n_blocks, int cells, int Q,
                             const int *level, int4 cur_lo, int4 cur_hi, int dir, Weights wt)
{
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    long long nc = n_blocks * cells;
    if (i >= nc) return;
    int b = (int)(i / cells);
    int L = level[b];
    int cl[8] = {cur_lo.x, cur_lo.y, cur_lo.z, cur_lo.w, cur_hi.x, cur_hi.y, cur_hi.z, cur_hi.w};
    int k = (L >= 0 && L < 8) ? cl[L] : 0;
    S *F = k ? F1 : F0;
    for (int q = 0; q < Q; q++) {
        if (dir == 0)
            stage[q * nc + i] = Store<T, S>::dec(F[q * ncm + i], (T)wt.w[q]);
        else
            F[q * ncm + i] = Store<T, S>::enc(stage[q * nc + i], (T)wt.w[q]);
    }
}

__global__ void k_init_gaps(int *gap, long long n, long long nmax)
{
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) gap[i] = (int)(nmax - 1 - i);
}

__global__ void k_mark_uniform(const int *ids, int n, int *ref)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && ref[ids[i]] == REF_UNREFINED) ref[ids[i]] = REF_MARK_REFINE;
}

__global__ void k_mark_list(const int *ids, int n, int *ref, int mark)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ref[ids[i]] = mark;
}

__global__ void k_freeze(int *ref, int n, int var)
{
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int r = ref[k];
    if (var == 0) {
        if (r == REF_REFINED) ref[k] = REF_REFINED_PERM;
        if (r == REF_REFINED_WCHILD) ref[k] = REF_REFINED_WCHILD_PERM;
    } else {
        if (r == REF_REFINED_PERM) ref[k] = REF_REFINED;
        if (r == REF_REFINED_WCHILD_PERM) ref[k] = REF_REFINED_WCHILD;
    }
}

}  // namespace agal

namespace {

// Level spacings and relaxation times in the precision's type, computed the way the reference
// does (mesh_init.cu:22-29,218-220; solver_lbm_init.cu:47-52).
template <typename T>
void compute_level_params(Ctx *c)
{
    float Lx = (float)c->P.L_c;
    T dx = (T)(Lx / c->P.Nx);
    T dxc = 4 * dx;
    c->dx = (double)dx;
    c->dx_cblock = (double)dxc;
    c->dxf_vec.assign(c->MAX_LEVELS, 0.0);
    c->tau_vec.assign(c->MAX_LEVELS, 0.0);
    for (int L = 0; L < c->MAX_LEVELS; L++) {
        T dxf = (T)((double)dx / std::pow(2.0, (double)(T)L));
        c->dxf_vec[L] = (double)dxf;
        c->tau_vec[L] = (double)(T)(c->P.v0 * 3.0 + 0.5 * (double)dxf);
    }
    // The control volume snapped to root-block edges (solver_lbm_init.cu:36-41).
    const double fr[6] = {c->P.S_FORCEVOLUME_Xm, c->P.S_FORCEVOLUME_XM, c->P.S_FORCEVOLUME_Ym,
                          c->P.S_FORCEVOLUME_YM, c->P.S_FORCEVOLUME_Zm, c->P.S_FORCEVOLUME_ZM};
    for (int k = 0; k < 6; k++)
        c->cv_box[k] = (double)(T)(4.0 * (double)(int)(fr[k] * (double)(c->Nxi[k / 2] / 4)) * c->dxf_vec[0]);
}

void free_ctx(Ctx *c)
{
    if (!c) return;
    void *ptrs[] = {c->F[0], c->F[1], c->M, c->cells_mask, c->X, c->nbr, c->nbr_child, c->nbr_slot, c->ref, c->level,
                    c->onb, c->blk_mask, c->part, c->simple, c->id_set, c->order, c->gap_set, c->iface, c->map_stamp,
                    c->first_child, c->scan_a, c->scan_b, c->partials, c->dscal, c->exc_idx, c->exc_val,
                    c->exc_count, c->exc_head, c->exc_next, c->cv_F};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free_geometry(c);
    free_links(c);
    free_predictor(c);
    if (c->hscal) (void)hipHostFree(c->hscal);
    for (auto &g : c->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    if (c->sched_start) (void)hipEventDestroy(c->sched_start);
    for (hipStream_t s : c->lstream)
        if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : c->timing.ev_pairs) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <typename P>
int dalloc(Ctx *c, P **p, size_t bytes)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        c->last_error = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        return AGAL_ERR_ALLOC;
    }
    *p = static_cast<P *>(v);
    return AGAL_OK;
}

// Root grid of Mesh::M_Init (mesh_init.cu:325-516): block IDs in x-fastest order over a padded
// index grid whose rim holds the face IDs (-1..-6, x overriding y overriding z on edges) or the
// periodic images.
int build_root(Ctx *c)
{
    const int DIM = c->DIM, QM = c->QMAX;
    const long long nmax = c->n_maxcblocks;
    const int bx = c->Nxi[0] / 4, by = c->Nxi[1] / 4;
    const int gh = DIM == 2 ? 1 : c->Nxi[2] / 4;
    const int ghm = DIM == 2 ? 0 : 1;
    const int px = bx + 2, py = by + 2, pz = gh + 2;
    std::vector<int> grid((size_t)px * py * pz, 0);
    auto G = [&](int i, int j, int k) -> int & { return grid[(size_t)i + px * ((size_t)j + py * k)]; };
    int id = 0;
    for (int k = ghm; k < gh * ghm + 1; k++)
        for (int j = 1; j < by + 1; j++)
            for (int i = 1; i < bx + 1; i++) G(i, j, k) = id++;
    if (DIM == 3)
        for (int j = 0; j < py; j++)
            for (int i = 0; i < px; i++) {
                if (c->P.PERIODIC_Z == 1) {
                    G(i, j, 0) = G(i, j, pz - 2);
                    G(i, j, pz - 1) = G(i, j, 1);
                } else {
                    G(i, j, 0) = -5;
                    G(i, j, pz - 1) = -6;
                }
            }
    for (int k = 0; k < gh + 2 * ghm; k++)
        for (int i = 0; i < px; i++) {
            if (c->P.PERIODIC_Y == 1) {
                G(i, 0, k) = G(i, py - 2, k);
                G(i, py - 1, k) = G(i, 1, k);
            } else {
                G(i, 0, k) = -3;
                G(i, py - 1, k) = -4;
            }
        }
    for (int k = 0; k < gh + 2 * ghm; k++)
        for (int j = 0; j < py; j++) {
            if (c->P.PERIODIC_X == 1) {
                G(0, j, k) = G(px - 2, j, k);
                G(px - 1, j, k) = G(1, j, k);
            } else {
                G(0, j, k) = -1;
                G(px - 1, j, k) = -2;
            }
        }
    const int nc = c->n_coarsecblocks;
    std::vector<int> nbr((size_t)QM * nc), nbr_child((size_t)QM * nc), onb(nc), ref(nc, REF_UNREFINED);
    std::vector<float> X((size_t)DIM * nc);
    for (int k = ghm; k < gh * ghm + 1; k++)
        for (int j = 1; j < by + 1; j++)
            for (int i = 1; i < bx + 1; i++) {
                int b = G(i, j, k);
                auto coord = [&](int n) {
                    if (c->precision == AGAL_FP64) return (float)((double)n * c->dx_cblock);
                    return (float)((float)n * (float)c->dx_cblock);
                };
                X[b + 0 * (size_t)nc] = coord(i - 1);
                X[b + 1 * (size_t)nc] = coord(j - 1);
                if (DIM == 3) X[b + 2 * (size_t)nc] = coord(k - 1);
                int ob = 0;
                for (int p = 0; p < QM; p++) {
                    int cx = DIM == 3 ? c3x(p) : c2x(p), cy = DIM == 3 ? c3y(p) : c2y(p), cz = DIM == 3 ? c3z(p) : 0;
                    int nb = G(i + cx, j + cy, ghm * (k + cz));
                    nbr[b + (size_t)p * nc] = nb;
                    if (nb < 0) ob = 1;
                    nbr_child[b + (size_t)p * nc] = nb < 0 ? nb : N_SKIPID;
                }
                onb[b] = ob;
            }
    for (int p = 0; p < QM; p++) {
        AGAL_HIP(c, hipMemcpy(c->nbr + p * nmax, nbr.data() + (size_t)p * nc, sizeof(int) * nc, hipMemcpyHostToDevice));
        AGAL_HIP(c, hipMemcpy(c->nbr_child + p * nmax, nbr_child.data() + (size_t)p * nc, sizeof(int) * nc, hipMemcpyHostToDevice));
    }
    for (int d = 0; d < DIM; d++)
        AGAL_HIP(c, hipMemcpy(c->X + d * nmax, X.data() + (size_t)d * nc, sizeof(float) * nc, hipMemcpyHostToDevice));
    AGAL_HIP(c, hipMemcpy(c->onb, onb.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    AGAL_HIP(c, hipMemcpy(c->ref, ref.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    return AGAL_OK;
}

// V_AVERAGE_ADVANCE (solver_lbm_init.cu:94-97): INTERFACE after linear interpolation, BLOCK
// after any other order.
static int average_advance(const Ctx *c)
{
    return c->P.interp_order == AGAL_INTERP_CUBIC ? AGAL_AVERAGE_BLOCK : AGAL_AVERAGE_INTERFACE;
}

int advance_rec(Ctx *c, int L)
{
    int st;
    const int ML = c->MAX_LEVELS, L0 = c->N_LEVEL_START;
    if (ML > 1 && ML != L0 + 1) {
        // Interpolation of level L and its first fused sub-step share a launch inside a coarse
        // step; the averages wait for the pending fixups of the levels they read.
        if (L == L0) {
            c->global_Liter = 0;
            if ((st = launch_interp_collide(c, L))) return st;
            if ((st = advance_rec(c, L + 1))) return st;
            if ((st = launch_fix_pending(c))) return st;
            if ((st = launch_average(c, L, average_advance(c)))) return st;
        } else {
            if (L < ML - 1) {
                if ((st = launch_interp_collide(c, L))) return st;
            } else {
                // The finest level's two sub-steps run back to back: the first leaves its
                // exception list for the second to apply as it loads (fine_pair 1, then 2).
                c->fine_pair = 1;
                st = launch_collide_stream(c, L);
                c->fine_pair = 0;
                if (st) return st;
            }
            if (L == ML - 1) c->global_Liter++;
            if (L < ML - 1) {
                if ((st = advance_rec(c, L + 1))) return st;
                if ((st = launch_fix_pending(c))) return st;
                if ((st = launch_average_interpolate(c, L))) return st;
            }
            if (L == ML - 1) c->fine_pair = 2;
            st = launch_collide_stream(c, L);
            c->fine_pair = 0;
            if (st) return st;
            if (L == ML - 1) c->global_Liter++;
            if (L < ML - 1) {
                if ((st = advance_rec(c, L + 1))) return st;
                if ((st = launch_fix_pending(c))) return st;
                if ((st = launch_average(c, L, average_advance(c)))) return st;
            }
        }
    } else {
        if ((st = launch_collide_stream(c, L))) return st;
    }
    return AGAL_OK;
}

// One coarse step from level L0 with the coarse levels' fused sub-steps folded into the finest
// level's launches (the held-back launches are issued before returning).
int advance_step(Ctx *c, int L0)
{
    int st;
    const bool cv = c->cv_F != nullptr;
    const int ML = c->MAX_LEVELS;
    const bool multi = ML > 1 && ML != L0 + 1;
    // Control-volume force (solver_lbm_advance.cu:45-49, :139-140): the first half on every level
    // (the single level only, without refinement) before the step.
    const int K_end = multi ? ML : L0 + 1;
    if (cv)
        for (int K = L0; K < K_end; K++)
            if ((st = launch_cv_force(c, K, 0))) return st;
    c->defer_coarse = true;
    st = advance_rec(c, L0);
    c->defer_coarse = false;
    const int st2 = launch_flush_deferred(c);
    if (st || st2) return st ? st : st2;
    if (cv) {
        // The second half (:60-68, :148-149): with refinement, after a full fine-to-coarse average
        // and the interface interpolations of every level.
        if (multi) {
            for (int K = ML - 2; K >= L0; K--)
                if ((st = launch_average(c, K, AGAL_AVERAGE_GRID))) return st;
            for (int K = L0; K < ML - 1; K++)
                if ((st = launch_interpolate(c, K, AGAL_INTERP_INTERFACE))) return st;
        }
        for (int K = L0; K < K_end; K++)
            if ((st = launch_cv_force(c, K, 1))) return st;
    }
    // The after-stream half of the momentum exchange, on the finished finest level.
    return launch_mea_post(c);
}

unsigned cur_mask(const Ctx *c)
{
    unsigned m = 0;
    for (int L = 0; L < c->MAX_LEVELS; L++) m |= (unsigned)(c->cur[L] & 1) << L;
    return m;
}

void set_cur_mask(Ctx *c, unsigned m)
{
    for (int L = 0; L < c->MAX_LEVELS; L++) c->cur[L] = (m >> L) & 1;
}

// One coarse step from level L0.  With graphs enabled the launch sequence of advance_rec is
// recorded once per (mesh epoch, starting buffer parity) and replayed; the host-side bookkeeping
// (buffer parity per level) is advanced exactly as a direct run would.
// `k` consecutive coarse steps (no rebuild between them) as one recorded graph: one graph launch,
// one wait between graph launches, per k steps.
int steps_once(Ctx *c, int L0, int k)
{
    int st;
    if (!c->use_graphs || c->timing.enabled || L0 != c->N_LEVEL_START) {
        for (int j = 0; j < k; j++)
            if ((st = advance_step(c, L0))) return st;
        return AGAL_OK;
    }
    const unsigned before = cur_mask(c);
    for (auto &g : c->graphs)
        if (g.exec && g.epoch == c->mesh_epoch && g.cur_before == before && g.nsteps == k) {
            AGAL_HIP(c, hipGraphLaunch(g.exec, c->stream));
            set_cur_mask(c, g.cur_after);
            c->global_Liter = 1 << (c->MAX_LEVELS - 1 - L0);
            return AGAL_OK;
        }
    // Record a new graph (evicting a stale or the oldest entry).
    Ctx::StepGraph *slot = &c->graphs[0];
    for (auto &g : c->graphs)
        if (!g.exec || g.epoch != c->mesh_epoch) {
            slot = &g;
            break;
        }
    if (slot->exec) {
        // The evicted graph may still be executing (it was launched asynchronously, possibly by
        // the previous coarse step): wait for the stream before destroying its executable.
        AGAL_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(slot->exec);
        slot->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    AGAL_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    // Record with the level scheduler: fork the level streams from the capture stream, join
    // every used one back before the capture ends.
    c->ev_next = 0;
    for (auto &d : c->deps) d = Ctx::ResDeps();
    for (int k = 0; k < Ctx::NSLOT; k++) {
        c->slot_joined[k] = false;
        for (int t = 0; t < Ctx::NSLOT; t++) c->seen[k][t] = -1;
    }
    st = AGAL_OK;
    hipError_t e = hipEventRecord(c->sched_start, c->stream);
    if (e != hipSuccess) st = hip_fail(c, e, "hipEventRecord");
    c->sched_on = (st == AGAL_OK);
    for (int j = 0; j < k && st == AGAL_OK; j++) st = advance_step(c, L0);
    c->sched_on = false;
    for (int k = 1; k < Ctx::NSLOT; k++) {  // joined also after a failure so that the capture can end
        if (!c->slot_joined[k]) continue;
        int ev;
        e = sched_event(c, ev);
        if (e == hipSuccess) e = hipEventRecord(c->evpool[ev], slot_stream(c, k));
        if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->evpool[ev], 0);
        if (e != hipSuccess && st == AGAL_OK) st = hip_fail(c, e, "join level stream");
    }
    e = hipStreamEndCapture(c->stream, &graph);
    if (st != AGAL_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        set_cur_mask(c, before);
        return st;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    const unsigned after = cur_mask(c);
    e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        slot->exec = nullptr;
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    slot->cur_before = before;
    slot->cur_after = after;
    slot->epoch = c->mesh_epoch;
    slot->nsteps = k;
    AGAL_HIP(c, hipGraphLaunch(slot->exec, c->stream));
    return AGAL_OK;
}

int step_once(Ctx *c, int L0) { return steps_once(c, L0, 1); }

int check_exceptions(Ctx *c);

// The field updates of M_Advance_RefineWithSolution before the criterion
// (src/mesh/mesh_advance.cu:204-217): the reference's S_Average / S_Interpolate refresh the
// macroscopic fields first (solver_lbm_kernels_comm.cu:143,354); the criterion reads them.
int prepare_fields(Ctx *c)
{
    int st;
    for (int L = c->MAX_LEVELS - 2; L >= c->N_LEVEL_START; L--) {
        if ((st = launch_refresh(c, L + 1))) return st;
        if ((st = launch_average(c, L, AGAL_AVERAGE_GRID))) return st;
    }
    for (int L = c->N_LEVEL_START; L < c->MAX_LEVELS - 1; L++) {
        if ((st = launch_refresh(c, L))) return st;
        if ((st = launch_interpolate(c, L, AGAL_INTERP_INTERFACE))) return st;
    }
    if (c->MAX_LEVELS == 1 && (st = launch_refresh(c, 0))) return st;
    return AGAL_OK;
}

int refine_with_solution_impl(Ctx *c, int out_counts[4])
{
    int st;
    if (!(c->MAX_LEVELS_INTERIOR > 1)) {
        if (out_counts) std::fill(out_counts, out_counts + 4, 0);
        return AGAL_OK;
    }
    if ((st = prepare_fields(c))) return st;
    if ((st = launch_ref_criteria(c))) return st;
    return refine_coarsen(c, out_counts);
}

// The plain criterion's labels without changing the refinement marks (the marks are saved and
// restored around the criterion).
int criterion_labels(Ctx *c)
{
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(c->scan_b, c->ref, sizeof(int) * idm, hipMemcpyDeviceToDevice, c->stream));
    int st = launch_ref_criteria(c, c->pred.label);
    AGAL_HIP(c, hipMemcpyAsync(c->ref, c->scan_b, sizeof(int) * idm, hipMemcpyDeviceToDevice, c->stream));
    return st;
}

// Two timing events that live until the owning scope ends, on every return path.
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~EventPair()
    {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// M_AdvanceLoop with the learned predictor (no reference counterpart).  Every `infer_every` coarse
// steps the features of all blocks are computed; during the first `train_steps` steps they are
// kept as snapshots, and at each solution-based rebuild the plain criterion's labels train the
// network (`adam_steps` Adam steps on batches sampled from the snapshots).  After that the
// network runs at every `infer_every` step: unrefined blocks it predicts are refined at once
// (a refine-only rebuild), and at each rebuild its predicted blocks are kept from coarsening while
// the plain criterion still refines what it flags (fallback).  The audit counts, at every rebuild,
// the blocks the criterion wants refined and those of them still unrefined (missed).
int advance_loop_learned_impl(Ctx *c, int i0, int n, const agal_learn_opts *o, agal_learn_stats *S)
{
    int st;
    const int ML = c->MAX_LEVELS, L0 = c->N_LEVEL_START, iter_mult = 1 << L0;
    const int every = std::max(1, o->infer_every);
    const int max_slots = std::max(1, c->P.P_REFINE / every + 1);
    const float thr = (float)o->threshold, lr = (float)o->lr, pw = (float)o->pos_weight;
    std::memset(S, 0, sizeof(*S));
    if (!c->pred.w || c->pred.batch_cap < o->batch)
        if ((st = predictor_setup(c, o->seed, o->batch))) return st;
    uint64_t rng = o->seed ^ 0x5DEECE66DULL;
    int n_snap = 0;
    auto ms_since = [](std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    EventPair ev;
    AGAL_HIP(c, hipEventCreate(&ev.e[0]));
    AGAL_HIP(c, hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    auto close_window = [&]() -> int {
        float ms = 0;
        AGAL_HIP(c, hipEventRecord(e1, c->stream));
        AGAL_HIP(c, hipEventSynchronize(e1));
        AGAL_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        S->step_ms += ms;
        return AGAL_OK;
    };
    AGAL_HIP(c, hipEventRecord(e0, c->stream));
    for (int i = i0; i < i0 + n; i++) {
        for (int j = 0; j < iter_mult; j++)
            if ((st = step_once(c, L0))) return st;
        const bool training = (i + 1) <= o->train_steps;
        const bool rebuild = ML > 1 && ML != L0 + 1 && (i + 1) % c->P.P_REFINE == 0 && i > 0 && c->MAX_LEVELS_INTERIOR > 1;
        const bool infer = (i + 1) % every == 0;
        if (!rebuild && !infer) continue;
        if ((st = close_window())) return st;
        auto t0 = std::chrono::steady_clock::now();
        if ((st = prepare_fields(c)) || (st = predictor_features(c))) return st;
        if (!rebuild) {
            if (training) {
                if (n_snap < max_slots && (st = predictor_snapshot(c, n_snap++, max_slots))) return st;
            } else if (o->use_predictions == 1 || o->use_predictions == 2) {
                int marked = 0;
                if (o->use_predictions == 1) {
                    if ((st = predictor_predict_blocks(c)) || (st = predictor_mark(c, thr, &marked))) return st;
                } else {
                    // Control: the plain criterion's refinement marks at the inference cadence.
                    if ((st = launch_ref_criteria(c)) || (st = marks_refine_only(c)) || (st = count_refine_marks(c, &marked)))
                        return st;
                }
                if (marked > 0) {
                    st = refine_coarsen(c, nullptr);
                    if (st && st != AGAL_ERR_POOL_EXHAUSTED) return st;
                    S->early_rebuilds++;
                    S->predicted_refinements += marked;
                }
            }
            AGAL_HIP(c, hipStreamSynchronize(c->stream));
            S->predictor_ms += ms_since(t0);
        } else {
            // Audit and labels from the plain criterion, then its marks.
            if ((st = criterion_labels(c))) return st;
            int jc[2] = {0, 0};
            if ((st = predictor_audit(c, jc))) return st;
            if (training) {
                S->flagged_train += jc[0];
                S->missed_train += jc[1];
                double loss = 0;
                if ((st = predictor_train_snapshots(c, n_snap, o->batch, o->adam_steps, lr, pw, rng, &loss))) return st;
                if (n_snap > 0) {
                    S->adam_steps_done += o->adam_steps;
                    S->last_loss = loss;
                }
                n_snap = 0;
            } else {
                S->flagged += jc[0];
                S->missed += jc[1];
            }
            AGAL_HIP(c, hipStreamSynchronize(c->stream));
            S->predictor_ms += ms_since(t0);
            auto t1 = std::chrono::steady_clock::now();
            if ((st = launch_ref_criteria(c))) return st;
            if (!training && o->use_predictions == 3 && (st = marks_dilate(c))) return st;
            if (!training && o->use_predictions == 1) {
                int marked = 0;
                if ((st = predictor_predict_blocks(c)) || (st = predictor_keep(c, thr)) ||
                    (st = predictor_mark(c, thr, &marked)))
                    return st;
                S->predicted_refinements += marked;
            }
            st = refine_coarsen(c, nullptr);
            if (st && st != AGAL_ERR_POOL_EXHAUSTED) return st;
            AGAL_HIP(c, hipStreamSynchronize(c->stream));
            S->rebuilds++;
            S->rebuild_ms += ms_since(t1);
        }
        AGAL_HIP(c, hipEventRecord(e0, c->stream));
    }
    if ((st = close_window())) return st;
    return check_exceptions(c);
}

// Mesh::M_Advance_RefineNearWall (src/mesh/mesh_advance.cu:53-121).  The benchmark-case hook
// Cu_RefineRegion marks nothing as shipped (inc/custom.h:17-43).  With a geometry, each wall level
// is voxelized and marked before its rebuild (M_Geometry_Voxelize_S1) and the boundary cells are
// found once the hierarchy is complete (M_Geometry_Voxelize_S2).  A cancelled refinement (pool
// exhausted) does not stop the sequence; its status is returned at the end.
int refine_near_wall_impl(Ctx *c)
{
    int st, status = AGAL_OK;
    const bool geom = c->geo_tri != nullptr;
    const int ML = c->MAX_LEVELS;
    auto keep = [&](int s) {
        if (s == AGAL_ERR_POOL_EXHAUSTED) {
            if (status == AGAL_OK) status = s;
            return AGAL_OK;
        }
        return s;
    };
    for (int L = 0; L < c->N_LEVEL_START; L++) {
        const int n = c->n_ids[L];
        if (n > 0) {
            hipLaunchKernelGGL(k_mark_uniform, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->id_set_L(L), n, c->ref);
            AGAL_HIP(c, hipGetLastError());
        }
        if ((st = keep(refine_coarsen(c, nullptr)))) return st;
    }
    for (int L = c->N_LEVEL_START; L < c->MAX_LEVELS_WALL; L++) {
        if (geom) {
            if ((st = voxelize_level(c, L))) return st;
            if ((st = mark_near_wall(c, L))) return st;
        }
        if (ML > 1 && L < c->MAX_LEVELS_WALL - 1)
            if ((st = keep(refine_coarsen(c, nullptr)))) return st;
        if ((st = launch_set_ic(c, L))) return st;
        c->ic_done = true;
    }
    if (geom) {
        if ((st = links_reset_counts(c))) return st;
        for (int L = c->N_LEVEL_START; L < c->MAX_LEVELS_WALL; L++)
            if ((st = check_masks_level(c, L))) return st;
        // Solid cells and masks changed: participation and fast-path flags of the blocks.
        if ((st = rebuild_level_lists(c, 0))) return st;
        // Ranks of the boundary blocks and, with G_BIN_MAKE_LINKS, their link fractions
        // (M_Geometry_Voxelize_S3, M_LinkLengthComputation; mesh_advance.cu:99-111).
        if (c->P.G_BIN_MAKE_LINKS == 1 && (st = compute_links(c))) return st;
    }
    if (ML > 1) {
        const int n = c->id_max[ML];
        if (n > 0) {
            hipLaunchKernelGGL(k_freeze, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->ref, n, 0);
            AGAL_HIP(c, hipGetLastError());
        }
    }
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return status;
}

// An overflowed exception list dropped links: the populations are wrong from then on, so the
// context refuses further steps (AGAL_ERR_STATE) once the overflow has been reported.
int check_exceptions(Ctx *c)
{
    if (c->exc_overflowed) return AGAL_ERR_OVERFLOW;
    std::vector<int> v(2 * (c->MAX_LEVELS + 1), 0);
    AGAL_HIP(c, hipMemcpyAsync(v.data(), c->exc_count, sizeof(int) * v.size(), hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    for (int L = 0; L <= c->MAX_LEVELS; L++)
        if (v[2 * L + 1]) {
            c->last_error = "stream-guard exception list of level " + std::to_string(L) + " overflowed (" +
                            std::to_string(c->exc_cap) + " entries): populations are invalid, the context refuses "
                            "further steps";
            c->exc_overflowed = true;
            return AGAL_ERR_OVERFLOW;
        }
    return AGAL_OK;
}

// Steps refuse to run on a context whose exception list overflowed.
static int usable(Ctx *c)
{
    if (!c->exc_overflowed) return AGAL_OK;
    c->last_error = "a stream-guard exception list overflowed earlier: the populations are invalid";
    return AGAL_ERR_STATE;
}

int collect_timing(Ctx *c)
{
    auto &tm = c->timing;
    if (tm.ev_pairs.empty()) return AGAL_OK;
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k + 1 < tm.ev_pairs.size(); k += 2) {
        float ms = 0;
        AGAL_HIP(c, hipEventElapsedTime(&ms, tm.ev_pairs[k], tm.ev_pairs[k + 1]));
        tm.fused_ms += ms;
        (void)hipEventDestroy(tm.ev_pairs[k]);
        (void)hipEventDestroy(tm.ev_pairs[k + 1]);
    }
    tm.ev_pairs.clear();
    return AGAL_OK;
}

}  // namespace

#define CTX_OR_ARG(ctx) \
    if (!(ctx)) return AGAL_ERR_ARG;

// ---- allocation-free download: staging area -----------------------------------------------------
// A file-scope device array, one copy per GPU that the library runs on, held until the process
// ends.  On the host `scratch` is only a stand-in symbol, never a device address: host code uses
// the address hipGetSymbolAddress returns on each GPU, looked up once per GPU (with that GPU
// current) when the first context on it is created, and kept in scratch_dev[].
static __device__ __align__(256) unsigned char scratch[AGAL_NOALLOC_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex scratch_mu;
static void *scratch_dev[SCRATCH_MAX_DEVICES];

// Device address of `scratch` on the current GPU (c->device), looked up on first use there.
static int scratch_lookup(Ctx *c)
{
    if (c->device < 0 || c->device >= SCRATCH_MAX_DEVICES) {
        c->last_error = "device index beyond the staging table";
        return AGAL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[c->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        scratch_dev[c->device] = ptr;
    }
    c->scratch = scratch_dev[c->device];
    return AGAL_OK;
}

extern "C" {

void agal_params_default(agal_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->velocity_set = AGAL_D3Q19;
    p->precision = AGAL_FP32;
    p->collision = AGAL_BGK;
    p->interp_order = AGAL_INTERP_LINEAR;
    p->Nx = 64;
    p->L_c = p->L_fy = p->L_fz = 1.0;
    p->MAX_LEVELS_WALL = 5;
    p->MAX_LEVELS_INTERIOR = 4;
    p->N_LEVEL_START = 0;
    p->v0 = 1.5625e-5;
    p->S_CRITERION = 0;
    p->N_REFINE_START = -3.0;
    p->N_REFINE_INC = 1.0;
    p->N_REFINE_MAX = 1.0;
    p->P_REFINE = 32;
    p->bc_preset = AGAL_BC_LDC3D;
    p->bc_velocity = 0.05;
    p->ic_rho = 1.0;
    p->s1 = 1.19;
    p->s2 = 1.4;
    p->s3 = 1.2;
    p->s4 = 1.2;
    p->s5 = 1.2;
    p->s6 = 1.98;
    p->G_NEAR_WALL_DISTANCE = 0.05;
    p->G_BIN_DENSITY = 1;
    p->G_BIN_LEVELS = 1;
    p->G_BIN_SPEC = 1;
    p->G_BIN_MAKE_LINKS = 0;
    p->S_BC_TYPE = 0;
    p->S_FORCE_TYPE = 0;
    p->S_FORCE_ORDER = 0;
    p->S_FORCEVOLUME_Xm = p->S_FORCEVOLUME_XM = 0.0;
    p->S_FORCEVOLUME_Ym = p->S_FORCEVOLUME_YM = 0.0;
    p->S_FORCEVOLUME_Zm = p->S_FORCEVOLUME_ZM = 0.0;
}

const char *agal_status_string(int s)
{
    switch (s) {
        case AGAL_OK: return "ok";
        case AGAL_ERR_ARG: return "invalid argument";
        case AGAL_ERR_HIP: return "HIP runtime error";
        case AGAL_ERR_ALLOC: return "allocation failed";
        case AGAL_ERR_CAPACITY: return "root grid exceeds the block pool";
        case AGAL_ERR_POOL_EXHAUSTED: return "refinement cancelled: block pool exhausted";
        case AGAL_ERR_STATE: return "call out of order";
        case AGAL_ERR_OVERFLOW: return "internal list overflow";
        case AGAL_ERR_NOT_BUILT: return "variant not built";
    }
    return "unknown status";
}

const char *agal_last_error(agal_ctx *ctx)
{
    if (!ctx) return "";
    return reinterpret_cast<Ctx *>(ctx)->last_error.c_str();
}

int agal_create(const agal_params *p, long long block_capacity, int device, agal_ctx **out)
{
    if (!p || !out || block_capacity <= 0) return AGAL_ERR_ARG;
    if (p->velocity_set != 9 && p->velocity_set != 19 && p->velocity_set != 27) return AGAL_ERR_ARG;
    if (p->precision != AGAL_FP32 && p->precision != AGAL_FP64 && p->precision != AGAL_FP16 &&
        p->precision != AGAL_BF16)
        return AGAL_ERR_ARG;
    if (p->collision != AGAL_BGK && p->collision != AGAL_MRT) return AGAL_ERR_NOT_BUILT;
    if (p->interp_order != AGAL_INTERP_LINEAR && p->interp_order != AGAL_INTERP_CUBIC) return AGAL_ERR_ARG;
    if (p->Nx <= 0 || p->Nx % 4 != 0) return AGAL_ERR_ARG;
    Ctx *c = new (std::nothrow) Ctx;
    if (!c) return AGAL_ERR_ALLOC;
    c->P = *p;
    c->device = device;
    c->Q = p->velocity_set;
    c->DIM = c->Q == 9 ? 2 : 3;
    c->QMAX = c->DIM == 2 ? 9 : 27;
    c->CELLS = c->DIM == 2 ? 16 : 64;
    c->NCH = c->DIM == 2 ? 4 : 8;
    c->precision = p->precision;
    c->csize = p->precision == AGAL_FP64 ? 8 : 4;
    c->tsize = (p->precision == AGAL_FP16 || p->precision == AGAL_BF16) ? 2 : c->csize;
    for (int q = 0; q < c->Q; q++)
        c->weight[q] = c->Q == 9 ? VelSet<9>::w(q) : (c->Q == 19 ? VelSet<19>::w(q) : VelSet<27>::w(q));
    float Lx = (float)p->L_c;
    float Ly = (float)(p->L_fy * Lx), Lz = (float)(p->L_fz * Lx);
    c->Nxi[0] = p->Nx;
    c->Nxi[1] = (int)(p->Nx * (Ly / Lx));
    c->Nxi[2] = c->DIM == 2 ? 1 : (int)(p->Nx * (Lz / Lx));
    c->MAX_LEVELS_WALL = p->MAX_LEVELS_WALL;
    c->MAX_LEVELS_INTERIOR = p->MAX_LEVELS_INTERIOR;
    c->N_LEVEL_START = p->N_LEVEL_START;
    c->MAX_LEVELS = std::max(p->MAX_LEVELS_WALL, p->MAX_LEVELS_INTERIOR);
    if (c->MAX_LEVELS == 1) c->N_LEVEL_START = 0;
    if (c->N_LEVEL_START > c->MAX_LEVELS - 1) c->N_LEVEL_START = c->MAX_LEVELS - 1;
    if (c->MAX_LEVELS < 1 || c->MAX_LEVELS > 8 || c->Nxi[1] % 4 || c->Nxi[2] % (c->DIM == 3 ? 4 : 1) || c->P.P_REFINE <= 0) {
        delete c;
        return AGAL_ERR_ARG;
    }
    c->n_maxcblocks = block_capacity;
    c->n_maxcells = block_capacity * c->CELLS;
    c->n_coarsecblocks = (c->Nxi[0] * c->Nxi[1] * c->Nxi[2]) / c->CELLS;
    if ((long long)c->Nxi[0] * c->Nxi[1] * c->Nxi[2] >= c->n_maxcells || c->n_maxcells >= (1LL << 31)) {
        delete c;
        return AGAL_ERR_CAPACITY;
    }
    if (p->precision == AGAL_FP64)
        compute_level_params<double>(c);
    else
        compute_level_params<float>(c);
    c->n_ids.assign(c->MAX_LEVELS + 1, 0);
    c->n_iface.assign(c->MAX_LEVELS + 1, 0);
    c->id_max.assign(c->MAX_LEVELS + 1, 0);
    c->cur.assign(c->MAX_LEVELS, 0);

    int st = AGAL_OK;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        st = hip_fail(c, e, "hipSetDevice");
        free_ctx(c);
        return st;
    }
    if ((st = scratch_lookup(c))) {
        free_ctx(c);
        return st;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int L = 0; L < c->MAX_LEVELS && e == hipSuccess; L++)
        e = hipStreamCreateWithFlags(&c->lstream[L], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->sched_start, hipEventDisableTiming);
    if (e != hipSuccess) {
        st = hip_fail(c, e, "hipStreamCreate");
        free_ctx(c);
        return st;
    }
    const long long nmax = c->n_maxcblocks, ncm = c->n_maxcells;
    const size_t ts = c->tsize, cs = c->csize;
    c->exc_cap = c->exc_cap_alloc = std::min<long long>(ncm, 1LL << 22);
#define ALLOC(ptr, bytes)                                        \
    if ((st = dalloc(c, &(ptr), (bytes))) != AGAL_OK) {          \
        free_ctx(c);                                             \
        return st;                                               \
    }
    ALLOC(c->F[0], (size_t)c->Q * ncm * ts);
    ALLOC(c->F[1], (size_t)c->Q * ncm * ts);
    ALLOC(c->M, (size_t)4 * ncm * cs);
    ALLOC(c->cells_mask, sizeof(int) * ncm);
    ALLOC(c->X, sizeof(float) * c->DIM * nmax);
    ALLOC(c->nbr, sizeof(int) * c->QMAX * nmax);
    ALLOC(c->nbr_child, sizeof(int) * c->QMAX * nmax);
    ALLOC(c->nbr_slot, sizeof(int) * c->QMAX * nmax);
    ALLOC(c->ref, sizeof(int) * nmax);
    ALLOC(c->level, sizeof(int) * nmax);
    ALLOC(c->onb, sizeof(int) * nmax);
    ALLOC(c->blk_mask, sizeof(int) * nmax);
    ALLOC(c->part, sizeof(int) * nmax);
    ALLOC(c->simple, sizeof(int) * nmax);
    ALLOC(c->id_set, sizeof(int) * c->MAX_LEVELS * nmax);
    ALLOC(c->order, sizeof(int) * c->MAX_LEVELS * nmax);
    ALLOC(c->iface, sizeof(int) * c->MAX_LEVELS * nmax);
    ALLOC(c->gap_set, sizeof(int) * nmax);
    ALLOC(c->map_stamp, sizeof(unsigned) * nmax);
    ALLOC(c->first_child, sizeof(int) * nmax);
    ALLOC(c->scan_a, sizeof(int) * nmax);
    ALLOC(c->scan_b, sizeof(int) * nmax);
    ALLOC(c->partials, sizeof(int) * ((nmax + 2047) / 2048 + 64));
    ALLOC(c->dscal, sizeof(int) * 64);
    // One exception list per level, and one more (slot MAX_LEVELS) that the first sub-step of each
    // finest-level pair writes for the second to apply as it loads.
    ALLOC(c->exc_idx, sizeof(int) * 3 * c->exc_cap * (c->MAX_LEVELS + 1));
    ALLOC(c->exc_val, cs * c->exc_cap * (c->MAX_LEVELS + 1));
    ALLOC(c->exc_count, sizeof(int) * 2 * (c->MAX_LEVELS + 1));
    ALLOC(c->exc_head, sizeof(int) * nmax);
    ALLOC(c->exc_next, sizeof(int) * 2 * c->exc_cap_alloc);
    if (c->P.S_FORCE_TYPE == 1) ALLOC(c->cv_F, cs * 6 * nmax);
#undef ALLOC
    e = hipHostMalloc((void **)&c->hscal, sizeof(int) * 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        st = hip_fail(c, e, "hipHostMalloc");
        free_ctx(c);
        return st;
    }
    std::memset(c->hscal, 0, sizeof(int) * 64);
    // Initial values (mesh_init.cu:284-304).
    auto fill = [&](int *p, long long n, int v) { return launch_fill(c, p, n, v); };
    if ((st = fill(c->cells_mask, ncm, CELL_INTERIOR)) || (st = fill(c->blk_mask, nmax, BLK_REGULAR)) ||
        (st = fill(c->nbr, c->QMAX * nmax, N_SKIPID)) || (st = fill(c->nbr_child, c->QMAX * nmax, N_SKIPID)) ||
        (st = fill(c->ref, nmax, REF_INACTIVE)) || (st = fill(c->level, nmax, 0)) || (st = fill(c->onb, nmax, 0)) ||
        (st = fill((int *)c->map_stamp, nmax, 0)) || (st = fill(c->exc_count, 2 * (c->MAX_LEVELS + 1), 0)) || (st = fill(c->exc_head, nmax, -1)) || (st = fill(c->part, nmax, 1)) ||
        (st = fill(c->simple, nmax, 0)) || (st = fill((int *)c->X, c->DIM * nmax, 0)) ||
        (c->cv_F && (st = fill((int *)c->cv_F, (long long)(cs / 4) * 6 * nmax, 0)))) {
        free_ctx(c);
        return st;
    }
    hipLaunchKernelGGL(k_init_gaps, dim3((unsigned)((nmax - c->n_coarsecblocks + 255) / 256 + 1)), dim3(256), 0,
                       c->stream, c->gap_set, nmax - c->n_coarsecblocks, nmax);
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        st = hip_fail(c, e, "init");
        free_ctx(c);
        return st;
    }
    if ((st = build_root(c))) {
        free_ctx(c);
        return st;
    }
    c->n_gaps = (int)(nmax - c->n_coarsecblocks);
    c->id_max[0] = c->id_max[c->MAX_LEVELS] = c->n_coarsecblocks;
    if ((st = rebuild_level_lists(c, c->n_coarsecblocks))) {
        free_ctx(c);
        return st;
    }
    *out = reinterpret_cast<agal_ctx *>(c);
    return AGAL_OK;
}

int agal_destroy(agal_ctx *ctx)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return AGAL_OK;
}

int agal_get_sizes(agal_ctx *ctx, agal_sizes *s)
{
    CTX_OR_ARG(ctx);
    if (!s) return AGAL_ERR_ARG;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    std::memset(s, 0, sizeof(*s));
    s->n_dim = c->DIM;
    s->n_q = c->Q;
    s->n_q_max = c->QMAX;
    s->cells_per_block = c->CELLS;
    s->n_children = c->NCH;
    s->max_levels = c->MAX_LEVELS;
    s->n_maxcblocks = c->n_maxcblocks;
    s->n_maxcells = c->n_maxcells;
    s->n_coarsecblocks = c->n_coarsecblocks;
    for (int L = 0; L < c->MAX_LEVELS; L++) s->n_ids[L] = c->n_ids[L];
    for (int L = 0; L <= c->MAX_LEVELS; L++) s->id_max[L] = c->id_max[L];
    s->n_gaps = c->n_gaps;
    return AGAL_OK;
}

int agal_sync(agal_ctx *ctx)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return check_exceptions(c);
}

int agal_set_ic(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    c->ic_done = true;
    return launch_set_ic(c, L);
}

int agal_advance(agal_ctx *ctx, int L0)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L0 < 0 || L0 >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    if (!c->ic_done) return AGAL_ERR_STATE;
    if (int st = usable(c)) return st;
    return step_once(c, L0);
}

int agal_collide_stream(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    if (int st = usable(c)) return st;
    return launch_collide_stream(c, L);
}

int agal_interpolate(agal_ctx *ctx, int L, int var)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS || (var != AGAL_INTERP_INTERFACE && var != AGAL_INTERP_ADDED)) return AGAL_ERR_ARG;
    if (L >= c->MAX_LEVELS - 1) return AGAL_OK;
    int st = launch_refresh(c, L);
    if (st) return st;
    return launch_interpolate(c, L, var);
}

int agal_average(agal_ctx *ctx, int L, int var)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS || var < 0 || var > 2) return AGAL_ERR_ARG;
    if (L >= c->MAX_LEVELS - 1) return AGAL_OK;
    int st = launch_refresh(c, L + 1);
    if (st) return st;
    return launch_average(c, L, var);
}

int agal_refresh(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    return launch_refresh(c, L);
}

int agal_ref_criteria(agal_ctx *ctx)
{
    CTX_OR_ARG(ctx);
    return launch_ref_criteria(reinterpret_cast<Ctx *>(ctx));
}

int agal_refine_coarsen(agal_ctx *ctx, int out_counts[4])
{
    CTX_OR_ARG(ctx);
    return refine_coarsen(reinterpret_cast<Ctx *>(ctx), out_counts);
}

int agal_refine_with_solution(agal_ctx *ctx, int out_counts[4])
{
    CTX_OR_ARG(ctx);
    return refine_with_solution_impl(reinterpret_cast<Ctx *>(ctx), out_counts);
}

int agal_refine_near_wall(agal_ctx *ctx)
{
    CTX_OR_ARG(ctx);
    return refine_near_wall_impl(reinterpret_cast<Ctx *>(ctx));
}

int agal_set_geometry(agal_ctx *ctx, const float *tri, long long n_tri)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    AGAL_HIP(c, hipSetDevice(c->device));
    return set_geometry(c, tri, n_tri);
}

int agal_voxelize(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    int st;
    if ((st = voxelize_level(c, L)) || (st = mark_near_wall(c, L))) return st;
    return AGAL_OK;
}

int agal_check_masks(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    if (!c->geo_tri) return AGAL_ERR_STATE;
    return check_masks_level(c, L);
}

int agal_forces(agal_ctx *ctx, double out[3])
{
    CTX_OR_ARG(ctx);
    if (!out) return AGAL_ERR_ARG;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->ic_done) return AGAL_ERR_STATE;
    return compute_forces(c, out);
}

int agal_download_links(agal_ctx *ctx, int *onb_solid, long long n_blocks, float *X_b, long long n_values,
                        long long info[3])
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (n_blocks < 0 || n_blocks > c->n_maxcblocks || n_values < 0) return AGAL_ERR_ARG;
    if (info) {
        info[0] = c->lk_n_solidb;
        info[1] = c->lk_corrections;
        info[2] = c->links_on ? 1 : 0;
    }
    if (onb_solid && n_blocks > 0) {
        if (c->lk_onb_solid)
            AGAL_HIP(c, hipMemcpy(onb_solid, c->lk_onb_solid, sizeof(int) * n_blocks, hipMemcpyDeviceToHost));
        else
            for (long long k = 0; k < n_blocks; k++) onb_solid[k] = -1;
    }
    if (X_b && n_values > 0) {
        if (!c->lk_X || n_values > 27 * c->lk_ncb) return AGAL_ERR_ARG;
        AGAL_HIP(c, hipMemcpy(X_b, c->lk_X, sizeof(float) * n_values, hipMemcpyDeviceToHost));
    }
    return AGAL_OK;
}

int agal_download_link_bins(agal_ctx *ctx, int *bin_start, long long n_start, int *faces, long long n_faces,
                            long long info[4])
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (info) {
        info[0] = c->lb_n_bins;
        info[1] = c->lb_n_ent;
        info[2] = c->lb_G;
        info[3] = c->lb_n_kept;
    }
    if (bin_start && n_start > 0) {
        if (!c->lb_start || n_start > c->lb_n_bins + 1) return AGAL_ERR_ARG;
        AGAL_HIP(c, hipMemcpy(bin_start, c->lb_start, sizeof(int) * n_start, hipMemcpyDeviceToHost));
    }
    if (faces && n_faces > 0) {
        if (!c->lb_faces || n_faces > c->lb_n_ent) return AGAL_ERR_ARG;
        AGAL_HIP(c, hipMemcpy(faces, c->lb_faces, sizeof(int) * n_faces, hipMemcpyDeviceToHost));
    }
    return AGAL_OK;
}

long long agal_icosphere(double cx, double cy, double cz, double radius, int subdivisions, double noise,
                         unsigned long long seed, float *out, long long capacity)
{
    return icosphere(cx, cy, cz, radius, subdivisions, noise, seed, out, capacity);
}

long long agal_load_stl(const char *path, float *out, long long capacity)
{
    if (!path) return -1;
    return load_stl_ascii(path, out, capacity);
}

int agal_predictor_init(agal_ctx *ctx, unsigned long long seed, int batch_capacity)
{
    CTX_OR_ARG(ctx);
    if (batch_capacity <= 0) return AGAL_ERR_ARG;
    return predictor_setup(reinterpret_cast<Ctx *>(ctx), seed, batch_capacity);
}

int agal_predictor_set(agal_ctx *ctx, const float *params)
{
    CTX_OR_ARG(ctx);
    if (!params) return AGAL_ERR_ARG;
    return predictor_set_weights(reinterpret_cast<Ctx *>(ctx), params);
}

int agal_predictor_get(agal_ctx *ctx, int which, float *out)
{
    CTX_OR_ARG(ctx);
    if (!out) return AGAL_ERR_ARG;
    return predictor_get(reinterpret_cast<Ctx *>(ctx), which, out);
}

int agal_predictor_forward(agal_ctx *ctx, const float *x, int n, float *z_out, float *p_out)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->pred.w) return AGAL_ERR_STATE;
    if (n < 0 || n > c->pred.batch_cap || (n > 0 && !x)) return AGAL_ERR_ARG;
    if (n == 0) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(c->pred.X, x, sizeof(float) * NFEAT * n, hipMemcpyHostToDevice, c->stream));
    int st = predictor_forward_dev(c, c->pred.X, n, c->pred.z3, c->pred.Y, nullptr, nullptr);
    if (st) return st;
    if (z_out) AGAL_HIP(c, hipMemcpyAsync(z_out, c->pred.z3, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    if (p_out) AGAL_HIP(c, hipMemcpyAsync(p_out, c->pred.Y, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_predictor_train_step(agal_ctx *ctx, const float *x, const float *y, int n, double lr, double pos_weight,
                              double *loss)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->pred.w) return AGAL_ERR_STATE;
    if (n <= 0 || n > c->pred.batch_cap || !x || !y) return AGAL_ERR_ARG;
    AGAL_HIP(c, hipMemcpyAsync(c->pred.X, x, sizeof(float) * NFEAT * n, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemcpyAsync(c->pred.Y, y, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    double l = 0;
    int st = predictor_train_dev(c, n, (float)lr, (float)pos_weight, &l);
    if (loss) *loss = l;
    return st;
}

int agal_predictor_features(agal_ctx *ctx, float *out, long long n_blocks)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->pred.w) return AGAL_ERR_STATE;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (n_blocks < 0 || n_blocks > c->n_maxcblocks || (n_blocks > 0 && !out)) return AGAL_ERR_ARG;
    int st = predictor_features(c);
    if (st) return st;
    const long long nc = std::min<long long>(n_blocks, idm);
    if (nc > 0)
        AGAL_HIP(c, hipMemcpyAsync(out, c->pred.feat, sizeof(float) * NFEAT * nc, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    if (n_blocks > nc) std::memset(out + NFEAT * nc, 0, sizeof(float) * NFEAT * (n_blocks - nc));
    return AGAL_OK;
}

int agal_criterion_labels(agal_ctx *ctx, int *out, long long n_blocks)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->pred.w) return AGAL_ERR_STATE;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (n_blocks < 0 || n_blocks > c->n_maxcblocks || (n_blocks > 0 && !out)) return AGAL_ERR_ARG;
    int st = criterion_labels(c);
    if (st) return st;
    const long long nc = std::min<long long>(n_blocks, idm);
    if (nc > 0) AGAL_HIP(c, hipMemcpyAsync(out, c->pred.label, sizeof(int) * nc, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    for (long long b = nc; b < n_blocks; b++) out[b] = -1;
    return AGAL_OK;
}

void agal_learn_opts_default(agal_learn_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->train_steps = 2000;
    o->infer_every = 25;
    o->batch = 4096;
    o->adam_steps = 8;
    o->lr = 1e-3;
    o->threshold = 0.5;
    o->pos_weight = 4.0;
    o->seed = 1;
    o->use_predictions = 1;
}

int agal_advance_loop_learned(agal_ctx *ctx, int i0, int n, const agal_learn_opts *opts, agal_learn_stats *stats)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!opts || !stats || n < 0 || opts->batch <= 0 || opts->infer_every <= 0 || opts->adam_steps < 0) return AGAL_ERR_ARG;
    if (!c->ic_done) return AGAL_ERR_STATE;
    if (int st = usable(c)) return st;
    return advance_loop_learned_impl(c, i0, n, opts, stats);
}

int agal_mark_uniform(agal_ctx *ctx, int L)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS) return AGAL_ERR_ARG;
    int n = c->n_ids[L];
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_mark_uniform, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->id_set_L(L), n, c->ref);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

int agal_mark_blocks(agal_ctx *ctx, const int *ids, int n, int mark)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (n < 0 || (n > 0 && !ids) || n > c->n_maxcblocks) return AGAL_ERR_ARG;
    for (int k = 0; k < n; k++)
        if (ids[k] < 0 || ids[k] >= c->n_maxcblocks) return AGAL_ERR_ARG;
    if (n == 0) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(c->scan_a, ids, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_mark_list, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->scan_a, n, c->ref, mark);
    AGAL_HIP(c, hipGetLastError());
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_freeze_refined(agal_ctx *ctx, int var)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int n = c->id_max[c->MAX_LEVELS];
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_freeze, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->ref, n, var);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

int agal_advance_loop(agal_ctx *ctx, int i0, int n, double *out_step_ms, double *out_rebuild_ms)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->ic_done) return AGAL_ERR_STATE;
    if (n < 0) return AGAL_ERR_ARG;
    if (int st = usable(c)) return st;
    EventPair ev;
    AGAL_HIP(c, hipEventCreate(&ev.e[0]));
    AGAL_HIP(c, hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    double ts = 0, tr = 0;
    int status = AGAL_OK, st;
    const int iter_mult = 1 << c->N_LEVEL_START;
    const int ML = c->MAX_LEVELS;
    AGAL_HIP(c, hipEventRecord(e0, c->stream));
    const bool adaptive = ML > 1 && ML != c->N_LEVEL_START + 1;
    auto rebuild_after = [&](int i) { return adaptive && (i + 1) % c->P.P_REFINE == 0 && i > 0; };
    for (int i = i0; i < i0 + n; i++) {
        if (iter_mult == 1) {
            // Up to graph_steps coarse steps with no rebuild between them run as one recorded graph.
            int k = 1;
            while (k < c->graph_steps && i + k < i0 + n && !rebuild_after(i + k - 1)) k++;
            if ((st = steps_once(c, c->N_LEVEL_START, k))) return st;
            i += k - 1;
        } else {
            for (int j = 0; j < iter_mult; j++)
                if ((st = step_once(c, c->N_LEVEL_START))) return st;
        }
        if (rebuild_after(i)) {
            float ms = 0;
            AGAL_HIP(c, hipEventRecord(e1, c->stream));
            AGAL_HIP(c, hipEventSynchronize(e1));
            AGAL_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            ts += ms;
            auto t0 = std::chrono::steady_clock::now();
            st = refine_with_solution_impl(c, nullptr);
            if (st == AGAL_ERR_POOL_EXHAUSTED)
                status = st;
            else if (st)
                return st;
            AGAL_HIP(c, hipStreamSynchronize(c->stream));
            tr += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            AGAL_HIP(c, hipEventRecord(e0, c->stream));
        }
    }
    {
        float ms = 0;
        AGAL_HIP(c, hipEventRecord(e1, c->stream));
        AGAL_HIP(c, hipEventSynchronize(e1));
        AGAL_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        ts += ms;
    }
    if ((st = check_exceptions(c))) return st;
    if (out_step_ms) *out_step_ms = ts;
    if (out_rebuild_ms) *out_rebuild_ms = tr;
    return status;
}

// ---- data exchange ---------------------------------------------------------------------------
// `fixed_stage`: a device staging area of at least Q * n_blocks * cells transfer values to use
// instead of a fresh allocation (the allocation-free download), else nullptr.
static int stage_pops(Ctx *c, void *host, long long n_blocks, long long stride, int dir, void *fixed_stage = nullptr)
{
    if (n_blocks < 0 || n_blocks > c->n_maxcblocks) return AGAL_ERR_ARG;
    if (n_blocks == 0) return AGAL_OK;
    const long long nc = n_blocks * c->CELLS;
    if (stride < nc) return AGAL_ERR_ARG;
    void *stage = fixed_stage;
    const size_t cs = c->csize;
    if (!stage) AGAL_HIP(c, hipMalloc(&stage, (size_t)c->Q * nc * cs));
    int cl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int L = 0; L < c->MAX_LEVELS; L++) cl[L] = c->cur[L];
    int4 lo = make_int4(cl[0], cl[1], cl[2], cl[3]), hi = make_int4(cl[4], cl[5], cl[6], cl[7]);
    Weights wt{};
    for (int q = 0; q < c->Q; q++) wt.w[q] = c->weight[q];
    int rc = AGAL_OK;
    auto body = [&]() -> int {
        const char *h = static_cast<const char *>(host);
        if (dir == 1)
            for (int q = 0; q < c->Q; q++)
                AGAL_HIP(c, hipMemcpyAsync((char *)stage + (size_t)q * nc * cs, h + (size_t)q * stride * cs, nc * cs,
                                           hipMemcpyHostToDevice, c->stream));
        const dim3 grid((unsigned)((nc + 255) / 256)), blk(256);
        switch (c->precision) {
            case AGAL_FP64:
                hipLaunchKernelGGL((k_stage_pops<double, double>), grid, blk, 0, c->stream, (double *)c->F[0],
                                   (double *)c->F[1], (double *)stage, c->n_maxcells, n_blocks, c->CELLS, c->Q, c->level,
                                   lo, hi, dir, wt);
                break;
            case AGAL_FP16:
                hipLaunchKernelGGL((k_stage_pops<float, _Float16>), grid, blk, 0, c->stream, (_Float16 *)c->F[0],
                                   (_Float16 *)c->F[1], (float *)stage, c->n_maxcells, n_blocks, c->CELLS, c->Q,
                                   c->level, lo, hi, dir, wt);
                break;
            case AGAL_BF16:
                hipLaunchKernelGGL((k_stage_pops<float, bf16s>), grid, blk, 0, c->stream, (bf16s *)c->F[0],
                                   (bf16s *)c->F[1], (float *)stage, c->n_maxcells, n_blocks, c->CELLS, c->Q, c->level,
                                   lo, hi, dir, wt);
                break;
            default:
                hipLaunchKernelGGL((k_stage_pops<float, float>), grid, blk, 0, c->stream, (float *)c->F[0],
                                   (float *)c->F[1], (float *)stage, c->n_maxcells, n_blocks, c->CELLS, c->Q, c->level,
                                   lo, hi, dir, wt);
        }
        AGAL_HIP(c, hipGetLastError());
        if (dir == 0)
            for (int q = 0; q < c->Q; q++)
                AGAL_HIP(c, hipMemcpyAsync((char *)host + (size_t)q * stride * cs, (char *)stage + (size_t)q * nc * cs,
                                           nc * cs, hipMemcpyDeviceToHost, c->stream));
        AGAL_HIP(c, hipStreamSynchronize(c->stream));
        return AGAL_OK;
    };
    rc = body();
    if (!fixed_stage) (void)hipFree(stage);
    return rc;
}

int agal_download_populations(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride)
{
    CTX_OR_ARG(ctx);
    if (!dst) return AGAL_ERR_ARG;
    return stage_pops(reinterpret_cast<Ctx *>(ctx), dst, n_blocks, dst_stride, 0);
}

int agal_download_populations_noalloc(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!dst || n_blocks < 0 || n_blocks > c->n_maxcblocks) return AGAL_ERR_ARG;
    const unsigned long long need = (unsigned long long)c->Q * (unsigned long long)n_blocks * c->CELLS * c->csize;
    if (need > AGAL_NOALLOC_SCRATCH_BYTES) {
        c->last_error = "agal_download_populations_noalloc: " + std::to_string(n_blocks) + " blocks need " +
                        std::to_string(need) + " staging bytes, above the fixed " +
                        std::to_string((unsigned long long)AGAL_NOALLOC_SCRATCH_BYTES);
        return AGAL_ERR_ARG;
    }
    if (!c->scratch) {
        c->last_error = "agal_download_populations_noalloc: no staging area on this device";
        return AGAL_ERR_STATE;
    }
    return stage_pops(c, dst, n_blocks, dst_stride, 0, c->scratch);
}

int agal_upload_populations(agal_ctx *ctx, const void *src, long long n_blocks, long long src_stride)
{
    CTX_OR_ARG(ctx);
    if (!src) return AGAL_ERR_ARG;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    c->ic_done = true;
    return stage_pops(c, const_cast<void *>(src), n_blocks, src_stride, 1);
}

int agal_download_macros(agal_ctx *ctx, void *dst, long long n_blocks, long long dst_stride)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!dst || n_blocks < 0 || n_blocks > c->n_maxcblocks) return AGAL_ERR_ARG;
    const long long nc = n_blocks * c->CELLS;
    if (dst_stride < nc) return AGAL_ERR_ARG;
    for (int k = 0; k < 4; k++)
        AGAL_HIP(c, hipMemcpyAsync((char *)dst + (size_t)k * dst_stride * c->csize,
                                   (char *)c->M + (size_t)k * c->n_maxcells * c->csize, nc * c->csize,
                                   hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_download_cell_masks(agal_ctx *ctx, int *dst, long long n_blocks)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!dst || n_blocks < 0 || n_blocks > c->n_maxcblocks) return AGAL_ERR_ARG;
    AGAL_HIP(c, hipMemcpyAsync(dst, c->cells_mask, sizeof(int) * n_blocks * c->CELLS, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_upload_cell_masks(agal_ctx *ctx, const int *src, long long n_blocks)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!src || n_blocks < 0 || n_blocks > c->n_maxcblocks) return AGAL_ERR_ARG;
    AGAL_HIP(c, hipMemcpyAsync(c->cells_mask, src, sizeof(int) * n_blocks * c->CELLS, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return rebuild_level_lists(c, 0);
}

static int *block_array(Ctx *c, int which, int &rows)
{
    rows = 1;
    switch (which) {
        case AGAL_BLK_NBR: rows = c->QMAX; return c->nbr;
        case AGAL_BLK_NBR_CHILD: rows = c->QMAX; return c->nbr_child;
        case AGAL_BLK_REF: return c->ref;
        case AGAL_BLK_LEVEL: return c->level;
        case AGAL_BLK_ONB: return c->onb;
        case AGAL_BLK_MASK: return c->blk_mask;
    }
    return nullptr;
}

int agal_download_block_ints(agal_ctx *ctx, int which, int *dst, long long n_blocks, long long dst_stride)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int rows;
    int *src = block_array(c, which, rows);
    if (!src || !dst || n_blocks < 0 || n_blocks > c->n_maxcblocks || (rows > 1 && dst_stride < n_blocks)) return AGAL_ERR_ARG;
    for (int r = 0; r < rows; r++)
        AGAL_HIP(c, hipMemcpyAsync(dst + r * dst_stride, src + r * c->n_maxcblocks, sizeof(int) * n_blocks,
                                   hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_upload_block_ints(agal_ctx *ctx, int which, const int *src, long long n_blocks, long long src_stride)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int rows;
    int *dstp = block_array(c, which, rows);
    if (!dstp || !src || n_blocks < 0 || n_blocks > c->n_maxcblocks || (rows > 1 && src_stride < n_blocks)) return AGAL_ERR_ARG;
    for (int r = 0; r < rows; r++)
        AGAL_HIP(c, hipMemcpyAsync(dstp + r * c->n_maxcblocks, src + r * src_stride, sizeof(int) * n_blocks,
                                   hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return rebuild_level_lists(c, 0);
}

int agal_download_coords(agal_ctx *ctx, float *dst, long long n_blocks, long long dst_stride)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!dst || n_blocks < 0 || n_blocks > c->n_maxcblocks || dst_stride < n_blocks) return AGAL_ERR_ARG;
    for (int d = 0; d < c->DIM; d++)
        AGAL_HIP(c, hipMemcpyAsync(dst + d * dst_stride, c->X + d * c->n_maxcblocks, sizeof(float) * n_blocks,
                                   hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_download_id_set(agal_ctx *ctx, int L, int *dst, int n)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (L < 0 || L >= c->MAX_LEVELS || n < 0 || n > c->n_ids[L] || (n > 0 && !dst)) return AGAL_ERR_ARG;
    if (n == 0) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(dst, c->id_set_L(L), sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_download_gap_set(agal_ctx *ctx, int *dst, int n)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (n < 0 || n > c->n_gaps || (n > 0 && !dst)) return AGAL_ERR_ARG;
    if (n == 0) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(dst, c->gap_set, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

int agal_import_pool(agal_ctx *ctx, const agal_pool *pl)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    const int ML = c->MAX_LEVELS, QM = c->QMAX;
    const long long nmax = c->n_maxcblocks;
    if (!pl || !pl->nbr || !pl->nbr_child || !pl->ref || !pl->level || !pl->onb || !pl->blk_mask || !pl->coords ||
        !pl->cell_mask || !pl->n_ids || (pl->n_gaps > 0 && !pl->gap_set))
        return AGAL_ERR_ARG;
    const long long nb = pl->n_blocks;
    if (nb < c->n_coarsecblocks || nb > nmax || pl->n_gaps < 0 || pl->n_gaps > nmax) return AGAL_ERR_ARG;
    long long total = 0;
    for (int L = 0; L < ML; L++) {
        if (pl->n_ids[L] < 0 || pl->n_ids[L] > nb) return AGAL_ERR_ARG;
        total += pl->n_ids[L];
    }
    if (total > 0 && !pl->id_sets) return AGAL_ERR_ARG;
    for (long long k = 0; k < total; k++)
        if (pl->id_sets[k] < 0 || pl->id_sets[k] >= nb) return AGAL_ERR_ARG;
    AGAL_HIP(c, hipSetDevice(c->device));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    int st;
    // IDs the context used beyond the imported range become inactive blocks without neighbours.
    const long long old = c->id_max[ML];
    if (old > nb) {
        if ((st = launch_fill(c, c->ref + nb, old - nb, REF_INACTIVE)) || (st = launch_fill(c, c->level + nb, old - nb, 0)) ||
            (st = launch_fill(c, c->onb + nb, old - nb, 0)) || (st = launch_fill(c, c->blk_mask + nb, old - nb, BLK_REGULAR)))
            return st;
        for (int p = 0; p < QM; p++)
            if ((st = launch_fill(c, c->nbr + p * nmax + nb, old - nb, N_SKIPID)) ||
                (st = launch_fill(c, c->nbr_child + p * nmax + nb, old - nb, N_SKIPID)))
                return st;
    }
    for (int p = 0; p < QM; p++) {
        AGAL_HIP(c, hipMemcpyAsync(c->nbr + p * nmax, pl->nbr + p * nb, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
        AGAL_HIP(c, hipMemcpyAsync(c->nbr_child + p * nmax, pl->nbr_child + p * nb, sizeof(int) * nb, hipMemcpyHostToDevice,
                                   c->stream));
    }
    AGAL_HIP(c, hipMemcpyAsync(c->ref, pl->ref, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemcpyAsync(c->level, pl->level, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemcpyAsync(c->onb, pl->onb, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemcpyAsync(c->blk_mask, pl->blk_mask, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    for (int d = 0; d < c->DIM; d++)
        AGAL_HIP(c, hipMemcpyAsync(c->X + d * nmax, pl->coords + d * nb, sizeof(float) * nb, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemcpyAsync(c->cells_mask, pl->cell_mask, sizeof(int) * nb * c->CELLS, hipMemcpyHostToDevice, c->stream));
    if (pl->n_gaps > 0)
        AGAL_HIP(c, hipMemcpyAsync(c->gap_set, pl->gap_set, sizeof(int) * pl->n_gaps, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    c->n_gaps = pl->n_gaps;
    for (int L = 1; L < ML; L++) c->id_max[L] = 0;
    c->id_max[0] = c->n_coarsecblocks;
    c->id_max[ML] = (int)nb;
    std::fill(c->cur.begin(), c->cur.end(), 0);
    c->pending_fix.clear();
    c->deferred.clear();
    if ((st = launch_fill(c, c->exc_count, 2 * (ML + 1), 0))) return st;
    if ((st = rebuild_level_lists(c, nb))) return st;
    if (pl->populations) {
        c->ic_done = true;
        if ((st = stage_pops(c, const_cast<void *>(pl->populations), nb, nb * c->CELLS, 1))) return st;
    }
    // The rebuilt lists must be the given ID sets.
    long long off = 0;
    for (int L = 0; L < ML; L++) {
        if (c->n_ids[L] != pl->n_ids[L]) {
            c->last_error = "agal_import_pool: level " + std::to_string(L) + " holds " + std::to_string(c->n_ids[L]) +
                            " active blocks by ref / level, the ID set " + std::to_string(pl->n_ids[L]);
            return AGAL_ERR_ARG;
        }
        std::vector<int> ids(c->n_ids[L]);
        if (!ids.empty())
            AGAL_HIP(c, hipMemcpy(ids.data(), c->id_set_L(L), sizeof(int) * ids.size(), hipMemcpyDeviceToHost));
        if (!std::equal(ids.begin(), ids.end(), pl->id_sets + off)) {
            c->last_error = "agal_import_pool: the ID set of level " + std::to_string(L) + " is not the ascending list of "
                            "its active blocks";
            return AGAL_ERR_ARG;
        }
        off += c->n_ids[L];
    }
    return AGAL_OK;
}

int agal_set_option(agal_ctx *ctx, int option, long long value)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (option == AGAL_OPT_GRAPHS) {
        c->use_graphs = value != 0;
        return AGAL_OK;
    }
    if (option == AGAL_OPT_MRT_MFMA) {
        c->mrt_mfma = value != 0;
        c->mesh_epoch++;  // recorded step graphs hold the other kernel
        return AGAL_OK;
    }
    if (option == AGAL_OPT_GRAPH_STEPS) {
        if (value < 1 || value > 64) return AGAL_ERR_ARG;
        c->graph_steps = (int)value;
        return AGAL_OK;
    }
    if (option == AGAL_OPT_LINK_MASKS) {
        c->link_masks = value != 0;
        return AGAL_OK;
    }
    if (option == AGAL_OPT_EXC_CAP) {
        // Entries per exception list, at most the allocated capacity.
        if (value < 1 || value > c->exc_cap_alloc) return AGAL_ERR_ARG;
        c->exc_cap = value;
        c->mesh_epoch++;  // recorded step graphs hold the old bound
        return AGAL_OK;
    }
    return AGAL_ERR_ARG;
}

int agal_timing_reset(agal_ctx *ctx, int enable)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int st = collect_timing(c);
    if (st) return st;
    c->timing.enabled = enable != 0;
    c->timing.fused_ms = 0;
    c->timing.fused_launches = 0;
    c->timing.fused_cells = 0;
    return AGAL_OK;
}

int agal_timing_get(agal_ctx *ctx, double *fused_ms_total, long long *fused_launches, double *fused_cells_total)
{
    CTX_OR_ARG(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int st = collect_timing(c);
    if (st) return st;
    if (fused_ms_total) *fused_ms_total = c->timing.fused_ms;
    if (fused_launches) *fused_launches = c->timing.fused_launches;
    if (fused_cells_total) *fused_cells_total = c->timing.fused_cells;
    return AGAL_OK;
}

}  // extern "C"
