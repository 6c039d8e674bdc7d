// agal_ctx.hpp -- device state of one mesh + LBM solver on one GPU.
//
// HBM layout (all structure-of-arrays, 256-byte aligned allocations):
//   F[0], F[1]   two population buffers, Q rows x n_maxcells (row q = direction q, physical order);
//                level L reads F[cur[L]] and writes F[cur[L]^1] in its fused step.
//   M            rho,u,v,w rows x n_maxcells, written only by explicit refresh calls.
//   cells_mask   n_maxcells ints.
//   nbr, nbr_child  Q_max rows x n_maxcblocks ints, the reference's velocity-ordered tables
//                (inc/mesh.h:397-399); nbr_slot is the same table in 3x3x3 slot order, one row of
//                Q_max ints per block, for the fused kernel's lane-parallel load.
//   ref, level, onb, blk_mask, part  n_maxcblocks ints; X  DIM rows x n_maxcblocks floats.
//   id_set       MAX_LEVELS rows x n_maxcblocks (sorted ascending per level), gap_set n_maxcblocks.
//   iface        per level: blocks of id_set[L] whose block mask is INTERFACE.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/agal_capi.h"

namespace agal {

// Learned refinement predictor: 12 features per block, a 12-64-64-1 network (flat parameter
// layout W1[12][64], b1[64], W2[64][64], b2[64], W3[64], b3).
constexpr int NFEAT = 12, MLP_H = 64;
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = NFEAT * MLP_H, MLP_OFF_W2 = MLP_OFF_B1 + MLP_H,
              MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H, MLP_OFF_W3 = MLP_OFF_B2 + MLP_H, MLP_OFF_B3 = MLP_OFF_W3 + MLP_H,
              MLP_NPARAM = MLP_OFF_B3 + 1;

struct Predictor {
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // master weights, Adam moments, gradient
    _Float16 *wh = nullptr;                                         // FP16 copy used by every pass
    int t = 0;                                                      // Adam steps taken
    float *stats = nullptr, *dist = nullptr, *feat = nullptr, *prob = nullptr;  // per block ID
    int *label = nullptr, *counts = nullptr;
    float *snap = nullptr;  // feature snapshots, snap_slots x n_maxcblocks x NFEAT
    int snap_slots = 0;
    int batch_cap = 0;      // rows of the training buffers below
    float *X = nullptr, *Y = nullptr, *Z1 = nullptr, *Z2 = nullptr, *z3 = nullptr, *dZ3 = nullptr, *dZ2 = nullptr,
          *dZ1 = nullptr, *loss = nullptr;
    int *ids = nullptr;
};

struct LevelTiming {
    bool enabled = false;
    double fused_ms = 0;
    long long fused_launches = 0;
    double fused_cells = 0;
    std::vector<hipEvent_t> ev_pairs;  // start/stop events of fused launches awaiting readout
};

struct Ctx {
    agal_params P{};
    int device = 0;
    void *scratch = nullptr;  // this GPU's copy of the allocation-free staging area (not owned)
    hipStream_t stream = nullptr;
    std::string last_error;

    // Geometry of the pool.
    int Q = 19, DIM = 3, QMAX = 27, CELLS = 64, NCH = 8;
    int precision = AGAL_FP32;
    size_t tsize = 4;   // bytes per stored population
    size_t csize = 4;   // bytes of the compute type (macroscopic fields, transfers)
    double weight[27] = {};
    long long n_maxcblocks = 0, n_maxcells = 0;
    int Nxi[3] = {0, 0, 0};
    int MAX_LEVELS = 1, MAX_LEVELS_INTERIOR = 1, MAX_LEVELS_WALL = 1, N_LEVEL_START = 0;
    int n_coarsecblocks = 0;
    double dx = 0, dx_cblock = 0;         // stored in the precision's type on the host side
    std::vector<double> dxf_vec, tau_vec; // exact values of the precision type, widened
    bool ic_done = false;

    // Host mirrors of the pool bookkeeping (inc/mesh.h:447-462).
    std::vector<int> n_ids, id_max, n_iface;
    int n_gaps = 0;
    std::vector<int> cur;  // current population buffer per level
    int global_Liter = 0;
    unsigned stamp = 0;    // generation counter of the rebuild's neighbourhood map

    // Device arrays.
    void *F[2] = {nullptr, nullptr};
    void *M = nullptr;
    int *cells_mask = nullptr;
    float *X = nullptr;
    int *nbr = nullptr, *nbr_child = nullptr, *nbr_slot = nullptr;
    int *ref = nullptr, *level = nullptr, *onb = nullptr, *blk_mask = nullptr, *part = nullptr;
    int *simple = nullptr;
    int *id_set = nullptr, *gap_set = nullptr, *iface = nullptr;
    // Per level, the same blocks as id_set in octree traversal order (the fused step's list).
    int *order = nullptr;
    // Rebuild scratch.
    unsigned *map_stamp = nullptr;
    int *first_child = nullptr, *scan_a = nullptr, *scan_b = nullptr, *partials = nullptr;
    int *dscal = nullptr;     // small device scalars (counters)
    int *hscal = nullptr;     // pinned host mirror
    // Exceptions of the stream guard: (initiator cell, partner cell, direction) + value, one list
    // of exc_cap entries per level.
    long long exc_cap = 0, exc_cap_alloc = 0;
    bool exc_overflowed = false;  // an overflow was reported: steps are refused
    int *exc_idx = nullptr;   // 3 ints per entry
    void *exc_val = nullptr;  // one value per entry
    int *exc_count = nullptr; // per level (and the extra slot MAX_LEVELS): [2L] count, [2L+1] overflow flag
    // The extra list (slot MAX_LEVELS) linked per block: chain heads per block ID (-1 when empty)
    // and two links per entry (through its initiator's and its partner's block).
    int *exc_head = nullptr, *exc_next = nullptr;

    // Triangle geometry: n_geo_tri triangles (9 floats each) binned on a G x G grid over the (y, z)
    // bounding box; bin (iy, iz) lists geo_bin_tri[geo_bin_start[iy + G iz] ...].
    float *geo_tri = nullptr;
    int *geo_bin_start = nullptr, *geo_bin_tri = nullptr;
    long long geo_n_tri = 0;
    int geo_G = 0;
    double geo_y0 = 0, geo_z0 = 0, geo_inv_dy = 0, geo_inv_dz = 0;

    // Face-cell links (G_BIN_MAKE_LINKS 1): bins of the link search, ranks of the blocks holding
    // boundary cells, the fractions [rank*64 + cell + q*lk_ncb] (27 directions), the linked blocks of
    // each level (lk_ids + lk_off[L], lk_n[L] of them) and the per-block momentum-exchange sums
    // (3 x n_maxcblocks values of the compute type).
    bool links_on = false;
    bool link_masks = false;  // links made to agree with the cell masks (AGAL_OPT_LINK_MASKS)
    int *lk_bcount = nullptr, *lk_onb_solid = nullptr, *lk_ids = nullptr, *lk_counter = nullptr;
    float *lk_X = nullptr;
    void *lk_blkF = nullptr;
    long long lk_n_solidb = 0, lk_ncb = 0, lk_corrections = 0;
    long long lk_off[8] = {};
    int lk_n[8] = {};
    int *lb_start = nullptr, *lb_faces = nullptr;
    int lb_G = 0, lb_level = 0;
    long long lb_n_bins = 0, lb_n_ent = 0, lb_n_kept = 0;

    // Control-volume force (S_FORCE_TYPE 1): the box snapped to root-block edges, in the precision's
    // type (widened), and the per-block sums (6 x n_maxcblocks values of the compute type: Fp, Fm
    // per direction).
    double cv_box[6] = {};
    void *cv_F = nullptr;

    Predictor pred;

    LevelTiming timing;

    // One coarse step recorded as a HIP graph, per starting buffer parity of the levels.  A graph
    // is valid for one mesh epoch (bumped whenever the level lists change).
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        unsigned cur_before = 0, cur_after = 0;
        unsigned epoch = 0;
        int nsteps = 1;  // coarse steps recorded in the graph
    };
    StepGraph graphs[8];
    unsigned mesh_epoch = 1;
    bool use_graphs = true;
    int graph_steps = 4;    // coarse steps per recorded graph at most (agal_advance_loop)
    bool mrt_mfma = false;  // MRT moment transform on matrix cores (FP32 compute, D3Q19 / D3Q27)

    // Scheduler used while a step is being recorded.  Resources are (level, buffer parity) pairs,
    // bit 2L+b; each kernel waits only for the last writer and the readers since of the resources
    // it touches.  Stream slot 0 is the context stream (the chain of the finest level's sub-cycle,
    // interpolation and averaging); slot 1+L hosts the fused step of a coarser level L, which can
    // then overlap the finer levels' sub-cycles in the recorded graph.
    struct ResDeps {
        int w_ev = -1, w_slot = -1;                // last writer (event, stream slot)
        std::vector<std::pair<int, int>> readers; // (event, stream slot) since the last write
    };
    static constexpr int NSLOT = 9, NRES = 16;
    bool sched_on = false;
    // Coarse fused sub-steps held back to share the next finest sub-step's launch: (level, buffer
    // read).  Only while a coarse step is being issued (defer_coarse).
    bool defer_coarse = false;
    std::vector<std::pair<int, int>> deferred;
    // Exception lists awaiting their fixup: (level, buffer the level's launch wrote).  Resolved
    // before anything else touches the level, or by workgroups of a fused launch that does not.
    std::vector<std::pair<int, int>> pending_fix;
    // 1 while the first, 2 while the second sub-step of a finest-level pair is issued, else 0.
    int fine_pair = 0;
    hipStream_t lstream[8] = {};
    bool slot_joined[NSLOT] = {};
    int seen[NSLOT][NSLOT] = {};  // latest event of slot t already waited for by slot s
    hipEvent_t sched_start = nullptr;
    std::vector<hipEvent_t> evpool;
    int ev_next = 0;
    ResDeps deps[NRES];

    int *id_set_L(int L) const { return id_set + (size_t)L * n_maxcblocks; }
    int *order_L(int L) const { return order + (size_t)L * n_maxcblocks; }
    int *iface_L(int L) const { return iface + (size_t)L * n_maxcblocks; }
};

// Records the HIP error text and returns AGAL_ERR_HIP.
int hip_fail(Ctx *c, hipError_t e, const char *what);
#define AGAL_HIP(c, call)                                   \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return ::agal::hip_fail((c), _e, #call); \
    } while (0)

// Stream for a kernel that reads resources `rmask` and writes `wmask` (bit 2L+b = buffer b of
// level L), placed on stream slot `slot`; inserts the waits it needs while recording.  Without a
// recording in progress this is the context stream.  sched_done() records the kernel's event.
inline unsigned res_bit(int L, int b) { return 1u << (2 * L + (b & 1)); }
hipStream_t sched_begin(Ctx *c, int slot, unsigned rmask, unsigned wmask);
int sched_done(Ctx *c, hipStream_t s, int slot, unsigned rmask, unsigned wmask);

// Kernel launchers (lbm_kernels.hip).
int launch_set_ic(Ctx *c, int L);
int launch_refresh(Ctx *c, int L);
int launch_collide_stream(Ctx *c, int L);
int launch_flush_deferred(Ctx *c);
int launch_interp_collide(Ctx *c, int L);
int launch_fix_pending(Ctx *c);
int launch_interpolate(Ctx *c, int L, int var);
int launch_average(Ctx *c, int L, int var);
// average(L) in the advance's mode followed by interpolate(L, INTERFACE), fused for linear
// interpolation (one launch, same values as the two kernels).
int launch_average_interpolate(Ctx *c, int L);

// Mesh kernels and the rebuild (amr_kernels.hip).
// `label` (may be null): per block ID, 1 when the criterion wants the block to have children, 0
// when not, -1 for inactive IDs (the learned predictor's training target).
int launch_ref_criteria(Ctx *c, int *label = nullptr);
int marks_refine_only(Ctx *c);
int count_refine_marks(Ctx *c, int *n);
int marks_dilate(Ctx *c);
int refine_coarsen(Ctx *c, int out_counts[4]);
int rebuild_level_lists(Ctx *c, long long range_hint);
int launch_fill(Ctx *c, int *p, long long n, int v);

// Geometry (geometry.hip).
int set_geometry(Ctx *c, const float *tri, long long n_tri);
void free_geometry(Ctx *c);
int voxelize_level(Ctx *c, int L);
int near_wall_passes(double near_wall_distance, int L, double dx_L);
int mark_near_wall(Ctx *c, int L);
int check_masks_level(Ctx *c, int L);
int compute_forces(Ctx *c, double out[3]);
void free_links(Ctx *c);
int links_reset_counts(Ctx *c);
int compute_links(Ctx *c);
int links_rebuild_lists(Ctx *c);
int launch_mea_post(Ctx *c);
// Control-volume force sums of level L, post 0 (start of a coarse step) or 1 (end).
int launch_cv_force(Ctx *c, int L, int post);
long long icosphere(double cx, double cy, double cz, double r, int subdiv, double noise, unsigned long long seed,
                    float *out, long long cap);
long long load_stl_ascii(const char *path, float *out, long long cap);

// Learned refinement predictor (predictor.hip).
void free_predictor(Ctx *c);
void predictor_init_weights(unsigned long long seed, float *w);
int predictor_setup(Ctx *c, unsigned long long seed, int batch_cap);
int predictor_set_weights(Ctx *c, const float *w_host);
int predictor_get(Ctx *c, int which, float *out);
int predictor_forward_dev(Ctx *c, const float *X, int n, float *z3, float *p, float *Z1, float *Z2);
int predictor_train_dev(Ctx *c, int n, float lr, float pos_weight, double *loss);
int predictor_features(Ctx *c);
int predictor_predict_blocks(Ctx *c);
int predictor_mark(Ctx *c, float thr, int *n_marked);
int predictor_keep(Ctx *c, float thr);
int predictor_audit(Ctx *c, int out[2]);
int predictor_snapshot(Ctx *c, int slot, int max_slots);
int predictor_train_snapshots(Ctx *c, int n_snap, int batch, int steps, float lr, float pos_weight, uint64_t &rng,
                              double *last_loss);

}  // namespace agal
