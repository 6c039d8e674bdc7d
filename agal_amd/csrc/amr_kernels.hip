// amr_kernels.hip -- refinement criterion and the refine/coarsen rebuild of the block pool on
// gfx950.
//
// The reference runs the rebuild (src/mesh/mesh_amr.cu:2161-2744) as ~20 Thrust calls plus
// hand-written kernels with a host round-trip after every count.  Here every list the rebuild
// needs (blocks marked for refinement in ID order, removed children grouped by level, the sorted
// per-level ID sets, the interface lists) comes from one stable compaction built on wavefront
// prefix sums (__shfl_up inside a 64-lane wave, LDS across the four waves of a workgroup, one
// small kernel across workgroups).  The bookkeeping -- which gap-stack IDs become which children,
// how the gap stack and ID sets are updated, which blocks get which refinement states, masks and
// neighbour lists -- is the reference's, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

namespace agal {

constexpr int AWG = 256;          // threads per workgroup of the 1-D kernels
constexpr int ITEMS = 8;          // elements per thread in the compaction
constexpr int CHUNK = AWG * ITEMS;

static inline int grid1(long long n) { return (int)std::max<long long>(1, (n + AWG - 1) / AWG); }

__global__ void k_fill(int *p, long long n, int v)
{
    long long i = (long long)blockIdx.x * AWG + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill(Ctx *c, int *p, long long n, int v)
{
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_fill, dim3(grid1(n)), dim3(AWG), 0, c->stream, p, n, v);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// ---------------------------------------------------------------------------------------------
// Stable compaction.  Pred: bool operator()(long long i) over an index range; Emit:
// void operator()(long long i, int rank).  `base` (device) is added to every rank and receives
// the running total, so consecutive compactions can append to one output.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int x)
{
    const int lane = threadIdx.x % 64;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across the workgroup; returns the exclusive prefix and
// the workgroup total.
__device__ __forceinline__ int wg_excl_scan(int x, int &total)
{
    __shared__ int s_w[AWG / 64];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    int incl = wave_incl_scan(x);
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < AWG / 64; k++) {
        if (k < wv) off += s_w[k];
        total += s_w[k];
    }
    __syncthreads();
    return off + incl - x;
}

template <class Pred>
__global__ __launch_bounds__(AWG) void k_count(Pred pred, long long n, int *partials)
{
    const long long b0 = (long long)blockIdx.x * CHUNK + (long long)threadIdx.x * ITEMS;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        if (b0 + k < n && pred(b0 + k)) cnt++;
    int total;
    wg_excl_scan(cnt, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the partial counts, offset by *base; *base += total.
__global__ __launch_bounds__(AWG) void k_scan_partials(int *partials, int nparts, int *base)
{
    int run = *base;
    for (int s = 0; s < nparts; s += AWG) {
        int i = s + threadIdx.x;
        int x = (i < nparts) ? partials[i] : 0;
        int total;
        int ex = wg_excl_scan(x, total);
        if (i < nparts) partials[i] = run + ex;
        run += total;
    }
    __syncthreads();
    if (threadIdx.x == 0) *base = run;
}

template <class Pred, class Emit>
__global__ __launch_bounds__(AWG) void k_emit(Pred pred, long long n, const int *partials, Emit emit)
{
    const long long b0 = (long long)blockIdx.x * CHUNK + (long long)threadIdx.x * ITEMS;
    unsigned flags = 0;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        if (b0 + k < n && pred(b0 + k)) {
            flags |= 1u << k;
            cnt++;
        }
    int total;
    int r = partials[blockIdx.x] + wg_excl_scan(cnt, total);
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        if (flags & (1u << k)) emit(b0 + k, r++);
}

template <class Pred, class Emit>
static int compact(Ctx *c, Pred pred, long long n, Emit emit, int *base)
{
    if (n <= 0) return AGAL_OK;
    int nparts = (int)((n + CHUNK - 1) / CHUNK);
    hipLaunchKernelGGL((k_count<Pred>), dim3(nparts), dim3(AWG), 0, c->stream, pred, n, c->partials);
    AGAL_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(AWG), 0, c->stream, c->partials, nparts, base);
    AGAL_HIP(c, hipGetLastError());
    hipLaunchKernelGGL((k_emit<Pred, Emit>), dim3(nparts), dim3(AWG), 0, c->stream, pred, n, c->partials, emit);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// Predicates and emitters.
struct PredActiveLevel {
    const int *ref, *level;
    int L;
    __device__ bool operator()(long long i) const
    {
        int r = ref[i];
        return r != REF_INACTIVE && r != REF_REMOVE && level[i] == L;
    }
};
struct PredIfaceInList {
    const int *list, *blk_mask;
    __device__ bool operator()(long long i) const { return blk_mask[list[i]] == BLK_INTERFACE; }
};
struct PredRefEq {
    const int *ref;
    int v;
    __device__ bool operator()(long long i) const { return ref[i] == v; }
};
struct PredRemoveLevel {
    const int *ref, *level;
    int L;
    __device__ bool operator()(long long i) const { return ref[i] == REF_REMOVE && level[i] == L; }
};
struct EmitIndex {
    int *out;
    __device__ void operator()(long long i, int r) const { out[r] = (int)i; }
};
struct EmitListEntry {
    const int *list;
    int *out;
    __device__ void operator()(long long i, int r) const { out[r] = list[i]; }
};
// Blocks of a list whose children are active blocks of level L (the traversal order below).
struct PredActiveChildren {
    const int *list, *nbr_child, *ref, *level;
    int L;
    __device__ bool operator()(long long i) const
    {
        const int fc = nbr_child[list[i]];
        if (fc < 0) return false;
        const int r = ref[fc];
        return r != REF_INACTIVE && r != REF_REMOVE && level[fc] == L;
    }
};
struct EmitChildren {
    const int *list, *nbr_child;
    int *out;
    int nch;
    __device__ void operator()(long long i, int r) const
    {
        const int fc = nbr_child[list[i]];
        for (int k = 0; k < nch; k++) out[(long long)nch * r + k] = fc + k;
    }
};
// first_child[kap] = gap_set[n_gaps-1-NCH*rank]: the reversed tail of the gap stack, every NCH-th
// entry (thrust::reverse_copy + Cu_ContractByFrac + scatter, mesh_amr.cu:2334-2352).
struct EmitFirstChild {
    const int *gap_set;
    int *first_child;
    int n_gaps, nch;
    __device__ void operator()(long long i, int r) const { first_child[i] = gap_set[n_gaps - 1 - nch * r]; }
};

// ---------------------------------------------------------------------------------------------
// Small per-block kernels of the rebuild (mesh_amr.cu:104-1923).
// ---------------------------------------------------------------------------------------------
__global__ void k_count_marks(const int *ref, int n, int *out)
{
    int i = blockIdx.x * AWG + threadIdx.x;
    int r = (i < n) ? ref[i] : -1;
    unsigned long long br = __ballot(r == REF_MARK_REFINE), bc = __ballot(r == REF_MARK_COARSEN);
    if (threadIdx.x % 64 == 0) {
        if (br) atomicAdd(&out[0], __popcll(br));
        if (bc) atomicAdd(&out[1], __popcll(bc));
    }
}

__global__ void k_cancel(int *ref, int n)
{
    int i = blockIdx.x * AWG + threadIdx.x;
    if (i < n && ref[i] == REF_MARK_REFINE) ref[i] = REF_UNREFINED;
}

// Cu_RefineCells_Prep_S1/S2 + unique/scatter: marked blocks and their neighbours.
__global__ void k_map_mark(const int *ref, const int *nbr, long long nmax, int qmax, int n, unsigned *map, unsigned stamp)
{
    int k = blockIdx.x * AWG + threadIdx.x;
    if (k >= n) return;
    int r = ref[k];
    if (r != REF_MARK_REFINE && r != REF_MARK_COARSEN) return;
    int self = nbr[k];
    if (self < 0) return;
    map[self] = stamp;
    for (int p = 1; p < qmax; p++) {
        int nb = nbr[self + p * nmax];
        if (nb >= 0) map[nb] = stamp;
    }
}

// max over gap_set[lo, hi) into out[0] (atomicMax; out preset to -1).
__global__ void k_max_range(const int *a, int lo, int hi, int *out)
{
    int i = lo + blockIdx.x * AWG + threadIdx.x;
    if (i < hi) atomicMax(out, a[i]);
}

// Cu_RefineCells_S2_1 (mesh_amr.cu:1600-1662).
__global__ void k_s2_1(const int *ref, const int *nbr, int *nbr_child, const unsigned *map, unsigned stamp,
                       const int *first_child, long long nmax, int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || map[kap] != stamp) return;
    for (int p = 0; p < qmax; p++) {
        int ip = nbr[kap + p * nmax];
        if (ip > -1 && ref[ip] == REF_MARK_REFINE) nbr_child[kap + p * nmax] = first_child[ip];
    }
}

// Cu_RefineCells_Q1_1 (mesh_amr.cu:1399-1424).
__global__ void k_q1_1(int *ref, const int *nbr_child, long long nmax, int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || ref[kap] != REF_UNREFINED) return;
    bool has = false;
    for (int p = 1; p < qmax; p++)
        if (nbr_child[kap + p * nmax] >= 0) has = true;
    if (has) ref[kap] = REF_UNREFINED_VIO;
}

// Cu_RefineCells_Q1_2 (mesh_amr.cu:1427-1488).  Children of a block marked for coarsening are
// never themselves marked for coarsening, so reading their states while parents are reverted
// is race-free.
__global__ void k_q1_2(int *ref, const int *nbr_child, int nch, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || ref[kap] != REF_MARK_COARSEN) return;
    int ic = nbr_child[kap];
    if (ic < 0) return;  // a coarsening mark on a block without children is ignored
    bool viol = false;
    for (int xc = 0; xc < nch; xc++) {
        int r = ref[ic + xc];
        if (r == REF_UNREFINED_VIO || r == REF_MARK_REFINE) viol = true;
    }
    if (viol) ref[kap] = REF_REFINED;
}

// Cu_RefineCells_S2_2 (mesh_amr.cu:1665-1728).
__global__ void k_s2_2(const int *ref, const int *nbr, int *nbr_child, const unsigned *map, unsigned stamp,
                       long long nmax, int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || map[kap] != stamp) return;
    for (int p = 1; p < qmax; p++) {
        int ip = nbr[kap + p * nmax];
        if (ip > -1 && ref[ip] == REF_MARK_COARSEN) nbr_child[kap + p * nmax] = N_SKIPID;
    }
}

// Cu_AddRemoveBlocks (mesh_amr.cu:1493-1597).  One thread per parent; children IDs are free
// gap-stack IDs, so no other thread reads their states as marks.
template <typename T>
__global__ void k_add_remove(int *ref, int *level, int *nbr_child, float *X, const int *first_child, long long nmax,
                             int dim, int nch, T dx_cblock, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    int r = ref[kap];
    if (r == REF_MARK_REFINE) {
        int fc = first_child[kap];
        nbr_child[kap] = fc;
        int lq = level[kap] + 1;
        T xs = (T)X[kap], ys = (T)X[kap + nmax], zs = dim == 3 ? (T)X[kap + 2 * nmax] : (T)0;
        T h = dx_cblock / (T)(1 << lq);
        for (int c = 0; c < nch; c++) {
            int id = fc + c;
            level[id] = lq;
            ref[id] = REF_NEW;
            X[id] = (float)(xs + (T)(c % 2) * h);
            X[id + nmax] = (float)(ys + (T)((c / 2) % 2) * h);
            if (dim == 3) X[id + 2 * nmax] = (float)(zs + (T)((c / 4) % 2) * h);
            nbr_child[id] = N_SKIPID;
        }
    } else if (r == REF_MARK_COARSEN) {
        int fc = nbr_child[kap];
        nbr_child[kap] = N_SKIPID;
        if (fc >= 0)
            for (int c = 0; c < nch; c++) ref[fc + c] = REF_REMOVE;
    }
}

// Cu_ConcatReverse (mesh_amr.cu:38-48): gap_set[at + k] = rem[nrem-1-k].
__global__ void k_gap_append(int *gap_set, int at, const int *rem, const int *nrem_ptr)
{
    int k = blockIdx.x * AWG + threadIdx.x;
    int nrem = *nrem_ptr;
    if (k < nrem) gap_set[at + k] = rem[nrem - 1 - k];
}

// Cu_CoarsenCells_S2 (mesh_amr.cu:1859-1885).
__global__ void k_coarsen_s2(const int *ref, int *nbr, int *nbr_child, long long nmax, int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || ref[kap] != REF_REMOVE) return;
    for (int p = 0; p < qmax; p++) {
        if (p > 0) nbr[kap + p * nmax] = N_SKIPID;
        nbr_child[kap + p * nmax] = N_SKIPID;
    }
}

// Cu_UpdateMasks_1 (mesh_amr.cu:946-979).
__global__ void k_update_masks_1(int *blk_mask, const int *ref, const int *nbr_child, long long nmax, int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n || ref[kap] == REF_INACTIVE) return;
    int m = BLK_REGULAR;
    if (nbr_child[kap] > -1)
        for (int p = 1; p < qmax; p++)
            if (nbr_child[kap + p * nmax] == N_SKIPID) m = BLK_INTERFACE;
    if (blk_mask[kap] > -1) blk_mask[kap] = m;
}

// Cu_RefineCells_S3 (mesh_amr.cu:1730-1753).
__global__ void k_s3(int *ref, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    int r = ref[kap];
    if (r == REF_MARK_REFINE) ref[kap] = REF_REFINED;
    if (r == REF_NEW || r == REF_MARK_COARSEN) ref[kap] = REF_UNREFINED;
    if (r == REF_REMOVE) ref[kap] = REF_INACTIVE;
    if (r == REF_UNREFINED_VIO) ref[kap] = REF_UNREFINED;
}

// Cu_RefineCells_S4 (mesh_amr.cu:1772-1833).  A child read while its own state flips between
// REFINED and REFINED_WCHILD counts as refined either way, so the in-place update is exact.
__global__ void k_s4(int *ref, const int *nbr_child, int nch, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    int r = ref[kap];
    if (r != REF_REFINED && r != REF_REFINED_WCHILD) return;
    int ic = nbr_child[kap];
    if (ic < 0) return;
    int nr = REF_REFINED;
    for (int xc = 0; xc < nch; xc++) {
        int rc = ref[ic + xc];
        if (rc == REF_REFINED || rc == REF_REFINED_WCHILD) nr = REF_REFINED_WCHILD;
    }
    ref[kap] = nr;
}

// Cu_UpdateConnectivity (mesh_amr.cu:154-944): the neighbour of child c in direction p is the
// child of the parent-level block containing that position, read from the parent's neighbour-
// child table (negative entries -- domain faces, N_SKIPID -- are copied).  One thread per
// (parent, child, direction).
__global__ void k_update_connectivity(int *nbr, const int *nbr_child, const unsigned *map, unsigned stamp,
                                      long long nmax, int dim, int qmax, int nch, int n)
{
    long long gid = (long long)blockIdx.x * AWG + threadIdx.x;
    long long per = (long long)nch * qmax;
    if (gid >= (long long)n * per) return;
    int kap = (int)(gid / per);
    int rem = (int)(gid % per);
    int c = rem / qmax, p = rem % qmax;
    if (map[kap] != stamp) return;
    int fc = nbr_child[kap];
    if (fc < 0) return;
    int cc[3] = {c % 2, (c / 2) % 2, (c / 4) % 2};
    int cv[3] = {dim == 3 ? c3x(p) : c2x(p), dim == 3 ? c3y(p) : c2y(p), dim == 3 ? c3z(p) : 0};
    int d[3], o[3];
    for (int a = 0; a < 3; a++) {
        int tg = cc[a] + cv[a];
        d[a] = tg < 0 ? -1 : (tg > 1 ? 1 : 0);
        o[a] = (tg + 2) % 2;
    }
    int pp = dim == 3 ? dir3(d[0], d[1], d[2]) : dir2(d[0], d[1]);
    int src = nbr_child[kap + pp * nmax];
    nbr[(fc + c) + p * nmax] = src < 0 ? src : src + o[0] + 2 * o[1] + 4 * o[2];
}

// Cu_UpdateBoundaries (mesh_amr.cu:104-152).
__global__ void k_update_boundaries(const int *nbr, int *nbr_child, const int *blk_mask, int *onb, long long nmax,
                                    int qmax, int n)
{
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    int on_b = 0;
    for (int p = 1; p < qmax; p++) {
        int nb = nbr[kap + p * nmax];
        if (nb < 0) {
            nbr_child[kap + p * nmax] = nb;
            if (nb != N_SKIPID) on_b = 1;
        }
    }
    if (blk_mask[kap] == BLK_SOLIDA) on_b = 1;
    onb[kap] = on_b;
}

// Cu_UpdateMasks_2 (mesh_amr.cu:981-1297) with one sub-block per block: interface (1) if any
// neighbour is missing, ghost (2) in the half of the block facing a missing neighbour.
template <int DIM>
__global__ void k_update_masks_2(int *cells_mask, const int *nbr, const int *ref, long long nmax, int n)
{
    constexpr int CELLS = DIM == 3 ? 64 : 16, QMAX = DIM == 3 ? 27 : 9;
    long long gid = (long long)blockIdx.x * AWG + threadIdx.x;
    if (gid >= (long long)n * CELLS) return;
    int kap = (int)(gid / CELLS), t = (int)(gid % CELLS);
    if (ref[kap] == REF_INACTIVE) return;
    int m = cells_mask[gid];
    if (m < 0) return;
    int ijk[3] = {t % 4, (t / 4) % 4, t / 16};
    m = 0;
    for (int p = 1; p < QMAX; p++) {
        if (nbr[kap + p * nmax] != N_SKIPID) continue;
        if (m < 1) m = 1;
        int cv[3] = {DIM == 3 ? c3x(p) : c2x(p), DIM == 3 ? c3y(p) : c2y(p), DIM == 3 ? c3z(p) : 0};
        bool outer = true;
        for (int a = 0; a < DIM; a++) {
            if (cv[a] == 1 && !(ijk[a] >= 2)) outer = false;
            if (cv[a] == -1 && !(ijk[a] < 2)) outer = false;
        }
        if (outer) m = 2;
    }
    cells_mask[gid] = m;
}

// Slot-ordered neighbour table, participation flag and solid-cell flag per block.
template <int DIM>
__global__ void k_block_aux(const int *nbr, const int *nbr_child, const int *blk_mask, const int *cells_mask,
                            int *nbr_slot, int *part, int *has_solid, long long nmax, int n)
{
    constexpr int CELLS = DIM == 3 ? 64 : 16, QMAX = DIM == 3 ? 27 : 9;
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    for (int S = 0; S < QMAX; S++) {
        int sx = S % 3 - 1, sy = (S / 3) % 3 - 1, sz = S / 9 - 1;
        int p = DIM == 3 ? dir3(sx, sy, sz) : dir2(sx, sy);
        nbr_slot[(long long)kap * QMAX + S] = nbr[kap + p * nmax];
    }
    part[kap] = (nbr_child[kap] < 0 || blk_mask[kap] != BLK_SOLID) ? 1 : 0;
    int hs = 0;
    for (int t = 0; t < CELLS; t++)
        if (cells_mask[(long long)kap * CELLS + t] == CELL_SOLID) hs = 1;
    has_solid[kap] = hs;
}

// simple: the block and every existing neighbour participate and hold no solid cell (the fused
// step then skips the guard reads).  2: in addition every neighbour slot that the velocity set
// links to exists (so no link leaves the domain or the level) and slot 13 is the block itself --
// the fused step's interior path.  max_nz: most nonzero components of a link (2 for D3Q19).
template <int DIM>
__global__ void k_block_simple(const int *nbr_slot, const int *part, const int *has_solid, int *simple, int n,
                               int max_nz)
{
    constexpr int QMAX = DIM == 3 ? 27 : 9;
    int kap = blockIdx.x * AWG + threadIdx.x;
    if (kap >= n) return;
    int s = part[kap] && !has_solid[kap];
    bool all = true;
    for (int S = 0; S < QMAX && s; S++) {
        int nb = nbr_slot[(long long)kap * QMAX + S];
        if (nb >= 0 && (!part[nb] || has_solid[nb])) s = 0;
        const int nz = (S % 3 != 1) + ((S / 3) % 3 != 1) + (DIM == 3 ? (S / 9 != 1) : 0);
        if (nb < 0 && nz <= max_nz) all = false;
    }
    if (s && all && nbr_slot[(long long)kap * QMAX + (QMAX - 1) / 2] == kap) s = 2;
    simple[kap] = s;
}

// ---------------------------------------------------------------------------------------------
// Refinement criterion: Cu_ComputeRefCriteria (solver_lbm_criterion.cu:438-592).
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T floor_log2_exact(T x)
{
    if (x != x) return x;
    if (x == (T)0) return -INFINITY;
    if (__builtin_isinf(x)) return x;
    if constexpr (sizeof(T) == 4)
        return (T)ilogbf(x);
    else
        return (T)ilogb(x);
}

template <typename T>
__device__ __forceinline__ T ex4(T d1, T d2, T d3, T d4)
{
    return (T)(+4.0) * d1 + (T)(-6.0) * d2 + (T)(+4.0) * d3 + (T)(-1.0) * d4;
}

template <typename T, int DIM>
__global__ __launch_bounds__(AWG) void k_criterion(int id_max, long long ncm, const T *__restrict__ M,
                                                   int *ref, const int *__restrict__ level,
                                                   const int *__restrict__ nbr, const int *__restrict__ cells_mask,
                                                   long long nmax, T dx_L, int crit, T NRS, T NRI, T NRM, int MLI,
                                                   int *__restrict__ label)
{
    constexpr int CELLS = DIM == 3 ? 64 : 16, QMAX = DIM == 3 ? 27 : 9, HB = DIM == 3 ? 216 : 36;
    constexpr int BPG = AWG / CELLS;
    __shared__ T s_u[BPG * HB], s_v[BPG * HB], s_w[BPG * HB];
    const int t = threadIdx.x % CELLS, g = threadIdx.x / CELLS;
    const int b = blockIdx.x * BPG + g;
    bool work = b < id_max;
    if (work) work = ref[b] != REF_INACTIVE;
    const int I = t % 4, J = (t / 4) % 4, K = DIM == 3 ? t / 16 : 0;
    T *su = s_u + g * HB, *sv = s_v + g * HB, *sw = s_w + g * HB;
    auto H = [](int i, int j, int k) { return DIM == 2 ? (i + 1) + 6 * (j + 1) : (i + 1) + 6 * (j + 1) + 36 * (k + 1); };
    if (work) {
        long long cell = (long long)b * CELLS + t;
        su[H(I, J, K)] = M[1 * ncm + cell];
        sv[H(I, J, K)] = M[2 * ncm + cell];
        sw[H(I, J, K)] = DIM == 3 ? M[3 * ncm + cell] : (T)0;
    }
    __syncthreads();
    // ExtrapolateToHalo (inc/util_math.h:54-112), applied to u, then v, then w.
    T *arrs[3] = {su, sv, sw};
    for (int a = 0; a < (DIM == 3 ? 3 : 2); a++) {
        T *s = arrs[a];
        if (work) {
            if (I == 0) {
                s[H(-1, J, K)] = ex4(s[H(0, J, K)], s[H(1, J, K)], s[H(2, J, K)], s[H(3, J, K)]);
                s[H(4, J, K)] = ex4(s[H(3, J, K)], s[H(2, J, K)], s[H(1, J, K)], s[H(0, J, K)]);
            }
            if (J == 0) {
                s[H(I, -1, K)] = ex4(s[H(I, 0, K)], s[H(I, 1, K)], s[H(I, 2, K)], s[H(I, 3, K)]);
                s[H(I, 4, K)] = ex4(s[H(I, 3, K)], s[H(I, 2, K)], s[H(I, 1, K)], s[H(I, 0, K)]);
            }
            if (DIM == 3 && K == 0) {
                s[H(I, J, -1)] = ex4(s[H(I, J, 0)], s[H(I, J, 1)], s[H(I, J, 2)], s[H(I, J, 3)]);
                s[H(I, J, 4)] = ex4(s[H(I, J, 3)], s[H(I, J, 2)], s[H(I, J, 1)], s[H(I, J, 0)]);
            }
        }
    }
    __syncthreads();
    T W = (T)0;
    if (work) {
        T uX = 0, uY = 0, uZ = 0, vX = 0, vY = 0, vZ = 0, wX = 0, wY = 0, wZ = 0;
        if (DIM == 2) {
            double d2 = 2.0 * (double)dx_L;
            uX = (T)((double)(su[H(I + 1, J, K)] - su[H(I - 1, J, K)]) / d2);
            uY = (T)((double)(su[H(I, J + 1, K)] - su[H(I, J - 1, K)]) / d2);
            vX = (T)((double)(sv[H(I + 1, J, K)] - sv[H(I - 1, J, K)]) / d2);
            vY = (T)((double)(sv[H(I, J + 1, K)] - sv[H(I, J - 1, K)]) / d2);
        } else {
            T d2 = 2 * dx_L;
            uX = (su[H(I + 1, J, K)] - su[H(I - 1, J, K)]) / d2;
            uY = (su[H(I, J + 1, K)] - su[H(I, J - 1, K)]) / d2;
            uZ = (su[H(I, J, K + 1)] - su[H(I, J, K - 1)]) / d2;
            vX = (sv[H(I + 1, J, K)] - sv[H(I - 1, J, K)]) / d2;
            vY = (sv[H(I, J + 1, K)] - sv[H(I, J - 1, K)]) / d2;
            vZ = (sv[H(I, J, K + 1)] - sv[H(I, J, K - 1)]) / d2;
            wX = (sw[H(I + 1, J, K)] - sw[H(I - 1, J, K)]) / d2;
            wY = (sw[H(I, J + 1, K)] - sw[H(I, J - 1, K)]) / d2;
            wZ = (sw[H(I, J, K + 1)] - sw[H(I, J, K - 1)]) / d2;
        }
        if (crit == 0)
            W = floor_log2_exact<T>(sqrt((wY - vZ) * (wY - vZ) + (uZ - wX) * (uZ - wX) + (vX - uY) * (vX - uY)));
        if (crit == 1)
            W = floor_log2_exact<T>(fabs((T)0.5 * (uX * uX + vY * vY + wZ * wZ + (T)2.0 * (uY * vX + uZ * wX + wY * vZ))));
        if (crit == 2)  // |grad u| (Frobenius norm); no counterpart in the reference
            W = floor_log2_exact<T>(sqrt(uX * uX + uY * uY + uZ * uZ + vX * vX + vY * vY + vZ * vZ + wX * wX + wY * wY + wZ * wZ));
        if (cells_mask[(long long)b * CELLS + t] == CELL_GHOST) W = (T)0.0;
    }
    // Block maximum (BlockwiseMaximum, inc/util.h:786-799): fmax ignores NaN operands.
#pragma unroll
    for (int d = 1; d < CELLS; d <<= 1) W = fmax(W, __shfl_xor(W, d, CELLS));
    if (work && t == 0) {
        if (W > NRM) W = NRM;
        int ref_kap = ref[b];
        int level_kap = level[b];
        int L_desired = MLI - 1;
        for (int p = 1; p <= MLI - 1; p++)
            if (W < NRS - NRI * (T)p) L_desired = (MLI - 1) - p;
        bool eligible = true;
        for (int p = 0; p < QMAX; p++)
            if (nbr[b + p * nmax] == N_SKIPID) eligible = false;
        if (eligible && level_kap != MLI - 1 && ref_kap == REF_UNREFINED && L_desired > level_kap)
            ref[b] = REF_MARK_REFINE;
        if (ref_kap == REF_REFINED && L_desired < level_kap + 1) ref[b] = REF_MARK_COARSEN;
        // Label of the learned predictor: the criterion wants this block to have children.
        if (label) label[b] = (eligible && level_kap != MLI - 1 && L_desired > level_kap) ? 1 : 0;
    } else if (label && b < id_max && t == 0) {
        label[b] = -1;  // inactive ID
    }
}

template <typename T>
static int criterion_t(Ctx *c, int *label)
{
    int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0) return AGAL_OK;
    const int cells = c->CELLS;
    int grid = (int)(((long long)idm * cells + AWG - 1) / AWG);
    T dx0 = (T)c->dxf_vec[0];
    if (c->DIM == 3)
        hipLaunchKernelGGL((k_criterion<T, 3>), dim3(grid), dim3(AWG), 0, c->stream, idm, c->n_maxcells,
                           static_cast<const T *>(c->M), c->ref, c->level, c->nbr, c->cells_mask, c->n_maxcblocks, dx0,
                           c->P.S_CRITERION, (T)c->P.N_REFINE_START, (T)c->P.N_REFINE_INC, (T)c->P.N_REFINE_MAX,
                           c->MAX_LEVELS_INTERIOR, label);
    else
        hipLaunchKernelGGL((k_criterion<T, 2>), dim3(grid), dim3(AWG), 0, c->stream, idm, c->n_maxcells,
                           static_cast<const T *>(c->M), c->ref, c->level, c->nbr, c->cells_mask, c->n_maxcblocks, dx0,
                           c->P.S_CRITERION, (T)c->P.N_REFINE_START, (T)c->P.N_REFINE_INC, (T)c->P.N_REFINE_MAX,
                           c->MAX_LEVELS_INTERIOR, label);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// Controls of the learned-refinement study (no reference counterpart).  Refine-only marks: a
// coarsening mark of the criterion reverts to the REFINED state it replaced.
__global__ void k_unmark_coarsen(int *ref, int n)
{
    const int i = blockIdx.x * AWG + threadIdx.x;
    if (i < n && ref[i] == REF_MARK_COARSEN) ref[i] = REF_REFINED;
}

// One-block dilation of the refinement marks: an unrefined block below the interior maximum level,
// with no missing neighbour (the criterion's eligibility), next to a block marked in `snap` (the
// marks before the pass) is marked too.  Same-level neighbours only.
__global__ void k_dilate_refine(const int *snap, int *ref, const int *level, const int *nbr, long long nmax, int n,
                                int mli)
{
    const int b = blockIdx.x * AWG + threadIdx.x;
    if (b >= n || snap[b] != REF_UNREFINED || level[b] >= mli - 1) return;
    bool near = false;
    for (int p = 0; p < 27; p++) {
        const int nb = nbr[b + p * nmax];
        if (nb == N_SKIPID) return;
        if (nb >= 0 && snap[nb] == REF_MARK_REFINE) near = true;
    }
    if (near) ref[b] = REF_MARK_REFINE;
}

int marks_refine_only(Ctx *c)
{
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_unmark_coarsen, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, idm);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

static int read_scalars(Ctx *c, int n);

int count_refine_marks(Ctx *c, int *n)
{
    *n = 0;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0) return AGAL_OK;
    int st;
    if ((st = launch_fill(c, c->dscal, 2, 0))) return st;
    hipLaunchKernelGGL(k_count_marks, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, idm, c->dscal);
    AGAL_HIP(c, hipGetLastError());
    if ((st = read_scalars(c, 1))) return st;
    *n = c->hscal[0];
    return AGAL_OK;
}

int marks_dilate(Ctx *c)
{
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0 || c->DIM != 3) return AGAL_OK;
    AGAL_HIP(c, hipMemcpyAsync(c->scan_b, c->ref, sizeof(int) * idm, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_dilate_refine, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->scan_b, c->ref, c->level, c->nbr,
                       c->n_maxcblocks, idm, c->MAX_LEVELS_INTERIOR);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

int launch_ref_criteria(Ctx *c, int *label)
{
    if (c->precision == AGAL_FP64) return criterion_t<double>(c, label);
    return criterion_t<float>(c, label);
}

// ---------------------------------------------------------------------------------------------
// Host orchestration.
// ---------------------------------------------------------------------------------------------
static int read_scalars(Ctx *c, int n)
{
    AGAL_HIP(c, hipMemcpyAsync(c->hscal, c->dscal, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

// Sorted per-level ID sets, interface lists, id_max, and the per-block auxiliary tables.
// The fused step's block list of every level in octree traversal order: level 0 in ID order (the
// root grid, x fastest), then level L as the children of level L-1's list, eight (four in 2D)
// consecutive IDs per parent in child order.  Neighbouring blocks then run close
// together on one XCD, where the partial cache lines that their pushes write into each other merge
// in L2; ID order drifts from that as refinement reuses freed IDs.  Only the order of the fused
// step's wavefronts changes, not what any of them computes.  A level whose traversal does not
// reproduce its ID set (blocks without an active parent) keeps the ID order.  (A Z-order of the
// root level measured 2.6% more bytes written per fused launch on the 3-level cavity.)
static int build_traversal_order(Ctx *c)
{
    int st;
    const int ML = c->MAX_LEVELS;
    if (c->n_ids[0] > 0)
        AGAL_HIP(c, hipMemcpyAsync(c->order_L(0), c->id_set_L(0), sizeof(int) * c->n_ids[0], hipMemcpyDeviceToDevice,
                                   c->stream));
    if ((st = launch_fill(c, c->dscal + 16, 16, 0))) return st;
    for (int L = 1; L < ML; L++) {
        st = compact(c, PredActiveChildren{c->order_L(L - 1), c->nbr_child, c->ref, c->level, L}, c->n_ids[L - 1],
                     EmitChildren{c->order_L(L - 1), c->nbr_child, c->order_L(L), c->NCH}, c->dscal + 16 + L);
        if (st) return st;
    }
    AGAL_HIP(c, hipMemcpyAsync(c->hscal + 16, c->dscal + 16, sizeof(int) * 16, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    for (int L = 1; L < ML; L++)
        if ((long long)c->hscal[16 + L] * c->NCH != c->n_ids[L] && c->n_ids[L] > 0)
            AGAL_HIP(c, hipMemcpyAsync(c->order_L(L), c->id_set_L(L), sizeof(int) * c->n_ids[L], hipMemcpyDeviceToDevice,
                                       c->stream));
    return AGAL_OK;
}

int rebuild_level_lists(Ctx *c, long long range_hint)
{
    int st;
    const int ML = c->MAX_LEVELS;
    long long range = range_hint;
    for (int L = 0; L <= ML; L++) range = std::max<long long>(range, c->id_max[L]);
    // dscal[0..ML): per-level counts (compaction bases); dscal[8..8+ML): interface counts.
    if ((st = launch_fill(c, c->dscal, 16, 0))) return st;
    for (int L = 0; L < ML; L++) {
        st = compact(c, PredActiveLevel{c->ref, c->level, L}, range, EmitIndex{c->id_set_L(L)}, c->dscal + L);
        if (st) return st;
    }
    if ((st = read_scalars(c, 16))) return st;
    for (int L = 0; L < ML; L++) c->n_ids[L] = c->hscal[L];
    // Interface lists and the last (largest) ID of every level.
    for (int L = 0; L < ML; L++) {
        st = compact(c, PredIfaceInList{c->id_set_L(L), c->blk_mask}, c->n_ids[L],
                     EmitListEntry{c->id_set_L(L), c->iface_L(L)}, c->dscal + 8 + L);
        if (st) return st;
    }
    std::vector<int> last(ML, -1);
    for (int L = 0; L < ML; L++)
        if (c->n_ids[L] > 0)
            AGAL_HIP(c, hipMemcpyAsync(&last[L], c->id_set_L(L) + c->n_ids[L] - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((st = read_scalars(c, 16))) return st;
    for (int L = 0; L < ML; L++) c->n_iface[L] = c->hscal[8 + L];
    if ((st = build_traversal_order(c))) return st;
    c->mesh_epoch++;  // recorded step graphs carry the old list sizes
    for (int L = 1; L < ML; L++) c->id_max[L] = c->n_ids[L] > 0 ? last[L] + 1 : 0;
    c->id_max[ML] = *std::max_element(c->id_max.begin(), c->id_max.begin() + ML);
    int n = c->id_max[ML];
    if (c->DIM == 3) {
        hipLaunchKernelGGL((k_block_aux<3>), dim3(grid1(n)), dim3(AWG), 0, c->stream, c->nbr, c->nbr_child, c->blk_mask,
                           c->cells_mask, c->nbr_slot, c->part, c->scan_a, c->n_maxcblocks, n);
        hipLaunchKernelGGL((k_block_simple<3>), dim3(grid1(n)), dim3(AWG), 0, c->stream, c->nbr_slot, c->part, c->scan_a,
                           c->simple, n, c->Q == 27 ? 3 : 2);
    } else {
        hipLaunchKernelGGL((k_block_aux<2>), dim3(grid1(n)), dim3(AWG), 0, c->stream, c->nbr, c->nbr_child, c->blk_mask,
                           c->cells_mask, c->nbr_slot, c->part, c->scan_a, c->n_maxcblocks, n);
        hipLaunchKernelGGL((k_block_simple<2>), dim3(grid1(n)), dim3(AWG), 0, c->stream, c->nbr_slot, c->part, c->scan_a,
                           c->simple, n, 2);
    }
    AGAL_HIP(c, hipGetLastError());
    // The linked blocks of the new level lists (their ranks and fractions stay with the block IDs).
    return links_rebuild_lists(c);
}

// Mesh::M_RefineAndCoarsenBlocks (src/mesh/mesh_amr.cu:2161-2744).
int refine_coarsen(Ctx *c, int out_counts[4])
{
    int st, status = AGAL_OK;
    const int ML = c->MAX_LEVELS, NCH = c->NCH, QM = c->QMAX;
    const long long nmax = c->n_maxcblocks;
    const int idm = c->id_max[ML];
    // Pre: count marks (two thrust::count_if, :2179-2185).
    if ((st = launch_fill(c, c->dscal, 16, 0))) return st;
    hipLaunchKernelGGL(k_count_marks, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, idm, c->dscal);
    AGAL_HIP(c, hipGetLastError());
    if ((st = read_scalars(c, 2))) return st;
    int n_ref = c->hscal[0], n_coa = c->hscal[1];
    long long idm_wnew = idm + (long long)NCH * n_ref;
    bool proceed_ref = n_ref > 0 && (long long)NCH * n_ref <= c->n_gaps;
    bool proceed_coa = n_coa > 0;
    if (!proceed_ref && n_ref > 0) {
        status = AGAL_ERR_POOL_EXHAUSTED;
        n_ref = 0;
        idm_wnew = idm;
        hipLaunchKernelGGL(k_cancel, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, idm);
        AGAL_HIP(c, hipGetLastError());
    }
    int n_added = 0, n_removed = 0;
    if (proceed_ref || proceed_coa) {
        c->stamp++;
        hipLaunchKernelGGL(k_map_mark, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr, nmax, QM, idm,
                           c->map_stamp, c->stamp);
        AGAL_HIP(c, hipGetLastError());
    }
    // S2: child IDs from the gap stack.
    if (proceed_ref) {
        if ((st = launch_fill(c, c->dscal, 16, 0))) return st;
        if ((st = launch_fill(c, c->dscal + 4, 1, -1))) return st;
        st = compact(c, PredRefEq{c->ref, REF_MARK_REFINE}, idm,
                     EmitFirstChild{c->gap_set, c->first_child, c->n_gaps, NCH}, c->dscal);
        if (st) return st;
        int lo = c->n_gaps - NCH * n_ref, hi = c->n_gaps;
        hipLaunchKernelGGL(k_max_range, dim3(grid1(hi - lo)), dim3(AWG), 0, c->stream, c->gap_set, lo, hi, c->dscal + 4);
        hipLaunchKernelGGL(k_s2_1, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr, c->nbr_child, c->map_stamp,
                           c->stamp, c->first_child, nmax, QM, idm);
        AGAL_HIP(c, hipGetLastError());
        if ((st = read_scalars(c, 8))) return st;
        idm_wnew = std::max<long long>(idm_wnew, (long long)c->hscal[4] + 1);
    }
    // S3: quality control on coarsening.
    if (proceed_coa) {
        hipLaunchKernelGGL(k_q1_1, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr_child, nmax, QM, idm);
        hipLaunchKernelGGL(k_q1_2, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr_child, NCH, idm);
        if ((st = launch_fill(c, c->dscal, 16, 0))) return st;
        hipLaunchKernelGGL(k_count_marks, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, idm, c->dscal);
        AGAL_HIP(c, hipGetLastError());
        if ((st = read_scalars(c, 2))) return st;
        n_coa = c->hscal[1];
        proceed_coa = n_coa > 0;
        if (proceed_coa) {
            hipLaunchKernelGGL(k_s2_2, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr, c->nbr_child,
                               c->map_stamp, c->stamp, nmax, QM, idm);
            AGAL_HIP(c, hipGetLastError());
        }
    }
    if (!(proceed_ref || proceed_coa)) {
        if (out_counts) {
            out_counts[0] = n_ref;
            out_counts[1] = 0;
            out_counts[2] = 0;
            out_counts[3] = 0;
        }
        return status;
    }
    // S4: add children / mark children of coarsened blocks for removal.
    if (c->precision == AGAL_FP64)
        hipLaunchKernelGGL((k_add_remove<double>), dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->level,
                           c->nbr_child, c->X, c->first_child, nmax, c->DIM, NCH, (double)c->dx_cblock, idm);
    else
        hipLaunchKernelGGL((k_add_remove<float>), dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->level,
                           c->nbr_child, c->X, c->first_child, nmax, c->DIM, NCH, (float)c->dx_cblock, idm);
    AGAL_HIP(c, hipGetLastError());
    // S5: removed children, grouped by level in ID order, pushed reversed on the gap stack.
    if (proceed_coa) {
        if ((st = launch_fill(c, c->dscal, 16, 0))) return st;
        for (int L = 1; L < ML; L++) {
            st = compact(c, PredRemoveLevel{c->ref, c->level, L}, idm_wnew, EmitIndex{c->scan_b}, c->dscal);
            if (st) return st;
        }
        int at = c->n_gaps - NCH * n_ref;
        hipLaunchKernelGGL(k_gap_append, dim3(grid1(idm_wnew)), dim3(AWG), 0, c->stream, c->gap_set, at, c->scan_b, c->dscal);
        hipLaunchKernelGGL(k_coarsen_s2, dim3(grid1(idm)), dim3(AWG), 0, c->stream, c->ref, c->nbr, c->nbr_child, nmax, QM, idm);
        AGAL_HIP(c, hipGetLastError());
        if ((st = read_scalars(c, 1))) return st;
        n_removed = c->hscal[0];
        c->n_gaps += n_removed;
    }
    // S6: new children join the ID sets (rebuilt sorted from the pool), gaps consumed.
    if (proceed_ref) {
        c->n_gaps -= NCH * n_ref;
        n_added = NCH * n_ref;
    }
    if ((st = rebuild_level_lists(c, idm_wnew))) return st;
    // S7: masks, interpolation to new blocks, state bookkeeping.
    const int idm_new = c->id_max[ML];
    hipLaunchKernelGGL(k_update_masks_1, dim3(grid1(idm_new)), dim3(AWG), 0, c->stream, c->blk_mask, c->ref,
                       c->nbr_child, nmax, QM, idm_new);
    AGAL_HIP(c, hipGetLastError());
    for (int L = 0; L < ML - 1; L++)
        if ((st = launch_interpolate(c, L, AGAL_INTERP_ADDED))) return st;
    hipLaunchKernelGGL(k_s3, dim3(grid1(idm_wnew)), dim3(AWG), 0, c->stream, c->ref, (int)idm_wnew);
    hipLaunchKernelGGL(k_s4, dim3(grid1(idm_wnew)), dim3(AWG), 0, c->stream, c->ref, c->nbr_child, NCH, (int)idm_wnew);
    AGAL_HIP(c, hipGetLastError());
    // S8: children's neighbour lists, boundary flags, cell masks.
    {
        long long nthreads = (long long)idm_new * NCH * QM;
        hipLaunchKernelGGL(k_update_connectivity, dim3(grid1(nthreads)), dim3(AWG), 0, c->stream, c->nbr, c->nbr_child,
                           c->map_stamp, c->stamp, nmax, c->DIM, QM, NCH, idm_new);
        hipLaunchKernelGGL(k_update_boundaries, dim3(grid1(idm_new)), dim3(AWG), 0, c->stream, c->nbr, c->nbr_child,
                           c->blk_mask, c->onb, nmax, QM, idm_new);
        long long ncell = (long long)idm_new * c->CELLS;
        if (c->DIM == 3)
            hipLaunchKernelGGL((k_update_masks_2<3>), dim3(grid1(ncell)), dim3(AWG), 0, c->stream, c->cells_mask, c->nbr,
                               c->ref, nmax, idm_new);
        else
            hipLaunchKernelGGL((k_update_masks_2<2>), dim3(grid1(ncell)), dim3(AWG), 0, c->stream, c->cells_mask, c->nbr,
                               c->ref, nmax, idm_new);
        AGAL_HIP(c, hipGetLastError());
    }
    // Interface lists depend on block masks (updated above): recompute lists and aux tables.
    if ((st = rebuild_level_lists(c, 0))) return st;
    if (out_counts) {
        out_counts[0] = n_ref;
        out_counts[1] = proceed_coa ? n_coa : 0;
        out_counts[2] = n_added;
        out_counts[3] = n_removed;
    }
    return status;
}

}  // namespace agal
