/*
 * oracle_capi.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * C wrappers around the CPU restatement in agal_oracle.hpp so that tests/ (via ctypes),
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg can drive it with the same call
 * sequence they use for the HIP library.  Signatures mirror include/agal_capi.h with an
 * `oracle_` prefix; populations are exchanged in physical order (entry q = direction q).
 */
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "agal_oracle.hpp"

using namespace agal_oracle;

namespace {
struct Handle {
    int precision = AGAL_FP32;
    Oracle<float> *f = nullptr;
    Oracle<double> *d = nullptr;
};

template <typename Fn>
int dispatch(void *h, Fn fn)
{
    if (!h) return AGAL_ERR_ARG;
    Handle *H = static_cast<Handle *>(h);
    if (H->f) return fn(*H->f);
    if (H->d) return fn(*H->d);
    return AGAL_ERR_STATE;
}
}  // namespace

extern "C" {

int oracle_create(const agal_params *p, long long capacity, int device, void **out)
{
    (void)device;
    if (!p || !out || capacity <= 0) return AGAL_ERR_ARG;
    if (p->velocity_set != 9 && p->velocity_set != 19 && p->velocity_set != 27) return AGAL_ERR_ARG;
    Handle *H = new (std::nothrow) Handle;
    if (!H) return AGAL_ERR_ALLOC;
    H->precision = p->precision;
    int st;
    try {
        if (p->precision == AGAL_FP64) {
            H->d = new Oracle<double>;
            st = H->d->init(*p, capacity);
        } else {
            H->f = new Oracle<float>;
            st = H->f->init(*p, capacity);
        }
    } catch (const std::bad_alloc &) {
        st = AGAL_ERR_ALLOC;
    }
    if (st != AGAL_OK) {
        delete H->f;
        delete H->d;
        delete H;
        return st;
    }
    *out = H;
    return AGAL_OK;
}

int oracle_destroy(void *h)
{
    if (!h) return AGAL_ERR_ARG;
    Handle *H = static_cast<Handle *>(h);
    delete H->f;
    delete H->d;
    delete H;
    return AGAL_OK;
}

int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// Options of the GPU library that change results; the others (graphs, graph length, exception-list
// cap) have no meaning here.
int oracle_set_option(void *h, int option, long long value)
{
    if (option != AGAL_OPT_LINK_MASKS) return AGAL_ERR_ARG;
    return dispatch(h, [&](auto &O) {
        O.link_masks = value != 0;
        return AGAL_OK;
    });
}

int oracle_sync(void *h) { return h ? AGAL_OK : AGAL_ERR_ARG; }

int oracle_get_sizes(void *h, agal_sizes *s)
{
    return dispatch(h, [&](auto &O) {
        std::memset(s, 0, sizeof(*s));
        s->n_dim = O.DIM;
        s->n_q = O.Q;
        s->n_q_max = O.Q_max;
        s->cells_per_block = O.M_CBLOCK;
        s->n_children = O.N_CHILDREN;
        s->max_levels = O.MAX_LEVELS;
        s->n_maxcblocks = O.n_maxcblocks;
        s->n_maxcells = O.n_maxcells;
        s->n_coarsecblocks = O.n_coarsecblocks;
        for (int L = 0; L < O.MAX_LEVELS; L++) s->n_ids[L] = O.n_ids[L];
        for (int L = 0; L <= O.MAX_LEVELS; L++) s->id_max[L] = O.id_max[L];
        s->n_gaps = O.n_gaps;
        return AGAL_OK;
    });
}

int oracle_set_ic(void *h, int L) { return dispatch(h, [&](auto &O) { O.set_ic(L); return AGAL_OK; }); }
int oracle_advance(void *h, int L) { return dispatch(h, [&](auto &O) { O.advance(L); return AGAL_OK; }); }
// Planted mistake of the restatement (0 = none); only the tests that show the pins catch
// mistakes set it.
int oracle_plant(void *h, int code) { return dispatch(h, [&](auto &O) { O.plant = code; return AGAL_OK; }); }
int oracle_collide(void *h, int L) { return dispatch(h, [&](auto &O) { O.collide(L); return AGAL_OK; }); }
int oracle_impose_bc(void *h, int L) { return dispatch(h, [&](auto &O) { O.impose_bc(L); return AGAL_OK; }); }
int oracle_stream(void *h, int L) { return dispatch(h, [&](auto &O) { O.stream(L); return AGAL_OK; }); }
int oracle_collide_stream(void *h, int L) { return dispatch(h, [&](auto &O) { O.collide_stream(L); return AGAL_OK; }); }
int oracle_interpolate(void *h, int L, int var) { return dispatch(h, [&](auto &O) { O.interpolate(L, var); return AGAL_OK; }); }
int oracle_average(void *h, int L, int var) { return dispatch(h, [&](auto &O) { O.average(L, var); return AGAL_OK; }); }
int oracle_refresh(void *h, int L) { return dispatch(h, [&](auto &O) { O.refresh(L); return AGAL_OK; }); }
int oracle_ref_criteria(void *h) { return dispatch(h, [&](auto &O) { O.ref_criteria(); return AGAL_OK; }); }
int oracle_refine_coarsen(void *h, int *counts) { return dispatch(h, [&](auto &O) { return O.refine_coarsen(counts); }); }
int oracle_refine_with_solution(void *h, int *counts) { return dispatch(h, [&](auto &O) { return O.refine_with_solution(counts); }); }
int oracle_refine_near_wall(void *h) { return dispatch(h, [&](auto &O) { return O.refine_near_wall(); }); }
int oracle_set_geometry(void *h, const float *tri, long long n_tri)
{
    if (!tri || n_tri <= 0) return AGAL_ERR_ARG;
    return dispatch(h, [&](auto &O) { O.set_geometry(tri, (long)n_tri); return AGAL_OK; });
}
int oracle_voxelize(void *h, int L)
{
    return dispatch(h, [&](auto &O) {
        O.voxelize(L);
        O.mark_near_wall(L);
        return AGAL_OK;
    });
}
int oracle_download_links(void *h, int *onb_solid, long long n_blocks, float *X_b, long long n_values,
                          long long *info)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        if (info) {
            info[0] = O.n_solidb;
            info[1] = O.n_link_corrections;
            info[2] = O.links ? 1 : 0;
        }
        if (onb_solid)
            for (long long k = 0; k < n_blocks; k++) onb_solid[k] = O.onb_solid.empty() ? -1 : O.onb_solid[k];
        if (X_b) {
            if (n_values > (long long)O.X_b.size()) return AGAL_ERR_ARG;
            std::copy(O.X_b.begin(), O.X_b.begin() + n_values, X_b);
        }
        return AGAL_OK;
    });
}

int oracle_download_link_bins(void *h, int *bin_start, long long n_start, int *faces, long long n_faces,
                              long long *info)
{
    return dispatch(h, [&](auto &O) {
        const auto &B = O.lbins;
        if (info) {
            info[0] = (long long)B.bin_start.size() - 1;
            info[1] = (long long)B.faces.size();
            info[2] = B.G;
            info[3] = (long long)B.filtered.size();
        }
        if (bin_start) {
            if (n_start > (long long)B.bin_start.size()) return AGAL_ERR_ARG;
            std::copy(B.bin_start.begin(), B.bin_start.begin() + n_start, bin_start);
        }
        if (faces) {
            if (n_faces > (long long)B.faces.size()) return AGAL_ERR_ARG;
            std::copy(B.faces.begin(), B.faces.begin() + n_faces, faces);
        }
        return AGAL_OK;
    });
}

int oracle_check_masks(void *h, int L) { return dispatch(h, [&](auto &O) { O.check_masks(L); return AGAL_OK; }); }
int oracle_forces(void *h, double *out) { return dispatch(h, [&](auto &O) { O.report_forces(out); return AGAL_OK; }); }
/* Crossing count of the +x ray from (px, py, pz) with n_tri triangles (the voxelizer's predicate). */
long long oracle_ray_count(const float *tri, long long n_tri, double px, double py, double pz)
{
    long long c = 0;
    for (long long k = 0; k < n_tri; k++) c += Oracle<double>::ray_crosses(tri + 9 * k, px, py, pz);
    return c;
}
/* Predictor features of block IDs [0, n_blocks) (zeros beyond the largest active ID). */
int oracle_features(void *h, float *out, long long n_blocks)
{
    return dispatch(h, [&](auto &O) {
        const long long idm = O.id_max[O.MAX_LEVELS];
        std::vector<float> f((size_t)12 * std::max<long long>(idm, 1));
        O.features(f.data());
        for (long long b = 0; b < n_blocks; b++)
            for (int k = 0; k < 12; k++) out[12 * b + k] = b < idm ? f[12 * b + k] : 0.f;
        return AGAL_OK;
    });
}
/* The plain criterion's labels; refinement marks left unchanged. */
int oracle_criterion_labels(void *h, int *out, long long n_blocks)
{
    return dispatch(h, [&](auto &O) {
        const long long idm = O.id_max[O.MAX_LEVELS];
        std::vector<int> saved = O.ref, lab(std::max<long long>(idm, 1), -1);
        O.ref_criteria(lab.data());
        O.ref = saved;
        for (long long b = 0; b < n_blocks; b++) out[b] = b < idm ? lab[b] : -1;
        return AGAL_OK;
    });
}
/* The network restated: logits of n rows with the FP16 copy of master parameters w. */
int oracle_mlp_forward(const float *w, const float *x, int n, float *z)
{
    MlpRef M;
    M.w.assign(w, w + MlpRef::N);
    const std::vector<float> wh = M.half_copy();
    for (int r = 0; r < n; r++) z[r] = MlpRef::forward_row(wh.data(), x + 12 * (size_t)r, nullptr, nullptr);
    return AGAL_OK;
}
/* One Adam step; w, m, v (AGAL_MLP_NPARAM each) and the step count t are updated in place, the
 * gradient goes to g (may be NULL). */
int oracle_mlp_train_step(float *w, float *m, float *v, int *t, const float *x, const float *y, int n, double lr,
                          double pos_weight, float *g, double *loss)
{
    MlpRef M;
    M.w.assign(w, w + MlpRef::N);
    M.m.assign(m, m + MlpRef::N);
    M.v.assign(v, v + MlpRef::N);
    M.t = *t;
    const double l = M.train_step(x, y, n, (float)lr, (float)pos_weight);
    std::copy(M.w.begin(), M.w.end(), w);
    std::copy(M.m.begin(), M.m.end(), m);
    std::copy(M.v.begin(), M.v.end(), v);
    if (g) std::copy(M.g.begin(), M.g.end(), g);
    *t = M.t;
    if (loss) *loss = l;
    return AGAL_OK;
}
float oracle_plog2(float v) { return Oracle<float>::plog2(v); }
int oracle_mark_uniform(void *h, int L) { return dispatch(h, [&](auto &O) { O.mark_uniform(L); return AGAL_OK; }); }
int oracle_freeze_refined(void *h, int var) { return dispatch(h, [&](auto &O) { O.freeze_refined(var); return AGAL_OK; }); }

int oracle_mark_blocks(void *h, const int *ids, int n, int mark)
{
    return dispatch(h, [&](auto &O) {
        for (int k = 0; k < n; k++) {
            if (ids[k] < 0 || ids[k] >= O.n_maxcblocks) return AGAL_ERR_ARG;
            O.ref[ids[k]] = mark;
        }
        return AGAL_OK;
    });
}

/* Wall time (ms) of the coarse steps and of the rebuilds, as agal_advance_loop reports. */
int oracle_advance_loop(void *h, int i0, int n, double *step_ms, double *rebuild_ms)
{
    return dispatch(h, [&](auto &O) {
        using clk = std::chrono::steady_clock;
        double ts = 0, tr = 0;
        int st = AGAL_OK;
        int iter_mult = 1 << O.N_LEVEL_START;
        for (int i = i0; i < i0 + n; i++) {
            auto a = clk::now();
            for (int j = 0; j < iter_mult; j++) O.advance(O.N_LEVEL_START);
            auto b = clk::now();
            ts += std::chrono::duration<double, std::milli>(b - a).count();
            if (O.MAX_LEVELS > 1 && (O.MAX_LEVELS != O.N_LEVEL_START + 1) && (i + 1) % O.P.P_REFINE == 0 && i > 0) {
                int s = O.refine_with_solution(nullptr);
                if (s != AGAL_OK) st = s;
                tr += std::chrono::duration<double, std::milli>(clk::now() - b).count();
            }
        }
        if (step_ms) *step_ms = ts;
        if (rebuild_ms) *rebuild_ms = tr;
        return st;
    });
}

/* Populations in physical order: dst[q*dst_stride + cell] = pre-collision f_q. */
int oracle_download_populations(void *h, void *dst, long long n_blocks, long long dst_stride)
{
    return dispatch(h, [&](auto &O) {
        using T = typename std::decay<decltype(O.F[0])>::type;
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        T *out = static_cast<T *>(dst);
        long long nc = n_blocks * O.M_CBLOCK;
        for (int q = 0; q < O.Q; q++)
            std::memcpy(out + q * dst_stride, O.Fp(O.V.pb[q]), sizeof(T) * nc);
        return AGAL_OK;
    });
}

int oracle_upload_populations(void *h, const void *src, long long n_blocks, long long src_stride)
{
    return dispatch(h, [&](auto &O) {
        using T = typename std::decay<decltype(O.F[0])>::type;
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        const T *in = static_cast<const T *>(src);
        long long nc = n_blocks * O.M_CBLOCK;
        for (int q = 0; q < O.Q; q++) {
            T *F = O.Fp(O.V.pb[q]);
            for (long long i = 0; i < nc; i++) F[i] = O.quant(in[q * src_stride + i], q);
        }
        return AGAL_OK;
    });
}

/* rho,u,v,w as stored by the last refresh/collide (the reference's fields N_Q..N_Q+3). */
int oracle_download_macros(void *h, void *dst, long long n_blocks, long long dst_stride)
{
    return dispatch(h, [&](auto &O) {
        using T = typename std::decay<decltype(O.F[0])>::type;
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        T *out = static_cast<T *>(dst);
        long long nc = n_blocks * O.M_CBLOCK;
        for (int k = 0; k < 4; k++) {
            if (k == 3 && O.DIM == 2) {
                for (long long i = 0; i < nc; i++) out[k * dst_stride + i] = 0;
                continue;
            }
            std::memcpy(out + k * dst_stride, O.Fp(O.Q + k), sizeof(T) * nc);
        }
        return AGAL_OK;
    });
}

int oracle_download_cell_masks(void *h, int *dst, long long n_blocks)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        std::memcpy(dst, O.cells_mask.data(), sizeof(int) * n_blocks * O.M_CBLOCK);
        return AGAL_OK;
    });
}

int oracle_upload_cell_masks(void *h, const int *src, long long n_blocks)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        std::memcpy(O.cells_mask.data(), src, sizeof(int) * n_blocks * O.M_CBLOCK);
        return AGAL_OK;
    });
}

int oracle_download_block_ints(void *h, int which, int *dst, long long n_blocks, long long dst_stride)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        const std::vector<int> *src = nullptr;
        int rows = 1;
        switch (which) {
            case AGAL_BLK_NBR: src = &O.nbr; rows = O.Q_max; break;
            case AGAL_BLK_NBR_CHILD: src = &O.nbr_child; rows = O.Q_max; break;
            case AGAL_BLK_REF: src = &O.ref; break;
            case AGAL_BLK_LEVEL: src = &O.level; break;
            case AGAL_BLK_ONB: src = &O.onb; break;
            case AGAL_BLK_MASK: src = &O.blk_mask; break;
            default: return AGAL_ERR_ARG;
        }
        for (int r = 0; r < rows; r++)
            std::memcpy(dst + r * dst_stride, src->data() + (size_t)r * O.n_maxcblocks, sizeof(int) * n_blocks);
        return AGAL_OK;
    });
}

int oracle_download_coords(void *h, float *dst, long long n_blocks, long long dst_stride)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        for (int d = 0; d < O.DIM; d++)
            std::memcpy(dst + d * dst_stride, O.X.data() + (size_t)d * O.n_maxcblocks, sizeof(float) * n_blocks);
        return AGAL_OK;
    });
}

int oracle_download_id_set(void *h, int L, int *dst, int n)
{
    return dispatch(h, [&](auto &O) {
        if (L < 0 || L >= O.MAX_LEVELS || n < 0 || n > O.n_ids[L]) return AGAL_ERR_ARG;
        std::memcpy(dst, O.idset(L), sizeof(int) * n);
        return AGAL_OK;
    });
}

int oracle_download_gap_set(void *h, int *dst, int n)
{
    return dispatch(h, [&](auto &O) {
        if (n < 0 || n > O.n_gaps) return AGAL_ERR_ARG;
        std::memcpy(dst, O.gap_set.data(), sizeof(int) * n);
        return AGAL_OK;
    });
}

/* Replace the whole block-pool state (tables for blocks [0, n_blocks), per-level ID sets given
 * concatenated in level order, the gap stack and the ID maxima) -- used to time the restatement on
 * a mesh produced elsewhere. */
int oracle_import_pool_arrays(void *h, long long n_blocks, const int *nbr, const int *nbr_child, const int *ref,
                       const int *level, const int *onb, const int *blk_mask, const float *coords,
                       const int *cell_mask, const int *n_ids, const int *id_max, const int *id_sets,
                       int n_gaps, const int *gap_set)
{
    return dispatch(h, [&](auto &O) {
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks || n_gaps < 0 || n_gaps > O.n_maxcblocks) return AGAL_ERR_ARG;
        for (int p = 0; p < O.Q_max; p++) {
            std::memcpy(&O.nbr[(size_t)p * O.n_maxcblocks], nbr + p * n_blocks, sizeof(int) * n_blocks);
            std::memcpy(&O.nbr_child[(size_t)p * O.n_maxcblocks], nbr_child + p * n_blocks, sizeof(int) * n_blocks);
        }
        std::memcpy(O.ref.data(), ref, sizeof(int) * n_blocks);
        std::memcpy(O.level.data(), level, sizeof(int) * n_blocks);
        std::memcpy(O.onb.data(), onb, sizeof(int) * n_blocks);
        std::memcpy(O.blk_mask.data(), blk_mask, sizeof(int) * n_blocks);
        for (int d = 0; d < O.DIM; d++)
            std::memcpy(&O.X[(size_t)d * O.n_maxcblocks], coords + d * n_blocks, sizeof(float) * n_blocks);
        std::memcpy(O.cells_mask.data(), cell_mask, sizeof(int) * n_blocks * O.M_CBLOCK);
        long long off = 0;
        for (int L = 0; L < O.MAX_LEVELS; L++) {
            O.n_ids[L] = n_ids[L];
            std::memcpy(&O.id_set[(size_t)L * O.n_maxcblocks], id_sets + off, sizeof(int) * n_ids[L]);
            off += n_ids[L];
        }
        for (int L = 0; L <= O.MAX_LEVELS; L++) O.id_max[L] = id_max[L];
        O.n_gaps = n_gaps;
        std::memcpy(O.gap_set.data(), gap_set, sizeof(int) * n_gaps);
        return AGAL_OK;
    });
}

/* agal_import_pool for the restatement: the tables of blocks [0, n_blocks), later IDs inactive,
 * the ID sets as given (checked against ref / level), populations in physical order. */
int oracle_import_pool(void *h, const agal_pool *pl)
{
    if (!pl) return AGAL_ERR_ARG;
    return dispatch(h, [&](auto &O) {
        using T = typename std::decay<decltype(O.F[0])>::type;
        const long long nb = pl->n_blocks;
        if (nb < O.n_coarsecblocks || nb > O.n_maxcblocks || pl->n_gaps < 0 || pl->n_gaps > O.n_maxcblocks) return AGAL_ERR_ARG;
        const long long old = O.id_max[O.MAX_LEVELS];
        for (long long b = nb; b < old; b++) {
            O.ref[b] = REF_INACTIVE;
            O.level[b] = 0;
            O.onb[b] = 0;
            O.blk_mask[b] = BLK_REGULAR;
            for (int p = 0; p < O.Q_max; p++) {
                O.nbr[(size_t)p * O.n_maxcblocks + b] = N_SKIPID;
                O.nbr_child[(size_t)p * O.n_maxcblocks + b] = N_SKIPID;
            }
        }
        for (int p = 0; p < O.Q_max; p++) {
            std::memcpy(&O.nbr[(size_t)p * O.n_maxcblocks], pl->nbr + p * nb, sizeof(int) * nb);
            std::memcpy(&O.nbr_child[(size_t)p * O.n_maxcblocks], pl->nbr_child + p * nb, sizeof(int) * nb);
        }
        std::memcpy(O.ref.data(), pl->ref, sizeof(int) * nb);
        std::memcpy(O.level.data(), pl->level, sizeof(int) * nb);
        std::memcpy(O.onb.data(), pl->onb, sizeof(int) * nb);
        std::memcpy(O.blk_mask.data(), pl->blk_mask, sizeof(int) * nb);
        for (int d = 0; d < O.DIM; d++)
            std::memcpy(&O.X[(size_t)d * O.n_maxcblocks], pl->coords + d * nb, sizeof(float) * nb);
        std::memcpy(O.cells_mask.data(), pl->cell_mask, sizeof(int) * nb * O.M_CBLOCK);
        long long off = 0;
        for (int L = 0; L < O.MAX_LEVELS; L++) {
            std::vector<int> expect;
            for (long long b = 0; b < nb; b++)
                if (O.ref[b] != REF_INACTIVE && O.level[b] == L) expect.push_back((int)b);
            if ((long long)expect.size() != pl->n_ids[L] || !std::equal(expect.begin(), expect.end(), pl->id_sets + off))
                return AGAL_ERR_ARG;
            O.n_ids[L] = pl->n_ids[L];
            std::memcpy(&O.id_set[(size_t)L * O.n_maxcblocks], pl->id_sets + off, sizeof(int) * pl->n_ids[L]);
            off += pl->n_ids[L];
            O.id_max[L] = (L == 0) ? O.n_coarsecblocks : (pl->n_ids[L] > 0 ? pl->id_sets[off - 1] + 1 : 0);
        }
        int m = 0;
        for (int L = 0; L < O.MAX_LEVELS; L++) m = std::max(m, O.id_max[L]);
        O.id_max[O.MAX_LEVELS] = m;
        O.n_gaps = pl->n_gaps;
        if (pl->n_gaps > 0) std::memcpy(O.gap_set.data(), pl->gap_set, sizeof(int) * pl->n_gaps);
        if (pl->populations) {
            const T *in = static_cast<const T *>(pl->populations);
            const long long nc = nb * O.M_CBLOCK;
            for (int q = 0; q < O.Q; q++) {
                T *F = O.Fp(O.V.pb[q]);
                for (long long i = 0; i < nc; i++) F[i] = O.quant(in[q * nc + i], q);
            }
        }
        return AGAL_OK;
    });
}

/* The restated velocity set: c (3 rows of 27), w, pb, conn_map (all of length Q_max). */
int oracle_velset(int Q, int *c, double *w, int *pb, int *conn_map)
{
    if (Q != 9 && Q != 19 && Q != 27) return AGAL_ERR_ARG;
    VelSet V = make_velset(Q);
    for (int d = 0; d < 3; d++)
        for (int p = 0; p < 27; p++) c[d * 27 + p] = V.c[d][p];
    for (int p = 0; p < 27; p++) {
        w[p] = V.w[p];
        pb[p] = V.pb[p];
        conn_map[p] = V.conn_map[p];
    }
    return V.Q_max;
}

int oracle_child_neighbor_rule(int Q, int c, int p, int *src_dir, int *off)
{
    if (Q != 9 && Q != 19 && Q != 27) return AGAL_ERR_ARG;
    VelSet V = make_velset(Q);
    child_neighbor_rule(V, c, p, *src_dir, *off);
    return AGAL_OK;
}

int oracle_cell_mask_rule(int Q, const int *skip, int t)
{
    VelSet V = make_velset(Q);
    return cell_mask_rule(V, skip, t);
}

int oracle_in_half_set(int dim, int p) { return in_half_set(dim, p) ? 1 : 0; }

/* Terms of the cubic interpolation's alpha_idx (dim 2 or 3): fills q[] and coef[] (room for 64)
 * in summation order and returns their number. */
int oracle_cubic_terms(int dim, int idx, int *q, double *coef)
{
    const int *start = (dim == 3) ? CUBIC3_START : CUBIC2_START;
    const unsigned char *qs = (dim == 3) ? CUBIC3_Q : CUBIC2_Q;
    const double *cs = (dim == 3) ? CUBIC3_C : CUBIC2_C;
    if (idx < 0 || idx >= (dim == 3 ? 64 : 16)) return -1;
    int n = 0;
    for (int k = start[idx]; k < start[idx + 1]; k++, n++) {
        q[n] = qs[k];
        coef[n] = cs[k];
    }
    return n;
}

/* The restatement's MRT basis: fills m[k*Q + p] and rate[k] (0 conserved, 1 omega, 2..7 s1..s6). */
int oracle_mrt_basis(int Q, int *m, int *rate)
{
    if (Q != 9 && Q != 19 && Q != 27) return AGAL_ERR_ARG;
    Oracle<double> O;
    O.V = make_velset(Q);
    O.Q = Q;
    for (int k = 0; k < Q; k++) {
        rate[k] = O.mrt_rate(k);
        for (int p = 0; p < Q; p++) m[k * Q + p] = O.mrt_m(k, p);
    }
    return AGAL_OK;
}

/* The 16-bit conversions of the restatement: mode 1 binary16, mode 2 bfloat16. */
int oracle_to16(int mode, const float *in, uint16_t *out, long long n)
{
    if (mode != 1 && mode != 2) return AGAL_ERR_ARG;
    for (long long i = 0; i < n; i++) out[i] = mode == 1 ? f16_from_f32(in[i]) : bf16_from_f32(in[i]);
    return AGAL_OK;
}

int oracle_from16(int mode, const uint16_t *in, float *out, long long n)
{
    if (mode != 1 && mode != 2) return AGAL_ERR_ARG;
    for (long long i = 0; i < n; i++) out[i] = mode == 1 ? f16_to_f32(in[i]) : bf16_to_f32(in[i]);
    return AGAL_OK;
}

/* Overwrite the stored rho,u,v,w fields (the reference's fields N_Q..N_Q+3). */
int oracle_upload_macros(void *h, const void *src, long long n_blocks, long long src_stride)
{
    return dispatch(h, [&](auto &O) {
        using T = typename std::decay<decltype(O.F[0])>::type;
        if (n_blocks < 0 || n_blocks > O.n_maxcblocks) return AGAL_ERR_ARG;
        const T *in = static_cast<const T *>(src);
        long long nc = n_blocks * O.M_CBLOCK;
        for (int k = 0; k < (O.DIM == 3 ? 4 : 3); k++)
            std::memcpy(O.Fp(O.Q + k), in + k * src_stride, sizeof(T) * nc);
        return AGAL_OK;
    });
}

/* Restates the tau / dx ladders (solver_lbm_init.cu:47-52) for tests. */
int oracle_get_level_params(void *h, double *dxf, double *tau, int n)
{
    return dispatch(h, [&](auto &O) {
        for (int L = 0; L < n && L < O.MAX_LEVELS; L++) {
            dxf[L] = (double)O.dxf_vec[L];
            tau[L] = (double)O.tau_vec[L];
        }
        return AGAL_OK;
    });
}
}
