// lbm_kernels.hip -- the per-level LBM kernels for gfx950.
//
// One 64-lane wavefront owns one 4x4x4 cell-block (3D) or four 4x4 blocks (2D); a 256-thread
// workgroup therefore covers 4 (3D) or 16 (2D) blocks of a level's sorted ID list.
//
// The fused step replaces the reference's three passes Cu_Collide -> Cu_ImposeBC -> Cu_Stream
// (src/solver_lbm/solver_lbm_kernels.cu:204-289, 691-887, 532-633) by one read of the Q
// pre-collision populations of a cell and one write of its Q streamed populations into the other
// buffer (push).  The reference's in-place swap exchanges a link only when the initiating cell is
// not solid and the partner's post-collision value is > 0 (kernels.cu:624); the push reproduces
// that rule exactly:
//   - every cell computes its post-collision values f*_q (BGK for cells that collide; for cells
//     that do not -- solid cells and blocks that are solid with children -- the swap semantics
//     of the reference amount to f*_q = f_{pb(q)});
//   - the domain-face hook modifies f*_q on links that leave the domain (Cu_ImposeBC);
//   - a link (x -> x+c_h, h in the reference's half set) streams when x collides and the
//     partner's f*_{pb(h)} > 0; otherwise both ends bounce back into their own opposite slot.
//     The partner knows its value, the initiator does not: the initiator always pushes, and a
//     partner whose value fails the guard records the link in an exception list that a one-
//     workgroup fix-up kernel resolves after the step (normally empty).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

namespace agal {

constexpr int WG = 256;

// Moments in the reference's order (p-order sums, then division): Cu_RefreshVariables (:110-159).
// The reference adds c * f for every component, zeros included; those terms are left out here.
// A sum that starts at +0 never becomes -0 under round-to-nearest, so adding 0 * f = +-0 leaves
// it unchanged: for finite populations the results are the same bits, at about a third fewer
// vector instructions.
template <typename T, int Q>
__device__ __forceinline__ void moments(const T *f, T &rho, T &u, T &v, T &w)
{
    using VS = VelSet<Q>;
    rho = 0;
    u = 0;
    v = 0;
    w = 0;
#pragma unroll
    for (int p = 0; p < Q; p++) {
        rho += f[p];
        if (VS::cx(p) > 0) u += f[p];
        if (VS::cx(p) < 0) u -= f[p];
        if (VS::cy(p) > 0) v += f[p];
        if (VS::cy(p) < 0) v -= f[p];
        if (VS::DIM == 3 && VS::cz(p) > 0) w += f[p];
        if (VS::DIM == 3 && VS::cz(p) < 0) w -= f[p];
    }
    // One correctly rounded division, then products (the CPU restatement does the same).
    const T r = (T)(1.0) / rho;
    u = u * r;
    v = v * r;
    if (VS::DIM == 3) w = w * r;
}

// c_p . u in the reference's left-to-right order (cx u + cy v) + cz w, without the terms of zero
// components.  A nonzero sum is unchanged by them; a zero sum may differ in sign only, which the
// equilibrium polynomial 1 + 3 c.u + 4.5 (c.u)^2 does not see.
template <typename T, int Q>
__device__ __forceinline__ T cdot(int p, T u, T v, T w)
{
    using VS = VelSet<Q>;
    T s = (T)0;
    bool first = true;
    if (VS::cx(p) != 0) {
        s = VS::cx(p) > 0 ? u : -u;
        first = false;
    }
    if (VS::cy(p) != 0) {
        const T t = VS::cy(p) > 0 ? v : -v;
        s = first ? t : s + t;
        first = false;
    }
    if (VS::DIM == 3 && VS::cz(p) != 0) {
        const T t = VS::cz(p) > 0 ? w : -w;
        s = first ? t : s + t;
    }
    return s;
}

// ---------------------------------------------------------------------------------------------
// Initial conditions: Cu_SetInitialConditions (solver_lbm_kernels.cu:13-77), uniform state hook.
// ---------------------------------------------------------------------------------------------
template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_set_ic(const int *__restrict__ ids, int n, S *__restrict__ F, long long ncm,
                                               const int *__restrict__ cells_mask, T rho, T u, T v, T w)
{
    using VS = VelSet<Q>;
    constexpr int CELLS = VS::CELLS;
    const int t = threadIdx.x % CELLS;
    const int bi = blockIdx.x * (WG / CELLS) + threadIdx.x / CELLS;
    if (bi >= n) return;
    const long long cell = (long long)ids[bi] * CELLS + t;
    const T udotu = u * u + v * v + w * w;
    const int m = cells_mask[cell];
#pragma unroll
    for (int p = 0; p < Q; p++) {
        T cdotu = cdot<T, Q>(p, u, v, w);
        T f = (T)(-1.0);
        if (m != CELL_SOLID) f = rho * (T)VS::w(p) * ((T)1.0 + (T)3.0 * cdotu + (T)4.5 * cdotu * cdotu - (T)1.5 * udotu);
        F[(long long)p * ncm + cell] = Store<T, S>::enc(f, (T)VS::w(p));
    }
}


template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_refresh(const int *__restrict__ ids, int n, const S *__restrict__ F,
                                                T *__restrict__ M, long long ncm)
{
    using VS = VelSet<Q>;
    constexpr int CELLS = VS::CELLS;
    const int t = threadIdx.x % CELLS;
    const int bi = blockIdx.x * (WG / CELLS) + threadIdx.x / CELLS;
    if (bi >= n) return;
    const long long cell = (long long)ids[bi] * CELLS + t;
    T f[Q];
#pragma unroll
    for (int p = 0; p < Q; p++) f[p] = Store<T, S>::dec(F[(long long)p * ncm + cell], (T)VS::w(p));
    T rho, u, v, w;
    moments<T, Q>(f, rho, u, v, w);
    M[0 * ncm + cell] = rho;
    M[1 * ncm + cell] = u;
    M[2 * ncm + cell] = v;
    M[3 * ncm + cell] = (VS::DIM == 3) ? w : (T)0;
}

// ---------------------------------------------------------------------------------------------
// Fused collide + boundary hook + stream of one level sub-step.
// ---------------------------------------------------------------------------------------------
template <typename T, typename S>
struct StepArgs {
    const int *ids;
    int n;
    const S *A;
    S *B;
    long long ncm;
    const int *nbr_slot;   // per block, Q_max neighbour IDs in 3x3(x3) slot order
    const int *part;       // block collides and streams (not solid-with-children)
    const int *simple;     // block and its neighbourhood: all participate, no solid cells
    const int *cells_mask;
    const int *onb;
    T omeg, omegp, U;
    int mrt;      // collision operator: 0 BGK, 1 MRT
    T ck[27];     // MRT: rate_k / D_k per moment row (0 for conserved rows)
    int bc_preset;
    int *exc_idx;
    T *exc_val;
    int *exc_count;
    long long exc_cap;
    // Finest-level pairs (3D): vmode 1 -- the exception list (exc_*) is the extra list that the
    // next sub-step applies as it loads, and each block with an entry gets PEND_FLAG in flags;
    // vmode 2 -- apply that list (va_*) to the loaded values of flagged blocks; 0 -- neither.
    int vmode;
    int *flags;
    const int *va_idx;
    const T *va_val;
    const int *va_count;
    // The extra list linked per block: x_head[block] is the last node 2k + s of the block's chain
    // (entry k, s = 0 through its initiator cell, 1 through its partner cell), -1 when empty;
    // x_next[node] the node before it.  Built in vmode 1, walked and emptied in vmode 2.
    int *x_head;
    int *x_next;
};

// Block flag bit: the block has entries in the finest level's extra exception list (bits 0-1 of
// the same word hold the simple / interior kind).
constexpr int PEND_FLAG = 4;

// Appends a guard exception (target cell, partner cell, direction, partner's stored value) to the
// launch's list; in vmode 1 also flags the two blocks for the next sub-step.
template <typename T, typename S>
__device__ __forceinline__ void record_exception(const StepArgs<T, S> &a, long long tgt, long long cell, int q, T fv)
{
    int k = atomicAdd(&a.exc_count[0], 1);
    if (k < a.exc_cap) {
        a.exc_idx[3 * (long long)k + 0] = (int)tgt;  // cell indices < 2^31
        a.exc_idx[3 * (long long)k + 1] = (int)cell;
        a.exc_idx[3 * (long long)k + 2] = q;
        a.exc_val[k] = fv;
    } else {
        a.exc_count[1] = 1;
    }
    if (a.vmode == 1) {
        const int bt = (int)(tgt >> 6), bc = (int)(cell >> 6);
        atomicOr(&a.flags[bt], PEND_FLAG);
        atomicOr(&a.flags[bc], PEND_FLAG);
        if (k < a.exc_cap) {
            a.x_next[2 * k] = atomicExch(&a.x_head[bt], 2 * k);
            if (bc != bt) a.x_next[2 * k + 1] = atomicExch(&a.x_head[bc], 2 * k + 1);
        }
    }
}

// Workgroup -> position in the level's block list.  Workgroups are dealt round-robin to the
// eight XCDs (each with its own L2); giving every XCD one contiguous range of the sorted ID list
// keeps a block and most of its neighbours -- whose pushes share cache lines -- on one L2.
__device__ __forceinline__ int xcd_local_wg(int nwg_padded)
{
    const int per = nwg_padded / 8;
    return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

// The same for the first nwg_padded (a multiple of 8) workgroups of a larger grid.
__device__ __forceinline__ int xcd_local_wg(int wg, int nwg_padded)
{
    const int per = nwg_padded / 8;
    return (wg % 8) * per + wg / 8;
}

// Post-collision values f*_p of one cell.  BGK: Cu_Collide (solver_lbm_kernels.cu:275-282).  MRT:
// f* = f - M^T (ck .* (M (f - feq))), the transforms summed in ascending row / column order with
// zero entries and conserved rows left out (ck_k = rate_k / D_k).  Cells that do not collide
// (solid cells, blocks that are solid with children) keep their values: the swap semantics of the
// reference amount to f*_q = f_pb(q) for them.
template <typename T, typename S, int Q>
__device__ __forceinline__ void collide_cell(const StepArgs<T, S> &a, const T *f, T *fs, T rho, T u, T v, T w,
                                             bool inert)
{
    using VS = VelSet<Q>;
    if (inert) {
#pragma unroll
        for (int p = 0; p < Q; p++) fs[p] = f[VS::pb(p)];
        return;
    }
    const T udotu = u * u + v * v + w * w;
    if (!a.mrt) {
        // BGK in fused multiply-adds, in the same order as the CPU restatement:
        //   b_p = (4.5 c.u) c.u + (1 - 1.5 u.u),  feq_p / (w_p rho) = 3 c.u + b_p,
        //   f*_p = (w_p rho omega) (feq_p / (w_p rho)) + (1 - omega) f_p.
        // A pair (p, pb(p)) shares |c.u|, b and the weight: c_pb.u = -(c_p.u) bit for bit (sums of
        // negated terms), so the pb direction takes -3 c.u.
        const T c1 = (T)(1.0) - (T)(1.5) * udotu;
        const T ro = rho * a.omeg;
        fs[0] = fmadd((T)VS::w(0) * ro, c1, f[0] * a.omegp);
#pragma unroll
        for (int p = 1; p < Q; p++) {
            if (!VS::half(p)) continue;
            const int q = VS::pb(p);
            const T cu = cdot<T, Q>(p, u, v, w);
            const T b = fmadd((T)(4.5) * cu, cu, c1);
            const T wro = (T)VS::w(p) * ro;
            fs[p] = fmadd(wro, fmadd((T)(3.0), cu, b), f[p] * a.omegp);
            fs[q] = fmadd(wro, fmadd((T)(-3.0), cu, b), f[q] * a.omegp);
        }
        return;
    }
    using MB = Mrt<Q>;
    // Full unrolling keeps d, r and the rates in registers (otherwise the FP64 D3Q27 variant indexes
    // them dynamically from scratch memory).
    T d[Q];
#pragma clang loop unroll(full)
    for (int p = 0; p < Q; p++) {
        T cdotu = cdot<T, Q>(p, u, v, w);
        d[p] = f[p] - (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
    }
    T r[Q];
#pragma clang loop unroll(full)
    for (int k = 0; k < Q; k++) {
        r[k] = (T)0;
        if (MB::rate(k) == 0) continue;
        T acc = (T)0;
        bool first = true;
#pragma clang loop unroll(full)
        for (int p = 0; p < Q; p++) {
            if (MB::m(k, p) == 0) continue;
            const T term = (T)MB::m(k, p) * d[p];
            acc = first ? term : acc + term;
            first = false;
        }
        r[k] = acc * a.ck[k];
    }
#pragma clang loop unroll(full)
    for (int p = 0; p < Q; p++) {
        T acc = (T)0;
        bool first = true;
#pragma clang loop unroll(full)
        for (int k = 0; k < Q; k++) {
            if (MB::rate(k) == 0 || MB::m(k, p) == 0) continue;
            const T term = (T)MB::m(k, p) * r[k];
            acc = first ? term : acc + term;
            first = false;
        }
        fs[p] = f[p] - acc;
    }
}

// MRT transform on matrix cores for one 3D block per wavefront (Q = 19, 27; FP32 compute):
// g = M^T (c . M d) for the non-equilibrium parts d of the 64 cells.  FP32 pools: d and c . M d are
// split into two bf16 terms each (hi + lo carry 16 significant bits; the products with the integer
// matrix are exact and the sums FP32), so the result agrees with the vector path to about 2^-17 of
// d, below the FP32 rounding of the populations themselves.  16-bit pools (F16): one binary16 term
// of 2^10 d and of 2^10 c . M d each (11 significant bits, the scale keeps small parts out of the
// subnormal range), products exact and sums FP32, on v_mfma_f32_16x16x32_f16: an error of about
// 2^-11 of the non-equilibrium part, below the binary16 rounding of the stored deviation f - w,
// which is larger than that part, and the correction goes back through LDS in binary16.  `lds` is
// this wavefront's region of 64 rows (36 floats; 80 bytes with 16-bit pools); only the wavefront
// itself touches it.
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4_t;

__constant__ const MfmaTab<19> kMfma19 = make_mfma_tab<19>();
__constant__ const MfmaTab<27> kMfma27 = make_mfma_tab<27>();

constexpr int MFMA_SPLITS = 2;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Next bf16 term of `rem` (round to nearest even, the hardware conversion); rem keeps the rest.
__device__ __forceinline__ __bf16 split_term(float &rem)
{
    const __bf16 h = (__bf16)rem;
    rem -= (float)h;
    return h;
}

template <int Q, bool F16>
__device__ __forceinline__ void mrt_mfma(const float *d, float *g, const float *ckrow, float *lds)
{
    using VT = typename std::conditional<F16, f16x8_t, bf16x8_t>::type;
    constexpr int SPLITS = F16 ? 1 : MFMA_SPLITS;
    const MfmaTab<Q> &tab = *(Q == 19 ? reinterpret_cast<const MfmaTab<Q> *>(&kMfma19)
                                      : reinterpret_cast<const MfmaTab<Q> *>(&kMfma27));
    const int l = threadIdx.x & 63, q = l >> 4, j = l & 15;
    const u32x4_t *fw = reinterpret_cast<const u32x4_t *>(F16 ? &tab.fw16[l][0][0] : &tab.fw[l][0][0]);
    const u32x4_t *bk = reinterpret_cast<const u32x4_t *>(F16 ? &tab.bk16[l][0][0] : &tab.bk[l][0][0]);
    const VT afw0 = __builtin_bit_cast(VT, fw[0]), afw1 = __builtin_bit_cast(VT, fw[1]);
    const VT abk0 = __builtin_bit_cast(VT, bk[0]), abk1 = __builtin_bit_cast(VT, bk[1]);
    VT *ldsv = reinterpret_cast<VT *>(lds);  // rows of 5 vectors (80 bytes)
    auto mfma = [](const VT &a, const VT &b, const f32x4_t &c) {
        if constexpr (F16)
            return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
        else
            return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    };
    auto term = [](float &rem) {
        if constexpr (F16) {
            return (_Float16)rem;
        } else {
            return split_term(rem);
        }
    };

    // Forward: one pass per term of d through a 64-row transpose region.
    f32x4_t C0[4], C1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) C0[t] = C1[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    float rem[32];
#pragma unroll
    for (int k = 0; k < 32; k++) rem[k] = k < Q ? d[k] : 0.f;
#pragma unroll
    for (int sp = 0; sp < SPLITS; sp++) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            VT v;
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = term(rem[8 * c + e]);
            ldsv[5 * l + c] = v;
        }
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const VT b = ldsv[5 * (16 * t + j) + q];
            C0[t] = mfma(afw0, b, C0[t]);
            C1[t] = mfma(afw1, b, C1[t]);
        }
        wave_lds_sync();
    }
    // Relaxation factors rate / D of the rows this lane holds (rows 4q + r and 16 + 4q + r, zero
    // beyond the velocity set), picked from the level's row table by the lane's quarter q.
    float ck0[4], ck1[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float a0[4], a1[4];
#pragma unroll
        for (int qq = 0; qq < 4; qq++) {
            a0[qq] = 4 * qq + r < Q ? ckrow[4 * qq + r] : 0.f;
            a1[qq] = 16 + 4 * qq + r < Q ? ckrow[16 + 4 * qq + r] : 0.f;
        }
        ck0[r] = q == 0 ? a0[0] : (q == 1 ? a0[1] : (q == 2 ? a0[2] : a0[3]));
        ck1[r] = q == 0 ? a1[0] : (q == 1 ? a1[1] : (q == 2 ? a1[2] : a1[3]));
    }
    // Back: B operand straight from the forward results of the same lane.
    f32x4_t G0[4], G1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        G0[t] = G1[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        float v[8];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            v[r] = C0[t][r] * ck0[r];
            v[4 + r] = C1[t][r] * ck1[r];
        }
#pragma unroll
        for (int sp = 0; sp < SPLITS; sp++) {
            VT b;
#pragma unroll
            for (int e = 0; e < 8; e++) b[e] = term(v[e]);
            G0[t] = mfma(abk0, b, G0[t]);
            G1[t] = mfma(abk1, b, G1[t]);
        }
    }
    if constexpr (F16) {
        // Transpose back in binary16 (the scaled correction, 11 significant bits, far below the
        // storage step of the deviation it corrects): rows of 40 halves (80 bytes) per cell.
        typedef __attribute__((ext_vector_type(4))) _Float16 f16x4_t;
        _Float16 *ldsh = reinterpret_cast<_Float16 *>(lds);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f16x4_t a, b;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                a[r] = (_Float16)G0[t][r];
                b[r] = (_Float16)G1[t][r];
            }
            *reinterpret_cast<f16x4_t *>(ldsh + 40 * (16 * t + j) + 4 * q) = a;
            *reinterpret_cast<f16x4_t *>(ldsh + 40 * (16 * t + j) + 16 + 4 * q) = b;
        }
        wave_lds_sync();
        const f16x8_t *row = reinterpret_cast<const f16x8_t *>(ldsh + 40 * l);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const f16x8_t h = row[c];
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (8 * c + e < Q) g[8 * c + e] = (float)h[e];
        }
        wave_lds_sync();
    } else {
        // Transpose back: G row p of cell 16t + j -> LDS row of that cell (36 floats), then own row.
#pragma unroll
        for (int t = 0; t < 4; t++) {
            *reinterpret_cast<f32x4_t *>(lds + 36 * (16 * t + j) + 4 * q) = G0[t];
            *reinterpret_cast<f32x4_t *>(lds + 36 * (16 * t + j) + 16 + 4 * q) = G1[t];
        }
        wave_lds_sync();
#pragma unroll
        for (int p = 0; p < Q; p++) g[p] = lds[36 * l + p];
        wave_lds_sync();
    }
}

// Population load of the fused step.  FP32/FP64 rows are read with non-temporal loads: they are
// used once, and keeping them out of L2 leaves it to the pushes, whose partial cache lines from
// neighbouring blocks merge there (measured: the FP32 fused step runs ~20% faster).  16-bit rows
// measured slightly slower that way and use ordinary loads.
template <typename S>
__device__ __forceinline__ S load_pop(const S *p)
{
    if constexpr (sizeof(S) >= 4)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <typename S>
__device__ __forceinline__ S load_pop(const __attribute__((address_space(1))) S *p)
{
    if constexpr (sizeof(S) >= 4)
        return __builtin_nontemporal_load(p);
    else if constexpr (std::is_same<S, bf16s>::value)
        return bf16s{*reinterpret_cast<const __attribute__((address_space(1))) uint16_t *>(p)};
    else
        return *p;
}

// A global-memory pointer that is the same on every lane, forced into scalar registers: a lane's
// element of it is then a scalar base plus a 32-bit offset, with no 64-bit vector address
// arithmetic per access.
template <typename S>
__device__ __forceinline__ const __attribute__((address_space(1))) S *scalar_row(const S *p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (const __attribute__((address_space(1))) S *)(((unsigned long long)hi << 32) | lo);
}

// Store of the fused step into a row (the same pointer on every lane) at element idx.  A 16-bit
// row holds fewer than 2^32 bytes (the pool has fewer than 2^31 cells, checked at creation), so the
// address is a scalar base plus a 32-bit byte offset, with no 64-bit vector address arithmetic per
// store; wider rows keep 64-bit addresses.
template <typename S>
__device__ __forceinline__ void store_pop(S *row, unsigned idx, S v)
{
    if constexpr (sizeof(S) == 2) {
        const unsigned long long r = reinterpret_cast<unsigned long long>(row);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)r), hi = __builtin_amdgcn_readfirstlane((unsigned)(r >> 32));
        auto *base = (__attribute__((address_space(1))) char *)(((unsigned long long)hi << 32) | lo);
        *(__attribute__((address_space(1))) uint16_t *)(base + idx * 2u) = __builtin_bit_cast(uint16_t, v);
    } else {
        row[idx] = v;
    }
}

// Store of a 16-bit value at a byte offset within a row (the same pointer on every lane).
template <typename S>
__device__ __forceinline__ void store_pop_bytes(S *row, unsigned byte_off, S v)
{
    static_assert(sizeof(S) == 2, "16-bit rows only");
    const unsigned long long r = reinterpret_cast<unsigned long long>(row);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)r), hi = __builtin_amdgcn_readfirstlane((unsigned)(r >> 32));
    auto *base = (__attribute__((address_space(1))) char *)(((unsigned long long)hi << 32) | lo);
    *(__attribute__((address_space(1))) uint16_t *)(base + byte_off) = __builtin_bit_cast(uint16_t, v);
}

// One link of the fused step for the cell `cell`, direction q >= 1, whose neighbour block in that
// direction is nbk (target cell tgt): boundary hook, swap guard, and the single store of f*_q
// either into the neighbour (the link streams) or into the cell's own opposite slot (bounce-back).
template <typename T, typename S, int Q>
__device__ __forceinline__ void push_link(const StepArgs<T, S> &a, int q, T fq, int nbk, long long tgt, long long cell,
                                          bool inert, bool simple, bool onb, T u, T v, T w)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    // Domain-face hook on links that leave the domain (Cu_ImposeBC, onb blocks, non-solid cells).
    if (onb && nbk < 0 && nbk != N_SKIPID)
        fq = bc_hook<T>(a.bc_preset, a.U, nbk, fq, u, v, w, (T)VS::w(q), (T)VS::cx(q), (T)VS::cy(q), (T)VS::cz(q));
    // In a 16-bit pool the guard below sees the stored value.
    fq = quant<T, S>(fq, (T)VS::w(q));
    bool push;
    bool exc = false;
    if (VS::half(q)) {
        // Initiator of the link x -> x + c_q.
        push = !inert && nbk >= 0;
        if (!simple && push) {
            const bool p_inert = !a.part[nbk] || a.cells_mask[tgt] == CELL_SOLID;
            if (p_inert) push = ST::dec(a.A[(long long)q * a.ncm + tgt], (T)VS::w(q)) > (T)0;
        }
    } else {
        // Partner of the link initiated by the cell x + c_q in direction pb(q).
        push = nbk >= 0;
        if (!simple && push) push = a.part[nbk] && a.cells_mask[tgt] != CELL_SOLID;
        if (push && !(fq > (T)0)) {
            push = false;
            exc = !inert;  // a colliding partner cannot tell the initiator in time
        }
    }
    if (exc) {
        record_exception(a, tgt, cell, q, fq);
        return;
    }
    const long long at = push ? (long long)q * a.ncm + tgt : (long long)VS::pb(q) * a.ncm + cell;
    a.B[at] = ST::enc(fq, (T)VS::w(q));
}

// Stream of an interior block (block flag simple == 2): every neighbour slot of the velocity set
// exists, and the block and its neighbours participate with no solid cell.  No link then leaves
// the domain or the level, no guard read is needed, half-set links always push, and the partner
// side pushes unless its stored value fails the guard (then the link goes to the exception list,
// as in push_link).  The target block comes from the wave's neighbour table (lane s holds slot s)
// through ds_bpermute; the target cell from a carry-free add of the link to the lane's (I,J,K)
// 2-bit fields.  Same stores, same values as push_link on these blocks, with about a fifth of its
// vector instructions.
template <typename T, typename S, int Q>
__device__ __forceinline__ void push_interior(const StepArgs<T, S> &a, const T *fs, int my_nb, int t, unsigned cell)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    // ds_bpermute byte steps to the neighbour slot along each axis, nonzero on the block's faces.
    const int xp = I == 3 ? 4 : 0, xm = I == 0 ? -4 : 0;
    const int yp = J == 3 ? 12 : 0, ym = J == 0 ? -12 : 0;
    const int zp = K == 3 ? 36 : 0, zm = K == 0 ? -36 : 0;
    // 16-bit rows: the target as a byte offset within the row (block << 7 | cell << 1), the same
    // field-wise sum on fields shifted up by one bit, so the store needs no further shift.
    constexpr int SH = sizeof(S) == 2 ? 1 : 0;
    const unsigned tl = ((unsigned)t & 0x15u) << SH, th = ((unsigned)t & 0x2Au) << SH;
    const int nb_src = (int)((unsigned)my_nb << (6 + SH));
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int cx = VS::cx(q), cy = VS::cy(q), cz = VS::cz(q);
        int sa = 13 * 4;
        if (cx > 0) sa += xp;
        if (cx < 0) sa += xm;
        if (cy > 0) sa += yp;
        if (cy < 0) sa += ym;
        if (cz > 0) sa += zp;
        if (cz < 0) sa += zm;
        const unsigned nbk = (unsigned)__builtin_amdgcn_ds_bpermute(sa, nb_src);
        // ((I+cx)&3) + 4((J+cy)&3) + 16((K+cz)&3): field-wise sum without carries between fields.
        const unsigned d = (unsigned)((cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4));
        const unsigned tgt_s = nbk + ((tl + ((d & 0x15u) << SH)) ^ th ^ ((d & 0x2Au) << SH));
        const S e = ST::enc(fs[q], (T)VS::w(q));
        bool push = true;
        if (!VS::half(q)) {
            // Partner of the link initiated by the cell x + c_q: the guard on the stored value.
            if (!ST::positive(e, (T)VS::w(q))) {
                push = false;
                record_exception(a, (long long)(tgt_s >> SH), (long long)cell, q, ST::dec(e, (T)VS::w(q)));
            }
        }
        if (push) {
            if constexpr (SH == 1)
                store_pop_bytes(a.B + (long long)q * a.ncm, tgt_s, e);
            else
                store_pop(a.B + (long long)q * a.ncm, tgt_s, e);
        }
    }
}

// Stream of a simple block that is not interior (block flag simple == 1): the block and its
// existing neighbours participate with no solid cell, but some neighbour slots are missing (a
// domain face, or a coarser region).  Same lookup and target arithmetic as push_interior; a link
// whose neighbour slot is missing bounces back into the cell's own opposite slot, after the face
// hook when the block lies on the domain boundary (onb) and the slot is a face -- push_link's
// rules for non-inert cells of simple blocks.
template <typename T, typename S, int Q>
__device__ __forceinline__ void push_simple(const StepArgs<T, S> &a, const T *fs, int my_nb, int t, unsigned cell,
                                            bool onb, T u, T v, T w)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const int xp = I == 3 ? 4 : 0, xm = I == 0 ? -4 : 0;
    const int yp = J == 3 ? 12 : 0, ym = J == 0 ? -12 : 0;
    const int zp = K == 3 ? 36 : 0, zm = K == 0 ? -36 : 0;
    const unsigned tl = (unsigned)t & 0x15u, th = (unsigned)t & 0x2Au;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int cx = VS::cx(q), cy = VS::cy(q), cz = VS::cz(q);
        int sa = 13 * 4;
        if (cx > 0) sa += xp;
        if (cx < 0) sa += xm;
        if (cy > 0) sa += yp;
        if (cy < 0) sa += ym;
        if (cz > 0) sa += zp;
        if (cz < 0) sa += zm;
        const int nbk = __builtin_amdgcn_ds_bpermute(sa, my_nb);
        const unsigned d = (unsigned)((cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4));
        const unsigned tgt = ((unsigned)nbk << 6) + ((tl + (d & 0x15u)) ^ th ^ (d & 0x2Au));
        T fq = fs[q];
        if (onb && nbk < 0 && nbk != N_SKIPID)
            fq = bc_hook<T>(a.bc_preset, a.U, nbk, fq, u, v, w, (T)VS::w(q), (T)VS::cx(q), (T)VS::cy(q), (T)VS::cz(q));
        const S e = ST::enc(fq, (T)VS::w(q));
        bool push = nbk >= 0;
        if (!VS::half(q) && push) {
            // Partner of the link initiated by the cell x + c_q: the guard on the stored value.
            if (!ST::positive(e, (T)VS::w(q))) {
                record_exception(a, (long long)tgt, (long long)cell, q, ST::dec(e, (T)VS::w(q)));
                continue;  // no store: the fixup places both values
            }
        }
        if (push)
            store_pop(a.B + (long long)q * a.ncm, tgt, e);
        else
            store_pop(a.B + (long long)VS::pb(q) * a.ncm, cell, e);
    }
}

// The extra exception list of the previous finest sub-step, applied to this block's loaded values:
// for an entry (initiator cell i, partner cell c, direction q) the fixup would move the value at
// (pb(q), c) -- the initiator's push -- to (q, i) and store the partner's value at (pb(q), c).
// The previous sub-step's output is this sub-step's input, read by no one else in between, so
// the same values are taken here as the loads' results instead.  No two entries share a slot (an
// entry's q is the partner's, non-initiating direction and pb(q) an initiating one), so the order
// of application is free.  The wavefront walks its block's chain (the entries with i or c in the
// block, linked when they were recorded), so the work per block is its own entries however long
// the list is; the chain is emptied for the next pair.
template <typename T, typename S, int Q>
__device__ __forceinline__ void apply_pending(const StepArgs<T, S> &a, T *f, int b, int t)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const unsigned my = ((unsigned)b << 6) | (unsigned)t;
    int e = __builtin_amdgcn_readfirstlane(a.x_head[b]);
    while (e >= 0) {
        const int k = e >> 1;
        const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane(a.va_idx[3 * (long long)k + 0]);
        const unsigned hc = (unsigned)__builtin_amdgcn_readfirstlane(a.va_idx[3 * (long long)k + 1]);
        const int hq = __builtin_amdgcn_readfirstlane(a.va_idx[3 * (long long)k + 2]);
        int qb = 0;
        T wq = 0;
#pragma unroll
        for (int p = 0; p < Q; p++)
            if (p == hq) {
                qb = VS::pb(p);
                wq = (T)VS::w(p);
            }
        // Through storage, as the fixup's store and this sub-step's load would take them.
        const T vi = ST::dec(a.A[(long long)qb * a.ncm + hc], wq);
        const T vc = ST::dec(ST::enc(a.va_val[k], wq), wq);
        const bool at_i = hi == my, at_c = hc == my;
#pragma unroll
        for (int p = 1; p < Q; p++) {
            if (at_i && p == hq) f[p] = vi;
            if (at_c && p == qb) f[p] = vc;
        }
        e = __builtin_amdgcn_readfirstlane(a.x_next[e]);
    }
    if (t == 0) a.x_head[b] = -1;
}

// 3D: one wavefront per block, so the block, its flags and its 27 neighbour IDs are wave-uniform
// (scalar registers); each lane picks its link's neighbour among at most eight of them with
// selects on its own cell position -- no LDS traffic and no per-direction waits.  Interior blocks
// take push_interior instead.
template <typename T, typename S, int Q, bool MF>
__device__ __forceinline__ void collide_stream_3d(const StepArgs<T, S> &a, int wave, float *lds)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const int t = threadIdx.x & 63;
    const int bi = wave;  // one block per wavefront
    if (bi >= a.n) return;  // whole wavefronts only; the kernel has no barrier
    const int b = __builtin_amdgcn_readfirstlane(a.ids[bi]);
    // Cell index within a population row (the pool holds fewer than 2^32 cells, checked at
    // creation), so rows are addressed as a row base plus a 32-bit index.
    const unsigned cell = (unsigned)b * 64u + (unsigned)t;
    // Lane s < 27 holds neighbour slot s (lanes above repeat slot 26): a load on every lane keeps
    // the block index out of divergent control flow, so it stays a scalar.
    const int my_nb = a.nbr_slot[(long long)b * 27 + (t < 27 ? t : 26)];
    const int flags_u = __builtin_amdgcn_readfirstlane(a.simple[b]);
    if (flags_u & LINK_FLAG) return;  // a linked block: its sub-step is k_link_step's
    const int simple_u = flags_u & 3;
    const int part_u = __builtin_amdgcn_readfirstlane(a.part[b]);
    const int onb_u = __builtin_amdgcn_readfirstlane(a.onb[b]);
    T f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++)  // row q of block b: a scalar base, lane t at offset t
        f[q] = ST::dec(load_pop(scalar_row(a.A + (long long)q * a.ncm + (long long)b * 64) + t), (T)VS::w(q));
    if (a.vmode == 2 && (flags_u & PEND_FLAG)) {
        apply_pending<T, S, Q>(a, f, b, t);
        if (t == 0) a.flags[b] = simple_u;  // this sub-step sets no flags: no other writer
    }
    const bool simple = simple_u != 0;
    int mask = CELL_INTERIOR;
    if (!simple) mask = a.cells_mask[cell];
    const bool inert = part_u == 0 || mask == CELL_SOLID;
    const bool onb = onb_u == 1 && mask != CELL_SOLID;

    T fs[Q];
    T rho, u, v, w;
    moments<T, Q>(f, rho, u, v, w);
    if constexpr (MF) {
        // MRT with the moment transform on matrix cores (whole wavefront, inert lanes carry zeros).
        // The non-equilibrium parts go in scaled by SC (16-bit pools: 2^10, see mrt_mfma) and the
        // correction comes back scaled by SC; the equilibrium in pairs (p, pb(p)) with fused
        // multiply-adds, SC folded into the density.
        constexpr float SC = ST::quantized ? 1024.0f : 1.0f;
        const T c1 = (T)(1.0) - (T)(1.5) * (u * u + v * v + w * w);
        const T rs = rho * (T)SC;
        float d[Q], g[Q];
        d[0] = inert ? 0.f : fmadd(f[0], (T)SC, -((T)VS::w(0) * rs * c1));
#pragma unroll
        for (int p = 1; p < Q; p++) {
            if (!VS::half(p)) continue;
            const int q = VS::pb(p);
            const T cu = cdot<T, Q>(p, u, v, w);
            const T b = fmadd((T)(4.5) * cu, cu, c1);
            const T wr = (T)VS::w(p) * rs;
            d[p] = inert ? 0.f : fmadd(f[p], (T)SC, -(wr * fmadd((T)(3.0), cu, b)));
            d[q] = inert ? 0.f : fmadd(f[q], (T)SC, -(wr * fmadd((T)(-3.0), cu, b)));
        }
        mrt_mfma<Q, ST::quantized>(d, g, a.ck, lds);
#pragma unroll
        for (int p = 0; p < Q; p++) fs[p] = inert ? f[VS::pb(p)] : fmadd((T)g[p], (T)(-1.0f / SC), f[p]);
    } else {
        collide_cell<T, S, Q>(a, f, fs, rho, u, v, w, inert);
    }

    store_pop(a.B, cell, ST::enc(fs[0], (T)VS::w(0)));
    if (simple_u == 2) {
        push_interior<T, S, Q>(a, fs, my_nb, t, cell);
        return;
    }
    if (simple_u == 1) {
        push_simple<T, S, Q>(a, fs, my_nb, t, cell, onb_u == 1, u, v, w);
        return;
    }
    int nbs[27];
#pragma unroll
    for (int s = 0; s < 27; s++) nbs[s] = __builtin_amdgcn_readlane(my_nb, s);
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int cx = VS::cx(q), cy = VS::cy(q), cz = VS::cz(q);
        // Does the link leave the block along x / y / z?
        const bool ex = cx > 0 ? I == 3 : (cx < 0 ? I == 0 : false);
        const bool ey = cy > 0 ? J == 3 : (cy < 0 ? J == 0 : false);
        const bool ez = cz > 0 ? K == 3 : (cz < 0 ? K == 0 : false);
        // Slot (dx+1) + 3(dy+1) + 9(dz+1) of the neighbour block; 13 is the block itself.
        const int sxyz = nbs[13 + cx + 3 * cy + 9 * cz], sxy = nbs[13 + cx + 3 * cy], sxz = nbs[13 + cx + 9 * cz],
                  syz = nbs[13 + 3 * cy + 9 * cz], sx = nbs[13 + cx], sy = nbs[13 + 3 * cy], sz = nbs[13 + 9 * cz];
        const int nbk = ex ? (ey ? (ez ? sxyz : sxy) : (ez ? sxz : sx)) : (ey ? (ez ? syz : sy) : (ez ? sz : b));
        const long long tgt = (long long)nbk * 64 + (((I + cx) & 3) + 4 * ((J + cy) & 3) + 16 * ((K + cz) & 3));
        push_link<T, S, Q>(a, q, fs[q], nbk, tgt, (long long)cell, inert, simple, onb, u, v, w);
    }
}

template <typename T, typename S, int Q>
__device__ __forceinline__ void collide_stream_2d(const StepArgs<T, S> &a, int wave);

// One launch of the fused step over the blocks of up to MAXLV levels: the levels' wavefront slots
// are consecutive ranges [start[k], start[k+1]) of the launch (one 3D block or four 2D blocks per
// slot), dealt to the XCDs in contiguous ranges.  Levels of one launch touch disjoint cells and
// their own exception lists, so one launch does what separate launches of each level would.
constexpr int MAXLV = 4;

template <typename T, typename S>
struct MultiStep {
    int nlv;
    int start[MAXLV + 1];
    StepArgs<T, S> lv[MAXLV];
};

// 2D: four 4x4 blocks per wavefront; neighbour IDs are distributed with ds_bpermute.
template <typename T, typename S, int Q>
__device__ __forceinline__ void collide_stream_2d(const StepArgs<T, S> &a, int wave)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    constexpr int CELLS = VS::CELLS, DIM = VS::DIM, QMAX = VS::QMAX;
    const int t = threadIdx.x % CELLS;
    const int lane = threadIdx.x % 64;
    const int gbase = lane - t;  // lane of cell 0 of this block within the wavefront
    const int bi = wave * (64 / CELLS) + lane / CELLS;
    const bool active = bi < a.n;
    const int b = active ? a.ids[bi] : 0;
    // Lane-parallel load of the block's neighbour table; distributed with ds_bpermute (every
    // __shfl below sits outside divergent control flow).
    const int my_nb = (active && t < QMAX) ? a.nbr_slot[(long long)b * QMAX + t] : N_SKIPID;
    const int I = t % 4, J = (t / 4) % 4, K = (DIM == 3) ? t / 16 : 0;
    const long long cell = (long long)b * CELLS + t;

    bool simple = true, inert = false, onb = false;
    int mask = CELL_INTERIOR;
    T f[Q];
    if (active) {
        simple = (a.simple[b] & 3) != 0;
        const bool part = a.part[b] != 0;
        if (!simple) mask = a.cells_mask[cell];
        inert = !part || mask == CELL_SOLID;
        onb = a.onb[b] == 1 && mask != CELL_SOLID;
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = ST::dec(load_pop(&a.A[(long long)q * a.ncm + cell]), (T)VS::w(q));
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = (T)1;
    }

    T fs[Q];
    T rho, u, v, w;
    moments<T, Q>(f, rho, u, v, w);
    collide_cell<T, S, Q>(a, f, fs, rho, u, v, w, inert);

    if (active) a.B[cell] = ST::enc(fs[0], (T)VS::w(0));
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int Ip = I + VS::cx(q), Jp = J + VS::cy(q), Kp = K + VS::cz(q);
        const int slot = (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) +
                         (DIM == 3 ? 9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1)) : 0);
        const int nbk = __shfl(my_nb, gbase + slot, 64);
        if (!active) continue;
        const long long tgt = (long long)nbk * CELLS + ((Ip & 3) + 4 * (Jp & 3) + (DIM == 3 ? 16 * (Kp & 3) : 0));
        push_link<T, S, Q>(a, q, fs[q], nbk, tgt, cell, inert, simple, onb, u, v, w);
    }
}

// Resolves guard exceptions: the initiator's pushed value returns to it (bounce-back) and the
// partner keeps its own post-collision value.  One workgroup per level of the fused launch, each on
// its level's list; it clears the list's counter when done.
template <typename T, typename S>
struct FixupArgs {
    long long ncm, exc_cap;
    S *B[MAXLV];
    const int *exc_idx[MAXLV];
    const T *exc_val[MAXLV];
    int *exc_count[MAXLV];
    int *va_count[MAXLV];  // the finest level: its extra list's counter, cleared here, else null
};

template <typename T, typename S, int Q>
__device__ __forceinline__ void fixup_level(const FixupArgs<T, S> &fa, int lv)
{
    using VS = VelSet<Q>;
    S *__restrict__ B = fa.B[lv];
    const int *__restrict__ exc_idx = fa.exc_idx[lv];
    const T *__restrict__ exc_val = fa.exc_val[lv];
    int *exc_count = fa.exc_count[lv];
    const long long ncm = fa.ncm;
    // The extra list was applied by the sub-step that this list comes from; it starts empty for the
    // next pair.
    if (fa.va_count[lv] && threadIdx.x == 0 && fa.va_count[lv][0] != 0) fa.va_count[lv][0] = 0;
    const int cnt = exc_count[0];
    if (cnt == 0) return;  // the usual case: nothing to resolve, nothing to reset (whole workgroup)
    const long long n = min((long long)cnt, fa.exc_cap);
    for (long long k = threadIdx.x; k < n; k += WG) {
        const long long i = (unsigned)exc_idx[3 * k + 0], c = (unsigned)exc_idx[3 * k + 1];
        const int q = exc_idx[3 * k + 2];
        int qb = 0;
        T wq = 0;
#pragma unroll
        for (int p = 0; p < Q; p++)
            if (p == q) {
                qb = VS::pb(p);
                wq = (T)VS::w(p);
            }
        // Rows q and pb(q) share one weight: a stored value moves between them unchanged.
        B[(long long)q * ncm + i] = B[(long long)qb * ncm + c];
        B[(long long)qb * ncm + c] = Store<T, S>::enc(exc_val[k], wq);
    }
    __syncthreads();
    if (threadIdx.x == 0) exc_count[0] = 0;
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_fixup(FixupArgs<T, S> fa)
{
    fixup_level<T, S, Q>(fa, blockIdx.x);
}

// ---------------------------------------------------------------------------------------------
// Coarse -> fine: Cu_Interpolate_Linear (solver_lbm_kernels_comm.cu:13-138).  The first CELLS
// lanes of each parent put the Q rescaled parent populations in LDS; then every lane takes one
// cell in each half of the children (children c and c + NCH/2), so a parent's children are
// written by half as many wavefronts with row-contiguous stores.  Half-width groups keep a level's
// interface parents within one round of resident wavefronts (a 3D parent is one 256-lane
// workgroup; 2D packs eight 4x4 parents).
// ---------------------------------------------------------------------------------------------
constexpr int IWG = 256;

// Interpolation coefficients of the reference's trilinear (bilinear) form, per parent, child
// (octant) and direction: the corner differences that multiply 1, Ix, Iy, Iz, IxIy, IxIz, IyIz,
// IxIyIz (1, Ix, Iy, IxIy in 2D), formed once from the parent's rescaled populations s (rows of
// `stride` per direction) with the reference's operand order, instead of once per child cell.
template <typename T, int Q>
struct InterpCoef {
    static constexpr int DIM = VelSet<Q>::DIM, NCH = VelSet<Q>::CHILDREN, N = DIM == 3 ? 8 : 4;
};

// Fills coef[(g * Q + p) * NCH + c][0..N) for the PPW parents of the workgroup (tasks strided over
// the workgroup's lanes).
template <typename T, int Q, int PPW, int NTHR>
__device__ __forceinline__ void interp_coefficients(const T *s_all, T *coef)
{
    using IC = InterpCoef<T, Q>;
    constexpr int CELLS = VelSet<Q>::CELLS, NCH = IC::NCH, N = IC::N;
    for (int task = threadIdx.x; task < PPW * Q * NCH; task += NTHR) {
        const int c = task % NCH, p = (task / NCH) % Q, g = task / (NCH * Q);
        const T *s = s_all + p * PPW * CELLS + g * CELLS;
        const int ci = c % 2, cj = (c / 2) % 2, ck = (c / 4) % 2;
        T *a = coef + (long long)task * N;
        if (IC::DIM == 2) {
            const int o = 2 * ci + 2 * 4 * cj;
            a[0] = s[0 + o];
            a[1] = s[1 + o] - s[0 + o];
            a[2] = s[4 + o] - s[0 + o];
            a[3] = s[5 + o] - s[4 + o] - s[1 + o] + s[0 + o];
        } else {
            const int o = 2 * ci + 2 * 4 * cj + 2 * 4 * 4 * ck;
            a[0] = s[0 + o];
            a[1] = s[1 + o] - s[0 + o];
            a[2] = s[4 + o] - s[0 + o];
            a[3] = s[16 + o] - s[0 + o];
            a[4] = s[5 + o] - s[1 + o] - s[4 + o] + s[0 + o];
            a[5] = s[17 + o] - s[1 + o] - s[16 + o] + s[0 + o];
            a[6] = s[20 + o] - s[4 + o] - s[16 + o] + s[0 + o];
            a[7] = s[21 + o] + s[1 + o] + s[4 + o] + s[16 + o] - s[5 + o] - s[17 + o] - s[20 + o] - s[0 + o];
        }
    }
}

// Value at child c's cell t in the reference's operation order (Cu_Interpolate_Linear,
// solver_lbm_kernels_comm.cu:13-138): s0 + a1 Ix + a2 Iy + a3 Iz + a4 Ix Iy + ... summed left to
// right, each product formed left to right; stored into every row.
template <typename T, typename S, int Q>
__device__ __forceinline__ void interp_child(const T *coef_g, int c, int t, S *__restrict__ Fc, long long ncm,
                                             long long ccell)
{
    using VS = VelSet<Q>;
    using IC = InterpCoef<T, Q>;
    constexpr int DIM = VS::DIM, NCH = IC::NCH, N = IC::N;
    const int I = t % 4, J = (t / 4) % 4, K = (DIM == 3) ? t / 16 : 0;
    const T Ix = (T)(-0.25) + (T)(0.5) * (T)I;
    const T Iy = (T)(-0.25) + (T)(0.5) * (T)J;
    const T Iz = (DIM == 3) ? (T)(-0.25) + (T)(0.5) * (T)K : (T)0;
#pragma unroll
    for (int p = 0; p < Q; p++) {
        const T *a = coef_g + (p * NCH + c) * N;
        T val;
        if (DIM == 2)
            val = a[0] + a[1] * Ix + a[2] * Iy + a[3] * Ix * Iy;
        else
            val = a[0] + a[1] * Ix + a[2] * Iy + a[3] * Iz + a[4] * Ix * Iy + a[5] * Ix * Iz + a[6] * Iy * Iz +
                  a[7] * Ix * Iy * Iz;
        Fc[(long long)p * ncm + ccell] = Store<T, S>::enc(val, (T)VS::w(p));
    }
}

// Arguments of the linear interpolation of one level's parent list into the children.
template <typename T, typename S>
struct InterpArgs {
    const int *ids;
    int n;
    const S *Fp;
    S *Fc;
    long long ncm;
    const int *nbr_child, *blk_mask, *ref, *cells_mask;
    T tau_ratio;
};

template <typename T, int Q>
struct InterpLds {  // LDS of one interpolation workgroup
    static constexpr int CELLS = VelSet<Q>::CELLS, NCH = VelSet<Q>::CHILDREN;
    static constexpr int PPW = IWG / (NCH / 2 * CELLS);
    static constexpr int ALL = Q * PPW * CELLS, COEF = PPW * Q * NCH * InterpCoef<T, Q>::N;
};

// Workgroup `wg` of the interpolation (IWG threads): its parents, their children's cells.
template <typename T, typename S, int Q, bool ADDED>
__device__ __forceinline__ void interp_parents(const InterpArgs<T, S> &ia, int wg, T *s_all, T *s_coef)
{
    using VS = VelSet<Q>;
    constexpr int CELLS = VS::CELLS, NCH = VS::CHILDREN, HALF = NCH / 2;
    constexpr int GROUP = HALF * CELLS;  // lanes per parent: one per cell of half the children
    constexpr int PPW = IWG / GROUP;     // parents per workgroup
    constexpr int NCOEF = InterpCoef<T, Q>::N;
    const long long ncm = ia.ncm;
    const int g = threadIdx.x / GROUP, r = threadIdx.x % GROUP;
    const int c = r / CELLS, t = r % CELLS;
    const int bi = wg * PPW + g;
    int b = -1;
    if (bi < ia.n) b = ia.ids[bi];
    bool work = false;
    int fc = 0;
    if (b >= 0) {
        fc = ia.nbr_child[b];
        work = fc >= 0 && (ADDED ? (ia.ref[b] == REF_MARK_REFINE) : (ia.blk_mask[b] == BLK_INTERFACE));
    }
    // The children's cell masks, fetched before the barrier.
    bool put[2] = {false, false};
    if (work)
#pragma unroll
        for (int h = 0; h < 2; h++)
            put[h] = ADDED || ia.cells_mask[(long long)(fc + c + h * HALF) * CELLS + t] == CELL_GHOST;
    T *sg = s_all + g * CELLS;
    if (work && c == 0) {
        const long long cell = (long long)b * CELLS + t;
        T f[Q];
#pragma unroll
        for (int p = 0; p < Q; p++) f[p] = Store<T, S>::dec(ia.Fp[(long long)p * ncm + cell], (T)VS::w(p));
        T rho, u, v, w;
        moments<T, Q>(f, rho, u, v, w);
        const T udotu = u * u + v * v + w * w;
#pragma unroll
        for (int p = 0; p < Q; p++) {
            T cdotu = cdot<T, Q>(p, u, v, w);
            T fe = (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
            sg[p * PPW * CELLS + t] = fe + (f[p] - fe) * ia.tau_ratio;
        }
    }
    __syncthreads();
    interp_coefficients<T, Q, PPW, IWG>(s_all, s_coef);
    __syncthreads();
    if (!work) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (put[h])
            interp_child<T, S, Q>(s_coef + g * Q * NCH * NCOEF, c + h * HALF, t, ia.Fc, ncm,
                                  (long long)(fc + c + h * HALF) * CELLS + t);
}

template <typename T, typename S, int Q, bool ADDED>
__global__ __launch_bounds__(IWG) void k_interp(InterpArgs<T, S> ia)
{
    using IL = InterpLds<T, Q>;
    __shared__ T s_all[IL::ALL];
    __shared__ T s_coef[IL::COEF];
    interp_parents<T, S, Q, ADDED>(ia, blockIdx.x, s_all, s_coef);
}

// Sum over the NG lanes of each lane group (NG consecutive lanes, NG = 8 in 3D, 4 in 2D) of the
// values v[p], left to right in lane order as the reference sums the children (comm.cu:330-339),
// for the directions p = k (mod NG) of lane k of the group: the wavefront writes its Q x 64 values
// to its own LDS rows (stride 65: the group reads fall in distinct banks) and each lane reads back
// the NG values of its own directions.  `lds` is this wavefront's Q * 65-float region.
template <typename T, int Q, int NG>
__device__ __forceinline__ void group_sums(const T *v, T *acc, T *lds)
{
    const int lane = threadIdx.x & 63, k = lane % NG, gl = lane - k;
    wave_lds_sync();  // earlier reads of this region are done
#pragma unroll
    for (int p = 0; p < Q; p++) lds[p * 65 + lane] = v[p];
    wave_lds_sync();
#pragma unroll
    for (int m = 0; m < (Q + NG - 1) / NG; m++) {
        const int p = k + NG * m;
        if (p < Q) {
            const T *row = lds + p * 65 + gl;
            T a = row[0];
#pragma unroll
            for (int j = 1; j < NG; j++) a = a + row[j];
            acc[m] = a;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Fine -> coarse then coarse -> fine in one launch: the pair "average(L) in INTERFACE mode, then
// linear interpolation(L) in INTERFACE mode" that the advance runs between the two sub-cycles of
// level L+1 (solver_lbm_advance.cu:108-112).  Both act on the INTERFACE parents of level L, and each
// reads and writes only one parent block and its children: the average writes parent cells whose
// first child cell is an interface cell (mask 1), the interpolation reads the parent block and
// writes the children's ghost cells (mask 2).  A workgroup therefore does both for its parents:
// the averaged values go to the pool and, as the pool will hold them, to LDS, from where the
// interpolation stage takes them.  Arithmetic and order are those of k_average and k_interp.
// ---------------------------------------------------------------------------------------------
template <typename T, typename S, int Q>
__global__ __launch_bounds__(IWG) void k_avg_interp(const int *__restrict__ ids, int n, S *__restrict__ Fp,
                                                    S *__restrict__ Fc, long long ncm, const int *__restrict__ nbr_child,
                                                    const int *__restrict__ blk_mask, const int *__restrict__ cells_mask,
                                                    T ratio_avg, T ratio_int)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    constexpr int CELLS = VS::CELLS, DIM = VS::DIM, NCH = VS::CHILDREN, HALF = NCH / 2;
    constexpr int GROUP = HALF * CELLS;  // lanes per parent
    constexpr int PPW = IWG / GROUP;     // parents per workgroup
    constexpr int NCOEF = InterpCoef<T, Q>::N;
    // The averaging stage's group-sum rows and the interpolation stage's rescaled parent and
    // coefficients share one region (a barrier separates the stages).
    constexpr int RED = (IWG / 64) * Q * 65, INTERP = Q * PPW * CELLS + PPW * Q * NCH * NCOEF;
    __shared__ T s_stage[RED > INTERP ? RED : INTERP];
    __shared__ T s_avg[Q * PPW * CELLS];  // averaged parent populations as stored
    __shared__ int s_done[PPW * CELLS];   // parent cell was averaged
    T *const s_red = s_stage;                        // per-wavefront rows of the group sums
    T *const s_all = s_stage;                        // rescaled parent populations for the interpolation
    T *const s_coef = s_stage + Q * PPW * CELLS;     // interpolation coefficients
    const int g = threadIdx.x / GROUP, r = threadIdx.x % GROUP;
    const int bi = blockIdx.x * PPW + g;
    int b = -1;
    if (bi < n) b = ids[bi];
    int fc = -1;
    bool work = false;
    if (b >= 0) {
        fc = nbr_child[b];
        work = fc >= 0 && blk_mask[b] == BLK_INTERFACE;
    }
    // ---- average: NCH lanes per parent cell, two parent cells per lane group (k_average) ----
    {
        const int k = r % NCH, ta0 = r / NCH;  // parent cells ta0 and ta0 + CELLS / 2
        const int off = (k & 1) + 4 * ((k >> 1) & 1) + 16 * ((k >> 2) & 1);
        bool avg[2];
        long long cb[2];
        int c0[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int ta = ta0 + h * (CELLS / 2);
            const int I = ta % 4, J = (ta / 4) % 4, K = (DIM == 3) ? ta / 16 : 0;
            const int cq = (I / 2) + 2 * (J / 2) + 4 * (K / 2);
            c0[h] = 2 * (I % 2) + 4 * (2 * (J % 2));
            if (DIM == 3) c0[h] += 4 * 4 * (2 * (K % 2));
            cb[h] = (long long)(fc + cq) * CELLS;
            avg[h] = work && cells_mask[cb[h] + c0[h]] == 1;  // uniform over the NCH lanes of the group
        }
        T fl[2][Q];
#pragma unroll
        for (int h = 0; h < 2; h++)
            if (avg[h])
#pragma unroll
                for (int p = 0; p < Q; p++) fl[h][p] = ST::dec(Fc[(long long)p * ncm + cb[h] + c0[h] + off], (T)VS::w(p));
        const T scale = (DIM == 2) ? (T)(0.25) : (T)(0.125);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int ta = ta0 + h * (CELLS / 2);
            T sv[Q];
            if (avg[h]) {
                const T *f = fl[h];
                T rho, u, v, w;
                moments<T, Q>(f, rho, u, v, w);
                const T udotu = (DIM == 3) ? u * u + v * v + w * w : u * u + v * v;
#pragma unroll
                for (int p = 0; p < Q; p++) {
                    T cdotu = cdot<T, Q>(p, u, v, w);
                    T fe = (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
                    sv[p] = fe + (f[p] - fe) * ratio_avg;
                }
            } else {
#pragma unroll
                for (int p = 0; p < Q; p++) sv[p] = (T)0;
            }
            const long long pcell = (long long)b * CELLS + ta;
            T acc[(Q + NCH - 1) / NCH];
            group_sums<T, Q, NCH>(sv, acc, s_red + (threadIdx.x >> 6) * Q * 65);
#pragma unroll
            for (int m = 0; m < (Q + NCH - 1) / NCH; m++) {
                const int p = k + NCH * m;
                if (avg[h] && p < Q) {
                    const T wq = (T)VelSet<Q>::w(p);
                    const S enc = ST::enc(scale * acc[m], wq);
                    Fp[(long long)p * ncm + pcell] = enc;
                    s_avg[p * PPW * CELLS + g * CELLS + ta] = ST::dec(enc, wq);
                }
            }
            if (k == 0) s_done[g * CELLS + ta] = avg[h] ? 1 : 0;
        }
    }
    // The children's cell masks for the interpolation stage, fetched before the barrier.
    const int c = r / CELLS, t = r % CELLS;
    bool put[2] = {false, false};
    if (work)
#pragma unroll
        for (int h = 0; h < 2; h++) put[h] = cells_mask[(long long)(fc + c + h * HALF) * CELLS + t] == CELL_GHOST;
    __syncthreads();
    // ---- interpolation: parent rescaled into LDS, then two child cells per lane (k_interp) ----
    T *sg = s_all + g * CELLS;
    if (work && c == 0) {
        const long long cell = (long long)b * CELLS + t;
        const bool done = s_done[g * CELLS + t] != 0;
        T f[Q];
#pragma unroll
        for (int p = 0; p < Q; p++)
            f[p] = done ? s_avg[p * PPW * CELLS + g * CELLS + t] : ST::dec(Fp[(long long)p * ncm + cell], (T)VS::w(p));
        T rho, u, v, w;
        moments<T, Q>(f, rho, u, v, w);
        const T udotu = u * u + v * v + w * w;
#pragma unroll
        for (int p = 0; p < Q; p++) {
            T cdotu = cdot<T, Q>(p, u, v, w);
            T fe = (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
            sg[p * PPW * CELLS + t] = fe + (f[p] - fe) * ratio_int;
        }
    }
    __syncthreads();
    interp_coefficients<T, Q, PPW, IWG>(s_all, s_coef);
    __syncthreads();
    if (!work) return;
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (put[h])
            interp_child<T, S, Q>(s_coef + g * Q * NCH * NCOEF, c + h * HALF, t, Fc, ncm,
                                  (long long)(fc + c + h * HALF) * CELLS + t);
}

// ---------------------------------------------------------------------------------------------
// Coarse -> fine, cubic: Cu_Interpolate_Cubic (solver_lbm_kernels_comm.cu:406-638).  Per parent
// and direction the 4^d monomial coefficients alpha_i = sum_q C[i][q] s[q] of the tensor-product
// Lagrange interpolant through the parent's cells are formed once (the reference's coefficient
// tables, summed term by term in its order), spread over all lanes of the parent's group, and
// kept in LDS; each child cell then evaluates the polynomial by Horner's rule at its own centre.
// Same workgroup shape as the linear kernel: one lane per child cell.
// ---------------------------------------------------------------------------------------------
#define AGAL_CUBIC_TABLE __constant__ static const
#include "cubic_table.inc"
#undef AGAL_CUBIC_TABLE

constexpr int CWG = 512;  // cubic interpolation: one lane per child cell of a parent

template <typename T, typename S, int Q, bool ADDED>
__global__ __launch_bounds__(CWG) void k_interp_cubic(const int *__restrict__ ids, int n, const S *__restrict__ Fp,
                                                      S *__restrict__ Fc, long long ncm, const int *__restrict__ nbr_child,
                                                      const int *__restrict__ blk_mask, const int *__restrict__ ref,
                                                      const int *__restrict__ cells_mask, T tau_ratio)
{
    using VS = VelSet<Q>;
    constexpr int CELLS = VS::CELLS, DIM = VS::DIM, NCH = VS::CHILDREN;
    constexpr int GROUP = NCH * CELLS;
    constexpr int PPW = CWG / GROUP;
    __shared__ T s_all[Q * PPW * CELLS];  // rescaled parent populations, [p][parent][cell]
    __shared__ T a_all[Q * PPW * CELLS];  // monomial coefficients, [p][parent][i]
    const int g = threadIdx.x / GROUP, r = threadIdx.x % GROUP;
    const int c = r / CELLS, t = r % CELLS;
    const int bi = blockIdx.x * PPW + g;
    int b = -1;
    if (bi < n) b = ids[bi];
    bool work = false;
    int fc = 0;
    if (b >= 0) {
        fc = nbr_child[b];
        work = fc >= 0 && (ADDED ? (ref[b] == REF_MARK_REFINE) : (blk_mask[b] == BLK_INTERFACE));
    }
    if (work && c == 0) {
        const long long cell = (long long)b * CELLS + t;
        T f[Q];
#pragma unroll
        for (int p = 0; p < Q; p++) f[p] = Store<T, S>::dec(Fp[(long long)p * ncm + cell], (T)VS::w(p));
        T rho, u, v, w;
        moments<T, Q>(f, rho, u, v, w);
        const T udotu = u * u + v * v + w * w;
#pragma unroll
        for (int p = 0; p < Q; p++) {
            T cdotu = cdot<T, Q>(p, u, v, w);
            T fe = (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
            s_all[(p * PPW + g) * CELLS + t] = fe + (f[p] - fe) * tau_ratio;
        }
    }
    __syncthreads();
    if (work) {
        const int *start = (DIM == 3) ? CUBIC3_START : CUBIC2_START;
        const unsigned char *qs = (DIM == 3) ? CUBIC3_Q : CUBIC2_Q;
        const double *cs = (DIM == 3) ? CUBIC3_C : CUBIC2_C;
        for (int k = r; k < Q * CELLS; k += GROUP) {
            const int p = k / CELLS, i = k % CELLS;
            const T *s = s_all + (p * PPW + g) * CELLS;
            T acc = (T)cs[start[i]] * s[qs[start[i]]];
            for (int m = start[i] + 1; m < start[i + 1]; m++) acc = acc + (T)cs[m] * s[qs[m]];
            a_all[(p * PPW + g) * CELLS + i] = acc;
        }
    }
    __syncthreads();
    if (!work) return;
    const long long ccell = (long long)(fc + c) * CELLS + t;
    if (!ADDED && cells_mask[ccell] != CELL_GHOST) return;
    const T M_ONE_TWELFTH = (T)(-0.083333333333333), ONE_SIXTH = (T)(0.166666666666667);
    const T TWO_THIRDS = (T)(0.666666666666667), ZERO_THIRDS = (T)(0.0);
    const int I = t % 4, J = (t / 4) % 4, K = (DIM == 3) ? t / 16 : 0;
    const T x = M_ONE_TWELFTH + ONE_SIXTH * (T)I;
    const T y = M_ONE_TWELFTH + ONE_SIXTH * (T)J;
    const T z = (DIM == 3) ? M_ONE_TWELFTH + ONE_SIXTH * (T)K : (T)0;
    const T ox = (c & 1) ? TWO_THIRDS : ZERO_THIRDS, oy = (c & 2) ? TWO_THIRDS : ZERO_THIRDS;
    const T oz = (c & 4) ? TWO_THIRDS : ZERO_THIRDS;
#pragma unroll 1
    for (int p = 0; p < Q; p++) {
        const T *al = a_all + (p * PPW + g) * CELLS;
        T res = 0;
#pragma unroll
        for (int mk = 0; mk < (DIM == 3 ? 4 : 1); mk++) {
            T gam = 0;
#pragma unroll
            for (int mj = 0; mj < 4; mj++) {
                T beta = 0;
#pragma unroll
                for (int mi = 0; mi < 4; mi++)
                    beta = al[(3 - mi) + 4 * (3 - mj) + (DIM == 3 ? 16 * (3 - mk) : 0)] + (x + ox) * beta;
                gam = beta + (y + oy) * gam;
            }
            res = (DIM == 3) ? gam + (z + oz) * res : gam;
        }
        Fc[(long long)p * ncm + ccell] = Store<T, S>::enc(res, (T)VS::w(p));
    }
}

// ---------------------------------------------------------------------------------------------
// Fine -> coarse: Cu_Average (solver_lbm_kernels_comm.cu:174-347).  NG consecutive lanes (NG = 8
// in 3D, 4 in 2D) serve one parent cell, each lane loading and rescaling one child cell of the
// cell's 2x2(x2) group; the group sum is formed in the reference's order from the NG lanes by
// shuffles and each lane stores the directions p = lane (mod NG).  Each lane group serves two
// parent cells (t and t + CELLS/2, both loaded before either is reduced), which halves the
// wavefronts per parent.  No LDS, no barriers.
// ---------------------------------------------------------------------------------------------
template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_average(const int *__restrict__ ids, int n, S *__restrict__ Fp,
                                                const S *__restrict__ Fc, long long ncm, const int *__restrict__ nbr_child,
                                                const int *__restrict__ blk_mask, const int *__restrict__ cells_mask,
                                                int ave_type, T tau_ratio)
{
    using VS = VelSet<Q>;
    constexpr int CELLS = VS::CELLS, DIM = VS::DIM;
    constexpr int NG = (DIM == 3) ? 8 : 4;     // child cells per parent cell
    constexpr int PER_BLOCK = CELLS / 2 * NG;   // lanes per parent block: two parent cells per lane group
    __shared__ T s_red[(WG / 64) * Q * 65];     // per-wavefront rows of the group sums
    const int k = threadIdx.x % NG;             // which child cell of the group
    const long long gid = (long long)blockIdx.x * WG + threadIdx.x;
    const long long bi = gid / PER_BLOCK;
    const int t0 = (int)((gid % PER_BLOCK) / NG);  // parent cells t0 and t0 + CELLS / 2
    bool work = bi < n;
    int b = 0, fc = -1;
    if (work) {
        b = ids[bi];
        fc = nbr_child[b];
        work = (fc > -1) && (ave_type == 2 || blk_mask[b] == BLK_INTERFACE);
    }
    const int off = (k & 1) + 4 * ((k >> 1) & 1) + 16 * ((k >> 2) & 1);
    bool avg[2];
    long long ccell[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int t = t0 + h * (CELLS / 2);
        const int I = t % 4, J = (t / 4) % 4, K = (DIM == 3) ? t / 16 : 0;
        // Child block of this cell's quadrant and the first cell of its 2x2(x2) group.
        const int c = (I / 2) + 2 * (J / 2) + 4 * (K / 2);
        int c0 = 2 * (I % 2) + 4 * (2 * (J % 2));
        if (DIM == 3) c0 += 4 * 4 * (2 * (K % 2));
        const long long cb = (long long)(fc + c) * CELLS;
        avg[h] = false;
        if (work) {
            int m = cells_mask[cb + c0];
            if (ave_type > 0 && m < 2) m = 1;
            avg[h] = (m == 1);  // uniform over the NG lanes of the group
        }
        ccell[h] = cb + c0 + off;
    }
    T fl[2][Q];
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (avg[h])
#pragma unroll
            for (int p = 0; p < Q; p++) fl[h][p] = Store<T, S>::dec(Fc[(long long)p * ncm + ccell[h]], (T)VS::w(p));
    const T scale = (DIM == 2) ? (T)(0.25) : (T)(0.125);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        T s[Q];
        if (avg[h]) {
            const T *f = fl[h];
            T rho, u, v, w;
            moments<T, Q>(f, rho, u, v, w);
            const T udotu = (DIM == 3) ? u * u + v * v + w * w : u * u + v * v;
#pragma unroll
            for (int p = 0; p < Q; p++) {
                T cdotu = cdot<T, Q>(p, u, v, w);
                T fe = (T)VS::w(p) * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
                s[p] = fe + (f[p] - fe) * tau_ratio;
            }
        } else {
#pragma unroll
            for (int p = 0; p < Q; p++) s[p] = (T)0;
        }
        const long long pcell = (long long)b * CELLS + t0 + h * (CELLS / 2);
        // s of child cell offsets 0,1,4,5,16,17,20,21 summed left to right (comm.cu:330-339)
        T acc[(Q + NG - 1) / NG];
        group_sums<T, Q, NG>(s, acc, s_red + (threadIdx.x >> 6) * Q * 65);
#pragma unroll
        for (int m = 0; m < (Q + NG - 1) / NG; m++) {
            const int p = k + NG * m;
            if (avg[h] && p < Q) Fp[(long long)p * ncm + pcell] = Store<T, S>::enc(scale * acc[m], (T)VelSet<Q>::w(p));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// The fused step's launch.  The first interp_wgs workgroups interpolate one level's interface
// parents into their children (k_interp, INTERFACE mode), the next fix_wgs resolve the exception
// lists of levels that an earlier launch left (k_fixup), and from `lead` on (those two counts
// rounded up to a multiple of 8, so that the XCD mapping below still deals contiguous ranges)
// step_wgs workgroups run the fused step of the MultiStep levels.  The few, latency-bound
// interpolation workgroups go first so that they run beside the fused step instead of after it.
// The three parts touch disjoint cells: a level's interpolation reads the level's current buffer
// and writes the finer level's ghost cells, and the launcher hands over only fixups of levels that
// none of the other parts touches.
// ---------------------------------------------------------------------------------------------
template <typename T, typename S>
struct StepExtras {
    int step_wgs, interp_wgs, fix_wgs, lead;
    InterpArgs<T, S> ia;
    FixupArgs<T, S> fa;
};

static_assert(IWG == WG, "interpolation and fused-step workgroups share one launch");

template <typename T, typename S, int Q, bool MF>
__global__ __launch_bounds__(WG) void k_collide_stream(MultiStep<T, S> m, StepExtras<T, S> x)
{
    using VS = VelSet<Q>;
    using IL = InterpLds<T, Q>;
    // One LDS region shared by the two workgroup kinds of the launch: the interpolation's parent
    // rows and coefficients, or the matrix-core MRT's per-wavefront transpose rows.
    constexpr int ROW = Store<T, S>::quantized ? 20 : 36;
    constexpr int ALL16 = (IL::ALL * (int)sizeof(T) + 15) / 16 * 16 / (int)sizeof(T);  // coefficients 16-byte aligned
    constexpr size_t INTERP_BYTES = sizeof(T) * (ALL16 + IL::COEF);
    constexpr size_t MRT_BYTES = MF ? sizeof(float) * (WG / 64) * 64 * ROW : 0;
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[INTERP_BYTES > MRT_BYTES ? INTERP_BYTES : MRT_BYTES];
    if ((int)blockIdx.x < x.lead) {
        const int w2 = blockIdx.x;
        if (w2 < x.interp_wgs) {
            T *s_all = reinterpret_cast<T *>(s_raw);
            T *s_coef = s_all + ALL16;
            interp_parents<T, S, Q, false>(x.ia, w2, s_all, s_coef);
        } else if (w2 < x.interp_wgs + x.fix_wgs) {
            fixup_level<T, S, Q>(x.fa, w2 - x.interp_wgs);
        }
        return;
    }
    const int wg = (int)blockIdx.x - x.lead;
    const int slot = __builtin_amdgcn_readfirstlane(xcd_local_wg(wg, x.step_wgs) * (WG / 64) + (int)(threadIdx.x >> 6));
    if (slot >= m.start[m.nlv]) return;  // whole wavefronts only; the fused step has no barrier
    int k = 0;
#pragma unroll
    for (int j = 1; j < MAXLV; j++)
        if (j < m.nlv && slot >= m.start[j]) k = j;
    k = __builtin_amdgcn_readfirstlane(k);
    const StepArgs<T, S> &a = m.lv[k];
    const int wave = slot - m.start[k];
    if constexpr (MF) {
        // One transpose region per wavefront for the matrix-core MRT: 64 rows of 36 floats, or of
        // 20 floats (80 bytes: the forward rows, and the back rows in binary16) with 16-bit pools.
        float *s_mrt = reinterpret_cast<float *>(s_raw);
        collide_stream_3d<T, S, Q, true>(a, wave, s_mrt + (threadIdx.x >> 6) * 64 * ROW);
    } else if constexpr (VS::DIM == 3) {
        collide_stream_3d<T, S, Q, false>(a, wave, nullptr);
    } else {
        collide_stream_2d<T, S, Q>(a, wave);
    }
}

// ---------------------------------------------------------------------------------------------
// Linked blocks (3D, face-cell links of a geometry): the sub-step of blocks holding boundary cells
// with cut-link fractions.  Cu_Collide, then Cu_ImposeBC with the interpolated bounce-back of
// Bouzidi, Firdaouss and Lallemand (2001) (solver_lbm_kernels.cu:796-875), then the push of the
// fused step (push_link), one wavefront per block; the fused step skips these blocks (LINK_FLAG).
// Both read the level's input buffer and write disjoint slots of its output buffer, and record
// guard exceptions into the same list, so they can run in either order.
//
// For a link of fraction q along c_p from boundary cell x:
//   q < 1/2:  f_p <- 2q f_p + (1 - 2q) f*_p(x - c_p);   q >= 1/2:  f_p <- f_p / 2q + (2q - 1)/2q f_pb,
// per half-set pair in the order hook(p), link(p), hook(pb), link(pb), so the pb link's second
// form sees the new f_p.  f*_p(x - c_p) is the post-collision value behind the cell: a lane of the
// same wavefront, or, across a block face, the collision of that cell recomputed here from the
// same inputs with the same code (so the same bits).  A missing block behind the cell leaves f_p
// unchanged (the reference reads out of bounds there).
//
// With `mea` the block's momentum-exchange sums before the boundary rules (Cu_ComputeForcesMEA
// post_step 0, solver_lbm_kernels_forces.cu:322-540) are formed here from the post-collision
// values and stored (assigned) per block; k_mea_post adds the after-stream part.
// ---------------------------------------------------------------------------------------------
template <typename T>
struct LinkArgs {
    const int *ids;       // the level's linked blocks
    int n;
    const int *onb_solid;  // block -> rank in the link arrays
    const float *X_b;      // fractions [rank*64 + cell + q*ncb], 27 directions
    long long ncb;
    int bc_type;          // S_BC_TYPE: 1 = interpolated bounce-back
    int mea;              // form the pre-boundary momentum-exchange sums (post_step 0)
    int order;            // S_FORCE_ORDER
    T *blkF;              // 3 x nmax block sums
    long long nmax;
    T dv;                 // cell volume dx^3 of the level
};

// Post-collision value f*_q of cell `tc` of block nbk (which exists), from the launch's input
// buffer: collide_cell on the same loads and flags as that cell's own wavefront.
template <typename T, typename S, int Q>
__device__ __forceinline__ T post_collision_at(const StepArgs<T, S> &a, int nbk, int tc, int q)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const long long cell = (long long)nbk * 64 + tc;
    T f[Q], fs[Q];
#pragma unroll
    for (int p = 0; p < Q; p++) f[p] = ST::dec(a.A[(long long)p * a.ncm + cell], (T)VS::w(p));
    const bool inert = a.part[nbk] == 0 || a.cells_mask[cell] == CELL_SOLID;
    T rho, u, v, w;
    moments<T, Q>(f, rho, u, v, w);
    collide_cell<T, S, Q>(a, f, fs, rho, u, v, w, inert);
    T r = (T)0;
#pragma unroll
    for (int p = 0; p < Q; p++)
        if (p == q) r = fs[p];
    return r;
}

// Reduction of one value per lane over the wavefront in the reference's shared-memory tree order
// (s = 32, 16, ..., 1: v[t] = v[t] + v[t + s]); lane 0 holds the result.
template <typename T>
__device__ __forceinline__ T tree_sum64(T v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    return v;
}

// One momentum-exchange contribution (side s of a half-set pair) added to acc:
// order 1: f c_s; order 2: ((1/2 + q) f + (1/2 - q) f_m) c_s, the literal 1/2 a double, so the term
// and the running sum are formed in double and rounded by the reference's +=.
template <typename T>
__device__ __forceinline__ void mea_add(T *acc, int order, T f, T fm, float dist, int cx, int cy, int cz)
{
    const int c[3] = {cx, cy, cz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (order == 1) {
            acc[d] += f * (T)c[d];
        } else {
            const double qd = (double)(T)dist;
            acc[d] = (T)((double)acc[d] + ((0.5 + qd) * (double)f + (0.5 - qd) * (double)fm) * (double)c[d]);
        }
    }
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_link_step(StepArgs<T, S> a, LinkArgs<T> la)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (WG / 64) + (int)(threadIdx.x >> 6);
    if (bi >= la.n) return;  // whole wavefronts
    const int b = __builtin_amdgcn_readfirstlane(la.ids[bi]);
    const unsigned cell = (unsigned)b * 64u + (unsigned)t;
    const int my_nb = a.nbr_slot[(long long)b * 27 + (t < 27 ? t : 26)];
    const int part_u = __builtin_amdgcn_readfirstlane(a.part[b]);
    const int onb_u = __builtin_amdgcn_readfirstlane(a.onb[b]);
    const int rank = __builtin_amdgcn_readfirstlane(la.onb_solid[b]);
    T f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = ST::dec(a.A[(long long)q * a.ncm + cell], (T)VS::w(q));
    const int mask = a.cells_mask[cell];
    const bool inert = part_u == 0 || mask == CELL_SOLID;
    T fs[Q];
    T rho, u, v, w;
    moments<T, Q>(f, rho, u, v, w);
    collide_cell<T, S, Q>(a, f, fs, rho, u, v, w, inert);

    int nbs[27];
#pragma unroll
    for (int s = 0; s < 27; s++) nbs[s] = __builtin_amdgcn_readlane(my_nb, s);
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    // Block of the cell x + c_q (slot table) and of x - c_q.
    auto slot_of = [&](int dx, int dy, int dz) {
        const int Ip = I + dx, Jp = J + dy, Kp = K + dz;
        return (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) + 9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1));
    };
    auto cell_of = [&](int dx, int dy, int dz) { return ((I + dx) & 3) + 4 * ((J + dy) & 3) + 16 * ((K + dz) & 3); };
    auto block_at = [&](int sl) {
        int r = nbs[0];
#pragma unroll
        for (int k = 1; k < 27; k++)
            if (sl == k) r = nbs[k];
        return r;
    };

    const bool linked = rank >= 0 && mask == CELL_BOUNDARY;
    const long long cb = (long long)rank * 64 + t;
    float xq[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) xq[q] = (linked && q > 0) ? la.X_b[cb + (long long)q * la.ncb] : -1.0f;
    const bool curved = la.bc_type == 1 && linked;
    const bool mea2 = la.mea && la.order != 1 && linked;

    // f*_q(x - c_q): within the block from lane (I,J,K) - c_q (every lane takes part), else recomputed.
    T fm[Q];
    unsigned long long need = 0;  // bit q: needed and across a block face
    fm[0] = (T)0;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int cx = VS::cx(q), cy = VS::cy(q), cz = VS::cz(q);
        const int src = cell_of(-cx, -cy, -cz);
        if constexpr (sizeof(T) == 8) {
            const long long bits = __builtin_bit_cast(long long, fs[q]);
            const int lo = __shfl((int)(bits & 0xffffffffLL), src, 64), hi = __shfl((int)(bits >> 32), src, 64);
            fm[q] = __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned)lo);
        } else {
            fm[q] = __shfl(fs[q], src, 64);
        }
        const bool inside = (unsigned)(I - cx) < 4u && (unsigned)(J - cy) < 4u && (unsigned)(K - cz) < 4u;
        const bool want = (curved && xq[q] > 0 && xq[q] < 0.5f) || (mea2 && xq[q] > 0);
        if (want && !inside) need |= 1ull << q;
    }
    unsigned long long missing = 0;  // bit q: the block behind is missing
    while (need) {
        const int q = __builtin_ctzll(need);
        need &= need - 1;
        int cx = 0, cy = 0, cz = 0;
#pragma unroll
        for (int p = 1; p < Q; p++)
            if (p == q) {
                cx = VS::cx(p);
                cy = VS::cy(p);
                cz = VS::cz(p);
            }
        const int nbk = block_at(slot_of(-cx, -cy, -cz));
        T val = (T)0;
        if (nbk >= 0)
            val = post_collision_at<T, S, Q>(a, nbk, cell_of(-cx, -cy, -cz), q);
        else
            missing |= 1ull << q;
#pragma unroll
        for (int p = 1; p < Q; p++)
            if (p == q) fm[p] = val;
    }

    // Momentum exchange before the boundary rules (post_step 0).
    if (la.mea) {
        T acc[3] = {(T)0, (T)0, (T)0};
        if (linked) {
#pragma unroll
            for (int p = 1; p < Q; p++) {
                if (!VS::half(p)) continue;
                const int pb = VS::pb(p);
                if (xq[p] > 0) mea_add<T>(acc, la.order, fs[p], fm[p], xq[p], VS::cx(p), VS::cy(p), VS::cz(p));
                if (xq[pb] > 0) mea_add<T>(acc, la.order, fs[pb], fm[pb], xq[pb], VS::cx(pb), VS::cy(pb), VS::cz(pb));
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T s = tree_sum64<T>(acc[d]);
            if (t == 0) la.blkF[(long long)d * la.nmax + b] = s * la.dv;
        }
    }

    // Targets of the pushes (blocks in direction +c_q).
    int nbq[Q];
#pragma unroll
    for (int q = 1; q < Q; q++) nbq[q] = block_at(slot_of(VS::cx(q), VS::cy(q), VS::cz(q)));
    nbq[0] = b;

    // Cu_ImposeBC: face hooks and the interpolated bounce-back, pair by pair.
    const bool onb = onb_u == 1 && mask != CELL_SOLID;
    if (onb) {
#pragma unroll
        for (int p = 1; p < Q; p++) {
            if (!VS::half(p)) continue;
            const int pb = VS::pb(p);
            T f_p = fs[p], f_q = fs[pb];
            if (nbq[p] < 0 && nbq[p] != N_SKIPID)
                f_p = bc_hook<T>(a.bc_preset, a.U, nbq[p], f_p, u, v, w, (T)VS::w(p), (T)VS::cx(p), (T)VS::cy(p), (T)VS::cz(p));
            if (curved) {
                const T dQ = (T)xq[p];
                if (xq[p] > 0 && xq[p] < 0.5f) {
                    if (!((missing >> p) & 1)) f_p = (T)(2.0) * dQ * f_p + ((T)(1.0) - (T)(2.0) * dQ) * fm[p];
                } else if (xq[p] >= 0.5f) {
                    f_p = ((T)(1.0) / ((T)(2.0) * dQ)) * f_p + (((T)(2.0) * dQ - (T)(1.0)) / ((T)(2.0) * dQ)) * f_q;
                }
            }
            if (nbq[pb] < 0 && nbq[pb] != N_SKIPID)
                f_q = bc_hook<T>(a.bc_preset, a.U, nbq[pb], f_q, u, v, w, (T)VS::w(pb), (T)VS::cx(pb), (T)VS::cy(pb),
                                 (T)VS::cz(pb));
            if (curved) {
                const T dQ = (T)xq[pb];
                if (xq[pb] > 0 && xq[pb] < 0.5f) {
                    if (!((missing >> pb) & 1)) f_q = (T)(2.0) * dQ * f_q + ((T)(1.0) - (T)(2.0) * dQ) * fm[pb];
                } else if (xq[pb] >= 0.5f) {
                    f_q = ((T)(1.0) / ((T)(2.0) * dQ)) * f_q + (((T)(2.0) * dQ - (T)(1.0)) / ((T)(2.0) * dQ)) * f_p;
                }
            }
            fs[p] = f_p;
            fs[pb] = f_q;
        }
    }

    // The push (hooks already applied).
    store_pop(a.B, cell, ST::enc(fs[0], (T)VS::w(0)));
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int nbk = nbq[q];
        const long long tgt = (long long)nbk * 64 + cell_of(VS::cx(q), VS::cy(q), VS::cz(q));
        push_link<T, S, Q>(a, q, fs[q], nbk, tgt, (long long)cell, inert, false, false, u, v, w);
    }
}

// Cu_ComputeForcesMEA post_step 1 on the finest level after its last sub-step of the coarse step
// (and its exception fixup): the same sums from the values that arrived -- the reference's slot p
// after the stream is this build's row pb(p) -- added to the block's entry.
template <typename T, typename S, int Q>
__global__ __launch_bounds__(WG) void k_mea_post(const int *__restrict__ ids, int n, const S *__restrict__ F,
                                                 long long ncm, const int *__restrict__ nbr_slot,
                                                 const int *__restrict__ cells_mask, LinkArgs<T> la)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (WG / 64) + (int)(threadIdx.x >> 6);
    if (bi >= n) return;
    const int b = __builtin_amdgcn_readfirstlane(ids[bi]);
    const int rank = __builtin_amdgcn_readfirstlane(la.onb_solid[b]);
    const long long cell = (long long)b * 64 + t;
    const int my_nb = nbr_slot[(long long)b * 27 + (t < 27 ? t : 26)];
    const bool linked = rank >= 0 && cells_mask[cell] == CELL_BOUNDARY;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const long long cb = (long long)rank * 64 + t;
    T acc[3] = {(T)0, (T)0, (T)0};
    // Slot s of cell (I,J,K) - c_s: row pb(s) there.
    auto behind = [&](int s_dir) -> T {
        const int cx = VS::cx(s_dir), cy = VS::cy(s_dir), cz = VS::cz(s_dir);
        const int Ip = I - cx, Jp = J - cy, Kp = K - cz;
        const int slot = (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) +
                         9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1));
        const int nbk = __shfl(my_nb, slot, 64);
        if (nbk < 0) return (T)0;
        const long long o = (long long)nbk * 64 + ((Ip & 3) + 4 * (Jp & 3) + 16 * (Kp & 3));
        return ST::dec(F[(long long)VS::pb(s_dir) * ncm + o], (T)VS::w(s_dir));
    };
#pragma unroll
    for (int p = 1; p < Q; p++) {
        if (!VS::half(p)) continue;
        const int pb = VS::pb(p);
        // every lane evaluates its shuffles; only linked lanes add
        const T fmp = behind(p), fmq = behind(pb);
        if (!linked) continue;
        const float xp = la.X_b[cb + (long long)p * la.ncb], xb = la.X_b[cb + (long long)pb * la.ncb];
        const T f_p = ST::dec(F[(long long)pb * ncm + cell], (T)VS::w(p));
        const T f_q = ST::dec(F[(long long)p * ncm + cell], (T)VS::w(pb));
        if (xp > 0) mea_add<T>(acc, la.order, f_p, fmp, xp, VS::cx(p), VS::cy(p), VS::cz(p));
        if (xb > 0) mea_add<T>(acc, la.order, f_q, fmq, xb, VS::cx(pb), VS::cy(pb), VS::cz(pb));
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T s = tree_sum64<T>(acc[d]);
        if (t == 0) la.blkF[(long long)d * la.nmax + b] += s * la.dv;
    }
}

// ---------------------------------------------------------------------------------------------
// Cu_ComputeForcesCV (solver_lbm_kernels_forces.cu:13-246): the control-volume force sums of one
// level, one wavefront per block, lane t = cell t (lanes past the block's cells hold zeros).  A
// cell takes part when its centre lies strictly inside the box and it is not solid.
//   POST 0 (start of a coarse step): Fp = the momentum of taking-part leaf cells; on the root level
//     Fm = sum over p >= 1 of f*_p c_p for the directions whose neighbour centre x + c_p dx lies
//     outside the box, f*_p the BGK collision with the root rate, formed here only (the pool is
//     not written).
//   POST 1 (end of the coarse step): Fm = the momentum of taking-part leaf cells, plus on the root
//     level f_p c_pb(p) for the directions whose upwind centre x + c_pb(p) dx lies outside.
// Per block, the lane values go through the reference's shared-memory tree (s = cells/2 ... 1),
// times dx^DIM; POST 0 assigns the entries [2d] (Fp) and [2d+1] (Fm), POST 1 adds to them.
// ---------------------------------------------------------------------------------------------
template <typename T>
struct CvArgs {
    T dx, dv;        // spacing and cell volume of the level
    T om, omp;       // root-level BGK rate dx/tau and 1 - rate
    T box[6];        // xm, xM, ym, yM, zm, zM
    int root;        // the level is N_LEVEL_START
    T *blkF;         // 6 x nmax block sums
    long long nmax;
};

template <typename T, int DIM>
__device__ __forceinline__ bool cv_inside(const CvArgs<T> &cv, T x, T y, T z)
{
    bool in = x > cv.box[0] && x < cv.box[1] && y > cv.box[2] && y < cv.box[3];
    if (DIM == 3) in = in && z > cv.box[4] && z < cv.box[5];
    return in;
}

template <typename T, int CELLS>
__device__ __forceinline__ T tree_sum_cells(T v)
{
#pragma unroll
    for (int s = CELLS / 2; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
    return v;
}

template <typename T, typename S, int Q, int POST>
__global__ __launch_bounds__(WG) void k_cv_force(const int *__restrict__ ids, int n, const S *__restrict__ F,
                                                 long long ncm, const float *__restrict__ X,
                                                 const int *__restrict__ cells_mask,
                                                 const int *__restrict__ nbr_child, CvArgs<T> cv)
{
    using VS = VelSet<Q>;
    using ST = Store<T, S>;
    constexpr int CELLS = VS::CELLS, DIM = VS::DIM;
    const int lane = threadIdx.x & 63;
    const int bi = blockIdx.x * (WG / 64) + (int)(threadIdx.x >> 6);
    if (bi >= n) return;
    const int b = __builtin_amdgcn_readfirstlane(ids[bi]);
    if (b < 0) return;
    const bool act = lane < CELLS;
    const int t = act ? lane : 0;
    const long long cell = (long long)b * CELLS + t;
    const bool leaf = nbr_child[b] < 0;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const T x = (T)X[b] + cv.dx * ((T)0.5 + (T)I);
    const T y = (T)X[b + cv.nmax] + cv.dx * ((T)0.5 + (T)J);
    const T z = DIM == 3 ? (T)X[b + 2 * cv.nmax] + cv.dx * ((T)0.5 + (T)K) : (T)0;
    const bool pV = act && cv_inside<T, DIM>(cv, x, y, z) && cells_mask[cell] != CELL_SOLID;
    T f[Q];
    T rho = (T)0, m[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int p = 0; p < Q; p++) {
        f[p] = ST::dec(F[(long long)p * ncm + cell], (T)VS::w(p));
        rho = rho + f[p];
        m[0] = m[0] + f[p] * (T)VS::cx(p);
        m[1] = m[1] + f[p] * (T)VS::cy(p);
        if (DIM == 3) m[2] = m[2] + f[p] * (T)VS::cz(p);
    }
    T Fp[3] = {(T)0, (T)0, (T)0}, Fm[3] = {(T)0, (T)0, (T)0};
    if (pV && leaf) {
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            if (POST == 0) Fp[d] = Fp[d] + m[d];
            else Fm[d] = Fm[d] + m[d];
        }
    }
    if (cv.root && pV) {
        const T u = m[0] / rho, v = m[1] / rho, w = m[2] / rho;
#pragma unroll
        for (int p = 1; p < Q; p++) {
            const int s = POST == 0 ? p : VS::pb(p);  // the direction whose far centre is tested
            const int c[3] = {VS::cx(s), VS::cy(s), VS::cz(s)};
            if (cv_inside<T, DIM>(cv, x + (T)c[0] * cv.dx, y + (T)c[1] * cv.dx, z + (T)c[2] * cv.dx)) continue;
            T fv = f[p];
            if (POST == 0) {
                const T udotu = u * u + v * v + w * w;
                const T cdotu = (T)c[0] * u + (T)c[1] * v + (T)c[2] * w;
                fv = fv * cv.omp +
                     ((T)VS::w(p) * rho * ((T)1 + (T)3 * cdotu + (T)4.5 * cdotu * cdotu - (T)1.5 * udotu)) * cv.om;
            }
#pragma unroll
            for (int d = 0; d < DIM; d++) Fm[d] = Fm[d] + fv * (T)c[d];
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const T sp = tree_sum_cells<T, CELLS>(Fp[d]), sm = tree_sum_cells<T, CELLS>(Fm[d]);
        if (lane == 0) {
            T &ep = cv.blkF[(long long)(2 * d) * cv.nmax + b];
            T &em = cv.blkF[(long long)(2 * d + 1) * cv.nmax + b];
            if (POST == 0) {
                ep = sp * cv.dv;
                em = sm * cv.dv;
            } else {
                ep += sp * cv.dv;
                em += sm * cv.dv;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------------------
template <typename T>
static T *Fbuf(Ctx *c, int k) { return static_cast<T *>(c->F[k]); }

static inline int nwg(long long nblocks, int cells) { return (int)((nblocks * cells + WG - 1) / WG); }

template <int Q>
static inline int interp_grid(long long nparents)
{
    constexpr int ppw = IWG / (VelSet<Q>::CHILDREN / 2 * VelSet<Q>::CELLS);
    return (int)((nparents + ppw - 1) / ppw);
}

template <int Q>
static inline int cubic_grid(long long nparents)
{
    constexpr int ppw = CWG / (VelSet<Q>::CHILDREN * VelSet<Q>::CELLS);
    return (int)((nparents + ppw - 1) / ppw);
}

template <typename T, typename S, int Q>
static int set_ic_t(Ctx *c, int L)
{
    int n = c->n_ids[L];
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL((k_set_ic<T, S, Q>), dim3(nwg(n, VelSet<Q>::CELLS)), dim3(WG), 0, c->stream, c->id_set_L(L), n,
                       Fbuf<S>(c, c->cur[L]), c->n_maxcells, c->cells_mask, (T)c->P.ic_rho, (T)c->P.ic_u,
                       (T)c->P.ic_v, (T)c->P.ic_w);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

template <typename T, typename S, int Q>
static int refresh_t(Ctx *c, int L)
{
    int n = c->n_ids[L];
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL((k_refresh<T, S, Q>), dim3(nwg(n, VelSet<Q>::CELLS)), dim3(WG), 0, c->stream, c->id_set_L(L),
                       n, Fbuf<S>(c, c->cur[L]), static_cast<T *>(c->M), c->n_maxcells);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// Fused-step arguments of level L reading population buffer rd (writing rd ^ 1).
template <typename T, typename S, int Q>
static StepArgs<T, S> step_args(Ctx *c, int L, int rd, int vmode = 0)
{
    StepArgs<T, S> a;
    a.ids = c->order_L(L);  // the level's blocks in traversal order
    a.n = c->n_ids[L];
    a.A = Fbuf<S>(c, rd);
    a.B = Fbuf<S>(c, rd ^ 1);
    a.ncm = c->n_maxcells;
    a.nbr_slot = c->nbr_slot;
    a.part = c->part;
    a.simple = c->simple;
    a.cells_mask = c->cells_mask;
    a.onb = c->onb;
    const T dx_L = (T)c->dxf_vec[L], tau_L = (T)c->tau_vec[L];
    a.omeg = dx_L / tau_L;
    a.omegp = (T)(1.0) - a.omeg;
    a.mrt = c->P.collision == AGAL_MRT;
    {
        const double s[6] = {c->P.s1, c->P.s2, c->P.s3, c->P.s4, c->P.s5, c->P.s6};
        for (int k = 0; k < 27; k++) a.ck[k] = (T)0;
        for (int k = 0; k < Q; k++) {
            const int rc = Mrt<Q>::rate(k);
            if (rc == 0) continue;
            const T rate = rc == 1 ? a.omeg : (T)s[rc - 2];
            a.ck[k] = (T)((double)rate / (double)mrt_norm<Q>(k));
        }
    }
    a.U = (T)c->P.bc_velocity;
    a.bc_preset = c->P.bc_preset;
    // Each level has its own exception list: levels share launches.
    a.exc_idx = c->exc_idx + (size_t)3 * c->exc_cap * L;
    a.exc_val = static_cast<T *>(c->exc_val) + (size_t)c->exc_cap * L;
    a.exc_count = c->exc_count + 2 * L;
    a.exc_cap = c->exc_cap;
    const int XL = c->MAX_LEVELS;  // the extra list's slot
    a.vmode = vmode;
    a.flags = c->simple;
    a.va_idx = c->exc_idx + (size_t)3 * c->exc_cap * XL;
    a.va_val = static_cast<T *>(c->exc_val) + (size_t)c->exc_cap * XL;
    a.va_count = c->exc_count + 2 * XL;
    a.x_head = c->exc_head;
    a.x_next = c->exc_next;
    if (vmode == 1) {
        a.exc_idx = c->exc_idx + (size_t)3 * c->exc_cap * XL;
        a.exc_val = static_cast<T *>(c->exc_val) + (size_t)c->exc_cap * XL;
        a.exc_count = c->exc_count + 2 * XL;
    }
    return a;
}

// The extra-list mode of level L in the fused launch being issued: only for the finest level of a
// 3D set inside a coarse step.
template <int Q>
static int pair_mode(const Ctx *c, int L)
{
    // Linked blocks recompute neighbours' collisions from the input buffer as it is stored, so
    // with links every sub-step resolves its exceptions by a fixup instead.
    return (VelSet<Q>::DIM == 3 && c->defer_coarse && L == c->MAX_LEVELS - 1 && !c->links_on) ? c->fine_pair : 0;
}

// The momentum-exchange force (force type 0) is evaluated on the finest level's last sub-step of a
// coarse step, every step on a single level (solver_lbm_advance.cu:124-131, :144-150).
static bool mea_active(const Ctx *c) { return c->links_on && c->P.S_FORCE_TYPE == 0 && c->DIM == 3; }
static bool mea_due(const Ctx *c, int L)
{
    if (!mea_active(c) || L != c->MAX_LEVELS - 1) return false;
    const bool multi = c->MAX_LEVELS > 1 && c->MAX_LEVELS != c->N_LEVEL_START + 1;
    return !multi || c->global_Liter == (1 << (c->MAX_LEVELS - 1)) - 1;
}

template <typename T>
static LinkArgs<T> link_args(Ctx *c, int L, bool mea)
{
    LinkArgs<T> la;
    la.ids = c->lk_ids + c->lk_off[L];
    la.n = c->lk_n[L];
    la.onb_solid = c->lk_onb_solid;
    la.X_b = c->lk_X;
    la.ncb = c->lk_ncb;
    la.bc_type = c->P.S_BC_TYPE;
    la.mea = mea ? 1 : 0;
    la.order = c->P.S_FORCE_ORDER;
    la.blkF = static_cast<T *>(c->lk_blkF);
    la.nmax = c->n_maxcblocks;
    const T dx = (T)c->dxf_vec[L];
    la.dv = (T)((T)((T)(T)1 * dx) * dx) * dx;  // dvf_vec (solver_lbm_init.cu:19-27)
    return la;
}

// The linked blocks' sub-step of level L reading buffer rd, on stream s.
template <typename T, typename S, int Q>
static int launch_link_step(Ctx *c, int L, int rd, hipStream_t s)
{
    if constexpr (VelSet<Q>::DIM == 3) {
        if (!c->links_on || c->lk_n[L] <= 0) return AGAL_OK;
        const LinkArgs<T> la = link_args<T>(c, L, mea_due(c, L));
        hipLaunchKernelGGL((k_link_step<T, S, Q>), dim3((unsigned)((la.n + WG / 64 - 1) / (WG / 64))), dim3(WG), 0, s,
                           step_args<T, S, Q>(c, L, rd, 0), la);
        AGAL_HIP(c, hipGetLastError());
    }
    return AGAL_OK;
}

template <typename T, typename S>
static InterpArgs<T, S> interp_args(Ctx *c, int L, const int *ids, int n, T ratio)
{
    InterpArgs<T, S> ia;
    ia.ids = ids;
    ia.n = n;
    ia.Fp = Fbuf<S>(c, c->cur[L]);
    ia.Fc = Fbuf<S>(c, c->cur[L + 1]);
    ia.ncm = c->n_maxcells;
    ia.nbr_child = c->nbr_child;
    ia.blk_mask = c->blk_mask;
    ia.ref = c->ref;
    ia.cells_mask = c->cells_mask;
    ia.tau_ratio = ratio;
    return ia;
}

// Exception lists waiting for their fixup: up to MAXLV of the pending levels in `mask` go into fa
// (with the buffer that the launch which left the list wrote) and leave the pending set.
template <typename T, typename S>
static int take_fixups(Ctx *c, unsigned mask, FixupArgs<T, S> &fa, unsigned &wmask)
{
    int n = 0;
    std::vector<std::pair<int, int>> keep;
    for (const auto &pf : c->pending_fix) {
        const int L = pf.first, buf = pf.second;
        if (!((mask >> L) & 1) || n == MAXLV) {
            keep.push_back(pf);
            continue;
        }
        fa.B[n] = Fbuf<S>(c, buf);
        fa.exc_idx[n] = c->exc_idx + (size_t)3 * c->exc_cap * L;
        fa.exc_val[n] = static_cast<T *>(c->exc_val) + (size_t)c->exc_cap * L;
        fa.exc_count[n] = c->exc_count + 2 * L;
        fa.va_count[n] = L == c->MAX_LEVELS - 1 ? c->exc_count + 2 * c->MAX_LEVELS : nullptr;
        wmask |= res_bit(L, buf);
        n++;
    }
    c->pending_fix = keep;
    fa.ncm = c->n_maxcells;
    fa.exc_cap = c->exc_cap;
    return n;
}

// Fixup launches for the pending levels in `mask`.
template <typename T, typename S, int Q>
static int launch_fixups(Ctx *c, unsigned mask)
{
    for (;;) {
        FixupArgs<T, S> fa;
        unsigned wmask = 0;
        const int n = take_fixups<T, S>(c, mask, fa, wmask);
        if (n == 0) return AGAL_OK;
        hipStream_t s = sched_begin(c, 0, 0, wmask);
        hipLaunchKernelGGL((k_fixup<T, S, Q>), dim3(n), dim3(WG), 0, s, fa);
        AGAL_HIP(c, hipGetLastError());
        int st = sched_done(c, s, 0, 0, wmask);
        if (st) return st;
    }
}

// Fused step of levels[j] reading buffers rd[j], up to MAXLV levels per launch.  With `ia`, the
// first launch also interpolates level interp_L's interface parents (reading the buffer those
// sub-steps read).  Exception lists: inside a coarse step (defer_coarse) a launch leaves its
// levels' lists pending; pending lists of levels the launch touches are resolved by a fixup
// launch before it, the others by workgroups of the launch itself.  Outside a coarse step the
// fixup follows at once.  A coarse level alone goes to its level stream (the recorded step's
// scheduler); launches that hold the finest level stay on the context stream.
template <typename T, typename S, int Q>
static int launch_fused(Ctx *c, const int *levels, const int *rd, int nl, const InterpArgs<T, S> *ia = nullptr,
                        int interp_L = -1)
{
    constexpr int CELLS = VelSet<Q>::CELLS, BPW = 64 / CELLS;  // blocks per wavefront
    int st;
    unsigned touch = 0;
    for (int j = 0; j < nl; j++) touch |= 1u << levels[j];
    if (ia) touch |= (1u << interp_L) | (1u << (interp_L + 1));
    if ((st = launch_fixups<T, S, Q>(c, touch))) return st;
    for (int base = 0; base < nl || (base == 0 && ia); base += MAXLV) {
        MultiStep<T, S> m;
        StepExtras<T, S> x;
        m.nlv = 0;
        long long waves = 0;
        double cells = 0;
        unsigned rmask = 0, wmask = 0;
        bool finest = false;
        int first = -1;
        for (int j = base; j < nl && j < base + MAXLV; j++) {
            const int L = levels[j], n = c->n_ids[L];
            rmask |= res_bit(L, rd[j]);
            wmask |= res_bit(L, rd[j] ^ 1);
            if (n <= 0) continue;
            finest = finest || L == c->MAX_LEVELS - 1;
            if (first < 0) first = L;
            m.lv[m.nlv] = step_args<T, S, Q>(c, L, rd[j], pair_mode<Q>(c, L));
            m.start[m.nlv] = (int)waves;
            waves += (n + BPW - 1) / BPW;
            cells += (double)n * CELLS;
            m.nlv++;
        }
        for (int k = m.nlv; k <= MAXLV; k++) m.start[k] = (int)waves;
        x.step_wgs = m.nlv ? (int)((waves + WG / 64 - 1) / (WG / 64) + 7) / 8 * 8 : 0;  // a multiple of the XCD count
        x.interp_wgs = 0;
        x.fix_wgs = 0;
        if (base == 0) {
            if (ia && ia->n > 0) {
                x.ia = *ia;
                x.interp_wgs = interp_grid<Q>(ia->n);
                rmask |= res_bit(interp_L, rd[0]);
                wmask |= res_bit(interp_L + 1, c->cur[interp_L + 1]);
            }
            if (c->defer_coarse && !c->timing.enabled) x.fix_wgs = take_fixups<T, S>(c, ~touch, x.fa, wmask);
        }
        x.lead = (x.interp_wgs + x.fix_wgs + 7) / 8 * 8;
        if (x.step_wgs + x.interp_wgs + x.fix_wgs == 0) continue;
        const int slot = (finest || m.nlv != 1 || x.interp_wgs || x.fix_wgs) ? 0 : 1 + first;
        hipStream_t s = sched_begin(c, slot, rmask, wmask);
        const int grid = x.lead + x.step_wgs;
        // The matrix-core MRT variant: FP32 compute, 3D sets of 19 / 27 directions, when enabled.
        constexpr bool MF_OK = sizeof(T) == 4 && (Q == 19 || Q == 27);
        const bool mf = MF_OK && c->P.collision == AGAL_MRT && c->mrt_mfma;
        auto kern = mf ? (void (*)(MultiStep<T, S>, StepExtras<T, S>))k_collide_stream<T, S, Q, MF_OK>
                       : (void (*)(MultiStep<T, S>, StepExtras<T, S>))k_collide_stream<T, S, Q, false>;
        if (c->timing.enabled && m.nlv) {
            // Events recorded by the dispatch itself: they bracket the kernel's execution.
            hipEvent_t e0 = nullptr, e1 = nullptr;
            AGAL_HIP(c, hipEventCreate(&e0));
            AGAL_HIP(c, hipEventCreate(&e1));
            hipExtLaunchKernelGGL(kern, dim3(grid), dim3(WG), 0, s, e0, e1, 0, m, x);
            AGAL_HIP(c, hipGetLastError());
            c->timing.ev_pairs.push_back(e0);
            c->timing.ev_pairs.push_back(e1);
            c->timing.fused_launches++;
            c->timing.fused_cells += cells;
        } else {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), 0, s, m, x);
            AGAL_HIP(c, hipGetLastError());
        }
        for (int j = base; j < nl && j < base + MAXLV; j++)
            if ((st = launch_link_step<T, S, Q>(c, levels[j], rd[j], s))) return st;
        if ((st = sched_done(c, s, slot, rmask, wmask))) return st;
        // A list left for the next finest sub-step (pair mode 1) is not pending a fixup.
        for (int j = base; j < nl && j < base + MAXLV; j++)
            if (pair_mode<Q>(c, levels[j]) != 1) c->pending_fix.push_back({levels[j], rd[j] ^ 1});
        if (!c->defer_coarse && (st = launch_fixups<T, S, Q>(c, ~0u))) return st;
    }
    return AGAL_OK;
}

// The fused step of level L.  Inside a coarse step (defer_coarse), a level above the finest is
// held back and joins the launch of the next finest-level sub-step: between the two the advance
// runs only kernels of finer levels (interpolation into level L+1's children, finer sub-steps),
// which touch none of level L's cells, so the order of results is unchanged.
template <typename T, typename S, int Q>
static int collide_stream_t(Ctx *c, int L)
{
    if (c->defer_coarse && L < c->MAX_LEVELS - 1) {
        c->deferred.push_back({L, c->cur[L]});
        c->cur[L] ^= 1;
        return AGAL_OK;
    }
    std::vector<int> lv{L}, rd{c->cur[L]};
    c->cur[L] ^= 1;
    for (const auto &d : c->deferred) {
        lv.push_back(d.first);
        rd.push_back(d.second);
    }
    c->deferred.clear();
    return launch_fused<T, S, Q>(c, lv.data(), rd.data(), (int)lv.size());
}

// Interpolation of level L (INTERFACE mode) followed by level L's fused sub-step: one launch
// inside a coarse step with linear interpolation (the interpolation reads the buffer the sub-step
// reads and writes level L+1's ghost cells, which the sub-step does not touch), held-back coarser
// sub-steps included; two launches otherwise.
template <typename T, typename S, int Q>
static int interp_collide_t(Ctx *c, int L)
{
    int st;
    // The timed pass (timing.enabled) keeps the interpolation apart, so that the events around a
    // fused launch time the fused step alone.
    if (!c->defer_coarse || c->timing.enabled || c->P.interp_order == AGAL_INTERP_CUBIC || L >= c->MAX_LEVELS - 1) {
        if ((st = launch_fixups<T, S, Q>(c, (1u << L) | (2u << L)))) return st;  // levels L, L+1
        if ((st = launch_interpolate(c, L, AGAL_INTERP_INTERFACE))) return st;
        return launch_collide_stream(c, L);
    }
    const T ratio = (T)c->tau_vec[L + 1] / (T)c->tau_vec[L];
    const InterpArgs<T, S> ia = interp_args<T, S>(c, L, c->iface_L(L), c->n_iface[L], ratio);
    std::vector<int> lv{L}, rd{c->cur[L]};
    c->cur[L] ^= 1;
    for (const auto &d : c->deferred) {
        lv.push_back(d.first);
        rd.push_back(d.second);
    }
    c->deferred.clear();
    return launch_fused<T, S, Q>(c, lv.data(), rd.data(), (int)lv.size(), &ia, L);
}

// Launches coarse sub-steps still held back, then resolves every pending exception list.
template <typename T, typename S, int Q>
static int flush_deferred_t(Ctx *c)
{
    int st;
    if (!c->deferred.empty()) {
        std::vector<int> lv, rd;
        for (const auto &d : c->deferred) {
            lv.push_back(d.first);
            rd.push_back(d.second);
        }
        c->deferred.clear();
        if ((st = launch_fused<T, S, Q>(c, lv.data(), rd.data(), (int)lv.size()))) return st;
    }
    return launch_fixups<T, S, Q>(c, ~0u);
}

template <typename T, typename S, int Q>
static int fix_pending_t(Ctx *c)
{
    return launch_fixups<T, S, Q>(c, ~0u);
}

template <typename T, typename S, int Q>
static int interpolate_t(Ctx *c, int L, int var)
{
    if (L >= c->MAX_LEVELS - 1) return AGAL_OK;
    const T ratio = (T)c->tau_vec[L + 1] / (T)c->tau_vec[L];
    const int *ids;
    int n;
    if (var == AGAL_INTERP_INTERFACE) {
        ids = c->iface_L(L);
        n = c->n_iface[L];
    } else {
        ids = c->id_set_L(L);
        n = c->n_ids[L];
    }
    if (n <= 0) return AGAL_OK;
    const unsigned rl = res_bit(L, c->cur[L]), wl = res_bit(L + 1, c->cur[L + 1]);
    hipStream_t s = sched_begin(c, 0, rl, wl);
    if (c->P.interp_order == AGAL_INTERP_CUBIC) {
        if (var == AGAL_INTERP_ADDED)
            hipLaunchKernelGGL((k_interp_cubic<T, S, Q, true>), dim3(cubic_grid<Q>(n)), dim3(CWG), 0, s, ids, n,
                               Fbuf<S>(c, c->cur[L]), Fbuf<S>(c, c->cur[L + 1]), c->n_maxcells, c->nbr_child,
                               c->blk_mask, c->ref, c->cells_mask, ratio);
        else
            hipLaunchKernelGGL((k_interp_cubic<T, S, Q, false>), dim3(cubic_grid<Q>(n)), dim3(CWG), 0, s, ids, n,
                               Fbuf<S>(c, c->cur[L]), Fbuf<S>(c, c->cur[L + 1]), c->n_maxcells, c->nbr_child,
                               c->blk_mask, c->ref, c->cells_mask, ratio);
    } else if (var == AGAL_INTERP_ADDED)
        hipLaunchKernelGGL((k_interp<T, S, Q, true>), dim3(interp_grid<Q>(n)), dim3(IWG), 0, s,
                           interp_args<T, S>(c, L, ids, n, ratio));
    else
        hipLaunchKernelGGL((k_interp<T, S, Q, false>), dim3(interp_grid<Q>(n)), dim3(IWG), 0, s,
                           interp_args<T, S>(c, L, ids, n, ratio));
    AGAL_HIP(c, hipGetLastError());
    return sched_done(c, s, 0, rl, wl);
}

template <typename T, typename S, int Q>
static int average_t(Ctx *c, int L, int var)
{
    if (L >= c->MAX_LEVELS - 1) return AGAL_OK;
    const T ratio = (T)c->tau_vec[L] / (T)c->tau_vec[L + 1];
    const int *ids;
    int n;
    if (var == AGAL_AVERAGE_INTERFACE || var == AGAL_AVERAGE_BLOCK) {  // both act on INTERFACE blocks only
        ids = c->iface_L(L);
        n = c->n_iface[L];
    } else {
        ids = c->id_set_L(L);
        n = c->n_ids[L];
    }
    if (n <= 0) return AGAL_OK;
    const unsigned rl = res_bit(L + 1, c->cur[L + 1]), wl = res_bit(L, c->cur[L]);
    hipStream_t s = sched_begin(c, 0, rl, wl);
    hipLaunchKernelGGL((k_average<T, S, Q>), dim3(nwg(n, VelSet<Q>::CELLS / 2 * (VelSet<Q>::DIM == 3 ? 8 : 4))), dim3(WG),
                       0, s, ids, n, Fbuf<S>(c, c->cur[L]), Fbuf<S>(c, c->cur[L + 1]), c->n_maxcells, c->nbr_child,
                       c->blk_mask, c->cells_mask, var, ratio);
    AGAL_HIP(c, hipGetLastError());
    return sched_done(c, s, 0, rl, wl);
}

template <typename T, typename S, int Q>
static int avg_interp_t(Ctx *c, int L)
{
    if (L >= c->MAX_LEVELS - 1) return AGAL_OK;
    if (c->P.interp_order == AGAL_INTERP_CUBIC) {  // cubic: BLOCK averaging and the cubic kernel
        int st = average_t<T, S, Q>(c, L, AGAL_AVERAGE_BLOCK);
        return st ? st : interpolate_t<T, S, Q>(c, L, AGAL_INTERP_INTERFACE);
    }
    const int n = c->n_iface[L];
    if (n <= 0) return AGAL_OK;
    const T ratio_avg = (T)c->tau_vec[L] / (T)c->tau_vec[L + 1];
    const T ratio_int = (T)c->tau_vec[L + 1] / (T)c->tau_vec[L];
    const unsigned m = res_bit(L, c->cur[L]) | res_bit(L + 1, c->cur[L + 1]);
    hipStream_t s = sched_begin(c, 0, m, m);
    hipLaunchKernelGGL((k_avg_interp<T, S, Q>), dim3(interp_grid<Q>(n)), dim3(IWG), 0, s, c->iface_L(L), n,
                       Fbuf<S>(c, c->cur[L]), Fbuf<S>(c, c->cur[L + 1]), c->n_maxcells, c->nbr_child, c->blk_mask,
                       c->cells_mask, ratio_avg, ratio_int);
    AGAL_HIP(c, hipGetLastError());
    return sched_done(c, s, 0, m, m);
}

#define AGAL_DISPATCH_Q(c, FN, T, S, ...)                                                  \
    do {                                                                                   \
        if ((c)->Q == 9) return FN<T, S, 9>(__VA_ARGS__);                                  \
        if ((c)->Q == 19) return FN<T, S, 19>(__VA_ARGS__);                                \
        if ((c)->Q == 27) return FN<T, S, 27>(__VA_ARGS__);                                \
    } while (0)
#define AGAL_DISPATCH(c, FN, ...)                                                          \
    do {                                                                                   \
        switch ((c)->precision) {                                                          \
            case AGAL_FP64: AGAL_DISPATCH_Q(c, FN, double, double, __VA_ARGS__); break;    \
            case AGAL_FP32: AGAL_DISPATCH_Q(c, FN, float, float, __VA_ARGS__); break;      \
            case AGAL_FP16: AGAL_DISPATCH_Q(c, FN, float, _Float16, __VA_ARGS__); break;   \
            case AGAL_BF16: AGAL_DISPATCH_Q(c, FN, float, bf16s, __VA_ARGS__); break;      \
        }                                                                                  \
        return AGAL_ERR_ARG;                                                               \
    } while (0)

template <typename T, typename S, int Q>
static int cv_force_t(Ctx *c, int L, int post)
{
    const int n = c->n_ids[L];
    if (n <= 0 || !c->cv_F) return AGAL_OK;
    CvArgs<T> cv;
    const T dx = (T)c->dxf_vec[L];
    cv.dx = dx;
    T dv = (T)1;
    for (int d = 0; d < VelSet<Q>::DIM; d++) dv = dv * dx;  // dvf_vec (solver_lbm_init.cu:19-27)
    cv.dv = dv;
    const int L0 = c->N_LEVEL_START;
    cv.om = (T)c->dxf_vec[L0] / (T)c->tau_vec[L0];
    cv.omp = (T)1 - cv.om;
    for (int k = 0; k < 6; k++) cv.box[k] = (T)c->cv_box[k];
    cv.root = L == L0;
    cv.blkF = static_cast<T *>(c->cv_F);
    cv.nmax = c->n_maxcblocks;
    const unsigned rm = res_bit(L, c->cur[L]);
    hipStream_t s = sched_begin(c, 0, rm, 0);
    const dim3 grid((unsigned)((n + WG / 64 - 1) / (WG / 64)));
    if (post == 0)
        hipLaunchKernelGGL((k_cv_force<T, S, Q, 0>), grid, dim3(WG), 0, s, c->id_set_L(L), n, Fbuf<S>(c, c->cur[L]),
                           c->n_maxcells, c->X, c->cells_mask, c->nbr_child, cv);
    else
        hipLaunchKernelGGL((k_cv_force<T, S, Q, 1>), grid, dim3(WG), 0, s, c->id_set_L(L), n, Fbuf<S>(c, c->cur[L]),
                           c->n_maxcells, c->X, c->cells_mask, c->nbr_child, cv);
    AGAL_HIP(c, hipGetLastError());
    return sched_done(c, s, 0, rm, 0);
}

template <typename T, typename S, int Q>
static int mea_post_t(Ctx *c)
{
    if constexpr (VelSet<Q>::DIM == 3) {
        const int L = c->MAX_LEVELS - 1;
        if (!mea_active(c) || c->lk_n[L] <= 0) return AGAL_OK;
        const LinkArgs<T> la = link_args<T>(c, L, true);
        const unsigned rm = res_bit(L, c->cur[L]);
        hipStream_t s = sched_begin(c, 0, rm, 0);
        hipLaunchKernelGGL((k_mea_post<T, S, Q>), dim3((unsigned)((la.n + WG / 64 - 1) / (WG / 64))), dim3(WG), 0, s, la.ids,
                           la.n, Fbuf<S>(c, c->cur[L]), c->n_maxcells, c->nbr_slot, c->cells_mask, la);
        AGAL_HIP(c, hipGetLastError());
        return sched_done(c, s, 0, rm, 0);
    }
    return AGAL_OK;
}

int launch_mea_post(Ctx *c) { AGAL_DISPATCH(c, mea_post_t, c); }
int launch_cv_force(Ctx *c, int L, int post) { AGAL_DISPATCH(c, cv_force_t, c, L, post); }
int launch_set_ic(Ctx *c, int L) { AGAL_DISPATCH(c, set_ic_t, c, L); }
int launch_refresh(Ctx *c, int L) { AGAL_DISPATCH(c, refresh_t, c, L); }
int launch_collide_stream(Ctx *c, int L) { AGAL_DISPATCH(c, collide_stream_t, c, L); }
int launch_flush_deferred(Ctx *c) { AGAL_DISPATCH(c, flush_deferred_t, c); }
int launch_interp_collide(Ctx *c, int L) { AGAL_DISPATCH(c, interp_collide_t, c, L); }
int launch_fix_pending(Ctx *c) { AGAL_DISPATCH(c, fix_pending_t, c); }
int launch_interpolate(Ctx *c, int L, int var) { AGAL_DISPATCH(c, interpolate_t, c, L, var); }
int launch_average(Ctx *c, int L, int var) { AGAL_DISPATCH(c, average_t, c, L, var); }
int launch_average_interpolate(Ctx *c, int L) { AGAL_DISPATCH(c, avg_interp_t, c, L); }

}  // namespace agal
