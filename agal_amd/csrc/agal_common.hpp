// agal_common.hpp -- compile-time velocity sets and small device helpers for the gfx950 kernels.
//
// Direction order, weights and opposite directions are the reference's: V_CONN_ID / V_CONN_MAP
// (inc/util.h:94-101), LBMw / LBMpb (src/solver_lbm/solver_lbm_init.cu:79-88).  They are
// constexpr here so every unrolled direction loop sees its velocity as an immediate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace agal {

constexpr int N_SKIPID = -999;
constexpr int REF_UNREFINED = 0, REF_UNREFINED_VIO = 9, REF_REFINED = 1, REF_REFINED_WCHILD = 2,
              REF_REFINED_PERM = 3, REF_REFINED_WCHILD_PERM = 10, REF_MARK_REFINE = 4,
              REF_MARK_COARSEN = 5, REF_NEW = 6, REF_REMOVE = 7, REF_INACTIVE = 8, REF_INDET_E = 11,
              REF_INDET_O = 12;
constexpr int CELL_INTERIOR = 0, CELL_INTERFACE = 1, CELL_GHOST = 2, CELL_SOLID = -1, CELL_BOUNDARY = -6;
// Block flag bit (beside the fused step's simple / interior kind in bits 0-1): the block holds
// boundary cells with face-cell links and takes the linked sub-step instead of the fused one.
constexpr int LINK_FLAG = 8;
constexpr int BLK_REGULAR = 0, BLK_INTERFACE = 1, BLK_SOLID = -3, BLK_SOLIDA = -2, BLK_SOLIDB = -1;

// D3Q27 order; D3Q19 is its first 19 entries (inc/util.h:95).
__host__ __device__ constexpr int c3x(int p)
{
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return t[p];
}
__host__ __device__ constexpr int c3y(int p)
{
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    return t[p];
}
__host__ __device__ constexpr int c3z(int p)
{
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    return t[p];
}
// D2Q9 (inc/util.h:94).
__host__ __device__ constexpr int c2x(int p)
{
    constexpr int t[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    return t[p];
}
__host__ __device__ constexpr int c2y(int p)
{
    constexpr int t[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    return t[p];
}

template <int Q>
struct VelSet;

template <>
struct VelSet<9> {
    static constexpr int Q = 9, DIM = 2, QMAX = 9, CELLS = 16, CHILDREN = 4;
    __host__ __device__ static constexpr int cx(int p) { return c2x(p); }
    __host__ __device__ static constexpr int cy(int p) { return c2y(p); }
    __host__ __device__ static constexpr int cz(int) { return 0; }
    __host__ __device__ static constexpr double w(int p) { return p == 0 ? 4.0 / 9.0 : (p < 5 ? 1.0 / 9.0 : 1.0 / 36.0); }
    __host__ __device__ static constexpr int pb(int p)
    {
        constexpr int t[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
        return t[p];
    }
    // Directions that initiate the swap in the reference stream (solver_lbm_kernels.cu:597).
    __host__ __device__ static constexpr bool half(int p) { return p == 1 || (p + 1) % 3 == 0; }
};

template <>
struct VelSet<19> {
    static constexpr int Q = 19, DIM = 3, QMAX = 27, CELLS = 64, CHILDREN = 8;
    __host__ __device__ static constexpr int cx(int p) { return c3x(p); }
    __host__ __device__ static constexpr int cy(int p) { return c3y(p); }
    __host__ __device__ static constexpr int cz(int p) { return c3z(p); }
    __host__ __device__ static constexpr double w(int p) { return p == 0 ? 1.0 / 3.0 : (p < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }
    __host__ __device__ static constexpr int pb(int p) { return p == 0 ? 0 : (p % 2 == 1 ? p + 1 : p - 1); }
    __host__ __device__ static constexpr bool half(int p) { return p == 26 || ((p - 1) % 2 == 0 && p < 25); }
};

template <>
struct VelSet<27> {
    static constexpr int Q = 27, DIM = 3, QMAX = 27, CELLS = 64, CHILDREN = 8;
    __host__ __device__ static constexpr int cx(int p) { return c3x(p); }
    __host__ __device__ static constexpr int cy(int p) { return c3y(p); }
    __host__ __device__ static constexpr int cz(int p) { return c3z(p); }
    __host__ __device__ static constexpr double w(int p)
    {
        return p == 0 ? 8.0 / 27.0 : (p < 7 ? 2.0 / 27.0 : (p < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
    __host__ __device__ static constexpr int pb(int p) { return p == 0 ? 0 : (p % 2 == 1 ? p + 1 : p - 1); }
    __host__ __device__ static constexpr bool half(int p) { return p == 26 || ((p - 1) % 2 == 0 && p < 25); }
};

// Velocity index of a neighbour offset (dx,dy,dz) in {-1,0,1}^3, D3Q27 order (V_CONN_MAP).
__host__ __device__ constexpr int dir3(int dx, int dy, int dz)
{
    constexpr int m[27] = {20, 12, 26, 10, 6, 15, 24, 17, 21, 8, 4, 13, 2, 0, 1, 14, 3, 7, 22, 18, 23, 16, 5, 9, 25, 11, 19};
    return m[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
}
// D2Q9 order.
__host__ __device__ constexpr int dir2(int dx, int dy)
{
    constexpr int m[9] = {7, 4, 8, 3, 0, 1, 6, 2, 5};
    return m[(dx + 1) + 3 * (dy + 1)];
}

// ---------------------------------------------------------------------------------------------
// MRT moment bases: integer rows m(k, p) over the velocity set, mutually orthogonal, so that
// M^-1 = M^T D^-1 with D_k = sum_p m(k,p)^2.  rate(k): 0 conserved (not relaxed), 1 the viscous
// rate omega_L, 2..7 the parameters s1..s6.
//   D3Q19: d'Humieres et al. (2002), the rows of the reference's legacy transform
//          (src/solver_old/solver_lbm_vel2mom_mrt_d3q19.cu:118-136) and its rate assignment
//          (solver_lbm_collide_mrt_d3q19.cu:132-146).
//   D2Q9 / D3Q27: products of the 1D orthogonal polynomials {1, c, 3c^2 - 2} per axis, with the
//          second-order block {P2x, P2y, P2z} rotated into its trace (bulk) and deviators.
// ---------------------------------------------------------------------------------------------
__host__ __device__ constexpr int mrt_p2(int c) { return 3 * c * c - 2; }

template <int Q>
struct Mrt;

template <>
struct Mrt<19> {
    __host__ __device__ static constexpr int m(int k, int p)
    {
        const int x = c3x(p), y = c3y(p), z = c3z(p), c2 = x * x + y * y + z * z;
        switch (k) {
            case 0: return 1;
            case 1: return 19 * c2 - 30;
            case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
            case 3: return x;
            case 4: return (5 * c2 - 9) * x;
            case 5: return y;
            case 6: return (5 * c2 - 9) * y;
            case 7: return z;
            case 8: return (5 * c2 - 9) * z;
            case 9: return 3 * x * x - c2;
            case 10: return (3 * c2 - 5) * (3 * x * x - c2);
            case 11: return y * y - z * z;
            case 12: return (3 * c2 - 5) * (y * y - z * z);
            case 13: return x * y;
            case 14: return y * z;
            case 15: return x * z;
            case 16: return (y * y - z * z) * x;
            case 17: return (z * z - x * x) * y;
            case 18: return (x * x - y * y) * z;
        }
        return 0;
    }
    __host__ __device__ static constexpr int rate(int k)
    {
        constexpr int r[19] = {0, 2, 3, 0, 4, 0, 4, 0, 4, 1, 6, 1, 6, 1, 1, 1, 7, 7, 7};
        return r[k];
    }
};

template <>
struct Mrt<27> {
    __host__ __device__ static constexpr int m(int k, int p)
    {
        const int x = c3x(p), y = c3y(p), z = c3z(p);
        const int px = mrt_p2(x), py = mrt_p2(y), pz = mrt_p2(z);
        switch (k) {
            case 0: return 1;
            case 1: return x;
            case 2: return y;
            case 3: return z;
            case 4: return x * y;
            case 5: return y * z;
            case 6: return x * z;
            case 7: return 2 * px - py - pz;
            case 8: return py - pz;
            case 9: return px + py + pz;
            case 10: return px * y;
            case 11: return px * z;
            case 12: return py * x;
            case 13: return py * z;
            case 14: return pz * x;
            case 15: return pz * y;
            case 16: return x * y * z;
            case 17: return px * py;
            case 18: return py * pz;
            case 19: return px * pz;
            case 20: return px * y * z;
            case 21: return py * x * z;
            case 22: return pz * x * y;
            case 23: return px * py * z;
            case 24: return py * pz * x;
            case 25: return px * pz * y;
            case 26: return px * py * pz;
        }
        return 0;
    }
    __host__ __device__ static constexpr int rate(int k)
    {
        constexpr int r[27] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 5, 5, 5, 6, 6, 6, 7};
        return r[k];
    }
};

template <>
struct Mrt<9> {
    __host__ __device__ static constexpr int m(int k, int p)
    {
        const int x = c2x(p), y = c2y(p);
        const int px = mrt_p2(x), py = mrt_p2(y);
        switch (k) {
            case 0: return 1;
            case 1: return x;
            case 2: return y;
            case 3: return x * y;
            case 4: return px - py;
            case 5: return px + py;
            case 6: return px * y;
            case 7: return py * x;
            case 8: return px * py;
        }
        return 0;
    }
    __host__ __device__ static constexpr int rate(int k)
    {
        constexpr int r[9] = {0, 0, 0, 1, 1, 2, 4, 4, 3};
        return r[k];
    }
};

template <int Q>
__host__ __device__ constexpr int mrt_norm(int k)
{
    int s = 0;
    for (int p = 0; p < Q; p++) s += Mrt<Q>::m(k, p) * Mrt<Q>::m(k, p);
    return s;
}

// ---------------------------------------------------------------------------------------------
// MRT on matrix cores (3D).  Per wavefront the 64 cells of a block are the N dimension of
// v_mfma_f32_16x16x32_bf16 products.  Lane l = 16q + j of that instruction holds A[j][8q..8q+7]
// and B[8q..8q+7][j] and receives C[4q+r][j] in result register r.  Forward m = M d: moment rows
// in two halves h of 16, K = direction.  Back g = M^T (c . m): K runs over moments in the order the
// forward result already occupies the lanes, K position 8q+s <-> moment 4q+s (s < 4) or
// 16+4q+(s-4) (s >= 4), so no data moves between the two products.  The integer entries of M are
// exact in bf16.
// ---------------------------------------------------------------------------------------------
__host__ __device__ constexpr uint16_t bf16_of_int(int v)  // exact for |v| < 256
{
    if (v == 0) return 0;
    const uint32_t sign = v < 0 ? 0x8000u : 0u;
    const uint32_t a = (uint32_t)(v < 0 ? -v : v);
    int e = 0;
    while ((a >> e) > 1u) e++;
    const uint32_t mant = (a << (7 - e)) & 0x7Fu;
    return (uint16_t)(sign | ((uint32_t)(127 + e) << 7) | mant);
}

__host__ __device__ constexpr uint16_t f16_of_int(int v)  // exact for |v| <= 2048
{
    if (v == 0) return 0;
    const uint32_t sign = v < 0 ? 0x8000u : 0u;
    const uint32_t a = (uint32_t)(v < 0 ? -v : v);
    int e = 0;
    while ((a >> e) > 1u) e++;
    const uint32_t mant = (a << (10 - e)) & 0x3FFu;
    return (uint16_t)(sign | ((uint32_t)(15 + e) << 10) | mant);
}

__host__ __device__ constexpr int mfma_moment_of(int q, int s) { return s < 4 ? 4 * q + s : 16 + 4 * q + (s - 4); }

template <int Q>
struct alignas(16) MfmaTab {
    uint16_t fw[64][2][8];  // forward A fragments (bf16 bits) per lane and moment half
    uint16_t bk[64][2][8];  // back A fragments per lane and direction half
    uint16_t fw16[64][2][8];  // the same fragments as binary16 bits
    uint16_t bk16[64][2][8];
};

template <int Q>
constexpr MfmaTab<Q> make_mfma_tab()
{
    MfmaTab<Q> t{};
    int M[32][32] = {};
    for (int k = 0; k < Q; k++)
        for (int p = 0; p < Q; p++) {
            M[k][p] = Mrt<Q>::m(k, p);
        }
    uint16_t Mb[32][32] = {};
    for (int k = 0; k < 32; k++)
        for (int p = 0; p < 32; p++) Mb[k][p] = bf16_of_int(M[k][p]);
    for (int l = 0; l < 64; l++) {
        const int q = l / 16, j = l % 16;
        for (int h = 0; h < 2; h++) {
            for (int s = 0; s < 8; s++) {
                const int i = 16 * h + j;
                t.fw[l][h][s] = Mb[i][8 * q + s];
                t.bk[l][h][s] = Mb[mfma_moment_of(q, s)][i];
                t.fw16[l][h][s] = f16_of_int(M[i][8 * q + s]);
                t.bk16[l][h][s] = f16_of_int(M[mfma_moment_of(q, s)][i]);
            }
        }
    }
    return t;
}

// Boundary-condition presets of the reference hook Cu_ImposeBC (inc/custom.h:66-123).
enum BcPreset { BC_NONE = 0, BC_LDC2D = 1, BC_LDC3D = 2, BC_FPSC2D = 3, BC_FPSC3D = 4 };

template <typename T>
__device__ __forceinline__ T bc_hook(int preset, T U, int nbr_id, T f, T u, T v, T w, T wp, T cxp, T cyp, T czp)
{
    bool wall = false, outflow = false;
    if (preset == BC_LDC2D) wall = (nbr_id == -4);
    if (preset == BC_LDC3D) wall = (nbr_id == -6);
    if (preset == BC_FPSC2D) {
        wall = (nbr_id == -1 || nbr_id == -3 || nbr_id == -4);
        outflow = (nbr_id == -2);
    }
    if (preset == BC_FPSC3D) {
        wall = (nbr_id == -1 || nbr_id == -3 || nbr_id == -4 || nbr_id == -5 || nbr_id == -6);
        outflow = (nbr_id == -2);
    }
    if (wall) {
        T cdotu = (T)(3.0) * (cxp * U);
        return f - (T)(2.0) * wp * cdotu;
    }
    if (outflow) {
        T cdotu = cxp * u + cyp * v + czp * w;
        cdotu = (T)(1.0) + (T)(4.5) * cdotu * cdotu - (T)(1.5) * (u * u + v * v + w * w);
        return -f + (T)(2.0) * wp * cdotu;
    }
    return f;
}

// ---------------------------------------------------------------------------------------------
// Population storage.  FP32 / FP64 pools store f itself.  The 16-bit pools store the deviation
// from the lattice weight, d = f - w_q, rounded to nearest-even, and every kernel computes in
// FP32 on f = (float)d + w_q.  A value written to a 16-bit pool is therefore always
// w_q + round16(f - w_q); writing back a decoded value reproduces it exactly.
// ---------------------------------------------------------------------------------------------
struct bf16s {
    uint16_t bits;
};

template <typename T, typename S>
struct Store {  // S == T: the pool holds f
    __device__ static T dec(S s, T) { return s; }
    __device__ static S enc(T f, T) { return f; }
    __device__ static bool positive(S s, T) { return s > (T)0; }  // dec(s, w) > 0
    static constexpr bool quantized = false;
};

// Fused multiply-add a * b + c in the operands' own type, one rounding (v_fma_f32 / v_fma_f64);
// __builtin_fma alone would promote float operands to double and round twice.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Binary16 bits of the largest half value at or below the negative float c (round toward -inf),
// for |c| in the normal binary16 range: the truncated magnitude, plus one unit when bits were cut.
constexpr uint16_t half_round_down_neg(float c)
{
    const uint32_t b = __builtin_bit_cast(uint32_t, c);
    const uint32_t e16 = ((b >> 23) & 0xffu) - 127u + 15u, m = b & 0x7fffffu;
    uint32_t h = 0x8000u | (e16 << 10) | (m >> 13);
    if (m & 0x1fffu) h += 1u;
    return (uint16_t)h;
}

template <>
struct Store<float, _Float16> {
    // (float)s + w: one mixed-precision FMA, s * 1 + w with the exact half-to-float conversion of
    // s and one rounding -- the conversion and the add.
    __device__ static float dec(_Float16 s, float w)
    {
        float r;
        const uint32_t bits = __builtin_bit_cast(uint16_t, s);
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(bits), "v"(w));
        return r;
    }
    __device__ static _Float16 enc(float f, float w) { return (_Float16)(f - w); }
    // dec(s, w) > 0 without the add: a rounded float sum has the sign of the exact sum (a nonzero
    // exact sum of these two is a multiple of 2^-133, far from underflow), so the test is
    // (float)s > -w, and with t the largest half at or below -w that is s > t, one half compare
    // (no half lies strictly between t and -w).
    __device__ static bool positive(_Float16 s, float w)
    {
        return s > __builtin_bit_cast(_Float16, half_round_down_neg(-w));
    }
    static constexpr bool quantized = true;
};

template <>
struct Store<float, bf16s> {
    __device__ static float dec(bf16s s, float w) { return __builtin_bit_cast(float, (uint32_t)s.bits << 16) + w; }
    // A plain conversion to __bf16 is gfx950's v_cvt_pk_bf16_f32: round to nearest, ties to even,
    // subnormals kept, overflow to infinity; a NaN stays a NaN.
    __device__ static bf16s enc(float f, float w) { return bf16s{__builtin_bit_cast(uint16_t, (__bf16)(f - w))}; }
    __device__ static bool positive(bf16s s, float w) { return __builtin_bit_cast(float, (uint32_t)s.bits << 16) > -w; }
    static constexpr bool quantized = true;
};

// The value a store followed by a load yields.
template <typename T, typename S>
__device__ __forceinline__ T quant(T f, T w)
{
    if constexpr (Store<T, S>::quantized)
        return Store<T, S>::dec(Store<T, S>::enc(f, w), w);
    else
        return f;
}

}  // namespace agal
