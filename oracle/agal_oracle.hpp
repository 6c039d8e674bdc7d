/*
 * agal_oracle.hpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of KhodrJ/AGAL's octree-AMR lattice-Boltzmann step, used only as the parity
 * checker of the HIP library (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
 * Nothing in agal_amd/ links, loads or calls it.
 *
 * It follows the reference expression for expression: the same slot convention (direction p of
 * a pre-collision population lives in slot pb(p) between steps), the same p-order sums, the same
 * guards (solid cells hold -1, the stream swaps only when the partner value is > 0) and the same
 * integer bookkeeping of the block pool (gap stack, sorted per-level ID sets, neighbour and
 * neighbour-child tables).  Each function cites the reference lines it restates.
 *
 * Pinning (see tests/): velocity tables against the reference's text (tests/golden), isotropy
 * closed forms, equilibrium fixed point and conservation of BGK, stream as a permutation and
 * single-population translation, Couette closed form for the moving-wall hook, linear
 * reproduction of interpolation, average of constants, criterion on analytic shear fields, child
 * connectivity and cell-mask rules against tables parsed from the reference source, and AMR
 * invariants plus brute-force neighbour reconstruction from block coordinates.
 *
 * Deliberate difference from the reference, kept identical in the HIP library:
 *   - floor(log2(x)) in the refinement criterion is evaluated exactly (ilogb) instead of through
 *     log2f, whose last-bit rounding differs between math libraries; the two differ only when x
 *     lies within one ulp below a power of two.
 *   - id_max of an empty level is 0 (the reference dereferences an empty thrust range there,
 *     src/mesh/mesh_amr.cu:2629); id_max is only a loop bound over inactive IDs.
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/agal_capi.h"

namespace agal_oracle {

// Coefficient tables of the cubic interpolation (the reference's Cu_Interp2D/3D terms, pinned
// against its text by tests/test_oracle_tables.py).
#define AGAL_CUBIC_TABLE static const
#include "../agal_amd/csrc/cubic_table.inc"
#undef AGAL_CUBIC_TABLE

constexpr int N_SKIPID = -999;

constexpr int REF_UNREFINED = 0, REF_UNREFINED_VIO = 9, REF_REFINED = 1, REF_REFINED_WCHILD = 2,
              REF_REFINED_PERM = 3, REF_REFINED_WCHILD_PERM = 10, REF_MARK_REFINE = 4,
              REF_MARK_COARSEN = 5, REF_NEW = 6, REF_REMOVE = 7, REF_INACTIVE = 8, REF_INDET_E = 11,
              REF_INDET_O = 12;
constexpr int CELL_INTERIOR = 0, CELL_INTERFACE = 1, CELL_GHOST = 2, CELL_SOLID = -1, CELL_BOUNDARY = -6;
constexpr int BLK_REGULAR = 0, BLK_INTERFACE = 1, BLK_SOLID = -3, BLK_SOLIDA = -2, BLK_SOLIDB = -1;
constexpr int INTERP_INTERFACE = 0, INTERP_ADDED = 1;
constexpr int AVERAGE_INTERFACE = 0, AVERAGE_BLOCK = 1, AVERAGE_GRID = 2;

/* Velocity sets: V_CONN_ID / V_CONN_MAP (inc/util.h:94-101), LBMw / LBMpb
 * (src/solver_lbm/solver_lbm_init.cu:79-88).  The neighbour tables of the mesh always use the
 * full 9 (2D) or 27 (3D) directions in D2Q9 / D3Q27 order. */
struct VelSet {
    int Q = 0, DIM = 0, Q_max = 0;
    int c[3][27] = {};
    double w[27] = {};
    int pb[27] = {};
    int conn_map[27] = {};
};

inline VelSet make_velset(int Q)
{
    VelSet V;
    static const int c2[3][9] = {{0, 1, 0, -1, 0, 1, -1, -1, 1},
                                 {0, 0, 1, 0, -1, 1, 1, -1, -1},
                                 {0, 0, 0, 0, 0, 0, 0, 0, 0}};
    static const int c3[3][27] = {
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1},
        {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1},
        {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1}};
    V.Q = Q;
    V.DIM = (Q == 9) ? 2 : 3;
    V.Q_max = (V.DIM == 2) ? 9 : 27;
    for (int d = 0; d < 3; d++)
        for (int p = 0; p < V.Q_max; p++)
            V.c[d][p] = (V.DIM == 2) ? c2[d][p] : c3[d][p];
    // Opposite direction: the unique index with c = -c_p.
    for (int p = 0; p < V.Q_max; p++)
        for (int q = 0; q < V.Q_max; q++)
            if (V.c[0][q] == -V.c[0][p] && V.c[1][q] == -V.c[1][p] && V.c[2][q] == -V.c[2][p])
                V.pb[p] = q;
    // conn_map[S] = velocity index of the 3x3(x3) slot S = (dx+1) + 3(dy+1) + 9(dz+1).
    for (int p = 0; p < V.Q_max; p++) {
        int S = (V.c[0][p] + 1) + 3 * (V.c[1][p] + 1) + (V.DIM == 3 ? 9 * (V.c[2][p] + 1) : 0);
        V.conn_map[S] = p;
    }
    if (Q == 9) {
        double w[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                       1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
        for (int p = 0; p < 9; p++) V.w[p] = w[p];
    } else if (Q == 19) {
        V.w[0] = 1.0 / 3.0;
        for (int p = 1; p < 7; p++) V.w[p] = 1.0 / 18.0;
        for (int p = 7; p < 19; p++) V.w[p] = 1.0 / 36.0;
    } else {
        V.w[0] = 8.0 / 27.0;
        for (int p = 1; p < 7; p++) V.w[p] = 2.0 / 27.0;
        for (int p = 7; p < 19; p++) V.w[p] = 1.0 / 54.0;
        for (int p = 19; p < 27; p++) V.w[p] = 1.0 / 216.0;
    }
    return V;
}

/* 16-bit population pools (no reference counterpart): the pool holds d = f - w_q rounded to
 * nearest-even in binary16 or bfloat16; a stored value reads back as (float)d + w_q.  The
 * conversions are written out so the restatement does not depend on compiler support for 16-bit
 * types; tests compare them with numpy's float16 and torch's bfloat16. */
inline uint32_t f32_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u;
}
inline float bits_f32(uint32_t u)
{
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

inline uint16_t f16_from_f32(float x)
{
    const uint32_t u = f32_bits(x);
    const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    const uint32_t a = u & 0x7FFFFFFFu;
    if (a > 0x7F800000u) return (uint16_t)(sign | 0x7E00u);  // NaN
    if (a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to infinity (>= 65520)
    if (a < 0x38800000u) {                                     // below 2^-14: subnormal or zero
        // |x| * 2^24 is exact in double; nearbyint rounds half to even.
        const double m = std::nearbyint((double)bits_f32(a) * 16777216.0);
        return (uint16_t)(sign | (uint16_t)m);                 // m == 1024 is the smallest normal
    }
    const uint32_t e = (a >> 23) - 127 + 15, mant = a & 0x7FFFFFu;
    uint32_t h = (e << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // carries into the exponent
    return (uint16_t)(sign | h);
}

inline float f16_to_f32(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    if (e == 0) {
        const float v = (float)((double)m / 16777216.0);
        return (h & 0x8000u) ? -v : v;
    }
    if (e == 31) return bits_f32(sign | 0x7F800000u | (m << 13));
    return bits_f32(sign | ((e - 15 + 127) << 23) | (m << 13));
}

inline uint16_t bf16_from_f32(float x)
{
    uint32_t u = f32_bits(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

inline float bf16_to_f32(uint16_t h) { return bits_f32((uint32_t)h << 16); }

/* Half set of directions that initiate a swap in Cu_Stream / pair in Cu_ImposeBC
 * (solver_lbm_kernels.cu:597,769). */
inline bool in_half_set(int dim, int p)
{
    if (dim == 2) return p == 1 || (p + 1) % 3 == 0;
    return p == 26 || ((p - 1) % 2 == 0 && p < 25);
}

/* Cu_NbrMap (inc/util_index.h:43-66): 3x3(x3) slot of the block holding local cell (I,J,K). */
inline int nbr_map(int dim, int I, int J, int K)
{
    int S = 0;
    if (I >= 4) S = 2;
    if (I >= 0 && I < 4) S = 1;
    if (J >= 4) S += 3 * 2;
    if (J >= 0 && J < 4) S += 3 * 1;
    if (dim == 3) {
        if (K >= 4) S += 9 * 2;
        if (K >= 0 && K < 4) S += 9 * 1;
    }
    return S;
}

/* Cu_NbrCellId (inc/util_index.h:88-98). */
inline int nbr_cell_id(int dim, int I, int J, int K)
{
    I = (4 + (I % 4)) % 4;
    J = (4 + (J % 4)) % 4;
    if (dim == 3) K = (4 + (K % 4)) % 4;
    return (dim == 2) ? I + 4 * J : I + 4 * J + 16 * K;
}

/* Neighbour of child c (= ci + 2cj + 4ck) in direction p, read from the parent's neighbour-child
 * table: the parent-level direction whose block contains the target position, and the child
 * offset inside that block (the table of Cu_UpdateConnectivity, src/mesh/mesh_amr.cu:154-944). */
inline void child_neighbor_rule(const VelSet &V, int c, int p, int &src_dir, int &off)
{
    int cc[3] = {c % 2, (c / 2) % 2, (c / 4) % 2};
    int d[3], o[3];
    for (int a = 0; a < 3; a++) {
        int tgt = cc[a] + V.c[a][p];
        d[a] = tgt < 0 ? -1 : (tgt > 1 ? 1 : 0);
        o[a] = (tgt + 2) % 2;
    }
    int S = (d[0] + 1) + 3 * (d[1] + 1) + (V.DIM == 3 ? 9 * (d[2] + 1) : 0);
    src_dir = V.conn_map[S];
    off = o[0] + 2 * o[1] + 4 * o[2];
}

/* Cell mask of local cell t of a block whose direction-p neighbour is missing when skip[p] != 0
 * (Cu_UpdateMasks_2, src/mesh/mesh_amr.cu:981-1297, one sub-block per block). */
inline int cell_mask_rule(const VelSet &V, const int *skip, int t)
{
    int ijk[3] = {t % 4, (t / 4) % 4, (V.DIM == 3) ? t / 16 : 0};
    int m = 0;
    for (int p = 1; p < V.Q_max; p++)
        if (skip[p]) m = 1;
    for (int p = 1; p < V.Q_max; p++) {
        if (!skip[p]) continue;
        bool outer = true;
        for (int a = 0; a < V.DIM; a++) {
            if (V.c[a][p] == 1 && !(ijk[a] >= 2)) outer = false;
            if (V.c[a][p] == -1 && !(ijk[a] < 2)) outer = false;
        }
        if (outer) m = 2;
    }
    return m;
}

/* Exact floor(log2(x)) for float/double x (see header note). */
template <typename T>
inline T floor_log2_exact(T x)
{
    if (std::isnan(x)) return x;
    if (x == T(0)) return -std::numeric_limits<T>::infinity();
    if (std::isinf(x)) return x;  // +inf (x >= 0 here)
    return T(std::ilogb(x));
}

// ---------------------------------------------------------------------------------------------
// Float geometry of the face-cell links, in the reference's operation order (ufloat_g_t = float,
// src/main.cu:61): vec3 arithmetic and face helpers (inc/util.h:273-560), triangle-box overlap by
// separating axes (inc/util_tribin.h:81-160).  Compiled without FMA contraction.
// ---------------------------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline float dotv(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 crossv(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 unitv(V3 a)
{
    const float onorm = 1.0f / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return a * onorm;
}
inline V3 face_vertex(const float *tri, long f, int k) { return V3{tri[9 * f + 3 * k], tri[9 * f + 3 * k + 1], tri[9 * f + 3 * k + 2]}; }
// FaceNormal / FaceNormalUnit (util.h:419-432), 3D.
inline V3 face_normal(V3 v1, V3 v2, V3 v3) { return crossv(v2 - v1, v3 - v1); }
inline V3 face_normal_unit(V3 v1, V3 v2, V3 v3) { return unitv(face_normal(v1, v2, v3)); }
// CheckPointInTriangleExtended (util.h:522-545): inside or within eps of every edge line.
inline bool point_in_triangle_ext(V3 vp, V3 v1, V3 v2, V3 v3, V3 n, float eps = std::numeric_limits<float>::epsilon())
{
    V3 ven = unitv(crossv(v2 - v1, n));
    if (dotv(v1 - vp, ven) <= -eps) return false;
    ven = unitv(crossv(v3 - v2, n));
    if (dotv(v2 - vp, ven) <= -eps) return false;
    ven = unitv(crossv(v1 - v3, n));
    if (dotv(v3 - vp, ven) <= -eps) return false;
    return true;
}
// TrianglePlaneCutsBox (util_tribin.h:81-94).
inline bool plane_cuts_box(V3 vm, V3 vM, V3 v1, V3 v2, V3 v3)
{
    const V3 n = face_normal(v1, v2, v3);
    const V3 c{n.x > 0 ? (vM.x - vm.x) : 0.0f, n.y > 0 ? (vM.y - vm.y) : 0.0f, n.z > 0 ? (vM.z - vm.z) : 0.0f};
    const float d = dotv(n, vm);
    const float d1 = dotv(n, c - v1);
    const float d2 = dotv(n, (vM - vm - c) - v1);
    return (d + d1) * (d + d2) <= 0;
}
// TriangleBinOverlap2D_AxisGap (util_tribin.h:96-128): true when `axis` separates the projected
// triangle from the projected box.
inline bool axis_gap_2d(int axis, float v1x, float v1y, float v2x, float v2y, float v3x, float v3y, float rmx, float rmy,
                        float rMx, float rMy)
{
    float ex = 0.0f, ey = 0.0f;
    if (axis == 0) { ex = 1.0f; ey = 0.0f; }
    if (axis == 1) { ex = 0.0f; ey = 1.0f; }
    if (axis == 2) { ex = v2y - v1y; ey = v1x - v2x; }
    if (axis == 3) { ex = v3y - v2y; ey = v2x - v3x; }
    if (axis == 4) { ex = v1y - v3y; ey = v3x - v1x; }
    float t = std::fmax(std::fmax(v1x * ex + v1y * ey, v2x * ex + v2y * ey), v3x * ex + v3y * ey);
    float r = std::fmin(std::fmin(std::fmin(rmx * ex + rmy * ey, rMx * ex + rmy * ey), rmx * ex + rMy * ey), rMx * ex + rMy * ey);
    bool b = t < r;
    t = std::fmin(std::fmin(v1x * ex + v1y * ey, v2x * ex + v2y * ey), v3x * ex + v3y * ey);
    r = std::fmax(std::fmax(std::fmax(rmx * ex + rmy * ey, rMx * ex + rmy * ey), rmx * ex + rMy * ey), rMx * ex + rMy * ey);
    return b || (r < t);
}
// TriangleBinOverlap3D (util_tribin.h:131-160): the plane test, then five axes in each of the XY,
// YZ and ZX projections.
inline bool tri_box_overlap(V3 vm, V3 vM, V3 v1, V3 v2, V3 v3)
{
    if (!plane_cuts_box(vm, vM, v1, v2, v3)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap_2d(a, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y, vm.x, vm.y, vM.x, vM.y)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap_2d(a, v1.y, v1.z, v2.y, v2.z, v3.y, v3.z, vm.y, vm.z, vM.y, vM.z)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap_2d(a, v1.z, v1.x, v2.z, v2.x, v3.z, v3.x, vm.z, vm.x, vM.z, vM.x)) return false;
    return true;
}

// Face bins of the link computation: Geometry::Bins with G_MakeBinsGPU_MD (geometry_bins_gpu.cu:
// 1043-1336), for the 3D velocity sets.  bin_start has n_bins + 1 offsets into faces.
struct LinkBins {
    int level = 0;         // n_bin_level_MD = min(MAX_LEVELS_WALL, G_BIN_LEVELS) - 1 (geometry_bins.h:201)
    int G = 0;             // bins per unit length: G_BIN_DENSITY * 2^level
    bool use_ray = false;  // the ray-indicator filter (geometry_bins_gpu.cu:1060-1065)
    std::vector<int> filtered;   // face ids that pass the filter, ascending (all when unused)
    std::vector<int> bin_start;  // CSR offsets, n_bins + 1
    std::vector<int> faces;      // per bin: (overlap rank, face) ascending
};

// Cu_ComputeVoxelRayIndicators_MD (geometry_bins_gpu.cu:22-157), 3D: the face is kept when the line
// through some level cell centre of its bounding box, along some half-set direction, meets it.
inline bool face_ray_indicator(const float *tri, long f, float dx_L, float dx_Lo2, float Nx, float Ny, float Nz,
                               const int conn[3][27])
{
    const V3 v1 = face_vertex(tri, f, 0), v2 = face_vertex(tri, f, 1), v3 = face_vertex(tri, f, 2);
    const V3 n = face_normal_unit(v1, v2, v3);
    const int ixmin = (int)(std::fmin(std::fmin(v1.x, v2.x), v3.x) * Nx);
    const int iymin = (int)(std::fmin(std::fmin(v1.y, v2.y), v3.y) * Ny);
    const int izmin = (int)(std::fmin(std::fmin(v1.z, v2.z), v3.z) * Nz);
    const int ixmax = (int)(std::fmax(std::fmax(v1.x, v2.x), v3.x) * Nx);
    const int iymax = (int)(std::fmax(std::fmax(v1.y, v2.y), v3.y) * Ny);
    const int izmax = (int)(std::fmax(std::fmax(v1.z, v2.z), v3.z) * Nz);
    for (int p = 1; p < 27; p++) {
        if (!(p == 26 || ((p - 1) % 2 == 0 && p < 25))) continue;
        const V3 ray{(float)conn[0][p], (float)conn[1][p], (float)conn[2][p]};
        for (int iz = izmin; iz <= izmax; iz++)
            for (int iy = iymin; iy <= iymax; iy++)
                for (int ix = ixmin; ix <= ixmax; ix++) {
                    const V3 vp{dx_Lo2 + dx_L * (float)ix, dx_Lo2 + dx_L * (float)iy, dx_Lo2 + dx_L * (float)iz};
                    const float d = dotv(v1 - vp, n) / dotv(ray, n);
                    const V3 vi = vp + ray * d;
                    if (point_in_triangle_ext(vi, v1, v2, v3, n)) return true;
                }
    }
    return false;
}

// The bins.  Cu_ComputeBoundingBoxLimits3D (geometry_bins_gpu.cu:421-561) gives each face its
// first (2 + G_BIN_SPEC)^3 overlapping bins, boxes grown by dx_L + FLT_EPSILON, scanning K, J, I
// from one bin below to one above the face's bounding box; the (bin, face) pairs sit at
// position rank * n_filtered + filtered index and are stably sorted by bin (thrust's radix sort).
// Pairs whose bin index falls outside [0, n_bins) are dropped (the reference scatters them out of
// bounds).
inline void make_link_bins(LinkBins &B, const float *tri, long n_faces, const agal_params &P, const int conn[3][27])
{
    const int n_bin_levels = std::min(P.G_BIN_LEVELS, P.MAX_LEVELS_WALL);
    B.level = std::min(P.MAX_LEVELS_WALL, n_bin_levels) - 1;
    B.G = P.G_BIN_DENSITY << B.level;
    B.use_ray = P.MAX_LEVELS_WALL == B.level + 1;
    const float Lx = (float)P.L_c, Ly = (float)(P.L_fy * P.L_c), Lz = (float)(P.L_fz * P.L_c);
    const int Nx = P.Nx, Ny = (int)(Nx * (Ly / Lx)), Nz = (int)(Nx * (Lz / Lx));
    float dxg = Lx / (float)Nx, Lx0g = Lx / (float)P.G_BIN_DENSITY, Ly0g = Ly / (float)P.G_BIN_DENSITY,
          Lz0g = Lz / (float)P.G_BIN_DENSITY;
    int NxL = Nx, NyL = Ny, NzL = Nz;
    for (int L = 1; L <= B.level; L++) {
        dxg *= 0.5f;
        Lx0g *= 0.5f;
        Ly0g *= 0.5f;
        Lz0g *= 0.5f;
        NxL *= 2;
        NyL *= 2;
        NzL *= 2;
    }
    B.filtered.clear();
    for (long f = 0; f < n_faces; f++)
        if (!B.use_ray || face_ray_indicator(tri, f, dxg, 0.5f * dxg, (float)NxL, (float)NyL, (float)NzL, conn))
            B.filtered.push_back((int)f);
    const long nf = (long)B.filtered.size();
    const long n_bins = (long)B.G * B.G * B.G;
    const int n_limits = (2 + P.G_BIN_SPEC) * (2 + P.G_BIN_SPEC) * (2 + P.G_BIN_SPEC);
    const float dx = dxg + std::numeric_limits<float>::epsilon();
    struct E {
        long bin, pos;
        int face;
    };
    std::vector<E> ents;
    for (long m = 0; m < nf; m++) {
        const int f = B.filtered[m];
        const V3 v1 = face_vertex(tri, f, 0), v2 = face_vertex(tri, f, 1), v3 = face_vertex(tri, f, 2);
        const V3 vBm{std::fmin(std::fmin(v1.x, v2.x), v3.x), std::fmin(std::fmin(v1.y, v2.y), v3.y),
                     std::fmin(std::fmin(v1.z, v2.z), v3.z)};
        const V3 vBM{std::fmax(std::fmax(v1.x, v2.x), v3.x), std::fmax(std::fmax(v1.y, v2.y), v3.y),
                     std::fmax(std::fmax(v1.z, v2.z), v3.z)};
        bool C = true;
        if ((vBm.x < -dx && vBM.x < -dx) || (vBm.x > Lx + dx && vBM.x > Lx + dx)) C = false;
        if ((vBm.y < -dx && vBM.y < -dx) || (vBm.y > Ly + dx && vBM.y > Ly + dx)) C = false;
        if ((vBm.z < -dx && vBM.z < -dx) || (vBm.z > Lz + dx && vBM.z > Lz + dx)) C = false;
        if (!C) continue;
        const int xl = (int)(vBm.x * (float)B.G) - 1, yl = (int)(vBm.y * (float)B.G) - 1, zl = (int)(vBm.z * (float)B.G) - 1;
        const int xL = (int)(vBM.x * (float)B.G) + 1, yL = (int)(vBM.y * (float)B.G) + 1, zL = (int)(vBM.z * (float)B.G) + 1;
        int counter = 0;
        for (int K = zl; K < zL + 1; K++)
            for (int J = yl; J < yL + 1; J++)
                for (int I = xl; I < xL + 1; I++) {
                    const V3 vm{(float)I * Lx0g - dx, (float)J * Ly0g - dx, (float)K * Lz0g - dx};
                    const V3 vM{(float)(I + 1) * Lx0g + dx, (float)(J + 1) * Ly0g + dx, (float)(K + 1) * Lz0g + dx};
                    if (tri_box_overlap(vm, vM, v1, v2, v3) && counter < n_limits) {
                        const long gid = (long)I + (long)B.G * J + (long)B.G * B.G * K;
                        if (gid >= 0 && gid < n_bins) ents.push_back(E{gid, (long)counter * nf + m, f});
                        counter++;
                    }
                }
    }
    std::sort(ents.begin(), ents.end(), [](const E &a, const E &b) { return a.bin != b.bin ? a.bin < b.bin : a.pos < b.pos; });
    B.bin_start.assign(n_bins + 1, 0);
    B.faces.resize(ents.size());
    for (size_t k = 0; k < ents.size(); k++) {
        B.bin_start[ents[k].bin + 1]++;
        B.faces[k] = ents[k].face;
    }
    for (long k = 0; k < n_bins; k++) B.bin_start[k + 1] += B.bin_start[k];
}

template <typename T>
struct Oracle {
    agal_params P{};
    VelSet V;
    int DIM = 3, Q = 19, Q_max = 27, M_CBLOCK = 64, N_CHILDREN = 8, N_U = 23;
    long n_maxcblocks = 0, n_maxcells = 0;
    int Nxi[3] = {0, 0, 0};
    int MAX_LEVELS = 1, MAX_LEVELS_INTERIOR = 1, MAX_LEVELS_WALL = 1, N_LEVEL_START = 0;
    T dx = 1, dx_cblock = 1;
    std::vector<T> dxf_vec, tau_vec;
    int n_coarsecblocks = 0;

    // Mesh-owned arrays (inc/mesh.h:367-462).  F holds N_Q populations then rho,u,v,w.
    std::vector<T> F;
    std::vector<int> cells_mask;
    std::vector<float> X;
    std::vector<int> blk_mask, nbr, nbr_child, onb, ref, level;
    std::vector<int> id_set, n_ids, id_max, gap_set;
    int n_gaps = 0;
    int global_Liter = 0;
    // 0: populations kept in T.  1 / 2: a binary16 / bfloat16 pool of f - w_q; F then holds the
    // value each stored population reads back as, and every write into the pool goes through
    // quant().  The pool is written where the build's kernels write it: after the fused
    // collide + boundary hook (before the swap, whose guard sees stored values), by the initial
    // condition, by interpolation and averaging, and by uploads.
    int store_mode = 0;
    // Planted mistake for the tests that show the pins catch it (0 = none, the restatement as
    // written).  1: interpolation without the tau rescale; 2: averaging without it; 3: linear
    // interpolation at child offsets +1/4 instead of -1/4; 4: averaging that drops one child cell
    // of the 2^d group; 5: the advance skips the mid-cycle interpolation; 6: the advance skips
    // the average after a level's second sub-cycle; 7: the root level averages before its finer
    // levels advance; 8: the stream swaps without the f_pb > 0 guard; 9: the control-volume inflow
    // tested and weighted along c_p instead of c_pb(p); 10: the control-volume outflow without the
    // collision.
    int plant = 0;

    T quant(T f, int dir) const
    {
        if (store_mode == 0) return f;
        const float w = (float)V.w[dir];
        const float d = (float)f - w;
        const float r = store_mode == 1 ? f16_to_f32(f16_from_f32(d)) : bf16_to_f32(bf16_from_f32(d));
        return (T)(r + w);
    }

    // ---------------------------------------------------------------------------------------
    // Mesh<..>::M_Init (src/mesh/mesh_init.cu:10-567) and Solver_LBM::S_Init (solver_lbm_init.cu).
    // ---------------------------------------------------------------------------------------
    int init(const agal_params &p, long long capacity)
    {
        P = p;
        V = make_velset(P.velocity_set);
        store_mode = (P.precision == AGAL_FP16) ? 1 : (P.precision == AGAL_BF16 ? 2 : 0);
        if (store_mode && sizeof(T) != sizeof(float)) return AGAL_ERR_ARG;
        Q = V.Q;
        DIM = V.DIM;
        Q_max = V.Q_max;
        M_CBLOCK = (DIM == 2) ? 16 : 64;
        N_CHILDREN = (DIM == 2) ? 4 : 8;
        N_U = Q + 4;
        // mesh_init.cu:22-29: Lx is a float member (inc/mesh.h:292).
        float Lx = (float)P.L_c;
        float Ly = (float)(P.L_fy * Lx);
        float Lz = (float)(P.L_fz * Lx);
        Nxi[0] = P.Nx;
        Nxi[1] = (int)(P.Nx * (Ly / Lx));
        Nxi[2] = (int)(P.Nx * (Lz / Lx));
        dx = (T)(Lx / P.Nx);
        dx_cblock = 4 * dx;
        MAX_LEVELS_WALL = P.MAX_LEVELS_WALL;
        MAX_LEVELS_INTERIOR = P.MAX_LEVELS_INTERIOR;
        N_LEVEL_START = P.N_LEVEL_START;
        if (DIM == 2) Nxi[2] = 1;
        MAX_LEVELS = std::max(MAX_LEVELS_WALL, MAX_LEVELS_INTERIOR);
        if (MAX_LEVELS == 1) N_LEVEL_START = 0;
        if (N_LEVEL_START > MAX_LEVELS - 1) N_LEVEL_START = MAX_LEVELS - 1;
        if (MAX_LEVELS > 8) return AGAL_ERR_ARG;

        n_maxcblocks = capacity;
        n_maxcells = n_maxcblocks * M_CBLOCK;
        n_coarsecblocks = (Nxi[0] * Nxi[1] * Nxi[2]) / M_CBLOCK;
        if ((long)Nxi[0] * Nxi[1] * Nxi[2] >= n_maxcells) return AGAL_ERR_CAPACITY;

        dxf_vec.assign(MAX_LEVELS, 0);
        for (int L = 0; L < MAX_LEVELS; L++)
            dxf_vec[L] = (T)((double)dx / std::pow(2.0, (double)(T)L));
        // solver_lbm_init.cu:47-52
        tau_vec.assign(MAX_LEVELS, 0);
        for (int L = 0; L < MAX_LEVELS; L++)
            tau_vec[L] = (T)(P.v0 * 3.0 + 0.5 * (double)dxf_vec[L]);
        // solver_lbm_init.cu:36-41: the control volume snapped to root-block edges.
        const double fr[6] = {P.S_FORCEVOLUME_Xm, P.S_FORCEVOLUME_XM, P.S_FORCEVOLUME_Ym,
                              P.S_FORCEVOLUME_YM, P.S_FORCEVOLUME_Zm, P.S_FORCEVOLUME_ZM};
        for (int k = 0; k < 6; k++) cv_box[k] = (T)(4.0 * (double)(int)(fr[k] * (double)(Nxi[k / 2] / 4)) * (double)dxf_vec[0]);

        F.assign((size_t)N_U * n_maxcells, T(0));
        cells_mask.assign(n_maxcells, CELL_INTERIOR);
        X.assign((size_t)DIM * n_maxcblocks, -1.0f);
        blk_F.assign((size_t)3 * n_maxcblocks, T(0));
        cv_F.assign(P.S_FORCE_TYPE == 1 ? (size_t)6 * n_maxcblocks : 0, T(0));
        blk_mask.assign(n_maxcblocks, BLK_REGULAR);
        nbr.assign((size_t)Q_max * n_maxcblocks, N_SKIPID);
        nbr_child.assign((size_t)Q_max * n_maxcblocks, N_SKIPID);
        ref.assign(n_maxcblocks, REF_INACTIVE);
        level.assign(n_maxcblocks, 0);
        onb.assign(n_maxcblocks, 0);
        id_set.assign((size_t)MAX_LEVELS * n_maxcblocks, 0);
        n_ids.assign(MAX_LEVELS + 1, 0);
        id_max.assign(MAX_LEVELS + 1, 0);
        gap_set.assign(n_maxcblocks, -1);

        // Root grid over a padded structured index grid (mesh_init.cu:325-501).
        int bx = Nxi[0] / 4, by = Nxi[1] / 4;
        int gh = (DIM == 2) ? 1 : Nxi[2] / 4;
        int ghm = (DIM == 2) ? 0 : 1;
        int px = bx + 2, py = by + 2, pz = gh + 2;
        std::vector<int> grid((size_t)px * py * pz, 0);
        auto G = [&](int i, int j, int k) -> int & { return grid[(size_t)i + px * ((size_t)j + py * k)]; };
        int ID_counter = 0;
        for (int k = 1 * ghm; k < gh * ghm + 2 - 1; k++)
            for (int j = 1; j < by + 2 - 1; j++)
                for (int i = 1; i < bx + 2 - 1; i++)
                    G(i, j, k) = ID_counter++;
        if (DIM == 3) {
            for (int j = 0; j < by + 2; j++)
                for (int i = 0; i < bx + 2; i++) {
                    if (P.PERIODIC_Z == 1) {
                        G(i, j, 0) = G(i, j, gh + 2 - 1 - 1);
                        G(i, j, gh + 2 - 1) = G(i, j, 1);
                    } else {
                        G(i, j, 0) = -5;
                        G(i, j, gh + 2 - 1) = -6;
                    }
                }
        }
        for (int k = 0; k < gh + 2 * ghm; k++)
            for (int i = 0; i < bx + 2; i++) {
                if (P.PERIODIC_Y == 1) {
                    G(i, 0, k) = G(i, by + 2 - 1 - 1, k);
                    G(i, by + 2 - 1, k) = G(i, 1, k);
                } else {
                    G(i, 0, k) = -3;
                    G(i, by + 2 - 1, k) = -4;
                }
            }
        for (int k = 0; k < gh + 2 * ghm; k++)
            for (int j = 0; j < by + 2; j++) {
                if (P.PERIODIC_X == 1) {
                    G(0, j, k) = G(bx + 2 - 1 - 1, j, k);
                    G(bx + 2 - 1, j, k) = G(1, j, k);
                } else {
                    G(0, j, k) = -1;
                    G(bx + 2 - 1, j, k) = -2;
                }
            }
        for (int k = 1 * ghm; k < gh * ghm + 2 - 1; k++)
            for (int j = 1; j < by + 2 - 1; j++)
                for (int i = 1; i < bx + 2 - 1; i++) {
                    int b = G(i, j, k);
                    if (b < 0) continue;
                    X[b + 0 * n_maxcblocks] = (float)((T)(i - 1) * dx_cblock);
                    X[b + 1 * n_maxcblocks] = (float)((T)(j - 1) * dx_cblock);
                    if (DIM == 3) X[b + 2 * n_maxcblocks] = (float)((T)(k - 1) * dx_cblock);
                    int on_b = 0;
                    for (int p = 0; p < Q_max; p++) {
                        int nb = G(i + V.c[0][p], j + V.c[1][p], ghm * (k + V.c[2][p]));
                        nbr[b + p * n_maxcblocks] = nb;
                        if (nb < 0) on_b = 1;
                        nbr_child[b + p * n_maxcblocks] = nb < 0 ? nb : N_SKIPID;
                    }
                    onb[b] = on_b;
                }
        for (int kap = 0; kap < n_coarsecblocks; kap++) ref[kap] = REF_UNREFINED;
        n_ids[0] = n_coarsecblocks;
        id_max[0] = id_max[MAX_LEVELS] = n_ids[0];
        for (int xi = 0; xi < n_ids[0]; xi++) id_set[xi] = xi;
        for (long xi = 0; xi < n_maxcblocks - n_ids[0]; xi++) gap_set[xi] = (int)((n_maxcblocks - 1) - xi);
        n_gaps = (int)(n_maxcblocks - n_ids[0]);
        return AGAL_OK;
    }

    T *Fp(int slot) { return &F[(size_t)slot * n_maxcells]; }
    const int *idset(int L) const { return &id_set[(size_t)L * n_maxcblocks]; }
    void cell_ijk(int t, int &I, int &J, int &K) const
    {
        I = t % 4;
        J = (t / 4) % 4;
        K = (DIM == 3) ? (t / 4) / 4 : 0;
    }

    // Cu_ComputeIC hook (inc/custom.h:48-60): a uniform state from the parameters.
    void ic_state(T &rho, T &u, T &v, T &w) const
    {
        rho = (T)P.ic_rho;
        u = (T)P.ic_u;
        v = (T)P.ic_v;
        w = (T)P.ic_w;
    }

    // Cu_ImposeBC hook (inc/custom.h:66-123) for the named presets.
    T bc_hook(int nbr_id, T f, T rho, T u, T v, T w, T wp, T cxp, T cyp, T czp) const
    {
        (void)rho;
        const T U = (T)P.bc_velocity;
        int pre = P.bc_preset;
        bool wall = false, outflow = false;
        if (pre == AGAL_BC_LDC2D) wall = (nbr_id == -4);
        if (pre == AGAL_BC_LDC3D) wall = (nbr_id == -6);
        if (pre == AGAL_BC_FPSC2D) { wall = (nbr_id == -1 || nbr_id == -3 || nbr_id == -4); outflow = (nbr_id == -2); }
        if (pre == AGAL_BC_FPSC3D) {
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

    // Equilibrium in the reference's operation order (solver_lbm_kernels.cu:282).
    T feq(int p, T rho, T u, T v, T w, T udotu) const
    {
        T cdotu = (T)V.c[0][p] * u + (T)V.c[1][p] * v + (T)V.c[2][p] * w;
        return (T)V.w[p] * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
    }

    // ---------------------------------------------------------------------------------------
    // Cu_SetInitialConditions (solver_lbm_kernels.cu:13-77).
    // ---------------------------------------------------------------------------------------
    void set_ic(int L)
    {
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            for (int t = 0; t < M_CBLOCK; t++) {
                T rho = 0, u = 0, v = 0, w = 0;
                ic_state(rho, u, v, w);
                T udotu = u * u + v * v + w * w;
                long cell = (long)b * M_CBLOCK + t;
                int valid_mask = cells_mask[cell];
                for (int p = 0; p < Q; p++) {
                    T cdotu = (T)V.c[0][p] * u + (T)V.c[1][p] * v + (T)V.c[2][p] * w;
                    T f_p = (T)(-1.0);
                    if (valid_mask != CELL_SOLID)
                        f_p = rho * (T)V.w[p] * ((T)1.0 + (T)3.0 * cdotu + (T)4.5 * cdotu * cdotu - (T)1.5 * udotu);
                    Fp(V.pb[p])[cell] = quant(f_p, p);
                }
                Fp(Q + 0)[cell] = rho;
                Fp(Q + 1)[cell] = u;
                Fp(Q + 2)[cell] = v;
                if (DIM == 3) Fp(Q + 3)[cell] = w;
            }
        }
    }

    // Moments of the pre-collision populations (read from slots pb(p)), p order.
    void moments(long cell, T &rho, T &u, T &v, T &w)
    {
        rho = 0; u = 0; v = 0; w = 0;
        for (int p = 0; p < Q; p++) {
            T f = Fp(V.pb[p])[cell];
            rho += f;
            u += V.c[0][p] * f;
            v += V.c[1][p] * f;
            if (DIM == 3) w += V.c[2][p] * f;
        }
        // One correctly rounded division, then products (as the GPU kernels).
        const T rinv = (T)(1.0) / rho;
        u = u * rinv;
        v = v * rinv;
        if (DIM == 3) w = w * rinv;
    }

    // Cu_RefreshVariables (solver_lbm_kernels.cu:110-159).
    void refresh(int L)
    {
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)b * M_CBLOCK + t;
                T rho, u, v, w;
                moments(cell, rho, u, v, w);
                Fp(Q + 0)[cell] = rho;
                Fp(Q + 1)[cell] = u;
                Fp(Q + 2)[cell] = v;
                if (DIM == 3) Fp(Q + 3)[cell] = w;
            }
        }
    }

    bool block_streams(int b) const
    {
        // solver_lbm_kernels.cu:236,570: skipped only when the block has children and is solid.
        return nbr_child[b] < 0 || blk_mask[b] != BLK_SOLID;
    }

    // Cu_Collide (solver_lbm_kernels.cu:204-289) + S_Collide (:464-510).
    void collide(int L)
    {
        if (n_ids[L] <= 0) return;
        if (MAX_LEVELS == 1) refresh(L);
        const int nL = n_ids[L];
        const int *ids = idset(L);
        const T dx_L = dxf_vec[L], tau_L = tau_vec[L];
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0 || !block_streams(b)) continue;
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)b * M_CBLOCK + t;
                int valid_mask = cells_mask[cell];
                T rho = 0, u = 0, v = 0, w = 0;
                T f_p[27];
                for (int p = 0; p < Q; p++) {
                    T f_pi = Fp(V.pb[p])[cell];
                    rho += f_pi;
                    u += V.c[0][p] * f_pi;
                    v += V.c[1][p] * f_pi;
                    if (DIM == 3) w += V.c[2][p] * f_pi;
                    f_p[p] = f_pi;
                }
                // One correctly rounded division, then products.
                const T rinv = (T)(1.0) / rho;
                u = u * rinv;
                v = v * rinv;
                if (DIM == 3) w = w * rinv;
                T udotu = u * u + v * v + w * w;
                Fp(Q + 0)[cell] = rho;
                Fp(Q + 1)[cell] = u;
                Fp(Q + 2)[cell] = v;
                if (DIM == 3) Fp(Q + 3)[cell] = w;
                T omeg = dx_L / tau_L;
                T omegp = (T)(1.0) - omeg;
                if (P.collision == AGAL_MRT) {
                    T fs[27];
                    mrt_collide(f_p, fs, rho, u, v, w, udotu, omeg);
                    for (int p = 0; p < Q; p++)
                        if (valid_mask != CELL_SOLID) Fp(p)[cell] = fs[p];
                    continue;
                }
                // The equilibrium polynomial in fused multiply-adds (std::fma, correctly rounded as
                // the GPU's): b = (4.5 c.u) c.u + (1 - 1.5 u.u), feq / (w rho) = 3 c.u + b,
                // f* = (w rho omega) (feq / (w rho)) + (1 - omega) f.
                const T c1 = (T)(1.0) - (T)(1.5) * udotu;
                const T ro = rho * omeg;
                for (int p = 0; p < Q; p++) {
                    T cdotu = (T)V.c[0][p] * u + (T)V.c[1][p] * v + (T)V.c[2][p] * w;
                    T b = std::fma((T)(4.5) * cdotu, cdotu, c1);
                    T f_pi = std::fma((T)V.w[p] * ro, std::fma((T)(3.0), cdotu, b), f_p[p] * omegp);
                    if (valid_mask != CELL_SOLID) Fp(p)[cell] = f_pi;
                }
            }
        }
    }

    // ---------------------------------------------------------------------------------------
    // MRT (CollisionOperator::MRT; the reference has it only as legacy D3Q19 code,
    // src/solver_old/solver_lbm_collide_mrt_d3q19.cu:14-1460).  f* = f - M^T D^-1 S M (f - feq)
    // with an integer orthogonal basis M: for D3Q19 the d'Humieres et al. (2002) rows of the
    // legacy transform (solver_lbm_vel2mom_mrt_d3q19.cu:118-136) and its rates (legacy collide
    // :132-146: s1 e, s2 eps, s3 q, s5 pi, s6 m, omega for the stresses); for D2Q9 / D3Q27 the
    // products of {1, c, 3c^2-2} per axis with the second-order block split into trace (s1) and
    // deviators (omega).  M (f - feq) equals m - m_eq of the legacy form, whose equilibrium
    // moments are those of feq.
    // ---------------------------------------------------------------------------------------
    int mrt_m(int k, int p) const
    {
        const int x = V.c[0][p], y = V.c[1][p], z = V.c[2][p];
        auto P2 = [](int c) { return 3 * c * c - 2; };
        if (Q == 19) {
            const int c2 = x * x + y * y + z * z;
            const int row[19] = {1, 19 * c2 - 30, (21 * c2 * c2 - 53 * c2 + 24) / 2, x, (5 * c2 - 9) * x, y,
                                 (5 * c2 - 9) * y, z, (5 * c2 - 9) * z, 3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
                                 y * y - z * z, (3 * c2 - 5) * (y * y - z * z), x * y, y * z, x * z,
                                 (y * y - z * z) * x, (z * z - x * x) * y, (x * x - y * y) * z};
            return row[k];
        }
        const int px = P2(x), py = P2(y), pz = P2(z);
        if (Q == 9) {
            const int row[9] = {1, x, y, x * y, px - py, px + py, px * y, py * x, px * py};
            return row[k];
        }
        const int row[27] = {1, x, y, z, x * y, y * z, x * z, 2 * px - py - pz, py - pz, px + py + pz,
                             px * y, px * z, py * x, py * z, pz * x, pz * y, x * y * z, px * py, py * pz, px * pz,
                             px * y * z, py * x * z, pz * x * y, px * py * z, py * pz * x, px * pz * y, px * py * pz};
        return row[k];
    }

    // 0: conserved, 1: omega, 2..7: s1..s6.
    int mrt_rate(int k) const
    {
        static const int r19[19] = {0, 2, 3, 0, 4, 0, 4, 0, 4, 1, 6, 1, 6, 1, 1, 1, 7, 7, 7};
        static const int r27[27] = {0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 5, 5, 5, 6, 6, 6, 7};
        static const int r9[9] = {0, 0, 0, 1, 1, 2, 4, 4, 3};
        return Q == 19 ? r19[k] : (Q == 27 ? r27[k] : r9[k]);
    }

    void mrt_collide(const T *f, T *fs, T rho, T u, T v, T w, T udotu, T omeg) const
    {
        const double s[6] = {P.s1, P.s2, P.s3, P.s4, P.s5, P.s6};
        T d[27], r[27];
        for (int p = 0; p < Q; p++) d[p] = f[p] - feq(p, rho, u, v, w, udotu);
        for (int k = 0; k < Q; k++) {
            r[k] = 0;
            const int rc = mrt_rate(k);
            if (rc == 0) continue;
            int norm = 0;
            for (int p = 0; p < Q; p++) norm += mrt_m(k, p) * mrt_m(k, p);
            const T rate = rc == 1 ? omeg : (T)s[rc - 2];
            const T ck = (T)((double)rate / (double)norm);
            T acc = 0;
            bool first = true;
            for (int p = 0; p < Q; p++) {
                const int m = mrt_m(k, p);
                if (m == 0) continue;
                const T term = (T)m * d[p];
                acc = first ? term : acc + term;
                first = false;
            }
            r[k] = acc * ck;
        }
        for (int p = 0; p < Q; p++) {
            T acc = 0;
            bool first = true;
            for (int k = 0; k < Q; k++) {
                const int m = mrt_m(k, p);
                if (mrt_rate(k) == 0 || m == 0) continue;
                const T term = (T)m * r[k];
                acc = first ? term : acc + term;
                first = false;
            }
            fs[p] = f[p] - acc;
        }
    }

    void load_block_nbrs(int b, int *s_nbr) const
    {
        // s_ID_nbr[S] = cblock_ID_nbr[b + V_CONN_MAP[S]*n] (solver_lbm_kernels.cu:579-583).
        for (int S = 0; S < Q_max; S++) s_nbr[S] = nbr[b + (size_t)V.conn_map[S] * n_maxcblocks];
    }

    // Post-collision populations of level L before the boundary hooks: what the interpolated
    // bounce-back reads "behind" a cell.  The reference's kernel reads them from the pool while
    // other blocks of the same launch overwrite their boundary cells' slots; here every read sees
    // the value the collision wrote.
    std::vector<T> post;
    T post_at(int p, long cell) const { return post[(size_t)p * n_maxcells + cell]; }

    bool bouzidi_on() const { return links && P.S_BC_TYPE == 1; }

    // Cu_ImposeBC (solver_lbm_kernels.cu:691-887): the preset hook on domain faces and, with
    // S_BC_TYPE 1, the linear interpolated bounce-back of Bouzidi, Firdaouss and Lallemand (2001)
    // on the cut links of BOUNDARY cells: for the fraction q of the link along c_p,
    // q < 1/2: f_p <- 2q f_p + (1 - 2q) f*_p(x - c_p);  q >= 1/2: f_p <- f_p / 2q + (2q - 1)/2q f_pb,
    // applied per half-set pair in the order hook(p), link(p), hook(pb), link(pb) -- so the pb
    // link's second form already sees the new f_p.  A missing block behind the cell (the reference
    // reads out of bounds there) leaves f_p unchanged.
    void impose_bc(int L)
    {
        if (n_ids[L] <= 0) return;
        const int nL = n_ids[L];
        const int *ids = idset(L);
        const bool bz = bouzidi_on();
        if (bz) {
            post.assign((size_t)Q * n_maxcells, T(0));
            for (int bi = 0; bi < nL; bi++) {
                const int b = ids[bi];
                for (int p = 0; p < Q; p++)
                    for (int t = 0; t < M_CBLOCK; t++) {
                        const long cell = (long)b * M_CBLOCK + t;
                        post[(size_t)p * n_maxcells + cell] = Fp(p)[cell];
                    }
            }
        }
        // f*_p at x - c_p; `ok` false when that block is missing.
        auto behind = [&](const int *s_nbr, int I, int J, int K, int p, bool &ok) -> T {
            const int pbp = V.pb[p];
            const int Ip = I + V.c[0][pbp], Jp = J + V.c[1][pbp], Kp = K + V.c[2][pbp];
            const int nb = s_nbr[nbr_map(DIM, Ip, Jp, Kp)];
            ok = nb >= 0;
            return ok ? post_at(p, (long)nb * M_CBLOCK + nbr_cell_id(DIM, Ip, Jp, Kp)) : (T)0;
        };
        auto link = [&](T f, T fm, bool ok, T fo, float dQ) -> T {
            if (dQ > 0 && dQ < 0.5f) {
                if (ok) f = (T)(2.0) * (T)dQ * f + ((T)(1.0) - (T)(2.0) * (T)dQ) * fm;
            } else if (dQ >= 0.5f) {
                f = ((T)(1.0) / ((T)(2.0) * (T)dQ)) * f + (((T)(2.0) * (T)dQ - (T)(1.0)) / ((T)(2.0) * (T)dQ)) * fo;
            }
            return f;
        };
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0 || onb[b] != 1) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)b * M_CBLOCK + t;
                int I, J, K;
                cell_ijk(t, I, J, K);
                int valid_mask = cells_mask[cell];
                T rho = Fp(Q + 0)[cell], u = Fp(Q + 1)[cell], v = Fp(Q + 2)[cell];
                T w = (DIM == 3) ? Fp(Q + 3)[cell] : (T)0;
                for (int p = 1; p < Q; p++) {
                    if (!in_half_set(DIM, p)) continue;
                    int pb = V.pb[p];
                    T f_p = Fp(p)[cell];
                    T f_q = Fp(pb)[cell];
                    {
                        int nb = s_nbr[nbr_map(DIM, I + V.c[0][p], J + V.c[1][p], (DIM == 3) ? K + V.c[2][p] : 0)];
                        if (nb < 0 && nb != N_SKIPID)
                            f_p = bc_hook(nb, f_p, rho, u, v, w, (T)V.w[p], (T)V.c[0][p], (T)V.c[1][p], (T)V.c[2][p]);
                    }
                    const bool curved = bz && valid_mask == CELL_BOUNDARY && onb_solid[b] > -1;
                    const long cb = curved ? (long)onb_solid[b] * M_CBLOCK + t : 0;
                    if (curved) {
                        bool ok;
                        const T fm = behind(s_nbr, I, J, K, p, ok);
                        f_p = link(f_p, fm, ok, f_q, link_x(cb, p));
                    }
                    {
                        int nb = s_nbr[nbr_map(DIM, I + V.c[0][pb], J + V.c[1][pb], (DIM == 3) ? K + V.c[2][pb] : 0)];
                        if (nb < 0 && nb != N_SKIPID)
                            f_q = bc_hook(nb, f_q, rho, u, v, w, (T)V.w[pb], (T)V.c[0][pb], (T)V.c[1][pb], (T)V.c[2][pb]);
                    }
                    if (curved) {
                        bool ok;
                        const T fm = behind(s_nbr, I, J, K, pb, ok);
                        f_q = link(f_q, fm, ok, f_p, link_x(cb, pb));
                    }
                    if (valid_mask != CELL_SOLID) {
                        Fp(p)[cell] = f_p;
                        Fp(pb)[cell] = f_q;
                    }
                }
            }
        }
    }

    // Cu_Stream (solver_lbm_kernels.cu:532-633): in-place swap over the half set.
    void stream(int L)
    {
        if (n_ids[L] <= 0) return;
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0 || !block_streams(b)) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)b * M_CBLOCK + t;
                int valid_mask = cells_mask[cell];
                int I, J, K;
                cell_ijk(t, I, J, K);
                for (int p = 1; p < Q; p++) {
                    if (!in_half_set(DIM, p)) continue;
                    T f_p = Fp(p)[cell];
                    T dQ = (T)(-1.0);
                    int Ip = I + V.c[0][p], Jp = J + V.c[1][p], Kp = (DIM == 3) ? K + V.c[2][p] : 0;
                    int nb = s_nbr[nbr_map(DIM, Ip, Jp, Kp)];
                    int nc = nbr_cell_id(DIM, Ip, Jp, Kp);
                    T f_pb = (T)(-1.0);
                    if (nb > -1) f_pb = Fp(V.pb[p])[(long)nb * M_CBLOCK + nc];
                    if (valid_mask != CELL_SOLID && (f_pb > 0 || (plant == 8 && nb > -1)) && dQ < 0) {
                        Fp(V.pb[p])[(long)nb * M_CBLOCK + nc] = f_p;
                        Fp(p)[cell] = f_pb;
                    }
                }
            }
        }
    }

    // Solver_LBM::S_Interpolate (inc/solver_lbm.h:186-200): linear or cubic by InterpOrder.
    void interpolate(int L, int var)
    {
        if (P.interp_order == AGAL_INTERP_CUBIC)
            interpolate_cubic(L, var);
        else
            interpolate_linear(L, var);
    }

    // Averaging mode of the advance (solver_lbm_init.cu:94-97): INTERFACE after linear
    // interpolation, BLOCK after any other order.
    int average_advance() const { return P.interp_order == AGAL_INTERP_CUBIC ? AVERAGE_BLOCK : AVERAGE_INTERFACE; }

    // ---------------------------------------------------------------------------------------
    // Cu_Interpolate_Cubic (solver_lbm_kernels_comm.cu:406-638) + S_Interpolate_Cubic (:643-659).
    // alpha_idx = sum of the terms C * s_F[q] that Cu_Interp2D/3D (inc/util_interp.h) list for case
    // idx, left to right: the monomial coefficients of the tensor-product Lagrange interpolant
    // through the parent's cell centres at nodes {0, 1/3, 2/3, 1} per axis (plus, in 3D, the
    // reference's ~1e-13 remainders of a numerical inversion, which its sums include).  Children
    // are evaluated by Horner's rule at x = -1/12 + I/6 (+ 2/3 for the upper child on the axis).
    // ---------------------------------------------------------------------------------------
    T cubic_alpha(int idx, const T *s) const
    {
        const int *start = (DIM == 3) ? CUBIC3_START : CUBIC2_START;
        const unsigned char *qs = (DIM == 3) ? CUBIC3_Q : CUBIC2_Q;
        const double *cs = (DIM == 3) ? CUBIC3_C : CUBIC2_C;
        T acc = (T)cs[start[idx]] * s[qs[start[idx]]];
        for (int k = start[idx] + 1; k < start[idx + 1]; k++) acc = acc + (T)cs[k] * s[qs[k]];
        return acc;
    }

    void interpolate_cubic(int L, int var)
    {
        if (L >= MAX_LEVELS - 1) return;
        refresh(L);
        if (n_ids[L] <= 0) return;
        const T tau_ratio = plant == 1 ? (T)1 : tau_vec[L + 1] / tau_vec[L];
        const T M_ONE_TWELFTH = (T)(-0.083333333333333), ONE_SIXTH = (T)(0.166666666666667);
        const T TWO_THIRDS = (T)(0.666666666666667), ZERO_THIRDS = (T)(0.0);
        const int NCH = N_CHILDREN;
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            int i_kap_bc = nbr_child[b];
            bool b1 = (var == INTERP_INTERFACE) && (blk_mask[b] == BLK_INTERFACE);
            bool b2 = (var == INTERP_ADDED) && (ref[b] == REF_MARK_REFINE);
            if (!(b1 || b2) || i_kap_bc < 0) continue;
            T s_F[64], alpha[64];
            for (int p = 0; p < Q; p++) {
                for (int t = 0; t < M_CBLOCK; t++) {
                    long cell = (long)b * M_CBLOCK + t;
                    T rho = Fp(Q + 0)[cell], u = Fp(Q + 1)[cell], v = Fp(Q + 2)[cell];
                    T w = (DIM == 3) ? Fp(Q + 3)[cell] : (T)0;
                    T udotu = u * u + v * v + w * w;
                    T f_p = Fp(V.pb[p])[cell];
                    T fe = feq(p, rho, u, v, w, udotu);
                    s_F[t] = fe + (f_p - fe) * tau_ratio;
                }
                for (int idx = 0; idx < M_CBLOCK; idx++) alpha[idx] = cubic_alpha(idx, s_F);
                for (int t = 0; t < M_CBLOCK; t++) {
                    int I, J, K;
                    cell_ijk(t, I, J, K);
                    const T x = M_ONE_TWELFTH + ONE_SIXTH * (T)I;
                    const T y = M_ONE_TWELFTH + ONE_SIXTH * (T)J;
                    const T z = (DIM == 3) ? M_ONE_TWELFTH + ONE_SIXTH * (T)K : (T)0;
                    for (int c = 0; c < NCH; c++) {
                        const T ox = (c & 1) ? TWO_THIRDS : ZERO_THIRDS, oy = (c & 2) ? TWO_THIRDS : ZERO_THIRDS;
                        const T oz = (c & 4) ? TWO_THIRDS : ZERO_THIRDS;
                        T res = 0;
                        for (int mk = 0; mk < (DIM == 3 ? 4 : 1); mk++) {
                            T gam = 0;
                            for (int mj = 0; mj < 4; mj++) {
                                T beta = 0;
                                for (int mi = 0; mi < 4; mi++)
                                    beta = alpha[(3 - mi) + 4 * (3 - mj) + (DIM == 3 ? 16 * (3 - mk) : 0)] + (x + ox) * beta;
                                gam = beta + (y + oy) * gam;
                            }
                            if (DIM == 3)
                                res = gam + (z + oz) * res;
                            else
                                res = gam;
                        }
                        long ccell = (long)(i_kap_bc + c) * M_CBLOCK + t;
                        if ((var == INTERP_INTERFACE && cells_mask[ccell] == CELL_GHOST) || var == INTERP_ADDED)
                            Fp(V.pb[p])[ccell] = quant(res, p);
                    }
                }
            }
        }
    }

    // ---------------------------------------------------------------------------------------
    // Cu_Interpolate_Linear (solver_lbm_kernels_comm.cu:13-138) + S_Interpolate_Linear (:141-155).
    // ---------------------------------------------------------------------------------------
    void interpolate_linear(int L, int var)
    {
        if (L >= MAX_LEVELS - 1) return;  // M_Interpolate guard (src/mesh/mesh_comm.cu:9)
        refresh(L);
        if (n_ids[L] <= 0) return;
        const T tau_ratio = plant == 1 ? (T)1 : tau_vec[L + 1] / tau_vec[L];
        const T off = plant == 3 ? (T)(0.25) : (T)(-0.25);
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            int i_kap_bc = nbr_child[b];
            bool b1 = (var == INTERP_INTERFACE) && (blk_mask[b] == BLK_INTERFACE);
            bool b2 = (var == INTERP_ADDED) && (ref[b] == REF_MARK_REFINE);
            if (!(b1 || b2) || i_kap_bc < 0) continue;
            T s_F[64];
            T Ixs[64], Iys[64], Izs[64];
            T rho[64], uu[64], vv[64], ww[64];
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                Ixs[t] = off + (T)(0.5) * (T)I;
                Iys[t] = off + (T)(0.5) * (T)J;
                Izs[t] = (DIM == 3) ? off + (T)(0.5) * (T)K : (T)0;
                long cell = (long)b * M_CBLOCK + t;
                rho[t] = Fp(Q + 0)[cell];
                uu[t] = Fp(Q + 1)[cell];
                vv[t] = Fp(Q + 2)[cell];
                ww[t] = (DIM == 3) ? Fp(Q + 3)[cell] : (T)0;
            }
            for (int p = 0; p < Q; p++) {
                for (int t = 0; t < M_CBLOCK; t++) {
                    long cell = (long)b * M_CBLOCK + t;
                    T udotu = uu[t] * uu[t] + vv[t] * vv[t] + ww[t] * ww[t];
                    T f_p = Fp(V.pb[p])[cell];
                    T fe = feq(p, rho[t], uu[t], vv[t], ww[t], udotu);
                    s_F[t] = fe + (f_p - fe) * tau_ratio;
                }
                for (int t = 0; t < M_CBLOCK; t++) {
                    T Ix = Ixs[t], Iy = Iys[t], Iz = Izs[t];
                    if (DIM == 2) {
                        for (int cj = 0; cj < 2; cj++)
                            for (int ci = 0; ci < 2; ci++) {
                                int c = ci + 2 * cj;
                                int s = 2 * ci + 2 * 4 * cj;
                                long ccell = (long)(i_kap_bc + c) * M_CBLOCK + t;
                                if ((var == 0 && cells_mask[ccell] == CELL_GHOST) || (var == 1))
                                    Fp(V.pb[p])[ccell] = quant(s_F[0 + s] + (s_F[1 + s] - s_F[0 + s]) * Ix +
                                                                   (s_F[4 + s] - s_F[0 + s]) * Iy +
                                                                   (s_F[5 + s] - s_F[4 + s] - s_F[1 + s] + s_F[0 + s]) * Ix * Iy,
                                                               p);
                            }
                    } else {
                        for (int ck = 0; ck < 2; ck++)
                            for (int cj = 0; cj < 2; cj++)
                                for (int ci = 0; ci < 2; ci++) {
                                    int c = ci + 2 * cj + 4 * ck;
                                    int s = 2 * ci + 2 * 4 * cj + 2 * 4 * 4 * ck;
                                    long ccell = (long)(i_kap_bc + c) * M_CBLOCK + t;
                                    if ((var == 0 && cells_mask[ccell] == CELL_GHOST) || (var == 1))
                                        Fp(V.pb[p])[ccell] = quant(
                                            s_F[0 + s] + (s_F[1 + s] - s_F[0 + s]) * Ix + (s_F[4 + s] - s_F[0 + s]) * Iy +
                                                (s_F[16 + s] - s_F[0 + s]) * Iz +
                                                (s_F[5 + s] - s_F[1 + s] - s_F[4 + s] + s_F[0 + s]) * Ix * Iy +
                                                (s_F[17 + s] - s_F[1 + s] - s_F[16 + s] + s_F[0 + s]) * Ix * Iz +
                                                (s_F[20 + s] - s_F[4 + s] - s_F[16 + s] + s_F[0 + s]) * Iy * Iz +
                                                (s_F[21 + s] + s_F[1 + s] + s_F[4 + s] + s_F[16 + s] - s_F[5 + s] -
                                                 s_F[17 + s] - s_F[20 + s] - s_F[0 + s]) * Ix * Iy * Iz,
                                            p);
                                }
                    }
                }
            }
        }
    }

    // Cu_Average (solver_lbm_kernels_comm.cu:174-347) + S_Average (:351-370).
    void average(int L, int var)
    {
        if (L >= MAX_LEVELS - 1) return;  // M_Average guard (src/mesh/mesh_comm.cu:21)
        refresh(L + 1);
        if (n_ids[L] <= 0) return;
        const T tau_ratio = plant == 2 ? (T)1 : tau_vec[L] / tau_vec[L + 1];
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            int i_kap_bc = nbr_child[b];
            bool b1 = i_kap_bc > -1;
            bool b2 = (var == AVERAGE_GRID || blk_mask[b] == BLK_INTERFACE);
            if (!(b1 && b2)) continue;
            int nck = (DIM == 3) ? 2 : 1;
            for (int ck = 0; ck < nck; ck++)
                for (int cj = 0; cj < 2; cj++)
                    for (int ci = 0; ci < 2; ci++) {
                        int c = ci + 2 * cj + 4 * ck;
                        long cb = (long)(i_kap_bc + c) * M_CBLOCK;
                        int s_mask[64];
                        T s_Fc[64];
                        for (int t = 0; t < M_CBLOCK; t++) {
                            s_mask[t] = cells_mask[cb + t];
                            if (var > 0 && s_mask[t] < 2) s_mask[t] = 1;
                        }
                        for (int p = 0; p < Q; p++) {
                            for (int t = 0; t < M_CBLOCK; t++) {
                                T rho = Fp(Q + 0)[cb + t], u = Fp(Q + 1)[cb + t], v = Fp(Q + 2)[cb + t];
                                T w = (DIM == 3) ? Fp(Q + 3)[cb + t] : (T)0;
                                T udotu = (DIM == 3) ? u * u + v * v + w * w : u * u + v * v;
                                T f_p = Fp(V.pb[p])[cb + t];
                                T cdotu = (DIM == 3) ? (T)V.c[0][p] * u + (T)V.c[1][p] * v + (T)V.c[2][p] * w
                                                     : (T)V.c[0][p] * u + (T)V.c[1][p] * v;
                                T fe = (T)V.w[p] * rho * ((T)(1.0) + (T)(3.0) * cdotu + (T)(4.5) * cdotu * cdotu - (T)(1.5) * udotu);
                                s_Fc[t] = fe + (f_p - fe) * tau_ratio;
                            }
                            for (int t = 0; t < M_CBLOCK; t++) {
                                int I, J, K;
                                cell_ijk(t, I, J, K);
                                int c0 = 2 * (I % 2) + 4 * (2 * (J % 2));
                                if (DIM == 3) c0 += 4 * 4 * (2 * (K % 2));
                                bool inq = (I >= 2 * ci) && (I < 2 + 2 * ci) && (J >= 2 * cj) && (J < 2 + 2 * cj);
                                if (DIM == 3) inq = inq && (K >= 2 * ck) && (K < 2 + 2 * ck);
                                if (s_mask[c0] == 1 && inq) {
                                    long pcell = (long)b * M_CBLOCK + t;
                                    if (DIM == 2)
                                        Fp(V.pb[p])[pcell] = quant((T)(0.25) * (s_Fc[c0 + 0 + 4 * 0] + s_Fc[c0 + 1 + 4 * 0] +
                                                                                s_Fc[c0 + 0 + 4 * 1] + s_Fc[c0 + 1 + 4 * 1]),
                                                                   p);
                                    else if (plant == 4)
                                        Fp(V.pb[p])[pcell] = quant((T)(0.125) * (s_Fc[c0 + 0 + 4 * 0 + 16 * 0] + s_Fc[c0 + 1 + 4 * 0 + 16 * 0] +
                                                                                 s_Fc[c0 + 0 + 4 * 1 + 16 * 0] + s_Fc[c0 + 1 + 4 * 1 + 16 * 0] +
                                                                                 s_Fc[c0 + 0 + 4 * 0 + 16 * 1] + s_Fc[c0 + 1 + 4 * 0 + 16 * 1] +
                                                                                 s_Fc[c0 + 0 + 4 * 1 + 16 * 1] + s_Fc[c0 + 0 + 4 * 1 + 16 * 1]),
                                                                   p);
                                    else
                                        Fp(V.pb[p])[pcell] = quant((T)(0.125) * (s_Fc[c0 + 0 + 4 * 0 + 16 * 0] + s_Fc[c0 + 1 + 4 * 0 + 16 * 0] +
                                                                                 s_Fc[c0 + 0 + 4 * 1 + 16 * 0] + s_Fc[c0 + 1 + 4 * 1 + 16 * 0] +
                                                                                 s_Fc[c0 + 0 + 4 * 0 + 16 * 1] + s_Fc[c0 + 1 + 4 * 0 + 16 * 1] +
                                                                                 s_Fc[c0 + 0 + 4 * 1 + 16 * 1] + s_Fc[c0 + 1 + 4 * 1 + 16 * 1]),
                                                                   p);
                                }
                            }
                        }
                    }
        }
    }

    void collide_stream(int L, bool mea = false)
    {
        collide(L);
        if (mea) forces_mea(L, 0);
        impose_bc(L);
        if (store_mode) store_level(L);
        stream(L);
        if (mea) forces_mea(L, 1);
    }

    // Writes every population of level L into the 16-bit pool (values already stored read back
    // unchanged).
    void store_level(int L)
    {
        const int nL = n_ids[L];
        const int *ids = idset(L);
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < nL; bi++) {
            int b = ids[bi];
            if (b < 0) continue;
            for (int s = 0; s < Q; s++) {
                T *F = Fp(s);
                for (int t = 0; t < M_CBLOCK; t++) {
                    long cell = (long)b * M_CBLOCK + t;
                    F[cell] = quant(F[cell], V.pb[s]);
                }
            }
        }
    }

    // Solver_LBM::S_Advance (solver_lbm_advance.cu:30-174).  The momentum-exchange force (force
    // type 0) brackets the finest level's last sub-step of a coarse step, and every step on a
    // single level (:124-126, :144-150); the control-volume force (type 1) brackets the coarse
    // step (:45-49, :60-68, :139-149).
    void advance(int L)
    {
        if (MAX_LEVELS > 1 && (MAX_LEVELS != N_LEVEL_START + 1)) {
            if (L == N_LEVEL_START) {
                global_Liter = 0;
                if (cv_on())
                    for (int K = N_LEVEL_START; K < MAX_LEVELS; K++) forces_cv(K, 0);
                interpolate(L, INTERP_INTERFACE);
                collide_stream(L);
                if (plant == 7) average(L, average_advance());
                advance(L + 1);
                if (plant != 7) average(L, average_advance());
                if (cv_on()) {
                    // solver_lbm_advance.cu:60-68: a full fine-to-coarse average and the interface
                    // interpolations of every level before the second half.
                    for (int K = MAX_LEVELS - 2; K >= N_LEVEL_START; K--) average(K, AVERAGE_GRID);
                    for (int K = N_LEVEL_START; K < MAX_LEVELS - 1; K++) interpolate(K, INTERP_INTERFACE);
                    for (int K = N_LEVEL_START; K < MAX_LEVELS; K++) forces_cv(K, 1);
                }
            } else {
                if (L < MAX_LEVELS - 1) interpolate(L, INTERP_INTERFACE);
                collide_stream(L);
                if (L == MAX_LEVELS - 1) global_Liter++;
                if (L < MAX_LEVELS - 1) {
                    advance(L + 1);
                    average(L, average_advance());
                    if (plant != 5) interpolate(L, INTERP_INTERFACE);
                }
                collide_stream(L, mea_on() && L == MAX_LEVELS - 1 && global_Liter == (1 << (MAX_LEVELS - 1)) - 1);
                if (L == MAX_LEVELS - 1) global_Liter++;
                if (L < MAX_LEVELS - 1) {
                    advance(L + 1);
                    if (plant != 6) average(L, average_advance());
                }
            }
        } else {
            if (cv_on()) forces_cv(L, 0);
            collide_stream(L, mea_on());
            if (cv_on()) forces_cv(L, 1);
        }
    }

    // ---------------------------------------------------------------------------------------
    // Cu_ComputeRefCriteria (solver_lbm_criterion.cu:438-592) with ExtrapolateToHalo
    // (inc/util_math.h:34-112) and BlockwiseMaximum (inc/util.h:786-799); dx of level 0 for all
    // levels as the caller passes L = 0 (src/mesh/mesh_advance.cu:228).
    // ---------------------------------------------------------------------------------------
    static T extrap4(T d1, T d2, T d3, T d4)
    {
        return (T)(+4.0) * d1 + (T)(-6.0) * d2 + (T)(+4.0) * d3 + (T)(-1.0) * d4;
    }

    // `label` (may be null): 1 when the criterion wants block b to have children, 0 when not, -1
    // for inactive IDs (the learned predictor's training target; no reference counterpart).
    void ref_criteria(int *label = nullptr)
    {
        const int idm = id_max[MAX_LEVELS];
        const T dx_L = dxf_vec[0];
        const T NRS = (T)P.N_REFINE_START, NRI = (T)P.N_REFINE_INC, NRM = (T)P.N_REFINE_MAX;
        const int MLI = MAX_LEVELS_INTERIOR;
#pragma omp parallel for schedule(static)
        for (int b = 0; b < idm; b++) {
            if (ref[b] == REF_INACTIVE) {
                if (label) label[b] = -1;
                continue;
            }
            T su[216], sv[216], sw[216];
            auto H = [&](int I, int J, int K) { return (DIM == 2) ? (I + 1) + 6 * (J + 1) : (I + 1) + 6 * (J + 1) + 36 * (K + 1); };
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                long cell = (long)b * M_CBLOCK + t;
                su[H(I, J, K)] = Fp(Q + 1)[cell];
                sv[H(I, J, K)] = Fp(Q + 2)[cell];
                sw[H(I, J, K)] = (DIM == 3) ? Fp(Q + 3)[cell] : (T)0;
            }
            auto halo = [&](T *s) {
                for (int t = 0; t < M_CBLOCK; t++) {
                    int I, J, K;
                    cell_ijk(t, I, J, K);
                    if (I == 0) {
                        s[H(-1, J, K)] = extrap4(s[H(0, J, K)], s[H(1, J, K)], s[H(2, J, K)], s[H(3, J, K)]);
                        s[H(4, J, K)] = extrap4(s[H(3, J, K)], s[H(2, J, K)], s[H(1, J, K)], s[H(0, J, K)]);
                    }
                    if (J == 0) {
                        s[H(I, -1, K)] = extrap4(s[H(I, 0, K)], s[H(I, 1, K)], s[H(I, 2, K)], s[H(I, 3, K)]);
                        s[H(I, 4, K)] = extrap4(s[H(I, 3, K)], s[H(I, 2, K)], s[H(I, 1, K)], s[H(I, 0, K)]);
                    }
                    if (DIM == 3 && K == 0) {
                        s[H(I, J, -1)] = extrap4(s[H(I, J, 0)], s[H(I, J, 1)], s[H(I, J, 2)], s[H(I, J, 3)]);
                        s[H(I, J, 4)] = extrap4(s[H(I, J, 3)], s[H(I, J, 2)], s[H(I, J, 1)], s[H(I, J, 0)]);
                    }
                }
            };
            halo(su);
            halo(sv);
            if (DIM == 3) halo(sw);
            T Wmax = 0;
            bool first = true;
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                T uX = 0, uY = 0, uZ = 0, vX = 0, vY = 0, vZ = 0, wX = 0, wY = 0, wZ = 0;
                if (DIM == 2) {
                    uX = (T)((double)(su[H(I + 1, J, K)] - su[H(I - 1, J, K)]) / (2.0 * (double)dx_L));
                    uY = (T)((double)(su[H(I, J + 1, K)] - su[H(I, J - 1, K)]) / (2.0 * (double)dx_L));
                    vX = (T)((double)(sv[H(I + 1, J, K)] - sv[H(I - 1, J, K)]) / (2.0 * (double)dx_L));
                    vY = (T)((double)(sv[H(I, J + 1, K)] - sv[H(I, J - 1, K)]) / (2.0 * (double)dx_L));
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
                T W = 0;
                if (P.S_CRITERION == 0)
                    W = floor_log2_exact<T>(std::sqrt((wY - vZ) * (wY - vZ) + (uZ - wX) * (uZ - wX) + (vX - uY) * (vX - uY)));
                if (P.S_CRITERION == 1)
                    W = floor_log2_exact<T>(std::fabs((T)0.5 * (uX * uX + vY * vY + wZ * wZ + (T)2.0 * (uY * vX + uZ * wX + wY * vZ))));
                // S_CRITERION 2: |grad u|, Frobenius norm (no reference counterpart; the build defines it).
                if (P.S_CRITERION == 2)
                    W = floor_log2_exact<T>(std::sqrt(uX * uX + uY * uY + uZ * uZ + vX * vX + vY * vY + vZ * vZ + wX * wX +
                                                      wY * wY + wZ * wZ));
                if (cells_mask[(long)b * M_CBLOCK + t] == CELL_GHOST) W = (T)0.0;
                // fmax semantics: a NaN operand is ignored.
                if (first) { Wmax = W; first = false; }
                else Wmax = std::fmax(Wmax, W);
            }
            if (Wmax > NRM) Wmax = NRM;
            int ref_kap = ref[b];
            int level_kap = level[b];
            int L_desired = MLI - 1;
            for (int p = 1; p <= MLI - 1; p++)
                if (Wmax < NRS - NRI * (T)p) L_desired = (MLI - 1) - p;
            bool eligible = true;
            for (int p = 0; p < Q_max; p++)
                if (nbr[b + (size_t)p * n_maxcblocks] == N_SKIPID) eligible = false;
            if (eligible && level_kap != MLI - 1 && ref_kap == REF_UNREFINED && L_desired > level_kap)
                ref[b] = REF_MARK_REFINE;
            if (ref_kap == REF_REFINED && L_desired < level_kap + 1)
                ref[b] = REF_MARK_COARSEN;
            if (label) label[b] = (eligible && level_kap != MLI - 1 && L_desired > level_kap) ? 1 : 0;
        }
    }

    // ---------------------------------------------------------------------------------------
    // Learned refinement predictor: per-block features (no reference counterpart).  The block
    // statistics use the criterion's halo-extrapolated central differences with the block's own
    // spacing: max |u|, max |rho - 1|, max |grad u|, max |vorticity| over non-ghost, non-solid cells.
    // ---------------------------------------------------------------------------------------
    static float plog2(float v)
    {
        int e;
        const float m = std::frexp(v, &e);
        return (float)(e - 1) + (m - 0.5f) * 2.0f;
    }

    void block_stats(std::vector<float> &stats)
    {
        const int idm = id_max[MAX_LEVELS];
        stats.assign((size_t)4 * std::max(idm, 0), 0.f);
        const T dx0 = dxf_vec[0];
        for (int b = 0; b < idm; b++) {
            if (ref[b] == REF_INACTIVE) continue;
            T su[216], sv[216], sw[216];
            auto H = [&](int I, int J, int K) { return (DIM == 2) ? (I + 1) + 6 * (J + 1) : (I + 1) + 6 * (J + 1) + 36 * (K + 1); };
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                long cell = (long)b * M_CBLOCK + t;
                su[H(I, J, K)] = Fp(Q + 1)[cell];
                sv[H(I, J, K)] = Fp(Q + 2)[cell];
                sw[H(I, J, K)] = (DIM == 3) ? Fp(Q + 3)[cell] : (T)0;
            }
            auto halo = [&](T *s) {
                for (int t = 0; t < M_CBLOCK; t++) {
                    int I, J, K;
                    cell_ijk(t, I, J, K);
                    if (I == 0) {
                        s[H(-1, J, K)] = extrap4(s[H(0, J, K)], s[H(1, J, K)], s[H(2, J, K)], s[H(3, J, K)]);
                        s[H(4, J, K)] = extrap4(s[H(3, J, K)], s[H(2, J, K)], s[H(1, J, K)], s[H(0, J, K)]);
                    }
                    if (J == 0) {
                        s[H(I, -1, K)] = extrap4(s[H(I, 0, K)], s[H(I, 1, K)], s[H(I, 2, K)], s[H(I, 3, K)]);
                        s[H(I, 4, K)] = extrap4(s[H(I, 3, K)], s[H(I, 2, K)], s[H(I, 1, K)], s[H(I, 0, K)]);
                    }
                    if (DIM == 3 && K == 0) {
                        s[H(I, J, -1)] = extrap4(s[H(I, J, 0)], s[H(I, J, 1)], s[H(I, J, 2)], s[H(I, J, 3)]);
                        s[H(I, J, 4)] = extrap4(s[H(I, J, 3)], s[H(I, J, 2)], s[H(I, J, 1)], s[H(I, J, 0)]);
                    }
                }
            };
            halo(su);
            halo(sv);
            if (DIM == 3) halo(sw);
            const T d2 = (T)2 * (dx0 / (T)(1 << level[b]));
            float mx[4] = {0.f, 0.f, 0.f, 0.f};
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)b * M_CBLOCK + t;
                const int m = cells_mask[cell];
                if (m == CELL_GHOST || m == CELL_SOLID) continue;
                int I, J, K;
                cell_ijk(t, I, J, K);
                const T u = su[H(I, J, K)], v = sv[H(I, J, K)], w = sw[H(I, J, K)], rho = Fp(Q)[cell];
                const T uX = (su[H(I + 1, J, K)] - su[H(I - 1, J, K)]) / d2, uY = (su[H(I, J + 1, K)] - su[H(I, J - 1, K)]) / d2;
                const T vX = (sv[H(I + 1, J, K)] - sv[H(I - 1, J, K)]) / d2, vY = (sv[H(I, J + 1, K)] - sv[H(I, J - 1, K)]) / d2;
                T uZ = 0, vZ = 0, wX = 0, wY = 0, wZ = 0;
                if (DIM == 3) {
                    uZ = (su[H(I, J, K + 1)] - su[H(I, J, K - 1)]) / d2;
                    vZ = (sv[H(I, J, K + 1)] - sv[H(I, J, K - 1)]) / d2;
                    wX = (sw[H(I + 1, J, K)] - sw[H(I - 1, J, K)]) / d2;
                    wY = (sw[H(I, J + 1, K)] - sw[H(I, J - 1, K)]) / d2;
                    wZ = (sw[H(I, J, K + 1)] - sw[H(I, J, K - 1)]) / d2;
                }
                const float st[4] = {
                    (float)std::sqrt(u * u + v * v + w * w), (float)std::fabs(rho - (T)1),
                    (float)std::sqrt(uX * uX + uY * uY + uZ * uZ + vX * vX + vY * vY + vZ * vZ + wX * wX + wY * wY + wZ * wZ),
                    (float)std::sqrt((wY - vZ) * (wY - vZ) + (uZ - wX) * (uZ - wX) + (vX - uY) * (vX - uY))};
                for (int k = 0; k < 4; k++) mx[k] = std::fmax(mx[k], st[k]);
            }
            for (int k = 0; k < 4; k++) stats[4 * (size_t)b + k] = mx[k];
        }
    }

    // Signed distance feature: to the nearest triangle centroid (negative for blocks holding solid
    // cells) with a geometry, else to the nearest domain face; root-block units, clamped to
    // [-8, 8], divided by 8.
    void block_distance(std::vector<float> &dist)
    {
        const int idm = id_max[MAX_LEVELS];
        dist.assign(std::max(idm, 0), 0.f);
        const float dx0 = (float)dxf_vec[0];
        const float Lx = (float)P.L_c, Ly = (float)(P.L_fy * Lx), Lz = (float)(P.L_fz * Lx);
        const long nt = (long)geo_tri.size() / 9;
        for (int b = 0; b < idm; b++) {
            if (ref[b] == REF_INACTIVE) continue;
            const float h = 2.0f * dx0 / (float)(1 << level[b]);
            const float xc = X[b] + h, yc = X[b + n_maxcblocks] + h, zc = DIM == 3 ? X[b + 2 * n_maxcblocks] + h : 0.f;
            float d;
            if (nt > 0) {
                float best = 3.0e38f;
                for (long k = 0; k < nt; k++) {
                    const float *q = &geo_tri[9 * k];
                    const float cx = (q[0] + q[3] + q[6]) * (1.0f / 3.0f) - xc;
                    const float cy = (q[1] + q[4] + q[7]) * (1.0f / 3.0f) - yc;
                    const float cz = (q[2] + q[5] + q[8]) * (1.0f / 3.0f) - zc;
                    best = std::fmin(best, cx * cx + cy * cy + cz * cz);
                }
                d = std::sqrt(best);
                if (blk_mask[b] == BLK_SOLID || blk_mask[b] == BLK_SOLIDB) d = -d;
            } else {
                d = std::fmin(std::fmin(xc, Lx - xc), std::fmin(yc, Ly - yc));
                if (DIM == 3) d = std::fmin(d, std::fmin(zc, Lz - zc));
            }
            d = d / (4.0f * dx0);
            dist[b] = std::fmin(std::fmax(d, -8.0f), 8.0f) * 0.125f;
        }
    }

    // Feature rows (12 per block ID below id_max; zeros for inactive IDs): own statistics as
    // plog2(stat + 2^-30) / 16, level / (MAX_LEVELS - 1), distance, neighbour maxima of the four
    // statistics, fraction of existing neighbours with children, neighbour minimum of distance.
    void features(float *out)
    {
        std::vector<float> st, dist;
        block_stats(st);
        block_distance(dist);
        const int idm = id_max[MAX_LEVELS];
        const float tiny = 9.313225746154785e-10f;
        for (int b = 0; b < idm; b++) {
            float *f = out + 12 * (size_t)b;
            if (ref[b] == REF_INACTIVE) {
                for (int k = 0; k < 12; k++) f[k] = 0.f;
                continue;
            }
            float own[4], mx[4];
            for (int k = 0; k < 4; k++) mx[k] = own[k] = plog2(st[4 * (size_t)b + k] + tiny) * 0.0625f;
            float dmin = dist[b];
            int n_nb = 0, n_ref = 0;
            for (int p = 1; p < Q_max; p++) {
                const int nb = nbr[b + (size_t)p * n_maxcblocks];
                if (nb < 0) continue;
                n_nb++;
                if (nbr_child[nb] >= 0) n_ref++;
                for (int k = 0; k < 4; k++) mx[k] = std::fmax(mx[k], plog2(st[4 * (size_t)nb + k] + tiny) * 0.0625f);
                dmin = std::fmin(dmin, dist[nb]);
            }
            for (int k = 0; k < 4; k++) {
                f[k] = own[k];
                f[6 + k] = mx[k];
            }
            f[4] = (float)level[b] / (float)(MAX_LEVELS > 1 ? MAX_LEVELS - 1 : 1);
            f[5] = dist[b];
            f[10] = n_nb > 0 ? (float)n_ref / (float)n_nb : 0.f;
            f[11] = dmin;
        }
    }

    // Cu_ComputeRefCriteria_Uniform (src/mesh/mesh_criterion.cu:67-85).
    void mark_uniform(int L)
    {
        for (int k = 0; k < n_ids[L]; k++) {
            int b = idset(L)[k];
            if (ref[b] == REF_UNREFINED) ref[b] = REF_MARK_REFINE;
        }
    }

    // Cu_FreezeRefined / Cu_UnfreezeRefined (src/mesh/mesh_amr.cu:1887-1923).
    void freeze_refined(int var)
    {
        for (int kap = 0; kap < id_max[MAX_LEVELS]; kap++) {
            if (var == 0) {
                if (ref[kap] == REF_REFINED) ref[kap] = REF_REFINED_PERM;
                if (ref[kap] == REF_REFINED_WCHILD) ref[kap] = REF_REFINED_WCHILD_PERM;
            } else {
                if (ref[kap] == REF_REFINED_PERM) ref[kap] = REF_REFINED;
                if (ref[kap] == REF_REFINED_WCHILD_PERM) ref[kap] = REF_REFINED_WCHILD;
            }
        }
    }

    int N(int b, int p) const { return nbr[b + (size_t)p * n_maxcblocks]; }
    int &Nn(int b, int p) { return nbr[b + (size_t)p * n_maxcblocks]; }
    int &NC(int b, int p) { return nbr_child[b + (size_t)p * n_maxcblocks]; }

    /* Velocity index of a direction vector (components in {-1,0,1}). */
    int dir_index(int dx_, int dy_, int dz_) const
    {
        int S = (dx_ + 1) + 3 * (dy_ + 1) + (DIM == 3 ? 9 * (dz_ + 1) : 0);
        return V.conn_map[S];
    }

    // ---------------------------------------------------------------------------------------
    // Mesh::M_RefineAndCoarsenBlocks (src/mesh/mesh_amr.cu:2161-2744), kernels at :26-1923.
    // Returns AGAL_ERR_POOL_EXHAUSTED when the refinement was cancelled.
    // ---------------------------------------------------------------------------------------
    int refine_coarsen(int out_counts[4])
    {
        int status = AGAL_OK;
        const int NCH = N_CHILDREN;
        int id_max_curr = id_max[MAX_LEVELS];
        int n_ref = 0, n_coa = 0;
        for (int k = 0; k < id_max_curr; k++) {
            if (ref[k] == REF_MARK_REFINE) n_ref++;
            if (ref[k] == REF_MARK_COARSEN) n_coa++;
        }
        int id_max_curr_wnew = id_max_curr + NCH * n_ref;
        bool proceed_ref = n_ref > 0 && NCH * n_ref <= n_gaps;
        bool proceed_coa = n_coa > 0;
        if (!proceed_ref && n_ref > 0) {
            status = AGAL_ERR_POOL_EXHAUSTED;
            n_ref = 0;
            id_max_curr_wnew = id_max_curr;
            for (int k = 0; k < id_max_curr_wnew; k++)   // Cu_RefineCells_Cancel
                if (ref[k] == REF_MARK_REFINE) ref[k] = REF_UNREFINED;
        }
        int n_added = 0, n_removed = 0;
        // S1: map of marked blocks and their neighbours (c_tmp_8).
        std::vector<int> map(n_maxcblocks, -1);
        if (proceed_ref || proceed_coa) {
            for (int k = 0; k < id_max_curr; k++) {
                if (ref[k] == REF_MARK_REFINE || ref[k] == REF_MARK_COARSEN) {
                    int self = N(k, 0);
                    if (self < 0) continue;
                    map[self] = self;
                    for (int p = 1; p < Q_max; p++) {
                        int nb = N(self, p);
                        if (nb >= 0) map[nb] = nb;
                    }
                }
            }
        }
        // S2: first-child IDs from the reversed tail of the gap stack.
        std::vector<int> first_child(n_maxcblocks, -1);
        if (proceed_ref) {
            std::vector<int> list1;
            for (int k = 0; k < id_max_curr; k++)
                if (ref[k] == REF_MARK_REFINE) list1.push_back(k);
            std::vector<int> tmp3(NCH * n_ref);
            for (int j = 0; j < NCH * n_ref; j++) tmp3[j] = gap_set[n_gaps - 1 - j];
            id_max_curr_wnew = std::max(id_max_curr_wnew, *std::max_element(tmp3.begin(), tmp3.end()) + 1);
            for (int k = 0; k < n_ref; k++) first_child[list1[k]] = tmp3[NCH * k];
            // Cu_RefineCells_S2_1
            for (int kap = 0; kap < id_max_curr; kap++) {
                if (map[kap] < 0) continue;
                for (int p = 0; p < Q_max; p++) {
                    int ip = N(kap, p);
                    if (ip > -1 && ref[ip] == REF_MARK_REFINE) NC(kap, p) = first_child[ip];
                }
            }
        }
        // S3: quality control on coarsening.
        if (proceed_coa) {
            for (int kap = 0; kap < id_max_curr; kap++) {   // Cu_RefineCells_Q1_1
                if (ref[kap] != REF_UNREFINED) continue;
                bool has_refined_nbr = false;
                for (int p = 1; p < Q_max; p++)
                    if (NC(kap, p) >= 0) has_refined_nbr = true;
                if (has_refined_nbr) ref[kap] = REF_UNREFINED_VIO;
            }
            std::vector<int> ref_snapshot(ref.begin(), ref.begin() + id_max_curr);
            for (int kap = 0; kap < id_max_curr; kap++) {   // Cu_RefineCells_Q1_2
                if (ref_snapshot[kap] != REF_MARK_COARSEN) continue;
                int i_c = NC(kap, 0);
                if (i_c < 0) continue;  // a coarsening mark on a block without children is ignored
                bool viol = false;
                for (int xc = 0; xc < NCH; xc++) {
                    int r = ref[i_c + xc];
                    if (r == REF_UNREFINED_VIO || r == REF_MARK_REFINE) viol = true;
                }
                if (viol) ref[kap] = REF_REFINED;
            }
            n_coa = 0;
            for (int k = 0; k < id_max_curr; k++)
                if (ref[k] == REF_MARK_COARSEN) n_coa++;
            proceed_coa = n_coa > 0;
            if (proceed_coa) {
                for (int kap = 0; kap < id_max_curr; kap++) {   // Cu_RefineCells_S2_2
                    if (map[kap] < 0) continue;
                    for (int p = 1; p < Q_max; p++) {
                        int ip = N(kap, p);
                        if (ip > -1 && ref[ip] == REF_MARK_COARSEN) NC(kap, p) = N_SKIPID;
                    }
                }
            }
        }
        // S4: Cu_AddRemoveBlocks.
        if (proceed_ref || proceed_coa) {
            for (int kap = 0; kap < id_max_curr; kap++) {
                if (ref[kap] == REF_MARK_REFINE) {
                    int fc = first_child[kap];
                    NC(kap, 0) = fc;
                    int level_q = level[kap] + 1;
                    T xs[3] = {(T)X[kap], (T)X[kap + n_maxcblocks], DIM == 3 ? (T)X[kap + 2 * n_maxcblocks] : (T)0};
                    for (int c = 0; c < NCH; c++) {
                        int id = fc + c;
                        int ci = c % 2, cj = (c / 2) % 2, ck = (c / 4) % 2;
                        level[id] = level_q;
                        ref[id] = REF_NEW;
                        T h = dx_cblock / (T)(1 << level_q);
                        X[id] = (float)(xs[0] + (T)ci * h);
                        X[id + n_maxcblocks] = (float)(xs[1] + (T)cj * h);
                        if (DIM == 3) X[id + 2 * n_maxcblocks] = (float)(xs[2] + (T)ck * h);
                        NC(id, 0) = N_SKIPID;
                    }
                } else if (ref[kap] == REF_MARK_COARSEN) {
                    int fc = NC(kap, 0);
                    NC(kap, 0) = N_SKIPID;
                    if (fc >= 0)
                        for (int c = 0; c < NCH; c++) ref[fc + c] = REF_REMOVE;
                }
            }
        }
        // S5: remove children of coarsened blocks from the ID sets, push them on the gap stack.
        if (proceed_coa) {
            std::vector<std::pair<int, int>> rem;  // (level, id), ascending id
            for (int k = 0; k < id_max_curr_wnew; k++)
                if (ref[k] == REF_REMOVE) rem.push_back({level[k], k});
            std::stable_sort(rem.begin(), rem.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
            int n_rem_prev = 0;
            for (int L = 1; L < MAX_LEVELS; L++) {
                int n_rem_L = 0;
                for (auto &r : rem)
                    if (r.first == L) n_rem_L++;
                if (n_rem_L > 0) {
                    std::vector<char> kill(n_maxcblocks, 0);
                    for (int k = n_rem_prev; k < n_rem_prev + n_rem_L; k++) kill[rem[k].second] = 1;
                    int *s = &id_set[(size_t)L * n_maxcblocks];
                    int w = 0;
                    for (int k = 0; k < n_ids[L]; k++)
                        if (!kill[s[k]]) s[w++] = s[k];
                    n_ids[L] -= n_rem_L;
                    n_rem_prev += n_rem_L;
                }
            }
            for (int k = 0; k < n_rem_prev; k++)   // Cu_ConcatReverse
                gap_set[n_gaps - NCH * n_ref + k] = rem[n_rem_prev - 1 - k].second;
            n_gaps += n_rem_prev;
            n_removed = n_rem_prev;
            for (int kap = 0; kap < id_max_curr; kap++) {   // Cu_CoarsenCells_S2
                if (ref[kap] != REF_REMOVE) continue;
                for (int p = 0; p < Q_max; p++) {
                    if (p > 0) Nn(kap, p) = N_SKIPID;
                    NC(kap, p) = N_SKIPID;
                }
            }
        }
        // S6: insert new children in the ID sets.
        if (proceed_ref) {
            std::vector<std::pair<int, int>> nw;
            for (int k = 0; k < id_max_curr_wnew; k++)
                if (ref[k] == REF_NEW) nw.push_back({level[k], k});
            std::stable_sort(nw.begin(), nw.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
            int prev = 0;
            for (int L = 1; L < MAX_LEVELS; L++) {
                int n_new_L = 0;
                for (auto &r : nw)
                    if (r.first == L) n_new_L++;
                if (n_new_L > 0) {
                    int *s = &id_set[(size_t)L * n_maxcblocks];
                    for (int k = 0; k < n_new_L; k++) s[n_ids[L] + k] = nw[prev + k].second;
                    n_ids[L] += n_new_L;
                    std::sort(s, s + n_ids[L]);
                }
                prev += n_new_L;
            }
            n_gaps -= NCH * n_ref;
            n_added = NCH * n_ref;
        }
        // S7: ID maxima, block masks, interpolation to new blocks, ref-state bookkeeping.
        if (proceed_ref || proceed_coa) {
            for (int L = 1; L < MAX_LEVELS; L++) {
                int *s = &id_set[(size_t)L * n_maxcblocks];
                id_max[L] = n_ids[L] > 0 ? *std::max_element(s, s + n_ids[L]) + 1 : 0;
            }
            id_max[MAX_LEVELS] = *std::max_element(id_max.begin(), id_max.begin() + MAX_LEVELS);
            for (int kap = 0; kap < id_max[MAX_LEVELS]; kap++) {   // Cu_UpdateMasks_1
                if (ref[kap] == REF_INACTIVE) continue;
                int mask_ID = BLK_REGULAR;
                if (NC(kap, 0) > -1)
                    for (int p = 1; p < Q_max; p++)
                        if (NC(kap, p) == N_SKIPID) mask_ID = BLK_INTERFACE;
                if (blk_mask[kap] > -1) blk_mask[kap] = mask_ID;
            }
            for (int L = 0; L < MAX_LEVELS - 1; L++) interpolate(L, INTERP_ADDED);
            for (int kap = 0; kap < id_max_curr_wnew; kap++) {   // Cu_RefineCells_S3
                int r = ref[kap];
                if (r == REF_MARK_REFINE) ref[kap] = REF_REFINED;
                if (r == REF_NEW || r == REF_MARK_COARSEN) ref[kap] = REF_UNREFINED;
                if (r == REF_REMOVE) ref[kap] = REF_INACTIVE;
                if (r == REF_UNREFINED_VIO) ref[kap] = REF_UNREFINED;
            }
            std::vector<int> newref(id_max_curr_wnew, -1);
            for (int kap = 0; kap < id_max_curr_wnew; kap++) {   // Cu_RefineCells_S4
                int r = ref[kap];
                if (r == REF_REFINED || r == REF_REFINED_WCHILD) {
                    int i_c = NC(kap, 0);
                    if (i_c < 0) continue;
                    int nr = REF_REFINED;
                    for (int xc = 0; xc < NCH; xc++) {
                        int rc = ref[i_c + xc];
                        if (rc == REF_REFINED || rc == REF_REFINED_WCHILD) nr = REF_REFINED_WCHILD;
                    }
                    newref[kap] = nr;
                }
            }
            for (int kap = 0; kap < id_max_curr_wnew; kap++)
                if (newref[kap] > -1) ref[kap] = newref[kap];
        }
        // S8: connectivity of children, boundary flags, cell masks.
        if (proceed_ref || proceed_coa) {
            const int idm = id_max[MAX_LEVELS];
            for (int kap = 0; kap < idm; kap++) {   // Cu_UpdateConnectivity
                if (map[kap] < 0 || NC(kap, 0) < 0) continue;
                int fc = NC(kap, 0);
                for (int c = 0; c < NCH; c++) {
                    for (int p = 0; p < Q_max; p++) {
                        int src_dir, off;
                        child_neighbor_rule(V, c, p, src_dir, off);
                        int src = NC(kap, src_dir);
                        Nn(fc + c, p) = src < 0 ? src : src + off;
                    }
                }
            }
            for (int kap = 0; kap < idm; kap++) {   // Cu_UpdateBoundaries
                int on_b = 0;
                for (int p = 1; p < Q_max; p++) {
                    int nb = N(kap, p);
                    if (nb < 0) {
                        NC(kap, p) = nb;
                        if (nb != N_SKIPID) on_b = 1;
                    }
                }
                if (blk_mask[kap] == BLK_SOLIDA) on_b = 1;
                onb[kap] = on_b;
            }
            update_cell_masks(idm);
        }
        if (out_counts) {
            out_counts[0] = n_ref;
            out_counts[1] = proceed_coa ? n_coa : 0;
            out_counts[2] = n_added;
            out_counts[3] = n_removed;
        }
        return status;
    }

    // Cu_UpdateMasks_2 (src/mesh/mesh_amr.cu:981-1297) with one sub-block per block (Nqx = 1).
    void update_cell_masks(int idm)
    {
        for (int kap = 0; kap < idm; kap++) {
            if (ref[kap] == REF_INACTIVE) continue;
            int skip[27];
            for (int p = 0; p < Q_max; p++) skip[p] = (N(kap, p) == N_SKIPID);
            for (int t = 0; t < M_CBLOCK; t++) {
                long cell = (long)kap * M_CBLOCK + t;
                if (cells_mask[cell] < 0) continue;
                cells_mask[cell] = cell_mask_rule(V, skip, t);
            }
        }
    }

    // Mesh::M_Advance_RefineWithSolution (src/mesh/mesh_advance.cu:184-243).
    int refine_with_solution(int out_counts[4])
    {
        if (!(MAX_LEVELS_INTERIOR > 1)) {
            if (out_counts) out_counts[0] = out_counts[1] = out_counts[2] = out_counts[3] = 0;
            return AGAL_OK;
        }
        for (int L = MAX_LEVELS - 2; L >= N_LEVEL_START; L--) average(L, AVERAGE_GRID);
        for (int L = N_LEVEL_START; L < MAX_LEVELS - 1; L++) interpolate(L, INTERP_INTERFACE);
        ref_criteria();
        return refine_coarsen(out_counts);
    }

    // Mesh::M_Advance_RefineNearWall (src/mesh/mesh_advance.cu:53-121) without geometry: the
    // near-wall hook Cu_RefineRegion returns false (inc/custom.h:41).
    // ---------------------------------------------------------------------------------------
    // Geometry.  The solid is a closed triangle surface; a cell is solid when the ray from its
    // centre along +x crosses the surface an odd number of times.  This classification stands in
    // for the reference's nearest-face classification and x-sweep fill (Cu_Voxelize_V1,
    // mesh_voxelizer_voxelize.cu:202-352; Cu_Voxelize_Propagate_*, mesh_voxelizer_propagate.cu:
    // 93-252): both classify cell centres as inside or outside the closed surface.  Here every
    // triangle is tested (no bins), so the binned GPU voxelizer is checked against it.
    // ---------------------------------------------------------------------------------------
    std::vector<float> geo_tri;  // 9 floats per triangle

    static bool edge_owns(double ey, double ez) { return ez < 0.0 || (ez == 0.0 && ey > 0.0); }

    // Crossing of the ray (p + s e_x, s > 0) with triangle t: the point's (y, z) projection lies in
    // the triangle's (counter-clockwise) projection, points on an edge belonging to the triangle
    // when the edge points down or, if level, to +y; then the plane's x there exceeds p.x.
    static int ray_crosses(const float *t, double px, double py, double pz)
    {
        double ax = t[0], ay = t[1], az = t[2];
        double bx = t[3], by = t[4], bz = t[5];
        double cx = t[6], cy = t[7], cz = t[8];
        const double area = (by - ay) * (cz - az) - (bz - az) * (cy - ay);
        if (area == 0.0) return 0;
        if (area < 0.0) {
            std::swap(bx, cx);
            std::swap(by, cy);
            std::swap(bz, cz);
        }
        const double d1 = (by - ay) * (pz - az) - (bz - az) * (py - ay);
        const double d2 = (cy - by) * (pz - bz) - (cz - bz) * (py - by);
        const double d3 = (ay - cy) * (pz - cz) - (az - cz) * (py - cy);
        const bool in1 = d1 > 0.0 || (d1 == 0.0 && edge_owns(by - ay, bz - az));
        const bool in2 = d2 > 0.0 || (d2 == 0.0 && edge_owns(cy - by, cz - bz));
        const bool in3 = d3 > 0.0 || (d3 == 0.0 && edge_owns(ay - cy, az - cz));
        if (!(in1 && in2 && in3)) return 0;
        const double ny = (bz - az) * (cx - ax) - (bx - ax) * (cz - az);
        const double nz = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
        const double nx = (by - ay) * (cz - az) - (bz - az) * (cy - ay);
        const double xh = ax - (ny * (py - ay) + nz * (pz - az)) / nx;
        return xh > px ? 1 : 0;
    }

    void set_geometry(const float *tri, long n_tri) { geo_tri.assign(tri, tri + 9 * n_tri); }

    // Solid cells of level L, then Cu_Voxelize_UpdateMasks (mesh_voxelizer_masks.cu:269-304): a
    // block holding a solid cell gets the SOLID block mask.
    void voxelize(int L)
    {
        const long nt = (long)geo_tri.size() / 9;
        const int *ids = idset(L);
        const double dxL = (double)dxf_vec[L];
#pragma omp parallel for schedule(dynamic, 4)
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            bool any = false;
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                const double px = (double)X[b] + ((double)I + 0.5) * dxL;
                const double py = (double)X[b + n_maxcblocks] + ((double)J + 0.5) * dxL;
                const double pz = (double)X[b + 2 * n_maxcblocks] + ((double)K + 0.5) * dxL;
                int count = 0;
                for (long k = 0; k < nt; k++) count += ray_crosses(&geo_tri[9 * k], px, py, pz);
                const long cell = (long)b * M_CBLOCK + t;
                if (count & 1) cells_mask[cell] = CELL_SOLID;
                if (cells_mask[cell] == CELL_SOLID) any = true;
            }
            if (any) blk_mask[b] = BLK_SOLID;
        }
    }

    bool any_nbr_skip(int b) const
    {
        for (int p = 0; p < Q_max; p++)
            if (nbr[b + (size_t)p * n_maxcblocks] == N_SKIPID) return true;
        return false;
    }

    // The block marks of M_Geometry_Voxelize_S1 (mesh_voxelizer.cu:82-125) with the kernels
    // Cu_MarkBlocks_MarkBoundary / MarkExterior / Propagate / Finalize (mesh_voxelizer_refine.cu:
    // 23-206, 278-414).  Each kernel reads only values it does not write, so sequential loops
    // give the parallel result.
    void mark_near_wall(int L)
    {
        const int idm = id_max[L];
        if (n_ids[L] <= 0 || idm <= 0) return;
        const bool hit_max = L == MAX_LEVELS - 1;
        const float R = (float)P.G_NEAR_WALL_DISTANCE / (float)std::pow(2.0, (double)L);
        const int nprop = (int)(R / (float)(4.0 * std::sqrt(2.0) * (double)dxf_vec[L])) + 1;
        std::vector<int> tmp(idm, -1);
        auto nb_of = [&](int b, int p) { return nbr[b + (size_t)p * n_maxcblocks]; };
        for (int k = 0; k < idm; k++) {  // MarkBoundary
            if (level[k] != L || blk_mask[k] != BLK_SOLID) continue;
            bool bnd = false;
            for (int p = 0; p < Q_max; p++) {
                const int ip = nb_of(k, p);
                if (ip > -1 && blk_mask[ip] == BLK_REGULAR) bnd = true;
            }
            if (!bnd || ref[k] != REF_UNREFINED) continue;
            blk_mask[k] = BLK_SOLIDB;
            if (!hit_max && !any_nbr_skip(k)) ref[k] = REF_MARK_REFINE;
        }
        for (int k = 0; k < idm; k++) {  // MarkExterior
            if (level[k] != L) continue;
            bool near = false;
            for (int p = 0; p < Q_max; p++) {
                const int ip = nb_of(k, p);
                if (ip > -1 && blk_mask[ip] == BLK_SOLIDB) near = true;
            }
            if (!near || ref[k] != REF_UNREFINED) continue;
            if (!hit_max && !any_nbr_skip(k)) ref[k] = REF_INDET_E;
            if (blk_mask[k] == BLK_REGULAR) blk_mask[k] = BLK_SOLIDA;
        }
        for (int j = 0; j < nprop; j++) {  // Propagate: even passes ref -> tmp, odd passes tmp -> ref
            const bool even = j % 2 == 0;
            for (int k = 0; k < idm; k++) {
                if (level[k] != L) continue;
                const int m = blk_mask[k];
                if (!(m > -1 || (j == 0 && m == BLK_SOLID))) continue;
                if (even ? ref[k] != REF_UNREFINED : tmp[k] != -1) continue;
                bool mark = false;
                for (int p = 0; p < Q_max; p++) {
                    const int ip = nb_of(k, p);
                    if (ip > -1 && (even ? ref[ip] == REF_INDET_E : tmp[ip] == REF_INDET_O)) mark = true;
                }
                if (!mark || any_nbr_skip(k)) continue;
                if (even)
                    tmp[k] = REF_INDET_O;
                else
                    ref[k] = REF_INDET_E;
            }
        }
        for (int k = 0; k < idm; k++)  // Finalize
            if (ref[k] == REF_INDET_E || tmp[k] == REF_INDET_O) ref[k] = REF_MARK_REFINE;
    }

    // Cu_MarkBlocks_CheckMasks_S1/S2 (mesh_voxelizer_masks.cu:21-226): non-solid cells of SOLIDB /
    // SOLIDA leaf blocks of level L with a solid neighbour cell become BOUNDARY cells; their
    // blocks are flagged as on a boundary.  With links made to agree with the cell masks
    // (link_masks), every leaf block is examined (REGULAR ones and SOLID ones holding fluid cells
    // too).
    void check_masks(int L)
    {
        if ((long)bcount.size() != n_maxcblocks) bcount.assign(n_maxcblocks, 0);
        const int *ids = idset(L);
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            const int bm = blk_mask[b];
            if (nbr_child[b] >= 0) continue;
            if (!link_masks && !(bm < 0 && bm != BLK_SOLID)) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            bool any = false;
            for (int t = 0; t < M_CBLOCK; t++) {
                const long cell = (long)b * M_CBLOCK + t;
                if (cells_mask[cell] == CELL_SOLID) continue;
                int I, J, K;
                cell_ijk(t, I, J, K);
                bool near = false;
                for (int o = 0; o < 27; o++) {
                    if (o == 13) continue;  // the cell itself
                    const int Ip = I + o % 3 - 1, Jp = J + (o / 3) % 3 - 1, Kp = K + o / 9 - 1;
                    const int nb = s_nbr[nbr_map(DIM, Ip, Jp, Kp)];
                    if (nb >= 0 && cells_mask[(long)nb * M_CBLOCK + nbr_cell_id(DIM, Ip, Jp, Kp)] == CELL_SOLID)
                        near = true;
                }
                if (near) {
                    cells_mask[cell] = CELL_BOUNDARY;
                    any = true;
                    bcount[b]++;
                }
            }
            if (any) onb[b] = 1;
        }
    }

    // ---------------------------------------------------------------------------------------
    // Face-cell links.  Mesh::M_Geometry_Voxelize_S3 (mesh_voxelizer.cu:177-240): blocks holding
    // boundary cells get consecutive ranks in ID order (cblock_ID_onb_solid), the link arrays hold
    // (ranks rounded up to a multiple of 128) x 64 cells x 27 directions, reset to -1.
    // ---------------------------------------------------------------------------------------
    std::vector<int> bcount;      // boundary cells per block (tmp_1)
    std::vector<int> onb_solid;   // block ID -> rank, -1 for blocks without boundary cells
    long n_solidb = 0, n_maxcells_b = 0;
    std::vector<float> X_b;       // link fraction of (rank, cell, direction), -1 = no cut
    LinkBins lbins;
    bool links = false;
    long n_link_corrections = 0;
    bool link_masks = false;      // links made to agree with the cell masks (AGAL_OPT_LINK_MASKS)
    std::vector<T> blk_F;         // cblock_f_Ff: per block, the x / y / z momentum-exchange sums
    // cblock_f_Ff of the control-volume force (S_FORCE_TYPE 1): per block, rows 2d / 2d+1 hold the
    // Fp / Fm sums of direction d; the box (xm, xM, ym, yM, zm, zM) snapped to root-block edges.
    std::vector<T> cv_F;
    T cv_box[6] = {0, 0, 0, 0, 0, 0};

    void rank_boundary_blocks()
    {
        const int idm = id_max[MAX_LEVELS];
        onb_solid.assign(n_maxcblocks, -1);
        long n = 0;
        for (int k = 0; k < idm; k++)
            if (bcount[k] > 0) onb_solid[k] = (int)n++;
        n_solidb = ((n + 128) / 128) * 128;
        n_maxcells_b = n_solidb * M_CBLOCK;
    }

    float link_x(long cell_b, int q) const { return X_b[(size_t)cell_b + (size_t)q * n_maxcells_b]; }

    // Cu_LinkLengthComputation (mesh_voxelizer_links.cu:25-199), 3D: for the cells of ranked
    // blocks of level L whose bin holds faces, the smallest... in the reference's comparison: a
    // candidate distance d in (0, dx + FLT_EPSILON) along c_q replaces the stored value when d is
    // below the stored *fraction* d_old / dx or none is stored, and the fraction d / dx is stored;
    // faces come in bin order.  Only BOUNDARY cells write their 26 fractions.
    void link_lengths(int L)
    {
        const int *ids = idset(L);
        const float dx_L = (float)dxf_vec[L];
        const float eps = std::numeric_limits<float>::epsilon();
        const long nt = (long)geo_tri.size() / 9;
        (void)nt;
        const int G = lbins.G;
#pragma omp parallel for schedule(dynamic, 4)
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            const int rank = onb_solid[b];
            if (rank < 0) continue;
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                // vec3<float>(X + I*dx + 0.5*dx): the 0.5*dx term is a double (mesh_voxelizer_links.cu:70-74).
                const V3 vp{(float)((double)(X[b] + (float)I * dx_L) + 0.5 * (double)dx_L),
                            (float)((double)(X[b + n_maxcblocks] + (float)J * dx_L) + 0.5 * (double)dx_L),
                            (float)((double)(X[b + 2 * n_maxcblocks] + (float)K * dx_L) + 0.5 * (double)dx_L)};
                const int bx = (int)(vp.x * (float)G), by = (int)(vp.y * (float)G), bz = (int)(vp.z * (float)G);
                const long gbin = (long)bx + (long)G * by + (long)G * G * bz;
                if (gbin < 0 || gbin >= (long)G * G * G) continue;
                const int f0 = lbins.bin_start[gbin], n_f = lbins.bin_start[gbin + 1] - f0;
                if (n_f <= 0) continue;
                float dQ[27];
                for (int q = 0; q < 27; q++) dQ[q] = -1.0f;
                for (int k = 0; k < n_f; k++) {
                    const int f = lbins.faces[f0 + k];
                    const V3 v1 = face_vertex(geo_tri.data(), f, 0), v2 = face_vertex(geo_tri.data(), f, 1),
                             v3 = face_vertex(geo_tri.data(), f, 2);
                    const V3 n = face_normal_unit(v1, v2, v3);
                    for (int q = 1; q < 27; q++) {
                        const V3 ray{(float)V.c[0][q], (float)V.c[1][q], (float)V.c[2][q]};
                        float d = dotv(ray, n);
                        if (d != 0.0f) {
                            d = dotv(v1 - vp, n) / d;
                            if (d > 0.0f && d < dx_L + eps) {
                                const V3 vi = vp + ray * d;
                                if (point_in_triangle_ext(vi, v1, v2, v3, n)) {
                                    const float dist_q = dQ[q];
                                    if (d < dist_q || dist_q < 0) dQ[q] = d / dx_L;
                                }
                            }
                        }
                    }
                }
                const long cb = (long)rank * M_CBLOCK + t;
                if (cells_mask[(long)b * M_CBLOCK + t] == CELL_BOUNDARY)
                    for (int q = 1; q < 27; q++) X_b[(size_t)cb + (size_t)q * n_maxcells_b] = dQ[q];
            }
        }
    }

    // Cu_LinkLengthValidation (mesh_voxelizer_links.cu:200-357), 3D, over the half set: a BOUNDARY
    // cell with no cut link towards a SOLID partner gets the fraction 1/2 (either side).  The
    // partner's cell mask is read at (partner's rank) x 64 + cell, the rank standing where a block
    // ID belongs, as the reference reads it (cells_ID_mask[nbr_kap_b*M_CBLOCK + nbr_kap_c] after
    // nbr_kap_b = cblock_ID_onb_solid[nbr_kap_b], :268-279); the -1 "corrections" are never written
    // (only values > 0 are stored).  Blocks run one after another here, in ID-set order.
    void link_validation(int L)
    {
        const int *ids = idset(L);
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            const int rank = onb_solid[b];
            if (rank < 0) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            for (int t = 0; t < M_CBLOCK; t++) {
                int I, J, K;
                cell_ijk(t, I, J, K);
                for (int p = 1; p < 27; p++) {
                    if (!(p == 26 || ((p - 1) % 2 == 0 && p < 25))) continue;
                    const int Ip = I + V.c[0][p], Jp = J + V.c[1][p], Kp = K + V.c[2][p];
                    int nb = s_nbr[nbr_map(3, Ip, Jp, Kp)];
                    const int nc = nbr_cell_id(3, Ip, Jp, Kp);
                    if (nb > -1) nb = onb_solid[nb];
                    if (nb <= -1) continue;
                    const long a = (long)rank * M_CBLOCK + t, o = (long)nb * M_CBLOCK + nc;
                    const float dQ_p = X_b[(size_t)a + (size_t)p * n_maxcells_b];
                    const float dQ_pb = X_b[(size_t)o + (size_t)V.pb[p] * n_maxcells_b];
                    const int mask_p = cells_mask[(long)b * M_CBLOCK + t];
                    const int mask_pb = cells_mask[(long)nb * M_CBLOCK + nc];
                    float c_p = -1.0f, c_pb = -1.0f;
                    if (dQ_p > 0 && mask_pb == CELL_BOUNDARY && dQ_pb > 0) n_link_corrections++;
                    if (dQ_pb > 0 && mask_p == CELL_BOUNDARY && dQ_p > 0) n_link_corrections++;
                    if (mask_p == CELL_BOUNDARY && dQ_p < 0 && mask_pb == CELL_SOLID) {
                        c_p = 0.5f;
                        n_link_corrections++;
                    }
                    if (mask_pb == CELL_BOUNDARY && dQ_pb < 0 && mask_p == CELL_SOLID) {
                        c_pb = 0.5f;
                        n_link_corrections++;
                    }
                    if (c_p > 0) X_b[(size_t)a + (size_t)p * n_maxcells_b] = c_p;
                    if (c_pb > 0) X_b[(size_t)o + (size_t)V.pb[p] * n_maxcells_b] = c_pb;
                }
            }
        }
    }

    // Links made to agree with the cell masks (AGAL_OPT_LINK_MASKS; no reference counterpart): a
    // BOUNDARY cell keeps its fraction towards p only when the partner cell (same-level neighbour,
    // mask at its block ID) is SOLID; a SOLID partner without a cut face gets 1/2.
    void link_mask_fix(int L)
    {
        const int *ids = idset(L);
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            const int rank = onb_solid[b];
            if (rank < 0) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            for (int t = 0; t < M_CBLOCK; t++) {
                if (cells_mask[(long)b * M_CBLOCK + t] != CELL_BOUNDARY) continue;
                int I, J, K;
                cell_ijk(t, I, J, K);
                for (int p = 1; p < 27; p++) {
                    const int Ip = I + V.c[0][p], Jp = J + V.c[1][p], Kp = K + V.c[2][p];
                    const int nb = s_nbr[nbr_map(3, Ip, Jp, Kp)];
                    const int nc = nbr_cell_id(3, Ip, Jp, Kp);
                    const bool solid = nb > -1 && cells_mask[(long)nb * M_CBLOCK + nc] == CELL_SOLID;
                    float &x = X_b[(size_t)rank * M_CBLOCK + t + (size_t)p * n_maxcells_b];
                    x = solid ? (x > 0.0f ? x : 0.5f) : -1.0f;
                }
            }
        }
    }

    // Mesh::M_LinkLengthComputation (mesh_voxelizer.cu:262-296) on the wall levels, after the bins
    // (Geometry::Bins, geometry_bins.h:197-203).
    void compute_links()
    {
        make_link_bins(lbins, geo_tri.data(), (long)geo_tri.size() / 9, P, V.c);
        X_b.assign((size_t)n_maxcells_b * 27, -1.0f);
        n_link_corrections = 0;
        for (int L = N_LEVEL_START; L < MAX_LEVELS_WALL; L++) {
            if (n_ids[L] <= 0) continue;
            link_lengths(L);
            link_validation(L);
            if (link_masks) link_mask_fix(L);
        }
        links = true;
    }

    // Mesh::M_Advance_RefineNearWall (src/mesh/mesh_advance.cu:53-121).  The benchmark-case hook
    // Cu_RefineRegion marks nothing as shipped (inc/custom.h:17-43); with a geometry each wall
    // level is voxelized and marked before its rebuild, and the boundary cells are found once the
    // hierarchy is complete.
    int refine_near_wall()
    {
        const bool geom = !geo_tri.empty();
        int status = AGAL_OK;
        auto keep = [&](int s) {
            if (s != AGAL_OK && status == AGAL_OK) status = s;
        };
        for (int L = 0; L < N_LEVEL_START; L++) {
            mark_uniform(L);
            keep(refine_coarsen(nullptr));
        }
        for (int L = N_LEVEL_START; L < MAX_LEVELS_WALL; L++) {
            if (geom) {
                voxelize(L);
                mark_near_wall(L);
            }
            if (MAX_LEVELS > 1 && L < MAX_LEVELS_WALL - 1) keep(refine_coarsen(nullptr));
            set_ic(L);
        }
        if (geom) {
            bcount.assign(n_maxcblocks, 0);
            for (int L = N_LEVEL_START; L < MAX_LEVELS_WALL; L++) check_masks(L);
            rank_boundary_blocks();
            if (P.G_BIN_MAKE_LINKS == 1 && DIM == 3 && n_solidb > 0) compute_links();
        }
        if (MAX_LEVELS > 1) freeze_refined(0);
        return status;
    }

    // Cu_ComputeForcesMEA (solver_lbm_kernels_forces.cu:322-540) on level L: for each BOUNDARY
    // cell of a ranked block and each half-set pair (p, pb), a link with fraction > 0 adds
    // f c (order 1) or ((1/2 + q) f + (1/2 - q) f_m) c (order 2), f being the cell's slot p (pb)
    // and f_m the same slot "behind" it, x - c_p (x + c_p).  Called after the collision
    // (post_step 0: slot p holds f*_p) and after the stream (post_step 1: slot p holds what arrived,
    // the reflected value on a wall link).  The per-cell sums (ufloat_t) are reduced over the block
    // by the reference's shared-memory tree (s = 32, 16, ..., 1), times the cell volume dx^3; step
    // 0 assigns the block's entry, step 1 adds to it.
    void forces_mea(int L, int post_step)
    {
        const int *ids = idset(L);
        const T dv = dxf_vec[L] * dxf_vec[L] * dxf_vec[L];
        const int order = P.S_FORCE_ORDER;
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            const int rank = onb_solid[b];
            if (rank < 0) continue;
            int s_nbr[27];
            load_block_nbrs(b, s_nbr);
            T acc[3][64];
            for (int t = 0; t < M_CBLOCK; t++) {
                acc[0][t] = acc[1][t] = acc[2][t] = (T)0;
                const long cell = (long)b * M_CBLOCK + t;
                if (cells_mask[cell] != CELL_BOUNDARY) continue;
                int I, J, K;
                cell_ijk(t, I, J, K);
                const long cb = (long)rank * M_CBLOCK + t;
                for (int p = 1; p < Q; p++) {
                    if (!in_half_set(DIM, p)) continue;
                    const int pb = V.pb[p];
                    const T f_p = Fp(p)[cell], f_q = Fp(pb)[cell];
                    for (int side = 0; side < 2; side++) {
                        const int s_dir = side == 0 ? p : pb;    // the link's direction
                        const float dist = link_x(cb, s_dir);
                        if (!(dist > 0)) continue;
                        const int bd = V.pb[s_dir];              // "behind": x + c_pb(s)
                        const int Ip = I + V.c[0][bd], Jp = J + V.c[1][bd], Kp = K + V.c[2][bd];
                        const int nb = s_nbr[nbr_map(DIM, Ip, Jp, Kp)];
                        const T f = side == 0 ? f_p : f_q;
                        T f_m = (T)0;
                        if (nb >= 0) f_m = Fp(s_dir)[(long)nb * M_CBLOCK + nbr_cell_id(DIM, Ip, Jp, Kp)];
                        for (int d = 0; d < 3; d++) {
                            if (order == 1) {
                                acc[d][t] += f * (T)V.c[d][s_dir];
                            } else {
                                // The literal 0.5 is a double: the term and the sum are formed in
                                // double and rounded to ufloat_t by the +=.
                                const double qd = (double)(T)dist;
                                acc[d][t] = (T)((double)acc[d][t] +
                                                ((0.5 + qd) * (double)f + (0.5 - qd) * (double)f_m) * (double)V.c[d][s_dir]);
                            }
                        }
                    }
                }
            }
            for (int st = M_CBLOCK / 2; st > 0; st >>= 1)
                for (int t = 0; t < st; t++)
                    for (int d = 0; d < 3; d++) acc[d][t] = acc[d][t] + acc[d][t + st];
            for (int d = 0; d < 3; d++) {
                T &e = blk_F[(size_t)d * n_maxcblocks + b];
                e = post_step == 0 ? acc[d][0] * dv : e + acc[d][0] * dv;
            }
        }
    }

    bool mea_on() const { return links && P.S_FORCE_TYPE == 0 && DIM == 3; }

    bool cv_on() const { return P.S_FORCE_TYPE == 1; }

    bool cv_inside(T x, T y, T z) const
    {
        // CheckPointInRegion2D / 3D (inc/util_tribin_old.h:96-110): strict inequalities.
        bool in = x > cv_box[0] && x < cv_box[1] && y > cv_box[2] && y < cv_box[3];
        if (DIM == 3) in = in && z > cv_box[4] && z < cv_box[5];
        return in;
    }

    // Cu_ComputeForcesCV (solver_lbm_kernels_forces.cu:13-246) over the blocks of level L.  A cell
    // takes part when its centre x = X_b + dx (1/2 + I) lies strictly inside the box and it is not
    // solid.  Per cell, in ufloat_t: the momentum sum_p f_p c_p (p ascending) goes to Fp (post 0)
    // or Fm (post 1) for leaf blocks.  On the root level, post 0 adds to Fm, for p >= 1 whose
    // neighbour centre x + c_p dx is outside the box, f*_p c_p with f*_p = f_p (1 - om) +
    // (w_p rho (1 + 3 c.u + 4.5 (c.u) c.u - 1.5 u.u)) om, u = momentum / rho, om = dx_0 / tau_0;
    // post 1 adds f_p c_pb(p) for p >= 1 whose upwind centre x + c_pb(p) dx is outside.  The cell
    // values are summed by the shared-memory tree (s = cells/2 ... 1) and times dx^DIM; post 0
    // assigns the block's entries, post 1 adds to them.
    void forces_cv(int L, int post_step)
    {
        const int *ids = idset(L);
        const T dxL = dxf_vec[L];
        T dv = (T)1;
        for (int d = 0; d < DIM; d++) dv *= dxL;
        const T om = dxf_vec[N_LEVEL_START] / tau_vec[N_LEVEL_START], omp = (T)1 - om;
        const bool root = L == N_LEVEL_START;
#pragma omp parallel for schedule(static)
        for (int bi = 0; bi < n_ids[L]; bi++) {
            const int b = ids[bi];
            if (b < 0) continue;
            const bool leaf = nbr_child[b] < 0;
            T sFp[3][64], sFm[3][64];
            for (int t = 0; t < M_CBLOCK; t++) {
                for (int d = 0; d < 3; d++) sFp[d][t] = sFm[d][t] = (T)0;
                const long cell = (long)b * M_CBLOCK + t;
                const int I = t % 4, J = (t / 4) % 4, K = (t / 4) / 4;
                const T x = (T)X[b + 0 * n_maxcblocks] + dxL * ((T)0.5 + (T)I);
                const T y = (T)X[b + 1 * n_maxcblocks] + dxL * ((T)0.5 + (T)J);
                const T z = DIM == 3 ? (T)X[b + 2 * n_maxcblocks] + dxL * ((T)0.5 + (T)K) : (T)0;
                const bool pV = cv_inside(x, y, z) && cells_mask[cell] != CELL_SOLID;
                T f[27], rho = (T)0, mom[3] = {(T)0, (T)0, (T)0};
                for (int p = 0; p < Q; p++) {
                    f[p] = Fp(V.pb[p])[cell];  // f_p sits in slot pb(p) between steps
                    rho += f[p];
                    for (int d = 0; d < DIM; d++) mom[d] += f[p] * (T)V.c[d][p];
                }
                if (pV && leaf)
                    for (int d = 0; d < DIM; d++) (post_step == 0 ? sFp : sFm)[d][t] += mom[d];
                if (!(root && pV)) continue;
                const T u = mom[0] / rho, v = mom[1] / rho, w = mom[2] / rho;
                for (int p = 1; p < Q; p++) {
                    const int s_dir = (post_step == 0 || plant == 9) ? p : V.pb[p];
                    const T cx = (T)V.c[0][s_dir], cy = (T)V.c[1][s_dir], cz = (T)V.c[2][s_dir];
                    if (cv_inside(x + cx * dxL, y + cy * dxL, z + cz * dxL)) continue;
                    T fv = f[p];
                    if (post_step == 0 && plant != 10) {
                        const T udotu = u * u + v * v + w * w;
                        const T cdotu = cx * u + cy * v + cz * w;
                        fv = fv * omp + ((T)V.w[p] * rho * ((T)1 + (T)3 * cdotu + (T)4.5 * cdotu * cdotu - (T)1.5 * udotu)) * om;
                    }
                    for (int d = 0; d < DIM; d++) sFm[d][t] += fv * (T)V.c[d][s_dir];
                }
            }
            for (int st = M_CBLOCK / 2; st > 0; st >>= 1)
                for (int t = 0; t < st; t++)
                    for (int d = 0; d < DIM; d++) {
                        sFp[d][t] = sFp[d][t] + sFp[d][t + st];
                        sFm[d][t] = sFm[d][t] + sFm[d][t + st];
                    }
            for (int d = 0; d < DIM; d++) {
                T &ep = cv_F[(size_t)(2 * d) * n_maxcblocks + b];
                T &em = cv_F[(size_t)(2 * d + 1) * n_maxcblocks + b];
                ep = post_step == 0 ? sFp[d][0] * dv : ep + sFp[d][0] * dv;
                em = post_step == 0 ? sFm[d][0] * dv : em + sFm[d][0] * dv;
            }
        }
    }

    // S_ReportForces (solver_lbm_compute.cu:165-218): the block sums of each row over IDs below
    // id_max, summed in FP64 in ID order.  Force type 0: the three momentum-exchange rows times
    // 1/dx of the finest level; force type 1: (Fp - Fm) per direction times 1/dx of the root level.
    // The sums are then reset.
    void report_forces(double out[3])
    {
        out[0] = out[1] = out[2] = 0.0;
        const int idm = id_max[MAX_LEVELS];
        auto row_sum = [&](const std::vector<T> &A, int r) {
            double acc = 0.0;
            for (int k = 0; k < idm; k++) acc += (double)A[(size_t)r * n_maxcblocks + k];
            return acc;
        };
        if (cv_on()) {
            const double factor = 1.0 / (double)dxf_vec[N_LEVEL_START];
            for (int d = 0; d < DIM; d++) out[d] = factor * (row_sum(cv_F, 2 * d) - row_sum(cv_F, 2 * d + 1));
            std::fill(cv_F.begin(), cv_F.end(), (T)0);
            return;
        }
        const double factor = 1.0 / (double)dxf_vec[MAX_LEVELS - 1];
        for (int d = 0; d < 3; d++) out[d] = factor * row_sum(blk_F, d);
        std::fill(blk_F.begin(), blk_F.end(), (T)0);
    }

    // The body of Mesh::M_AdvanceLoop (src/mesh/mesh_advance.cu:421-463) for coarse steps [i0,i0+n).
    int advance_loop(int i0, int n)
    {
        int iter_mult = 1 << N_LEVEL_START;
        int st = AGAL_OK;
        for (int i = i0; i < i0 + n; i++) {
            for (int j = 0; j < iter_mult; j++) advance(N_LEVEL_START);
            if (MAX_LEVELS > 1 && (MAX_LEVELS != N_LEVEL_START + 1) && (i + 1) % P.P_REFINE == 0 && i > 0) {
                int s = refine_with_solution(nullptr);
                if (s != AGAL_OK) st = s;
            }
        }
        return st;
    }
};

// ---------------------------------------------------------------------------------------------
// The learned predictor's network, restated: 12-64-64-1, ReLU, sigmoid output, FP16 copies of the
// FP32 master weights in every pass, FP32 sums in index order, weighted binary cross-entropy on
// logits, Adam with bias correction.  Parameter layout: W1[12][64], b1[64], W2[64][64], b2[64],
// W3[64], b3.
// ---------------------------------------------------------------------------------------------
struct MlpRef {
    static constexpr int F = 12, H = 64;
    static constexpr int OW1 = 0, OB1 = F * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H,
                         N = OB3 + 1;
    std::vector<float> w, m, v, g;
    int t = 0;

    std::vector<float> half_copy() const
    {
        std::vector<float> h(N);
        for (int q = 0; q < N; q++) h[q] = f16_to_f32(f16_from_f32(w[q]));
        return h;
    }

    // z3 of one row; z1, z2 (64 each) when given.
    static float forward_row(const float *wh, const float *x, float *z1, float *z2)
    {
        float h1[H];
        for (int j = 0; j < H; j++) {
            float z = wh[OB1 + j];
            for (int i = 0; i < F; i++) z += x[i] * wh[OW1 + i * H + j];
            if (z1) z1[j] = z;
            h1[j] = z > 0.f ? z : 0.f;
        }
        float z3 = wh[OB3];
        for (int k = 0; k < H; k++) {
            float z = wh[OB2 + k];
            for (int j = 0; j < H; j++) z += h1[j] * wh[OW2 + j * H + k];
            if (z2) z2[k] = z;
            z3 += (z > 0.f ? z : 0.f) * wh[OW3 + k];
        }
        return z3;
    }

    // One Adam step on n rows; returns the mean loss.
    double train_step(const float *X, const float *Y, int n, float lr, float pos_weight)
    {
        const std::vector<float> wh = half_copy();
        std::vector<float> Z1((size_t)n * H), Z2((size_t)n * H), dZ1((size_t)n * H), dZ2((size_t)n * H), dZ3(n);
        double loss = 0;
        for (int r = 0; r < n; r++) {
            const float z = forward_row(wh.data(), X + (size_t)F * r, &Z1[(size_t)H * r], &Z2[(size_t)H * r]);
            const float y = Y[r], wgt = y > 0.5f ? pos_weight : 1.0f;
            const float sp_pos = std::fmax(z, 0.f) + std::log1p(std::exp(-std::fabs(z)));
            const float sp_neg = sp_pos - z;
            loss += wgt * (y * sp_neg + (1.0f - y) * sp_pos);
            dZ3[r] = wgt * (1.0f / (1.0f + std::exp(-z)) - y) / (float)n;
            for (int k = 0; k < H; k++) dZ2[(size_t)H * r + k] = Z2[(size_t)H * r + k] > 0.f ? dZ3[r] * wh[OW3 + k] : 0.f;
            for (int j = 0; j < H; j++) {
                float s = 0.f;
                for (int k = 0; k < H; k++) s += wh[OW2 + j * H + k] * dZ2[(size_t)H * r + k];
                dZ1[(size_t)H * r + j] = Z1[(size_t)H * r + j] > 0.f ? s : 0.f;
            }
        }
        g.assign(N, 0.f);
        for (int q = 0; q < N; q++) {
            float s = 0.f;
            if (q < OB1) {
                const int i = q / H, j = q % H;
                for (int r = 0; r < n; r++) s += X[(size_t)F * r + i] * dZ1[(size_t)H * r + j];
            } else if (q < OW2) {
                for (int r = 0; r < n; r++) s += dZ1[(size_t)H * r + (q - OB1)];
            } else if (q < OB2) {
                const int j = (q - OW2) / H, k = (q - OW2) % H;
                for (int r = 0; r < n; r++) {
                    const float z = Z1[(size_t)H * r + j];
                    s += (z > 0.f ? z : 0.f) * dZ2[(size_t)H * r + k];
                }
            } else if (q < OW3) {
                for (int r = 0; r < n; r++) s += dZ2[(size_t)H * r + (q - OB2)];
            } else if (q < OB3) {
                const int k = q - OW3;
                for (int r = 0; r < n; r++) {
                    const float z = Z2[(size_t)H * r + k];
                    s += (z > 0.f ? z : 0.f) * dZ3[r];
                }
            } else {
                for (int r = 0; r < n; r++) s += dZ3[r];
            }
            g[q] = s;
        }
        t++;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float bc1 = (float)(1.0 - std::pow(0.9, (double)t)), bc2 = (float)(1.0 - std::pow(0.999, (double)t));
        if (m.empty()) m.assign(N, 0.f);
        if (v.empty()) v.assign(N, 0.f);
        for (int q = 0; q < N; q++) {
            m[q] = b1 * m[q] + (1.0f - b1) * g[q];
            v[q] = b2 * v[q] + (1.0f - b2) * g[q] * g[q];
            w[q] = w[q] - lr * (m[q] / bc1) / (std::sqrt(v[q] / bc2) + eps);
        }
        return loss / n;
    }
};

}  // namespace agal_oracle
