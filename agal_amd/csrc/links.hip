// links.hip -- face-cell links of a triangle geometry on gfx950: the face bins of the link search,
// the ranks of the blocks that hold boundary cells and the cut-link fractions of those cells.
//
// Reference path (G_BIN_MAKE_LINKS 1):
//   Geometry::Bins::G_MakeBinsGPU_MD (src/geometry/geometry_bins_gpu.cu:1043-1336) with
//     Cu_ComputeVoxelRayIndicators_MD (:22-157) and Cu_ComputeBoundingBoxLimits3D (:421-561) and
//     the triangle / box test TriangleBinOverlap3D (inc/util_tribin.h:81-160);
//   Mesh::M_Geometry_Voxelize_S3 (src/mesh/mesh_voxelizer.cu:177-240): consecutive ranks, in block-ID
//     order, of the blocks holding boundary cells (cblock_ID_onb_solid);
//   Mesh::M_LinkLengthComputation (mesh_voxelizer.cu:262-296): Cu_LinkLengthComputation and
//     Cu_LinkLengthValidation (mesh_voxelizer_links.cu:25-357).
// The linked blocks' sub-step and the momentum-exchange force are in lbm_kernels.hip.
//
// The float geometry keeps the reference's operation order (this build does not contract FMAs),
// so bins and fractions are bit-identical to a serial evaluation.  The (bin, rank, face) keys of
// the bins are sorted with hipCUB's radix sort; scans use hipCUB as well.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <string>
#include <vector>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

namespace agal {

constexpr int KWG = 256;

// ---------------------------------------------------------------------------------------------
// Float vector helpers (inc/util.h:273-560): same operations, same order.
// ---------------------------------------------------------------------------------------------
struct G3 {
    float x, y, z;
};
__device__ __forceinline__ G3 gsub(G3 a, G3 b) { return G3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ G3 gadd(G3 a, G3 b) { return G3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ G3 gmul(G3 a, float s) { return G3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float gdot(G3 a, G3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ G3 gcross(G3 a, G3 b)
{
    return G3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ G3 gunit(G3 a)
{
    const float onorm = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return gmul(a, onorm);
}
__device__ __forceinline__ G3 vertex(const float *tri, long long f, int k)
{
    return G3{tri[9 * f + 3 * k], tri[9 * f + 3 * k + 1], tri[9 * f + 3 * k + 2]};
}
// CheckPointInTriangleExtended (util.h:522-545).
__device__ __forceinline__ bool in_triangle_ext(G3 vp, G3 v1, G3 v2, G3 v3, G3 n)
{
    const float eps = FLT_EPSILON;
    G3 ven = gunit(gcross(gsub(v2, v1), n));
    if (gdot(gsub(v1, vp), ven) <= -eps) return false;
    ven = gunit(gcross(gsub(v3, v2), n));
    if (gdot(gsub(v2, vp), ven) <= -eps) return false;
    ven = gunit(gcross(gsub(v1, v3), n));
    if (gdot(gsub(v3, vp), ven) <= -eps) return false;
    return true;
}
// TriangleBinOverlap2D_AxisGap (util_tribin.h:96-128): `axis` separates the projections.
__device__ __forceinline__ bool axis_gap(int axis, float v1x, float v1y, float v2x, float v2y, float v3x, float v3y,
                                         float rmx, float rmy, float rMx, float rMy)
{
    float ex = 0.0f, ey = 0.0f;
    if (axis == 0) { ex = 1.0f; ey = 0.0f; }
    if (axis == 1) { ex = 0.0f; ey = 1.0f; }
    if (axis == 2) { ex = v2y - v1y; ey = v1x - v2x; }
    if (axis == 3) { ex = v3y - v2y; ey = v2x - v3x; }
    if (axis == 4) { ex = v1y - v3y; ey = v3x - v1x; }
    float t = fmaxf(fmaxf(v1x * ex + v1y * ey, v2x * ex + v2y * ey), v3x * ex + v3y * ey);
    float r = fminf(fminf(fminf(rmx * ex + rmy * ey, rMx * ex + rmy * ey), rmx * ex + rMy * ey), rMx * ex + rMy * ey);
    const bool b = t < r;
    t = fminf(fminf(v1x * ex + v1y * ey, v2x * ex + v2y * ey), v3x * ex + v3y * ey);
    r = fmaxf(fmaxf(fmaxf(rmx * ex + rmy * ey, rMx * ex + rmy * ey), rmx * ex + rMy * ey), rMx * ex + rMy * ey);
    return b || (r < t);
}
// TrianglePlaneCutsBox + the fifteen projected axes (util_tribin.h:81-160).
__device__ __forceinline__ bool tri_box(G3 vm, G3 vM, G3 v1, G3 v2, G3 v3)
{
    const G3 n = gcross(gsub(v2, v1), gsub(v3, v1));
    const G3 c{n.x > 0 ? (vM.x - vm.x) : 0.0f, n.y > 0 ? (vM.y - vm.y) : 0.0f, n.z > 0 ? (vM.z - vm.z) : 0.0f};
    const float d = gdot(n, vm);
    const float d1 = gdot(n, gsub(c, v1));
    const float d2 = gdot(n, gsub(gsub(gsub(vM, vm), c), v1));
    if (!((d + d1) * (d + d2) <= 0)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap(a, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y, vm.x, vm.y, vM.x, vM.y)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap(a, v1.y, v1.z, v2.y, v2.z, v3.y, v3.z, vm.y, vm.z, vM.y, vM.z)) return false;
    for (int a = 0; a < 5; a++)
        if (axis_gap(a, v1.z, v1.x, v2.z, v2.x, v3.z, v3.x, vm.z, vm.x, vM.z, vM.x)) return false;
    return true;
}

// ---------------------------------------------------------------------------------------------
// Bins.  One thread per face.
// ---------------------------------------------------------------------------------------------
struct BinGeom {
    int G, n_limits;
    float Lx, Ly, Lz, dx;           // domain lengths; box growth dx_L + FLT_EPSILON
    float Lx0g, Ly0g, Lz0g;         // bin widths
    float dxL, dxLo2, NxL, NyL, NzL;  // ray-indicator lattice of the bin level
    int use_ray;
};

// Cu_ComputeVoxelRayIndicators_MD (geometry_bins_gpu.cu:22-157), 3D: a face is kept when a line
// through a cell centre of its bounding box along a half-set direction meets it.
__global__ void k_face_keep(const float *__restrict__ tri, long long n_faces, BinGeom g, int *__restrict__ keep)
{
    const long long f = (long long)blockIdx.x * KWG + threadIdx.x;
    if (f >= n_faces) return;
    if (!g.use_ray) {
        keep[f] = 1;
        return;
    }
    using V = VelSet<27>;
    const G3 v1 = vertex(tri, f, 0), v2 = vertex(tri, f, 1), v3 = vertex(tri, f, 2);
    const G3 n = gunit(gcross(gsub(v2, v1), gsub(v3, v1)));
    const int ixmin = (int)(fminf(fminf(v1.x, v2.x), v3.x) * g.NxL);
    const int iymin = (int)(fminf(fminf(v1.y, v2.y), v3.y) * g.NyL);
    const int izmin = (int)(fminf(fminf(v1.z, v2.z), v3.z) * g.NzL);
    const int ixmax = (int)(fmaxf(fmaxf(v1.x, v2.x), v3.x) * g.NxL);
    const int iymax = (int)(fmaxf(fmaxf(v1.y, v2.y), v3.y) * g.NyL);
    const int izmax = (int)(fmaxf(fmaxf(v1.z, v2.z), v3.z) * g.NzL);
    int found = 0;
    for (int p = 1; p < 27 && !found; p++) {
        if (!V::half(p)) continue;
        const G3 ray{(float)V::cx(p), (float)V::cy(p), (float)V::cz(p)};
        for (int iz = izmin; iz <= izmax && !found; iz++)
            for (int iy = iymin; iy <= iymax && !found; iy++)
                for (int ix = ixmin; ix <= ixmax && !found; ix++) {
                    const G3 vp{g.dxLo2 + g.dxL * (float)ix, g.dxLo2 + g.dxL * (float)iy, g.dxLo2 + g.dxL * (float)iz};
                    const float d = gdot(gsub(v1, vp), n) / gdot(ray, n);
                    const G3 vi = gadd(vp, gmul(ray, d));
                    if (in_triangle_ext(vi, v1, v2, v3, n)) found = 1;
                }
    }
    keep[f] = found;
}

// Cu_ComputeBoundingBoxLimits3D (geometry_bins_gpu.cu:421-561) for the kept face of compacted
// index m: its first n_limits overlapping bins (boxes grown by dx), scanning K, J, I from one bin
// below to one above the face's bounding box.  Pass 0 counts the bins inside [0, n_bins); pass 1
// writes the key (bin, rank, m) of each at off[m] + k.  A bin index outside the grid takes a rank
// and is dropped (the reference scatters it out of bounds).
template <int PASS>
__global__ void k_face_bins(const float *__restrict__ tri, const int *__restrict__ kept, long long nf, BinGeom g,
                            int *__restrict__ count, const int *__restrict__ off, unsigned long long *__restrict__ keys)
{
    const long long m = (long long)blockIdx.x * KWG + threadIdx.x;
    if (m >= nf) return;
    const int f = kept[m];
    const G3 v1 = vertex(tri, f, 0), v2 = vertex(tri, f, 1), v3 = vertex(tri, f, 2);
    const G3 vBm{fminf(fminf(v1.x, v2.x), v3.x), fminf(fminf(v1.y, v2.y), v3.y), fminf(fminf(v1.z, v2.z), v3.z)};
    const G3 vBM{fmaxf(fmaxf(v1.x, v2.x), v3.x), fmaxf(fmaxf(v1.y, v2.y), v3.y), fmaxf(fmaxf(v1.z, v2.z), v3.z)};
    const float dx = g.dx;
    int emitted = 0;
    bool C = true;
    if ((vBm.x < -dx && vBM.x < -dx) || (vBm.x > g.Lx + dx && vBM.x > g.Lx + dx)) C = false;
    if ((vBm.y < -dx && vBM.y < -dx) || (vBm.y > g.Ly + dx && vBM.y > g.Ly + dx)) C = false;
    if ((vBm.z < -dx && vBM.z < -dx) || (vBm.z > g.Lz + dx && vBM.z > g.Lz + dx)) C = false;
    if (C) {
        const float Gf = (float)g.G;
        const int xl = (int)(vBm.x * Gf) - 1, yl = (int)(vBm.y * Gf) - 1, zl = (int)(vBm.z * Gf) - 1;
        const int xL = (int)(vBM.x * Gf) + 1, yL = (int)(vBM.y * Gf) + 1, zL = (int)(vBM.z * Gf) + 1;
        const long long n_bins = (long long)g.G * g.G * g.G;
        const unsigned long long span = (unsigned long long)g.n_limits * (unsigned long long)nf;
        int counter = 0;
        for (int K = zl; K < zL + 1 && counter < g.n_limits; K++)
            for (int J = yl; J < yL + 1 && counter < g.n_limits; J++)
                for (int I = xl; I < xL + 1 && counter < g.n_limits; I++) {
                    const G3 vm{(float)I * g.Lx0g - dx, (float)J * g.Ly0g - dx, (float)K * g.Lz0g - dx};
                    const G3 vM{(float)(I + 1) * g.Lx0g + dx, (float)(J + 1) * g.Ly0g + dx, (float)(K + 1) * g.Lz0g + dx};
                    if (!tri_box(vm, vM, v1, v2, v3)) continue;
                    const long long gid = (long long)I + (long long)g.G * J + (long long)g.G * g.G * K;
                    if (gid >= 0 && gid < n_bins) {
                        if (PASS == 1)
                            keys[off[m] + emitted] = (unsigned long long)gid * span +
                                                     (unsigned long long)counter * (unsigned long long)nf + (unsigned long long)m;
                        emitted++;
                    }
                    counter++;
                }
    }
    if (PASS == 0) count[m] = emitted;
}

// Sorted keys -> face ids, and the start of every bin (first key at or above bin * span).
__global__ void k_bins_decode(const unsigned long long *__restrict__ keys, long long n_ent, const int *__restrict__ kept,
                              long long nf, int *__restrict__ faces)
{
    const long long k = (long long)blockIdx.x * KWG + threadIdx.x;
    if (k < n_ent) faces[k] = kept[keys[k] % (unsigned long long)nf];
}

__global__ void k_bins_start(const unsigned long long *__restrict__ keys, long long n_ent, long long n_bins,
                             unsigned long long span, int *__restrict__ start)
{
    const long long g = (long long)blockIdx.x * KWG + threadIdx.x;
    if (g > n_bins) return;
    const unsigned long long lo = (unsigned long long)g * span;
    long long a = 0, b = n_ent;  // first index with keys[i] >= lo
    while (a < b) {
        const long long mid = (a + b) / 2;
        if (keys[mid] < lo)
            a = mid + 1;
        else
            b = mid;
    }
    start[g] = (int)a;
}

// ---------------------------------------------------------------------------------------------
// Ranks of the blocks holding boundary cells (M_Geometry_Voxelize_S3).
// ---------------------------------------------------------------------------------------------
__global__ void k_rank_flags(const int *__restrict__ bcount, int n, int *__restrict__ flag)
{
    const int k = blockIdx.x * KWG + threadIdx.x;
    if (k < n) flag[k] = bcount[k] > 0 ? 1 : 0;
}

__global__ void k_rank_write(const int *__restrict__ bcount, const int *__restrict__ scan, long long nmax, int idm,
                             int *__restrict__ onb_solid)
{
    const long long k = (long long)blockIdx.x * KWG + threadIdx.x;
    if (k >= nmax) return;
    onb_solid[k] = (k < idm && bcount[k] > 0) ? scan[k] : -1;
}

// ---------------------------------------------------------------------------------------------
// Cu_LinkLengthComputation (mesh_voxelizer_links.cu:25-199), 3D, one wavefront per block of the
// level: for each cell of a ranked block whose bin holds faces, every face of the bin in bin order
// and every direction q of the 27: a plane distance d in (0, dx + FLT_EPSILON) along c_q whose point
// lies in the (extended) triangle replaces the stored value when d is below the stored *fraction*
// or none is stored, and d / dx is stored.  BOUNDARY cells write their 26 fractions.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(KWG) void k_link_lengths(const int *__restrict__ ids, int n, const float *__restrict__ X,
                                                      long long nmax, float dx_L, const int *__restrict__ onb_solid,
                                                      const int *__restrict__ cells_mask, const float *__restrict__ tri,
                                                      int G, const int *__restrict__ bin_start,
                                                      const int *__restrict__ bin_faces, float *__restrict__ X_b,
                                                      long long ncb)
{
    using V = VelSet<27>;
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (KWG / 64) + (threadIdx.x >> 6);
    if (bi >= n) return;  // whole wavefronts
    const int b = ids[bi];
    const int rank = onb_solid[b];
    if (rank < 0) return;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    // vec3<float>(X + I*dx + 0.5*dx): the 0.5*dx term is a double (mesh_voxelizer_links.cu:70-74).
    const G3 vp{(float)((double)(X[b] + (float)I * dx_L) + 0.5 * (double)dx_L),
                (float)((double)(X[b + nmax] + (float)J * dx_L) + 0.5 * (double)dx_L),
                (float)((double)(X[b + 2 * nmax] + (float)K * dx_L) + 0.5 * (double)dx_L)};
    const int bx = (int)(vp.x * (float)G), by = (int)(vp.y * (float)G), bz = (int)(vp.z * (float)G);
    const long long gbin = (long long)bx + (long long)G * by + (long long)G * G * bz;
    if (gbin < 0 || gbin >= (long long)G * G * G) return;
    const int f0 = bin_start[gbin], n_f = bin_start[gbin + 1] - f0;
    if (n_f <= 0) return;
    float dQ[27];
#pragma unroll
    for (int q = 0; q < 27; q++) dQ[q] = -1.0f;
    const float lim = dx_L + FLT_EPSILON;
    for (int k = 0; k < n_f; k++) {
        const int f = bin_faces[f0 + k];
        const G3 v1 = vertex(tri, f, 0), v2 = vertex(tri, f, 1), v3 = vertex(tri, f, 2);
        const G3 nrm = gunit(gcross(gsub(v2, v1), gsub(v3, v1)));
        const float num = gdot(gsub(v1, vp), nrm);
#pragma unroll
        for (int q = 1; q < 27; q++) {
            const G3 ray{(float)V::cx(q), (float)V::cy(q), (float)V::cz(q)};
            float d = gdot(ray, nrm);
            if (d != 0.0f) {
                d = num / d;
                if (d > 0.0f && d < lim) {
                    const G3 vi = gadd(vp, gmul(ray, d));
                    if (in_triangle_ext(vi, v1, v2, v3, nrm)) {
                        const float dist_q = dQ[q];
                        if (d < dist_q || dist_q < 0) dQ[q] = d / dx_L;
                    }
                }
            }
        }
    }
    if (cells_mask[(long long)b * 64 + t] != CELL_BOUNDARY) return;
    const long long cb = (long long)rank * 64 + t;
#pragma unroll
    for (int q = 1; q < 27; q++) X_b[cb + (long long)q * ncb] = dQ[q];
}

// Cu_LinkLengthValidation (mesh_voxelizer_links.cu:200-357), one wavefront per block of the level,
// over the half set of the 27 directions: a BOUNDARY cell whose link to a SOLID partner is not cut
// gets the fraction 1/2 (either side).  The partner's cell mask is read at (partner block's rank)
// x 64 + cell, the rank standing in for the block ID, as the reference reads it (:268-279); the
// "-1" corrections are counted but never stored (only values > 0 are, :346-354).  Each slot is read
// and written by one thread only: its own half-set slots by the cell, its other slots by the one
// partner that links to it.
__global__ __launch_bounds__(KWG) void k_link_validation(const int *__restrict__ ids, int n,
                                                         const int *__restrict__ nbr_slot,
                                                         const int *__restrict__ onb_solid,
                                                         const int *__restrict__ cells_mask, float *__restrict__ X_b,
                                                         long long ncb, int *__restrict__ corrections)
{
    using V = VelSet<27>;
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (KWG / 64) + (threadIdx.x >> 6);
    if (bi >= n) return;
    const int b = ids[bi];
    const int rank = onb_solid[b];
    if (rank < 0) return;
    const int my_nb = nbr_slot[(long long)b * 27 + (t < 27 ? t : 26)];
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const long long a = (long long)rank * 64 + t;
    const int mask_p = cells_mask[(long long)b * 64 + t];
    int ncorr = 0;
    for (int p = 1; p < 27; p++) {
        if (!V::half(p)) continue;
        const int Ip = I + V::cx(p), Jp = J + V::cy(p), Kp = K + V::cz(p);
        const int slot = (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) +
                         9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1));
        int nb = __shfl(my_nb, slot, 64);
        const int nc = (Ip & 3) + 4 * (Jp & 3) + 16 * (Kp & 3);
        if (nb > -1) nb = onb_solid[nb];
        if (nb <= -1) continue;
        const long long o = (long long)nb * 64 + nc;
        const int pb = V::pb(p);
        const float dQ_p = X_b[a + (long long)p * ncb];
        const float dQ_pb = X_b[o + (long long)pb * ncb];
        const int mask_pb = cells_mask[(long long)nb * 64 + nc];
        float c_p = -1.0f, c_pb = -1.0f;
        if (dQ_p > 0 && mask_pb == CELL_BOUNDARY && dQ_pb > 0) ncorr++;
        if (dQ_pb > 0 && mask_p == CELL_BOUNDARY && dQ_p > 0) ncorr++;
        if (mask_p == CELL_BOUNDARY && dQ_p < 0 && mask_pb == CELL_SOLID) {
            c_p = 0.5f;
            ncorr++;
        }
        if (mask_pb == CELL_BOUNDARY && dQ_pb < 0 && mask_p == CELL_SOLID) {
            c_pb = 0.5f;
            ncorr++;
        }
        if (c_p > 0) X_b[a + (long long)p * ncb] = c_p;
        if (c_pb > 0) X_b[o + (long long)pb * ncb] = c_pb;
    }
    if (ncorr) atomicAdd(corrections, ncorr);
}

// Links made to agree with the cell masks (AGAL_OPT_LINK_MASKS; no counterpart in the reference):
// for each BOUNDARY cell of a linked block and each of the 26 directions, the fraction stays only
// when the partner cell -- the same-level neighbour, its mask read at its block ID -- is SOLID, and a
// SOLID partner without a cut face gets 1/2.  One wavefront per block; a cell writes only its own
// slots.
__global__ __launch_bounds__(KWG) void k_link_mask_fix(const int *__restrict__ ids, int n,
                                                       const int *__restrict__ nbr_slot,
                                                       const int *__restrict__ onb_solid,
                                                       const int *__restrict__ cells_mask, float *__restrict__ X_b,
                                                       long long ncb)
{
    using V = VelSet<27>;
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (KWG / 64) + (threadIdx.x >> 6);
    if (bi >= n) return;
    const int b = ids[bi];
    const int rank = onb_solid[b];
    if (rank < 0) return;
    const int my_nb = nbr_slot[(long long)b * 27 + (t < 27 ? t : 26)];
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const long long a = (long long)rank * 64 + t;
    const bool bnd = cells_mask[(long long)b * 64 + t] == CELL_BOUNDARY;
    for (int p = 1; p < 27; p++) {
        const int Ip = I + V::cx(p), Jp = J + V::cy(p), Kp = K + V::cz(p);
        const int slot = (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) +
                         9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1));
        const int nb = __shfl(my_nb, slot, 64);
        if (!bnd) continue;
        const int nc = (Ip & 3) + 4 * (Jp & 3) + 16 * (Kp & 3);
        const bool solid = nb > -1 && cells_mask[(long long)nb * 64 + nc] == CELL_SOLID;
        const float x = X_b[a + (long long)p * ncb];
        X_b[a + (long long)p * ncb] = solid ? (x > 0.0f ? x : 0.5f) : -1.0f;
    }
}

__global__ void k_fill_float(float *a, long long n, float v)
{
    const long long i = (long long)blockIdx.x * KWG + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
static inline unsigned kblk(long long n) { return (unsigned)((n + KWG - 1) / KWG); }

void free_links(Ctx *c)
{
    for (void *p : {(void *)c->lk_onb_solid, (void *)c->lk_X, (void *)c->lk_ids, (void *)c->lk_bcount, (void *)c->lk_blkF,
                    (void *)c->lk_counter, (void *)c->lb_start, (void *)c->lb_faces})
        if (p) (void)hipFree(p);
    c->lk_onb_solid = c->lk_ids = c->lk_bcount = c->lk_counter = nullptr;
    c->lk_X = nullptr;
    c->lk_blkF = nullptr;
    c->lb_start = c->lb_faces = nullptr;
    c->links_on = false;
    c->lk_n_solidb = c->lk_ncb = 0;
}

// Exclusive sum of n ints (device to device) with hipCUB; total returned in *total.
static int exclusive_sum(Ctx *c, const int *in, int *out, long long n, long long *total)
{
    *total = 0;
    if (n <= 0) return AGAL_OK;
    size_t bytes = 0;
    AGAL_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, c->stream));
    void *tmp = nullptr;
    AGAL_HIP(c, hipMalloc(&tmp, std::max<size_t>(bytes, 16)));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, c->stream);
    int last_in = 0, last_out = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&last_in, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&last_out, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(c, e, "exclusive_sum");
    *total = (long long)last_in + last_out;
    return AGAL_OK;
}

// Boundary-cell counts per block, filled by check_masks_level; reset before the first level.
int links_reset_counts(Ctx *c)
{
    if (!c->lk_bcount) AGAL_HIP(c, hipMalloc(&c->lk_bcount, sizeof(int) * c->n_maxcblocks));
    AGAL_HIP(c, hipMemsetAsync(c->lk_bcount, 0, sizeof(int) * c->n_maxcblocks, c->stream));
    return AGAL_OK;
}

// Geometry::Bins with G_MakeBinsGPU_MD.  The bin level is min(MAX_LEVELS_WALL, G_BIN_LEVELS) - 1;
// the ray-indicator filter is used when G_BIN_LEVELS reaches MAX_LEVELS_WALL
// (geometry_bins.h:134-139, 199-203; geometry_bins_gpu.cu:1060-1065).
static int make_bins(Ctx *c)
{
    const agal_params &P = c->P;
    const long long n_faces = c->geo_n_tri;
    const int n_bin_levels = std::min(P.G_BIN_LEVELS, P.MAX_LEVELS_WALL);
    const int level = std::min(P.MAX_LEVELS_WALL, n_bin_levels) - 1;
    if (level < 0 || P.G_BIN_DENSITY <= 0 || P.G_BIN_SPEC < 0) {
        c->last_error = "link bins need G_BIN_DENSITY > 0, G_BIN_LEVELS > 0 and G_BIN_SPEC >= 0";
        return AGAL_ERR_ARG;
    }
    BinGeom g{};
    g.G = P.G_BIN_DENSITY << level;
    g.n_limits = (2 + P.G_BIN_SPEC) * (2 + P.G_BIN_SPEC) * (2 + P.G_BIN_SPEC);
    g.Lx = (float)P.L_c;
    g.Ly = (float)(P.L_fy * P.L_c);
    g.Lz = (float)(P.L_fz * P.L_c);
    const int Nx = P.Nx, Ny = (int)(Nx * (g.Ly / g.Lx)), Nz = (int)(Nx * (g.Lz / g.Lx));
    float dxg = g.Lx / (float)Nx;
    g.Lx0g = g.Lx / (float)P.G_BIN_DENSITY;
    g.Ly0g = g.Ly / (float)P.G_BIN_DENSITY;
    g.Lz0g = g.Lz / (float)P.G_BIN_DENSITY;
    int NxL = Nx, NyL = Ny, NzL = Nz;
    for (int L = 1; L <= level; L++) {
        dxg *= 0.5f;
        g.Lx0g *= 0.5f;
        g.Ly0g *= 0.5f;
        g.Lz0g *= 0.5f;
        NxL *= 2;
        NyL *= 2;
        NzL *= 2;
    }
    g.dxL = dxg;
    g.dxLo2 = 0.5f * dxg;
    g.NxL = (float)NxL;
    g.NyL = (float)NyL;
    g.NzL = (float)NzL;
    g.dx = dxg + FLT_EPSILON;
    g.use_ray = P.MAX_LEVELS_WALL == level + 1;
    const long long n_bins = (long long)g.G * g.G * g.G;
    if (n_bins > (1LL << 30)) {
        c->last_error = "too many link bins";
        return AGAL_ERR_ARG;
    }
    int *keep = nullptr, *scan = nullptr, *kept = nullptr, *count = nullptr, *off = nullptr;
    unsigned long long *keys = nullptr, *keys_sorted = nullptr;
    void *tmp = nullptr;
    int st = AGAL_OK;
    auto done = [&](int s) {
        for (void *p : {(void *)keep, (void *)scan, (void *)kept, (void *)count, (void *)off, (void *)keys,
                        (void *)keys_sorted, tmp})
            if (p) (void)hipFree(p);
        return s;
    };
    auto hip = [&](hipError_t e, const char *what) { return e == hipSuccess ? AGAL_OK : hip_fail(c, e, what); };
    if ((st = hip(hipMalloc(&keep, sizeof(int) * n_faces), "hipMalloc"))) return done(st);
    if ((st = hip(hipMalloc(&scan, sizeof(int) * n_faces), "hipMalloc"))) return done(st);
    hipLaunchKernelGGL(k_face_keep, dim3(kblk(n_faces)), dim3(KWG), 0, c->stream, c->geo_tri, n_faces, g, keep);
    if ((st = hip(hipGetLastError(), "k_face_keep"))) return done(st);
    long long nf = 0;
    if ((st = exclusive_sum(c, keep, scan, n_faces, &nf))) return done(st);
    // Compacted face list (kept[m] = face) on the host side of a small array.
    std::vector<int> hkeep(n_faces), hscan(n_faces), hkept(std::max<long long>(nf, 1));
    if ((st = hip(hipMemcpy(hkeep.data(), keep, sizeof(int) * n_faces, hipMemcpyDeviceToHost), "hipMemcpy"))) return done(st);
    for (long long f = 0, m = 0; f < n_faces; f++)
        if (hkeep[f]) hkept[m++] = (int)f;
    if ((st = hip(hipMalloc(&kept, sizeof(int) * std::max<long long>(nf, 1)), "hipMalloc"))) return done(st);
    if ((st = hip(hipMemcpy(kept, hkept.data(), sizeof(int) * std::max<long long>(nf, 1), hipMemcpyHostToDevice), "hipMemcpy")))
        return done(st);
    long long n_ent = 0;
    if (nf > 0) {
        if ((st = hip(hipMalloc(&count, sizeof(int) * nf), "hipMalloc"))) return done(st);
        if ((st = hip(hipMalloc(&off, sizeof(int) * nf), "hipMalloc"))) return done(st);
        hipLaunchKernelGGL(k_face_bins<0>, dim3(kblk(nf)), dim3(KWG), 0, c->stream, c->geo_tri, kept, nf, g, count,
                           (const int *)nullptr, (unsigned long long *)nullptr);
        if ((st = hip(hipGetLastError(), "k_face_bins<0>"))) return done(st);
        if ((st = exclusive_sum(c, count, off, nf, &n_ent))) return done(st);
    }
    const unsigned long long span = (unsigned long long)g.n_limits * (unsigned long long)std::max<long long>(nf, 1);
    if (n_ent > 0) {
        if ((st = hip(hipMalloc(&keys, sizeof(unsigned long long) * n_ent), "hipMalloc"))) return done(st);
        if ((st = hip(hipMalloc(&keys_sorted, sizeof(unsigned long long) * n_ent), "hipMalloc"))) return done(st);
        hipLaunchKernelGGL(k_face_bins<1>, dim3(kblk(nf)), dim3(KWG), 0, c->stream, c->geo_tri, kept, nf, g, count, off,
                           keys);
        if ((st = hip(hipGetLastError(), "k_face_bins<1>"))) return done(st);
        size_t bytes = 0;
        if ((st = hip(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys, keys_sorted, (int)n_ent, 0, 64, c->stream),
                      "SortKeys")))
            return done(st);
        if ((st = hip(hipMalloc(&tmp, std::max<size_t>(bytes, 16)), "hipMalloc"))) return done(st);
        if ((st = hip(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys, keys_sorted, (int)n_ent, 0, 64, c->stream),
                      "SortKeys")))
            return done(st);
    }
    if (c->lb_start) (void)hipFree(c->lb_start);
    if (c->lb_faces) (void)hipFree(c->lb_faces);
    c->lb_start = c->lb_faces = nullptr;
    if ((st = hip(hipMalloc(&c->lb_start, sizeof(int) * (n_bins + 1)), "hipMalloc"))) return done(st);
    if ((st = hip(hipMalloc(&c->lb_faces, sizeof(int) * std::max<long long>(n_ent, 1)), "hipMalloc"))) return done(st);
    hipLaunchKernelGGL(k_bins_start, dim3(kblk(n_bins + 1)), dim3(KWG), 0, c->stream, keys_sorted, n_ent, n_bins, span,
                       c->lb_start);
    if (n_ent > 0)
        hipLaunchKernelGGL(k_bins_decode, dim3(kblk(n_ent)), dim3(KWG), 0, c->stream, keys_sorted, n_ent, kept,
                           std::max<long long>(nf, 1), c->lb_faces);
    if ((st = hip(hipGetLastError(), "bins decode"))) return done(st);
    if ((st = hip(hipStreamSynchronize(c->stream), "make_bins"))) return done(st);
    c->lb_G = g.G;
    c->lb_level = level;
    c->lb_n_bins = n_bins;
    c->lb_n_ent = n_ent;
    c->lb_n_kept = nf;
    return done(AGAL_OK);
}

// The linked blocks of each level (ranked blocks of the level's ID set), concatenated per level,
// and the LINK_FLAG bit of their block flags (the fused step leaves them to the linked sub-step).
__global__ void k_link_list_flag(const int *__restrict__ ids, int n, const int *__restrict__ onb_solid,
                                 int *__restrict__ flag)
{
    const int k = blockIdx.x * KWG + threadIdx.x;
    if (k < n) flag[k] = onb_solid[ids[k]] >= 0 ? 1 : 0;
}

__global__ void k_link_list_emit(const int *__restrict__ ids, int n, const int *__restrict__ flag,
                                 const int *__restrict__ scan, int *__restrict__ out, int *__restrict__ simple)
{
    const int k = blockIdx.x * KWG + threadIdx.x;
    if (k < n && flag[k]) {
        out[scan[k]] = ids[k];
        simple[ids[k]] |= LINK_FLAG;
    }
}

int links_rebuild_lists(Ctx *c)
{
    if (!c->links_on) return AGAL_OK;
    int st;
    long long pos = 0;
    for (int L = 0; L < c->MAX_LEVELS; L++) {
        c->lk_off[L] = pos;
        c->lk_n[L] = 0;
        const int n = c->n_ids[L];
        if (n <= 0) continue;
        hipLaunchKernelGGL(k_link_list_flag, dim3(kblk(n)), dim3(KWG), 0, c->stream, c->id_set_L(L), n, c->lk_onb_solid,
                           c->scan_a);
        long long cnt = 0;
        if ((st = exclusive_sum(c, c->scan_a, c->scan_b, n, &cnt))) return st;
        hipLaunchKernelGGL(k_link_list_emit, dim3(kblk(n)), dim3(KWG), 0, c->stream, c->id_set_L(L), n, c->scan_a,
                           c->scan_b, c->lk_ids + pos, c->simple);
        AGAL_HIP(c, hipGetLastError());
        c->lk_n[L] = (int)cnt;
        pos += cnt;
    }
    return AGAL_OK;
}

// M_Geometry_Voxelize_S3 and M_LinkLengthComputation on the wall levels, after the boundary masks.
int compute_links(Ctx *c)
{
    int st;
    if (c->DIM != 3 || !c->geo_tri || !c->lk_bcount) {
        c->last_error = "links need a 3D geometry whose boundary cells have been found";
        return AGAL_ERR_STATE;
    }
    const long long nmax = c->n_maxcblocks;
    const int idm = c->id_max[c->MAX_LEVELS];
    // Ranks (thrust::copy_if + scatter, mesh_voxelizer.cu:211-219).
    if (!c->lk_onb_solid) AGAL_HIP(c, hipMalloc(&c->lk_onb_solid, sizeof(int) * nmax));
    hipLaunchKernelGGL(k_rank_flags, dim3(kblk(idm)), dim3(KWG), 0, c->stream, c->lk_bcount, idm, c->scan_a);
    long long nsolid = 0;
    if ((st = exclusive_sum(c, c->scan_a, c->scan_b, idm, &nsolid))) return st;
    hipLaunchKernelGGL(k_rank_write, dim3(kblk(nmax)), dim3(KWG), 0, c->stream, c->lk_bcount, c->scan_b, nmax, idm,
                       c->lk_onb_solid);
    AGAL_HIP(c, hipGetLastError());
    c->lk_n_solidb = ((nsolid + 128) / 128) * 128;  // mesh_voxelizer.cu:186
    c->lk_ncb = c->lk_n_solidb * 64;
    if (nsolid == 0) return AGAL_OK;
    if ((st = make_bins(c))) return st;
    if (c->lk_X) (void)hipFree(c->lk_X);
    c->lk_X = nullptr;
    AGAL_HIP(c, hipMalloc(&c->lk_X, sizeof(float) * 27 * c->lk_ncb));
    hipLaunchKernelGGL(k_fill_float, dim3(kblk(27 * c->lk_ncb)), dim3(KWG), 0, c->stream, c->lk_X, 27 * c->lk_ncb, -1.0f);
    if (!c->lk_counter) AGAL_HIP(c, hipMalloc(&c->lk_counter, sizeof(int)));
    AGAL_HIP(c, hipMemsetAsync(c->lk_counter, 0, sizeof(int), c->stream));
    for (int L = c->N_LEVEL_START; L < c->MAX_LEVELS_WALL; L++) {
        const int n = c->n_ids[L];
        if (n <= 0) continue;
        hipLaunchKernelGGL(k_link_lengths, dim3((unsigned)((n + 3) / 4)), dim3(KWG), 0, c->stream, c->id_set_L(L), n,
                           c->X, nmax, (float)c->dxf_vec[L], c->lk_onb_solid, c->cells_mask, c->geo_tri, c->lb_G,
                           c->lb_start, c->lb_faces, c->lk_X, c->lk_ncb);
        hipLaunchKernelGGL(k_link_validation, dim3((unsigned)((n + 3) / 4)), dim3(KWG), 0, c->stream, c->id_set_L(L), n,
                           c->nbr_slot, c->lk_onb_solid, c->cells_mask, c->lk_X, c->lk_ncb, c->lk_counter);
        if (c->link_masks)
            hipLaunchKernelGGL(k_link_mask_fix, dim3((unsigned)((n + 3) / 4)), dim3(KWG), 0, c->stream, c->id_set_L(L), n,
                               c->nbr_slot, c->lk_onb_solid, c->cells_mask, c->lk_X, c->lk_ncb);
        AGAL_HIP(c, hipGetLastError());
    }
    if (!c->lk_ids) AGAL_HIP(c, hipMalloc(&c->lk_ids, sizeof(int) * nmax));
    if (!c->lk_blkF) {
        AGAL_HIP(c, hipMalloc(&c->lk_blkF, c->csize * 3 * nmax));
        AGAL_HIP(c, hipMemsetAsync(c->lk_blkF, 0, c->csize * 3 * nmax, c->stream));
    }
    int hc = 0;
    AGAL_HIP(c, hipMemcpyAsync(&hc, c->lk_counter, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    c->lk_corrections = hc;
    c->links_on = true;
    c->mesh_epoch++;  // recorded step graphs do not hold the linked sub-steps
    return links_rebuild_lists(c);
}

// S_ReportForces (solver_lbm_compute.cu:165-218): the per-block sums of each row over IDs below
// id_max in FP64; the rows read are then reset.  One workgroup, lanes striding over the IDs, then a
// tree: a fixed order, not the serial one.
template <typename T, int R>
__global__ void k_report_forces(T *blkF, long long nmax, int idm, double *out)
{
    __shared__ double sh[R][KWG];
    double a[R];
    for (int d = 0; d < R; d++) a[d] = 0.0;
    for (int k = threadIdx.x; k < idm; k += KWG)
        for (int d = 0; d < R; d++) {
            a[d] += (double)blkF[(long long)d * nmax + k];
            blkF[(long long)d * nmax + k] = (T)0;
        }
    for (int d = 0; d < R; d++) sh[d][threadIdx.x] = a[d];
    __syncthreads();
    for (int w = KWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < R; d++) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < R; d++) out[d] = sh[d][0];
}

// Force type 0: the momentum-exchange rows times 1/dx of the finest level.  Force type 1: the
// control-volume rows, (sum Fp - sum Fm) per direction times 1/dx of the root level; the whole
// array is cleared afterwards (Cu_ResetToValue over 6 x n_maxcblocks).
int compute_forces(Ctx *c, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    const int type = c->P.S_FORCE_TYPE;
    if (type == 0 && c->DIM != 3) return AGAL_ERR_NOT_BUILT;
    if (type == 0 && !c->links_on) {
        c->last_error = "the momentum-exchange force needs face-cell links (G_BIN_MAKE_LINKS 1)";
        return AGAL_ERR_STATE;
    }
    if (type != 0 && !(type == 1 && c->cv_F)) {
        c->last_error = "S_FORCE_TYPE must be 0 (momentum exchange) or 1 (control volume)";
        return AGAL_ERR_STATE;
    }
    double *d = nullptr;
    AGAL_HIP(c, hipMalloc(&d, sizeof(double) * 6));
    const int idm = c->id_max[c->MAX_LEVELS];
    const bool f64 = c->precision == AGAL_FP64;
    if (type == 0) {
        if (f64)
            hipLaunchKernelGGL((k_report_forces<double, 3>), dim3(1), dim3(KWG), 0, c->stream, (double *)c->lk_blkF,
                               c->n_maxcblocks, idm, d);
        else
            hipLaunchKernelGGL((k_report_forces<float, 3>), dim3(1), dim3(KWG), 0, c->stream, (float *)c->lk_blkF,
                               c->n_maxcblocks, idm, d);
    } else {
        if (f64)
            hipLaunchKernelGGL((k_report_forces<double, 6>), dim3(1), dim3(KWG), 0, c->stream, (double *)c->cv_F,
                               c->n_maxcblocks, idm, d);
        else
            hipLaunchKernelGGL((k_report_forces<float, 6>), dim3(1), dim3(KWG), 0, c->stream, (float *)c->cv_F,
                               c->n_maxcblocks, idm, d);
    }
    double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && type == 1)
        e = hipMemsetAsync(c->cv_F, 0, c->csize * 6 * (size_t)c->n_maxcblocks, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "compute_forces");
    if (type == 0) {
        const double factor = 1.0 / c->dxf_vec[c->MAX_LEVELS - 1];
        for (int k = 0; k < 3; k++) out[k] = factor * h[k];
    } else {
        const double factor = 1.0 / c->dxf_vec[c->N_LEVEL_START];
        for (int k = 0; k < c->DIM; k++) out[k] = factor * (h[2 * k] - h[2 * k + 1]);
    }
    return AGAL_OK;
}

}  // namespace agal
