// geometry.hip -- triangle geometry, voxelization, near-wall refinement marks and the
// momentum-exchange force, for gfx950.
//
// Reference path: Mesh::M_Advance_RefineNearWall (src/mesh/mesh_advance.cu:53-121) with
// M_Geometry_Voxelize_S1 (src/mesh/mesh_voxelizer.cu:33-136), the marking kernels
// Cu_MarkBlocks_* (src/mesh/mesh_voxelizer_refine.cu:23-414) and Cu_Voxelize_UpdateMasks
// (mesh_voxelizer_masks.cu:269-304); forces Cu_ComputeForcesMEA (solver_lbm_kernels_forces.cu:
// 322-540) reported by S_ReportForces (solver_lbm_compute.cu:165-218).
//
// The voxelizer is this build's own: a cell is solid when the ray from its centre along +x crosses
// the closed surface an odd number of times.  Triangles are binned by their (y, z) bounding boxes
// so each cell tests only the triangles of its own bin; every crossing test runs in FP64 with a
// tie rule on shared edges (each point on an edge belongs to exactly one of its triangles), so the
// count is exact for closed meshes.  Block masks, the boundary / exterior marks and their
// propagation follow the reference's kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

namespace agal {

constexpr int GWG = 256;

// ---------------------------------------------------------------------------------------------
// Ray-parity crossing test, shared by the kernel and the host (no FMA contraction: built with
// -ffp-contract=off).  Returns 1 when the ray (p + t e_x, t > 0) crosses triangle (a, b, c).
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline bool edge_owns(double ey, double ez)
{
    // For a counter-clockwise (y, z) triangle, a point on the edge with direction (ey, ez) belongs
    // to the triangle when the edge is a "top" or "left" edge.
    return ez < 0.0 || (ez == 0.0 && ey > 0.0);
}

__host__ __device__ inline int ray_crosses(const float *t, double px, double py, double pz)
{
    double ay = t[1], az = t[2], by = t[4], bz = t[5], cy = t[7], cz = t[8];
    double ax = t[0], bx = t[3], cx = t[6];
    const double area = (by - ay) * (cz - az) - (bz - az) * (cy - ay);
    if (area == 0.0) return 0;
    if (area < 0.0) {  // make (a, b, c) counter-clockwise in (y, z)
        double s;
        s = by; by = cy; cy = s;
        s = bz; bz = cz; cz = s;
        s = bx; bx = cx; cx = s;
    }
    const double d1 = (by - ay) * (pz - az) - (bz - az) * (py - ay);
    const double d2 = (cy - by) * (pz - bz) - (cz - bz) * (py - by);
    const double d3 = (ay - cy) * (pz - cz) - (az - cz) * (py - cy);
    const bool in1 = d1 > 0.0 || (d1 == 0.0 && edge_owns(by - ay, bz - az));
    const bool in2 = d2 > 0.0 || (d2 == 0.0 && edge_owns(cy - by, cz - bz));
    const bool in3 = d3 > 0.0 || (d3 == 0.0 && edge_owns(ay - cy, az - cz));
    if (!(in1 && in2 && in3)) return 0;
    // Plane x where it meets the ray: n . (X - a) = 0.
    const double ny = (bz - az) * (cx - ax) - (bx - ax) * (cz - az);
    const double nz = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    const double nx = (by - ay) * (cz - az) - (bz - az) * (cy - ay);
    const double xh = ax - (ny * (py - ay) + nz * (pz - az)) / nx;
    return xh > px ? 1 : 0;
}

// One thread per cell of the level's blocks (one wavefront per 3D block).
__global__ __launch_bounds__(GWG) void k_voxelize(const int *__restrict__ ids, int n, const float *__restrict__ X,
                                                  long long nmax, double dx_L, int *__restrict__ cells_mask,
                                                  int *__restrict__ blk_mask, const float *__restrict__ tri,
                                                  const int *__restrict__ bin_start, const int *__restrict__ bin_tri,
                                                  int G, double y0, double z0, double inv_dy, double inv_dz)
{
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (GWG / 64) + (threadIdx.x >> 6);
    if (bi >= n) return;  // whole wavefronts
    const int b = ids[bi];
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    const double px = (double)X[b] + ((double)I + 0.5) * dx_L;
    const double py = (double)X[b + nmax] + ((double)J + 0.5) * dx_L;
    const double pz = (double)X[b + 2 * nmax] + ((double)K + 0.5) * dx_L;
    int count = 0;
    const double fy = (py - y0) * inv_dy, fz = (pz - z0) * inv_dz;
    if (fy >= 0.0 && fz >= 0.0 && fy < (double)G && fz < (double)G) {
        const int bin = (int)fy + G * (int)fz;
        for (int k = bin_start[bin]; k < bin_start[bin + 1]; k++) count += ray_crosses(tri + 9LL * bin_tri[k], px, py, pz);
    }
    const long long cell = (long long)b * 64 + t;
    const bool solid = (count & 1) != 0;
    if (solid) cells_mask[cell] = CELL_SOLID;
    const bool any = __any(solid || cells_mask[cell] == CELL_SOLID);
    if (t == 0 && any) blk_mask[b] = BLK_SOLID;  // Cu_Voxelize_UpdateMasks (mesh_voxelizer_masks.cu:269-304)
}

__device__ inline bool nbr_has_skip(const int *nbr, long long nmax, int b)
{
    for (int p = 0; p < 27; p++)
        if (nbr[b + p * nmax] == N_SKIPID) return true;
    return false;
}

// Cu_MarkBlocks_MarkBoundary (mesh_voxelizer_refine.cu:23-106): an unrefined SOLID block of level L
// next to a REGULAR block lies on the solid's boundary and becomes SOLIDB; below the finest level,
// and when none of its neighbours is missing, it is marked for refinement.  Only SOLID -> SOLIDB is
// written while REGULAR is read, so the result does not depend on thread order.
__global__ void k_mark_boundary(int idm, long long nmax, int L, bool hit_max, int *ref, int *blk_mask, const int *nbr,
                                const int *level)
{
    const int kap = blockIdx.x * GWG + threadIdx.x;
    if (kap >= idm || level[kap] != L || blk_mask[kap] != BLK_SOLID) return;
    bool boundary = false;
    for (int p = 0; p < 27; p++) {
        const int ip = nbr[kap + p * nmax];
        if (ip > -1 && blk_mask[ip] == BLK_REGULAR) boundary = true;
    }
    if (!boundary || ref[kap] != REF_UNREFINED) return;
    blk_mask[kap] = BLK_SOLIDB;
    if (!hit_max && !nbr_has_skip(nbr, nmax, kap)) ref[kap] = REF_MARK_REFINE;
}

// Cu_MarkBlocks_MarkExterior (:108-206): unrefined level-L blocks next to a SOLIDB block get the
// even intermediate mark (when eligible, below the finest level); REGULAR ones become SOLIDA.
__global__ void k_mark_exterior(int idm, long long nmax, int L, bool hit_max, int *ref, int *blk_mask, const int *nbr,
                                const int *level)
{
    const int kap = blockIdx.x * GWG + threadIdx.x;
    if (kap >= idm || level[kap] != L) return;
    bool near = false;
    for (int p = 0; p < 27; p++) {
        const int ip = nbr[kap + p * nmax];
        if (ip > -1 && blk_mask[ip] == BLK_SOLIDB) near = true;
    }
    if (!near || ref[kap] != REF_UNREFINED) return;
    if (!hit_max && !nbr_has_skip(nbr, nmax, kap)) ref[kap] = REF_INDET_E;
    if (blk_mask[kap] == BLK_REGULAR) blk_mask[kap] = BLK_SOLIDA;
}

// Cu_MarkBlocks_Propagate (:208-394): one layer of the near-wall band.  Even passes read the even
// marks in `ref` and write odd marks into `tmp` (resting value -1); odd passes read `tmp` and
// write even marks into `ref`.  Candidates are level-L blocks with a non-negative block mask (on
// the first pass also SOLID blocks) and no missing neighbour.
__global__ void k_mark_propagate(int idm, long long nmax, int L, int jprop, int *ref, const int *blk_mask,
                                 const int *nbr, const int *level, int *tmp)
{
    const int kap = blockIdx.x * GWG + threadIdx.x;
    if (kap >= idm || level[kap] != L) return;
    const int m = blk_mask[kap];
    if (!(m > -1 || (jprop == 0 && m == BLK_SOLID))) return;
    const bool even = (jprop % 2) == 0;
    if (even ? ref[kap] != REF_UNREFINED : tmp[kap] != -1) return;
    bool mark = false;
    for (int p = 0; p < 27; p++) {
        const int ip = nbr[kap + p * nmax];
        if (ip > -1 && (even ? ref[ip] == REF_INDET_E : tmp[ip] == REF_INDET_O)) mark = true;
    }
    if (!mark || nbr_has_skip(nbr, nmax, kap)) return;
    if (even)
        tmp[kap] = REF_INDET_O;
    else
        ref[kap] = REF_INDET_E;
}

// Cu_MarkBlocks_Finalize (:396-414): intermediate marks become refinement marks.
__global__ void k_mark_finalize(int idm, int *ref, int *tmp)
{
    const int kap = blockIdx.x * GWG + threadIdx.x;
    if (kap < idm && (ref[kap] == REF_INDET_E || tmp[kap] == REF_INDET_O)) {
        ref[kap] = REF_MARK_REFINE;
        tmp[kap] = -1;
    }
}

// Cu_MarkBlocks_CheckMasks_S1/S2 (mesh_voxelizer_masks.cu:21-226, the neighbour-table variant):
// in leaf blocks of level L whose block mask is SOLIDB or SOLIDA, every non-solid cell with a solid
// cell among its 26 neighbours becomes a BOUNDARY cell, and the block is flagged as on a boundary.
// With all_leaf (links made to agree with the cell masks) every leaf block is examined, so that
// fluid cells beside a solid cell are BOUNDARY cells also in REGULAR blocks (no solid cell of their
// own) and in blocks classified SOLID that still hold fluid cells (the reference skips both).  Only SOLID masks are read
// and only non-solid cells are written, so one pass is race-free.
__global__ __launch_bounds__(GWG) void k_check_masks(const int *__restrict__ ids, int n, const int *__restrict__ nbr_slot,
                                                     const int *__restrict__ nbr_child, const int *__restrict__ blk_mask,
                                                     int *__restrict__ cells_mask, int *__restrict__ onb,
                                                     int *__restrict__ bcount, int all_leaf)
{
    const int t = threadIdx.x & 63;
    const int bi = blockIdx.x * (GWG / 64) + (threadIdx.x >> 6);
    if (bi >= n) return;  // whole wavefronts
    const int b = ids[bi];
    const int bm = blk_mask[b];
    if (nbr_child[b] >= 0) return;  // wave-uniform
    if (!all_leaf && !(bm < 0 && bm != BLK_SOLID)) return;
    const int my_nb = t < 27 ? nbr_slot[(long long)b * 27 + t] : 0;
    const long long cell = (long long)b * 64 + t;
    const int I = t & 3, J = (t >> 2) & 3, K = t >> 4;
    bool near = false;
    for (int s = 0; s < 27; s++) {
        const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
        const int Ip = I + dx, Jp = J + dy, Kp = K + dz;
        const int slot = (Ip < 0 ? 0 : (Ip >= 4 ? 2 : 1)) + 3 * (Jp < 0 ? 0 : (Jp >= 4 ? 2 : 1)) +
                         9 * (Kp < 0 ? 0 : (Kp >= 4 ? 2 : 1));
        const int nbk = __shfl(my_nb, slot, 64);
        if (nbk >= 0 && cells_mask[(long long)nbk * 64 + ((Ip & 3) + 4 * (Jp & 3) + 16 * (Kp & 3))] == CELL_SOLID)
            near = true;
    }
    const bool bnd = near && cells_mask[cell] != CELL_SOLID;
    if (bnd) cells_mask[cell] = CELL_BOUNDARY;
    // The block's count of boundary cells (tmp_1, mesh_voxelizer_masks.cu:151-175).
    const int nb = __popcll(__ballot(bnd));
    if (nb > 0 && t == 0) {
        onb[b] = 1;
        if (bcount) bcount[b] = nb;
    }
}

__global__ void k_fill_int(int *a, long long n, int v)
{
    const long long i = (long long)blockIdx.x * GWG + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
static inline unsigned nblk(long long n) { return (unsigned)((n + GWG - 1) / GWG); }

void free_geometry(Ctx *c)
{
    if (c->geo_tri) (void)hipFree(c->geo_tri);
    if (c->geo_bin_start) (void)hipFree(c->geo_bin_start);
    if (c->geo_bin_tri) (void)hipFree(c->geo_bin_tri);
    c->geo_tri = nullptr;
    c->geo_bin_start = c->geo_bin_tri = nullptr;
    c->geo_n_tri = 0;
}

// Uploads n_tri triangles (x1 y1 z1 x2 y2 z2 x3 y3 z3 each) and bins them by their (y, z) extent.
int set_geometry(Ctx *c, const float *tri, long long n_tri)
{
    if (c->DIM != 3) {
        c->last_error = "geometry is built for 3D velocity sets only";
        return AGAL_ERR_NOT_BUILT;
    }
    if (n_tri <= 0 || !tri) return AGAL_ERR_ARG;
    if (c->geo_tri) (void)hipFree(c->geo_tri);
    if (c->geo_bin_start) (void)hipFree(c->geo_bin_start);
    if (c->geo_bin_tri) (void)hipFree(c->geo_bin_tri);
    c->geo_tri = nullptr;
    c->geo_bin_start = c->geo_bin_tri = nullptr;
    // (y, z) bins over the geometry's bounding box.
    double ymin = 1e300, ymax = -1e300, zmin = 1e300, zmax = -1e300;
    for (long long k = 0; k < n_tri; k++)
        for (int v = 0; v < 3; v++) {
            ymin = std::min(ymin, (double)tri[9 * k + 3 * v + 1]);
            ymax = std::max(ymax, (double)tri[9 * k + 3 * v + 1]);
            zmin = std::min(zmin, (double)tri[9 * k + 3 * v + 2]);
            zmax = std::max(zmax, (double)tri[9 * k + 3 * v + 2]);
        }
    const int G = std::max(1, std::min(256, (int)std::sqrt((double)n_tri)));
    const double pad = 1e-6 * std::max(1.0, std::max(ymax - ymin, zmax - zmin));
    ymin -= pad;
    zmin -= pad;
    ymax += pad;
    zmax += pad;
    c->geo_G = G;
    c->geo_y0 = ymin;
    c->geo_z0 = zmin;
    c->geo_inv_dy = (double)G / (ymax - ymin);
    c->geo_inv_dz = (double)G / (zmax - zmin);
    std::vector<std::vector<int>> bins((size_t)G * G);
    for (long long k = 0; k < n_tri; k++) {
        double y0 = 1e300, y1 = -1e300, z0 = 1e300, z1 = -1e300;
        for (int v = 0; v < 3; v++) {
            y0 = std::min(y0, (double)tri[9 * k + 3 * v + 1]);
            y1 = std::max(y1, (double)tri[9 * k + 3 * v + 1]);
            z0 = std::min(z0, (double)tri[9 * k + 3 * v + 2]);
            z1 = std::max(z1, (double)tri[9 * k + 3 * v + 2]);
        }
        const int iy0 = std::max(0, std::min(G - 1, (int)std::floor((y0 - ymin) * c->geo_inv_dy)));
        const int iy1 = std::max(0, std::min(G - 1, (int)std::floor((y1 - ymin) * c->geo_inv_dy)));
        const int iz0 = std::max(0, std::min(G - 1, (int)std::floor((z0 - zmin) * c->geo_inv_dz)));
        const int iz1 = std::max(0, std::min(G - 1, (int)std::floor((z1 - zmin) * c->geo_inv_dz)));
        for (int iz = iz0; iz <= iz1; iz++)
            for (int iy = iy0; iy <= iy1; iy++) bins[(size_t)iy + (size_t)G * iz].push_back((int)k);
    }
    std::vector<int> start((size_t)G * G + 1, 0), list;
    for (size_t i = 0; i < bins.size(); i++) {
        start[i] = (int)list.size();
        list.insert(list.end(), bins[i].begin(), bins[i].end());
    }
    start[bins.size()] = (int)list.size();
    AGAL_HIP(c, hipMalloc(&c->geo_tri, sizeof(float) * 9 * n_tri));
    AGAL_HIP(c, hipMalloc(&c->geo_bin_start, sizeof(int) * start.size()));
    AGAL_HIP(c, hipMalloc(&c->geo_bin_tri, sizeof(int) * std::max<size_t>(1, list.size())));
    AGAL_HIP(c, hipMemcpy(c->geo_tri, tri, sizeof(float) * 9 * n_tri, hipMemcpyHostToDevice));
    AGAL_HIP(c, hipMemcpy(c->geo_bin_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    if (!list.empty()) AGAL_HIP(c, hipMemcpy(c->geo_bin_tri, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice));
    c->geo_n_tri = n_tri;
    return AGAL_OK;
}

int voxelize_level(Ctx *c, int L)
{
    if (!c->geo_tri) {
        c->last_error = "no geometry set";
        return AGAL_ERR_STATE;
    }
    const int n = c->n_ids[L];
    if (n > 0) {
        hipLaunchKernelGGL(k_voxelize, dim3((unsigned)((n + 3) / 4)), dim3(GWG), 0, c->stream, c->id_set_L(L), n, c->X,
                           c->n_maxcblocks, c->dxf_vec[L], c->cells_mask, c->blk_mask, c->geo_tri, c->geo_bin_start,
                           c->geo_bin_tri, c->geo_G, c->geo_y0, c->geo_z0, c->geo_inv_dy, c->geo_inv_dz);
        AGAL_HIP(c, hipGetLastError());
    }
    return AGAL_OK;
}

// Number of propagation passes of the near-wall band on level L (mesh_voxelizer.cu:39-41), in the
// geometry's single precision.
int near_wall_passes(double near_wall_distance, int L, double dx_L)
{
    const float R = (float)near_wall_distance / (float)std::pow(2.0, (double)L);
    return (int)(R / (float)(4.0 * std::sqrt(2.0) * dx_L)) + 1;
}

// Block marks of level L after its voxelization (M_Geometry_Voxelize_S1, mesh_voxelizer.cu:82-125):
// boundary and exterior marks, `passes` propagation passes and the finalization.  The scratch row
// scan_b carries the odd marks (resting value -1).
int mark_near_wall(Ctx *c, int L)
{
    const int idm = c->id_max[L];
    if (c->n_ids[L] <= 0 || idm <= 0) return AGAL_OK;
    const bool hit_max = L == c->MAX_LEVELS - 1;
    const int nprop = near_wall_passes(c->P.G_NEAR_WALL_DISTANCE, L, c->dxf_vec[L]);
    const long long nmax = c->n_maxcblocks;
    int *tmp = c->scan_b;
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(idm)), dim3(GWG), 0, c->stream, tmp, (long long)idm, -1);
    hipLaunchKernelGGL(k_mark_boundary, dim3(nblk(idm)), dim3(GWG), 0, c->stream, idm, nmax, L, hit_max, c->ref,
                       c->blk_mask, c->nbr, c->level);
    hipLaunchKernelGGL(k_mark_exterior, dim3(nblk(idm)), dim3(GWG), 0, c->stream, idm, nmax, L, hit_max, c->ref,
                       c->blk_mask, c->nbr, c->level);
    for (int j = 0; j < nprop; j++)
        hipLaunchKernelGGL(k_mark_propagate, dim3(nblk(idm)), dim3(GWG), 0, c->stream, idm, nmax, L, j, c->ref,
                           c->blk_mask, c->nbr, c->level, tmp);
    hipLaunchKernelGGL(k_mark_finalize, dim3(nblk(idm)), dim3(GWG), 0, c->stream, idm, c->ref, tmp);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// Boundary-cell masks of level L (M_Geometry_Voxelize_S2, mesh_voxelizer.cu:139-174).
int check_masks_level(Ctx *c, int L)
{
    const int n = c->n_ids[L];
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_check_masks, dim3((unsigned)((n + 3) / 4)), dim3(GWG), 0, c->stream, c->id_set_L(L), n,
                       c->nbr_slot, c->nbr_child, c->blk_mask, c->cells_mask, c->onb, c->lk_bcount,
                       c->link_masks ? 1 : 0);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// ---------------------------------------------------------------------------------------------
// Icosphere and STL import (host only).
// ---------------------------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

long long icosphere(double cx, double cy, double cz, double r, int subdiv, double noise, unsigned long long seed,
                    float *out, long long cap)
{
    if (subdiv < 0 || subdiv > 8 || r <= 0) return -1;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> V = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<int> F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                          3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto norm = [](double *v) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= l;
        v[1] /= l;
        v[2] /= l;
    };
    for (size_t i = 0; i < V.size(); i += 3) norm(&V[i]);
    for (int s = 0; s < subdiv; s++) {
        std::vector<int> G;
        std::vector<std::pair<long long, int>> cache;  // edge key -> midpoint index
        auto mid = [&](int a, int b) {
            const long long key = (long long)std::min(a, b) * 1000000007LL + std::max(a, b);
            for (auto &e : cache)
                if (e.first == key) return e.second;
            double m[3] = {(V[3 * a] + V[3 * b]) / 2, (V[3 * a + 1] + V[3 * b + 1]) / 2, (V[3 * a + 2] + V[3 * b + 2]) / 2};
            norm(m);
            const int idx = (int)(V.size() / 3);
            V.insert(V.end(), m, m + 3);
            cache.push_back({key, idx});
            return idx;
        };
        for (size_t f = 0; f < F.size(); f += 3) {
            const int a = F[f], b = F[f + 1], c = F[f + 2];
            const int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const int nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            G.insert(G.end(), nf, nf + 12);
        }
        F.swap(G);
    }
    const long long nt = (long long)F.size() / 3;
    if (!out) return nt;
    if (cap < nt) return -2;
    // Radii with uniform seeded noise in [-noise, +noise) per vertex, in vertex order.
    uint64_t st = seed;
    std::vector<double> rad(V.size() / 3);
    for (size_t i = 0; i < rad.size(); i++) {
        const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
        rad[i] = r * (1.0 + noise * (2.0 * u - 1.0));
    }
    for (long long f = 0; f < nt; f++)
        for (int v = 0; v < 3; v++) {
            const int i = F[3 * f + v];
            out[9 * f + 3 * v + 0] = (float)(cx + rad[i] * V[3 * i + 0]);
            out[9 * f + 3 * v + 1] = (float)(cy + rad[i] * V[3 * i + 1]);
            out[9 * f + 3 * v + 2] = (float)(cz + rad[i] * V[3 * i + 2]);
        }
    return nt;
}

// ASCII STL (Geometry::G_ImportSTL_ASCII, src/geometry/geometry_import.cu:11-86): the three
// "vertex" lines of every facet, read as floats; facets with two equal vertices are dropped.
long long load_stl_ascii(const char *path, float *out, long long cap)
{
    FILE *fp = std::fopen(path, "r");
    if (!fp) return -1;
    long long nt = 0;
    int nv = 0;
    float v[9];
    char word[256];
    while (std::fscanf(fp, "%255s", word) == 1) {
        if (std::strcmp(word, "vertex") != 0) continue;
        if (std::fscanf(fp, "%f %f %f", &v[3 * nv], &v[3 * nv + 1], &v[3 * nv + 2]) != 3) {
            std::fclose(fp);
            return -3;
        }
        if (++nv < 3) continue;
        nv = 0;
        auto same = [&](int i, int j) { return v[3 * i] == v[3 * j] && v[3 * i + 1] == v[3 * j + 1] && v[3 * i + 2] == v[3 * j + 2]; };
        if (same(0, 1) || same(0, 2) || same(1, 2)) continue;
        if (out) {
            if (nt >= cap) {
                std::fclose(fp);
                return -2;
            }
            std::memcpy(out + 9 * nt, v, sizeof(v));
        }
        nt++;
    }
    std::fclose(fp);
    return nv == 0 ? nt : -3;
}

}  // namespace agal
