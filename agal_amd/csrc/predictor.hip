// predictor.hip -- learned refinement predictor for gfx950: per-block features, a 12-64-64-1
// multilayer perceptron with FP16 weights and FP32 accumulation, its training with Adam on the
// plain criterion's labels, and inference over all blocks.
//
// No counterpart in the reference.  The plain criterion (Cu_ComputeRefCriteria,
// solver_lbm_criterion.cu:438-592) stays the fallback and the referee: its label for a block is
// "the criterion wants this block to have children" (the refine / keep decision of its marks).
//
// Parameters, one flat FP32 master array of NPARAM values (row-major):
//   W1 [12][64], b1 [64], W2 [64][64], b2 [64], W3 [64], b3
// and an FP16 copy used by every forward and backward pass.  z1 = b1 + x W1, h1 = relu(z1),
// z2 = b2 + h1 W2, h2 = relu(z2), z3 = b3 + h2 . W3, p = sigmoid(z3).  Every dot product is summed
// in index order in FP32, so the CPU restatement reproduces z3 bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "agal_common.hpp"
#include "agal_ctx.hpp"

namespace agal {

constexpr int PWG = 256;

// Piecewise-linear log2 (exact at powers of two), the same bits on every IEEE machine:
// v = m 2^e with m in [0.5, 1) -> (e - 1) + 2 (m - 0.5).
__host__ __device__ inline float plog2(float v)
{
    int e;
    const float m = frexpf(v, &e);
    return (float)(e - 1) + (m - 0.5f) * 2.0f;
}

// ---------------------------------------------------------------------------------------------
// Per-block statistics of the stored macroscopic fields: max |u|, max |rho - 1|, max |grad u|
// (Frobenius), max |vorticity|, over the block's non-ghost, non-solid cells.  The velocity
// gradient uses the criterion's stencil: central differences on a 6^3 tile whose halo is
// extrapolated from inside the block (ExtrapolateToHalo, inc/util_math.h:54-112), with the block's
// own spacing.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T pex4(T d1, T d2, T d3, T d4)
{
    return (T)(+4.0) * d1 + (T)(-6.0) * d2 + (T)(+4.0) * d3 + (T)(-1.0) * d4;
}

template <typename T, int DIM>
__global__ __launch_bounds__(PWG) void k_block_stats(int id_max, long long ncm, const T *__restrict__ M,
                                                     const int *__restrict__ ref, const int *__restrict__ level,
                                                     const int *__restrict__ cells_mask, T dx0,
                                                     float *__restrict__ stats)
{
    constexpr int CELLS = DIM == 3 ? 64 : 16, HB = DIM == 3 ? 216 : 36, BPG = PWG / CELLS;
    __shared__ T s_u[BPG * HB], s_v[BPG * HB], s_w[BPG * HB];
    const int t = threadIdx.x % CELLS, g = threadIdx.x / CELLS;
    const int b = blockIdx.x * BPG + g;
    bool work = b < id_max;
    if (work) work = ref[b] != REF_INACTIVE;
    const int I = t % 4, J = (t / 4) % 4, K = DIM == 3 ? t / 16 : 0;
    T *su = s_u + g * HB, *sv = s_v + g * HB, *sw = s_w + g * HB;
    auto H = [](int i, int j, int k) { return DIM == 2 ? (i + 1) + 6 * (j + 1) : (i + 1) + 6 * (j + 1) + 36 * (k + 1); };
    const long long cell = (long long)b * CELLS + t;
    T rho = 1, u = 0, v = 0, w = 0;
    if (work) {
        rho = M[cell];
        u = M[1 * ncm + cell];
        v = M[2 * ncm + cell];
        w = DIM == 3 ? M[3 * ncm + cell] : (T)0;
        su[H(I, J, K)] = u;
        sv[H(I, J, K)] = v;
        sw[H(I, J, K)] = w;
    }
    __syncthreads();
    T *arrs[3] = {su, sv, sw};
    for (int a = 0; a < (DIM == 3 ? 3 : 2); a++) {
        T *s = arrs[a];
        if (work) {
            if (I == 0) {
                s[H(-1, J, K)] = pex4(s[H(0, J, K)], s[H(1, J, K)], s[H(2, J, K)], s[H(3, J, K)]);
                s[H(4, J, K)] = pex4(s[H(3, J, K)], s[H(2, J, K)], s[H(1, J, K)], s[H(0, J, K)]);
            }
            if (J == 0) {
                s[H(I, -1, K)] = pex4(s[H(I, 0, K)], s[H(I, 1, K)], s[H(I, 2, K)], s[H(I, 3, K)]);
                s[H(I, 4, K)] = pex4(s[H(I, 3, K)], s[H(I, 2, K)], s[H(I, 1, K)], s[H(I, 0, K)]);
            }
            if (DIM == 3 && K == 0) {
                s[H(I, J, -1)] = pex4(s[H(I, J, 0)], s[H(I, J, 1)], s[H(I, J, 2)], s[H(I, J, 3)]);
                s[H(I, J, 4)] = pex4(s[H(I, J, 3)], s[H(I, J, 2)], s[H(I, J, 1)], s[H(I, J, 0)]);
            }
        }
    }
    __syncthreads();
    float st[4] = {0.f, 0.f, 0.f, 0.f};
    if (work) {
        const int m = cells_mask[cell];
        if (m != CELL_GHOST && m != CELL_SOLID) {
            const T d2 = (T)2 * (dx0 / (T)(1 << level[b]));
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
            st[0] = (float)sqrt(u * u + v * v + w * w);
            st[1] = (float)fabs(rho - (T)1);
            st[2] = (float)sqrt(uX * uX + uY * uY + uZ * uZ + vX * vX + vY * vY + vZ * vZ + wX * wX + wY * wY + wZ * wZ);
            st[3] = (float)sqrt((wY - vZ) * (wY - vZ) + (uZ - wX) * (uZ - wX) + (vX - uY) * (vX - uY));
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int d = 1; d < CELLS; d <<= 1) st[k] = fmaxf(st[k], __shfl_xor(st[k], d, CELLS));
    if (b < id_max && t == 0)
        for (int k = 0; k < 4; k++) stats[4LL * b + k] = st[k];
}

// Signed distance feature of a block: geometry set -> distance from the block centre to the
// nearest triangle centroid, negative for blocks holding solid cells; otherwise the distance to the
// nearest domain face.  In root-block units, clamped to [-8, 8], divided by 8.
__global__ void k_block_distance(int id_max, const int *__restrict__ ref, const int *__restrict__ level,
                                 const int *__restrict__ blk_mask, const float *__restrict__ X, long long nmax,
                                 int dim, float dx0, float Lx, float Ly, float Lz, const float *__restrict__ tri,
                                 long long n_tri, float *__restrict__ dist)
{
    const int b = blockIdx.x * PWG + threadIdx.x;
    if (b >= id_max) return;
    if (ref[b] == REF_INACTIVE) {
        dist[b] = 0.f;
        return;
    }
    const float h = 2.0f * dx0 / (float)(1 << level[b]);  // half the block edge
    const float xc = X[b] + h, yc = X[b + nmax] + h, zc = dim == 3 ? X[b + 2 * nmax] + h : 0.f;
    float d;
    if (tri) {
        float best = 3.0e38f;
        for (long long k = 0; k < n_tri; k++) {
            const float *q = tri + 9 * k;
            const float cx = (q[0] + q[3] + q[6]) * (1.0f / 3.0f) - xc;
            const float cy = (q[1] + q[4] + q[7]) * (1.0f / 3.0f) - yc;
            const float cz = (q[2] + q[5] + q[8]) * (1.0f / 3.0f) - zc;
            best = fminf(best, cx * cx + cy * cy + cz * cz);
        }
        d = sqrtf(best);
        const int m = blk_mask[b];
        if (m == BLK_SOLID || m == BLK_SOLIDB) d = -d;
    } else {
        d = fminf(fminf(xc, Lx - xc), fminf(yc, Ly - yc));
        if (dim == 3) d = fminf(d, fminf(zc, Lz - zc));
    }
    d = d / (4.0f * dx0);
    dist[b] = fminf(fmaxf(d, -8.0f), 8.0f) * 0.125f;
}

// Feature vector of every block ID (zeros for inactive IDs):
//   0-3  plog2(stat + 2^-30) / 16 of the block's max |u|, |rho-1|, |grad u|, |vorticity|
//   4    level / max(1, MAX_LEVELS - 1)
//   5    signed distance feature
//   6-9  maxima of 0-3 over the block and its same-level neighbours
//   10   fraction of the existing neighbours that have children
//   11   minimum of 5 over the block and its neighbours
__global__ void k_features(int id_max, const int *__restrict__ ref, const int *__restrict__ level,
                           const int *__restrict__ nbr, const int *__restrict__ nbr_child, long long nmax, int qmax,
                           int max_levels, const float *__restrict__ stats, const float *__restrict__ dist,
                           float *__restrict__ feat)
{
    const int b = blockIdx.x * PWG + threadIdx.x;
    if (b >= id_max) return;
    float *f = feat + (long long)NFEAT * b;
    if (ref[b] == REF_INACTIVE) {
        for (int k = 0; k < NFEAT; k++) f[k] = 0.f;
        return;
    }
    const float tiny = 9.313225746154785e-10f;  // 2^-30
    float own[4], mx[4];
    for (int k = 0; k < 4; k++) mx[k] = own[k] = plog2(stats[4LL * b + k] + tiny) * 0.0625f;
    float dmin = dist[b];
    int n_nb = 0, n_ref = 0;
    for (int p = 1; p < qmax; p++) {
        const int nb = nbr[b + p * nmax];
        if (nb < 0) continue;
        n_nb++;
        if (nbr_child[nb] >= 0) n_ref++;
        for (int k = 0; k < 4; k++) mx[k] = fmaxf(mx[k], plog2(stats[4LL * nb + k] + tiny) * 0.0625f);
        dmin = fminf(dmin, dist[nb]);
    }
    for (int k = 0; k < 4; k++) {
        f[k] = own[k];
        f[6 + k] = mx[k];
    }
    f[4] = (float)level[b] / (float)(max_levels > 1 ? max_levels - 1 : 1);
    f[5] = dist[b];
    f[10] = n_nb > 0 ? (float)n_ref / (float)n_nb : 0.f;
    f[11] = dmin;
}

// ---------------------------------------------------------------------------------------------
// The network.  Weights are staged in LDS as FP32 values of the FP16 copy; one thread per sample.
// ---------------------------------------------------------------------------------------------
__device__ inline void stage_weights(const _Float16 *__restrict__ wh, float *s_w)
{
    for (int q = threadIdx.x; q < MLP_NPARAM; q += blockDim.x) s_w[q] = (float)wh[q];
    __syncthreads();
}

__device__ inline float sigmoidf_dev(float z) { return 1.0f / (1.0f + expf(-z)); }

// Forward pass over n feature rows; optional outputs z1, z2 (n x 64) for training.
__global__ __launch_bounds__(PWG) void k_mlp_forward(const float *__restrict__ X, int n, const _Float16 *__restrict__ wh,
                                                     float *__restrict__ z3_out, float *__restrict__ p_out,
                                                     float *__restrict__ Z1, float *__restrict__ Z2)
{
    __shared__ float s_w[MLP_NPARAM];
    stage_weights(wh, s_w);
    const int r = blockIdx.x * PWG + threadIdx.x;
    if (r >= n) return;
    const float *W1 = s_w + MLP_OFF_W1, *B1 = s_w + MLP_OFF_B1, *W2 = s_w + MLP_OFF_W2, *B2 = s_w + MLP_OFF_B2;
    const float *W3 = s_w + MLP_OFF_W3;
    float x[NFEAT];
    for (int i = 0; i < NFEAT; i++) x[i] = X[(long long)NFEAT * r + i];
    float h1[MLP_H];
    for (int j = 0; j < MLP_H; j++) {
        float z = B1[j];
        for (int i = 0; i < NFEAT; i++) z += x[i] * W1[i * MLP_H + j];
        if (Z1) Z1[(long long)MLP_H * r + j] = z;
        h1[j] = z > 0.f ? z : 0.f;
    }
    float z3 = s_w[MLP_OFF_B3];
    for (int k = 0; k < MLP_H; k++) {
        float z = B2[k];
        for (int j = 0; j < MLP_H; j++) z += h1[j] * W2[j * MLP_H + k];
        if (Z2) Z2[(long long)MLP_H * r + k] = z;
        z3 += (z > 0.f ? z : 0.f) * W3[k];
    }
    if (z3_out) z3_out[r] = z3;
    if (p_out) p_out[r] = sigmoidf_dev(z3);
}

// Loss and output gradient of the weighted binary cross-entropy on logits:
//   loss_r = w_r (y_r softplus(-z) + (1 - y_r) softplus(z)),  w_r = pos_weight if y_r = 1 else 1,
//   dL/dz_r = w_r (sigmoid(z) - y_r) / n.
__global__ void k_mlp_output_grad(const float *__restrict__ z3, const float *__restrict__ Y, int n, float pos_weight,
                                  float *__restrict__ dZ3, float *__restrict__ loss)
{
    const int r = blockIdx.x * PWG + threadIdx.x;
    if (r >= n) return;
    const float z = z3[r], y = Y[r];
    const float wgt = y > 0.5f ? pos_weight : 1.0f;
    const float sp_pos = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z)));  // softplus(z)
    const float sp_neg = sp_pos - z;                                // softplus(-z)
    loss[r] = wgt * (y * sp_neg + (1.0f - y) * sp_pos);
    dZ3[r] = wgt * (sigmoidf_dev(z) - y) / (float)n;
}

// Hidden-layer deltas per sample: dz2_k = dz3 W3_k [z2_k > 0]; dz1_j = [z1_j > 0] sum_k W2_jk dz2_k.
__global__ __launch_bounds__(PWG) void k_mlp_backward(int n, const _Float16 *__restrict__ wh,
                                                      const float *__restrict__ Z1, const float *__restrict__ Z2,
                                                      const float *__restrict__ dZ3, float *__restrict__ dZ2,
                                                      float *__restrict__ dZ1)
{
    __shared__ float s_w[MLP_NPARAM];
    stage_weights(wh, s_w);
    const int r = blockIdx.x * PWG + threadIdx.x;
    if (r >= n) return;
    const float *W2 = s_w + MLP_OFF_W2, *W3 = s_w + MLP_OFF_W3;
    const float d3 = dZ3[r];
    float d2[MLP_H];
    for (int k = 0; k < MLP_H; k++) {
        d2[k] = Z2[(long long)MLP_H * r + k] > 0.f ? d3 * W3[k] : 0.f;
        dZ2[(long long)MLP_H * r + k] = d2[k];
    }
    for (int j = 0; j < MLP_H; j++) {
        float s = 0.f;
        for (int k = 0; k < MLP_H; k++) s += W2[j * MLP_H + k] * d2[k];
        dZ1[(long long)MLP_H * r + j] = Z1[(long long)MLP_H * r + j] > 0.f ? s : 0.f;
    }
}

// Parameter gradients, one thread per parameter, samples summed in order.
__global__ void k_mlp_wgrad(int n, const float *__restrict__ X, const float *__restrict__ Z1,
                            const float *__restrict__ Z2, const float *__restrict__ dZ3, const float *__restrict__ dZ2,
                            const float *__restrict__ dZ1, float *__restrict__ grad)
{
    const int q = blockIdx.x * PWG + threadIdx.x;
    if (q >= MLP_NPARAM) return;
    float s = 0.f;
    if (q < MLP_OFF_B1) {  // W1[i][j]
        const int i = q / MLP_H, j = q % MLP_H;
        for (int r = 0; r < n; r++) s += X[(long long)NFEAT * r + i] * dZ1[(long long)MLP_H * r + j];
    } else if (q < MLP_OFF_W2) {  // b1[j]
        const int j = q - MLP_OFF_B1;
        for (int r = 0; r < n; r++) s += dZ1[(long long)MLP_H * r + j];
    } else if (q < MLP_OFF_B2) {  // W2[j][k]
        const int j = (q - MLP_OFF_W2) / MLP_H, k = (q - MLP_OFF_W2) % MLP_H;
        for (int r = 0; r < n; r++) {
            const float z = Z1[(long long)MLP_H * r + j];
            s += (z > 0.f ? z : 0.f) * dZ2[(long long)MLP_H * r + k];
        }
    } else if (q < MLP_OFF_W3) {  // b2[k]
        const int k = q - MLP_OFF_B2;
        for (int r = 0; r < n; r++) s += dZ2[(long long)MLP_H * r + k];
    } else if (q < MLP_OFF_B3) {  // W3[k]
        const int k = q - MLP_OFF_W3;
        for (int r = 0; r < n; r++) {
            const float z = Z2[(long long)MLP_H * r + k];
            s += (z > 0.f ? z : 0.f) * dZ3[r];
        }
    } else {  // b3
        for (int r = 0; r < n; r++) s += dZ3[r];
    }
    grad[q] = s;
}

// Adam (Kingma & Ba 2015) with bias correction; the FP16 copy follows the master weights.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, _Float16 *__restrict__ wh, float lr, float b1, float b2,
                       float eps, float bc1, float bc2)
{
    const int q = blockIdx.x * PWG + threadIdx.x;
    if (q >= MLP_NPARAM) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    const float wq = w[q] - lr * (mq / bc1) / (sqrtf(vq / bc2) + eps);
    w[q] = wq;
    wh[q] = (_Float16)wq;
}

__global__ void k_to_half(const float *__restrict__ w, _Float16 *__restrict__ wh)
{
    const int q = blockIdx.x * PWG + threadIdx.x;
    if (q < MLP_NPARAM) wh[q] = (_Float16)w[q];
}

// Training batch rows gathered from a feature snapshot: X[r] = snap[ids[r]], Y[r] = label[ids[r]].
__global__ void k_gather_batch(const float *__restrict__ snap, const int *__restrict__ label,
                               const int *__restrict__ ids, int n, float *__restrict__ X, float *__restrict__ Y)
{
    const int r = blockIdx.x * PWG + threadIdx.x;
    if (r >= n) return;
    const int b = ids[r];
    for (int k = 0; k < NFEAT; k++) X[(long long)NFEAT * r + k] = snap[(long long)NFEAT * b + k];
    Y[r] = (float)label[b];
}

// Refinement marks from predictions: an unrefined, eligible block below the finest interior level
// whose predicted probability exceeds the threshold is marked for refinement.
__global__ void k_mark_predicted(int id_max, const float *__restrict__ prob, float thr, int *__restrict__ ref,
                                 const int *__restrict__ level, const int *__restrict__ nbr, long long nmax, int qmax,
                                 int mli, int *__restrict__ count)
{
    const int b = blockIdx.x * PWG + threadIdx.x;
    if (b >= id_max || ref[b] != REF_UNREFINED || level[b] >= mli - 1 || !(prob[b] > thr)) return;
    for (int p = 0; p < qmax; p++)
        if (nbr[b + p * nmax] == N_SKIPID) return;
    ref[b] = REF_MARK_REFINE;
    atomicAdd(count, 1);
}

// Predicted blocks are kept refined: coarsening marks on blocks predicted to need children are
// withdrawn.
__global__ void k_keep_predicted(int id_max, const float *__restrict__ prob, float thr, int *__restrict__ ref)
{
    const int b = blockIdx.x * PWG + threadIdx.x;
    if (b < id_max && ref[b] == REF_MARK_COARSEN && prob[b] > thr) ref[b] = REF_REFINED;
}

// Audit counts of the plain criterion against the current mesh: [0] blocks the criterion wants
// refined, [1] of those still unrefined (missed).
__global__ void k_audit(int id_max, const int *__restrict__ label, const int *__restrict__ nbr_child,
                        int *__restrict__ counts)
{
    const int b = blockIdx.x * PWG + threadIdx.x;
    if (b >= id_max || label[b] != 1) return;
    atomicAdd(&counts[0], 1);
    if (nbr_child[b] < 0) atomicAdd(&counts[1], 1);
}

// ---------------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------------
static inline unsigned pgrid(long long n) { return (unsigned)((n + PWG - 1) / PWG); }

static uint64_t mix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double unit01(uint64_t &s) { return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0); }

void free_predictor(Ctx *c)
{
    Predictor &P = c->pred;
    void *ptrs[] = {P.w, P.m, P.v, P.g, P.wh, P.stats, P.dist, P.feat, P.prob, P.label, P.snap, P.X, P.Y, P.ids,
                    P.Z1, P.Z2, P.z3, P.dZ3, P.dZ2, P.dZ1, P.loss, P.counts};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    P = Predictor{};
}

template <typename Tp>
static int palloc(Ctx *c, Tp **p, size_t bytes)
{
    if (*p) return AGAL_OK;
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        c->last_error = std::string("hipMalloc(predictor): ") + hipGetErrorString(e);
        return AGAL_ERR_ALLOC;
    }
    *p = static_cast<Tp *>(v);
    return AGAL_OK;
}

// He-uniform initialisation from a splitmix64 stream (weights in parameter order, biases zero).
void predictor_init_weights(unsigned long long seed, float *w)
{
    uint64_t s = seed;
    auto fill = [&](int off, int fan_in, int count) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int q = 0; q < count; q++) w[off + q] = (float)((2.0 * unit01(s) - 1.0) * a);
    };
    std::fill(w, w + MLP_NPARAM, 0.0f);
    fill(MLP_OFF_W1, NFEAT, NFEAT * MLP_H);
    fill(MLP_OFF_W2, MLP_H, MLP_H * MLP_H);
    fill(MLP_OFF_W3, MLP_H, MLP_H);
}

int predictor_setup(Ctx *c, unsigned long long seed, int batch_cap)
{
    Predictor &P = c->pred;
    const long long nmax = c->n_maxcblocks;
    int st;
    if ((st = palloc(c, &P.w, sizeof(float) * MLP_NPARAM)) || (st = palloc(c, &P.m, sizeof(float) * MLP_NPARAM)) ||
        (st = palloc(c, &P.v, sizeof(float) * MLP_NPARAM)) || (st = palloc(c, &P.g, sizeof(float) * MLP_NPARAM)) ||
        (st = palloc(c, &P.wh, sizeof(_Float16) * MLP_NPARAM)) || (st = palloc(c, &P.stats, sizeof(float) * 4 * nmax)) ||
        (st = palloc(c, &P.dist, sizeof(float) * nmax)) || (st = palloc(c, &P.feat, sizeof(float) * NFEAT * nmax)) ||
        (st = palloc(c, &P.prob, sizeof(float) * nmax)) || (st = palloc(c, &P.label, sizeof(int) * nmax)) ||
        (st = palloc(c, &P.counts, sizeof(int) * 4)))
        return st;
    if (batch_cap > P.batch_cap) {
        for (void **p : {(void **)&P.X, (void **)&P.Y, (void **)&P.ids, (void **)&P.Z1, (void **)&P.Z2, (void **)&P.z3,
                         (void **)&P.dZ3, (void **)&P.dZ2, (void **)&P.dZ1, (void **)&P.loss}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        const size_t B = (size_t)batch_cap;
        if ((st = palloc(c, &P.X, sizeof(float) * NFEAT * B)) || (st = palloc(c, &P.Y, sizeof(float) * B)) ||
            (st = palloc(c, &P.ids, sizeof(int) * B)) || (st = palloc(c, &P.Z1, sizeof(float) * MLP_H * B)) ||
            (st = palloc(c, &P.Z2, sizeof(float) * MLP_H * B)) || (st = palloc(c, &P.z3, sizeof(float) * B)) ||
            (st = palloc(c, &P.dZ3, sizeof(float) * B)) || (st = palloc(c, &P.dZ2, sizeof(float) * MLP_H * B)) ||
            (st = palloc(c, &P.dZ1, sizeof(float) * MLP_H * B)) || (st = palloc(c, &P.loss, sizeof(float) * B)))
            return st;
        P.batch_cap = batch_cap;
    }
    std::vector<float> w(MLP_NPARAM);
    predictor_init_weights(seed, w.data());
    return predictor_set_weights(c, w.data());
}

int predictor_set_weights(Ctx *c, const float *w_host)
{
    Predictor &P = c->pred;
    if (!P.w) return AGAL_ERR_STATE;
    AGAL_HIP(c, hipMemcpyAsync(P.w, w_host, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
    AGAL_HIP(c, hipMemsetAsync(P.m, 0, sizeof(float) * MLP_NPARAM, c->stream));
    AGAL_HIP(c, hipMemsetAsync(P.v, 0, sizeof(float) * MLP_NPARAM, c->stream));
    hipLaunchKernelGGL(k_to_half, dim3(pgrid(MLP_NPARAM)), dim3(PWG), 0, c->stream, P.w, P.wh);
    AGAL_HIP(c, hipGetLastError());
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    P.t = 0;
    return AGAL_OK;
}

// which: 0 master FP32 weights, 1 the FP16 copy (as FP32 values), 2 last gradient, 3 Adam m, 4 Adam v.
int predictor_get(Ctx *c, int which, float *out)
{
    Predictor &P = c->pred;
    if (!P.w) return AGAL_ERR_STATE;
    if (which == 1) {
        std::vector<_Float16> h(MLP_NPARAM);
        AGAL_HIP(c, hipMemcpyAsync(h.data(), P.wh, sizeof(_Float16) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
        AGAL_HIP(c, hipStreamSynchronize(c->stream));
        for (int q = 0; q < MLP_NPARAM; q++) out[q] = (float)h[q];
        return AGAL_OK;
    }
    const float *src = which == 0 ? P.w : which == 2 ? P.g : which == 3 ? P.m : which == 4 ? P.v : nullptr;
    if (!src) return AGAL_ERR_ARG;
    AGAL_HIP(c, hipMemcpyAsync(out, src, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

// Forward on device rows.
int predictor_forward_dev(Ctx *c, const float *X, int n, float *z3, float *p, float *Z1, float *Z2)
{
    if (n <= 0) return AGAL_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3(pgrid(n)), dim3(PWG), 0, c->stream, X, n, c->pred.wh, z3, p, Z1, Z2);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

// One Adam step on the n rows already in P.X / P.Y; the batch loss (mean) is returned in *loss.
int predictor_train_dev(Ctx *c, int n, float lr, float pos_weight, double *loss)
{
    Predictor &P = c->pred;
    if (n <= 0 || n > P.batch_cap) return AGAL_ERR_ARG;
    int st;
    if ((st = predictor_forward_dev(c, P.X, n, P.z3, nullptr, P.Z1, P.Z2))) return st;
    hipLaunchKernelGGL(k_mlp_output_grad, dim3(pgrid(n)), dim3(PWG), 0, c->stream, P.z3, P.Y, n, pos_weight, P.dZ3,
                       P.loss);
    hipLaunchKernelGGL(k_mlp_backward, dim3(pgrid(n)), dim3(PWG), 0, c->stream, n, P.wh, P.Z1, P.Z2, P.dZ3, P.dZ2,
                       P.dZ1);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(pgrid(MLP_NPARAM)), dim3(PWG), 0, c->stream, n, P.X, P.Z1, P.Z2, P.dZ3, P.dZ2,
                       P.dZ1, P.g);
    P.t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)P.t)), bc2 = (float)(1.0 - std::pow(0.999, (double)P.t));
    hipLaunchKernelGGL(k_adam, dim3(pgrid(MLP_NPARAM)), dim3(PWG), 0, c->stream, P.w, P.m, P.v, P.g, P.wh, lr, b1, b2,
                       1e-8f, bc1, bc2);
    AGAL_HIP(c, hipGetLastError());
    if (loss) {
        std::vector<float> h(n);
        AGAL_HIP(c, hipMemcpyAsync(h.data(), P.loss, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        AGAL_HIP(c, hipStreamSynchronize(c->stream));
        double s = 0;
        for (float x : h) s += x;
        *loss = s / n;
    }
    return AGAL_OK;
}

// Features of every block ID below id_max from the stored macroscopic fields.
int predictor_features(Ctx *c)
{
    Predictor &P = c->pred;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm <= 0) return AGAL_OK;
    const int cells = c->CELLS;
    const int grid = (int)(((long long)idm * cells + PWG - 1) / PWG);
    if (c->precision == AGAL_FP64) {
        if (c->DIM == 3)
            hipLaunchKernelGGL((k_block_stats<double, 3>), dim3(grid), dim3(PWG), 0, c->stream, idm, c->n_maxcells,
                               static_cast<const double *>(c->M), c->ref, c->level, c->cells_mask, c->dxf_vec[0], P.stats);
        else
            hipLaunchKernelGGL((k_block_stats<double, 2>), dim3(grid), dim3(PWG), 0, c->stream, idm, c->n_maxcells,
                               static_cast<const double *>(c->M), c->ref, c->level, c->cells_mask, c->dxf_vec[0], P.stats);
    } else {
        if (c->DIM == 3)
            hipLaunchKernelGGL((k_block_stats<float, 3>), dim3(grid), dim3(PWG), 0, c->stream, idm, c->n_maxcells,
                               static_cast<const float *>(c->M), c->ref, c->level, c->cells_mask, (float)c->dxf_vec[0],
                               P.stats);
        else
            hipLaunchKernelGGL((k_block_stats<float, 2>), dim3(grid), dim3(PWG), 0, c->stream, idm, c->n_maxcells,
                               static_cast<const float *>(c->M), c->ref, c->level, c->cells_mask, (float)c->dxf_vec[0],
                               P.stats);
    }
    const float Lx = (float)c->P.L_c, Ly = (float)(c->P.L_fy * Lx), Lz = (float)(c->P.L_fz * Lx);
    hipLaunchKernelGGL(k_block_distance, dim3(pgrid(idm)), dim3(PWG), 0, c->stream, idm, c->ref, c->level, c->blk_mask,
                       c->X, c->n_maxcblocks, c->DIM, (float)c->dxf_vec[0], Lx, Ly, Lz, c->geo_tri, c->geo_n_tri,
                       P.dist);
    hipLaunchKernelGGL(k_features, dim3(pgrid(idm)), dim3(PWG), 0, c->stream, idm, c->ref, c->level, c->nbr,
                       c->nbr_child, c->n_maxcblocks, c->QMAX, c->MAX_LEVELS, P.stats, P.dist, P.feat);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

int predictor_predict_blocks(Ctx *c)
{
    const int idm = c->id_max[c->MAX_LEVELS];
    return predictor_forward_dev(c, c->pred.feat, idm, nullptr, c->pred.prob, nullptr, nullptr);
}

int predictor_mark(Ctx *c, float thr, int *n_marked)
{
    Predictor &P = c->pred;
    const int idm = c->id_max[c->MAX_LEVELS];
    AGAL_HIP(c, hipMemsetAsync(P.counts, 0, sizeof(int) * 4, c->stream));
    if (idm > 0)
        hipLaunchKernelGGL(k_mark_predicted, dim3(pgrid(idm)), dim3(PWG), 0, c->stream, idm, P.prob, thr, c->ref,
                           c->level, c->nbr, c->n_maxcblocks, c->QMAX, c->MAX_LEVELS_INTERIOR, P.counts);
    AGAL_HIP(c, hipGetLastError());
    int h = 0;
    AGAL_HIP(c, hipMemcpyAsync(&h, P.counts, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    if (n_marked) *n_marked = h;
    return AGAL_OK;
}

int predictor_keep(Ctx *c, float thr)
{
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm > 0)
        hipLaunchKernelGGL(k_keep_predicted, dim3(pgrid(idm)), dim3(PWG), 0, c->stream, idm, c->pred.prob, thr, c->ref);
    AGAL_HIP(c, hipGetLastError());
    return AGAL_OK;
}

int predictor_audit(Ctx *c, int out[2])
{
    Predictor &P = c->pred;
    const int idm = c->id_max[c->MAX_LEVELS];
    AGAL_HIP(c, hipMemsetAsync(P.counts, 0, sizeof(int) * 4, c->stream));
    if (idm > 0)
        hipLaunchKernelGGL(k_audit, dim3(pgrid(idm)), dim3(PWG), 0, c->stream, idm, P.label, c->nbr_child, P.counts);
    AGAL_HIP(c, hipGetLastError());
    AGAL_HIP(c, hipMemcpyAsync(out, P.counts, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    return AGAL_OK;
}

// Training on stored snapshots: `n_rows` rows sampled uniformly (seeded) from the pairs
// (snapshot k, active block) with label 0 / 1, one Adam step per batch of `batch` rows.
int predictor_train_snapshots(Ctx *c, int n_snap, int batch, int steps, float lr, float pos_weight, uint64_t &rng,
                              double *last_loss)
{
    Predictor &P = c->pred;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (n_snap <= 0 || idm <= 0) return AGAL_OK;
    std::vector<int> lab(idm);
    AGAL_HIP(c, hipMemcpyAsync(lab.data(), P.label, sizeof(int) * idm, hipMemcpyDeviceToHost, c->stream));
    AGAL_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int> valid;
    for (int b = 0; b < idm; b++)
        if (lab[b] >= 0) valid.push_back(b);
    if (valid.empty()) return AGAL_OK;
    std::vector<int> ids(batch);
    int st;
    for (int s = 0; s < steps; s++) {
        const int k = (int)(mix64(rng) % (uint64_t)n_snap);
        for (int r = 0; r < batch; r++) ids[r] = valid[mix64(rng) % valid.size()];
        AGAL_HIP(c, hipMemcpyAsync(P.ids, ids.data(), sizeof(int) * batch, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_gather_batch, dim3(pgrid(batch)), dim3(PWG), 0, c->stream,
                           P.snap + (size_t)k * NFEAT * c->n_maxcblocks, P.label, P.ids, batch, P.X, P.Y);
        AGAL_HIP(c, hipGetLastError());
        if ((st = predictor_train_dev(c, batch, lr, pos_weight, s + 1 == steps ? last_loss : nullptr))) return st;
    }
    return AGAL_OK;
}

int predictor_snapshot(Ctx *c, int slot, int max_slots)
{
    Predictor &P = c->pred;
    int st;
    if (!P.snap) {
        if ((st = palloc(c, &P.snap, sizeof(float) * NFEAT * (size_t)c->n_maxcblocks * max_slots))) return st;
        P.snap_slots = max_slots;
    }
    if (slot >= P.snap_slots) return AGAL_ERR_ARG;
    const int idm = c->id_max[c->MAX_LEVELS];
    if (idm > 0)
        AGAL_HIP(c, hipMemcpyAsync(P.snap + (size_t)slot * NFEAT * c->n_maxcblocks, P.feat, sizeof(float) * NFEAT * idm,
                                   hipMemcpyDeviceToDevice, c->stream));
    return AGAL_OK;
}

}  // namespace agal
