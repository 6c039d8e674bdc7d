/*
 * agal/mesh.hpp -- the mesh class of the header-library host API.
 *
 * agal::Mesh replaces Mesh<ufloat_t,ufloat_g_t,AP> (inc/mesh.h:137-765) on the host side.  The
 * block pool, the populations and every kernel live in the HIP library behind the C ABI; this
 * class owns one context, keeps the parser's parameters and the reference's method names
 * (M_AdvanceLoop, M_RefineAndCoarsenBlocks, M_Restart, M_Print_VTHB, ...), and holds host mirrors
 * of the device arrays after M_RetrieveFromGPU (inc/mesh.h:367-462), which the output, restart
 * and probe routines read.  The solver is attached as in the reference: a Solver constructed on
 * the mesh registers itself (inc/solver.h:73-80).
 */
#ifndef AGAL_MESH_HPP
#define AGAL_MESH_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../agal_capi.h"
#include "api.hpp"
#include "geometry.hpp"
#include "parser.hpp"
#include "solver.hpp"

namespace agal {

// Mesh-level selectors (inc/mesh.h:48-70).
constexpr int V_MESH_REF_NW_CASES = 0;
constexpr int V_MESH_REF_NW_GEOMETRY = 1;
constexpr int V_MESH_REF_UNIFORM = 2;
constexpr int V_MESH_REF_SOLUTION = 3;
constexpr int V_MESH_RESTART_SAVE = 0;
constexpr int V_MESH_RESTART_LOAD = 1;

// A failing C-ABI call.
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &what) : std::runtime_error(what + ": " + agal_status_string(st)), status(st) {}
};

// Host copies of the device arrays (M_RetrieveFromGPU), for block IDs [0, n_blocks).
struct HostPool {
    agal_sizes s{};
    long long n_blocks = 0;       // id_max over all levels
    std::vector<int> n_ids, id_max;
    int n_gaps = 0;
    std::vector<std::vector<int>> id_set;
    std::vector<int> gap_set;
    std::vector<int> nbr, nbr_child;             // n_q_max rows x n_blocks (velocity order)
    std::vector<int> ref, level, onb, blk_mask;  // n_blocks
    std::vector<float> X;                        // n_dim rows x n_blocks (lower corner of a block)
    std::vector<int> cell_mask;                  // n_blocks * cells_per_block
    std::vector<double> macros;                  // rho, u, v, w rows x n_cells
    std::vector<unsigned char> f;                // n_q rows x n_cells, physical order, `fsize` bytes each
    int fsize = 4;                               // 8 for FP64 pools, else 4 (decoded FP32)

    long long n_cells() const { return n_blocks * s.cells_per_block; }
    double F(int q, long long cell) const
    {
        const size_t i = (size_t)q * n_cells() + cell;
        if (fsize == 8) {
            double v;
            std::memcpy(&v, &f[i * 8], 8);
            return v;
        }
        float v;
        std::memcpy(&v, &f[i * 4], 4);
        return v;
    }
};

class Mesh {
public:
    Parser *parser = nullptr;
    agal_params P{};
    Api api;
    agal_ctx *ctx = nullptr;
    long long n_maxcblocks = 0;
    int device = 0;
    Solver *solver = nullptr;
    Geometry *geometry = nullptr;

    // Parameters of the run (mesh_init.cu:30-68).
    int MAX_LEVELS = 1, MAX_LEVELS_WALL = 1, MAX_LEVELS_INTERIOR = 1, N_LEVEL_START = 0;
    int N_DIM = 3, N_Q = 19, N_Q_max = 27, M_CBLOCK = 64, N_CHILDREN = 8;
    int N_ITER_TOTAL = 0, P_REFINE = 1, N_RESTART = 0, P_OUTPUT = 1, N_OUTPUT_START = 0;
    int N_PRINT_LEVELS_VTHB = 0, N_PROBE = 0, N_PROBE_DENSITY = 4, N_PROBE_FREQUENCY = 32;
    int N_PROBE_FORCE = 0, N_PROBE_F_FREQUENCY = 16, N_PROBE_F_START = 0;
    int N_PROBE_AVE = 0, N_PROBE_AVE_FREQUENCY = 1, N_PROBE_AVE_START = 0;
    double V_PROBE_TOL = 1e-4;
    std::string output_dir;
    double dx = 0;                  // root spacing (the library's: (float)L_c / Nx)
    std::vector<double> dxf_vec;    // spacing per level
    int Nxi[3] = {0, 0, 0};
    bool verbose = true;

    HostPool host;
    // Probes (mesh_init.cu:388-393, mesh_output.cu:112-209): root blocks on a sparse lattice,
    // their previous velocities, and the running mean of rho, u, v, w on the root level.
    std::vector<int> id_set_probed;
    std::vector<double> U_probed_tn, U_mean;
    int n_iters_ave = 0;
    bool converged = false;

    // `base` carries what the reference fixes at compile time (velocity set, precision, boundary
    // hook, initial state; agal_params_default() gives the parser defaults); the parser's file
    // overrides the rest.  `capacity` is the block pool (the reference sizes it from free memory,
    // mesh_init.cu:155-169).
    Mesh(Parser *parser_, const agal_params &base, long long capacity, Api api_ = Api::Linked(), int device_ = 0)
        : parser(parser_), api(api_), n_maxcblocks(capacity), device(device_)
    {
        P = parser ? parser->ToParams(base) : base;
        if (parser) {
            N_ITER_TOTAL = parser->Int("N_ITER_TOTAL");
            N_RESTART = parser->Int("N_RESTART");
            P_OUTPUT = std::max(1, parser->Int("P_OUTPUT"));
            N_OUTPUT_START = parser->Int("N_OUTPUT_START");
            N_PRINT_LEVELS_VTHB = parser->Int("N_PRINT_LEVELS_VTHB");
            N_PROBE = parser->Int("N_PROBE");
            N_PROBE_DENSITY = std::max(1, parser->Int("N_PROBE_DENSITY"));
            N_PROBE_FREQUENCY = std::max(1, parser->Int("N_PROBE_FREQUENCY"));
            V_PROBE_TOL = parser->Dbl("V_PROBE_TOL");
            N_PROBE_FORCE = parser->Int("N_PROBE_FORCE");
            N_PROBE_F_FREQUENCY = std::max(1, parser->Int("N_PROBE_F_FREQUENCY"));
            N_PROBE_F_START = parser->Int("N_PROBE_F_START");
            N_PROBE_AVE = parser->Int("N_PROBE_AVE");
            N_PROBE_AVE_FREQUENCY = std::max(1, parser->Int("N_PROBE_AVE_FREQUENCY"));
            N_PROBE_AVE_START = parser->Int("N_PROBE_AVE_START");
            output_dir = parser->Str("P_DIR_NAME");
        }
        M_Init();
    }

    ~Mesh() { M_Dest(); }
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;

    void check(int st, const char *what) const
    {
        if (st != AGAL_OK) throw Error(st, what);
    }

    // Mesh::M_Init (mesh_init.cu:10-567): the context (pool, root grid, gap stack) and the
    // host-side parameters.
    int M_Init()
    {
        check(api.create(&P, n_maxcblocks, device, &ctx), "agal_create");
        agal_sizes s{};
        check(api.get_sizes(ctx, &s), "agal_get_sizes");
        MAX_LEVELS = s.max_levels;
        MAX_LEVELS_WALL = P.MAX_LEVELS_WALL;
        MAX_LEVELS_INTERIOR = P.MAX_LEVELS_INTERIOR;
        N_LEVEL_START = std::min(P.N_LEVEL_START, MAX_LEVELS - 1);
        if (MAX_LEVELS == 1) N_LEVEL_START = 0;
        N_DIM = s.n_dim;
        N_Q = s.n_q;
        N_Q_max = s.n_q_max;
        M_CBLOCK = s.cells_per_block;
        N_CHILDREN = s.n_children;
        P_REFINE = std::max(1, P.P_REFINE);
        const float Lx = (float)P.L_c;
        dx = (double)(Lx / (float)P.Nx);
        dxf_vec.resize(MAX_LEVELS);
        for (int L = 0; L < MAX_LEVELS; L++) dxf_vec[L] = dx / std::pow(2.0, L);
        Nxi[0] = P.Nx;
        Nxi[1] = (int)(P.Nx * ((float)(P.L_fy * Lx) / Lx));
        Nxi[2] = N_DIM == 2 ? 1 : (int)(P.Nx * ((float)(P.L_fz * Lx) / Lx));
        // Probed root blocks (mesh_init.cu:388-393): every N_PROBE_DENSITY-th root block per axis.
        const int bx = Nxi[0] / 4, by = Nxi[1] / 4, bz = N_DIM == 2 ? 1 : Nxi[2] / 4;
        id_set_probed.clear();
        for (int k = 0, id = 0; k < bz; k++)
            for (int j = 0; j < by; j++)
                for (int i = 0; i < bx; i++, id++)
                    if (i % N_PROBE_DENSITY == 0 && j % N_PROBE_DENSITY == 0 && (N_DIM == 2 || k % N_PROBE_DENSITY == 0))
                        id_set_probed.push_back(id);
        U_probed_tn.assign((size_t)N_DIM * id_set_probed.size() * M_CBLOCK, 0.0);
        U_mean.assign((size_t)(N_DIM + 1) * s.n_coarsecblocks * M_CBLOCK, 0.0);
        return 0;
    }

    int M_Dest()
    {
        if (ctx) api.destroy(ctx);
        ctx = nullptr;
        return 0;
    }

    agal_sizes Sizes() const
    {
        agal_sizes s{};
        check(api.get_sizes(ctx, &s), "agal_get_sizes");
        return s;
    }

    int M_AddGeometry(Geometry *g)
    {
        geometry = g;
        if (g && N_DIM == 3) check(api.set_geometry(ctx, g->tri.data(), g->n_faces()), "agal_set_geometry");
        return 0;
    }
    int M_AddSolver(Solver *s)
    {
        solver = s;
        return 0;
    }

    // ---- forwarding to the solver (mesh_comm.cu:4-26, mesh_advance.cu:12-18) --------------------
    int M_Interpolate(int i_dev, int L, int var) { return need_solver()->S_Interpolate(i_dev, L, var); }
    int M_Average(int i_dev, int L, int var) { return need_solver()->S_Average(i_dev, L, var); }
    int M_Advance(int i_dev, int L, double *tmp) { return need_solver()->S_Advance(i_dev, L, tmp); }
    int M_ComputeProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out)
    {
        return need_solver()->S_ComputeProperties(i_dev, i_Q, i_kap, dx_L, out);
    }
    int M_ComputeOutputProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out)
    {
        return need_solver()->S_ComputeOutputProperties(i_dev, i_Q, i_kap, dx_L, out);
    }
    int M_ReportForces(int i_dev, int L, int i, double t, int START)
    {
        return need_solver()->S_ReportForces(i_dev, L, i, t, START);
    }

    // M_ComputeRefCriteria (mesh_criterion.cu:90-119): uniform marks of level L, or the solver's
    // solution criterion.
    int M_ComputeRefCriteria(int i_dev, int L, int var)
    {
        if (var == V_MESH_REF_UNIFORM) check(api.mark_uniform(ctx, L), "agal_mark_uniform");
        if (var == V_MESH_REF_SOLUTION) return need_solver()->S_ComputeRefCriteria(i_dev, L, 1);
        return 0;
    }

    // M_RefineAndCoarsenBlocks (mesh_amr.cu:2161-2744).  Returns AGAL_ERR_POOL_EXHAUSTED when the
    // refinement was cancelled for lack of free block IDs (the reference goes on the same way).
    int M_RefineAndCoarsenBlocks(int var)
    {
        (void)var;
        int counts[4];
        const int st = api.refine_coarsen(ctx, counts);
        if (st != AGAL_ERR_POOL_EXHAUSTED) check(st, "agal_refine_coarsen");
        return st;
    }
    int M_FreezeRefinedCells(int var)
    {
        check(api.freeze_refined(ctx, var), "agal_freeze_refined");
        return 0;
    }
    // M_Advance_RefineNearWall (mesh_advance.cu:53-121): uniform levels below N_LEVEL_START, the
    // geometry's near-wall levels, the initial condition of every level, the freeze.
    int M_Advance_RefineNearWall()
    {
        const int st = api.refine_near_wall(ctx);
        if (st != AGAL_ERR_POOL_EXHAUSTED) check(st, "agal_refine_near_wall");
        return st;
    }

    // M_Advance_Step (mesh_advance.cu:155-182): iter_mult coarse steps of the root level.
    int M_Advance_Step(int i, int iter_s, int iter_mult)
    {
        (void)i;
        (void)iter_s;
        for (int j = 0; j < iter_mult; j++) M_Advance(0, N_LEVEL_START, nullptr);
        return 0;
    }
    // M_Advance_RefineWithSolution (mesh_advance.cu:184-243).
    int M_Advance_RefineWithSolution(int i, int iter_s)
    {
        (void)iter_s;
        if (!(MAX_LEVELS_INTERIOR > 1 && (N_PROBE_AVE == 0 || (i + 1) <= N_PROBE_AVE_START))) return 0;
        int counts[4];
        const int st = api.refine_with_solution(ctx, counts);
        if (st != AGAL_ERR_POOL_EXHAUSTED) check(st, "agal_refine_with_solution");
        return st;
    }

    // M_AdvanceLoop (mesh_advance.cu:385-466).  Stretches of coarse steps with no host-side event
    // between them run through the library's own loop (steps plus the solution-based rebuilds
    // every P_REFINE steps, recorded HIP graphs); forces, probes, averages and output happen at
    // the reference's iterations.  Returns the last iteration index run.
    int M_AdvanceLoop()
    {
        const int iter_mult = 1 << N_LEVEL_START;
        int iter_s = 0;
        M_Advance_RefineNearWall();
        if (N_RESTART == 1) M_Advance_LoadRestartFile(iter_s);
        if (N_ITER_TOTAL == 1) {
            need_solver()->S_SetIC(0, 0);
            M_Advance_PrintData(0, 0);
            return 0;
        }
        const int end = iter_s + N_ITER_TOTAL;
        auto start_event = [&](int i) { return N_PROBE_FORCE == 1 && i % N_PROBE_F_FREQUENCY == 0 && i > N_PROBE_F_START; };
        auto end_event = [&](int i) {
            return ((i + 1) % P_OUTPUT == 0 && i > iter_s + N_OUTPUT_START) ||
                   (N_PROBE == 1 && i % N_PROBE_FREQUENCY == 0 && i > 0) ||
                   (N_PROBE_AVE == 1 && i % N_PROBE_AVE_FREQUENCY == 0 && i > N_PROBE_AVE_START);
        };
        int i = iter_s;
        while (i < end && !converged) {
            if (start_event(i)) M_Advance_PrintForces(i, iter_s, 1);
            int j = i;  // run steps i..j, then the events at the end of j
            while (j + 1 < end && !end_event(j) && !start_event(j + 1)) j++;
            if (iter_mult == 1) {
                double sms = 0, rms = 0;
                const int st = api.advance_loop(ctx, i, j - i + 1, &sms, &rms);
                if (st != AGAL_ERR_POOL_EXHAUSTED) check(st, "agal_advance_loop");
            } else {
                for (int k = i; k <= j; k++) {
                    M_Advance_Step(k, iter_s, iter_mult);
                    if (MAX_LEVELS > 1 && MAX_LEVELS != N_LEVEL_START + 1 && (k + 1) % P_REFINE == 0 && k > 0)
                        M_Advance_RefineWithSolution(k, iter_s);
                }
            }
            if (N_PROBE == 1 && j % N_PROBE_FREQUENCY == 0 && j > 0) M_Advance_Probe(j, iter_s);
            if (N_PROBE_AVE == 1 && j % N_PROBE_AVE_FREQUENCY == 0 && j > N_PROBE_AVE_START) M_Advance_ProbeAverage(j, iter_s);
            if ((j + 1) % P_OUTPUT == 0 && j > iter_s + N_OUTPUT_START) M_Advance_PrintData(j, iter_s);
            i = j + 1;
        }
        return i - 1;
    }

    int M_Advance_LoadRestartFile(int &iter_s)
    {
        if (M_Restart(0, V_MESH_RESTART_LOAD, &iter_s) != 0) throw std::runtime_error("restart load failed");
        M_LoadToGPU();
        iter_s += 1;
        return 0;
    }

    // M_Advance_PrintForces (mesh_advance.cu:347-364).
    int M_Advance_PrintForces(int i, int iter_s, int START)
    {
        (void)iter_s;
        return M_ReportForces(0, -1, i, (double)i * dxf_vec[N_LEVEL_START], START);
    }

    // M_Advance_PrintData (mesh_advance.cu:295-345): grid average, interpolation to the ghost
    // cells, macroscopic refresh, retrieve, the restart file at the end of the run, output.
    int M_Advance_PrintData(int i, int iter_s)
    {
        PrepareOutputFields();
        M_RetrieveFromGPU();
        if ((i + 1) % (iter_s + N_ITER_TOTAL) == 0 && i > iter_s) M_Restart(0, V_MESH_RESTART_SAVE, &i);
        if (N_PRINT_LEVELS_VTHB > 0) M_Print_VTHB(0, i);
        return 0;
    }

    void PrepareOutputFields()
    {
        Solver *s = need_solver();
        if (MAX_LEVELS > 1 && N_Q > 1) {
            for (int L = MAX_LEVELS - 2; L >= 0; L--) s->S_Average(0, L, AGAL_AVERAGE_GRID);
            for (int L = 0; L < MAX_LEVELS - 1; L++) s->S_Interpolate(0, L, AGAL_INTERP_INTERFACE);
        }
        for (int L = 0; L < MAX_LEVELS; L++) s->S_RefreshVariables(0, L);
    }

    // M_Advance_Probe (mesh_advance.cu:245-266) with M_CheckConvergence (mesh_output.cu:157-209).
    int M_Advance_Probe(int i, int iter_s)
    {
        (void)i;
        (void)iter_s;
        for (int L = 0; L < MAX_LEVELS; L++) need_solver()->S_RefreshVariables(0, L);
        M_RetrieveFromGPU();
        const double ctol = M_CheckConvergence(0);
        if (ctol < V_PROBE_TOL) converged = true;
        return 0;
    }

    // M_Advance_ProbeAverage (mesh_advance.cu:268-290) with M_UpdateMeanVelocities
    // (mesh_output.cu:112-155).
    int M_Advance_ProbeAverage(int i, int iter_s)
    {
        (void)i;
        (void)iter_s;
        PrepareOutputFields();
        M_RetrieveFromGPU();
        M_UpdateMeanVelocities(0, n_iters_ave);
        n_iters_ave++;
        return 0;
    }

    // Relative change of the probed root blocks' velocities since the last check:
    // sqrt(sum |u - u_prev|^2 / sum |u|^2); stores the current velocities (mesh_output.cu:157-209).
    double M_CheckConvergence(int i_dev)
    {
        std::vector<double> out((size_t)M_CBLOCK * 8);
        double sum = 0, norm = 0;
        const size_t np = id_set_probed.size();
        for (size_t k = 0; k < np; k++) {
            M_ComputeProperties(i_dev, 0, id_set_probed[k], dxf_vec[0], out.data());
            for (int t = 0; t < M_CBLOCK; t++)
                for (int d = 0; d < N_DIM; d++) {
                    const double u = out[t + (1 + d) * M_CBLOCK];
                    double &prev = U_probed_tn[k * M_CBLOCK + t + d * np * M_CBLOCK];
                    sum += (u - prev) * (u - prev);
                    norm += u * u;
                    prev = u;
                }
        }
        return std::sqrt(sum / norm);  // NaN while the probes see no flow: never below the tolerance
    }

    // Running mean of rho, u, v, (w) on the root level over the N_iters_ave + 1 samples so far.
    int M_UpdateMeanVelocities(int i_dev, int N_iters_ave)
    {
        std::vector<double> out((size_t)M_CBLOCK * 8);
        const long long n0 = host.n_ids.empty() ? 0 : host.n_ids[0];
        for (long long kap = 0; kap < n0; kap++) {
            const int b = host.id_set[0][kap];
            M_ComputeProperties(i_dev, 0, b, dxf_vec[0], out.data());
            for (int t = 0; t < M_CBLOCK; t++)
                for (int d = 0; d < N_DIM + 1; d++) {
                    double &m = U_mean[(size_t)b * M_CBLOCK + t + (size_t)d * n0 * M_CBLOCK];
                    m = ((double)N_iters_ave * m + out[t + d * M_CBLOCK]) / ((double)N_iters_ave + 1.0);
                }
        }
        return 0;
    }

    // ---- data exchange (mesh_output.cu:4-82) --------------------------------------------------
    int M_RetrieveFromGPU()
    {
        HostPool &h = host;
        h.s = Sizes();
        const int ML = h.s.max_levels;
        h.n_blocks = h.s.id_max[ML];
        h.n_ids.assign(h.s.n_ids, h.s.n_ids + ML);
        h.id_max.assign(h.s.id_max, h.s.id_max + ML + 1);
        h.n_gaps = h.s.n_gaps;
        h.id_set.assign(ML, {});
        for (int L = 0; L < ML; L++) {
            h.id_set[L].resize(h.n_ids[L]);
            if (h.n_ids[L] > 0) check(api.download_id_set(ctx, L, h.id_set[L].data(), h.n_ids[L]), "download_id_set");
        }
        h.gap_set.resize(h.n_gaps);
        if (h.n_gaps > 0) check(api.download_gap_set(ctx, h.gap_set.data(), h.n_gaps), "download_gap_set");
        const long long nb = h.n_blocks, nc = h.n_cells();
        auto ints = [&](int which, std::vector<int> &v, int rows) {
            v.resize((size_t)rows * nb);
            if (nb > 0) check(api.download_block_ints(ctx, which, v.data(), nb, nb), "download_block_ints");
        };
        ints(AGAL_BLK_NBR, h.nbr, h.s.n_q_max);
        ints(AGAL_BLK_NBR_CHILD, h.nbr_child, h.s.n_q_max);
        ints(AGAL_BLK_REF, h.ref, 1);
        ints(AGAL_BLK_LEVEL, h.level, 1);
        ints(AGAL_BLK_ONB, h.onb, 1);
        ints(AGAL_BLK_MASK, h.blk_mask, 1);
        h.X.resize((size_t)h.s.n_dim * nb);
        if (nb > 0) check(api.download_coords(ctx, h.X.data(), nb, nb), "download_coords");
        h.cell_mask.resize(nc);
        if (nb > 0) check(api.download_cell_masks(ctx, h.cell_mask.data(), nb), "download_cell_masks");
        h.fsize = P.precision == AGAL_FP64 ? 8 : 4;
        h.f.resize((size_t)h.s.n_q * nc * h.fsize);
        if (nb > 0) check(api.download_populations(ctx, h.f.data(), nb, nc), "download_populations");
        h.macros.assign((size_t)4 * nc, 0.0);
        if (nb > 0) {
            if (h.fsize == 8) {
                check(api.download_macros(ctx, h.macros.data(), nb, nc), "download_macros");
            } else {
                std::vector<float> m((size_t)4 * nc);
                check(api.download_macros(ctx, m.data(), nb, nc), "download_macros");
                std::copy(m.begin(), m.end(), h.macros.begin());
            }
        }
        return 0;
    }

    // M_LoadToGPU (mesh_output.cu:4-42): the whole host pool (topology, masks, ID sets, gap stack,
    // populations) into the context.
    int M_LoadToGPU()
    {
        const HostPool &h = host;
        std::vector<int> sets;
        for (const auto &v : h.id_set) sets.insert(sets.end(), v.begin(), v.end());
        agal_pool pool{};
        pool.n_blocks = h.n_blocks;
        pool.nbr = h.nbr.data();
        pool.nbr_child = h.nbr_child.data();
        pool.ref = h.ref.data();
        pool.level = h.level.data();
        pool.onb = h.onb.data();
        pool.blk_mask = h.blk_mask.data();
        pool.coords = h.X.data();
        pool.cell_mask = h.cell_mask.data();
        pool.n_ids = h.n_ids.data();
        pool.id_sets = sets.data();
        pool.n_gaps = h.n_gaps;
        pool.gap_set = h.gap_set.data();
        pool.populations = h.f.empty() ? nullptr : h.f.data();
        check(api.import_pool(ctx, &pool), "agal_import_pool");
        return 0;
    }

    // ---- restart (mesh_restart.cu:8-225) --------------------------------------------------------
    // The reference's binary layout, little-endian, in its order: sizes (n_maxcblocks, n_ids[ML+1],
    // n_gaps, id_max[ML+1], cblocks_id_max, cells_id_max as 8 bytes, n_ids_probed, iter), the ID
    // sets and the gap stack, cell masks, block coordinates (compute type), block masks, the
    // neighbour and child tables interleaved per direction, onb, ref, level, the populations
    // (row p holds slot p, i.e. the population of direction pb(p)), the probed velocities and the
    // root-level means.  The host pool must be current (M_RetrieveFromGPU) before a save.
    int M_Restart(int i_dev, int var, int *iter)
    {
        (void)i_dev;
        const std::string filename = output_dir + "restart.bin";
        if (var == V_MESH_RESTART_SAVE) return WriteRestart(filename, iter ? *iter : 0);
        if (var == V_MESH_RESTART_LOAD) return ReadRestart(filename, iter);
        return 1;
    }

    int WriteRestart(const std::string &filename, int iter) const;
    int ReadRestart(const std::string &filename, int *iter);

    // ---- output ---------------------------------------------------------------------------------
    // M_Print_VTHB (mesh_print_vthb.cu:24-242): a vtkOverlappingAMR file (out_<iter+1>.vthb) and
    // one image-data piece per block of the first N_PRINT_LEVELS_VTHB levels, written as VTK XML
    // without VTK.  The host pool must be current.
    int M_Print_VTHB(int i_dev, int iter);

    int PrintVTHB(const std::string &dir, const std::string &name, int n_levels, bool binary = true);

    Solver *need_solver() const
    {
        if (!solver) throw std::runtime_error("no solver attached to the mesh");
        return solver;
    }

    // Pre-collision populations in slot order (the reference's storage): slot p = direction pb(p).
    // D2Q9 (inc/util.h:94): 1 +x, 2 +y, 3 -x, 4 -y, 5..8 the diagonals counter-clockwise, so the
    // opposite is two steps round each group; D3Q19 / D3Q27 (util.h:95) pair 2k-1 with 2k.
    static int pb(int Q, int p)
    {
        if (p == 0) return 0;
        if (Q == 9) return p <= 4 ? (p + 1) % 4 + 1 : (p - 3) % 4 + 5;
        return (p % 2 == 1) ? p + 1 : p - 1;
    }
};

inline Solver::Solver(Mesh *mesh_) : mesh(mesh_)
{
    if (mesh) mesh->M_AddSolver(this);
}

}  // namespace agal

#include "output.hpp"

#endif  // AGAL_MESH_HPP
