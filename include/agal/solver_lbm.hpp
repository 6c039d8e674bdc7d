/*
 * agal/solver_lbm.hpp -- the lattice-Boltzmann solver of the header-library host API.
 *
 * agal::Solver_LBM implements the ten routines of agal::Solver (inc/solver.h:62-71) as the
 * reference's Solver_LBM does (inc/solver_lbm.h:82-217): the step, interpolation, averaging,
 * refresh, criterion and initial condition run on the GPU through the C ABI; the output
 * properties (S_ComputeProperties / S_ComputeOutputProperties, solver_lbm_compute.cu:4-163) are
 * computed on the host from the mesh's retrieved arrays; S_ReportForces reduces the force of the
 * configured S_FORCE_TYPE (0: momentum exchange, 1: control volume) on the GPU and prints / records
 * the drag and lift coefficients (solver_lbm_compute.cu:165-218).
 */
#ifndef AGAL_SOLVER_LBM_HPP
#define AGAL_SOLVER_LBM_HPP

#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "mesh.hpp"
#include "solver.hpp"

namespace agal {

class Solver_LBM : public Solver {
public:
    // Reference diameter and inflow speed of the force coefficients (the reference hard-codes
    // Dp = 1/64, uin = 0.05, solver_lbm_compute.cu:206-207).
    double Dp = 1.0 / 64.0, uin = 0.05;
    double last_force[3] = {0, 0, 0};
    double last_cd = 0, last_cl = 0;
    bool print = true;

    explicit Solver_LBM(Mesh *mesh_) : Solver(mesh_) {}

    int S_SetIC(int i_dev, int L) override
    {
        (void)i_dev;
        mesh->check(mesh->api.set_ic(mesh->ctx, L), "agal_set_ic");
        return 0;
    }
    // S_Interpolate (solver_lbm_kernels_comm.cu:141-155, 643-659): refresh of level L, then the
    // interpolation of its order.
    int S_Interpolate(int i_dev, int L, int var) override
    {
        (void)i_dev;
        mesh->check(mesh->api.interpolate(mesh->ctx, L, var), "agal_interpolate");
        return 0;
    }
    int S_Average(int i_dev, int L, int var) override
    {
        (void)i_dev;
        mesh->check(mesh->api.average(mesh->ctx, L, var), "agal_average");
        return 0;
    }
    // One coarse step of level L with the recursive sub-cycling of the finer levels
    // (solver_lbm_advance.cu:30-174).  `tmp` (the reference's per-kernel timers) is not used.
    int S_Advance(int i_dev, int L, double *tmp) override
    {
        (void)i_dev;
        (void)tmp;
        mesh->check(mesh->api.advance(mesh->ctx, L), "agal_advance");
        return 0;
    }
    int S_ComputeRefCriteria(int i_dev, int L, int var) override
    {
        (void)i_dev;
        (void)L;
        (void)var;
        mesh->check(mesh->api.ref_criteria(mesh->ctx), "agal_ref_criteria");
        return 0;
    }
    int S_RefreshVariables(int i_dev, int L, int var = 0) override
    {
        (void)i_dev;
        (void)var;
        mesh->check(mesh->api.refresh(mesh->ctx, L), "agal_refresh");
        return 0;
    }
    // The reference's geometry debug drawing (solver_lbm_kernels_debug.cu) has no counterpart.
    int S_Debug(int i_dev, int L, int var) override
    {
        (void)i_dev;
        (void)L;
        (void)var;
        return AGAL_ERR_NOT_BUILT;
    }

    // rho, u, v, w of the 4^d cells of block i_kap into out[cell + k*M_CBLOCK] (k = 0..3), from the
    // retrieved macroscopic fields.  i_Q (the reference's cell-block quadrant) is 0: one quadrant.
    int S_ComputeProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out) override
    {
        (void)i_dev;
        (void)i_Q;
        (void)dx_L;
        const HostPool &h = mesh->host;
        const int C = mesh->M_CBLOCK;
        const long long nc = h.n_cells();
        for (int t = 0; t < C; t++) {
            const long long cell = (long long)i_kap * C + t;
            out[t + 0 * C] = h.macros[0 * nc + cell];
            out[t + 1 * C] = h.macros[1 * nc + cell];
            out[t + 2 * C] = h.macros[2 * nc + cell];
            out[t + 3 * C] = mesh->N_DIM == 3 ? h.macros[3 * nc + cell] : 0.0;
        }
        return 0;
    }

    // S_ComputeProperties, then the vorticity (out rows 4..6) and the Q-criterion (row 7, 3D) from
    // central differences over the neighbouring cells (one-sided where the neighbour cell's block
    // does not exist), divided by dx_L (solver_lbm_compute.cu:31-163).
    int S_ComputeOutputProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out) override
    {
        S_ComputeProperties(i_dev, i_Q, i_kap, dx_L, out);
        const HostPool &h = mesh->host;
        const int C = mesh->M_CBLOCK, D = mesh->N_DIM;
        const long long nb = h.n_blocks, nc = h.n_cells();
        // Velocity index of the neighbour block across each face: -x, +x, -y, +y, -z, +z.
        const int face_dir[6] = {D == 3 ? 2 : 3, 1, D == 3 ? 4 : 4, D == 3 ? 3 : 2, 6, 5};
        auto vel = [&](long long cell, int a) { return a < D ? h.macros[(size_t)(1 + a) * nc + cell] : 0.0; };
        for (int t = 0; t < C; t++) {
            const int I[3] = {t % 4, (t / 4) % 4, D == 3 ? t / 16 : 0};
            double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // g[a][b] = d u_a / d x_b (cell units)
            const double uc[3] = {out[t + C], out[t + 2 * C], out[t + 3 * C]};
            for (int b = 0; b < 3; b++) {
                double um[3] = {0, 0, 0}, up[3] = {0, 0, 0};
                int nbm = N_SKIPID_, nbp = N_SKIPID_;
                if (b < D) {
                    for (int s = 0; s < 2; s++) {
                        int J[3] = {I[0], I[1], I[2]};
                        J[b] += s ? 1 : -1;
                        int blk = i_kap;
                        if (J[b] < 0 || J[b] > 3) blk = h.nbr[(size_t)face_dir[2 * b + s] * nb + i_kap];
                        J[b] = (J[b] + 4) % 4;
                        (s ? nbp : nbm) = blk;
                        if (blk > -1) {
                            const long long cell = (long long)blk * C + J[0] + 4 * J[1] + 16 * J[2];
                            for (int a = 0; a < 3; a++) (s ? up : um)[a] = vel(cell, a);
                        }
                    }
                }
                for (int a = 0; a < 3; a++) {
                    double d = 0.5 * (up[a] - um[a]);
                    if (nbm < 0) d = up[a] - uc[a];
                    if (nbp < 0) d = uc[a] - um[a];
                    g[a][b] = d;
                }
            }
            for (int r = 4; r < 8; r++) out[t + r * C] = 0.0;
            if (D == 3) out[t + 4 * C] = (g[2][1] - g[1][2]) / dx_L;
            if (D == 3) out[t + 5 * C] = (g[0][2] - g[2][0]) / dx_L;
            out[t + 6 * C] = (g[1][0] - g[0][1]) / dx_L;
            if (D == 3)
                out[t + 7 * C] = -0.5 *
                                 (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2] +
                                  2.0 * (g[0][1] * g[1][0] + g[0][2] * g[2][0] + g[2][1] * g[1][2])) /
                                 (dx_L * dx_L);
        }
        return 0;
    }

    // With START == 1 and L == -1: the momentum-exchange force on the geometry after the last
    // coarse step (FP64 reduction on the GPU), C_D = 2 F_x / (uin^2 Dp) and C_L = 2 F_y / (uin^2 Dp)
    // printed and appended to <P_DIR_NAME>forces.txt as "iter t C_D C_L C_D,sphere C_L,sphere" with
    // the sphere normalisation 8 F / (uin^2 pi Dp^2).
    int S_ReportForces(int i_dev, int L, int i, double t, int START) override
    {
        (void)i_dev;
        if (!(START == 1 && L == -1)) return 0;
        double F[3] = {0, 0, 0};
        mesh->check(mesh->api.forces(mesh->ctx, F), "agal_forces");
        for (int d = 0; d < 3; d++) last_force[d] = F[d];
        const double pi = 3.14159265358979323846;
        last_cd = 2.0 * F[0] / (uin * uin * Dp);
        last_cl = 2.0 * F[1] / (uin * uin * Dp);
        const double cds = 8.0 * F[0] / (uin * uin * (pi * Dp * Dp)), cls = 8.0 * F[1] / (uin * uin * (pi * Dp * Dp));
        if (print) std::printf("Report:\nCD: %g   %g\nCL: %g\n", last_cd, cds, last_cl);
        std::ofstream fo(mesh->output_dir + "forces.txt", std::ios::app);
        if (fo) fo << i << " " << t << " " << last_cd << " " << last_cl << " " << cds << " " << cls << "\n";
        return 0;
    }

private:
    static constexpr int N_SKIPID_ = AGAL_N_SKIPID;
};

}  // namespace agal

#endif  // AGAL_SOLVER_LBM_HPP
