// hostapi_check.cpp -- drives the header-library host API (include/agal/*.hpp) for the tests.
//
// Modes (each prints one JSON object on stdout):
//   parser   <input_dir>                    the parsed maps of <input_dir>/input.txt
//   text     <string>                       the parsed maps of an input given as text
//   geometry <stl> <primitive list> <l>     face counts / areas of the geometry front-end
//   run      <backend> <oracle.so> <outdir> <tag> <n_iter> <restart> <precision>
//            a lid-driven cavity through Parser -> Mesh -> Solver_LBM -> M_AdvanceLoop on the
//            HIP library ("gpu") or on a library with the same C ABI under the prefix oracle_
//            ("oracle"); writes <outdir>/<tag>_f.bin (final populations, physical order) and the
//            loop's .vthb output and restart file under <outdir>/
//   props    <backend> <oracle.so> <outdir>  vorticity / Q-criterion of prescribed linear fields
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/agal/mesh.hpp"
#include "../../include/agal/solver_lbm.hpp"

using namespace agal;

static void print_maps(const Parser &p)
{
    std::printf("{\"int\": {");
    bool first = true;
    for (auto &kv : p.params_int) {
        std::printf("%s\"%s\": %d", first ? "" : ", ", kv.first.c_str(), kv.second);
        first = false;
    }
    std::printf("}, \"dbl\": {");
    first = true;
    for (auto &kv : p.params_dbl) {
        std::printf("%s\"%s\": %.17g", first ? "" : ", ", kv.first.c_str(), kv.second);
        first = false;
    }
    std::printf("}, \"str\": {");
    first = true;
    for (auto &kv : p.params_str) {
        std::printf("%s\"%s\": \"%s\"", first ? "" : ", ", kv.first.c_str(), kv.second.c_str());
        first = false;
    }
    std::printf("}, \"invalid\": [");
    for (size_t k = 0; k < p.invalid_keys.size(); k++) std::printf("%s\"%s\"", k ? ", " : "", p.invalid_keys[k].c_str());
    std::printf("], \"found\": %s}\n", p.file_found ? "true" : "false");
}

static double area(const std::vector<float> &t, size_t k)
{
    const float *a = &t[k * 9], *b = a + 3, *c = a + 6;
    const double u[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
    const double v[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
    const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    return 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
}

static double max_edge(const std::vector<float> &t)
{
    double m = 0;
    for (size_t k = 0; k < t.size() / 9; k++)
        for (int e = 0; e < 3; e++) {
            const float *a = &t[k * 9 + 3 * e], *b = &t[k * 9 + 3 * ((e + 1) % 3)];
            const double l = std::sqrt((double)(b[0] - a[0]) * (b[0] - a[0]) + (double)(b[1] - a[1]) * (b[1] - a[1]) +
                                       (double)(b[2] - a[2]) * (b[2] - a[2]));
            m = std::max(m, l);
        }
    return m;
}

static double total_area(const std::vector<float> &t)
{
    double s = 0;
    for (size_t k = 0; k < t.size() / 9; k++) s += area(t, k);
    return s;
}

static Api backend(const std::string &which, const std::string &oracle_so)
{
    if (which == "gpu") return Api::Linked();
    return Api::Load(oracle_so, "oracle_");
}

static agal_params cavity_base(int precision)
{
    agal_params b;
    agal_params_default(&b);
    b.velocity_set = AGAL_D3Q19;
    b.precision = precision;
    b.bc_preset = AGAL_BC_LDC3D;
    b.bc_velocity = 0.05;
    return b;
}

static std::string cavity_input(const std::string &outdir, int n_iter, int restart)
{
    std::ostringstream s;
    // Re = U N / (v0 N) = 0.05 * 16 / (v0 * 16): v0 = U / Re, Re 50.
    s << "Nx 16\nMAX_LEVELS_WALL 1\nMAX_LEVELS_INTERIOR 2\nN_LEVEL_START 0\nv0 0.001\n"
      << "S_CRITERION 0\nN_REFINE_START -2\nN_REFINE_INC 1\nN_REFINE_MAX 1\nP_REFINE 20\n"
      << "N_ITER_TOTAL " << n_iter << "\nN_RESTART " << restart << "\nP_OUTPUT 30\nN_OUTPUT_START 0\n"
      << "N_PRINT_LEVELS_VTHB 2\nN_PROBE 1\nN_PROBE_FREQUENCY 10\nV_PROBE_TOL 1e-12\n"
      << "P_DIR_NAME " << outdir << "/\n";
    return s.str();
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "parser" && argc >= 3) {
            Parser p(std::string(argv[2]) + "/");
            print_maps(p);
            return 0;
        }
        if (mode == "text" && argc >= 3) {
            print_maps(Parser::FromString(argv[2]));
            return 0;
        }
        if (mode == "geometry" && argc >= 5) {
            Geometry g;
            g.G_ImportSTL_ASCII(argv[2]);
            const long long n_stl = g.n_faces();
            const double a0 = total_area(g.tri), e0 = max_edge(g.tri);
            const float l = (float)std::atof(argv[4]);
            g.G_RefineFaces_Length(l);
            const double a1 = total_area(g.tri), e1 = max_edge(g.tri);
            Geometry prim;
            prim.G_ImportBoundariesFromTextFile(argv[3]);
            Geometry box;
            box.G_AddPrism(0.1, 0.2, 0.3, 0.5, 0.6, 0.9);
            Geometry sph;
            sph.G_AddSphere(8, 12, 0.5, 0.5, 0.5, 0.25);
            std::printf("{\"stl_faces\": %lld, \"stl_area\": %.17g, \"stl_max_edge\": %.17g, \"refined_faces\": %lld, "
                        "\"refined_area\": %.17g, \"refined_max_edge\": %.17g, \"primitive_faces\": %lld, "
                        "\"primitive_area\": %.17g, \"box_faces\": %lld, \"box_area\": %.17g, \"sphere_faces\": %lld, "
                        "\"sphere_area\": %.17g}\n",
                        n_stl, a0, e0, g.n_faces(), a1, e1, prim.n_faces(), total_area(prim.tri), box.n_faces(),
                        total_area(box.tri), sph.n_faces(), total_area(sph.tri));
            return 0;
        }
        if (mode == "run" && argc >= 9) {
            const std::string outdir = argv[4], tag = argv[5];
            const int n_iter = std::atoi(argv[6]), restart = std::atoi(argv[7]), prec = std::atoi(argv[8]);
            Parser parser = Parser::FromString(cavity_input(outdir, n_iter, restart), outdir + "/");
            Mesh mesh(&parser, cavity_base(prec), 4096, backend(argv[2], argv[3]));
            Solver_LBM solver(&mesh);
            solver.print = false;
            const int last = mesh.M_AdvanceLoop();
            // Final state and a round trip of the whole pool through M_LoadToGPU.
            for (int L = 0; L < mesh.MAX_LEVELS; L++) solver.S_RefreshVariables(0, L);
            mesh.M_RetrieveFromGPU();
            HostPool before = mesh.host;
            Mesh copy(&parser, cavity_base(prec), 4096, backend(argv[2], argv[3]));
            copy.host = before;
            copy.M_LoadToGPU();
            copy.M_RetrieveFromGPU();
            const bool roundtrip = copy.host.f == before.f && copy.host.nbr == before.nbr &&
                                   copy.host.nbr_child == before.nbr_child && copy.host.ref == before.ref &&
                                   copy.host.id_set == before.id_set && copy.host.gap_set == before.gap_set &&
                                   copy.host.cell_mask == before.cell_mask && copy.host.X == before.X;
            std::ofstream fo(outdir + "/" + tag + "_f.bin", std::ios::binary);
            fo.write(reinterpret_cast<const char *>(before.f.data()), (std::streamsize)before.f.size());
            std::printf("{\"last_iter\": %d, \"n_blocks\": %lld, \"n_ids\": [", last, before.n_blocks);
            for (size_t L = 0; L < before.n_ids.size(); L++) std::printf("%s%d", L ? ", " : "", before.n_ids[L]);
            std::printf("], \"n_gaps\": %d, \"roundtrip\": %s, \"probes\": %zu, \"converged\": %s}\n", before.n_gaps,
                        roundtrip ? "true" : "false", mesh.id_set_probed.size(), mesh.converged ? "true" : "false");
            return 0;
        }
        if (mode == "props" && argc >= 5) {
            // A uniform 16^3 root grid; prescribed macroscopic fields replace the retrieved ones.
            Parser parser = Parser::FromString("Nx 16\nMAX_LEVELS_WALL 1\nMAX_LEVELS_INTERIOR 1\n", std::string(argv[4]) + "/");
            Mesh mesh(&parser, cavity_base(AGAL_FP64), 128, backend(argv[2], argv[3]));
            Solver_LBM solver(&mesh);
            solver.S_SetIC(0, 0);
            mesh.M_RetrieveFromGPU();
            HostPool &h = mesh.host;
            const long long nc = h.n_cells();
            const double dx = mesh.dxf_vec[0];
            // u = (a x + b y, c x - a y, e z): grad u = [[a, b, 0], [c, -a, 0], [0, 0, e]]
            const double a = 0.3, b = -0.2, c = 0.5, e = 0.1;
            for (long long blk = 0; blk < h.n_blocks; blk++)
                for (int t = 0; t < 64; t++) {
                    const double x = h.X[blk] + (t % 4 + 0.5) * dx, y = h.X[h.n_blocks + blk] + ((t / 4) % 4 + 0.5) * dx,
                                 z = h.X[2 * h.n_blocks + blk] + (t / 16 + 0.5) * dx;
                    const long long cell = blk * 64 + t;
                    h.macros[1 * nc + cell] = a * x + b * y;
                    h.macros[2 * nc + cell] = c * x - a * y;
                    h.macros[3 * nc + cell] = e * z;
                }
            double err_w = 0, err_q = 0;
            std::vector<double> out(64 * 8);
            const double wx = 0, wy = 0, wz = c - b;
            const double q = -0.5 * (a * a + a * a + e * e + 2.0 * (b * c));
            for (long long blk = 0; blk < h.n_blocks; blk++) {
                solver.S_ComputeOutputProperties(0, 0, (int)blk, dx, out.data());
                for (int t = 0; t < 64; t++) {
                    err_w = std::max(err_w, std::fabs(out[t + 4 * 64] - wx));
                    err_w = std::max(err_w, std::fabs(out[t + 5 * 64] - wy));
                    err_w = std::max(err_w, std::fabs(out[t + 6 * 64] - wz));
                    err_q = std::max(err_q, std::fabs(out[t + 7 * 64] - q));
                }
            }
            mesh.PrintVTHB(std::string(argv[4]) + "/", "props", 1, false);
            std::printf("{\"err_vorticity\": %.3g, \"err_q\": %.3g, \"q\": %.17g, \"wz\": %.17g}\n", err_w, err_q, q, wz);
            return 0;
        }
    } catch (const std::exception &ex) {
        std::printf("{\"error\": \"%s\"}\n", ex.what());
        return 1;
    }
    std::fprintf(stderr, "usage: see the header of hostapi_check.cpp\n");
    return 2;
}
