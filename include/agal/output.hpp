/*
 * agal/output.hpp -- restart files and vtkOverlappingAMR output of agal::Mesh (included by
 * agal/mesh.hpp).
 *
 * Restart: the reference's binary layout (src/mesh/mesh_restart.cu:8-225), so a file written by
 * one run restarts another with the same parameters.  Output: the .vthb / .vti XML files VTK's
 * vtkXMLUniformGridAMRWriter produces for the reference's M_Print_VTHB (mesh_print_vthb.cu:24-242),
 * written directly (no VTK in this build): one <Block> per level with its spacing, one image-data
 * piece per block with the reference's cell arrays (Ref ID, On Boundary?, Cell Mask, Block Mask,
 * Density, Velocity, Vorticity, Q-Criterion).
 */
#ifndef AGAL_OUTPUT_HPP
#define AGAL_OUTPUT_HPP

#include <sys/stat.h>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace agal {

namespace detail {

template <typename T>
void put(std::ostream &o, const T &v)
{
    o.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <typename T>
void put_n(std::ostream &o, const T *p, size_t n)
{
    if (n) o.write(reinterpret_cast<const char *>(p), (std::streamsize)(sizeof(T) * n));
}
template <typename T>
bool get(std::istream &i, T &v)
{
    return static_cast<bool>(i.read(reinterpret_cast<char *>(&v), sizeof(T)));
}
template <typename T>
bool get_n(std::istream &i, T *p, size_t n)
{
    return n == 0 || static_cast<bool>(i.read(reinterpret_cast<char *>(p), (std::streamsize)(sizeof(T) * n)));
}

inline std::string base64(const unsigned char *d, size_t n)
{
    static const char *tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((n + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = (uint32_t)d[i] << 16 | (uint32_t)d[i + 1] << 8 | d[i + 2];
        out += tab[v >> 18 & 63];
        out += tab[v >> 12 & 63];
        out += tab[v >> 6 & 63];
        out += tab[v & 63];
    }
    if (i < n) {
        uint32_t v = (uint32_t)d[i] << 16;
        if (i + 1 < n) v |= (uint32_t)d[i + 1] << 8;
        out += tab[v >> 18 & 63];
        out += tab[v >> 12 & 63];
        out += (i + 1 < n) ? tab[v >> 6 & 63] : '=';
        out += '=';
    }
    return out;
}

// A VTK XML DataArray of doubles: inline "binary" (base64 of a UInt64 byte count and the data) or
// ascii.
inline void data_array(std::ostream &o, const char *name, int comps, const std::vector<double> &v, bool binary)
{
    o << "        <DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"" << comps << "\" format=\""
      << (binary ? "binary" : "ascii") << "\">\n          ";
    if (binary) {
        std::vector<unsigned char> buf(8 + v.size() * 8);
        const uint64_t nb = v.size() * 8;
        std::memcpy(buf.data(), &nb, 8);
        if (!v.empty()) std::memcpy(buf.data() + 8, v.data(), nb);
        o << base64(buf.data(), buf.size());
    } else {
        char tmp[32];
        for (size_t k = 0; k < v.size(); k++) {
            std::snprintf(tmp, sizeof(tmp), "%.17g", v[k]);
            o << tmp << (k + 1 < v.size() ? " " : "");
        }
    }
    o << "\n        </DataArray>\n";
}

inline void make_dir(const std::string &d)
{
    if (!d.empty()) ::mkdir(d.c_str(), 0755);
}

}  // namespace detail

inline int Mesh::WriteRestart(const std::string &filename, int iter) const
{
    using namespace detail;
    const HostPool &h = host;
    if (h.s.max_levels != MAX_LEVELS) throw std::runtime_error("restart save: the host pool is not current");
    std::ofstream o(filename, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + filename);
    const int ML = MAX_LEVELS;
    const int cblocks_id_max = (int)h.n_blocks;
    const int64_t cells_id_max = (int64_t)h.n_blocks * M_CBLOCK;
    const int n_probed = (int)id_set_probed.size();
    std::vector<int> n_ids(h.n_ids);
    n_ids.push_back(0);
    put(o, (int)n_maxcblocks);
    put_n(o, n_ids.data(), ML + 1);
    put(o, h.n_gaps);
    put_n(o, h.id_max.data(), ML + 1);
    put(o, cblocks_id_max);
    put(o, cells_id_max);
    put(o, n_probed);
    put(o, iter);
    for (int L = 0; L < ML; L++) put_n(o, h.id_set[L].data(), h.id_set[L].size());
    put_n(o, h.gap_set.data(), h.gap_set.size());
    put_n(o, h.cell_mask.data(), (size_t)cells_id_max);
    const bool dbl = P.precision == AGAL_FP64;
    auto put_real = [&](double v) {
        if (dbl)
            put(o, v);
        else
            put(o, (float)v);
    };
    for (int d = 0; d < N_DIM; d++)
        for (int b = 0; b < cblocks_id_max; b++) put_real((double)h.X[(size_t)d * h.n_blocks + b]);
    put_n(o, h.blk_mask.data(), (size_t)cblocks_id_max);
    for (int p = 0; p < N_Q_max; p++) {
        put_n(o, &h.nbr[(size_t)p * h.n_blocks], (size_t)cblocks_id_max);
        put_n(o, &h.nbr_child[(size_t)p * h.n_blocks], (size_t)cblocks_id_max);
    }
    put_n(o, h.onb.data(), (size_t)cblocks_id_max);
    put_n(o, h.ref.data(), (size_t)cblocks_id_max);
    put_n(o, h.level.data(), (size_t)cblocks_id_max);
    for (int p = 0; p < N_Q; p++) {
        const int q = pb(N_Q, p);  // slot p holds the pre-collision population of direction pb(p)
        const unsigned char *row = &h.f[(size_t)q * cells_id_max * h.fsize];
        if (h.fsize == 8 && dbl)
            put_n(o, row, (size_t)cells_id_max * 8);
        else if (h.fsize == 4 && !dbl)
            put_n(o, row, (size_t)cells_id_max * 4);
        else
            for (int64_t c = 0; c < cells_id_max; c++) put_real(h.F(q, c));
    }
    for (int d = 0; d < N_DIM; d++)
        for (int k = 0; k < n_probed * M_CBLOCK; k++) put_real(U_probed_tn[(size_t)d * n_probed * M_CBLOCK + k]);
    const int n0 = h.n_ids.empty() ? 0 : h.n_ids[0];
    for (int d = 0; d < N_DIM + 1; d++)
        for (int k = 0; k < n0 * M_CBLOCK; k++) put_real(U_mean[(size_t)d * n0 * M_CBLOCK + k]);
    return o.good() ? 0 : 1;
}

// Returns 1 (as the reference) when the file's pool does not fit this mesh's capacity.
inline int Mesh::ReadRestart(const std::string &filename, int *iter)
{
    using namespace detail;
    std::ifstream in(filename, std::ios::binary);
    if (!in) return 1;
    const int ML = MAX_LEVELS;
    HostPool &h = host;
    h.s = Sizes();
    int n_max_old = 0, cblocks_id_max = 0, n_probed = 0, it = 0;
    int64_t cells_id_max = 0;
    std::vector<int> n_ids(ML + 1), id_max(ML + 1);
    bool ok = get(in, n_max_old) && get_n(in, n_ids.data(), ML + 1) && get(in, h.n_gaps) &&
              get_n(in, id_max.data(), ML + 1) && get(in, cblocks_id_max) && get(in, cells_id_max) && get(in, n_probed) &&
              get(in, it);
    if (!ok) return 1;
    if (id_max[ML] > n_maxcblocks || cblocks_id_max != id_max[ML] || cells_id_max != (int64_t)cblocks_id_max * M_CBLOCK ||
        n_probed != (int)id_set_probed.size())
        return 1;
    if (iter) *iter = it;
    h.n_blocks = cblocks_id_max;
    h.n_ids.assign(n_ids.begin(), n_ids.begin() + ML);
    h.id_max = id_max;
    h.id_set.assign(ML, {});
    for (int L = 0; L < ML; L++) {
        h.id_set[L].resize(h.n_ids[L]);
        ok = ok && get_n(in, h.id_set[L].data(), h.id_set[L].size());
    }
    // The gap stack: the IDs a larger pool adds go below the file's entries (they are taken
    // last); a smaller pool drops the file's lowest entries (its highest IDs, never used).
    std::vector<int> gaps((size_t)std::max(0, h.n_gaps));
    ok = ok && get_n(in, gaps.data(), gaps.size());
    const long long grow = n_maxcblocks - n_max_old;
    if (grow >= 0) {
        h.gap_set.resize((size_t)(grow + h.n_gaps));
        for (long long k = 0; k < grow; k++) h.gap_set[k] = (int)(n_maxcblocks - 1 - k);
        std::copy(gaps.begin(), gaps.end(), h.gap_set.begin() + grow);
    } else {
        if (h.n_gaps + grow < 0) return 1;
        h.gap_set.assign(gaps.begin() + (-grow), gaps.end());
    }
    h.n_gaps = (int)h.gap_set.size();
    const long long nb = h.n_blocks, nc = nb * M_CBLOCK;
    h.cell_mask.resize(nc);
    ok = ok && get_n(in, h.cell_mask.data(), (size_t)nc);
    const bool dbl = P.precision == AGAL_FP64;
    auto get_real = [&](double &v) {
        if (dbl) return get(in, v);
        float f;
        const bool r = get(in, f);
        v = f;
        return r;
    };
    h.X.resize((size_t)N_DIM * nb);
    for (size_t k = 0; k < h.X.size() && ok; k++) {
        double v;
        ok = get_real(v);
        h.X[k] = (float)v;
    }
    h.blk_mask.resize(nb);
    ok = ok && get_n(in, h.blk_mask.data(), (size_t)nb);
    h.nbr.resize((size_t)N_Q_max * nb);
    h.nbr_child.resize((size_t)N_Q_max * nb);
    for (int p = 0; p < N_Q_max && ok; p++)
        ok = get_n(in, &h.nbr[(size_t)p * nb], (size_t)nb) && get_n(in, &h.nbr_child[(size_t)p * nb], (size_t)nb);
    h.onb.resize(nb);
    h.ref.resize(nb);
    h.level.resize(nb);
    ok = ok && get_n(in, h.onb.data(), (size_t)nb) && get_n(in, h.ref.data(), (size_t)nb) &&
         get_n(in, h.level.data(), (size_t)nb);
    h.fsize = dbl ? 8 : 4;
    h.f.assign((size_t)N_Q * nc * h.fsize, 0);
    for (int p = 0; p < N_Q && ok; p++) {
        unsigned char *row = &h.f[(size_t)pb(N_Q, p) * nc * h.fsize];
        ok = get_n(in, row, (size_t)nc * h.fsize);
    }
    for (size_t k = 0; k < U_probed_tn.size() && ok; k++) ok = get_real(U_probed_tn[k]);
    const size_t n_mean = (size_t)(N_DIM + 1) * (h.n_ids.empty() ? 0 : h.n_ids[0]) * M_CBLOCK;
    U_mean.assign(n_mean, 0.0);
    for (size_t k = 0; k < n_mean && ok; k++) ok = get_real(U_mean[k]);
    h.macros.assign((size_t)4 * nc, 0.0);
    h.s.n_gaps = h.n_gaps;
    return ok ? 0 : 1;
}

inline int Mesh::M_Print_VTHB(int i_dev, int iter)
{
    (void)i_dev;
    const std::string name = "out_" + std::to_string(iter + 1);
    return PrintVTHB(output_dir, name, N_PRINT_LEVELS_VTHB);
}

// `dir` + `name`.vthb and the directory `dir` + `name` holding `name`_<level>_<k>.vti, for the
// first n_levels levels that hold blocks.  The host pool must be current; the velocity gradients
// come from the solver's S_ComputeOutputProperties.
inline int Mesh::PrintVTHB(const std::string &dir, const std::string &name, int n_levels, bool binary)
{
    using namespace detail;
    const HostPool &h = host;
    if (h.s.max_levels != MAX_LEVELS) throw std::runtime_error("vthb output: the host pool is not current");
    Solver *sol = need_solver();
    int nz_levels = 1;
    for (int L = 1; L < MAX_LEVELS; L++)
        if (h.n_ids[L] > 0) nz_levels++;
    const int levels = std::min(n_levels, nz_levels);
    make_dir(dir);
    make_dir(dir + name);
    std::ofstream top(dir + name + ".vthb");
    if (!top) throw std::runtime_error("cannot write " + dir + name + ".vthb");
    top << "<?xml version=\"1.0\"?>\n<VTKFile type=\"vtkOverlappingAMR\" version=\"1.1\" byte_order=\"LittleEndian\" "
           "header_type=\"UInt64\">\n  <vtkOverlappingAMR origin=\"0 0 0\" grid_description=\""
        << (N_DIM == 2 ? "XY" : "XYZ") << "\">\n";
    const int C = M_CBLOCK;
    std::vector<double> u((size_t)C * 8);
    char buf[640];
    for (int L = 0; L < levels; L++) {
        const double hL = dxf_vec[L], hz = N_DIM == 2 ? dx : hL;
        std::snprintf(buf, sizeof(buf), "    <Block level=\"%d\" spacing=\"%.17g %.17g %.17g\">\n", L, hL, hL, hz);
        top << buf;
        for (int k = 0; k < h.n_ids[L]; k++) {
            const int b = h.id_set[L][k];
            double o[3] = {0, 0, 0};
            int lo[3] = {0, 0, 0};
            for (int d = 0; d < N_DIM; d++) {
                o[d] = h.X[(size_t)d * h.n_blocks + b];
                lo[d] = (int)std::lround(o[d] / hL);
            }
            const std::string piece = name + "_" + std::to_string(L) + "_" + std::to_string(k) + ".vti";
            std::snprintf(buf, sizeof(buf), "      <DataSet index=\"%d\" amr_box=\"%d %d %d %d %d %d\" file=\"%s/%s\"/>\n", k,
                          lo[0], lo[0] + 3, lo[1], lo[1] + 3, lo[2], lo[2] + (N_DIM == 3 ? 3 : 0), name.c_str(),
                          piece.c_str());
            top << buf;
            std::fill(u.begin(), u.end(), 0.0);
            sol->S_ComputeOutputProperties(0, 0, b, hL, u.data());
            std::vector<double> ref(C, h.ref[b]), onb(C, h.onb[b]), cm(C), bm(2 * C), rho(C), vel(3 * C), vor(3 * C), qc(C);
            for (int t = 0; t < C; t++) {
                cm[t] = h.cell_mask[(size_t)b * C + t];
                bm[2 * t] = h.blk_mask[b];
                bm[2 * t + 1] = 0.0;
                rho[t] = u[t];
                for (int d = 0; d < 3; d++) {
                    vel[3 * t + d] = u[t + (1 + d) * C];
                    vor[3 * t + d] = u[t + (4 + d) * C];
                }
                qc[t] = u[t + 7 * C];
            }
            std::ofstream v(dir + name + "/" + piece);
            if (!v) throw std::runtime_error("cannot write " + dir + name + "/" + piece);
            const int ez = N_DIM == 3 ? 4 : 0;
            std::snprintf(buf, sizeof(buf),
                          "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" "
                          "header_type=\"UInt64\">\n  <ImageData WholeExtent=\"0 4 0 4 0 %d\" Origin=\"%.17g %.17g %.17g\" "
                          "Spacing=\"%.17g %.17g %.17g\">\n",
                          ez, o[0], o[1], o[2], hL, hL, hz);
            v << buf;
            std::snprintf(buf, sizeof(buf), "    <Piece Extent=\"0 4 0 4 0 %d\">\n      <PointData>\n      </PointData>\n"
                                            "      <CellData Scalars=\"Density\" Vectors=\"Velocity\">\n", ez);
            v << buf;
            data_array(v, "Ref ID", 1, ref, binary);
            data_array(v, "On Boundary?", 1, onb, binary);
            data_array(v, "Cell Mask", 1, cm, binary);
            data_array(v, "Block Mask", 2, bm, binary);
            data_array(v, "Density", 1, rho, binary);
            data_array(v, "Velocity", 3, vel, binary);
            data_array(v, "Vorticity", 3, vor, binary);
            data_array(v, "Q-Criterion", 1, qc, binary);
            v << "      </CellData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n";
        }
        top << "    </Block>\n";
    }
    top << "  </vtkOverlappingAMR>\n</VTKFile>\n";
    return top.good() ? 0 : 1;
}

}  // namespace agal

#endif  // AGAL_OUTPUT_HPP
