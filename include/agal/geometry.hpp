/*
 * agal/geometry.hpp -- the geometry front-end of the header-library host API.
 *
 * Replaces Geometry<ufloat_t,ufloat_g_t,AP> (inc/geometry.h:35-205) for what the step needs: a
 * closed triangle surface in domain coordinates, 9 floats per triangle, built from an ASCII STL
 * file (G_ImportSTL_ASCII, src/geometry/geometry_import.cu:10-83), the seeded noisy icosphere, or
 * the reference's primitive list file (G_ImportBoundariesFromTextFile, geometry_add.cu:539), with
 * long faces split (G_RefineFaces_Length, geometry_refine.cu:178-238).  The mesh hands the
 * triangles to the library (agal_set_geometry), which bins them for the voxelizer.
 */
#ifndef AGAL_GEOMETRY_HPP
#define AGAL_GEOMETRY_HPP

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../agal_capi.h"
#include "parser.hpp"

namespace agal {

class Geometry {
public:
    std::vector<float> tri;  // 9 floats per triangle: x1 y1 z1 x2 y2 z2 x3 y3 z3

    Geometry() = default;
    // G_Init (geometry_init.cu:10-31): G_LOADTYPE 1 reads G_FILENAME (an ASCII STL) from the input
    // directory; 0 reads the primitive list "geometry.txt" there.
    explicit Geometry(const Parser *parser)
    {
        if (!parser) return;
        const std::string dir = parser->Str("I_DIR_NAME");
        if (parser->Int("G_LOADTYPE") == 1)
            G_ImportSTL_ASCII(dir + "geometry/" + parser->Str("G_FILENAME"));
        else
            G_ImportBoundariesFromTextFile(dir + "geometry.txt");
    }

    long long n_faces() const { return (long long)tri.size() / 9; }

    int G_ImportSTL_ASCII(const std::string &path)
    {
        long long n = agal_load_stl(path.c_str(), nullptr, 0);
        if (n < 0) throw std::runtime_error("cannot read STL file " + path);
        std::vector<float> t((size_t)n * 9);
        if (n > 0 && agal_load_stl(path.c_str(), t.data(), n) != n) throw std::runtime_error("STL read failed: " + path);
        tri.insert(tri.end(), t.begin(), t.end());
        return 0;
    }

    int G_AddIcosphere(double cx, double cy, double cz, double r, int subdivisions, double noise, unsigned long long seed)
    {
        long long n = agal_icosphere(cx, cy, cz, r, subdivisions, noise, seed, nullptr, 0);
        if (n < 0) throw std::invalid_argument("icosphere arguments");
        std::vector<float> t((size_t)n * 9);
        agal_icosphere(cx, cy, cz, r, subdivisions, noise, seed, t.data(), n);
        tri.insert(tri.end(), t.begin(), t.end());
        return 0;
    }

    // Primitive list (the reference's geometry.txt, geometry_add.cu:539-600): one primitive per
    // line, '#' comments.  3D primitives built here: "SPHERE N1 N2 cx cy cz R" (G_AddSphere,
    // geometry_add.cu:281-316), "PRISM CENTER cx cy cz lx ly lz" and
    // "PRISM CORNER ax bx ay by az bz" (an axis-aligned box, 12 outward triangles; geometry_add.cu:
    // 587-608).
    int G_ImportBoundariesFromTextFile(const std::string &path)
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot read primitive list " + path);
        return G_ImportBoundariesFromText(in);
    }

    int G_ImportBoundariesFromText(std::istream &in)
    {
        for (std::string line; std::getline(in, line);) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ss(line);
            std::string kind;
            if (!(ss >> kind)) continue;
            if (kind == "SPHERE") {
                int n1 = 0, n2 = 0;
                double cx, cy, cz, r;
                if (!(ss >> n1 >> n2 >> cx >> cy >> cz >> r)) throw std::runtime_error("bad SPHERE line: " + line);
                G_AddSphere(n1, n2, cx, cy, cz, r);
            } else if (kind == "PRISM") {
                std::string mode;
                double a[6];
                if (!(ss >> mode >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5])) throw std::runtime_error("bad PRISM line: " + line);
                if (mode == "CENTER")
                    G_AddPrism(a[0] - a[3] / 2, a[1] - a[4] / 2, a[2] - a[5] / 2, a[0] + a[3] / 2, a[1] + a[4] / 2, a[2] + a[5] / 2);
                else  // CORNER ax bx ay by az bz
                    G_AddPrism(a[0], a[2], a[4], a[1], a[3], a[5]);
            } else {
                throw std::runtime_error("primitive not built in 3D: " + kind);
            }
        }
        return 0;
    }

    // G_AddSphere (geometry_add.cu:281-316): an N1 x N2 node grid, polar angle pi j/(N1-1), azimuth
    // 2 pi k/(N2-1) (the seam column repeats the first), two triangles per grid cell with the
    // reference's vertex order; the pole rows give zero-area triangles, kept as the reference does.
    int G_AddSphere(int N1, int N2, double cx, double cy, double cz, double R)
    {
        if (N1 < 2 || N2 < 2 || !(R > 0)) throw std::invalid_argument("sphere resolution / radius");
        const float pi = 3.14159265358979323846f;
        const float d1 = (float)N1 - 1.0f, d2 = (float)N2 - 1.0f;
        auto node = [&](int j, int k, float *o) {
            const float t = (pi / d1) * (float)j, ph = (2 * pi / d2) * (float)k;
            o[0] = (float)R * std::sin(t) * std::cos(ph) + (float)cx;
            o[1] = (float)R * std::sin(t) * std::sin(ph) + (float)cy;
            o[2] = (float)R * std::cos(t) + (float)cz;
        };
        for (int j = 0; j < N1 - 1; j++)
            for (int k = 0; k < N2 - 1; k++) {
                float a[3], b[3], c[3], d[3];
                node(j, k, a);
                node(j, k + 1, b);
                node(j + 1, k + 1, c);
                node(j + 1, k, d);
                push(a, c, b);  // face_1 (j,k), face_2 (j+1,k+1), face_3 (j,k+1)
                push(a, d, c);  // face_1 (j,k), face_2 (j+1,k), face_3 (j+1,k+1)
            }
        return 0;
    }

    int G_AddPrism(double x0, double y0, double z0, double x1, double y1, double z1)
    {
        const float v[8][3] = {{(float)x0, (float)y0, (float)z0}, {(float)x1, (float)y0, (float)z0},
                               {(float)x1, (float)y1, (float)z0}, {(float)x0, (float)y1, (float)z0},
                               {(float)x0, (float)y0, (float)z1}, {(float)x1, (float)y0, (float)z1},
                               {(float)x1, (float)y1, (float)z1}, {(float)x0, (float)y1, (float)z1}};
        // Outward-oriented faces (two triangles each).
        static const int f[12][3] = {{0, 3, 2}, {0, 2, 1}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                                     {3, 7, 6}, {3, 6, 2}, {0, 4, 7}, {0, 7, 3}, {1, 2, 6}, {1, 6, 5}};
        for (auto &t : f) push(v[t[0]], v[t[1]], v[t[2]]);
        return 0;
    }

    // The face-length limit of G_RefineFaces_Length (geometry_refine.cu:180-189): 0.95 G_BIN_SPEC
    // times the shortest domain length over the finest bin density 2^(G_BIN_LEVELS-1) G_BIN_DENSITY,
    // in single precision.
    static float FaceLengthLimit(const Parser &p)
    {
        const float Lx = (float)p.Dbl("L_c"), Ly = (float)(p.Dbl("L_fy") * Lx), Lz = (float)(p.Dbl("L_fz") * Lx);
        const float mult = 0.95f * (float)p.Int("G_BIN_SPEC");
        const float finest = std::pow(2.0f, (float)(p.Int("G_BIN_LEVELS") - 1)) * (float)p.Int("G_BIN_DENSITY");
        return mult * std::min(Lx, std::min(Ly, Lz)) / finest;
    }

    // G_RefineFaces_Length (geometry_refine.cu:10-238): a face whose longest edge exceeds l_spec is
    // split at its edge midpoints into four, (v1,m12,m31), (m12,v2,m23), (m31,m23,v3),
    // (m12,m23,m31), recursively until every piece is short enough.  The first piece takes the
    // original face's place, the others are appended in that order; faces are visited from the
    // last to the first.  Orientation and area are kept.
    int G_RefineFaces_Length(float l_spec)
    {
        if (!(l_spec > 0)) throw std::invalid_argument("l_spec");
        const long long n0 = n_faces();
        for (long long p = n0 - 1; p >= 0; p--) {
            float v[9];
            for (int k = 0; k < 9; k++) v[k] = tri[(size_t)p * 9 + k];
            bool replaced_first = false;
            refine_face(0, l_spec, v, replaced_first, p);
        }
        return 0;
    }

private:
    void refine_face(int nref, float l_spec, const float *v, bool &replaced_first, long long first)
    {
        auto len = [](const float *a, const float *b) {
            return std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) + (b[2] - a[2]) * (b[2] - a[2]));
        };
        const float lM = std::max(len(v, v + 3), std::max(len(v + 3, v + 6), len(v + 6, v)));
        if (lM > l_spec) {
            float m1[3], m2[3], m3[3];
            for (int d = 0; d < 3; d++) {
                m1[d] = 0.5f * (v[d] + v[3 + d]);
                m2[d] = 0.5f * (v[3 + d] + v[6 + d]);
                m3[d] = 0.5f * (v[6 + d] + v[d]);
            }
            const float a[9] = {v[0], v[1], v[2], m1[0], m1[1], m1[2], m3[0], m3[1], m3[2]};
            const float b[9] = {m1[0], m1[1], m1[2], v[3], v[4], v[5], m2[0], m2[1], m2[2]};
            const float c[9] = {m3[0], m3[1], m3[2], m2[0], m2[1], m2[2], v[6], v[7], v[8]};
            const float d[9] = {m1[0], m1[1], m1[2], m2[0], m2[1], m2[2], m3[0], m3[1], m3[2]};
            refine_face(nref + 1, l_spec, a, replaced_first, first);
            refine_face(nref + 1, l_spec, b, replaced_first, first);
            refine_face(nref + 1, l_spec, c, replaced_first, first);
            refine_face(nref + 1, l_spec, d, replaced_first, first);
        } else if (nref > 0) {
            if (!replaced_first) {
                for (int k = 0; k < 9; k++) tri[(size_t)first * 9 + k] = v[k];
                replaced_first = true;
            } else {
                tri.insert(tri.end(), v, v + 9);
            }
        }
    }

    void push(const float *a, const float *b, const float *c)
    {
        tri.insert(tri.end(), a, a + 3);
        tri.insert(tri.end(), b, b + 3);
        tri.insert(tri.end(), c, c + 3);
    }
};

}  // namespace agal

#endif  // AGAL_GEOMETRY_HPP
