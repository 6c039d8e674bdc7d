/*
 * agal/parser.hpp -- the input.txt reader of the header-library host API.
 *
 * Replaces Parser (inc/parser.h:13-188): key-value lines "NAME VALUE", '#' starts a comment line,
 * three typed maps with defaults, integer values may be written "A*Nx" (A times the file's Nx),
 * unknown keys are reported and skipped.  The key set and the default values are the reference's
 * (parser.h:62-144): they are the file format a user of the reference already has.  ToParams()
 * maps the keys the octree-AMR step reads onto the C ABI's agal_params.
 */
#ifndef AGAL_PARSER_HPP
#define AGAL_PARSER_HPP

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../agal_capi.h"

namespace agal {

class Parser {
public:
    std::string input_file_directory;
    std::map<std::string, int> params_int;
    std::map<std::string, double> params_dbl;
    std::map<std::string, std::string> params_str;
    // Keys of the file that are not parameters (the reference prints them and goes on).
    std::vector<std::string> invalid_keys;
    bool file_found = false;

    // Reads `dir + "input.txt"` (parser.h:185-188: the directory string is used as a prefix).
    explicit Parser(const std::string &dir) : input_file_directory(dir) { ReadInputFile(); }
    // Defaults only, then the text `content` as if it were the file (for generated inputs).
    static Parser FromString(const std::string &content, const std::string &dir = "")
    {
        Parser p(dir, 0);
        std::istringstream in(content);
        p.file_found = true;
        p.Read(in);
        return p;
    }

    // "A*Nx" -> A * Nx, "A" -> A (parser.h:34-44).  Nx must come second and "1*Nx" be written
    // out; the factor is an integer.
    static int ParseNxFactor(const std::string &s, int Nx)
    {
        const std::string::size_type star = s.find('*');
        const int A = std::stoi(s.substr(0, star));
        return star == std::string::npos ? A : A * Nx;
    }

    int ReadInputFile()
    {
        std::ifstream in(input_file_directory + "input.txt");
        file_found = static_cast<bool>(in);
        Read(in);
        return 0;
    }

    int PrintMaps() const
    {
        for (const auto &kv : params_int) std::printf("%s %d\n", kv.first.c_str(), kv.second);
        for (const auto &kv : params_dbl) std::printf("%s %g\n", kv.first.c_str(), kv.second);
        for (const auto &kv : params_str) std::printf("%s %s\n", kv.first.c_str(), kv.second.c_str());
        return 0;
    }

    int Int(const std::string &k) const { return params_int.at(k); }
    double Dbl(const std::string &k) const { return params_dbl.at(k); }
    const std::string &Str(const std::string &k) const { return params_str.at(k); }

    // The step's parameters.  Fields that the reference fixes at compile time (velocity set,
    // precision, boundary hook, initial state: src/main.cu:48-64, inc/custom.h) come from `base`,
    // which agal_params_default() or the caller fills; the file sets the rest.  S_COLLISION 0 is
    // BGK; 1 (LES-BGK in the reference) is not built; S_INTERP 0/1 linear, 2 cubic
    // (InterpOrder, inc/solver_lbm.h:25-30).
    agal_params ToParams(agal_params base) const
    {
        base.Nx = Int("Nx");
        base.L_c = Dbl("L_c");
        base.L_fy = Dbl("L_fy");
        base.L_fz = Dbl("L_fz");
        base.PERIODIC_X = Int("PERIODIC_X");
        base.PERIODIC_Y = Int("PERIODIC_Y");
        base.PERIODIC_Z = Int("PERIODIC_Z");
        base.MAX_LEVELS_WALL = Int("MAX_LEVELS_WALL");
        base.MAX_LEVELS_INTERIOR = Int("MAX_LEVELS_INTERIOR");
        base.N_LEVEL_START = Int("N_LEVEL_START");
        base.v0 = Dbl("v0");
        base.S_CRITERION = Int("S_CRITERION");
        base.N_REFINE_START = Dbl("N_REFINE_START");
        base.N_REFINE_INC = Dbl("N_REFINE_INC");
        base.N_REFINE_MAX = Dbl("N_REFINE_MAX");
        base.P_REFINE = Int("P_REFINE");
        base.s1 = Dbl("s1");
        base.s2 = Dbl("s2");
        base.s3 = Dbl("s3");
        base.s4 = Dbl("s4");
        base.s5 = Dbl("s5");
        base.s6 = Dbl("s6");
        base.G_NEAR_WALL_DISTANCE = Dbl("G_NEAR_WALL_DISTANCE");
        base.G_BIN_DENSITY = Int("G_BIN_DENSITY");
        base.G_BIN_LEVELS = Int("G_BIN_LEVELS");
        base.G_BIN_SPEC = Int("G_BIN_SPEC");
        base.G_BIN_MAKE_LINKS = Int("G_BIN_MAKE_LINKS");
        base.S_BC_TYPE = Int("S_BC_TYPE");
        base.S_FORCE_TYPE = Int("S_FORCE_TYPE");
        base.S_FORCE_ORDER = Int("S_FORCE_ORDER");
        base.S_FORCEVOLUME_Xm = Dbl("S_FORCEVOLUME_Xm");
        base.S_FORCEVOLUME_XM = Dbl("S_FORCEVOLUME_XM");
        base.S_FORCEVOLUME_Ym = Dbl("S_FORCEVOLUME_Ym");
        base.S_FORCEVOLUME_YM = Dbl("S_FORCEVOLUME_YM");
        base.S_FORCEVOLUME_Zm = Dbl("S_FORCEVOLUME_Zm");
        base.S_FORCEVOLUME_ZM = Dbl("S_FORCEVOLUME_ZM");
        const int coll = Int("S_COLLISION");
        if (coll == 0) base.collision = AGAL_BGK;
        else if (coll == 2) base.collision = AGAL_MRT;
        else throw std::invalid_argument("S_COLLISION " + std::to_string(coll) + " is not built");
        base.interp_order = Int("S_INTERP") == 2 ? AGAL_INTERP_CUBIC : AGAL_INTERP_LINEAR;
        return base;
    }

private:
    Parser(const std::string &dir, int) : input_file_directory(dir) { SetDefaults(); }

    void SetDefaults()
    {
        // Integer keys (parser.h:62-111).
        static const std::pair<const char *, int> ints[] = {
            {"MAX_LEVELS_WALL", 5}, {"MAX_LEVELS_INTERIOR", 4}, {"N_ITER_TOTAL", 640}, {"N_LEVEL_START", 0},
            {"Nx", 64}, {"N_RESTART", 1}, {"G_LOADTYPE", 0}, {"G_PRINT", 1}, {"G_VIS_TREE", 0},
            {"G_BIN_DENSITY", 1}, {"G_BIN_APPROACH", 0}, {"G_BIN_LEVELS", 1}, {"G_BIN_SPEC", 1},
            {"G_BIN_MAKE_LINKS", 0}, {"PERIODIC_X", 0}, {"PERIODIC_Y", 0}, {"PERIODIC_Z", 0},
            {"S_COLLISION", 0}, {"S_INTERP", 1}, {"S_INTERP_HYBRID", 0}, {"S_AVERAGE", 0}, {"S_CRITERION", 0},
            {"S_LES", 0}, {"S_FORCE_ORDER", 0}, {"S_FORCE_TYPE", 0}, {"S_BC_TYPE", 0}, {"P_REFINE", 32},
            {"N_PROBE", 0}, {"N_PROBE_DENSITY", 4}, {"N_PROBE_FREQUENCY", 32}, {"N_PROBE_FORCE", 0},
            {"N_PROBE_F_START", 0}, {"N_PROBE_F_FREQUENCY", 16}, {"N_PROBE_AVE", 0}, {"N_PROBE_AVE_FREQUENCY", 1},
            {"N_PROBE_AVE_START", 6400}, {"N_PRINT_LEVELS", 2}, {"N_PRINT_LEVELS_VTHB", 0},
            {"N_PRINT_LEVELS_PATCH", 2}, {"N_PRINT_LEVELS_IMAGE", 2}, {"P_OUTPUT", 64}, {"N_OUTPUT", 10},
            {"N_OUTPUT_START", 0}, {"VOL_I_MIN", 0}, {"VOL_I_MAX", 64}, {"VOL_J_MIN", 0}, {"VOL_J_MAX", 64},
            {"VOL_K_MIN", 0}, {"VOL_K_MAX", 64}, {"USE_CPU", 1}};
        // Floating-point keys (parser.h:112-135).
        static const std::pair<const char *, double> dbls[] = {
            {"L_c", 1.0}, {"L_fy", 1.0}, {"L_fz", 1.0}, {"v0", 1.5625e-5}, {"G_NEAR_WALL_DISTANCE", 0.05},
            {"s1", 1.19}, {"s2", 1.40}, {"s3", 1.20}, {"s4", 1.20}, {"s5", 1.20}, {"s6", 1.98},
            {"V_PROBE_TOL", 1e-4}, {"N_REFINE_START", -3.0}, {"N_REFINE_INC", 1.0}, {"N_REFINE_MAX", 1.0},
            {"S_FORCEVOLUME_Xm", 0.0}, {"S_FORCEVOLUME_XM", 0.0}, {"S_FORCEVOLUME_Ym", 0.0},
            {"S_FORCEVOLUME_YM", 0.0}, {"S_FORCEVOLUME_Zm", 0.0}, {"S_FORCEVOLUME_ZM", 0.0}};
        for (const auto &kv : ints) params_int[kv.first] = kv.second;
        for (const auto &kv : dbls) params_dbl[kv.first] = kv.second;
        // String keys (parser.h:136-141); the input directory defaults to the one given.
        params_str["I_DIR_NAME"] = input_file_directory;
        params_str["P_DIR_NAME"] = "../out/TEST_UPLOAD/";
        params_str["G_FILENAME"] = "bunny_ascii.stl";
    }

    // Two passes as parser.h:143-181: Nx first (it scales the "A*Nx" values), then every key.
    void Read(std::istream &in)
    {
        if (params_int.empty()) SetDefaults();
        std::vector<std::pair<std::string, std::string>> kv;
        int Nx = -1;
        for (std::string line; std::getline(in, line);) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ss(line);
            std::string name, val;
            if (!(ss >> name)) continue;  // blank line
            ss >> val;
            kv.emplace_back(name, val);
            if (name == "Nx") Nx = std::stoi(val);
        }
        for (const auto &e : kv) {
            if (params_int.count(e.first))
                params_int[e.first] = ParseNxFactor(e.second, Nx);
            else if (params_dbl.count(e.first))
                params_dbl[e.first] = std::stod(e.second);
            else if (params_str.count(e.first))
                params_str[e.first] = e.second;
            else {
                invalid_keys.push_back(e.first);
                std::fprintf(stderr, "Error: Invalid input (%s) skipped...\n", e.first.c_str());
            }
        }
    }
};

}  // namespace agal

#endif  // AGAL_PARSER_HPP
