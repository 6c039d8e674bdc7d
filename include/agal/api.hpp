/*
 * agal/api.hpp -- the C ABI of include/agal_capi.h as a table of function pointers, for the
 * header-library classes (agal/mesh.hpp, agal/solver_lbm.hpp).
 *
 * Api::Linked() takes the agal_* symbols of the HIP library the program links (libagal.so).
 * Api::Load(path, prefix) opens a library with dlopen and binds `prefix` + name, for programs that
 * choose the library at run time; any library exporting the same signatures under another prefix
 * binds the same way.
 */
#ifndef AGAL_API_HPP
#define AGAL_API_HPP

#include <dlfcn.h>

#include <stdexcept>
#include <string>

#include "../agal_capi.h"

namespace agal {

struct Api {
    int (*create)(const agal_params *, long long, int, agal_ctx **) = nullptr;
    int (*destroy)(agal_ctx *) = nullptr;
    int (*get_sizes)(agal_ctx *, agal_sizes *) = nullptr;
    int (*sync)(agal_ctx *) = nullptr;
    int (*set_ic)(agal_ctx *, int) = nullptr;
    int (*advance)(agal_ctx *, int) = nullptr;
    int (*interpolate)(agal_ctx *, int, int) = nullptr;
    int (*average)(agal_ctx *, int, int) = nullptr;
    int (*refresh)(agal_ctx *, int) = nullptr;
    int (*ref_criteria)(agal_ctx *) = nullptr;
    int (*refine_coarsen)(agal_ctx *, int *) = nullptr;
    int (*refine_with_solution)(agal_ctx *, int *) = nullptr;
    int (*mark_uniform)(agal_ctx *, int) = nullptr;
    int (*mark_blocks)(agal_ctx *, const int *, int, int) = nullptr;
    int (*freeze_refined)(agal_ctx *, int) = nullptr;
    int (*advance_loop)(agal_ctx *, int, int, double *, double *) = nullptr;
    int (*refine_near_wall)(agal_ctx *) = nullptr;
    int (*set_geometry)(agal_ctx *, const float *, long long) = nullptr;
    int (*forces)(agal_ctx *, double *) = nullptr;
    int (*download_populations)(agal_ctx *, void *, long long, long long) = nullptr;
    int (*upload_populations)(agal_ctx *, const void *, long long, long long) = nullptr;
    int (*download_macros)(agal_ctx *, void *, long long, long long) = nullptr;
    int (*download_cell_masks)(agal_ctx *, int *, long long) = nullptr;
    int (*download_block_ints)(agal_ctx *, int, int *, long long, long long) = nullptr;
    int (*download_coords)(agal_ctx *, float *, long long, long long) = nullptr;
    int (*download_id_set)(agal_ctx *, int, int *, int) = nullptr;
    int (*download_gap_set)(agal_ctx *, int *, int) = nullptr;
    int (*import_pool)(agal_ctx *, const agal_pool *) = nullptr;
    void *handle = nullptr;  // dlopen handle (Load only)

    static Api Linked()
    {
        Api a;
        a.create = agal_create;
        a.destroy = agal_destroy;
        a.get_sizes = agal_get_sizes;
        a.sync = agal_sync;
        a.set_ic = agal_set_ic;
        a.advance = agal_advance;
        a.interpolate = agal_interpolate;
        a.average = agal_average;
        a.refresh = agal_refresh;
        a.ref_criteria = agal_ref_criteria;
        a.refine_coarsen = agal_refine_coarsen;
        a.refine_with_solution = agal_refine_with_solution;
        a.mark_uniform = agal_mark_uniform;
        a.mark_blocks = agal_mark_blocks;
        a.freeze_refined = agal_freeze_refined;
        a.advance_loop = agal_advance_loop;
        a.refine_near_wall = agal_refine_near_wall;
        a.set_geometry = agal_set_geometry;
        a.forces = agal_forces;
        a.download_populations = agal_download_populations;
        a.upload_populations = agal_upload_populations;
        a.download_macros = agal_download_macros;
        a.download_cell_masks = agal_download_cell_masks;
        a.download_block_ints = agal_download_block_ints;
        a.download_coords = agal_download_coords;
        a.download_id_set = agal_download_id_set;
        a.download_gap_set = agal_download_gap_set;
        a.import_pool = agal_import_pool;
        return a;
    }

    static Api Load(const std::string &path, const std::string &prefix = "agal_")
    {
        Api a;
        a.handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!a.handle) throw std::runtime_error("dlopen(" + path + "): " + dlerror());
        auto bind = [&](auto &fn, const char *name) {
            void *s = dlsym(a.handle, (prefix + name).c_str());
            if (!s) throw std::runtime_error("missing symbol " + prefix + name + " in " + path);
            fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(s);
        };
        bind(a.create, "create");
        bind(a.destroy, "destroy");
        bind(a.get_sizes, "get_sizes");
        bind(a.sync, "sync");
        bind(a.set_ic, "set_ic");
        bind(a.advance, "advance");
        bind(a.interpolate, "interpolate");
        bind(a.average, "average");
        bind(a.refresh, "refresh");
        bind(a.ref_criteria, "ref_criteria");
        bind(a.refine_coarsen, "refine_coarsen");
        bind(a.refine_with_solution, "refine_with_solution");
        bind(a.mark_uniform, "mark_uniform");
        bind(a.mark_blocks, "mark_blocks");
        bind(a.freeze_refined, "freeze_refined");
        bind(a.advance_loop, "advance_loop");
        bind(a.refine_near_wall, "refine_near_wall");
        bind(a.set_geometry, "set_geometry");
        bind(a.forces, "forces");
        bind(a.download_populations, "download_populations");
        bind(a.upload_populations, "upload_populations");
        bind(a.download_macros, "download_macros");
        bind(a.download_cell_masks, "download_cell_masks");
        bind(a.download_block_ints, "download_block_ints");
        bind(a.download_coords, "download_coords");
        bind(a.download_id_set, "download_id_set");
        bind(a.download_gap_set, "download_gap_set");
        bind(a.import_pool, "import_pool");
        return a;
    }
};

}  // namespace agal

#endif  // AGAL_API_HPP
