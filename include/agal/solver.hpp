/*
 * agal/solver.hpp -- the solver plug-in interface of the header-library host API.
 *
 * The reference's abstract Solver<ufloat_t,ufloat_g_t,AP> (inc/solver.h:33-87): ten pure virtual
 * routines that the mesh calls, with the same names, arguments and meaning; the constructor
 * registers the solver with its mesh (solver.h:73-80).  The precision is a run-time field of the
 * mesh here, so the interface is not a template and dx_L is a double.
 */
#ifndef AGAL_SOLVER_HPP
#define AGAL_SOLVER_HPP

namespace agal {

class Mesh;

class Solver {
public:
    Mesh *mesh = nullptr;

    explicit Solver(Mesh *mesh_);
    virtual ~Solver() = default;

    virtual int S_SetIC(int i_dev, int L) = 0;
    virtual int S_Interpolate(int i_dev, int L, int var) = 0;
    virtual int S_Average(int i_dev, int L, int var) = 0;
    virtual int S_Advance(int i_dev, int L, double *tmp) = 0;
    virtual int S_ComputeProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out_u) = 0;
    virtual int S_ComputeOutputProperties(int i_dev, int i_Q, int i_kap, double dx_L, double *out_u) = 0;
    virtual int S_ReportForces(int i_dev, int L, int i, double t, int START) = 0;
    virtual int S_ComputeRefCriteria(int i_dev, int L, int var) = 0;
    virtual int S_RefreshVariables(int i_dev, int L, int var = 0) = 0;
    virtual int S_Debug(int i_dev, int L, int var) = 0;
};

}  // namespace agal

#endif  // AGAL_SOLVER_HPP
