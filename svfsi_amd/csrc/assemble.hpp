// assemble.hpp -- kernel-side parameter blocks of the assembly kernels.
#pragma once

#include "internal.hpp"

namespace svfsi {

// Equation parameters passed by value to the assembly kernels.
struct AsmParams {
    double am, af, beta, gam, dt;
    double rho, E, nu;
    double fb[3];
    double mu, bstab;
    double C10, Kpen;
    int volType;
    int pad;
    int viscType;  // 698 constant, 697 Carreau-Yasuda, 696 Casson
    int pad2;
    double mu_o, lam, va, vn;
};

// TET4 4-point rule and shape values (GETGIP/GETGNN, svFSI/NN.f:266-275,
// 655-671), computed on the host with the reference's formulas.
struct Tet4Rule {
    double w[4];
    double N[4][4]; // N[g][a]
};

// TRI3 3-point face rule (svFSI/NN.f:416-422, 1100-1110).
struct Tri3Rule {
    double w[3];
    double N[3][3];
};

Tet4Rule make_tet4_rule();
Tri3Rule make_tri3_rule();
AsmParams to_asm(const svfsi_eq_params& p);

int launch_asm_tet4(Mesh& M, int phys, const AsmParams& P, const Tet4Rule& rule, int tDof,
                    const double* Ag, const double* Yg, const double* Dg, const double* prop,
                    double* val, double* R, hipStream_t s);
int launch_face_traction(const Mesh& M, const Mesh::FaceSet& F, int dof, const double* h,
                         const Tri3Rule& rule, double* R, hipStream_t s);
int launch_face_backflow(const Mesh& M, const Mesh::FaceSet& F, const AsmParams& P, int tDof,
                         const double* Yg, const Tri3Rule& rule, double* val, double* R,
                         hipStream_t s);
int launch_lelas_dkde(const Mesh& M, const AsmParams& P, const Tet4Rule& rule, const double* lam,
                      const double* u, double* g, hipStream_t s);
int launch_fluid_dmu(const Mesh& M, const AsmParams& P, const Tet4Rule& rule, int tDof, const double* Ag,
                     const double* Yg, const double* muEl, const double* lam, const double* du, double* g,
                     hipStream_t s);

} // namespace svfsi

namespace svfsi {

// Quadrature table of a TET10 rule (GETGIP/GETGNN, svFSI/NN.f:277-307,
// 674-715): per Gauss point w, N(10), Nxi(10,3) -> 41 doubles.
constexpr int kTet10Rec = 41;
// returns nG; xiOut (optional, nG x 3) receives the points
int make_tet10_rule(int quad, std::vector<double>& tab, std::vector<double>* xiOut = nullptr);
// Taylor-Hood (P2P1) FLUID quadrature: the 15 points of the tet10 rule (the
// momentum loop) then the 4 points of the tet4 rule (the continuity loop),
// each as w, Nw(10), Nwxi(10,3) of the tet10 basis and Nq(4) of the tet4
// basis -> 45 doubles per point.
constexpr int kThRec = 45;
int make_th_rule(std::vector<double>& tab);
// CONSTRUCT_FLUID with nFs = 2 on a tet10 mesh, colour by colour, then
// THOOD_ValRC on the mid-edge nodes.
// Gather assembly's second pass: every Val block and R row summed from the
// stored element blocks (M.ke) and rows (M.re) in element-id order.
int launch_gather(Mesh& M, int dof, double* val, double* R, hipStream_t s);
int launch_asm_fluid_th(Mesh& M, const AsmParams& P, int tDof, const double* Ag, const double* Yg,
                        const double* muEl, double* val, double* R, hipStream_t s);
// kern: one of SVFSI_STRUCT_*.  The MFMA kernels keep material-point
// records in M.srec.
int launch_asm_struct(Mesh& M, const AsmParams& P, int nG, const double* rule, int tDof,
                      const double* Ag, const double* Dg, double* val, double* R, int kern, hipStream_t s);

} // namespace svfsi
