// The C++ binding SEMO adds to its tree to run the pressure-based path on
// the GPU: SEMO's own equation member functions, with the reference's
// signatures and return values, forwarding to libsemo_gpu's C ABI
// (include/semo_gpu.h).  Linked in place of src/solvers/eqMomentum.cpp,
// eqPressure.cpp and eqVolfrac.cpp, it leaves SEMO's
// incompressiblePressureBased (src/solvers/pressureBased.cpp:6-483) and the
// rest of its call graph untouched: boundary values, EOS, sources and the
// time step stay SEMO's CPU code, and each equation call copies the fields it
// reads to the device and the fields it writes back.  (gpuBridgeStep.cpp
// replaces the whole step instead, with the fields resident on the device.)
//
// Configuration (SEMO has no dictionary entry for it): SEMO_GPU_DEVICE
// (default 0), SEMO_GPU_PRECISION = mixed | fp64 (default mixed: FP32 AMG
// inside FP64 Krylov), SEMO_GPU_TOL (relative residual, default 1e-9, AMGCL's
// pressure tolerance in solveAMGCL.cpp:967).
#include <mpi.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "gpuBridge.h"
#include "solvers/build.h"

using std::string;
using std::vector;

namespace semo_bridge {
namespace {

semo_ctx* g_ctx = nullptr;
size_t g_cells = 0;
double g_dt = -1.0;

int p_type(const string& t) {             // reconIncom.cpp:600-637
    if (t == "fixedValue") return SEMO_BC_FIXED_VALUE;
    if (t == "zeroGradient") return SEMO_BC_ZERO_GRADIENT;
    if (t == "switch") return SEMO_BC_SWITCH;
    return -1;
}
int u_type(const string& t) {             // reconIncom.cpp:642-852
    if (t == "fixedValue") return SEMO_BC_FIXED_VALUE;
    if (t == "zeroGradient") return SEMO_BC_ZERO_GRADIENT;
    if (t == "slip") return SEMO_BC_SLIP;
    if (t == "noSlip") return SEMO_BC_NO_SLIP;
    if (t == "surfaceNormalFixedValue") return SEMO_BC_SURFACE_NORMAL_FIXED_VALUE;
    if (t == "inletOutlet") return SEMO_BC_INLET_OUTLET;
    return -1;
}
int vf_type(const string& t) {            // reconIncom.cpp:882-1019
    if (t == "fixedValue") return SEMO_BC_FIXED_VALUE;
    if (t == "zeroGradient") return SEMO_BC_ZERO_GRADIENT;
    if (t == "inletOutlet") return SEMO_BC_INLET_OUTLET;
    if (t == "function") return SEMO_BC_FUNCTION;
    return -1;
}

void fail(const string& msg) {
    std::cerr << "| #Error : libsemo_gpu bridge: " << msg << std::endl;
    MPI_Abort(MPI_COMM_WORLD, EXIT_FAILURE);
}

// AoS mesh -> flat arrays (internal faces first, then the patches, as SEMO's
// loaders order them)
void upload(SEMO_Mesh_Builder& mesh) {
    vector<double> pts;
    vector<int32_t> own, nbr, fptr{0}, fpts, cptr{0}, cpts, pst, pnf;
    for (auto& p : mesh.points) { pts.push_back(p.x); pts.push_back(p.y); pts.push_back(p.z); }
    int nif = 0;
    for (auto& f : mesh.faces) {
        bool internal = f.getType() == SEMO_Types::INTERNAL_FACE;
        if (internal && (size_t)nif != own.size()) fail("internal faces must come first");
        own.push_back(f.owner);
        if (internal) { nbr.push_back(f.neighbour); ++nif; }
        for (int q : f.points) fpts.push_back(q);
        fptr.push_back((int32_t)fpts.size());
    }
    for (auto& c : mesh.cells) {
        for (int q : c.points) cpts.push_back(q);
        cptr.push_back((int32_t)cpts.size());
    }
    for (auto& b : mesh.boundary) {
        if (b.neighbProcNo != -1) fail("one rank only: processor patches are not supported");
        pst.push_back(b.startFace); pnf.push_back(b.nFaces);
    }
    semo_mesh_desc d{(int32_t)mesh.points.size(), (int32_t)mesh.cells.size(), (int32_t)mesh.faces.size(),
                     nif, (int32_t)mesh.boundary.size(), pts.data(), own.data(), nbr.data(),
                     fptr.data(), fpts.data(), cptr.data(), cpts.data(), pst.data(), pnf.data()};
    check(semo_mesh_upload(g_ctx, &d, /*reorder=*/1));
    g_cells = mesh.cells.size();
}

// ./VF.csv as reconIncomZeroOrder reads it (reconIncom.cpp:929-985): a
// header line, then comma-separated rows whose columns 3-6 are VF, x, y, z
void vf_points(vector<double>& pts) {
    std::ifstream in("./VF.csv");
    if (!in.is_open()) fail("Could not open the file - './VF.csv'");
    string line;
    std::getline(in, line);
    while (std::getline(in, line)) {
        vector<double> row;
        std::istringstream ls(line);
        string item;
        while (std::getline(ls, item, ',')) row.push_back(std::stod(item));
        if (row.size() < 7) fail("VF.csv row with fewer than 7 columns");
        for (int k = 3; k < 7; ++k) pts.push_back(row[k]);
    }
}

void set_case(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls, vector<SEMO_Species>& species) {
    semo_case cs{};
    for (int s = 0; s < 2; ++s) { cs.rho[s] = species[s].rho; cs.mu[s] = species[s].mu; }
    cs.sigma = species[0].sigma;
    for (int a = 0; a < 3; ++a) cs.g[a] = controls.gravityAcceleration[a];
    cs.dt = controls.timeStep;
    cs.urf_mom = controls.momVelURF; cs.urf_p = controls.prePreURF; cs.urf_pu = controls.preVelURF;
    cs.limit = controls.maxU;
    cs.urf_vf = controls.vofVofURF;
    cs.n_vof = controls.iterVofMax;
    cs.adjust_dt = controls.adjustTimeStep == "yes" ? 1 : 0;
    cs.max_co = controls.maxCo; cs.max_dt = controls.maxTimeStep;
    vector<semo_patch_bc> bcs(mesh.boundary.size());
    bool function = false;
    for (size_t b = 0; b < bcs.size(); ++b) {
        const SEMO_Boundary& B = mesh.boundary[b];
        semo_patch_bc& p = bcs[b];
        p.p_type = p_type(B.type[controls.P]);
        p.u_type = u_type(B.type[controls.U]);
        p.vf_type = vf_type(B.type[controls.VF[0]]);
        if (p.p_type < 0 || p.u_type < 0 || p.vf_type < 0) fail("not defined B.C. on patch " + B.name);
        p.p_value = B.var[controls.P];
        p.u_value[0] = B.var[controls.U]; p.u_value[1] = B.var[controls.V]; p.u_value[2] = B.var[controls.W];
        p.vf_value = B.var[controls.VF[0]];
        function = function || p.vf_type == SEMO_BC_FUNCTION;
    }
    check(semo_set_case(g_ctx, &cs, bcs.data()));
    // turbulence (transport.cpp:38-72): LES models by name, laminar / none
    int les = 0;
    if (controls.turbType == "LES") {
        if (controls.turbLESModel == "WALE") les = SEMO_LES_WALE;
        else if (controls.turbLESModel == "smagorinsky") les = SEMO_LES_SMAGORINSKY;
        else if (controls.turbLESModel == "dynamicSmagorinsky") les = SEMO_LES_DYNAMIC_SMAGORINSKY;
        else fail("not defined LES model");
    } else if (controls.turbType != "laminar" && controls.turbType != "none") {
        fail(controls.turbType == "RANS" ? "not yet defined RANS model" : "not defined turb type");
    }
    check(semo_set_les(g_ctx, les));
    if (function) {
        vector<double> pts;
        vf_points(pts);
        check(semo_set_vf_function_points(g_ctx, (int32_t)(pts.size() / 4), pts.data()));
    }
    g_dt = controls.timeStep;
}

}  // namespace

void check(int rc) {
    if (rc == SEMO_OK) return;
    fail(g_ctx ? semo_ctx_last_error(g_ctx) : "status " + std::to_string(rc));
}

semo_ctx* context(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls, vector<SEMO_Species>& species) {
    if (!g_ctx) {
        const char* dev = std::getenv("SEMO_GPU_DEVICE");
        check(semo_ctx_create(dev ? std::atoi(dev) : 0, nullptr, &g_ctx));
        upload(mesh);
        set_case(mesh, controls, species);
        const char* pr = std::getenv("SEMO_GPU_PRECISION");
        const char* tl = std::getenv("SEMO_GPU_TOL");
        double tol = tl ? std::atof(tl) : 1e-9;
        semo_solver_params sp{};
        sp.precision = (pr && !std::strcmp(pr, "fp64")) ? SEMO_PREC_FP64 : SEMO_PREC_MIXED;
        sp.tol_p = tol; sp.tol_u = tol; sp.maxit_p = 1000; sp.maxit_u = 1000;
        check(semo_set_solver(g_ctx, &sp));
    }
    if (mesh.cells.size() != g_cells) fail("the mesh changed after the first call");
    if (controls.timeStep != g_dt) set_case(mesh, controls, species);
    return g_ctx;
}

void put(SEMO_Mesh_Builder& mesh, int field, int var) {
    vector<double> v(mesh.cells.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = mesh.cells[i].var[var];
    check(semo_field_set(g_ctx, field, v.data(), (int64_t)v.size()));
}

void get(SEMO_Mesh_Builder& mesh, int field, int var) {
    vector<double> v(mesh.cells.size());
    check(semo_field_get(g_ctx, field, v.data(), (int64_t)v.size()));
    for (size_t i = 0; i < v.size(); ++i) mesh.cells[i].var[var] = v[i];
}

namespace {
// what the face and cell assemblies read, as SEMO holds it at the call
void put_state(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& c) {
    const std::pair<int, int> f[] = {
        {SEMO_F_P, c.P}, {SEMO_F_U, c.U}, {SEMO_F_V, c.V}, {SEMO_F_W, c.W}, {SEMO_F_VF, c.VF[0]},
        {SEMO_F_RHO, c.Rho}, {SEMO_F_MU, c.mu}, {SEMO_F_OLDU, c.oldU}, {SEMO_F_OLDV, c.oldV},
        {SEMO_F_OLDW, c.oldW}, {SEMO_F_QM0, c.Qm[0]}, {SEMO_F_QM1, c.Qm[1]}, {SEMO_F_QM2, c.Qm[2]},
        {SEMO_F_QM3, c.Qm[3]}, {SEMO_F_OLDVF, c.oldVF[0]}, {SEMO_F_KAPPA, c.kappa},
        {SEMO_F_SRCG0, c.sourceGravity[0]}, {SEMO_F_SRCG1, c.sourceGravity[1]},
        {SEMO_F_SRCG2, c.sourceGravity[2]}, {SEMO_F_SRCS0, c.sourceSurfaceTension[0]},
        {SEMO_F_SRCS1, c.sourceSurfaceTension[1]}, {SEMO_F_SRCS2, c.sourceSurfaceTension[2]},
        {SEMO_F_C, c.C}};
    for (auto& q : f) put(mesh, q.first, q.second);
}
}  // namespace

}  // namespace semo_bridge

using namespace semo_bridge;

// src/solvers/eqMomentum.cpp:9: assemble, solve, under-relax and clip U, V, W;
// returns |dU| / |U|
double SEMO_Solvers_Builder::calcMomentumEqs(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls,
                                             vector<SEMO_Species>& species) {
    semo_ctx* ctx = context(mesh, controls, species);
    put_state(mesh, controls);
    double norm = 0.0;
    check(semo_calc_momentum_eqs(ctx, &norm, nullptr));
    get(mesh, SEMO_F_U, controls.U); get(mesh, SEMO_F_V, controls.V); get(mesh, SEMO_F_W, controls.W);
    return norm;
}

// src/solvers/eqPressure.cpp:8 (iterNonOrthogonality is unused there too):
// assemble, solve p', two Green-Gauss passes, correct P and U, V, W
double SEMO_Solvers_Builder::calcPressureEq(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls,
                                            vector<SEMO_Species>& species, int) {
    semo_ctx* ctx = context(mesh, controls, species);
    put_state(mesh, controls);
    double norm = 0.0;
    check(semo_calc_pressure_eq(ctx, &norm, nullptr));
    get(mesh, SEMO_F_P, controls.P);
    get(mesh, SEMO_F_U, controls.U); get(mesh, SEMO_F_V, controls.V); get(mesh, SEMO_F_W, controls.W);
    return norm;
}

// src/solvers/eqVolfrac.cpp:9: MSTACS face values, the VOF solve, VF update;
// returns sqrt(sum dalpha^2).  (The device call also runs the mixture EOS
// and transport, which SEMO's loop runs right after it: pressureBased.cpp:115.)
double SEMO_Solvers_Builder::calcVolfracEq(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls,
                                           vector<SEMO_Species>& species) {
    semo_ctx* ctx = context(mesh, controls, species);
    put_state(mesh, controls);
    double norm = 0.0;
    check(semo_calc_volfrac_eq(ctx, &norm, nullptr));
    get(mesh, SEMO_F_VF, controls.VF[0]);
    return norm;
}
