// The whole-step seam: SEMO_Solvers_Builder::incompressiblePressureBased
// (src/solvers/pressureBased.cpp:6-483) as one call of
// semo_incompressible_pressure_based, the fields resident on the device
// between steps.  Linked in place of src/solvers/pressureBased.cpp (with
// gpuBridge.cpp in place of the equation files).  The first call uploads
// the state SEMO holds (P, U, V, W, VF); after every step the fields SEMO's
// output and diagnostics read (P, U, V, W, VF, Rho, mu) and, with
// adjustTimeStep, the step's time step come back.  SEMO code that changes
// cell values between steps (AMR, user functions) is outside this seam.
#include <vector>

#include "gpuBridge.h"
#include "solvers/build.h"

using std::vector;
using namespace semo_bridge;

namespace {
bool g_loaded = false;
}

void SEMO_Solvers_Builder::incompressiblePressureBased(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls,
                                                       vector<SEMO_Species>& species) {
    semo_ctx* ctx = context(mesh, controls, species);
    if (!g_loaded) {
        put(mesh, SEMO_F_P, controls.P); put(mesh, SEMO_F_U, controls.U); put(mesh, SEMO_F_V, controls.V);
        put(mesh, SEMO_F_W, controls.W); put(mesh, SEMO_F_VF, controls.VF[0]);
        g_loaded = true;
    }
    vector<double> norms(2 * (size_t)controls.iterPBsMax);
    check(semo_incompressible_pressure_based(ctx, controls.iterPBsMax, norms.data()));
    get(mesh, SEMO_F_P, controls.P); get(mesh, SEMO_F_U, controls.U); get(mesh, SEMO_F_V, controls.V);
    get(mesh, SEMO_F_W, controls.W); get(mesh, SEMO_F_VF, controls.VF[0]);
    get(mesh, SEMO_F_RHO, controls.Rho); get(mesh, SEMO_F_MU, controls.mu);
    if (controls.adjustTimeStep == "yes") {
        double dt = 0.0, co = 0.0;
        check(semo_calc_time_step(ctx, 0, &dt, &co));      // the step's time step, unchanged
        controls.timeStep = dt;
    }
}
