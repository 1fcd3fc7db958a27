// The C++ binding of libsemo_gpu into SEMO: shared helpers of the two seams
// (gpuBridge.cpp: SEMO's equation member functions; gpuBridgeStep.cpp: the
// whole incompressiblePressureBased step).  Compiled inside SEMO's tree
// (include paths: SEMO's src/ and this repository's include/).
#pragma once

#include <vector>

#include "controls/build.h"
#include "mesh/build.h"
#include "semo_gpu.h"
#include "species/species.h"

namespace semo_bridge {

// SEMO's error behaviour for a failed call: the message on cerr, MPI_Abort
void check(int rc);

// The library context of this process, created on first use: mesh uploaded
// once (RCM on, geometry and LSQ set-up on the device), case and solver set;
// the case is set again when SEMO's time step changed.
semo_ctx* context(SEMO_Mesh_Builder& mesh, SEMO_Controls_Builder& controls, std::vector<SEMO_Species>& species);

// cell.var[var] -> device field / device field -> cell.var[var]
void put(SEMO_Mesh_Builder& mesh, int field, int var);
void get(SEMO_Mesh_Builder& mesh, int field, int var);

}  // namespace semo_bridge
