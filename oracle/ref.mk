# Builds the reference SEMO hot path in place from the reference checkout
# (never copied) plus our own driver (ref_driver.cpp) into oracle/_ref/.
#
#   make -f oracle/ref.mk REF=/root/reference -j8
#
# Toolchain notes (recorded so the build is reproducible):
#  * MPI: MPICH 3.3.2 headers and shared libraries that ship in this image under
#    /opt/conda (the reference uses the MPI C++ bindings).  The libraries are
#    linked by path, without -L/opt/conda/lib, whose old libstdc++ breaks the link;
#    run-time symlinks to the image's libmpi/libmpicxx/libgfortran/libquadmath go
#    into oracle/_ref/lib, which the driver's rpath names.
#  * zlib: the system libz of this image (the reference's prebuilt lib/zlib
#    binaries are never used).
#  * math.cpp uses the Intel compiler's built-in __int64 type and a redeclared
#    default argument; g++ accepts both with -D__int64="long long" -fpermissive.
#  * Every solveAMGCL call is intercepted with -Wl,--wrap (see ref_driver.cpp).
#  * semo_initial: SEMO's initial-field utility (src/utility/initial.cpp) built
#    as it is, for the region-initialiser fixtures of tests/golden;
#    semo_compressvf likewise from src/utility/compressVolumeFraction.cpp.
#  * semo_gpu_eq / semo_gpu_step: the same driver and reference objects with
#    the C++ bridge of semo_mesh_amd/bridge linked in place of the equation
#    files (eq: SEMO's own incompressiblePressureBased, its equation calls on
#    the GPU) or of those and pressureBased.cpp too (step: the whole step on
#    the GPU, fields resident), against semo_mesh_amd/libsemo_gpu.so.
REF ?= /root/reference
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref
OBJ := $(OUT)/obj
MPIDIR := /opt/conda
CXX := g++
CXXFLAGS := -std=c++17 -O2 -fpermissive -D__int64="long long" -w \
            -I$(MPIDIR)/include -I$(REF)/lib/amgcl -I$(REF)/lib -I$(REF)/lib/zlib -I$(REF)/src

SRCS := mesh/build mesh/geometric math/math math/gradient math/RCM mpi/build \
        controls/build load/load load/loadFiles load/loadOpenFoam load/loadBinary \
        load/loadCompress load/loadBase64 \
        solvers/reconIncom solvers/eqMomentum solvers/eqPressure solvers/eqVolfrac \
        solvers/solveAMGCL solvers/solvePETSc solvers/timestep solvers/eos solvers/transport \
        solvers/NVD solvers/cellVarMinMax solvers/surfaceNormalVelocity solvers/pressureBased \
        solvers/source/gravity solvers/source/surfaceTension solvers/source/curvature \
        turbulenceModels/LES solvers/flux solvers/linearSolver solvers/RHS solvers/source solvers/viscousFlux solvers/densityBased solvers/norm solvers/reconComp \
        save/save save/saveAscii save/saveBinary save/saveCompress save/saveBase64 save/saveMkdirs save/saveGnuplot save/saveCellData
OBJS := $(foreach s,$(SRCS),$(OBJ)/$(subst /,_,$(s)).o)

SOLVE1 := _ZN20SEMO_Solvers_Builder10solveAMGCLENSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEER17SEMO_Mesh_BuilderiRSt6vectorIiSaIiEESB_RS8_IdSaIdEESE_SE_
SOLVE3 := $(SOLVE1)SE_SE_SE_SE_

BRIDGE := $(HERE)../semo_mesh_amd/bridge
GPULIB := $(HERE)../semo_mesh_amd/libsemo_gpu.so
EQ_OBJS := $(OBJ)/solvers_eqMomentum.o $(OBJ)/solvers_eqPressure.o $(OBJ)/solvers_eqVolfrac.o
BOBJS_EQ := $(filter-out $(EQ_OBJS),$(OBJS))
BOBJS_STEP := $(filter-out $(OBJ)/solvers_pressureBased.o,$(BOBJS_EQ))
MPILIBS := $(MPIDIR)/lib/libmpicxx.so.12 $(MPIDIR)/lib/libmpi.so.12

all: $(OUT)/semo_ref $(OUT)/rcm_ref $(OUT)/lib/.stamp $(OUT)/semo_gpu_eq $(OUT)/semo_gpu_step $(OUT)/semo_initial \
     $(OUT)/semo_compressvf

$(OBJ)/.dir:
	mkdir -p $(OBJ) && touch $@

define objrule
$(OBJ)/$(subst /,_,$(1)).o: $(REF)/src/$(1).cpp $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -c $$< -o $$@
endef
$(foreach s,$(SRCS),$(eval $(call objrule,$(s))))

$(OBJ)/ref_driver.o: $(HERE)ref_driver.cpp $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/semo_ref: $(OBJS) $(OBJ)/ref_driver.o
	$(CXX) -o $@ $^ -Wl,--wrap=$(SOLVE1) -Wl,--wrap=$(SOLVE3) \
	  $(MPIDIR)/lib/libmpicxx.so.12 $(MPIDIR)/lib/libmpi.so.12 -lz \
	  -Wl,-rpath,'$$ORIGIN/lib' -Wl,--allow-shlib-undefined

$(OBJ)/gpuBridge.o: $(BRIDGE)/gpuBridge.cpp $(BRIDGE)/gpuBridge.h $(HERE)../include/semo_gpu.h $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -I$(HERE)../include -I$(BRIDGE) -c $< -o $@

$(OBJ)/gpuBridgeStep.o: $(BRIDGE)/gpuBridgeStep.cpp $(BRIDGE)/gpuBridge.h $(HERE)../include/semo_gpu.h $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -I$(HERE)../include -I$(BRIDGE) -c $< -o $@

$(OUT)/semo_gpu_eq: $(BOBJS_EQ) $(OBJ)/gpuBridge.o $(OBJ)/ref_driver.o $(GPULIB)
	$(CXX) -o $@ $(BOBJS_EQ) $(OBJ)/gpuBridge.o $(OBJ)/ref_driver.o -Wl,--wrap=$(SOLVE1) -Wl,--wrap=$(SOLVE3) \
	  $(GPULIB) $(MPILIBS) -lz -Wl,-rpath,'$$ORIGIN/lib:$$ORIGIN/../../semo_mesh_amd' -Wl,--allow-shlib-undefined

$(OUT)/semo_gpu_step: $(BOBJS_STEP) $(OBJ)/gpuBridge.o $(OBJ)/gpuBridgeStep.o $(OBJ)/ref_driver.o $(GPULIB)
	$(CXX) -o $@ $(BOBJS_STEP) $(OBJ)/gpuBridge.o $(OBJ)/gpuBridgeStep.o $(OBJ)/ref_driver.o \
	  -Wl,--wrap=$(SOLVE1) -Wl,--wrap=$(SOLVE3) \
	  $(GPULIB) $(MPILIBS) -lz -Wl,-rpath,'$$ORIGIN/lib:$$ORIGIN/../../semo_mesh_amd' -Wl,--allow-shlib-undefined

# SEMO's initial-field utility (src/utility/initial.cpp, its own main) as it
# is, for the region-initialiser fixtures
$(OBJ)/utility_initial.o: $(REF)/src/utility/initial.cpp $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/semo_initial: $(OBJS) $(OBJ)/utility_initial.o
	$(CXX) -o $@ $^ $(MPILIBS) -lz -ldl -Wl,-rpath,'$$ORIGIN/lib' -Wl,--allow-shlib-undefined

# SEMO's interface-compression utility (src/utility/compressVolumeFraction.cpp,
# its own main) as it is, for the compression fixtures
$(OBJ)/utility_compressVolumeFraction.o: $(REF)/src/utility/compressVolumeFraction.cpp $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/semo_compressvf: $(OBJS) $(OBJ)/utility_compressVolumeFraction.o
	$(CXX) -o $@ $^ $(MPILIBS) -lz -ldl -Wl,-rpath,'$$ORIGIN/lib' -Wl,--allow-shlib-undefined

$(OBJ)/rcm_ref_main.o: $(HERE)rcm_ref_main.cpp $(OBJ)/.dir
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/rcm_ref: $(OBJ)/math_RCM.o $(OBJ)/rcm_ref_main.o
	$(CXX) -o $@ $^ $(MPIDIR)/lib/libmpicxx.so.12 $(MPIDIR)/lib/libmpi.so.12 \
	  -Wl,-rpath,'$$ORIGIN/lib' -Wl,--allow-shlib-undefined

$(OUT)/lib/.stamp:
	mkdir -p $(OUT)/lib
	for l in libmpi.so.12 libmpicxx.so.12 libgfortran.so.4 libquadmath.so.0; do \
	  ln -sf $(MPIDIR)/lib/$$l $(OUT)/lib/$$l; done
	touch $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
