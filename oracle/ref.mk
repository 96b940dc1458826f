# Builds the reference KIWI (read-only sources under $(REF)) into oracle/_ref/ so that the oracle
# can be checked against the real thing.  Only plain g++ on the reference's own source files; the
# reference's CMake build is not used.  Outputs land in oracle/_ref/ only (git-ignored).
#
#   make -f oracle/ref.mk            -> oracle/_ref/kiwi (the reference CLI)
#                                       oracle/_ref/libkiwi_ref.so (reference kernel math + our C shim)
#                                       oracle/_ref/libkiwi_voro.so (the vendored voro++ + our C shim)
#                                       oracle/_ref/libkiwi_pts.so (the reference's library objects +
#                                       our C shim over its coverage-grid buffer)
#
# The outputs are git-ignored but travel with the working tree to the GPU machine.  There the GPU tests
# load libkiwi_voro.so (tessellation of their inputs) and libkiwi_pts.so (the reference's coverage
# buffer as the checker) on the host; libkiwi_hip, smoke() and bench.py never load anything from here.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CXXFLAGS ?= -O2 -std=c++17 -fopenmp -fPIC -w
INC      := -I$(REF)/include -I$(REF)/third_party -I$(REF)/third_party/voro/src
DEFS     := -DKIWI_VERSION=\"0.1.0\" -DKIWI_SOURCE_DIR=\"$(REF)\"

VORO_SRC := cell.cc common.cc container.cc unitcell.cc v_compute.cc c_loops.cc v_base.cc wall.cc \
            pre_container.cc container_prd.cc
KIWI_SRC := $(wildcard $(REF)/src/*.cpp)
MATH_SRC := $(REF)/src/kernel_integral.cpp $(REF)/src/kernel_integral_generic.cpp \
            $(REF)/src/kernel_spec_data.cpp

VORO_OBJ := $(patsubst %.cc,$(OUT)/obj/voro_%.o,$(VORO_SRC))
KIWI_OBJ := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/kiwi_%.o,$(KIWI_SRC))
MATH_OBJ := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/kiwi_%.o,$(MATH_SRC))

all: $(OUT)/kiwi $(OUT)/libkiwi_ref.so $(OUT)/libkiwi_voro.so $(OUT)/libkiwi_pts.so

$(OUT)/obj/voro_%.o: $(REF)/third_party/voro/src/%.cc
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/obj/kiwi_%.o: $(REF)/src/%.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) $(INC) $(DEFS) -c $< -o $@

$(OUT)/kiwi: $(KIWI_OBJ) $(VORO_OBJ)
	$(CXX) $(CXXFLAGS) $^ -o $@

$(OUT)/libkiwi_ref.so: oracle/ref_shim.cpp $(MATH_OBJ)
	$(CXX) $(CXXFLAGS) $(INC) -shared $^ -o $@

$(OUT)/libkiwi_voro.so: oracle/voro_shim.cpp $(VORO_OBJ)
	$(CXX) $(CXXFLAGS) $(INC) -shared $^ -o $@

$(OUT)/libkiwi_pts.so: oracle/pts_shim.cpp $(filter-out $(OUT)/obj/kiwi_main.o,$(KIWI_OBJ)) $(VORO_OBJ)
	$(CXX) $(CXXFLAGS) $(INC) $(DEFS) -shared $^ -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
