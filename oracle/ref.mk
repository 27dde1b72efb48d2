# Compiles the reference's own C11 hot kernels, from the sources where they
# lie under /root/reference (never copied), into oracle/_ref/ for use as a
# checker of the oracle restatement.  Only src/common/c_kernels.c and its
# header are needed; nothing else of the reference is built.
REF ?= /root/reference
CC ?= gcc
CFLAGS = -O2 -std=c11 -fPIC -ffp-contract=off -fno-fast-math -I$(REF)/include

OUT = _ref/libref_ckernels.so

all: $(OUT)

$(OUT): $(REF)/src/common/c_kernels.c $(REF)/include/vibe/common/c_kernels.h
	mkdir -p _ref
	$(CC) $(CFLAGS) -shared -o $@ $(REF)/src/common/c_kernels.c -lm

clean:
	rm -f $(OUT)

# The reference's .vibebin writer (src/io/binary_writer.cpp, plus
# src/grid/geometry.cpp for grid::Grid), linked with this directory's
# ref_vibebin.cpp driver; writes the golden files of tests/golden/.
CXX ?= g++
CXXFLAGS = -O1 -std=c++20 -I$(REF)/include
VBIN = _ref/ref_vibebin

all: $(VBIN)

$(VBIN): ref_vibebin.cpp $(REF)/src/io/binary_writer.cpp $(REF)/src/grid/geometry.cpp
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -o $@ ref_vibebin.cpp $(REF)/src/io/binary_writer.cpp $(REF)/src/grid/geometry.cpp

# The reference's grid generators, sponge profile, isothermal reference state
# energy budget, CFL number and stability report (src/grid/geometry.cpp,
# src/dyn/damping.cpp, src/dyn/reference_state.cpp, src/dyn/diagnostics.cpp,
# src/dyn/state.cpp, src/dyn/advection.cpp, src/time/timestep_control.cpp) and
# nest operators (src/grid/nest.cpp)
# with this directory's ref_setup.cpp driver; prints the golden JSON of
# tests/golden/setup_reference.json.
SETUP = _ref/ref_setup

all: $(SETUP)

$(SETUP): ref_setup.cpp $(REF)/src/grid/geometry.cpp $(REF)/src/dyn/damping.cpp $(REF)/src/dyn/reference_state.cpp \
	  $(REF)/src/dyn/diagnostics.cpp $(REF)/src/dyn/state.cpp $(REF)/src/dyn/advection.cpp \
	  $(REF)/src/time/timestep_control.cpp $(REF)/src/grid/nest.cpp
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -o $@ ref_setup.cpp $(REF)/src/grid/geometry.cpp $(REF)/src/dyn/damping.cpp \
	  $(REF)/src/dyn/reference_state.cpp $(REF)/src/dyn/diagnostics.cpp $(REF)/src/dyn/state.cpp \
	  $(REF)/src/dyn/advection.cpp $(REF)/src/time/timestep_control.cpp $(REF)/src/grid/nest.cpp

# The reference's literal discrete equations (src/dyn/equations.cpp with
# src/dyn/advection.cpp in the same translation unit -- advection.hpp
# declares advection_from_string inline and advection.cpp defines it -- and
# the unmodified coriolis, damping, tendency, state, reference_state and
# geometry sources) with this directory's ref_equations.cpp driver; prints
# the golden arrays of tests/golden/equations_reference.npz.
EQNS = _ref/ref_equations

all: $(EQNS)

$(EQNS): ref_equations.cpp $(REF)/src/dyn/equations.cpp $(REF)/src/dyn/advection.cpp $(REF)/src/dyn/coriolis.cpp \
	  $(REF)/src/dyn/damping.cpp $(REF)/src/dyn/tendency.cpp $(REF)/src/dyn/state.cpp \
	  $(REF)/src/dyn/reference_state.cpp $(REF)/src/grid/geometry.cpp
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -ffp-contract=off -c -o _ref/eqadv.o -include $(REF)/src/dyn/advection.cpp $(REF)/src/dyn/equations.cpp
	$(CXX) $(CXXFLAGS) -ffp-contract=off -o $@ ref_equations.cpp _ref/eqadv.o $(REF)/src/dyn/coriolis.cpp \
	  $(REF)/src/dyn/damping.cpp $(REF)/src/dyn/tendency.cpp $(REF)/src/dyn/state.cpp \
	  $(REF)/src/dyn/reference_state.cpp $(REF)/src/grid/geometry.cpp
