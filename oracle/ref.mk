# TEST INFRASTRUCTURE ONLY: builds the parts of the reference that compile
# from their own sources with g++ (no CUDA): the dimension-independent
# primitives in common/, the reference's own known-answer test, and the
# host-only 3D headers (CPU one-step reference, Philox polarity draws, brick
# sizing, centre placement). Outputs go to oracle/_ref/ only; the sources are
# read where they lie.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CXXFLAGS := -std=c++17 -O2 -ffp-contract=off -fno-fast-math
REF3D_HDRS := $(REF)/include/pf3d/reference.hpp $(REF)/include/pf3d/rng.cuh \
              $(REF)/include/pf3d/params.cuh $(REF)/include/pf3d/palmieri_initializer.hpp \
              $(REF)/include/pf3d/initializer.hpp

all: $(OUT)/common_primitives_cpu_test $(OUT)/libref_common.so \
     $(OUT)/model_one_step_cpu_test $(OUT)/libref3d.so

$(OUT)/common_primitives_cpu_test: $(REF)/tests/common_primitives_cpu_test.cpp $(REF)/common/philox.cuh $(REF)/common/model_coefficients.cuh
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -I$(REF)/common -o $@ $<

$(OUT)/libref_common.so: $(HERE)ref_common_shim.cpp $(REF)/common/philox.cuh $(REF)/common/model_coefficients.cuh
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -fPIC -shared -I$(REF)/common -o $@ $<

$(OUT)/model_one_step_cpu_test: $(REF)/tests/3d/model_one_step_cpu_test.cpp $(REF)/include/pf3d/reference.hpp
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -I$(REF)/include -o $@ $<

$(OUT)/libref3d.so: $(HERE)ref3d_shim.cpp $(REF3D_HDRS)
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -fPIC -shared -I$(REF)/common -I$(REF)/include/pf3d -o $@ $<

.PHONY: all
