# TEST INFRASTRUCTURE ONLY: builds the parts of the reference that compile
# from their own sources with g++ (no CUDA): the dimension-independent
# primitives in common/ and the reference's own known-answer test. Outputs
# go to oracle/_ref/ only; the sources are read where they lie.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CXXFLAGS := -std=c++17 -O2 -ffp-contract=off -fno-fast-math

all: $(OUT)/common_primitives_cpu_test $(OUT)/libref_common.so

$(OUT)/common_primitives_cpu_test: $(REF)/tests/common_primitives_cpu_test.cpp $(REF)/common/philox.cuh $(REF)/common/model_coefficients.cuh
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -I$(REF)/common -o $@ $<

$(OUT)/libref_common.so: $(HERE)ref_common_shim.cpp $(REF)/common/philox.cuh $(REF)/common/model_coefficients.cuh
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -fPIC -shared -I$(REF)/common -o $@ $<

.PHONY: all
