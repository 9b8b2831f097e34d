# Builds the reference's CPU-compilable pieces from their own sources into
# oracle/_ref/ (checker use only).  The reference's GPU hot path is CUDA and
# METIS-dependent and is not buildable here.
REF ?= /root/reference
OUT := _ref

all: $(OUT)/gpart_verify_csr_ref $(OUT)/capacity_ref

$(OUT):
	mkdir -p $(OUT)

$(OUT)/gpart_verify_csr_ref: $(REF)/tools/verify_csr.cpp | $(OUT)
	g++ -O2 -std=c++17 -fopenmp -o $@ $(REF)/tools/verify_csr.cpp

$(OUT)/capacity_ref: ref_capacity.cpp $(REF)/include/partition_capacity.hpp | $(OUT)
	g++ -O2 -std=c++17 -I$(REF)/include -o $@ ref_capacity.cpp
