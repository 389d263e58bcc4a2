# Builds the reference's standalone least-squares solver (Ax_b.cpp) from its
# own sources under /root/reference, with a C entry point, into oracle/_ref/.
# Test infrastructure only; nothing is copied into this repository.
REF ?= /root/reference
CXX ?= g++
all: _ref/libref_axb.so
_ref/libref_axb.so: ref_axb.cpp $(REF)/Ax_b.cpp $(REF)/Ax_b.h
	mkdir -p _ref
	$(CXX) -O2 -std=c++11 -ffp-contract=off -fPIC -shared -w -I$(REF) $(REF)/Ax_b.cpp ref_axb.cpp -o $@
.PHONY: all
