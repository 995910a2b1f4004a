# Compiles the reference's field arithmetic from its own sources (no stand-in
# headers are needed for these two files) into oracle/_ref/libref_field.so.
# Test infrastructure only; run in the build container where /root/reference
# exists.  Usage: make -f oracle/ref_field.mk
REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
all: $(OUT)/libref_field.so
$(OUT)/libref_field.so: oracle/ref_field_shim.cpp $(REF)/src/linear_gkr/prime_field.cpp $(REF)/src/linear_gkr/polynomial.cpp
	mkdir -p $(OUT)
	$(CXX) -O3 -mavx2 -mbmi2 -fPIC -shared -I$(REF)/include -o $@ oracle/ref_field_shim.cpp $(REF)/src/linear_gkr/prime_field.cpp $(REF)/src/linear_gkr/polynomial.cpp
.PHONY: all
