# Compiles parts of the reference from their own sources (no stand-in headers
# are needed for these files) into oracle/_ref/:
#  * libref_field.so: prime_field.cpp + polynomial.cpp behind a C shim;
#  * libref_fftgkr.so: the fft_gkr program (src/VPD/fft_circuit_GKR.cpp,
#    main renamed) with the same two files, linked so that every
#    quadratic_poly::eval / linear_poly::eval call is recorded.
# Test infrastructure only; run in the build container where /root/reference
# exists.  Usage: make -f oracle/ref_field.mk
REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
FIELD := $(REF)/src/linear_gkr/prime_field.cpp $(REF)/src/linear_gkr/polynomial.cpp
QEVAL := _ZNK14quadratic_poly4evalERKN11prime_field13field_elementE
LEVAL := _ZNK11linear_poly4evalERKN11prime_field13field_elementE
all: $(OUT)/libref_field.so $(OUT)/libref_fftgkr.so
$(OUT)/libref_field.so: oracle/ref_field_shim.cpp $(FIELD)
	mkdir -p $(OUT)
	$(CXX) -O3 -mavx2 -mbmi2 -fPIC -shared -I$(REF)/include -o $@ oracle/ref_field_shim.cpp $(FIELD)
$(OUT)/libref_fftgkr.so: oracle/ref_fftgkr_shim.cpp $(REF)/src/VPD/fft_circuit_GKR.cpp $(FIELD)
	mkdir -p $(OUT)
	$(CXX) -O3 -mavx2 -mbmi2 -fPIC -shared -I$(REF)/include -I$(REF)/src -Wl,--wrap=$(QEVAL) -Wl,--wrap=$(LEVAL) \
		-o $@ oracle/ref_fftgkr_shim.cpp $(FIELD)
.PHONY: all
