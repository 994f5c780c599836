# Builds oracle/_ref/libhfemref.so: the reference's element path, compiled from
# its own sources under REF (read-only) plus the harness ref_harness.cpp, linked
# against the image's MKL for BLAS/LAPACK.  Test infrastructure only.
REF ?= /root/reference
MKLDIR ?= /opt/conda/lib
CXX ?= g++
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref/libhfemref.so
all: $(OUT)
$(OUT): $(HERE)ref_harness.cpp
	mkdir -p $(HERE)_ref
	$(CXX) -O2 -fPIC -shared -w -I$(REF) -o $@ $(HERE)ref_harness.cpp -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)
clean:
	rm -f $(OUT)
.PHONY: all clean
