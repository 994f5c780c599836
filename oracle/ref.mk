# Builds oracle/_ref/libhfemref.so: the reference's element path, compiled from
# its own sources under REF (read-only) plus the harness ref_harness.cpp, linked
# against the image's MKL for BLAS/LAPACK.  Test infrastructure only.
REF ?= /root/reference
MKLDIR ?= /opt/conda/lib
CXX ?= g++
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref/libhfemref.so
IO := $(HERE)_ref/libhfemrefio.so
all: $(OUT) $(IO)
$(OUT): $(HERE)ref_harness.cpp
	mkdir -p $(HERE)_ref
	$(CXX) -O2 -fPIC -shared -w -fopenmp -I$(REF) -o $@ $(HERE)ref_harness.cpp -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)
# the reference's binary readers and writers (readbinaryfiles.cpp, ioutilities.cpp)
$(IO): $(HERE)ref_io.cpp
	mkdir -p $(HERE)_ref
	$(CXX) -O2 -fPIC -shared -w -I$(REF) -o $@ $(HERE)ref_io.cpp -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)
clean:
	rm -f $(OUT) $(IO)
.PHONY: all clean
