# Builds oracle/_ref/libref.so from the curtree sources where they lie
# (REF, default /root/reference).  Only the reference's CPU code builds here:
# src/utility.cpp (readMBR) and src/utility.h (nested_loop_query).  The CUDA
# files need nvcc and CUDA Thrust and are not built.
REF ?= /root/reference
HIPCC ?= /opt/rocm/bin/hipcc

all: _ref/libref.so

_ref/libref.so: ref_wrap.cpp $(REF)/src/utility.cpp $(REF)/src/utility.h
	@mkdir -p _ref
	g++ -O2 -fPIC -c $(REF)/src/utility.cpp -o _ref/utility.o
	$(HIPCC) -O2 -fPIC --offload-arch=gfx950 -I$(REF)/src -c ref_wrap.cpp -o _ref/ref_wrap.o
	$(HIPCC) -shared -fPIC --offload-arch=gfx950 -o $@ _ref/ref_wrap.o _ref/utility.o

clean:
	rm -rf _ref
