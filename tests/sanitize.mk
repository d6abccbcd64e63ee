# Sanitized CPU builds for tests/test_sanitized.py (test infrastructure; no
# GPU needed to build or run them):
#   liboracle_sw.so  the FP64 oracle with AddressSanitizer + UBSan (gcc);
#   libanuga_sw.so   the library with AddressSanitizer + UBSan on its host
#                    code only (mesh validation and packing, neighbour words,
#                    geometry checks, Hilbert order); the GPU code is built as
#                    usual and not instrumented.
# Usage: make -f tests/sanitize.mk OUT=<dir> <dir>/liboracle_sw.so <dir>/libanuga_sw.so
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT ?= /tmp/anuga_sanitized
CC ?= gcc
HIPCC ?= /opt/rocm/bin/hipcc
CSRC := $(ROOT)/anuga_cuda_amd/csrc

$(OUT)/liboracle_sw.so: $(ROOT)/oracle/sw_oracle.c $(ROOT)/oracle/sw_oracle.h
	mkdir -p $(OUT)
	$(CC) -O1 -g -fPIC -shared -std=c11 -ffp-contract=off -fno-fast-math -fno-omit-frame-pointer \
	    -fsanitize=address,undefined -fno-sanitize-recover=undefined -o $@ $(ROOT)/oracle/sw_oracle.c -lm

$(OUT)/libanuga_sw.so: $(CSRC)/sw_capi.hip $(CSRC)/mesh_build.hip $(CSRC)/sw_kernels.h $(CSRC)/sw_device.h \
                       $(CSRC)/sw_step_vertex.h $(ROOT)/include/anuga_sw.h
	mkdir -p $(OUT)
	$(HIPCC) --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -shared -ffp-contract=off -fno-fast-math \
	    -Xclang -target-feature -Xclang -packed-fp32-ops \
	    -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
	    -Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer -shared-libsan \
	    -o $@ $(CSRC)/sw_capi.hip $(CSRC)/mesh_build.hip
