# The host side of libpsim_amd.so (C++17 model, JSON reader, C ABI) under AddressSanitizer, for the
# CPU tests (no GPU code runs there; device code is built as usual, its host stubs instrumented):
#   make -f asan.mk -C psim_amd/csrc
#   ASAN_OPTIONS=detect_leaks=0 LD_PRELOAD=$(clang's libclang_rt.asan-x86_64.so) PSIM_AMD_LIB=$PWD/psim_amd/_build/asan/libpsim_amd.so \
#       python -m pytest tests/test_host.py tests/test_surrogate_host.py tests/test_analysis.py -m "not gpu"
HIPCC ?= /opt/rocm/bin/hipcc
CLANGXX ?= /opt/rocm/llvm/bin/clang++
ARCH ?= gfx950
OUT = ../_build/asan
COMMON = -O1 -g -std=c++17 -fPIC -ffp-contract=off -I../../include -I.
HIPFLAGS = --offload-arch=$(ARCH) -Xclang -target-feature -Xclang -packed-fp32-ops

all: $(OUT)/libpsim_amd.so

$(OUT)/engine.o: engine.hip
	mkdir -p $(OUT)
	$(HIPCC) $(HIPFLAGS) $(COMMON) -Xarch_host -fsanitize=address -c engine.hip -o $@

$(OUT)/surrogate.o: surrogate.hip
	mkdir -p $(OUT)
	$(HIPCC) $(HIPFLAGS) $(COMMON) -Xarch_host -fsanitize=address -c surrogate.hip -o $@

$(OUT)/host_model.o: host_model.cpp
	mkdir -p $(OUT)
	$(CLANGXX) $(COMMON) -pthread -fsanitize=address -c host_model.cpp -o $@

$(OUT)/capi_host.o: capi_host.cpp
	mkdir -p $(OUT)
	$(CLANGXX) $(COMMON) -fsanitize=address -c capi_host.cpp -o $@

$(OUT)/libpsim_amd.so: $(OUT)/engine.o $(OUT)/surrogate.o $(OUT)/host_model.o $(OUT)/capi_host.o
	$(HIPCC) $(HIPFLAGS) -shared -pthread -Xarch_host -fsanitize=address -shared-libsan -o $@ $^

.PHONY: all
