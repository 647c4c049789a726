# AddressSanitizer build of libadsr.so's host code (CPU tests only): the
# host objects are instrumented by clang against the shared clang ASan
# runtime, which the test loads with LD_PRELOAD; device code is built as
# usual (the sanitizer applies to the host side of the .hip file only).
#   make -f asan.mk      ->  _build/asan/libadsr.so
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
ROCM ?= /opt/rocm
HIPCC ?= $(ROCM)/bin/hipcc
CLANGXX ?= $(ROCM)/llvm/bin/clang++
ARCH ?= gfx950
BUILD := $(HERE)_build/asan
OUT := $(BUILD)/libadsr.so

CXXFLAGS = -O1 -g -std=c++17 -fPIC -march=x86-64-v3 -ffp-contract=off -fno-fast-math \
           -fno-omit-frame-pointer -fsanitize=address -shared-libasan
HIPFLAGS = -O1 -g -std=c++17 -fPIC --offload-arch=$(ARCH) -ffp-contract=off -fno-fast-math \
           -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result \
           -Xarch_host -fno-omit-frame-pointer -Xarch_host -fsanitize=address \
           -Xarch_host -shared-libasan

all: $(OUT)

$(BUILD)/adsr_host.o: $(HERE)adsr_host.cpp $(HERE)adsr_internal.hpp $(HERE)../../include/adsr.h
	@mkdir -p $(BUILD)
	$(CLANGXX) $(CXXFLAGS) -c $< -o $@

$(BUILD)/adsr_device.o: $(HERE)adsr_device.hip $(HERE)adsr_internal.hpp $(HERE)adsr_logf.hpp $(HERE)../../include/adsr.h
	@mkdir -p $(BUILD)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(OUT): $(BUILD)/adsr_host.o $(BUILD)/adsr_device.o
	$(HIPCC) -shared -fPIC --offload-arch=$(ARCH) -Xarch_host -fsanitize=address \
	    -Xarch_host -shared-libasan -o $@ $^ -lpthread

.PHONY: all
