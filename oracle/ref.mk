# Builds oracle/_ref/libtcr_ref.so: the reference's own CPU ops
# (torchCompactRadius 0.5.5, src/torchCompactRadius/cppSrc/
# neighborhoodDynamic_cpu.cpp and hashing_cpu.cpp, compiled where they lie)
# plus oracle/ref_driver.cpp, against the installed torch headers and
# libraries.  Test infrastructure only: the output stays in oracle/_ref/.
# Flags follow the reference's CppExtension (setup.py:127-137): -O2, OpenMP,
# the limited Python API, no -march (so no FMA contraction is possible).
REF ?= /root/reference/src/torchCompactRadius/cppSrc
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref
PY ?= python3
TORCH_INC := $(shell $(PY) -c "from torch.utils.cpp_extension import include_paths; print(' '.join('-I'+p for p in include_paths()))")
TORCH_LIB := $(shell $(PY) -c "from torch.utils.cpp_extension import library_paths; print(library_paths()[0])")
CXX11_ABI := $(shell $(PY) -c "import torch; print(int(torch._C._GLIBCXX_USE_CXX11_ABI))")
PY_INC := $(shell $(PY) -c "import sysconfig; print(sysconfig.get_paths()['include'])")
CXXFLAGS = -O2 -std=c++17 -fPIC -fopenmp -Wno-sign-compare -Wno-unused-variable -DPy_LIMITED_API=0x03090000 \
           -D_GLIBCXX_USE_CXX11_ABI=$(CXX11_ABI) -DTORCH_EXTENSION_NAME=torchCompactRadius_cpu \
           -DTORCH_API_INCLUDE_EXTENSION_H -I$(REF) $(TORCH_INC) -I$(PY_INC)

all: $(OUT)/libtcr_ref.so

$(OUT)/neighborhoodDynamic_cpu.o: $(REF)/neighborhoodDynamic_cpu.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/hashing_cpu.o: $(REF)/hashing_cpu.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/ref_driver.o: $(HERE)ref_driver.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libtcr_ref.so: $(OUT)/neighborhoodDynamic_cpu.o $(OUT)/hashing_cpu.o $(OUT)/ref_driver.o
	$(CXX) -shared -fopenmp -o $@ $^ -L$(TORCH_LIB) -Wl,-rpath,$(TORCH_LIB) -lc10 -ltorch_cpu -ltorch

.PHONY: all
