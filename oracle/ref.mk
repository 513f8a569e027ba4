# Builds the reference converter (HongYue1/KBSRC 1.3.0) from its own sources, in place under
# /root/reference, into oracle/_ref/. Only a checker and a CPU baseline: nothing in the product links it.
#   make -f oracle/ref.mk            (from the repository root)
# Same flags as the reference's portable preset (CMakeLists.txt:30-31,57-59): -ffp-contract=off everywhere,
# a baseline x86-64 engine (namespace base) and an x86-64-v3 engine (namespace avx2) picked at run time.
REF ?= /root/reference
OUT := oracle/_ref
CXX ?= g++
CXXFLAGS := -std=c++20 -O3 -DNDEBUG -ffp-contract=off -Wall -Wextra -Wno-ignored-attributes -pthread
INC := -I$(REF)/include -I$(REF)/src
ENGINE := resampler design lpc

BASE_OBJ := $(ENGINE:%=$(OUT)/obj/base_%.o)
AVX2_OBJ := $(ENGINE:%=$(OUT)/obj/avx2_%.o)
LIB_OBJ := $(OUT)/obj/dispatch.o $(OUT)/obj/parallel.o $(BASE_OBJ) $(AVX2_OBJ)

all: $(OUT)/kbsrc $(OUT)/kbsrc_fenv_driver

$(OUT)/obj/base_%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -DKBSRC_ISA=base -c $< -o $@
$(OUT)/obj/avx2_%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -DKBSRC_ISA=avx2 -march=x86-64-v3 -mtune=generic -c $< -o $@
$(OUT)/obj/dispatch.o: $(REF)/src/dispatch.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -DKBSRC_VERSION='"1.3.0"' -DKBSRC_ENGINE_AVX2 -c $< -o $@
$(OUT)/obj/parallel.o: $(REF)/src/parallel.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@
$(OUT)/obj/wav.o: $(REF)/src/io/wav.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@
$(OUT)/obj/main.o: $(REF)/app/main.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@
$(OUT)/obj/fenv_driver.o: $(REF)/tests/fenv_driver.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/kbsrc: $(OUT)/obj/main.o $(OUT)/obj/wav.o $(LIB_OBJ)
	$(CXX) $(CXXFLAGS) $^ -o $@
$(OUT)/kbsrc_fenv_driver: $(OUT)/obj/fenv_driver.o $(LIB_OBJ)
	$(CXX) $(CXXFLAGS) $^ -o $@

clean:
	rm -rf $(OUT)
.PHONY: all clean
